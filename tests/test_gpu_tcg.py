"""Truncated conjugate gradient (the reference's polalg = 3: epolar3tcg1 and
epolar3tcg2, epolar3tcg.f:66-621) on the GPU against the oracle's
restatement, over every combination of order, preconditioner, guess and
peek step, in FP64 and in mixed precision."""
import itertools

import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


@pytest.mark.parametrize("order,prec,guess,peek", list(itertools.product([1, 2], [True, False], [True, False],
                                                                         [True, False])))
def test_tcg_matches_oracle_fp64(ctx, order, prec, guess, peek):
    s = thp.tiny_water(27, 20.3, 24)
    ref = O.induce_tcg(O.Oracle(s), order=order, prec=prec, guess=guess, peek=peek, omega=0.8)
    solver = thp.Solver(ctx, s, FP64)
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.set_tcg(order, prec, guess, peek, 0.8)
    mu, ep, st = solver.newinduce_pme2()
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    rms = _rms_debye(mu, ref["mu"])
    print(f"tcg{order} prec={prec} guess={guess} peek={peek}: erel {erel:.1e} rmsD {rms:.1e}", flush=True)
    assert st.status == thp.THP_OK and st.niter == order
    assert erel < 1e-11 and rms < 1e-9
    solver.close()


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_tcg_defaults_on_a_larger_box(ctx, order, prec):
    # the reference's defaults (preconditioner, no guess, peek with omega 1)
    s = thp.serving_system(1500, 0.0, seed=44)
    ref = O.induce_tcg(O.Oracle(s), order=order)
    solver = thp.Solver(ctx, s, prec)
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.set_tcg(order)
    mu, ep, st = solver.newinduce_pme2()
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    rms = _rms_debye(mu, ref["mu"])
    print(f"tcg{order} prec={prec}: erel {erel:.1e} rmsD {rms:.1e}", flush=True)
    assert st.status == thp.THP_OK and st.niter == order
    if prec == FP64:
        assert erel < 1e-11 and rms < 1e-9
    else:
        assert erel < 1e-6 and rms < 1e-5
    # a second solve of the same system gives the same result
    mu2, ep2, _ = solver.newinduce_pme2()
    assert ep2 == ep and np.array_equal(mu2, mu)
    solver.close()


def test_tcg_options_are_checked(ctx):
    s = thp.tiny_water(27, 20.3, 24)
    solver = thp.Solver(ctx, s, FP64)
    with pytest.raises(thp.ThpError):
        solver.set_tcg(3)
    with pytest.raises(thp.ThpError):
        solver.set_tcg(2, omega=float("nan"))
    # PCG again after a truncated solve
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.newinduce_pme2()
    solver.set_solver(thp.THP_POLALG_PCG)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    ref = O.Oracle(s).induce(poleps=1e-8)
    assert st.status == thp.THP_OK and abs(ep - ref["ep"]) < 1e-11 * abs(ref["ep"])
    solver.close()


@pytest.mark.parametrize("prec", [thp.THP_PREC_FP64, thp.THP_PREC_MIXED])
def test_tcg_omega_fit_matches_oracle(ctx, prec):
    # tcgomegafit (kpolar.f:1059-1326) on the device: the omega the oracle's
    # bisection reaches, then a solve with it within the bisection's
    # tolerance of the PCG energy
    from test_oracle_pins import _virial_case
    s, _, _ = _virial_case(0.4, nfft=16)   # p-scales 0.5: the peek step moves the energy
    w_ref, done_ref, e = O.tcg_omega_fit(O.Oracle(s), order=2, prec=True)
    solver = thp.Solver(ctx, s, prec)
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.set_tcg(order=2, precond=True, guess=False, peek=True, omega=1.0)
    w, done = solver.tcg_omega_fit()
    print(f"omega fit prec={prec}: device {w} ({done}) oracle {w_ref} ({done_ref})", flush=True)
    assert done == done_ref and done
    assert w == w_ref
    _, ep, _ = solver.newinduce_pme2()
    assert abs(ep - e["epcg"]) <= e["cvg"] * (1 + 1e-6)
    solver.set_solver(thp.THP_POLALG_PCG)
    with pytest.raises(thp.ThpError):
        solver.tcg_omega_fit()      # needs the truncated CG with the peek step
    solver.close()


@pytest.mark.parametrize("order,prec,guess,peek", list(itertools.product([1, 2], [True, False], [True, False],
                                                                         [True, False])))
def test_tcg_gradient_matches_oracle_fp64(ctx, order, prec, guess, peek):
    # epolar1tcg: the adjoint of the truncated CG with the gradient kernels'
    # permanent-field and dipole terms, against the oracle's (pinned by
    # differences of the truncated energy); local frames carry the torques
    from test_oracle_pins import _virial_case
    s, fr, _ = _virial_case(0.4)
    ep, grad = O.epolar1_tcg(s, fr, order, prec, guess, peek, 0.8)
    sf = s.with_multipoles(np.zeros((s.n, 13)))
    sf.frames = dict(axistype=fr[0], zaxis=fr[1], xaxis=fr[2], yaxis=fr[3], pole=fr[4])
    solver = thp.Solver(ctx, sf, FP64)
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.set_tcg(order, prec, guess, peek, 0.8)
    _, e_dev, st = solver.newinduce_pme2()
    assert st.status == thp.THP_OK and abs(e_dev - ep) < 1e-11 * abs(ep)
    dep = solver.epolar1()
    gerr = np.abs(dep - grad).max() / np.abs(grad).max()
    print(f"tcg{order} prec={prec} guess={guess} peek={peek}: gradient {gerr:.1e}", flush=True)
    assert gerr < 1e-9
    # the solve's dipoles are kept: the gradient again is the same
    assert np.array_equal(solver.epolar1(), dep)
    solver.close()


def test_tcg_gradient_mixed_and_refusals(ctx):
    from test_oracle_pins import _virial_case
    s, fr, _ = _virial_case(0.4)
    ep, grad = O.epolar1_tcg(s, fr)
    sf = s.with_multipoles(np.zeros((s.n, 13)))
    sf.frames = dict(axistype=fr[0], zaxis=fr[1], xaxis=fr[2], yaxis=fr[3], pole=fr[4])
    solver = thp.Solver(ctx, sf, MIXED)
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.set_tcg(2)
    solver.newinduce_pme2()
    dep = solver.epolar1()
    gerr = np.abs(dep - grad).max() / np.abs(grad).max()
    print(f"tcg2 mixed: gradient {gerr:.1e}", flush=True)
    assert gerr < 2e-5
    # no virial for the truncated CG, and the no-allocation entry refuses it
    with pytest.raises(thp.ThpError):
        solver.epolar1_vir()
    from test_gpu_boundary import _Hip
    hip = _Hip()
    dep_d = hip.alloc(24 * s.n)
    try:
        assert solver.lib.thp_epolar1_noalloc(solver.h, dep_d, None, None) == -1  # THP_ERR_ARG
        assert b"truncated-CG" in solver.lib.thp_last_error(ctx.h)
    finally:
        hip.free_all()
    solver.close()
