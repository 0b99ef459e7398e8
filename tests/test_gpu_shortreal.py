"""Short-range real-space solves (the reference's use_polarshortreal with
ewaldshortcut: newinduce_shortreal, newinduce_shortreal.f:16-610) on the GPU
against the oracle at the short cutoff: PCG and Jacobi/DIIS, FP64 and mixed,
the direct-field entry point, and the return to the full solve."""
import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED
SHORT = 5.0  # ewaldshortcut default (cutoffs.f:59)


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


def _systems():
    return {"water81": lambda: thp.tiny_water(27, 20.3, 24),
            "serving1500": lambda: thp.serving_system(1500, 0.0, seed=44)}


@pytest.mark.parametrize("name", ["water81", "serving1500"])
@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("jacobi", [False, True])
def test_short_range_solve_matches_oracle(ctx, name, prec, jacobi):
    s = _systems()[name]()
    ref = O.induce_shortreal(O.Oracle(s, cut=SHORT), poleps=1e-8, jacobi=jacobi)
    solver = thp.Solver(ctx, s, prec)
    if jacobi:
        solver.set_solver(thp.THP_POLALG_JACOBI_DIIS)
    solver.set_shortreal(SHORT)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    rms = _rms_debye(mu, ref["mu"])
    print(f"{name} prec={prec} jacobi={jacobi}: niter {st.niter}/{ref['niter']} erel {erel:.1e} rmsD {rms:.1e}",
          flush=True)
    assert st.status == thp.THP_OK
    if prec == FP64:
        assert st.niter == ref["niter"] and erel < 1e-11 and rms < 1e-9
    else:
        assert abs(st.niter - ref["niter"]) <= 1 and erel < 1e-6 and rms < 1e-5
    solver.close()


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-11), (MIXED, 2e-5)])
def test_short_range_direct_field_and_back_to_full(ctx, prec, tol):
    s = thp.tiny_water(27, 20.3, 24)
    short = O.Oracle(s, cut=SHORT)
    full = O.Oracle(s)
    solver = thp.Solver(ctx, s, prec)
    solver.set_shortreal(SHORT)
    ef = solver.efld0_direct()
    ref = short.efld0_direct()
    assert np.linalg.norm(ef - ref) < tol * np.linalg.norm(ref)
    solver.newinduce_pme2(poleps=1e-8)
    # the full solve again, as if the mode had never been set
    solver.set_shortreal(0.0)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    r = full.induce(poleps=1e-8)
    assert st.status == thp.THP_OK and abs(ep - r["ep"]) < (1e-11 if prec == FP64 else 1e-6) * abs(r["ep"])
    solver.close()


def test_short_range_options_are_checked(ctx):
    s = thp.tiny_water(27, 20.3, 24)
    solver = thp.Solver(ctx, s, FP64)
    with pytest.raises(thp.ThpError):
        solver.set_shortreal(s.cut + 1.0)
    with pytest.raises(thp.ThpError):
        solver.set_shortreal(-1.0)
    # the truncated CG's omega fit is for the full solve
    solver.set_solver(thp.THP_POLALG_TCG)
    solver.set_shortreal(SHORT)
    with pytest.raises(thp.ThpError):
        solver.tcg_omega_fit()
    solver.close()


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (MIXED, 2e-5)])
def test_short_range_gradient_and_virial_match_oracle(ctx, prec, tol):
    # epolar1c with use_polarshortreal (epolar1.f:1347-1356): the real-space
    # terms within the short cutoff and the frames' torques, after a
    # short-range solve; the oracle's form is pinned by differences of the
    # short-range energy
    import dataclasses
    from test_oracle_pins import _virial_case
    short = 3.8
    s, fr, xyz = _virial_case(0.4)
    ep, grad, res = O.epolar1_short(s, fr, short, poleps=1e-10)
    rp = O.rotpole(*xyz.T, *fr)
    s2 = dataclasses.replace(s, cut=short).with_multipoles(rp)
    D = O.epolar1_virial(s2, rp, res["mu"][:, 0], res["mu"][:, 1], fr, real_only=True)
    vref = 0.5 * (D + D.T)
    sf = s.with_multipoles(np.zeros((s.n, 13)))
    sf.frames = dict(axistype=fr[0], zaxis=fr[1], xaxis=fr[2], yaxis=fr[3], pole=fr[4])
    solver = thp.Solver(ctx, sf, prec)
    solver.set_shortreal(short)
    _, e_dev, st = solver.newinduce_pme2(poleps=1e-10)
    assert st.status == 0 and abs(e_dev - ep) < 1e-6 * abs(ep)
    dep, vir = solver.epolar1_vir()
    gerr = np.abs(dep - grad).max() / np.abs(grad).max()
    verr = np.abs(vir - vref).max() / np.abs(vref).max()
    print(f"short-range gradient prec={prec}: {gerr:.2e}, virial {verr:.2e}", flush=True)
    assert gerr < tol and verr < tol
    # the full range again: a short solve's dipoles have no full gradient
    solver.set_shortreal(0.0)
    with pytest.raises(thp.ThpError):
        solver.epolar1()
    solver.close()


@pytest.mark.parametrize("order,prec_tcg,guess,peek", [(1, True, False, False), (2, True, True, True),
                                                       (2, False, False, True), (1, False, True, True)])
def test_short_range_tcg_energy_and_gradient_match_oracle(ctx, order, prec_tcg, guess, peek):
    # polalgshort = 3 (epolar1.f:29-31): epolar1tcg1shortreal and
    # epolar1tcg2*shortreal, the truncated CG on the short-range operators
    # with its own gradient (the adjoint of the truncated steps), FP64 and
    # mixed; the oracle's gradient is pinned by differences of its energy
    from test_oracle_pins import _virial_case
    s, fr, _ = _virial_case(0.4)
    short = 3.8
    ep, grad = O.epolar1_tcg(s, fr, order, prec_tcg, guess, peek, 0.8, shortcut=short)
    sf = s.with_multipoles(np.zeros((s.n, 13)))
    sf.frames = dict(axistype=fr[0], zaxis=fr[1], xaxis=fr[2], yaxis=fr[3], pole=fr[4])
    for prec, etol, gtol in ((FP64, 1e-11, 1e-9), (MIXED, 1e-6, 2e-5)):
        solver = thp.Solver(ctx, sf, prec)
        solver.set_solver(thp.THP_POLALG_TCG)
        solver.set_tcg(order, prec_tcg, guess, peek, 0.8)
        solver.set_shortreal(short)
        _, e_dev, st = solver.newinduce_pme2()
        dep = solver.epolar1()
        erel = abs(e_dev - ep) / abs(ep)
        gerr = np.abs(dep - grad).max() / np.abs(grad).max()
        print(f"short tcg{order} prec_tcg={prec_tcg} guess={guess} peek={peek} prec={prec}: "
              f"erel {erel:.1e} gradient {gerr:.1e}", flush=True)
        assert st.status == thp.THP_OK and st.niter == order
        assert erel < etol and gerr < gtol
        solver.close()
