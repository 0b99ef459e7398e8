"""Jacobi/DIIS solver (the reference's polalg = 2: inducejac_pme2,
newinduce_pme2.f:638-890, with diis.f) on the GPU against the oracle's
restatement: iteration counts, E_pol and dipoles, a run long enough to
restart the DIIS subspace, and agreement with the PCG solution."""
import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


def _systems():
    return {"water81": lambda: thp.tiny_water(27, 20.3, 24),
            "serving1500": lambda: thp.serving_system(1500, 0.0, seed=44)}


@pytest.mark.parametrize("name", ["water81", "serving1500"])
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_jacobi_diis_matches_oracle(ctx, name, prec):
    s = _systems()[name]()
    ref = O.induce_jacobi_diis(O.Oracle(s), poleps=1e-8)
    solver = thp.Solver(ctx, s, prec)
    solver.set_solver(thp.THP_POLALG_JACOBI_DIIS)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    rms = _rms_debye(mu, ref["mu"])
    print(f"{name} prec={prec}: niter {st.niter}/{ref['niter']} erel {erel:.1e} rmsD {rms:.1e}", flush=True)
    assert st.status == thp.THP_OK
    if prec == FP64:
        assert st.niter == ref["niter"] and erel < 1e-11 and rms < 1e-9
    else:
        assert abs(st.niter - ref["niter"]) <= 1 and erel < 1e-6 and rms < 1e-5
    # the same solution as the PCG, to the tolerance
    solver.set_solver(thp.THP_POLALG_PCG)
    mu_pcg, ep_pcg, _ = solver.newinduce_pme2(poleps=1e-8)
    assert abs(ep - ep_pcg) < 1e-6 * abs(ep_pcg) and _rms_debye(mu, mu_pcg) < 1e-5
    solver.close()


def test_jacobi_diis_subspace_restart(ctx):
    # 26 iterations at 1e-12: the subspace fills at iteration 25 and keeps
    # its ten newest vectors
    s = thp.serving_system(1500, 0.0, seed=44)
    ref = O.induce_jacobi_diis(O.Oracle(s), poleps=1e-12)
    assert ref["niter"] > O.NDISMX
    solver = thp.Solver(ctx, s, FP64)
    solver.set_solver(thp.THP_POLALG_JACOBI_DIIS)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-12)
    print(f"restart: niter {st.niter}/{ref['niter']} erel {abs(ep - ref['ep']) / abs(ref['ep']):.1e}", flush=True)
    assert st.status == thp.THP_OK and abs(st.niter - ref["niter"]) <= 1
    assert abs(ep - ref["ep"]) < 1e-11 * abs(ref["ep"]) and _rms_debye(mu, ref["mu"]) < 1e-9
    with pytest.raises(thp.ThpError):
        solver.set_solver(4)
    solver.close()
