"""Induced-dipole predictor on the GPU against the oracle's restatement
(ulspred and the dipole history, newinduce_pme2.f:165-178, 253-264,
910-1010) along a short synthetic trajectory: every step's E_pol, dipoles
and iteration count, for the Gear, ASPC and least-squares predictors."""
import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED
METHODS = {"GEAR": thp.THP_PRED_GEAR, "ASPC": thp.THP_PRED_ASPC, "LSQR": thp.THP_PRED_LSQR}


def _trajectory(s, steps, seed=6, v=0.004, acc=0.0005):
    """atoms on smooth paths x(t) = x0 + v t + a t^2 / 2 (A per step); the
    largest move over the run stays inside the lists' skin"""
    rng = np.random.default_rng(seed)
    vel = rng.normal(size=(3, s.n)) * v
    ac = rng.normal(size=(3, s.n)) * acc
    for t in range(steps):
        yield tuple(c + vel[k] * t + 0.5 * ac[k] * t * t for k, c in enumerate((s.x, s.y, s.z)))


@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("method", ["GEAR", "ASPC", "LSQR"])
def test_predictor_along_a_trajectory_matches_oracle(ctx, method, prec):
    s = thp.serving_system(1500, 0.0, seed=44)
    solver = thp.Solver(ctx, s, prec)
    solver.set_predictor(METHODS[method])
    pred = O.Predictor(method)
    nits = []
    for t, (x, y, z) in enumerate(_trajectory(s, 10)):
        solver.update_coords(x, y, z)
        mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
        guess = pred.predict()
        assert (guess is None) == (t < O.MAXUALT)
        ref = O.Oracle(s.with_coords(x, y, z)).induce(poleps=1e-8, mu0=guess)
        pred.push(ref["mu"])
        erel = abs(ep - ref["ep"]) / abs(ref["ep"])
        rms = float(np.sqrt(np.mean((mu - ref["mu"]) ** 2)) * thp.DEBYE)
        print(f"{method} prec={prec} step {t}: niter {st.niter}/{ref['niter']} erel {erel:.1e} rmsD {rms:.1e}",
              flush=True)
        assert st.status == thp.THP_OK
        if prec == FP64:
            assert erel < 1e-11 and rms < 1e-9
            # the least-squares fit is ill-conditioned: its guess may move the
            # stopping iteration by one
            assert abs(st.niter - ref["niter"]) <= (1 if method == "LSQR" else 0)
        else:
            assert erel < 1e-6 and rms < 1e-5 and abs(st.niter - ref["niter"]) <= 1
        nits.append(st.niter)
    # predicted starts take fewer iterations than the cold start of step 0
    assert max(nits[O.MAXUALT:]) < nits[0]
    solver.close()


def test_predictor_history_restarts_after_reset(ctx):
    s = thp.serving_system(900, 0.0, seed=45)
    solver = thp.Solver(ctx, s, MIXED)
    solver.set_predictor(thp.THP_PRED_ASPC)
    steps = list(_trajectory(s, 9))
    for x, y, z in steps[:8]:
        solver.update_coords(x, y, z)
        solver.newinduce_pme2(poleps=1e-8)
    solver.update_coords(*steps[8])
    _, ep_pred, st_pred = solver.newinduce_pme2(poleps=1e-8)
    # a cleared history falls back to the cold guess alpha*E: the same
    # solve as a fresh system at these coordinates, up to the FP32 rounding
    # of a different atom order (the fresh system sorts the moved atoms)
    solver.set_predictor(thp.THP_PRED_ASPC)
    _, ep_cold, st_cold = solver.newinduce_pme2(poleps=1e-8)
    cold = thp.Solver(ctx, s.with_coords(*steps[8]), MIXED)
    _, ep_ref, st_ref = cold.newinduce_pme2(poleps=1e-8)
    assert abs(st_cold.niter - st_ref.niter) <= 1 and abs(ep_cold - ep_ref) <= 1e-7 * abs(ep_ref)
    assert st_pred.niter < st_cold.niter
    assert abs(ep_pred - ep_ref) <= 1e-6 * abs(ep_ref)
    with pytest.raises(thp.ThpError):
        solver.set_predictor(7)
    solver.close()
    cold.close()
