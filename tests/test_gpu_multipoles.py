"""New global-frame multipoles between solves, as an MD driver passes them
after rotpole each step: through thp_system_update_multipoles on one system
and with a serving request, against fresh systems and the oracle."""
import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


def _new_multipoles(s, seed=3):
    """dipoles scaled by 1.1 and a random symmetric traceless change of the
    quadrupoles (rpole layout q, d(3), Q(3x3) row-major)"""
    rng = np.random.default_rng(seed)
    rp = np.array(s.rpole, dtype=np.float64).copy()
    rp[:, 1:4] *= 1.1
    dq = rng.normal(0.0, 0.02, size=(s.n, 3, 3))
    dq = 0.5 * (dq + dq.transpose(0, 2, 1))
    dq -= np.eye(3)[None] * (np.trace(dq, axis1=1, axis2=2) / 3.0)[:, None, None]
    rp[:, 4:13] += dq.reshape(s.n, 9)
    return rp


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_update_multipoles_equals_a_fresh_system(ctx, prec):
    s = thp.serving_system(1500, 0.1, seed=46)
    rp2 = _new_multipoles(s)
    solver = thp.Solver(ctx, s, prec)
    solver.newinduce_pme2(poleps=1e-8)
    solver.update_multipoles(rp2)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    fresh = thp.Solver(ctx, s.with_multipoles(rp2), prec)
    mu_f, ep_f, st_f = fresh.newinduce_pme2(poleps=1e-8)
    assert st.niter == st_f.niter and ep == ep_f and np.array_equal(mu, mu_f)
    ref = O.Oracle(s.with_multipoles(rp2)).induce(poleps=1e-8)
    tol_e, tol_mu = (1e-11, 1e-9) if prec == FP64 else (1e-6, 1e-5)
    assert abs(ep - ref["ep"]) < tol_e * abs(ref["ep"]) and _rms_debye(mu, ref["mu"]) < tol_mu
    fresh.close()
    solver.close()


def test_serving_request_with_multipoles(ctx):
    from thp.serve import Server
    s = thp.serving_system(2000, 0.1, seed=47)
    rp2 = _new_multipoles(s, seed=4)
    rng = np.random.default_rng(5)
    x, y, z = (v + rng.normal(0.0, 0.01, s.n) for v in (s.x, s.y, s.z))
    with Server(ctx, max_systems=2, warm_start=0) as srv:
        sid = srv.add(s)
        assert srv.poll(srv.submit(sid)).status == thp.THP_OK
        r = srv.poll(srv.submit(sid, x, y, z, rpole=rp2))
        mu = srv.dipoles(sid)
        again = srv.poll(srv.submit(sid))  # unchanged coordinates and multipoles
    one = thp.Solver(ctx, s.with_coords(x, y, z).with_multipoles(rp2), MIXED)
    mu1, e1, _ = one.newinduce_pme2(poleps=1e-8)
    one.close()
    assert r.status == again.status == thp.THP_OK
    assert abs(r.ep - e1) <= 1e-7 * abs(e1) and _rms_debye(mu, mu1) < 1e-6
    assert abs(again.ep - r.ep) <= 1e-7 * abs(r.ep)
