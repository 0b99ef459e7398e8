"""Local-frame multipoles rotated on the device (rotpole + chkpole,
rotpole.f:24-339, chkpole.f:20-98) against the oracle's restatement: a
water box with every axis type and chiral sites, solved in FP64 and mixed;
the rotation follows coordinate updates; explicit global multipoles
replace the frames."""
import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


def _frames(s, seed=3):
    """AMOEBA-water-like frames (O bisector, H Z-then-X) and, on some
    molecules, Z-Bisect, 3-Fold, Z-Only and chiral Z-then-X sites whose y
    atom is the next molecule's oxygen; random local multipoles, symmetric
    under y -> -y except at chiral sites"""
    rng = np.random.default_rng(seed)
    n, nmol = s.n, s.n // 3
    axt = np.zeros(n, np.int32)
    fz, fx, fy = (np.full(n, -1, np.int32) for _ in range(3))
    for m in range(nmol):
        o, h1, h2 = 3 * m, 3 * m + 1, 3 * m + 2
        other = 3 * ((m + 1) % nmol)
        kind = m % 3
        axt[o] = [thp.THP_AXIS_BISECTOR, thp.THP_AXIS_Z_BISECT, thp.THP_AXIS_3_FOLD][kind]
        fz[o], fx[o] = h1, h2
        if kind:
            fy[o] = other
        axt[h1], fz[h1], fx[h1] = thp.THP_AXIS_Z_THEN_X, o, h2
        fy[h1] = other if m % 2 == 0 else -1
        if m % 4 == 3:
            axt[h2], fz[h2] = thp.THP_AXIS_Z_ONLY, o
        else:
            axt[h2], fz[h2], fx[h2] = thp.THP_AXIS_Z_THEN_X, o, h1
    pole = np.zeros((n, 13))
    for i in range(n):
        q = rng.normal(0, 0.1, (3, 3))
        q = q + q.T
        q -= np.trace(q) / 3 * np.eye(3)
        d = rng.normal(0, 0.1, 3)
        if axt[i] == thp.THP_AXIS_Z_ONLY:
            d[:2] = 0.0
            q = np.diag([-0.5, -0.5, 1.0]) * rng.normal(0, 0.1)
        if not (axt[i] == thp.THP_AXIS_Z_THEN_X and fy[i] >= 0):
            d[1] = 0.0
            q[0, 1] = q[1, 0] = q[1, 2] = q[2, 1] = 0.0
        pole[i] = np.concatenate([[s.rpole[i, 0]], d, q.reshape(-1)])
    return dict(axistype=axt, zaxis=fz, xaxis=fx, yaxis=fy, pole=pole)


def _oracle_system(s, fr):
    rp = O.rotpole(s.x, s.y, s.z, fr["axistype"], fr["zaxis"], fr["xaxis"], fr["yaxis"], fr["pole"])
    return s.with_multipoles(rp)


def _moved(s, seed=4):
    """every molecule turned a few degrees about its oxygen and shifted"""
    rng = np.random.default_rng(seed)
    xyz = np.stack([s.x, s.y, s.z], axis=1).copy()
    for m in range(s.n // 3):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        t = np.deg2rad(rng.uniform(-6, 6))
        k = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        r = np.eye(3) + np.sin(t) * k + (1 - np.cos(t)) * (k @ k)
        o = xyz[3 * m].copy()
        xyz[3 * m:3 * m + 3] = (xyz[3 * m:3 * m + 3] - o) @ r.T + o + rng.normal(0, 0.02, 3)
    return s.with_coords(*xyz.T)


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_frames_match_oracle_and_follow_the_coordinates(ctx, prec):
    s = thp.tiny_water(27, 20.3, 24)
    fr = _frames(s)
    ref = O.Oracle(_oracle_system(s, fr)).induce(poleps=1e-8)
    sf = s.with_multipoles(np.zeros((s.n, 13)))   # the device ignores rpole when frames are set
    sf.frames = fr
    solver = thp.Solver(ctx, sf, prec)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    tol_e, tol_mu = (1e-11, 1e-9) if prec == FP64 else (1e-6, 1e-5)
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    print(f"prec={prec}: erel {erel:.1e} rmsD {_rms_debye(mu, ref['mu']):.1e}", flush=True)
    assert st.status == thp.THP_OK and erel < tol_e and _rms_debye(mu, ref["mu"]) < tol_mu
    # the molecules turn: the device rotates the multipoles again
    s2 = _moved(s)
    ref2 = O.Oracle(_oracle_system(s2, fr)).induce(poleps=1e-8)
    solver.update_coords(s2.x, s2.y, s2.z)
    mu2, ep2, st2 = solver.newinduce_pme2(poleps=1e-8)
    erel2 = abs(ep2 - ref2["ep"]) / abs(ref2["ep"])
    assert st2.status == thp.THP_OK and erel2 < tol_e and _rms_debye(mu2, ref2["mu"]) < tol_mu
    assert abs(ep2 - ep) > 1e-3 * abs(ep)   # the turn changed the answer
    solver.close()


def test_global_multipoles_replace_the_frames(ctx):
    s = thp.tiny_water(27, 20.3, 24)
    fr = _frames(s)
    sf = s.with_multipoles(np.zeros((s.n, 13)))
    sf.frames = fr
    solver = thp.Solver(ctx, sf, FP64)
    g = _oracle_system(s, fr).rpole
    solver.update_multipoles(g)
    s2 = _moved(s)
    solver.update_coords(s2.x, s2.y, s2.z)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    ref = O.Oracle(s2.with_multipoles(g)).induce(poleps=1e-8)   # the old global multipoles, not rotated
    assert st.status == thp.THP_OK and abs(ep - ref["ep"]) < 1e-11 * abs(ref["ep"])
    solver.close()


def test_frames_are_checked(ctx):
    s = thp.tiny_water(27, 20.3, 24)
    solver = thp.Solver(ctx, s, FP64)
    fr = _frames(s)
    bad = dict(fr, zaxis=fr["zaxis"].copy())
    bad["zaxis"][0] = 0                      # the site itself
    with pytest.raises(thp.ThpError):
        solver.set_frames(**bad)
    bad = dict(fr, axistype=fr["axistype"].copy())
    bad["axistype"][0] = 9
    with pytest.raises(thp.ThpError):
        solver.set_frames(**bad)
    bad = dict(fr, yaxis=fr["yaxis"].copy(), axistype=fr["axistype"].copy())
    bad["axistype"][3] = thp.THP_AXIS_3_FOLD
    bad["yaxis"][3] = -1                     # 3-Fold needs its y atom
    with pytest.raises(thp.ThpError):
        solver.set_frames(**bad)
    solver.close()
