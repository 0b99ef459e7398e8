"""Pin the FP64 CPU oracle against mathematics it does not restate.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand, a wrong normalisation)
fails: closed forms, finite-difference identities, a brute-force neighbor
search, the textbook Ewald structure-factor sum, alpha-invariance of the
full Ewald field, a dense linear solve and the energy decomposition.
"""
import dataclasses
import math

import numpy as np
import pytest

import ewald_ref as E
import thp
from oracle import oracle as O


# ------------------------------------------------------------ B-splines
@pytest.mark.parametrize("w", [0.0, 1e-9, 0.1234, 0.5, 0.77, 1.0 - 1e-12])
def test_bsplgen_matches_cardinal_bspline(w):
    # bsplgen (pmestuff.f:67-156): values and 3 derivatives at the 5 nodes
    np.testing.assert_allclose(O.bsplgen(w), E.bspline_nodes(w), rtol=0, atol=1e-13)


def test_bsplgen_partition_of_unity_and_symmetry():
    for w in np.linspace(0, 0.999, 37):
        th = O.bsplgen(w)
        assert abs(th[:, 0].sum() - 1.0) < 1e-14       # sum of values = 1
        assert abs(th[:, 1].sum()) < 1e-14             # sum of first derivatives = 0
        th2 = O.bsplgen(1.0 - w) if w > 0 else None
        if th2 is not None:                            # mirror symmetry
            np.testing.assert_allclose(th[::-1, 0], th2[:, 0], atol=1e-14)


# ----------------------------------------------------------- PME moduli
@pytest.mark.parametrize("K", [18, 24, 64, 108])
def test_bsmod_raw_is_dft_modulus(K):
    # moduli/dftmod (kewald.f:672-770) before the zeta factor: |DFT of the
    # spline samples|^2; the samples M5(4..0) sit at indices 1..5
    samples = np.zeros(K)
    for j in range(1, 6):
        samples[j] = E.cardinal_bspline(5 - j, 5)
    ref = np.abs(np.fft.fft(samples)) ** 2
    got = O.bsmod(K, zeta=False)
    tiny = ref < 1e-7
    np.testing.assert_allclose(got[~tiny], ref[~tiny], rtol=1e-12)
    # Euler exponential-spline fix (kewald.f:768-772): vanishing moduli are
    # replaced by the mean of their neighbours (odd order, even K: Nyquist)
    for i in np.nonzero(tiny)[0]:
        assert abs(got[i] - 0.5 * (ref[i - 1] + ref[(i + 1) % K])) < 1e-14


@pytest.mark.parametrize("K", [24, 64])
def test_bsmod_zeta_factor_closed_form(K):
    raw, full = O.bsmod(K, zeta=False), O.bsmod(K)
    for i in range(K):
        k = i if i < K // 2 else i - K     # dftmod's wave index (kewald.f:786-788)
        if abs(k) == K // 2:
            continue                       # S(5) = 0 exactly at the Nyquist index
        z = E.zeta_closed(k, K, 5)
        assert abs(full[i] / raw[i] - z * z) < 1e-8 * z * z


# ------------------------------------------------------------------ FFT
def test_fft3d_matches_numpy():
    rng = np.random.default_rng(3)
    g = rng.normal(size=(6, 10, 8)) + 1j * rng.normal(size=(6, 10, 8))
    np.testing.assert_allclose(O.fft3d(g, -1), np.fft.fftn(g), rtol=0, atol=1e-11)
    np.testing.assert_allclose(O.fft3d(g, +1), np.fft.ifftn(g) * g.size, rtol=0, atol=1e-11)


# -------------------------------------------------------------- damping
@pytest.mark.parametrize("r", [0.9, 2.5, 4.0, 6.9])
def test_dampewald_recursion_is_the_radial_derivative(r):
    # dampewald (damping.f:17-56): bn_{k} = -(1/r) d/dr bn_{k-1}, bn_0 = erfc(a r)/r
    a = 0.4
    b = O.dampewald(a, r)
    assert abs(b[0] - math.erfc(a * r) / r) < 1e-15
    h = 1e-3 * r
    for k, idx in [(1, 2), (2, 4), (3, 6)]:
        prev = lambda x: O.dampewald(a, x)[2 * (k - 1)]  # noqa: E731
        d1 = (prev(r + h) - prev(r - h)) / (2 * h)
        d2 = (prev(r + h / 2) - prev(r - h / 2)) / h
        fd = -(4 * d2 - d1) / 3 / r  # Richardson: O(h^4)
        assert abs(b[idx] - fd) < 1e-8 * abs(b[idx])


def test_thole_lambdas_follow_the_field_hierarchy():
    # lambda_{k+2} = lambda_k - r/k d lambda_k/dr for a radially smeared
    # interaction; pins damptholed's lambda5/lambda7 (damping.f:142-215)
    pdi, pdk, a = 0.837 ** (1 / 6), 0.496 ** (1 / 6), 0.39
    for r in [0.6, 0.96, 1.5, 2.2, 3.0]:
        lam = lambda x: O.dampthole(pdi, pdk, a, x, 7)  # noqa: E731
        l3, l5, l7 = lam(r)[2], lam(r)[4], lam(r)[6]
        h = 1e-5
        d3 = (lam(r + h)[2] - lam(r - h)[2]) / (2 * h)
        d5 = (lam(r + h)[4] - lam(r - h)[4]) / (2 * h)
        assert abs(l5 - (l3 - r / 3 * d3)) < 1e-8
        assert abs(l7 - (l5 - r / 5 * d5)) < 1e-8
        # lambda3 is the charge of Thole's density rho(u) = 3a/(4 pi) exp(-a u^3)
        # inside the scaled radius u = r/(alpha_i alpha_k)^(1/6)
        u = r / (pdi * pdk)
        s = np.linspace(0, u, 20001)
        enc = np.trapezoid(4 * np.pi * s ** 2 * 3 * a / (4 * np.pi) * np.exp(-a * s ** 3), s)
        assert abs(l3 - enc) < 1e-7


# --------------------------------------------------------- neighbor list
def brute_pairs(s, mb):
    """O(N^2) enumeration with the reference's operation order:
    xr = xi - xk, orthogonal image, r2 = (x*x + y*y) + z*z, r2 <= mb^2."""
    x, y, z = s.x, s.y, s.z
    L = s.box
    out = []
    for i in range(s.n - 1):
        d = [x[i] - x[i + 1:], y[i] - y[i + 1:], z[i] - z[i + 1:]]
        for c in range(3):
            v = d[c]
            big = np.abs(v) > 0.5 * L[c]
            cel = np.copysign(L[c], v)
            v2 = np.where(big, v - cel * np.floor(np.abs(v) / L[c]), v)
            big2 = big & (np.abs(v2) > 0.5 * L[c])
            d[c] = np.where(big2, v2 - np.copysign(L[c], v2), v2)
        r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        ks = np.nonzero(r2 <= mb * mb)[0] + i + 1
        out.extend((i, k) for k in ks)
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def test_nblist_equals_brute_force_pair_set():
    s = thp.water_1944()
    ref = brute_pairs(s, s.cut + s.buffer)
    got = O.Oracle(s).nblist_pairs()
    assert len(got) == len(ref) == 891263
    assert np.array_equal(got, ref)


# ------------------------------------------------ Ewald: textbook reference
def _few_site_system(seed, K):
    """eight sites with charges, dipoles and quadrupoles, no exclusions"""
    return thp.make_system(8, 16.0, n_solute=8, blob_radius=4.0, seed=seed, nfft=(K, K, K),
                           cut=7.9, buffer=0.05)


def test_pme_recip_field_matches_structure_factor_sum():
    # efld0_recip (newinduce_pme.f:1671-1917) against the Ewald sum over
    # reciprocal vectors; the difference is PME interpolation error, so it
    # must be small on a fine grid and shrink fast under refinement (a sign
    # or factor error in any moment would not)
    err = {}
    for K in (48, 96):
        s = _few_site_system(5, K)
        cphi = O.Oracle(s).efld0_recip()
        pos = np.stack([s.x, s.y, s.z], 1)
        q, d, Q = E.rpole_to_parts(s.rpole)
        ref = E.recip_grad_multipoles(pos, q, d, Q, s.box, s.aewald, mmax=14)
        err[K] = np.abs(cphi[:, 1:4] - ref).max() / np.abs(ref).max()
    assert err[96] < 1e-5 and err[48] / err[96] > 8, err


def test_pme_recip_dipole_field_matches_structure_factor_sum():
    # tmatxbrecip (newinduce_pme.f:1919-2128), d and p grids separately
    rng = np.random.default_rng(2)
    err = {}
    for K in (48, 96):
        s = _few_site_system(7, K)
        mu = rng.normal(size=(s.n, 2, 3)) * 0.2 if K == 48 else mu  # noqa: F821
        got = O.Oracle(s).tmatxbrecip(mu)
        pos = np.stack([s.x, s.y, s.z], 1)
        zq, zQ = np.zeros(s.n), np.zeros((s.n, 3, 3))
        e = 0.0
        for rhs in range(2):
            ref = E.recip_grad_multipoles(pos, zq, mu[:, rhs], zQ, s.box, s.aewald, mmax=14)
            e = max(e, np.abs(got[:, rhs] - ref).max() / np.abs(ref).max())
        err[K] = e
    assert err[96] < 2e-6 and err[48] / err[96] > 8, err


def _alpha_case():
    # twelve sites in a 20 A box, cutoff 9.9 A, 64^3 grid (0.31 A spacing)
    return thp.make_system(12, 20.0, n_solute=12, blob_radius=5.0, seed=11, nfft=(64, 64, 64),
                           cut=9.9, buffer=0.05)


def test_total_permanent_field_is_alpha_invariant():
    # direct + reciprocal + self (newinduce_pme2.f:111-147) does not depend
    # on the Ewald coefficient; the residual spread is PME interpolation
    # error (~1e-4 here), while a wrong self term alone would move the
    # field by ~1e-1 of its scale between these alphas
    s = _alpha_case()
    fields = [O.Oracle(s, aewald=a).efld0_total() for a in (0.40, 0.45, 0.50)]
    scale = np.abs(fields[0]).max()
    for f in fields[1:]:
        assert np.abs(f - fields[0]).max() < 3e-4 * scale


def test_total_dipole_matvec_is_alpha_invariant():
    # tmatxb + tmatxbrecip - self (newinduce_pme2.f:406-424)
    s = _alpha_case()
    rng = np.random.default_rng(4)
    mu = rng.normal(size=(s.n, 2, 3)) * 0.3
    outs = [O.Oracle(s, aewald=a).matvec(mu) for a in (0.40, 0.45, 0.50)]
    scale = np.abs(outs[0]).max()
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() < 1e-5 * scale


# ------------------------------------------ scaled (excluded) pairs
# A scaled pair enters the direct sum as dmp_n = bn_n - (1 - s lambda_n) rr_n
# (newinduce_pme.f:1197-1230 for the d and p scales, :1564-1574 for u;
# kscalfactor.f:925-949), while the reciprocal sum and the self term carry
# every pair in full.  Two consequences that no line of the oracle states:
# (1) the total is independent of the Ewald coefficient for any s, and
# (2) the nearest image of a scaled pair adds s times its Thole-damped bare
# interaction to the field of all other images.  A one-water box, its field
# compared with a brute-force lattice sum over the images, checks (2); a
# water box with d, p, u scales below 1 checks (1) at three coefficients.

def _bare_field(R, q, d, Q, lam=(1.0, 1.0, 1.0)):
    """field at R = x_i - x_k of point multipoles at x_k (textbook; rpole's
    convention phi = q/r + d.R/r^3 + 3 R.Q.R/r^5), each rank's radial factor
    times the Thole factor lam = (lambda3, lambda5, lambda7)"""
    r2 = (R * R).sum(-1)
    r = np.sqrt(r2)
    rr3 = lam[0] / (r * r2)
    rr5 = lam[1] * 3.0 / (r * r2 * r2)
    rr7 = lam[2] * 15.0 / (r * r2 * r2 * r2)
    dR = (d * R).sum(-1)
    QR = np.einsum("...ab,...b->...a", Q, R)
    RQR = (QR * R).sum(-1)
    return (q * rr3 + dR * rr5 + RQR * rr7)[..., None] * R - rr3[..., None] * d - 2.0 * rr5[..., None] * QR


def _image_field(pos, rpole, L, nshell=10):
    """field at each site of every periodic image (cells n != 0) of all sites
    in a cubic box, summed over a sphere of cells, plus the tin-foil surface
    term 4 pi M / 3V that the Ewald sum implies (M = the cell's dipole)"""
    q, d, Q = E.rpole_to_parts(rpole)
    g = np.arange(-nshell, nshell + 1)
    cells = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    cells = cells[((cells * cells).sum(1) <= nshell * nshell) & np.any(cells != 0, axis=1)]
    out = np.zeros((len(pos), 3))
    for i in range(len(pos)):
        for k in range(len(pos)):
            R = pos[i] - pos[k] - cells * L
            out[i] += _bare_field(R, q[k], d[k], Q[k]).sum(0)
    M = (q[:, None] * pos).sum(0) + d.sum(0)
    return out + 4.0 * math.pi / (3.0 * L ** 3) * M


def _one_water(scales):
    """one water in a 24 A box (cutoff 9.9 A, 64^3 grid); its three pairs
    carry the given (d, p, u) scales"""
    s = thp.make_system(3, 24.0, seed=3, nfft=(64, 64, 64), cut=9.9, buffer=0.05)
    s.scaled_dpu = np.tile(np.asarray(scales, dtype=np.float64), (len(s.scaled_ik), 1))
    return s


def _home_field(s, rpole, order):
    """Thole-damped bare field of the other sites of the molecule (nearest
    image) at each site; lambda from the pinned damptholed"""
    pos = np.stack([s.x, s.y, s.z], 1)
    q, d, Q = E.rpole_to_parts(rpole)
    pd = s.polarity ** (1.0 / 6.0)
    out = np.zeros((s.n, 3))
    for i in range(s.n):
        for k in range(s.n):
            if k == i:
                continue
            R = pos[i] - pos[k]
            lam = O.dampthole(pd[i], pd[k], min(s.thole[i], s.thole[k]), np.linalg.norm(R), order)
            out[i] += _bare_field(R[None], q[k], d[k][None], Q[k][None], (lam[2], lam[4], lam[6] if order > 5 else 1.0))[0]
    return out


def _isolated_field_error(dsc, psc):
    """max |oracle total permanent field - (images + s * damped home field)|
    over both right-hand sides, and the home field's scale"""
    s = _one_water((dsc, psc, 1.0))
    pos = np.stack([s.x, s.y, s.z], 1)
    img = _image_field(pos, s.rpole, s.box[0])
    home = _home_field(s, s.rpole, 7)
    ef = O.Oracle(s).efld0_total()
    err = max(np.abs(ef[:, 0] - (img + dsc * home)).max(), np.abs(ef[:, 1] - (img + psc * home)).max())
    return err, np.abs(home).max()


def _isolated_matvec_error(usc):
    """the same for the dipole field of the matvec: mu/alpha - h equals the
    images' field plus u times the damped home field"""
    s = _one_water((0.0, 0.0, usc))
    pos = np.stack([s.x, s.y, s.z], 1)
    mu = np.random.default_rng(5).normal(size=(s.n, 2, 3)) * 0.2
    h = O.Oracle(s).matvec(mu)
    err = 0.0
    for rhs in range(2):
        rp = np.zeros((s.n, 13))
        rp[:, 1:4] = mu[:, rhs]
        want = _image_field(pos, rp, s.box[0]) + usc * _home_field(s, rp, 5)
        err = max(err, np.abs(mu[:, rhs] / s.polarity[:, None] - h[:, rhs] - want).max())
    return err


def _scaled_alpha_spread():
    """spread of the total permanent field and of the full matvec over three
    Ewald coefficients, for 27 waters (cutoff 9.9 A, 64^3 grid) whose
    intramolecular pairs carry d = 0, p = 0.5, u = 0.7; relative to scale"""
    s = thp.tiny_water(27, 20.3, 64, cut=9.9, buffer=0.05)
    s.scaled_dpu = np.tile([0.0, 0.5, 0.7], (len(s.scaled_ik), 1))
    mu = np.random.default_rng(4).normal(size=(s.n, 2, 3)) * 0.3
    f = [O.Oracle(s, aewald=a).efld0_total() for a in (0.40, 0.45, 0.50)]
    m = [O.Oracle(s, aewald=a).matvec(mu) for a in (0.40, 0.45, 0.50)]
    fs = max(np.abs(x - f[0]).max() for x in f[1:]) / np.abs(f[0]).max()
    ms = max(np.abs(x - m[0]).max() for x in m[1:]) / np.abs(m[0]).max()
    return fs, ms


# bars: PME interpolation error of the home pairs' smooth part at 0.375 A
# grid spacing (observed 2.4e-6 field, 6e-7 matvec, 5e-5 and 2e-6 spreads)
ISO_TOL = 1e-5
ALPHA_FIELD_TOL, ALPHA_MATVEC_TOL = 1.5e-4, 5e-6


def test_isolated_molecule_field_is_the_lattice_sum_of_its_images():
    # d = p = 0: the molecule's own pairs drop out of the total field
    err, scale = _isolated_field_error(0.0, 0.0)
    assert err < ISO_TOL and scale > 0.1, (err, scale)


def test_isolated_molecule_scaled_field_adds_the_scaled_thole_field():
    # d = 0.3 and p = 0.6 on the same pairs: each right-hand side adds its
    # own scale times the damped bare field
    err, _ = _isolated_field_error(0.3, 0.6)
    assert err < ISO_TOL, err


@pytest.mark.parametrize("usc", [0.0, 0.7])
def test_isolated_molecule_matvec_adds_the_u_scaled_dipole_field(usc):
    err = _isolated_matvec_error(usc)
    assert err < ISO_TOL, err


def test_total_field_and_matvec_are_alpha_invariant_with_scaled_pairs():
    fs, ms = _scaled_alpha_spread()
    assert fs < ALPHA_FIELD_TOL and ms < ALPHA_MATVEC_TOL, (fs, ms)


@pytest.mark.parametrize("form", [1, 2])
def test_scaled_pair_pins_catch_planted_mistakes(form):
    # form 1 scales the Ewald part by s as well; form 2 uses (1 - s) lambda
    # in place of 1 - s lambda.  Each must break one of the checks above by
    # a wide margin (the mistake is planted in the oracle, on the CPU).
    lib = O.load()
    lib.orc_plant_scaled_form(form)
    try:
        miss_iso = max(_isolated_field_error(0.0, 0.0)[0], _isolated_field_error(0.3, 0.6)[0],
                       _isolated_matvec_error(0.0), _isolated_matvec_error(0.7)) / ISO_TOL
        fs, ms = _scaled_alpha_spread()
        miss_alpha = max(fs / ALPHA_FIELD_TOL, ms / ALPHA_MATVEC_TOL)
    finally:
        lib.orc_plant_scaled_form(0)
    assert max(miss_iso, miss_alpha) > 30, (miss_iso, miss_alpha)
    if form == 1:
        assert miss_alpha > 30, miss_alpha   # not alpha-invariant
    else:
        assert miss_iso > 30, miss_iso       # alpha-invariant, wrong home pairs


# ----------------------------------------------------- solver and energy
def test_pcg_solution_matches_dense_solve():
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    n6 = 6 * s.n
    A = np.zeros((n6, n6))
    e = np.zeros(n6)
    for j in range(n6):
        e[:] = 0.0
        e[j] = 1.0
        A[:, j] = orc.matvec(e.reshape(s.n, 2, 3)).reshape(-1)
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()       # (alpha^-1 - T) symmetric
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0          # positive definite here
    res = orc.induce(poleps=1e-12, politer=200)
    mu = np.linalg.solve(A, res["ef"].reshape(-1)).reshape(s.n, 2, 3)
    assert np.abs(res["mu"] - mu).max() < 1e-9 * np.abs(mu).max()


def test_energy_identity_matches_pairwise_decomposition():
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    res = orc.induce(poleps=1e-10)
    er, erc, es = orc.epolar_decomp(res["mu"][:, 0])
    total = er + erc + es
    assert abs(total - res["ep"]) < 1e-10 * abs(res["ep"])


def test_cg_energy_functional_decreases():
    s = thp.water_1944()
    res = O.Oracle(s).induce(poleps=1e-8)
    e = res["ene_hist"][:, 0]
    assert np.all(np.diff(e) <= 1e-9 * np.abs(e).max())
    assert res["status"] == 0 and res["niter"] == 22


# ------------------------------------------------------------ dipole predictor
def test_ldlt_solve_matches_dense_solve():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 6):
        m = rng.normal(size=(n, n))
        a = m @ m.T + n * np.eye(n)
        b = rng.normal(size=n)
        np.testing.assert_allclose(O.ldlt_solve(a, b), np.linalg.solve(a, b), rtol=1e-12, atol=1e-14)


def _poly_history(degree, t, n=40, seed=4, h=0.25):
    """dipoles u(s) = sum_j a_j s^j with random vectors a_j at the steps
    s = (t-1) h .. (t-7) h (newest first), and u(t h)"""
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(degree + 1, n, 2, 3))
    u = lambda s: sum(a[j] * s ** j for j in range(degree + 1))
    return [u((t - k) * h) for k in range(1, O.MAXUALT + 1)], u(t * h)


def test_gear_predictor_is_exact_for_quintic_histories():
    # the binomial coefficients are the sixth finite difference: polynomials
    # of degree 5 are extrapolated exactly
    hist, target = _poly_history(5, 0.3)
    p = O.Predictor("GEAR")
    for h in reversed(hist):
        p.push(h)
    np.testing.assert_allclose(p.predict(), target, rtol=1e-9, atol=1e-9)
    assert O.Predictor("GEAR").predict() is None  # nothing stored yet


def test_aspc_predictor_is_consistent_and_exact_for_linear_histories():
    b = np.array(O.PRED_COEF["ASPC"])
    k = np.arange(1, O.MAXUALT + 1)
    assert abs(b.sum() - 1.0) < 1e-15 and abs((k * b).sum()) < 1e-14  # zeroth and first moments
    hist, target = _poly_history(1, 2.0)
    p = O.Predictor("ASPC")
    for h in reversed(hist):
        p.push(h)
    np.testing.assert_allclose(p.predict(), target, rtol=1e-12, atol=1e-12)


def test_lsqr_predictor_recovers_the_gear_recurrence():
    # a degree-5 history satisfies the sixth-difference recurrence exactly,
    # so the least-squares fit of the newest entry on the six older ones is
    # the Gear coefficient set
    hist, target = _poly_history(5, 1.0)
    p = O.Predictor("LSQR")
    for h in reversed(hist):
        p.push(h)
    np.testing.assert_allclose(p.coefficients()[:6], O.PRED_COEF["GEAR"][:6], rtol=1e-7)
    assert p.coefficients()[6] == 0.0
    np.testing.assert_allclose(p.predict(), target, rtol=1e-8, atol=1e-8)


# ------------------------------------------------------------ Jacobi / DIIS
def test_jacobi_diis_first_step_is_the_jacobi_iterate():
    # one vector: B = [[0, 1], [1, e.e]] gives c = (-e.e, 1), so the
    # extrapolation returns the Jacobi iterate itself
    rng = np.random.default_rng(8)
    n = 12
    T = rng.normal(size=(6 * n, 6 * n)) * 0.02
    T = -(T + T.T) / 2
    alpha = rng.uniform(0.5, 1.5, size=(n, 1, 1))
    ef = rng.normal(size=(n, 2, 3))
    hfun = lambda mu: (T @ mu.reshape(-1)).reshape(n, 2, 3)
    mu1, it, _, _ = O.jacobi_diis(hfun, ef, alpha, poleps=1e300, politer=5)
    assert it == 1
    np.testing.assert_allclose(mu1, alpha * (ef - hfun(alpha * ef)), rtol=1e-14, atol=1e-15)


def test_jacobi_diis_with_restarts_matches_dense_solve():
    # slowly converging Jacobi map (spectral radius near 1): the subspace
    # fills up and restarts several times on the way
    rng = np.random.default_rng(0)
    n = 30
    M = rng.normal(size=(6 * n, 6 * n)) * 0.05
    T = -(M + M.T) / 2
    alpha = np.ones((n, 1, 1))
    ef = rng.normal(size=(n, 2, 3))
    hfun = lambda mu: (T @ mu.reshape(-1)).reshape(n, 2, 3)
    mu, it, rr, restarts = O.jacobi_diis(hfun, ef, alpha, poleps=1e-12, politer=300)
    assert restarts >= 2 and rr < 1e-12
    ref = np.linalg.solve(np.eye(6 * n) + T, ef.reshape(-1)).reshape(n, 2, 3)
    assert np.abs(mu - ref).max() < 1e-10


def test_jacobi_diis_solution_matches_dense_solve():
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    n6 = 6 * s.n
    A = np.zeros((n6, n6))
    e = np.zeros(n6)
    for j in range(n6):
        e[:] = 0.0
        e[j] = 1.0
        A[:, j] = orc.matvec(e.reshape(s.n, 2, 3)).reshape(-1)
    mu = np.linalg.solve(A, orc.efld0_total().reshape(-1)).reshape(s.n, 2, 3)
    jd = O.induce_jacobi_diis(orc, poleps=1e-14, politer=200)
    # the stopping rule bounds the Jacobi increment (1e-14 e A), not the
    # error, which is larger by about 1 / (1 - spectral radius); the
    # oracle's threaded sums vary in the last bits from run to run, and with
    # them the iterate the rule stops at (errors of 2e-14 to 5e-13 seen)
    assert np.abs(jd["mu"] - mu).max() < 1e-9 * np.abs(mu).max()


# ------------------------------------------------- truncated conjugate gradient
def _dense(orc, n):
    A = np.zeros((6 * n, 6 * n))
    e = np.zeros(6 * n)
    for j in range(6 * n):
        e[:] = 0.0
        e[j] = 1.0
        A[:, j] = orc.matvec(e.reshape(n, 2, 3)).reshape(-1)
    return A


@pytest.mark.parametrize("order", [1, 2])
def test_tcg_with_guess_is_the_pcg_iterate(order):
    # from mu0 = alpha E with the diagonal preconditioner, TCG of order k is
    # the k-th iterate of the preconditioned conjugate gradient
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    cg = orc.induce(poleps=1e-300, politer=order)
    assert cg["niter"] == order
    t = O.induce_tcg(orc, order=order, prec=True, guess=True, peek=False)
    assert np.abs(t["mu"] - cg["mu"]).max() < 1e-12 * np.abs(cg["mu"]).max()
    assert abs(t["ep"] - cg["ep"]) < 1e-12 * abs(cg["ep"])


@pytest.mark.parametrize("order", [1, 2])
def test_tcg_is_exact_when_the_operator_has_order_distinct_eigenvalues(order):
    # CG reaches the solution in as many steps as the operator has distinct
    # eigenvalues: TCG1 solves lambda I, TCG2 a two-eigenvalue operator
    rng = np.random.default_rng(21 + order)
    n = 10
    Q, _ = np.linalg.qr(rng.normal(size=(3 * n, 3 * n)))
    lam = np.where(np.arange(3 * n) % 2 == 0, 1.7, 3.1) if order == 2 else np.full(3 * n, 2.3)
    A = (Q * lam) @ Q.T
    afun = lambda mu: np.einsum("ij,rj->ri", A, mu.transpose(1, 0, 2).reshape(2, -1)).reshape(2, n, 3).transpose(1, 0, 2)
    ef = rng.normal(size=(n, 2, 3))
    mu, _ = O.tcg(afun, ef, np.ones((n, 1, 1)), order=order, prec=False, guess=False, peek=False)
    ref = np.stack([np.linalg.solve(A, ef[:, r, :].reshape(-1)).reshape(n, 3) for r in range(2)], axis=1)
    assert np.abs(mu - ref).max() < 1e-12 * np.abs(ref).max()
    if order == 2:
        # one step is not enough for two eigenvalues
        one, _ = O.tcg(afun, ef, np.ones((n, 1, 1)), order=1, prec=False, guess=False, peek=False)
        assert np.abs(one - ref).max() > 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", [True, False])
def test_tcg_peek_is_a_scaled_step_along_the_preconditioned_residual(order, prec):
    # mu_peek - mu = omega alpha (E - A mu): the residual computed afresh
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    ef = orc.efld0_total()
    alpha = np.asarray(s.polarity, dtype=np.float64)[:, None, None]
    mu, _ = O.tcg(orc.matvec, ef, alpha, order=order, prec=prec, guess=True, peek=False)
    mup, _ = O.tcg(orc.matvec, ef, alpha, order=order, prec=prec, guess=True, peek=True, omega=0.7)
    step = 0.7 * alpha * (ef - orc.matvec(mu))
    assert np.abs((mup - mu) - step).max() < 1e-10 * np.abs(step).max()


# ------------------------------------------------- short-range real-space solve
def test_pcg_restatement_reproduces_the_oracle_pcg():
    # the numpy PCG used for the short-range solve, on the full operator, is
    # the C++ oracle's PCG (pinned above against a dense solve)
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s)
    ref = orc.induce(poleps=1e-10)
    ef = orc.efld0_total()
    alpha = np.asarray(s.polarity, dtype=np.float64)[:, None, None]
    mu, it, rr = O.pcg(orc.matvec, ef, alpha, alpha * ef, 1e-10, 500)
    assert it == ref["niter"]
    assert np.abs(mu - ref["mu"]).max() < 1e-12 * np.abs(ref["mu"]).max()


@pytest.mark.parametrize("jacobi", [False, True])
def test_short_range_solve_matches_dense_solve(jacobi):
    # at a 5 A cutoff without reciprocal or self terms: the solution of
    # (alpha^-1 - T_short) mu = E_short
    s = thp.tiny_water(27, 20.3, 24)
    orc = O.Oracle(s, cut=5.0)
    n6 = 6 * s.n
    A = np.zeros((n6, n6))
    e = np.zeros(n6)
    for j in range(n6):
        e[:] = 0.0
        e[j] = 1.0
        A[:, j] = orc.tmatxb_pme(e.reshape(s.n, 2, 3), dodiag=True).reshape(-1)
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    mu = np.linalg.solve(A, orc.efld0_direct().reshape(-1)).reshape(s.n, 2, 3)
    res = O.induce_shortreal(orc, poleps=1e-12, jacobi=jacobi)
    assert np.abs(res["mu"] - mu).max() < 1e-9 * np.abs(mu).max()
    # and it differs from the full solve
    full = O.Oracle(s).induce(poleps=1e-12)
    assert np.abs(full["mu"] - mu).max() > 1e-4 * np.abs(mu).max()


# ------------------------------------------------------ local frames (rotpole)
def _frame_case(seed=5, nmol=12):
    """waters with every axis type: O bisector or Z-Bisect or 3-Fold, H
    Z-then-X (some with a y atom, so chiral) or Z-Only; random local
    multipoles (axially symmetric for Z-Only)"""
    rng = np.random.default_rng(seed)
    xyz = []
    for m in range(nmol):
        o = rng.uniform(0, 12, 3)
        r = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        h1 = o + r @ np.array([0.9572, 0.0, 0.0])
        ang = np.deg2rad(104.52)
        h2 = o + r @ np.array([0.9572 * np.cos(ang), 0.9572 * np.sin(ang), 0.0])
        xyz += [o, h1, h2]
    xyz = np.array(xyz)
    n = len(xyz)
    axt = np.zeros(n, np.int32)
    fz, fx, fy = (np.full(n, -1, np.int32) for _ in range(3))
    pole = np.zeros((n, 13))
    for m in range(nmol):
        o, h1, h2 = 3 * m, 3 * m + 1, 3 * m + 2
        other = 3 * ((m + 1) % nmol)
        kind = m % 3
        axt[o], fz[o], fx[o] = [O.AXIS_BISECTOR, O.AXIS_Z_BISECT, O.AXIS_3_FOLD][kind], h1, h2
        if kind:
            fy[o] = other if kind == 2 else h2
            if kind == 1:
                fx[o], fy[o] = h2, other
        axt[h1], fz[h1], fx[h1] = O.AXIS_Z_THEN_X, o, h2
        fy[h1] = other if m % 2 == 0 else -1
        if m % 4 == 3:
            axt[h2], fz[h2] = O.AXIS_Z_ONLY, o
        else:
            axt[h2], fz[h2], fx[h2], fy[h2] = O.AXIS_Z_THEN_X, o, h1, other
    for i in range(n):
        q = rng.normal(0, 0.1, (3, 3))
        q = q + q.T
        q -= np.trace(q) / 3 * np.eye(3)
        d = rng.normal(0, 0.1, 3)
        if axt[i] == O.AXIS_Z_ONLY:
            d[:2] = 0.0
            q = np.diag([-0.5, -0.5, 1.0]) * rng.normal(0, 0.1)
        if not (axt[i] == O.AXIS_Z_THEN_X and fy[i] >= 0):
            # sites without a chirality check have parameters symmetric
            # under y -> -y, as force-field parameters do
            d[1] = 0.0
            q[0, 1] = q[1, 0] = q[1, 2] = q[2, 1] = 0.0
        pole[i] = np.concatenate([[rng.normal(0, 0.3)], d, q.reshape(-1)])
    return xyz, axt, fz, fx, fy, pole


def _transform(rp, m):
    out = rp.copy()
    out[:, 1:4] = rp[:, 1:4] @ m.T
    out[:, 4:13] = np.einsum("ij,njk,lk->nil", m, rp[:, 4:13].reshape(-1, 3, 3), m).reshape(-1, 9)
    return out


def test_rotpole_is_covariant_under_rigid_rotation():
    xyz, axt, fz, fx, fy, pole = _frame_case()
    rp = O.rotpole(*xyz.T, axt, fz, fx, fy, pole)
    r = np.linalg.qr(np.random.default_rng(9).normal(size=(3, 3)))[0]
    if np.linalg.det(r) < 0:
        r[:, 0] = -r[:, 0]
    xyz2 = xyz @ r.T + np.array([3.0, -1.0, 2.0])
    rp2 = O.rotpole(*xyz2.T, axt, fz, fx, fy, pole)
    np.testing.assert_allclose(rp2, _transform(rp, r), atol=1e-12)


def test_rotpole_frames_point_at_their_atoms():
    xyz, axt, fz, fx, fy, pole = _frame_case()
    for i in range(len(xyz)):
        a = O.rotmat(xyz, i, axt[i], fz[i], fx[i], fy[i])
        assert np.abs(a.T @ a - np.eye(3)).max() < 1e-13 and abs(np.linalg.det(a) - 1.0) < 1e-13
        if axt[i] == O.AXIS_Z_THEN_X:
            dz, dx = xyz[fz[i]] - xyz[i], xyz[fx[i]] - xyz[i]
            np.testing.assert_allclose(a[:, 2], dz / np.linalg.norm(dz), atol=1e-14)
            assert abs(np.dot(a[:, 1], dx)) < 1e-13 and np.dot(a[:, 0], dx) > 0   # x atom in the xz half-plane
        if axt[i] == O.AXIS_BISECTOR:
            u = [(xyz[k] - xyz[i]) / np.linalg.norm(xyz[k] - xyz[i]) for k in (fz[i], fx[i])]
            assert abs(np.dot(a[:, 2], u[0]) - np.dot(a[:, 2], u[1])) < 1e-13   # z bisects the two bonds


def test_rotpole_mirror_image_needs_the_chirality_check():
    # the mirror image of the geometry gives the mirror image of the global
    # multipoles only with chkpole's inversion of the y components
    xyz, axt, fz, fx, fy, pole = _frame_case()
    m = np.diag([-1.0, 1.0, 1.0])
    rp = O.rotpole(*xyz.T, axt, fz, fx, fy, pole)
    rpm = O.rotpole(*(xyz @ m.T).T, axt, fz, fx, fy, pole)
    np.testing.assert_allclose(rpm, _transform(rp, m), atol=1e-12)
    chiral = (axt == O.AXIS_Z_THEN_X) & (fy >= 0)
    assert chiral.sum() >= 4
    no_y = fy.copy()
    no_y[chiral] = -1   # the same frames without the check
    rpn = O.rotpole(*(xyz @ m.T).T, axt, fz, fx, no_y, pole)
    assert np.abs(rpn[chiral] - _transform(rp, m)[chiral]).max() > 1e-3


# ------------------------------------------ polarization energy gradient
# The energy gradient (epolar1.f: epreal1c, eprecip1 and the self and
# torque terms of epolar1c) rests on one identity.  With A = 1/alpha - T
# symmetric, u_d = A^-1 E_d and u_p = A^-1 E_p, the derivative of
# E_pol = -f/2 u_d . E_p along any coordinate is
#     dE_pol = -f/2 (u_p . dE_d + u_d . dE_p + u_p . dT u_d),
# i.e. the derivative of G = -f/2 (u_p . E_d + u_d . E_p + u_p . T u_d)
# with both dipole sets held fixed.  The check below differentiates both
# sides numerically on a small water box whose multipoles follow local
# frames (so the permanent fields move with the orientations as well), the
# left side re-solving the dipoles at every displaced geometry.
KCOUL = 332.063714  # electric, MOD_units.f:81


def _water_frames(s, seed=11):
    """O bisector (H1, H2), H Z-then-X (O, other H); random local multipoles
    symmetric under y -> -y (no site has a chirality check)"""
    rng = np.random.default_rng(seed)
    n = s.n
    axt = np.zeros(n, np.int32)
    fz, fx, fy = (np.full(n, -1, np.int32) for _ in range(3))
    pole = np.zeros((n, 13))
    for m in range(n // 3):
        o, h1, h2 = 3 * m, 3 * m + 1, 3 * m + 2
        axt[o], fz[o], fx[o] = O.AXIS_BISECTOR, h1, h2
        axt[h1], fz[h1], fx[h1] = O.AXIS_Z_THEN_X, o, h2
        axt[h2], fz[h2], fx[h2] = O.AXIS_Z_THEN_X, o, h1
    for i in range(n):
        q = rng.normal(0, 0.1, (3, 3))
        q = q + q.T
        q -= np.trace(q) / 3 * np.eye(3)
        q[0, 1] = q[1, 0] = q[1, 2] = q[2, 1] = 0.0
        d = rng.normal(0, 0.1, 3)
        d[1] = 0.0
        pole[i] = np.concatenate([[s.rpole[i, 0]], d, q.reshape(-1)])
    return axt, fz, fx, fy, pole


def _at(s, fr, xyz):
    rp = O.rotpole(*xyz.T, *fr)
    return O.Oracle(s.with_coords(*xyz.T).with_multipoles(rp))


def _epol(s, fr, xyz):
    return _at(s, fr, xyz).induce(poleps=1e-12)["ep"]


def _g_fixed(s, fr, xyz, ud, up):
    orc = _at(s, fr, xyz)
    ef = orc.efld0_total()
    alpha = np.asarray(s.polarity)[:, None]
    tud = ud / alpha - orc.matvec(np.stack([ud, ud], axis=1))[:, 0]   # T u_d
    return -0.5 * KCOUL * (np.sum(up * ef[:, 0]) + np.sum(ud * ef[:, 1]) + np.sum(up * tud))


def _d4(f, h):
    """four-point central difference"""
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def test_polarization_gradient_identity_with_rotating_frames():
    s = thp.tiny_water(27, 10.0, 12, seed=4)
    # intramolecular p-scales of 0.5 (d-scales stay 0) so that u_d, u_p and
    # E_d, E_p differ: exchanging d and p in G then changes its derivative
    dpu = np.array(s.scaled_dpu, dtype=np.float64).reshape(-1, 3).copy()
    dpu[:, 1] = 0.5
    s = dataclasses.replace(s, scaled_dpu=dpu)
    fr = _water_frames(s)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    ref = _at(s, fr, xyz).induce(poleps=1e-12)
    ud, up = ref["mu"][:, 0], ref["mu"][:, 1]
    ef = _at(s, fr, xyz).efld0_total()
    # the energy itself: E_pol = -f/2 u_d . E_p
    assert abs(ref["ep"] + 0.5 * KCOUL * np.sum(ud * ef[:, 1])) < 1e-10 * abs(ref["ep"])
    h = 2e-4
    # the real-space sums stop at the cutoff: an atom with a partner within
    # a few steps of it would see a jump, so the checked atoms keep clear
    box = s.box[0]

    def clear(a):
        d = xyz - xyz[a]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        r[a] = 0.0
        return np.min(np.abs(r - s.cut)) > 50 * h

    atoms = [a for a in range(s.n) if clear(a)]
    picks = [atoms[0], atoms[1], atoms[2], atoms[len(atoms) // 2], atoms[-2], atoms[-1]]
    assert len(set(picks)) == 6 and {p % 3 for p in picks} >= {0, 1, 2}
    for atom, axis in zip(picks, (0, 1, 2, 0, 2, 1)):
        def move(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return x2
        de = _d4(lambda dx: _epol(s, fr, move(dx)), h)
        dg = _d4(lambda dx: _g_fixed(s, fr, move(dx), ud, up), h)
        assert abs(de) > 1e-3                      # a real force component
        assert abs(de - dg) < 1e-8 * max(1.0, abs(de)), (atom, axis, de, dg)


def _rot(axis, ang):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)


def test_real_space_polarization_forces_and_torques_match_differences():
    # epolar1_real against central differences of the real-space part of G
    # (efld0_direct's fields and tmatxb's dipole field, dipoles and global
    # multipoles held fixed): forces along coordinates, torques along
    # rotations of one site's multipole about its centre
    s = thp.tiny_water(27, 10.0, 12, seed=4)
    dpu = np.array(s.scaled_dpu, dtype=np.float64).reshape(-1, 3).copy()
    dpu[:, 1] = 0.5
    s = dataclasses.replace(s, scaled_dpu=dpu)
    fr = _water_frames(s)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    rp = O.rotpole(*xyz.T, *fr)
    s = s.with_multipoles(rp)
    mu = O.Oracle(s).induce(poleps=1e-10)["mu"]
    ud, up = mu[:, 0], mu[:, 1]
    F, T = O.epolar1_real(s, rp, ud, up)

    def g_real(xyz2, rp2):
        orc = O.Oracle(s.with_coords(*xyz2.T).with_multipoles(rp2))
        ef = orc.efld0_direct()
        tud = -orc.tmatxb_pme(np.stack([ud, ud], axis=1), dodiag=False)[:, 0]
        return -0.5 * KCOUL * (np.sum(up * ef[:, 0]) + np.sum(ud * ef[:, 1]) + np.sum(up * tud))

    h = 2e-4
    box = s.box[0]

    def clear(a):
        d = xyz - xyz[a]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        r[a] = 0.0
        return np.min(np.abs(r - s.cut)) > 50 * h

    atoms = [a for a in range(s.n) if clear(a)]
    for atom, axis in zip([atoms[0], atoms[1], atoms[2], atoms[-1]], (0, 1, 2, 1)):
        def g_at(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return g_real(x2, rp)
        fd = -_d4(g_at, h)
        assert abs(F[atom, axis] - fd) < 1e-7 * max(1.0, abs(fd)), (atom, axis, F[atom, axis], fd)
    for atom, e in ((0, (1.0, 0.3, -0.2)), (4, (0.0, 1.0, 0.5)), (8, (-0.4, 0.1, 1.0))):
        def g_rot(th):
            m = _rot(e, th)
            rp2 = rp.copy()
            rp2[atom] = _transform(rp[atom:atom + 1], m)[0]
            return g_real(xyz, rp2)
        dg = _d4(g_rot, 1e-4)
        ee = np.asarray(e) / np.linalg.norm(e)
        assert abs(dg + T[atom] @ ee) < 1e-7 * max(1.0, abs(dg)), (atom, dg, -T[atom] @ ee)


def test_reciprocal_and_self_polarization_forces_and_torques_match_differences():
    # epolar1_recip_self against central differences of the reciprocal and
    # self parts of G (efld0_recip's field, tmatxbrecip's dipole field and
    # the self fields, dipoles and global multipoles held fixed); its grid
    # reproduces tmatxbrecip's first derivatives
    s = thp.tiny_water(27, 10.0, 12, seed=4)
    fr = _water_frames(s)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    rp = O.rotpole(*xyz.T, *fr)
    s = s.with_multipoles(rp)
    mu = O.Oracle(s).induce(poleps=1e-10)["mu"]
    ud, up = mu[:, 0], mu[:, 1]
    th, nodes = O._pme_stencils(s)
    D1 = O._pme_dipole_potential(s, ud, th, nodes, O._qfac(s))[0]
    np.testing.assert_allclose(D1, O.Oracle(s).tmatxbrecip(mu)[:, 0], rtol=0, atol=1e-12 * np.abs(D1).max())
    F, T = O.epolar1_recip_self(s, rp, ud, up)
    cself = 4.0 * s.aewald ** 3 / (3.0 * np.sqrt(np.pi))

    def g_rs(xyz2, rp2):
        orc = O.Oracle(s.with_coords(*xyz2.T).with_multipoles(rp2))
        e = -orc.efld0_recip()[:, 1:4] + cself * rp2[:, 1:4]
        tud = -orc.tmatxbrecip(np.stack([ud, ud], axis=1))[:, 0]
        return -0.5 * KCOUL * (np.sum(up * e) + np.sum(ud * e) + np.sum(up * tud))

    h = 1e-4
    for atom, axis in ((0, 0), (1, 1), (2, 2), (40, 0), (77, 2)):
        def g_at(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return g_rs(x2, rp)
        fd = -_d4(g_at, h)
        assert abs(F[atom, axis] - fd) < 1e-7 * max(1.0, abs(fd)), (atom, axis, F[atom, axis], fd)
    for atom, e in ((0, (1.0, 0.3, -0.2)), (4, (0.0, 1.0, 0.5)), (8, (-0.4, 0.1, 1.0))):
        def g_rot(t):
            rp2 = rp.copy()
            rp2[atom] = _transform(rp[atom:atom + 1], _rot(e, t))[0]
            return g_rs(xyz, rp2)
        dg = _d4(g_rot, 1e-4)
        ee = np.asarray(e) / np.linalg.norm(e)
        assert abs(dg + T[atom] @ ee) < 1e-7 * max(1.0, abs(dg)), (atom, dg, -T[atom] @ ee)


def test_polarization_gradient_matches_differences_of_the_solved_energy():
    # the whole oracle gradient (real, reciprocal and self terms, torques
    # carried to the frame atoms) against central differences of E_pol with
    # the multipoles rotated and the dipoles re-solved at every geometry
    s = thp.tiny_water(27, 10.0, 12, seed=4)
    dpu = np.array(s.scaled_dpu, dtype=np.float64).reshape(-1, 3).copy()
    dpu[:, 1] = 0.5
    s = dataclasses.replace(s, scaled_dpu=dpu)
    fr = _water_frames(s)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    ep, grad, _ = O.epolar1(s, fr, poleps=1e-12)
    h = 2e-4
    box = s.box[0]

    def clear(a):
        d = xyz - xyz[a]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        r[a] = 0.0
        return np.min(np.abs(r - s.cut)) > 50 * h

    atoms = [a for a in range(s.n) if clear(a)]
    picks = [atoms[0], atoms[1], atoms[2], atoms[len(atoms) // 2], atoms[-1]]
    for atom, axis in zip(picks, (0, 1, 2, 1, 0)):
        def e_at(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return _epol(s, fr, x2)
        fd = _d4(e_at, h)
        assert abs(grad[atom, axis] - fd) < 1e-6 * max(1.0, abs(fd)), (atom, axis, grad[atom, axis], fd)


def test_closed_form_torque_forces_match_the_frame_jacobian():
    # every axis type (with and without chirality atoms), random torques:
    # the closed form against central differences of rotmat; the forces of
    # a torque sum to zero and their moment about the origin is the torque
    xyz, axt, fz, fx, fy, pole = _frame_case()
    tq = np.random.default_rng(3).normal(size=(len(xyz), 3))
    a = O.torque_forces_closed(xyz, axt, fz, fx, fy, tq)
    b = O.torque_forces(xyz, axt, fz, fx, fy, tq)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-7 * np.abs(b).max())
    for i in range(len(xyz)):
        one = np.zeros_like(tq)
        one[i] = tq[i]
        f = O.torque_forces_closed(xyz, axt, fz, fx, fy, one)
        assert np.abs(f.sum(0)).max() < 1e-10 * max(1.0, np.abs(f).max())
        if axt[i] != O.AXIS_Z_ONLY:   # Z-Only leaves the spin about z to the global axis
            np.testing.assert_allclose(np.cross(xyz, f).sum(0), tq[i], rtol=0, atol=1e-9 * np.abs(tq[i]).max())


# ------------------------------------------------ polarization virial
# epolar1c's vir (epolar1.f:326-334, 1152-1160, 2332-2340, 2478-2526) is the
# derivative of the polarization energy along a homogeneous strain of the
# positions and the box.  Three checks that the oracle's formula
# (real-space pairs, the reciprocal sum's dependence on the box and on the
# fractional moments, the torques' share) does not state:
#  * the diagonal against central differences of the re-solved E_pol in a
#    strained box (frames turning with their atoms, or multipoles fixed);
#  * with frames, a rigid rotation changes nothing, so D is symmetric,
#    although its real, reciprocal and torque shares are not;
#  * the whole tensor, off-diagonal terms included, does not depend on the
#    Ewald coefficient, while its real and reciprocal shares change by
#    amounts of the order of the total between the coefficients tried.

def _virial_case(aewald, nmol=27, box=10.0, nfft=16):
    s = thp.tiny_water(nmol, box, nfft, seed=4)
    dpu = np.array(s.scaled_dpu, dtype=np.float64).reshape(-1, 3).copy()
    dpu[:, 1] = 0.5
    dpu[:, 2] = 0.7
    s = dataclasses.replace(s, scaled_dpu=dpu, aewald=aewald)
    fr = _water_frames(s)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    return s, fr, xyz


def _virial_of(s, fr, xyz, frames=True):
    rp = O.rotpole(*xyz.T, *fr)
    s2 = s.with_multipoles(rp)
    mu = O.Oracle(s2).induce(poleps=1e-12)["mu"]
    return O.epolar1_virial(s2, rp, mu[:, 0], mu[:, 1], fr if frames else None, parts=True)


def _strain_diagonal_fd(s, fr, xyz, frames):
    rp0 = O.rotpole(*xyz.T, *fr)

    def epol(a, h):
        f = np.ones(3)
        f[a] += h
        x2 = xyz * f
        s2 = dataclasses.replace(s.with_coords(*x2.T), box=np.asarray(s.box) * f)
        rp = O.rotpole(*x2.T, *fr) if frames else rp0
        return O.Oracle(s2.with_multipoles(rp)).induce(poleps=1e-12)["ep"]

    return np.array([_d4(lambda h, a=a: epol(a, h), 1e-4) for a in range(3)])


@pytest.mark.parametrize("frames", [True, False])
def test_polarization_virial_diagonal_matches_strain_differences(frames):
    # an Ewald coefficient large enough (1.2 / A at a 4.4 A cutoff) that no
    # pair crossing the cutoff under the strain moves the energy
    s, fr, xyz = _virial_case(1.2)
    D, parts = _virial_of(s, fr, xyz, frames)
    fd = _strain_diagonal_fd(s, fr, xyz, frames)
    print(f"virial diagonal frames={frames}: {np.diag(D)} differences {fd}")
    assert np.abs(np.diag(D) - fd).max() < 1e-7 * np.abs(fd).max()
    asym = np.abs(D - D.T).max() / np.abs(D).max()
    if frames:
        assert asym < 1e-8
        for k in ("real", "recip", "torque"):
            assert np.abs(parts[k] - parts[k].T).max() > 1e-2 * np.abs(D).max()
    else:
        assert asym > 1e-2     # fixed global multipoles feel the rotation


@pytest.mark.parametrize("frames", [True, False])
def test_polarization_virial_is_alpha_invariant(frames):
    # a 14 A box, 6.4 A cutoff and 64^3 grid: the tensor agrees between
    # alpha = 0.7 and 0.85 / A to the PME interpolation error (~3e-4 of its
    # scale), while its real-space share moves by about half of it
    Ds, reals = [], []
    for a in (0.7, 0.85):
        s, fr, xyz = _virial_case(a, box=14.0, nfft=64)
        D, parts = _virial_of(s, fr, xyz, frames)
        Ds.append(D)
        reals.append(parts["real"])
    scale = np.abs(Ds[0]).max()
    assert np.abs(Ds[1] - Ds[0]).max() < 1e-3 * scale
    assert np.abs(reals[1] - reals[0]).max() > 0.2 * scale


def test_planted_virial_mistakes_fail_the_pins(monkeypatch):
    # the strain differences catch a dropped 1/m^2 term of the reciprocal
    # weights and a reversed real-space separation
    s, fr, xyz = _virial_case(1.2)
    fd = _strain_diagonal_fd(s, fr, xyz, True)
    good = O._virial_kweights

    def no_inverse_m2(s_, qf):
        W = good(s_, qf)
        # keep -delta_ab and the pi^2/alpha^2 part only
        n1, n2, n3 = (int(v) for v in s_.nfft)
        ms = [np.where(np.arange(K) + 1 > (K + 1) // 2, np.arange(K) - K, np.arange(K)) / float(L)
              for K, L in zip((n1, n2, n3), s_.box)]
        h3, h2, h1 = np.meshgrid(ms[2], ms[1], ms[0], indexing="ij")
        hv = (h1, h2, h3)
        hsq = h1 ** 2 + h2 ** 2 + h3 ** 2
        for a in range(3):
            for b in range(3):
                with np.errstate(divide="ignore", invalid="ignore"):
                    W[a, b] -= np.where(hsq > 0, qf * 2.0 * hv[a] * hv[b] / np.where(hsq > 0, hsq, 1.0), 0.0)
        return W

    monkeypatch.setattr(O, "_virial_kweights", no_inverse_m2)
    D1, _ = _virial_of(s, fr, xyz, True)
    monkeypatch.setattr(O, "_virial_kweights", good)
    good_pairs = O._epolar1_real_pairs

    def reversed_r(*a, **k):
        F, T, (j, m, R, fp) = good_pairs(*a, **k)
        return F, T, (j, m, -R, fp)

    monkeypatch.setattr(O, "_epolar1_real_pairs", reversed_r)
    D2, _ = _virial_of(s, fr, xyz, True)
    for D in (D1, D2):
        assert np.abs(np.diag(D) - fd).max() > 1e-3 * np.abs(fd).max()


def test_short_range_gradient_matches_differences_of_the_short_energy():
    # epolar1c with use_polarshortreal: the real-space terms within the
    # short cutoff, with the frames' torques, against central differences
    # of the re-solved short-range energy (newinduce_shortreal)
    s, fr, xyz = _virial_case(0.4)
    short = 3.8
    ep, grad, _ = O.epolar1_short(s, fr, short, poleps=1e-12)

    def e_short(x2):
        s2 = dataclasses.replace(s.with_coords(*x2.T), cut=short)
        rp = O.rotpole(*x2.T, *fr)
        return O.induce_shortreal(O.Oracle(s2.with_multipoles(rp)), poleps=1e-12)["ep"]

    h = 2e-4
    box = s.box[0]

    def clear(a):
        d = xyz - xyz[a]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        r[a] = 0.0
        return np.min(np.abs(r - short)) > 50 * h

    atoms = [a for a in range(s.n) if clear(a)]
    for atom, axis in zip([atoms[0], atoms[1], atoms[2], atoms[-1]], (0, 1, 2, 1)):
        def e_at(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return e_short(x2)
        fd = _d4(e_at, h)
        assert abs(fd) > 1e-2                      # a real force component
        assert abs(grad[atom, axis] - fd) < 1e-6 * max(1.0, abs(fd)), (atom, axis, grad[atom, axis], fd)


def test_tcg_omega_fit_bisects_a_linear_energy_toward_the_converged_one():
    # omegafit (kpolar.f:1059-1326): the truncated CG energy is linear in
    # the peek step's omega (checked at a third omega), and the bisection
    # lands within its tolerance of the PCG energy, next to the root of that
    # line, which no line of tcg_omega_fit computes
    # intramolecular p-scales of 0.5: with E_d = E_p and no guess the peek
    # direction is orthogonal to the field (the CG residual property) and
    # the energy does not depend on omega at all
    s, _, _ = _virial_case(0.4, nfft=16)
    orc = O.Oracle(s)
    for order, prec in ((2, True), (1, True), (2, False)):
        w, done, e = O.tcg_omega_fit(orc, order=order, prec=prec)
        e37 = O.induce_tcg(orc, order, prec, False, True, 0.37)["ep"]
        assert abs(e37 - (e["e0"] + 0.37 * (e["e1"] - e["e0"]))) < 1e-11 * abs(e37)
        assert abs(e["e1"] - e["e0"]) > 1e-3 * abs(e["e0"])
        root = (e["epcg"] - e["e0"]) / (e["e1"] - e["e0"])
        print(f"order {order} prec {prec}: omega {w} fitted {done}, root {root:.6f}")
        if -5.0 < root < 5.0:
            assert done
            ew = O.induce_tcg(orc, order, prec, False, True, w)["ep"]
            assert abs(ew - e["epcg"]) <= e["cvg"] * (1 + 1e-9)
            assert abs(w - root) <= e["cvg"] / abs(e["e1"] - e["e0"]) * (1 + 1e-9)


@pytest.mark.parametrize("order,prec,guess,peek", [(2, True, False, True), (1, True, True, False),
                                                   (2, False, True, True), (1, False, False, True)])
def test_tcg_gradient_matches_differences_of_the_truncated_energy(order, prec, guess, peek):
    # epolar1tcg (epolar1tcg.f): the truncated CG's dipoles are not
    # stationary, so its gradient carries the adjoint terms; against central
    # differences of the truncated energy re-solved at displaced coordinates
    # (frames rotated with them), for atoms clear of the real-space cutoff
    s, fr, xyz = _virial_case(0.4)
    ep, grad = O.epolar1_tcg(s, fr, order, prec, guess, peek, 0.8)

    def e_tcg(x2):
        rp = O.rotpole(*x2.T, *fr)
        orc = O.Oracle(s.with_coords(*x2.T).with_multipoles(rp))
        return O.induce_tcg(orc, order, prec, guess, peek, 0.8)["ep"]

    assert abs(e_tcg(xyz) - ep) < 1e-12 * abs(ep)
    h = 2e-4
    box = s.box[0]

    def clear(a):
        d = xyz - xyz[a]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        r[a] = 0.0
        return np.min(np.abs(r - s.cut)) > 50 * h

    atoms = [a for a in range(s.n) if clear(a)]
    # the stationary-dipole gradient (the same dipoles, no adjoint terms)
    # misses the truncation's share: the check below must be able to tell
    rp = O.rotpole(*xyz.T, *fr)
    s2 = s.with_multipoles(rp)
    mu = O.induce_tcg(O.Oracle(s2), order, prec, guess, peek, 0.8)["mu"]
    F1, T1 = O.epolar1_real(s2, rp, mu[:, 0], mu[:, 1])
    F2, T2 = O.epolar1_recip_self(s2, rp, mu[:, 0], mu[:, 1])
    naive = -(F1 + F2 + O.torque_forces_closed(xyz, *fr[:4], T1 + T2))
    for atom, axis in zip([atoms[0], atoms[1], atoms[len(atoms) // 2], atoms[-1]], (0, 1, 2, 1)):
        def e_at(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return e_tcg(x2)
        fd = _d4(e_at, h)
        assert abs(fd) > 1e-2
        assert abs(grad[atom, axis] - fd) < 1e-6 * max(1.0, abs(fd)), (atom, axis, grad[atom, axis], fd)
        assert abs(naive[atom, axis] - fd) > 1e-4 * max(1.0, abs(fd))


@pytest.mark.parametrize("order,prec,guess,peek", [(1, True, False, False), (2, True, True, True),
                                                   (2, False, False, True)])
def test_short_range_tcg_gradient_matches_differences_of_the_short_energy(order, prec, guess, peek):
    # epolar1tcg1shortreal / epolar1tcg2*shortreal: the truncated CG on the
    # short-range operators (direct field and real-space matvec within the
    # short cutoff, no reciprocal or self terms), against central
    # differences of the re-solved short-range truncated energy
    s, fr, xyz = _virial_case(0.4)
    short = 3.8
    ep, grad = O.epolar1_tcg(s, fr, order, prec, guess, peek, 0.8, shortcut=short)

    def e_short(x2):
        s2 = dataclasses.replace(s.with_coords(*x2.T), cut=short)
        rp = O.rotpole(*x2.T, *fr)
        return O.induce_tcg(O.Oracle(s2.with_multipoles(rp)), order, prec, guess, peek, 0.8, short=True)["ep"]

    assert abs(e_short(xyz) - ep) < 1e-12 * abs(ep)
    # the short energy is not the full one
    full = O.induce_tcg(O.Oracle(s.with_multipoles(O.rotpole(*xyz.T, *fr))), order, prec, guess, peek, 0.8)["ep"]
    assert abs(full - ep) > 1e-3 * abs(ep)
    h = 2e-4
    box = s.box[0]

    def clear(a):
        d = xyz - xyz[a]
        d -= box * np.round(d / box)
        r = np.linalg.norm(d, axis=1)
        r[a] = 0.0
        return np.min(np.abs(r - short)) > 50 * h

    atoms = [a for a in range(s.n) if clear(a)]
    for atom, axis in zip([atoms[0], atoms[1], atoms[2], atoms[-1]], (0, 1, 2, 1)):
        def e_at(dx):
            x2 = xyz.copy()
            x2[atom, axis] += dx
            return e_short(x2)
        fd = _d4(e_at, h)
        assert abs(fd) > 1e-2
        assert abs(grad[atom, axis] - fd) < 1e-6 * max(1.0, abs(fd)), (atom, axis, grad[atom, axis], fd)
