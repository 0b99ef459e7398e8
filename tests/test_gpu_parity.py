"""GPU parity of the HIP path against the FP64 CPU oracle (through the C ABI).

Bars: neighbor lists bit-exact as pair sets; FP64 path within 1e-11
relative of the oracle (only summation order differs); mixed path within
1e-6 relative in E_pol and 1e-5 Debye RMS in the induced dipoles
per-kernel fields within 2e-5 relative.
"""
import ctypes as C
import dataclasses
import os
import re
import subprocess

import numpy as np
import pytest

import thp
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def golden81():
    g = np.load(os.path.join(GOLD, "water81.npz"))
    s = thp.tiny_water(27, 20.3, 24)
    assert np.array_equal(s.x, g["x"])
    return s, g


# ------------------------------------------------------------ neighbor lists
def _shifted(s, seed=3):
    """same system with every atom moved by random whole box lengths, which
    exercises the floor branch of the minimum image"""
    rng = np.random.default_rng(seed)
    sh = rng.integers(-2, 3, size=(s.n, 3)) * s.box[None, :]
    return s.with_coords(s.x + sh[:, 0], s.y + sh[:, 1], s.z + sh[:, 2])


@pytest.mark.parametrize("case", ["water81", "water1944", "solute", "serving", "shifted"])
def test_nblist_pair_set_bit_exact(ctx, case):
    if case == "water81":
        s, g = golden81()
        ref = g["pairs"]
    else:
        s = {"water1944": thp.water_1944,
             "solute": lambda: thp.make_system(3000, 34.0, n_solute=1200, blob_radius=11.0, seed=9, nfft=(36,) * 3),
             "serving": lambda: thp.serving_system(800, 0.2, seed=5),
             "shifted": lambda: _shifted(thp.water_1944())}[case]()
        ref = O.Oracle(s).nblist_pairs()
    solver = thp.Solver(ctx, s, FP64)
    solver.nblist()
    got = solver.nblist_pairs()
    assert got.shape == ref.shape and np.array_equal(got, ref)


def within_cut(s, pairs, cut):
    """the rows of `pairs` with r^2 <= cut^2 in the reference's operation
    order (xr = xi - xk, orthogonal image with floor, (x*x + y*y) + z*z;
    image.f:113-140, newinduce_pme.f:1160)"""
    i, k = pairs[:, 0], pairs[:, 1]
    d = []
    for c, L in zip((s.x, s.y, s.z), s.box):
        v = c[i] - c[k]
        big = np.abs(v) > 0.5 * L
        v2 = np.where(big, v - np.copysign(L, v) * np.floor(np.abs(v) / L), v)
        big2 = big & (np.abs(v2) > 0.5 * L)
        d.append(np.where(big2, v2 - np.copysign(L, v2), v2))
    r2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    return pairs[r2 <= cut * cut]


@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("case", ["water1944", "solute", "serving", "shifted"])
def test_solver_pair_set_bit_exact(ctx, case, prec):
    # the pairs the field pass and the matvec actually use (compacted lists
    # plus the scaled pairs within the cutoff) against the oracle's list
    # entries within the cutoff
    s = {"water1944": thp.water_1944,
         "solute": lambda: thp.make_system(3000, 34.0, n_solute=1200, blob_radius=11.0, seed=9, nfft=(36,) * 3),
         "serving": lambda: thp.serving_system(800, 0.2, seed=5),
         "shifted": lambda: _shifted(thp.water_1944())}[case]()
    ref = within_cut(s, O.Oracle(s).nblist_pairs(), s.cut)
    solver = thp.Solver(ctx, s, prec)
    got = solver.pairs_in_cut()
    print(f"{case} prec={prec} pairs={len(got)} ref={len(ref)}", flush=True)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    # and after the atoms moved within the lists' skin (lists kept, the
    # compacted entries rewritten for the new coordinates)
    rng = np.random.default_rng(1)
    s2 = s.with_coords(*(c + rng.normal(0, 0.05, s.n) for c in (s.x, s.y, s.z)))
    solver.update_coords(s2.x, s2.y, s2.z)
    got2 = solver.pairs_in_cut()
    ref2 = within_cut(s2, O.Oracle(s2).nblist_pairs(), s.cut)
    assert got2.shape == ref2.shape and np.array_equal(got2, ref2)
    solver.close()


# ---------------------------------------------------------- path entry points
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-11), (MIXED, 2e-5)])
def test_field_and_matvec_entry_points_golden(ctx, prec, tol):
    s, g = golden81()
    solver = thp.Solver(ctx, s, prec)
    assert rel(solver.efld0_direct(), g["ef_direct"]) < tol
    assert rel(solver.efld0_recip(), g["cphi"]) < tol
    assert rel(solver.tmatxb_pme(g["mu_probe"]), g["tmatxb"]) < tol
    assert rel(solver.tmatxbrecip(g["mu_probe"]), g["tmatxbrecip"]) < tol


# grid lengths whose transforms take every radix of the grid FFT (8, 4, 2,
# 3, 5), odd lengths and a different length per axis
@pytest.mark.parametrize("nfft", [(30, 30, 30), (36, 36, 36), (45, 45, 45), (50, 50, 50), (30, 36, 45)])
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-11), (MIXED, 2e-5)])
def test_reciprocal_terms_on_mixed_radix_grids(ctx, nfft, prec, tol):
    s = thp.make_system(699, 19.5, seed=13, nfft=nfft, name=f"water-699-{nfft}")
    orc = O.Oracle(s)
    solver = thp.Solver(ctx, s, prec)
    mu = np.random.default_rng(2).normal(size=(s.n, 2, 3)) * 0.05
    assert rel(solver.tmatxbrecip(mu), orc.tmatxbrecip(mu)) < tol
    assert rel(solver.efld0_recip(), orc.efld0_recip()) < tol


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-11), (MIXED, 2e-5)])
def test_field_and_matvec_water1944(ctx, prec, tol):
    s = thp.water_1944()
    orc = O.Oracle(s)
    solver = thp.Solver(ctx, s, prec)
    rng = np.random.default_rng(0)
    mu = rng.normal(size=(s.n, 2, 3)) * 0.1
    assert rel(solver.tmatxb_pme(mu), orc.tmatxb_pme(mu)) < tol
    assert rel(solver.tmatxb_pme(mu, dodiag=False), orc.tmatxb_pme(mu, dodiag=False)) < tol
    assert rel(solver.tmatxbrecip(mu), orc.tmatxbrecip(mu)) < tol
    assert rel(solver.efld0_recip(), orc.efld0_recip()) < tol
    # the 648-water box is three times liquid density: FP32 pair arithmetic
    # loses more at the overlapping contacts, so its permanent field is
    # checked at 2e-5 only in FP64
    assert rel(solver.efld0_direct(), orc.efld0_direct()) < (tol if prec == FP64 else 2e-5 * 10)


# The mixed path's real-space matvec either recomputes the pair coefficients
# per entry (FP32 systems with one Thole value, the default) or stores them
# per solve (THP_MATVEC_RC=0, or systems whose Thole values differ): both
# kinds against the oracle, on water and on the uneven solute box.
def _thole_varied(s):
    th = s.thole.copy()
    th[::7] = 0.41
    return dataclasses.replace(s, thole=th, name=s.name + "-thole")


def _some_unpolarizable(s):
    """every 7th water without polarizability: pdamp 0, so its atoms' pairs
    are undamped (damptholed's damp = 0 branch) and their diagonal takes
    tinypol (whole molecules, so that no polarizable atom shares a molecule
    with an unpolarizable one)"""
    pol = s.polarity.copy()
    mol = np.arange(s.n) // 3
    pol[mol % 7 == 3] = 0.0
    return dataclasses.replace(s, polarity=pol, name=s.name + "-pol0")


@pytest.mark.parametrize("case", ["water1944", "solute", "thole", "pol0"])
@pytest.mark.parametrize("kind", ["recomputed", "stored"])
def test_matvec_coefficient_kinds_match_oracle(ctx, ctx_stored, case, kind):
    s = {"water1944": thp.water_1944,
         "solute": lambda: thp.make_system(3000, 34.0, n_solute=1200, blob_radius=11.0, seed=9, nfft=(36,) * 3),
         "thole": lambda: _thole_varied(thp.water_1944()),
         "pol0": lambda: _some_unpolarizable(thp.serving_system(2400, 0.0, seed=17))}[case]()
    orc = O.Oracle(s)
    solver = thp.Solver(ctx if kind == "recomputed" else ctx_stored, s, MIXED)
    mu = np.random.default_rng(4).normal(size=(s.n, 2, 3)) * 0.1
    for dodiag in (True, False):
        e = rel(solver.tmatxb_pme(mu, dodiag=dodiag), orc.tmatxb_pme(mu, dodiag=dodiag))
        print(f"{case} {kind} dodiag={dodiag} rel={e:.2e}", flush=True)
        assert e < 2e-5
    solver.close()
    check_solve(ctx if kind == "recomputed" else ctx_stored, s, MIXED, orc.induce(poleps=1e-8))


def numpy_grid_uind(s, fuind, fuinp):
    """grid_uind_site (pmestuff.f:280-400) written out with numpy"""
    K1, K2, K3 = s.nfft
    g = np.zeros((K3, K2, K1, 2))
    for i in range(s.n):
        th, g0 = [], []
        for d, (c, K, L) in enumerate(zip((s.x[i], s.y[i], s.z[i]), (K1, K2, K3), s.box)):
            w = c / L
            fr = K * (w - np.round(w) + 0.5)
            ifr = int(fr - 1e-8)
            th.append(O.bsplgen(fr - ifr))
            g0.append(ifr - 4)
        for mk in range(5):
            for mj in range(5):
                for mi in range(5):
                    t, u, v = th[0][mi], th[1][mj], th[2][mk]
                    vd = fuind[i, 0] * t[1] * u[0] * v[0] + fuind[i, 1] * t[0] * u[1] * v[0] + fuind[i, 2] * t[0] * u[0] * v[1]
                    vp = fuinp[i, 0] * t[1] * u[0] * v[0] + fuinp[i, 1] * t[0] * u[1] * v[0] + fuinp[i, 2] * t[0] * u[0] * v[1]
                    g[(g0[2] + mk) % K3, (g0[1] + mj) % K2, (g0[0] + mi) % K1] += (vd, vp)
    return g


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (MIXED, 2e-6)])
def test_grid_uind_and_fphi_uind(ctx, prec, tol):
    s, g = golden81()
    rng = np.random.default_rng(5)
    fd, fp = rng.normal(size=(s.n, 3)), rng.normal(size=(s.n, 3))
    solver = thp.Solver(ctx, s, prec)
    grid = solver.grid_uind_site(fd, fp)
    ref = numpy_grid_uind(s, fd, fp)
    assert rel(grid, ref) < tol
    # gather from the same grid: fphi_uind_site against an explicit sum
    p1, p2 = solver.fphi_uind_site(ref)
    K1, K2, K3 = s.nfft
    for i in range(0, s.n, 7):
        th, g0 = [], []
        for c, K, L in zip((s.x[i], s.y[i], s.z[i]), (K1, K2, K3), s.box):
            w = c / L
            fr = K * (w - np.round(w) + 0.5)
            ifr = int(fr - 1e-8)
            th.append(O.bsplgen(fr - ifr))
            g0.append(ifr - 4)
        e1 = np.zeros(3)
        e2 = np.zeros(3)
        for mk in range(5):
            for mj in range(5):
                for mi in range(5):
                    q = ref[(g0[2] + mk) % K3, (g0[1] + mj) % K2, (g0[0] + mi) % K1]
                    t, u, v = th[0][mi], th[1][mj], th[2][mk]
                    d = np.array([t[1] * u[0] * v[0], t[0] * u[1] * v[0], t[0] * u[0] * v[1]])
                    e1 += q[0] * d
                    e2 += q[1] * d
        assert np.allclose(p1[i], e1, rtol=tol * 10, atol=1e-12) and np.allclose(p2[i], e2, rtol=tol * 10, atol=1e-12)


# ------------------------------------------------------------------ solves
def check_solve(ctx, s, prec, ref=None):
    ref = ref or O.Oracle(s).induce(poleps=1e-8)
    solver = thp.Solver(ctx, s, prec)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
    erel = abs(ep - ref["ep"]) / abs(ref["ep"])
    rms = float(np.sqrt(np.mean((mu - ref["mu"]) ** 2)) * thp.DEBYE)
    print(f"{s.name} prec={prec} ep={ep:.10f} ref={ref['ep']:.10f} rel={erel:.2e} muRMS_D={rms:.2e} "
          f"niter={st.niter}/{ref['niter']}", flush=True)
    assert st.status == 0
    if prec == FP64:
        assert erel < 1e-11 and rms < 1e-9 and st.niter == ref["niter"]
    else:
        assert erel < 1e-6 and rms < 1e-5 and abs(st.niter - ref["niter"]) <= 1
    return mu, ep, st


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_induce_golden81(ctx, prec):
    s, g = golden81()
    check_solve(ctx, s, prec, dict(ep=float(g["ep"]), mu=g["mu"], niter=int(g["niter"])))


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_induce_water1944(ctx, prec):
    g = np.load(os.path.join(GOLD, "water1944_summary.npz"))
    s = thp.water_1944()
    ref = O.Oracle(s).induce(poleps=1e-8)
    assert abs(ref["ep"] - float(g["ep"])) < 1e-11 * abs(ref["ep"])
    check_solve(ctx, s, prec, ref)


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_induce_water23559_full_size(ctx, prec):
    check_solve(ctx, thp.water_23559(), prec)


@pytest.fixture(scope="module")
def solute96000_ref():
    s = thp.solute_96000()
    return s, O.Oracle(s).induce(poleps=1e-8)


# The uneven-density configuration at its full size: 96,000 atoms in a 100 A
# box with a 20,000-atom solute blob, 108^3 grid
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_induce_solute96000_full_size(ctx, solute96000_ref, prec):
    s, ref = solute96000_ref
    check_solve(ctx, s, prec, ref)


# The large configuration at its full size: 1,068,000 atoms, 220 A box,
# 240^3 grid (five one-axis FFT passes: the plane kernel's LDS does not hold
# a 240^2 plane).  An oracle solve takes minutes here, so the mixed path is
# held to the FP64 path, which the oracle pins at every other size, with
# the same bars as against the oracle.
def test_water1068000_mixed_matches_fp64_full_size(ctx):
    s = thp.water_1068000()
    out = {}
    for prec in (FP64, MIXED):
        solver = thp.Solver(ctx, s, prec)
        mu, ep, st = solver.newinduce_pme2(poleps=1e-8)
        assert st.status == 0
        out[prec] = (mu, ep, st.niter)
        solver.close()
    (mu64, ep64, it64), (mu32, ep32, it32) = out[FP64], out[MIXED]
    erel = abs(ep32 - ep64) / abs(ep64)
    rms = float(np.sqrt(np.mean((mu32 - mu64) ** 2)) * thp.DEBYE)
    print(f"water-1068000 ep64={ep64:.8f} ep32={ep32:.8f} rel={erel:.2e} muRMS_D={rms:.2e} niter={it64}/{it32}",
          flush=True)
    assert erel < 1e-6 and rms < 1e-5 and abs(it32 - it64) <= 1


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_induce_solute_blob(ctx, prec):
    s = thp.make_system(6000, 40.0, n_solute=2000, blob_radius=14.0, seed=21, nfft=(40,) * 3, name="solute-6000")
    check_solve(ctx, s, prec)


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_results_are_deterministic(ctx, prec):
    s = thp.water_1944()
    solver = thp.Solver(ctx, s, prec)
    a = solver.newinduce_pme2()
    b = solver.newinduce_pme2()
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    other = thp.Solver(ctx, s, prec)
    c = other.newinduce_pme2()
    assert np.array_equal(a[0], c[0]) and a[1] == c[1]


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_split_and_interleaved_dipoles_agree(ctx, prec):
    # thp_induce (uind, uinp apart, as the reference passes them) and
    # thp_induce_mu (one [n][2][3] array) return the same solve
    solver = thp.Solver(ctx, thp.water_1944(), prec)
    mu, ep, st = solver.newinduce_pme2()
    uind, uinp, ep2, st2 = solver.newinduce_pme2_split()
    assert np.array_equal(mu[:, 0], uind) and np.array_equal(mu[:, 1], uinp)
    assert ep == ep2 and st.niter == st2.niter


def test_iterations_queued_past_convergence_change_nothing(ctx):
    # a mixed solve queues the last solve's iteration count before its first
    # read of the solve state: after a tight solve (many iterations) a loose
    # one stops where a fresh solver stops, with the same dipoles
    s = thp.water_1944()
    ref = thp.Solver(ctx, s, MIXED).newinduce_pme2(poleps=1e-4)
    solver = thp.Solver(ctx, s, MIXED)
    tight = solver.newinduce_pme2(poleps=1e-10)
    assert tight[2].niter > ref[2].niter + 3
    loose = solver.newinduce_pme2(poleps=1e-4)
    assert loose[2].niter == ref[2].niter
    assert np.array_equal(loose[0], ref[0]) and loose[1] == ref[1]
    # politer caps the queued iterations, and the cap is reported
    capped = solver.newinduce_pme2(poleps=1e-10, politer=3)
    assert capped[2].niter == 3 and capped[2].status == -6


def test_warm_start_converges_faster(ctx):
    s = thp.water_1944()
    solver = thp.Solver(ctx, s, MIXED)
    mu, ep, st = solver.newinduce_pme2()
    rng = np.random.default_rng(8)
    x, y, z = (v + rng.normal(0, 0.01, s.n) for v in (s.x, s.y, s.z))
    solver.update_coords(x, y, z)
    cold = solver.newinduce_pme2()
    warm = solver.newinduce_pme2(mu0=cold[0] + 1e-4)
    assert warm[2].niter < cold[2].niter
    ref = O.Oracle(s.with_coords(x, y, z)).induce()
    assert abs(warm[1] - ref["ep"]) < 1e-6 * abs(ref["ep"])


def test_coordinate_updates_keep_lists_consistent(ctx):
    s = thp.water_23559()
    solver = thp.Solver(ctx, s, FP64)
    solver.newinduce_pme2()
    rng = np.random.default_rng(11)
    cases = [
        ("inside the union skin", (v + rng.normal(0, 0.01, s.n) for v in (s.x, s.y, s.z))),
        ("beyond the skin", (v + rng.normal(0, 0.1, s.n) for v in (s.x, s.y, s.z))),
        ("rigid shift beyond the buffer", (v + 1.2 for v in (s.x, s.y, s.z))),
    ]
    for label, xyz in cases:
        x, y, z = xyz
        solver.update_coords(x, y, z)
        s2 = s.with_coords(x, y, z)
        orc = O.Oracle(s2)
        mu, ep, st = solver.newinduce_pme2()
        ref = orc.induce()
        assert abs(ep - ref["ep"]) < 1e-11 * abs(ref["ep"]), label
        assert np.array_equal(solver.nblist_pairs(), orc.nblist_pairs()), label


# ---------------------------------------------------------- failure paths
def test_nonconvergence_is_reported(ctx):
    solver = thp.Solver(ctx, thp.water_1944(), FP64)
    mu, ep, st = solver.newinduce_pme2(poleps=1e-8, politer=3)
    assert st.niter == 3 and st.status == -6


def test_box_too_small_is_rejected(ctx):
    s = thp.make_system(300, 15.0, seed=1, nfft=(16,) * 3)
    with pytest.raises(thp.ThpError) as e:
        thp.Solver(ctx, s, FP64)
    assert e.value.code == -5


def test_nonfinite_coordinates_report_nan(ctx):
    s, _ = golden81()
    x = s.x.copy()
    x[5] = np.nan
    solver = thp.Solver(ctx, s.with_coords(x, s.y, s.z), FP64)
    with pytest.raises(thp.ThpError) as e:
        solver.newinduce_pme2(politer=20)
    assert e.value.code == -7


# ------------------------------------------------------------ host layers
def test_fortran_host_matches_oracle():
    exe = os.path.join(ROOT, "tinker-hp_amd", "fortran", "polar_driver")
    out = subprocess.run([exe, "1944", "18.643", "24", "1", "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = dict(re.findall(r"(\w+)=\s*(\S+)", out.stdout))
    ref = O.Oracle(thp.water_1944()).induce()
    assert abs(float(vals["EP"]) - ref["ep"]) < 1e-11 * abs(ref["ep"])
    assert int(vals["NITER"]) == ref["niter"]
    assert int(vals["NPAIRS"]) == 891263


def test_device_pointer_variants(ctx):
    # device buffers come from the same HIP runtime libthp uses
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    H2D, D2H = 1, 2
    s, g = golden81()
    solver = thp.Solver(ctx, s, FP64)
    nbytes = s.n * 6 * 8
    d_mu, d_out = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_mu), nbytes) == 0 and hip.hipMalloc(C.byref(d_out), nbytes) == 0
    try:
        mu = np.ascontiguousarray(g["mu_probe"])
        assert hip.hipMemcpy(d_mu, mu.ctypes.data, nbytes, H2D) == 0
        assert solver.lib.thp_tmatxb_dev(solver.h, d_mu, d_out, 1) == 0
        out = np.zeros_like(mu)
        assert hip.hipMemcpy(out.ctypes.data, d_out, nbytes, D2H) == 0
        assert rel(out, g["tmatxb"]) < 1e-11
        zero = np.zeros_like(mu)
        assert hip.hipMemcpy(d_mu, zero.ctypes.data, nbytes, H2D) == 0
        opts = thp.SolveOpts(poleps=1e-8, politer=100, guess=0)
        ep, st = C.c_double(0), thp.SolveStats()
        assert solver.lib.thp_induce_dev(solver.h, C.byref(opts), d_mu, C.byref(ep), C.byref(st)) == 0
        got = np.zeros_like(mu)
        assert hip.hipMemcpy(got.ctypes.data, d_mu, nbytes, D2H) == 0
        assert abs(ep.value - float(g["ep"])) < 1e-11 * abs(float(g["ep"]))
        assert rel(got, g["mu"]) < 1e-10
    finally:
        hip.hipFree(d_mu)
        hip.hipFree(d_out)


def test_kernel_timing_over_a_batch(ctx):
    """The batched timing hook covers every system of the launch: its pair
    count is the sum of the systems' counts, and a system given twice is
    refused."""
    s1, _ = golden81()
    s2 = thp.water_1944()
    a, b = thp.Solver(ctx, s1, MIXED), thp.Solver(ctx, s2, MIXED)
    ms_a, pa = a.time_kernel(0, reps=2)
    ms_b, pb = b.time_kernel(0, reps=2)
    for k in range(4):
        ms, pairs = thp.time_kernel_batch([a, b], k, reps=2)
        assert ms > 0.0
        assert pairs == pa + pb
    with pytest.raises(thp.ThpError):
        thp.time_kernel_batch([a, a], 0, reps=1)
    a.close()
    b.close()
