"""The drop-in boundary on the GPU, against the oracle.

* The Fortran host: every reference-named routine of module thp_polar
  (nblist, efld0_direct, efld0_recip, tmatxb_pme with and without the
  diagonal, tmatxbrecip, grid_uind_site, fphi_uind_site, newinduce_pme2,
  polar_gradient, polar_set_frames) called by polar_driver --shims in the
  order of a Tinker-HP step, each result against the oracle.
* The device-pointer entry points ordered on a caller's HIP stream
  (efld0_direct, efld0_recip, tmatxb_pme, tmatxbrecip, grid_uind_site,
  fphi_uind_site), against the host entry points and the oracle.
* Device-memory accounting of a system and of the serving loop.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import thp
from oracle import oracle as O
from test_gpu_frames import _frames

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FP64, MIXED = thp.THP_PREC_FP64, thp.THP_PREC_MIXED
TOL = {FP64: 1e-11, MIXED: 2e-5}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _stencil(s, i):
    th, g0 = [], []
    for c, K, L in zip((s.x[i], s.y[i], s.z[i]), s.nfft, s.box):
        w = c / L
        fr = K * (w - np.round(w) + 0.5)
        ifr = int(fr - 1e-8)
        th.append(O.bsplgen(fr - ifr))
        g0.append(ifr - 4)
    return th, g0


def numpy_grid_uind(s, fuind, fuinp):
    """grid_uind_site (pmestuff.f:280-400) written out with numpy"""
    K1, K2, K3 = s.nfft
    g = np.zeros((K3, K2, K1, 2))
    for i in range(s.n):
        th, g0 = _stencil(s, i)
        for mk in range(5):
            for mj in range(5):
                for mi in range(5):
                    t, u, v = th[0][mi], th[1][mj], th[2][mk]
                    w = np.array([t[1] * u[0] * v[0], t[0] * u[1] * v[0], t[0] * u[0] * v[1]])
                    g[(g0[2] + mk) % K3, (g0[1] + mj) % K2, (g0[0] + mi) % K1] += (fuind[i] @ w, fuinp[i] @ w)
    return g


def numpy_fphi_uind(s, g):
    """fphi_uind_site (pmestuff.f:746-1037), first derivatives, with numpy"""
    K1, K2, K3 = s.nfft
    p1, p2 = np.zeros((s.n, 3)), np.zeros((s.n, 3))
    for i in range(s.n):
        th, g0 = _stencil(s, i)
        for mk in range(5):
            for mj in range(5):
                for mi in range(5):
                    q = g[(g0[2] + mk) % K3, (g0[1] + mj) % K2, (g0[0] + mi) % K1]
                    t, u, v = th[0][mi], th[1][mj], th[2][mk]
                    d = np.array([t[1] * u[0] * v[0], t[0] * u[1] * v[0], t[0] * u[0] * v[1]])
                    p1[i] += q[0] * d
                    p2[i] += q[1] * d
    return p1, p2


# ------------------------------------------------------------ Fortran host
def _write_input(path, s, prec, frames, mu, fuind, fuinp, poleps):
    ik = np.asarray(s.scaled_ik, dtype=np.int32).reshape(-1, 2) + 1
    with open(path, "wb") as f:
        np.array([s.n, len(ik), *s.nfft, prec, 1 if frames else 0], dtype="<i4").tofile(f)
        np.array([*s.box, s.aewald, s.cut, s.buffer, poleps], dtype="<f8").tofile(f)
        for a in (s.x, s.y, s.z, s.rpole, s.polarity, s.thole):
            np.ascontiguousarray(a, dtype="<f8").tofile(f)
        if len(ik):
            ik.astype("<i4").tofile(f)
            np.ascontiguousarray(s.scaled_dpu, dtype="<f8").tofile(f)
        if frames:
            for k in ("axistype", "zaxis", "xaxis", "yaxis"):
                v = np.asarray(frames[k], dtype=np.int64)
                (v if k == "axistype" else v + 1).astype("<i4").tofile(f)
            np.ascontiguousarray(frames["pole"], dtype="<f8").tofile(f)
        for a in (mu, fuind, fuinp):
            np.ascontiguousarray(a, dtype="<f8").tofile(f)


def _read_output(path, s):
    n, (K1, K2, K3) = s.n, s.nfft
    with open(path, "rb") as f:
        npairs = int(np.fromfile(f, "<i8", 1)[0])
        niter = int(np.fromfile(f, "<i4", 1)[0])
        take = lambda *shape: np.fromfile(f, "<f8", int(np.prod(shape))).reshape(shape)  # noqa: E731
        out = dict(npairs=npairs, niter=niter, ef=take(n, 2, 3), cphi=take(n, 10), efi=take(n, 2, 3),
                   efi0=take(n, 2, 3), dipfield=take(n, 2, 3), qgrid=take(K3, K2, K1, 2), phi1=take(n, 3),
                   phi2=take(n, 3), uind=take(n, 3), uinp=take(n, 3), ep=float(take(1)[0]), dep=take(n, 3),
                   torque=take(n, 3))
    return out


@pytest.mark.parametrize("framed", [False, True])
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_fortran_reference_routines_match_oracle(prec, framed):
    s = thp.tiny_water(27, 20.3, 24, seed=2)
    fr = _frames(s, seed=4) if framed else None
    rng = np.random.default_rng(6)
    mu = rng.normal(size=(s.n, 2, 3)) * 0.1
    fuind, fuinp = rng.normal(size=(s.n, 3)), rng.normal(size=(s.n, 3))
    exe = os.path.join(ROOT, "tinker-hp_amd", "fortran", "polar_driver")
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        _write_input(fin, s, prec, fr, mu, fuind, fuinp, 1e-10)
        print(f"polar_driver --shims prec={prec} framed={framed}", flush=True)
        r = subprocess.run([exe, "--shims", fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        got = _read_output(fout, s)
    # the oracle sees the multipoles the device rotated from the frames
    if framed:
        rp = O.rotpole(s.x, s.y, s.z, fr["axistype"], fr["zaxis"], fr["xaxis"], fr["yaxis"], fr["pole"])
        s = s.with_multipoles(rp)
    orc = O.Oracle(s)
    tol = TOL[prec]
    assert got["npairs"] == len(orc.nblist_pairs())
    assert rel(got["ef"], orc.efld0_direct()) < tol
    assert rel(got["cphi"], orc.efld0_recip()) < tol
    assert rel(got["efi"], orc.tmatxb_pme(mu)) < tol
    assert rel(got["efi0"], orc.tmatxb_pme(mu, dodiag=False)) < tol
    assert rel(got["dipfield"], orc.tmatxbrecip(mu)) < tol
    ref_grid = numpy_grid_uind(s, fuind, fuinp)
    assert rel(got["qgrid"], ref_grid) < (1e-12 if prec == FP64 else 2e-6)
    p1, p2 = numpy_fphi_uind(s, got["qgrid"])
    assert rel(got["phi1"], p1) < tol and rel(got["phi2"], p2) < tol
    ref = orc.induce(poleps=1e-10)
    assert abs(got["ep"] - ref["ep"]) < (1e-11 if prec == FP64 else 1e-6) * abs(ref["ep"])
    mu_got = np.stack([got["uind"], got["uinp"]], axis=1)
    assert float(np.sqrt(np.mean((mu_got - ref["mu"]) ** 2))) * thp.DEBYE < (1e-9 if prec == FP64 else 1e-5)
    if prec == FP64:
        assert got["niter"] == ref["niter"]
    gtol = 1e-9 if prec == FP64 else 1e-5
    if framed:
        # torques carried to the frame atoms on the device
        frt = (fr["axistype"], fr["zaxis"], fr["xaxis"], fr["yaxis"], fr["pole"])
        _, grad, _ = O.epolar1(thp.tiny_water(27, 20.3, 24, seed=2), frt, poleps=1e-10)
        assert np.abs(got["dep"] - grad).max() < gtol * np.abs(grad).max()
    else:
        # fixed-orientation forces and the site torques, for the host
        f1, t1 = O.epolar1_real(s, s.rpole, ref["mu"][:, 0], ref["mu"][:, 1])
        f2, t2 = O.epolar1_recip_self(s, s.rpole, ref["mu"][:, 0], ref["mu"][:, 1])
        f, t = f1 + f2, t1 + t2
        assert np.abs(-got["dep"] - f).max() < gtol * np.abs(f).max()
        assert np.abs(got["torque"] - t).max() < gtol * np.abs(t).max()


# --------------------------------------------- device entries on a stream
class _Hip:
    """the HIP runtime's C API, for device buffers and a caller stream"""

    def __init__(self):
        h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        self.h = h
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def free_all(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []

    def mem_free(self):
        f, t = C.c_size_t(0), C.c_size_t(0)
        assert self.h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_device_entry_points_ordered_on_the_callers_stream(ctx, prec):
    H2D, D2H = 1, 2
    hip = _Hip()
    s = thp.tiny_water(27, 20.3, 24, seed=3)
    solver = thp.Solver(ctx, s, prec)
    lib, h = solver.lib, solver.h
    rng = np.random.default_rng(8)
    mu = np.ascontiguousarray(rng.normal(size=(s.n, 2, 3)) * 0.1)
    fuind, fuinp = (np.ascontiguousarray(rng.normal(size=(s.n, 3))) for _ in range(2))
    K1, K2, K3 = s.nfft
    # host entry points first (the same kernels, synchronous)
    want = dict(ef=solver.efld0_direct(), cphi=solver.efld0_recip(), efi=solver.tmatxb_pme(mu),
                dip=solver.tmatxbrecip(mu), grid=solver.grid_uind_site(fuind, fuinp))
    want["phi1"], want["phi2"] = solver.fphi_uind_site(want["grid"])
    stream = C.c_void_p()
    assert hip.h.hipStreamCreate(C.byref(stream)) == 0
    try:
        nv, nc, ng, n3 = s.n * 6 * 8, s.n * 10 * 8, K1 * K2 * K3 * 2 * 8, s.n * 3 * 8
        d = {k: hip.alloc(b) for k, b in dict(mu=nv, ef=nv, cphi=nc, efi=nv, dip=nv, fd=n3, fp=n3, grid=ng,
                                                  p1=n3, p2=n3).items()}
        big = hip.alloc(256 << 20)
        junk = np.full(s.n * 6, 1e3)
        for k in ("mu", "efi", "dip"):
            assert hip.h.hipMemcpy(d[k], junk.ctypes.data, nv, H2D) == 0
        # the caller's stream is busy (a 256 MiB fill) before its input
        # copies: the library must wait for them, not read the junk
        assert hip.h.hipMemsetAsync(big, 0, 256 << 20, stream) == 0
        for k, a in (("mu", mu), ("fd", fuind), ("fp", fuinp)):
            assert hip.h.hipMemcpyAsync(d[k], a.ctypes.data, a.nbytes, H2D, stream) == 0
        print(f"stream-ordered entries prec={prec}", flush=True)
        assert lib.thp_efld0_direct_dev(h, d["ef"], stream) == 0
        assert lib.thp_efld0_recip_dev(h, d["cphi"], stream) == 0
        assert lib.thp_tmatxb_dev_async(h, d["mu"], d["efi"], 1, stream) == 0
        assert lib.thp_tmatxb_recip_dev(h, d["mu"], d["dip"], stream) == 0
        assert lib.thp_grid_uind_dev(h, d["fd"], d["fp"], d["grid"], stream) == 0
        assert lib.thp_fphi_uind_dev(h, d["grid"], d["p1"], d["p2"], stream) == 0
        got = {k: np.zeros_like(want[k]) for k in want}
        for k, dk in (("ef", "ef"), ("cphi", "cphi"), ("efi", "efi"), ("dip", "dip"), ("grid", "grid"),
                      ("phi1", "p1"), ("phi2", "p2")):
            assert hip.h.hipMemcpyAsync(got[k].ctypes.data, d[dk], got[k].nbytes, D2H, stream) == 0
        assert hip.h.hipStreamSynchronize(stream) == 0
        for k in want:
            assert rel(got[k], want[k]) < 1e-12, k
        # and against the oracle
        orc = O.Oracle(s)
        tol = TOL[prec]
        assert rel(got["ef"], orc.efld0_direct()) < tol
        assert rel(got["cphi"], orc.efld0_recip()) < tol
        assert rel(got["efi"], orc.tmatxb_pme(mu)) < tol
        assert rel(got["dip"], orc.tmatxbrecip(mu)) < tol
        assert rel(got["grid"], numpy_grid_uind(s, fuind, fuinp)) < (1e-12 if prec == FP64 else 2e-6)
        # a null stream: the call returns with the results in place
        out = np.zeros_like(mu)
        assert lib.thp_tmatxb_dev_async(h, d["mu"], d["efi"], 0, None) == 0
        assert hip.h.hipMemcpy(out.ctypes.data, d["efi"], nv, D2H) == 0
        assert rel(out, orc.tmatxb_pme(mu, dodiag=False)) < tol
    finally:
        hip.h.hipStreamSynchronize(stream)
        hip.free_all()
        hip.h.hipStreamDestroy(stream)
        solver.close()


# ------------------------------------------------------- memory accounting
def test_footprint_is_stable_over_repeated_calls(ctx):
    s = thp.water_1944()
    solver = thp.Solver(ctx, s, MIXED)
    solver.newinduce_pme2()
    solver.epolar1()
    mu = np.random.default_rng(1).normal(size=(s.n, 2, 3)) * 0.1
    solver.tmatxb_pme(mu)
    solver.efld0_recip()
    fp = solver.footprint()
    for _ in range(3):
        solver.newinduce_pme2()
        solver.epolar1()
        solver.tmatxb_pme(mu)
        solver.efld0_recip()
        solver.tmatxbrecip(mu)
    assert solver.footprint() == fp
    solver.close()


def test_footprint_accounts_for_the_device_memory_of_a_system(ctx):
    hip = _Hip()
    s = thp.water_23559()
    free0 = hip.mem_free()
    solver = thp.Solver(ctx, s, MIXED)
    solver.newinduce_pme2()
    delta = free0 - hip.mem_free()
    fp = solver.footprint()
    print(f"footprint {fp / 2**20:.1f} MiB, device memory taken {delta / 2**20:.1f} MiB", flush=True)
    # the runtime rounds allocations up (and may reuse cached memory), so
    # the footprint can sit below the figure the device reports, not far
    assert fp >= 0.8 * delta
    solver.close()


def test_serving_resident_bytes_follow_the_lists():
    from thp.serve import Server
    ctx = thp.Context(0)
    s = thp.serving_system(6000, 0.2, seed=3)
    # the same system on its own: its footprint less its own grid and
    # spreading tiles (serving systems take theirs from the loop's pools)
    # plus the caller-order multipoles a serving system keeps
    one = thp.Solver(ctx, s, MIXED)
    one.newinduce_pme2()
    K = int(np.prod(s.nfft))
    bricks = 1
    for k in s.nfft:
        b = next(B for B in range(8, 3, -1) if k % B == 0 or k % B >= 4)
        bricks *= -(-k // b)
    own = 2 * K * 4 + bricks * 12 ** 3 * 2 * 4
    want = one.footprint() - own + 13 * s.n * 8
    one.close()
    # no grid autotuning: the serving system keeps the bricks of `one`
    with Server(ctx, max_systems=2, autotune=0) as srv:
        sid = srv.add(s)
        r = srv.poll(srv.submit(sid))
        assert r.status == 0
        st = srv.stats()
        print(f"resident {st.resident_bytes} want {want}", flush=True)
        assert st.resident_systems == 1
        assert abs(st.resident_bytes - want) <= 0.02 * want
    ctx.close()
