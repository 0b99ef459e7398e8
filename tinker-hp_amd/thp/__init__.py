"""Python host binding of libthp, the MI355X-native AMOEBA induced-dipole solver.

The product is the C ABI in include/thp.h (libthp.so, hand-written HIP
kernels for gfx950).  This module binds it with ctypes and mirrors the
reference routine names of the polarization path:

    Solver.nblist()          <- nblist / mlistcell      (v1.2/source/nblist.f:17-247)
    Solver.efld0_direct()    <- efld0_direct            (v1.2/source/newinduce_pme.f:948)
    Solver.efld0_recip()     <- efld0_recip             (v1.2/source/newinduce_pme.f:1671)
    Solver.tmatxb_pme(mu)    <- tmatxb_pme              (v1.2/source/newinduce_pme.f:1392)
    Solver.tmatxbrecip(mu)   <- tmatxbrecip             (v1.2/source/newinduce_pme.f:1919)
    Solver.grid_uind_site()  <- grid_uind_site          (v1.2/source/pmestuff.f:280)
    Solver.fphi_uind_site()  <- fphi_uind_site          (v1.2/source/pmestuff.f:746)
    Solver.newinduce_pme2()  <- newinduce_pme2          (v1.2/source/newinduce_pme2.f:16)

There is no CPU fallback: every call goes to the GPU through libthp and
raises ThpError when the library, the device or a kernel fails.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libthp.so")

THP_OK = 0
THP_PREC_FP64 = 0
THP_PREC_MIXED = 1
# induced-dipole solvers (polalg)
THP_POLALG_PCG, THP_POLALG_JACOBI_DIIS, THP_POLALG_TCG = 1, 2, 3
# local-frame axis types (rotpole.f)
THP_AXIS_NONE, THP_AXIS_Z_ONLY, THP_AXIS_Z_THEN_X, THP_AXIS_BISECTOR, THP_AXIS_Z_BISECT, THP_AXIS_3_FOLD = range(6)
# induced-dipole predictors (polpred)
THP_PRED_NONE, THP_PRED_GEAR, THP_PRED_ASPC, THP_PRED_LSQR = 0, 1, 2, 3
ERRORS = {
    -1: "THP_ERR_ARG", -2: "THP_ERR_HIP", -3: "THP_ERR_FFT", -4: "THP_ERR_NOMEM",
    -5: "THP_ERR_LIST", -6: "THP_ERR_NOCONV", -7: "THP_ERR_NAN", -8: "THP_ERR_STATE",
    -9: "THP_ERR_BUSY",
}
COULOMB = 332.063714  # v1.2/source/MOD_units.f:81
DEBYE = 4.80321       # v1.2/source/MOD_units.f:82


class ThpError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SystemDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("box", C.c_double * 3),
        ("x", _dp), ("y", _dp), ("z", _dp), ("rpole", _dp), ("polarity", _dp), ("thole", _dp),
        ("nscaled", C.c_int), ("scaled_ik", _ip), ("scaled_dpu", _dp),
        ("aewald", C.c_double), ("cut", C.c_double), ("buffer", C.c_double),
        ("nfft", C.c_int * 3), ("order", C.c_int), ("precision", C.c_int),
    ]


class SolveOpts(C.Structure):
    _fields_ = [("poleps", C.c_double), ("politer", C.c_int), ("guess", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [
        ("niter", C.c_int), ("status", C.c_int), ("resnrm", C.c_double), ("ep", C.c_double),
        ("n_indefinite", C.c_int), ("t_field_ms", C.c_double), ("t_solve_ms", C.c_double),
    ]


class ServeOpts(C.Structure):
    _fields_ = [
        ("max_systems", C.c_int), ("hbm_budget", C.c_int64), ("iters_per_graph", C.c_int),
        ("autotune", C.c_int), ("poleps", C.c_double), ("politer", C.c_int),
        ("warm_start", C.c_int), ("use_graphs", C.c_int),
    ]


class Result(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64), ("system_id", C.c_int), ("status", C.c_int), ("niter", C.c_int),
        ("ep", C.c_double), ("latency_ms", C.c_double), ("resnrm", C.c_double),
        ("queue_ms", C.c_double), ("warm", C.c_int),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("solves", C.c_int64), ("chunks", C.c_int64), ("mean_active", C.c_double),
        ("graph_launches", C.c_int64), ("evictions", C.c_int64), ("resident_bytes", C.c_int64),
        ("resident_systems", C.c_int), ("pair_segs", C.c_int), ("batch_builds", C.c_int64),
        ("host_ms", C.c_double), ("wait_ms", C.c_double), ("chunk_ms", C.c_double),
        ("admit_ms", C.c_double), ("setup_ms", C.c_double), ("coords_ms", C.c_double),
        ("lists_ms", C.c_double),
    ]


_lib = None


def load_lib(path: str | None = None) -> C.CDLL:
    """Load libthp.so (built by __graft_entry__.build()); fail loudly if absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ThpError(-8, f"{p} not built; run __graft_entry__.build()")
    lib = C.CDLL(p)
    vp = C.c_void_p
    sig = {
        "thp_init": (C.c_int, [C.POINTER(vp), C.c_int]),
        "thp_finalize": (C.c_int, [vp]),
        "thp_last_error": (C.c_char_p, [vp]),
        "thp_version": (C.c_char_p, []),
        "thp_system_create": (C.c_int, [vp, C.POINTER(SystemDesc), C.POINTER(vp)]),
        "thp_system_update_coords": (C.c_int, [vp, _dp, _dp, _dp]),
        "thp_system_update_multipoles": (C.c_int, [vp, _dp]),
        "thp_system_destroy": (C.c_int, [vp]),
        "thp_system_footprint": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "thp_nblist": (C.c_int, [vp, C.c_int]),
        "thp_nblist_get": (C.c_int, [vp, C.POINTER(C.c_int64), _ip, C.c_int64]),
        "thp_pairs_in_cut": (C.c_int, [vp, C.POINTER(C.c_int64), _ip, C.c_int64]),
        "thp_efld0_direct": (C.c_int, [vp, _dp]),
        "thp_efld0_recip": (C.c_int, [vp, _dp]),
        "thp_epolar1": (C.c_int, [vp, _dp, _dp]),
        "thp_epolar1_vir": (C.c_int, [vp, _dp, _dp, _dp]),
        "thp_tmatxb": (C.c_int, [vp, _dp, _dp, C.c_int]),
        "thp_tmatxb_recip": (C.c_int, [vp, _dp, _dp]),
        "thp_grid_uind": (C.c_int, [vp, _dp, _dp, _dp]),
        "thp_fphi_uind": (C.c_int, [vp, _dp, _dp, _dp]),
        "thp_pme_convolve": (C.c_int, [vp]),
        "thp_induce": (C.c_int, [vp, C.POINTER(SolveOpts), _dp, _dp, _dp, C.POINTER(SolveStats)]),
        "thp_induce_mu": (C.c_int, [vp, C.POINTER(SolveOpts), _dp, _dp, C.POINTER(SolveStats)]),
        "thp_set_predictor": (C.c_int, [vp, C.c_int]),
        "thp_set_solver": (C.c_int, [vp, C.c_int]),
        "thp_set_tcg": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
        "thp_set_shortreal": (C.c_int, [vp, C.c_double]),
        "thp_system_set_frames": (C.c_int, [vp, _ip, _ip, _ip, _ip, _dp]),
        "thp_tmatxb_dev": (C.c_int, [vp, vp, vp, C.c_int]),
        "thp_induce_dev": (C.c_int, [vp, C.POINTER(SolveOpts), vp, _dp, C.POINTER(SolveStats)]),
        "thp_efld0_direct_dev": (C.c_int, [vp, vp, vp]),
        "thp_efld0_recip_dev": (C.c_int, [vp, vp, vp]),
        "thp_tmatxb_dev_async": (C.c_int, [vp, vp, vp, C.c_int, vp]),
        "thp_tmatxb_recip_dev": (C.c_int, [vp, vp, vp, vp]),
        "thp_grid_uind_dev": (C.c_int, [vp, vp, vp, vp, vp]),
        "thp_fphi_uind_dev": (C.c_int, [vp, vp, vp, vp, vp]),
        "thp_epolar1_noalloc": (C.c_int, [vp, vp, vp, vp]),
        "thp_serve_set_frames": (C.c_int, [vp, C.c_int, _ip, _ip, _ip, _ip, _dp]),
        "thp_grid_shape": (C.c_int, [vp, _ip, C.c_int, _ip, _ip]),
        "thp_tcg_omega_fit": (C.c_int, [vp, _dp, _ip]),
        "thp_serve_start": (C.c_int, [vp, C.POINTER(ServeOpts)]),
        "thp_serve_add_system": (C.c_int, [vp, C.POINTER(SystemDesc), _ip]),
        "thp_serve_submit": (C.c_int, [vp, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_uint64)]),
        "thp_serve_submit_multipoles": (C.c_int, [vp, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint64)]),
        "thp_serve_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(Result), C.c_double]),
        "thp_serve_get_dipoles": (C.c_int, [vp, C.c_int, _dp]),
        "thp_serve_stats_get": (C.c_int, [vp, C.POINTER(ServeStats)]),
        "thp_serve_stop": (C.c_int, [vp]),
        "thp_time_kernel": (C.c_int, [vp, C.c_int, C.c_int, _dp, C.POINTER(C.c_int64)]),
        "thp_time_kernel_batch": (C.c_int, [C.POINTER(vp), C.c_int, C.c_int, C.c_int, _dp, C.POINTER(C.c_int64)]),
        "thp_make_system": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint64,
                                      _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp]),
        "thp_pick_grid": (C.c_int, [C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


# --------------------------------------------------------------- systems
@dataclass
class System:
    """Inputs of one polarization solve (caller order, FP64)."""
    n: int
    box: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    rpole: np.ndarray        # (n, 13)
    polarity: np.ndarray
    thole: np.ndarray
    scaled_ik: np.ndarray    # (m, 2) int32
    scaled_dpu: np.ndarray   # (m, 3)
    aewald: float = 0.4
    cut: float = 7.0
    buffer: float = 2.0
    nfft: tuple = (24, 24, 24)
    order: int = 5
    name: str = ""
    extra: dict = field(default_factory=dict)
    # local-frame multipoles: dict(axistype, zaxis, xaxis, yaxis, pole), each
    # solver rotates them on the device (rpole then serves only as storage)
    frames: object = None

    def with_coords(self, x, y, z):
        s = System(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        s.x, s.y, s.z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
        return s

    def with_multipoles(self, rpole):
        s = System(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        s.rpole = np.ascontiguousarray(rpole, dtype=np.float64).reshape(self.n, 13)
        return s

    def desc(self, precision: int) -> SystemDesc:
        d = SystemDesc()
        d.n = self.n
        for k in range(3):
            d.box[k] = float(self.box[k])
            d.nfft[k] = int(self.nfft[k])
        self._keep = [np.ascontiguousarray(a) for a in (
            self.x, self.y, self.z, self.rpole, self.polarity, self.thole)]
        self._keep_i = np.ascontiguousarray(self.scaled_ik, dtype=np.int32)
        self._keep_s = np.ascontiguousarray(self.scaled_dpu, dtype=np.float64)
        d.x, d.y, d.z, d.rpole, d.polarity, d.thole = (_d(a) for a in self._keep)
        d.nscaled = len(self._keep_i)
        d.scaled_ik = _i(self._keep_i)
        d.scaled_dpu = _d(self._keep_s)
        d.aewald, d.cut, d.buffer = self.aewald, self.cut, self.buffer
        d.order = self.order
        d.precision = precision
        return d


def make_system(n: int, box: float, n_solute: int = 0, blob_radius: float = 0.0, seed: int = 1,
                nfft=None, cut: float = 7.0, buffer: float = 2.0, aewald: float = 0.4,
                name: str = "") -> System:
    """Seeded water box (and optional solute blob) from libthp's generator."""
    lib = load_lib()
    x, y, z = (np.zeros(n) for _ in range(3))
    rpole = np.zeros((n, 13))
    pol, th = np.zeros(n), np.zeros(n)
    ik = np.zeros((3 * n, 2), dtype=np.int32)
    dpu = np.zeros((3 * n, 3))
    ns = C.c_int(0)
    rc = lib.thp_make_system(n, box, n_solute, blob_radius, seed, _d(x), _d(y), _d(z), _d(rpole),
                             _d(pol), _d(th), C.byref(ns), _i(ik), _d(dpu))
    if rc != THP_OK:
        raise ThpError(rc, "thp_make_system")
    if nfft is None:
        k = lib.thp_pick_grid(box)
        nfft = (k, k, k)
    m = ns.value
    return System(n=n, box=np.array([box, box, box], dtype=np.float64), x=x, y=y, z=z, rpole=rpole,
                  polarity=pol, thole=th, scaled_ik=ik[:m].copy(), scaled_dpu=dpu[:m].copy(),
                  aewald=aewald, cut=cut, buffer=buffer, nfft=tuple(int(v) for v in nfft), name=name)


# The workloads (water recipe, seeded); sizes as named by the benchmark.
def water_1944(seed=1):
    """648 waters, 18.643 A box, 24^3 grid (the CPU-runnable case)."""
    return make_system(1944, 18.643, seed=seed, nfft=(24, 24, 24), name="water-1944")


def water_23559(seed=1):
    """7,853 waters, 62.23 A box, 64^3 grid (the 24k-atom headline box)."""
    return make_system(23559, 62.23, seed=seed, nfft=(64, 64, 64), name="water-23559")


def solute_96000(seed=1):
    """96,000 atoms, 100 A box, 20,000-atom solute blob of radius 30 A, 108^3."""
    return make_system(96000, 100.0, n_solute=20000, blob_radius=30.0, seed=seed,
                       nfft=(108, 108, 108), name="solute-96000")


def water_1068000(seed=1):
    """356,000 waters, 220 A box, 240^3 grid."""
    return make_system(1068000, 220.0, seed=seed, nfft=(240, 240, 240), name="water-1068000")


def serving_system(n: int, solute_frac: float, seed: int) -> System:
    """Serving-stream system: water box at 0.1 atoms/A^3 with a solute blob.
    Buffer min(2, L/2 - cut - 0.05) keeps cut+buffer inside half the box."""
    L = (n / 0.1) ** (1.0 / 3.0)
    ns = int(round(solute_frac * n))
    R = (3.0 * ns / (4.0 * np.pi * 0.177)) ** (1.0 / 3.0) if ns > 0 else 0.0
    buf = min(2.0, 0.5 * L - 7.0 - 0.05)
    if buf <= 0:
        raise ValueError(f"box {L:.2f} too small for a 7 A cutoff")
    return make_system(n, L, n_solute=ns, blob_radius=R, seed=seed, buffer=buf,
                       name=f"serve-{n}-{seed}")


def tiny_water(nmol: int, box: float, nfft: int, seed=1, cut=None, buffer=None):
    """Small water box for parity tests (cut+buffer fitted to the box)."""
    c = cut if cut is not None else min(7.0, 0.5 * box - 0.6)
    b = buffer if buffer is not None else min(2.0, 0.5 * box - c - 0.05)
    return make_system(3 * nmol, box, seed=seed, nfft=(nfft,) * 3, cut=c, buffer=b,
                       name=f"water-{3 * nmol}")


# ---------------------------------------------------------------- solver
class Context:
    def __init__(self, device: int = 0):
        self.lib = load_lib()
        self.h = C.c_void_p()
        self._solvers = weakref.WeakSet()
        rc = self.lib.thp_init(C.byref(self.h), device)
        if rc != THP_OK:
            msg = self.lib.thp_last_error(self.h).decode() if self.h else ""
            raise ThpError(rc, msg)

    def check(self, rc, what=""):
        if rc != THP_OK:
            raise ThpError(rc, f"{what}: {self.lib.thp_last_error(self.h).decode()}")

    def grid_shape(self, nfft, precision=THP_PREC_MIXED):
        """(largest brick edge, most FFT lines per workgroup) the serving
        loop's autotuner picked for this grid on this context, or None."""
        nf = np.ascontiguousarray(nfft, dtype=np.int32).reshape(3)
        b, f = C.c_int(0), C.c_int(0)
        rc = self.lib.thp_grid_shape(self.h, _i(nf), precision, C.byref(b), C.byref(f))
        return (b.value, f.value) if rc == THP_OK else None

    def close(self):
        if self.h:
            for s in list(self._solvers):
                s.close()
            self.lib.thp_finalize(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """One resident system on the GPU; methods mirror the reference routines."""

    def __init__(self, ctx: Context, system: System, precision: int = THP_PREC_FP64):
        self.ctx, self.sys, self.precision = ctx, system, precision
        self.lib = ctx.lib
        self.h = C.c_void_p()
        d = system.desc(precision)
        ctx.check(self.lib.thp_system_create(ctx.h, C.byref(d), C.byref(self.h)), "thp_system_create")
        ctx._solvers.add(self)
        if system.frames is not None:
            self.set_frames(**system.frames)

    def set_frames(self, axistype, zaxis, xaxis, yaxis, pole):
        """Local-frame multipoles (rotpole + chkpole on the device each step):
        axis types THP_AXIS_*, frame atoms (0-based, -1 for none), pole (n, 13)
        in the local frame."""
        n = self.sys.n
        arrs = [np.ascontiguousarray(v, dtype=np.int32).reshape(n) for v in (axistype, zaxis, xaxis, yaxis)]
        pl = np.ascontiguousarray(pole, dtype=np.float64).reshape(n, 13)
        self._c(self.lib.thp_system_set_frames(self.h, *(_i(v) for v in arrs), _d(pl)), "set_frames")

    def close(self):
        # a finalized context has already released its systems
        if self.h and self.ctx.h:
            self.lib.thp_system_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        self.ctx.check(rc, what)

    def footprint(self) -> int:
        b = C.c_int64(0)
        self._c(self.lib.thp_system_footprint(self.h, C.byref(b)), "footprint")
        return b.value

    def update_coords(self, x, y, z):
        x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
        self._c(self.lib.thp_system_update_coords(self.h, _d(x), _d(y), _d(z)), "update_coords")

    def update_multipoles(self, rpole):
        """New global-frame multipoles rpole (n, 13), e.g. after rotpole."""
        rp = np.ascontiguousarray(rpole, dtype=np.float64).reshape(self.sys.n, 13)
        self._c(self.lib.thp_system_update_multipoles(self.h, _d(rp)), "update_multipoles")

    def nblist(self, force_rebuild=True):
        self._c(self.lib.thp_nblist(self.h, 1 if force_rebuild else 0), "nblist")

    def nblist_pairs(self) -> np.ndarray:
        cnt = C.c_int64(0)
        self._c(self.lib.thp_nblist_get(self.h, C.byref(cnt), None, 0), "nblist_get")
        ik = np.zeros((cnt.value, 2), dtype=np.int32)
        self._c(self.lib.thp_nblist_get(self.h, C.byref(cnt), _i(ik), cnt.value), "nblist_get")
        return ik

    def pairs_in_cut(self) -> np.ndarray:
        """(i, k) rows, i < k, of the pairs the solver uses (within the cutoff)"""
        cnt = C.c_int64(0)
        self._c(self.lib.thp_pairs_in_cut(self.h, C.byref(cnt), None, 0), "pairs_in_cut")
        ik = np.zeros((cnt.value, 2), dtype=np.int32)
        self._c(self.lib.thp_pairs_in_cut(self.h, C.byref(cnt), _i(ik), cnt.value), "pairs_in_cut")
        return ik

    def efld0_direct(self) -> np.ndarray:
        ef = np.zeros((self.sys.n, 2, 3))
        self._c(self.lib.thp_efld0_direct(self.h, _d(ef)), "efld0_direct")
        return ef

    def efld0_recip(self) -> np.ndarray:
        cphi = np.zeros((self.sys.n, 10))
        self._c(self.lib.thp_efld0_recip(self.h, _d(cphi)), "efld0_recip")
        return cphi

    def epolar1_vir(self):
        """Gradient dep (n, 3) and the polarization virial (3, 3), kcal/mol
        (epolar1c's dep and vir)."""
        dep = np.zeros((self.sys.n, 3))
        vir = np.zeros((3, 3))
        self._c(self.lib.thp_epolar1_vir(self.h, _d(dep), None, _d(vir)), "epolar1_vir")
        return dep, vir

    def epolar1(self, with_torque: bool = False):
        """Polarization energy gradient dep (n, 3), kcal/mol/A, for the
        dipoles of the last solve (epolar1c); with_torque also returns the
        site torques (n, 3), which dep leaves out for global multipoles."""
        dep = np.zeros((self.sys.n, 3))
        tq = np.zeros((self.sys.n, 3)) if with_torque else None
        self._c(self.lib.thp_epolar1(self.h, _d(dep), _d(tq) if tq is not None else None), "epolar1")
        return (dep, tq) if with_torque else dep

    def tmatxb_pme(self, mu, dodiag=True) -> np.ndarray:
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(self.sys.n, 2, 3)
        efi = np.zeros_like(mu)
        self._c(self.lib.thp_tmatxb(self.h, _d(mu), _d(efi), 1 if dodiag else 0), "tmatxb")
        return efi

    def tmatxbrecip(self, mu) -> np.ndarray:
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(self.sys.n, 2, 3)
        out = np.zeros_like(mu)
        self._c(self.lib.thp_tmatxb_recip(self.h, _d(mu), _d(out)), "tmatxb_recip")
        return out

    def grid_uind_site(self, fuind, fuinp) -> np.ndarray:
        fd = np.ascontiguousarray(fuind, dtype=np.float64).reshape(self.sys.n, 3)
        fp = np.ascontiguousarray(fuinp, dtype=np.float64).reshape(self.sys.n, 3)
        k1, k2, k3 = self.sys.nfft
        g = np.zeros((k3, k2, k1, 2))
        self._c(self.lib.thp_grid_uind(self.h, _d(fd), _d(fp), _d(g)), "grid_uind")
        return g

    def fphi_uind_site(self, qgrid=None):
        p1 = np.zeros((self.sys.n, 3))
        p2 = np.zeros((self.sys.n, 3))
        g = None if qgrid is None else np.ascontiguousarray(qgrid, dtype=np.float64)
        self._c(self.lib.thp_fphi_uind(self.h, None if g is None else _d(g), _d(p1), _d(p2)), "fphi_uind")
        return p1, p2

    def pme_convolve(self):
        self._c(self.lib.thp_pme_convolve(self.h), "pme_convolve")

    def set_solver(self, polalg: int):
        """THP_POLALG_PCG (1), THP_POLALG_JACOBI_DIIS (2) or THP_POLALG_TCG (3)."""
        self._c(self.lib.thp_set_solver(self.h, int(polalg)), "set_solver")

    def set_shortreal(self, shortcut: float):
        """Short-range real-space mode (use_polarshortreal) with this cutoff;
        0 restores the full solve."""
        self._c(self.lib.thp_set_shortreal(self.h, float(shortcut)), "set_shortreal")

    def set_tcg(self, order: int = 2, precond: bool = True, guess: bool = False, peek: bool = True,
                omega: float = 1.0):
        """Options of the truncated CG (the reference's defaults)."""
        self._c(self.lib.thp_set_tcg(self.h, int(order), int(bool(precond)), int(bool(guess)), int(bool(peek)),
                                     float(omega)), "set_tcg")

    def tcg_omega_fit(self):
        """Fit the peek step's omega to the PCG energy at poleps 1e-10
        (tcgomegafit); returns (omega in use, fitted)."""
        w, f = C.c_double(0), C.c_int(0)
        self._c(self.lib.thp_tcg_omega_fit(self.h, C.byref(w), C.byref(f)), "tcg_omega_fit")
        return w.value, bool(f.value)

    def set_predictor(self, method: int):
        """THP_PRED_NONE/GEAR/ASPC/LSQR; clears the stored solves."""
        self._c(self.lib.thp_set_predictor(self.h, int(method)), "set_predictor")

    def newinduce_pme2(self, poleps=1e-8, politer=500, mu0=None):
        """Induced dipoles uind/uinp (n,3) each, E_pol (kcal/mol) and stats."""
        n = self.sys.n
        guess = 0
        if mu0 is not None:
            # a copy: the solve writes its dipoles over the starting ones
            mu = np.array(mu0, dtype=np.float64, order="C").reshape(n, 2, 3)
            guess = 1
        else:
            mu = np.empty((n, 2, 3))
        o = SolveOpts(poleps=poleps, politer=politer, guess=guess)
        st = SolveStats()
        ep = C.c_double(0)
        self._c(self.lib.thp_induce_mu(self.h, C.byref(o), _d(mu), C.byref(ep), C.byref(st)), "induce")
        return mu, ep.value, st

    def newinduce_pme2_split(self, poleps=1e-8, politer=500):
        """The same solve through thp_induce (the reference's separate uind
        and uinp arrays): (uind, uinp, E_pol, stats)."""
        n = self.sys.n
        uind, uinp = np.empty((n, 3)), np.empty((n, 3))
        o = SolveOpts(poleps=poleps, politer=politer, guess=0)
        st = SolveStats()
        ep = C.c_double(0)
        self._c(self.lib.thp_induce(self.h, C.byref(o), _d(uind), _d(uinp), C.byref(ep), C.byref(st)),
                "induce")
        return uind, uinp, ep.value, st

    induce = newinduce_pme2

    def time_kernel(self, kernel: int, reps: int = 20):
        """(ms per launch, unscaled pairs within the cutoff) of a path kernel:
        0 = real-space matvec, 1 = spreading, 2 = gather, 3 = permanent-field
        pass."""
        ms = C.c_double(0)
        pairs = C.c_int64(0)
        self._c(self.lib.thp_time_kernel(self.h, kernel, reps, C.byref(ms), C.byref(pairs)), "time_kernel")
        return ms.value, pairs.value


def time_kernel_batch(solvers, kernel: int, reps: int = 20):
    """(ms per launch, unscaled pairs within the cutoff summed over the
    systems) of one launch of a path kernel over the systems of `solvers`
    (distinct, one context, one precision), laid out as a serving batch."""
    if not solvers:
        raise ValueError("no systems")
    lib = solvers[0].lib
    hs = (C.c_void_p * len(solvers))(*[s.h for s in solvers])
    ms = C.c_double(0)
    pairs = C.c_int64(0)
    solvers[0].ctx.check(lib.thp_time_kernel_batch(hs, len(solvers), kernel, reps, C.byref(ms), C.byref(pairs)),
                         "time_kernel_batch")
    return ms.value, pairs.value
