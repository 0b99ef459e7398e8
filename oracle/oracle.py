"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker of the HIP product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp,
                            C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_set_coords": (None, [vp, _dp, _dp, _dp]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_nblist": (C.c_long, [vp]),
        "orc_nblist_pairs": (None, [vp, _ip]),
        "orc_efld0_direct": (None, [vp, _dp]),
        "orc_efld0_recip": (None, [vp, _dp]),
        "orc_efld0_total": (None, [vp, _dp]),
        "orc_tmatxb": (None, [vp, _dp, _dp, C.c_int]),
        "orc_tmatxb_recip": (None, [vp, _dp, _dp]),
        "orc_matvec_full": (None, [vp, _dp, _dp]),
        "orc_induce": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip, _dp, _dp]),
        "orc_epolar_decomp": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_bsplgen": (None, [C.c_double, C.c_int, _dp]),
        "orc_bsmod": (None, [C.c_int, C.c_int, _dp]),
        "orc_bsmod_raw": (None, [C.c_int, C.c_int, _dp]),
        "orc_fft3d": (None, [C.c_int, C.c_int, C.c_int, _dp, C.c_int]),
        "orc_dampewald": (None, [C.c_double, C.c_double, _dp]),
        "orc_damptholed": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
        "orc_erfc": (C.c_double, [C.c_double]),
        "orc_plant_scaled_form": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def set_threads(n: int):
    load().orc_set_threads(int(n))


class Oracle:
    """FP64 restatement of the reference CPU path for one system."""

    def __init__(self, system, aewald=None, cut=None, buffer=None, nfft=None):
        lib = load()
        s = system
        self.n = s.n
        self.sys = s
        self._arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (s.box, s.x, s.y, s.z, s.rpole, s.polarity, s.thole)]
        self._ik = np.ascontiguousarray(s.scaled_ik, dtype=np.int32).reshape(-1)
        self._dpu = np.ascontiguousarray(s.scaled_dpu, dtype=np.float64).reshape(-1)
        nf = nfft if nfft is not None else s.nfft
        self.h = lib.orc_create(s.n, *(_d(a) for a in self._arrs), len(self._ik) // 2,
                                self._ik.ctypes.data_as(_ip), _d(self._dpu),
                                aewald if aewald is not None else s.aewald,
                                cut if cut is not None else s.cut,
                                buffer if buffer is not None else s.buffer,
                                int(nf[0]), int(nf[1]), int(nf[2]), s.order)
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.lib = lib

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def set_coords(self, x, y, z):
        self._xyz = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        self.lib.orc_set_coords(self.h, *(_d(v) for v in self._xyz))

    def nblist_pairs(self) -> np.ndarray:
        cnt = self.lib.orc_nblist(self.h)
        ik = np.zeros((cnt, 2), dtype=np.int32)
        self.lib.orc_nblist_pairs(self.h, ik.ctypes.data_as(_ip))
        return ik

    def efld0_direct(self):
        ef = np.zeros((self.n, 2, 3))
        self.lib.orc_efld0_direct(self.h, _d(ef))
        return ef

    def efld0_recip(self):
        c = np.zeros((self.n, 10))
        self.lib.orc_efld0_recip(self.h, _d(c))
        return c

    def efld0_total(self):
        ef = np.zeros((self.n, 2, 3))
        self.lib.orc_efld0_total(self.h, _d(ef))
        return ef

    def tmatxb_pme(self, mu, dodiag=True):
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(self.n, 2, 3)
        out = np.zeros_like(mu)
        self.lib.orc_tmatxb(self.h, _d(mu), _d(out), 1 if dodiag else 0)
        return out

    def tmatxbrecip(self, mu):
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(self.n, 2, 3)
        out = np.zeros_like(mu)
        self.lib.orc_tmatxb_recip(self.h, _d(mu), _d(out))
        return out

    def matvec(self, mu):
        mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(self.n, 2, 3)
        out = np.zeros_like(mu)
        self.lib.orc_matvec_full(self.h, _d(mu), _d(out))
        return out

    def induce(self, poleps=1e-8, politer=500, mu0=None):
        mu = np.zeros((self.n, 2, 3))
        ef = np.zeros((self.n, 2, 3))
        m0 = np.ascontiguousarray(mu0 if mu0 is not None else mu, dtype=np.float64)
        ep = C.c_double(0)
        nit = C.c_int(0)
        rn = C.c_double(0)
        hist = np.zeros((politer, 2))
        st = self.lib.orc_induce(self.h, poleps, politer, 1 if mu0 is not None else 0, _d(m0), _d(mu),
                                 _d(ef), C.byref(ep), C.byref(nit), C.byref(rn), _d(hist))
        return dict(mu=mu, ef=ef, ep=ep.value, niter=nit.value, resnrm=rn.value, status=st,
                    ene_hist=hist[:nit.value] if nit.value > 0 else hist[:0])

    def epolar_decomp(self, uind):
        uind = np.ascontiguousarray(uind, dtype=np.float64).reshape(self.n, 3)
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        self.lib.orc_epolar_decomp(self.h, _d(uind), C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value


def bsplgen(w, order=5):
    th = np.zeros((order, 4))
    load().orc_bsplgen(w, order, _d(th))
    return th


def bsmod(nfft, order=5, zeta=True):
    b = np.zeros(nfft)
    (load().orc_bsmod if zeta else load().orc_bsmod_raw)(nfft, order, _d(b))
    return b


def fft3d(grid, sign):
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    n3, n2, n1 = g.shape
    load().orc_fft3d(n1, n2, n3, g.ctypes.data_as(_dp), sign)
    return g


def dampewald(aewald, r):
    d = np.zeros(7)
    load().orc_dampewald(aewald, r, _d(d))
    return d


def dampthole(pdi, pdk, pgamma, r, order=7):
    d = np.zeros(9)
    load().orc_damptholed(pdi, pdk, pgamma, r, order, _d(d))
    return d


# ------------------------------------------------------------ dipole predictor
# Gear binomial and ASPC (Kolafa 2004) coefficients of kpolar.f:405-423,
# newest history entry first.
PRED_COEF = {
    "GEAR": (6.0, -15.0, 20.0, -15.0, 6.0, -1.0, 0.0),
    "ASPC": (22.0 / 7.0, -55.0 / 14.0, 55.0 / 21.0, -22.0 / 21.0, 5.0 / 21.0, -1.0 / 42.0, 0.0),
}
MAXUALT = 7  # MOD_uprior.f:34


def ldlt_solve(a, b):
    """x with a x = b for symmetric positive definite a through the
    symmetric decomposition a = L D L^T, L unit lower triangular (cholesky.f,
    Martin, Peters and Wilkinson 1965)."""
    a = np.array(a, dtype=np.float64)
    n = len(b)
    L = np.eye(n)
    d = np.zeros(n)
    for i in range(n):
        for j in range(i):
            s = a[i, j]
            for k in range(j):
                s -= L[i, k] * d[k] * L[j, k]
            L[i, j] = s / d[j]
        s = a[i, i]
        for k in range(i):
            s -= L[i, k] * L[i, k] * d[k]
        d[i] = s
    y = np.array(b, dtype=np.float64)
    for i in range(n):
        for k in range(i):
            y[i] -= L[i, k] * y[k]
    x = y / d
    for i in reversed(range(n)):
        for k in range(i + 1, n):
            x[i] -= L[k, i] * x[k]
    return x


class Predictor:
    """History of converged dipoles and their extrapolation into the next
    solve's guess: ulspred (newinduce_pme2.f:910-1010), its use
    (newinduce_pme2.f:165-178) and the history update (:253-264).  The guess
    applies the coefficients fitted on the d dipoles (bpred) to both d and p,
    as the reference does."""

    def __init__(self, method: str):
        assert method in ("GEAR", "ASPC", "LSQR")
        self.method = method
        self.hist = []  # newest first, each (n, 2, 3)

    def push(self, mu):
        self.hist.insert(0, np.array(mu, dtype=np.float64).reshape(-1, 2, 3))
        del self.hist[MAXUALT:]

    def coefficients(self):
        if self.method in PRED_COEF:
            return np.array(PRED_COEF[self.method])
        # least-squares fit of the newest d dipoles on the six older ones
        u = [h[:, 0, :].reshape(-1) for h in self.hist]
        c = np.array([[float(np.dot(u[k], u[m])) for m in range(MAXUALT)] for k in range(MAXUALT)])
        b = c[0, 1:]
        a = c[1:, 1:]
        coef = np.zeros(MAXUALT)
        coef[:-1] = ldlt_solve(a, b) if np.max(np.triu(a)) != 0.0 else b
        return coef

    def predict(self):
        """Guess (n, 2, 3), or None while fewer than seven solves are stored."""
        if len(self.hist) < MAXUALT:
            return None
        b = self.coefficients()
        return sum(b[k] * self.hist[k] for k in range(MAXUALT - 1))


# ------------------------------------------------------------ Jacobi / DIIS
NDISMX = 25  # inducejac_pme2 (newinduce_pme2.f:664)


def jacobi_diis(hfun, ef, alpha, poleps, politer, mu0=None):
    """Jacobi iterations mu <- alpha (ef - h(mu)) with DIIS extrapolation:
    inducejac_pme2 (newinduce_pme2.f:638-890) with diis, makeb and extrap
    (diis.f), the subspace restarting from its ten newest vectors when it
    reaches NDISMX.  The small linear system goes to numpy's LAPACK gesv,
    the routine the reference calls.  Vectors are (n, 2, 3); the norm is
    sqrt(sum over one rhs / (3 n)) as in the reference."""
    shape = ef.shape
    n = shape[0]
    mu = alpha * ef if mu0 is None else np.array(mu0, dtype=np.float64).reshape(shape)
    xd, ed = [], []  # subspace vectors, oldest first
    bmat = np.zeros((NDISMX + 1, NDISMX + 1))
    nmat = 1
    rr = np.inf
    it = 0
    restarts = 0
    for it in range(1, politer + 1):
        munew = alpha * (ef - hfun(mu))
        rnorm = ((munew - mu) ** 2).sum(axis=(0, 2))
        xd.append(munew.reshape(-1).copy())
        ed.append((munew - mu).reshape(-1))
        if nmat >= NDISMX:  # restart: drop the ten oldest vectors
            bmat[1:nmat - 10, 1:nmat - 10] = bmat[11:nmat, 11:nmat].copy()
            xd, ed = xd[10:], ed[10:]
            nmat -= 10
            restarts += 1
        # makeb for vector nmat (1-based), Lagrange row and column 0
        if nmat == 1:
            bmat[0, 0] = 0.0
            bmat[0, 1] = bmat[1, 0] = 1.0
            bmat[1, 1] = float(np.dot(ed[0], ed[0]))
        else:
            bmat[nmat, 0] = bmat[0, nmat] = 1.0
            for j in range(1, nmat):
                bmat[nmat, j] = bmat[j, nmat] = float(np.dot(ed[j - 1], ed[nmat - 1]))
            bmat[nmat, nmat] = float(np.dot(ed[nmat - 1], ed[nmat - 1]))
        nmat += 1
        rhs = np.zeros(nmat)
        rhs[0] = 1.0
        cex = np.linalg.solve(bmat[:nmat, :nmat], rhs)
        mu = sum(cex[j + 1] * xd[j] for j in range(nmat - 1)).reshape(shape)
        rr = float(np.max(np.sqrt(rnorm / (3.0 * n))))
        if rr < poleps:
            break
    return mu, it, rr, restarts


def induce_jacobi_diis(orc, poleps=1e-8, politer=500, mu0=None):
    """The reference's polalg = 2 solve of one system: h(mu) = -E_real(mu)
    without the diagonal + the reciprocal dipole field - term mu
    (newinduce_pme2.f:739-755).  Returns dict(mu, ep, niter, resnrm)."""
    s = orc.sys
    ef = orc.efld0_total()
    alpha = np.asarray(s.polarity, dtype=np.float64)[:, None, None]
    term = (4.0 / 3.0) * s.aewald ** 3 / np.sqrt(np.pi)
    hfun = lambda mu: orc.tmatxb_pme(mu, dodiag=False) + orc.tmatxbrecip(mu) - term * mu
    mu, it, rr, _ = jacobi_diis(hfun, ef, alpha, poleps, politer, mu0)
    f = 332.063714  # electric (MOD_units.f:81)
    ep = -0.5 * f * float(np.sum(mu[:, 0, :] * ef[:, 1, :]))
    return dict(mu=mu, ep=ep, niter=it, resnrm=rr)


def tcg(afun, ef, alpha, order=2, prec=True, guess=False, peek=True, omega=1.0):
    """Truncated conjugate gradient of order 1 or 2: epolar3tcg1 and
    epolar3tcg2 (epolar3tcg.f:66-311 and 320-621).  afun applies the
    polarization matrix A = alpha^-1 - T (tmatxb with the diagonal, the
    reciprocal dipole field, minus the self term); alpha multiplies like
    diagvec (newinduce_pme2.f:1017-1038).  From mu0 = alpha ef (tcgguess) or
    zero: r0 = ef - A mu0, preconditioned by alpha (tcgprec); one (order 1)
    or two (order 2) conjugate-gradient steps in closed form; then the
    optional peek step omega alpha r (tcgpeek), whose scalars are those of
    the first right-hand side for both, as in the reference.  Vectors are
    (n, 2, 3).  Returns (mu, scalars)."""
    ef = np.asarray(ef, dtype=np.float64)

    def dot(a, b):
        return np.einsum("nrc,nrc->r", a, b)

    def col(v):
        return np.asarray(v)[None, :, None]

    def pre(v):
        return alpha * v if not prec else v  # the alpha of the peek step

    mu0 = alpha * ef if guess else np.zeros_like(ef)
    oldr0 = ef - afun(mu0) if guess else ef.copy()
    r0 = alpha * oldr0 if prec else oldr0.copy()
    oldTr0 = afun(r0)
    Tr0 = alpha * oldTr0 if prec else oldTr0.copy()
    n0 = dot(oldr0, r0)
    t1 = dot(r0, oldTr0)
    t4 = n0 / t1
    sc = dict(n0=n0, t1=t1, t4=t4)
    if order == 1:
        mu = col(t4) * r0
        if guess:
            mu = mu + mu0
        if peek:
            mu = mu + (omega * pre(r0) - omega * t4[0] * pre(Tr0))
        return mu, sc
    if order != 2:
        raise ValueError("tcg order is 1 or 2")
    mupeek = omega * pre(r0) - omega * t4[0] * pre(Tr0) if peek else 0.0
    oldT2r0 = afun(Tr0)
    T2r0 = alpha * oldT2r0 if prec else oldT2r0
    np1 = dot(oldTr0, Tr0)
    t2 = n0 * np1 / (t1 * t1)
    p2 = col(t2) * Tr0 - col(t4) * T2r0
    t3 = t1 * dot(oldTr0, p2)
    g1 = (t1 * t1 - n0 * np1) / t3
    mu = col(g1 * t2 + t4) * r0 - col(g1 * t4) * Tr0
    if guess:
        mu = mu + mu0
    if peek:
        mupeek = mupeek - omega * g1[0] * t2[0] * pre(Tr0) + omega * g1[0] * t4[0] * pre(T2r0)
        mu = mu + mupeek
    sc.update(np1=np1, t2=t2, t3=t3, g1=g1)
    return mu, sc


def induce_tcg(orc, order=2, prec=True, guess=False, peek=True, omega=1.0, short=False):
    """The reference's polalg = 3 energy of one system (epolar3tcg1/2):
    dipoles from tcg() on the full permanent field, E_pol = -1/2 f
    ef_p . mu_d.  short: the short-range form of the RESPA1 inner steps
    (epolar1tcg1shortreal.f:100-312 and its order-2 siblings): the direct
    field and the real-space matvec with its diagonal at the oracle's
    cutoff, which the caller sets to the short cutoff, no reciprocal or self
    terms.  Returns dict(mu, ep)."""
    s = orc.sys
    if short:
        ef = orc.efld0_direct()
        afun = lambda v: orc.tmatxb_pme(v, dodiag=True)
    else:
        ef = orc.efld0_total()
        afun = orc.matvec
    alpha = np.asarray(s.polarity, dtype=np.float64)[:, None, None]
    mu, sc = tcg(afun, ef, alpha, order, prec, guess, peek, omega)
    f = 332.063714  # electric (MOD_units.f:81)
    ep = -0.5 * f * float(np.sum(mu[:, 0, :] * ef[:, 1, :]))
    return dict(mu=mu, ep=ep, scalars=sc)


def tcg_omega_fit(orc, order=2, prec=True, guess=False, omega=1.0):
    """omegafit (kpolar.f:1059-1326): E_CG from the PCG solve at poleps
    1e-10; the truncated CG energy with the peek step, linear in omega,
    from its values at omega = 0 and 1 (the reference's cnp and cr terms);
    bisection over [-5, 5] toward E_CG, 20 steps, within max(1e-4 |E_CG|,
    1e-6) kcal/mol.  Returns (omega, fitted, dict of the energies)."""
    epcg = orc.induce(poleps=1e-10)["ep"]
    e0 = induce_tcg(orc, order, prec, guess, True, 0.0)["ep"]
    e1 = induce_tcg(orc, order, prec, guess, True, 1.0)["ep"]
    er = e1 - e0
    cvg = abs(epcg * 1e-4)
    if cvg <= 1e-6:
        cvg = 1e-6
    o1, o2 = -5.0, 5.0
    ep1, ep2 = e0 + o1 * er, e0 + o2 * er
    w, done = omega, False
    if abs(er) > cvg:
        for _ in range(20):
            if abs(epcg - ep1) <= cvg:
                w, done = o1, True
                break
            if abs(epcg - ep2) <= cvg:
                w, done = o2, True
                break
            if abs(epcg - ep1) >= abs(epcg - ep2):
                o1 = o1 + (o2 - o1) * 0.5
                ep1 = e0 + o1 * er
            else:
                o2 = o2 - (o2 - o1) * 0.5
                ep2 = e0 + o2 * er
    return w, done, dict(epcg=epcg, e0=e0, e1=e1, cvg=cvg)


def pcg(afun, ef, diag, mu0, poleps, politer):
    """The preconditioned conjugate gradient of inducepcg_pme2 and
    inducepcg_shortreal (newinduce_pme2.f:285-636; newinduce_shortreal.f:
    186-425) on an operator afun: residual norm sqrt(r.(diag r) / 3n) per
    right-hand side, stop when the larger is below poleps.  Vectors are
    (n, 2, 3).  Returns (mu, iterations, residual norm)."""
    n = ef.shape[0]
    mu = np.array(mu0, dtype=np.float64)
    res = ef - afun(mu)
    zr = diag * res
    pp = zr.copy()
    ggold = (res * zr).sum(axis=(0, 2))
    it, rr = 0, np.inf
    for it in range(1, politer + 1):
        h = afun(pp)
        gg = (pp * h).sum(axis=(0, 2))
        if np.any(gg == 0.0):
            break
        a = ggold / gg
        mu = mu + a[None, :, None] * pp
        res = res - a[None, :, None] * h
        zr = diag * res
        ggnew = (res * zr).sum(axis=(0, 2))
        rr = float(np.sqrt(ggnew / (3.0 * n)).max())
        pp = zr + (ggnew / ggold)[None, :, None] * pp
        ggold = ggnew
        if rr < poleps:
            break
    return mu, it, rr


TINYPOL = 1.0e-5  # tinypol (v1.2/source/MOD_polar.f:36)


def induce_shortreal(orc, poleps=1e-8, politer=500, jacobi=False):
    """newinduce_shortreal (newinduce_shortreal.f:16-185) of one system: the
    direct permanent field and the real-space matvec only (no reciprocal
    field, no self term) at the oracle's cutoff, which the caller sets to
    the short cutoff; guess alpha ef; PCG (inducepcg_shortreal) or
    Jacobi/DIIS (inducejac_shortreal).  Returns dict(mu, ep, niter,
    resnrm)."""
    s = orc.sys
    ef = orc.efld0_direct()
    alpha = np.asarray(s.polarity, dtype=np.float64)[:, None, None]
    if jacobi:
        hfun = lambda mu: orc.tmatxb_pme(mu, dodiag=False)
        mu, it, rr, _ = jacobi_diis(hfun, ef, alpha, poleps, politer)
    else:
        diag = np.where(alpha == 0.0, TINYPOL, alpha)
        mu, it, rr = pcg(lambda v: orc.tmatxb_pme(v, dodiag=True), ef, diag, alpha * ef, poleps, politer)
    f = 332.063714  # electric (MOD_units.f:81)
    ep = -0.5 * f * float(np.sum(mu[:, 0, :] * ef[:, 1, :]))
    return dict(mu=mu, ep=ep, niter=it, resnrm=rr)


AXIS_NONE, AXIS_Z_ONLY, AXIS_Z_THEN_X, AXIS_BISECTOR, AXIS_Z_BISECT, AXIS_3_FOLD = range(6)


def rotmat(xyz, i, axt, iz, ix, iy):
    """Rotation matrix of site i (rotmat, rotpole.f:61-254): columns are the
    local x, y, z axes in the global frame.  Z-Only takes the global x (or
    y) axis where the reference draws a random vector."""
    a = np.eye(3)
    if axt == AXIS_NONE:
        return a
    unit = lambda v: v / np.linalg.norm(v)
    d1 = unit(xyz[iz] - xyz[i])
    if axt == AXIS_Z_ONLY:
        z = d1
        u = np.array([1.0, 0.0, 0.0]) if abs(z[0]) < 0.866 else np.array([0.0, 1.0, 0.0])
    elif axt == AXIS_Z_THEN_X:
        z = d1
        u = xyz[ix] - xyz[i]
    elif axt == AXIS_BISECTOR:
        d2 = unit(xyz[ix] - xyz[i])
        z = unit(d1 + d2)
        u = d2
    elif axt == AXIS_Z_BISECT:
        z = d1
        u = unit(unit(xyz[ix] - xyz[i]) + unit(xyz[iy] - xyz[i]))
    else:
        d2 = unit(xyz[ix] - xyz[i])
        z = unit(d1 + d2 + unit(xyz[iy] - xyz[i]))
        u = d2
    x = unit(u - np.dot(u, z) * z)
    a[:, 2] = z
    a[:, 0] = x
    a[:, 1] = np.cross(z, x)
    return a


def rotpole(x, y, z, axistype, zaxis, xaxis, yaxis, pole):
    """Global-frame multipoles (n, 13) from local ones: chkpole
    (chkpole.f:20-98: a Z-then-X site with a y-axis atom has its y
    components inverted while the parallelepiped volume is negative), then
    rotsite (rotpole.f:261-339): d = A d_local, Q = A Q_local A^T."""
    xyz = np.stack([x, y, z], axis=1).astype(np.float64)
    n = len(x)
    out = np.zeros((n, 13))
    for i in range(n):
        p = np.array(pole[i], dtype=np.float64)
        t = int(axistype[i])
        if t == AXIS_Z_THEN_X and yaxis[i] >= 0:
            ad, bd, cd = xyz[i] - xyz[yaxis[i]], xyz[zaxis[i]] - xyz[yaxis[i]], xyz[xaxis[i]] - xyz[yaxis[i]]
            vol = np.dot(ad, np.cross(bd, cd))
            if vol < 0.0:
                p[[2, 5, 7, 9, 11]] = -p[[2, 5, 7, 9, 11]]
        a = rotmat(xyz, i, t, zaxis[i], xaxis[i], yaxis[i])
        out[i, 0] = p[0]
        out[i, 1:4] = a @ p[1:4]
        out[i, 4:13] = (a @ p[4:13].reshape(3, 3) @ a.T).reshape(-1)
    return out


# ------------------------------------------- polarization energy gradient
# Real-space part of the polarization gradient of epolar1c
# (v1.2/source/epolar1.f:1215-2543, epreal1c), written from the identity
# dE_pol = -f/2 (u_p . dE_d + u_d . dE_p + u_p . dT u_d) at fixed dipoles.
# Every pair interaction derives from one radial kernel psi whose radial
# ladder dmp3, dmp5, dmp7, dmp9 (Ewald bn_n less the Thole-damped, scaled
# rr_n: dmp_n = bn_n - (1 - s lambda_n) rr_n) obeys d dmp_n / dr =
# -r dmp_(n+2), so the Cartesian derivative tensors of psi are
#   T1_a = -R_a G1,  T2_ab = -d_ab G1 + R_a R_b G2,
#   T3_abc = (d_ab R_c + d_ac R_b + d_bc R_a) G2 - R_a R_b R_c G3,
#   T4_abcd = (d_ab d_cd + d_ac d_bd + d_ad d_bc) G2 - (six d_.. R_. R_.) G3
#             + R_a R_b R_c R_d G4
# (G1..G4 = dmp3..dmp9, R = field point - source).  A multipole (c, d, Q)
# has the potential c psi - d . grad psi + Q : grad grad psi (Q as stored
# in rpole, the reference's traceless Q/3), a dipole u the potential
# -u . grad psi.

KCOUL = 332.063714  # electric, v1.2/source/MOD_units.f:81


def _thole_lambdas(r, pdi, pdk, thl):
    """lambda3, 5, 7, 9 of dampthole (damping.f:74-215) and their
    continuation to order 9 (1 - (1 + x + 18/35 x^2 + 9/35 x^3) e^-x)"""
    damp = pdi * pdk
    ok = (damp != 0.0) & (thl != 0.0)
    x = np.where(ok, thl * (r / np.where(ok, damp, 1.0)) ** 3, np.inf)
    live = x < 50.0
    e = np.where(live, np.exp(-np.where(live, x, 0.0)), 0.0)
    xs = np.where(live, x, 0.0)
    return (1.0 - e, 1.0 - (1.0 + xs) * e, 1.0 - (1.0 + xs + 0.6 * xs ** 2) * e,
            1.0 - (1.0 + xs + 18.0 / 35.0 * xs ** 2 + 9.0 / 35.0 * xs ** 3) * e)


def _ladder(r, aewald, lam, scale):
    """dmp3, dmp5, dmp7, dmp9 of the pairs for one set of scale factors"""
    from scipy.special import erfc
    r2 = r * r
    ra = aewald * r
    ex = np.exp(-ra * ra)
    bn = [erfc(ra) / r]
    afac = 1.0 / (np.sqrt(np.pi) * aewald)
    for i in range(1, 5):
        afac *= 2.0 * aewald * aewald
        bn.append(((2 * i - 1) * bn[-1] + afac * ex) / r2)
    rr = [1.0 / (r2 * r), 3.0 / (r2 * r2 * r), 15.0 / (r2 ** 3 * r), 105.0 / (r2 ** 4 * r)]
    return [bn[n + 1] - (1.0 - scale * lam[n]) * rr[n] for n in range(4)]


def _psi_tensors(R, G):
    """T1..T4 of the kernel at separations R (P, 3)"""
    G1, G2, G3, G4 = (g[:, None] for g in G)
    I = np.eye(3)
    T1 = -R * G1
    RR = np.einsum("pa,pb->pab", R, R)
    T2 = -I[None] * G1[:, :, None] + RR * G2[:, :, None]
    dR = np.einsum("ab,pc->pabc", I, R)
    T3 = (dR + dR.transpose(0, 1, 3, 2) + dR.transpose(0, 3, 1, 2)) * G2[:, :, None, None] \
        - np.einsum("pab,pc->pabc", RR, R) * G3[:, :, None, None]
    dd = np.einsum("ab,cd->abcd", I, I)
    dd = dd + dd.transpose(0, 2, 1, 3) + dd.transpose(0, 3, 2, 1)
    dRR = np.einsum("ab,pcd->pabcd", I, RR)
    six = (dRR + dRR.transpose(0, 1, 3, 2, 4) + dRR.transpose(0, 1, 3, 4, 2)
           + dRR.transpose(0, 3, 4, 1, 2) + dRR.transpose(0, 3, 1, 4, 2) + dRR.transpose(0, 3, 1, 2, 4))
    T4 = dd[None] * G2[:, :, None, None, None] - six * G3[:, :, None, None, None] \
        + np.einsum("pab,pcd->pabcd", RR, RR) * G4[:, :, None, None, None]
    return T1, T2, T3, T4


_EPS = np.zeros((3, 3, 3))
_EPS[0, 1, 2] = _EPS[1, 2, 0] = _EPS[2, 0, 1] = 1.0
_EPS[0, 2, 1] = _EPS[2, 1, 0] = _EPS[1, 0, 2] = -1.0


def epolar1_real(system, rpole, ud, up, pairs=None):
    """Real-space polarization forces and torques (kcal/mol/A and
    kcal/mol) on every site for global multipoles `rpole` (n, 13) and the
    dipole sets ud, up (n, 3): F = -dG/dr and T with dG/dtheta = -T for a
    rotation of a site's multipole about its centre, where
    G = -f/2 (u_p . E_d + u_d . E_p + u_p . T u_d) over the real-space
    pairs within the cutoff (scaled pairs with their d, p, u factors)."""
    F, T, _ = _epolar1_real_pairs(system, rpole, ud, up, pairs)
    return F, T


def _epolar1_real_pairs(system, rpole, ud, up, pairs=None):
    """epolar1_real, also returning the ordered pairs' separations and
    forces (field point j, source m, R = r_j - r_m, force on j from the
    pair, kcal/mol/A) for the virial"""
    s = system
    n = s.n
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    box = np.asarray(s.box, dtype=np.float64)
    if pairs is None:
        pairs = Oracle(s).nblist_pairs()
    i, k = pairs[:, 0].astype(np.int64), pairs[:, 1].astype(np.int64)
    d = xyz[i] - xyz[k]
    d -= box * np.round(d / box)
    r = np.linalg.norm(d, axis=1)
    keep = r <= s.cut
    i, k, d, r = i[keep], k[keep], d[keep], r[keep]
    # both orders: field point j, source m, R = r_j - r_m
    j = np.concatenate([i, k])
    m = np.concatenate([k, i])
    R = np.concatenate([d, -d])
    r = np.concatenate([r, r])
    sc = np.ones((len(j), 3))
    lut = {}
    for (a, b), v in zip(np.asarray(s.scaled_ik).reshape(-1, 2), np.asarray(s.scaled_dpu).reshape(-1, 3)):
        lut[(int(a), int(b))] = v
        lut[(int(b), int(a))] = v
    for q in range(len(j)):
        v = lut.get((int(j[q]), int(m[q])))
        if v is not None:
            sc[q] = v
    pol = np.asarray(s.polarity, dtype=np.float64)
    th = np.asarray(s.thole, dtype=np.float64)
    pd = pol ** (1.0 / 6.0)
    thl = np.minimum(th[j], th[m])
    thl = np.where(thl == 0.0, np.maximum(th[j], th[m]), thl)
    lam = _thole_lambdas(r, pd[j], pd[m], thl)
    c = rpole[:, 0]
    dp = rpole[:, 1:4]
    Q = rpole[:, 4:13].reshape(n, 3, 3)
    fp = np.zeros((len(j), 3))
    T = np.zeros((n, 3))
    for which, (u_at_j, u_src, s_idx) in enumerate(((up, up, 0), (ud, ud, 1))):
        # u_p with the d-scaled fields (s_idx 0), u_d with the p-scaled (1)
        T1, T2, T3, T4 = _psi_tensors(R, _ladder(r, s.aewald, lam, sc[:, s_idx]))
        # A: dipole at j in the field gradient of the multipole at m
        ddphi = (c[m][:, None, None] * T2 - np.einsum("pc,pabc->pab", dp[m], T3)
                 + np.einsum("pcd,pabcd->pab", Q[m], T4))
        fp -= np.einsum("pa,pab->pb", u_at_j[j], ddphi)
        # B: multipole at j in the field of the dipole at m
        dphi = -np.einsum("pb,pab->pa", u_src[m], T2)
        d2phi = -np.einsum("pc,pabc->pab", u_src[m], T3)
        d3phi = -np.einsum("pd,pabcd->pabc", u_src[m], T4)
        E = -dphi
        fB = c[j][:, None] * E - np.einsum("pa,pab->pb", dp[j], d2phi) - np.einsum("pac,pacb->pb", Q[j], d3phi)
        fp += fB
        tq = np.cross(dp[j], E)
        A = np.einsum("pab,pbc->pac", Q[j], d2phi) - np.einsum("pab,pbc->pac", d2phi, Q[j])
        tq -= np.einsum("cdb,pdb->pc", _EPS, A)
        np.add.at(T, j, tq)
    # C: the mutual term, u-scaled, both dipole sets at j
    T1, T2, T3, T4 = _psi_tensors(R, _ladder(r, s.aewald, lam, sc[:, 2]))
    for u_at_j, u_src in ((up, ud), (ud, up)):
        d2phi = -np.einsum("pc,pabc->pab", u_src[m], T3)
        fp -= np.einsum("pa,pab->pb", u_at_j[j], d2phi)
    F = np.zeros((n, 3))
    np.add.at(F, j, fp)
    return 0.5 * KCOUL * F, 0.5 * KCOUL * T, (j, m, R, 0.5 * KCOUL * fp)


# Reciprocal part (eprecip1, v1.2/source/epolar1.f:397-1203).  The PME
# bilinear forms are differentiated through the B-spline stencils of both
# their field points and their sources, which gives the same assembly as the
# real-space pairs with PME potential derivatives in place of the pair
# tensors (no scale factors in reciprocal space, so E_d = E_p there).  The
# grid conventions are those of the C++ oracle (bspline_fill_site,
# grid_uind_site, make_qfac, unnormalised transforms both ways); the
# orthorhombic box turns fractional derivatives into Cartesian ones by the
# factors a_k = nfft_k / L_k.

def _pme_stencils(s):
    """per axis: B-spline values and 3 derivatives (n, order, 4) and the
    grid nodes (n, order) of every site (bspline_fill_site, node_index)"""
    order = s.order
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    th, nodes = [], []
    for d in range(3):
        L, K = float(s.box[d]), int(s.nfft[d])
        w = xyz[:, d] / L
        anint = np.sign(w) * np.floor(np.abs(w) + 0.5)
        fr = K * (w - anint + 0.5)
        ifr = np.trunc(fr - 1.0e-8).astype(np.int64)
        ww = fr - ifr
        th.append(np.stack([bsplgen(v, order) for v in ww]))
        g1 = (ifr - order)[:, None] + ((order + 1) // 2 + 1) - (order - 1) // 2 + np.arange(order)[None, :]
        g1 = np.where(g1 < 1, g1 + K, g1)
        nodes.append(g1 - 1)
    return th, nodes


def _qfac(s):
    """make_qfac: exp(-pi^2 h^2 / alpha^2) / (pi V h^2 bsmod1 bsmod2 bsmod3)"""
    n1, n2, n3 = (int(v) for v in s.nfft)
    L = [float(v) for v in s.box]
    ms = []
    for K, Lk in zip((n1, n2, n3), L):
        k = np.arange(K)
        m = np.where(k + 1 > (K + 1) // 2, k - K, k)
        ms.append(m / Lk)
    h3, h2, h1 = np.meshgrid(ms[2], ms[1], ms[0], indexing="ij")
    hsq = h1 ** 2 + h2 ** 2 + h3 ** 2
    b1, b2, b3 = (bsmod(K, s.order) for K in (n1, n2, n3))
    bb = b3[:, None, None] * b2[None, :, None] * b1[None, None, :]
    term = -(np.pi / s.aewald) ** 2 * hsq
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(term > -50.0, np.exp(term) / (np.pi * L[0] * L[1] * L[2] * hsq * bb), 0.0)
    q[0, 0, 0] = 0.0
    return q


def _pme_dipole_potential(s, u, th, nodes, q):
    """Cartesian derivatives 1..3 of the PME potential of the dipoles u at
    every site: D1 (n, 3), D2 (n, 3, 3), D3 (n, 3, 3, 3)"""
    n1, n2, n3 = (int(v) for v in s.nfft)
    a = np.array([n1 / s.box[0], n2 / s.box[1], n3 / s.box[2]])
    fu = u * a[None, :]
    grid = np.zeros((n3, n2, n1))
    t, uu, v = th[0], th[1], th[2]
    for i in range(s.n):
        w = (fu[i, 0] * np.einsum("a,b,c->cba", t[i, :, 1], uu[i, :, 0], v[i, :, 0])
             + fu[i, 1] * np.einsum("a,b,c->cba", t[i, :, 0], uu[i, :, 1], v[i, :, 0])
             + fu[i, 2] * np.einsum("a,b,c->cba", t[i, :, 0], uu[i, :, 0], v[i, :, 1]))
        grid[np.ix_(nodes[2][i], nodes[1][i], nodes[0][i])] += w
    pot = np.real(np.fft.ifftn(q * np.fft.fftn(grid))) * grid.size
    D1 = np.zeros((s.n, 3))
    D2 = np.zeros((s.n, 3, 3))
    D3 = np.zeros((s.n, 3, 3, 3))
    for i in range(s.n):
        P = pot[np.ix_(nodes[2][i], nodes[1][i], nodes[0][i])]  # [z, y, x]

        def der(ax):
            o = [ax.count(0), ax.count(1), ax.count(2)]
            return np.einsum("cba,a,b,c->", P, t[i, :, o[0]], uu[i, :, o[1]], v[i, :, o[2]]) * np.prod(a ** o)

        for p in range(3):
            D1[i, p] = der((p,))
            for r in range(3):
                D2[i, p, r] = der((p, r))
                for w3 in range(3):
                    D3[i, p, r, w3] = der((p, r, w3))
    return D1, D2, D3


def epolar1_recip_self(system, rpole, ud, up):
    """Reciprocal and self parts of the polarization forces and torques,
    with the conventions of epolar1_real.  The multipole potential's second
    derivatives come from efld0_recip (cphi); the dipole sets' derivatives
    up to third order from the grid above.  The Ewald self field
    (4 alpha^3 / 3 sqrt(pi)) d of the permanent dipoles (efld0_total) gives
    each site the torque f/2 C d x (u_d + u_p); the dipoles' own self field
    depends on no coordinate."""
    s = system
    n = s.n
    c = rpole[:, 0]
    dp = rpole[:, 1:4]
    Q = rpole[:, 4:13].reshape(n, 3, 3)
    cphi = Oracle(s.with_multipoles(rpole)).efld0_recip()
    PhiM = np.zeros((n, 3, 3))
    for (p, r), col in zip(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)), range(4, 10)):
        PhiM[:, p, r] = PhiM[:, r, p] = cphi[:, col]
    th, nodes = _pme_stencils(s)
    q = _qfac(s)
    Dd = _pme_dipole_potential(s, ud, th, nodes, q)
    Dp = _pme_dipole_potential(s, up, th, nodes, q)
    F = np.zeros((n, 3))
    T = np.zeros((n, 3))
    # A: both dipole sets at their sites in the multipoles' field gradient
    F -= np.einsum("pa,pab->pb", ud + up, PhiM)
    # B: the multipoles in the field of both dipole sets, and their torques
    D1 = Dd[0] + Dp[0]
    D2 = Dd[1] + Dp[1]
    D3 = Dd[2] + Dp[2]
    E = -D1
    F += c[:, None] * E - np.einsum("pa,pab->pb", dp, D2) - np.einsum("pac,pacb->pb", Q, D3)
    T += np.cross(dp, E)
    A = np.einsum("pab,pbc->pac", Q, D2) - np.einsum("pab,pbc->pac", D2, Q)
    T -= np.einsum("cdb,pdb->pc", _EPS, A)
    # C: the mutual term
    F -= np.einsum("pa,pab->pb", up, Dd[1]) + np.einsum("pa,pab->pb", ud, Dp[1])
    # self torque
    cself = 4.0 * s.aewald ** 3 / (3.0 * np.sqrt(np.pi))
    T += cself * np.cross(dp, ud + up)
    return 0.5 * KCOUL * F, 0.5 * KCOUL * T


def torque_forces(xyz, axistype, zaxis, xaxis, yaxis, torque, h=1e-6):
    """Forces on the frame atoms from the torques on the sites' multipoles
    (the product that torque.f forms in closed form per axis type), by brute
    force: a site's frame turns by theta(r), and an atom a of the frame
    takes F_a = sum_sites T . d theta / d r_a, with d theta the axial vector
    of dR R^T from central differences of rotmat."""
    xyz = np.asarray(xyz, dtype=np.float64)
    F = np.zeros_like(xyz)
    for i in range(len(xyz)):
        t = int(axistype[i])
        if t == AXIS_NONE:
            continue
        R0 = rotmat(xyz, i, t, zaxis[i], xaxis[i], yaxis[i])
        atoms = {i, int(zaxis[i])} | {int(v) for v in (xaxis[i], yaxis[i]) if v >= 0}
        for a in atoms:
            for c in range(3):
                xp, xm = xyz.copy(), xyz.copy()
                xp[a, c] += h
                xm[a, c] -= h
                dR = (rotmat(xp, i, t, zaxis[i], xaxis[i], yaxis[i]) - rotmat(xm, i, t, zaxis[i], xaxis[i], yaxis[i])) / (2 * h)
                W = dR @ R0.T
                dth = np.array([W[2, 1], W[0, 2], W[1, 0]])
                F[a, c] += torque[i] @ dth
    return F


def epolar1(system, frames, poleps=1e-10):
    """E_pol and its gradient (-forces, kcal/mol/A) for local multipoles on
    frames (axistype, zaxis, xaxis, yaxis, pole): rotpole, the solve, the
    real, reciprocal and self forces and torques, and the torques carried to
    the frame atoms in closed form (epolar1c, epolar1.f:57-376; torque.f)."""
    s = system
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    rp = rotpole(*xyz.T, *frames)
    s2 = s.with_multipoles(rp)
    res = Oracle(s2).induce(poleps=poleps)
    ud, up = res["mu"][:, 0], res["mu"][:, 1]
    F1, T1 = epolar1_real(s2, rp, ud, up)
    F2, T2 = epolar1_recip_self(s2, rp, ud, up)
    F = F1 + F2 + torque_forces_closed(xyz, frames[0], frames[1], frames[2], frames[3], T1 + T2)
    return res["ep"], -F, res


def torque_forces_closed(xyz, axistype, zaxis, xaxis, yaxis, torque):
    """torque_forces in closed form (the product torque.f forms per axis
    type).  A frame (x, y, z) built from z = unit(v) and the xz-plane vector
    u turns by omega = z x dz + (y . dx) z, with y . dx = (y . du -
    (u . z)(y . dz)) / |u - (u . z) z|, so T . omega = g_z . dz + g_u . du
    with g_z = T x z - (T . z)(u . z) y / |u_perp| and g_u = (T . z) y /
    |u_perp|; dz = (1 - z z^T) dv / |v|, and v and u are sums of bond
    vectors or of unit bond vectors (a unit vector e = unit(r_a - r_i)
    passes a gradient g on as (1 - e e^T) g / |r_a - r_i| to r_a and its
    negative to r_i).  Scaling u changes nothing, so Z-Bisect uses the
    unnormalised sum of its two unit vectors."""
    xyz = np.asarray(xyz, dtype=np.float64)
    F = np.zeros_like(xyz)
    unit = lambda v: v / np.linalg.norm(v)

    def through_unit(a, i, g):
        # gradient g on unit(r_a - r_i) -> forces on a and i
        d = xyz[a] - xyz[i]
        r = np.linalg.norm(d)
        e = d / r
        h = (g - e * np.dot(e, g)) / r
        F[a] += h
        F[i] -= h

    for i in range(len(xyz)):
        t = int(axistype[i])
        if t == AXIS_NONE:
            continue
        iz, ix, iy = int(zaxis[i]), int(xaxis[i]), int(yaxis[i])
        R = rotmat(xyz, i, t, iz, ix, iy)
        x, y, z = R[:, 0], R[:, 1], R[:, 2]
        Tq = np.asarray(torque[i], dtype=np.float64)
        # v (z = unit(v)) as unit-vector or bond-vector terms, and u
        if t == AXIS_Z_ONLY:
            v_bonds, v_units = [iz], []
            u = np.array([1.0, 0.0, 0.0]) if abs(z[0]) < 0.866 else np.array([0.0, 1.0, 0.0])
            u_bonds, u_units = [], []
        elif t == AXIS_Z_THEN_X:
            v_bonds, v_units = [iz], []
            u, u_bonds, u_units = xyz[ix] - xyz[i], [ix], []
        elif t == AXIS_BISECTOR:
            v_bonds, v_units = [], [iz, ix]
            u, u_bonds, u_units = unit(xyz[ix] - xyz[i]), [], [ix]
        elif t == AXIS_Z_BISECT:
            v_bonds, v_units = [iz], []
            u = unit(xyz[ix] - xyz[i]) + unit(xyz[iy] - xyz[i])
            u_bonds, u_units = [], [ix, iy]
        else:  # 3-Fold
            v_bonds, v_units = [], [iz, ix, iy]
            u, u_bonds, u_units = unit(xyz[ix] - xyz[i]), [], [ix]
        v = sum((xyz[a] - xyz[i] for a in v_bonds), np.zeros(3)) + \
            sum((unit(xyz[a] - xyz[i]) for a in v_units), np.zeros(3))
        uperp = np.linalg.norm(u - np.dot(u, z) * z)
        tz = np.dot(Tq, z)
        g_z = np.cross(Tq, z) - tz * np.dot(u, z) / uperp * y
        g_u = tz / uperp * y
        g_v = (g_z - z * np.dot(z, g_z)) / np.linalg.norm(v)
        for a in v_bonds:
            F[a] += g_v
            F[i] -= g_v
        for a in v_units:
            through_unit(a, i, g_v)
        for a in u_bonds:
            F[a] += g_u
            F[i] -= g_u
        for a in u_units:
            through_unit(a, i, g_u)
    return F


# Polarization virial (epolar1c's vir: v1.2/source/epolar1.f:2332-2340 for
# the real-space pairs, :1152-1160 for the reciprocal sum, :2478-2526 for
# the torques carried to the frame atoms; :326-334 accumulates the self
# field's share).  The virial is the derivative of G (the form whose
# coordinate derivative is the gradient, with both dipole sets held fixed)
# along the homogeneous strain r -> (1 + eps) r of every position and of
# the box: D_ab = dG/d eps_ab, reported symmetrised as vir = (D + D^T)/2.
#  * real space: D_ab = sum over pairs of dU/dR_a R_b (R the minimum-image
#    separation, which scales with the box);
#  * reciprocal space, at fixed fractional positions: G_recip =
#    f/2 (E(U, M) + E(U_p, U_d)) with E(X, Y) = sum_m qfac Re(F_X* F_Y) the
#    PME interaction of two site sets (U = U_d + U_p).  qfac carries 1/V
#    and m^2 = |h^-T k|^2, so dqfac/d eps_ab = qfac (-delta_ab + 2 m_a m_b
#    (1/m^2 + pi^2/alpha^2)); and Cartesian moments held fixed turn into
#    fractional ones as if d -> (1 - eps) d and Q -> (1 - eps) Q (1 - eps)^T,
#    which adds -sum_i (d_b dphi/dr_a + 2 (Q grad grad phi)_ba) for each set
#    in the other's potential;
#  * the self terms depend on the moments only;
#  * with local frames the multipoles turn with their frame atoms: the
#    forces the torques put on the frame atoms add sum (r_a - r_site) x
#    gradient, per site.

def _pme_site_grid(s, th, nodes, q, d, Q):
    """B-spline grid [z, y, x] of site multipoles (charges q (n,), Cartesian
    dipoles d (n, 3), quadrupoles Q (n, 3, 3) in the rpole convention), such
    that sum(grid * pot) = sum_i q phi + d . grad phi + Q : grad grad phi for
    any interpolated potential grid pot (orthorhombic box)"""
    n1, n2, n3 = (int(v) for v in s.nfft)
    a = np.array([n1 / s.box[0], n2 / s.box[1], n3 / s.box[2]])
    grid = np.zeros((n3, n2, n1))
    t, u, v = th[0], th[1], th[2]
    for i in range(s.n):
        def w(o):
            return np.einsum("a,b,c->cba", t[i, :, o[0]], u[i, :, o[1]], v[i, :, o[2]]) * np.prod(a ** np.array(o))
        g = q[i] * w((0, 0, 0))
        for c in range(3):
            o = [0, 0, 0]
            o[c] = 1
            g = g + d[i, c] * w(o)
            for e in range(3):
                o2 = list(o)
                o2[e] += 1
                g = g + Q[i, c, e] * w(o2)
        grid[np.ix_(nodes[2][i], nodes[1][i], nodes[0][i])] += g
    return grid


def _virial_kweights(s, qf):
    """dqfac/d eps_ab = qfac (-delta_ab + 2 m_a m_b (1/m^2 + pi^2/alpha^2)),
    shape (3, 3, K3, K2, K1)"""
    n1, n2, n3 = (int(v) for v in s.nfft)
    ms = []
    for K, Lk in zip((n1, n2, n3), s.box):
        k = np.arange(K)
        ms.append(np.where(k + 1 > (K + 1) // 2, k - K, k) / float(Lk))
    h3, h2, h1 = np.meshgrid(ms[2], ms[1], ms[0], indexing="ij")
    hv = (h1, h2, h3)
    hsq = h1 ** 2 + h2 ** 2 + h3 ** 2
    with np.errstate(divide="ignore"):
        fac = np.where(hsq > 0, 2.0 * (1.0 / np.where(hsq > 0, hsq, 1.0) + (np.pi / s.aewald) ** 2), 0.0)
    W = np.zeros((3, 3) + qf.shape)
    for a in range(3):
        for b in range(3):
            W[a, b] = qf * ((-1.0 if a == b else 0.0) + fac * hv[a] * hv[b])
    return W


def epolar1_short(system, frames, shortcut, poleps=1e-10):
    """The short-range polarization energy and gradient (epolar1c with
    use_polarshortreal: newinduce_shortreal's dipoles, and epreal1c's terms
    within ewaldshortcut, mode SHORTEWALD, epolar1.f:1347-1356 and
    switch.f:60-62; no reciprocal or self terms), frames carried to the
    atoms as in epolar1.  Returns (ep, gradient, dipoles)."""
    import dataclasses
    s = dataclasses.replace(system, cut=shortcut)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    rp = rotpole(*xyz.T, *frames)
    s2 = s.with_multipoles(rp)
    res = induce_shortreal(Oracle(s2), poleps=poleps)
    ud, up = res["mu"][:, 0], res["mu"][:, 1]
    F1, T1 = epolar1_real(s2, rp, ud, up)
    F = F1 + torque_forces_closed(xyz, frames[0], frames[1], frames[2], frames[3], T1)
    return res["ep"], -F, res


def epolar1_virial(system, rpole, ud, up, frames=None, parts=False, real_only=False):
    """D_ab = dG/d eps_ab (kcal/mol) for global multipoles rpole (n, 13) and
    the dipole sets ud, up (n, 3); with frames (axistype, zaxis, xaxis,
    yaxis, pole) the multipoles turn with their frame atoms.  The reference's
    vir is (D + D^T) / 2.  parts: also the real, reciprocal and torque
    shares.  real_only: the short-range virial (the system's cut is the
    short cutoff; no reciprocal or self terms)."""
    s = system
    n = s.n
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    F1, T1, (j, m, R, fp) = _epolar1_real_pairs(s, rpole, ud, up)
    D_real = -0.5 * np.einsum("pa,pb->ab", fp, R)
    if real_only:
        D_torque = np.zeros((3, 3))
        if frames is not None:
            for i in range(n):
                if int(frames[0][i]) == AXIS_NONE:
                    continue
                Ti = np.zeros((n, 3))
                Ti[i] = T1[i]
                Fi = torque_forces_closed(xyz, frames[0], frames[1], frames[2], frames[3], Ti)
                D_torque -= np.einsum("ka,kb->ab", Fi, xyz - xyz[i])
        D = D_real + D_torque
        if parts:
            return D, dict(real=D_real, recip=np.zeros((3, 3)), torque=D_torque, ksum=np.zeros((3, 3)))
        return D
    th, nodes = _pme_stencils(s)
    qf = _qfac(s)
    c = rpole[:, 0]
    dp = rpole[:, 1:4]
    Q = rpole[:, 4:13].reshape(n, 3, 3)
    zero, zero33 = np.zeros(n), np.zeros((n, 3, 3))
    FM = np.fft.fftn(_pme_site_grid(s, th, nodes, c, dp, Q))
    FD = np.fft.fftn(_pme_site_grid(s, th, nodes, zero, ud, zero33))
    FP = np.fft.fftn(_pme_site_grid(s, th, nodes, zero, up, zero33))
    cross = np.real(np.conj(FD + FP) * FM) + np.real(np.conj(FP) * FD)
    Dk = np.einsum("abzyx,zyx->ab", _virial_kweights(s, qf), cross)
    cphi = Oracle(s.with_multipoles(rpole)).efld0_recip()
    Dd = _pme_dipole_potential(s, ud, th, nodes, qf)
    Dp = _pme_dipole_potential(s, up, th, nodes, qf)
    Dm = -(np.einsum("ib,ia->ab", ud + up, cphi[:, 1:4]) + np.einsum("ib,ia->ab", dp, Dd[0] + Dp[0])
           + 2.0 * np.einsum("ibe,iae->ab", Q, Dd[1] + Dp[1])
           + np.einsum("ib,ia->ab", up, Dd[0]) + np.einsum("ib,ia->ab", ud, Dp[0]))
    D_recip = 0.5 * KCOUL * (Dk + Dm)
    D_torque = np.zeros((3, 3))
    if frames is not None:
        _, T2 = epolar1_recip_self(s, rpole, ud, up)
        T = T1 + T2
        for i in range(n):
            if int(frames[0][i]) == AXIS_NONE:
                continue
            Ti = np.zeros((n, 3))
            Ti[i] = T[i]
            Fi = torque_forces_closed(xyz, frames[0], frames[1], frames[2], frames[3], Ti)
            D_torque -= np.einsum("ka,kb->ab", Fi, xyz - xyz[i])
    D = D_real + D_recip + D_torque
    if parts:
        return D, dict(real=D_real, recip=D_recip, torque=D_torque, ksum=0.5 * KCOUL * Dk)
    return D


# Gradient of the truncated-CG energy (epolar1tcg: epolar1tcg.f:16-119 and
# epolar1tcg2.f, for tcgorder 1 and 2 with the tcgprec, tcgguess and
# tcgpeek options).  E_TCG = -f/2 <mu_d, E_p> with mu_d an explicit function
# of the permanent fields and of the polarization matrix A = 1/alpha - T
# (truncated CG is not stationary in the dipoles).  Reverse-mode
# differentiation of that function gives
#   dE = <efb, dE_d> + <Epb, dE_p> + sum_k <lam_k, dA v_k>
# with adjoint vectors efb, Epb and one (lam, v) pair per application of A
# in the forward pass; dA = -dT, and each term is one of the derivatives
# the PCG gradient already takes (a dipole set in a permanent field, two
# dipole sets in T).

def _tcg_adjoint(A1, efd, Ep, alpha, order, prec, guess, peek, omega, f):
    """(mu_d, efb, Epb, pairs) of the right-hand side d; A1 applies A to an
    (n, 3) vector, alpha (n, 1)"""
    P = (lambda v: alpha * v) if prec else (lambda v: v)
    Q = (lambda v: v) if prec else (lambda v: alpha * v)   # the peek step's alpha
    dot = lambda a, b: float(np.sum(a * b))
    # forward
    if guess:
        m0 = alpha * efd
        q0 = A1(m0)
        r0 = efd - q0
    else:
        r0 = efd.copy()
    z0 = P(r0)
    w1 = A1(z0)
    y1 = P(w1)
    n0, t1 = dot(r0, z0), dot(z0, w1)
    t4 = n0 / t1
    if order == 2:
        w2 = A1(y1)
        y2 = P(w2)
        np1 = dot(w1, y1)
        t2 = n0 * np1 / t1 ** 2
        p2 = t2 * y1 - t4 * y2
        s3 = dot(w1, p2)
        t3 = t1 * s3
        num = t1 ** 2 - n0 * np1
        g1 = num / t3
        cz, cy = g1 * t2 + t4, -g1 * t4
    else:
        cz, cy = t4, 0.0
    mu = cz * z0 + cy * y1
    if guess:
        mu = mu + m0
    if peek:
        pk = omega * Q(z0) - omega * t4 * Q(y1)
        if order == 2:
            pk = pk - omega * g1 * t2 * Q(y1) + omega * g1 * t4 * Q(y2)
        mu = mu + pk
    # reverse
    mub = -0.5 * f * Ep
    Epb = -0.5 * f * mu
    zeros = np.zeros_like(efd)
    z0b, y1b, y2b, w1b, w2b, r0b = (zeros.copy() for _ in range(6))
    t4b = g1b = t2b = t1b = n0b = np1b = 0.0
    z0b += cz * mub
    y1b += cy * mub
    czb, cyb = dot(mub, z0), dot(mub, y1)
    m0b = mub.copy() if guess else None
    if peek:
        qm = Q(mub)
        z0b += omega * qm
        y1b += -omega * t4 * qm
        t4b += -omega * dot(qm, y1)
        if order == 2:
            y1b += -omega * g1 * t2 * qm
            y2b += omega * g1 * t4 * qm
            g1b += -omega * t2 * dot(qm, y1) + omega * t4 * dot(qm, y2)
            t2b += -omega * g1 * dot(qm, y1)
            t4b += omega * g1 * dot(qm, y2)
    pairs = []
    if order == 2:
        g1b += czb * t2 - cyb * t4
        t2b += czb * g1
        t4b += czb - cyb * g1
        # g1 = num / t3, num = t1^2 - n0 np1, t3 = t1 s3
        t1b += g1b * 2.0 * t1 / t3
        n0b += -g1b * np1 / t3
        np1b += -g1b * n0 / t3
        t3b = -g1b * num / t3 ** 2
        t1b += t3b * s3
        s3b = t3b * t1
        w1b += s3b * p2
        p2b = s3b * w1
        t2b += dot(p2b, y1)
        y1b += t2 * p2b
        t4b += -dot(p2b, y2)
        y2b += -t4 * p2b
        # t2 = n0 np1 / t1^2
        n0b += t2b * np1 / t1 ** 2
        np1b += t2b * n0 / t1 ** 2
        t1b += -2.0 * t2b * n0 * np1 / t1 ** 3
        # np1 = <w1, y1>
        w1b += np1b * y1
        y1b += np1b * w1
        # y2 = P w2, w2 = A y1
        w2b += P(y2b)
        y1b += A1(w2b)
        pairs.append((w2b, y1))
    else:
        t4b += czb
    # t4 = n0 / t1, t1 = <z0, w1>, n0 = <r0, z0>
    n0b += t4b / t1
    t1b += -t4b * n0 / t1 ** 2
    z0b += t1b * w1
    w1b += t1b * z0
    r0b += n0b * z0
    z0b += n0b * r0
    # y1 = P w1, w1 = A z0, z0 = P r0
    w1b += P(y1b)
    z0b += A1(w1b)
    pairs.append((w1b, z0))
    r0b += P(z0b)
    if guess:
        efb = r0b.copy()
        q0b = -r0b
        m0b = m0b + A1(q0b)
        pairs.append((q0b, m0))
        efb = efb + alpha * m0b
    else:
        efb = r0b
    return mu, efb, Epb, pairs


def epolar1_tcg(system, frames, order=2, prec=True, guess=False, peek=True, omega=1.0, shortcut=None):
    """E_TCG and its gradient (kcal/mol/A) for local multipoles on frames:
    the adjoint of the truncated CG (_tcg_adjoint), then the gradient
    pieces of epolar1_real and epolar1_recip_self combined with it, the
    torques carried to the frame atoms.  shortcut: the short-range form
    (epolar1tcg1shortreal, epolar1tcg2*shortreal): real-space terms within
    shortcut only.  Returns (ep, gradient)."""
    import dataclasses
    s = system if shortcut is None else dataclasses.replace(system, cut=shortcut)
    xyz = np.stack([s.x, s.y, s.z], axis=1)
    rp = rotpole(*xyz.T, *frames)
    s2 = s.with_multipoles(rp)
    orc = Oracle(s2)
    alpha = np.asarray(s.polarity, dtype=np.float64)[:, None]
    if shortcut is None:
        ef = orc.efld0_total()
        A1 = lambda v: orc.matvec(np.stack([v, v], axis=1))[:, 0]
    else:
        ef = orc.efld0_direct()
        A1 = lambda v: orc.tmatxb_pme(np.stack([v, v], axis=1), dodiag=True)[:, 0]
    f = KCOUL
    mu, efb, Epb, pairs = _tcg_adjoint(A1, ef[:, 0], ef[:, 1], alpha, order, prec, guess, peek, omega, f)
    ep = -0.5 * f * float(np.sum(mu * ef[:, 1]))

    def FG(ud, up):
        F1, T1 = epolar1_real(s2, rp, ud, up)
        if shortcut is not None:
            return F1, T1
        F2, T2 = epolar1_recip_self(s2, rp, ud, up)
        return F1 + F2, T1 + T2

    z = np.zeros_like(efb)
    F, T = FG(z, efb)
    F2, T2 = FG(Epb, z)
    F, T = F + F2, T + T2
    for lam, v in pairs:
        Fa, Ta = FG(v, lam)
        Fb, Tb = FG(z, lam)
        Fc, Tc = FG(v, z)
        F = F - (Fa - Fb - Fc)
        T = T - (Ta - Tb - Tc)
    # dE/dr = (2/f) F-combination; forces and torques of E are -(2/f) times
    FE, TE = -(2.0 / f) * F, -(2.0 / f) * T
    Ftot = FE + torque_forces_closed(xyz, frames[0], frames[1], frames[2], frames[3], TE)
    return ep, -Ftot
