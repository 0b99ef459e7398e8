"""Textbook references used to pin the oracle (test code only).

* the order-p cardinal B-spline and its derivatives in closed form
  (Schoenberg; Essmann et al., J. Chem. Phys. 103, 8577 (1995), eq. 4.1);
* the Ewald reciprocal-space potential gradient of point multipoles by a
  direct structure-factor sum over reciprocal lattice vectors (no grid,
  no B-splines), for an orthorhombic cell;
* the optimal-influence zeta factor of the smooth PME moduli in closed
  form through the Hurwitz zeta function.
"""
from math import comb, factorial, pi

import numpy as np
from scipy.special import zeta as hurwitz


def cardinal_bspline(x, p, deriv=0):
    """M_p^{(deriv)}(x) = 1/(p-1-deriv)! sum_k (-1)^k C(p,k) (x-k)_+^(p-1-deriv)."""
    e = p - 1 - deriv
    s = 0.0
    for k in range(p + 1):
        t = x - k
        if t > 0:
            s += (-1) ** k * comb(p, k) * t ** e
    return s / factorial(e)


def bspline_nodes(w, p=5):
    """theta[m][level] for nodes m = 0..p-1 at fractional offset w in [0,1):
    node m sits at distance w + p-1-m from the spline origin."""
    th = np.zeros((p, 4))
    for m in range(p):
        for lev in range(4):
            th[m, lev] = cardinal_bspline(w + p - 1 - m, p, lev)
    return th


def zeta_closed(k, K, p):
    """zeta of dftmod for wave index k: S(2p)/S(p), S(n) = sum_j (f/(f+pi j))^n,
    f = pi k/K, summed in closed form with Hurwitz zeta functions."""
    if k == 0:
        return 1.0
    # S(n) is even in f (substitute j -> -j), so use x = |f|/pi in (0, 1/2]
    x = abs(k) / K

    def S(n):
        # sum_{j in Z} 1/(x+j)^n = zeta(n, x) + (-1)^n zeta(n, 1-x)
        return x ** n * (hurwitz(n, x) + (-1) ** n * hurwitz(n, 1.0 - x))

    return S(2 * p) / S(p)


def recip_grad_multipoles(pos, q, d, Q, box, alpha, mmax):
    """grad phi_recip at every site from charges q, dipoles d and quadrupoles Q
    (3x3, the convention phi = q/R + d.R/R^3 + 3 R.Q.R/R^5), by the Ewald
    structure-factor sum over |m_i| <= mmax (orthorhombic box)."""
    V = box[0] * box[1] * box[2]
    rng = np.arange(-mmax, mmax + 1)
    M = np.array(np.meshgrid(rng, rng, rng, indexing="ij")).reshape(3, -1).T
    M = M[np.any(M != 0, axis=1)]
    H = M / box[None, :]                      # reciprocal vectors h
    h2 = (H ** 2).sum(1)
    pref = np.exp(-pi ** 2 * h2 / alpha ** 2) / h2 / (pi * V)
    phase = np.exp(-2j * pi * pos @ H.T)      # (n, nh)
    hd = d @ H.T
    hQh = np.einsum("hi,nij,hj->nh", H, Q, H)
    S = ((q[:, None] - 2j * pi * hd - 4 * pi ** 2 * hQh) * phase).sum(0)  # (nh,)
    grad = np.zeros((len(pos), 3))
    for c in range(3):
        grad[:, c] = np.real((pref * S * 2j * pi * H[:, c])[None, :] * np.conj(phase)).sum(1)
    return grad


def rpole_to_parts(rpole):
    q = rpole[:, 0].copy()
    d = rpole[:, 1:4].copy()
    Q = rpole[:, 4:13].reshape(-1, 3, 3).copy()
    return q, d, Q
