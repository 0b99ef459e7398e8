// Polarization energy gradient (epolar1c, v1.2/source/epolar1.f:57-376;
// epreal1c :1215-2543; eprecip1 :397-1203; torque.f) after a solve.
//
// With A = 1/alpha - T symmetric, u_d = A^-1 E_d and u_p = A^-1 E_p, the
// derivative of E_pol = -f/2 u_d . E_p is the derivative of
//   G = -f/2 (u_p . E_d + u_d . E_p + u_p . T u_d)
// with both dipole sets held fixed.  Every term is a bilinear form between
// a dipole set and a source set, so the force on atom i is gathered per
// atom, without atomics:
//   A  the dipoles at i in the field gradient of the permanent multipoles,
//   B  the multipole at i in the field, field gradient and field curvature
//      of the dipole sets (and its torque),
//   C  each dipole set at i in the field gradient of the other (u-scale).
// Real space: every pair term derives from one damped kernel psi whose
// radial ladder G1..G4 = dmp3, dmp5, dmp7, dmp9 (Ewald bn_n less the scaled,
// Thole-damped rr_n) obeys dG_n/dr = -r G_(n+1); the Cartesian derivative
// tensors of psi are T1 = -R G1, T2 = -I G1 + R R G2, T3, T4 (R = field
// point - source), contracted in closed form below.  Reciprocal space: the
// same contractions with the PME potential derivatives of the multipole grid
// (second order, from the permanent-field gather) and of the two induced
// grids (to third order, gathered here with a fourth B-spline level).  The
// Ewald self field of the permanent dipoles leaves the torque
// (4 alpha^3 / 3 sqrt(pi)) d x (u_d + u_p).  Torques on sites with local
// frames are carried to the frame atoms by omega = z x dz + (y . dx) z.
// Forces and torques accumulate unscaled in F, T ([3][npad], sorted order)
// and k_ep1_finish applies f/2.

#pragma once

namespace thp {

constexpr double kCoulombFp = 332.063714;  // electric, MOD_units.f:81

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 v3(double a, double b, double c) { return V3{a, b, c}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// symmetric 3x3 (stored traceless Q/3 as in rpole: xx, xy, xz, yy, yz, zz)
struct S3 {
  double xx, xy, xz, yy, yz, zz;
};
__device__ __forceinline__ V3 mul(const S3& q, V3 v) {
  return v3(q.xx * v.x + q.xy * v.y + q.xz * v.z, q.xy * v.x + q.yy * v.y + q.yz * v.z,
            q.xz * v.x + q.yz * v.y + q.zz * v.z);
}

struct Site {
  double c;
  V3 d;
  S3 q;
  V3 ud, up;
};
__device__ __forceinline__ Site load_site(const SysView& sv, int i) {
  const size_t np = sv.npad;
  const double* rp = sv.rpole;
  Site s;
  s.c = rp[i];
  s.d = v3(rp[np + i], rp[2 * np + i], rp[3 * np + i]);
  s.q = S3{rp[4 * np + i], rp[5 * np + i], rp[6 * np + i], rp[7 * np + i], rp[8 * np + i], rp[9 * np + i]};
  s.ud = v3(sv.mu[i], sv.mu[np + i], sv.mu[2 * np + i]);
  s.up = v3(sv.mu[3 * np + i], sv.mu[4 * np + i], sv.mu[5 * np + i]);
  return s;
}

// w . (grad grad phi) of a multipole source (c, d, Q) at R, as a vector
__device__ __forceinline__ V3 w_ddphi_mpole(V3 w, double c, V3 d, const S3& q, V3 R, const double* G) {
  const double wR = dot(w, R), dR = dot(d, R);
  const V3 QR = mul(q, R);
  const double RQR = dot(R, QR);
  V3 o = c * ((-G[0]) * w + (wR * G[1]) * R);
  o = o - G[1] * (dR * w + dot(w, d) * R + wR * d) + (wR * dR * G[2]) * R;
  o = o + (2.0 * G[1]) * mul(q, w) - G[2] * (RQR * w + (2.0 * dot(w, QR)) * R + (2.0 * wR) * QR) +
      (wR * RQR * G[3]) * R;
  return o;
}

// w . (grad grad phi) of a dipole source u at R
__device__ __forceinline__ V3 w_ddphi_dip(V3 w, V3 u, V3 R, const double* G) {
  const double wR = dot(w, R), uR = dot(u, R);
  return (wR * uR * G[2]) * R - G[1] * (uR * w + dot(w, u) * R + wR * u);
}

// force and torque on the multipole of site i in the field of a dipole u
// at R (E, its gradient and curvature from the ladder)
__device__ __forceinline__ void mpole_in_dipole(const Site& si, V3 u, V3 R, const double* G, V3& F, V3& T) {
  const double uR = dot(u, R);
  const V3 E = (uR * G[1]) * R - G[0] * u;
  const V3 QR = mul(si.q, R), Qu = mul(si.q, u);
  const double RQR = dot(R, QR);
  // d . grad grad phi and Q : grad grad grad phi of the dipole's potential
  const V3 dd = w_ddphi_dip(si.d, u, R, G);
  const V3 qd = (2.0 * G[1]) * Qu - G[2] * ((2.0 * uR) * QR + (2.0 * dot(u, QR)) * R + RQR * u) + (RQR * uR * G[3]) * R;
  F = F + si.c * E - dd + qd;
  T = T + cross(si.d, E) + (2.0 * G[1]) * (cross(Qu, R) + cross(QR, u)) - (2.0 * G[2] * uR) * cross(QR, R);
}

// G1..G4 of one pair for the scale factors s[0..nl): Ewald bn1..bn4 less
// (1 - s lambda_n) rr_n; lambda9 continues dampthole's ladder
// (1 - (1 + x + 18/35 x^2 + 9/35 x^3) e^-x)
__device__ __forceinline__ void ep1_ladders(double r2, double aew, double pdi, double pdk, double thi, double thk,
                                            const double* s, int nl, double G[3][4]) {
  const double r = sqrt(r2);
  const double rr1 = 1.0 / r;
  const double ir2 = rr1 * rr1;
  const double ra = aew * r;
  const double ex = exp(-ra * ra);
  double bn[5];
  bn[0] = erfc(ra) * rr1;
  double afac = 1.0 / (1.7724538509055160273 * aew);
#pragma unroll
  for (int q = 1; q <= 4; ++q) {
    afac *= 2.0 * aew * aew;
    bn[q] = (double(2 * q - 1) * bn[q - 1] + afac * ex) * ir2;
  }
  const double rr3 = ir2 * rr1;
  const double rr[4] = {rr3, 3.0 * rr3 * ir2, 15.0 * rr3 * ir2 * ir2, 105.0 * rr3 * ir2 * ir2 * ir2};
  double lam[4] = {1.0, 1.0, 1.0, 1.0};
  const double damp = pdi * pdk;
  double pg = fmin(thi, thk);
  if (pg == 0.0) pg = fmax(thi, thk);
  if (damp != 0.0 && pg != 0.0) {
    const double rd = r / damp;
    const double x = pg * rd * rd * rd;
    if (x < 50.0) {
      const double e = exp(-x);
      lam[0] = 1.0 - e;
      lam[1] = 1.0 - (1.0 + x) * e;
      lam[2] = 1.0 - (1.0 + x + 0.6 * x * x) * e;
      lam[3] = 1.0 - (1.0 + x + (18.0 / 35.0) * x * x + (9.0 / 35.0) * x * x * x) * e;
    }
  }
  for (int l = 0; l < nl; ++l)
#pragma unroll
    for (int q = 0; q < 4; ++q) G[l][q] = bn[q + 1] - (1.0 - s[l] * lam[q]) * rr[q];
}

// terms A, B, C of one pair; Gd, Gp, Gu the ladders of the d, p and u
// scales; TERMS: bit 0 the permanent-field terms A and B, bit 1 the dipole
// term C (the truncated CG's gradient takes them with different dipole
// sets; a compile-time mask, so the full gradient's code is unchanged)
constexpr int kTermsField = 1, kTermsDipole = 2, kTermsAll = 3;
template <int TERMS>
__device__ __forceinline__ void ep1_pair(const Site& si, const Site& sk, V3 R, const double* Gd, const double* Gp,
                                         const double* Gu, V3& F, V3& T) {
  if (TERMS & kTermsField) {
    F = F - w_ddphi_mpole(si.up, sk.c, sk.d, sk.q, R, Gd) - w_ddphi_mpole(si.ud, sk.c, sk.d, sk.q, R, Gp);
    mpole_in_dipole(si, sk.up, R, Gd, F, T);
    mpole_in_dipole(si, sk.ud, R, Gp, F, T);
  }
  if (TERMS & kTermsDipole) F = F - w_ddphi_dip(si.up, sk.ud, R, Gu) - w_ddphi_dip(si.ud, sk.up, R, Gu);
}

// Virial shares (the strain derivative of G, symmetrised, unscaled like F:
// xx, yy, zz, xy, xz, yz).  A pair adds -1/2 f (x) R per direction it is
// walked in (R = r_i - r_k, f its force on i); each kernel sums its shares
// over a workgroup in a fixed order into vpart[slot][6].
__device__ __forceinline__ void vir_pair(V3 f, double rx, double ry, double rz, double* vw) {
  vw[0] -= 0.5 * f.x * rx;
  vw[1] -= 0.5 * f.y * ry;
  vw[2] -= 0.5 * f.z * rz;
  vw[3] -= 0.25 * (f.x * ry + f.y * rx);
  vw[4] -= 0.25 * (f.x * rz + f.z * rx);
  vw[5] -= 0.25 * (f.y * rz + f.z * ry);
}

// Real-space part: a workgroup of S wavefronts per 64-atom block, each
// wavefront walking one contiguous segment of every lane's block-union
// partners (within the cutoff) and every S-th of its scaled pairs (with
// their d, p, u factors); the S partial forces and torques are summed in
// LDS in a fixed order.  FP64 throughout (one launch per gradient).
template <int S, bool VIR, int TERMS>
__global__ void __launch_bounds__(64 * S) k_ep1_real(const SysView* views, const Work* work, double* F, double* T,
                                                     double* vpart) {
  __shared__ double red[S][6][64];
  __shared__ double vred[VIR ? S : 1][6][64];
  double vw[6] = {0, 0, 0, 0, 0, 0};
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = w.blk * kBlock + lane;
  const bool ok = i < sv.n;
  const size_t np = sv.npad;
  V3 f = v3(0, 0, 0), t = v3(0, 0, 0);
  if (ok) {
  const Site si = load_site(sv, i);
  const double xi = sv.x[i], yi = sv.y[i], zi = sv.z[i];
  const double pdi = sv.pdamp[i], thi = sv.thole[i];
  const int cnt = sv.lcount[i];
  const int* UL = sv.ulist + (size_t)w.blk * sv.umax;
  const double one[1] = {1.0};
  for (int s = (cnt * wv) / S; s < (cnt * (wv + 1)) / S; ++s) {
    const int k = UL[sv.llist[(size_t)s * np + i]];
    const double xr = image_exact(xi - sv.x[k], sv.box[0]);
    const double yr = image_exact(yi - sv.y[k], sv.box[1]);
    const double zr = image_exact(zi - sv.z[k], sv.box[2]);
    const double r2 = r2_exact(xr, yr, zr);
    if (!(r2 <= sv.cut2)) continue;
    double G[3][4];
    ep1_ladders(r2, sv.aewald, pdi, sv.pdamp[k], thi, sv.thole[k], one, 1, G);
    const V3 f0 = f;
    ep1_pair<TERMS>(si, load_site(sv, k), v3(xr, yr, zr), G[0], G[0], G[0], f, t);
    if (VIR) vir_pair(f - f0, xr, yr, zr, vw);
  }
  for (int e = sv.excl_off[i] + wv; e < sv.excl_off[i + 1]; e += S) {
    const int k = sv.excl_k[e];
    const double xr = image_exact(xi - sv.x[k], sv.box[0]);
    const double yr = image_exact(yi - sv.y[k], sv.box[1]);
    const double zr = image_exact(zi - sv.z[k], sv.box[2]);
    const double r2 = r2_exact(xr, yr, zr);
    if (!(r2 <= sv.cut2)) continue;
    const double* sc = sv.scale_tab + sv.excl_code[e] * 3;
    double G[3][4];
    ep1_ladders(r2, sv.aewald, pdi, sv.pdamp[k], thi, sv.thole[k], sc, 3, G);
    const V3 f0 = f;
    ep1_pair<TERMS>(si, load_site(sv, k), v3(xr, yr, zr), G[0], G[1], G[2], f, t);
    if (VIR) vir_pair(f - f0, xr, yr, zr, vw);
  }
  }
  red[wv][0][lane] = f.x;
  red[wv][1][lane] = f.y;
  red[wv][2][lane] = f.z;
  red[wv][3][lane] = t.x;
  red[wv][4][lane] = t.y;
  red[wv][5][lane] = t.z;
  if (VIR)
#pragma unroll
    for (int c = 0; c < 6; ++c) vred[wv][c][lane] = vw[c];
  __syncthreads();
  if (VIR && wv == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = vred[0][c][lane];
      for (int q = 1; q < S; ++q) v += vred[q][c][lane];
      v = wave_sum(v);
      if (lane == 0) vpart[(size_t)w.blk * 6 + c] = v;
    }
  }
  if (wv == 0 && ok) {
    double v[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      v[c] = red[0][c][lane];
      for (int q = 1; q < S; ++q) v[c] += red[q][c][lane];
    }
    F[i] = v[0];
    F[np + i] = v[1];
    F[2 * np + i] = v[2];
    T[i] = v[3];
    T[np + i] = v[4];
    T[2 * np + i] = v[5];
  }
}

// Reciprocal and self parts, one thread per atom, added to F and T.  The
// grid holds the convolved induced grids (u_d real, u_p imaginary); cphi
// the Cartesian potential derivatives of the multipole grid ([10][npad]).
template <typename real, bool VIR, int TERMS>
__global__ void __launch_bounds__(64) k_ep1_recip(const SysView* views, const Work* work, const double* cphi,
                                                   double* F, double* T, double* vpart) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  const int i0 = w.blk * kBlock + threadIdx.x;
  const bool ok = i0 < sv.n;
  if (!VIR && !ok) return;
  // lanes past the last atom evaluate atom 0 and add nothing
  const int i = ok ? i0 : 0;
  const size_t np = sv.npad;
  const int n1 = sv.nfft[0], n2 = sv.nfft[1], n3 = sv.nfft[2];
  real th[3][4][kOrder];
  const double xyz[3] = {sv.x[i], sv.y[i], sv.z[i]};
  int g[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) g[d] = bspline_axis<real, 4>(xyz[d], 1.0 / sv.box[d], sv.nfft[d], th[d]);
  // fractional derivatives of both grids: index o = (ox, oy, oz), ox+oy+oz <= 3
  double pd[4][4][4], pp[4][4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int c = 0; c < 4; ++c) pd[a][b][c] = pp[a][b][c] = 0.0;
  const real* grid = static_cast<const real*>(sv.grid);
  for (int mk = 0; mk < kOrder; ++mk) {
    int k = g[2] + mk;
    if (k >= n3) k -= n3;
    double yd[4][4], yp[4][4];  // [ox][oy] partial sums over the x and y nodes
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) yd[a][b] = yp[a][b] = 0.0;
    for (int mj = 0; mj < kOrder; ++mj) {
      int j = g[1] + mj;
      if (j >= n2) j -= n2;
      const real* row = grid + ((size_t)k * n2 + j) * n1 * 2;
      double xd[4] = {0, 0, 0, 0}, xp[4] = {0, 0, 0, 0};
#pragma unroll
      for (int mi = 0; mi < kOrder; ++mi) {
        int ii = g[0] + mi;
        if (ii >= n1) ii -= n1;
        const double vd = double(row[2 * ii]), vp = double(row[2 * ii + 1]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          xd[a] += vd * double(th[0][a][mi]);
          xp[a] += vp * double(th[0][a][mi]);
        }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b + a < 4; ++b) {
          yd[a][b] += xd[a] * double(th[1][b][mj]);
          yp[a][b] += xp[a] * double(th[1][b][mj]);
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b + a < 4; ++b)
#pragma unroll
        for (int c = 0; c + b + a < 4; ++c) {
          pd[a][b][c] += yd[a][b] * double(th[2][c][mk]);
          pp[a][b][c] += yp[a][b] * double(th[2][c][mk]);
        }
  }
  // Cartesian derivatives (orthorhombic: a_k^o per axis) of the sum of
  // both grids (term B) and of each grid's second derivatives (term C)
  const double af[3] = {sv.afrac[0], sv.afrac[1], sv.afrac[2]};
  auto cart = [&](double (*p)[4][4], int o0, int o1, int o2) {
    double s = p[o0][o1][o2];
    for (int q = 0; q < o0; ++q) s *= af[0];
    for (int q = 0; q < o1; ++q) s *= af[1];
    for (int q = 0; q < o2; ++q) s *= af[2];
    return s;
  };
  auto ord = [](int a, int b, int c, int& o0, int& o1, int& o2) {
    o0 = (a == 0) + (b == 0) + (c == 0);
    o1 = (a == 1) + (b == 1) + (c == 1);
    o2 = (a == 2) + (b == 2) + (c == 2);
  };
  double D1[3], D1d[3], D1p[3], D2s[3][3], D2d[3][3], D2p[3][3], D3[3][3][3];
  for (int a = 0; a < 3; ++a) {
    int o0 = (a == 0), o1 = (a == 1), o2 = (a == 2);
    D1d[a] = cart(pd, o0, o1, o2);
    D1p[a] = cart(pp, o0, o1, o2);
    D1[a] = D1d[a] + D1p[a];
    for (int b = 0; b < 3; ++b) {
      o0 = (a == 0) + (b == 0);
      o1 = (a == 1) + (b == 1);
      o2 = (a == 2) + (b == 2);
      D2d[a][b] = cart(pd, o0, o1, o2);
      D2p[a][b] = cart(pp, o0, o1, o2);
      D2s[a][b] = D2d[a][b] + D2p[a][b];
      for (int c = 0; c < 3; ++c) {
        ord(a, b, c, o0, o1, o2);
        D3[a][b][c] = cart(pd, o0, o1, o2) + cart(pp, o0, o1, o2);
      }
    }
  }
  const Site si = load_site(sv, i);
  const double qm[3][3] = {{si.q.xx, si.q.xy, si.q.xz}, {si.q.xy, si.q.yy, si.q.yz}, {si.q.xz, si.q.yz, si.q.zz}};
  const double dm[3] = {si.d.x, si.d.y, si.d.z};
  const double ud[3] = {si.ud.x, si.ud.y, si.ud.z}, up[3] = {si.up.x, si.up.y, si.up.z};
  // the multipole grid's second derivatives: xx, yy, zz, xy, xz, yz
  const double c4 = cphi[4 * np + i], c5 = cphi[5 * np + i], c6 = cphi[6 * np + i], c7 = cphi[7 * np + i],
               c8 = cphi[8 * np + i], c9 = cphi[9 * np + i];
  const double PM[3][3] = {{c4, c7, c8}, {c7, c5, c9}, {c8, c9, c6}};
  double f[3], t[3];
  constexpr bool tf = (TERMS & kTermsField) != 0, tu = (TERMS & kTermsDipole) != 0;
  for (int b = 0; b < 3; ++b) {
    const double E = -D1[b];
    double v = tf ? si.c * E : 0.0;
    for (int a = 0; a < 3; ++a) {
      if (tf) {
        v -= (ud[a] + up[a]) * PM[a][b];        // A
        v -= dm[a] * D2s[a][b];                  // B, dipole
        for (int c = 0; c < 3; ++c) v -= qm[a][c] * D3[a][c][b];  // B, quadrupole
      }
      if (tu) v -= up[a] * D2d[a][b] + ud[a] * D2p[a][b];  // C
    }
    f[b] = v;
  }
  // torques: d x E, the quadrupole's -eps_cdb (Q D2 - D2 Q)_db, the self term
  const V3 E = v3(-D1[0], -D1[1], -D1[2]);
  const V3 dE = cross(si.d, E);
  double A[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += qm[a][c] * D2s[c][b] - D2s[a][c] * qm[c][b];
      A[a][b] = s;
    }
  const V3 self = sv.term_self * cross(si.d, si.ud + si.up);
  t[0] = tf ? dE.x - (A[1][2] - A[2][1]) + self.x : 0.0;
  t[1] = tf ? dE.y - (A[2][0] - A[0][2]) + self.y : 0.0;
  t[2] = tf ? dE.z - (A[0][1] - A[1][0]) + self.z : 0.0;
  if (ok)
    for (int c = 0; c < 3; ++c) {
      F[(size_t)c * np + i] += f[c];
      T[(size_t)c * np + i] += t[c];
    }
  if (VIR) {
    // the moments held fixed in Cartesian form turn into fractional ones as
    // the box strains: -(u_b dphi_M/dr_a + d_b dphi_U/dr_a + 2 (Q ddphi_U)_ba
    // + u_p,b dphi_Ud/dr_a + u_d,b dphi_Up/dr_a), U = U_d + U_p
    const double M1[3] = {cphi[1 * np + i], cphi[2 * np + i], cphi[3 * np + i]};
    double v[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double q2 = 0.0;
        for (int e = 0; e < 3; ++e) q2 += qm[b][e] * D2s[a][e];
        v[a][b] = -((ud[b] + up[b]) * M1[a] + dm[b] * D1[a] + 2.0 * q2 + up[b] * D1d[a] + ud[b] * D1p[a]);
      }
    const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double x = ok ? 0.5 * (v[pr[c][0]][pr[c][1]] + v[pr[c][1]][pr[c][0]]) : 0.0;
      x = wave_sum(x);
      if (threadIdx.x == 0) vpart[(size_t)w.blk * 6 + c] = x;
    }
  }
}

// Torques of the sites with local frames carried to their frame atoms
// (torque.f): z = unit(v), u the xz-plane vector; T . omega = g_z . dz +
// g_u . du with g_z = T x z - (T . z)(u . z) y / |u_perp|, g_u =
// (T . z) y / |u_perp|, dz = (1 - z z^T) dv / |v|; v and u are sums of
// bond vectors or unit bond vectors (a unit vector e = unit(r_a - r_i)
// passes g on as (1 - e e^T) g / |r_a - r_i|).  The site itself takes minus
// the sum of what it passes on.  frame_forces lists site s's contributions
// as (atom, force) in a fixed order, at most 5.
constexpr int kFrameContrib = 5;

__device__ __forceinline__ int frame_forces(const SysView& sv, const int* frames, const double* T, int s,
                                            int* atom, V3* g_out) {
  const size_t np = sv.npad;
  const int4 fr = reinterpret_cast<const int4*>(frames)[s];
  const int axt = fr.x, iz = fr.y, ix = fr.z, iy = fr.w;
  if (axt == kAxisNone) return 0;
  const V3 ri = v3(sv.x[s], sv.y[s], sv.z[s]);
  auto pos = [&](int k) { return v3(sv.x[k], sv.y[k], sv.z[k]); };
  auto unit = [](V3 a) { return (1.0 / sqrt(dot(a, a))) * a; };
  // frame (as k_rotpole builds it) and the v, u of its axis type
  V3 z, u, v;
  const V3 dz1 = pos(iz) - ri;
  if (axt == kAxisZOnly) {
    v = dz1;
    z = unit(v);
    u = fabs(z.x) < 0.866 ? v3(1, 0, 0) : v3(0, 1, 0);
  } else if (axt == kAxisZThenX) {
    v = dz1;
    z = unit(v);
    u = pos(ix) - ri;
  } else if (axt == kAxisBisector) {
    v = unit(dz1) + unit(pos(ix) - ri);
    z = unit(v);
    u = unit(pos(ix) - ri);
  } else if (axt == kAxisZBisect) {
    v = dz1;
    z = unit(v);
    u = unit(pos(ix) - ri) + unit(pos(iy) - ri);
  } else {  // 3-Fold
    v = unit(dz1) + unit(pos(ix) - ri) + unit(pos(iy) - ri);
    z = unit(v);
    u = unit(pos(ix) - ri);
  }
  const V3 up = u - dot(u, z) * z;
  const double nup = sqrt(dot(up, up));
  const V3 x = (1.0 / nup) * up;
  const V3 y = cross(z, x);
  const V3 Tq = v3(T[s], T[np + s], T[2 * np + s]);
  const double tz = dot(Tq, z);
  const V3 gz = cross(Tq, z) - (tz * dot(u, z) / nup) * y;
  const V3 gu = (tz / nup) * y;
  const V3 gv = (1.0 / sqrt(dot(v, v))) * (gz - dot(z, gz) * z);
  int m = 0;
  V3 fi = v3(0, 0, 0);
  auto via_unit = [&](int a, V3 g) {
    const V3 d = pos(a) - ri;
    const double r = sqrt(dot(d, d));
    const V3 e = (1.0 / r) * d;
    const V3 h = (1.0 / r) * (g - dot(e, g) * e);
    atom[m] = a;
    g_out[m++] = h;
    fi = fi - h;
  };
  auto via_bond = [&](int a, V3 g) {
    atom[m] = a;
    g_out[m++] = g;
    fi = fi - g;
  };
  if (axt == kAxisZOnly) {
    via_bond(iz, gv);
  } else if (axt == kAxisZThenX) {
    via_bond(iz, gv);
    via_bond(ix, gu);
  } else if (axt == kAxisBisector) {
    via_unit(iz, gv);
    via_unit(ix, gv);
    via_unit(ix, gu);
  } else if (axt == kAxisZBisect) {
    via_bond(iz, gv);
    via_unit(ix, gu);
    via_unit(iy, gu);
  } else {
    via_unit(iz, gv);
    via_unit(ix, gv);
    via_unit(iy, gv);
    via_unit(ix, gu);
  }
  atom[m] = s;
  g_out[m++] = fi;
  return m;
}

// One thread per atom a gathers the frame forces of every site whose frame
// holds a (tinv[tinv_off[a] .. tinv_off[a + 1]), sites in ascending order,
// a itself included when it has a frame) and adds them to F[a]: a fixed
// order of sums, no atomics, the same bits on every run.
// With VIR each atom also adds -g (x) (r_a - r_site) for what it receives
// (the multipoles turn with their frames as the box strains); the site's
// own share sits at zero separation.
template <bool VIR>
__global__ void __launch_bounds__(256) k_ep1_torque(const SysView* views, const int* frames, const int* tinv_off,
                                                    const int* tinv, const double* T, double* F, double* vpart) {
  const SysView& sv = views[0];
  const int a = blockIdx.x * 256 + threadIdx.x;
  if (!VIR && a >= sv.n) return;
  const size_t np = sv.npad;
  V3 acc = v3(0, 0, 0);
  double vw[6] = {0, 0, 0, 0, 0, 0};
  if (a < sv.n) {
    for (int e = tinv_off[a]; e < tinv_off[a + 1]; ++e) {
      int atom[kFrameContrib];
      V3 g[kFrameContrib];
      const int s = tinv[e];
      const int m = frame_forces(sv, frames, T, s, atom, g);
      for (int q = 0; q < m; ++q)
        if (atom[q] == a) {
          acc = acc + g[q];
          if (VIR) {
            const double rx = sv.x[a] - sv.x[s], ry = sv.y[a] - sv.y[s], rz = sv.z[a] - sv.z[s];
            vw[0] -= g[q].x * rx;
            vw[1] -= g[q].y * ry;
            vw[2] -= g[q].z * rz;
            vw[3] -= 0.5 * (g[q].x * ry + g[q].y * rx);
            vw[4] -= 0.5 * (g[q].x * rz + g[q].z * rx);
            vw[5] -= 0.5 * (g[q].y * rz + g[q].z * ry);
          }
        }
    }
    F[a] += acc.x;
    F[np + a] += acc.y;
    F[2 * np + a] += acc.z;
  }
  if (VIR) {
    __shared__ double wred[4][6];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double x = wave_sum(vw[c]);
      if (lane == 0) wred[wv][c] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      const int c = threadIdx.x;
      vpart[(size_t)blockIdx.x * 6 + c] = ((wred[0][c] + wred[1][c]) + wred[2][c]) + wred[3][c];
    }
  }
}

// Reciprocal k-sum of the virial: sum_m dqfac/d eps_ab(m) (Re(F_U* F_M) +
// Re(F_Up* F_Ud)) with dqfac/d eps_ab = qfac (-delta_ab + 2 m_a m_b (1/m^2 +
// pi^2/alpha^2)), m the reciprocal vector of the grid index.  spec_m holds
// the multipole grid's spectrum, spec_u that of the induced grid (U_d in the
// real part, U_p in the imaginary part of the grid, so F_Ud = (G(m) +
// conj G(-m)) / 2 and F_Up = (G(m) - conj G(-m)) / 2i).  One system; the
// workgroups take chunks of kVirChunk points, sums in a fixed order.
constexpr int kVirChunk = 2048;
template <typename real>
__global__ void __launch_bounds__(256) k_vir_ksum(const SysView* views, const void* spec_m, const void* spec_u,
                                                  double* vpart) {
  const SysView& sv = views[0];
  const int K0 = sv.nfft[0], K1 = sv.nfft[1], K2 = sv.nfft[2];
  const size_t npts = (size_t)K0 * K1 * K2;
  const real* SM = static_cast<const real*>(spec_m);
  const real* SU = static_cast<const real*>(spec_u);
  const real* Q = static_cast<const real*>(sv.qfac);
  const double pa2 = (M_PI / sv.aewald) * (M_PI / sv.aewald);
  double vw[6] = {0, 0, 0, 0, 0, 0};
  const size_t p0 = (size_t)blockIdx.x * kVirChunk, p1 = min(npts, p0 + kVirChunk);
  for (size_t p = p0 + threadIdx.x; p < p1; p += 256) {
    const int x = int(p % K0), y = int((p / K0) % K1), z = int(p / ((size_t)K0 * K1));
    const double q = double(Q[p]);
    if (q == 0.0) continue;
    const int nx = x ? K0 - x : 0, ny = y ? K1 - y : 0, nz = z ? K2 - z : 0;
    const size_t pn = ((size_t)nz * K1 + ny) * K0 + nx;
    const double a = double(SU[2 * p]), b = double(SU[2 * p + 1]);
    const double c = double(SU[2 * pn]), d = double(SU[2 * pn + 1]);
    const double udr = 0.5 * (a + c), udi = 0.5 * (b - d);
    const double upr = 0.5 * (b + d), upi = -0.5 * (a - c);
    const double mr = double(SM[2 * p]), mi = double(SM[2 * p + 1]);
    const double cross = (udr + upr) * mr + (udi + upi) * mi + (upr * udr + upi * udi);
    const double m[3] = {double(x + 1 > (K0 + 1) / 2 ? x - K0 : x) / sv.box[0],
                         double(y + 1 > (K1 + 1) / 2 ? y - K1 : y) / sv.box[1],
                         double(z + 1 > (K2 + 1) / 2 ? z - K2 : z) / sv.box[2]};
    const double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    const double fac = 2.0 * (1.0 / m2 + pa2);
    const double qc = q * cross;
    vw[0] += qc * (fac * m[0] * m[0] - 1.0);
    vw[1] += qc * (fac * m[1] * m[1] - 1.0);
    vw[2] += qc * (fac * m[2] * m[2] - 1.0);
    vw[3] += qc * fac * m[0] * m[1];
    vw[4] += qc * fac * m[0] * m[2];
    vw[5] += qc * fac * m[1] * m[2];
  }
  __shared__ double wred[4][6];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double s = wave_sum(vw[c]);
    if (lane == 0) wred[wv][c] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int c = threadIdx.x;
    vpart[(size_t)blockIdx.x * 6 + c] = ((wred[0][c] + wred[1][c]) + wred[2][c]) + wred[3][c];
  }
}

// f/2 on the forces and torques
// (sf = +1: forces; sf = -1: the gradient)
__global__ void __launch_bounds__(256) k_ep1_finish(int n, size_t np, double* F, double* T, double sf) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 3; ++c) {
    F[(size_t)c * np + i] *= sf * 0.5 * kCoulombFp;
    T[(size_t)c * np + i] *= 0.5 * kCoulombFp;
  }
}

}  // namespace thp
