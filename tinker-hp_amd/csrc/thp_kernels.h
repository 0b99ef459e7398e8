// HIP kernels of the AMOEBA induced-dipole solve for gfx950 (CDNA4).
//
// Design (one GPU, wave64):
//  * atoms live in a spatially sorted order; 64 consecutive atoms form a
//    block = one wavefront.  Per-atom arrays are component-major [c][npad]
//    so that a wavefront touches 64 consecutive words per component.
//  * the neighbor list is full (both directions), one list per atom in
//    slot-major layout nb[s*npad + i] so that a wavefront reading slot s of
//    its 64 atoms issues one coalesced 256-byte load.  Pair kernels give
//    each atom to one lane and walk its list with no atomics and no
//    reaction term; S waves of a workgroup split the list and are summed in
//    LDS in a fixed order, so results are run-to-run deterministic.
//  * pair and grid arithmetic run in `real` (double for the plain FP64
//    path, float for the mixed path); per-atom fields, CG vectors and CG
//    scalars are FP64 everywhere.
//  * reductions use per-block partials plus a last-block finalisation that
//    sums them in a fixed order and writes the FP64 scalars to the device
//    SolveState, so the PCG loop needs no host synchronisation.
#pragma once
#include "thp_internal.h"

namespace thp {

// ------------------------------------------------------------------ utils
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// Minimum image and squared distance evaluated with the reference's
// operation order and no contraction (v1.2/source/image.f:113-140,
// v1.2/source/nblist.f:214-216), so that r^2 is bit-identical to the CPU.
__device__ __forceinline__ double image_exact(double xr, double cell) {
#pragma clang fp contract(off)
  const double cell2 = 0.5 * cell;
  if (fabs(xr) > cell2) {
    const double cel = copysign(cell, xr);
    const double fl = floor(fabs(xr) / cell);
    const double prod = cel * fl;
    xr = xr - prod;
    if (fabs(xr) > cell2) xr = xr - copysign(cell, xr);
  }
  return xr;
}
__device__ __forceinline__ double r2_exact(double xr, double yr, double zr) {
#pragma clang fp contract(off)
  const double a = xr * xr;
  const double b = yr * yr;
  const double c = zr * zr;
  return (a + b) + c;
}

__device__ __forceinline__ double* vsel(const SysView& sv, int which) {
  switch (which) {
    case 0: return sv.mu;
    case 1: return sv.pp;
    case 2: return sv.h;
    case 3: return sv.hreal;
    case 4: return sv.ef;
    case 5: return sv.res;
    default: return sv.zr;
  }
}
enum : int { V_MU = 0, V_PP = 1, V_H = 2, V_HREAL = 3, V_EF = 4, V_RES = 5, V_ZR = 6 };

template <typename T> __device__ __forceinline__ T rsqrt_t(T v);
template <> __device__ __forceinline__ float rsqrt_t(float v) { return rsqrtf(v); }
template <> __device__ __forceinline__ double rsqrt_t(double v) { return rsqrt(v); }
template <typename T> __device__ __forceinline__ T erfc_t(T v);
template <> __device__ __forceinline__ float erfc_t(float v) { return erfcf(v); }
template <> __device__ __forceinline__ double erfc_t(double v) { return erfc(v); }
template <typename T> __device__ __forceinline__ T exp_t(T v);
template <> __device__ __forceinline__ float exp_t(float v) { return expf(v); }
template <> __device__ __forceinline__ double exp_t(double v) { return exp(v); }
template <typename T> __device__ __forceinline__ T sqrt_t(T v);
template <> __device__ __forceinline__ float sqrt_t(float v) { return sqrtf(v); }
template <> __device__ __forceinline__ double sqrt_t(double v) { return sqrt(v); }

// last-block detection: returns true in exactly one block per (kernel,
// system) after every block of that system has published its partials.
__device__ __forceinline__ bool last_block(unsigned int* counter, int nblocks) {
  __shared__ int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int prev = atomicAdd(counter, 1u);
    is_last = (prev == (unsigned int)(nblocks - 1));
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    if (threadIdx.x == 0) *counter = 0u;
  }
  return is_last != 0;
}

// fixed-order sum of partial[q][0..nblk)[r] by one 64-lane wave
__device__ __forceinline__ double sum_partials(const double* part, int nblk, int r) {
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) acc += part[b * 2 + r];
  return wave_sum(acc);
}

// ===================================================== neighbor lists
// Block bounding boxes in a frame local to the block's first atom (minimum
// image), stored as centre + half extent.  One wavefront per block.
__global__ void __launch_bounds__(64) k_bbox(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int lane = threadIdx.x;
  const int i0 = w.blk * kBlock;
  const int i = i0 + lane;
  const bool ok = i < sv.n;
  const double x0 = sv.x[i0], y0 = sv.y[i0], z0 = sv.z[i0];
  double dx = 0, dy = 0, dz = 0;
  if (ok) {
    dx = sv.x[i] - x0;
    dy = sv.y[i] - y0;
    dz = sv.z[i] - z0;
    dx -= sv.box[0] * rint(dx / sv.box[0]);
    dy -= sv.box[1] * rint(dy / sv.box[1]);
    dz -= sv.box[2] * rint(dz / sv.box[2]);
  }
  const double inf = 1e300;
  const double mnx = wave_min(ok ? dx : inf), mxx = wave_max(ok ? dx : -inf);
  const double mny = wave_min(ok ? dy : inf), mxy = wave_max(ok ? dy : -inf);
  const double mnz = wave_min(ok ? dz : inf), mxz = wave_max(ok ? dz : -inf);
  if (lane < 6) {
    double v;
    switch (lane) {
      case 0: v = x0 + 0.5 * (mnx + mxx); break;
      case 1: v = y0 + 0.5 * (mny + mxy); break;
      case 2: v = z0 + 0.5 * (mnz + mxz); break;
      case 3: v = 0.5 * (mxx - mnx); break;
      case 4: v = 0.5 * (mxy - mny); break;
      default: v = 0.5 * (mxz - mnz); break;
    }
    sv.bbox[w.blk * 6 + lane] = v;
  }
}

// Candidate block pairs: block J is a candidate of block I when the
// periodic gap between their boxes is within cut+buffer (with a small
// relative slack; the per-atom test below is the exact one).
// One 256-thread workgroup per block I; the list of J is written in
// increasing order.  FILL=false only counts.
template <bool FILL>
__global__ void __launch_bounds__(256) k_blockpairs(const SysView* views, const Work* work,
                                                    int* bcount, int* blist, int maxbl) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const double* bb = sv.bbox;
  const int I = w.blk;
  const double cx = bb[I * 6], cy = bb[I * 6 + 1], cz = bb[I * 6 + 2];
  const double hx = bb[I * 6 + 3], hy = bb[I * 6 + 4], hz = bb[I * 6 + 5];
  const double lim = sqrt(sv.mbuf2) * (1.0 + 1e-9) + 1e-9;
  const double lim2 = lim * lim;
  __shared__ int wave_cnt[4];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j0 = 0; j0 < sv.nblk; j0 += 256) {
    const int J = j0 + threadIdx.x;
    bool cand = false;
    if (J < sv.nblk) {
      double dx = bb[J * 6] - cx, dy = bb[J * 6 + 1] - cy, dz = bb[J * 6 + 2] - cz;
      dx -= sv.box[0] * rint(dx / sv.box[0]);
      dy -= sv.box[1] * rint(dy / sv.box[1]);
      dz -= sv.box[2] * rint(dz / sv.box[2]);
      const double gx = fmax(0.0, fabs(dx) - hx - bb[J * 6 + 3]);
      const double gy = fmax(0.0, fabs(dy) - hy - bb[J * 6 + 4]);
      const double gz = fmax(0.0, fabs(dz) - hz - bb[J * 6 + 5]);
      cand = gx * gx + gy * gy + gz * gz <= lim2;
    }
    const unsigned long long m = __ballot(cand);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_cnt[wv] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int q = 0; q < wv; ++q) off += wave_cnt[q];
    if (FILL && cand && off + before < maxbl) blist[(size_t)I * maxbl + off + before] = J;
    __syncthreads();
    if (threadIdx.x == 0) base += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) bcount[I] = base;
}

// Verlet list of each atom: all k != i with r^2 <= (cut+buffer)^2, k in
// increasing order.  One wavefront per block I; the 64 atoms of each
// candidate block are staged in LDS and tested by every lane.
template <bool FILL>
__global__ void __launch_bounds__(64) k_nblist_atoms(const SysView* views, const Work* work,
                                                      const int* bcount, const int* blist, int maxbl) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int lane = threadIdx.x;
  const int i = w.blk * kBlock + lane;
  const bool ok = i < sv.n;
  __shared__ double sx[kBlock], sy[kBlock], sz[kBlock];
  const double xi = ok ? sv.x[i] : 0.0, yi = ok ? sv.y[i] : 0.0, zi = ok ? sv.z[i] : 0.0;
  int cnt = 0;
  const int nbl = bcount[w.blk];
  for (int q = 0; q < nbl; ++q) {
    const int J = blist[(size_t)w.blk * maxbl + q];
    const int k0 = J * kBlock;
    const int kk = k0 + lane;
    __syncthreads();
    sx[lane] = kk < sv.n ? sv.x[kk] : 0.0;
    sy[lane] = kk < sv.n ? sv.y[kk] : 0.0;
    sz[lane] = kk < sv.n ? sv.z[kk] : 0.0;
    __syncthreads();
    if (!ok) continue;
    const int mmax = min(kBlock, sv.n - k0);
    for (int m = 0; m < mmax; ++m) {
      const int k = k0 + m;
      if (k == i) continue;
      const double xr = image_exact(xi - sx[m], sv.box[0]);
      const double yr = image_exact(yi - sy[m], sv.box[1]);
      const double zr = image_exact(zi - sz[m], sv.box[2]);
      if (r2_exact(xr, yr, zr) <= sv.mbuf2) {
        if (FILL) {
          if (cnt < sv.maxnb) sv.nb[(size_t)cnt * sv.npad + i] = k;
          else *sv.err = THP_ERR_LIST;
        }
        ++cnt;
      }
    }
  }
  if (ok) sv.nbcount[i] = cnt;
}

// Tight list: entries of the Verlet list with r^2 <= cut^2 (the test of
// efld0_direct / tmatxb_pme, newinduce_pme.f:1160 and :1516), tagged with
// the code of the pair's (d,p,u) scale triple (0 = all ones).  One thread
// per atom.
__global__ void __launch_bounds__(64) k_tight(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const double xi = sv.x[i], yi = sv.y[i], zi = sv.z[i];
  const int cnt = sv.nbcount[i];
  const int e0 = sv.excl_off[i], e1 = sv.excl_off[i + 1];
  int t = 0;
  for (int s = 0; s < cnt; ++s) {
    const int k = sv.nb[(size_t)s * sv.npad + i];
    const double xr = image_exact(sv.x[k] - xi, sv.box[0]);
    const double yr = image_exact(sv.y[k] - yi, sv.box[1]);
    const double zr = image_exact(sv.z[k] - zi, sv.box[2]);
    if (r2_exact(xr, yr, zr) <= sv.cut2) {
      int code = 0;
      for (int e = e0; e < e1; ++e)
        if (sv.excl_k[e] == k) {
          code = sv.excl_code[e];
          break;
        }
      sv.tnb[(size_t)t * sv.npad + i] = k | (code << kCodeShift);
      ++t;
    }
  }
  sv.tcount[i] = t;
}

// ===================================================== pair kernels
// Ewald real-space coefficients bn_1..bn_3 (dampewald, damping.f:17-56)
// and Thole lambda_3/5/7 (damptholed, damping.f:142-215).
template <typename real, int NBN>
__device__ __forceinline__ void ewald_bn(real r, real r2, real aewald, real rr1, real* bn) {
  const real ra = aewald * r;
  const real bn0 = erfc_t(ra) * rr1;
  const real expterm = exp_t(-ra * ra);
  const real aesq2 = real(2) * aewald * aewald;
  real afac = real(1) / (real(1.7724538509055160273) * aewald);
  const real ir2 = rr1 * rr1;
  real prev = bn0;
#pragma unroll
  for (int q = 1; q <= NBN; ++q) {
    afac *= aesq2;
    prev = (real(2 * q - 1) * prev + afac * expterm) * ir2;
    bn[q - 1] = prev;
  }
  (void)r2;
}

template <typename real, int ORDER7>
__device__ __forceinline__ void thole_l(real r, real pdi, real pdk, real thi, real thk, real* l) {
  l[0] = l[1] = l[2] = real(1);
  real damp = pdi * pdk;
  real pg = fmin(thi, thk);
  if (pg == real(0)) pg = fmax(thi, thk);
  if (damp != real(0) && pg != real(0)) {
    const real rd = r / damp;
    damp = pg * (rd * rd * rd);
    if (damp < real(50)) {
      const real e = exp_t(-damp);
      l[0] = real(1) - e;
      l[1] = real(1) - e * (real(1) + damp);
      if (ORDER7) l[2] = real(1) - e * (real(1) + damp + real(0.6) * damp * damp);
    }
  }
}

// displacement x_k - x_i evaluated in FP64 (minimum image), then rounded
// to the pair arithmetic type
template <typename real>
__device__ __forceinline__ void disp(const SysView& sv, int k, double xi, double yi, double zi,
                                     real& xr, real& yr, real& zr) {
  double dx = sv.x[k] - xi, dy = sv.y[k] - yi, dz = sv.z[k] - zi;
  dx = image_exact(dx, sv.box[0]);
  dy = image_exact(dy, sv.box[1]);
  dz = image_exact(dz, sv.box[2]);
  xr = real(dx);
  yr = real(dy);
  zr = real(dz);
}

// efld0_direct: permanent field (d and p) at each atom from its tight
// list.  newinduce_pme.f:1116-1264; field at i due to k (fid/fip).
template <typename real, int S>
__global__ void __launch_bounds__(64 * S) k_efld0(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = w.blk * kBlock + lane;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (i < sv.n) {
    const double xi = sv.x[i], yi = sv.y[i], zi = sv.z[i];
    const real pdi = real(sv.pdamp[i]), thi = real(sv.thole[i]);
    const real aew = real(sv.aewald);
    const int cnt = sv.tcount[i];
    const int s0 = (cnt * wv) / S, s1 = (cnt * (wv + 1)) / S;
    const size_t np = sv.npad;
    for (int s = s0; s < s1; ++s) {
      const int e = sv.tnb[(size_t)s * np + i];
      const int k = e & (kMaxAtoms - 1);
      const int code = e >> kCodeShift;
      real xr, yr, zr;
      disp<real>(sv, k, xi, yi, zi, xr, yr, zr);
      const real r2 = xr * xr + yr * yr + zr * zr;
      const real rr1 = rsqrt_t(r2);
      const real r = r2 * rr1;
      const real rr2 = rr1 * rr1;
      const real rr3 = rr2 * rr1;
      const real rr5 = real(3) * rr2 * rr3;
      const real rr7 = real(5) * rr2 * rr5;
      const real ck = real(sv.rpole[0 * np + k]);
      const real dkx = real(sv.rpole[1 * np + k]), dky = real(sv.rpole[2 * np + k]), dkz = real(sv.rpole[3 * np + k]);
      const real qkxx = real(sv.rpole[4 * np + k]), qkxy = real(sv.rpole[5 * np + k]);
      const real qkxz = real(sv.rpole[6 * np + k]), qkyy = real(sv.rpole[7 * np + k]);
      const real qkyz = real(sv.rpole[8 * np + k]), qkzz = real(sv.rpole[9 * np + k]);
      const real dkr = dkx * xr + dky * yr + dkz * zr;
      const real qkx = qkxx * xr + qkxy * yr + qkxz * zr;
      const real qky = qkxy * xr + qkyy * yr + qkyz * zr;
      const real qkz = qkxz * xr + qkyz * yr + qkzz * zr;
      const real qkr = qkx * xr + qky * yr + qkz * zr;
      real bn[3];
      ewald_bn<real, 3>(r, r2, aew, rr1, bn);
      real l[3];
      thole_l<real, 1>(r, pdi, real(sv.pdamp[k]), thi, real(sv.thole[k]), l);
      real dsc = real(1), psc = real(1);
      if (code) {
        dsc = real(sv.scale_tab[code * 3 + 0]);
        psc = real(sv.scale_tab[code * 3 + 1]);
      }
#pragma unroll
      for (int pass = 0; pass < 2; ++pass) {
        const real sc = pass == 0 ? dsc : psc;
        const real dmp3 = bn[0] - (real(1) - sc * l[0]) * rr3;
        const real dmp5 = bn[1] - (real(1) - sc * l[1]) * rr5;
        const real dmp7 = bn[2] - (real(1) - sc * l[2]) * rr7;
        const real ai = dmp3 * ck - dmp5 * dkr + dmp7 * qkr;
        const real d5 = real(2) * dmp5;
        acc[pass * 3 + 0] += double(-xr * ai - dmp3 * dkx + d5 * qkx);
        acc[pass * 3 + 1] += double(-yr * ai - dmp3 * dky + d5 * qky);
        acc[pass * 3 + 2] += double(-zr * ai - dmp3 * dkz + d5 * qkz);
      }
    }
  }
  __shared__ double red[S > 1 ? S : 1][6][64];
  if (S > 1) {
#pragma unroll
    for (int c = 0; c < 6; ++c) red[wv][c][lane] = acc[c];
    __syncthreads();
  }
  if (wv == 0 && i < sv.n) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = acc[c];
      for (int q = 1; q < S; ++q) v += red[q][c][lane];
      sv.ef[(size_t)c * sv.npad + i] = v;
    }
  }
}

// tmatxb_pme: out = diag(1/alpha) in - E_real(in), the Thole-damped Ewald
// real-space dipole field over the tight list (newinduce_pme.f:1516-1661).
template <typename real, int S>
__global__ void __launch_bounds__(64 * S) k_tmatxb(const SysView* views, const Work* work, int vin,
                                                    int vout, int dodiag, int check_conv) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  if (check_conv && sv.st->conv) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = w.blk * kBlock + lane;
  const double* in = vsel(sv, vin);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const size_t np = sv.npad;
  if (i < sv.n) {
    const double xi = sv.x[i], yi = sv.y[i], zi = sv.z[i];
    const real pdi = real(sv.pdamp[i]), thi = real(sv.thole[i]);
    const real aew = real(sv.aewald);
    const int cnt = sv.tcount[i];
    const int s0 = (cnt * wv) / S, s1 = (cnt * (wv + 1)) / S;
    for (int s = s0; s < s1; ++s) {
      const int e = sv.tnb[(size_t)s * np + i];
      const int k = e & (kMaxAtoms - 1);
      const int code = e >> kCodeShift;
      real xr, yr, zr;
      disp<real>(sv, k, xi, yi, zi, xr, yr, zr);
      const real r2 = xr * xr + yr * yr + zr * zr;
      const real rr1 = rsqrt_t(r2);
      const real r = r2 * rr1;
      const real rr2 = rr1 * rr1;
      const real rr3 = rr2 * rr1;
      const real rr5 = real(3) * rr2 * rr3;
      real bn[2];
      ewald_bn<real, 2>(r, r2, aew, rr1, bn);
      real l[3];
      thole_l<real, 0>(r, pdi, real(sv.pdamp[k]), thi, real(sv.thole[k]), l);
      real usc = real(1);
      if (code) usc = real(sv.scale_tab[code * 3 + 2]);
      const real dmp3 = bn[0] - (real(1) - usc * l[0]) * rr3;
      const real dmp5 = bn[1] - (real(1) - usc * l[1]) * rr5;
#pragma unroll
      for (int rhs = 0; rhs < 2; ++rhs) {
        const real ux = real(in[(size_t)(rhs * 3 + 0) * np + k]);
        const real uy = real(in[(size_t)(rhs * 3 + 1) * np + k]);
        const real uz = real(in[(size_t)(rhs * 3 + 2) * np + k]);
        const real ur = dmp5 * (xr * ux + yr * uy + zr * uz);
        // efi -= dlocal . u_k with dlocal = -dmp3 I + dmp5 r r^T
        acc[rhs * 3 + 0] += double(dmp3 * ux - ur * xr);
        acc[rhs * 3 + 1] += double(dmp3 * uy - ur * yr);
        acc[rhs * 3 + 2] += double(dmp3 * uz - ur * zr);
      }
    }
  }
  __shared__ double red[S > 1 ? S : 1][6][64];
  if (S > 1) {
#pragma unroll
    for (int c = 0; c < 6; ++c) red[wv][c][lane] = acc[c];
    __syncthreads();
  }
  if (wv == 0 && i < sv.n) {
    const double pol = sv.pol[i];
    const double ipol = pol == 0.0 ? 1.0 / kTinyPol : 1.0 / pol;
    double* out = vsel(sv, vout);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = acc[c];
      for (int q = 1; q < S; ++q) v += red[q][c][lane];
      if (dodiag) v += in[(size_t)c * np + i] * ipol;
      out[(size_t)c * np + i] = v;
    }
  }
}

// ===================================================== PME
// Grid position and B-spline coefficients of one site along one axis
// (bspline_fill_site, pmestuff.f:13-50; bsplgen, pmestuff.f:67-156).
// The node index follows grid_*_site: first node 0-based = ifr - 4 for
// order 5, wrapped by +nfft when negative.
template <typename real, int LEVELS>
__device__ __forceinline__ int bspline_axis(double xc, double rc, int nfft, real th[LEVELS][kOrder]) {
  double w = xc * rc;
  const double fr = double(nfft) * ((w - round(w)) + 0.5);
  int ifr = int(fr - 1.0e-8);
  ifr = min(max(ifr, 0), nfft - 1);  // in range for finite input; guards the grid index
  const real ww = real(fr - double(ifr));
  // order-k B-spline values in t[k][1..k]
  real t[kOrder + 1][kOrder + 1];
  t[2][2] = ww;
  t[2][1] = real(1) - ww;
  t[3][3] = real(0.5) * ww * t[2][2];
  t[3][2] = real(0.5) * ((real(1) + ww) * t[2][1] + (real(2) - ww) * t[2][2]);
  t[3][1] = real(0.5) * (real(1) - ww) * t[2][1];
#pragma unroll
  for (int i = 4; i <= kOrder; ++i) {
    const int k = i - 1;
    const real denom = real(1) / real(k);
    t[i][i] = denom * ww * t[k][k];
#pragma unroll
    for (int j = 1; j <= i - 2; ++j)
      t[i][i - j] = denom * ((ww + real(j)) * t[k][i - j - 1] + (real(i - j) - ww) * t[k][i - j]);
    t[i][1] = denom * (real(1) - ww) * t[k][1];
  }
#pragma unroll
  for (int m = 1; m <= kOrder; ++m) th[0][m - 1] = t[kOrder][m];
  if constexpr (LEVELS >= 2) {
    // first derivative from the order-4 values
    real a[kOrder + 1];
#pragma unroll
    for (int m = 1; m <= kOrder - 1; ++m) a[m] = t[kOrder - 1][m];
    a[kOrder] = a[kOrder - 1];
#pragma unroll
    for (int m = kOrder - 1; m >= 2; --m) a[m] = a[m - 1] - a[m];
    a[1] = -a[1];
#pragma unroll
    for (int m = 1; m <= kOrder; ++m) th[1][m - 1] = a[m];
  }
  if constexpr (LEVELS >= 3) {
    real a[kOrder + 1];
#pragma unroll
    for (int m = 1; m <= kOrder - 2; ++m) a[m] = t[kOrder - 2][m];
    a[kOrder - 1] = a[kOrder - 2];
#pragma unroll
    for (int m = kOrder - 2; m >= 2; --m) a[m] = a[m - 1] - a[m];
    a[1] = -a[1];
    a[kOrder] = a[kOrder - 1];
#pragma unroll
    for (int m = kOrder - 1; m >= 2; --m) a[m] = a[m - 1] - a[m];
    a[1] = -a[1];
#pragma unroll
    for (int m = 1; m <= kOrder; ++m) th[2][m - 1] = a[m];
  }
  int g0 = ifr - 4;
  if (g0 < 0) g0 += nfft;
  return g0;
}

template <typename real> struct Cplx { real re, im; };

__device__ __forceinline__ void atomic_add_real(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_real(float* p, float v) { unsafeAtomicAdd(p, v); }

// zero a system's grid (a fill kernel, safe inside captured graphs)
template <typename real>
__global__ void __launch_bounds__(256) k_zero_grid(const SysView* views, int sys, size_t count) {
  real* g = static_cast<real*>(views[sys].grid);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (size_t)gridDim.x * blockDim.x)
    g[q] = real(0);
}

// grid_uind_site (pmestuff.f:280-400): fractional d dipoles into the real
// part and p dipoles into the imaginary part; one thread per atom.
template <typename real>
__global__ void __launch_bounds__(64) k_spread_uind(const SysView* views, const Work* work, int vin,
                                                     int frac_in, int check_conv) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  if (check_conv && sv.st->conv) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const size_t np = sv.npad;
  const double* in = vsel(sv, vin);
  real fd[3], fp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double sc = frac_in ? 1.0 : sv.afrac[c];
    fd[c] = real(sc * in[(size_t)c * np + i]);
    fp[c] = real(sc * in[(size_t)(3 + c) * np + i]);
  }
  real t1[2][kOrder], t2[2][kOrder], t3[2][kOrder];
  const int g1 = bspline_axis<real, 2>(sv.x[i], 1.0 / sv.box[0], sv.nfft[0], t1);
  const int g2 = bspline_axis<real, 2>(sv.y[i], 1.0 / sv.box[1], sv.nfft[1], t2);
  const int g3 = bspline_axis<real, 2>(sv.z[i], 1.0 / sv.box[2], sv.nfft[2], t3);
  real* grid = static_cast<real*>(sv.grid);
  const int n1 = sv.nfft[0], n2 = sv.nfft[1], n3 = sv.nfft[2];
#pragma unroll
  for (int mk = 0; mk < kOrder; ++mk) {
    int k = g3 + mk;
    if (k >= n3) k -= n3;
    const real v0 = t3[0][mk], v1 = t3[1][mk];
#pragma unroll
    for (int mj = 0; mj < kOrder; ++mj) {
      int j = g2 + mj;
      if (j >= n2) j -= n2;
      const real u0 = t2[0][mj], u1 = t2[1][mj];
      const real term01 = fd[1] * u1 * v0 + fd[2] * u0 * v1;
      const real term11 = fd[0] * u0 * v0;
      const real term02 = fp[1] * u1 * v0 + fp[2] * u0 * v1;
      const real term12 = fp[0] * u0 * v0;
      real* row = grid + ((size_t)k * n2 + j) * n1 * 2;
#pragma unroll
      for (int mi = 0; mi < kOrder; ++mi) {
        int ii = g1 + mi;
        if (ii >= n1) ii -= n1;
        const real t0 = t1[0][mi], tt1 = t1[1][mi];
        atomic_add_real(row + 2 * ii, term01 * t0 + term11 * tt1);
        atomic_add_real(row + 2 * ii + 1, term02 * t0 + term12 * tt1);
      }
    }
  }
}

// grid_mpole_site (pmestuff.f:163-274) after cmp_to_fmp_site; real part only.
template <typename real>
__global__ void __launch_bounds__(64) k_spread_mpole(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const size_t np = sv.npad;
  // cmp order (m, dx,dy,dz, qxx,qyy,qzz, 2qxy, 2qxz, 2qyz) from the stored
  // (q, d, qxx,qxy,qxz,qyy,qyz,qzz)
  double cmp[10];
  cmp[0] = sv.rpole[0 * np + i];
  cmp[1] = sv.rpole[1 * np + i];
  cmp[2] = sv.rpole[2 * np + i];
  cmp[3] = sv.rpole[3 * np + i];
  cmp[4] = sv.rpole[4 * np + i];
  cmp[5] = sv.rpole[7 * np + i];
  cmp[6] = sv.rpole[9 * np + i];
  cmp[7] = 2.0 * sv.rpole[5 * np + i];
  cmp[8] = 2.0 * sv.rpole[6 * np + i];
  cmp[9] = 2.0 * sv.rpole[8 * np + i];
  real fmp[10];
  fmp[0] = real(sv.ctf[0][0] * cmp[0]);
#pragma unroll
  for (int j = 1; j < 4; ++j) {
    double a = 0.0;
#pragma unroll
    for (int k = 1; k < 4; ++k) a += sv.ctf[j][k] * cmp[k];
    fmp[j] = real(a);
  }
#pragma unroll
  for (int j = 4; j < 10; ++j) {
    double a = 0.0;
#pragma unroll
    for (int k = 4; k < 10; ++k) a += sv.ctf[j][k] * cmp[k];
    fmp[j] = real(a);
  }
  real t1[3][kOrder], t2[3][kOrder], t3[3][kOrder];
  const int g1 = bspline_axis<real, 3>(sv.x[i], 1.0 / sv.box[0], sv.nfft[0], t1);
  const int g2 = bspline_axis<real, 3>(sv.y[i], 1.0 / sv.box[1], sv.nfft[1], t2);
  const int g3 = bspline_axis<real, 3>(sv.z[i], 1.0 / sv.box[2], sv.nfft[2], t3);
  real* grid = static_cast<real*>(sv.grid);
  const int n1 = sv.nfft[0], n2 = sv.nfft[1], n3 = sv.nfft[2];
#pragma unroll
  for (int mk = 0; mk < kOrder; ++mk) {
    int k = g3 + mk;
    if (k >= n3) k -= n3;
    const real v0 = t3[0][mk], v1 = t3[1][mk], v2 = t3[2][mk];
#pragma unroll
    for (int mj = 0; mj < kOrder; ++mj) {
      int j = g2 + mj;
      if (j >= n2) j -= n2;
      const real u0 = t2[0][mj], u1 = t2[1][mj], u2 = t2[2][mj];
      const real term0 = fmp[0] * u0 * v0 + fmp[2] * u1 * v0 + fmp[3] * u0 * v1 + fmp[5] * u2 * v0 +
                         fmp[6] * u0 * v2 + fmp[9] * u1 * v1;
      const real term1 = fmp[1] * u0 * v0 + fmp[7] * u1 * v0 + fmp[8] * u0 * v1;
      const real term2 = fmp[4] * u0 * v0;
      real* row = grid + ((size_t)k * n2 + j) * n1 * 2;
#pragma unroll
      for (int mi = 0; mi < kOrder; ++mi) {
        int ii = g1 + mi;
        if (ii >= n1) ii -= n1;
        atomic_add_real(row + 2 * ii, term0 * t1[0][mi] + term1 * t1[1][mi] + term2 * t1[2][mi]);
      }
    }
  }
}

// scalar-sum convolution between the forward and backward transforms
// (qfac, newinduce_pme.f:1790-1861)
template <typename real>
__global__ void __launch_bounds__(256) k_conv(const SysView* views, int sys, size_t npts, int check_conv) {
  const SysView& sv = views[sys];
  if (check_conv && sv.st->conv) return;
  real* g = static_cast<real*>(sv.grid);
  const real* q = static_cast<const real*>(sv.qfac);
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npts; p += (size_t)gridDim.x * blockDim.x) {
    const real f = q[p];
    g[2 * p] *= f;
    g[2 * p + 1] *= f;
  }
}

// fphi_uind_site first derivatives (pmestuff.f:746-1037, fdip_phi(2:4) of
// both grids) and the fractional->Cartesian field (newinduce_pme.f:2103-2110).
// Accumulation of the gathered sums is FP64.
template <typename real>
__device__ __forceinline__ void gather_uind(const SysView& sv, int i, double p1[3], double p2[3]) {
  real t1[2][kOrder], t2[2][kOrder], t3[2][kOrder];
  const int g1 = bspline_axis<real, 2>(sv.x[i], 1.0 / sv.box[0], sv.nfft[0], t1);
  const int g2 = bspline_axis<real, 2>(sv.y[i], 1.0 / sv.box[1], sv.nfft[1], t2);
  const int g3 = bspline_axis<real, 2>(sv.z[i], 1.0 / sv.box[2], sv.nfft[2], t3);
  const real* grid = static_cast<const real*>(sv.grid);
  const int n1 = sv.nfft[0], n2 = sv.nfft[1], n3 = sv.nfft[2];
  double a[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int mk = 0; mk < kOrder; ++mk) {
    int k = g3 + mk;
    if (k >= n3) k -= n3;
    const real v0 = t3[0][mk], v1 = t3[1][mk];
    real tu00_1 = 0, tu10_1 = 0, tu01_1 = 0, tu00_2 = 0, tu10_2 = 0, tu01_2 = 0;
#pragma unroll
    for (int mj = 0; mj < kOrder; ++mj) {
      int j = g2 + mj;
      if (j >= n2) j -= n2;
      const real u0 = t2[0][mj], u1 = t2[1][mj];
      const real* row = grid + ((size_t)k * n2 + j) * n1 * 2;
      real t0_1 = 0, t1_1 = 0, t0_2 = 0, t1_2 = 0;
#pragma unroll
      for (int mi = 0; mi < kOrder; ++mi) {
        int ii = g1 + mi;
        if (ii >= n1) ii -= n1;
        const real q1 = row[2 * ii], q2 = row[2 * ii + 1];
        t0_1 += q1 * t1[0][mi];
        t1_1 += q1 * t1[1][mi];
        t0_2 += q2 * t1[0][mi];
        t1_2 += q2 * t1[1][mi];
      }
      tu00_1 += t0_1 * u0;
      tu10_1 += t1_1 * u0;
      tu01_1 += t0_1 * u1;
      tu00_2 += t0_2 * u0;
      tu10_2 += t1_2 * u0;
      tu01_2 += t0_2 * u1;
    }
    a[0] += double(tu10_1 * v0);
    a[1] += double(tu01_1 * v0);
    a[2] += double(tu00_1 * v1);
    a[3] += double(tu10_2 * v0);
    a[4] += double(tu01_2 * v0);
    a[5] += double(tu00_2 * v1);
  }
  p1[0] = a[0];
  p1[1] = a[1];
  p1[2] = a[2];
  p2[0] = a[3];
  p2[1] = a[4];
  p2[2] = a[5];
}

// MODE 0: write the reciprocal dipole field of vin into vout (frac_out: the
//         fractional derivatives fdip_phi(2:4) instead).
// MODE 1: PCG start (newinduce_pme2.f:406-438): h = hreal + dip - term*mu,
//         res = ef - h, zr = diag*res, pp = zr, ggold = sum res.zr.
// MODE 2: PCG iteration (:452-473): h = hreal + dip - term*pp, gg = pp.h.
template <typename real, int MODE>
__global__ void __launch_bounds__(64) k_gather_uind(const SysView* views, const Work* work, int vin, int vout,
                                                     int frac_out) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  if (MODE == 2 && sv.st->conv) return;
  const int i = w.blk * kBlock + threadIdx.x;
  const size_t np = sv.npad;
  const double* in = vsel(sv, vin);
  double q[2] = {0.0, 0.0};
  if (i < sv.n) {
    double p1[3], p2[3];
    gather_uind<real>(sv, i, p1, p2);
    double dip[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double sc = (MODE == 0 && frac_out) ? 1.0 : sv.afrac[c];
      dip[c] = sc * p1[c];
      dip[3 + c] = sc * p2[c];
    }
    if (MODE == 0) {
      double* out = vsel(sv, vout);
#pragma unroll
      for (int c = 0; c < 6; ++c) out[(size_t)c * np + i] = dip[c];
    } else {
      const double term = sv.term_self;
      const double diag = sv.pol[i] == 0.0 ? kTinyPol : sv.pol[i];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const size_t o = (size_t)c * np + i;
        const double u = in[o];
        const double h = sv.hreal[o] + dip[c] - term * u;
        sv.h[o] = h;
        if (MODE == 1) {
          const double r = sv.ef[o] - h;
          const double z = diag * r;
          sv.res[o] = r;
          sv.zr[o] = z;
          sv.pp[o] = z;
          q[c / 3] += r * z;
        } else {
          q[c / 3] += u * h;
        }
      }
    }
  }
  if (MODE == 0) return;
  const double s0 = wave_sum(q[0]), s1 = wave_sum(q[1]);
  double* part = sv.partial;  // slot 0
  if (threadIdx.x == 0) {
    part[w.blk * 2] = s0;
    part[w.blk * 2 + 1] = s1;
  }
  if (last_block(&sv.st->counter[0], sv.nblk)) {
    const double a0 = sum_partials(part, sv.nblk, 0), a1 = sum_partials(part, sv.nblk, 1);
    if (threadIdx.x == 0) {
      SolveState* st = sv.st;
      if (MODE == 1) {
        st->ggold[0] = a0;
        st->ggold[1] = a1;
      } else {
        st->gg[0] = a0;
        st->gg[1] = a1;
        if (a0 <= 0.0 || a1 <= 0.0) st->n_indef += 1;
        if (a0 == 0.0 || a1 == 0.0) st->conv = kZeroCurv;  // reference returns early
      }
    }
  }
}

// fphi_mpole_site (pmestuff.f:453-631, potential and derivatives through
// second order) + fphi_to_cphi_site (:1190-1215).  When add_to_ef, the
// permanent field becomes ef = E_direct - cphi(2:4) + term*d for both rhs
// (newinduce_pme2.f:134-147); cphi (optional) receives [10][npad].
template <typename real>
__global__ void __launch_bounds__(64) k_gather_mpole(const SysView* views, const Work* work, double* cphi_out,
                                                      int add_to_ef) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const size_t np = sv.npad;
  real t1[3][kOrder], t2[3][kOrder], t3[3][kOrder];
  const int g1 = bspline_axis<real, 3>(sv.x[i], 1.0 / sv.box[0], sv.nfft[0], t1);
  const int g2 = bspline_axis<real, 3>(sv.y[i], 1.0 / sv.box[1], sv.nfft[1], t2);
  const int g3 = bspline_axis<real, 3>(sv.z[i], 1.0 / sv.box[2], sv.nfft[2], t3);
  const real* grid = static_cast<const real*>(sv.grid);
  const int n1 = sv.nfft[0], n2 = sv.nfft[1], n3 = sv.nfft[2];
  double tuv[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int mk = 0; mk < kOrder; ++mk) {
    int k = g3 + mk;
    if (k >= n3) k -= n3;
    const real v0 = t3[0][mk], v1 = t3[1][mk], v2 = t3[2][mk];
    real tu00 = 0, tu10 = 0, tu01 = 0, tu20 = 0, tu11 = 0, tu02 = 0;
#pragma unroll
    for (int mj = 0; mj < kOrder; ++mj) {
      int j = g2 + mj;
      if (j >= n2) j -= n2;
      const real u0 = t2[0][mj], u1 = t2[1][mj], u2 = t2[2][mj];
      const real* row = grid + ((size_t)k * n2 + j) * n1 * 2;
      real t0 = 0, tt1 = 0, tt2 = 0;
#pragma unroll
      for (int mi = 0; mi < kOrder; ++mi) {
        int ii = g1 + mi;
        if (ii >= n1) ii -= n1;
        const real tq = row[2 * ii];
        t0 += tq * t1[0][mi];
        tt1 += tq * t1[1][mi];
        tt2 += tq * t1[2][mi];
      }
      tu00 += t0 * u0;
      tu10 += tt1 * u0;
      tu01 += t0 * u1;
      tu20 += tt2 * u0;
      tu11 += tt1 * u1;
      tu02 += t0 * u2;
    }
    tuv[0] += double(tu00 * v0);
    tuv[1] += double(tu10 * v0);
    tuv[2] += double(tu01 * v0);
    tuv[3] += double(tu00 * v1);
    tuv[4] += double(tu20 * v0);
    tuv[5] += double(tu02 * v0);
    tuv[6] += double(tu00 * v2);
    tuv[7] += double(tu11 * v0);
    tuv[8] += double(tu10 * v1);
    tuv[9] += double(tu01 * v1);
  }
  double c[10];
  c[0] = sv.ftc[0][0] * tuv[0];
#pragma unroll
  for (int j = 1; j < 4; ++j) {
    double a = 0.0;
#pragma unroll
    for (int k = 1; k < 4; ++k) a += sv.ftc[j][k] * tuv[k];
    c[j] = a;
  }
#pragma unroll
  for (int j = 4; j < 10; ++j) {
    double a = 0.0;
#pragma unroll
    for (int k = 4; k < 10; ++k) a += sv.ftc[j][k] * tuv[k];
    c[j] = a;
  }
  if (cphi_out) {
#pragma unroll
    for (int j = 0; j < 10; ++j) cphi_out[(size_t)j * np + i] = c[j];
  }
  if (add_to_ef) {
    const double term = sv.term_self;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double d = -c[1 + j] + term * sv.rpole[(size_t)(1 + j) * np + i];
      sv.ef[(size_t)j * np + i] += d;
      sv.ef[(size_t)(3 + j) * np + i] += d;
    }
  }
}

// ===================================================== PCG vector kernels
// direct guess mu = alpha*ef (polgsf = 0, newinduce_pme2.f:180-188)
__global__ void __launch_bounds__(64) k_guess(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const double a = sv.pol[i];
#pragma unroll
  for (int c = 0; c < 6; ++c) sv.mu[(size_t)c * sv.npad + i] = a * sv.ef[(size_t)c * sv.npad + i];
}

// pcg_b (newinduce_pme2.f:504-541): alpha = ggold/gg; mu += alpha pp;
// res -= alpha h; zr = diag*res; partial ggnew and ene.  Also clears the
// system's PME grid for the next spreading pass.
template <typename real>
__global__ void __launch_bounds__(64) k_cg_update1(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  if (sv.st->conv) return;
  const int i = w.blk * kBlock + threadIdx.x;
  const size_t np = sv.npad;
  const SolveState* st = sv.st;
  const double alpha[2] = {st->ggold[0] / st->gg[0], st->ggold[1] / st->gg[1]};
  double gn[2] = {0, 0}, en[2] = {0, 0};
  if (i < sv.n) {
    const double diag = sv.pol[i] == 0.0 ? kTinyPol : sv.pol[i];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      const double a = alpha[c / 3];
      const double m = sv.mu[o] + a * sv.pp[o];
      const double r = sv.res[o] - a * sv.h[o];
      const double z = diag * r;
      sv.mu[o] = m;
      sv.res[o] = r;
      sv.zr[o] = z;
      gn[c / 3] += r * z;
      en[c / 3] += -0.5 * (m * (r + sv.ef[o]));
    }
  }
  // clear this block's share of the grid (it was consumed by the gather)
  {
    const size_t total = (size_t)2 * sv.nfft[0] * sv.nfft[1] * sv.nfft[2];
    const size_t chunk = (total + sv.nblk - 1) / sv.nblk;
    const size_t b0 = (size_t)w.blk * chunk, b1 = min(total, b0 + chunk);
    real* g = static_cast<real*>(sv.grid);
    for (size_t q = b0 + threadIdx.x; q < b1; q += 64) g[q] = real(0);
  }
  double* part = sv.partial + (size_t)2 * sv.nblk;  // slot 1 (ggnew), slot 2 (ene)
  const double s0 = wave_sum(gn[0]), s1 = wave_sum(gn[1]);
  const double e0 = wave_sum(en[0]), e1 = wave_sum(en[1]);
  if (threadIdx.x == 0) {
    part[w.blk * 2] = s0;
    part[w.blk * 2 + 1] = s1;
    part[2 * sv.nblk + w.blk * 2] = e0;
    part[2 * sv.nblk + w.blk * 2 + 1] = e1;
  }
  if (last_block(&sv.st->counter[1], sv.nblk)) {
    const double a0 = sum_partials(part, sv.nblk, 0), a1 = sum_partials(part, sv.nblk, 1);
    const double b0 = sum_partials(part + 2 * sv.nblk, sv.nblk, 0);
    const double b1 = sum_partials(part + 2 * sv.nblk, sv.nblk, 1);
    if (threadIdx.x == 0) {
      sv.st->ggnew[0] = a0;
      sv.st->ggnew[1] = a1;
      sv.st->ene[0] = b0;
      sv.st->ene[1] = b1;
    }
  }
}

// pcg_newDirection + convergence (newinduce_pme2.f:543-568):
// pp = zr + ggnew/ggold pp; last block: ggold = ggnew, residual norm
// sqrt(ggnew/(3 npolar)), stop when max over rhs < poleps.
__global__ void __launch_bounds__(64) k_cg_update2(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  SolveState* st = sv.st;
  if (st->conv) return;
  const int i = w.blk * kBlock + threadIdx.x;
  const size_t np = sv.npad;
  const double beta[2] = {st->ggnew[0] / st->ggold[0], st->ggnew[1] / st->ggold[1]};
  if (i < sv.n) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      sv.pp[o] = sv.zr[o] + beta[c / 3] * sv.pp[o];
    }
  }
  if (last_block(&st->counter[2], sv.nblk)) {
    if (threadIdx.x == 0) {
      const double g0 = st->ggnew[0], g1 = st->ggnew[1];
      st->ggold[0] = g0;
      st->ggold[1] = g1;
      st->niter += 1;
      const double n3 = 3.0 * double(sv.n);
      const double rn = fmax(sqrt(g0 / n3), sqrt(g1 / n3));
      st->resnrm = rn;
      if (!isfinite(g0) || !isfinite(g1)) st->conv = kNonFinite;
      else if (rn < st->poleps) st->conv = kConverged;
      else if (st->niter >= st->politer) st->conv = kMaxIter;
    }
  }
}

// E_pol = -1/2 f sum mu_d . E_p (epolar0c: real + recip + self)
__global__ void __launch_bounds__(64) k_energy(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  const SysView& sv = views[w.sys];
  const int i = w.blk * kBlock + threadIdx.x;
  const size_t np = sv.npad;
  double e = 0.0;
  if (i < sv.n) {
#pragma unroll
    for (int c = 0; c < 3; ++c) e += sv.mu[(size_t)c * np + i] * sv.ef[(size_t)(3 + c) * np + i];
  }
  const double s = wave_sum(e);
  double* part = sv.partial + (size_t)6 * sv.nblk;  // slot 3
  if (threadIdx.x == 0) {
    part[w.blk * 2] = s;
    part[w.blk * 2 + 1] = 0.0;
  }
  if (last_block(&sv.st->counter[3], sv.nblk)) {
    const double a = sum_partials(part, sv.nblk, 0);
    if (threadIdx.x == 0) sv.st->ep = -0.5 * kCoulomb * a;
  }
}

// permutation helpers between caller order and sorted order, [n][6] <-> [6][npad]
__global__ void k_scatter_in(double* dst, const double* src, const int* iperm, int n, int npad, int ncomp) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int s = iperm[a];
  for (int c = 0; c < ncomp; ++c) dst[(size_t)c * npad + s] = src[(size_t)a * ncomp + c];
}
__global__ void k_gather_out(double* dst, const double* src, const int* iperm, int n, int npad, int ncomp) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int s = iperm[a];
  for (int c = 0; c < ncomp; ++c) dst[(size_t)a * ncomp + c] = src[(size_t)c * npad + s];
}

}  // namespace thp
