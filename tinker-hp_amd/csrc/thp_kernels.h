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
//  * reductions write per-block FP64 partials; the next kernel re-sums them
//    in a fixed order (a kernel boundary orders the accesses), so the PCG
//    scalars live on the device, need no atomics or fences, and are
//    bit-identical from run to run.
#pragma once
#include "thp_internal.h"

namespace thp {

// ------------------------------------------------------------------ utils
// Arrays reached through a SysView are generic pointers; casting them to the
// global address space lets the compiler emit global_load/global_store
// instead of flat accesses.  Hot kernels copy the SysView fields they need
// into locals once, so loop bodies hold no reloads of the view.
#define GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const GLOBAL T* gp(const T* p) {
  return (const GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ GLOBAL T* gw(T* p) {
  return (GLOBAL T*)p;
}

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
  const double ax = fabs(xr);
  if (ax > cell2) {
    if (__builtin_expect(ax < 0.9999 * cell, 1)) {
      // floor(|xr|/cell) is exactly 0 here, so the reference's first step
      // leaves xr unchanged and only the second shift applies
      xr = xr - copysign(cell, xr);
    } else {
      const double cel = copysign(cell, xr);
      const double fl = floor(ax / cell);
      const double prod = cel * fl;
      xr = xr - prod;
      if (fabs(xr) > cell2) xr = xr - copysign(cell, xr);
    }
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

__device__ __forceinline__ bool gated_out(const SolveState* st, int gate) {
  if (gate == kGateNone) return false;
  if (gate == kGateIter) return st->conv != kRunning || st->in_setup;
  return !st->in_setup;
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

// fixed-order sum of partial[q][0..nblk)[r] by one 64-lane wave
__device__ __forceinline__ double sum_partials(const double* part, int nblk, int r) {
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) acc += part[b * 2 + r];
  return wave_sum(acc);
}

// fixed-order sum of partial[0..count)[r] by a whole workgroup (any
// multiple of 64 threads); every thread receives the sum
__device__ __forceinline__ double block_sum_partials(const double* part, int count, int r, double* scratch) {
  const GLOBAL double* P = gp(part);
  double acc = 0.0;
  for (int b = threadIdx.x; b < count; b += blockDim.x) acc += P[b * 2 + r];
  acc = wave_sum(acc);
  const int nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = acc;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < nw; ++q) s += scratch[q];
  return s;
}

// partial-sum buffers: gather kernels write gat_blocks pairs (gg, init),
// vector kernels write red_blocks pairs (rz ping-pong, ep)
__device__ __forceinline__ double* part_gg(const SysView& sv) { return sv.partial; }
__device__ __forceinline__ double* part_init(const SysView& sv) { return sv.partial + (size_t)2 * sv.gat_blocks; }
__device__ __forceinline__ double* part_rz(const SysView& sv, int parity) {
  return sv.partial + (size_t)4 * sv.gat_blocks + (size_t)2 * sv.red_blocks * parity;
}
__device__ __forceinline__ double* part_ep(const SysView& sv) {
  return sv.partial + (size_t)4 * sv.gat_blocks + (size_t)4 * sv.red_blocks;
}
inline size_t partial_doubles(int gat_blocks, int red_blocks) { return (size_t)4 * gat_blocks + (size_t)6 * red_blocks; }

// sum of two values over a 256-thread block, written as partial `slot`
__device__ __forceinline__ void block_partial2(double a, double b, double* part, int slot) {
  __shared__ double red[2][4];
  a = wave_sum(a);
  b = wave_sum(b);
  // a previous call's thread 0 may still be reading red: wait for it before
  // the wavefronts overwrite their slots
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    double s0 = 0.0, s1 = 0.0;
    for (int q = 0; q < nw; ++q) {
      s0 += red[0][q];
      s1 += red[1][q];
    }
    part[slot * 2] = s0;
    part[slot * 2 + 1] = s1;
  }
}

// ===================================================== neighbor lists
// Block bounding boxes in a frame local to the block's first atom (minimum
// image), stored as centre + half extent.  One wavefront per block.
__global__ void __launch_bounds__(64) k_bbox(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
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
  if (w.sys < 0) return;
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
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  const int n = sv.n, maxnb = sv.maxnb;
  const size_t np = sv.npad;
  const double bx = sv.box[0], by = sv.box[1], bz = sv.box[2], mbuf2 = sv.mbuf2;
  const GLOBAL double* X = gp(sv.x);
  const GLOBAL double* Y = gp(sv.y);
  const GLOBAL double* Z = gp(sv.z);
  GLOBAL int* NB = gw(sv.nb);
  const GLOBAL int* BL = gp(blist) + (size_t)w.blk * maxbl;
  const int lane = threadIdx.x;
  const int i = w.blk * kBlock + lane;
  const bool ok = i < n;
  __shared__ double sx[kBlock], sy[kBlock], sz[kBlock];
  const double xi = ok ? X[i] : 0.0, yi = ok ? Y[i] : 0.0, zi = ok ? Z[i] : 0.0;
  int cnt = 0;
  bool overflow = false;
  const int nbl = gp(bcount)[w.blk];
  for (int q = 0; q < nbl; ++q) {
    const int k0 = BL[q] * kBlock;
    const int kk = k0 + lane;
    __syncthreads();
    sx[lane] = kk < n ? X[kk] : 0.0;
    sy[lane] = kk < n ? Y[kk] : 0.0;
    sz[lane] = kk < n ? Z[kk] : 0.0;
    __syncthreads();
    if (!ok) continue;
    const int mmax = min(kBlock, n - k0);
    for (int m = 0; m < mmax; ++m) {
      const int k = k0 + m;
      const double xr = image_exact(xi - sx[m], bx);
      const double yr = image_exact(yi - sy[m], by);
      const double zr = image_exact(zi - sz[m], bz);
      if (r2_exact(xr, yr, zr) <= mbuf2 && k != i) {
        if (FILL) {
          if (cnt < maxnb) NB[(size_t)cnt * np + i] = k;
          else overflow = true;
        }
        ++cnt;
      }
    }
  }
  if (ok) gw(sv.nbcount)[i] = cnt;
  if (FILL && overflow) *gw(sv.err) = THP_ERR_LIST;
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

// Field at i due to the permanent multipoles of k (fid and fip of
// efld0_direct, newinduce_pme.f:1116-1264), added to acc[0..5].
template <typename real>
__device__ __forceinline__ void efld0_pair(const SysView& sv, int k, real xr, real yr, real zr, real pdi,
                                           real thi, real dsc, real psc, double* acc, real usc = real(1),
                                           real* ucoef = nullptr) {
  const size_t np = sv.npad;
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
  ewald_bn<real, 3>(r, r2, real(sv.aewald), rr1, bn);
  real l[3];
  thole_l<real, 1>(r, pdi, real(sv.pdamp[k]), thi, real(sv.thole[k]), l);
  if (ucoef) {  // the matvec's coefficients of this pair (u-scale)
    ucoef[0] = bn[0] - (real(1) - usc * l[0]) * rr3;
    ucoef[1] = bn[1] - (real(1) - usc * l[1]) * rr5;
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

// Scaled (excluded) pairs of atom i in the permanent field, evaluated in
// FP64 in every precision mode: their bn - (1 - s*lambda)/r^n terms cancel
// strongly at bonded distances.  Mirrors the reference's separate
// correction pass over scaled pairs (GPU/source/tmatxb_pmecu.f:568-706).
// Also writes each scaled pair's matvec coefficients (u-scale) to scoef,
// zero beyond the cutoff; the matvec applies them in FP64.
__device__ __forceinline__ void scaled_pairs(const SysView& sv, int i, double xi, double yi, double zi, double* acc) {
  const int e0 = sv.excl_off[i], e1 = sv.excl_off[i + 1];
  const double pdi = sv.pdamp[i], thi = sv.thole[i];
  for (int e = e0; e < e1; ++e) {
    const int k = sv.excl_k[e];
    const double xr = image_exact(sv.x[k] - xi, sv.box[0]);
    const double yr = image_exact(sv.y[k] - yi, sv.box[1]);
    const double zr = image_exact(sv.z[k] - zi, sv.box[2]);
    if (!(r2_exact(xr, yr, zr) <= sv.cut2)) {
      sv.scoef[2 * e] = sv.scoef[2 * e + 1] = 0.0;
      continue;
    }
    const double* sc = sv.scale_tab + sv.excl_code[e] * 3;
    efld0_pair<double>(sv, k, xr, yr, zr, pdi, thi, sc[0], sc[1], acc, sc[2], sv.scoef + 2 * e);
  }
}

// Scaled pairs of every atom of the launch, added to the permanent field
// of the unscaled pass that precedes it.  A kernel of its own keeps the
// FP64 damping code out of the register budget of the pair kernel.
__global__ void __launch_bounds__(64) k_efld0_scaled(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  if (sv.excl_off[i] == sv.excl_off[i + 1]) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  scaled_pairs(sv, i, sv.x[i], sv.y[i], sv.z[i], acc);
  const size_t np = sv.npad;
#pragma unroll
  for (int c = 0; c < 6; ++c) sv.ef[(size_t)c * np + i] += acc[c];
}

// ===================================================== block-union lists
// For each 64-atom block, the union U of all atoms within cut + skin of at
// least one atom of the block (unscaled pairs), in increasing order, and
// for each atom its partners as 16-bit indices into U.  The pair kernels
// apply the exact cutoff test r^2 <= cut^2 (newinduce_pme.f:1160) to every
// entry, so the lists stay valid while no atom moves more than skin/2.
// Built from the block candidate list by 4 wavefronts: each tests the
// block's 64 atoms against a quarter of the candidate blocks (staged in
// LDS) and stores its lanes' 64-bit hit masks in LDS; wavefront 0 then walks
// the candidates in order, ORs the masks across lanes and assigns local
// indices by popcount.  Overflow of the preallocated capacities is
// reported through the error flag (the host then grows them and rebuilds).
__device__ __forceinline__ unsigned long long wave_or64(unsigned long long m) {
  unsigned lo = (unsigned)m, hi = (unsigned)(m >> 32);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo |= __shfl_xor(lo, off, 64);
    hi |= __shfl_xor(hi, off, 64);
  }
  return ((unsigned long long)hi << 32) | lo;
}

constexpr int kUnionWaves = 4;
constexpr int kUnionCand = 96;  // candidate blocks per LDS pass

__global__ void __launch_bounds__(64 * kUnionWaves) k_union(const SysView* views, const Work* work,
                                                            const int* bcount, const int* blist, int maxbl) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  const int n = sv.n, umax = sv.umax, lmax = sv.lmax;
  const size_t np = sv.npad;
  const double bx = sv.box[0], by = sv.box[1], bz = sv.box[2], lim2 = sv.ulim2;
  const GLOBAL double* X = gp(sv.x);
  const GLOBAL double* Y = gp(sv.y);
  const GLOBAL double* Z = gp(sv.z);
  const GLOBAL double* BB = gp(sv.bbox);
  const GLOBAL int* BL = gp(blist) + (size_t)w.blk * maxbl;
  GLOBAL int* UL = gw(sv.ulist) + (size_t)w.blk * umax;
  GLOBAL unsigned short* LL = gw(sv.llist);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = w.blk * kBlock + lane;
  const bool ok = i < n;
  __shared__ double sx[kUnionWaves][kBlock], sy[kUnionWaves][kBlock], sz[kUnionWaves][kBlock];
  __shared__ unsigned long long smask[kUnionCand][kBlock];
  __shared__ int s_k0[kUnionCand];
  const double xi = ok ? X[i] : 0.0, yi = ok ? Y[i] : 0.0, zi = ok ? Z[i] : 0.0;
  const int e0 = ok ? gp(sv.excl_off)[i] : 0, e1 = ok ? gp(sv.excl_off)[i + 1] : 0;
  int ex[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) ex[e] = e0 + e < e1 ? gp(sv.excl_k)[e0 + e] : -1;
  // this block's current bounding box: candidates farther than cut + skin
  // are skipped
  const double cxI = BB[w.blk * 6], cyI = BB[w.blk * 6 + 1], czI = BB[w.blk * 6 + 2];
  const double hxI = BB[w.blk * 6 + 3], hyI = BB[w.blk * 6 + 4], hzI = BB[w.blk * 6 + 5];
  int ucount = 0, lcount = 0;
  bool overflow = false;
  const int nbl = gp(bcount)[w.blk];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int p0 = 0; p0 < nbl; p0 += kUnionCand) {
    const int pn = min(kUnionCand, nbl - p0);
    // phase 1: masks of candidates p0+wv, p0+wv+4, ...
    for (int q = wv; q < pn; q += kUnionWaves) {
      const int J = BL[p0 + q];
      const int k0 = J * kBlock;
      unsigned long long mask = 0ull;
      double gx = BB[J * 6] - cxI, gy = BB[J * 6 + 1] - cyI, gz = BB[J * 6 + 2] - czI;
      gx -= bx * rint(gx / bx);
      gy -= by * rint(gy / by);
      gz -= bz * rint(gz / bz);
      gx = fmax(0.0, fabs(gx) - hxI - BB[J * 6 + 3]);
      gy = fmax(0.0, fabs(gy) - hyI - BB[J * 6 + 4]);
      gz = fmax(0.0, fabs(gz) - hzI - BB[J * 6 + 5]);
      const bool near = gx * gx + gy * gy + gz * gz <= lim2 * (1.0 + 1e-9) + 1e-9;
      if (near) {
        const int kk = k0 + lane;
        sx[wv][lane] = kk < n ? X[kk] : 0.0;
        sy[wv][lane] = kk < n ? Y[kk] : 0.0;
        sz[wv][lane] = kk < n ? Z[kk] : 0.0;
        // a single wavefront writes and reads its own staging row
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (ok) {
          const int mmax = min(kBlock, n - k0);
          for (int m = 0; m < mmax; ++m) {
            const double xr = image_exact(sx[wv][m] - xi, bx);
            const double yr = image_exact(sy[wv][m] - yi, by);
            const double zr = image_exact(sz[wv][m] - zi, bz);
            if (r2_exact(xr, yr, zr) <= lim2) mask |= 1ull << m;
          }
          if (k0 <= i && i < k0 + kBlock) mask &= ~(1ull << (i - k0));
#pragma unroll
          for (int e = 0; e < 8; ++e)
            if (ex[e] >= k0 && ex[e] < k0 + kBlock) mask &= ~(1ull << (ex[e] - k0));
          for (int e = e0 + 8; e < e1; ++e) {
            const int kx = gp(sv.excl_k)[e];
            if (kx >= k0 && kx < k0 + kBlock) mask &= ~(1ull << (kx - k0));
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      smask[q][lane] = mask;
      if (lane == 0) s_k0[q] = k0;
    }
    __syncthreads();
    // phase 2: wavefront 0 assigns union slots and local indices in order
    if (wv == 0) {
      for (int q = 0; q < pn; ++q) {
        const unsigned long long mask = smask[q][lane];
        const unsigned long long U = wave_or64(mask);
        if (U == 0ull) continue;
        if ((U >> lane) & 1ull) {
          const int pos = ucount + __popcll(U & below);
          if (pos < umax) UL[pos] = s_k0[q] + lane;
          else overflow = true;
        }
        unsigned long long mm = mask;
        while (mm) {
          const int m = __ffsll((long long)mm) - 1;
          mm &= mm - 1ull;
          const int loc = ucount + __popcll(U & ((1ull << m) - 1ull));
          if (lcount < lmax) LL[(size_t)lcount * np + i] = (unsigned short)loc;
          else overflow = true;
          ++lcount;
        }
        ucount += __popcll(U);
      }
    }
    __syncthreads();
  }
  if (wv == 0) {
    if (ok) gw(sv.lcount)[i] = lcount;
    if (lane == 0) gw(sv.ucount)[w.blk] = ucount;
    if (overflow) *gw(sv.err) = THP_ERR_LIST;
  }
}

// four values of the pair type, one 16- or 32-byte LDS access
template <typename real>
struct __attribute__((aligned(4 * sizeof(real)))) Quad {
  real a, b, c, d;
};

// (dmp3, dmp5) of one list entry, as a two-wide vector (one 8- or 16-byte
// access)
template <typename real>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <typename real>
using Coef2 = typename Vec2<real>::type;
template <typename real>
__device__ __forceinline__ Coef2<real> coef2(real a, real b) {
  Coef2<real> v;
  v.x = a;
  v.y = b;
  return v;
}

// Permanent field over the block-union lists (efld0_direct,
// newinduce_pme.f:1116-1264) and the per-solve compacted lists of the
// matvec.  A workgroup of S wavefronts owns one 64-atom block; each lane's
// local list is split in S contiguous segments, one per wavefront.  A first
// pass counts each segment's entries within the cutoff (the exact FP64
// test of efld0_direct), so that every wavefront knows where its entries
// start in the lane's compacted list; the second pass evaluates the field
// of those entries and writes each one, in list order, with its matvec
// coefficients (dmp3, dmp5).  Union atoms are staged in LDS in chunks; the
// S partial fields are summed in LDS in a fixed order.
template <typename real>
struct Efld0Stage {
  // positions stay FP64 so that the cutoff test and the displacements are
  // exact before rounding to the pair type
  static constexpr int SMEM = 48 * 1024;
  static constexpr int CHUNK = SMEM / (24 + 12 * (int)sizeof(real));  // xyz, 10 moments, pdamp, thole
  static constexpr int CHUNK_POS = SMEM / 24;                           // xyz only (counting pass)
};

//
// TAB (FP32 systems that carry the Ewald table of the recomputed matvec
// coefficients): G1 = r^3 bn1 and G2 = r^5 bn2 come from the table's cubic
// pieces, G3 = r^7 bn3 = 5 G2 + (8/sqrt(pi)) x^5 exp(-x^2) from the
// recursion of dampewald (damping.f:17-56, x = aewald r), and the Thole
// term from exp2 of t = a (r / (pdamp_i pdamp_k))^3 with the reciprocal
// pdamp of each atom (damptholed, damping.f:142-215):
//   dmp3 = (G1 - e) / r^3
//   dmp5 = (G2 - 3 (1 + t) e) / r^5
//   dmp7 = (G3 - 15 (1 + t + 0.6 t^2) e) / r^7,  e = exp(-t) for t < 50.
// Without TAB the pair type's erfc and exp form bn1..bn3 as the reference
// does.  Each kind of kernel skips the systems of the other kind.
constexpr float kG3Coef = 4.51351666838205f;  // 8 / sqrt(pi)
constexpr float kLog2e = 1.44269504088896f;
template <typename real, int S, bool TAB>
__global__ void __launch_bounds__(64 * S) k_efld0(const SysView* views, const Work* work, int gate) {
  static_assert(!TAB || sizeof(real) == 4, "the table path is FP32");
  using ST = Efld0Stage<real>;
  constexpr int C = ST::CHUNK;
  constexpr int CP = ST::CHUNK_POS;
  constexpr int NB = 8;  // list entries fetched per batch
  static_assert(S * 6 * 64 * (int)sizeof(double) <= ST::SMEM, "the cross-wavefront sum fits the staging buffer");
  __shared__ __attribute__((aligned(16))) unsigned char smem[ST::SMEM];
  __shared__ int s_kept[S][64];
  __shared__ Quad<float> s_gt[TAB ? 2 * kGtabN : 1];
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  if (TAB != (sv.rc_sqa > 0.0)) return;
  if (TAB) {
    // visible after the barrier that ends the counting pass
    typedef float F4 __attribute__((ext_vector_type(4)));
    const GLOBAL F4* GT = reinterpret_cast<const GLOBAL F4*>(gp(sv.gtab));
    for (int t = threadIdx.x; t < 2 * kGtabN; t += 64 * S) {
      const F4 v = GT[t];
      s_gt[t] = Quad<float>{v.x, v.y, v.z, v.w};
    }
  }
  const int n = sv.n;
  const size_t np = sv.npad;
  const double bx = sv.box[0], by = sv.box[1], bz = sv.box[2], cut2 = sv.cut2;
  const real aew = real(sv.aewald);
  const GLOBAL double* X = gp(sv.x);
  const GLOBAL double* Y = gp(sv.y);
  const GLOBAL double* Z = gp(sv.z);
  const GLOBAL double* PD = gp(sv.pdamp);
  const GLOBAL double* TH = gp(sv.thole);
  const GLOBAL double* RP = gp(sv.rpole);
  const GLOBAL unsigned short* LL = gp(sv.llist);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = w.blk * kBlock + lane;
  const bool ok = i < n;
  const int ucount = gp(sv.ucount)[w.blk];
  const GLOBAL int* UL = gp(sv.ulist) + (size_t)w.blk * sv.umax;
  const double xi = ok ? X[i] : 0.0, yi = ok ? Y[i] : 0.0, zi = ok ? Z[i] : 0.0;
  // TAB stages 1 / pdamp (1e15 for an atom without polarizability: its
  // pairs then reach t = inf and stay undamped, as damp = 0 leaves them in
  // damptholed)
  const real pdi = ok ? (TAB ? (PD[i] > 0.0 ? real(1.0 / PD[i]) : real(1e15)) : real(PD[i])) : real(0);
  const real thi = ok ? real(TH[i]) : real(0);
  const float aws = float(sv.aewald) * kGtabScale;
  const int cnt = ok ? gp(sv.lcount)[i] : 0;
  const int s_begin = (cnt * wv) / S;
  const int s_end = (cnt * (wv + 1)) / S;

  // ---- pass 1: entries of this wavefront's segment within the cutoff
  {
    double* px = reinterpret_cast<double*>(smem);
    double* py = px + CP;
    double* pz = py + CP;
    int s = s_begin, kept = 0;
    for (int c0 = 0; c0 < ucount; c0 += CP) {
      const int cn = min(CP, ucount - c0);
      __syncthreads();
      for (int q = threadIdx.x; q < cn; q += 64 * S) {
        const int kk = UL[c0 + q];
        px[q] = X[kk];
        py[q] = Y[kk];
        pz[q] = Z[kk];
      }
      __syncthreads();
      const int c1 = c0 + cn;
      while (s < s_end) {
        int loc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) loc[b] = s + b < s_end ? int(LL[(size_t)(s + b) * np + i]) : 0x7fffffff;
        bool done = false;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (loc[b] >= c1) {  // end of this segment or of the staged chunk
            done = true;
            break;
          }
          ++s;
          const int q = loc[b] - c0;
          const double dx = image_exact(px[q] - xi, bx);
          const double dy = image_exact(py[q] - yi, by);
          const double dz = image_exact(pz[q] - zi, bz);
          kept += r2_exact(dx, dy, dz) <= cut2 ? 1 : 0;
        }
        if (done) break;
      }
    }
    s_kept[wv][lane] = kept;
  }
  __syncthreads();
  int t = 0;  // this segment's first position in the lane's compacted list
  for (int q = 0; q < wv; ++q) t += s_kept[q][lane];
  if (wv == 0 && ok) {
    int tot = 0;
    for (int q = 0; q < S; ++q) tot += s_kept[q][lane];
    gw(sv.tcount)[i] = tot;
  }

  // ---- pass 2: field of the kept entries, compacted list written in order
  double* s_pos = reinterpret_cast<double*>(smem);  // [3][C]
  real* s_m = reinterpret_cast<real*>(s_pos + 3 * C);  // [10][C]
  real* s_pd = s_m + 10 * C;
  real* s_th = s_pd + C;
  GLOBAL unsigned short* TL = gw(sv.tlist);
  GLOBAL Coef2<real>* TC = gw(static_cast<Coef2<real>*>(sv.tcoef));
  // systems whose matvec recomputes the coefficients keep indices only,
  // their last pack padded with the union's count: the matvec stages a
  // sentinel there (an atom far away with zero dipoles)
  const bool keep_coef = sv.rc_sqa == 0.0;
  const unsigned short pad_ix = keep_coef ? 0 : (unsigned short)ucount;
  int s = s_begin;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int c0 = 0; c0 < ucount; c0 += C) {
    const int cn = min(C, ucount - c0);
    __syncthreads();
    for (int q = threadIdx.x; q < cn; q += 64 * S) {
      const int kk = UL[c0 + q];
      s_pos[0 * C + q] = X[kk];
      s_pos[1 * C + q] = Y[kk];
      s_pos[2 * C + q] = Z[kk];
      s_pd[q] = TAB ? (PD[kk] > 0.0 ? real(1.0 / PD[kk]) : real(1e15)) : real(PD[kk]);
      s_th[q] = real(TH[kk]);
#pragma unroll
      for (int m = 0; m < 10; ++m) s_m[m * C + q] = real(RP[(size_t)m * np + kk]);
    }
    __syncthreads();
    const int c1 = c0 + cn;
    while (s < s_end) {
      int loc[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) loc[b] = s + b < s_end ? int(LL[(size_t)(s + b) * np + i]) : 0x7fffffff;
      bool done = false;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (loc[b] >= c1) {
          done = true;
          break;
        }
        ++s;
        const int q = loc[b] - c0;
        const double dx = image_exact(s_pos[0 * C + q] - xi, bx);
        const double dy = image_exact(s_pos[1 * C + q] - yi, by);
        const double dz = image_exact(s_pos[2 * C + q] - zi, bz);
        if (!(r2_exact(dx, dy, dz) <= cut2)) continue;  // list entry within the skin only
        const real xr = real(dx), yr = real(dy), zr = real(dz);
        const real r2 = xr * xr + yr * yr + zr * zr;
        real dmp3, dmp5, dmp7;
        if constexpr (TAB) {
          // r^2 of a pair within the cutoff is far from the denormal range
          const float rr1 = __builtin_amdgcn_rsqf(r2);
          const float r = r2 * rr1;
          const float rr2 = rr1 * rr1;
          const float rr3 = rr2 * rr1;
          const float x = r * aws;
          const int j = min(int(x), kGtabN - 1);
          const float fr = __builtin_amdgcn_fractf(x);
          const Quad<float> P = s_gt[2 * j];
          const Quad<float> Q = s_gt[2 * j + 1];
          const float G1 = ((P.d * fr + P.c) * fr + P.b) * fr + P.a;
          const float G2 = ((Q.d * fr + Q.c) * fr + Q.b) * fr + Q.a;
          const float xa = x * (1.0f / kGtabScale);
          const float xa2 = xa * xa;
          const float G3 = 5.0f * G2 + kG3Coef * (xa2 * xa2 * xa) * __builtin_amdgcn_exp2f(-kLog2e * xa2);
          float pg = fminf(thi, s_th[q]);
          if (pg == 0.0f) pg = fmaxf(thi, s_th[q]);
          const float u = r * pdi * s_pd[q];
          const float t = pg * (u * u * u);  // inf (or NaN for pg = 0) when an atom is unpolarizable
          const bool damped = pg > 0.0f && t < 50.0f;
          // the polynomial factors take t clamped, so that e = 0 keeps them finite
          const float tc = fminf(t, 50.0f);
          const float e = damped ? __builtin_amdgcn_exp2f(-kLog2e * tc) : 0.0f;
          dmp3 = rr3 * (G1 - e);
          dmp5 = (rr3 * rr2) * (G2 - 3.0f * (1.0f + tc) * e);
          dmp7 = (rr3 * rr2 * rr2) * (G3 - 15.0f * ((1.0f + tc) + 0.6f * (tc * tc)) * e);
        } else {
          const real rr1 = rsqrt_t(r2);
          const real r = r2 * rr1;
          const real rr2 = rr1 * rr1;
          const real rr3 = rr2 * rr1;
          const real rr5 = real(3) * rr2 * rr3;
          const real rr7 = real(5) * rr2 * rr5;
          real bn[3];
          ewald_bn<real, 3>(r, r2, aew, rr1, bn);
          real l[3];
          thole_l<real, 1>(r, pdi, s_pd[q], thi, s_th[q], l);
          // unscaled pairs: dmp_n = bn_n - (1 - lambda_n) rr_n, the same for d and p
          dmp3 = bn[0] - (real(1) - l[0]) * rr3;
          dmp5 = bn[1] - (real(1) - l[1]) * rr5;
          dmp7 = bn[2] - (real(1) - l[2]) * rr7;
        }
        const real ck = s_m[0 * C + q], dkx = s_m[1 * C + q], dky = s_m[2 * C + q], dkz = s_m[3 * C + q];
        const real qkxx = s_m[4 * C + q], qkxy = s_m[5 * C + q], qkxz = s_m[6 * C + q], qkyy = s_m[7 * C + q],
                   qkyz = s_m[8 * C + q], qkzz = s_m[9 * C + q];
        const real dkr = dkx * xr + dky * yr + dkz * zr;
        const real qkx = qkxx * xr + qkxy * yr + qkxz * zr;
        const real qky = qkxy * xr + qkyy * yr + qkyz * zr;
        const real qkz = qkxz * xr + qkyz * yr + qkzz * zr;
        const real qkr = qkx * xr + qky * yr + qkz * zr;
        const size_t at = ((size_t)(t / kPack) * np + i) * kPack + t % kPack;
        TL[at] = (unsigned short)loc[b];
        if (keep_coef) TC[at] = coef2<real>(dmp3, dmp5);
        ++t;
        const real ai = dmp3 * ck - dmp5 * dkr + dmp7 * qkr;
        const real d5 = real(2) * dmp5;
        acc[0] += double(-xr * ai - dmp3 * dkx + d5 * qkx);
        acc[1] += double(-yr * ai - dmp3 * dky + d5 * qky);
        acc[2] += double(-zr * ai - dmp3 * dkz + d5 * qkz);
      }
      if (done) break;
    }
  }
  // the last wavefront ends at the lane's total: pad its last pack
  if (wv == S - 1 && ok) {
    for (; t % kPack; ++t) {
      const size_t at = ((size_t)(t / kPack) * np + i) * kPack + t % kPack;
      TL[at] = pad_ix;
      if (keep_coef) TC[at] = coef2<real>(real(0), real(0));
    }
  }
  // unscaled pairs give identical d and p fields
  acc[3] = acc[0];
  acc[4] = acc[1];
  acc[5] = acc[2];
  // fixed-order sum over the S wavefronts (the staging buffer is free once
  // every wavefront has passed the barrier)
  double* red = reinterpret_cast<double*>(smem);
  if (S > 1) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; ++c) red[(wv * 6 + c) * 64 + lane] = acc[c];
    __syncthreads();
  }
  if (wv == 0 && ok) {
    GLOBAL double* OUT = gw(sv.ef);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = acc[c];
      for (int q = 1; q < S; ++q) v += red[(q * 6 + c) * 64 + lane];
      OUT[(size_t)c * np + i] = v;
    }
  }
}

// Real-space dipole field over the block-union lists with the per-solve
// coefficients of the permanent-field pass (tmatxb_pme,
// newinduce_pme.f:1548-1608: efi -= dlocal . u_k with dlocal = -dmp3 I +
// dmp5 r r^T).  Union atoms are staged in LDS as positions relative to the
// block centre plus their dipoles.  Each lane walks its atom's compacted
// list one lane pack at a time (kPack entries: a 16-byte index load and
// kPack coefficient pairs, contiguous per lane and coalesced across the
// wavefront) with the next packs already in flight; an entry costs one LDS
// gather and ~25 flops.  A pack's contributions are summed in the pair type
// and each pack sum is added to the FP64 field.  Blocks whose union could
// reach across half the box take the minimum image of each displacement.
// Unions larger than one LDS chunk are walked chunk by chunk, with the
// entries outside the staged chunk masked.

template <typename real>
struct ClStage {
  // per union atom: (x, y, z, ud_x), (ud_y, ud_z, up_x, up_y), up_z
  static constexpr int BYTES = 9 * (int)sizeof(real);
  static constexpr int CHUNK = (sizeof(real) == 4 ? 48 * 1024 : 64 * 1024) / BYTES;  // atoms per LDS chunk
};
// Recomputed coefficients: per union atom two quads and (up_z, Thole
// weight), 40 bytes, then the Ewald table
struct RcStage {
  static constexpr int TAB_BYTES = kGtabN * 8 * (int)sizeof(float);
  static constexpr int BYTES = 48 * 1024;
  static constexpr int CHUNK = ((BYTES - TAB_BYTES) / 40) & ~1;  // even: the table after it stays 16-byte aligned
};
struct __attribute__((aligned(8))) RcTail {
  float upz, w;
};
// Thole weight of an atom, w = sqrt(a log2 e) / pdamp^3, so that
// w_i w_k r^3 is a (r / (pdamp_i pdamp_k))^3 in base 2 (damptholed,
// damping.f:142-215).  An atom without polarizability is undamped: its
// weight 1e15 puts w_i w_k r^3 past the cut of 72 (and, with r below 60,
// below the FP32 range) for any partner
__device__ __forceinline__ float rc_weight(double pdamp, double sqa) {
  const float p = float(pdamp);
  return p > 0.0f ? float(sqa) / (p * p * p) : 1.0e15f;
}
constexpr float kTholeCut2 = 72.1347520444f;  // 50 log2 e: damptholed's damp < 50 in base 2
constexpr float k3Ln2 = 2.0794415416798f;     // 3 ln 2

// one lane pack of the compacted list (indices only when the coefficients
// are recomputed)
template <typename real, bool RC>
struct LanePack {
  unsigned int ix[kPack / 2];  // two 16-bit union positions per word, low half first
  Coef2<real> cf[RC ? 1 : kPack];
};

template <typename real, bool RC>
__device__ __forceinline__ void load_pack(const GLOBAL unsigned short* TL, const GLOBAL Coef2<real>* TC, size_t np,
                                          int i, int b, LanePack<real, RC>& p) {
  typedef unsigned int U4 __attribute__((ext_vector_type(4)));
  typedef real R4 __attribute__((ext_vector_type(4)));
  static_assert(kPack == 8, "one 16-byte index load per pack");
  const size_t at = (size_t)b * np + i;
  const U4 v = reinterpret_cast<const GLOBAL U4*>(TL)[at];
  p.ix[0] = v.x;
  p.ix[1] = v.y;
  p.ix[2] = v.z;
  p.ix[3] = v.w;
  if (!RC) {
    const GLOBAL R4* c = reinterpret_cast<const GLOBAL R4*>(TC + at * kPack);
#pragma unroll
    for (int k = 0; k < kPack / 2; ++k) {
      const R4 u = c[k];
      p.cf[2 * k] = coef2<real>(u.x, u.y);
      p.cf[2 * k + 1] = coef2<real>(u.z, u.w);
    }
  }
}

// geometry of the lane's atom and of the staged union
template <typename real>
struct ClGeom {
  const Quad<real>* s_a;
  const Quad<real>* s_b;
  const real* s_c;
  const RcTail* s_t;        // recomputed coefficients: (up_z, w) per union atom
  const Quad<float>* s_g;   // and the Ewald table
  real pix, piy, piz;  // the lane's atom relative to the block centre
  real lx, ly, lz, ilx, ily, ilz;
  float wi, aw;  // the lane's Thole weight; aewald * kGtabScale
};

// Entries of one pack.  With RC the coefficients of the unscaled pair
// (tmatxb_pme, newinduce_pme.f:1564-1574) are formed from r:
//   dmp3 = bn1 - (1 - lambda3) / r^3 = (G1 - e) / r^3
//   dmp5 = bn2 - 3 (1 - lambda5) / r^5 = (G2 - 3 (1 + t) e) / r^5
// with G1 = r^3 bn1, G2 = r^5 bn2 (dampewald, damping.f:17-56) from the
// cubic table in x = aewald r, and e = exp(-t), t = a u^3 (damptholed,
// damping.f:142-215; e = 0 from t = 50 on, as there).  Entries outside
// [c0, c0 + cn) (another chunk, or padding) add nothing.
template <typename real, bool MASK, bool WRAP, bool RC>
__device__ __forceinline__ void apply_pack(const LanePack<real, RC>& p, int c0, int cn, const ClGeom<real>& g,
                                           double* acc) {
  real f0 = real(0), f1 = real(0), f2 = real(0), f3 = real(0), f4 = real(0), f5 = real(0);
#pragma unroll
  for (int e = 0; e < kPack; ++e) {
    const unsigned int word = p.ix[e >> 1];
    int q = int((e & 1) ? (word >> 16) : (word & 0xffffu));
    bool in = true;
    if (MASK) {
      q -= c0;
      in = (unsigned int)q < (unsigned int)cn;
      q = in ? q : 0;
    }
    const Quad<real> A = g.s_a[q];
    const Quad<real> B = g.s_b[q];
    real dx = A.a - g.pix, dy = A.b - g.piy, dz = A.c - g.piz;
    if (WRAP) {
      dx -= g.lx * rint(dx * g.ilx);
      dy -= g.ly * rint(dy * g.ily);
      dz -= g.lz * rint(dz * g.ilz);
    }
    real upz, d3, d5;
    if constexpr (RC) {
      const RcTail T = g.s_t[q];
      upz = T.upz;
      const float r2 = dx * dx + dy * dy + dz * dz;
      // r^2 of a pair is far from the denormal range: the bare rsq
      const float rr1 = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * rr1;
      // x < kGtabN within the cutoff (the host checks aewald * cut); the
      // clamp keeps masked and sentinel entries inside the table
      const float x = r * g.aw;
      const int j = min(int(x), kGtabN - 1);
      const float fr = __builtin_amdgcn_fractf(x);  // x - j wherever j is not clamped
      // (G2 from the table rather than from G1 and one more exp2: measured
      // 3% faster over a 16-system launch)
      const Quad<float> P = g.s_g[2 * j];
      const Quad<float> Q = g.s_g[2 * j + 1];
      const float G1 = ((P.d * fr + P.c) * fr + P.b) * fr + P.a;
      const float G2 = ((Q.d * fr + Q.c) * fr + Q.b) * fr + Q.a;
      const float tq = g.wi * T.w * (r2 * r);
      const float ex = tq < kTholeCut2 ? __builtin_amdgcn_exp2f(-tq) : 0.0f;
      const float rr2 = rr1 * rr1;
      const float rr3 = rr2 * rr1;
      d3 = rr3 * (G1 - ex);
      d5 = (rr3 * rr2) * (G2 - ex * (3.0f + k3Ln2 * tq));
    } else {
      upz = g.s_c[q];
      d3 = p.cf[e].x;
      d5 = p.cf[e].y;
    }
    if (MASK) {
      d3 = in ? d3 : real(0);
      d5 = in ? d5 : real(0);
    }
    const real urd = d5 * (dx * A.d + dy * B.a + dz * B.b);
    const real urp = d5 * (dx * B.c + dy * B.d + dz * upz);
    f0 = (f0 + d3 * A.d) - urd * dx;
    f1 = (f1 + d3 * B.a) - urd * dy;
    f2 = (f2 + d3 * B.b) - urd * dz;
    f3 = (f3 + d3 * B.c) - urp * dx;
    f4 = (f4 + d3 * B.d) - urp * dy;
    f5 = (f5 + d3 * upz) - urp * dz;
  }
  acc[0] += double(f0);
  acc[1] += double(f1);
  acc[2] += double(f2);
  acc[3] += double(f3);
  acc[4] += double(f4);
  acc[5] += double(f5);
}

// whole union staged: packs [b0, b1) of the lane, NP per step, the next
// NP in flight (loads past the segment re-read its last pack).  With RC the
// padding entries reach the sentinel staged past the union (exact zeros),
// except in blocks that wrap, where the sentinel's image could land near
// the atom: there every entry is tested against the union's cn atoms.
template <typename real, int NP, bool WRAP, bool RC>
__device__ __forceinline__ void walk_packs(const GLOBAL unsigned short* TL, const GLOBAL Coef2<real>* TC, size_t np,
                                           int i, int b0, int b1, int cn, const ClGeom<real>& g, double* acc) {
  if (b0 >= b1) return;
  LanePack<real, RC> cur[NP], nxt[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) load_pack<real, RC>(TL, TC, np, i, min(b0 + k, b1 - 1), cur[k]);
  for (int b = b0; b < b1; b += NP) {
#pragma unroll
    for (int k = 0; k < NP; ++k) load_pack<real, RC>(TL, TC, np, i, min(b + NP + k, b1 - 1), nxt[k]);
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (k == 0 || b + k < b1) apply_pack<real, RC && WRAP, WRAP, RC>(cur[k], 0, cn, g, acc);
#pragma unroll
    for (int k = 0; k < NP; ++k) cur[k] = nxt[k];
  }
}

// one chunk [c0, c0 + cn) of a union larger than the LDS stage.  Lists are
// sorted by union position, so the walk stops at the first pack that starts
// beyond the chunk; the last pack walked may straddle into the next chunk
// and is walked again there (masked).
template <typename real, bool WRAP, bool RC>
__device__ __forceinline__ void walk_chunk(const GLOBAL unsigned short* TL, const GLOBAL Coef2<real>* TC, size_t np,
                                           int i, int b1, int c0, int cn, int& resume, const ClGeom<real>& g,
                                           double* acc) {
  int last = -1;
  for (int b = resume; b < b1; ++b) {
    LanePack<real, RC> p;
    load_pack<real, RC>(TL, TC, np, i, b, p);
    if (int(p.ix[0] & 0xffffu) >= c0 + cn) break;
    apply_pack<real, true, WRAP, RC>(p, c0, cn, g, acc);
    last = b;
  }
  if (last >= 0) resume = last;
}

// RC: the coefficients are recomputed per entry (systems with rc_sqa > 0);
// each instance skips the systems of the other kind.
// Waves per SIMD asked of RC: the LDS stage allows three workgroups per CU,
// so four waves per workgroup can reach three waves per SIMD (without the
// bound the register count allows two) and eight can reach four (two
// workgroups).
constexpr int rc_waves_per_simd(int S) { return S == 8 ? 4 : S >= 4 ? 3 : 1; }
template <typename real, int S, int NB, bool RC>
__global__ void __launch_bounds__(64 * S, RC ? rc_waves_per_simd(S) : 1) k_tmatxb_cl(const SysView* views, const Work* work, int vin, int vout,
                                                       int dodiag, int gate) {
  static_assert(!RC || sizeof(real) == 4, "recomputed coefficients are FP32");
  constexpr int C = RC ? RcStage::CHUNK : ClStage<real>::CHUNK;
  static_assert(NB % kPack == 0, "whole lane packs per step");
  constexpr int NP = NB / kPack;
  // the staging arrays and the final cross-wavefront sum share one buffer
  constexpr int STAGE_BYTES = RC ? RcStage::BYTES : C * (2 * (int)sizeof(Quad<real>) + (int)sizeof(real));
  constexpr int RED_BYTES = S * 6 * 64 * (int)sizeof(double);
  __shared__ __attribute__((aligned(16))) unsigned char smem[STAGE_BYTES > RED_BYTES ? STAGE_BYTES : RED_BYTES];
  Quad<real>* s_a = reinterpret_cast<Quad<real>*>(smem);
  Quad<real>* s_b = s_a + C;
  real* s_c = reinterpret_cast<real*>(s_b + C);
  RcTail* s_t = RC ? reinterpret_cast<RcTail*>(s_b + C) : nullptr;
  Quad<float>* s_g = RC ? reinterpret_cast<Quad<float>*>(smem + 40 * C) : nullptr;
  double* red = reinterpret_cast<double*>(smem);
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  if (RC != (sv.rc_sqa > 0.0)) return;
  const int n = sv.n;
  const size_t np = sv.npad;
  const double bx = sv.box[0], by = sv.box[1], bz = sv.box[2];
  const double ibx = 1.0 / bx, iby = 1.0 / by, ibz = 1.0 / bz;
  const GLOBAL double* X = gp(sv.x);
  const GLOBAL double* Y = gp(sv.y);
  const GLOBAL double* Z = gp(sv.z);
  const GLOBAL double* PD = gp(sv.pdamp);
  const double sqa = sv.rc_sqa;
  const double* in_g = vsel(sv, vin);
  const GLOBAL double* IN = gp(in_g);
  const GLOBAL unsigned short* TL = gp(sv.tlist);
  const GLOBAL Coef2<real>* TC = gp(static_cast<const Coef2<real>*>(sv.tcoef));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = w.blk * kBlock + lane;
  const bool ok = i < n;
  const int ucount = gp(sv.ucount)[w.blk];
  const GLOBAL int* UL = gp(sv.ulist) + (size_t)w.blk * sv.umax;
  const GLOBAL double* BB = gp(sv.bbox) + (size_t)w.blk * 6;
  const double cx = BB[0], cy = BB[1], cz = BB[2];
  if (RC) {
    // the Ewald table, visible after the first staging barrier
    typedef float F4 __attribute__((ext_vector_type(4)));
    const GLOBAL F4* GT = reinterpret_cast<const GLOBAL F4*>(gp(sv.gtab));
    for (int t = threadIdx.x; t < 2 * kGtabN; t += 64 * S) {
      const F4 v = GT[t];
      s_g[t] = Quad<float>{v.x, v.y, v.z, v.w};
    }
  }
  // union atoms lie within cut + skin of a block atom; the block's atoms
  // moved at most skin/2 since its box was taken.  Away from the wrap, the
  // image nearest the block centre is the pair's image, and rounding
  // d / L through the reciprocal cannot change it.
  const double lim = sv.cut + 2.0 * sv.skin;
  const bool wrap = !(BB[3] + lim < 0.5 * bx && BB[4] + lim < 0.5 * by && BB[5] + lim < 0.5 * bz);
  ClGeom<real> g;
  g.s_a = s_a;
  g.s_b = s_b;
  g.s_c = s_c;
  g.s_t = s_t;
  g.s_g = s_g;
  g.lx = real(bx);
  g.ly = real(by);
  g.lz = real(bz);
  g.ilx = real(ibx);
  g.ily = real(iby);
  g.ilz = real(ibz);
  g.wi = RC && ok ? rc_weight(PD[i], sqa) : 0.0f;
  g.aw = float(sv.aewald) * kGtabScale;
  const double xi = ok ? X[i] : 0.0, yi = ok ? Y[i] : 0.0, zi = ok ? Z[i] : 0.0;
  {
    double d = xi - cx, e = yi - cy, f = zi - cz;
    d -= bx * rint(d * ibx);
    e -= by * rint(e * iby);
    f -= bz * rint(f * ibz);
    g.pix = real(d);
    g.piy = real(e);
    g.piz = real(f);
  }
  const int cnt = ok ? gp(sv.tcount)[i] : 0;
  const int ex0 = ok ? gp(sv.excl_off)[i] : 0, ex1 = ok ? gp(sv.excl_off)[i + 1] : 0;
  // partner and coefficients of this wavefront's first scaled pair, loaded
  // ahead of the walk
  const bool ex_first = ex0 + wv < ex1;
  const int kx_first = ex_first ? gp(sv.excl_k)[ex0 + wv] : 0;
  const double d3_first = ex_first ? gp(sv.scoef)[2 * (ex0 + wv)] : 0.0;
  const double d5_first = ex_first ? gp(sv.scoef)[2 * (ex0 + wv) + 1] : 0.0;
  const int npk = (cnt + kPack - 1) / kPack;
  const int b0 = (npk * wv) / S, b1 = (npk * (wv + 1)) / S;
  int resume = b0;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int c0 = 0; c0 < ucount; c0 += C) {
    const int cn = min(C, ucount - c0);
    __syncthreads();
    // a thread stages its union atoms SB at a time with every load of a
    // batch in flight before the batch's LDS writes (the tail of the last
    // batch re-reads the chunk's last atom and writes nothing)
    constexpr int SB = 4;
    for (int q0 = threadIdx.x; q0 < cn; q0 += 64 * S * SB) {
      int kk[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) kk[u] = UL[c0 + min(q0 + u * 64 * S, cn - 1)];
      double px[SB], py[SB], pz[SB], mu[SB][6], pdk[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) {
        px[u] = X[kk[u]];
        py[u] = Y[kk[u]];
        pz[u] = Z[kk[u]];
#pragma unroll
        for (int c = 0; c < 6; ++c) mu[u][c] = IN[(size_t)c * np + kk[u]];
        pdk[u] = RC ? PD[kk[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < SB; ++u) {
        const int q = q0 + u * 64 * S;
        if (q < cn) {
          double d = px[u] - cx, e = py[u] - cy, f = pz[u] - cz;
          d -= bx * rint(d * ibx);
          e -= by * rint(e * iby);
          f -= bz * rint(f * ibz);
          s_a[q] = Quad<real>{real(d), real(e), real(f), real(mu[u][0])};
          s_b[q] = Quad<real>{real(mu[u][1]), real(mu[u][2]), real(mu[u][3]), real(mu[u][4])};
          if (RC)
            s_t[q] = RcTail{float(mu[u][5]), rc_weight(pdk[u], sqa)};
          else
            s_c[q] = real(mu[u][5]);
        }
      }
    }
    // the padding entries' sentinel: zero dipoles and weight, 1.7e3 A away
    // (every term it adds is an exact zero)
    if (RC && threadIdx.x == 0 && ucount < C) {
      s_a[cn] = Quad<real>{real(1.0e3), real(1.0e3), real(1.0e3), real(0)};
      s_b[cn] = Quad<real>{real(0), real(0), real(0), real(0)};
      s_t[cn] = RcTail{0.0f, 0.0f};
    }
    __syncthreads();
    if (RC ? ucount < C : ucount <= C) {
      if (wrap)
        walk_packs<real, NP, true, RC>(TL, TC, np, i, b0, b1, cn, g, acc);
      else
        walk_packs<real, NP, false, RC>(TL, TC, np, i, b0, b1, cn, g, acc);
    } else {
      if (wrap)
        walk_chunk<real, true, RC>(TL, TC, np, i, b1, c0, cn, resume, g, acc);
      else
        walk_chunk<real, false, RC>(TL, TC, np, i, b1, c0, cn, resume, g, acc);
    }
  }
  // scaled pairs with their per-solve FP64 coefficients (tmatxb u-scale),
  // dealt round the wavefronts (entries ex0 + wv, ex0 + wv + S, ...) so that
  // no wavefront walks a long chain of dependent loads.  Pairs beyond the
  // cutoff have zero coefficients and add exact zeros.
  if (ok) {
    for (int e = ex0 + wv; e < ex1; e += S) {
      const bool first = e == ex0 + wv;
      const double d3 = first ? d3_first : sv.scoef[2 * e], d5 = first ? d5_first : sv.scoef[2 * e + 1];
      const int kx = first ? kx_first : sv.excl_k[e];
      const double xr = image_exact(X[kx] - xi, bx);
      const double yr = image_exact(Y[kx] - yi, by);
      const double zr = image_exact(Z[kx] - zi, bz);
#pragma unroll
      for (int rhs = 0; rhs < 2; ++rhs) {
        const double ux = IN[(size_t)(rhs * 3 + 0) * np + kx], uy = IN[(size_t)(rhs * 3 + 1) * np + kx],
                     uz = IN[(size_t)(rhs * 3 + 2) * np + kx];
        const double ur = d5 * (xr * ux + yr * uy + zr * uz);
        acc[rhs * 3 + 0] += d3 * ux - ur * xr;
        acc[rhs * 3 + 1] += d3 * uy - ur * yr;
        acc[rhs * 3 + 2] += d3 * uz - ur * zr;
      }
    }
  }
  // fixed-order sum over the S wavefronts (the staging buffer is free once
  // every wavefront has passed the barrier)
  if (S > 1) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; ++c) red[(wv * 6 + c) * 64 + lane] = acc[c];
    __syncthreads();
  }
  if (wv == 0 && ok) {
    const double pol = gp(sv.pol)[i];
    const double ipol = pol == 0.0 ? 1.0 / kTinyPol : 1.0 / pol;
    GLOBAL double* OUT = gw(vsel(sv, vout));
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = acc[c];
      for (int q = 1; q < S; ++q) v += red[(q * 6 + c) * 64 + lane];
      if (dodiag) v += IN[(size_t)c * np + i] * ipol;
      OUT[(size_t)c * np + i] = v;
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
  if constexpr (LEVELS >= 4) {
    // third derivative: the order-2 values differenced three times
    real a[kOrder + 1];
#pragma unroll
    for (int m = 1; m <= kOrder - 3; ++m) a[m] = t[kOrder - 3][m];
    a[kOrder - 2] = a[kOrder - 3];
#pragma unroll
    for (int m = kOrder - 3; m >= 2; --m) a[m] = a[m - 1] - a[m];
    a[1] = -a[1];
    a[kOrder - 1] = a[kOrder - 2];
#pragma unroll
    for (int m = kOrder - 2; m >= 2; --m) a[m] = a[m - 1] - a[m];
    a[1] = -a[1];
    a[kOrder] = a[kOrder - 1];
#pragma unroll
    for (int m = kOrder - 1; m >= 2; --m) a[m] = a[m - 1] - a[m];
    a[1] = -a[1];
#pragma unroll
    for (int m = 1; m <= kOrder; ++m) th[3][m - 1] = a[m];
  }
  int g0 = ifr - 4;
  if (g0 < 0) g0 += nfft;
  return g0;
}


// ---------------------------------------------------------------------
// Bricks and bins.  The grid is cut into bricks of B[d] <= 8 points per
// axis (every brick at least 4 wide, so a 5-point B-spline support spans
// at most two consecutive bricks).  Once per solve each atom gets its first
// grid node g0 and its B-spline values theta (3 levels x 5 nodes x 3 axes)
// and is binned by the brick that holds g0; bins are sorted by atom index.
// Spreading and gathering then work brick by brick (below).
constexpr int kTheta = 48;  // per-atom theta row: [axis][level][node] = 3*3*5 = 45, padded

template <typename real>
__device__ __forceinline__ const real* theta_row(const SysView& sv, int i) {
  return static_cast<const real*>(sv.theta) + (size_t)i * kTheta;
}
__device__ __forceinline__ int th_idx(int axis, int level, int node) { return axis * 15 + level * 5 + node; }

// per-solve: g0, theta and home bin of every atom; bin counts
template <typename real>
__global__ void __launch_bounds__(64) k_pme_prep(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const double xyz[3] = {sv.x[i], sv.y[i], sv.z[i]};
  real* th = static_cast<real*>(sv.theta) + (size_t)i * kTheta;
  int g0[3], bin[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    real t[3][kOrder];
    g0[d] = bspline_axis<real, 3>(xyz[d], 1.0 / sv.box[d], sv.nfft[d], t);
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int m = 0; m < kOrder; ++m) th[th_idx(d, l, m)] = t[l][m];
    bin[d] = g0[d] / sv.brick[d];
  }
  const int b = (bin[2] * sv.nbrick[1] + bin[1]) * sv.nbrick[0] + bin[0];
  sv.g0[4 * i] = g0[0];
  sv.g0[4 * i + 1] = g0[1];
  sv.g0[4 * i + 2] = g0[2];
  sv.g0[4 * i + 3] = b;
  atomicAdd(&sv.bin_count[b], 1);
}

// exclusive scan of the bin counts (one workgroup per system)
__global__ void __launch_bounds__(1024) k_pme_scan(const SysView* views, const int* sys_ids, int gate) {
  const SysView& sv = views[sys_ids[blockIdx.x]];
  if (gated_out(sv.st, gate)) return;
  const int nb = sv.nbrick[0] * sv.nbrick[1] * sv.nbrick[2];
  __shared__ int carry;
  __shared__ int wsum[16];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int q = base + threadIdx.x;
    const int v = q < nb ? sv.bin_count[q] : 0;
    // inclusive scan within the wave
    int x = v;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int pre = carry;
    for (int u = 0; u < wv; ++u) pre += wsum[u];
    if (q < nb) {
      sv.bin_start[q] = pre + x - v;
      sv.bin_count[q] = 0;  // reused as the fill cursor
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int u = 0; u < 16; ++u) t += wsum[u];
      carry += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sv.bin_start[nb] = carry;
}

__global__ void __launch_bounds__(64) k_pme_fill(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const int b = sv.g0[4 * i + 3];
  const int pos = atomicAdd(&sv.bin_count[b], 1);
  sv.bin_list[sv.bin_start[b] + pos] = i;
}

// sort each bin by atom index (bitonic, in LDS) so that spreading order
// does not depend on the fill order
constexpr int kMaxBin = 2048;
__global__ void __launch_bounds__(256) k_pme_binsort(const SysView* views, const Work* bwork, int gate) {
  const Work w = bwork[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int b = w.blk;
  const int s0 = sv.bin_start[b], cnt = sv.bin_start[b + 1] - s0;
  if (cnt <= 1) return;
  if (cnt > kMaxBin) {
    if (threadIdx.x == 0) *sv.err = THP_ERR_LIST;
    return;
  }
  __shared__ int key[kMaxBin];
  int n2 = 1;
  while (n2 < cnt) n2 <<= 1;
  for (int q = threadIdx.x; q < n2; q += blockDim.x) key[q] = q < cnt ? sv.bin_list[s0 + q] : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int q = threadIdx.x; q < n2; q += blockDim.x) {
        const int partner = q ^ stride;
        if (partner > q) {
          const bool up = (q & size) == 0;
          const int a = key[q], c = key[partner];
          if ((a > c) == up) {
            key[q] = c;
            key[partner] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) sv.bin_list[s0 + q] = key[q];
}

// an int carried in a slot of the grid's real type (bit copy)
template <typename real>
__device__ __forceinline__ real int_as_real(int v) {
  if constexpr (sizeof(real) == 4) return __int_as_float(v);
  else return __longlong_as_double((long long)v);
}
template <typename real>
__device__ __forceinline__ int real_as_int(real v) {
  if constexpr (sizeof(real) == 4) return __float_as_int(v);
  else return (int)__double_as_longlong(v);
}

// ---------------------------------------------------------------------
// Tile spreading.  One workgroup per brick takes the atoms whose first
// spline node lies in the brick (its "home" bin, sorted by atom index); the
// 5x5x5 support of such an atom lies inside the brick's tile of
// (B+4)^3 points.  Each wavefront owns a private LDS copy of the tile and
// adds whole stencils into it, atom after atom, one stencil point per lane
// (two passes of 64 lanes for 125 points): program order within a
// wavefront makes the read-add-write sequences of consecutive atoms exact,
// with no atomics.  The wavefront tiles are summed in a fixed order and
// written to the brick's slot of the tile buffer; k_spread_merge then adds,
// for every grid point, its own brick's tile and the overhang of the
// previous brick along each axis (at most 8 tiles, fixed order).  Bricks
// are at least 4 points wide, so no tile reaches two bricks ahead.

// W wavefronts (private tiles) per workgroup: more tiles shorten one
// brick's atom loop, fewer tiles cost less LDS and less zeroing and
// summing per brick, which pays when a batch brings many bricks
// home atoms staged in LDS at a time (16 to 32, for five workgroups per CU,
// measured no faster)
constexpr int kSpreadChunk = 64;
template <typename real, int KIND, int W>
__global__ void __launch_bounds__(64 * W) k_spread_tile(const SysView* views, const Work* bwork, int vin, int frac_in,
                                                        int gate) {
  constexpr int NL = KIND == 0 ? 2 : 3;   // spline levels used
  constexpr int NF = KIND == 0 ? 6 : 10;  // moments per atom
  // staged per atom: the spline values of (axis, node) with their levels
  // side by side (TW wide), and the moments with the tile offset after
  // them (FW wide), so that a lane reads its weights in a few vector reads
  constexpr int TW = KIND == 0 ? 2 : 4;
  constexpr int FW = KIND == 0 ? 8 : 12;
  typedef real TV __attribute__((ext_vector_type(TW)));
  typedef real R4 __attribute__((ext_vector_type(4)));
  // a wavefront's tile, plus a slot the idle lanes of the second pass write
  __shared__ __attribute__((aligned(16))) real s_tile[W][kTile3 * 2 + 4];
  __shared__ __attribute__((aligned(32))) real s_th[kSpreadChunk][3][kOrder][TW];
  __shared__ __attribute__((aligned(32))) real s_f[kSpreadChunk][FW];
  const Work w = bwork[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int nb0 = sv.nbrick[0], nb1 = sv.nbrick[1];
  const int tb = w.blk;
  const int t0 = tb % nb0, t1 = (tb / nb0) % nb1, t2 = tb / (nb0 * nb1);
  const int sx = t0 * sv.brick[0], sy = t1 * sv.brick[1], sz = t2 * sv.brick[2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t np = sv.npad;
  const GLOBAL double* IN = gp(vsel(sv, vin));
  const GLOBAL double* RP = gp(sv.rpole);
  const GLOBAL real* THETA = gp(static_cast<const real*>(sv.theta));
  const GLOBAL int* G0 = gp(sv.g0);
  const int a0 = gp(sv.bin_start)[tb], a1 = gp(sv.bin_start)[tb + 1];
  const GLOBAL int* BLIST = gp(sv.bin_list);
  const double af0 = frac_in ? 1.0 : sv.afrac[0], af1 = frac_in ? 1.0 : sv.afrac[1], af2 = frac_in ? 1.0 : sv.afrac[2];
  real* my = s_tile[wv];
  typedef real RV __attribute__((ext_vector_type(16 / sizeof(real))));  // 16-byte vector
  constexpr int NV = kTile3 * 2 * (int)sizeof(real) / 16;
  for (int p = lane; p < NV; p += 64) reinterpret_cast<RV*>(my)[p] = RV(real(0));
  // stencil point of this lane in each pass (lanes past the 125th point in
  // the second pass take the last point's weights and write the spare slot)
  int pmx[2], pmy[2], pmz[2], poff[2];
#pragma unroll
  for (int ps = 0; ps < 2; ++ps) {
    const int l = min(lane + 64 * ps, kOrder * kOrder * kOrder - 1);
    pmx[ps] = l % kOrder;
    pmy[ps] = (l / kOrder) % kOrder;
    pmz[ps] = l / (kOrder * kOrder);
    poff[ps] = (pmz[ps] * kTile + pmy[ps]) * kTile + pmx[ps];
  }
  for (int c0 = a0; c0 < a1; c0 += kSpreadChunk) {
    const int cnt = min(kSpreadChunk, a1 - c0);
    __syncthreads();
    for (int q = threadIdx.x; q < cnt; q += 64 * W) {
      const int i = BLIST[c0 + q];
      const int lx = G0[4 * i] - sx, ly = G0[4 * i + 1] - sy, lz = G0[4 * i + 2] - sz;
      s_f[q][NF] = int_as_real<real>((lz * kTile + ly) * kTile + lx);
      const GLOBAL real* th = THETA + (size_t)i * kTheta;
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
          for (int m = 0; m < kOrder; ++m) s_th[q][d][m][l] = th[th_idx(d, l, m)];
      if (KIND == 0) {
        s_f[q][0] = real(af0 * IN[0 * np + i]);
        s_f[q][1] = real(af1 * IN[1 * np + i]);
        s_f[q][2] = real(af2 * IN[2 * np + i]);
        s_f[q][3] = real(af0 * IN[3 * np + i]);
        s_f[q][4] = real(af1 * IN[4 * np + i]);
        s_f[q][5] = real(af2 * IN[5 * np + i]);
      } else {
        double cmp[10];
        cmp[0] = RP[0 * np + i];
        cmp[1] = RP[1 * np + i];
        cmp[2] = RP[2 * np + i];
        cmp[3] = RP[3 * np + i];
        cmp[4] = RP[4 * np + i];
        cmp[5] = RP[7 * np + i];
        cmp[6] = RP[9 * np + i];
        cmp[7] = 2.0 * RP[5 * np + i];
        cmp[8] = 2.0 * RP[6 * np + i];
        cmp[9] = 2.0 * RP[8 * np + i];
        s_f[q][0] = real(sv.ctf[0][0] * cmp[0]);
#pragma unroll
        for (int j = 1; j < 4; ++j) {
          double s = 0.0;
#pragma unroll
          for (int kk = 1; kk < 4; ++kk) s += sv.ctf[j][kk] * cmp[kk];
          s_f[q][j] = real(s);
        }
#pragma unroll
        for (int j = 4; j < 10; ++j) {
          double s = 0.0;
#pragma unroll
          for (int kk = 4; kk < 10; ++kk) s += sv.ctf[j][kk] * cmp[kk];
          s_f[q][j] = real(s);
        }
      }
    }
    __syncthreads();
    // The two passes of an atom touch distinct points: both are read before
    // either is written.  The next atom's weights are read from LDS while
    // this atom's tile reads are in flight, so an atom costs one LDS round
    // trip.  Consecutive atoms of a wavefront may share points and follow
    // each other in program order.
    const bool live1 = lane < kOrder * kOrder * kOrder - 64;
    auto weights = [&](int q, real (&g)[2][2], int& t0, int& t1) {
      const R4* fv = reinterpret_cast<const R4*>(s_f[q]);
      real f[FW];
#pragma unroll
      for (int c = 0; c < FW / 4; ++c) {
        const R4 t = fv[c];
        f[4 * c] = t.x;
        f[4 * c + 1] = t.y;
        f[4 * c + 2] = t.z;
        f[4 * c + 3] = t.w;
      }
      const int base = real_as_int<real>(f[NF]);
#pragma unroll
      for (int ps = 0; ps < 2; ++ps) {
        const TV tx = *reinterpret_cast<const TV*>(s_th[q][0][pmx[ps]]);
        const TV ty = *reinterpret_cast<const TV*>(s_th[q][1][pmy[ps]]);
        const TV tz = *reinterpret_cast<const TV*>(s_th[q][2][pmz[ps]]);
        if constexpr (KIND == 0) {
          const real ta = tx.x, tb1 = tx.y;
          const real u0 = ty.x, u1 = ty.y;
          const real v0 = tz.x, v1 = tz.y;
          g[ps][0] = ta * (f[1] * u1 * v0 + f[2] * u0 * v1) + tb1 * (f[0] * u0 * v0);
          g[ps][1] = ta * (f[4] * u1 * v0 + f[5] * u0 * v1) + tb1 * (f[3] * u0 * v0);
        } else {
          const real t0v = tx.x, t1v = tx.y, t2v = tx.z;
          const real u0 = ty.x, u1 = ty.y, u2 = ty.z;
          const real v0 = tz.x, v1 = tz.y, v2 = tz.z;
          const real term0 =
              f[0] * u0 * v0 + f[2] * u1 * v0 + f[3] * u0 * v1 + f[5] * u2 * v0 + f[6] * u0 * v2 + f[9] * u1 * v1;
          const real term1 = f[1] * u0 * v0 + f[7] * u1 * v0 + f[8] * u0 * v1;
          const real term2 = f[4] * u0 * v0;
          g[ps][0] = term0 * t0v + term1 * t1v + term2 * t2v;
          g[ps][1] = real(0);
        }
      }
      t0 = 2 * (base + poff[0]);
      t1 = live1 ? 2 * (base + poff[1]) : 2 * kTile3;
    };
    if (wv < cnt) {
      real g[2][2];
      int t0, t1;
      weights(wv, g, t0, t1);
      for (int q = wv; q < cnt; q += W) {
        typedef real R2 __attribute__((ext_vector_type(2)));
        R2 a0 = *reinterpret_cast<const R2*>(my + t0);
        R2 a1 = *reinterpret_cast<const R2*>(my + t1);
        real gn[2][2];
        int n0, n1;
        weights(min(q + W, cnt - 1), gn, n0, n1);
        a0.x += g[0][0];
        a0.y += g[0][1];
        a1.x += g[1][0];
        a1.y += g[1][1];
        *reinterpret_cast<R2*>(my + t0) = a0;
        *reinterpret_cast<R2*>(my + t1) = a1;
#pragma unroll
        for (int ps = 0; ps < 2; ++ps) {
          g[ps][0] = gn[ps][0];
          g[ps][1] = gn[ps][1];
        }
        t0 = n0;
        t1 = n1;
      }
    }
  }
  __syncthreads();
  // fixed-order sum of the wavefront tiles into the brick's tile slot
  GLOBAL RV* OUT = reinterpret_cast<GLOBAL RV*>(gw(static_cast<real*>(sv.tiles)) + (size_t)tb * kTile3 * 2);
  for (int p = threadIdx.x; p < NV; p += 64 * W) {
    RV v = reinterpret_cast<const RV*>(s_tile[0])[p];
#pragma unroll
    for (int q = 1; q < W; ++q) v += reinterpret_cast<const RV*>(s_tile[q])[p];
    OUT[p] = v;
  }
}

// grid point = own brick's tile + overhang of the previous brick per axis
// (fixed order); KIND 1 clears the imaginary part
template <typename real, int KIND>
__global__ void __launch_bounds__(256) k_spread_merge(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int K0 = sv.nfft[0], K1 = sv.nfft[1], K2 = sv.nfft[2];
  const int B0 = sv.brick[0], B1 = sv.brick[1], B2 = sv.brick[2];
  const int nb0 = sv.nbrick[0], nb1 = sv.nbrick[1], nb2 = sv.nbrick[2];
  const size_t npts = (size_t)K0 * K1 * K2;
  const GLOBAL real* T = gp(static_cast<const real*>(sv.tiles));
  GLOBAL real* G = gw(static_cast<real*>(sv.grid));
  const size_t p0 = (size_t)w.blk * kConvChunk, p1 = min(npts, p0 + kConvChunk);
  for (size_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int x = int(p % K0), y = int((p / K0) % K1), z = int(p / ((size_t)K0 * K1));
    const int bx = min(x / B0, nb0 - 1), by = min(y / B1, nb1 - 1), bz = min(z / B2, nb2 - 1);
    const int ox = x - bx * B0, oy = y - by * B1, oz = z - bz * B2;
    real re = real(0), im = real(0);
#pragma unroll
    for (int dz = 0; dz < 2; ++dz) {
      if (dz && oz >= kOrder - 1) break;
      const int cz = dz ? (bz + nb2 - 1) % nb2 : bz;
      const int tz = dz ? oz + (cz == nb2 - 1 ? K2 - (nb2 - 1) * B2 : B2) : oz;
#pragma unroll
      for (int dy = 0; dy < 2; ++dy) {
        if (dy && oy >= kOrder - 1) break;
        const int cy = dy ? (by + nb1 - 1) % nb1 : by;
        const int ty = dy ? oy + (cy == nb1 - 1 ? K1 - (nb1 - 1) * B1 : B1) : oy;
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          if (dx && ox >= kOrder - 1) break;
          const int cx = dx ? (bx + nb0 - 1) % nb0 : bx;
          const int tx = dx ? ox + (cx == nb0 - 1 ? K0 - (nb0 - 1) * B0 : B0) : ox;
          const size_t tile = ((size_t)cz * nb1 + cy) * nb0 + cx;
          const size_t t = (tile * kTile3 + ((size_t)(tz * kTile + ty) * kTile + tx)) * 2;
          re += T[t];
          if (KIND == 0) im += T[t + 1];
        }
      }
    }
    G[2 * p] = re;
    G[2 * p + 1] = im;
  }
}

// ===================================================== 3-D FFT
// Batched multi-size complex FFT of the PME grids, one axis per launch: the
// unnormalised forward then backward transforms of fft2d_frontmpi/backmpi
// (fft_mpi.f90:65-160), whose normalisation sits in qfac.  A workgroup
// transforms a group of lines of one system's grid along one axis in LDS
// with self-sorting (Stockham) radix-8/4/2/3/5 stages; systems of a launch
// may have different grid sizes.  The z pass fuses the forward transform,
// the scalar-sum product with qfac (newinduce_pme.f:1790-1861) and the
// backward transform, so a convolution is five passes over the grid.
// Lines of a group are consecutive along the fastest other index (y for the
// x pass, x for the y and z passes), so every global access of a group is
// a run of consecutive complex values.

template <typename real>
__device__ __forceinline__ Coef2<real> cmul(Coef2<real> a, Coef2<real> b) {
  return coef2<real>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// s * i * a
template <typename real>
__device__ __forceinline__ Coef2<real> cmul_si(Coef2<real> a, real s) {
  return coef2<real>(-s * a.y, s * a.x);
}

// X_k = sum_q x_q w^(qk), w = exp(s 2 pi i / R)
template <typename real, int R>
__device__ __forceinline__ void dft_small(Coef2<real>* x, real s) {
  if constexpr (R == 2) {
    const Coef2<real> a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = a - b;
  } else if constexpr (R == 4) {
    const Coef2<real> t0 = x[0] + x[2], t1 = x[0] - x[2], t2 = x[1] + x[3], t3 = x[1] - x[3];
    const Coef2<real> it3 = cmul_si<real>(t3, s);
    x[0] = t0 + t2;
    x[2] = t0 - t2;
    x[1] = t1 + it3;
    x[3] = t1 - it3;
  } else if constexpr (R == 3) {
    const real h = real(0.86602540378443864676);  // sin(2 pi / 3)
    const Coef2<real> t1 = x[1] + x[2], t2 = x[1] - x[2];
    const Coef2<real> m = x[0] - real(0.5) * t1;
    const Coef2<real> it2 = cmul_si<real>(h * t2, s);
    x[0] = x[0] + t1;
    x[1] = m + it2;
    x[2] = m - it2;
  } else {
    static_assert(R == 5, "radix 2, 3, 4 or 5");
    const real c1 = real(0.30901699437494742410), c2 = real(-0.80901699437494742410);  // cos(2 pi/5), cos(4 pi/5)
    const real s1 = real(0.95105651629515357212), s2 = real(0.58778525229247312917);   // sin(2 pi/5), sin(4 pi/5)
    const Coef2<real> t1 = x[1] + x[4], t2 = x[2] + x[3], t3 = x[1] - x[4], t4 = x[2] - x[3];
    const Coef2<real> a1 = x[0] + c1 * t1 + c2 * t2, a2 = x[0] + c2 * t1 + c1 * t2;
    const Coef2<real> b1 = cmul_si<real>(s1 * t3 + s2 * t4, s), b2 = cmul_si<real>(s2 * t3 - s1 * t4, s);
    x[0] = x[0] + t1 + t2;
    x[1] = a1 + b1;
    x[4] = a1 - b1;
    x[2] = a2 + b2;
    x[3] = a2 - b2;
  }
}

// radix-8 DFT as two radix-4 DFTs: even outputs from x_q + x_(q+4), odd
// outputs from (x_q - x_(q+4)) w^q
template <typename real>
__device__ __forceinline__ void dft8(Coef2<real>* x, real s) {
  const real r = real(0.70710678118654752440);  // 1 / sqrt(2)
  Coef2<real> e[4], o[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    e[q] = x[q] + x[q + 4];
    o[q] = x[q] - x[q + 4];
  }
  // w = exp(s i pi / 4): w^1 = (1 + s i) r, w^2 = s i, w^3 = (-1 + s i) r
  o[1] = coef2<real>(r * (o[1].x - s * o[1].y), r * (o[1].y + s * o[1].x));
  o[2] = cmul_si<real>(o[2], s);
  o[3] = coef2<real>(r * (-o[3].x - s * o[3].y), r * (-o[3].y + s * o[3].x));
  dft_small<real, 4>(e, s);
  dft_small<real, 4>(o, s);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    x[2 * m] = e[m];
    x[2 * m + 1] = o[m];
  }
}

// one Stockham stage of radix R over the group's lines (line a starts at
// a * kp in LDS, its points ps apart): butterfly i of a line reads
// src[i + q K/R], applies the twiddles w^(q k K/(pR)) with k = i mod p, and
// writes dst[(i - k) R + k + q p]
template <typename real, int R>
__device__ __forceinline__ void fft_stage(const Coef2<real>* src, Coef2<real>* dst, const Coef2<real>* tw, int K,
                                          int kp, int ps, int p, int nl, int tid, int tpl_shift, real s) {
  const int T = K / R;
  const int step = K / (p * R);
  const int a = tid >> tpl_shift;
  if (a >= nl) return;
  const Coef2<real>* in = src + a * kp;
  Coef2<real>* out = dst + a * kp;
  for (int i = tid & ((1 << tpl_shift) - 1); i < T; i += 1 << tpl_shift) {
    const int k = i % p;
    Coef2<real> x[R];
    x[0] = in[i * ps];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      Coef2<real> w = tw[q * k * step];
      w.y *= -s;  // tw holds exp(-2 pi i m / K): conjugated for the backward sign
      x[q] = cmul<real>(in[(i + q * T) * ps], coef2<real>(w.x, w.y));
    }
    if constexpr (R == 8) dft8<real>(x, s);
    else dft_small<real, R>(x, s);
    const int j = (i - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) out[(j + q * p) * ps] = x[q];
  }
}

// whole transform of the group's lines; returns the buffer holding it
template <typename real>
__device__ int fft_lines_lds(Coef2<real>* b0, Coef2<real>* b1, int cur, const Coef2<real>* tw, int K, int kp, int nl,
                             int tid, int tpl_shift, real s, int ps = 1) {
  int rem = K, p = 1;
  while (rem > 1) {
    const int R = rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : 5;
    const Coef2<real>* src = cur ? b1 : b0;
    Coef2<real>* dst = cur ? b0 : b1;
    switch (R) {
      case 8: fft_stage<real, 8>(src, dst, tw, K, kp, ps, p, nl, tid, tpl_shift, s); break;
      case 4: fft_stage<real, 4>(src, dst, tw, K, kp, ps, p, nl, tid, tpl_shift, s); break;
      case 2: fft_stage<real, 2>(src, dst, tw, K, kp, ps, p, nl, tid, tpl_shift, s); break;
      case 3: fft_stage<real, 3>(src, dst, tw, K, kp, ps, p, nl, tid, tpl_shift, s); break;
      default: fft_stage<real, 5>(src, dst, tw, K, kp, ps, p, nl, tid, tpl_shift, s); break;
    }
    __syncthreads();
    cur ^= 1;
    p *= R;
    rem /= R;
  }
  return cur;
}

// MODE 0 forward, 1 backward, 2 forward * qfac * backward (z pass only)
// MODE 0 forward, 1 backward, 2 the z pass of the convolution (forward,
// times qfac, backward), 3 the same keeping the spectrum (the forward
// transform of the whole grid, before qfac) in `spec`, grid layout
template <typename real, int AXIS, int MODE>
__global__ void __launch_bounds__(256) k_fft_axis(const SysView* views, const Work* work, int gate, void* spec) {
  constexpr bool CONV = MODE == 2 || MODE == 3;
  static_assert(!CONV || AXIS == 2, "the convolution is fused into the z pass");
  __shared__ Coef2<real> s_buf[2][kFftBuf];
  __shared__ Coef2<real> s_tw[kFftMaxLen];
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int K0 = sv.nfft[0], K1 = sv.nfft[1], K2 = sv.nfft[2];
  const int K = sv.nfft[AXIS];
  const int kp = K + 1;  // padded line stride in LDS
  const int fl = fft_group_lines(K, sv.fft_cap);
  const int lg = 31 - __clz(fl);
  const int inner_n = AXIS == 0 ? K1 : K0;
  const int outer_n = AXIS == 2 ? K1 : K2;
  const int ngi = (inner_n + fl - 1) / fl;
  if (w.blk >= outer_n * ngi) return;
  const int outer = w.blk / ngi;
  const int in0 = (w.blk - outer * ngi) * fl;
  const int nl = min(fl, inner_n - in0);
  // complex element (line a, position t) of the group: base + t * ts + a * as
  size_t base, ts;
  if (AXIS == 0) {
    base = ((size_t)outer * K1 + in0) * K0;
    ts = 1;
  } else if (AXIS == 1) {
    base = (size_t)outer * K1 * K0 + in0;
    ts = (size_t)K0;
  } else {
    base = (size_t)outer * K0 + in0;
    ts = (size_t)K0 * K1;
  }
  GLOBAL Coef2<real>* G = gw(static_cast<Coef2<real>*>(sv.grid));
  const GLOBAL Coef2<real>* TW =
      gp(static_cast<const Coef2<real>*>(sv.twid)) + (AXIS == 0 ? 0 : AXIS == 1 ? K0 : K0 + K1);
  const int tid = threadIdx.x;
  // element u of this thread: LDS slot and grid index.  The x pass reads
  // its group as one contiguous run; the y and z passes read runs of fl
  // consecutive values, one per line.
  constexpr int PER = kFftBuf / 256;  // elements per thread, at most
  const int ra = tid & (fl - 1), rt = tid >> lg, rstep = 256 >> lg;
  auto elem = [&](int u, int& slot, size_t& g) {
    if (AXIS == 0) {
      const int f = tid + 256 * u;
      const int a = f / K, t = f - a * K;
      slot = a * kp + t;
      g = base + f;
      return f < nl * K;
    } else {
      const int t = rt + u * rstep;
      slot = ra * kp + t;
      g = base + t * ts + ra;
      return ra < nl && t < K;
    }
  };
  // every global load of the group (values, twiddles, qfac) is in flight
  // before the first LDS write; the loads are unconditional (elements past
  // the group re-read its first one) so that none waits for another
  Coef2<real> v[PER];
  real qv[CONV ? PER : 1];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    int slot;
    size_t g;
    v[u] = G[elem(u, slot, g) ? g : base];
  }
  if constexpr (CONV) {
    const GLOBAL real* Q = gp(static_cast<const real*>(sv.qfac));
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int slot;
      size_t g;
      qv[u] = Q[elem(u, slot, g) ? g : base];
    }
  }
  static_assert(kFftMaxLen <= 2 * 256, "two twiddles per thread");
  Coef2<real> tv[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) tv[j] = TW[min(tid + 256 * j, K - 1)];
#pragma unroll
  for (int j = 0; j < 2; ++j)
    if (tid + 256 * j < K) s_tw[tid + 256 * j] = tv[j];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    int slot;
    size_t g;
    if (elem(u, slot, g)) s_buf[0][slot] = v[u];
  }
  __syncthreads();
  const int tpl_shift = 8 - lg;  // threads per line = 256 / fl
  int cur = fft_lines_lds<real>(s_buf[0], s_buf[1], 0, s_tw, K, kp, nl, tid, tpl_shift, MODE == 1 ? real(1) : real(-1));
  if (CONV) {
    if constexpr (MODE == 3) {
      GLOBAL Coef2<real>* SP = gw(static_cast<Coef2<real>*>(spec));
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        int slot;
        size_t g;
        if (elem(u, slot, g)) SP[g] = s_buf[cur][slot];
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int slot;
      size_t g;
      if (elem(u, slot, g)) s_buf[cur][slot] *= qv[u];
    }
    __syncthreads();
    cur = fft_lines_lds<real>(s_buf[0], s_buf[1], cur, s_tw, K, kp, nl, tid, tpl_shift, real(1));
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    int slot;
    size_t g;
    if (elem(u, slot, g)) G[g] = s_buf[cur][slot];
  }
}

// The x and y passes of one z plane of one grid in LDS, for batches whose
// planes fit (FP32 grids up to about 75 x 75 points per plane): MODE 0 the
// forward x then y transforms, MODE 1 the backward y then x transforms.
// With the fused z pass a convolution is then three passes over the grid.
// The plane is row-major in LDS with rows padded to K0 + 1 values; the x
// lines are its rows, the y lines its columns (points K0 + 1 apart).
template <typename real, int MODE, int NT>
__global__ void __launch_bounds__(NT) k_fft_plane(const SysView* views, const Work* work, int gate) {
  extern __shared__ __attribute__((aligned(16))) unsigned char dsm[];
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int K0 = sv.nfft[0], K1 = sv.nfft[1], K2 = sv.nfft[2];
  if (w.blk >= K2) return;
  const int kp = K0 + 1, cap = K1 * kp;
  Coef2<real>* b0 = reinterpret_cast<Coef2<real>*>(dsm);
  Coef2<real>* b1 = b0 + cap;
  Coef2<real>* tw = b1 + cap;  // twiddles of the x then the y axis
  GLOBAL Coef2<real>* G = gw(static_cast<Coef2<real>*>(sv.grid)) + (size_t)w.blk * K0 * K1;
  const GLOBAL Coef2<real>* TW = gp(static_cast<const Coef2<real>*>(sv.twid));
  const int tid = threadIdx.x;
  const int npts = K0 * K1;
  // loads: a batch of unconditional loads in flight per thread (indices past
  // the plane re-read its first value), then the LDS writes
  constexpr int PB = 8;
  for (int f0 = tid; f0 < npts; f0 += NT * PB) {
    Coef2<real> v[PB];
#pragma unroll
    for (int u = 0; u < PB; ++u) {
      const int f = f0 + NT * u;
      v[u] = G[f < npts ? f : 0];
    }
#pragma unroll
    for (int u = 0; u < PB; ++u) {
      const int f = f0 + NT * u;
      if (f < npts) {
        const int y = f / K0;
        b0[y * kp + (f - y * K0)] = v[u];
      }
    }
  }
  for (int m = tid; m < K0 + K1; m += NT) tw[m] = TW[m];
  __syncthreads();
  auto lines = [&](int cur, int axis, real s) {
    const int K = axis == 0 ? K0 : K1, nl = axis == 0 ? K1 : K0;
    int fl = 1;
    while (fl < nl) fl <<= 1;
    const int tpl_shift = (31 - __clz(NT)) - (31 - __clz(fl));  // threads per line = NT / fl
    return axis == 0 ? fft_lines_lds<real>(b0, b1, cur, tw, K, kp, nl, tid, tpl_shift, s, 1)
                     : fft_lines_lds<real>(b0, b1, cur, tw + K0, K, 1, nl, tid, tpl_shift, s, kp);
  };
  int cur = 0;
  if (MODE == 0) {
    cur = lines(cur, 0, real(-1));
    cur = lines(cur, 1, real(-1));
  } else {
    cur = lines(cur, 1, real(1));
    cur = lines(cur, 0, real(1));
  }
  const Coef2<real>* res = cur ? b1 : b0;
  for (int f = tid; f < npts; f += NT) {
    const int y = f / K0;
    G[f] = res[y * kp + (f - y * K0)];
  }
}

// ----------------------------------------------------- register FFT, 64 points
// For grids of 64 x 64 x 64 points (FP32) the transforms run in registers.
// P lanes (1, 2 or 4, neighbours in the wave) hold one 64-point line, lane L
// the points P n + L (n = 0 .. 64/P - 1).  With n = P n1 + L and k = k1 +
// (64/P) k2 a line's DFT is a (64/P)-point DFT in each lane (a four-step
// R1 x R2 split with compile-time indices and twiddles), the twiddles
// exp(s 2 pi i L k1 / 64), and a P-point DFT across the lanes (radix-2
// stages over DPP quad permutes; for P = 4 lane 3 is turned by s i between
// the two); lane L then holds the frequencies k1 + (64/P) rev(L), rev
// swapping the two bits of L when P = 4.  The same steps in reverse order
// take the lanes from that arrangement back to the points P n + L.  A wave
// holds 64/P lines that lie side by side in memory (consecutive x), so its
// loads and stores are runs of 512/P bytes.

// calls f(std::integral_constant<int, i>) for i = 0 .. N-1
template <typename F, int... I>
__device__ __forceinline__ void sfor_seq(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_seq(f, std::make_integer_sequence<int, N>{});
}

// cos or sin of a in [0, 2 pi) at compile time (Taylor series after
// reduction to [-pi, pi], to double precision)
constexpr double ct_trig(double a, bool sine) {
  constexpr double pi = 3.14159265358979323846;
  if (a > pi) a -= 2.0 * pi;
  double term = sine ? a : 1.0, sum = term;
  for (int n = 1; n < 30; ++n) {
    const double d = sine ? (2.0 * n) * (2.0 * n + 1.0) : (2.0 * n - 1.0) * (2.0 * n);
    term *= -a * a / d;
    sum += term;
  }
  return sum;
}
// exp(S 2 pi i M / K)
template <int K, int M, int S>
__device__ __forceinline__ Coef2<float> ct_twiddle() {
  constexpr double a = 2.0 * 3.14159265358979323846 * double(M % K) / double(K);
  constexpr float c = float(ct_trig(a, false)), sn = float(S) * float(ct_trig(a, true));
  return coef2<float>(c, sn);
}

template <int R, int S>
__device__ __forceinline__ void dft_reg(Coef2<float>* a) {
  if constexpr (R == 8) dft8<float>(a, float(S));
  else dft_small<float, R>(a, float(S));
}

// in-place DFT of x[R1 R2] with exp(S 2 pi i n k / (R1 R2)), natural order:
// n = R2 n1 + n2, k = k1 + R1 k2
template <int R1, int R2, int S>
__device__ __forceinline__ void fft_reg(Coef2<float>* x) {
  constexpr int N = R1 * R2;
  Coef2<float> t[N];
  sfor<R2>([&](auto n2) {
    Coef2<float> a[R1];
    sfor<R1>([&](auto n1) { a[n1] = x[R2 * n1 + n2]; });
    dft_reg<R1, S>(a);
    sfor<R1>([&](auto k1) {
      constexpr int m = decltype(n2)::value * decltype(k1)::value;
      if constexpr (m == 0) t[R1 * n2 + k1] = a[k1];
      else t[R1 * n2 + k1] = cmul<float>(a[k1], ct_twiddle<N, m, S>());
    });
  });
  sfor<R1>([&](auto k1) {
    Coef2<float> a[R2];
    sfor<R2>([&](auto n2) { a[n2] = t[R1 * n2 + k1]; });
    dft_reg<R2, S>(a);
    sfor<R2>([&](auto k2) { x[k1 + R1 * k2] = a[k2]; });
  });
}

// value of lane L ^ 1 (kQuadXor1) or L ^ 2 (kQuadXor2) of each quad
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E;  // DPP quad_perm [1,0,3,2], [2,3,0,1]
template <int CTRL>
__device__ __forceinline__ float quad_perm(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
// radix-2 across lanes L and L ^ X: the lane with (L & X) == 0 keeps own +
// partner, the other partner - own
template <int M, int CTRL, int X>
__device__ __forceinline__ void lane_stage(Coef2<float>* v, int L) {
  const float sg = (L & X) ? -1.0f : 1.0f;
  sfor<M>([&](auto k) {
    const float px = quad_perm<CTRL>(v[k].x), py = quad_perm<CTRL>(v[k].y);
    v[k].x = fmaf(sg, v[k].x, px);
    v[k].y = fmaf(sg, v[k].y, py);
  });
}
// lane 3 of the quad times S i
template <int M, int S>
__device__ __forceinline__ void lane_rot3(Coef2<float>* v, int L) {
  const bool r = L == 3;
  sfor<M>([&](auto k) {
    const float nx = -float(S) * v[k].y, ny = float(S) * v[k].x;
    v[k].x = r ? nx : v[k].x;
    v[k].y = r ? ny : v[k].y;
  });
}
// v[k] *= exp(S 2 pi i L k / 64), L < P; the lane's twiddles are picked by
// its index bits (selects, no branches)
template <int P, int S>
__device__ __forceinline__ void lane_twiddle(Coef2<float>* v, int L) {
  constexpr int M = 64 / P;
  const bool b0 = (L & 1) != 0, b1 = (L & 2) != 0;
  sfor<M>([&](auto k) {
    constexpr int kk = decltype(k)::value;
    if constexpr (kk != 0) {
      const Coef2<float> w1 = ct_twiddle<64, kk, S>(), w2 = ct_twiddle<64, 2 * kk, S>(), w3 = ct_twiddle<64, 3 * kk, S>();
      float c, sn;
      if constexpr (P == 2) {
        c = b0 ? w1.x : 1.0f;
        sn = b0 ? w1.y : 0.0f;
      } else {
        c = b1 ? (b0 ? w3.x : w2.x) : (b0 ? w1.x : 1.0f);
        sn = b1 ? (b0 ? w3.y : w2.y) : (b0 ? w1.y : 0.0f);
      }
      v[k] = cmul<float>(v[k], coef2<float>(c, sn));
    }
  });
}
template <int P>
__device__ __forceinline__ int lane_rev(int L) {
  return P == 4 ? 2 * (L & 1) + (L >> 1) : L;
}
// the in-lane split of a 64/P-point DFT
template <int P, int S>
__device__ __forceinline__ void lane_dft(Coef2<float>* v) {
  if constexpr (P == 1) fft_reg<8, 8, S>(v);
  else if constexpr (P == 2) fft_reg<8, 4, S>(v);
  else fft_reg<4, 4, S>(v);
}
// DFT (sign S) of a line held by P lanes: points P n + L in, frequencies
// k + (64/P) rev(L) out
template <int P, int S>
__device__ __forceinline__ void line_fft64(Coef2<float>* v, int L) {
  constexpr int M = 64 / P;
  lane_dft<P, S>(v);
  if constexpr (P > 1) lane_twiddle<P, S>(v, L);
  if constexpr (P == 2) lane_stage<M, kQuadXor1, 1>(v, L);
  if constexpr (P == 4) {
    lane_stage<M, kQuadXor2, 2>(v, L);
    lane_rot3<M, S>(v, L);
    lane_stage<M, kQuadXor1, 1>(v, L);
  }
}
// the same DFT in the other arrangement: values at k + (64/P) rev(L) in,
// points P n + L out
template <int P, int S>
__device__ __forceinline__ void line_fft64_rev(Coef2<float>* v, int L) {
  constexpr int M = 64 / P;
  if constexpr (P == 2) lane_stage<M, kQuadXor1, 1>(v, L);
  if constexpr (P == 4) {
    lane_stage<M, kQuadXor1, 1>(v, L);
    lane_rot3<M, S>(v, L);
    lane_stage<M, kQuadXor2, 2>(v, L);
  }
  if constexpr (P > 1) lane_twiddle<P, S>(v, L);
  lane_dft<P, S>(v);
}

// The z pass of the convolution for 64^3 FP32 grids: a wave holds the z
// lines of 64/P consecutive x of one y row; forward transform, (MODE 3:
// the spectrum kept in `spec`), times qfac, backward transform into the
// points the lines were loaded from.
template <int P, int MODE>
__global__ void __launch_bounds__(64) k_fft_z64(const SysView* views, const Work* work, int gate, void* spec) {
  constexpr int M = 64 / P;
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int K0 = sv.nfft[0], K1 = sv.nfft[1];
  const int nch = (K0 + M - 1) / M;
  if (w.blk >= K1 * nch) return;
  const int y = w.blk / nch;
  const int L = threadIdx.x & (P - 1);
  const int x = (w.blk - y * nch) * M + threadIdx.x / P;
  const bool ok = x < K0;  // the P lanes of a line share x
  const size_t plane = (size_t)K0 * K1;
  const size_t base = (size_t)y * K0 + (ok ? x : 0);
  GLOBAL Coef2<float>* G = gw(static_cast<Coef2<float>*>(sv.grid)) + base;
  const GLOBAL float* Q = gp(static_cast<const float*>(sv.qfac)) + base;
  Coef2<float> v[M];
  sfor<M>([&](auto n) { v[n] = G[(P * n + L) * plane]; });
  line_fft64<P, -1>(v, L);
  const int kq = M * lane_rev<P>(L);
  if constexpr (MODE == 3) {
    GLOBAL Coef2<float>* SP = gw(static_cast<Coef2<float>*>(spec)) + base;
    if (ok) sfor<M>([&](auto k) { SP[(k + kq) * plane] = v[k]; });
  }
  sfor<M>([&](auto k) { v[k] *= Q[(k + kq) * plane]; });
  line_fft64_rev<P, 1>(v, L);
  if (ok) sfor<M>([&](auto n) { G[(P * n + L) * plane] = v[n]; });
}

// The x and y passes of one z plane of a 64^3 FP32 grid in a workgroup of
// P waves: line c (P lanes) loads the points P n + L of column x = c (the y
// line) and transforms it; the plane goes through LDS, where line c takes
// the points of row y = c (the x line) and transforms that; back through
// LDS for the store of the columns.  MODE 0 forward, 1 backward (the sign
// of both transforms).
template <int P, int MODE>
__global__ void __launch_bounds__(64 * P) k_fft_plane64(const SysView* views, const Work* work, int gate) {
  constexpr int M = 64 / P;
  __shared__ Coef2<float> s_t[64 * 65];
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  if (w.blk >= sv.nfft[2]) return;
  constexpr int S = MODE == 0 ? -1 : 1;
  const int L = threadIdx.x & (P - 1), c = threadIdx.x / P;
  const int kq = M * lane_rev<P>(L);
  GLOBAL Coef2<float>* G = gw(static_cast<Coef2<float>*>(sv.grid)) + (size_t)w.blk * 4096 + c;
  Coef2<float> v[M];
  sfor<M>([&](auto n) { v[n] = G[64 * (P * n + L)]; });
  line_fft64<P, S>(v, L);
  sfor<M>([&](auto k) { s_t[65 * (k + kq) + c] = v[k]; });
  __syncthreads();
  sfor<M>([&](auto n) { v[n] = s_t[65 * c + P * n + L]; });
  line_fft64<P, S>(v, L);
  __syncthreads();
  sfor<M>([&](auto k) { s_t[65 * c + k + kq] = v[k]; });
  __syncthreads();
  sfor<M>([&](auto n) { v[n] = s_t[65 * (P * n + L) + c]; });
  sfor<M>([&](auto n) { G[64 * (P * n + L)] = v[n]; });
}

// Tile gather: fphi_uind_site first derivatives (pmestuff.f:746-1037,
// fdip_phi(2:4) of both grids) and the fractional->Cartesian field
// (newinduce_pme.f:2103-2110).  One wavefront per brick loads the brick's
// (B+4)^3 tile of the grid into LDS once; each lane then gathers one home
// atom's 5x5x5 stencil from LDS (rows summed in the grid type, rows
// accumulated in FP64).  MODE 0 writes the field (fdip_phi or, with
// frac_out, the fractional derivatives); MODE 1 also starts the PCG
// (res = ef - h, zr, pp, newinduce_pme2.f:406-438) and MODE 2 also forms
// p.h (newinduce_pme2.f:452-473), with one partial pair per brick.
// Without RECIP (the short-range solve of newinduce_shortreal, with no
// reciprocal field) the kernel only carries out the PCG steps.
template <typename real, int MODE, bool RECIP = true>
__global__ void __launch_bounds__(64) k_gather_tile(const SysView* views, const Work* bwork, int vin, int vout,
                                                     int frac_out, int gate) {
  __shared__ real s_g[kTile3 * 2];
  const Work w = bwork[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int K0 = sv.nfft[0], K1 = sv.nfft[1], K2 = sv.nfft[2];
  const int B0 = sv.brick[0], B1 = sv.brick[1], B2 = sv.brick[2];
  const int nb0 = sv.nbrick[0], nb1 = sv.nbrick[1], nb2 = sv.nbrick[2];
  const int tb = w.blk;
  const int t0 = tb % nb0, t1 = (tb / nb0) % nb1, t2 = tb / (nb0 * nb1);
  const int sx = t0 * B0, sy = t1 * B1, sz = t2 * B2;
  const int e0 = (t0 == nb0 - 1 ? K0 - sx : B0) + kOrder - 1;
  const int e1 = (t1 == nb1 - 1 ? K1 - sy : B1) + kOrder - 1;
  const int e2 = (t2 == nb2 - 1 ? K2 - sz : B2) + kOrder - 1;
  const int lane = threadIdx.x;
  // tile load: a batch of loads in flight per lane before its LDS writes
  // (points outside the brick's extent are never read and are left as 0)
  if (RECIP) {
    typedef real R2 __attribute__((ext_vector_type(2)));
    const GLOBAL R2* GRID2 = reinterpret_cast<const GLOBAL R2*>(gp(static_cast<const real*>(sv.grid)));
    R2* s_g2 = reinterpret_cast<R2*>(s_g);
    static_assert(kTile3 % 64 == 0, "whole passes over the tile");
    constexpr int PER = kTile3 / 64;
    constexpr int BATCH = sizeof(real) == 4 ? PER : PER / 3;
    static_assert(PER % BATCH == 0, "whole batches");
    for (int k0 = 0; k0 < PER; k0 += BATCH) {
      R2 v[BATCH];
#pragma unroll
      for (int k = 0; k < BATCH; ++k) {
        const int p = lane + 64 * (k0 + k);
        const int ox = p % kTile, oy = (p / kTile) % kTile, oz = p / (kTile * kTile);
        int gx = sx + ox, gy = sy + oy, gz = sz + oz;
        if (gx >= K0) gx -= K0;
        if (gy >= K1) gy -= K1;
        if (gz >= K2) gz -= K2;
        const bool in = ox < e0 && oy < e1 && oz < e2;
        const size_t g = in ? ((size_t)gz * K1 + gy) * K0 + gx : 0;
        v[k] = GRID2[g];
        if (!in) v[k] = R2{real(0), real(0)};
      }
#pragma unroll
      for (int k = 0; k < BATCH; ++k) s_g2[lane + 64 * (k0 + k)] = v[k];
    }
  }
  __syncthreads();
  const size_t np = sv.npad;
  const GLOBAL real* THETA = gp(static_cast<const real*>(sv.theta));
  const GLOBAL int* G0 = gp(sv.g0);
  const int a0 = gp(sv.bin_start)[tb], a1 = gp(sv.bin_start)[tb + 1];
  const GLOBAL int* BLIST = gp(sv.bin_list);
  double qs[2] = {0.0, 0.0};
  for (int c0 = a0; c0 < a1; c0 += 64) {
    if (c0 + lane >= a1) break;
    const int i = BLIST[c0 + lane];
    const int lx = G0[4 * i] - sx, ly = G0[4 * i + 1] - sy, lz = G0[4 * i + 2] - sz;
    const GLOBAL real* th = THETA + (size_t)i * kTheta;
    real tx0[kOrder], tx1[kOrder], ty0[kOrder], ty1[kOrder], tz0[kOrder], tz1[kOrder];
#pragma unroll
    for (int m = 0; m < kOrder; ++m) {
      tx0[m] = th[th_idx(0, 0, m)];
      tx1[m] = th[th_idx(0, 1, m)];
      ty0[m] = th[th_idx(1, 0, m)];
      ty1[m] = th[th_idx(1, 1, m)];
      tz0[m] = th[th_idx(2, 0, m)];
      tz1[m] = th[th_idx(2, 1, m)];
    }
    double a[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int mz = 0; mz < (RECIP ? kOrder : 0); ++mz) {
      real v0 = tz0[0], v1 = tz1[0];
#pragma unroll
      for (int m = 1; m < kOrder; ++m)
        if (mz == m) {
          v0 = tz0[m];
          v1 = tz1[m];
        }
#pragma unroll
      for (int my = 0; my < kOrder; ++my) {
        const real u0 = ty0[my], u1 = ty1[my];
        const real* row = s_g + 2 * (((lz + mz) * kTile + (ly + my)) * kTile + lx);
        real t0_1 = 0, t1_1 = 0, t0_2 = 0, t1_2 = 0;
#pragma unroll
        for (int mx = 0; mx < kOrder; ++mx) {
          const real q1 = row[2 * mx], q2 = row[2 * mx + 1];
          t0_1 += q1 * tx0[mx];
          t1_1 += q1 * tx1[mx];
          t0_2 += q2 * tx0[mx];
          t1_2 += q2 * tx1[mx];
        }
        a[0] += double(t1_1 * u0 * v0);
        a[1] += double(t0_1 * u1 * v0);
        a[2] += double(t0_1 * u0 * v1);
        a[3] += double(t1_2 * u0 * v0);
        a[4] += double(t0_2 * u1 * v0);
        a[5] += double(t0_2 * u0 * v1);
      }
    }
    double dip[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double sc = (MODE == 0 && frac_out) ? 1.0 : sv.afrac[c];
      dip[c] = sc * a[c];
      dip[3 + c] = sc * a[3 + c];
    }
    if (MODE == 0) {
      double* out = vsel(sv, vout);
#pragma unroll
      for (int c = 0; c < 6; ++c) out[(size_t)c * np + i] = dip[c];
    } else {
      const GLOBAL double* IN = gp(vsel(sv, vin));
      const double term = sv.term_self;
      const double pl = gp(sv.pol)[i];
      const double diag = pl == 0.0 ? kTinyPol : pl;
      const GLOBAL double* HR = gp(sv.hreal);
      GLOBAL double* H = gw(sv.h);
      // every load before the first store (the compiler cannot tell the
      // output vectors from the inputs)
      double uv[6], hv[6], ev[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const size_t o = (size_t)c * np + i;
        uv[c] = IN[o];
        hv[c] = HR[o];
        ev[c] = MODE == 1 ? gp(sv.ef)[o] : 0.0;
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const size_t o = (size_t)c * np + i;
        const double u = uv[c];
        const double h = hv[c] + dip[c] - term * u;
        H[o] = h;
        if (MODE == 1) {
          const double rr = ev[c] - h;
          const double z = diag * rr;
          gw(sv.res)[o] = rr;
          gw(sv.pp)[o] = z;
          qs[c / 3] += rr * z;
        } else {
          qs[c / 3] += u * h;
        }
      }
    }
  }
  if (MODE == 0) return;
  block_partial2(qs[0], qs[1], MODE == 1 ? part_init(sv) : part_gg(sv), w.blk);
}

// fphi_mpole_site (pmestuff.f:453-631, potential and derivatives through
// second order) + fphi_to_cphi_site (:1190-1215).  When add_to_ef, the
// permanent field becomes ef = E_direct - cphi(2:4) + term*d for both rhs
// (newinduce_pme2.f:134-147); cphi (optional) receives [10][npad].
template <typename real>
__global__ void __launch_bounds__(64) k_gather_mpole(const SysView* views, const Work* work, double* cphi_out,
                                                      int add_to_ef, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const size_t np = sv.npad;
  real t1[3][kOrder], t2[3][kOrder], t3[3][kOrder];
  const real* th = theta_row<real>(sv, i);
#pragma unroll
  for (int l = 0; l < 3; ++l)
#pragma unroll
    for (int m = 0; m < kOrder; ++m) {
      t1[l][m] = th[th_idx(0, l, m)];
      t2[l][m] = th[th_idx(1, l, m)];
      t3[l][m] = th[th_idx(2, l, m)];
    }
  const int g1 = sv.g0[4 * i], g2 = sv.g0[4 * i + 1], g3 = sv.g0[4 * i + 2];
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

// ===================================================== dipole predictor
// History of the converged dipoles of the last kMaxUalt solves and its
// extrapolation into the next guess (ulspred, newinduce_pme2.f:910-1010;
// the guess, :165-178; the history update, :253-264).  hist holds
// kMaxUalt slots of [6][npad] FP64 dipoles used as a ring; entry k counts
// back from the newest (k = 0).

// copy the solve's dipoles into history slot `slot`
__global__ void __launch_bounds__(256) k_pred_push(const SysView* views, double* hist, int slot) {
  const SysView& sv = views[0];
  const size_t m = (size_t)6 * sv.npad;
  const GLOBAL double* src = gp(sv.mu);
  GLOBAL double* dst = gw(hist) + (size_t)slot * m;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (size_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

// mu = sum_k coef[k] entry k over the kMaxUalt - 1 newest entries, the same
// coefficients for d and p (newinduce_pme2.f:170-176)
__global__ void __launch_bounds__(256) k_pred_apply(const SysView* views, const double* hist, const double* coef,
                                                    int newest) {
  const SysView& sv = views[0];
  const size_t m = (size_t)6 * sv.npad;
  double c[kMaxUalt - 1];
#pragma unroll
  for (int k = 0; k < kMaxUalt - 1; ++k) c[k] = coef[k];
  const GLOBAL double* H = gp(hist);
  GLOBAL double* MU = gw(sv.mu);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (size_t)gridDim.x * blockDim.x) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxUalt - 1; ++k) v += c[k] * H[(size_t)((newest - k + kMaxUalt) % kMaxUalt) * m + q];
    MU[q] = v;
  }
}

// least-squares predictor, normal equations: per 256-atom block, the sums
// c(k, m) = u_k . u_m over the block's atoms and the three d components,
// 0 <= k <= m < kMaxUalt, packed by rows of the upper triangle
__global__ void __launch_bounds__(256) k_pred_normal(const SysView* views, const double* hist, int newest,
                                                     double* part) {
  const SysView& sv = views[0];
  const size_t m = (size_t)6 * sv.npad;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double acc[kPredPairs];
#pragma unroll
  for (int p = 0; p < kPredPairs; ++p) acc[p] = 0.0;
  if (i < sv.n) {
    const GLOBAL double* H = gp(hist);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double u[kMaxUalt];
#pragma unroll
      for (int k = 0; k < kMaxUalt; ++k) u[k] = H[(size_t)((newest - k + kMaxUalt) % kMaxUalt) * m + (size_t)c * sv.npad + i];
      int p = 0;
#pragma unroll
      for (int k = 0; k < kMaxUalt; ++k)
#pragma unroll
        for (int q = k; q < kMaxUalt; ++q) acc[p++] += u[k] * u[q];
    }
  }
  __shared__ double red[kPredPairs][4];
#pragma unroll
  for (int p = 0; p < kPredPairs; ++p) {
    const double v = wave_sum(acc[p]);
    if ((threadIdx.x & 63) == 0) red[p][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kPredPairs)
    part[(size_t)blockIdx.x * kPredPairs + threadIdx.x] =
        ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
}

// the fit of the newest entry on the kMaxUalt - 1 older ones: normal
// equations summed over the block partials in block order, solved by the
// symmetric decomposition A = L D L^T of cholesky.f (Martin, Peters and
// Wilkinson 1965) when some A(k, m) is positive (newinduce_pme2.f:986-997);
// coef[kMaxUalt - 1] = 0
__global__ void __launch_bounds__(64) k_pred_solve(const double* part, int nblocks, double* coef) {
  __shared__ double c[kPredPairs];
  for (int p = threadIdx.x; p < kPredPairs; p += 64) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * kPredPairs + p];
    c[p] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  constexpr int N = kMaxUalt - 1;
  auto C = [&](int k, int q) {
    if (k > q) {
      const int t = k;
      k = q;
      q = t;
    }
    return c[k * kMaxUalt - k * (k - 1) / 2 + (q - k)];
  };
  double A[N][N], x[N];
  double amax = 0.0;
  for (int j = 0; j < N; ++j) {
    x[j] = C(0, j + 1);
    for (int l = 0; l < N; ++l) A[j][l] = C(j + 1, l + 1);
    for (int l = j; l < N; ++l) amax = fmax(amax, A[j][l]);
  }
  if (amax != 0.0) {
    double L[N][N], D[N];
    for (int r = 0; r < N; ++r) {
      for (int j = 0; j < r; ++j) {
        double s = A[r][j];
        for (int k = 0; k < j; ++k) s -= L[r][k] * D[k] * L[j][k];
        L[r][j] = s / D[j];
      }
      double s = A[r][r];
      for (int k = 0; k < r; ++k) s -= L[r][k] * L[r][k] * D[k];
      D[r] = s;
    }
    for (int r = 0; r < N; ++r)
      for (int k = 0; k < r; ++k) x[r] -= L[r][k] * x[k];
    for (int r = 0; r < N; ++r) x[r] /= D[r];
    for (int r = N - 1; r >= 0; --r)
      for (int k = r + 1; k < N; ++k) x[r] -= L[k][r] * x[k];
  }
  for (int j = 0; j < N; ++j) coef[j] = x[j];
  coef[N] = 0.0;
}

// ===================================================== Jacobi / DIIS solver
// Jacobi iterations with DIIS extrapolation, the reference's polalg = 2
// (inducejac_pme2, newinduce_pme2.f:638-890; diis, makeb and extrap,
// diis.f).  The subspace's Jacobi iterates x_j and increments e_j live in
// two device rings of kDiisMax [6][npad] FP64 vectors: subspace vector j
// (1 = oldest) sits in slot (first + j - 1) mod kDiisMax.  The subspace size
// and ring start do not depend on the data, so the host tracks them and
// passes them in; Pulay's matrix and the extrapolation coefficients stay on
// the device.

// Jacobi step of every atom: h = -E_real(mu) (no diagonal) + the reciprocal
// dipole field (in sv.h) - term mu, munew = alpha (ef - h)
// (newinduce_pme2.f:739-760), stored with its increment e = munew - mu in
// ring slot `slot`.  Per 256-atom block: the increment norms of both rhs
// and the products e_j . e of the nvec - 1 older subspace vectors and e . e
// (makeb, diis.f).
__global__ void __launch_bounds__(256) k_jac_step(const SysView* views, double* xr, double* er, int slot, int first,
                                                  int nvec, double* part) {
  const SysView& sv = views[0];
  const size_t np = sv.npad, m = (size_t)6 * np;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double rn0 = 0.0, rn1 = 0.0;
  double dots[kDiisMax];
#pragma unroll
  for (int j = 0; j < kDiisMax; ++j) dots[j] = 0.0;
  if (i < sv.n) {
    const double a = gp(sv.pol)[i], term = sv.term_self;
    double e[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      const double u = gp(sv.mu)[o];
      const double h = gp(sv.hreal)[o] + gp(sv.h)[o] - term * u;
      const double nw = a * (gp(sv.ef)[o] - h);
      e[c] = nw - u;
      gw(xr)[(size_t)slot * m + o] = nw;
      gw(er)[(size_t)slot * m + o] = e[c];
    }
    rn0 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2];
    rn1 = (e[3] * e[3] + e[4] * e[4]) + e[5] * e[5];
#pragma unroll
    for (int j = 0; j < kDiisMax - 1; ++j) {
      if (j < nvec - 1) {
        const GLOBAL double* ej = gp(er) + (size_t)((first + j) % kDiisMax) * m + i;
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) s += ej[(size_t)c * np] * e[c];
        dots[j] = s;
      }
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += e[c] * e[c];
#pragma unroll
    for (int j = 0; j < kDiisMax; ++j)
      if (j == nvec - 1) dots[j] = s;
  }
  __shared__ double red[kDiisParts][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  {
    const double v0 = wave_sum(rn0), v1 = wave_sum(rn1);
    if (lane == 0) {
      red[0][wv] = v0;
      red[1][wv] = v1;
    }
  }
#pragma unroll
  for (int j = 0; j < kDiisMax; ++j) {
    if (j < nvec) {
      const double v = wave_sum(dots[j]);
      if (lane == 0) red[2 + j][wv] = v;
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < 2 + nvec; p += 256)
    part[(size_t)blockIdx.x * kDiisParts + p] = ((red[p][0] + red[p][1]) + red[p][2]) + red[p][3];
}

// Pulay's matrix and the extrapolation coefficients (diis, makeb: diis.f;
// the dgesv call of newinduce_pme2.f:807-812), then the convergence test of
// the Jacobi increment, max over rhs of sqrt(|e|^2 / (3n)) < poleps
// (:836-846).  One workgroup; the small system is solved by Gaussian
// elimination with partial pivoting, as LAPACK's dgesv.
__global__ void __launch_bounds__(64) k_jac_solve(const SysView* views, DiisState* ds, const double* part,
                                                  int nblocks, int nvec, int shifted, int it) {
  const SysView& sv = views[0];
  __shared__ double sums[kDiisParts];
  __shared__ double A[kDiisMax + 1][kDiisMax + 1], x[kDiisMax + 1];
  for (int p = threadIdx.x; p < 2 + nvec; p += 64) {
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * kDiisParts + p];
    sums[p] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double(*B)[kDiisMax + 1] = ds->B;
  if (shifted) {  // the ten oldest vectors left the subspace
    for (int j = 1; j < nvec; ++j)
      for (int k = 1; k < nvec; ++k) B[j][k] = B[j + 10][k + 10];
  }
  if (nvec == 1) {
    B[0][0] = 0.0;
    B[0][1] = B[1][0] = 1.0;
    B[1][1] = sums[2];
  } else {
    B[nvec][0] = B[0][nvec] = 1.0;
    for (int j = 1; j < nvec; ++j) B[nvec][j] = B[j][nvec] = sums[2 + j - 1];
    B[nvec][nvec] = sums[2 + nvec - 1];
  }
  const int nm = nvec + 1;
  for (int r = 0; r < nm; ++r) {
    for (int k = 0; k < nm; ++k) A[r][k] = B[r][k];
    x[r] = r == 0 ? 1.0 : 0.0;
  }
  for (int k = 0; k < nm; ++k) {
    int piv = k;
    for (int r = k + 1; r < nm; ++r)
      if (fabs(A[r][k]) > fabs(A[piv][k])) piv = r;
    if (piv != k) {
      for (int q = 0; q < nm; ++q) {
        const double t = A[k][q];
        A[k][q] = A[piv][q];
        A[piv][q] = t;
      }
      const double t = x[k];
      x[k] = x[piv];
      x[piv] = t;
    }
    for (int r = k + 1; r < nm; ++r) {
      const double f = A[r][k] / A[k][k];
      for (int q = k + 1; q < nm; ++q) A[r][q] -= f * A[k][q];
      x[r] -= f * x[k];
    }
  }
  for (int r = nm - 1; r >= 0; --r) {
    double s = x[r];
    for (int q = r + 1; q < nm; ++q) s -= A[r][q] * x[q];
    x[r] = s / A[r][r];
  }
  for (int r = 0; r < nm; ++r) ds->cex[r] = x[r];
  SolveState* st = sv.st;
  const double n3 = 3.0 * double(sv.n);
  const double rr = fmax(sqrt(sums[0] / n3), sqrt(sums[1] / n3));
  st->niter = it;
  st->resnrm = rr;
  if (!(rr == rr) || rr > 1e300) st->conv = kNonFinite;
  else if (rr < st->poleps) st->conv = kConverged;
  else if (it >= st->politer) st->conv = kMaxIter;
}

// mu = sum_j cex[j] x_j over the subspace, oldest first (extrap, diis.f)
__global__ void __launch_bounds__(256) k_jac_extrap(const SysView* views, const double* xr, const DiisState* ds,
                                                    int first, int nvec) {
  const SysView& sv = views[0];
  const size_t m = (size_t)6 * sv.npad;
  const GLOBAL double* X = gp(xr);
  GLOBAL double* MU = gw(sv.mu);
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (size_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int j = 1; j <= nvec; ++j) v += ds->cex[j] * X[(size_t)((first + j - 1) % kDiisMax) * m + q];
    MU[q] = v;
  }
}

// ===================================================== PCG vector kernels
// Reductions: each reducing kernel writes one pair of FP64 partial sums
// per block (G = red_blocks blocks per system); the kernels that need a
// scalar re-sum the G partials in a fixed order themselves.  A kernel
// boundary orders these accesses, so no fences or atomics are needed, and
// every block obtains bit-identical scalars.

// direct guess mu = alpha*ef (polgsf = 0, newinduce_pme2.f:180-188)
__global__ void __launch_bounds__(64) k_guess(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate) || sv.st->warm) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const double a = sv.pol[i];
#pragma unroll
  for (int c = 0; c < 6; ++c) sv.mu[(size_t)c * sv.npad + i] = a * sv.ef[(size_t)c * sv.npad + i];
}

// pcg_b (newinduce_pme2.f:496-541) for iteration it = niter + 1 (the
// iteration number lives in the solve state, so one launch can carry
// systems at different iterations and captured launches can be replayed):
// alpha = ggold/gg with ggold = sum of rz partials of iteration it-1;
// mu += alpha pp; res -= alpha h; partial ggnew = sum res (diag res)
// (into rz buffer it&1).  The reference's ene (newinduce_pme2.f:527-538)
// feeds only its polprt printout (:546-548) and is not formed here.
template <typename real>
__global__ void __launch_bounds__(256) k_cg_update1(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  SolveState* st = sv.st;
  if (st->conv || st->in_setup) return;
  const int it = st->niter + 1;
  const int G = sv.red_blocks, P = sv.gat_blocks;
  __shared__ double sc[4];
  __shared__ double scratch[8];
  {
    const double* po = it == 1 ? part_init(sv) : part_rz(sv, (it - 1) & 1);
    const int npo = it == 1 ? P : G;
    const double* pg = part_gg(sv);
    const double go0 = block_sum_partials(po, npo, 0, scratch), go1 = block_sum_partials(po, npo, 1, scratch);
    const double g0 = block_sum_partials(pg, P, 0, scratch), g1 = block_sum_partials(pg, P, 1, scratch);
    if (threadIdx.x == 0) {
      sc[0] = go0 / g0;
      sc[1] = go1 / g1;
      sc[2] = g0;
      sc[3] = g1;
      if (w.blk == 0) {  // the only writer of these fields in this kernel
        st->gg[0] = g0;
        st->gg[1] = g1;
        st->it_cur = it;
        if (g0 <= 0.0 || g1 <= 0.0) st->n_indef += 1;
      }
    }
  }
  __syncthreads();
  if (sc[2] == 0.0 || sc[3] == 0.0) {  // the reference returns before updating mu
    if (threadIdx.x == 0) st->conv = kZeroCurv;
    return;
  }
  const double alpha[2] = {sc[0], sc[1]};
  const size_t np = sv.npad;
  const int n = sv.n;
  GLOBAL double* MU = gw(sv.mu);
  GLOBAL double* RES = gw(sv.res);
  const GLOBAL double* PP = gp(sv.pp);
  const GLOBAL double* H = gp(sv.h);
  const GLOBAL double* POL = gp(sv.pol);
  // the preconditioned residual z = diag r is not stored: k_cg_update2
  // forms it again from the stored r (the same product, the same value)
  double gn[2] = {0, 0};
  for (int i = w.blk * blockDim.x + threadIdx.x; i < n; i += G * blockDim.x) {
    const double pl = POL[i];
    const double diag = pl == 0.0 ? kTinyPol : pl;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      const double a = alpha[c / 3];
      const double m = MU[o] + a * PP[o];
      const double r = RES[o] - a * H[o];
      const double z = diag * r;
      MU[o] = m;
      RES[o] = r;
      gn[c / 3] += r * z;
    }
  }
  block_partial2(gn[0], gn[1], part_rz(sv, it & 1), w.blk);
}

// pcg_newDirection + convergence (newinduce_pme2.f:543-568) for the
// iteration started by k_cg_update1 (it_cur): pp = zr + ggnew/ggold pp with zr = diag res; residual norm sqrt(ggnew/(3 npolar)),
// stop when the max over both rhs is below poleps.
__global__ void __launch_bounds__(256) k_cg_update2(const SysView* views, const Work* work) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  SolveState* st = sv.st;
  if (st->conv || st->in_setup) return;
  const int it = st->it_cur;
  const int G = sv.red_blocks;
  __shared__ double sc[3];
  __shared__ double scratch[8];
  {
    const double* pn = part_rz(sv, it & 1);
    const double* po = it == 1 ? part_init(sv) : part_rz(sv, (it - 1) & 1);
    const int npo = it == 1 ? sv.gat_blocks : G;
    const double gn0 = block_sum_partials(pn, G, 0, scratch), gn1 = block_sum_partials(pn, G, 1, scratch);
    const double go0 = block_sum_partials(po, npo, 0, scratch), go1 = block_sum_partials(po, npo, 1, scratch);
    if (threadIdx.x == 0) {
      sc[0] = gn0 / go0;
      sc[1] = gn1 / go1;
      const double n3 = 3.0 * double(sv.n);
      const double rn = fmax(sqrt(gn0 / n3), sqrt(gn1 / n3));
      int conv = kRunning;
      if (!isfinite(gn0) || !isfinite(gn1)) conv = kNonFinite;
      else if (rn < st->poleps) conv = kConverged;
      else if (it >= st->politer) conv = kMaxIter;
      sc[2] = conv;
      if (w.blk == 0) {  // the only writer (and reader) of these fields
        st->ggnew[0] = gn0;
        st->ggnew[1] = gn1;
        st->resnrm = rn;
        st->niter = it;
      }
    }
  }
  __syncthreads();
  const double beta[2] = {sc[0], sc[1]};
  const size_t np = sv.npad;
  const int n = sv.n;
  GLOBAL double* PP = gw(sv.pp);
  const GLOBAL double* RES = gp(sv.res);
  const GLOBAL double* POL = gp(sv.pol);
  for (int i = w.blk * blockDim.x + threadIdx.x; i < n; i += G * blockDim.x) {
    const double pl = POL[i];
    const double diag = pl == 0.0 ? kTinyPol : pl;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      PP[o] = diag * RES[o] + beta[c / 3] * PP[o];
    }
  }
  // every block computed the same decision; the flag is read by the next
  // iteration's kernels (a block of this kernel that sees it set skips the
  // direction update, which is no longer used)
  if (threadIdx.x == 0 && sc[2] != kRunning) st->conv = int(sc[2]);
}

// ===================================================== local frames (rotpole)
// rotmat + rotsite (rotpole.f:24-339) with chkpole (chkpole.f:20-98) for
// every site whose frame is set: the rotation matrix from the positions of
// the site and its frame atoms (columns: the local x, y, z axes in the
// global frame; no minimum image, as without use_polymer), then
// d = A d_local and Q = A Q_local A^T.  chkpole inverts the y components of
// a Z-then-X site with a y-axis atom while the parallelepiped volume of
// (site, z, x) about y is negative: applied here from the volume at each
// call, it leaves the same multipoles as the reference's in-place sign
// flips.  Z-Only takes the global x axis (or y, when z lies near x) for
// the local x axis where the reference draws a random one: Z-Only
// parameters are symmetric about z, so the result does not depend on it.
// frames[q] = (axis type, z, x, y) in sorted slots, -1 for none; the local
// multipoles are [10][npad] in the layout of rpole.
enum : int { kAxisNone = 0, kAxisZOnly = 1, kAxisZThenX = 2, kAxisBisector = 3, kAxisZBisect = 4, kAxis3Fold = 5 };

__device__ __forceinline__ void unit3(double& x, double& y, double& z) {
  const double r = sqrt(x * x + y * y + z * z);
  x /= r;
  y /= r;
  z /= r;
}

__device__ __forceinline__ void rotate_site(const SysView& sv, const int* frames, const double* pole_local, int i) {
  const size_t np = sv.npad;
  const int4 f = reinterpret_cast<const int4*>(frames)[i];
  const int axt = f.x, iz = f.y, ix = f.z, iy = f.w;
  const double xi = sv.x[i], yi = sv.y[i], zi = sv.z[i];
  auto vec = [&](int k, double& dx, double& dy, double& dz) {
    dx = sv.x[k] - xi;
    dy = sv.y[k] - yi;
    dz = sv.z[k] - zi;
  };
  // a[r][c]: column 0 the local x axis, column 2 the local z axis
  double a[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  double zx = 0, zy = 0, zz = 1, xx = 1, xy = 0, xz = 0;
  if (axt != kAxisNone) {
    double d1x, d1y, d1z;
    vec(iz, d1x, d1y, d1z);
    unit3(d1x, d1y, d1z);
    double ux = 0, uy = 0, uz = 0;  // the vector the x axis is taken from
    if (axt == kAxisZOnly) {
      zx = d1x, zy = d1y, zz = d1z;
      if (fabs(zx) < 0.866) ux = 1.0;
      else uy = 1.0;
    } else if (axt == kAxisZThenX) {
      zx = d1x, zy = d1y, zz = d1z;
      vec(ix, ux, uy, uz);
    } else if (axt == kAxisBisector) {
      double d2x, d2y, d2z;
      vec(ix, d2x, d2y, d2z);
      unit3(d2x, d2y, d2z);
      zx = d1x + d2x, zy = d1y + d2y, zz = d1z + d2z;
      unit3(zx, zy, zz);
      ux = d2x, uy = d2y, uz = d2z;
    } else if (axt == kAxisZBisect) {
      zx = d1x, zy = d1y, zz = d1z;
      double d2x, d2y, d2z, d3x, d3y, d3z;
      vec(ix, d2x, d2y, d2z);
      unit3(d2x, d2y, d2z);
      vec(iy, d3x, d3y, d3z);
      unit3(d3x, d3y, d3z);
      ux = d2x + d3x, uy = d2y + d3y, uz = d2z + d3z;
      unit3(ux, uy, uz);
    } else {  // 3-Fold
      double d2x, d2y, d2z, d3x, d3y, d3z;
      vec(ix, d2x, d2y, d2z);
      unit3(d2x, d2y, d2z);
      vec(iy, d3x, d3y, d3z);
      unit3(d3x, d3y, d3z);
      zx = d1x + d2x + d3x, zy = d1y + d2y + d3y, zz = d1z + d2z + d3z;
      unit3(zx, zy, zz);
      ux = d2x, uy = d2y, uz = d2z;
    }
    const double dot = ux * zx + uy * zy + uz * zz;
    xx = ux - dot * zx, xy = uy - dot * zy, xz = uz - dot * zz;
    unit3(xx, xy, xz);
    a[0][2] = zx, a[1][2] = zy, a[2][2] = zz;
    a[0][0] = xx, a[1][0] = xy, a[2][0] = xz;
  }
  // y = z x x (rotpole.f: a(1,2) = a(3,1) a(2,3) - a(2,1) a(3,3), ...)
  a[0][1] = a[2][0] * a[1][2] - a[1][0] * a[2][2];
  a[1][1] = a[0][0] * a[2][2] - a[2][0] * a[0][2];
  a[2][1] = a[1][0] * a[0][2] - a[0][0] * a[1][2];
  double p[10];
#pragma unroll
  for (int c = 0; c < 10; ++c) p[c] = pole_local[(size_t)c * np + i];
  if (axt == kAxisZThenX && iy >= 0) {
    const double xad = xi - sv.x[iy], yad = yi - sv.y[iy], zad = zi - sv.z[iy];
    const double xbd = sv.x[iz] - sv.x[iy], ybd = sv.y[iz] - sv.y[iy], zbd = sv.z[iz] - sv.z[iy];
    const double xcd = sv.x[ix] - sv.x[iy], ycd = sv.y[ix] - sv.y[iy], zcd = sv.z[ix] - sv.z[iy];
    const double c1 = ybd * zcd - zbd * ycd, c2 = ycd * zad - zcd * yad, c3 = yad * zbd - zad * ybd;
    const double vol = xad * c1 + xbd * c2 + xcd * c3;
    if (vol < 0.0) {
      p[2] = -p[2];  // d_y
      p[5] = -p[5];  // Q_xy
      p[8] = -p[8];  // Q_yz
    }
  }
  // rpole layout: q, d (3), Qxx, Qxy, Qxz, Qyy, Qyz, Qzz
  const double q2[3][3] = {{p[4], p[5], p[6]}, {p[5], p[7], p[8]}, {p[6], p[8], p[9]}};
  double out[10];
  out[0] = p[0];
#pragma unroll
  for (int r = 0; r < 3; ++r) out[1 + r] = (p[1] * a[r][0] + p[2] * a[r][1]) + p[3] * a[r][2];
  double g[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = r; c < 3; ++c) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int m = 0; m < 3; ++m) s += a[r][k] * a[c][m] * q2[k][m];
      g[r][c] = s;
    }
  out[4] = g[0][0], out[5] = g[0][1], out[6] = g[0][2], out[7] = g[1][1], out[8] = g[1][2], out[9] = g[2][2];
  double* RP = const_cast<double*>(sv.rpole);
#pragma unroll
  for (int c = 0; c < 10; ++c) RP[(size_t)c * np + i] = out[c];
}

// one system, one thread per site
__global__ void __launch_bounds__(256) k_rotpole(const SysView* views, const int* frames, const double* pole_local) {
  const SysView& sv = views[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= sv.n) return;
  rotate_site(sv, frames, pole_local, i);
}

// a batch (the serving loop's set-up): the sites of every block of the
// systems that have frames, after k_load_coords has placed the coordinates
__global__ void __launch_bounds__(64) k_rotpole_batch(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (!sv.frames || gated_out(sv.st, gate)) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  rotate_site(sv, sv.frames, sv.pole_local, i);
}

// ===================================================== truncated CG (polalg 3)
// epolar3tcg1/epolar3tcg2 (epolar3tcg.f:66-311, 320-621) on one system.  The
// polarization matrix A = alpha^-1 - T is applied by the matvec with its
// diagonal (into hreal), the reciprocal dipole field (into h) and the self
// term: A v = hreal + h - term v.  Vectors: res = r0 before and zr = r0
// after the preconditioner (diagvec: times the polarity), pp = T r0 after
// it, vec[0] = A r0 before it, vec[1] = A (T r0) after it; mu holds the
// guess alpha ef (tcgguess) and finally the dipoles.  Dot products are
// per-block partials summed in a fixed order; sc holds the scalars.
enum : int { kTcgN0 = 0, kTcgT1 = 2, kTcgNp1 = 4, kTcgT4 = 6, kTcgT2 = 8, kTcgSp = 10, kTcgG1 = 12, kTcgScalars = 14 };

// r0 = ef - A mu0 (or ef), its preconditioned copy, n0 = r0 . M r0
__global__ void __launch_bounds__(256) k_tcg_start(const SysView* views, int guess, int prec, double* part) {
  const SysView& sv = views[0];
  const size_t np = sv.npad;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double n0[2] = {0.0, 0.0};
  if (i < sv.n) {
    const double a = gp(sv.pol)[i], term = sv.term_self;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      const double h = guess ? gp(sv.hreal)[o] + gp(sv.h)[o] - term * gp(sv.mu)[o] : 0.0;
      const double oldr = gp(sv.ef)[o] - h;
      const double r = prec ? a * oldr : oldr;
      gw(sv.res)[o] = oldr;
      gw(sv.zr)[o] = r;
      n0[c / 3] += oldr * r;
    }
  }
  block_partial2(n0[0], n0[1], part, blockIdx.x);
}

// A r0 and its preconditioned copy T r0; t1 = r0 . A r0, np1 = A r0 . T r0
__global__ void __launch_bounds__(256) k_tcg_apply1(const SysView* views, int prec, double* vec, double* part_t1,
                                                    double* part_np1) {
  const SysView& sv = views[0];
  const size_t np = sv.npad;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double t1[2] = {0.0, 0.0}, np1[2] = {0.0, 0.0};
  if (i < sv.n) {
    const double a = gp(sv.pol)[i], term = sv.term_self;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      const double r = gp(sv.zr)[o];
      const double ar = gp(sv.hreal)[o] + gp(sv.h)[o] - term * r;
      const double tr = prec ? a * ar : ar;
      gw(vec)[o] = ar;
      gw(sv.pp)[o] = tr;
      t1[c / 3] += r * ar;
      np1[c / 3] += ar * tr;
    }
  }
  block_partial2(t1[0], t1[1], part_t1, blockIdx.x);
  block_partial2(np1[0], np1[1], part_np1, blockIdx.x);
}

// n0, t1, np1 summed; t4 = n0 / t1, t2 = n0 np1 / t1^2 (per right-hand side)
__global__ void __launch_bounds__(64) k_tcg_scalars1(const double* part_n0, const double* part_t1,
                                                      const double* part_np1, int nblk, double* sc) {
  double v[6];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    v[r] = sum_partials(part_n0, nblk, r);
    v[2 + r] = sum_partials(part_t1, nblk, r);
    v[4 + r] = sum_partials(part_np1, nblk, r);
  }
  if (threadIdx.x == 0) {
    for (int r = 0; r < 2; ++r) {
      const double n0 = v[r], t1 = v[2 + r], np1 = v[4 + r];
      sc[kTcgN0 + r] = n0;
      sc[kTcgT1 + r] = t1;
      sc[kTcgNp1 + r] = np1;
      sc[kTcgT4 + r] = n0 / t1;
      sc[kTcgT2 + r] = n0 * np1 / (t1 * t1);
    }
  }
}

// A (T r0), preconditioned into vec[1]; P2 = t2 T r0 - t4 T2 r0 and
// A r0 . P2 (epolar3tcg.f:500-526)
__global__ void __launch_bounds__(256) k_tcg_apply2(const SysView* views, int prec, double* vec, const double* sc,
                                                    double* part) {
  const SysView& sv = views[0];
  const size_t np = sv.npad;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double sp[2] = {0.0, 0.0};
  if (i < sv.n) {
    const double a = gp(sv.pol)[i], term = sv.term_self;
    const double t2[2] = {sc[kTcgT2], sc[kTcgT2 + 1]}, t4[2] = {sc[kTcgT4], sc[kTcgT4 + 1]};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const size_t o = (size_t)c * np + i;
      const double tr = gp(sv.pp)[o];
      const double a2 = gp(sv.hreal)[o] + gp(sv.h)[o] - term * tr;
      const double t2r = prec ? a * a2 : a2;
      gw(vec)[np * 6 + o] = t2r;
      const double p2 = t2[c / 3] * tr - t4[c / 3] * t2r;
      sp[c / 3] += gp(vec)[o] * p2;
    }
  }
  block_partial2(sp[0], sp[1], part, blockIdx.x);
}

// t3 = t1 (A r0 . P2), g1 = (t1^2 - n0 np1) / t3
__global__ void __launch_bounds__(64) k_tcg_scalars2(const double* part, int nblk, double* sc) {
  const double s0 = sum_partials(part, nblk, 0), s1 = sum_partials(part, nblk, 1);
  if (threadIdx.x == 0) {
    const double s[2] = {s0, s1};
    for (int r = 0; r < 2; ++r) {
      const double t1 = sc[kTcgT1 + r], n0 = sc[kTcgN0 + r], np1 = sc[kTcgNp1 + r];
      sc[kTcgSp + r] = s[r];
      sc[kTcgG1 + r] = (t1 * t1 - n0 * np1) / (t1 * s[r]);
    }
  }
}

// the dipoles (epolar3tcg.f:263-300, 527-590): order 1: t4 r0; order 2:
// (g1 t2 + t4) r0 - g1 t4 T r0; plus the guess; plus the peek step, whose
// scalars are those of the first right-hand side for both, and which
// multiplies by the polarity itself when the preconditioner is off
__global__ void __launch_bounds__(256) k_tcg_final(const SysView* views, const double* vec, const double* sc, int order,
                                                   int prec, int guess, int peek, double omega) {
  const SysView& sv = views[0];
  const size_t np = sv.npad;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= sv.n) return;
  const double a = gp(sv.pol)[i];
  const double pa = prec ? 1.0 : a;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int r = c / 3;
    const size_t o = (size_t)c * np + i;
    const double r0 = gp(sv.zr)[o], tr = gp(sv.pp)[o];
    const double t4 = sc[kTcgT4 + r];
    double m;
    if (order == 1) {
      m = t4 * r0;
    } else {
      const double g1 = sc[kTcgG1 + r], t2 = sc[kTcgT2 + r];
      m = (g1 * t2 + t4) * r0 - g1 * t4 * tr;
    }
    if (guess) m += gp(sv.mu)[o];
    if (peek) {
      double pk = omega * (pa * r0) - omega * sc[kTcgT4] * (pa * tr);
      if (order == 2)
        pk = (pk - omega * sc[kTcgG1] * sc[kTcgT2] * (pa * tr)) + omega * sc[kTcgG1] * sc[kTcgT4] * (pa * gp(vec)[np * 6 + o]);
      m += pk;
    }
    gw(sv.mu)[o] = m;
  }
}

// a truncated solve has no convergence test: it is complete after its steps
__global__ void __launch_bounds__(64) k_tcg_done(const SysView* views, int order) {
  if (threadIdx.x != 0) return;
  SolveState* st = views[0].st;
  st->niter = order;
  st->it_cur = order;
  st->conv = kConverged;
  st->resnrm = 0.0;
}


// E_pol = -1/2 f sum mu_d . E_p (epolar0c: real + recip + self); partials
// into the energy buffer, summed by k_finish.
__global__ void __launch_bounds__(256) k_energy(const SysView* views, const Work* work, int only_done) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (only_done && (sv.st->conv == kRunning || sv.st->finalized || sv.st->in_setup)) return;
  const size_t np = sv.npad;
  double e = 0.0;
  for (int i = w.blk * blockDim.x + threadIdx.x; i < sv.n; i += sv.red_blocks * blockDim.x) {
#pragma unroll
    for (int c = 0; c < 3; ++c) e += sv.mu[(size_t)c * np + i] * sv.ef[(size_t)(3 + c) * np + i];
  }
  block_partial2(e, 0.0, part_ep(sv), w.blk);
}

// one wave per system: final scalars into the SolveState
__global__ void __launch_bounds__(64) k_finish(const SysView* views, const int* sys_ids, int only_done) {
  const SysView& sv = views[sys_ids[blockIdx.x]];
  if (only_done && (sv.st->conv == kRunning || sv.st->finalized || sv.st->in_setup)) return;
  const double e = sum_partials(part_ep(sv), sv.red_blocks, 0);
  if (threadIdx.x == 0) {
    sv.st->ep = -0.5 * kCoulomb * e;
    sv.st->finalized = 1;
  }
}

// start of a solve: fresh SolveState of every system of the launch
// (warm[s] != 0: mu already holds the starting dipoles); a kernel rather
// than a host copy, so that it can sit in a captured sequence
__global__ void k_init_state(const SysView* views, const int* sys_ids, int nsys, double poleps, int politer,
                             const int* warm, const int* req) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  if (req && !req[q]) return;  // only the requested systems start a solve
  SolveState* st = views[sys_ids[q]].st;
  SolveState z;
  z.ggold[0] = z.ggold[1] = z.gg[0] = z.gg[1] = z.ggnew[0] = z.ggnew[1] = 0.0;
  z.resnrm = 0.0;
  z.poleps = poleps;
  z.ep = 0.0;
  z.niter = 0;
  z.it_cur = 0;
  z.politer = politer;
  z.conv = kRunning;
  z.n_indef = 0;
  z.warm = warm ? warm[q] : 0;
  z.finalized = 0;
  z.in_setup = req ? 1 : 0;
  z.stale = 0;
  *st = z;
}

// end of the set-up phase: the systems set up in this launch iterate from now
__global__ void k_end_setup(const SysView* views, const int* sys_ids, int nsys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  SolveState* st = views[sys_ids[q]].st;
  if (st->in_setup) st->in_setup = 0;
}

// caller-order coordinates (and multipoles, for serving systems) into the
// sorted slots, with the skin check of the block-union lists (stale: the
// host rebuilds the lists and repeats the set-up)
__global__ void __launch_bounds__(64) k_load_coords(const SysView* views, const Work* work, int gate) {
  const Work w = work[blockIdx.x];
  if (w.sys < 0) return;
  const SysView& sv = views[w.sys];
  if (gated_out(sv.st, gate)) return;
  const int i = w.blk * kBlock + threadIdx.x;
  if (i >= sv.n) return;
  const int n = sv.n;
  const size_t np = sv.npad;
  const int a = sv.perm[i];
  const double x = sv.raw[a], y = sv.raw[(size_t)n + a], z = sv.raw[2 * (size_t)n + a];
  const_cast<double*>(sv.x)[i] = x;
  const_cast<double*>(sv.y)[i] = y;
  const_cast<double*>(sv.z)[i] = z;
  const double dx = x - sv.refu[i], dy = y - sv.refu[np + i], dz = z - sv.refu[2 * np + i];
  if (dx * dx + dy * dy + dz * dz > sv.lim_u2) sv.st->stale = 1;
  if (sv.raw_rpole && !sv.frames) {
    // the 10 used components of rpole(13) (q, d, Qxx Qxy Qxz Qyy Qyz Qzz);
    // systems with local frames rotate theirs after this pass
    // (k_rotpole_batch)
    const GLOBAL double* r = gp(sv.raw_rpole) + (size_t)a * 13;
    GLOBAL double* RP = gw(const_cast<double*>(sv.rpole));
    RP[0 * np + i] = r[0];
    RP[1 * np + i] = r[1];
    RP[2 * np + i] = r[2];
    RP[3 * np + i] = r[3];
    RP[4 * np + i] = r[4];
    RP[5 * np + i] = r[5];
    RP[6 * np + i] = r[6];
    RP[7 * np + i] = r[8];
    RP[8 * np + i] = r[9];
    RP[9 * np + i] = r[12];
  }
}

// bin counters of the brick bins, zeroed before k_pme_prep counts into them
__global__ void k_zero_bins(const SysView* views, const Work* bwork, int nwork, int gate) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nwork) return;
  const Work w = bwork[q];
  if (w.sys < 0) return;
  if (gated_out(views[w.sys].st, gate)) return;
  views[w.sys].bin_count[w.blk] = 0;
}

// solve states and error flags of every system of the launch into out[]
__global__ void k_collect(const SysView* views, const int* sys_ids, int nsys, StateOut* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nsys) return;
  const SysView& sv = views[sys_ids[q]];
  out[q].st = *sv.st;
  out[q].err = *sv.err;
  out[q].pad = 0;
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
// a grid of the pair type to FP64 and back (the FP64 grids of the entry points)
template <typename real>
__global__ void k_grid_widen(double* dst, const real* src, size_t count) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) dst[q] = (double)src[q];
}
template <typename real>
__global__ void k_grid_narrow(real* dst, const double* src, size_t count) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) dst[q] = (real)src[q];
}

}  // namespace thp
