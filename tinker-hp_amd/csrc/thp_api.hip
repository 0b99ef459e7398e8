// libthp host side: context, systems, the path entry points of the C ABI
// and the PCG driver.  Kernels live in thp_kernels.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>

#include "thp_internal.h"
#include "thp_kernels.h"
#include "thp_forces.h"

using namespace thp;

// Scratch of thp_epolar1_noalloc: the forces and torques before they are
// scaled and gathered ([6][npad] doubles, npad <= THP_NOALLOC_MAX_ATOMS),
// as a file-scope device array, so that the entry point allocates nothing.
// Each device holds its own copy until the process ends; host code uses
// only the per-device address thp_init resolves (ctx->scratch), never the
// array's name.
static __device__ __align__(256) double scratch[(size_t)6 * THP_NOALLOC_MAX_ATOMS];

thp_status thp::fail(thp_ctx* ctx, thp_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

size_t thp::grid_points(const thp_sys* s) { return (size_t)s->nfft[0] * s->nfft[1] * s->nfft[2]; }

// Stream-ordered entry points: the library's stream first waits for the
// work the caller queued on `caller`, and at the end `caller` waits for
// the library's work, so nothing blocks the host.  A null caller makes the
// call synchronous.
thp_status thp::order_in(thp_ctx* ctx, hipStream_t caller) {
  if (!caller || caller == ctx->stream) return THP_OK;
  hipError_t e = hipEventRecord(ctx->ev_in, caller);
  if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->ev_in, 0);
  if (e != hipSuccess) return fail(ctx, THP_ERR_HIP, "ordering behind the caller's stream: %s", hipGetErrorString(e));
  return THP_OK;
}

thp_status thp::order_out(thp_ctx* ctx, hipStream_t caller) {
  hipError_t e = hipSuccess;
  if (!caller) e = hipStreamSynchronize(ctx->stream);
  else if (caller != ctx->stream) {
    e = hipEventRecord(ctx->ev_out, ctx->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(caller, ctx->ev_out, 0);
  }
  if (e != hipSuccess) return fail(ctx, THP_ERR_HIP, "ordering the caller's stream behind the call: %s", hipGetErrorString(e));
  return THP_OK;
}


namespace {

#define HIPCHK(ctx, call)                                                                 \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((ctx), THP_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                   \
  } while (0)

#define RET(call)                      \
  do {                                 \
    thp_status r_ = (call);            \
    if (r_ != THP_OK) return r_;       \
  } while (0)

template <typename T>
thp_status dalloc(thp_sys* s, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    return fail(s->ctx, THP_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  s->allocs[q] = bytes;
  s->bytes += (int64_t)bytes;
  *p = static_cast<T*>(q);
  return THP_OK;
}

// releases an allocation of the system and takes it off the footprint
void dfree(thp_sys* s, void* p) {
  if (!p) return;
  auto it = s->allocs.find(p);
  if (it != s->allocs.end()) {
    s->bytes -= (int64_t)it->second;
    s->allocs.erase(it);
  }
  (void)hipFree(p);
}

// ------------------------------------------------ PME tables (host, FP64)
// bspline (kewald.f:720-745)
void bspline_coef(double x, int n, double* c) {
  c[0] = 1.0 - x;
  c[1] = x;
  for (int k = 3; k <= n; ++k) {
    const double denom = 1.0 / double(k - 1);
    c[k - 1] = x * c[k - 2] * denom;
    for (int i = 1; i <= k - 2; ++i)
      c[k - i - 1] = ((x + double(i)) * c[k - i - 2] + (double(k - i) - x) * c[k - i - 1]) * denom;
    c[0] = (1.0 - x) * c[0] * denom;
  }
}

// moduli + dftmod with the Euler-spline fix and zeta factor (kewald.f:672-812)
std::vector<double> bsmod_table(int nfft, int order) {
  double arr[16];
  bspline_coef(0.0, order, arr);
  std::vector<double> bsarray(std::max(nfft, order + 2), 0.0), bsmod(nfft);
  for (int i = 2; i <= order + 1; ++i) bsarray[i - 1] = arr[i - 2];
  const double pi = M_PI;
  const double factor = 2.0 * pi / nfft;
  for (int i = 0; i < nfft; ++i) {
    double s1 = 0, s2 = 0;
    for (int j = 0; j < nfft; ++j) {
      const double arg = factor * double((long)i * j);
      s1 += bsarray[j] * std::cos(arg);
      s2 += bsarray[j] * std::sin(arg);
    }
    bsmod[i] = s1 * s1 + s2 * s2;
  }
  const double eps = 1.0e-7;
  if (bsmod[0] < eps) bsmod[0] = 0.5 * bsmod[1];
  for (int i = 1; i < nfft - 1; ++i)
    if (bsmod[i] < eps) bsmod[i] = 0.5 * (bsmod[i - 1] + bsmod[i + 1]);
  if (bsmod[nfft - 1] < eps) bsmod[nfft - 1] = 0.5 * bsmod[nfft - 2];
  for (int i = 0; i < nfft; ++i) {
    int k = i;
    if (i + 1 > nfft / 2) k -= nfft;
    double zeta = 1.0;
    if (k != 0) {
      double s1 = 1.0, s2 = 1.0;
      const double f = pi * double(k) / double(nfft);
      for (int j = 1; j <= 50; ++j) {
        const double a = f / (f + pi * j), b = f / (f - pi * j);
        s1 += std::pow(a, order) + std::pow(b, order);
        s2 += std::pow(a, 2 * order) + std::pow(b, 2 * order);
      }
      zeta = s2 / s1;
    }
    bsmod[i] *= zeta * zeta;
  }
  return bsmod;
}

// scalar-sum table qfac (newinduce_pme.f:1790-1844), periodic boxes
std::vector<double> qfac_table(const double box[3], const int nf[3], double aewald, int order) {
  const auto b1 = bsmod_table(nf[0], order), b2 = bsmod_table(nf[1], order), b3 = bsmod_table(nf[2], order);
  std::vector<double> q((size_t)nf[0] * nf[1] * nf[2], 0.0);
  const double pterm = (M_PI / aewald) * (M_PI / aewald);
  const double volterm = M_PI * box[0] * box[1] * box[2];
  const int h1n = (nf[0] + 1) / 2, h2n = (nf[1] + 1) / 2, h3n = (nf[2] + 1) / 2;
  for (int k3 = 0; k3 < nf[2]; ++k3) {
    const int m3 = k3 + 1 > h3n ? k3 - nf[2] : k3;
    const double r3 = m3 / box[2];
    for (int k2 = 0; k2 < nf[1]; ++k2) {
      const int m2 = k2 + 1 > h2n ? k2 - nf[1] : k2;
      const double r2 = m2 / box[1];
      for (int k1 = 0; k1 < nf[0]; ++k1) {
        const int m1 = k1 + 1 > h1n ? k1 - nf[0] : k1;
        if (m1 == 0 && m2 == 0 && m3 == 0) continue;
        const double r1 = m1 / box[0];
        const double hsq = r1 * r1 + r2 * r2 + r3 * r3;
        const double term = -pterm * hsq;
        double e = 0.0;
        if (term > -50.0) e = std::exp(term) / (volterm * hsq * b1[k1] * b2[k2] * b3[k3]);
        q[((size_t)k3 * nf[1] + k2) * nf[0] + k1] = e;
      }
    }
  }
  return q;
}

// cart_to_frac / frac_to_cart (pmestuff.f:1122-1298) for an orthorhombic cell
void conversion_mats(const double box[3], const int nf[3], double ctf[10][10], double ftc[10][10]) {
  double a[3][3] = {};
  for (int k = 0; k < 3; ++k) a[k][k] = nf[k] / box[k];
  static const int qi1[6] = {0, 1, 2, 0, 0, 1};
  static const int qi2[6] = {0, 1, 2, 1, 2, 2};
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j) ctf[i][j] = ftc[i][j] = 0.0;
  ctf[0][0] = ftc[0][0] = 1.0;
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) ctf[i][j] = ftc[i][j] = a[i - 1][j - 1];
  for (int i1 = 0; i1 < 3; ++i1) {
    const int k = qi1[i1];
    for (int i2 = 0; i2 < 6; ++i2) {
      const int i = qi1[i2], j = qi2[i2];
      ctf[i1 + 4][i2 + 4] = a[k][i] * a[k][j];
      ftc[i1 + 4][i2 + 4] = i2 < 3 ? a[k][i] * a[k][i] : 2.0 * a[k][i] * a[k][j];
    }
  }
  for (int i1 = 3; i1 < 6; ++i1) {
    const int k = qi1[i1], m = qi2[i1];
    for (int i2 = 0; i2 < 6; ++i2) {
      const int i = qi1[i2], j = qi2[i2];
      ctf[i1 + 4][i2 + 4] = a[k][i] * a[m][j] + a[k][j] * a[m][i];
      ftc[i1 + 4][i2 + 4] = i2 < 3 ? a[k][i] * a[m][i] : a[k][i] * a[m][j] + a[m][i] * a[k][j];
    }
  }
}

// Spatial order: recursive median bisection of the wrapped coordinates
// along the longest extent, with left halves holding a multiple of 64 atoms,
// so that every 64-atom block (one wavefront) is a compact brick.  Ties
// are broken by the original index, so the order is deterministic.
void bisect(std::vector<int>& idx, size_t lo, size_t hi, const std::vector<double>& w) {
  const size_t n = hi - lo;
  if (n <= (size_t)kBlock) return;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (size_t q = lo; q < hi; ++q)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], w[3 * idx[q] + d]);
      mx[d] = std::max(mx[d], w[3 * idx[q] + d]);
    }
  int d = 0;
  for (int e = 1; e < 3; ++e)
    if (mx[e] - mn[e] > mx[d] - mn[d]) d = e;
  size_t nl = (size_t)kBlock * std::max<size_t>(1, (n + kBlock) / (2 * kBlock));
  auto less = [&](int a, int b) { return w[3 * a + d] < w[3 * b + d] || (w[3 * a + d] == w[3 * b + d] && a < b); };
  std::nth_element(idx.begin() + lo, idx.begin() + lo + nl, idx.begin() + hi, less);
  bisect(idx, lo, lo + nl, w);
  bisect(idx, lo + nl, hi, w);
}

std::vector<int> spatial_order(int n, const double* x, const double* y, const double* z, const double box[3]) {
  std::vector<double> w((size_t)3 * n);
  for (int i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; ++d) {
      const double v = p[d] - box[d] * std::floor(p[d] / box[d]);
      w[3 * i + d] = std::isfinite(v) ? v : 0.0;  // ordering only; keeps the comparator a strict order
    }
  }
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  bisect(perm, 0, (size_t)n, w);
  return perm;
}

// Cubic Hermite table of the Ewald functions of the recomputed matvec
// coefficients, in x = aewald r (dampewald, damping.f:17-56, at aewald = 1):
//   G1(x) = r^3 bn1 = erfc(x) + (2/sqrt(pi)) x exp(-x^2)
//   G2(x) = r^5 bn2 = 3 G1(x) + (4/sqrt(pi)) x^3 exp(-x^2)
// with G1' = -(4/sqrt(pi)) x^2 exp(-x^2) and G2' = -(8/sqrt(pi)) x^4 exp(-x^2).
// Interval j (x in [j h, (j+1) h), h = 1/kGtabScale) holds the coefficients
// of p(f) = c0 + c1 f + c2 f^2 + c3 f^3, f = x/h - j, for G1 and then G2,
// computed in FP64 and stored in FP32 (largest error 1.5e-7 on G2).
std::vector<float> ewald_gtab() {
  const double h = 1.0 / kGtabScale, s = 2.0 / std::sqrt(M_PI);
  auto g1 = [&](double x) { return std::erfc(x) + s * x * std::exp(-x * x); };
  auto g2 = [&](double x) { return 3.0 * g1(x) + 2.0 * s * x * x * x * std::exp(-x * x); };
  auto d1 = [&](double x) { return -2.0 * s * x * x * std::exp(-x * x); };
  auto d2 = [&](double x) { return -4.0 * s * x * x * x * x * std::exp(-x * x); };
  std::vector<float> t((size_t)kGtabN * 8);
  for (int j = 0; j < kGtabN; ++j) {
    const double x0 = j * h, x1 = x0 + h;
    for (int f = 0; f < 2; ++f) {
      const double y0 = f ? g2(x0) : g1(x0), y1 = f ? g2(x1) : g1(x1);
      const double e0 = (f ? d2(x0) : d1(x0)) * h, e1 = (f ? d2(x1) : d1(x1)) * h;
      float* c = &t[(size_t)j * 8 + 4 * f];
      c[0] = float(y0);
      c[1] = float(e0);
      c[2] = float(3.0 * (y1 - y0) - 2.0 * e0 - e1);
      c[3] = float(2.0 * (y0 - y1) + e0 + e1);
    }
  }
  return t;
}

// spreading brick edge: the largest B in bmax..4 (bmax <= 8) with K % B == 0
// or >= 4, so every brick (the last one included) is at least 4 points
// wide; when no B up to bmax qualifies, the largest one up to 8
int pick_brick(int K, int bmax) {
  for (int B = bmax; B >= 4; --B)
    if (K % B == 0 || K % B >= 4) return B;
  for (int B = 8; B > bmax; --B)
    if (K % B == 0 || K % B >= 4) return B;
  return 4;
}

thp_status upload_view(thp_sys* s) {
  ++s->view_version;
  HIPCHK(s->ctx, hipMemcpyAsync(s->d_view, &s->view, sizeof(SysView), hipMemcpyHostToDevice, s->ctx->stream));
  return THP_OK;
}

thp_status sync(thp_sys* s) {
  HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return THP_OK;
}

thp_status check_err_flag(thp_sys* s) {
  int err = 0;
  HIPCHK(s->ctx, hipMemcpyAsync(&err, s->view.err, sizeof(int), hipMemcpyDeviceToHost, s->ctx->stream));
  RET(sync(s));
  if (err) return fail(s->ctx, err, "device reported error %d (neighbor-list capacity)", err);
  return THP_OK;
}

// ---------------------------------------------------------- kernel launch
inline dim3 g1(int n) { return dim3((unsigned)n); }

thp_status launch_check(thp_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, THP_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
  return THP_OK;
}

// direct permanent field + the pair coefficients and compacted lists that
// the matvec of the same coordinates uses.  The pass carries ~300 flops and
// two exponentials per entry: eight wavefronts per block beat the matvec's
// tuned shape.
thp_status pair_efld0(thp_ctx* ctx, const Batch& b, int gate) {
  constexpr int S = kEfld0Segs;
  if (b.prec == THP_PREC_FP64) {
    hipLaunchKernelGGL((k_efld0<double, S, false>), g1(b.n_blk), dim3(64 * S), 0, ctx->stream, b.d_views, b.d_blk, gate);
  } else {
    // systems with the Ewald table and systems without take separate
    // kernels, each skipping the other kind (as the matvec)
    const int kinds = b.rc ? b.rc : 3;
    if (kinds & 1)
      hipLaunchKernelGGL((k_efld0<float, S, true>), g1(b.n_blk), dim3(64 * S), 0, ctx->stream, b.d_views, b.d_blk, gate);
    if (kinds & 2)
      hipLaunchKernelGGL((k_efld0<float, S, false>), g1(b.n_blk), dim3(64 * S), 0, ctx->stream, b.d_views, b.d_blk, gate);
  }
  hipLaunchKernelGGL(k_efld0_scaled, g1(b.n_blk), dim3(64), 0, ctx->stream, b.d_views, b.d_blk, gate);
  return launch_check(ctx, "k_efld0");
}

// spreading tiles of the home atoms (two wavefronts per brick), then the
// merge into the grid
template <typename real, int KIND>
void launch_spread(thp_ctx* ctx, const Batch& b, int vin, int frac, int gate) {
  hipStream_t st = ctx->stream;
  // one tile count for every launch: the order of a grid point's sums, and
  // so a request's result, does not depend on the batch it shares
  hipLaunchKernelGGL((k_spread_tile<real, KIND, 2>), g1(b.n_brick), dim3(128), 0, st, b.d_views, b.d_brick, vin, frac,
                     gate);
  hipLaunchKernelGGL((k_spread_merge<real, KIND>), g1(b.n_conv), dim3(256), 0, st, b.d_views, b.d_conv, gate);
}

// grid_uind_site
thp_status spread_uind(thp_ctx* ctx, const Batch& b, int vin, int frac, int gate) {
  if (b.prec == THP_PREC_FP64) launch_spread<double, 0>(ctx, b, vin, frac, gate);
  else launch_spread<float, 0>(ctx, b, vin, frac, gate);
  return launch_check(ctx, "k_spread_tile<uind>");
}

// grid_mpole_site (real part, imaginary part cleared)
thp_status spread_mpole(thp_ctx* ctx, const Batch& b, int gate) {
  if (b.prec == THP_PREC_FP64) launch_spread<double, 1>(ctx, b, 0, 0, gate);
  else launch_spread<float, 1>(ctx, b, 0, 0, gate);
  return launch_check(ctx, "k_spread_tile<mpole>");
}

// the grids' convolution with qfac: x and y forward passes, the fused z
// pass (forward, qfac, backward), y and x backward passes
// (spec: one system's batch only; its spectrum is kept there, see MODE 3)
template <int P>
void launch_plane64_p(thp_ctx* ctx, const Batch& b, int gate, int mode) {
  if (mode == 0) hipLaunchKernelGGL((k_fft_plane64<P, 0>), g1(b.n_fft), dim3(64 * P), 0, ctx->stream, b.d_views, b.d_fft, gate);
  else hipLaunchKernelGGL((k_fft_plane64<P, 1>), g1(b.n_fft), dim3(64 * P), 0, ctx->stream, b.d_views, b.d_fft, gate);
}
void launch_plane64(thp_ctx* ctx, const Batch& b, int gate, int mode) {
  if (ctx->fft_pp == 4) launch_plane64_p<4>(ctx, b, gate, mode);
  else if (ctx->fft_pp == 2) launch_plane64_p<2>(ctx, b, gate, mode);
  else launch_plane64_p<1>(ctx, b, gate, mode);
}
template <int P>
void launch_z64_p(thp_ctx* ctx, const Batch& b, int gate, void* spec) {
  if (spec) hipLaunchKernelGGL((k_fft_z64<P, 3>), g1(b.n_fft), dim3(64), 0, ctx->stream, b.d_views, b.d_fft, gate, spec);
  else hipLaunchKernelGGL((k_fft_z64<P, 2>), g1(b.n_fft), dim3(64), 0, ctx->stream, b.d_views, b.d_fft, gate, spec);
}
void launch_z64(thp_ctx* ctx, const Batch& b, int gate, void* spec) {
  if (ctx->fft_zp == 4) launch_z64_p<4>(ctx, b, gate, spec);
  else if (ctx->fft_zp == 2) launch_z64_p<2>(ctx, b, gate, spec);
  else launch_z64_p<1>(ctx, b, gate, spec);
}

template <typename real>
void launch_convolve(thp_ctx* ctx, const Batch& b, int gate, void* spec) {
  const dim3 g = g1(b.n_fft), t(256);
  hipStream_t st = ctx->stream;
  void* none = nullptr;
  auto zpass = [&]() {
    if (spec) hipLaunchKernelGGL((k_fft_axis<real, 2, 3>), g, t, 0, st, b.d_views, b.d_fft, gate, spec);
    else hipLaunchKernelGGL((k_fft_axis<real, 2, 2>), g, t, 0, st, b.d_views, b.d_fft, gate, none);
  };
  if constexpr (sizeof(real) == 4) {
    if (b.fft_r64 && ctx->reg_fft) {
      // 64^3 grids: the transforms in registers, three passes (P lanes per
      // line in the plane and z passes: ctx->fft_pp, ctx->fft_zp)
      launch_plane64(ctx, b, gate, 0);
      launch_z64(ctx, b, gate, spec);
      launch_plane64(ctx, b, gate, 1);
      return;
    }
  }
  if (b.fft_plane && ctx->plane_fft && sizeof(real) == 4) {
    // x and y fused per plane, then the z pass: three passes over the grid
    hipLaunchKernelGGL((k_fft_plane<real, 0, kPlaneThreads>), g, dim3(kPlaneThreads), b.fft_plane_lds, st, b.d_views, b.d_fft, gate);
    zpass();
    hipLaunchKernelGGL((k_fft_plane<real, 1, kPlaneThreads>), g, dim3(kPlaneThreads), b.fft_plane_lds, st, b.d_views, b.d_fft, gate);
    return;
  }
  hipLaunchKernelGGL((k_fft_axis<real, 0, 0>), g, t, 0, st, b.d_views, b.d_fft, gate, none);
  hipLaunchKernelGGL((k_fft_axis<real, 1, 0>), g, t, 0, st, b.d_views, b.d_fft, gate, none);
  zpass();
  hipLaunchKernelGGL((k_fft_axis<real, 1, 1>), g, t, 0, st, b.d_views, b.d_fft, gate, none);
  hipLaunchKernelGGL((k_fft_axis<real, 0, 1>), g, t, 0, st, b.d_views, b.d_fft, gate, none);
}

thp_status convolve(thp_ctx* ctx, const Batch& b, int gate, void* spec = nullptr) {
  if (b.prec == THP_PREC_FP64) launch_convolve<double>(ctx, b, gate, spec);
  else launch_convolve<float>(ctx, b, gate, spec);
  return launch_check(ctx, "k_fft_axis");
}

template <int MODE>
thp_status launch_gather_uind(thp_ctx* ctx, const Batch& b, int vin, int vout, int frac_out, int gate) {
  const dim3 g = g1(b.n_gat), t(64);
  if (b.real_only) {
    // the PCG steps of a short-range solve, without the reciprocal field
    if (b.prec == THP_PREC_FP64) hipLaunchKernelGGL((k_gather_tile<double, MODE, false>), g, t, 0, ctx->stream, b.d_views, b.d_gat, vin, vout, frac_out, gate);
    else hipLaunchKernelGGL((k_gather_tile<float, MODE, false>), g, t, 0, ctx->stream, b.d_views, b.d_gat, vin, vout, frac_out, gate);
  } else {
    if (b.prec == THP_PREC_FP64) hipLaunchKernelGGL((k_gather_tile<double, MODE>), g, t, 0, ctx->stream, b.d_views, b.d_gat, vin, vout, frac_out, gate);
    else hipLaunchKernelGGL((k_gather_tile<float, MODE>), g, t, 0, ctx->stream, b.d_views, b.d_gat, vin, vout, frac_out, gate);
  }
  return launch_check(ctx, "k_gather_tile");
}

// permanent field: ef = E_direct - grad phi_recip + self (newinduce_pme2.f:111-147)
thp_status field(thp_ctx* ctx, const Batch& b, double* cphi_dev, int add_to_ef, int do_direct, int gate,
                 void* spec = nullptr) {
  if (do_direct) RET(pair_efld0(ctx, b, gate));
  if (b.real_only && add_to_ef) return THP_OK;  // short range: the direct field alone
  RET(spread_mpole(ctx, b, gate));
  RET(convolve(ctx, b, gate, spec));
  if (b.prec == THP_PREC_FP64) hipLaunchKernelGGL((k_gather_mpole<double>), g1(b.n_blk), dim3(64), 0, ctx->stream, b.d_views, b.d_blk, cphi_dev, add_to_ef, gate);
  else hipLaunchKernelGGL((k_gather_mpole<float>), g1(b.n_blk), dim3(64), 0, ctx->stream, b.d_views, b.d_blk, cphi_dev, add_to_ef, gate);
  return launch_check(ctx, "k_gather_mpole");
}

// host array [n][ncomp] in caller order into a sorted device array
// [ncomp][npad] (ncomp <= 10), through the system's staging buffer
thp_status scatter_in(thp_sys* s, double* dst_dev, const double* host, int ncomp) {
  HIPCHK(s->ctx, hipMemcpyAsync(s->stage, host, sizeof(double) * s->n * ncomp, hipMemcpyHostToDevice, s->ctx->stream));
  hipLaunchKernelGGL(k_scatter_in, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, dst_dev, s->stage, s->iperm_dev, s->n, s->npad, ncomp);
  RET(launch_check(s->ctx, "k_scatter_in"));
  return sync(s);
}

// per-atom device array (sorted, [ncomp][npad]) into the pinned host stage
// in caller order ([n][ncomp]); returns once the copy has landed
thp_status gather_to_hstage(thp_sys* s, const double* src_dev, int ncomp) {
  const size_t len = (size_t)s->n * ncomp;
  if (len > s->hstage_len) {
    if (s->hstage) (void)hipHostFree(s->hstage);
    s->hstage = nullptr;
    s->hstage_len = 0;
    HIPCHK(s->ctx, hipHostMalloc((void**)&s->hstage, sizeof(double) * len, hipHostMallocDefault));
    s->hstage_len = len;
  }
  hipLaunchKernelGGL(k_gather_out, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, s->stage, src_dev, s->iperm_dev, s->n, s->npad, ncomp);
  RET(launch_check(s->ctx, "k_gather_out"));
  HIPCHK(s->ctx, hipMemcpyAsync(s->hstage, s->stage, sizeof(double) * len, hipMemcpyDeviceToHost, s->ctx->stream));
  return sync(s);
}

thp_status gather_out(thp_sys* s, double* host, const double* src_dev, int ncomp) {
  RET(gather_to_hstage(s, src_dev, ncomp));
  std::memcpy(host, s->hstage, sizeof(double) * s->n * ncomp);
  return THP_OK;
}

// bounding boxes and candidate block pairs (cut + buffer); valid until an
// atom moves more than buffer/2
thp_status build_block_lists(thp_sys* s) {
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_bbox, g1(s->nblk), dim3(64), 0, st, s->d_view, s->d_work);
  RET(launch_check(s->ctx, "k_bbox"));
  hipLaunchKernelGGL((k_blockpairs<false>), g1(s->nblk), dim3(256), 0, st, s->d_view, s->d_work, s->bcount, (int*)nullptr, 0);
  RET(launch_check(s->ctx, "k_blockpairs"));
  std::vector<int> bc(s->nblk);
  HIPCHK(s->ctx, hipMemcpyAsync(bc.data(), s->bcount, sizeof(int) * s->nblk, hipMemcpyDeviceToHost, st));
  RET(sync(s));
  const int maxbl = *std::max_element(bc.begin(), bc.end());
  if (maxbl > s->maxbl) {
    dfree(s, s->blist);
    s->blist = nullptr;
    RET(dalloc(s, &s->blist, (size_t)s->nblk * maxbl));
    s->maxbl = maxbl;
  }
  hipLaunchKernelGGL((k_blockpairs<true>), g1(s->nblk), dim3(256), 0, st, s->d_view, s->d_work, s->bcount, s->blist, s->maxbl);
  RET(launch_check(s->ctx, "k_blockpairs"));
  s->bx0 = s->hx;
  s->by0 = s->hy;
  s->bz0 = s->hz;
  s->list_valid = true;
  s->tight_valid = false;
  s->verlet_valid = false;
  return THP_OK;
}

// block-union lists for the current coordinates (cut + skin, unscaled
// pairs); single pass, capacities grown and the pass repeated on overflow
thp_status build_union(thp_sys* s) {
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_bbox, g1(s->nblk), dim3(64), 0, st, s->d_view, s->d_work);
  RET(launch_check(s->ctx, "k_bbox"));
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (s->view.umax == 0) {
      s->view.umax = 2048;
      s->view.lmax = 384;
      RET(dalloc(s, &s->view.ulist, (size_t)s->nblk * s->view.umax));
      RET(dalloc(s, &s->view.llist, (size_t)s->view.lmax * s->npad));
      RET(upload_view(s));
    }
    hipLaunchKernelGGL(k_union, g1(s->nblk), dim3(64 * kUnionWaves), 0, st, s->d_view, s->d_work, (const int*)s->bcount,
                       (const int*)s->blist, s->maxbl);
    RET(launch_check(s->ctx, "k_union"));
    int err = 0;
    HIPCHK(s->ctx, hipMemcpyAsync(&err, s->view.err, sizeof(int), hipMemcpyDeviceToHost, st));
    RET(sync(s));
    if (!err) {
      s->ux0 = s->hx;
      s->uy0 = s->hy;
      s->uz0 = s->hz;
      s->coef_valid = false;
      {
        // reference coordinates of the skin check of k_load_coords
        std::vector<double> r((size_t)3 * s->npad, 0.0);
        std::copy(s->hx.begin(), s->hx.end(), r.begin());
        std::copy(s->hy.begin(), s->hy.end(), r.begin() + s->npad);
        std::copy(s->hz.begin(), s->hz.end(), r.begin() + 2 * (size_t)s->npad);
        HIPCHK(s->ctx, hipMemcpy((void*)s->view.refu, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice));
      }
      // compacted-list storage for the longest list
      std::vector<int> lc(s->npad);
      HIPCHK(s->ctx, hipMemcpy(lc.data(), s->view.lcount, sizeof(int) * s->npad, hipMemcpyDeviceToHost));
      int cmax = 0;
      for (int q = 0; q < s->n; ++q) cmax = std::max(cmax, lc[q]);
      static_assert(32 % kPack == 0, "list storage holds whole lane packs");
      cmax = std::max(32, (cmax + 31) / 32 * 32);
      if (cmax > s->view.cmax) {
        dfree(s, s->view.tcoef);
        dfree(s, s->view.tlist);
        s->view.tcoef = nullptr;
        s->view.tlist = nullptr;
        RET(dalloc(s, &s->view.tlist, (size_t)cmax * s->npad));
        const size_t words = (size_t)cmax * s->npad * 2;
        if (s->view.rc_sqa > 0.0) {
          // indices only: the matvec recomputes the coefficients
        } else if (s->prec == THP_PREC_FP64) {
          double* t = nullptr;
          RET(dalloc(s, &t, words));
          s->view.tcoef = t;
        } else {
          float* t = nullptr;
          RET(dalloc(s, &t, words));
          s->view.tcoef = t;
        }
        s->view.cmax = cmax;
        RET(upload_view(s));
      }
      return THP_OK;
    }
    // grow to the sizes the pass counted, then repeat it
    std::vector<int> uc(s->nblk), lc(s->npad);
    HIPCHK(s->ctx, hipMemcpy(uc.data(), s->view.ucount, sizeof(int) * s->nblk, hipMemcpyDeviceToHost));
    HIPCHK(s->ctx, hipMemcpy(lc.data(), s->view.lcount, sizeof(int) * s->npad, hipMemcpyDeviceToHost));
    int umax = 0, lmax = 0;
    for (int b = 0; b < s->nblk; ++b) umax = std::max(umax, uc[b]);
    for (int q = 0; q < s->n; ++q) lmax = std::max(lmax, lc[q]);
    if (umax > 65535) return fail(s->ctx, THP_ERR_LIST, "block union of %d atoms exceeds 16-bit indexing", umax);
    dfree(s, s->view.ulist);
    dfree(s, s->view.llist);
    s->view.umax = std::max(s->view.umax, (umax + 63) / 64 * 64);
    s->view.lmax = std::max(s->view.lmax, (lmax + 31) / 32 * 32);
    RET(dalloc(s, &s->view.ulist, (size_t)s->nblk * s->view.umax));
    RET(dalloc(s, &s->view.llist, (size_t)s->view.lmax * s->npad));
    HIPCHK(s->ctx, hipMemsetAsync(s->view.err, 0, sizeof(int), st));
    RET(upload_view(s));
  }
  return fail(s->ctx, THP_ERR_LIST, "block-union list capacity");
}

thp_status ensure_pme(thp_sys* s) {
  if (!s->pme_valid) {
    RET(b_pme_prep(s->ctx, s->one, kGateNone));
    RET(check_err_flag(s));
    s->pme_valid = true;
  }
  return THP_OK;
}

}  // namespace

// ================================================ batched launch sequences
namespace thp {

thp_status ensure_lists(thp_sys* s) {
  if (!s->list_valid) RET(build_block_lists(s));
  if (!s->tight_valid) {
    RET(build_union(s));
    s->tight_valid = true;
  }
  return THP_OK;
}

// pair coefficients for the current coordinates (the permanent-field pass
// writes them; ef is rewritten as a side effect)
thp_status ensure_coef(thp_sys* s) {
  RET(ensure_lists(s));
  if (!s->coef_valid) {
    RET(pair_efld0(s->ctx, s->one, kGateNone));
    s->coef_valid = true;
  }
  return THP_OK;
}

// New coordinates (caller order).  Host work: sorted copy and the
// displacement checks that decide which lists stay valid.  With `pinned`
// (3*npad doubles of pinned host memory, untouched until the stream passes
// the copies) the upload is enqueued without waiting; otherwise it is
// synchronous.
// the dipoles must be a PCG or Jacobi/DIIS solve (full Ewald) of the
// current coordinates and multipoles
thp_status epolar1_check(thp_sys* s) {
  thp_ctx* ctx = s->ctx;
  if (!s->mu_valid)
    return fail(ctx, THP_ERR_STATE, "thp_epolar1: no solve of the current coordinates and multipoles");
  // a short-range solve has the short-range gradient (the real-space terms
  // within the short cutoff: epolar1c with use_polarshortreal,
  // epolar1.f:1347-1356, mode SHORTEWALD of switch.f:60-62), a full solve
  // the full one
  if (s->mu_short != (s->one.real_only != 0))
    return fail(ctx, THP_ERR_ARG, "thp_epolar1: the last solve's range is not the system's current range");
  return THP_OK;
}

// The gradient's launch sequence on the library's stream: the multipole
// grid's potential derivatives into cphi, the two induced grids of the
// solved dipoles, real-space and reciprocal forces and torques into F and T
// ([3][npad], sorted order, written for every atom), the frame forces of
// the torques, then F scaled to the gradient and T to kcal/mol.
// With vpart (virial_slots(s) x 6 doubles) and spec (two grid spectra of
// the system's real type) the virial shares are written too: the real-space
// pairs' (one slot per block), the reciprocal moments' (one per block), the
// frame forces' (one per 256 atoms, with frames) and the reciprocal k-sum's
// (one per kVirChunk grid points), in that order.
template <typename real>
thp_status launch_ep1_recip(thp_ctx* ctx, const Batch& b, const double* cphi, double* F, double* T, double* vpart,
                            int terms) {
  const dim3 g = g1(b.n_blk), t(64);
  hipStream_t hs = ctx->stream;
  if (vpart) hipLaunchKernelGGL((k_ep1_recip<real, true, kTermsAll>), g, t, 0, hs, b.d_views, b.d_blk, cphi, F, T, vpart);
  else if (terms == kTermsAll) hipLaunchKernelGGL((k_ep1_recip<real, false, kTermsAll>), g, t, 0, hs, b.d_views, b.d_blk, cphi, F, T, vpart);
  else if (terms == kTermsField) hipLaunchKernelGGL((k_ep1_recip<real, false, kTermsField>), g, t, 0, hs, b.d_views, b.d_blk, cphi, F, T, vpart);
  else hipLaunchKernelGGL((k_ep1_recip<real, false, kTermsDipole>), g, t, 0, hs, b.d_views, b.d_blk, cphi, F, T, vpart);
  return launch_check(ctx, "k_ep1_recip");
}

thp_status epolar1_launch(thp_sys* s, double* F, double* T, double* vpart = nullptr, void* spec = nullptr,
                          int terms = kTermsAll, bool finish = true) {
  thp_ctx* ctx = s->ctx;
  hipStream_t hs = ctx->stream;
  const Batch& b = s->one;
  const bool vir = vpart != nullptr;
  // short range: the real-space terms within the short cutoff (cut2) only
  const bool recip = !b.real_only;
  const size_t gbytes = grid_points(s) * 2 * (s->prec == THP_PREC_FP64 ? sizeof(double) : sizeof(float));
  void* spec_m = vir ? spec : nullptr;
  void* spec_u = vir ? static_cast<char*>(spec) + gbytes : nullptr;
  if (recip) {
    RET(field(ctx, b, s->cphi, 0, 0, kGateNone, spec_m));
    RET(spread_uind(ctx, b, V_MU, 0, kGateNone));
    RET(convolve(ctx, b, kGateNone, spec_u));
  }
  const size_t nb = (size_t)s->nblk, nt = (size_t)(s->n + 255) / 256;
  double* v_real = vpart;
  double* v_recip = vir ? vpart + 6 * nb : nullptr;
  double* v_torque = vir ? vpart + 12 * nb : nullptr;
  double* v_ksum = vir ? vpart + 12 * nb + 6 * nt : nullptr;
  // the virial takes every term; the truncated CG's gradient one mask at a time
  if (vir && terms != kTermsAll) return fail(ctx, THP_ERR_ARG, "epolar1_launch: the virial needs every term");
  const dim3 gb = g1(b.n_blk);
  if (vir) hipLaunchKernelGGL((k_ep1_real<4, true, kTermsAll>), gb, dim3(256), 0, hs, b.d_views, b.d_blk, F, T, v_real);
  else if (terms == kTermsAll) hipLaunchKernelGGL((k_ep1_real<4, false, kTermsAll>), gb, dim3(256), 0, hs, b.d_views, b.d_blk, F, T, v_real);
  else if (terms == kTermsField) hipLaunchKernelGGL((k_ep1_real<4, false, kTermsField>), gb, dim3(256), 0, hs, b.d_views, b.d_blk, F, T, v_real);
  else hipLaunchKernelGGL((k_ep1_real<4, false, kTermsDipole>), gb, dim3(256), 0, hs, b.d_views, b.d_blk, F, T, v_real);
  RET(launch_check(ctx, "k_ep1_real"));
  if (recip) {
    if (s->prec == THP_PREC_FP64) RET(launch_ep1_recip<double>(ctx, b, s->cphi, F, T, v_recip, terms));
    else RET(launch_ep1_recip<float>(ctx, b, s->cphi, F, T, v_recip, terms));
  }
  const dim3 ga((s->n + 255) / 256);
  if (!finish) return THP_OK;  // the raw forces and torques of the chosen terms
  if (s->has_frames) {
    if (vir)
      hipLaunchKernelGGL((k_ep1_torque<true>), ga, dim3(256), 0, hs, s->d_view, s->d_frames, s->d_tinv_off, s->d_tinv, T, F, v_torque);
    else
      hipLaunchKernelGGL((k_ep1_torque<false>), ga, dim3(256), 0, hs, s->d_view, s->d_frames, s->d_tinv_off, s->d_tinv, T, F, v_torque);
    RET(launch_check(ctx, "k_ep1_torque"));
  }
  if (vir && recip) {
    const dim3 gk((unsigned)((grid_points(s) + kVirChunk - 1) / kVirChunk));
    if (s->prec == THP_PREC_FP64)
      hipLaunchKernelGGL((k_vir_ksum<double>), gk, dim3(256), 0, hs, s->d_view, spec_m, spec_u, v_ksum);
    else
      hipLaunchKernelGGL((k_vir_ksum<float>), gk, dim3(256), 0, hs, s->d_view, spec_m, spec_u, v_ksum);
    RET(launch_check(ctx, "k_vir_ksum"));
  }
  hipLaunchKernelGGL(k_ep1_finish, ga, dim3(256), 0, hs, s->n, s->npad, F, T, -1.0);
  return launch_check(ctx, "k_ep1_finish");
}

// Gradient of the truncated-CG energy (epolar1tcg, epolar1tcg.f:16-119;
// epolar1tcg2.f): E_TCG = -f/2 <mu_d, E_p> with mu_d an explicit function of
// the permanent fields and of A = 1/alpha - T.  The adjoint of that function
// (reverse-mode differentiation of the truncated CG's steps, in FP64 on the
// host, A applied on the device as the solve applies it) gives
//   dE = <efb, dE_d> + <Epb, dE_p> - sum_k <lam_k, dT v_k>,
// and the device's gradient kernels evaluate the first two terms as the
// permanent-field terms with the dipole sets (u_d, u_p) = (Epb, efb) and
// each pair as the dipole term with (u_d, u_p) = (v_k, lam_k).  Returns the
// forces -dE/dr at fixed orientations and the torques ([3][npad] each,
// sorted order, kcal/mol/A and kcal/mol).
thp_status tcg_apply(thp_sys* s, int vin);

thp_status epolar1_tcg_terms(thp_sys* s, std::vector<double>& Fo, std::vector<double>& To) {
  thp_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const size_t np = s->npad;
  const int n = s->n;
  typedef std::vector<double> Vec;  // [3][np], rhs d
  auto down = [&](const double* dev, size_t comp0, Vec& out) -> thp_status {
    out.assign(3 * np, 0.0);
    HIPCHK(ctx, hipMemcpyAsync(out.data(), dev + comp0 * np, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, st));
    return sync(s);
  };
  // the solve's fields (total: direct, reciprocal and self) and dipoles
  Vec efd, Ep, alpha(3 * np, 0.0);
  RET(down(s->view.ef, 0, efd));
  RET(down(s->view.ef, 3, Ep));
  {
    std::vector<double> pol(np);
    HIPCHK(ctx, hipMemcpy(pol.data(), s->view.pol, sizeof(double) * np, hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < n; ++i) alpha[c * np + i] = pol[i];
  }
  std::vector<double> mu_keep(6 * np);
  HIPCHK(ctx, hipMemcpy(mu_keep.data(), s->view.mu, sizeof(double) * 6 * np, hipMemcpyDeviceToHost));
  // A v on the device (both right-hand sides carry v; A is the same)
  const double term = s->view.term_self;
  auto A1 = [&](const Vec& v, Vec& out) -> thp_status {
    std::vector<double> w(6 * np);
    std::copy(v.begin(), v.end(), w.begin());
    std::copy(v.begin(), v.end(), w.begin() + 3 * np);
    HIPCHK(ctx, hipMemcpy(s->view.pp, w.data(), sizeof(double) * 6 * np, hipMemcpyHostToDevice));
    RET(tcg_apply(s, V_PP));
    Vec hr, h;
    RET(down(s->view.hreal, 0, hr));
    RET(down(s->view.h, 0, h));
    out.assign(3 * np, 0.0);
    for (size_t q = 0; q < 3 * np; ++q) out[q] = hr[q] + h[q] - term * v[q];
    return THP_OK;
  };
  auto dot = [&](const Vec& a, const Vec& b) {
    double t = 0.0;
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < n; ++i) t += a[c * np + i] * b[c * np + i];
    return t;
  };
  auto axpy = [&](Vec& y, double a, const Vec& x) {
    for (size_t q = 0; q < y.size(); ++q) y[q] += a * x[q];
  };
  auto mulal = [&](const Vec& v) {
    Vec o(v.size());
    for (size_t q = 0; q < v.size(); ++q) o[q] = alpha[q] * v[q];
    return o;
  };
  const bool prec = s->mu_tcg[1], guess = s->mu_tcg[2], peek = s->mu_tcg[3];
  const int order = s->mu_tcg[0];
  const double om = s->mu_tcg_omega, f = kCoulombFp;
  auto P = [&](const Vec& v) { return prec ? mulal(v) : v; };
  auto Qm = [&](const Vec& v) { return prec ? v : mulal(v); };
  // forward (right-hand side d)
  Vec m0, q0, r0;
  if (guess) {
    m0 = mulal(efd);
    RET(A1(m0, q0));
    r0 = efd;
    axpy(r0, -1.0, q0);
  } else {
    r0 = efd;
  }
  const Vec z0 = P(r0);
  Vec w1, w2, y2;
  RET(A1(z0, w1));
  const Vec y1 = P(w1);
  const double n0 = dot(r0, z0), t1 = dot(z0, w1), t4 = n0 / t1;
  double t2 = 0, np1 = 0, s3 = 0, t3 = 0, num = 0, g1 = 0, cz = t4, cy = 0;
  Vec p2;
  if (order == 2) {
    RET(A1(y1, w2));
    y2 = P(w2);
    np1 = dot(w1, y1);
    t2 = n0 * np1 / (t1 * t1);
    p2 = y1;
    for (auto& x : p2) x *= t2;
    axpy(p2, -t4, y2);
    s3 = dot(w1, p2);
    t3 = t1 * s3;
    num = t1 * t1 - n0 * np1;
    g1 = num / t3;
    cz = g1 * t2 + t4;
    cy = -g1 * t4;
  }
  Vec mu(3 * np, 0.0);
  axpy(mu, cz, z0);
  if (order == 2) axpy(mu, cy, y1);
  if (guess) axpy(mu, 1.0, m0);
  if (peek) {
    axpy(mu, om, Qm(z0));
    axpy(mu, -om * t4, Qm(y1));
    if (order == 2) {
      axpy(mu, -om * g1 * t2, Qm(y1));
      axpy(mu, om * g1 * t4, Qm(y2));
    }
  }
  // reverse
  Vec mub(3 * np, 0.0), Epb(3 * np, 0.0);
  axpy(mub, -0.5 * f, Ep);
  axpy(Epb, -0.5 * f, mu);
  Vec z0b(3 * np, 0.0), y1b(3 * np, 0.0), y2b(3 * np, 0.0), w1b(3 * np, 0.0), w2b(3 * np, 0.0), r0b(3 * np, 0.0);
  double t4b = 0, g1b = 0, t2b = 0, t1b = 0, n0b = 0, np1b = 0;
  axpy(z0b, cz, mub);
  if (order == 2) axpy(y1b, cy, mub);
  const double czb = dot(mub, z0), cyb = order == 2 ? dot(mub, y1) : 0.0;
  if (peek) {
    const Vec qm = Qm(mub);
    axpy(z0b, om, qm);
    axpy(y1b, -om * t4, qm);
    t4b += -om * dot(qm, y1);
    if (order == 2) {
      axpy(y1b, -om * g1 * t2, qm);
      axpy(y2b, om * g1 * t4, qm);
      g1b += -om * t2 * dot(qm, y1) + om * t4 * dot(qm, y2);
      t2b += -om * g1 * dot(qm, y1);
      t4b += om * g1 * dot(qm, y2);
    }
  }
  std::vector<std::pair<Vec, Vec>> pairs;  // (lam, v)
  if (order == 2) {
    g1b += czb * t2 - cyb * t4;
    t2b += czb * g1;
    t4b += czb - cyb * g1;
    t1b += g1b * 2.0 * t1 / t3;
    n0b += -g1b * np1 / t3;
    np1b += -g1b * n0 / t3;
    const double t3b = -g1b * num / (t3 * t3);
    t1b += t3b * s3;
    const double s3b = t3b * t1;
    axpy(w1b, s3b, p2);
    Vec p2b = w1;
    for (auto& x : p2b) x *= s3b;
    t2b += dot(p2b, y1);
    axpy(y1b, t2, p2b);
    t4b += -dot(p2b, y2);
    axpy(y2b, -t4, p2b);
    n0b += t2b * np1 / (t1 * t1);
    np1b += t2b * n0 / (t1 * t1);
    t1b += -2.0 * t2b * n0 * np1 / (t1 * t1 * t1);
    axpy(w1b, np1b, y1);
    axpy(y1b, np1b, w1);
    axpy(w2b, 1.0, P(y2b));
    Vec aw2;
    RET(A1(w2b, aw2));
    axpy(y1b, 1.0, aw2);
    pairs.emplace_back(w2b, y1);
  } else {
    t4b += czb;
  }
  n0b += t4b / t1;
  t1b += -t4b * n0 / (t1 * t1);
  axpy(z0b, t1b, w1);
  axpy(w1b, t1b, z0);
  axpy(r0b, n0b, z0);
  axpy(z0b, n0b, r0);
  axpy(w1b, 1.0, P(y1b));
  Vec aw1;
  RET(A1(w1b, aw1));
  axpy(z0b, 1.0, aw1);
  pairs.emplace_back(w1b, z0);
  axpy(r0b, 1.0, P(z0b));
  Vec efb = r0b;
  if (guess) {
    Vec q0b = r0b;
    for (auto& x : q0b) x = -x;
    Vec aq;
    RET(A1(q0b, aq));
    Vec m0b = mub;
    axpy(m0b, 1.0, aq);
    pairs.emplace_back(q0b, m0);
    axpy(efb, 1.0, mulal(m0b));
  }
  // the gradient terms: dE/dr = F_raw(fields; Epb, efb) - sum F_raw(dipole; v, lam)
  if (!s->grad_f) {
    RET(dalloc(s, &s->grad_f, 3 * np));
    RET(dalloc(s, &s->grad_t, 3 * np));
  }
  Fo.assign(3 * np, 0.0);
  To.assign(3 * np, 0.0);
  auto run = [&](const Vec& ud, const Vec& up, int terms, double sign) -> thp_status {
    std::vector<double> w(6 * np);
    std::copy(ud.begin(), ud.end(), w.begin());
    std::copy(up.begin(), up.end(), w.begin() + 3 * np);
    HIPCHK(ctx, hipMemcpy(s->view.mu, w.data(), sizeof(double) * 6 * np, hipMemcpyHostToDevice));
    RET(epolar1_launch(s, s->grad_f, s->grad_t, nullptr, nullptr, terms, false));
    Vec Fr, Tr;
    RET(down(s->grad_f, 0, Fr));
    RET(down(s->grad_t, 0, Tr));
    // forces -dE/dr and torques -dE/dtheta: minus the raw combination
    for (size_t q = 0; q < 3 * np; ++q) {
      Fo[q] -= sign * Fr[q];
      To[q] -= sign * Tr[q];
    }
    return THP_OK;
  };
  thp_status rc = run(Epb, efb, kTermsField, 1.0);
  for (size_t k = 0; rc == THP_OK && k < pairs.size(); ++k) rc = run(pairs[k].second, pairs[k].first, kTermsDipole, -1.0);
  // the solve's dipoles back in place
  HIPCHK(ctx, hipMemcpy(s->view.mu, mu_keep.data(), sizeof(double) * 6 * np, hipMemcpyHostToDevice));
  return rc;
}

// slots of the virial shares written by epolar1_launch
size_t virial_slots(const thp_sys* s) {
  return 2 * (size_t)s->nblk + (size_t)(s->n + 255) / 256 + (grid_points(s) + kVirChunk - 1) / kVirChunk;
}

// Grid-shape autotuner of the serving loop, once per grid and precision on
// a context: the spreading brick edge (bricks set the spreading tiles, the
// merge and the gather's workgroups) and the FFT lines per workgroup,
// measured with HIP events on scratch copies of the request's system
// launched together as the loop launches its batches (enough copies for
// about 300,000 atoms, at most 12: one system alone favours small bricks
// that lose once a batch fills the device).  Bricks change the order of a
// grid point's sums, so the pick holds for every later system of that grid
// on the context, solved alone or in a batch; line groups change no
// arithmetic.
thp_status tune_grid_shape(thp_ctx* ctx, const thp_system_desc* d) {
  const std::array<int, 4> key{d->nfft[0], d->nfft[1], d->nfft[2], d->precision};
  if (ctx->grid_shape.count(key)) return THP_OK;
  const int copies = std::max(1, std::min(12, (int)std::lround(300000.0 / d->n)));
  // average ms of the batched launches of `kernels` (thp_time_kernel_batch
  // numbering) over scratch copies made with the given shapes
  auto measure = [&](int bmax, int fcap, std::initializer_list<int> kernels, double* ms) -> thp_status {
    std::vector<thp_sys*> sys;
    thp_status rc = THP_OK;
    ctx->force_brick = bmax;
    ctx->force_fft = fcap;
    for (int q = 0; q < copies && rc == THP_OK; ++q) {
      thp_sys* s = nullptr;
      rc = create_system(ctx, d, &s, true);
      if (rc == THP_OK) sys.push_back(s);
    }
    ctx->force_brick = ctx->force_fft = 0;
    *ms = 0.0;
    for (int k : kernels) {
      if (rc != THP_OK) break;
      double t = 0.0;
      rc = thp_time_kernel_batch(sys.data(), (int)sys.size(), k, 4, &t, nullptr);
      *ms += t;
    }
    for (thp_sys* s : sys) thp_system_destroy(s);
    return rc;
  };
  // brick edges: the distinct brick triples of bmax = 8, 6, 4, timed over a
  // spreading and a gather
  int best_b = 8;
  double best = 1e30;
  std::vector<std::array<int, 3>> seen;
  for (int bmax : {8, 6, 4}) {
    const std::array<int, 3> tri{pick_brick(d->nfft[0], bmax), pick_brick(d->nfft[1], bmax),
                                 pick_brick(d->nfft[2], bmax)};
    if (std::find(seen.begin(), seen.end(), tri) != seen.end()) continue;
    seen.push_back(tri);
    double ms = 0;
    RET(measure(bmax, 32, {1, 2}, &ms));
    if (ms < best) {
      best = ms;
      best_b = bmax;
    }
  }
  // FFT lines per workgroup: the distinct groupings of cap 32, 16, 8, timed
  // over a grid convolution
  int best_f = 32;
  best = 1e30;
  std::vector<std::array<int, 3>> fseen;
  for (int fcap : {32, 16, 8}) {
    const std::array<int, 3> g{fft_group_lines(d->nfft[0], fcap), fft_group_lines(d->nfft[1], fcap),
                               fft_group_lines(d->nfft[2], fcap)};
    if (std::find(fseen.begin(), fseen.end(), g) != fseen.end()) continue;
    fseen.push_back(g);
    double ms = 0;
    RET(measure(best_b, fcap, {6}, &ms));
    if (ms < best) {
      best = ms;
      best_f = fcap;
    }
  }
  ctx->grid_shape[key] = {best_b, best_f};
  return THP_OK;
}

// local-frame multipoles to the global frame on the device (rotpole)
thp_status rotate_frames(thp_sys* s) {
  hipLaunchKernelGGL(k_rotpole, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, s->d_view, s->d_frames,
                     s->d_pole_local);
  RET(launch_check(s->ctx, "k_rotpole"));
  return sync(s);
}

thp_status update_coords(thp_sys* s, const double* x, const double* y, const double* z, double* pinned) {
  const int n = s->n, np = s->npad;
  // sorted copy in ordinary host memory (pinned memory is only written,
  // never read back, from the host)
  std::vector<double>& hx = s->hx;
  std::vector<double>& hy = s->hy;
  std::vector<double>& hz = s->hz;
  hx.resize(n);
  hy.resize(n);
  hz.resize(n);
  double maxd2 = 0.0, maxu2 = 0.0;
  for (int q = 0; q < n; ++q) {
    const int a = s->perm[q];
    hx[q] = x[a];
    hy[q] = y[a];
    hz[q] = z[a];
    if (s->list_valid) {
      const double dx = hx[q] - s->bx0[q], dy = hy[q] - s->by0[q], dz = hz[q] - s->bz0[q];
      maxd2 = std::max(maxd2, dx * dx + dy * dy + dz * dz);
    }
    if (s->tight_valid) {
      const double dx = hx[q] - s->ux0[q], dy = hy[q] - s->uy0[q], dz = hz[q] - s->uz0[q];
      maxu2 = std::max(maxu2, dx * dx + dy * dy + dz * dz);
    }
  }
  std::vector<double> local;
  double* buf = pinned;
  if (!buf) {
    local.resize((size_t)3 * np);
    buf = local.data();
  }
  for (int c = 0; c < 3; ++c) {
    const std::vector<double>& src = c == 0 ? hx : c == 1 ? hy : hz;
    double* dst = buf + (size_t)c * np;
    std::memcpy(dst, src.data(), sizeof(double) * n);
    std::memset(dst + n, 0, sizeof(double) * (np - n));
  }
  hipStream_t st = s->ctx->stream;
  HIPCHK(s->ctx, hipMemcpyAsync((void*)s->view.x, buf, sizeof(double) * np, hipMemcpyHostToDevice, st));
  HIPCHK(s->ctx, hipMemcpyAsync((void*)s->view.y, buf + np, sizeof(double) * np, hipMemcpyHostToDevice, st));
  HIPCHK(s->ctx, hipMemcpyAsync((void*)s->view.z, buf + 2 * (size_t)np, sizeof(double) * np, hipMemcpyHostToDevice, st));
  if (!pinned) RET(sync(s));
  // block candidates (cut + buffer) stay valid while no atom moved more than
  // (buffer - skin)/2 since their build, the block-union lists (cut + skin)
  // while no atom moved more than skin/2
  const double lim_b = 0.5 * (s->buffer - s->skin), lim_u = 0.5 * s->skin;
  if (s->list_valid && maxd2 > lim_b * lim_b) s->list_valid = false;
  if (!s->list_valid || maxu2 > lim_u * lim_u) s->tight_valid = false;
  s->verlet_valid = false;
  s->pme_valid = false;
  s->coef_valid = false;
  s->mu_valid = false;
  // the global multipoles follow the frames (rotpole each step)
  if (s->has_frames) RET(rotate_frames(s));
  return THP_OK;
}

// per coordinate set: B-spline values, grid nodes and brick bins
thp_status b_pme_prep(thp_ctx* ctx, const Batch& b, int gate) {
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_zero_bins, g1((b.n_brick + 255) / 256), dim3(256), 0, st, b.d_views, b.d_brick, b.n_brick, gate);
  RET(launch_check(ctx, "k_zero_bins"));
  if (b.prec == THP_PREC_FP64) hipLaunchKernelGGL((k_pme_prep<double>), g1(b.n_blk), dim3(64), 0, st, b.d_views, b.d_blk, gate);
  else hipLaunchKernelGGL((k_pme_prep<float>), g1(b.n_blk), dim3(64), 0, st, b.d_views, b.d_blk, gate);
  RET(launch_check(ctx, "k_pme_prep"));
  hipLaunchKernelGGL(k_pme_scan, g1(b.nsys), dim3(1024), 0, st, b.d_views, (const int*)b.d_ids, gate);
  RET(launch_check(ctx, "k_pme_scan"));
  hipLaunchKernelGGL(k_pme_fill, g1(b.n_blk), dim3(64), 0, st, b.d_views, b.d_blk, gate);
  RET(launch_check(ctx, "k_pme_fill"));
  hipLaunchKernelGGL(k_pme_binsort, g1(b.n_brick), dim3(256), 0, st, b.d_views, b.d_brick, gate);
  return launch_check(ctx, "k_pme_binsort");
}

// caller-order coordinates of the requested systems into their sorted slots
thp_status b_load_coords(thp_ctx* ctx, const Batch& b, int gate) {
  hipLaunchKernelGGL(k_load_coords, g1(b.n_blk), dim3(64), 0, ctx->stream, b.d_views, b.d_blk, gate);
  RET(launch_check(ctx, "k_load_coords"));
  // local frames at the new coordinates (systems without frames exit)
  hipLaunchKernelGGL(k_rotpole_batch, g1(b.n_blk), dim3(64), 0, ctx->stream, b.d_views, b.d_blk, gate);
  return launch_check(ctx, "k_rotpole_batch");
}

thp_status b_end_setup(thp_ctx* ctx, const Batch& b) {
  hipLaunchKernelGGL(k_end_setup, g1((b.nsys + 63) / 64), dim3(64), 0, ctx->stream, b.d_views, (const int*)b.d_ids, b.nsys);
  return launch_check(ctx, "k_end_setup");
}

thp_status b_init_state(thp_ctx* ctx, const Batch& b, double poleps, int politer, const int* d_warm,
                        const int* d_req) {
  hipLaunchKernelGGL(k_init_state, g1((b.nsys + 63) / 64), dim3(64), 0, ctx->stream, b.d_views, (const int*)b.d_ids,
                     b.nsys, poleps, politer, d_warm, d_req);
  return launch_check(ctx, "k_init_state");
}

template <typename real, int NB, bool RC>
void launch_tmatxb_nb(const Batch& b, int vin, int vout, int dodiag, int gate, hipStream_t st) {
  const dim3 g = g1(b.n_blk);
  switch (b.segs) {
    case 1: hipLaunchKernelGGL((k_tmatxb_cl<real, 1, NB, RC>), g, dim3(64), 0, st, b.d_views, b.d_blk, vin, vout, dodiag, gate); break;
    case 2: hipLaunchKernelGGL((k_tmatxb_cl<real, 2, NB, RC>), g, dim3(128), 0, st, b.d_views, b.d_blk, vin, vout, dodiag, gate); break;
    case 8: hipLaunchKernelGGL((k_tmatxb_cl<real, 8, NB, RC>), g, dim3(512), 0, st, b.d_views, b.d_blk, vin, vout, dodiag, gate); break;
    case 16: hipLaunchKernelGGL((k_tmatxb_cl<real, 16, NB, RC>), g, dim3(1024), 0, st, b.d_views, b.d_blk, vin, vout, dodiag, gate); break;
    default: hipLaunchKernelGGL((k_tmatxb_cl<real, 4, NB, RC>), g, dim3(256), 0, st, b.d_views, b.d_blk, vin, vout, dodiag, gate); break;
  }
}

template <typename real, bool RC>
void launch_tmatxb(const Batch& b, int vin, int vout, int dodiag, int gate, hipStream_t st) {
  if (b.nb == 16) launch_tmatxb_nb<real, 16, RC>(b, vin, vout, dodiag, gate, st);
  else launch_tmatxb_nb<real, 8, RC>(b, vin, vout, dodiag, gate, st);
}

// real-space matvec on stream st; needs the compacted lists of the current
// coordinates (written by the permanent-field pass).  Systems that
// recompute their coefficients and systems that store them take separate
// kernels, each skipping the other kind; a batch launches those it holds.
thp_status b_pair_tmatxb_on(thp_ctx* ctx, const Batch& b, int vin, int vout, int dodiag, int gate, hipStream_t st) {
  if (b.prec == THP_PREC_FP64) {
    launch_tmatxb<double, false>(b, vin, vout, dodiag, gate, st);
  } else {
    const int kinds = b.rc ? b.rc : 3;  // not recorded: both kernels, each skips the other kind
    if (kinds & 1) launch_tmatxb<float, true>(b, vin, vout, dodiag, gate, st);
    if (kinds & 2) launch_tmatxb<float, false>(b, vin, vout, dodiag, gate, st);
  }
  return launch_check(ctx, "k_tmatxb_cl");
}

thp_status b_pair_tmatxb(thp_ctx* ctx, const Batch& b, int vin, int vout, int dodiag, int gate) {
  return b_pair_tmatxb_on(ctx, b, vin, vout, dodiag, gate, ctx->stream);
}

// The real-space matvec of a vector runs on the side stream while the main
// stream spreads and transforms the same vector (the two read it and write
// different arrays); join_matvec makes the main stream wait for it before
// the gather that adds the two parts.  A system's own launch sequence (its
// batch of one, always eager) and the serving loop's batches fork.  Eager
// launches use the context's fork and join events; a batch being captured
// as a graph uses the events its capture was given (Batch::cap_fork,
// cap_join), which serve that capture only and are destroyed with its
// graphs, so no event is recorded both inside a capture and in eager
// launches.  Without events for a capture the matvec runs in line.
inline bool capturing(const thp_ctx* ctx) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}
inline bool forks(const thp_ctx* ctx, const Batch& b) {
  if (!ctx->side || b.real_only) return false;
  if (b.fork_ok && b.nsys == 1) return true;
  if (!ctx->fork_batch) return false;
  return capturing(ctx) ? b.cap_fork != nullptr : true;
}

thp_status fork_matvec(thp_ctx* ctx, const Batch& b, int vin, int vout, int dodiag, int gate) {
  if (!forks(ctx, b)) return b_pair_tmatxb(ctx, b, vin, vout, dodiag, gate);
  const bool cap = b.cap_fork && capturing(ctx);
  hipEvent_t ef = cap ? b.cap_fork : ctx->ev_fork, ej = cap ? b.cap_join : ctx->ev_join;
  HIPCHK(ctx, hipEventRecord(ef, ctx->stream));
  HIPCHK(ctx, hipStreamWaitEvent(ctx->side, ef, 0));
  RET(b_pair_tmatxb_on(ctx, b, vin, vout, dodiag, gate, ctx->side));
  HIPCHK(ctx, hipEventRecord(ej, ctx->side));
  return THP_OK;
}

thp_status join_matvec(thp_ctx* ctx, const Batch& b) {
  if (!forks(ctx, b)) return THP_OK;
  const bool cap = b.cap_fork && capturing(ctx);
  HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, cap ? b.cap_join : ctx->ev_join, 0));
  return THP_OK;
}

// Field, guess and first residual (newinduce_pme2.f:111-196, 406-440);
// the solve states are initialised and the B-spline tables current.
// permanent field and the starting dipoles (newinduce_pme2.f:111-196)
thp_status b_setup_field(thp_ctx* ctx, const Batch& b, int gate) {
  RET(field(ctx, b, nullptr, 1, 1, gate));
  hipLaunchKernelGGL(k_guess, g1(b.n_blk), dim3(64), 0, ctx->stream, b.d_views, b.d_blk, gate);
  return launch_check(ctx, "k_guess");
}

thp_status b_setup(thp_ctx* ctx, const Batch& b, int gate) {
  RET(b_setup_field(ctx, b, gate));
  RET(fork_matvec(ctx, b, 0 /*mu*/, 3 /*hreal*/, 1, gate));
  if (!b.real_only) {
    RET(spread_uind(ctx, b, 0, 0, gate));
    RET(convolve(ctx, b, gate));
  }
  RET(join_matvec(ctx, b));
  return launch_gather_uind<1>(ctx, b, 0, 3, 0, gate);
}

// One PCG iteration of every running system (newinduce_pme2.f:442-569);
// each system's iteration number lives in its SolveState.
thp_status b_iteration(thp_ctx* ctx, const Batch& b) {
  RET(fork_matvec(ctx, b, 1 /*pp*/, 3 /*hreal*/, 1, kGateIter));
  if (!b.real_only) {
    RET(spread_uind(ctx, b, 1, 0, kGateIter));
    RET(convolve(ctx, b, kGateIter));
  }
  RET(join_matvec(ctx, b));
  RET(launch_gather_uind<2>(ctx, b, 1, 3, 0, kGateIter));
  hipStream_t st = ctx->stream;
  if (b.prec == THP_PREC_FP64) hipLaunchKernelGGL((k_cg_update1<double>), g1(b.n_red), dim3(256), 0, st, b.d_views, b.d_red);
  else hipLaunchKernelGGL((k_cg_update1<float>), g1(b.n_red), dim3(256), 0, st, b.d_views, b.d_red);
  RET(launch_check(ctx, "k_cg_update1"));
  hipLaunchKernelGGL(k_cg_update2, g1(b.n_red), dim3(256), 0, st, b.d_views, b.d_red);
  return launch_check(ctx, "k_cg_update2");
}

// E_pol of the finished systems (only_done) or of all
thp_status b_finish(thp_ctx* ctx, const Batch& b, int only_done) {
  hipLaunchKernelGGL(k_energy, g1(b.n_red), dim3(256), 0, ctx->stream, b.d_views, b.d_red, only_done);
  RET(launch_check(ctx, "k_energy"));
  hipLaunchKernelGGL(k_finish, g1(b.nsys), dim3(64), 0, ctx->stream, b.d_views, (const int*)b.d_ids, only_done);
  return launch_check(ctx, "k_finish");
}

thp_status b_collect(thp_ctx* ctx, const Batch& b, StateOut* d_out) {
  hipLaunchKernelGGL(k_collect, g1((b.nsys + 63) / 64), dim3(64), 0, ctx->stream, b.d_views, (const int*)b.d_ids, b.nsys, d_out);
  return launch_check(ctx, "k_collect");
}

thp_status download_mu(thp_sys* s, double* host) { return gather_out(s, host, s->view.mu, 6); }

// After device-side coordinate loading: bring the sorted coordinates back to
// the host copy and decide which lists stay valid (as update_coords does).
thp_status sync_host_coords(thp_sys* s) {
  const int n = s->n;
  std::vector<double> x(n), y(n), z(n);
  HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
  HIPCHK(s->ctx, hipMemcpy(x.data(), s->view.x, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(y.data(), s->view.y, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(z.data(), s->view.z, sizeof(double) * n, hipMemcpyDeviceToHost));
  double maxd2 = 0.0, maxu2 = 0.0;
  for (int q = 0; q < n; ++q) {
    if (s->list_valid) {
      const double dx = x[q] - s->bx0[q], dy = y[q] - s->by0[q], dz = z[q] - s->bz0[q];
      maxd2 = std::max(maxd2, dx * dx + dy * dy + dz * dz);
    }
    if (s->tight_valid) {
      const double dx = x[q] - s->ux0[q], dy = y[q] - s->uy0[q], dz = z[q] - s->uz0[q];
      maxu2 = std::max(maxu2, dx * dx + dy * dy + dz * dz);
    }
  }
  s->hx.swap(x);
  s->hy.swap(y);
  s->hz.swap(z);
  s->mu_valid = false;
  const double lim_b = 0.5 * (s->buffer - s->skin), lim_u = 0.5 * s->skin;
  if (s->list_valid && maxd2 > lim_b * lim_b) s->list_valid = false;
  if (!s->list_valid || maxu2 > lim_u * lim_u) s->tight_valid = false;
  s->verlet_valid = false;
  s->pme_valid = false;
  s->coef_valid = false;
  return THP_OK;
}

// Work table in XCD order.  MI355X hands workgroup k to XCD k % 8 (each XCD
// has its own L2): the list is cut into 8 contiguous ranges (spatially
// compact runs of blocks or bricks of the same systems) and interleaved so
// that range x runs on XCD x and its neighbouring data stay in that L2.
// Padding entries (sys = -1) keep the ranges aligned; kernels skip them.
std::vector<Work> xcd_order(const std::vector<Work>& list) {
  constexpr size_t kXcd = 8;
  const size_t n = list.size();
  const size_t per = (n + kXcd - 1) / kXcd;
  std::vector<Work> t(per * kXcd, Work{-1, 0});
  for (size_t x = 0; x < kXcd; ++x)
    for (size_t j = 0; j < per; ++j) {
      const size_t src = x * per + j;
      if (src < n) t[j * kXcd + x] = list[src];
    }
  return t;
}

// unscaled pairs within the cutoff (i<k), counted on the host from the
// block-union lists: the unit of the matvec's algorithmic FLOP count
thp_status count_pairs(thp_sys* s, int64_t* out) {
  int64_t cnt = 0;
  std::vector<int> lc(s->npad);
  HIPCHK(s->ctx, hipMemcpy(lc.data(), s->view.lcount, sizeof(int) * s->npad, hipMemcpyDeviceToHost));
  std::vector<unsigned short> ll((size_t)s->view.lmax * s->npad);
  std::vector<int> ul((size_t)s->nblk * s->view.umax);
  HIPCHK(s->ctx, hipMemcpy(ll.data(), s->view.llist, sizeof(unsigned short) * ll.size(), hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(ul.data(), s->view.ulist, sizeof(int) * ul.size(), hipMemcpyDeviceToHost));
  const double c2 = s->cut * s->cut;
  for (int q = 0; q < s->n; ++q) {
    const int b = q / kBlock;
    for (int t2 = 0; t2 < lc[q]; ++t2) {
      const int k = ul[(size_t)b * s->view.umax + ll[(size_t)t2 * s->npad + q]];
      if (k <= q) continue;
      double d[3] = {s->hx[k] - s->hx[q], s->hy[k] - s->hy[q], s->hz[k] - s->hz[q]};
      for (int c = 0; c < 3; ++c) d[c] -= s->box[c] * std::rint(d[c] / s->box[c]);
      if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= c2) ++cnt;
    }
  }
  *out = cnt;
  return THP_OK;
}

void append_work(const thp_sys* s, int index, std::vector<Work>& blk, std::vector<Work>& red, std::vector<Work>& gat,
                 std::vector<Work>& brick, std::vector<Work>& conv, std::vector<Work>& fft) {
  for (int q = 0; q < s->nblk; ++q) blk.push_back({index, q});
  for (int q = 0; q < s->view.red_blocks; ++q) red.push_back({index, q});
  for (int q = 0; q < s->view.gat_blocks; ++q) gat.push_back({index, q});
  for (int q = 0; q < s->nbricks; ++q) brick.push_back({index, q});
  const int nconv = (int)((grid_points(s) + kConvChunk - 1) / kConvChunk);
  for (int q = 0; q < nconv; ++q) conv.push_back({index, q});
  // line groups of the three axes, one entry per z plane, the register z pass
  const int fc = s->view.fft_cap;
  const int ng = fft_work_count(s->nfft, fc);
  for (int q = 0; q < ng; ++q) fft.push_back({index, q});
}

}  // namespace thp

// ====================================================================== API
extern "C" {

const char* thp_version(void) { return "thp 0.1 (gfx950)"; }

thp_status thp_init(thp_ctx** out, int device) {
  if (!out) return THP_ERR_ARG;
  thp_ctx* ctx = new thp_ctx;
  ctx->device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    *out = ctx;
    return fail(ctx, THP_ERR_HIP, "no HIP device %d (found %d)", device, ndev);
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    *out = ctx;
    return fail(ctx, THP_ERR_HIP, "cannot initialise device %d", device);
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
  // the fused plane FFT needs more than the default 64 KB of dynamic LDS;
  // THP_FFT_PLANE=0 keeps the five one-axis passes
  ctx->plane_fft = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_plane<float, 0, kPlaneThreads>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kPlaneLds) == hipSuccess &&
                   hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_plane<float, 1, kPlaneThreads>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kPlaneLds) == hipSuccess;
  if (const char* e = getenv("THP_FFT_PLANE"))
    if (atoi(e) == 0) ctx->plane_fft = false;
  if (const char* e = getenv("THP_FORK_BATCH")) ctx->fork_batch = atoi(e) != 0;
  if (const char* e = getenv("THP_FFT_REG"))
    if (atoi(e) == 0) ctx->reg_fft = false;
  // lanes per line of the register FFT's plane and z passes (1, 2 or 4)
  if (const char* e = getenv("THP_FFT_PP")) ctx->fft_pp = atoi(e);
  if (const char* e = getenv("THP_FFT_ZP")) ctx->fft_zp = atoi(e);
  // a second stream (default priority) for the real-space matvec beside the
  // reciprocal chain; THP_FORK=0 keeps every launch on one stream
  const char* fe = getenv("THP_FORK");
  if (!(fe && atoi(fe) == 0)) {
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
      *out = ctx;
      return fail(ctx, THP_ERR_HIP, "cannot create the side stream of device %d", device);
    }
  }
  {
    // the no-allocation entry point's scratch on this device (the device is
    // current here)
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      *out = ctx;
      return fail(ctx, THP_ERR_HIP, "cannot resolve the scratch array on device %d: %s", device, hipGetErrorString(e));
    }
    ctx->scratch = static_cast<double*>(ptr);
  }
  if (hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming) != hipSuccess) {
    *out = ctx;
    return fail(ctx, THP_ERR_HIP, "cannot create the ordering events of device %d", device);
  }
  // THP_SOLVE_CHUNK=k: PCG iterations of a one-system solve between host
  // reads of the convergence flag, FP64 and mixed alike (default 1 for FP64,
  // the reference's read after every iteration, and 4 for mixed)
  if (const char* e = getenv("THP_SOLVE_CHUNK")) {
    const int v = atoi(e);
    if (v >= 1 && v <= 16) ctx->solve_chunk[0] = ctx->solve_chunk[1] = v;
  }
  // THP_MATVEC_RC=0: every system stores its matvec coefficients per solve
  if (const char* e = getenv("THP_MATVEC_RC"))
    if (atoi(e) == 0) ctx->rc_matvec = false;
  // THP_SEGS forces the pair kernels' wavefronts per 64-atom block
  if (const char* e = getenv("THP_SEGS")) {
    const int v = atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) ctx->pair_segs = v;
  }
  // THP_NB forces the matvec's list entries per prefetched batch
  if (const char* e = getenv("THP_NB")) {
    const int v = atoi(e);
    if (v == 8 || v == 16) ctx->pair_nb = v;
  }
  *out = ctx;
  return THP_OK;
}

thp_status thp_finalize(thp_ctx* ctx) {
  if (!ctx) return THP_ERR_ARG;
  if (ctx->serve) thp_serve_stop(ctx);
  // systems the caller did not destroy are released here; their handles die
  // with the context
  while (!ctx->systems.empty()) thp_system_destroy(*ctx->systems.begin());
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->ev_in) (void)hipEventDestroy(ctx->ev_in);
  if (ctx->ev_out) (void)hipEventDestroy(ctx->ev_out);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return THP_OK;
}

const char* thp_last_error(thp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

thp_status thp_grid_shape(thp_ctx* ctx, const int* nfft, int precision, int* brick_max, int* fft_lines_max) {
  if (!ctx || !nfft || !brick_max || !fft_lines_max) return THP_ERR_ARG;
  const auto it = ctx->grid_shape.find({nfft[0], nfft[1], nfft[2], precision});
  if (it == ctx->grid_shape.end()) return THP_ERR_STATE;
  *brick_max = it->second.first;
  *fft_lines_max = it->second.second;
  return THP_OK;
}

thp_status thp_system_create(thp_ctx* ctx, const thp_system_desc* d, thp_sys** out) {
  return create_system(ctx, d, out, true);
}

}  // extern "C"

thp_status thp::create_system(thp_ctx* ctx, const thp_system_desc* d, thp_sys** out, bool own_grid) {
  if (!ctx || !d || !out) return THP_ERR_ARG;
  if (d->n <= 0 || d->n >= kMaxAtoms) return fail(ctx, THP_ERR_ARG, "n = %d out of range", d->n);
  if (d->order != kOrder) return fail(ctx, THP_ERR_ARG, "only B-spline order %d is implemented", kOrder);
  for (int k = 0; k < 3; ++k) {
    if (!(d->box[k] > 0)) return fail(ctx, THP_ERR_ARG, "box edge %d must be positive", k);
    if (d->nfft[k] < 2 * kOrder) return fail(ctx, THP_ERR_ARG, "nfft[%d] = %d too small", k, d->nfft[k]);
    // the grid FFT takes 2,3,5-smooth lengths (kewald.f:67-73 picks only those)
    int r = d->nfft[k];
    for (int f : {2, 3, 5})
      while (r % f == 0) r /= f;
    if (r != 1 || d->nfft[k] > kFftMaxLen)
      return fail(ctx, THP_ERR_ARG, "nfft[%d] = %d: the FFT takes 2,3,5-smooth lengths up to %d", k, d->nfft[k], kFftMaxLen);
  }
  if (!(d->cut > 0) || d->buffer < 0 || !(d->aewald > 0)) return fail(ctx, THP_ERR_ARG, "bad cut/buffer/aewald");
  // nblist.f:76-83: (cut + buffer) must not exceed half the box
  const double mb = d->cut + d->buffer;
  const double minhalf = 0.5 * std::min({d->box[0], d->box[1], d->box[2]});
  if (mb * mb > minhalf * minhalf)
    return fail(ctx, THP_ERR_LIST, "cutoff + buffer %.3f exceeds half the box %.3f", mb, minhalf);
  if (d->precision != THP_PREC_FP64 && d->precision != THP_PREC_MIXED) return fail(ctx, THP_ERR_ARG, "bad precision");
  if (!d->x || !d->y || !d->z || !d->rpole || !d->polarity || !d->thole) return fail(ctx, THP_ERR_ARG, "null array");
  if (d->nscaled < 0 || (d->nscaled > 0 && (!d->scaled_ik || !d->scaled_dpu))) return fail(ctx, THP_ERR_ARG, "bad scaled pairs");
  HIPCHK(ctx, hipSetDevice(ctx->device));

  thp_sys* s = new thp_sys;
  s->ctx = ctx;
  s->own_grid = own_grid;
  ctx->systems.insert(s);
  s->n = d->n;
  s->npad = (d->n + kBlock - 1) / kBlock * kBlock;
  s->nblk = s->npad / kBlock;
  s->prec = d->precision;
  for (int k = 0; k < 3; ++k) {
    s->box[k] = d->box[k];
    s->nfft[k] = d->nfft[k];
  }
  s->aewald = d->aewald;
  s->cut = d->cut;
  s->buffer = d->buffer;
  s->order = d->order;
  const int n = s->n, np = s->npad;

  s->perm = spatial_order(n, d->x, d->y, d->z, d->box);
  s->iperm.assign(n, 0);
  for (int q = 0; q < n; ++q) s->iperm[s->perm[q]] = q;

  auto cleanup = [&](thp_status r) {
    thp_system_destroy(s);
    return r;
  };
  // ---- per-atom arrays in sorted order
  std::vector<double> hx(np), hy(np), hz(np), rp((size_t)10 * np, 0.0), pol(np, 0.0), pd(np, 0.0), th(np, 0.0);
  for (int q = 0; q < np; ++q) {
    const int a = q < n ? s->perm[q] : -1;
    if (a < 0) {
      hx[q] = hy[q] = hz[q] = 0.0;
      continue;
    }
    hx[q] = d->x[a];
    hy[q] = d->y[a];
    hz[q] = d->z[a];
    const double* r = d->rpole + (size_t)a * 13;
    const double v[10] = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[8], r[9], r[12]};
    for (int c = 0; c < 10; ++c) rp[(size_t)c * np + q] = v[c];
    pol[q] = d->polarity[a];
    pd[q] = std::pow(d->polarity[a], 1.0 / 6.0);  // kpolar.f:235
    th[q] = d->thole[a];
  }
  s->hx.assign(hx.begin(), hx.begin() + n);
  s->hy.assign(hy.begin(), hy.begin() + n);
  s->hz.assign(hz.begin(), hz.begin() + n);

  // ---- scaled pairs -> per-atom CSR with codes into a table of (d,p,u)
  std::map<std::tuple<double, double, double>, int> codes;
  std::vector<double> tab = {1.0, 1.0, 1.0};
  codes[{1.0, 1.0, 1.0}] = 0;
  std::vector<std::vector<std::pair<int, int>>> ex(n);
  for (int m = 0; m < d->nscaled; ++m) {
    const int a = d->scaled_ik[2 * m], b = d->scaled_ik[2 * m + 1];
    if (a < 0 || b < 0 || a >= n || b >= n || a == b) return cleanup(fail(ctx, THP_ERR_ARG, "bad scaled pair %d", m));
    const auto key = std::make_tuple(d->scaled_dpu[3 * m], d->scaled_dpu[3 * m + 1], d->scaled_dpu[3 * m + 2]);
    auto it = codes.find(key);
    int code;
    if (it == codes.end()) {
      code = (int)codes.size();
      if (code >= 127) return cleanup(fail(ctx, THP_ERR_ARG, "too many distinct scale triples"));
      codes[key] = code;
      tab.push_back(std::get<0>(key));
      tab.push_back(std::get<1>(key));
      tab.push_back(std::get<2>(key));
    } else {
      code = it->second;
    }
    const int sa = s->iperm[a], sb = s->iperm[b];
    ex[sa].push_back({sb, code});
    ex[sb].push_back({sa, code});
  }
  std::vector<int> eoff(n + 1, 0), ek, ec;
  for (int q = 0; q < n; ++q) {
    std::sort(ex[q].begin(), ex[q].end());
    eoff[q + 1] = eoff[q] + (int)ex[q].size();
    for (auto& e : ex[q]) {
      ek.push_back(e.first);
      ec.push_back(e.second);
    }
  }

  SysView& v = s->view;
  memset(&v, 0, sizeof v);
  v.n = n;
  v.npad = np;
  v.nblk = s->nblk;
  v.red_blocks = std::max(1, std::min(256, (n + 255) / 256));
  // grid shapes: the autotuner's while it measures, else the ones it picked
  // for this grid on this context, else the largest bricks and line groups
  int bmax = 8, fcap = 32;
  {
    const auto it = ctx->grid_shape.find({s->nfft[0], s->nfft[1], s->nfft[2], s->prec});
    if (it != ctx->grid_shape.end()) {
      bmax = it->second.first;
      fcap = it->second.second;
    }
    if (ctx->force_brick) bmax = ctx->force_brick;
    if (ctx->force_fft) fcap = ctx->force_fft;
  }
  v.fft_cap = fcap;
  for (int k = 0; k < 3; ++k) {
    v.brick[k] = pick_brick(s->nfft[k], bmax);
    v.nbrick[k] = (s->nfft[k] + v.brick[k] - 1) / v.brick[k];
  }
  s->nbricks = v.nbrick[0] * v.nbrick[1] * v.nbrick[2];
  v.gat_blocks = s->nbricks;  // one gather workgroup (and partial pair) per brick
  v.prec = s->prec;
  for (int k = 0; k < 3; ++k) {
    v.box[k] = s->box[k];
    v.nfft[k] = s->nfft[k];
    v.afrac[k] = s->nfft[k] / s->box[k];
  }
  v.aewald = s->aewald;
  v.cut2 = s->cut * s->cut;
  v.cut = s->cut;
  s->skin = std::min(0.5, 0.5 * s->buffer);
  v.skin = s->skin;
  v.lim_u2 = 0.25 * s->skin * s->skin;
  v.ulim2 = (s->cut + s->skin) * (s->cut + s->skin);
  v.mbuf2 = mb * mb;
  v.term_self = (4.0 / 3.0) * s->aewald * s->aewald * s->aewald / std::sqrt(M_PI);
  v.ncode = (int)codes.size();
  conversion_mats(s->box, s->nfft, v.ctf, v.ftc);

  hipStream_t st = ctx->stream;
  double *dx, *dy, *dz, *drp, *dpol, *dpd, *dth, *dtab;
  int *deoff, *dek, *dec;
#define ALLOC(p, c) do { thp_status r_ = dalloc(s, &(p), (c)); if (r_ != THP_OK) return cleanup(r_); } while (0)
#define H2D(dst, src, bytes) do { if (hipMemcpyAsync((dst), (src), (bytes), hipMemcpyHostToDevice, st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "H2D copy failed")); } while (0)
  ALLOC(dx, np); ALLOC(dy, np); ALLOC(dz, np);
  ALLOC(drp, (size_t)10 * np); ALLOC(dpol, np); ALLOC(dpd, np); ALLOC(dth, np);
  ALLOC(dtab, tab.size()); ALLOC(deoff, n + 1); ALLOC(dek, ek.size()); ALLOC(dec, ec.size());
  H2D(dx, hx.data(), sizeof(double) * np);
  H2D(dy, hy.data(), sizeof(double) * np);
  H2D(dz, hz.data(), sizeof(double) * np);
  H2D(drp, rp.data(), sizeof(double) * rp.size());
  H2D(dpol, pol.data(), sizeof(double) * np);
  H2D(dpd, pd.data(), sizeof(double) * np);
  H2D(dth, th.data(), sizeof(double) * np);
  H2D(dtab, tab.data(), sizeof(double) * tab.size());
  H2D(deoff, eoff.data(), sizeof(int) * (n + 1));
  if (!ek.empty()) {
    H2D(dek, ek.data(), sizeof(int) * ek.size());
    H2D(dec, ec.data(), sizeof(int) * ec.size());
  }
  v.x = dx; v.y = dy; v.z = dz; v.rpole = drp; v.pol = dpol; v.pdamp = dpd; v.thole = dth;
  {
    // matvec coefficients recomputed per entry instead of stored per solve:
    // FP32 systems with one Thole value a > 0 (AMOEBA uses 0.39 throughout)
    // (and aewald * cut inside the table of the Ewald functions)
    bool one_a = s->prec != THP_PREC_FP64 && ctx->rc_matvec && n > 0 && d->thole[0] > 0.0 &&
                 d->aewald * d->cut < 0.98 * kGtabN / kGtabScale;
    for (int q = 1; one_a && q < n; ++q) one_a = d->thole[q] == d->thole[0];
    if (one_a) {
      const std::vector<float> gt = ewald_gtab();
      float* dg;
      ALLOC(dg, gt.size());
      H2D(dg, gt.data(), sizeof(float) * gt.size());
      if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
      v.gtab = dg;
      v.rc_sqa = std::sqrt(d->thole[0] * M_LOG2E);
    }
  }
  v.scale_tab = dtab; v.excl_off = deoff; v.excl_k = dek; v.excl_code = dec;
  {
    double* sco;
    ALLOC(sco, 2 * std::max<size_t>(1, ek.size()));
    v.scoef = sco;
  }
  ALLOC(s->iperm_dev, n);
  H2D(s->iperm_dev, s->iperm.data(), sizeof(int) * n);
  {
    // device-side coordinate loading (caller order -> sorted slots)
    double *raw, *refu;
    int* pd;
    ALLOC(raw, (size_t)3 * n);
    ALLOC(refu, (size_t)3 * np);
    ALLOC(pd, np);
    H2D(raw, d->x, sizeof(double) * n);
    H2D(raw + n, d->y, sizeof(double) * n);
    H2D(raw + 2 * (size_t)n, d->z, sizeof(double) * n);
    std::vector<int> pp(np, 0);
    for (int q = 0; q < n; ++q) pp[q] = s->perm[q];
    H2D(pd, pp.data(), sizeof(int) * np);
    if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
    v.raw = raw;
    v.refu = refu;
    v.perm = pd;
    v.raw_rpole = nullptr;
    if (!own_grid) {
      // serving systems take new multipoles with a request
      double* rr;
      ALLOC(rr, (size_t)13 * n);
      H2D(rr, d->rpole, sizeof(double) * 13 * n);
      if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
      v.raw_rpole = rr;
    }
  }
  ALLOC(v.nbcount, np); ALLOC(v.lcount, np); ALLOC(v.tcount, np); ALLOC(v.ucount, s->nblk); ALLOC(v.bbox, (size_t)s->nblk * 6);
  ALLOC(s->bcount, s->nblk);
  ALLOC(v.ef, (size_t)6 * np); ALLOC(v.mu, (size_t)6 * np); ALLOC(v.res, (size_t)6 * np);
  ALLOC(v.h, (size_t)6 * np); ALLOC(v.hreal, (size_t)6 * np); ALLOC(v.pp, (size_t)6 * np);
  ALLOC(v.zr, (size_t)6 * np); ALLOC(v.partial, partial_doubles(v.gat_blocks, v.red_blocks));
  ALLOC(s->cphi, (size_t)10 * np);
  ALLOC(s->stage, (size_t)10 * n);
  ALLOC(v.st, 1); ALLOC(v.err, 1);
  if (hipMemsetAsync(v.err, 0, sizeof(int), st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "memset"));
  {
    // no solve in progress: every PCG kernel skips the system until
    // k_init_state starts one
    SolveState idle;
    memset(&idle, 0, sizeof idle);
    idle.conv = kConverged;
    idle.finalized = 1;
    if (hipMemcpy(v.st, &idle, sizeof idle, hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(ctx, THP_ERR_HIP, "solve state"));
  }
  for (double* vec : {v.ef, v.mu, v.res, v.h, v.hreal, v.pp, v.zr})
    if (hipMemsetAsync(vec, 0, sizeof(double) * 6 * np, st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "memset"));

  // ---- PME grid, scalar-sum table and FFT twiddles
  const size_t npts = grid_points(s);
  const auto q = qfac_table(s->box, s->nfft, s->aewald, s->order);
  if (s->prec == THP_PREC_FP64) {
    double *g = nullptr, *qf;
    if (own_grid) ALLOC(g, 2 * npts);
    ALLOC(qf, npts);
    H2D(qf, q.data(), sizeof(double) * npts);
    v.grid = g;
    v.qfac = qf;
  } else {
    float *g = nullptr, *qf;
    if (own_grid) ALLOC(g, 2 * npts);
    ALLOC(qf, npts);
    std::vector<float> qs(q.begin(), q.end());
    H2D(qf, qs.data(), sizeof(float) * npts);
    if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
    v.grid = g;
    v.qfac = qf;
  }
  {
    // twiddles exp(-2 pi i m / K) of the three axes, from FP64
    std::vector<double> tw;
    for (int ax = 0; ax < 3; ++ax)
      for (int m = 0; m < s->nfft[ax]; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)s->nfft[ax];
        tw.push_back(std::cos(ang));
        tw.push_back(std::sin(ang));
      }
    if (s->prec == THP_PREC_FP64) {
      double* t;
      ALLOC(t, tw.size());
      H2D(t, tw.data(), sizeof(double) * tw.size());
      v.twid = t;
    } else {
      float* t;
      ALLOC(t, tw.size());
      std::vector<float> tf(tw.begin(), tw.end());
      H2D(t, tf.data(), sizeof(float) * tf.size());
      if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
      v.twid = t;
    }
  }

  // ---- brick bins and spreading tiles
  if (own_grid) {
    const size_t tw = (size_t)s->nbricks * kTile3 * 2;
    if (s->prec == THP_PREC_FP64) {
      double* t;
      ALLOC(t, tw);
      v.tiles = t;
    } else {
      float* t;
      ALLOC(t, tw);
      v.tiles = t;
    }
  }
  if (s->prec == THP_PREC_FP64) {
    double* th;
    ALLOC(th, (size_t)np * kTheta);
    v.theta = th;
  } else {
    float* th;
    ALLOC(th, (size_t)np * kTheta);
    v.theta = th;
  }
  ALLOC(v.g0, (size_t)4 * np);
  ALLOC(v.bin_count, s->nbricks);
  ALLOC(v.bin_start, s->nbricks + 1);
  ALLOC(v.bin_list, n);
  {
    std::vector<Work> bw(s->nbricks);
    for (int b = 0; b < s->nbricks; ++b) bw[b] = {0, b};
    ALLOC(s->d_brickwork, s->nbricks);
    H2D(s->d_brickwork, bw.data(), sizeof(Work) * s->nbricks);
    if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
  }

  // ---- view + work tables (natural order for the list builds, XCD order
  // for the path kernels)
  ALLOC(s->d_view, 1);
  std::vector<Work> wk(s->nblk);
  for (int b = 0; b < s->nblk; ++b) wk[b] = {0, b};
  ALLOC(s->d_work, s->nblk);
  H2D(s->d_work, wk.data(), sizeof(Work) * s->nblk);
  ALLOC(s->d_redwork, v.red_blocks);
  H2D(s->d_redwork, wk.data(), sizeof(Work) * v.red_blocks);
  {
    const std::vector<Work> wx = xcd_order(wk);
    std::vector<Work> bn(s->nbricks);
    for (int b = 0; b < s->nbricks; ++b) bn[b] = {0, b};
    const std::vector<Work> bx = xcd_order(bn);
    ALLOC(s->d_workx, wx.size());
    ALLOC(s->d_brickx, bx.size());
    H2D(s->d_workx, wx.data(), sizeof(Work) * wx.size());
    H2D(s->d_brickx, bx.data(), sizeof(Work) * bx.size());
    s->nworkx = (int)wx.size();
    s->nbrickx = (int)bx.size();
    if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
  }
  s->d_gatwork = s->d_brickx;  // gather workgroups are bricks
  {
    const int nconv = (int)((npts + kConvChunk - 1) / kConvChunk);
    std::vector<Work> cw(nconv);
    for (int b = 0; b < nconv; ++b) cw[b] = {0, b};
    ALLOC(s->d_convwork, nconv);
    H2D(s->d_convwork, cw.data(), sizeof(Work) * nconv);
    s->nconv = nconv;
    const int fc = v.fft_cap;
    const int nfftw = fft_work_count(s->nfft, fc);
    std::vector<Work> fw(nfftw);
    for (int b = 0; b < nfftw; ++b) fw[b] = {0, b};
    ALLOC(s->d_fftwork, nfftw);
    H2D(s->d_fftwork, fw.data(), sizeof(Work) * nfftw);
    s->nfftw = nfftw;
    if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
  }
  ALLOC(s->d_sysid, 2);
  {
    const int ids[2] = {0, 1};  // {system index 0, warm-start flag 1}
    H2D(s->d_sysid, ids, sizeof ids);
    if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
  }
  s->d_one = s->d_sysid + 1;
  {
    Batch& b = s->one;
    b.prec = s->prec;
    b.nsys = 1;
    b.rc = v.rc_sqa > 0.0 ? 1 : 2;
    b.segs = ctx->pair_segs > 0 ? ctx->pair_segs : kDefaultSegs;
    b.nb = ctx->pair_nb > 0 ? ctx->pair_nb : 8;
    b.fork_ok = 1;
    b.d_views = s->d_view;
    b.d_blk = s->d_workx;
    b.n_blk = s->nworkx;
    b.d_red = s->d_redwork;
    b.n_red = v.red_blocks;
    b.d_gat = s->d_gatwork;
    b.n_gat = s->nbrickx;
    b.d_brick = s->d_brickx;
    b.n_brick = s->nbrickx;
    b.d_conv = s->d_convwork;
    b.n_conv = s->nconv;
    b.d_ids = s->d_sysid;
    b.d_fft = s->d_fftwork;
    b.n_fft = s->nfftw;
    const int* nf = s->nfft;
    plane_fft_setup(b, &nf, 1, s->prec);
  }
  for (int e = 0; e < 4; ++e)
    if (hipEventCreate(&s->ev[e]) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "event"));
  if (upload_view(s) != THP_OK) return cleanup(THP_ERR_HIP);
  if (hipStreamSynchronize(st) != hipSuccess) return cleanup(fail(ctx, THP_ERR_HIP, "sync"));
#undef ALLOC
#undef H2D
  *out = s;
  return THP_OK;
}

extern "C" {

thp_status thp_system_destroy(thp_sys* s) {
  if (!s) return THP_ERR_ARG;
  if (s->ctx) {
    if (s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
    s->ctx->systems.erase(s);
  }
  for (auto& a : s->allocs) (void)hipFree(a.first);
  if (s->hstage) (void)hipHostFree(s->hstage);
  for (auto& e : s->ev)
    if (e) (void)hipEventDestroy(e);
  delete s;
  return THP_OK;
}

thp_status thp_system_footprint(const thp_sys* s, int64_t* bytes) {
  if (!s || !bytes) return THP_ERR_ARG;
  *bytes = s->bytes;
  return THP_OK;
}

thp_status thp_system_update_coords(thp_sys* s, const double* x, const double* y, const double* z) {
  if (!s || !x || !y || !z) return THP_ERR_ARG;
  return update_coords(s, x, y, z, nullptr);
}

thp_status thp_system_set_frames(thp_sys* s, const int* axistype, const int* zaxis, const int* xaxis, const int* yaxis,
                                 const double* pole) {
  if (!s || !axistype || !zaxis || !xaxis || !yaxis || !pole) return THP_ERR_ARG;
  const int n = s->n;
  const size_t np = s->npad;
  std::vector<int> fr((size_t)4 * np, -1);
  std::vector<double> pl((size_t)10 * np, 0.0);
  for (int q = 0; q < (int)np; ++q) fr[(size_t)4 * q] = kAxisNone;
  for (int i = 0; i < n; ++i) {
    const int t = axistype[i];
    if (t < kAxisNone || t > kAxis3Fold) return fail(s->ctx, THP_ERR_ARG, "atom %d: axis type %d", i, t);
    // frame atoms each type needs (rotpole.f: z for all but None, x for all
    // but None and Z-Only, y for Z-Bisect and 3-Fold, optional for Z-then-X)
    const bool nz = t != kAxisNone, nx = t >= kAxisZThenX, ny = t == kAxisZBisect || t == kAxis3Fold;
    const int f[3] = {zaxis[i], xaxis[i], yaxis[i]};
    const bool need[3] = {nz, nx, ny};
    const bool allowed[3] = {nz, nx, ny || t == kAxisZThenX};
    for (int k = 0; k < 3; ++k) {
      if (f[k] >= 0 && (f[k] >= n || f[k] == i || !allowed[k]))
        return fail(s->ctx, THP_ERR_ARG, "atom %d: frame atom %d (axis %d) invalid for axis type %d", i, f[k], k, t);
      if (need[k] && f[k] < 0) return fail(s->ctx, THP_ERR_ARG, "atom %d: axis type %d needs frame atom %d", i, t, k);
    }
    const int q = s->iperm[i];
    fr[(size_t)4 * q] = t;
    for (int k = 0; k < 3; ++k) fr[(size_t)4 * q + 1 + k] = f[k] >= 0 && allowed[k] ? s->iperm[f[k]] : -1;
    const double* r = pole + (size_t)i * 13;
    const double v[10] = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[8], r[9], r[12]};
    for (int c = 0; c < 10; ++c) pl[(size_t)c * np + q] = v[c];
  }
  // inverse map for the gradient: for each atom, the sites (ascending)
  // whose frame forces reach it, the atom itself first if it has a frame
  std::vector<int> toff((size_t)np + 1, 0), tinv;
  {
    std::vector<std::vector<int>> by_atom(np);
    for (int q = 0; q < (int)np; ++q) {
      if (fr[(size_t)4 * q] == kAxisNone) continue;
      int tg[4] = {q, fr[(size_t)4 * q + 1], fr[(size_t)4 * q + 2], fr[(size_t)4 * q + 3]};
      for (int k = 0; k < 4; ++k) {
        bool dup = tg[k] < 0;
        for (int j = 0; j < k && !dup; ++j) dup = tg[j] == tg[k];
        if (!dup) by_atom[tg[k]].push_back(q);
      }
    }
    for (int a = 0; a < (int)np; ++a) {
      toff[a + 1] = toff[a] + (int)by_atom[a].size();
      tinv.insert(tinv.end(), by_atom[a].begin(), by_atom[a].end());
    }
  }
  if (!s->d_frames) {
    RET(dalloc(s, &s->d_frames, (size_t)4 * np));
    RET(dalloc(s, &s->d_pole_local, (size_t)10 * np));
    RET(dalloc(s, &s->d_tinv_off, (size_t)np + 1));
    RET(dalloc(s, &s->d_tinv, (size_t)4 * np));
  }
  HIPCHK(s->ctx, hipMemcpy(s->d_frames, fr.data(), sizeof(int) * fr.size(), hipMemcpyHostToDevice));
  HIPCHK(s->ctx, hipMemcpy(s->d_tinv_off, toff.data(), sizeof(int) * toff.size(), hipMemcpyHostToDevice));
  if (!tinv.empty())
    HIPCHK(s->ctx, hipMemcpy(s->d_tinv, tinv.data(), sizeof(int) * tinv.size(), hipMemcpyHostToDevice));
  HIPCHK(s->ctx, hipMemcpy(s->d_pole_local, pl.data(), sizeof(double) * pl.size(), hipMemcpyHostToDevice));
  s->has_frames = true;
  s->pme_valid = false;
  s->coef_valid = false;
  s->mu_valid = false;
  s->view.frames = s->d_frames;
  s->view.pole_local = s->d_pole_local;
  RET(upload_view(s));
  return rotate_frames(s);
}

thp_status thp_system_update_multipoles(thp_sys* s, const double* rpole) {
  if (!s || !rpole) return THP_ERR_ARG;
  s->mu_valid = false;
  if (s->has_frames) {
    // global multipoles given outright replace the frames
    s->has_frames = false;
    s->view.frames = nullptr;
    s->view.pole_local = nullptr;
    RET(upload_view(s));
  }
  const int n = s->n;
  const size_t np = s->npad;
  std::vector<double> rp((size_t)10 * np, 0.0);
  for (int q = 0; q < n; ++q) {
    const double* r = rpole + (size_t)s->perm[q] * 13;
    const double v[10] = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[8], r[9], r[12]};
    for (int c = 0; c < 10; ++c) rp[(size_t)c * np + q] = v[c];
  }
  HIPCHK(s->ctx, hipMemcpy(const_cast<double*>(s->view.rpole), rp.data(), sizeof(double) * rp.size(),
                           hipMemcpyHostToDevice));
  if (s->view.raw_rpole)
    HIPCHK(s->ctx, hipMemcpy(const_cast<double*>(s->view.raw_rpole), rpole, sizeof(double) * 13 * n,
                             hipMemcpyHostToDevice));
  return THP_OK;
}

thp_status thp_nblist(thp_sys* s, int force_rebuild) {
  if (!s) return THP_ERR_ARG;
  if (s->verlet_valid && !force_rebuild) return THP_OK;
  // the Verlet list is the pair set of the current coordinates (mlistcell):
  // the block candidates are rebuilt from them first
  RET(build_block_lists(s));
  // per-atom Verlet list (mlistcell): every k with r^2 <= (cut+buffer)^2
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL((k_nblist_atoms<false>), g1(s->nblk), dim3(64), 0, st, s->d_view, s->d_work, (const int*)s->bcount, (const int*)s->blist, s->maxbl);
  RET(launch_check(s->ctx, "k_nblist_atoms"));
  std::vector<int> cnt(s->npad);
  HIPCHK(s->ctx, hipMemcpyAsync(cnt.data(), s->view.nbcount, sizeof(int) * s->npad, hipMemcpyDeviceToHost, st));
  RET(sync(s));
  int maxnb = 0;
  for (int q = 0; q < s->n; ++q) maxnb = std::max(maxnb, cnt[q]);
  maxnb = std::max(32, (maxnb + 31) / 32 * 32);
  if (maxnb > s->maxnb) {
    dfree(s, s->view.nb);
    s->view.nb = nullptr;
    RET(dalloc(s, &s->view.nb, (size_t)maxnb * s->npad));
    s->maxnb = maxnb;
    s->view.maxnb = maxnb;
    RET(upload_view(s));
  }
  hipLaunchKernelGGL((k_nblist_atoms<true>), g1(s->nblk), dim3(64), 0, st, s->d_view, s->d_work, (const int*)s->bcount, (const int*)s->blist, s->maxbl);
  RET(launch_check(s->ctx, "k_nblist_atoms"));
  RET(check_err_flag(s));
  s->verlet_valid = true;
  return THP_OK;
}

thp_status thp_nblist_get(thp_sys* s, int64_t* npairs, int* ik, int64_t cap) {
  if (!s || !npairs) return THP_ERR_ARG;
  RET(thp_nblist(s, 0));
  std::vector<int> cnt(s->npad), nb((size_t)s->maxnb * s->npad);
  HIPCHK(s->ctx, hipMemcpy(cnt.data(), s->view.nbcount, sizeof(int) * s->npad, hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(nb.data(), s->view.nb, sizeof(int) * nb.size(), hipMemcpyDeviceToHost));
  std::vector<std::pair<int, int>> pairs;
  for (int q = 0; q < s->n; ++q) {
    const int a = s->perm[q];
    for (int t = 0; t < cnt[q]; ++t) {
      const int b = s->perm[nb[(size_t)t * s->npad + q]];
      if (a < b) pairs.push_back({a, b});
    }
  }
  std::sort(pairs.begin(), pairs.end());
  *npairs = (int64_t)pairs.size();
  if (!ik) return THP_OK;
  if (cap < (int64_t)pairs.size()) return fail(s->ctx, THP_ERR_ARG, "pair buffer too small (%lld < %zu)", (long long)cap, pairs.size());
  for (size_t m = 0; m < pairs.size(); ++m) {
    ik[2 * m] = pairs[m].first;
    ik[2 * m + 1] = pairs[m].second;
  }
  return THP_OK;
}

thp_status thp_pairs_in_cut(thp_sys* s, int64_t* npairs, int* ik, int64_t cap) {
  if (!s || !npairs) return THP_ERR_ARG;
  // the compacted lists of the current coordinates (the permanent-field
  // pass writes them) and the scaled pairs' coefficients (zero beyond the
  // cutoff)
  RET(ensure_coef(s));
  RET(sync(s));
  const int np = s->npad;
  const SysView& v = s->view;
  const size_t packs = (size_t)v.cmax / kPack;
  std::vector<int> tcount(np), ulist((size_t)s->nblk * v.umax), eoff(s->n + 1);
  std::vector<unsigned short> tlist(packs * np * kPack);
  HIPCHK(s->ctx, hipMemcpy(tcount.data(), v.tcount, sizeof(int) * np, hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(ulist.data(), v.ulist, sizeof(int) * ulist.size(), hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(tlist.data(), v.tlist, sizeof(unsigned short) * tlist.size(), hipMemcpyDeviceToHost));
  HIPCHK(s->ctx, hipMemcpy(eoff.data(), v.excl_off, sizeof(int) * (s->n + 1), hipMemcpyDeviceToHost));
  const int nex = eoff[s->n];
  std::vector<int> ek(std::max(nex, 1));
  std::vector<double> sco(2 * (size_t)std::max(nex, 1));
  if (nex > 0) {
    HIPCHK(s->ctx, hipMemcpy(ek.data(), v.excl_k, sizeof(int) * nex, hipMemcpyDeviceToHost));
    HIPCHK(s->ctx, hipMemcpy(sco.data(), v.scoef, sizeof(double) * 2 * nex, hipMemcpyDeviceToHost));
  }
  // every entry as seen from each of its two atoms; a pair counts only if
  // both atoms hold it
  std::vector<std::pair<int, int>> half[2];
  for (int q = 0; q < s->n; ++q) {
    const int b = q / kBlock;
    for (int e = 0; e < tcount[q]; ++e) {
      const int u = tlist[((size_t)(e / kPack) * np + q) * kPack + e % kPack];
      const int k = ulist[(size_t)b * v.umax + u];
      if (k < 0 || k >= s->n || k == q) return fail(s->ctx, THP_ERR_LIST, "compacted entry %d of atom %d: %d", e, q, k);
      const int a = s->perm[q], c = s->perm[k];
      half[a < c ? 0 : 1].push_back(a < c ? std::make_pair(a, c) : std::make_pair(c, a));
    }
    for (int m = eoff[q]; m < eoff[q + 1]; ++m) {
      if (sco[2 * (size_t)m] == 0.0 && sco[2 * (size_t)m + 1] == 0.0) continue;  // beyond the cutoff
      const int a = s->perm[q], c = s->perm[ek[m]];
      half[a < c ? 0 : 1].push_back(a < c ? std::make_pair(a, c) : std::make_pair(c, a));
    }
  }
  for (auto& h : half) std::sort(h.begin(), h.end());
  if (half[0] != half[1])
    return fail(s->ctx, THP_ERR_LIST, "the compacted lists are not symmetric (%zu vs %zu entries)", half[0].size(),
                half[1].size());
  const auto& pairs = half[0];
  *npairs = (int64_t)pairs.size();
  if (!ik) return THP_OK;
  if (cap < (int64_t)pairs.size())
    return fail(s->ctx, THP_ERR_ARG, "pair buffer too small (%lld < %zu)", (long long)cap, pairs.size());
  for (size_t m = 0; m < pairs.size(); ++m) {
    ik[2 * m] = pairs[m].first;
    ik[2 * m + 1] = pairs[m].second;
  }
  return THP_OK;
}

thp_status thp_efld0_direct(thp_sys* s, double* ef) {
  if (!s || !ef) return THP_ERR_ARG;
  RET(ensure_lists(s));
  RET(pair_efld0(s->ctx, s->one, kGateNone));
  s->coef_valid = true;
  return gather_out(s, ef, s->view.ef, 6);
}

thp_status thp_epolar1(thp_sys* s, double* dep, double* torque) { return thp_epolar1_vir(s, dep, torque, nullptr); }

thp_status thp_epolar1_vir(thp_sys* s, double* dep, double* torque, double* vir) {
  if (!s || !dep) return THP_ERR_ARG;
  RET(epolar1_check(s));
  RET(ensure_lists(s));
  RET(ensure_pme(s));
  if (s->mu_polalg == THP_POLALG_TCG) {
    // the truncated CG's own gradient; no virial for it here
    if (vir) return fail(s->ctx, THP_ERR_ARG, "thp_epolar1_vir: no virial for truncated-CG solves");
    std::vector<double> Fo, To;
    RET(epolar1_tcg_terms(s, Fo, To));
    const size_t np = s->npad;
    HIPCHK(s->ctx, hipMemcpy(s->grad_f, Fo.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice));
    HIPCHK(s->ctx, hipMemcpy(s->grad_t, To.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice));
    const dim3 ga((s->n + 255) / 256);
    if (s->has_frames) {
      hipLaunchKernelGGL((k_ep1_torque<false>), ga, dim3(256), 0, s->ctx->stream, s->d_view, s->d_frames, s->d_tinv_off,
                         s->d_tinv, s->grad_t, s->grad_f, (double*)nullptr);
      RET(launch_check(s->ctx, "k_ep1_torque"));
    }
    RET(gather_out(s, dep, s->grad_f, 3));
    for (size_t q = 0; q < (size_t)3 * s->n; ++q) dep[q] = -dep[q];  // gradient = -force
    if (torque) RET(gather_out(s, torque, s->grad_t, 3));
    return THP_OK;
  }
  const size_t np = s->npad;
  if (!s->grad_f) {
    RET(dalloc(s, &s->grad_f, 3 * np));
    RET(dalloc(s, &s->grad_t, 3 * np));
  }
  const size_t slots = virial_slots(s);
  if (vir && !s->grad_vir) {
    RET(dalloc(s, &s->grad_vir, 6 * slots));
    char* sp = nullptr;
    RET(dalloc(s, &sp, 2 * grid_points(s) * 2 * (s->prec == THP_PREC_FP64 ? sizeof(double) : sizeof(float))));
    s->vir_spec = sp;
  }
  RET(epolar1_launch(s, s->grad_f, s->grad_t, vir ? s->grad_vir : nullptr, vir ? s->vir_spec : nullptr));
  RET(gather_out(s, dep, s->grad_f, 3));
  if (torque) RET(gather_out(s, torque, s->grad_t, 3));
  if (vir) {
    // the shares in their fixed order, FP64, times f/2
    std::vector<double> part(6 * slots);
    HIPCHK(s->ctx, hipMemcpy(part.data(), s->grad_vir, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
    const size_t nb = (size_t)s->nblk, nt = (size_t)(s->n + 255) / 256;
    double v[6] = {0, 0, 0, 0, 0, 0};
    const bool recip = !s->one.real_only;
    for (size_t q = 0; q < slots; ++q) {
      if (!s->has_frames && q >= 2 * nb && q < 2 * nb + nt) continue;                  // no frame forces
      if (!recip && ((q >= nb && q < 2 * nb) || q >= 2 * nb + nt)) continue;           // short range
      for (int c = 0; c < 6; ++c) v[c] += part[6 * q + c];
    }
    const double f2 = 0.5 * kCoulombFp;
    const int idx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) vir[3 * a + b] = f2 * v[idx[a][b]];
  }
  return THP_OK;
}

thp_status thp_efld0_recip(thp_sys* s, double* cphi) {
  if (!s || !cphi) return THP_ERR_ARG;
  RET(ensure_pme(s));
  RET(field(s->ctx, s->one, s->cphi, 0, 0, kGateNone));
  return gather_out(s, cphi, s->cphi, 10);
}

thp_status thp_tmatxb(thp_sys* s, const double* mu, double* efi, int dodiag) {
  if (!s || !mu || !efi) return THP_ERR_ARG;
  RET(ensure_coef(s));
  RET(scatter_in(s, s->view.pp, mu, 6));
  RET(b_pair_tmatxb(s->ctx, s->one, 1 /*pp*/, 3 /*hreal*/, dodiag, 0));
  return gather_out(s, efi, s->view.hreal, 6);
}

thp_status thp_tmatxb_dev(thp_sys* s, const double* mu, double* efi, int dodiag) {
  if (!s || !mu || !efi) return THP_ERR_ARG;
  RET(ensure_coef(s));
  hipLaunchKernelGGL(k_scatter_in, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, s->view.pp, mu, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_scatter_in"));
  RET(b_pair_tmatxb(s->ctx, s->one, 1, 3, dodiag, 0));
  hipLaunchKernelGGL(k_gather_out, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, efi, s->view.hreal, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_gather_out"));
  return sync(s);
}

thp_status thp_tmatxb_recip(thp_sys* s, const double* mu, double* dipfield) {
  if (!s || !mu || !dipfield) return THP_ERR_ARG;
  RET(ensure_pme(s));
  RET(scatter_in(s, s->view.pp, mu, 6));
  RET(spread_uind(s->ctx, s->one, 1, 0, 0));
  RET(convolve(s->ctx, s->one, 0));
  RET(launch_gather_uind<0>(s->ctx, s->one, 1, 3, 0, kGateNone));
  return gather_out(s, dipfield, s->view.hreal, 6);
}

thp_status thp_grid_uind(thp_sys* s, const double* fuind, const double* fuinp, double* qgrid) {
  if (!s || !fuind || !fuinp) return THP_ERR_ARG;
  std::vector<double> f((size_t)s->n * 6);
  for (int a = 0; a < s->n; ++a)
    for (int c = 0; c < 3; ++c) {
      f[(size_t)a * 6 + c] = fuind[(size_t)a * 3 + c];
      f[(size_t)a * 6 + 3 + c] = fuinp[(size_t)a * 3 + c];
    }
  RET(ensure_pme(s));
  RET(scatter_in(s, s->view.pp, f.data(), 6));
  RET(spread_uind(s->ctx, s->one, 1, 1, 0));
  if (qgrid) {
    const size_t cnt = 2 * grid_points(s);
    if (s->prec == THP_PREC_FP64) {
      HIPCHK(s->ctx, hipMemcpyAsync(qgrid, s->view.grid, sizeof(double) * cnt, hipMemcpyDeviceToHost, s->ctx->stream));
      RET(sync(s));
    } else {
      std::vector<float> g(cnt);
      HIPCHK(s->ctx, hipMemcpyAsync(g.data(), s->view.grid, sizeof(float) * cnt, hipMemcpyDeviceToHost, s->ctx->stream));
      RET(sync(s));
      for (size_t q = 0; q < cnt; ++q) qgrid[q] = g[q];
    }
  }
  return sync(s);
}

thp_status thp_pme_convolve(thp_sys* s) {
  if (!s) return THP_ERR_ARG;
  RET(convolve(s->ctx, s->one, 0));
  return sync(s);
}

thp_status thp_fphi_uind(thp_sys* s, const double* qgrid, double* phi1, double* phi2) {
  if (!s || !phi1 || !phi2) return THP_ERR_ARG;
  RET(ensure_pme(s));
  if (qgrid) {
    const size_t cnt = 2 * grid_points(s);
    if (s->prec == THP_PREC_FP64) {
      HIPCHK(s->ctx, hipMemcpyAsync(s->view.grid, qgrid, sizeof(double) * cnt, hipMemcpyHostToDevice, s->ctx->stream));
      RET(sync(s));
    } else {
      std::vector<float> g(qgrid, qgrid + cnt);
      HIPCHK(s->ctx, hipMemcpyAsync(s->view.grid, g.data(), sizeof(float) * cnt, hipMemcpyHostToDevice, s->ctx->stream));
      RET(sync(s));
    }
  }
  RET(launch_gather_uind<0>(s->ctx, s->one, 1, 3, 1, kGateNone));
  std::vector<double> out((size_t)s->n * 6);
  RET(gather_out(s, out.data(), s->view.hreal, 6));
  for (int a = 0; a < s->n; ++a)
    for (int c = 0; c < 3; ++c) {
      phi1[(size_t)a * 3 + c] = out[(size_t)a * 6 + c];
      phi2[(size_t)a * 3 + c] = out[(size_t)a * 6 + 3 + c];
    }
  return THP_OK;
}

// ------------------------------- device pointers, ordered on the caller's stream
// Each takes device arrays in the layouts of the host entry points and a
// stream (hipStream_t; null: the call returns once the device is done).
// With a stream the call enqueues its work behind what the caller queued
// there and makes the stream wait for it, without blocking the host (a list
// or B-spline rebuild due to moved atoms still waits for the device).

#define THP_NEED_DEV_ARGS(cond) \
  do {                          \
    if (!s || !(cond)) return THP_ERR_ARG; \
  } while (0)

}  // extern "C"

namespace {
inline dim3 atoms_grid(const thp_sys* s) { return dim3((s->n + 255) / 256); }

thp_status grid_out_dev(thp_sys* s, double* qgrid) {
  const size_t cnt = 2 * grid_points(s);
  hipStream_t st = s->ctx->stream;
  if (s->prec == THP_PREC_FP64) {
    HIPCHK(s->ctx, hipMemcpyAsync(qgrid, s->view.grid, sizeof(double) * cnt, hipMemcpyDeviceToDevice, st));
  } else {
    hipLaunchKernelGGL((k_grid_widen<float>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, qgrid,
                       static_cast<const float*>(s->view.grid), cnt);
  }
  return launch_check(s->ctx, "grid copy");
}

thp_status grid_in_dev(thp_sys* s, const double* qgrid) {
  const size_t cnt = 2 * grid_points(s);
  hipStream_t st = s->ctx->stream;
  if (s->prec == THP_PREC_FP64) {
    HIPCHK(s->ctx, hipMemcpyAsync(s->view.grid, qgrid, sizeof(double) * cnt, hipMemcpyDeviceToDevice, st));
  } else {
    hipLaunchKernelGGL((k_grid_narrow<float>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                       static_cast<float*>(s->view.grid), qgrid, cnt);
  }
  return launch_check(s->ctx, "grid copy");
}
}  // namespace

extern "C" {

thp_status thp_efld0_direct_dev(thp_sys* s, double* ef, void* stream) {
  THP_NEED_DEV_ARGS(ef);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(s->ctx, cs));
  RET(ensure_lists(s));
  RET(pair_efld0(s->ctx, s->one, kGateNone));
  s->coef_valid = true;
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, s->ctx->stream, ef, s->view.ef, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_gather_out"));
  return order_out(s->ctx, cs);
}

thp_status thp_efld0_recip_dev(thp_sys* s, double* cphi, void* stream) {
  THP_NEED_DEV_ARGS(cphi);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(s->ctx, cs));
  RET(ensure_pme(s));
  RET(field(s->ctx, s->one, s->cphi, 0, 0, kGateNone));
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, s->ctx->stream, cphi, s->cphi, s->iperm_dev, s->n, s->npad, 10);
  RET(launch_check(s->ctx, "k_gather_out"));
  return order_out(s->ctx, cs);
}

thp_status thp_tmatxb_dev_async(thp_sys* s, const double* mu, double* efi, int dodiag, void* stream) {
  THP_NEED_DEV_ARGS(mu && efi);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(s->ctx, cs));
  RET(ensure_coef(s));
  hipLaunchKernelGGL(k_scatter_in, atoms_grid(s), dim3(256), 0, s->ctx->stream, s->view.pp, mu, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_scatter_in"));
  RET(b_pair_tmatxb(s->ctx, s->one, V_PP, V_HREAL, dodiag, 0));
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, s->ctx->stream, efi, s->view.hreal, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_gather_out"));
  return order_out(s->ctx, cs);
}

thp_status thp_tmatxb_recip_dev(thp_sys* s, const double* mu, double* dipfield, void* stream) {
  THP_NEED_DEV_ARGS(mu && dipfield);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(s->ctx, cs));
  RET(ensure_pme(s));
  hipLaunchKernelGGL(k_scatter_in, atoms_grid(s), dim3(256), 0, s->ctx->stream, s->view.pp, mu, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_scatter_in"));
  RET(spread_uind(s->ctx, s->one, V_PP, 0, 0));
  RET(convolve(s->ctx, s->one, 0));
  RET(launch_gather_uind<0>(s->ctx, s->one, V_PP, V_HREAL, 0, kGateNone));
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, s->ctx->stream, dipfield, s->view.hreal, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_gather_out"));
  return order_out(s->ctx, cs);
}

thp_status thp_grid_uind_dev(thp_sys* s, const double* fuind, const double* fuinp, double* qgrid, void* stream) {
  THP_NEED_DEV_ARGS(fuind && fuinp);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(s->ctx, cs));
  RET(ensure_pme(s));
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_scatter_in, atoms_grid(s), dim3(256), 0, st, s->view.pp, fuind, s->iperm_dev, s->n, s->npad, 3);
  hipLaunchKernelGGL(k_scatter_in, atoms_grid(s), dim3(256), 0, st, s->view.pp + (size_t)3 * s->npad, fuinp,
                     s->iperm_dev, s->n, s->npad, 3);
  RET(launch_check(s->ctx, "k_scatter_in"));
  RET(spread_uind(s->ctx, s->one, V_PP, 1, 0));
  if (qgrid) RET(grid_out_dev(s, qgrid));
  return order_out(s->ctx, cs);
}

thp_status thp_fphi_uind_dev(thp_sys* s, const double* qgrid, double* phi1, double* phi2, void* stream) {
  THP_NEED_DEV_ARGS(phi1 && phi2);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(s->ctx, cs));
  RET(ensure_pme(s));
  if (qgrid) RET(grid_in_dev(s, qgrid));
  RET(launch_gather_uind<0>(s->ctx, s->one, V_PP, V_HREAL, 1, kGateNone));
  hipStream_t st = s->ctx->stream;
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, st, phi1, s->view.hreal, s->iperm_dev, s->n, s->npad, 3);
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, st, phi2, s->view.hreal + (size_t)3 * s->npad,
                     s->iperm_dev, s->n, s->npad, 3);
  RET(launch_check(s->ctx, "k_gather_out"));
  return order_out(s->ctx, cs);
}

thp_status thp_epolar1_noalloc(thp_sys* s, double* dep, double* torque, void* stream) {
  THP_NEED_DEV_ARGS(dep);
  thp_ctx* ctx = s->ctx;
  if (s->npad > THP_NOALLOC_MAX_ATOMS)
    return fail(ctx, THP_ERR_ARG, "thp_epolar1_noalloc: %d atoms (%d padded) exceed the scratch of %d atoms", s->n,
                s->npad, THP_NOALLOC_MAX_ATOMS);
  if (!ctx->scratch) return fail(ctx, THP_ERR_STATE, "thp_epolar1_noalloc: no scratch on device %d", ctx->device);
  RET(epolar1_check(s));
  // truncated-CG dipoles are not stationary: their gradient takes the host
  // adjoint of thp_epolar1, which allocates
  if (s->mu_polalg == THP_POLALG_TCG)
    return fail(ctx, THP_ERR_ARG, "thp_epolar1_noalloc: not for truncated-CG solves");
  // a solve of the current coordinates left the lists and B-spline tables
  // in place; rebuilding them here would allocate
  if (!s->list_valid || !s->tight_valid || (!s->pme_valid && !s->one.real_only))
    return fail(ctx, THP_ERR_STATE, "thp_epolar1_noalloc: lists or B-spline tables are not current");
  hipStream_t cs = static_cast<hipStream_t>(stream);
  RET(order_in(ctx, cs));
  double* F = ctx->scratch;
  double* T = F + (size_t)3 * s->npad;
  RET(epolar1_launch(s, F, T));
  hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, ctx->stream, dep, F, s->iperm_dev, s->n, s->npad, 3);
  if (torque)
    hipLaunchKernelGGL(k_gather_out, atoms_grid(s), dim3(256), 0, ctx->stream, torque, T, s->iperm_dev, s->n, s->npad, 3);
  RET(launch_check(ctx, "k_gather_out"));
  return order_out(ctx, cs);
}

#undef THP_NEED_DEV_ARGS

}  // extern "C"

// ============================================================ PCG driver
namespace thp {

// One solve of one system: the reference reads the residual norm back on
// the host after every iteration (newinduce_pme2gpu.f:1040-1041), and so
// does this one-system path.
// guess of the next solve from the stored solves (newinduce_pme2.f:165-178):
// the least-squares coefficients are fitted on the device first
thp_status predict_guess(thp_sys* s) {
  thp_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  if (s->pred == THP_PRED_LSQR) {
    hipLaunchKernelGGL(k_pred_normal, g1(s->pred_blocks), dim3(256), 0, st, s->d_view, s->pred_hist, s->pred_newest,
                       s->pred_part);
    hipLaunchKernelGGL(k_pred_solve, g1(1), dim3(64), 0, st, s->pred_part, s->pred_blocks, s->pred_coef);
  }
  const int nb = (int)std::min<size_t>(1024, ((size_t)6 * s->npad + 255) / 256);
  hipLaunchKernelGGL(k_pred_apply, g1(nb), dim3(256), 0, st, s->d_view, s->pred_hist, s->pred_coef, s->pred_newest);
  return launch_check(ctx, "k_pred_apply");
}

// store the solve's dipoles as the newest history entry (newinduce_pme2.f:253-264)
thp_status push_history(thp_sys* s) {
  const int slot = s->pred_n == 0 ? 0 : (s->pred_newest + 1) % kMaxUalt;
  const int nb = (int)std::min<size_t>(1024, ((size_t)6 * s->npad + 255) / 256);
  hipLaunchKernelGGL(k_pred_push, g1(nb), dim3(256), 0, s->ctx->stream, s->d_view, s->pred_hist, slot);
  RET(launch_check(s->ctx, "k_pred_push"));
  s->pred_newest = slot;
  s->pred_n = std::min(s->pred_n + 1, kMaxUalt);
  return THP_OK;
}

// One Jacobi/DIIS iteration (inducejac_pme2, newinduce_pme2.f:731-846)
// with the host's mirror of the subspace bookkeeping of diis.f: the new
// vector is stored at subspace position nmat, and a full subspace drops its
// ten oldest vectors first.
thp_status jacobi_iteration(thp_sys* s, int it, int* nmat, int* first) {
  thp_ctx* ctx = s->ctx;
  const Batch& b = s->one;
  hipStream_t st = ctx->stream;
  RET(b_pair_tmatxb(ctx, b, V_MU, V_HREAL, 0, kGateNone));
  if (!b.real_only) {
    RET(spread_uind(ctx, b, V_MU, 0, kGateNone));
    RET(convolve(ctx, b, kGateNone));
  }
  // without the reciprocal terms the gather writes a zero field into h
  RET(launch_gather_uind<0>(ctx, b, V_MU, V_H, 0, kGateNone));
  const int slot = (*first + *nmat - 1) % kDiisMax;
  int shifted = 0;
  if (*nmat >= kDiisMax) {
    *first = (*first + 10) % kDiisMax;
    *nmat -= 10;
    shifted = 1;
  }
  const int nvec = *nmat;
  hipLaunchKernelGGL(k_jac_step, g1(s->diis_blocks), dim3(256), 0, st, s->d_view, s->diis_x, s->diis_e, slot, *first,
                     nvec, s->diis_part);
  hipLaunchKernelGGL(k_jac_solve, g1(1), dim3(64), 0, st, s->d_view, s->diis, s->diis_part, s->diis_blocks, nvec, shifted,
                     it);
  const int nb = (int)std::min<size_t>(1024, ((size_t)6 * s->npad + 255) / 256);
  hipLaunchKernelGGL(k_jac_extrap, g1(nb), dim3(256), 0, st, s->d_view, s->diis_x, s->diis, *first, nvec);
  *nmat += 1;
  return launch_check(ctx, "k_jac_extrap");
}

// A v for the truncated CG: the matvec with its diagonal into hreal, the
// reciprocal dipole field into h (the kernels add them and the self term)
thp_status tcg_apply(thp_sys* s, int vin) {
  thp_ctx* ctx = s->ctx;
  const Batch& b = s->one;
  RET(b_pair_tmatxb(ctx, b, vin, V_HREAL, 1, kGateNone));
  if (!b.real_only) {
    RET(spread_uind(ctx, b, vin, 0, kGateNone));
    RET(convolve(ctx, b, kGateNone));
  }
  // short range: the gather writes a zero field into h (A = 1/alpha - T
  // within the short cutoff, no self term)
  return launch_gather_uind<0>(ctx, b, vin, V_H, 0, kGateNone);
}

// epolar3tcg1 / epolar3tcg2 after the permanent field (and the guess
// alpha ef in mu) are in place
thp_status tcg_solve(thp_sys* s) {
  thp_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  const int nb = s->tcg_blocks;
  double* part_n0 = s->tcg_part;
  double* part_t1 = s->tcg_part + (size_t)2 * nb;
  double* part_np1 = s->tcg_part + (size_t)4 * nb;
  double* part_sp = s->tcg_part + (size_t)6 * nb;
  if (s->tcg_guess) RET(tcg_apply(s, V_MU));
  hipLaunchKernelGGL(k_tcg_start, g1(nb), dim3(256), 0, st, s->d_view, s->tcg_guess, s->tcg_prec, part_n0);
  RET(tcg_apply(s, V_ZR));
  hipLaunchKernelGGL(k_tcg_apply1, g1(nb), dim3(256), 0, st, s->d_view, s->tcg_prec, s->tcg_vec, part_t1, part_np1);
  hipLaunchKernelGGL(k_tcg_scalars1, g1(1), dim3(64), 0, st, part_n0, part_t1, part_np1, nb, s->tcg_sc);
  if (s->tcg_order == 2) {
    RET(tcg_apply(s, V_PP));
    hipLaunchKernelGGL(k_tcg_apply2, g1(nb), dim3(256), 0, st, s->d_view, s->tcg_prec, s->tcg_vec, s->tcg_sc, part_sp);
    hipLaunchKernelGGL(k_tcg_scalars2, g1(1), dim3(64), 0, st, part_sp, nb, s->tcg_sc);
  }
  hipLaunchKernelGGL(k_tcg_final, g1(nb), dim3(256), 0, st, s->d_view, s->tcg_vec, s->tcg_sc, s->tcg_order,
                     s->tcg_prec, s->tcg_guess, s->tcg_peek, s->tcg_omega);
  hipLaunchKernelGGL(k_tcg_done, g1(1), dim3(64), 0, st, s->d_view, s->tcg_order);
  return launch_check(ctx, "k_tcg");
}

thp_status run_solve(thp_sys* s, const thp_solve_opts* o, bool warm, thp_solve_stats* stats) {
  hipStream_t st = s->ctx->stream;
  thp_ctx* ctx = s->ctx;
  const Batch& b = s->one;
  HIPCHK(ctx, hipEventRecord(s->ev[0], st));
  RET(ensure_lists(s));
  RET(ensure_pme(s));
  const bool tcg = s->polalg == THP_POLALG_TCG;
  // the truncated CG starts from alpha ef or zero, whatever opts->guess
  // and the predictor say (epolar3tcg.f:194-218)
  if (tcg) warm = false;
  // with a predictor and a full history the extrapolated guess replaces
  // opts->guess, as use_pred precedes polgsf in the reference
  // (the short-range solve keeps no predictor history here: the
  // reference's is a separate one, udshortalt)
  const bool short_range = s->short_cut > 0.0;
  if (!tcg && !short_range && s->pred != THP_PRED_NONE && s->pred_n == kMaxUalt) {
    RET(predict_guess(s));
    warm = true;
  }
  RET(b_init_state(ctx, b, o->poleps, o->politer, warm ? s->d_one : nullptr, nullptr));
  const bool jacobi = s->polalg == THP_POLALG_JACOBI_DIIS;
  if (jacobi || tcg) RET(b_setup_field(ctx, b, kGateNone));
  else RET(b_setup(ctx, b, kGateNone));
  s->coef_valid = true;
  HIPCHK(ctx, hipEventRecord(s->ev[1], st));
  SolveState fin{};
  if (tcg) RET(tcg_solve(s));
  if (jacobi) {
    // Jacobi/DIIS: a host step after every iteration
    int conv = 0, nmat = 1, first = 0;
    for (int it = 1; it <= o->politer; ++it) {
      RET(jacobi_iteration(s, it, &nmat, &first));
      HIPCHK(ctx, hipMemcpyAsync(&conv, &s->view.st->conv, sizeof(int), hipMemcpyDeviceToHost, st));
      RET(sync(s));
      if (conv) break;
    }
  } else if (!tcg) {
    // PCG: k iterations between host reads of the solve state.  The
    // kernels of an iteration skip a solve that has stopped (converged,
    // politer reached, non-finite), so iterations queued past its end
    // change nothing.  FP64 reads after every iteration, as the reference
    // tests convergence after each (THP_SOLVE_CHUNK changes that).  Mixed
    // solves queue the last solve's iteration count before the first read
    // and finish inside every read (k_energy and k_finish act only on a
    // stopped solve), so that the read that sees the stop also brings the
    // energy.
    const bool fp64 = s->prec == THP_PREC_FP64;
    const int chunk = ctx->solve_chunk[fp64 ? 0 : 1];
    int k = fp64 ? chunk : std::max(chunk, std::min(s->last_niter, kFirstChunkMax));
    for (int done = 0; done < o->politer; k = chunk) {
      k = std::min(k, o->politer - done);
      for (int q = 0; q < k; ++q) RET(b_iteration(ctx, b));
      done += k;
      if (!fp64) {
        RET(b_finish(ctx, b, 1));
        HIPCHK(ctx, hipEventRecord(s->ev[2], st));
      }
      HIPCHK(ctx, hipMemcpyAsync(&fin, s->view.st, sizeof fin, hipMemcpyDeviceToHost, st));
      RET(sync(s));
      if (fin.conv != kRunning) break;
    }
  }
  const bool finished = fin.finalized != 0;  // the loop's last read came after the finish
  if (!finished) RET(b_finish(ctx, b, 0));
  if (s->pred != THP_PRED_NONE && !short_range) RET(push_history(s));
  if (!finished) {
    HIPCHK(ctx, hipEventRecord(s->ev[2], st));
    HIPCHK(ctx, hipMemcpyAsync(&fin, s->view.st, sizeof fin, hipMemcpyDeviceToHost, st));
  }
  RET(sync(s));
  if (!tcg && !jacobi) s->last_niter = fin.niter;
  s->mu_valid = true;
  s->mu_polalg = s->polalg;
  s->mu_short = short_range;
  s->mu_tcg[0] = s->tcg_order;
  s->mu_tcg[1] = s->tcg_prec;
  s->mu_tcg[2] = s->tcg_guess;
  s->mu_tcg[3] = s->tcg_peek;
  s->mu_tcg_omega = s->tcg_omega;
  if (stats) {
    float t1 = 0, t2 = 0;
    (void)hipEventElapsedTime(&t1, s->ev[0], s->ev[1]);
    (void)hipEventElapsedTime(&t2, s->ev[1], s->ev[2]);
    stats->niter = fin.niter;
    stats->resnrm = fin.resnrm;
    stats->ep = fin.ep;
    stats->n_indefinite = fin.n_indef;
    stats->t_field_ms = t1;
    stats->t_solve_ms = t2;
    stats->status = fin.conv == kConverged || fin.conv == kZeroCurv ? THP_OK : fin.conv == kNonFinite ? THP_ERR_NAN : THP_ERR_NOCONV;
  }
  return THP_OK;
}

}  // namespace thp

extern "C" {

thp_status thp_set_solver(thp_sys* s, int polalg) {
  using namespace thp;
  if (!s) return THP_ERR_ARG;
  if (polalg != THP_POLALG_PCG && polalg != THP_POLALG_JACOBI_DIIS && polalg != THP_POLALG_TCG)
    return fail(s->ctx, THP_ERR_ARG, "unknown polalg %d (1: PCG, 2: Jacobi/DIIS, 3: truncated CG)", polalg);
  if (polalg == THP_POLALG_TCG && !s->tcg_vec) {
    s->tcg_blocks = (s->n + 255) / 256;
    RET(dalloc(s, &s->tcg_vec, (size_t)2 * 6 * s->npad));
    RET(dalloc(s, &s->tcg_part, (size_t)8 * s->tcg_blocks));
    RET(dalloc(s, &s->tcg_sc, (size_t)kTcgScalars));
  }
  if (polalg == THP_POLALG_JACOBI_DIIS && !s->diis) {
    s->diis_blocks = (s->n + 255) / 256;
    RET(dalloc(s, &s->diis_x, (size_t)kDiisMax * 6 * s->npad));
    RET(dalloc(s, &s->diis_e, (size_t)kDiisMax * 6 * s->npad));
    RET(dalloc(s, &s->diis_part, (size_t)s->diis_blocks * kDiisParts));
    RET(dalloc(s, &s->diis, 1));
  }
  s->polalg = polalg;
  return THP_OK;
}

thp_status thp_set_shortreal(thp_sys* s, double shortcut) {
  using namespace thp;
  if (!s) return THP_ERR_ARG;
  if (!(shortcut >= 0.0) || shortcut > s->cut)
    return fail(s->ctx, THP_ERR_ARG, "short-range cutoff %g must lie in (0, %g] (0 turns the mode off)", shortcut, s->cut);
  const bool on = shortcut > 0.0;
  s->short_cut = shortcut;
  s->one.real_only = on ? 1 : 0;
  // the block-union lists (built for the full cutoff plus the skin) hold
  // every short-range pair: only the exact cutoff test and the self term
  // change, and the per-solve coefficients are rebuilt
  s->view.cut2 = on ? shortcut * shortcut : s->cut * s->cut;
  s->view.term_self = on ? 0.0 : (4.0 / 3.0) * s->aewald * s->aewald * s->aewald / std::sqrt(M_PI);
  s->coef_valid = false;
  RET(upload_view(s));
  return sync(s);
}

thp_status thp_set_tcg(thp_sys* s, int order, int precond, int guess, int peek, double omega) {
  using namespace thp;
  if (!s) return THP_ERR_ARG;
  if (order != 1 && order != 2) return fail(s->ctx, THP_ERR_ARG, "tcg order %d (1 or 2)", order);
  if (!std::isfinite(omega)) return fail(s->ctx, THP_ERR_ARG, "tcg omega must be finite");
  s->tcg_order = order;
  s->tcg_prec = precond ? 1 : 0;
  s->tcg_guess = guess ? 1 : 0;
  s->tcg_peek = peek ? 1 : 0;
  s->tcg_omega = omega;
  return THP_OK;
}

thp_status thp_set_predictor(thp_sys* s, int method) {
  using namespace thp;
  if (!s) return THP_ERR_ARG;
  if (method < THP_PRED_NONE || method > THP_PRED_LSQR) return fail(s->ctx, THP_ERR_ARG, "unknown predictor %d", method);
  if (method != THP_PRED_NONE && !s->pred_hist) {
    s->pred_blocks = (s->n + 255) / 256;
    RET(dalloc(s, &s->pred_hist, (size_t)kMaxUalt * 6 * s->npad));
    RET(dalloc(s, &s->pred_coef, (size_t)kMaxUalt));
    RET(dalloc(s, &s->pred_part, (size_t)s->pred_blocks * kPredPairs));
  }
  // Gear binomial and ASPC coefficients, newest entry first (kpolar.f:405-423)
  static const double gear[kMaxUalt] = {6.0, -15.0, 20.0, -15.0, 6.0, -1.0, 0.0};
  static const double aspc[kMaxUalt] = {22.0 / 7.0, -55.0 / 14.0, 55.0 / 21.0, -22.0 / 21.0, 5.0 / 21.0, -1.0 / 42.0, 0.0};
  if (method == THP_PRED_GEAR || method == THP_PRED_ASPC)
    HIPCHK(s->ctx, hipMemcpy(s->pred_coef, method == THP_PRED_GEAR ? gear : aspc, sizeof gear, hipMemcpyHostToDevice));
  s->pred = method;
  s->pred_n = 0;
  s->pred_newest = 0;
  return THP_OK;
}

thp_status thp_induce(thp_sys* s, const thp_solve_opts* o, double* uind, double* uinp, double* ep,
                      thp_solve_stats* stats) {
  if (!s || !o || !uind || !uinp) return THP_ERR_ARG;
  if (!(o->poleps > 0) || o->politer <= 0) return fail(s->ctx, THP_ERR_ARG, "bad poleps/politer");
  const bool warm = o->guess == 1;
  if (warm) {
    std::vector<double> m((size_t)s->n * 6);
    for (int a = 0; a < s->n; ++a)
      for (int c = 0; c < 3; ++c) {
        m[(size_t)a * 6 + c] = uind[(size_t)a * 3 + c];
        m[(size_t)a * 6 + 3 + c] = uinp[(size_t)a * 3 + c];
      }
    RET(scatter_in(s, s->view.mu, m.data(), 6));
  }
  thp_solve_stats local;
  thp_solve_stats* sp = stats ? stats : &local;
  RET(run_solve(s, o, warm, sp));
  RET(gather_to_hstage(s, s->view.mu, 6));
  const double* m = s->hstage;
  for (int a = 0; a < s->n; ++a)
    for (int c = 0; c < 3; ++c) {
      uind[(size_t)a * 3 + c] = m[(size_t)a * 6 + c];
      uinp[(size_t)a * 3 + c] = m[(size_t)a * 6 + 3 + c];
    }
  if (ep) *ep = sp->ep;
  return sp->status == THP_ERR_NAN ? fail(s->ctx, THP_ERR_NAN, "non-finite residual") : THP_OK;
}

thp_status thp_tcg_omega_fit(thp_sys* s, double* omega, int* fitted) {
  if (!s || !omega) return THP_ERR_ARG;
  if (s->polalg != THP_POLALG_TCG || !s->tcg_peek)
    return fail(s->ctx, THP_ERR_STATE, "thp_tcg_omega_fit: the truncated CG with the peek step is not selected");
  if (s->short_cut > 0.0) return fail(s->ctx, THP_ERR_STATE, "thp_tcg_omega_fit: not in short-range mode");
  // the reference energy: the PCG solve at poleps 1e-10 (kpolar.f:1107-1110)
  thp_solve_opts o{1e-10, 500, 0};
  thp_solve_stats st{};
  s->polalg = THP_POLALG_PCG;
  const thp_status rc = run_solve(s, &o, false, &st);
  s->polalg = THP_POLALG_TCG;
  RET(rc);
  if (st.status != THP_OK) return fail(s->ctx, st.status, "thp_tcg_omega_fit: the reference PCG solve failed");
  const double epcg = st.ep;
  // E_TCG(omega) = -f/2 <mu_d(omega), E_p> is linear in omega: the peek
  // step adds omega times a fixed vector.  Its two coefficients, -f/2 <munp,
  // E_p> and -f/2 <peek direction, E_p> (cnp and cr of omegafit), from the
  // solves at omega = 0 and 1.
  const double keep = s->tcg_omega;
  double e[2];
  for (int k = 0; k < 2; ++k) {
    s->tcg_omega = double(k);
    thp_solve_stats t{};
    const thp_status r2 = run_solve(s, &o, false, &t);
    if (r2 != THP_OK) {
      s->tcg_omega = keep;
      return r2;
    }
    e[k] = t.ep;
  }
  s->tcg_omega = keep;
  // omegafit (kpolar.f:1059-1326): bisection over [-5, 5] toward the PCG
  // energy, 20 steps, within max(1e-4 |E_CG|, 1e-6) kcal/mol; no fit when the
  // peek step's energy is itself that small
  const double enp = e[0], er = e[1] - e[0];
  double cvg = std::fabs(epcg * 1e-4);
  if (cvg <= 1e-6) cvg = 1e-6;
  double o1 = -5.0, o2 = 5.0;
  double ep1 = enp + o1 * er, ep2 = enp + o2 * er;
  bool cvged = false;
  double w = keep;
  if (std::fabs(er) > cvg) {
    for (int it = 0; it < 20; ++it) {
      if (std::fabs(epcg - ep1) <= cvg) {
        w = o1;
        cvged = true;
        break;
      } else if (std::fabs(epcg - ep2) <= cvg) {
        w = o2;
        cvged = true;
        break;
      }
      if (std::fabs(epcg - ep1) >= std::fabs(epcg - ep2)) {
        o1 = o1 + (o2 - o1) * 0.5;
        ep1 = enp + o1 * er;
      } else {
        o2 = o2 - (o2 - o1) * 0.5;
        ep2 = enp + o2 * er;
      }
    }
  }
  if (cvged) s->tcg_omega = w;
  *omega = s->tcg_omega;
  if (fitted) *fitted = cvged ? 1 : 0;
  s->mu_valid = false;  // the last solve was a fitting solve
  return THP_OK;
}

thp_status thp_time_kernel(thp_sys* s, int kernel, int reps, double* ms, int64_t* pairs) {
  if (!s) return THP_ERR_ARG;
  return thp_time_kernel_batch(&s, 1, kernel, reps, ms, pairs);
}

thp_status thp_time_kernel_batch(thp_sys* const* systems, int nsys, int kernel, int reps, double* ms, int64_t* pairs) {
  if (!systems || nsys <= 0 || !ms || reps <= 0 || kernel < 0 || kernel > 6) return THP_ERR_ARG;
  thp_sys* s0 = systems[0];
  if (!s0) return THP_ERR_ARG;
  thp_ctx* ctx = s0->ctx;
  if ((kernel == 4 || kernel == 5) && nsys != 1) return fail(ctx, THP_ERR_ARG, "list rebuilds are timed one system at a time");
  for (int q = 0; q < nsys; ++q) {
    if (!systems[q] || systems[q]->ctx != ctx || systems[q]->prec != s0->prec)
      return fail(ctx, THP_ERR_ARG, "batch members must share one context and one precision");
    for (int r = 0; r < q; ++r)
      if (systems[r] == systems[q]) return fail(ctx, THP_ERR_ARG, "system %d appears twice in the batch", q);
  }
  for (int q = 0; q < nsys; ++q) {
    RET(ensure_coef(systems[q]));
    RET(ensure_pme(systems[q]));
  }
  hipStream_t st = ctx->stream;
  // a batch of one is the system's own; several get views and work tables
  // laid out as the serving loop lays them out (XCD-ordered blocks)
  Batch b = s0->one;
  void* d_mem = nullptr;
  if (nsys > 1) {
    std::vector<Work> blk, red, gat, brick, conv, fft;
    for (int q = 0; q < nsys; ++q) append_work(systems[q], q, blk, red, gat, brick, conv, fft);
    blk = xcd_order(blk);
    brick = xcd_order(brick);
    gat = xcd_order(gat);
    std::vector<SysView> views(nsys);
    for (int q = 0; q < nsys; ++q) views[q] = systems[q]->view;
    const size_t nt = blk.size() + red.size() + gat.size() + brick.size() + conv.size() + fft.size();
    const size_t vbytes = ((sizeof(SysView) * nsys + 255) / 256) * 256;
    const size_t ibytes = ((sizeof(int) * nsys + 255) / 256) * 256;
    HIPCHK(ctx, hipMalloc(&d_mem, vbytes + ibytes + sizeof(Work) * nt));
    SysView* d_views = static_cast<SysView*>(d_mem);
    int* d_ids = reinterpret_cast<int*>(static_cast<char*>(d_mem) + vbytes);
    Work* d_tab = reinterpret_cast<Work*>(static_cast<char*>(d_mem) + vbytes + ibytes);
    std::vector<Work> tab;
    tab.reserve(nt);
    auto put = [&](const std::vector<Work>& v, Work** dptr, int* cnt) {
      *dptr = d_tab + tab.size();
      *cnt = (int)v.size();
      tab.insert(tab.end(), v.begin(), v.end());
    };
    b.nsys = nsys;
    b.rc = 0;
    for (int q = 0; q < nsys; ++q) b.rc |= systems[q]->view.rc_sqa > 0.0 ? 1 : 2;
    put(blk, &b.d_blk, &b.n_blk);
    put(red, &b.d_red, &b.n_red);
    put(gat, &b.d_gat, &b.n_gat);
    put(brick, &b.d_brick, &b.n_brick);
    put(conv, &b.d_conv, &b.n_conv);
    put(fft, &b.d_fft, &b.n_fft);
    b.d_views = d_views;
    b.d_ids = d_ids;
    std::vector<const int*> nfs(nsys);
    for (int q = 0; q < nsys; ++q) nfs[q] = systems[q]->nfft;
    plane_fft_setup(b, nfs.data(), nsys, s0->prec);
    std::vector<int> ids(nsys);
    for (int q = 0; q < nsys; ++q) ids[q] = q;
    thp_status rc = THP_OK;
    if (hipMemcpy(d_views, views.data(), sizeof(SysView) * nsys, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_ids, ids.data(), sizeof(int) * nsys, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tab, tab.data(), sizeof(Work) * nt, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(ctx, THP_ERR_HIP, "batch tables");
    if (rc != THP_OK) {
      (void)hipFree(d_mem);
      return rc;
    }
  }
  auto launch = [&]() -> thp_status {
    if (kernel == 0) return b_pair_tmatxb(ctx, b, 1, 3, 1, 0);
    if (kernel == 1) return spread_uind(ctx, b, 1, 0, 0);
    if (kernel == 3) return pair_efld0(ctx, b, kGateNone);  // rewrites the same lists
    if (kernel == 4) {
      // the solver's lists from scratch: bounding boxes, candidate block
      // pairs and block-union lists (host round trips for the counts
      // included), then the compacted lists of the permanent-field pass
      RET(build_block_lists(s0));
      RET(build_union(s0));
      s0->tight_valid = true;
      return pair_efld0(ctx, b, kGateNone);
    }
    if (kernel == 5) {
      s0->verlet_valid = false;
      return thp_nblist(s0, 1);  // the reference's Verlet list (mlistcell)
    }
    if (kernel == 6) return convolve(ctx, b, kGateNone);  // forward FFT, qfac, backward FFT
    return launch_gather_uind<0>(ctx, b, 1, 3, 0, kGateNone);
  };
  thp_status rc = launch();  // warm
  if (rc == THP_OK && hipEventRecord(s0->ev[0], st) != hipSuccess) rc = fail(ctx, THP_ERR_HIP, "event");
  for (int r = 0; rc == THP_OK && r < reps; ++r) rc = launch();
  float t = 0;
  if (rc == THP_OK && (hipEventRecord(s0->ev[1], st) != hipSuccess || hipEventSynchronize(s0->ev[1]) != hipSuccess ||
                       hipEventElapsedTime(&t, s0->ev[0], s0->ev[1]) != hipSuccess))
    rc = fail(ctx, THP_ERR_HIP, "kernel timing");
  if (d_mem) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_mem);
  }
  RET(rc);
  *ms = double(t) / reps;
  if (pairs) {
    *pairs = 0;
    for (int q = 0; q < nsys; ++q) {
      int64_t c = 0;
      RET(count_pairs(systems[q], &c));
      *pairs += c;
    }
  }
  return THP_OK;
}

thp_status thp_induce_mu(thp_sys* s, const thp_solve_opts* o, double* mu, double* ep, thp_solve_stats* stats) {
  if (!s || !o || !mu) return THP_ERR_ARG;
  if (!(o->poleps > 0) || o->politer <= 0) return fail(s->ctx, THP_ERR_ARG, "bad poleps/politer");
  const bool warm = o->guess == 1;
  if (warm) RET(scatter_in(s, s->view.mu, mu, 6));
  thp_solve_stats local;
  thp_solve_stats* sp = stats ? stats : &local;
  RET(run_solve(s, o, warm, sp));
  RET(gather_out(s, mu, s->view.mu, 6));
  if (ep) *ep = sp->ep;
  return sp->status == THP_ERR_NAN ? fail(s->ctx, THP_ERR_NAN, "non-finite residual") : THP_OK;
}

thp_status thp_induce_dev(thp_sys* s, const thp_solve_opts* o, double* mu_dev, double* ep, thp_solve_stats* stats) {
  if (!s || !o || !mu_dev) return THP_ERR_ARG;
  const bool warm = o->guess == 1;
  if (warm) {
    hipLaunchKernelGGL(k_scatter_in, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, s->view.mu, mu_dev, s->iperm_dev, s->n, s->npad, 6);
    RET(launch_check(s->ctx, "k_scatter_in"));
  }
  thp_solve_stats local;
  thp_solve_stats* sp = stats ? stats : &local;
  RET(run_solve(s, o, warm, sp));
  hipLaunchKernelGGL(k_gather_out, dim3((s->n + 255) / 256), dim3(256), 0, s->ctx->stream, mu_dev, s->view.mu, s->iperm_dev, s->n, s->npad, 6);
  RET(launch_check(s->ctx, "k_gather_out"));
  RET(sync(s));
  if (ep) *ep = sp->ep;
  return THP_OK;
}

}  // extern "C"
