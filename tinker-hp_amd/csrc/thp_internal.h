// Internal definitions shared by the libthp host and device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <array>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/thp.h"

namespace thp {

constexpr int kBlock = 64;  // atoms per block = one wavefront
constexpr int kMaxUalt = 7;  // stored solves of the dipole predictor (MOD_uprior.f:34)
constexpr int kPredPairs = kMaxUalt * (kMaxUalt + 1) / 2;
constexpr int kDiisMax = 25;              // DIIS subspace (ndismx, newinduce_pme2.f:664)
constexpr int kDiisParts = 2 + kDiisMax;  // per-block partials of a Jacobi step
constexpr int kPack = 8;    // compacted list entries per lane pack (16 B of indices)
// Ewald functions of the recomputed matvec coefficients, tabulated in
// x = aewald * r over [0, kGtabN / kGtabScale): per interval the cubic
// coefficients of G1 = r^3 bn1 and then of G2 = r^5 bn2
constexpr int kGtabN = 256;
constexpr float kGtabScale = 32.0f;  // intervals per unit of x
constexpr int kFirstChunkMax = 48;  // PCG iterations queued before a mixed solve's first state read, at most
constexpr int kOrder = 5;   // PME B-spline order handled by the kernels
constexpr double kCoulomb = 332.063714;  // v1.2/source/MOD_units.f:81
constexpr double kTinyPol = 1.0e-5;      // v1.2/source/MOD_polar.f:36
constexpr int kMaxAtoms = 1 << 24;

// PCG termination codes held in SolveState::conv
enum : int { kRunning = 0, kConverged = 1, kMaxIter = 2, kNonFinite = 3, kZeroCurv = 4 };

// Which systems of a launch a kernel works on: all, those iterating (a
// solve in progress and its set-up done), or those in set-up.
enum : int { kGateNone = 0, kGateIter = 1, kGateSetup = 2 };

// Device-resident state of one PCG solve (all scalars FP64 on the device).
// Pulay's matrix (Lagrange row and column 0) and the extrapolation
// coefficients of the Jacobi/DIIS solver
struct DiisState {
  double B[kDiisMax + 1][kDiisMax + 1];
  double cex[kDiisMax + 1];
};

struct SolveState {
  double ggold[2];
  double gg[2];
  double ggnew[2];
  double resnrm;
  double poleps;
  double ep;
  int niter;    // completed iterations (written by k_cg_update2 block 0)
  int it_cur;   // iteration in progress (written by k_cg_update1 block 0)
  int politer;
  int conv;
  int n_indef;
  int warm;     // 1: mu holds the starting dipoles (no direct guess)
  int finalized;  // 1 once k_finish has written ep of a finished solve
  int in_setup;   // 1 between k_init_state (requested) and k_end_setup
  int stale;      // 1 if the coordinates moved past the block-union lists' skin
};

// Everything a kernel needs about one system; lives in device memory so a
// batched launch can reach many systems through one array.
struct SysView {
  int n, npad, nblk;
  int red_blocks;  // blocks per system in the PCG vector kernels (partials)
  int gat_blocks;  // blocks per system in the PME gather kernels (partials)
  int prec;
  double box[3];
  double aewald, cut2, mbuf2, term_self;
  double cut, skin;   // real-space cutoff; skin of the block-union lists
  double ulim2;       // (cut + skin)^2 of the block-union lists
  double afrac[3];  // nfft_k / box_k (orthorhombic fractional scaling)
  int nfft[3];
  int fft_cap;        // FFT lines per workgroup, at most (fft_group_lines)
  int ncode;
  double ctf[10][10];  // Cartesian -> fractional multipoles (cart_to_frac)
  double ftc[10][10];  // fractional -> Cartesian potential (frac_to_cart)
  // sorted-order per-atom arrays
  const double* x;
  const double* y;
  const double* z;
  // device-side coordinate loading: caller-order coordinates [3][n], the
  // sorted slot -> caller index map, and the sorted coordinates at the last
  // block-union build [3][npad] with the squared displacement they allow
  const double* raw;
  const double* raw_rpole;  // [n][13] caller-order multipoles (serving systems), else null
  // local frames (thp_system_set_frames): [npad] x (axis type, z, x, y) in
  // sorted slots and the local multipoles [10][npad]; null without frames
  const int* frames;
  const double* pole_local;
  const int* perm;
  const double* refu;
  double lim_u2;
  const double* rpole;  // [10][npad]: q, dx,dy,dz, qxx,qxy,qxz,qyy,qyz,qzz
  const double* pol;
  const double* pdamp;
  const double* thole;
  // matvec coefficients recomputed per entry (FP32 systems with one Thole
  // value a > 0): sqrt(a log2 e), else 0 (coefficients stored per solve);
  // gtab is the cubic table of the Ewald functions of the recomputation
  double rc_sqa;
  const float* gtab;  // [kGtabN][8]
  const double* scale_tab;  // [ncode][3] d,p,u
  const int* excl_off;      // [n+1]
  const int* excl_k;
  const int* excl_code;
  double* scoef;            // [nexcl][2] per-solve (dmp3, dmp5) of the scaled pairs, matvec u-scale
  // Verlet list (cut + buffer), full, sorted k per atom: nb[s*npad + i]
  int maxnb;
  int* nb;
  int* nbcount;
  double* bbox;       // [nblk][6] block bounding boxes (centre, half extent)
  // block-union lists (cut, unscaled pairs)
  int umax, lmax;
  int* ulist;              // [nblk][umax] union atoms of each block
  int* ucount;             // [nblk]
  unsigned short* llist;   // [lmax][npad] local partner indices into the block's union
  int* lcount;             // [npad]
  int cmax;  // longest local list, rounded up to whole lane packs
  // per-solve compacted lists, written by the permanent-field pass: the
  // in-cutoff entries of llist, in order, with the coefficients (dmp3, dmp5)
  // of the real-space dipole field (pair type), in lane packs of kPack
  // entries: entry e of
  // atom i sits at ((e / kPack) * npad + i) * kPack + e % kPack, and the
  // last pack of an atom is padded with index 0 and zero coefficients
  // (with recomputed coefficients: no tcoef, padding index = the block's
  // union count)
  unsigned short* tlist;   // [cmax / kPack][npad][kPack]
  void* tcoef;             // [cmax / kPack][npad][kPack] coefficient pairs of tlist, or null
  int* tcount;             // [npad]
  // PME
  int brick[3];       // spreading brick edge per axis (4..8 points)
  int nbrick[3];
  void* theta;        // [npad][48] B-spline values per atom (real type)
  int* g0;            // [npad][4] first grid node per axis + home brick
  int* bin_count;     // [nbricks]
  int* bin_start;     // [nbricks + 1]
  int* bin_list;      // [n] atoms grouped by home brick, sorted
  void* grid;         // complex interleaved, x fastest, real type per prec
  void* tiles;        // [nbricks][12^3][2] spreading tiles (real type), scratch
  const void* qfac;   // K1*K2*K3 real
  const void* twid;   // [K1 + K2 + K3] complex exp(-2 pi i m / K) per axis (real type)
  // solver vectors [6][npad] FP64 (component-major)
  double* ef;
  double* mu;
  double* res;
  double* h;
  double* hreal;
  double* pp;
  double* zr;
  double* partial;  // partial sums, see part_* in the kernels
  SolveState* st;
  int* err;         // device error flag (list overflow ...)
};

struct Work {  // one unit of a launch: a block (brick, chunk ...) of one system
  int sys;
  int blk;
};

constexpr int kConvChunk = 512;  // grid points per workgroup of the spreading merge
// grid FFT: complex values per LDS buffer of a workgroup, longest axis
constexpr int kFftBuf = 2048;
constexpr int kFftMaxLen = 512;

// lines per FFT workgroup along an axis of length K: the largest power of
// two up to `cap` (a power of two, at most 32; the grid autotuner picks it
// per grid) whose padded lines (K + 1 values each) fit one LDS buffer.  The
// grouping does not change any line's arithmetic.
__host__ __device__ inline int fft_group_lines(int K, int cap = 32) {
  int fl = cap;
  while (fl > 1 && fl * (K + 1) > kFftBuf) fl >>= 1;
  return fl;
}

// workgroups of one FFT pass along `axis` of a K0 x K1 x K2 grid
__host__ __device__ inline int fft_groups(const int* nfft, int axis, int cap = 32) {
  const int fl = fft_group_lines(nfft[axis], cap);
  const int inner = axis == 0 ? nfft[1] : nfft[0];
  const int outer = axis == 2 ? nfft[1] : nfft[2];
  return outer * ((inner + fl - 1) / fl);
}
// work entries one system takes in the FFT launches: the line groups of
// the three axes, one entry per z plane, and for 64^3 grids the register
// z pass's waves (runs of 16 x per grid row)
inline int fft_work_count(const int* nfft, int cap) {
  int ng = std::max(std::max(fft_groups(nfft, 0, cap), nfft[2]), std::max(fft_groups(nfft, 1, cap), fft_groups(nfft, 2, cap)));
  if (nfft[0] == 64 && nfft[1] == 64 && nfft[2] == 64) ng = std::max(ng, nfft[1] * ((nfft[0] + 15) / 16));
  return ng;
}
constexpr int kTile = 12;                      // spreading tile edge: brick (<= 8) + order - 1
constexpr int kTile3 = kTile * kTile * kTile;  // points per tile
constexpr int kDefaultSegs = 4;   // matvec wavefronts per block unless tuned
constexpr int kEfld0Segs = 8;     // permanent-field pass wavefronts per block

// A set of systems launched together.  Every kernel reaches its systems
// through d_views[w.sys] and walks one of the work tables; a single system
// is a batch of one.  All systems of a batch share one precision.
struct Batch {
  int prec = THP_PREC_FP64;
  int nsys = 0;
  int segs = 4;                 // wavefronts per 64-atom block in the pair kernels
  int nb = 8;                   // list entries per prefetched batch in the matvec (8 or 16)
  int rc = 0;                   // matvec kinds held: bit 0 recomputed coefficients, bit 1 stored (0: unknown, both)
  SysView* d_views = nullptr;   // [nsys]
  Work* d_blk = nullptr;        // 64-atom blocks
  int n_blk = 0;
  Work* d_red = nullptr;        // PCG vector-kernel blocks
  int n_red = 0;
  Work* d_gat = nullptr;        // PME gather blocks
  int n_gat = 0;
  Work* d_brick = nullptr;      // spreading bricks
  int n_brick = 0;
  Work* d_conv = nullptr;       // grid chunks of kConvChunk points
  int n_conv = 0;
  Work* d_fft = nullptr;        // FFT line groups (the most of the three axes)
  int n_fft = 0;
  int* d_ids = nullptr;         // [nsys] = 0..nsys-1
  int real_only = 0;            // short-range solves: no reciprocal terms
  int fft_plane = 0;            // x and y FFT passes fused per z plane (plane_fft_setup)
  int fft_plane_lds = 0;        // its dynamic LDS bytes
  int fft_r64 = 0;              // every grid 64^3 FP32: the register-FFT kernels
  // the real-space matvec may run on the context's side stream beside the
  // reciprocal chain (one-system launch sequences only, never captured)
  int fork_ok = 0;
  // fork and join events of the graph being captured for this batch (owned
  // by the serving loop's batch buffers, used in that capture only)
  hipEvent_t cap_fork = nullptr, cap_join = nullptr;
};

// Whether a batch of these grids takes the fused x-y plane FFT kernel: FP32
// grids whose two padded planes and twiddles fit kPlaneLds bytes of LDS.
constexpr int kPlaneLds = 80 * 1024;
constexpr int kPlaneThreads = 256;  // threads per plane workgroup (at least as many as the lines)
inline int plane_fft_bytes(const int* nfft, int prec) {
  return prec == THP_PREC_FP64 ? 0 : (2 * nfft[1] * (nfft[0] + 1) + nfft[0] + nfft[1]) * 2 * (int)sizeof(float);
}
inline void plane_fft_setup(Batch& b, const int* const* nffts, int m, int prec) {
  int need = 0;
  bool ok = prec != THP_PREC_FP64 && m > 0;
  for (int q = 0; ok && q < m; ++q) {
    const int bytes = plane_fft_bytes(nffts[q], prec);
    // one thread or more per line of either pass
    ok = bytes > 0 && bytes <= kPlaneLds && nffts[q][0] <= kPlaneThreads && nffts[q][1] <= kPlaneThreads;
    need = bytes > need ? bytes : need;
  }
  b.fft_plane = ok ? 1 : 0;
  b.fft_plane_lds = ok ? need : 0;
  bool r64 = prec != THP_PREC_FP64 && m > 0;
  for (int q = 0; r64 && q < m; ++q) r64 = nffts[q][0] == 64 && nffts[q][1] == 64 && nffts[q][2] == 64;
  b.fft_r64 = r64 ? 1 : 0;
}

}  // namespace thp

struct thp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void* serve = nullptr;
  int n_cu = 0;
  int pair_segs = 0;           // 0: default; else forced waves per block (1..16)
  int pair_nb = 0;             // 0: default; else forced matvec entries per batch (8, 16)
  bool plane_fft = false;       // the fused x-y plane FFT kernel may take its LDS
  bool reg_fft = true;          // 64^3 FP32 grids take the register-FFT kernels (THP_FFT_REG=0: not)
  int fft_pp = 1, fft_zp = 2;   // lanes per 64-point line in its plane and z passes
  hipStream_t side = nullptr;   // the real-space matvec beside the reciprocal chain
  bool fork_batch = true;       // batches of several systems fork too (THP_FORK_BATCH=0: not)
  bool rc_matvec = true;        // FP32 systems with one Thole value recompute matvec coefficients (THP_MATVEC_RC=0: not)
  int solve_chunk[2] = {1, 4};  // one-system PCG iterations per state read after the first: FP64, mixed (THP_SOLVE_CHUNK)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // ordering of the stream-ordered entry points against the caller's stream
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // grid shapes picked by the serving loop's autotuner per grid and
  // precision: (K0, K1, K2, prec) -> (largest brick edge, FFT lines per
  // workgroup at most); systems created on this context take them
  std::map<std::array<int, 4>, std::pair<int, int>> grid_shape;
  int force_brick = 0, force_fft = 0;  // set while the autotuner measures
  // this device's address of the file-scope scratch array of the entry point
  // that allocates nothing (thp_epolar1_noalloc), resolved by thp_init
  double* scratch = nullptr;
  std::set<thp_sys*> systems;  // live systems; thp_finalize releases what is left
};

struct thp_sys {
  thp_ctx* ctx = nullptr;
  int n = 0, npad = 0, nblk = 0;
  int prec = THP_PREC_FP64;
  double box[3] = {0, 0, 0};
  double aewald = 0, cut = 0, buffer = 0;
  int nfft[3] = {0, 0, 0};
  int order = 5;
  bool own_grid = true;  // false: the grid and FFT plans come from a batch (serving)
  std::vector<int> perm;   // sorted slot -> original index
  std::vector<int> iperm;  // original index -> sorted slot
  int* iperm_dev = nullptr;
  double* cphi = nullptr;  // [10][npad] scratch for efld0_recip
  // dipole predictor (thp_set_predictor)
  int pred = 0;                  // THP_PRED_*
  int pred_n = 0;                // stored solves, at most kMaxUalt (nualt)
  int pred_newest = 0;           // ring slot of the newest stored solve
  double* pred_hist = nullptr;   // [kMaxUalt][6][npad]
  double* pred_coef = nullptr;   // [kMaxUalt], newest entry first
  double* pred_part = nullptr;   // [pred_blocks][kPredPairs]
  int pred_blocks = 0;
  // solver (thp_set_solver): 1 PCG, 2 Jacobi/DIIS, 3 truncated CG
  int polalg = 1;
  double* diis_x = nullptr;     // [kDiisMax][6][npad] Jacobi iterates
  double* diis_e = nullptr;     // [kDiisMax][6][npad] their increments
  double* diis_part = nullptr;  // [diis_blocks][kDiisParts]
  thp::DiisState* diis = nullptr;
  int diis_blocks = 0;
  // short-range real-space solves (thp_set_shortreal): the cutoff, 0 = off
  double short_cut = 0.0;
  // local-frame multipoles (thp_system_set_frames), rotated on the device
  int* d_frames = nullptr;         // [npad] x (axis type, z, x, y), sorted slots
  double* d_pole_local = nullptr;  // [10][npad], the rpole layout
  int* d_tinv_off = nullptr;       // [npad + 1] CSR offsets of d_tinv
  int* d_tinv = nullptr;           // per atom, the sites whose frame forces reach it
  bool has_frames = false;
  // truncated conjugate gradient (thp_set_tcg; initprm.f:447-455 defaults)
  int tcg_order = 2, tcg_prec = 1, tcg_guess = 0, tcg_peek = 1;
  double tcg_omega = 1.0;
  double* tcg_vec = nullptr;   // [2][6][npad]: A r0, then A (T r0) preconditioned
  double* tcg_part = nullptr;  // [4][tcg_blocks][2]: n0, t1, np1, A r0 . P2
  double* tcg_sc = nullptr;    // scalars (kTcg* slots)
  int tcg_blocks = 0;
  std::vector<double> hx, hy, hz;       // sorted-order coordinates (host copy)
  std::vector<double> bx0, by0, bz0;    // coordinates at the last block-list build
  std::vector<double> ux0, uy0, uz0;    // coordinates at the last block-union build
  double skin = 0.5;                    // skin of the block-union lists
  bool list_valid = false;
  bool tight_valid = false;   // block-union lists match the coordinates
  bool verlet_valid = false;  // per-atom Verlet list matches the coordinates
  bool pme_valid = false;   // B-spline values and brick bins match the coordinates
  bool coef_valid = false;  // pair coefficients match the coordinates and lists
  int nbricks = 0;
  thp::Work* d_brickwork = nullptr;  // [nbricks]
  int maxnb = 0;
  // device allocations and their sizes; bytes is their sum
  std::unordered_map<void*, size_t> allocs;
  int64_t bytes = 0;
  // staging for host arrays copied in and out in caller order ([n][10]
  // doubles: the widest array of the entry points)
  double* stage = nullptr;
  // pinned host mirror of stage: results come back through it in one copy
  // at full rate (grown on first need, freed with the system)
  double* hstage = nullptr;
  size_t hstage_len = 0;  // doubles
  // iterations of the last PCG solve; the next mixed solve queues as many
  // (at most kFirstChunkMax) before its first read of the solve state
  int last_niter = 0;
  // dipoles in mu are the result of a solve of the current coordinates and
  // multipoles (the gradient needs them), and the solver that made them
  bool mu_valid = false;
  int mu_polalg = 1;
  bool mu_short = false;
  // the truncated-CG options of the last solve (order, prec, guess, peek)
  // and its omega: the gradient differentiates that solve
  int mu_tcg[4] = {2, 1, 0, 1};
  double mu_tcg_omega = 1.0;
  // forces and torques of the gradient ([3][npad] each), made on first use
  double* grad_f = nullptr;
  double* grad_t = nullptr;
  // virial shares of the gradient and the two grid spectra they need,
  // made on the first call that asks for the virial
  double* grad_vir = nullptr;
  char* vir_spec = nullptr;
  thp::SysView view;           // host copy of the view
  thp::SysView* d_view = nullptr;
  int view_version = 0;        // bumped whenever the device view is rewritten
  thp::Work* d_work = nullptr;  // single-system work table [nblk]
  thp::Work* d_redwork = nullptr;  // vector-kernel work table [red_blocks]
  thp::Work* d_gatwork = nullptr;  // gather work table [gat_blocks]
  int* d_sysid = nullptr;          // {0, 1}
  int* d_one = nullptr;            // d_sysid + 1
  thp::Work* d_convwork = nullptr; // grid-chunk work table
  thp::Work* d_fftwork = nullptr;  // FFT line-group work table
  int nfftw = 0;
  thp::Work* d_workx = nullptr;    // block table in XCD order (path kernels)
  thp::Work* d_brickx = nullptr;   // brick table in XCD order
  int nworkx = 0, nbrickx = 0;
  int nconv = 0;
  thp::Batch one;                  // this system as a batch of one
  // block neighbor lists
  double* bbox = nullptr;  // [nblk][6] centre + half-extent
  int* bcount = nullptr;   // [nblk]
  int* blist = nullptr;    // [nblk][maxbl]
  int maxbl = 0;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

// ------------------------------------------------ host helpers (thp_api.hip)
namespace thp {
thp_status fail(thp_ctx* ctx, thp_status code, const char* fmt, ...);
thp_status create_system(thp_ctx* ctx, const thp_system_desc* d, thp_sys** out, bool own_grid);
// the serving loop's grid-shape autotuner (brick edge, FFT lines per
// workgroup) for the grid of `d`, once per grid and precision on ctx
thp_status tune_grid_shape(thp_ctx* ctx, const thp_system_desc* d);
thp_status ensure_lists(thp_sys* s);
thp_status update_coords(thp_sys* s, const double* x, const double* y, const double* z, double* pinned);
size_t grid_points(const thp_sys* s);
// work table reordered so that workgroup k's entry runs on XCD k % 8 next to
// its spatial neighbours (padding entries have sys = -1)
std::vector<Work> xcd_order(const std::vector<Work>& list);
// per-system work tables with sys = `index` (appended)
void append_work(const thp_sys* s, int index, std::vector<Work>& blk, std::vector<Work>& red,
                 std::vector<Work>& gat, std::vector<Work>& brick, std::vector<Work>& conv,
                 std::vector<Work>& fft);
// batched launch sequences (enqueued on ctx->stream)
// d_req (optional): per-system request flags; requested systems start a
// solve in set-up state (kGateSetup kernels work on them until b_end_setup)
thp_status b_init_state(thp_ctx* ctx, const Batch& b, double poleps, int politer, const int* d_warm,
                        const int* d_req);
thp_status b_load_coords(thp_ctx* ctx, const Batch& b, int gate);
thp_status b_pme_prep(thp_ctx* ctx, const Batch& b, int gate);
thp_status b_setup(thp_ctx* ctx, const Batch& b, int gate);
thp_status b_end_setup(thp_ctx* ctx, const Batch& b);
thp_status b_iteration(thp_ctx* ctx, const Batch& b);
thp_status b_finish(thp_ctx* ctx, const Batch& b, int only_done);
thp_status b_pair_tmatxb(thp_ctx* ctx, const Batch& b, int vin, int vout, int dodiag, int check_conv);
// copy each system's SolveState and error flag into out[nsys] (device)
struct StateOut {
  SolveState st;
  int err;
  int pad;
};
thp_status b_collect(thp_ctx* ctx, const Batch& b, StateOut* d_out);
// the system's dipoles [n][2][3] in caller order into host memory (synchronous)
thp_status download_mu(thp_sys* s, double* host);
// stream-ordered calls: the library's stream waits for the caller's work
// (order_in), the caller's stream for the library's (order_out); a null
// caller makes the call synchronous
thp_status order_in(thp_ctx* ctx, hipStream_t caller);
thp_status order_out(thp_ctx* ctx, hipStream_t caller);
// host copy of the device's sorted coordinates and list validity (synchronous)
thp_status sync_host_coords(thp_sys* s);
}  // namespace thp
