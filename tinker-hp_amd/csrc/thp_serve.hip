// Serving loop: continuous batching of independent polarization solves.
//
// One worker thread owns the device work of a context while the loop runs.
// Per pass it (1) admits queued requests into free slots, creating the
// device data of systems that are not resident (evicting idle ones when the
// HBM budget is short) and uploading the requests' coordinates as given,
// (2) runs one launch sequence over the batch of member systems: the
// set-up of the admitted solves (coordinates into sorted order, B-spline
// tables, permanent field and pair coefficients, guess or warm start, first
// residual; kernels gated to the systems in set-up), then
// `iters_per_graph` PCG iterations of every solve in flight, then the
// energies of the solves that finished, and (3) reads every solve state
// back in one copy and retires the finished solves.  The sequence is
// captured once per batch layout as a HIP graph and replayed, so a pass
// costs the host a few copies and one launch.  Kernels walk work tables of
// (system, unit) entries, so solves at different stages share each launch;
// reciprocal grids of equal size sit side by side in a pool, and one set of
// FFT launches transforms the grids of every size in the batch.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "thp_internal.h"

using namespace thp;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// host copy of a registered system, enough to (re)create its device data
struct ServeSystem {
  int id = 0;
  thp_system_desc desc{};
  std::vector<double> x, y, z, rpole, pol, thole, dpu;
  std::vector<int> ik;
  // local frames (thp_serve_set_frames): axis types, frame atoms and local
  // multipoles [n][13], caller order; empty for global multipoles
  std::vector<int> fr_axt, fr_z, fr_x, fr_y;
  std::vector<double> pole_local;
  bool has_frames = false;
  thp_sys* s = nullptr;      // device data when resident
  int64_t bytes = 0;         // device footprint once created (estimate before)
  bool busy = false;         // a solve of it is in flight
  bool has_mu = false;       // device mu holds its last solution
  uint64_t last_use = 0;
  int member = -1;           // position in the iteration batch, -1 if not a member
  int idle_passes = 0;       // passes without a solve in flight while a member
  uint64_t gen = 0;          // counts the creations of its device data
};

struct ServeSystem;

struct Request {
  uint64_t ticket = 0;
  int sid = 0;
  ServeSystem* sys = nullptr;  // set on submission (systems never move: unique_ptr)
  double* xyz = nullptr;  // pinned [3][n] caller-order coordinates, or null (unchanged)
  size_t xyz_cap = 0;     // doubles held by xyz (a pool buffer)
  double* rp = nullptr;   // pinned [n][13] caller-order multipoles, or null (unchanged)
  size_t rp_cap = 0;
  Clock::time_point t_submit;
  // its coordinates (and multipoles) already copied into the system's device
  // buffers on the upload stream, for device data of generation pre_gen
  bool pre = false;
  uint64_t pre_gen = 0;
};

struct InFlight {
  Request req;
  ServeSystem* sys = nullptr;
  Clock::time_point t_start;
  int warm = 0;
  bool restart = false;  // set-up repeated after a list rebuild (stale coordinates)
  bool setup_seen = false;  // its set-up's pass has been read back, not stale
};

struct DipoleCopy {  // thp_serve_get_dipoles handed to the worker
  int sid = 0;
  double* out = nullptr;
  bool done = false;
  thp_status rc = THP_OK;
};

// reciprocal grids of one size, side by side, with their spreading tiles
struct GridPool {
  int K[3] = {0, 0, 0};
  int cap = 0;
  void* base = nullptr;
  size_t bytes_per_grid = 0;
  void* tiles = nullptr;       // spreading tiles of each slot, after the grids
  size_t bytes_per_tiles = 0;
};

// device and pinned host buffers of one batch layout
struct BatchBuf {
  Batch b;
  SysView* views = nullptr;
  Work* tables = nullptr;
  int* ids = nullptr;
  int* warm = nullptr;
  size_t cap_views = 0, cap_tables = 0;
  // pinned staging
  SysView* h_views = nullptr;
  Work* h_tables = nullptr;
  int* h_ids = nullptr;
  int* h_warm = nullptr;
  size_t hcap_views = 0, hcap_tables = 0;
  int* req = nullptr;    // per-system set-up request flags (device)
  int* h_req = nullptr;  // pinned staging of req
  // graphs of the current layout (set-up + iterations, iterations only),
  // each as a run of segment graphs launched back to back
  std::vector<hipGraphExec_t> exec[2];
  std::vector<hipGraph_t> graph[2];
  // fork and join events of each captured segment (the matvec on the side
  // stream): used only inside that capture, destroyed after its graphs
  std::vector<hipEvent_t> events;
  int graph_iters = 0;
};

struct Serve {
  thp_ctx* ctx = nullptr;
  thp_serve_opts o{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> queue;
  std::unordered_map<uint64_t, thp_result> done;
  std::vector<std::unique_ptr<ServeSystem>> systems;
  std::deque<DipoleCopy*> copies;
  uint64_t next_ticket = 1;
  bool stop = false;
  thp_status fatal = THP_OK;
  std::string fatal_msg;
  thp_serve_stats stats{};
  double active_sum = 0.0;
  std::thread worker;
  // worker-owned
  std::vector<InFlight> active;
  std::atomic<int64_t> resident{0};  // device bytes of resident systems (written by the worker)
  int64_t budget = 0;
  uint64_t use_clock = 0;
  std::map<std::tuple<int, int, int>, GridPool> pools;
  int pool_generation = 0;
  BatchBuf setup, iter;
  // The iteration batch: systems with a solve in flight plus recently idle
  // ones (they leave after kIdlePasses passes without work), so that the
  // layout, its tables and its captured graph survive the usual
  // finish-and-resubmit turnover.  Idle members cost only early exits.
  std::vector<ServeSystem*> members;
  bool iter_dirty = true;
  int layout_age = 0;  // passes run with the current iteration-batch layout
  int iter_generation = -1;
  std::vector<int> iter_versions;  // view versions the iteration batch was built from
  StateOut* d_out = nullptr;
  StateOut* h_out = nullptr;
  int out_cap = 0;
  bool graphs_ok = true;
  std::map<std::pair<int, int>, int> tuned;  // (log2 atoms, systems) -> pair waves per block
  std::multimap<size_t, double*> pinned_pool;  // free request coordinate buffers, by capacity (under mu)
  std::vector<InFlight> restarts;              // set-ups to repeat after list rebuilds (worker)
  // systems with device data, in admission order (worker); the worker never
  // walks `systems`, which other threads extend under `mu`
  std::vector<ServeSystem*> residents;
  // Early uploads: while a pass runs, the coordinates of the next queued
  // request of a system are copied into its device buffers on a second
  // stream (default priority) when nothing queued can read them before
  // that request's set-up: the system is idle, or its running solve's set-up
  // has been read back and was not stale (a stale set-up is repeated from
  // the same buffers).  The next pass waits on ev_up before its launches.
  hipStream_t up = nullptr;
  hipEvent_t ev_up = nullptr;
  bool up_pending = false;
};

// request coordinate buffers: pinned, recycled through the pool
double* take_pinned(Serve* sv, size_t doubles, size_t* cap) {
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->pinned_pool.lower_bound(doubles);
    if (it != sv->pinned_pool.end() && it->first <= 2 * doubles) {
      double* p = it->second;
      *cap = it->first;
      sv->pinned_pool.erase(it);
      return p;
    }
  }
  double* p = nullptr;
  if (hipHostMalloc((void**)&p, sizeof(double) * doubles, hipHostMallocDefault) != hipSuccess) return nullptr;
  *cap = doubles;
  return p;
}

void give_pinned(Serve* sv, double* p, size_t cap) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->pinned_pool.emplace(cap, p);
}

#define SRET(call)                   \
  do {                               \
    thp_status r_ = (call);          \
    if (r_ != THP_OK) return r_;     \
  } while (0)

#define SHIP(ctx, call)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return fail((ctx), THP_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// device footprint of a system before it exists: per-atom arrays and
// B-spline tables (~1.8 kB per atom), block-union and compacted lists at
// liquid density (~3 kB per atom with FP32 coefficients, ~4.5 kB with FP64),
// the scalar-sum table; replaced by the real figure once the system and its
// lists exist
int64_t estimate_bytes(const thp_system_desc& d) {
  const int64_t np = (d.n + kBlock - 1) / kBlock * kBlock;
  const int64_t pts = (int64_t)d.nfft[0] * d.nfft[1] * d.nfft[2];
  const int64_t real = d.precision == THP_PREC_FP64 ? 8 : 4;
  const int64_t lists = d.precision == THP_PREC_FP64 ? 4608 : 3072;
  return np * (1792 + lists) + pts * real + (256 << 10);
}

// the resident figure follows the system's device memory (lists grow)
void refresh_bytes(Serve* sv, ServeSystem* ss) {
  if (!ss->s) return;
  sv->resident += ss->s->bytes - ss->bytes;
  ss->bytes = ss->s->bytes;
}

thp_status grow_device(thp_ctx* ctx, void** p, size_t* cap, size_t need, size_t elem) {
  if (need <= *cap) return THP_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  const size_t c = std::max(need, *cap * 2);
  SHIP(ctx, hipMalloc(p, c * elem));
  *cap = c;
  return THP_OK;
}

thp_status grow_pinned(thp_ctx* ctx, void** p, size_t* cap, size_t need, size_t elem) {
  if (need <= *cap) return THP_OK;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  const size_t c = std::max(need, *cap * 2);
  SHIP(ctx, hipHostMalloc(p, c * elem, hipHostMallocDefault));
  *cap = c;
  return THP_OK;
}

void drop_graph(BatchBuf& bb) {
  for (int v = 0; v < 2; ++v) {
    for (hipGraphExec_t x : bb.exec[v]) (void)hipGraphExecDestroy(x);
    for (hipGraph_t g : bb.graph[v]) (void)hipGraphDestroy(g);
    bb.exec[v].clear();
    bb.graph[v].clear();
  }
  for (hipEvent_t e : bb.events) (void)hipEventDestroy(e);
  bb.events.clear();
  bb.graph_iters = 0;
}

void free_batchbuf(BatchBuf& bb) {
  drop_graph(bb);
  for (void* p : {(void*)bb.views, (void*)bb.tables, (void*)bb.ids, (void*)bb.warm, (void*)bb.req})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)bb.h_views, (void*)bb.h_tables, (void*)bb.h_ids, (void*)bb.h_warm, (void*)bb.h_req})
    if (p) (void)hipHostFree(p);
  bb = BatchBuf();
}

void free_pool(GridPool& gp) {
  if (gp.base) (void)hipFree(gp.base);
  gp = GridPool();
}

// Build a batch over `members` (views, work tables, grid slots)
// into bb.  Grids are scratch between kernels of one pass, so each batch
// takes pool slots 0..m-1 of each grid size afresh.
thp_status build_batch(Serve* sv, BatchBuf& bb, const std::vector<thp_sys*>& members, const std::vector<int>& warm) {
  thp_ctx* ctx = sv->ctx;
  const int m = (int)members.size();
  std::vector<Work> blk, red, gat, brick, conv, fft;
  for (int q = 0; q < m; ++q) append_work(members[q], q, blk, red, gat, brick, conv, fft);
  blk = xcd_order(blk);
  brick = xcd_order(brick);
  gat = xcd_order(gat);
  // grid slots per size
  std::map<std::tuple<int, int, int>, std::vector<int>> groups;
  for (int q = 0; q < m; ++q) groups[{members[q]->nfft[0], members[q]->nfft[1], members[q]->nfft[2]}].push_back(q);
  const int prec = members[0]->prec;
  const size_t real = prec == THP_PREC_FP64 ? 8 : 4;
  // the layout changes: its graphs go first, before any buffer they
  // reference is freed or rewritten
  drop_graph(bb);
  for (auto& g : groups) {
    GridPool& gp = sv->pools[g.first];
    const int need = (int)g.second.size();
    if (gp.cap < need) {
      // grow: nothing in flight reads the old grids (the stream is idle
      // between passes), and every graph that captured them is dropped
      SHIP(ctx, hipStreamSynchronize(ctx->stream));
      free_pool(gp);
      gp.K[0] = std::get<0>(g.first);
      gp.K[1] = std::get<1>(g.first);
      gp.K[2] = std::get<2>(g.first);
      gp.cap = std::max(need, 2);
      while (gp.cap < need) gp.cap *= 2;
      gp.bytes_per_grid = 2 * real * (size_t)gp.K[0] * gp.K[1] * gp.K[2];
      const thp_sys* s0 = members[g.second[0]];
      gp.bytes_per_tiles = ((2 * real * (size_t)s0->nbricks * kTile3 + 255) / 256) * 256;
      const size_t grids = ((gp.bytes_per_grid * gp.cap + 255) / 256) * 256;
      SHIP(ctx, hipMalloc(&gp.base, grids + gp.bytes_per_tiles * gp.cap));
      gp.tiles = static_cast<char*>(gp.base) + grids;
      ++sv->pool_generation;
      drop_graph(sv->setup);
      drop_graph(sv->iter);
      sv->iter_dirty = true;
    }
  }
  // views
  SRET(grow_pinned(ctx, (void**)&bb.h_views, &bb.hcap_views, m, sizeof(SysView)));
  const size_t nt = blk.size() + red.size() + gat.size() + brick.size() + conv.size() + fft.size();
  SRET(grow_pinned(ctx, (void**)&bb.h_tables, &bb.hcap_tables, nt, sizeof(Work)));
  size_t cap_ids = bb.cap_views;
  SRET(grow_device(ctx, (void**)&bb.views, &bb.cap_views, m, sizeof(SysView)));
  SRET(grow_device(ctx, (void**)&bb.tables, &bb.cap_tables, nt, sizeof(Work)));
  if (!bb.ids || cap_ids < bb.cap_views) {
    for (void* p : {(void*)bb.ids, (void*)bb.warm, (void*)bb.req})
      if (p) (void)hipFree(p);
    for (void* p : {(void*)bb.h_ids, (void*)bb.h_warm, (void*)bb.h_req})
      if (p) (void)hipHostFree(p);
    SHIP(ctx, hipMalloc((void**)&bb.ids, bb.cap_views * sizeof(int)));
    SHIP(ctx, hipMalloc((void**)&bb.warm, bb.cap_views * sizeof(int)));
    SHIP(ctx, hipMalloc((void**)&bb.req, bb.cap_views * sizeof(int)));
    SHIP(ctx, hipHostMalloc((void**)&bb.h_ids, bb.cap_views * sizeof(int), hipHostMallocDefault));
    SHIP(ctx, hipHostMalloc((void**)&bb.h_warm, bb.cap_views * sizeof(int), hipHostMallocDefault));
    SHIP(ctx, hipHostMalloc((void**)&bb.h_req, bb.cap_views * sizeof(int), hipHostMallocDefault));
  }
  Batch& b = bb.b;
  b = Batch();
  b.prec = prec;
  b.nsys = m;
  for (int q = 0; q < m; ++q) b.rc |= members[q]->view.rc_sqa > 0.0 ? 1 : 2;
  for (auto& g : groups) {
    GridPool& gp = sv->pools[g.first];
    for (size_t slot = 0; slot < g.second.size(); ++slot) {
      const int q = g.second[slot];
      bb.h_views[q] = members[q]->view;
      bb.h_views[q].grid = static_cast<char*>(gp.base) + slot * gp.bytes_per_grid;
      bb.h_views[q].tiles = static_cast<char*>(gp.tiles) + slot * gp.bytes_per_tiles;
    }
  }
  for (int q = 0; q < m; ++q) {
    bb.h_ids[q] = q;
    bb.h_warm[q] = q < (int)warm.size() ? warm[q] : 0;
  }
  size_t off = 0;
  auto put = [&](const std::vector<Work>& v, Work** dptr, int* cnt) {
    std::copy(v.begin(), v.end(), bb.h_tables + off);
    *dptr = bb.tables + off;
    *cnt = (int)v.size();
    off += v.size();
  };
  put(blk, &b.d_blk, &b.n_blk);
  put(red, &b.d_red, &b.n_red);
  put(gat, &b.d_gat, &b.n_gat);
  put(brick, &b.d_brick, &b.n_brick);
  put(conv, &b.d_conv, &b.n_conv);
  put(fft, &b.d_fft, &b.n_fft);
  b.d_views = bb.views;
  b.d_ids = bb.ids;
  std::vector<const int*> nfs(m);
  for (int q = 0; q < m; ++q) nfs[q] = members[q]->nfft;
  plane_fft_setup(b, nfs.data(), m, prec);
  hipStream_t st = ctx->stream;
  SHIP(ctx, hipMemcpyAsync(bb.views, bb.h_views, sizeof(SysView) * m, hipMemcpyHostToDevice, st));
  SHIP(ctx, hipMemcpyAsync(bb.tables, bb.h_tables, sizeof(Work) * nt, hipMemcpyHostToDevice, st));
  SHIP(ctx, hipMemcpyAsync(bb.ids, bb.h_ids, sizeof(int) * m, hipMemcpyHostToDevice, st));
  SHIP(ctx, hipMemcpyAsync(bb.warm, bb.h_warm, sizeof(int) * m, hipMemcpyHostToDevice, st));
  return THP_OK;
}

void drop_member(Serve* sv, ServeSystem* ss) {
  if (ss->member < 0) return;
  sv->members.erase(sv->members.begin() + ss->member);
  ss->member = -1;
  for (size_t q = 0; q < sv->members.size(); ++q) sv->members[q]->member = (int)q;
  sv->iter_dirty = true;
}

thp_status release_system(Serve* sv, ServeSystem* ss) {
  if (!ss->s) return THP_OK;
  // an early upload may still be writing the system's buffers
  SHIP(sv->ctx, hipStreamSynchronize(sv->up));
  drop_member(sv, ss);
  sv->resident -= ss->bytes;
  thp_system_destroy(ss->s);
  ss->s = nullptr;
  sv->residents.erase(std::find(sv->residents.begin(), sv->residents.end(), ss));
  ss->bytes = estimate_bytes(ss->desc);
  ss->has_mu = false;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stats.evictions += 1;
    sv->stats.resident_systems = (int)sv->residents.size();
  }
  return THP_OK;
}

// make a system resident within the budget; false in *ok when it cannot
// fit now (idle residents are released, least recently used first)
thp_status admit(Serve* sv, ServeSystem* ss, const std::vector<ServeSystem*>& pinned, bool* ok) {
  *ok = true;
  if (ss->s) return THP_OK;
  const int64_t need = ss->bytes;
  if (need > sv->budget) return fail(sv->ctx, THP_ERR_NOMEM, "system of %d atoms needs %lld bytes, over the budget %lld",
                                     ss->desc.n, (long long)need, (long long)sv->budget);
  while (sv->resident + need > sv->budget) {
    ServeSystem* victim = nullptr;
    for (ServeSystem* c : sv->residents) {
      if (!c->s || c->busy || c == ss) continue;
      if (std::find(pinned.begin(), pinned.end(), c) != pinned.end()) continue;
      if (!victim || c->last_use < victim->last_use) victim = c;
    }
    if (!victim) {
      *ok = false;
      return THP_OK;
    }
    SRET(release_system(sv, victim));
  }
  if (sv->o.autotune) SRET(tune_grid_shape(sv->ctx, &ss->desc));
  thp_sys* s = nullptr;
  SRET(create_system(sv->ctx, &ss->desc, &s, false));
  if (ss->has_frames) {
    // the frames ride with the device data: the set-up rotates the local
    // multipoles at every request's coordinates (k_rotpole_batch)
    const thp_status rf = thp_system_set_frames(s, ss->fr_axt.data(), ss->fr_z.data(), ss->fr_x.data(),
                                                ss->fr_y.data(), ss->pole_local.data());
    if (rf != THP_OK) {
      thp_system_destroy(s);
      return rf;
    }
  }
  ss->s = s;
  ++ss->gen;
  ss->bytes = s->bytes;
  sv->resident += ss->bytes;
  ss->has_mu = false;
  sv->residents.push_back(ss);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stats.resident_systems = (int)sv->residents.size();
  }
  return THP_OK;
}

int bucket_log2(int64_t v) {
  int b = 0;
  while ((int64_t(1) << (b + 1)) <= v) ++b;
  return b;
}

// pair kernel shape for the iteration batch: measured once per (size,
// batch) class over the candidates and remembered in the context
thp_status tune_pair(Serve* sv, Batch& b, const std::vector<thp_sys*>& members) {
  thp_ctx* ctx = sv->ctx;
  int64_t atoms = 0;
  for (thp_sys* s : members) atoms += s->n;
  const auto key = std::make_pair(bucket_log2(atoms), bucket_log2(b.nsys));
  auto it = sv->tuned.find(key);
  if (it != sv->tuned.end()) {
    b.segs = it->second / 100;
    b.nb = it->second % 100;
    return THP_OK;
  }
  if ((ctx->pair_segs > 0 && ctx->pair_nb > 0) || !sv->o.autotune) {
    b.segs = ctx->pair_segs > 0 ? ctx->pair_segs : kDefaultSegs;
    b.nb = ctx->pair_nb > 0 ? ctx->pair_nb : 8;
    sv->tuned[key] = b.segs * 100 + b.nb;
    return THP_OK;
  }
  // every candidate rewrites hreal from pp, which the next iteration does
  // again before anything reads hreal
  static const int cand_s[] = {2, 4, 8, 16};
  static const int cand_nb[] = {8, 16};
  hipEvent_t e0, e1;
  SHIP(ctx, hipEventCreate(&e0));
  SHIP(ctx, hipEventCreate(&e1));
  float best = 1e30f;
  int best_s = kDefaultSegs, best_nb = 8;
  for (int cs : cand_s) {
    if (ctx->pair_segs > 0 && cs != ctx->pair_segs) continue;
    for (int cn : cand_nb) {
      if (ctx->pair_nb > 0 && cn != ctx->pair_nb) continue;
      b.segs = cs;
      b.nb = cn;
      SRET(b_pair_tmatxb(ctx, b, 1, 3, 1, kGateIter));  // warm
      SHIP(ctx, hipEventRecord(e0, ctx->stream));
      for (int r = 0; r < 3; ++r) SRET(b_pair_tmatxb(ctx, b, 1, 3, 1, kGateIter));
      SHIP(ctx, hipEventRecord(e1, ctx->stream));
      SHIP(ctx, hipEventSynchronize(e1));
      float t = 0;
      SHIP(ctx, hipEventElapsedTime(&t, e0, e1));
      if (t < best) {
        best = t;
        best_s = cs;
        best_nb = cn;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b.segs = best_s;
  b.nb = best_nb;
  sv->tuned[key] = best_s * 100 + best_nb;
  return THP_OK;
}

// The launch sequence of one pass over the member batch: with set-up
// (variant 1) the requested systems go through their set-up first; then
// `iters` PCG iterations of everything in flight, the energies of the
// solves that finished, and the state collection.  The sequence is cut
// into segments of at most about thirty launches (the set-up; two
// iterations at a time, the last with the energies and the collection),
// each captured as a graph of its own: replaying one graph of ~80 kernels
// left the device idle for 0.5-1 ms after its 32nd kernel, while graphs
// this short replay back to back without a gap.
constexpr int kItersPerSegment = 2;

int pass_segments(int variant, int iters) { return (variant ? 1 : 0) + (iters + kItersPerSegment - 1) / kItersPerSegment; }

thp_status enqueue_segment(Serve* sv, int variant, int iters, int seg) {
  thp_ctx* ctx = sv->ctx;
  Batch& b = sv->iter.b;
  if (variant) {
    if (seg == 0) {
      SRET(b_init_state(ctx, b, sv->o.poleps, sv->o.politer, sv->iter.warm, sv->iter.req));
      SRET(b_load_coords(ctx, b, kGateSetup));
      SRET(b_pme_prep(ctx, b, kGateSetup));
      SRET(b_setup(ctx, b, kGateSetup));
      return b_end_setup(ctx, b);
    }
    --seg;
  }
  const int k0 = seg * kItersPerSegment, k1 = std::min(iters, k0 + kItersPerSegment);
  for (int k = k0; k < k1; ++k) SRET(b_iteration(ctx, b));
  if (k1 < iters) return THP_OK;
  SRET(b_finish(ctx, b, 1));
  return b_collect(ctx, b, sv->d_out);
}

thp_status enqueue_pass(Serve* sv, int variant, int iters) {
  for (int s = 0; s < pass_segments(variant, iters); ++s) SRET(enqueue_segment(sv, variant, iters, s));
  return THP_OK;
}

// Run one pass, replaying the layout's captured graph of the variant when
// capture works (the graph reads the request flags and coordinates from
// the batch's device buffers, written before the launch).
thp_status run_pass(Serve* sv, int variant, int iters) {
  thp_ctx* ctx = sv->ctx;
  BatchBuf& bb = sv->iter;
  if (bb.graph_iters != iters) drop_graph(bb);
  // capture a layout once it has survived a pass: with membership changing
  // every pass (streams of many distinct systems) eager launches are cheaper
  if (sv->o.use_graphs && sv->graphs_ok && sv->layout_age >= 1) {
    if (bb.exec[variant].empty()) {
      bool ok = true;
      for (int s = 0; ok && s < pass_segments(variant, iters); ++s) {
        // the segment's own fork and join events when batches fork
        if (ctx->fork_batch && ctx->side) {
          hipEvent_t ev[2] = {nullptr, nullptr};
          for (hipEvent_t& x : ev)
            if (hipEventCreateWithFlags(&x, hipEventDisableTiming) == hipSuccess) bb.events.push_back(x);
          if (ev[0] && ev[1]) {
            bb.b.cap_fork = ev[0];
            bb.b.cap_join = ev[1];
          }
        }
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        thp_status r = e == hipSuccess ? enqueue_segment(sv, variant, iters, s) : THP_ERR_HIP;
        hipGraph_t g = nullptr;
        if (e == hipSuccess && hipStreamEndCapture(ctx->stream, &g) != hipSuccess) r = THP_ERR_HIP;
        bb.b.cap_fork = bb.b.cap_join = nullptr;
        hipGraphExec_t x = nullptr;
        if (r == THP_OK && g && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess) {
          bb.graph[variant].push_back(g);
          bb.exec[variant].push_back(x);
        } else {
          if (g) (void)hipGraphDestroy(g);
          ok = false;
        }
      }
      if (ok) {
        bb.graph_iters = iters;
      } else {
        // capture is not available for this sequence: launch eagerly from now on
        for (hipGraphExec_t x : bb.exec[variant]) (void)hipGraphExecDestroy(x);
        for (hipGraph_t g : bb.graph[variant]) (void)hipGraphDestroy(g);
        bb.exec[variant].clear();
        bb.graph[variant].clear();
        (void)hipGetLastError();
        sv->graphs_ok = false;
      }
    }
    if (!bb.exec[variant].empty()) {
      for (hipGraphExec_t x : bb.exec[variant]) SHIP(ctx, hipGraphLaunch(x, ctx->stream));
      sv->stats.graph_launches += 1;
      return THP_OK;
    }
  }
  return enqueue_pass(sv, variant, iters);
}

thp_status handle_copies(Serve* sv, std::deque<DipoleCopy*>& pending) {
  for (auto it = pending.begin(); it != pending.end();) {
    DipoleCopy* c = *it;
    ServeSystem* ss = nullptr;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (c->sid >= 0 && c->sid < (int)sv->systems.size()) ss = sv->systems[c->sid].get();
    }
    if (ss && ss->busy) {
      ++it;
      continue;
    }
    thp_status rc = THP_OK;
    if (!ss) rc = THP_ERR_ARG;
    else if (!ss->s || !ss->has_mu) rc = fail(sv->ctx, THP_ERR_STATE, "system %d has no resident solution", c->sid);
    else rc = download_mu(ss->s, c->out);
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      c->rc = rc;
      c->done = true;
    }
    sv->cv_done.notify_all();
    it = pending.erase(it);
  }
  return THP_OK;
}

void publish(Serve* sv, const InFlight& f, const StateOut& so, Clock::time_point now) {
  thp_result r{};
  r.ticket = f.req.ticket;
  r.system_id = f.req.sid;
  r.niter = so.st.niter;
  r.ep = so.st.ep;
  r.resnrm = so.st.resnrm;
  r.latency_ms = ms_between(f.req.t_submit, now);
  r.queue_ms = ms_between(f.req.t_submit, f.t_start);
  r.warm = f.warm;
  if (so.err) r.status = so.err;
  else if (so.st.conv == kConverged || so.st.conv == kZeroCurv) r.status = THP_OK;
  else if (so.st.conv == kNonFinite) r.status = THP_ERR_NAN;
  else r.status = THP_ERR_NOCONV;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->done[r.ticket] = r;
  sv->stats.solves += 1;
}

void publish_error(Serve* sv, const Request& q, thp_status code, Clock::time_point t_start) {
  thp_result r{};
  r.ticket = q.ticket;
  r.system_id = q.sid;
  r.status = code;
  const auto now = Clock::now();
  r.latency_ms = ms_between(q.t_submit, now);
  r.queue_ms = ms_between(q.t_submit, t_start);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->done[r.ticket] = r;
    sv->stats.solves += 1;
  }
  sv->cv_done.notify_all();
}

constexpr int kIdlePasses = 4;

// iteration-batch membership: add systems with a solve in flight, drop
// members idle for kIdlePasses passes
void update_members(Serve* sv) {
  for (auto& f : sv->active)
    if (f.sys->member < 0) {
      f.sys->member = (int)sv->members.size();
      sv->members.push_back(f.sys);
      sv->iter_dirty = true;
    }
  std::vector<ServeSystem*> keep;
  for (ServeSystem* m : sv->members) {
    m->idle_passes = m->busy ? 0 : m->idle_passes + 1;
    if (m->idle_passes > kIdlePasses) {
      m->member = -1;
      sv->iter_dirty = true;
    } else {
      keep.push_back(m);
    }
  }
  if (keep.size() != sv->members.size()) {
    sv->members.swap(keep);
    for (size_t q = 0; q < sv->members.size(); ++q) sv->members[q]->member = (int)q;
  }
}

// Copy the coordinates of queued requests into their systems' device
// buffers on the upload stream while the launched pass runs (see Serve::up).
// Only the first queued request of a system qualifies.
thp_status early_uploads(Serve* sv) {
  std::vector<const ServeSystem*> ready;  // running solves past a clean set-up
  for (const InFlight& f : sv->active)
    if (f.setup_seen) ready.push_back(f.sys);
  bool any = false;
  std::lock_guard<std::mutex> lk(sv->mu);
  std::vector<const ServeSystem*> seen;
  for (Request& q : sv->queue) {
    ServeSystem* ss = q.sys;
    if (std::find(seen.begin(), seen.end(), ss) != seen.end()) continue;
    seen.push_back(ss);
    if (q.pre || !ss->s || (!q.xyz && !q.rp)) continue;
    if (ss->busy && std::find(ready.begin(), ready.end(), ss) == ready.end()) continue;
    const size_t n = (size_t)ss->s->n;
    if (q.xyz)
      SHIP(sv->ctx, hipMemcpyAsync((void*)ss->s->view.raw, q.xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, sv->up));
    if (q.rp)
      SHIP(sv->ctx, hipMemcpyAsync((void*)ss->s->view.raw_rpole, q.rp, sizeof(double) * 13 * n, hipMemcpyHostToDevice,
                                   sv->up));
    q.pre = true;
    q.pre_gen = ss->gen;
    any = true;
  }
  if (any) {
    SHIP(sv->ctx, hipEventRecord(sv->ev_up, sv->up));
    sv->up_pending = true;
  }
  return THP_OK;
}

// One pass of the loop: admission (residency, coordinate uploads, lists of
// new systems), one launch sequence over the member batch, one wait, then
// the finished solves retire and solves whose coordinates outran their
// lists get the lists rebuilt and their set-up repeated in the next pass.
// Returns a fatal status only for device failures.
thp_status pass(Serve* sv, std::vector<Request>& starting) {
  thp_ctx* ctx = sv->ctx;
  const auto t_pass = Clock::now();
  // ---- admission
  std::vector<ServeSystem*> pinned;
  for (auto& f : sv->active) pinned.push_back(f.sys);
  for (auto& f : sv->restarts) pinned.push_back(f.sys);
  std::vector<Request> deferred;
  std::vector<InFlight> admitted;
  std::vector<std::pair<double*, size_t>> uploads;  // pinned buffers in flight this pass
  for (auto& q : starting) {
    ServeSystem* ss = q.sys;
    bool ok = true;
    const thp_status ra = admit(sv, ss, pinned, &ok);
    if (ra == THP_ERR_HIP) return ra;
    if (ra != THP_OK) {
      give_pinned(sv, q.xyz, q.xyz_cap);
      give_pinned(sv, q.rp, q.rp_cap);
      publish_error(sv, q, ra, t_pass);
      continue;
    }
    if (!ok) {
      deferred.push_back(std::move(q));
      continue;
    }
    pinned.push_back(ss);
    ss->last_use = ++sv->use_clock;
    const auto tc0 = Clock::now();
    thp_status rc = THP_OK;
    const bool pre = q.pre && q.pre_gen == ss->gen;  // copied during an earlier pass
    if (pre) {
      if (q.xyz) uploads.emplace_back(q.xyz, q.xyz_cap);
      if (q.rp) uploads.emplace_back(q.rp, q.rp_cap);
      q.xyz = q.rp = nullptr;
    }
    if (q.xyz) {
      const size_t n = (size_t)ss->s->n;
      if (hipMemcpyAsync((void*)ss->s->view.raw, q.xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream) !=
          hipSuccess)
        return fail(ctx, THP_ERR_HIP, "coordinate upload");
      uploads.emplace_back(q.xyz, q.xyz_cap);
      q.xyz = nullptr;
    }
    if (q.rp) {
      const size_t n = (size_t)ss->s->n;
      if (hipMemcpyAsync((void*)ss->s->view.raw_rpole, q.rp, sizeof(double) * 13 * n, hipMemcpyHostToDevice,
                         ctx->stream) != hipSuccess)
        return fail(ctx, THP_ERR_HIP, "multipole upload");
      uploads.emplace_back(q.rp, q.rp_cap);
      q.rp = nullptr;
    }
    const auto tc1 = Clock::now();
    rc = ensure_lists(ss->s);  // new systems only; others keep theirs until the skin check fails
    refresh_bytes(sv, ss);
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->stats.coords_ms += ms_between(tc0, tc1);
      sv->stats.lists_ms += ms_between(tc1, Clock::now());
    }
    if (rc != THP_OK) {
      if (rc == THP_ERR_HIP) return rc;
      publish_error(sv, q, rc, t_pass);
      continue;
    }
    InFlight f;
    f.req = std::move(q);
    f.sys = ss;
    f.t_start = t_pass;
    f.warm = sv->o.warm_start && ss->has_mu ? 1 : 0;
    ss->busy = true;
    admitted.push_back(std::move(f));
  }
  if (!deferred.empty()) {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto it = deferred.rbegin(); it != deferred.rend(); ++it) sv->queue.push_front(std::move(*it));
  }
  for (auto& f : sv->restarts) admitted.push_back(std::move(f));
  sv->restarts.clear();
  for (auto& f : admitted) sv->active.push_back(std::move(f));
  const auto t_set = Clock::now();
  if (sv->active.empty()) {
    for (auto& u : uploads) give_pinned(sv, u.first, u.second);
    return THP_OK;
  }
  // ---- the member batch: layout, request flags
  update_members(sv);
  std::vector<thp_sys*> mem;
  for (ServeSystem* m : sv->members) mem.push_back(m->s);
  if (!sv->iter_dirty) {
    for (size_t q = 0; q < mem.size(); ++q)
      if (q >= sv->iter_versions.size() || sv->iter_versions[q] != mem[q]->view_version) sv->iter_dirty = true;
  }
  if (sv->out_cap < (int)mem.size()) {
    SHIP(ctx, hipStreamSynchronize(ctx->stream));
    if (sv->d_out) (void)hipFree(sv->d_out);
    if (sv->h_out) (void)hipHostFree(sv->h_out);
    sv->out_cap = std::max((int)mem.size(), 2 * sv->out_cap);
    SHIP(ctx, hipMalloc((void**)&sv->d_out, sizeof(StateOut) * sv->out_cap));
    SHIP(ctx, hipHostMalloc((void**)&sv->h_out, sizeof(StateOut) * sv->out_cap, hipHostMallocDefault));
    sv->iter_dirty = true;  // the graphs hold the old pointer
  }
  if (sv->iter_dirty || sv->iter_generation != sv->pool_generation) {
    SRET(build_batch(sv, sv->iter, mem, {}));
    SRET(tune_pair(sv, sv->iter.b, mem));
    sv->iter_dirty = false;
    sv->iter_generation = sv->pool_generation;
    sv->iter_versions.clear();
    for (thp_sys* s : mem) sv->iter_versions.push_back(s->view_version);
    sv->stats.batch_builds += 1;
    sv->layout_age = 0;
  } else {
    ++sv->layout_age;
  }
  BatchBuf& bb = sv->iter;
  const int m = (int)mem.size();
  for (int q = 0; q < m; ++q) {
    bb.h_req[q] = 0;
    bb.h_warm[q] = 0;
  }
  for (auto& f : admitted) {
    bb.h_req[f.sys->member] = 1;
    bb.h_warm[f.sys->member] = f.warm;
  }
  const int variant = admitted.empty() ? 0 : 1;
  if (variant) {
    SHIP(ctx, hipMemcpyAsync(bb.req, bb.h_req, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
    SHIP(ctx, hipMemcpyAsync(bb.warm, bb.h_warm, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
  }
  // ---- one launch sequence for the pass, after the early uploads
  if (sv->up_pending) {
    SHIP(ctx, hipStreamWaitEvent(ctx->stream, sv->ev_up, 0));
    sv->up_pending = false;
  }
  const auto t_chunk = Clock::now();
  const int iters = std::max(1, sv->o.iters_per_graph);
  SRET(run_pass(sv, variant, iters));
  sv->stats.chunks += 1;
  sv->active_sum += (double)sv->active.size();
  sv->stats.pair_segs = bb.b.segs;
  SHIP(ctx, hipMemcpyAsync(sv->h_out, sv->d_out, sizeof(StateOut) * m, hipMemcpyDeviceToHost, ctx->stream));
  SRET(early_uploads(sv));
  const auto t_wait = Clock::now();
  SHIP(ctx, hipStreamSynchronize(ctx->stream));
  const auto t_done = Clock::now();
  for (auto& u : uploads) give_pinned(sv, u.first, u.second);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stats.host_ms += ms_between(t_pass, t_wait);
    sv->stats.wait_ms += ms_between(t_wait, t_done);
    sv->stats.admit_ms += ms_between(t_pass, t_set);
    sv->stats.setup_ms += ms_between(t_set, t_chunk);
    sv->stats.chunk_ms += ms_between(t_chunk, t_wait);
  }
  // ---- retire finished solves; rebuild lists of stale ones
  const auto now = Clock::now();
  std::vector<InFlight> still;
  for (auto& f : sv->active) {
    const StateOut& so = sv->h_out[f.sys->member];
    if (so.st.stale) {
      // coordinates moved past the skin of the lists the set-up used
      thp_status rc = sync_host_coords(f.sys->s);
      if (rc == THP_OK) rc = ensure_lists(f.sys->s);
      refresh_bytes(sv, f.sys);
      if (rc == THP_ERR_HIP) return rc;
      if (rc != THP_OK) {
        f.sys->busy = false;
        publish_error(sv, f.req, rc, f.t_start);
        continue;
      }
      f.restart = true;
      sv->restarts.push_back(std::move(f));
      continue;
    }
    f.setup_seen = true;
    if (so.err == 0 && so.st.conv == kRunning) {
      still.push_back(std::move(f));
      continue;
    }
    f.sys->busy = false;
    f.sys->has_mu = so.err == 0 && (so.st.conv == kConverged || so.st.conv == kZeroCurv);
    f.sys->last_use = ++sv->use_clock;
    publish(sv, f, so, now);
  }
  if (still.size() != sv->active.size()) {
    sv->active.swap(still);
    sv->cv_done.notify_all();
  }
  return THP_OK;
}

void worker_main(Serve* sv) {
  thp_ctx* ctx = sv->ctx;
  (void)hipSetDevice(ctx->device);
  std::deque<DipoleCopy*> copies;
  for (;;) {
    std::vector<Request> starting;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] {
        return sv->stop || !sv->active.empty() || !sv->restarts.empty() || !sv->copies.empty() || !copies.empty() ||
               (!sv->queue.empty() && (int)(sv->active.size() + sv->restarts.size()) < sv->o.max_systems);
      });
      if (sv->stop) break;
      while (!sv->copies.empty()) {
        copies.push_back(sv->copies.front());
        sv->copies.pop_front();
      }
      // FIFO admission into free slots; one solve per system at a time
      int free_slots = sv->o.max_systems - (int)(sv->active.size() + sv->restarts.size());
      std::vector<int> taken;
      for (auto it = sv->queue.begin(); it != sv->queue.end() && free_slots > 0;) {
        ServeSystem* ss = it->sys;
        if (ss->busy || std::find(taken.begin(), taken.end(), it->sid) != taken.end()) {
          taken.push_back(it->sid);  // keeps later requests of this system behind it
          ++it;
          continue;
        }
        taken.push_back(it->sid);
        starting.push_back(std::move(*it));
        it = sv->queue.erase(it);
        --free_slots;
      }
    }
    thp_status rc = handle_copies(sv, copies);
    if (rc == THP_OK) rc = pass(sv, starting);
    if (rc != THP_OK) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->fatal = rc;
      sv->fatal_msg = ctx->err;
      sv->stop = true;
      break;
    }
    if (sv->active.empty() && sv->restarts.empty() && starting.empty() && copies.empty()) {
      // nothing ran in this pass: wait for work instead of spinning
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait_for(lk, std::chrono::milliseconds(1));
    }
  }
  // cancel what did not finish
  const auto now = Clock::now();
  for (auto* list : {&sv->active, &sv->restarts}) {
    for (auto& f : *list) {
      f.sys->busy = false;
      publish_error(sv, f.req, sv->fatal != THP_OK ? sv->fatal : THP_ERR_STATE, f.t_start);
    }
    list->clear();
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  for (auto& q : sv->queue) {
    if (q.xyz) sv->pinned_pool.emplace(q.xyz_cap, q.xyz);
    if (q.rp) sv->pinned_pool.emplace(q.rp_cap, q.rp);
    thp_result r{};
    r.ticket = q.ticket;
    r.system_id = q.sid;
    r.status = sv->fatal != THP_OK ? sv->fatal : THP_ERR_STATE;
    r.latency_ms = ms_between(q.t_submit, now);
    sv->done[r.ticket] = r;
  }
  sv->queue.clear();
  for (DipoleCopy* c : copies) {
    c->rc = THP_ERR_STATE;
    c->done = true;
  }
  for (DipoleCopy* c : sv->copies) {
    c->rc = THP_ERR_STATE;
    c->done = true;
  }
  sv->copies.clear();
  sv->cv_done.notify_all();
}

Serve* serve_of(thp_ctx* ctx) { return ctx ? static_cast<Serve*>(ctx->serve) : nullptr; }

}  // namespace

extern "C" {

thp_status thp_serve_start(thp_ctx* ctx, const thp_serve_opts* opts) {
  if (!ctx || !opts) return THP_ERR_ARG;
  if (ctx->serve) return fail(ctx, THP_ERR_STATE, "serving loop already running");
  if (opts->max_systems < 1 || opts->iters_per_graph < 1 || !(opts->poleps > 0) || opts->politer < 1 ||
      opts->hbm_budget < 0)
    return fail(ctx, THP_ERR_ARG, "bad serve options");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, THP_ERR_HIP, "hipSetDevice");
  Serve* sv = new Serve;
  sv->ctx = ctx;
  sv->o = *opts;
  if (opts->hbm_budget > 0) sv->budget = opts->hbm_budget;
  else {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete sv;
      return fail(ctx, THP_ERR_HIP, "hipMemGetInfo");
    }
    sv->budget = (int64_t)(0.8 * (double)fr);
  }
  if (hipStreamCreateWithFlags(&sv->up, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&sv->ev_up, hipEventDisableTiming) != hipSuccess) {
    if (sv->up) (void)hipStreamDestroy(sv->up);
    delete sv;
    return fail(ctx, THP_ERR_HIP, "cannot create the serving loop's upload stream");
  }
  ctx->serve = sv;
  sv->worker = std::thread(worker_main, sv);
  return THP_OK;
}

thp_status thp_serve_add_system(thp_ctx* ctx, const thp_system_desc* d, int* system_id) {
  Serve* sv = serve_of(ctx);
  if (!sv) return ctx ? fail(ctx, THP_ERR_STATE, "serving loop not started") : THP_ERR_ARG;
  if (!d || !system_id || d->n <= 0 || d->n >= kMaxAtoms || !d->x || !d->y || !d->z || !d->rpole || !d->polarity ||
      !d->thole || d->nscaled < 0 || (d->nscaled > 0 && (!d->scaled_ik || !d->scaled_dpu)))
    return THP_ERR_ARG;
  auto ss = std::make_unique<ServeSystem>();
  const int n = d->n;
  ss->x.assign(d->x, d->x + n);
  ss->y.assign(d->y, d->y + n);
  ss->z.assign(d->z, d->z + n);
  ss->rpole.assign(d->rpole, d->rpole + (size_t)13 * n);
  ss->pol.assign(d->polarity, d->polarity + n);
  ss->thole.assign(d->thole, d->thole + n);
  if (d->nscaled > 0) {
    ss->ik.assign(d->scaled_ik, d->scaled_ik + (size_t)2 * d->nscaled);
    ss->dpu.assign(d->scaled_dpu, d->scaled_dpu + (size_t)3 * d->nscaled);
  }
  ss->desc = *d;
  ss->desc.x = ss->x.data();
  ss->desc.y = ss->y.data();
  ss->desc.z = ss->z.data();
  ss->desc.rpole = ss->rpole.data();
  ss->desc.polarity = ss->pol.data();
  ss->desc.thole = ss->thole.data();
  ss->desc.scaled_ik = ss->ik.empty() ? nullptr : ss->ik.data();
  ss->desc.scaled_dpu = ss->dpu.empty() ? nullptr : ss->dpu.data();
  ss->bytes = estimate_bytes(ss->desc);
  std::lock_guard<std::mutex> lk(sv->mu);
  if (!sv->systems.empty() && sv->systems[0]->desc.precision != d->precision)
    return THP_ERR_ARG;  // one precision per loop
  ss->id = (int)sv->systems.size();
  *system_id = ss->id;
  sv->systems.push_back(std::move(ss));
  return THP_OK;
}

thp_status thp_serve_set_frames(thp_ctx* ctx, int sid, const int* axistype, const int* zaxis, const int* xaxis,
                                const int* yaxis, const double* pole) {
  Serve* sv = serve_of(ctx);
  if (!sv) return ctx ? fail(ctx, THP_ERR_STATE, "serving loop not started") : THP_ERR_ARG;
  if (!axistype || !zaxis || !xaxis || !yaxis || !pole) return THP_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sid < 0 || sid >= (int)sv->systems.size()) return THP_ERR_ARG;
  ServeSystem* ss = sv->systems[sid].get();
  // before the system's first request: its device data does not exist yet
  if (ss->s || ss->busy || ss->has_frames)
    return fail(ctx, THP_ERR_STATE, "thp_serve_set_frames: system %d has been requested already", sid);
  const int n = ss->desc.n;
  for (int i = 0; i < n; ++i) {
    if (axistype[i] < 0 || axistype[i] > 5) return fail(ctx, THP_ERR_ARG, "atom %d: axis type %d", i, axistype[i]);
    for (const int* f : {zaxis, xaxis, yaxis})
      if (f[i] >= n || f[i] == i) return fail(ctx, THP_ERR_ARG, "atom %d: frame atom %d", i, f[i]);
  }
  ss->fr_axt.assign(axistype, axistype + n);
  ss->fr_z.assign(zaxis, zaxis + n);
  ss->fr_x.assign(xaxis, xaxis + n);
  ss->fr_y.assign(yaxis, yaxis + n);
  ss->pole_local.assign(pole, pole + (size_t)13 * n);
  ss->has_frames = true;
  // frames and inverse map on the device
  ss->bytes += (int64_t)(4 * sizeof(int) + 10 * sizeof(double) + 5 * sizeof(int)) * ((n + kBlock - 1) / kBlock * kBlock);
  return THP_OK;
}

thp_status thp_serve_submit_multipoles(thp_ctx* ctx, int sid, const double* x, const double* y, const double* z,
                                       const double* rpole, uint64_t* ticket) {
  Serve* sv = serve_of(ctx);
  if (!sv) return ctx ? THP_ERR_STATE : THP_ERR_ARG;
  if (!ticket || (x && (!y || !z))) return THP_ERR_ARG;
  Request q;
  q.sid = sid;
  q.t_submit = Clock::now();
  size_t n = 0;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return sv->fatal != THP_OK ? sv->fatal : THP_ERR_STATE;
    if (sid < 0 || sid >= (int)sv->systems.size()) return THP_ERR_ARG;
    n = (size_t)sv->systems[sid]->desc.n;
    q.sys = sv->systems[sid].get();
    // a system with local frames takes its multipoles from its frames
    if (rpole && q.sys->has_frames) return THP_ERR_ARG;
  }
  // caller-order arrays into pinned buffers on the caller's thread; the
  // worker uploads them as they are
  if (x) {
    q.xyz = take_pinned(sv, 3 * n, &q.xyz_cap);
    if (!q.xyz) return THP_ERR_NOMEM;
    std::memcpy(q.xyz, x, sizeof(double) * n);
    std::memcpy(q.xyz + n, y, sizeof(double) * n);
    std::memcpy(q.xyz + 2 * n, z, sizeof(double) * n);
  }
  if (rpole) {
    q.rp = take_pinned(sv, 13 * n, &q.rp_cap);
    if (!q.rp) {
      give_pinned(sv, q.xyz, q.xyz_cap);
      return THP_ERR_NOMEM;
    }
    std::memcpy(q.rp, rpole, sizeof(double) * 13 * n);
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    q.ticket = sv->next_ticket++;
    *ticket = q.ticket;
    sv->queue.push_back(std::move(q));
  }
  sv->cv_work.notify_one();
  return THP_OK;
}

thp_status thp_serve_submit(thp_ctx* ctx, int sid, const double* x, const double* y, const double* z,
                            uint64_t* ticket) {
  return thp_serve_submit_multipoles(ctx, sid, x, y, z, nullptr, ticket);
}

thp_status thp_serve_poll(thp_ctx* ctx, uint64_t ticket, thp_result* result, double wait_ms) {
  Serve* sv = serve_of(ctx);
  if (!sv) return ctx ? THP_ERR_STATE : THP_ERR_ARG;
  if (!result) return THP_ERR_ARG;
  std::unique_lock<std::mutex> lk(sv->mu);
  // ticket 0: any finished request (the oldest ticket among them)
  auto find = [&] {
    if (ticket != 0) return sv->done.find(ticket);
    auto best = sv->done.end();
    for (auto it = sv->done.begin(); it != sv->done.end(); ++it)
      if (best == sv->done.end() || it->first < best->first) best = it;
    return best;
  };
  auto ready = [&] { return find() != sv->done.end() || sv->stop; };
  if (wait_ms < 0) sv->cv_done.wait(lk, ready);
  else if (wait_ms > 0) sv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms), ready);
  auto it = find();
  if (it == sv->done.end()) return sv->stop ? (sv->fatal != THP_OK ? sv->fatal : THP_ERR_STATE) : 0;
  *result = it->second;
  sv->done.erase(it);
  return 1;
}

thp_status thp_serve_get_dipoles(thp_ctx* ctx, int sid, double* mu) {
  Serve* sv = serve_of(ctx);
  if (!sv) return ctx ? THP_ERR_STATE : THP_ERR_ARG;
  if (!mu) return THP_ERR_ARG;
  DipoleCopy c;
  c.sid = sid;
  c.out = mu;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return THP_ERR_STATE;
    if (sid < 0 || sid >= (int)sv->systems.size()) return THP_ERR_ARG;
    sv->copies.push_back(&c);
  }
  sv->cv_work.notify_one();
  std::unique_lock<std::mutex> lk(sv->mu);
  sv->cv_done.wait(lk, [&] { return c.done; });
  return c.rc;
}

thp_status thp_serve_stats_get(thp_ctx* ctx, thp_serve_stats* out) {
  Serve* sv = serve_of(ctx);
  if (!sv) return ctx ? THP_ERR_STATE : THP_ERR_ARG;
  if (!out) return THP_ERR_ARG;
  std::lock_guard<std::mutex> lk(sv->mu);
  *out = sv->stats;
  out->mean_active = sv->stats.chunks ? sv->active_sum / (double)sv->stats.chunks : 0.0;
  out->resident_bytes = sv->resident.load();
  return THP_OK;
}

thp_status thp_serve_stop(thp_ctx* ctx) {
  if (!ctx) return THP_ERR_ARG;
  Serve* sv = serve_of(ctx);
  if (!sv) return THP_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(sv->up);
  free_batchbuf(sv->setup);
  free_batchbuf(sv->iter);
  for (auto& kv : sv->pools) free_pool(kv.second);
  if (sv->d_out) (void)hipFree(sv->d_out);
  if (sv->h_out) (void)hipHostFree(sv->h_out);
  for (auto& s : sv->systems) {
    if (s->s) thp_system_destroy(s->s);
  }
  for (auto& kv : sv->pinned_pool) (void)hipHostFree(kv.second);
  (void)hipEventDestroy(sv->ev_up);
  (void)hipStreamDestroy(sv->up);
  const thp_status fatal = sv->fatal;
  if (fatal != THP_OK) ctx->err = sv->fatal_msg;
  ctx->serve = nullptr;
  delete sv;
  return fatal;
}

}  // extern "C"
