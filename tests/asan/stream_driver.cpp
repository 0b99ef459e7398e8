// Poisson stream of mixed-size solves through the serving loop, for a
// host-side AddressSanitizer build of libthp (device code unchanged).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <random>
#include <thread>
#include <vector>
#include "../../include/thp.h"

struct Sys {
  int n;
  std::vector<double> x, y, z, rp, pol, th, dpu;
  std::vector<int> ik;
  thp_system_desc d;
};

int main(int argc, char** argv) {
  const int nsys = argc > 1 ? atoi(argv[1]) : 60;
  const int nreq = argc > 2 ? atoi(argv[2]) : 1200;
  const double rate = argc > 3 ? atof(argv[3]) : 500.0;
  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  thp_ctx* ctx = nullptr;
  if (thp_init(&ctx, 0) != THP_OK) { printf("init failed\n"); return 2; }
  thp_serve_opts o{};
  o.max_systems = 16; o.hbm_budget = 0; o.iters_per_graph = 4; o.autotune = 1;
  o.poleps = 1e-8; o.politer = 100; o.warm_start = 1; o.use_graphs = 1;
  if (thp_serve_start(ctx, &o) != THP_OK) { printf("serve_start: %s\n", thp_last_error(ctx)); return 2; }
  std::vector<Sys> sys(nsys);
  std::vector<int> sid(nsys);
  for (int k = 0; k < nsys; ++k) {
    Sys& s = sys[k];
    const int n = (int)std::lround(std::exp(std::log(500.0) + U(rng) * (std::log(30000.0) - std::log(500.0))));
    const double L = std::cbrt(n / 0.1);
    const int ns = (int)std::lround(0.3 * U(rng) * n);
    const double R = ns > 0 ? std::cbrt(3.0 * ns / (4.0 * M_PI * 0.177)) : 0.0;
    s.n = n;
    s.x.resize(n); s.y.resize(n); s.z.resize(n); s.rp.resize(13 * (size_t)n); s.pol.resize(n); s.th.resize(n);
    s.ik.resize(6 * (size_t)n); s.dpu.resize(9 * (size_t)n);
    int m = 0;
    if (thp_make_system(n, L, ns, R, 1000 + k, s.x.data(), s.y.data(), s.z.data(), s.rp.data(), s.pol.data(),
                        s.th.data(), &m, s.ik.data(), s.dpu.data()) != THP_OK) { printf("make_system\n"); return 2; }
    const int K = thp_pick_grid(L);
    thp_system_desc& d = s.d;
    d = thp_system_desc{};
    d.n = n; d.box[0] = d.box[1] = d.box[2] = L;
    d.x = s.x.data(); d.y = s.y.data(); d.z = s.z.data(); d.rpole = s.rp.data(); d.polarity = s.pol.data();
    d.thole = s.th.data(); d.nscaled = m; d.scaled_ik = s.ik.data(); d.scaled_dpu = s.dpu.data();
    d.aewald = 0.4; d.cut = 7.0; d.buffer = std::min(2.0, 0.5 * L - 7.0 - 0.05);
    d.nfft[0] = d.nfft[1] = d.nfft[2] = K; d.order = 5; d.precision = THP_PREC_MIXED;
    if (thp_serve_add_system(ctx, &d, &sid[k]) != THP_OK) { printf("add_system: %s\n", thp_last_error(ctx)); return 2; }
  }
  printf("systems registered: %d\n", nsys); fflush(stdout);
  std::vector<uint64_t> warm(nsys);
  for (int k = 0; k < nsys; ++k) thp_serve_submit(ctx, sid[k], nullptr, nullptr, nullptr, &warm[k]);
  int bad = 0;
  for (int k = 0; k < nsys; ++k) {
    thp_result r{};
    if (thp_serve_poll(ctx, warm[k], &r, 600000.0) != 1 || r.status != THP_OK) ++bad;
  }
  printf("warm-up solves done, failed %d\n", bad); fflush(stdout);
  std::normal_distribution<double> N(0.0, 0.01);
  std::exponential_distribution<double> gap(rate);
  std::vector<uint64_t> tickets(nreq);
  auto t0 = std::chrono::steady_clock::now();
  double t_next = 0.0;
  for (int q = 0; q < nreq; ++q) {
    const int k = (int)(U(rng) * nsys) % nsys;
    const Sys& s = sys[k];
    std::vector<double> x(s.x), y(s.y), z(s.z);
    for (int i = 0; i < s.n; ++i) { x[i] += N(rng); y[i] += N(rng); z[i] += N(rng); }
    t_next += gap(rng);
    const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (t_next > now) std::this_thread::sleep_for(std::chrono::duration<double>(t_next - now));
    if (thp_serve_submit(ctx, sid[k], x.data(), y.data(), z.data(), &tickets[q]) != THP_OK) {
      printf("submit failed: %s\n", thp_last_error(ctx)); return 2;
    }
    if (q % 200 == 199) { printf("submitted %d\n", q + 1); fflush(stdout); }
  }
  int failed = 0;
  double worst = 0.0;
  for (int q = 0; q < nreq; ++q) {
    thp_result r{};
    if (thp_serve_poll(ctx, tickets[q], &r, 600000.0) != 1 || r.status != THP_OK) ++failed;
    worst = std::max(worst, r.latency_ms);
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  printf("stream done: %d requests, %d failed, %.1f solves/s, worst latency %.1f ms\n", nreq, failed, nreq / dt, worst);
  fflush(stdout);
  thp_serve_stop(ctx);
  thp_finalize(ctx);
  printf("finalized\n");
  return failed ? 1 : 0;
}
