"""Serving loop front end: continuous batching of independent solves.

`Server` wraps thp_serve_* (a worker thread inside libthp owns the device):
systems are registered once (`add`), solves are requested at new
coordinates (`submit`) and collected by ticket (`poll`).  `bench_serving`
drives it closed-loop on copies of one workload and reports solves/s and
p50/p99 latency; `poisson_stream` replays a Poisson arrival stream of
mixed-size systems.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import (THP_OK, THP_PREC_MIXED, Result, ServeOpts, ServeStats, System, ThpError, _d, _i,
               serving_system)


class Server:
    def __init__(self, ctx, max_systems=8, hbm_budget=0, iters_per_graph=4, autotune=1, poleps=1e-8,
                 politer=500, warm_start=1, use_graphs=0):
        self.ctx, self.lib = ctx, ctx.lib
        o = ServeOpts(max_systems=max_systems, hbm_budget=hbm_budget, iters_per_graph=iters_per_graph,
                      autotune=autotune, poleps=poleps, politer=politer, warm_start=warm_start,
                      use_graphs=use_graphs)
        ctx.check(self.lib.thp_serve_start(ctx.h, C.byref(o)), "thp_serve_start")
        self.systems: list[System] = []
        self._open = True

    def add(self, system: System, precision: int = THP_PREC_MIXED) -> int:
        d = system.desc(precision)
        sid = C.c_int(-1)
        self.ctx.check(self.lib.thp_serve_add_system(self.ctx.h, C.byref(d), C.byref(sid)), "thp_serve_add_system")
        self.systems.append(system)
        if system.frames is not None:
            # local-frame multipoles, rotated at every request's coordinates
            f = system.frames
            n = system.n
            arrs = [np.ascontiguousarray(f[k], dtype=np.int32).reshape(n) for k in ("axistype", "zaxis", "xaxis", "yaxis")]
            pole = np.ascontiguousarray(f["pole"], dtype=np.float64).reshape(n, 13)
            self.ctx.check(self.lib.thp_serve_set_frames(self.ctx.h, sid.value, *(_i(v) for v in arrs), _d(pole)),
                           "thp_serve_set_frames")
        return sid.value

    def submit(self, sid: int, x=None, y=None, z=None, rpole=None) -> int:
        """Request a solve at new coordinates and, optionally, new global-frame
        multipoles rpole (n, 13); None keeps the current ones."""
        t = C.c_uint64(0)
        xyz = (None, None, None)
        if x is not None:
            x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
            xyz = (_d(x), _d(y), _d(z))
        if rpole is None:
            rc = self.lib.thp_serve_submit(self.ctx.h, sid, *xyz, C.byref(t))
        else:
            rp = np.ascontiguousarray(rpole, dtype=np.float64).reshape(self.systems[sid].n, 13)
            rc = self.lib.thp_serve_submit_multipoles(self.ctx.h, sid, *xyz, _d(rp), C.byref(t))
        self.ctx.check(rc, "thp_serve_submit")
        return t.value

    def poll(self, ticket: int, wait_ms: float = -1.0):
        r = Result()
        rc = self.lib.thp_serve_poll(self.ctx.h, ticket, C.byref(r), wait_ms)
        if rc < 0:
            raise ThpError(rc, "thp_serve_poll")
        return r if rc == 1 else None

    def dipoles(self, sid: int) -> np.ndarray:
        n = self.systems[sid].n
        mu = np.zeros((n, 2, 3))
        self.ctx.check(self.lib.thp_serve_get_dipoles(self.ctx.h, sid, _d(mu)), "thp_serve_get_dipoles")
        return mu

    def stats(self) -> ServeStats:
        s = ServeStats()
        self.ctx.check(self.lib.thp_serve_stats_get(self.ctx.h, C.byref(s)), "thp_serve_stats_get")
        return s

    def stop(self):
        if self._open:
            self._open = False
            rc = self.lib.thp_serve_stop(self.ctx.h)
            if rc != THP_OK:
                raise ThpError(rc, self.lib.thp_last_error(self.ctx.h).decode())

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()


def _pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def bench_serving(ctx, system: System, steps: int = 20, warmup: int = 2, poleps: float = 1e-8, copies: int = 8,
                  jitter: float = 0.01, warm_start: int = 1, iters_per_graph: int = 4, use_graphs: int = 1,
                  variants: int = 4, depth: int = 2, seed: int = 11, keep_last: bool = False) -> dict:
    """Closed loop over `copies` resident systems with the geometry of
    `system` (each registered as its own system): every system keeps `depth`
    requests outstanding (they run one after the other, each warm-started
    from the previous one) and each result triggers the next request of that
    system.  Requests cycle through `variants` coordinate sets per system,
    each the reference geometry perturbed by N(0, jitter), generated before
    the timed region.  `warmup` rounds per system are not timed.  Returns
    solves/s over the timed requests and latency percentiles (submit ->
    result, so with depth 2 a request's latency includes the solve ahead of
    it).  keep_last adds "last": per copy, the coordinates, E_pol and
    dipoles of its last solve (for checks against the oracle)."""
    rng = np.random.default_rng(seed)
    srv = Server(ctx, max_systems=copies, iters_per_graph=iters_per_graph, poleps=poleps,
                 warm_start=warm_start, use_graphs=use_graphs)
    try:
        sids = [srv.add(system) for _ in range(copies)]
        coords = [[tuple(np.ascontiguousarray(v + rng.normal(0.0, jitter, system.n))
                         for v in (system.x, system.y, system.z)) for _ in range(variants)] for _ in sids]
        nxt = [0] * copies
        last_ticket, last_ep = [None] * copies, [None] * copies

        def submit(k):
            xyz = coords[k][nxt[k] % variants]
            nxt[k] += 1
            last_ticket[k] = srv.submit(sids[k], *xyz)
            return last_ticket[k]

        # untimed rounds: device data creation, lists, plans, tuning
        for _ in range(max(1, warmup)):
            ts = [submit(k) for k in range(copies)]
            for t in ts:
                r = srv.poll(t)
                if r.status != THP_OK:
                    raise ThpError(r.status, "serving warm-up solve")
        total = steps * copies
        lat, nit, eps = [], [], []
        owner = {}
        st0 = srv.stats()
        t0 = time.perf_counter()
        issued = 0
        for _ in range(max(1, depth)):
            for k in range(copies):
                if issued < total:
                    owner[submit(k)] = k
                    issued += 1
        while owner:
            r = srv.poll(0)  # any finished request
            k = owner.pop(r.ticket)
            if r.status != THP_OK:
                raise ThpError(r.status, "serving solve")
            lat.append(r.latency_ms)
            nit.append(r.niter)
            eps.append(r.ep)
            if r.ticket == last_ticket[k]:
                last_ep[k] = r.ep
            if issued < total:
                owner[submit(k)] = k
                issued += 1
        dt = time.perf_counter() - t0
        st = srv.stats()
        d = _stats_delta(st0, st)
        out = dict(solves_per_s=total / dt, ms_per_solve=1e3 * dt / total, solves=total, copies=copies,
                   niter=float(np.mean(nit)), ep=float(eps[0]), p50_ms=_pct(lat, 50), p99_ms=_pct(lat, 99),
                   pair_segs=st.pair_segs, warm_start=warm_start, iters_per_graph=iters_per_graph,
                   graphs=int(st.graph_launches > 0), depth=depth, **d)
        if keep_last:
            out["last"] = [dict(xyz=coords[k][(nxt[k] - 1) % variants], ep=last_ep[k], mu=srv.dipoles(sids[k]))
                           for k in range(copies)]
        return out
    finally:
        srv.stop()


def _stats_delta(a, b):
    """Worker figures between two stats snapshots (the timed region)."""
    passes = b.chunks - a.chunks
    active = b.mean_active * b.chunks - a.mean_active * a.chunks
    out = dict(passes=passes, mean_active=active / passes if passes else 0.0,
               batch_builds=b.batch_builds - a.batch_builds, graph_launches=b.graph_launches - a.graph_launches)
    for f in ("host_ms", "wait_ms", "chunk_ms", "admit_ms", "setup_ms", "coords_ms", "lists_ms"):
        out["worker_" + f] = round(getattr(b, f) - getattr(a, f), 1)
    return out


def poisson_stream(ctx, n_requests=2000, rate_per_s=100.0, resubmits=10, n_min=500, n_max=30000, max_systems=16,
                   poleps=1e-8, seed=5, jitter=0.01, warm_start=1, iters_per_graph=4, use_graphs=1,
                   more_rates=(), more_requests=None):
    """Poisson arrivals of solves of independent mixed-size systems: sizes
    log-uniform in [n_min, n_max], 0-30% solute; each system is requested
    `resubmits` times at coordinates perturbed by N(0, jitter).  Systems are
    registered and solved once before the stream (device data resident, as
    for a server that has seen them), and every request's coordinates are
    generated before the clock starts.  Returns throughput and latency
    percentiles (submit -> result) at rate_per_s; with more_rates, the same
    systems then take one stream of `more_requests` requests at each of those
    rates, reported under "rates" (the capacity sweep)."""
    rng = np.random.default_rng(seed)
    n_sys = max(1, n_requests // resubmits)
    srv = Server(ctx, max_systems=max_systems, iters_per_graph=iters_per_graph, poleps=poleps,
                 warm_start=warm_start, use_graphs=use_graphs)
    try:
        systems = []
        for k in range(n_sys):
            n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
            systems.append(serving_system(n, float(rng.uniform(0.0, 0.3)), seed=1000 + k))
        sids = [srv.add(s) for s in systems]
        for r in [srv.poll(t) for t in [srv.submit(sid) for sid in sids]]:
            if r.status != THP_OK:
                raise ThpError(r.status, "stream warm-up solve")

        def replay(rate, count):
            order = np.repeat(np.arange(n_sys), resubmits)
            rng.shuffle(order)
            order = order[:count]
            coords = [tuple(np.ascontiguousarray(v + rng.normal(0.0, jitter, systems[k].n))
                            for v in (systems[k].x, systems[k].y, systems[k].z)) for k in order]
            gaps = rng.exponential(1.0 / rate, len(order))
            tickets = []
            st0 = srv.stats()
            t0 = time.perf_counter()
            t_next = t0
            for q, k in enumerate(order):
                t_next += gaps[q]
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets.append(srv.submit(sids[k], *coords[q]))
            t_sub = time.perf_counter() - t0
            lat, bad, nit = [], 0, []
            for t in tickets:
                r = srv.poll(t)
                if r.status != THP_OK:
                    bad += 1
                lat.append(r.latency_ms)
                nit.append(r.niter)
            dt = time.perf_counter() - t0
            st = srv.stats()
            atoms = sum(systems[k].n for k in order)
            return dict(requests=len(order), systems=n_sys, failed=bad, seconds=round(dt, 3),
                        solves_per_s=len(order) / dt, offered_rate=rate, submit_seconds=round(t_sub, 3),
                        p50_ms=_pct(lat, 50), p99_ms=_pct(lat, 99), niter=float(np.mean(nit)),
                        evictions=st.evictions, mean_atoms=atoms / len(order),
                        max_systems=max_systems, **_stats_delta(st0, st))

        out = replay(rate_per_s, n_requests)
        if more_rates:
            out["rates"] = [replay(r, more_requests or n_requests) for r in more_rates]
        return out
    finally:
        srv.stop()
