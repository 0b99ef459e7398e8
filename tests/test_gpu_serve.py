"""Serving loop on the GPU: batched results against one-system solves and
the FP64 oracle, the same result alone and in the mix, warm starts,
continuous admission, HBM-budget eviction, captured-graph replay."""
import numpy as np
import pytest

import thp
from thp.serve import Server

pytestmark = pytest.mark.gpu

MIXED = thp.THP_PREC_MIXED
E_TOL = 1e-6      # relative polarization energy vs the FP64 oracle
MU_TOL_D = 1e-5   # RMS dipole difference, Debye


def _rms_debye(a, b):
    return float(np.sqrt(np.mean((np.asarray(a) - np.asarray(b)) ** 2)) * thp.DEBYE)


def _systems():
    return [
        thp.water_1944(seed=3),
        thp.make_system(6000, 40.0, n_solute=1500, blob_radius=13.0, seed=22, nfft=(40,) * 3, name="solute-6000"),
        thp.serving_system(900, 0.2, seed=31),
        thp.serving_system(3000, 0.0, seed=32),
    ]


@pytest.fixture(scope="module")
def oracle_results():
    from oracle import oracle as O
    O.set_threads(8)
    out = []
    for s in _systems():
        orc = O.Oracle(s)
        orc.nblist_pairs()
        out.append(orc.induce(poleps=1e-8))
    return out


def _check(res, mu, ref):
    assert res.status == thp.THP_OK
    assert abs(res.ep - ref["ep"]) <= E_TOL * abs(ref["ep"]), (res.ep, ref["ep"])
    assert _rms_debye(mu, ref["mu"]) <= MU_TOL_D


def test_batched_solves_match_oracle_and_one_system_path(ctx, oracle_results):
    systems = _systems()
    with Server(ctx, max_systems=4, warm_start=0) as srv:
        sids = [srv.add(s) for s in systems]
        tickets = [srv.submit(sid) for sid in sids]
        results = [srv.poll(t) for t in tickets]
        mus = [srv.dipoles(sid) for sid in sids]
    for s, r, mu, ref in zip(systems, results, mus, oracle_results):
        _check(r, mu, ref)
        one = thp.Solver(ctx, s, MIXED)
        m1, e1, st1 = one.newinduce_pme2()
        one.close()
        assert abs(r.ep - e1) <= 1e-8 * abs(e1)
        assert abs(r.niter - st1.niter) <= 1
        assert _rms_debye(mu, m1) <= 1e-6


def test_result_is_the_same_alone_and_in_the_mix(ctx):
    systems = _systems()
    with Server(ctx, max_systems=1, warm_start=0) as srv:
        sid = srv.add(systems[1])
        alone = srv.poll(srv.submit(sid))
        mu_alone = srv.dipoles(sid)
    with Server(ctx, max_systems=4, warm_start=0) as srv:
        sids = [srv.add(s) for s in systems]
        tickets = [srv.submit(s) for s in sids]
        mix = [srv.poll(t) for t in tickets]
        mu_mix = srv.dipoles(sids[1])
        stats = srv.stats()
    assert stats.mean_active > 1.0  # the solves really shared launches
    r = mix[1]
    assert r.status == alone.status == thp.THP_OK
    assert abs(r.ep - alone.ep) <= 1e-8 * abs(alone.ep), (r.ep, alone.ep)
    assert abs(r.niter - alone.niter) <= 1
    assert _rms_debye(mu_mix, mu_alone) <= 1e-6


def test_warm_starts_and_coordinate_updates(ctx):
    s = thp.serving_system(3000, 0.1, seed=40)
    rng = np.random.default_rng(4)
    with Server(ctx, max_systems=2, warm_start=1) as srv:
        sid = srv.add(s)
        first = srv.poll(srv.submit(sid))
        assert first.status == thp.THP_OK and first.warm == 0
        x, y, z = (v + rng.normal(0.0, 0.01, s.n) for v in (s.x, s.y, s.z))
        warm = srv.poll(srv.submit(sid, x, y, z))
        mu_warm = srv.dipoles(sid)
    assert warm.status == thp.THP_OK and warm.warm == 1
    assert warm.niter < first.niter
    moved = s.with_coords(x, y, z)
    one = thp.Solver(ctx, moved, MIXED)
    m1, e1, _ = one.newinduce_pme2()
    one.close()
    assert abs(warm.ep - e1) <= E_TOL * abs(e1)
    assert _rms_debye(mu_warm, m1) <= MU_TOL_D


def test_continuous_admission_more_requests_than_slots(ctx):
    systems = [thp.serving_system(n, 0.1, seed=50 + k) for k, n in enumerate((600, 1500, 2500, 800, 4000))]
    rng = np.random.default_rng(9)
    with Server(ctx, max_systems=2, warm_start=1) as srv:
        sids = [srv.add(s) for s in systems]
        tickets = []
        for rep in range(3):
            for sid, s in zip(sids, systems):
                x, y, z = (v + rng.normal(0.0, 0.01, s.n) for v in (s.x, s.y, s.z))
                tickets.append((sid, srv.submit(sid, x, y, z), (x, y, z)))
        results = [(sid, srv.poll(t), xyz) for sid, t, xyz in tickets]
        stats = srv.stats()
    assert all(r.status == thp.THP_OK for _, r, _ in results)
    assert stats.solves == len(tickets)
    # the last request of each system against a one-system solve at its coordinates
    last = {}
    for sid, r, xyz in results:
        last[sid] = (r, xyz)
    for sid, (r, xyz) in last.items():
        one = thp.Solver(ctx, systems[sid].with_coords(*xyz), MIXED)
        _, e1, _ = one.newinduce_pme2()
        one.close()
        assert abs(r.ep - e1) <= E_TOL * abs(e1)


def test_hbm_budget_evicts_idle_systems(ctx):
    systems = [thp.serving_system(3000, 0.0, seed=60 + k) for k in range(4)]
    # footprint of one resident serving system (its grid lives in the pool)
    with Server(ctx, max_systems=1) as srv:
        srv.poll(srv.submit(srv.add(systems[0])))
        fp = srv.stats().resident_bytes
    assert fp > 0
    with Server(ctx, max_systems=1, hbm_budget=int(2.5 * fp), warm_start=1) as srv:
        sids = [srv.add(s) for s in systems]
        res = [srv.poll(srv.submit(sid)) for sid in sids + sids]
        stats = srv.stats()
    assert all(r.status == thp.THP_OK for r in res)
    assert stats.evictions > 0
    assert stats.resident_systems <= 2
    # a re-admitted system lost its dipoles and starts cold again
    assert res[4].warm == 0


def test_system_over_budget_is_an_error_result(ctx):
    s = thp.serving_system(3000, 0.0, seed=70)
    with Server(ctx, max_systems=1, hbm_budget=1 << 20) as srv:
        sid = srv.add(s)
        r = srv.poll(srv.submit(sid))
    assert r.status == -4  # THP_ERR_NOMEM


@pytest.mark.parametrize("iters", [1, 3])
def test_graph_replay_matches_eager(ctx, iters):
    systems = _systems()[:3]
    out = {}
    for graphs in (0, 1):
        with Server(ctx, max_systems=3, warm_start=0, use_graphs=graphs, iters_per_graph=iters, autotune=0) as srv:
            sids = [srv.add(s) for s in systems]
            rounds = []
            for _ in range(3):  # the graph is replayed across rounds and chunks
                ts = [srv.submit(sid) for sid in sids]
                rounds.append([srv.poll(t) for t in ts])
            mus = [srv.dipoles(sid) for sid in sids]
            out[graphs] = (rounds, mus, srv.stats())
    (r0, m0, _), (r1, m1, st1) = out[0], out[1]
    assert st1.graph_launches > 0
    for a_round, b_round in zip(r0, r1):
        for a, b in zip(a_round, b_round):
            assert a.status == b.status == thp.THP_OK
            assert a.ep == b.ep and a.niter == b.niter
    for a, b in zip(m0, m1):
        assert np.array_equal(a, b)


def test_coordinates_past_the_skin_rebuild_lists(ctx):
    # the lists allow 0.25 A of motion (skin/2) before the block-union lists
    # and 0.75 A before the block candidates must be rebuilt; the serving
    # loop detects it on the device, rebuilds and repeats the set-up
    s = thp.serving_system(3000, 0.1, seed=41)
    rng = np.random.default_rng(12)
    moves = [(s.x + 0.3, s.y, s.z), (s.x + 1.2, s.y - 0.9, s.z),
             tuple(v + rng.normal(0.0, 0.2, s.n) for v in (s.x, s.y, s.z))]
    with Server(ctx, max_systems=2, warm_start=1) as srv:
        sid = srv.add(s)
        assert srv.poll(srv.submit(sid)).status == thp.THP_OK
        results = []
        for x, y, z in moves:
            r = srv.poll(srv.submit(sid, x, y, z))
            results.append((r, srv.dipoles(sid)))
    for (r, mu), (x, y, z) in zip(results, moves):
        assert r.status == thp.THP_OK
        one = thp.Solver(ctx, s.with_coords(x, y, z), MIXED)
        m1, e1, _ = one.newinduce_pme2()
        one.close()
        assert abs(r.ep - e1) <= E_TOL * abs(e1), (r.ep, e1)
        assert _rms_debye(mu, m1) <= MU_TOL_D


def test_poisson_stream_of_mixed_sizes_matches_one_system_solves(ctx):
    # a short stream as the serving benchmark runs it: 20 systems of 500 to
    # 30,000 atoms (log-uniform, 0-30% solute), each requested 3 times at
    # jittered coordinates (warm starts), Poisson arrivals, captured graphs;
    # every result against a one-system solve at the same coordinates, the
    # last dipoles of each system too, and a sample against the oracle
    import time
    from oracle import oracle as O
    rng = np.random.default_rng(21)
    sizes = [500, 30000] + [int(round(np.exp(rng.uniform(np.log(500), np.log(30000))))) for _ in range(18)]
    systems = [thp.serving_system(n, float(rng.uniform(0.0, 0.3)), seed=900 + k) for k, n in enumerate(sizes)]
    order = np.repeat(np.arange(len(systems)), 3)
    rng.shuffle(order)
    coords = [tuple(v + rng.normal(0.0, 0.01, systems[k].n) for v in (systems[k].x, systems[k].y, systems[k].z))
              for k in order]
    gaps = rng.exponential(1.0 / 400.0, len(order))
    with Server(ctx, max_systems=8, warm_start=1, use_graphs=1) as srv:
        sids = [srv.add(s) for s in systems]
        tickets = []
        t_next = time.perf_counter()
        for q, k in enumerate(order):
            t_next += gaps[q]
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(sids[k], *coords[q]))
        results = [srv.poll(t) for t in tickets]
        last_mu = {int(k): srv.dipoles(sids[k]) for k in set(order.tolist())}
        stats = srv.stats()
    print(f"stream: {len(results)} solves, mean active {stats.mean_active:.2f}, graphs {stats.graph_launches}",
          flush=True)
    # the autotuner picked grid shapes for every grid of the stream; the
    # one-system solves below are created with them
    shapes = {tuple(s.nfft): ctx.grid_shape(s.nfft) for s in systems}
    print(f"grid shapes: {shapes}", flush=True)
    assert all(v is not None and 4 <= v[0] <= 8 and v[1] in (8, 16, 32) for v in shapes.values())
    assert all(r.status == thp.THP_OK for r in results)
    assert stats.mean_active > 1.0
    last_q = {}
    for q, k in enumerate(order):
        last_q[int(k)] = q
    for q, (k, r) in enumerate(zip(order, results)):
        one = thp.Solver(ctx, systems[k].with_coords(*coords[q]), MIXED)
        m1, e1, _ = one.newinduce_pme2()
        one.close()
        assert abs(r.ep - e1) <= E_TOL * abs(e1), (sizes[k], r.ep, e1)
        if last_q[int(k)] == q:
            assert _rms_debye(last_mu[int(k)], m1) <= MU_TOL_D
    O.set_threads(8)
    for k in (0, 1, 7):  # 500 and 30,000 atoms and one between
        q = last_q[k]
        ref = O.Oracle(systems[k].with_coords(*coords[q])).induce(poleps=1e-8)
        assert abs(results[q].ep - ref["ep"]) <= E_TOL * abs(ref["ep"])
        assert _rms_debye(last_mu[k], ref["mu"]) <= MU_TOL_D


def test_local_frames_enter_the_loop_unrotated(ctx):
    # dhfr2 (every axis type of a protein, chiral sites, polarization
    # groups) and a framed water box are registered with local-frame
    # multipoles; each request's set-up rotates them at the request's
    # coordinates on the device.  Every result against a one-system solve
    # with the same frames at the same coordinates.
    import os
    import thp.tinker as tinker
    from test_oracle_pins import _water_frames
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ds = tinker.load(os.path.join(gold, "dhfr2.xyz.gz"), os.path.join(gold, "dhfr2.key"))
    w = thp.water_1944(seed=5)
    fr = _water_frames(w)
    w = w.with_multipoles(np.zeros((w.n, 13)))
    w.frames = dict(axistype=fr[0], zaxis=fr[1], xaxis=fr[2], yaxis=fr[3], pole=fr[4])
    systems = [ds, w]
    rng = np.random.default_rng(17)
    coords = [[tuple(v + rng.normal(0.0, 0.03, s.n) for v in (s.x, s.y, s.z)) for _ in range(3)] for s in systems]
    with Server(ctx, max_systems=4, warm_start=1, use_graphs=1) as srv:
        sids = [srv.add(s) for s in systems]
        with pytest.raises(thp.ThpError):
            srv.submit(sids[1], *coords[1][0], rpole=np.zeros((w.n, 13)))
        results, mus = [[], []], [[], []]
        for r in range(3):
            tickets = [srv.submit(sid, *coords[k][r]) for k, sid in enumerate(sids)]
            for k, t in enumerate(tickets):
                results[k].append(srv.poll(t))
                mus[k].append(srv.dipoles(sids[k]))
        with pytest.raises(thp.ThpError) as e:
            srv.ctx.check(srv.lib.thp_serve_set_frames(ctx.h, sids[1], *(thp._i(np.zeros(w.n, np.int32)) for _ in range(4)),
                                                       thp._d(np.zeros((w.n, 13)))), "thp_serve_set_frames")
        assert e.value.code == -8
    for k, s in enumerate(systems):
        for r in range(3):
            res = results[k][r]
            assert res.status == thp.THP_OK
            one = thp.Solver(ctx, s.with_coords(*coords[k][r]), MIXED)
            m1, e1, _ = one.newinduce_pme2()
            one.close()
            print(f"frames in the loop: {s.n} atoms request {r} ep={res.ep:.9f} one-system {e1:.9f}", flush=True)
            assert abs(res.ep - e1) <= E_TOL * abs(e1), (s.n, r, res.ep, e1)
            assert _rms_debye(mus[k][r], m1) <= MU_TOL_D
    # the rotation follows the coordinates: the three requests' energies differ
    for k in range(2):
        eps = [r.ep for r in results[k]]
        assert max(eps) - min(eps) > 1e-4 * abs(eps[0])
