"""The batched and served workloads at their full stated sizes.

Batched: 64 road-like instances, n log-uniform in [10k, 500k], K uniform in
{2..8}, 1k-50k changes (40 % insertions), Pref random in [1, 4], in ONE
persistent launch replayed as a HIP graph.  Every instance's K trees equal
canonical Dijkstra on its updated graph (the reference's property, README
lines on canonical trees; src/mospTest.cu:249-298 for the pipeline), its
MOSP tree equals Dijkstra on the combined graph of those trees, and the
batched outputs equal each instance's own fused launch.

Served: 4 road grids of 1400 x 1400 (K = 3) and 4 R-MAT scale-20 graphs
(K = 4), Poisson-arriving requests of 100-100k changes (50 % insertions,
10 % disconnecting deletions) served with overlapping launches.  Every
request's result is identical in the mix and alone, and equals canonical
Dijkstra (trees) and Dijkstra on the combined graph (MOSP tree) of its
graph's state right after it, replayed on the host.

Oracle checks run in a thread pool (the oracle is C called through ctypes,
which releases the GIL).  Tolerance: none (integer outputs, bit-exact).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
from oracle_lib import i32

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def M():
    import mosp_amd
    return mosp_amd


def canonical_outputs(og: O.Graph, K: int, pref, source: int = 0):
    """(K trees by Dijkstra, combined distances, MOSP tree) of graph og."""
    dist = np.zeros((K, og.n), np.int64)
    par = np.zeros((K, og.n), np.int32)
    for k in range(K):
        dist[k], par[k] = O.dijkstra(og, k, source)
    comb = O.combined_graph(par.reshape(-1), og.n, K, source, list(pref))
    assert comb is not None
    cd, cp = O.dijkstra(comb, 0, source)
    return dist, par, cd, cp


def assert_canonical(got, og: O.Graph, K: int, pref, tag: str):
    d, p, cd, cp = canonical_outputs(og, K, pref)
    for k in range(K):
        assert np.array_equal(got["dist"][k], d[k]), f"{tag} obj{k} distances"
        assert np.array_equal(got["parents"][k], p[k]), f"{tag} obj{k} parents"
    assert np.array_equal(got["comb_dist"], cd), f"{tag} combined distances"
    assert np.array_equal(got["comb_parent"], cp), f"{tag} MOSP tree"
    return True


def test_batched_64_uneven_instances_full_size():
    from mosp_amd import batched as B
    t0 = time.perf_counter()
    specs = B.road_specs(count=64)
    built = []
    for i, spec in enumerate(specs):
        built.append(B.build_instance(spec))
        if i % 8 == 7:
            say(f"built {i + 1}/64 instances ({time.perf_counter() - t0:.1f} s)")
    sessions = [s for _, s in built]
    total_n = sum(s.n for s in sessions)
    changes = sum(h[1].n_ins + h[1].n_del for h, _ in built)
    assert len(sessions) == 64 and 10_000 <= min(s.n for s in sessions)
    assert max(s.n for s in sessions) <= 500_000 and {s.K for s in sessions} <= set(range(2, 9))
    bu = B.BatchedUpdate(sessions)
    ms = bu.run(use_graph=True)
    ms = bu.run(use_graph=True)  # a replay of the captured graph
    bu.check()
    sizes, mode = bu.schedule()
    say(f"batched: {total_n} vertices, {changes} changes, {ms:.2f} ms per pass, "
        f"graph mode {mode}, groups {min(sizes)}-{max(sizes)} workgroups")
    assert mode in (1, 2)
    results = [bu.results(i) for i in range(64)]
    # Each instance's own fused launch writes the same outputs.
    for i, s in enumerate(sessions):
        s.run()
        fused = s.results()
        for key in ("dist", "parents", "comb_dist", "comb_parent"):
            assert np.array_equal(fused[key], results[i][key]), f"instance {i} fused vs batched {key}"
    say(f"fused launches agree ({time.perf_counter() - t0:.1f} s)")

    def check(i: int) -> bool:
        (g, b, _, _), s = built[i]
        og = O.Graph(g.n, g.K, g.row, g.col, g.w)
        ob = O.Batch(b.K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to)
        upd, _ = O.apply_batch(og, ob)
        return assert_canonical(results[i], upd, specs[i].K, specs[i].pref, f"instance {i}")

    with cf.ThreadPoolExecutor(WORKERS) as pool:
        done = 0
        for ok in pool.map(check, range(64)):
            assert ok
            done += 1
            if done % 16 == 0:
                say(f"oracle: {done}/64 instances canonical ({time.perf_counter() - t0:.1f} s)")
    bu.close()
    for s in sessions:
        s.close()


def served_graphs():
    Mm = M()
    out = []
    for i in range(4):
        out.append((Mm.gen_road_grid(1400, 1400, 3, seed=40 + i), 3, (1, 1, 1)))
    for i in range(4):
        out.append((Mm.gen_rmat(20, 16, 4, seed=50 + i), 4, (1, 2, 3, 4)))
    return out


def test_served_8_resident_graphs_full_size():
    from mosp_amd import serve as S
    Mm = M()
    t0 = time.perf_counter()
    graphs = served_graphs()
    say(f"graphs: {[g.n for g, _, _ in graphs]} vertices ({time.perf_counter() - t0:.1f} s)")
    trees = [Mm.initial_trees(g, K) for g, K, _ in graphs]
    say(f"initial trees on the GPU ({time.perf_counter() - t0:.1f} s)")
    reqs = S.poisson_requests([g for g, _, _ in graphs], 24, rate=150.0, seed=77,
                              size_min=100, size_max=100_000)
    assert {r.gid for r in reqs} == set(range(8))

    def serve(**kw):
        srv = S.Server(max_batch=100_000, keep_results=True, **kw)
        for (g, K, pref), (d, p) in zip(graphs, trees):
            srv.add_graph(g, d, p, K=K, pref=pref)
        srv.start()
        return srv

    mix = serve(slots=4, max_coalesce=4, cu_policy="greedy")
    mix_reqs = [S.Request(r.rid, r.gid, r.batch, arrival=r.arrival) for r in reqs]
    stats = mix.play(mix_reqs)
    mix.close()
    say(f"mix: p50 {stats.p50_ms:.1f} ms, p99 {stats.p99_ms:.1f} ms, "
        f"{stats.changes_per_s / 1e6:.2f} M changes/s, mean coalesced {stats.mean_coalesced:.2f}")
    assert max(r.coalesced for r in mix_reqs) > 1 or max(r.slot for r in mix_reqs) > 0, \
        "nothing overlapped on the GPU"
    alone = serve(slots=1, max_coalesce=1)
    alone_reqs = [S.Request(r.rid, r.gid, r.batch) for r in reqs]
    for r in alone_reqs:
        alone.submit(r)
        alone.wait(r)
    alone.close()
    say(f"alone done ({time.perf_counter() - t0:.1f} s)")
    for a, b in zip(mix_reqs, alone_reqs):
        for key in ("dist", "parents", "comb_dist", "comb_parent"):
            assert np.array_equal(a.result[key], b.result[key]), \
                f"request {a.rid} (graph {a.gid}) mix vs alone: {key}"

    # The host replay: each graph's state after each of its requests.
    def replay(gid: int) -> int:
        g, K, pref = graphs[gid]
        og = O.Graph(g.n, g.K, g.row, g.col, g.w)
        checked = 0
        for r in [x for x in alone_reqs if x.gid == gid]:
            b = r.batch
            og, _ = O.apply_batch(og, O.Batch(g.K, b.ins_from, b.ins_to, b.ins_w,
                                              b.del_from, b.del_to))
            assert_canonical(r.result, og, K, pref, f"request {r.rid} (graph {gid})")
            checked += 1
        return checked

    with cf.ThreadPoolExecutor(min(8, WORKERS)) as pool:
        counts = list(pool.map(replay, range(8)))
    say(f"oracle: {sum(counts)} requests canonical ({time.perf_counter() - t0:.1f} s)")
    assert sum(counts) == len(reqs)
    disconnected = sum(int((r.result["dist"] >= O.INF // 2).any()) for r in alone_reqs)
    say(f"{disconnected} requests left some vertex unreachable")
