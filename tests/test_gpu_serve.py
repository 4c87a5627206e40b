"""Served updates (mosp_amd.serve): resident graphs, continuous batching.

Every request's result -- the K updated trees, combined distances and MOSP
tree of its graph right after it -- must be the same when requests overlap
on the GPU (several slots, several requests per launch) as when each runs
alone, and equal the oracle's pipeline replayed request by request on the
host (host apply + sequential SOSP update + combined graph + Dijkstra).
Tolerance: none (integer outputs, bit-exact).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import i32

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def S():
    from mosp_amd import serve
    return serve


def M():
    import mosp_amd
    return mosp_amd


def small_graphs():
    """Host graphs + K + Pref of a small served mix."""
    Mm = M()
    out = []
    for (w, h, K, pref, seed) in [(25, 25, 3, (1, 1, 1), 31), (30, 20, 2, (1, 3), 32),
                                  (18, 18, 4, (2, 1, 4, 3), 33)]:
        g = gen.grid_graph(w, h, K, 100, 0.3, seed=seed)
        out.append((Mm.HostGraph(g.n, g.K, i32(g.row), i32(g.col), i32(g.w)), K, pref))
    r = Mm.gen_rmat(9, 8, 3, seed=34)
    out.append((r, 3, (1, 2, 2)))
    return out


def make_server(graphs, **kw):
    """A started server holding `graphs` with their canonical initial trees."""
    srv = S().Server(**kw)
    for g, K, pref in graphs:
        og = O.Graph(g.n, g.K, g.row, g.col, g.w)
        d, p = gen.initial_trees(og, K)
        srv.add_graph(g, d, p, K=K, pref=pref)
    srv.start()
    return srv


def requests(graphs, count=36, seed=5):
    return S().poisson_requests([g for g, _, _ in graphs], count, rate=400.0, seed=seed,
                                size_min=10, size_max=300)


def fresh(reqs):
    return [S().Request(r.rid, r.gid, r.batch, arrival=r.arrival) for r in reqs]


def same(a, b, tag):
    for key in ("dist", "parents", "comb_dist", "comb_parent"):
        assert np.array_equal(a[key], b[key]), f"{tag}: {key}"


def test_mix_equals_alone_equals_oracle():
    graphs = small_graphs()
    reqs = requests(graphs)
    # In the mix: 3 slots with fixed CU shares, up to 3 requests per launch,
    # every request queued at once (so launches coalesce).
    mix = make_server(graphs, slots=3, max_coalesce=3, max_batch=400, keep_results=True)
    mix_reqs = fresh(reqs)
    for r in mix_reqs:
        r.arrival = 0.0
    say(f"mix: {mix.play(mix_reqs)}")
    mix.close()
    assert max(r.coalesced for r in mix_reqs) > 1, "no launch was coalesced"
    # Greedy CU grants, Poisson arrivals at 400/s.
    # (Launches take the idle CUs; a lone request gets them all.)
    greedy = make_server(graphs, slots=3, max_coalesce=2, max_batch=400, keep_results=True,
                         cu_policy="greedy")
    greedy_reqs = fresh(reqs)
    say(f"greedy: {greedy.play(greedy_reqs)}")
    greedy.close()
    # Alone: one slot, one request per launch, each submitted after the last is done.
    alone = make_server(graphs, slots=1, max_coalesce=1, max_batch=400, keep_results=True)
    alone_reqs = fresh(reqs)
    for r in alone_reqs:
        alone.submit(r)
        alone.wait(r)
        assert r.coalesced == 1
    alone.close()
    for a, g, b in zip(mix_reqs, greedy_reqs, alone_reqs):
        same(a.result, b.result, f"request {a.rid} (graph {a.gid}) mix vs alone")
        same(g.result, b.result, f"request {g.rid} (graph {g.gid}) greedy vs alone")
    # Oracle: replay every graph's requests in order on the host.
    for gid, (g, K, pref) in enumerate(graphs):
        og = O.Graph(g.n, g.K, g.row, g.col, g.w)
        d, p = gen.initial_trees(og, K)
        for r in [x for x in alone_reqs if x.gid == gid]:
            b = r.batch
            ob = O.Batch(g.K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to)
            want = O.mosp_update(og, ob, d, p, K, 0, list(pref))
            assert want["ok"]
            same(r.result, want, f"request {r.rid} (graph {gid}) vs oracle")
            og, d, p = want["updated"], want["dist"], want["parents"]


def test_budget_admission_limits_coalescing():
    graphs = small_graphs()[:3]
    reqs = requests(graphs, count=16, seed=7)
    probe = make_server(graphs, slots=2, max_coalesce=4, max_batch=400)
    resident = probe.counters()[0].resident_bytes
    probe.close()
    foot = max(S().footprint(g.n, K) for g, K, _ in graphs)
    # Room for exactly one update per slot, whatever max_coalesce says.
    srv = make_server(graphs, slots=2, max_coalesce=4, max_batch=400,
                      hbm_budget=resident + 2 * foot + 1024)
    run = fresh(reqs)
    say(f"budgeted: {srv.play(run)}")
    assert all(r.coalesced == 1 for r in run)
    assert all(not np.isnan(r.done) for r in run)
    srv.close()
    small = S().Server(slots=2, hbm_budget=resident + foot // 2)
    for g, K, pref in graphs:
        og = O.Graph(g.n, g.K, g.row, g.col, g.w)
        d, p = gen.initial_trees(og, K)
        small.add_graph(g, d, p, K=K, pref=pref)
    with pytest.raises(RuntimeError):
        small.start()
    small.close()


def test_request_errors_are_reported():
    graphs = small_graphs()[:1]
    srv = make_server(graphs, slots=1, max_batch=50)
    big = requests(graphs, count=1, seed=9)[0]
    z = np.zeros(0, np.int32)
    with pytest.raises(RuntimeError):  # more changes than max_batch
        srv.submit(S().Request(0, 0, M().Batch(3, np.zeros(60, np.int32),
                                                 np.ones(60, np.int32),
                                                 np.ones(180, np.int32), z, z)))
    with pytest.raises(RuntimeError):  # unknown graph
        srv.submit(S().Request(1, 5, big.batch))
    srv.close()


def test_malformed_request_is_refused_alone():
    """A request with a vertex outside the graph or a weight below 1 is
    refused at submission, alone: the requests around it complete, the
    server keeps serving, and their results equal the same requests served
    one at a time (which test_mix_equals_alone_equals_oracle ties to the
    oracle)."""
    graphs = small_graphs()[:2]
    srv = make_server(graphs, slots=2, max_batch=400, keep_results=True)
    good = requests(graphs, count=4, seed=11)
    g0 = graphs[0][0]
    z = np.zeros(0, np.int32)
    bad = [M().Batch(3, i32([0]), i32([g0.n]), i32([1, 1, 1]), z, z),        # head outside
           M().Batch(3, i32([1]), i32([2]), i32([1, 0, 1]), z, z),           # weight 0
           M().Batch(3, z, z, z, i32([-1]), i32([2]))]                       # tail outside
    srv.submit(good[0])
    for i, b in enumerate(bad):
        with pytest.raises(RuntimeError, match="invalid request"):
            srv.submit(S().Request(100 + i, 0, b))
    for r in good[1:]:
        srv.submit(r)
    for r in good:
        srv.wait(r)
    srv.close()
    alone = make_server(graphs, slots=1, max_coalesce=1, max_batch=400, keep_results=True)
    again = fresh(good)
    for r in again:
        alone.submit(r)
        alone.wait(r)
    alone.close()
    for a, b in zip(good, again):
        same(a.result, b.result, f"request {a.rid} beside the refused ones")
