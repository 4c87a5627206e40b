"""Parity of the HIP path (libmosp_hip.so via its C ABI) with the oracle.

Bit-exact for every integer output: K updated trees (distances + lowest-id
parents), the combined graph's distances (units of 1/L) and MOSP tree, and
the MOSP path costs.  Cases follow the reference's oracle suite
(src/mospTest.cu): thesis known answers, count-to-infinity regressions,
graph families x batch kinds, large-weight fallback with ties, the packing
boundary, determinism, plus the benchmark workloads.
"""
from __future__ import annotations

import json
import pathlib
import sys

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import INF, Batch, i32

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def M():
    import mosp_amd
    return mosp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mgraph(g: O.Graph):
    return M().HostGraph(g.n, g.K, i32(g.row), i32(g.col), i32(g.w))


def mbatch(b: Batch):
    return M().Batch(b.K, i32(b.ins_from), i32(b.ins_to), i32(b.ins_w), i32(b.del_from),
                     i32(b.del_to))


def check_pipeline(g: O.Graph, b: Batch, pref=(), K=0, fused=True, tag=""):
    Ku = K or g.K
    d0, p0 = gen.initial_trees(g, Ku)
    say(f"case {tag} n={g.n} m={g.m} K={Ku} ins={b.n_ins} del={b.n_del} fused={fused}")
    got = M().update_host(mgraph(g), mbatch(b), d0, p0, K=Ku, pref=list(pref), fused=fused)
    want = O.mosp_update(g, b, d0, p0, Ku, 0, list(pref))
    assert want["ok"]
    upd = got["updated"]
    assert np.array_equal(upd.row, want["updated"].row) and np.array_equal(upd.col,
                                                                           want["updated"].col)
    for k in range(Ku):
        assert np.array_equal(got["dist"][k], want["dist"][k]), f"{tag} obj{k} distances"
        assert np.array_equal(got["parents"][k], want["parents"][k]), f"{tag} obj{k} parents"
    assert np.array_equal(got["comb_dist"], want["comb_dist"]), f"{tag} combined distances"
    assert np.array_equal(got["comb_parent"], want["comb_parent"]), f"{tag} MOSP tree"
    assert got["combined_edges"] == want["combined_edges"]
    return got, want


# ----------------------------------------------------------- known answers --

def test_thesis_known_answer():
    t = json.loads((GOLDEN / "thesis_example.json").read_text())
    g = O.graph_from_edges(t["n"], t["K"], [(e[0], e[1], e[2:]) for e in t["edges"]])
    b = Batch(3, i32([x[0] for x in t["insert"]]), i32([x[1] for x in t["insert"]]),
              i32([w for x in t["insert"] for w in x[2:]]), i32([x[0] for x in t["delete"]]),
              i32([x[1] for x in t["delete"]]))
    d0, p0 = gen.initial_trees(g, 3)
    for fused in (True, False):
        for exp in t["mosp"]:
            r = M().update_host(mgraph(g), mbatch(b), d0, p0, pref=exp["pref"], fused=fused)
            assert r["parents"].tolist() == t["expected_trees"]
            path = [6]
            while r["comb_parent"][path[-1]] >= 0:
                path.append(int(r["comb_parent"][path[-1]]))
            assert path == exp["path_to_u7"]
            costs = M().path_costs(r["updated"], r["comb_parent"], 0)
            assert costs[6].tolist() == exp["cost_of_u7"]


def test_regression_fixtures_single_objective():
    cases = json.loads((GOLDEN / "regressions.json").read_text())["cases"]
    for c in cases:
        n, K = c["n"], c["K"]
        upd = M().HostGraph(n, K, i32(c["updated_row"]), i32(c["updated_col"]),
                            i32(c["updated_w"]))
        b = Batch(K, i32(c["ins_from"]), i32(c["ins_to"]), i32(c["ins_w"]), i32(c["del_from"]),
                  i32(c["del_to"]))
        mask = np.array(c["increase_mask"], np.uint32)
        dg = M().DeviceGraph(upd)
        ws = M().Workspace(n)
        for k in range(K):
            say(f"regression {c['case']} obj{k}")
            cf, ct = O.changed_lists(b, mask, k)
            maxw = max(int(np.max(np.asarray(c["w"]).reshape(-1, K)[:, k])),
                       int(np.max(b.ins_w.reshape(-1, K)[:, k])) if b.n_ins else 1)
            d, p, st = M().sosp_update(dg, k, cf, ct, b.ins_to, 0, 7, maxw, ws,
                                       np.array(c["init_dist"][k]), np.array(c["init_parent"][k]))
            assert d.tolist() == c["expected_dist"][k] and p.tolist() == c["expected_parent"][k]


# ------------------------------------------------- graph families x batches --

GRAPHS = {
    "random-tiny": lambda s: gen.random_graph(12, 18, 3, 50, s),
    "random-small": lambda s: gen.random_graph(60, 180, 2, 50, s),
    "random-medium": lambda s: gen.random_graph(800, 2400, 3, 100, s),
    "grid-small": lambda s: gen.grid_graph(8, 8, 2, 50, 0.1, s),
    "grid-medium": lambda s: gen.grid_graph(40, 30, 3, 100, 0.15, s),
}
CHANGES = {"random-tiny": 6, "random-small": 20, "random-medium": 200, "grid-small": 12,
           "grid-medium": 150}


@pytest.mark.parametrize("gname", list(GRAPHS))
@pytest.mark.parametrize("kind", gen.BATCH_KINDS)
def test_families_and_batch_kinds(gname, kind):
    for rep in range(2):
        seed = 1000 * rep + len(gname) + 31 * gen.BATCH_KINDS.index(kind)
        g = GRAPHS[gname](seed)
        b = gen.make_batch(kind, g, CHANGES[gname], seed)
        pref = [] if rep == 0 else [1] + [2] * (g.K - 1)
        check_pipeline(g, b, pref, fused=(rep == 0), tag=f"{gname}/{kind}/{rep}")


STEP1_MODES = [
    # (label, environment): Step 1's candidates for the invalidated vertices
    # pulled over their in-edges, or pushed from the valid vertices (epoch
    # 0's pile then read from the far marks), under both near-far loops.
    ("pull", {"MOSP_INIT_PUSH": "1"}),
    ("push", {"MOSP_INIT_PUSH": "2"}),
    ("push-synchronous", {"MOSP_INIT_PUSH": "2", "MOSP_ASYNC": "0"}),
    ("push-capped-lists", {"MOSP_INIT_PUSH": "2", "MOSP_LIST_CAP": "16"}),
]


@pytest.mark.parametrize("label,env", STEP1_MODES, ids=[m[0] for m in STEP1_MODES])
@pytest.mark.parametrize("kind", gen.BATCH_KINDS)
def test_step1_push_and_pull(monkeypatch, label, env, kind):
    """Every batch kind on grids (no hubs) with each way of seeding the
    invalidated vertices, fused and serial schedules, against the oracle."""
    for key, value in env.items():
        monkeypatch.setenv(key, value)
    for rep in range(2):
        seed = 500 + 7 * gen.BATCH_KINDS.index(kind) + rep
        g = gen.grid_graph(40, 30, 3, 100, 0.15, seed)
        b = gen.make_batch(kind, g, 150, seed)
        pref = [] if rep == 0 else [1, 2, 2]
        check_pipeline(g, b, pref, fused=(rep == 0), tag=f"{label}/{kind}/{rep}")


def test_fused_equals_serial_and_is_deterministic():
    g = gen.grid_graph(60, 50, 3, 100, 0.2, 5)
    b = gen.uniform_batch(g, 400, 50, 6, wmax=100)
    d0, p0 = gen.initial_trees(g, 3)
    runs = [M().update_host(mgraph(g), mbatch(b), d0, p0, fused=f) for f in (True, True, False)]
    for r in runs[1:]:
        for key in ("dist", "parents", "comb_dist", "comb_parent"):
            assert np.array_equal(r[key], runs[0][key]), key


def test_objective_subset_and_skewed_pref():
    g = gen.random_graph(300, 1200, 4, 60, 11)
    b = gen.uniform_batch(g, 80, 50, 12, wmax=60)
    check_pipeline(g, b, pref=[3, 1], K=2, tag="k2-of-4")
    check_pipeline(g, b, pref=[1, 2, 3, 4], tag="pref-1234")


# ------------------------------------------------------ engine entry points --

@pytest.mark.parametrize("seed", range(6))
def test_from_scratch_matches_dijkstra(seed):
    g = gen.random_graph(500, 1500 + 100 * seed, 2, 1 if seed == 0 else 40, seed)
    # a few vertices with no in-edges -> unreachable
    dg = M().DeviceGraph(mgraph(g))
    ws = M().Workspace(g.n)
    for k in range(g.K):
        for src in (0, 7):
            say(f"from-scratch seed={seed} k={k} src={src}")
            w = g.w.reshape(-1, g.K)[:, k]
            d, p, st = M().from_scratch(dg, k, src, M().default_delta(g.m, g.n, int(w.sum())),
                                        int(w.max()), ws)
            rd, rp = O.dijkstra(g, k, src)
            assert np.array_equal(d, rd) and np.array_equal(p, rp)


def test_combined_standalone_and_path_costs():
    g = gen.grid_graph(30, 25, 4, 50, 0.1, 3)
    d0, p0 = gen.initial_trees(g, 4)
    ws = M().Workspace(g.n)
    for pref in ([], [1, 2, 3, 4], [5, 5, 1, 5]):
        say(f"combined pref={pref}")
        d, p, st = M().combined_sosp(p0, 0, pref, ws)
        comb = O.combined_graph(p0, g.n, 4, 0, pref)
        rd, rp = O.dijkstra(comb, 0, 0)
        assert np.array_equal(d, rd) and np.array_equal(p, rp)
        assert st["n_edges"] == comb.m
        costs = M().path_costs(mgraph(g), p, 0)
        want = O.path_costs(g, p, 0)
        assert np.array_equal(costs, want)


def test_grid_sizes_do_not_change_results():
    """The persistent kernel with 1, 3 and the default number of workgroups."""
    g = gen.grid_graph(40, 40, 1, 80, 0.2, 9)
    b = gen.uniform_batch(g, 120, 50, 10, wmax=80)
    upd, mask = O.apply_batch(g, b)
    d0, p0 = O.dijkstra(g, 0, 0)
    rd, rp = O.dijkstra(upd, 0, 0)
    cf, ct = O.changed_lists(b, mask, 0)
    dg = M().DeviceGraph(mgraph(upd))
    for grid in (1, 3, 0):
        ws = M().Workspace(upd.n, 1, grid)
        say(f"grid={grid} -> {ws.grid} workgroups")
        d, p, _ = M().sosp_update(dg, 0, cf, ct, b.ins_to, 0, 50, 80, ws, d0, p0)
        assert np.array_equal(d, rd) and np.array_equal(p, rp)


def test_packing_boundary():
    """n = 2^17 - 1 path with W = 2^30 + 2^14 (src/mospTest.cu:723-815): the
    packed format is used at its limit and over-bound candidates must drop."""
    n = (1 << 17) - 1
    W = (1 << 30) + (1 << 14)

    def path(back):
        edges = [(u, u + 1, [W]) for u in range(n - 1)]
        if back:
            edges.append((n - 1, 1, [W]))
        return O.graph_from_edges(n, 1, edges)

    for name, back, ins in (("pull", False, (n - 1, 1, W)), ("push", True, (n - 2, n - 1, W - 1)),
                            ("from-scratch", True, None)):
        say(f"packing-boundary/{name}")
        orig = path(back)
        b = Batch(1, i32([ins[0]] if ins else []), i32([ins[1]] if ins else []),
                  i32([ins[2]] if ins else []), i32([]), i32([]))
        upd, mask = O.apply_batch(orig, b)
        d0, p0 = O.dijkstra(orig, 0, 0)
        rd, rp = O.dijkstra(upd, 0, 0)
        dg = M().DeviceGraph(mgraph(upd))
        ws = M().Workspace(n)
        delta = M().default_delta(upd.m, n, W * n)
        if ins:
            d, p, st = M().sosp_update(dg, 0, [], [], b.ins_to, 0, delta, W, ws, d0, p0)
        else:
            d, p, st = M().from_scratch(dg, 0, 0, delta, W, ws)
        assert st["packed_parents"] == 1
        assert np.array_equal(d, rd) and np.array_equal(p, rp)


def test_large_weights_fallback_with_ties(monkeypatch):
    """(n-1) * maxW does not fit next to the parent ids
    (src/mospTest.cu:597-700).  The launch first packs words that hold
    distances up to their own limit (here far above every distance: packed
    parents); with MOSP_PACK_LIMIT=1 it meets larger ones, writes nothing
    and is redone with distance-only words and the lowest-id parent
    recovery.  Both equal the oracle, ties included."""
    for W, tag in ((2**31 - 1, "int-max"), (2_000_000_000, "ties")):
        wmin = W if tag == "ties" else 1
        g = gen.grid_graph(320, 320, 2, W, 0.1, 13, wmin=wmin)
        b = gen.uniform_batch(g, 2000, 50, 15, wmax=W, wmin=wmin)
        for limit, packed in ((None, 1), ("1", 0)):
            if limit is None:
                monkeypatch.delenv("MOSP_PACK_LIMIT", raising=False)
            else:
                monkeypatch.setenv("MOSP_PACK_LIMIT", limit)
            for fused in (True, False):
                got, _ = check_pipeline(g, b, fused=fused,
                                        tag=f"large-weights/{tag} limit={limit} fused={fused}")
                assert all(s["packed_parents"] == packed for s in got["stats"])


def test_optimistic_packing_redo_mid_search(monkeypatch):
    """The optimistic packing meets a larger distance only during the search
    (every input distance fits): the launch stops before writing and is
    redone with distance-only words.  From-scratch searches (limit: half the
    longest distance) and updates (limit: the old trees' longest distance,
    below the updated graph's), both kernel variants; results equal the
    oracle's."""
    g = gen.grid_graph(60, 50, 2, 1000, 0.1, 21)
    b = gen.make_batch("deletions-only", g, 300, 22)  # distances only grow
    upd, _ = O.apply_batch(g, b)
    d0, _ = gen.initial_trees(g, 2)
    old_longest = int(d0[d0 < INF // 2].max())
    final = [O.dijkstra(upd, k, 0) for k in range(2)]
    new_longest = max(int(d[d < INF // 2].max()) for d, _ in final)
    assert new_longest > old_longest
    dg = M().DeviceGraph(mgraph(upd))
    ws = M().Workspace(upd.n)
    for hubs in ("0", "1"):
        monkeypatch.setenv("MOSP_HUBS", hubs)
        for k in range(2):
            rd, rp = final[k]
            w = upd.w.reshape(-1, 2)[:, k]
            delta = M().default_delta(upd.m, upd.n, int(w.sum()))
            monkeypatch.setenv("MOSP_PACK_LIMIT", str(int(rd[rd < INF // 2].max()) // 2))
            d, p, st = M().from_scratch(dg, k, 0, delta, 1000, ws)
            say(f"redo from-scratch hubs={hubs} k={k} packed={st['packed_parents']}")
            assert st["packed_parents"] == 0
            assert np.array_equal(d, rd) and np.array_equal(p, rp)
        monkeypatch.setenv("MOSP_PACK_LIMIT", str(old_longest))
        for fused in (True, False):
            got, _ = check_pipeline(g, b, pref=[1, 1], fused=fused,
                                    tag=f"redo update hubs={hubs} fused={fused}")
            assert all(s["packed_parents"] == 0 for s in got["stats"])
    monkeypatch.delenv("MOSP_PACK_LIMIT")
    ws.close()
    dg.close()


def test_edge_cases():
    # empty batch; insert into the source; duplicate arcs; isolated vertices
    g = O.graph_from_edges(6, 2, [(0, 1, [3, 1]), (0, 1, [1, 5]), (1, 2, [1, 1]),
                                  (2, 0, [1, 1]), (3, 4, [2, 2])])
    empty = Batch(2, i32([]), i32([]), i32([]), i32([]), i32([]))
    check_pipeline(g, empty, tag="empty")
    b = Batch(2, i32([2, 4]), i32([0, 5]), i32([1, 1, 9, 9]), i32([0]), i32([1]))
    check_pipeline(g, b, tag="insert-into-source")


def test_invalid_input_trees_are_reported():
    """A parent outside [-1, n) or a negative distance in the old trees is an
    error (the kernel reads such a parent as 'none', never as an index); a
    valid update afterwards is unaffected."""
    g = gen.grid_graph(12, 12, 2, 50, 0.1, 5)
    b = gen.make_batch("uniform-unsafe", g, 40, 6)
    d0, p0 = gen.initial_trees(g, 2)
    for k, v, bad in ((0, 7, g.n + 7), (1, 9, -3), (1, 11, 2**31 - 1)):
        p = p0.copy()
        p[k, v] = bad
        say(f"case bad parent obj{k} v={v} value={bad}")
        with pytest.raises(RuntimeError, match="not a vertex id"):
            M().update_host(mgraph(g), mbatch(b), d0, p, K=2)
    d = d0.copy()
    d[0, 5] = -5
    with pytest.raises(RuntimeError, match="input distance"):
        M().update_host(mgraph(g), mbatch(b), d, p0, K=2)
    check_pipeline(g, b, tag="after invalid inputs")


# -------------------------------------------------- power-law (R-MAT) graphs --

def rmat(scale, edge_factor, K, seed, wmax=100):
    mg = M().gen_rmat(scale, edge_factor, K, seed=seed, wmax=wmax)
    return O.Graph(mg.n, K, i32(mg.row), i32(mg.col), i32(mg.w))


def max_degree(g: O.Graph) -> int:
    return int(max(np.diff(g.row).max(), np.bincount(g.col, minlength=g.n).max()))


@pytest.mark.parametrize("kind", gen.BATCH_KINDS)
def test_rmat_hub_vertices(kind):
    """Power-law graphs: vertices with hundreds to thousands of edges are
    relaxed, pulled and scanned for parents by whole waves in the kernel
    variant the engine picks for graphs with vertices of more than 32 edges;
    results equal the oracle's for every batch kind."""
    i = gen.BATCH_KINDS.index(kind)
    g = rmat(12, 16, 3, seed=41 + i)
    assert max_degree(g) > 32
    b = gen.make_batch(kind, g, 800, 500 + i)
    check_pipeline(g, b, pref=[1, 2, 3], tag=f"rmat/{kind}")


def test_kernel_variants_agree(monkeypatch):
    """Both kernel variants (forced with MOSP_HUBS) give identical outputs on
    a hub graph and on a grid."""
    for tag, g in (("rmat", rmat(11, 16, 2, seed=77)), ("grid", gen.grid_graph(50, 40, 2, 80,
                                                                               0.2, 78))):
        b = gen.uniform_batch(g, 600, 50, 79)
        d0, p0 = gen.initial_trees(g, 2)
        runs = []
        for forced in ("0", "1"):
            monkeypatch.setenv("MOSP_HUBS", forced)
            say(f"variants {tag} MOSP_HUBS={forced}")
            runs.append(M().update_host(mgraph(g), mbatch(b), d0, p0, pref=[1, 2], fused=True))
        for key in ("dist", "parents", "comb_dist", "comb_parent"):
            assert np.array_equal(runs[0][key], runs[1][key]), f"{tag}: {key}"


def test_hub_tiers(monkeypatch):
    """Hub variant, push tiers: vertices of more out-edges than the hub
    threshold (MOSP_HUB_EDGES, lowered here so that a small graph has many)
    are relaxed in the next iteration by every workgroup of their objective
    over one range of all listed hubs' edges; a full hub list (512 entries)
    hands them to the workgroup's list, a full one of those to the wave's.
    With 2 workgroups for 3 objectives, teams of waves share each hub's
    edges instead.  Every threshold and grid gives the oracle's results."""
    g = rmat(12, 16, 3, seed=45)
    assert int(np.diff(g.row).max()) > 512
    monkeypatch.setenv("MOSP_HUBS", "1")
    for kind in ("uniform-unsafe", "disconnecting"):
        b = gen.make_batch(kind, g, 700, 46)
        for edges, grid in (("64", None), ("300", None), ("1000", None), ("64", "2"),
                            ("200", "3"), ("1000000", None)):
            monkeypatch.setenv("MOSP_HUB_EDGES", edges)
            if grid is None:
                monkeypatch.delenv("MOSP_GRID", raising=False)
            else:
                monkeypatch.setenv("MOSP_GRID", grid)
            check_pipeline(g, b, pref=[1, 2, 3], tag=f"hub tiers {kind} edges={edges} grid={grid}")


def test_workgroup_shapes_and_queue_caps(monkeypatch):
    """The persistent kernels' two workgroup shapes (MOSP_WG_THREADS: 1024
    threads, one workgroup per CU; 512, two per CU; 0, by kernel variant)
    and smaller LDS push queues (MOSP_QUEUE_CAP) give the oracle's results:
    fused and serial updates on a grid and an R-MAT graph."""
    cases = (("grid", gen.grid_graph(60, 50, 3, 100, 0.2, 61)), ("rmat", rmat(12, 16, 3, seed=62)))
    for tag, g in cases:
        b = gen.uniform_batch(g, 500, 50, 63)
        for threads, cap in ((512, 0), (512, 256), (1024, 1024), (0, 2048)):
            monkeypatch.setenv("MOSP_WG_THREADS", str(threads))
            monkeypatch.setenv("MOSP_QUEUE_CAP", str(cap))
            for fused in (True, False):
                check_pipeline(g, b, pref=[1, 2, 3], fused=fused,
                               tag=f"shape {tag} threads={threads} queue={cap} fused={fused}")


def test_full_lists_fall_back_to_sweeps(monkeypatch):
    """Near / far lists capped far below n (MOSP_LIST_CAP): appends past the
    cap are lost, and the iteration after a full list relaxes every vertex
    (near frontier) or splits every vertex with a far mark (far pile).
    Results stay exact: from-scratch searches (whose local hops re-list most
    of a small graph every iteration), and updates with the combined graph,
    in both kernel variants."""
    monkeypatch.setenv("MOSP_LIST_CAP", "16")
    g = gen.random_graph(500, 1800, 2, 40, 2)
    dg = M().DeviceGraph(mgraph(g))
    ws = M().Workspace(g.n)
    for k in range(g.K):
        for src in (0, 7):
            say(f"capped lists: from-scratch k={k} src={src}")
            w = g.w.reshape(-1, g.K)[:, k]
            d, p, st = M().from_scratch(dg, k, src, M().default_delta(g.m, g.n, int(w.sum())),
                                        int(w.max()), ws)
            rd, rp = O.dijkstra(g, k, src)
            assert np.array_equal(d, rd) and np.array_equal(p, rp)
    for forced in ("0", "1"):
        monkeypatch.setenv("MOSP_HUBS", forced)
        grid = gen.grid_graph(40, 30, 3, 90, 0.2, 81)
        check_pipeline(grid, gen.uniform_batch(grid, 400, 50, 82), pref=[1, 2, 1],
                       tag=f"capped-lists grid hubs={forced}")
        hub = rmat(10, 16, 2, seed=83)
        check_pipeline(hub, gen.make_batch("uniform-unsafe", hub, 300, 84), pref=[2, 1],
                       tag=f"capped-lists rmat hubs={forced}")


ASYNC_KNOBS = [
    # (label, environment, expected events) -- every path of the asynchronous
    # near-far loop forced on a small graph, and the iteration-synchronous
    # loop it replaces.
    ("default", {}, None),
    ("iteration-synchronous", {"MOSP_ASYNC": "0"}, None),
    ("carried-words", {"MOSP_ASYNC_FLAGS": "1"}, None),
    ("carried-words-hungry", {"MOSP_ASYNC_FLAGS": "3", "MOSP_ASYNC_KEEP": "16"}, None),
    ("heavy-spills", {"MOSP_ASYNC_SPILL": "64,32", "MOSP_ASYNC_FLAGS": "0"}, None),
    # Two workgroups keep long local queues and spill at least 32 entries at
    # a time into 32-entry pool segments: every second spill to a segment in
    # a phase finds it full.
    ("tiny-pools", {"MOSP_ASYNC_SPILL": "64,32", "MOSP_POOL_CAP": "32", "MOSP_GRID": "2"},
     "short_spills"),
    # 256-entry local queue kept nearly full, one lane per vertex (batches of
    # up to 1024 entries): appends past the room are dropped, and the next
    # phase lists every vertex under the threshold again.
    ("tiny-local-queue", {"MOSP_QUEUE_CAP": "256", "MOSP_LANE_BITS": "0",
                          "MOSP_ASYNC_SPILL": "250,200"}, "sweep_phases"),
    ("two-workgroups", {"MOSP_GRID": "2"}, None),
    ("one-workgroup", {"MOSP_GRID": "1"}, None),
    ("512-thread", {"MOSP_WG_THREADS": "512", "MOSP_GRID": "7"}, None),
    # The combined graph built as a CSR (count, scan, fill) instead of a mask
    # over the graph's out-edges.
    ("combined-csr", {"MOSP_COMB_MASK": "0"}, None),
]


@pytest.mark.parametrize("label,env,event", ASYNC_KNOBS, ids=[k[0] for k in ASYNC_KNOBS])
def test_async_near_far_paths(monkeypatch, label, env, event):
    """The asynchronous near-far loop (graphs without hubs, single-update
    kernel) under every spill, pool, queue and grid setting, against the
    oracle: K updated trees, the combined graph's distances and MOSP tree.
    Settings that must take a rare path (short spills, dropped near entries
    and the sweep phase after them) assert that they did."""
    for key in ("MOSP_ASYNC", "MOSP_ASYNC_FLAGS", "MOSP_ASYNC_KEEP", "MOSP_ASYNC_SPILL",
                "MOSP_POOL_CAP", "MOSP_QUEUE_CAP", "MOSP_LANE_BITS", "MOSP_GRID",
                "MOSP_WG_THREADS", "MOSP_COMB_MASK"):
        monkeypatch.delenv(key, raising=False)
    for key, value in env.items():
        monkeypatch.setenv(key, value)
    g = gen.grid_graph(70, 60, 3, 90, 0.2, 301)
    b = gen.uniform_batch(g, 900, 50, 302)
    d0, p0 = gen.initial_trees(g, 3)
    want = O.mosp_update(g, b, d0, p0, 3, 0, [1, 3, 2])
    assert want["ok"]
    say(f"async paths: {label} n={g.n} m={g.m} env={env}")
    sess = M().Session(mgraph(g), mbatch(b), d0, p0, pref=[1, 3, 2])
    try:
        for rep in range(2):
            sess.run()
            got = sess.results()
            for k in range(3):
                assert np.array_equal(got["dist"][k], want["dist"][k]), f"{label} obj{k} dist"
                assert np.array_equal(got["parents"][k], want["parents"][k]), f"{label} obj{k}"
            assert np.array_equal(got["comb_dist"], want["comb_dist"]), f"{label} combined"
            assert np.array_equal(got["comb_parent"], want["comb_parent"]), f"{label} MOSP tree"
            ev = sess.ws.async_events()
            assert ev["async"] == (env.get("MOSP_ASYNC") != "0"), ev
            if event is not None:
                assert ev[event] > 0, f"{label}: expected {event} ({ev})"
    finally:
        sess.close()
    # From-scratch search of the same engine (single objective, source list).
    dg = M().DeviceGraph(mgraph(g))
    ws = M().Workspace(g.n)
    try:
        w = g.w.reshape(-1, g.K)[:, 1]
        d, p, _ = M().from_scratch(dg, 1, 5, M().default_delta(g.m, g.n, int(w.sum())),
                                   int(w.max()), ws)
        rd, rp = O.dijkstra(g, 1, 5)
        assert np.array_equal(d, rd) and np.array_equal(p, rp), f"{label} from scratch"
    finally:
        ws.close()
        dg.close()


@pytest.mark.parametrize("mask", ["1", "0"])
def test_repeated_edges(monkeypatch, mask):
    """Rows that hold the same (u, v) edge more than once, with other
    weights: the trees take the lightest copy, and the combined graph has one
    edge per distinct parent (the first copy in the parent's row when the
    combined graph is a mask over the graph's edges; mask "0" builds its
    CSR).  Fused pipeline against the oracle, edge count included."""
    monkeypatch.setenv("MOSP_COMB_MASK", mask)
    base = gen.grid_graph(45, 40, 3, 60, 0.15, 411)
    rng = np.random.default_rng(412)
    edges = []
    for u in range(base.n):
        for e in range(int(base.row[u]), int(base.row[u + 1])):
            w = [base.weight(e, k) for k in range(3)]
            edges.append((u, int(base.col[e]), w))
            if rng.random() < 0.3:  # a copy, lighter or heavier per objective
                edges.append((u, int(base.col[e]),
                              [max(1, x + int(rng.integers(-20, 21))) for x in w]))
    g = O.graph_from_edges(base.n, 3, edges)
    assert g.m > base.m
    for kind, seed in (("safe", 413), ("mixed", 414)):
        b = gen.uniform_batch(g, 400, 50, seed, safe=(kind == "safe"))
        check_pipeline(g, b, pref=[2, 1, 3], tag=f"repeated edges {kind} mask={mask}")


def test_rmat_large_weights_fallback():
    """R-MAT scale 17, weights up to 2^31 - 1: (n - 1) * maxW does not fit next
    to an 18-bit parent id; the words packed up to their own limit (46 bits)
    hold every distance, over hub vertices (the power-law stress workload in
    miniature).  The distance-only words and parent recovery on hub graphs
    are covered by test_full_lists_fall_back_to_sweeps and the redo tests."""
    g = rmat(17, 16, 2, seed=91, wmax=2**31 - 1)
    assert max_degree(g) > 32
    b = gen.uniform_batch(g, 4000, 50, 92, wmax=2**31 - 1)
    got, _ = check_pipeline(g, b, pref=[1, 2], tag="rmat-large-weights")
    assert all(s["packed_parents"] == 1 for s in got["stats"])  # optimistic packing held


# ------------------------------------------------------ benchmark workloads --

def test_small_grid_30x30_against_cpu_path():
    """30x30 grid, K=3, 100 changes (50% inserts), Pref (1,1,1)."""
    mg = M().gen_road_grid(30, 30, 3, seed=1, drop=0.0, diag=0.0)
    assert mg.n == 900 and mg.m == 3480
    mb = M().gen_batch(mg, 100, 50.0, seed=2, safe=True)
    assert mb.n_ins == 50 and mb.n_del == 50
    g = O.Graph(mg.n, 3, mg.row, mg.col, mg.w)
    b = Batch(3, mb.ins_from, mb.ins_to, mb.ins_w, mb.del_from, mb.del_to)
    check_pipeline(g, b, pref=[1, 1, 1], tag="grid30x30")
    check_pipeline(g, b, pref=[1, 1, 1], fused=False, tag="grid30x30-serial")


def test_road_grid_1400x1400_full_size():
    """1400x1400 road-like grid (1.96M vertices), K=3, 50k safe changes:
    every tree equals canonical Dijkstra on the updated graph, the MOSP tree
    equals Dijkstra on the host combined graph, and the GPU initial trees
    equal Dijkstra on the original graph."""
    mg = M().gen_road_grid(1400, 1400, 3, seed=7)
    mb = M().gen_batch(mg, 50_000, 50.0, seed=8, safe=True)
    g = O.Graph(mg.n, 3, mg.row, mg.col, mg.w)
    say(f"road1400 n={g.n} m={g.m} ins={mb.n_ins} del={mb.n_del}")
    d0, p0 = M().initial_trees(mg, 3)
    for k in range(3):
        rd, rp = O.dijkstra(g, k, 0)
        assert np.array_equal(d0[k], rd) and np.array_equal(p0[k], rp), f"initial tree {k}"
    r = M().update_host(mg, mb, d0, p0, fused=True)
    upd = O.Graph(mg.n, 3, r["updated"].row, r["updated"].col, r["updated"].w)
    for k in range(3):
        rd, rp = O.dijkstra(upd, k, 0)
        assert np.array_equal(r["dist"][k], rd) and np.array_equal(r["parents"][k], rp)
        assert (rd < INF // 2).all(), "safe batch keeps every vertex reachable"
    comb = O.combined_graph(r["parents"], g.n, 3, 0, [])
    cd, cp = O.dijkstra(comb, 0, 0)
    assert np.array_equal(r["comb_dist"], cd) and np.array_equal(r["comb_parent"], cp)
    say(f"road1400 gpu_compute_ms={r['timings_ms']['gpu_compute']:.3f} stats={r['stats'][0]}")


def test_rmat_scale22_k5_full_size():
    """R-MAT scale 22 (4.19M vertices, ~60M arcs after removing self-loops
    and duplicates), K=5 weights U[1, 2^24] -- the distance-only words and
    the lowest-id parent recovery on every objective -- 100,000 changes with
    deletions drawn from all arcs (some disconnect vertices), Pref (1..5):
    every tree equals canonical Dijkstra on the updated graph, the MOSP tree
    equals Dijkstra on the host combined graph, and the GPU initial tree of
    objective 0 equals Dijkstra on the original graph."""
    K, pref = 5, [1, 2, 3, 4, 5]
    mg = M().gen_rmat(22, 16, K, seed=31, wmax=2**24)
    mb = M().gen_batch(mg, 100_000, 50.0, seed=32, safe=False, wmax=2**24)
    g = O.Graph(mg.n, K, mg.row, mg.col, mg.w)
    say(f"rmat22 n={g.n} m={g.m} ins={mb.n_ins} del={mb.n_del}")
    d0, p0 = M().initial_trees(mg, K)
    rd, rp = O.dijkstra(g, 0, 0)
    assert np.array_equal(d0[0], rd) and np.array_equal(p0[0], rp), "initial tree 0"
    r = M().update_host(mg, mb, d0, p0, pref=pref, fused=True)
    assert all(s["packed_parents"] == 1 for s in r["stats"])  # optimistic packing held
    del g, mg
    upd = O.Graph(r["updated"].n, K, r["updated"].row, r["updated"].col, r["updated"].w)
    unreachable = 0
    for k in range(K):
        say(f"rmat22 objective {k}")
        rd, rp = O.dijkstra(upd, k, 0)
        assert np.array_equal(r["dist"][k], rd), f"objective {k} distances"
        assert np.array_equal(r["parents"][k], rp), f"objective {k} parents"
        unreachable = max(unreachable, int((rd >= INF // 2).sum()))
    assert unreachable > 0, "the batch disconnects some vertices"
    comb = O.combined_graph(r["parents"], upd.n, K, 0, pref)
    cd, cp = O.dijkstra(comb, 0, 0)
    assert np.array_equal(r["comb_dist"], cd) and np.array_equal(r["comb_parent"], cp)
    say(f"rmat22 gpu_compute_ms={r['timings_ms']['gpu_compute']:.3f} stats={r['stats'][0]}")


def test_session_repeats_out_of_place():
    mg = M().gen_road_grid(200, 150, 3, seed=21)
    mb = M().gen_batch(mg, 3000, 50.0, seed=22, safe=True)
    d0, p0 = M().initial_trees(mg, 3)
    ref = M().update_host(mg, mb, d0, p0)
    s = M().Session(mg, mb, d0, p0)
    for i in range(3):
        s.run()
        r = s.results()
        for key in ("dist", "parents", "comb_dist", "comb_parent"):
            assert np.array_equal(r[key], ref[key]), (i, key)
    s.close()
