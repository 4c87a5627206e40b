"""Pins of the oracle (CPU restatement) before it is trusted as the checker.

Each oracle function is checked against something that does not restate it:
brute force on tiny graphs, the reference's known-answer thesis example,
the reference's own host code (oracle/_ref) on the same inputs, closed
forms, or reference-generated regression fixtures.
"""
from __future__ import annotations

import json
import os
import pathlib

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import INF, Batch, i32

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


# ---------------------------------------------------------------- helpers --

def brute_force_tree(g: O.Graph, k: int, s: int):
    """Bellman-Ford to the fixed point, then parent = lowest-id tight
    in-neighbour (the canonical rule, README.md:142-146)."""
    d = [INF] * g.n
    d[s] = 0
    for _ in range(g.n):
        changed = False
        for u in range(g.n):
            if d[u] >= INF // 2:
                continue
            for e in range(g.row[u], g.row[u + 1]):
                v = int(g.col[e])
                c = d[u] + g.weight(e, k)
                if c < d[v]:
                    d[v] = c
                    changed = True
        if not changed:
            break
    p = [-1] * g.n
    for u in range(g.n):
        if d[u] >= INF // 2:
            continue
        for e in range(g.row[u], g.row[u + 1]):
            v = int(g.col[e])
            if v != s and d[u] + g.weight(e, k) == d[v] and (p[v] < 0 or u < p[v]):
                p[v] = u
    return np.array(d, np.int64), np.array(p, np.int32)


def python_apply(g: O.Graph, b: Batch):
    """List-of-rows restatement of applyChangeBatch (csrGraph.cu:727-894)."""
    rows = [[[int(g.col[e]), [g.weight(e, k) for k in range(g.K)], True]
             for e in range(g.row[u], g.row[u + 1])] for u in range(g.n)]
    for u, v in zip(b.del_from, b.del_to):
        for ent in rows[u]:
            if ent[2] and ent[0] == v:
                ent[2] = False
                break
    w = np.asarray(b.ins_w).reshape(-1, g.K)
    for i, (u, v) in enumerate(zip(b.ins_from, b.ins_to)):
        for ent in rows[u]:
            if ent[2] and ent[0] == v:
                ent[1] = list(w[i])
                break
        else:
            rows[u].append([int(v), list(w[i]), True])
    edges = [(u, ent[0], ent[1]) for u in range(g.n) for ent in rows[u] if ent[2]]
    upd = O.graph_from_edges(g.n, g.K, edges)
    mask = []
    for u, v in zip(b.ins_from, b.ins_to):
        orig = [e for e in range(g.row[u], g.row[u + 1]) if g.col[e] == v]
        fin = [e for e in range(upd.row[u], upd.row[u + 1]) if upd.col[e] == v]
        mk = 0
        if orig and fin:
            for k in range(g.K):
                if upd.weight(fin[0], k) > g.weight(orig[0], k):
                    mk |= 1 << k
        mask.append(mk)
    return upd, np.array(mask, np.uint32)


def thesis():
    t = json.loads((GOLDEN / "thesis_example.json").read_text())
    g = O.graph_from_edges(t["n"], t["K"], [(e[0], e[1], e[2:]) for e in t["edges"]])
    b = Batch(t["K"], i32([x[0] for x in t["insert"]]), i32([x[1] for x in t["insert"]]),
              i32([w for x in t["insert"] for w in x[2:]]),
              i32([x[0] for x in t["delete"]]), i32([x[1] for x in t["delete"]]))
    return t, g, b


GRAPHS = {
    "random-tiny": lambda s: gen.random_graph(12, 18, 3, 50, s),
    "random-small": lambda s: gen.random_graph(60, 180, 2, 50, s),
    "random-medium": lambda s: gen.random_graph(800, 2400, 3, 100, s),
    "grid-small": lambda s: gen.grid_graph(8, 8, 2, 50, 0.1, s),
    "grid-medium": lambda s: gen.grid_graph(40, 30, 3, 100, 0.15, s),
}
CHANGES = {"random-tiny": 6, "random-small": 20, "random-medium": 200,
           "grid-small": 12, "grid-medium": 150}


# ------------------------------------------------------------------ tests --

@pytest.mark.parametrize("seed", range(12))
def test_dijkstra_matches_brute_force(seed):
    g = gen.random_graph(9 + seed % 5, 14, 2, 6 if seed % 2 else 60, seed)
    for k in range(g.K):
        for s in (0, g.n - 1):
            d, p = O.dijkstra(g, k, s)
            bd, bp = brute_force_tree(g, k, s)
            assert np.array_equal(d, bd) and np.array_equal(p, bp)


def test_dijkstra_unreachable_and_ties():
    # 0 -> 1 (w 2), 0 -> 2 (w 1), 2 -> 1 (w 1): tie at 1 goes to parent 0.
    g = O.graph_from_edges(4, 1, [(0, 1, [2]), (0, 2, [1]), (2, 1, [1])])
    d, p = O.dijkstra(g, 0, 0)
    assert d.tolist() == [0, 2, 1, INF] and p.tolist() == [-1, 0, 0, -1]


def test_thesis_known_answer():
    t, g, b = thesis()
    upd, mask = O.apply_batch(g, b)
    rev = O.transpose(upd)
    init_d, init_p = gen.initial_trees(g, 3)
    for k in range(3):
        d, p = O.dijkstra(upd, k, 0)
        assert p.tolist() == t["expected_trees"][k]
        ch = O.changed_lists(b, mask, k)
        sd, sp = O.sequential_update(upd, rev, k, 0, ch[0], ch[1], b.ins_to,
                                     init_d[k], init_p[k])
        assert sp.tolist() == t["expected_trees"][k] and np.array_equal(sd, d)
    parents = np.array(t["expected_trees"], np.int32)
    for exp in t["mosp"]:
        comb = O.combined_graph(parents, 7, 3, 0, exp["pref"])
        cd, cp = O.dijkstra(comb, 0, 0)
        path = [6]
        while cp[path[-1]] >= 0:
            path.append(int(cp[path[-1]]))
        assert path == exp["path_to_u7"]
        costs = O.path_costs(upd, cp, 0)
        assert costs[6].tolist() == exp["cost_of_u7"]


def test_pipeline_thesis():
    t, g, b = thesis()
    d0, p0 = gen.initial_trees(g, 3)
    for exp in t["mosp"]:
        r = O.mosp_update(g, b, d0, p0, 3, 0, exp["pref"])
        assert r["ok"]
        assert r["parents"].tolist() == t["expected_trees"]
        path = [6]
        while r["comb_parent"][path[-1]] >= 0:
            path.append(int(r["comb_parent"][path[-1]]))
        assert path == exp["path_to_u7"]


def test_pipeline_rejects_trees_of_another_source():
    g = gen.grid_graph(4, 4, 2, 9, 0.0, 7)
    d0, p0 = gen.initial_trees(g, 2, 0)
    b = Batch(2, *(np.zeros(0, np.int32) for _ in range(5)))
    assert O.mosp_update(g, b, d0, p0, 2, 0, [])["ok"]
    assert not O.mosp_update(g, b, d0, p0, 2, 1, [])["ok"]


@pytest.mark.parametrize("seed", range(10))
def test_apply_matches_python_restatement(seed):
    g = gen.random_graph(15, 40, 2, 9, seed)
    b = gen.uniform_batch(g, 30, 60, seed + 100, wmax=12)
    # Re-insert some deleted pairs and repeat inserts of one pair.
    b = Batch(2, np.concatenate([b.ins_from, b.del_from[:3], b.ins_from[:2]]).astype(np.int32),
              np.concatenate([b.ins_to, b.del_to[:3], b.ins_to[:2]]).astype(np.int32),
              np.concatenate([b.ins_w, np.arange(1, 11, dtype=np.int32)]),
              b.del_from, b.del_to)
    upd, mask = O.apply_batch(g, b)
    pu, pm = python_apply(g, b)
    assert np.array_equal(upd.row, pu.row) and np.array_equal(upd.col, pu.col)
    assert np.array_equal(upd.w, pu.w) and np.array_equal(mask, pm)


@needs_ref
@pytest.mark.parametrize("seed", range(8))
def test_apply_and_dijkstra_match_reference_code(seed):
    g = gen.random_graph(200, 700, 3, 30, seed)
    b = gen.uniform_batch(g, 120, 50, seed + 7, wmax=40)
    ours, mask = O.apply_batch(g, b)
    theirs, rmask = O.ref_apply_batch(g, b)
    for a, c in ((ours.row, theirs.row), (ours.col, theirs.col), (ours.w, theirs.w),
                 (mask, rmask)):
        assert np.array_equal(a, c)
    for k in range(3):
        d, p = O.dijkstra(ours, k, 0)
        rd, rp = O.ref_dijkstra(ours, k, 0)
        assert np.array_equal(d, rd) and np.array_equal(p, rp)


@pytest.mark.parametrize("gname", list(GRAPHS))
@pytest.mark.parametrize("kind", gen.BATCH_KINDS)
def test_sequential_update_is_canonical_dijkstra(gname, kind):
    for rep in range(2):
        seed = 1000 * rep + len(gname) + 31 * gen.BATCH_KINDS.index(kind)
        g = GRAPHS[gname](seed)
        b = gen.make_batch(kind, g, CHANGES[gname], seed)
        upd, mask = O.apply_batch(g, b)
        rev = O.transpose(upd)
        d0, p0 = gen.initial_trees(g, g.K)
        for k in range(g.K):
            ch = O.changed_lists(b, mask, k)
            d, p = O.sequential_update(upd, rev, k, 0, ch[0], ch[1], b.ins_to, d0[k], p0[k])
            rd, rp = O.dijkstra(upd, k, 0)
            assert np.array_equal(d, rd) and np.array_equal(p, rp), (gname, kind, k)
            assert O.check_tree(rev, k, 0, d, p, rd, rp).tolist() == [0, 0, 0, 0]


def test_regression_fixtures():
    """Count-to-infinity instances of src/mospTest.cu:550-595, generated by the
    reference's own generators (tests/golden/make_regressions.py)."""
    cases = json.loads((GOLDEN / "regressions.json").read_text())["cases"]
    for c in cases:
        n, K = c["n"], c["K"]
        g = O.Graph(n, K, i32(c["row"]), i32(c["col"]), i32(c["w"]))
        b = Batch(K, i32(c["ins_from"]), i32(c["ins_to"]), i32(c["ins_w"]),
                  i32(c["del_from"]), i32(c["del_to"]))
        upd, mask = O.apply_batch(g, b)
        assert upd.row.tolist() == c["updated_row"] and upd.col.tolist() == c["updated_col"]
        assert upd.w.tolist() == c["updated_w"] and mask.tolist() == c["increase_mask"]
        rev = O.transpose(upd)
        for k in range(K):
            d0, p0 = O.dijkstra(g, k, 0)
            assert d0.tolist() == c["init_dist"][k] and p0.tolist() == c["init_parent"][k]
            ch = O.changed_lists(b, mask, k)
            d, p = O.sequential_update(upd, rev, k, 0, ch[0], ch[1], b.ins_to, d0, p0)
            assert d.tolist() == c["expected_dist"][k]
            assert p.tolist() == c["expected_parent"][k]


@needs_ref
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_sequential_update_matches_reference_files(tmp_path, mode):
    """Our restatement vs the reference's sequentialSOSPUpdate (file based) on
    batches from the reference's own generateChangeBatch (Uniform, Targeted,
    Reweight, Increase)."""
    g = gen.grid_graph(12, 10, 2, 50, 0.1, 17 + mode)
    prefix = str(tmp_path / "g" / "graphCsr")
    O.write_csr(prefix, g)
    ip, dp = str(tmp_path / "c" / "insert.txt"), str(tmp_path / "c" / "delete.txt")
    os.makedirs(tmp_path / "c", exist_ok=True)
    assert O.ref().ref_generate_change_batch_files(prefix.encode(), 40, 50.0, mode, 1, 50,
                                                   9 + mode, 0, 0, 0, ip.encode(),
                                                   dp.encode())
    b = O.read_batch(ip, dp, g.K)
    upd, mask = O.apply_batch(g, b)
    rev = O.transpose(upd)
    for k in range(g.K):
        d0, p0 = O.dijkstra(g, k, 0)
        din, tin = str(tmp_path / f"i{k}" / "d.txt"), str(tmp_path / f"i{k}" / "t.txt")
        O.write_tree(din, tin, d0, p0)
        dout, tout = str(tmp_path / f"o{k}" / "d.txt"), str(tmp_path / f"o{k}" / "t.txt")
        assert O.ref().ref_sequential_update_files(prefix.encode(), din.encode(), tin.encode(),
                                                   ip.encode(), dp.encode(), k, 0,
                                                   dout.encode(), tout.encode())
        rd, rp = O.read_tree(dout, tout, g.n)
        ch = O.changed_lists(b, mask, k)
        d, p = O.sequential_update(upd, rev, k, 0, ch[0], ch[1], b.ins_to, d0, p0)
        assert np.array_equal(d, rd) and np.array_equal(p, rp)


def test_combined_graph_closed_form():
    rng = np.random.default_rng(5)
    n, K = 50, 4
    parents = rng.integers(-1, n, size=(K, n)).astype(np.int32)
    parents[:, 0] = -1
    pref = [1, 2, 3, 4]
    L = 12
    comb = O.combined_graph(parents, n, K, 0, pref)
    got = {(u, int(comb.col[e])): int(comb.w[e])
           for u in range(n) for e in range(comb.row[u], comb.row[u + 1])}
    want = {}
    for v in range(1, n):
        for p in set(int(x) for x in parents[:, v] if x >= 0):
            want[(p, v)] = L * (K + 1) - sum(L // pref[j] for j in range(K) if parents[j, v] == p)
    assert got == want
    # Equal trees, Pref all ones: every edge is in all K trees -> weight 1.
    same = np.tile(parents[:1], (K, 1))
    c2 = O.combined_graph(same, n, K, 0, [])
    assert set(c2.w.tolist()) <= {1}


def test_preference_scale():
    assert O.preference_scale([], 3) == 1
    assert O.preference_scale([4, 1, 4], 3) == 4
    assert O.preference_scale([1, 2, 3, 4, 5], 5) == 60
    assert O.preference_scale([1, 2], 3) == 0
    assert O.preference_scale([0, 1, 1], 3) == 0
    assert O.preference_scale([1 << 20, 3], 2) == 0  # L > 2^20
    assert O.combined_graph(np.zeros((2, 3), np.int32), 3, 2, 0, [1, 0]) is None


def test_path_costs_property():
    g = gen.grid_graph(9, 7, 3, 30, 0.1, 4)
    d, p = O.dijkstra(g, 1, 0)
    costs = O.path_costs(g, p, 0)
    for v in range(g.n):
        if v == 0:
            assert costs[v].tolist() == [0, 0, 0]
        elif p[v] < 0:
            assert (costs[v] == INF).all()
        else:
            e = next(e for e in range(g.row[p[v]], g.row[p[v] + 1]) if g.col[e] == v)
            assert costs[v].tolist() == [costs[p[v]][k] + g.weight(e, k) for k in range(3)]
    assert (costs[:, 1] == np.where(d < INF // 2, d, INF)).all()
    bad = p.copy()
    bad[5] = 60  # not an in-neighbour of 5 -> a tree edge is missing
    assert O.path_costs(g, bad, 0) is None or not np.array_equal(O.path_costs(g, bad, 0), costs)


def test_packing_closed_form():
    # packing-boundary case (src/mospTest.cu:723-731): b = 17, bound 2^47-2^15.
    assert O.choose_packing((1 << 17) - 1, (1 << 30) + (1 << 14)) == (17, (1 << 47) - (1 << 15))
    # road grid, n = 1.96M, W = 100: b = 21, packed.
    assert O.choose_packing(1_960_000, 100) == (21, 1_959_999 * 100)
    # R-MAT scale 22 with W = 2^24: fallback (bound 7.0e13 > 2^41 - 2).
    assert O.choose_packing(1 << 22, 1 << 24)[0] == 0
    # (n-1) * W beyond 62 bits is rejected.
    assert O.choose_packing(2**31 - 1, 2**31 - 1) is None
    assert O.choose_packing(1, 5) == (1, 5)


def test_default_delta_closed_form():
    assert O.default_delta(10, 5, 100) == 160        # 32 * 10 / 2
    assert O.default_delta(0, 5, 100) == 1
    assert O.default_delta(1000, 10, 1000) == 1        # 32 * 1 / 100 -> 0 -> 1


def test_check_tree_flags_planted_mistakes():
    g = gen.grid_graph(6, 6, 1, 5, 0.0, 2)
    rev = O.transpose(g)
    d, p = O.dijkstra(g, 0, 0)
    assert O.check_tree(rev, 0, 0, d, p, d, p).tolist() == [0, 0, 0, 0]
    d2 = d.copy()
    d2[7] += 1
    assert O.check_tree(rev, 0, 0, d2, p, d, p)[0] >= 1
    p2 = p.copy()
    v = next(v for v in range(1, g.n) if p[v] != v - 1 and v - 1 >= 0 and (v % 6) != 0)
    p2[v] = (v - 1)
    assert O.check_tree(rev, 0, 0, d, p2, d, p)[3] == 1


def test_canonicalize_tree():
    g = gen.grid_graph(7, 7, 1, 3, 0.0, 8)
    d, p = O.dijkstra(g, 0, 0)
    # Make ties resolve to the HIGHEST id, then canonicalize back.
    hi = p.copy()
    for v in range(1, g.n):
        tight = [u for u in range(g.n) for e in range(g.row[u], g.row[u + 1])
                 if g.col[e] == v and d[u] + g.weight(e, 0) == d[v]]
        hi[v] = max(tight)
    assert O.canonicalize(g, 0, 0, d, hi).tolist() == p.tolist()


def test_alg_bytes_hand_count():
    # 0 -> 1 -> 2, 0 -> 2 (w 5); delete (1,2): I = {2}, C = {2}, D = {2}.
    g = O.graph_from_edges(3, 1, [(0, 1, [1]), (1, 2, [1]), (0, 2, [5])])
    b = Batch(1, i32([]), i32([]), i32([]), i32([1]), i32([2]))
    upd, mask = O.apply_batch(g, b)
    rev = O.transpose(upd)
    d0, p0 = O.dijkstra(g, 0, 0)
    d1, _ = O.dijkstra(upd, 0, 0)
    ch = O.changed_lists(b, mask, 0)
    got = O.alg_bytes(upd, rev, 1, 0, p0, d0, d1, [ch], [], 2)
    # 28*3 + C{2}: 8 + 16*indeg(2)=1 -> 24 + D{2}: 8 + 16*0 -> 8
    # + combined 4*1*3 + 4*4 + 8*2 + 8*3 + 16*2 + 12*3
    assert got == 84 + 24 + 8 + (12 + 16 + 16 + 24 + 32 + 36)
