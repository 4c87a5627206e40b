"""CPU tests of the product's host file contract and host algorithms: the
pair-based batch application against the reference's own applyChangeBatch
(oracle/_ref) on multigraphs, the MOSPCSR2 binary cache against the
reference's writer and reader, cache reuse and invalidation
(src/mospTest.cu:924-983), the parallel text writers, the product's
canonicalizeTree and host mospPathCosts against the oracle and the thesis
known answer, and the upload's input checks.  No GPU compute is called."""
from __future__ import annotations

import json
import os
import pathlib
import subprocess
import time

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import Batch, i32

ROOT = pathlib.Path(__file__).resolve().parent.parent
PREP_BIN = ROOT / "mosp-cuda_amd" / "bin" / "mosp_prep"
GOLDEN = ROOT / "tests" / "golden"


def M():
    import mosp_amd
    return mosp_amd


def multigraph_case(seed: int):
    """Few vertices, many parallel arcs, self-loops; a batch that deletes
    more copies than exist, deletes absent arcs, inserts the same pair many
    times (raising and lowering weights) and appends new pairs in a scattered
    file order."""
    rng = np.random.default_rng(seed)
    n, K = 9, 3
    m = 70
    us = rng.integers(0, n, m)
    vs = rng.integers(0, 4, m)  # few distinct targets: many duplicates per row
    ws = rng.integers(1, 9, (m, K))
    g = gen._csr(n, K, us, vs, ws)
    pairs = [(int(rng.integers(0, n)), int(rng.integers(0, 6))) for _ in range(30)]
    dels = [pairs[int(i)] for i in rng.integers(0, len(pairs), 45)]
    ins = [pairs[int(i)] for i in rng.integers(0, len(pairs), 60)]
    b = Batch(K, i32([p[0] for p in ins]), i32([p[1] for p in ins]),
              i32(rng.integers(1, 9, (len(ins), K)).reshape(-1)),
              i32([p[0] for p in dels]), i32([p[1] for p in dels]))
    return g, b


@pytest.mark.parametrize("seed", range(12))
def test_host_apply_equals_oracle_on_multigraphs(seed):
    g, b = multigraph_case(seed)
    u, mask = M().apply_batch(M().HostGraph(g.n, g.K, g.row, g.col, g.w),
                              M().Batch(g.K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to))
    ou, om = O.apply_batch(g, b)
    assert np.array_equal(u.row, ou.row) and np.array_equal(u.col, ou.col)
    assert np.array_equal(u.w, ou.w) and np.array_equal(mask, om)
    assert mask.any(), "the case should raise some weights"


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", range(12))
def test_host_apply_equals_reference_code_on_multigraphs(seed):
    g, b = multigraph_case(seed + 100)
    u, mask = M().apply_batch(M().HostGraph(g.n, g.K, g.row, g.col, g.w),
                              M().Batch(g.K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to))
    ru, rm = O.ref_apply_batch(g, b)
    assert np.array_equal(u.row, ru.row) and np.array_equal(u.col, ru.col)
    assert np.array_equal(u.w, ru.w) and np.array_equal(mask, rm)


def test_host_apply_rejects_out_of_range_endpoints():
    g = gen.random_graph(10, 10, 2, 9, 1)
    b = M().Batch(2, i32([0]), i32([10]), i32([1, 1]), i32([]), i32([]))
    with pytest.raises(RuntimeError):
        M().apply_batch(M().HostGraph(g.n, 2, g.row, g.col, g.w), b)


def write_text_graph(prefix: pathlib.Path, g) -> None:
    (prefix.parent / (prefix.name + "RowPtr.txt")).write_text("".join(f"{x}\n" for x in g.row))
    (prefix.parent / (prefix.name + "ColInd.txt")).write_text("".join(f"{x}\n" for x in g.col))
    w = np.asarray(g.w).reshape(-1, g.K)
    (prefix.parent / (prefix.name + "Values.txt")).write_text(
        "".join(" ".join(map(str, r)) + "\n" for r in w))


def prep(*args, timeout=120):
    return subprocess.run([str(PREP_BIN), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


@pytest.mark.skipif(not (PREP_BIN.exists() and O.ref_available()), reason="tools not built")
def test_binary_cache_is_the_reference_format(tmp_path):
    """Our saveCsrGraphBinary writes the bytes the reference's writes, and each
    side's reader accepts the other's file."""
    g = gen.random_graph(500, 3000, 4, 1000, 7)
    pre = tmp_path / "g"
    write_text_graph(pre, g)
    ours, theirs = tmp_path / "ours.bin", tmp_path / "theirs.bin"
    r = prep("cache", pre, ours)
    assert r.returncode == 0, r.stdout + r.stderr
    assert O.ref().ref_save_cache(str(pre).encode(), str(theirs).encode()) == 1
    assert ours.read_bytes() == theirs.read_bytes()
    assert O.ref().ref_load_cache_ok(str(ours).encode(), str(pre).encode()) == 1
    r = prep("check", pre, theirs)
    assert r.returncode == 0 and "Note" not in r.stdout


@pytest.mark.skipif(not PREP_BIN.exists(), reason="mosp_prep not built")
def test_binary_cache_reuse_and_invalidation(tmp_path):
    """loadCsrGraph: writes the cache when absent, reads it while the text
    files are unchanged, and rebuilds it (with the reference's note) when they
    change or the cache is damaged."""
    g = gen.grid_graph(40, 30, 3, 50, 0.1, 3)
    pre = tmp_path / "g"
    write_text_graph(pre, g)
    cache = tmp_path / "c" / "graph.bin"
    first = prep("check", pre, cache)
    assert first.returncode == 0 and cache.exists() and "Note" not in first.stdout
    plain = prep("check", pre)
    assert first.stdout.strip() == plain.stdout.strip()
    stamp = cache.stat().st_mtime_ns
    again = prep("check", pre, cache)
    assert again.returncode == 0 and "Note" not in again.stdout
    assert again.stdout.strip() == plain.stdout.strip() and cache.stat().st_mtime_ns == stamp
    # The text changes (same content, new modification time): rebuilt.
    time.sleep(0.01)
    vals = tmp_path / "gValues.txt"
    vals.write_text(vals.read_text())
    os.utime(vals, ns=(stamp + 10**9, stamp + 10**9))
    changed = prep("check", pre, cache)
    assert changed.returncode == 0 and "Note: binary cache" in changed.stdout
    assert changed.stdout.splitlines()[-1] == plain.stdout.strip()
    assert prep("check", pre, cache).stdout.strip() == plain.stdout.strip()
    # A damaged cache (truncated arrays, then a bad magic) is rejected and rebuilt.
    data = cache.read_bytes()
    cache.write_bytes(data[:-8])
    r = prep("check", pre, cache)
    assert "Note: binary cache" in r.stdout and cache.read_bytes() == data
    cache.write_bytes(b"X" + data[1:])
    r = prep("check", pre, cache)
    assert "Note: binary cache" in r.stdout and cache.read_bytes() == data
    # A cache of another graph is not taken for this one.
    other = tmp_path / "h"
    write_text_graph(other, gen.grid_graph(20, 20, 3, 50, 0.1, 4))
    r = prep("check", other, cache)
    assert "Note: binary cache" in r.stdout


@pytest.mark.skipif(not PREP_BIN.exists(), reason="mosp_prep not built")
def test_text_writers_format_and_parallel_blocks(tmp_path):
    """writeCsrGraph (mosp_prep road) writes exactly the generator's arrays in
    the reference's text layout; 400 x 400 vertices take the multi-threaded
    block path (more than 2^17 lines)."""
    pre = tmp_path / "g"
    r = prep("road", pre, 400, 400, 3, 9)
    assert r.returncode == 0, r.stdout + r.stderr
    g = M().gen_road_grid(400, 400, 3, seed=9)
    assert (tmp_path / "gRowPtr.txt").read_text() == "".join(f"{x}\n" for x in g.row)
    assert (tmp_path / "gColInd.txt").read_text() == "".join(f"{x}\n" for x in g.col)
    w = g.w.reshape(-1, 3)
    assert (tmp_path / "gValues.txt").read_text() == "".join(f"{a} {b} {c}\n" for a, b, c in w)


@pytest.mark.skipif(not PREP_BIN.exists(), reason="mosp_prep not built")
def test_mtx2csr_symmetric_pattern(tmp_path):
    """Matrix Market input: symmetric entries give both directions, self-loops
    and duplicates are dropped, rows are sorted, weights lie in [wmin, wmax]."""
    mtx = tmp_path / "g.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% c\n5 5 6\n"
                   "1 2\n2 3\n3 3\n2 1\n5 4\n1 2\n")
    r = prep("mtx2csr", mtx, tmp_path / "g", 2, 3, 7, 1)
    assert r.returncode == 0, r.stdout + r.stderr
    row = [int(x) for x in (tmp_path / "gRowPtr.txt").read_text().split()]
    col = [int(x) for x in (tmp_path / "gColInd.txt").read_text().split()]
    vals = [[int(x) for x in ln.split()] for ln in (tmp_path / "gValues.txt").read_text().splitlines()]
    assert row == [0, 1, 3, 4, 5, 6]
    assert col == [1, 0, 2, 1, 4, 3]
    assert len(vals) == 6 and all(len(v) == 2 and all(3 <= x <= 7 for x in v) for v in vals)


@pytest.mark.parametrize("seed", range(6))
def test_product_canonicalize_tree_equals_oracle(seed):
    """canonicalizeTree (the product's parallel min pass) on tie-heavy graphs
    whose trees carry the highest-id tight parents: the result is the oracle's
    canonical tree, and Dijkstra's."""
    g = gen.grid_graph(30, 25, 2, 2, 0.05, seed)  # weights in [1, 2]: many ties
    for k in range(2):
        d, p = O.dijkstra(g, k, 0)
        high = p.copy()
        src = np.repeat(np.arange(g.n), np.diff(g.row))
        w = g.w.reshape(-1, 2)[:, k]
        tight = (d[src] < O.INF // 2) & (d[src] + w == d[g.col]) & (g.col != 0)
        for u, v in zip(src[tight], g.col[tight]):
            high[v] = max(high[v], u)
        assert (high != p).any()
        mg = M().HostGraph(g.n, 2, g.row, g.col, g.w)
        got = M().canonicalize_tree(mg, k, 0, d, high)
        assert np.array_equal(got, O.canonicalize(g, k, 0, d, high))
        assert np.array_equal(got, p)


def test_product_path_costs_known_answer_and_oracle():
    """Host mospPathCosts on the thesis example (src/mospTest.cu:454-542):
    the MOSP paths to u7 cost (15,3,20) and (15,24,7); on random trees it
    equals the oracle, including vertices cut off from the source and a tree
    edge missing from the graph."""
    t = json.loads((GOLDEN / "thesis_example.json").read_text())
    g = O.graph_from_edges(t["n"], t["K"], [(u, v, w) for u, v, *w in t["edges"]])
    b = Batch(3, i32([e[0] for e in t["insert"]]), i32([e[1] for e in t["insert"]]),
              i32([x for e in t["insert"] for x in e[2:]]), i32([e[0] for e in t["delete"]]),
              i32([e[1] for e in t["delete"]]))
    upd, _ = O.apply_batch(g, b)
    mg = M().HostGraph(upd.n, 3, upd.row, upd.col, upd.w)
    for case in t["mosp"]:
        path = case["path_to_u7"]
        parent = np.full(upd.n, -1, np.int32)
        for child, par in zip(path[:-1], path[1:]):
            parent[child] = par
        costs = M().path_costs_host(mg, parent, 0).reshape(upd.n, 3)
        assert costs[6].tolist() == case["cost_of_u7"]
    rng = np.random.default_rng(5)
    for seed in range(5):
        rg = gen.random_graph(200, 600, 3, 30, seed)
        d, p = O.dijkstra(rg, 1, 0)
        p = p.copy()
        cut = rng.choice(np.arange(1, 200), 10, replace=False)
        p[cut[:5]] = -1                 # subtrees cut off from the source
        p[cut[5]], p[cut[6]] = cut[6], cut[5]  # a two-vertex cycle
        want = O.path_costs(rg, p, 0)
        got = M().path_costs_host(M().HostGraph(rg.n, 3, rg.row, rg.col, rg.w), p, 0)
        assert got is not None and np.array_equal(got.reshape(rg.n, 3), want)
    # A reachable vertex whose tree edge is not in the graph.
    p = O.dijkstra(rg, 0, 0)[1].copy()
    v = 50
    while p[v] < 0:
        v += 1
    missing = next(u for u in range(rg.n) if u != v and v not in
                   set(rg.col[rg.row[u]:rg.row[u + 1]].tolist()) and p[u] >= 0 and u != p[v]
                   and not _descends(p, u, v))
    p[v] = missing
    assert O.path_costs(rg, p, 0) is None
    assert M().path_costs_host(M().HostGraph(rg.n, 3, rg.row, rg.col, rg.w), p, 0) is None


def _descends(parent, u: int, v: int) -> bool:
    """u lies in v's subtree (giving v parent u would close a cycle)."""
    x, hops = u, 0
    while x >= 0 and hops <= len(parent):
        if x == v:
            return True
        x, hops = parent[x], hops + 1
    return False


@pytest.mark.parametrize("bad", ["row_start", "row_decrease", "column", "weight"])
def test_device_graph_upload_rejects_bad_arrays_before_the_gpu(bad):
    """mosp_device_graph_upload checks the caller's arrays on the host before
    any copy: a bad column would otherwise be an out-of-bounds device access
    in the reverse build.  The call returns NULL with a message."""
    g = gen.random_graph(50, 100, 2, 9, 3)
    row, col, w = g.row.copy(), g.col.copy(), g.w.copy()
    if bad == "row_start":
        row[0] = 1
    elif bad == "row_decrease":
        row[10], row[11] = row[11] + 1, row[10]
    elif bad == "column":
        col[7] = 50
    else:
        w[3] = 0
    L = M().lib()
    h = L.mosp_device_graph_upload(g.n, 2, row, col, w)
    assert not h
    assert "invalid graph" in M().last_error()
