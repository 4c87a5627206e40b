"""Change batches applied in HBM (mosp_device_graph_apply) against the host
apply (the library's applyChangeBatch) and the oracle's: identical updated
CSR -- row order and edge-major weights -- and weight-increase masks, for
every batch kind, duplicate arcs, repeated and cancelling changes, and hub
rows longer than a wave.  The reverse CSR built on the device is checked
through an update on the applied graph, which must equal the update on the
host-applied, uploaded graph.
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


def M():
    import mosp_amd
    return mosp_amd


def mgraph(g):
    return M().HostGraph(g.n, g.K, i32(g.row), i32(g.col), i32(g.w))


def mbatch(b):
    return M().Batch(b.K, i32(b.ins_from), i32(b.ins_to), i32(b.ins_w), i32(b.del_from),
                     i32(b.del_to))


def check_apply(g, b, tag, spare=None):
    dg = M().DeviceGraph(mgraph(g))
    out = spare if spare is not None else M().DeviceGraph()
    mask = dg.apply(mbatch(b), out)
    got = out.download()
    want, wmask = O.apply_batch(g, b)
    host, hmask = M().apply_batch(mgraph(g), mbatch(b))
    say(f"apply {tag}: n={g.n} m={g.m}->{want.m} ins={b.n_ins} del={b.n_del}")
    assert np.array_equal(got.row, want.row), f"{tag} row pointers"
    if not np.array_equal(got.col, want.col):
        e = int(np.nonzero(got.col != want.col)[0][0])
        u = int(np.searchsorted(want.row, e, side="right") - 1)
        lo, hi = int(want.row[u]), int(want.row[u + 1])
        say(f"first mismatch at edge {e}, row {u}: got {got.col[lo:hi].tolist()} want "
            f"{want.col[lo:hi].tolist()} old {g.col[g.row[u]:g.row[u + 1]].tolist()} "
            f"dels {[int(t) for f, t in zip(b.del_from, b.del_to) if f == u]} "
            f"ins {[int(t) for f, t in zip(b.ins_from, b.ins_to) if f == u]}")
        bad = np.nonzero(got.col != want.col)[0]
        rows = np.unique(np.searchsorted(want.row, bad, side="right") - 1)
        say(f"{len(bad)} mismatching edges in {len(rows)} rows; first rows {rows[:12].tolist()}")
    assert np.array_equal(got.col, want.col), f"{tag} columns"
    assert np.array_equal(got.w, want.w), f"{tag} weights"
    assert np.array_equal(mask, wmask), f"{tag} increase mask"
    assert np.array_equal(hmask, wmask) and np.array_equal(host.col, want.col)
    dg.close()
    return out


@pytest.mark.parametrize("kind", gen.BATCH_KINDS)
def test_batch_kinds(kind):
    g = gen.random_graph(400, 1600, 3, 40, seed=5)
    b = gen.make_batch(kind, g, 300, seed=6)
    check_apply(g, b, kind)


def test_grid_and_rmat_like_hubs():
    g = gen.grid_graph(30, 30, 2, 50, 0.3, seed=3)
    check_apply(g, gen.make_batch("uniform-unsafe", g, 400, seed=4), "grid")
    # hub rows longer than a wave: vertex 0..3 connect to most vertices
    n, K = 600, 2
    rng = np.random.default_rng(8)
    us, vs = [], []
    for hub in range(4):
        for v in rng.choice(n, 400, replace=False):
            if v != hub:
                us.append(hub)
                vs.append(int(v))
    for u in range(4, n):
        us.append(u)
        vs.append((u + 1) % n)
    ws = rng.integers(1, 30, size=(len(us), K))
    hub_graph = gen._csr(n, K, us, vs, ws)
    b = gen.uniform_batch(hub_graph, 500, 50.0, seed=9)
    # aim deletions and insertions at the hubs, deep in their rows
    b.del_from[: 100] = rng.integers(0, 4, 100)
    b.del_to[: 100] = hub_graph.col[hub_graph.row[b.del_from[:100]] + 300]
    b.ins_from[: 100] = rng.integers(0, 4, 100)
    check_apply(hub_graph, b, "hubs")


def check_reverse(dg, g, tag):
    """The device in-CSR of every objective against the host transpose:
    identical in-row pointers, and each in-row holds the same (source,
    weight) entries (their order inside a row is unspecified)."""
    deg = np.bincount(g.col, minlength=g.n)
    want_row = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    w = np.asarray(g.w).reshape(-1, g.K) if g.m else np.zeros((0, g.K), np.int64)
    for k in range(g.K):
        row, col, wk = dg.download_reverse(k)
        assert np.array_equal(row.astype(np.int64), want_row), f"{tag} obj{k} in-row pointers"
        tgt = np.repeat(np.arange(g.n), np.diff(row))
        got = np.lexsort((wk, col, tgt))
        want = np.lexsort((w[:, k], src, g.col))
        assert np.array_equal(col[got], src[want]), f"{tag} obj{k} in-row sources"
        assert np.array_equal(wk[got], w[want, k]), f"{tag} obj{k} in-row weights"


def test_reverse_csr_matches_host_transpose():
    """The device in-CSR at upload and after an apply, by both methods (one
    atomic per entry when no vertex has more than 32 edges, else the
    bucketed transpose): in-hubs with thousands of in-edges (direct at
    upload, transpose after it), a graph whose rows are nearly all empty,
    and the R-MAT generator's hubs."""
    rng = np.random.default_rng(21)
    n, K = 5000, 2
    us, vs = [], []
    for u in range(n):                      # every vertex points at the two in-hubs
        us += [u, u]
        vs += [0 if u != 0 else 1, 7 if u != 7 else 8]
    for u in rng.choice(n, 300, replace=False):   # a few scattered rows, most rows small
        for v in rng.choice(n, 5, replace=False):
            if v != u and v not in (0, 7):
                us.append(int(u))
                vs.append(int(v))
    ws = rng.integers(1, 1000, size=(len(us), K))
    in_hubs = gen._csr(n, K, us, vs, ws)
    # nearly all rows empty, one row of 40 edges (so the transpose runs)
    su = [19990, 5, 5, 19999] + [12000] * 40
    sv = [3, 19990, 4, 0] + list(range(100, 140))
    sparse_rows = gen._csr(20000, 1, su, sv, rng.integers(1, 9, size=(len(su), 1)))
    mg = M().gen_rmat(14, 16, 3, seed=17, wmax=100)
    rmat = O.Graph(mg.n, 3, i32(mg.row), i32(mg.col), i32(mg.w))
    for tag, g in (("in-hubs", in_hubs), ("empty rows", sparse_rows), ("rmat14", rmat)):
        say(f"case reverse {tag} n={g.n} m={g.m}")
        dg = M().DeviceGraph(mgraph(g))
        check_reverse(dg, g, tag + " upload")
        b = gen.uniform_batch(g, min(400, max(g.m // 2, 2)), 50.0, seed=22)
        spare = M().DeviceGraph()
        dg.apply(mbatch(b), spare)
        updated = spare.download()
        check_reverse(spare, O.Graph(updated.n, updated.K, i32(updated.row), i32(updated.col),
                                     i32(updated.w)), tag + " apply")
        spare.close()
        dg.close()


def test_duplicates_repeats_and_cancelling_changes():
    K = 2
    # duplicate arcs 0->1 (x3) and 2->3 (x2)
    us = [0, 0, 0, 0, 1, 2, 2, 2, 3]
    vs = [1, 2, 1, 1, 2, 3, 3, 0, 0]
    ws = [[5, 5], [6, 6], [7, 7], [8, 8], [1, 1], [2, 2], [3, 3], [4, 4], [9, 9]]
    g = gen._csr(4, K, us, vs, ws)
    b = O.Batch(K,
                ins_from=i32([0, 0, 2, 3, 3, 1]), ins_to=i32([1, 3, 3, 1, 1, 2]),
                ins_w=i32([9, 1, 2, 2, 1, 1, 4, 4, 5, 6, 1, 9]),
                del_from=i32([0, 0, 2, 1, 3]), del_to=i32([1, 1, 3, 2, 2]))
    check_apply(g, b, "duplicates")
    check_apply(g, O.Batch(K, i32([]), i32([]), i32([]), i32([]), i32([])), "empty")


def test_spare_graph_is_reused_and_chains():
    g = gen.random_graph(300, 900, 2, 30, seed=11)
    spare = M().DeviceGraph()
    cur = g
    for step in range(4):
        b = gen.make_batch(gen.BATCH_KINDS[step % len(gen.BATCH_KINDS)], cur, 120, seed=20 + step)
        check_apply(cur, b, f"chain {step}", spare=spare)
        cur, _ = O.apply_batch(cur, b)
    spare.close()


def test_update_on_applied_graph_equals_uploaded():
    """Same trees from the device-applied graph (device reverse CSR) as from
    the host-applied, uploaded graph."""
    g = gen.grid_graph(40, 40, 3, 60, 0.3, seed=21)
    b = gen.make_batch("uniform-safe", g, 500, seed=22)
    d0, p0 = gen.initial_trees(g, 3)
    Mm = M()
    dg = Mm.DeviceGraph(mgraph(g))
    applied = Mm.DeviceGraph()
    mask = dg.apply(mbatch(b), applied)
    s = Mm.Session(mgraph(g), mbatch(b), d0, p0)
    assert np.array_equal(mask, s.mask)
    s.run()
    want = s.results()
    s.dg.close()
    s.dg = applied  # same Session, graph from the device apply
    s.run()
    got = s.results()
    for key in ("dist", "parents", "comb_dist", "comb_parent"):
        assert np.array_equal(got[key], want[key]), key
    s.close()
    dg.close()


def test_invalid_apply_is_rejected():
    g = gen.random_graph(20, 40, 1, 9, seed=1)
    dg = M().DeviceGraph(mgraph(g))
    bad = O.Batch(1, i32([0]), i32([20]), i32([1]), i32([]), i32([]))
    with pytest.raises(RuntimeError):
        dg.apply(mbatch(bad), M().DeviceGraph())
    with pytest.raises(RuntimeError):
        dg.apply(mbatch(O.Batch(1, i32([]), i32([]), i32([]), i32([]), i32([]))), dg)
    dg.close()
