"""CPU tests of the product's host layer: the C-ABI library loads and exports
every declared symbol, batch application / generators / helpers agree with
the oracle, the CLI rejects malformed inputs like the reference's readers
(src/mospTest.cu:827-919), and the roofline byte count matches the oracle's.
No GPU compute is called here."""
from __future__ import annotations

import pathlib
import re
import subprocess

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import Batch, i32

ROOT = pathlib.Path(__file__).resolve().parent.parent
MOSP_BIN = ROOT / "mosp-cuda_amd" / "bin" / "mosp"
PREP_BIN = ROOT / "mosp-cuda_amd" / "bin" / "mosp_prep"
LEGACY_BIN = ROOT / "mosp-cuda_amd" / "bin" / "mosp_legacy"


def M():
    import mosp_amd
    return mosp_amd


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "mosp_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mosp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = M().lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(M().EXPORTS)


@pytest.mark.parametrize("seed", range(8))
def test_apply_batch_matches_oracle(seed):
    g = gen.random_graph(120, 400, 3, 9, seed)
    b = gen.uniform_batch(g, 150, 60, seed + 50, wmax=12)
    b = Batch(3, np.concatenate([b.ins_from, b.del_from[:5]]).astype(np.int32),
              np.concatenate([b.ins_to, b.del_to[:5]]).astype(np.int32),
              np.concatenate([b.ins_w, np.arange(1, 16, dtype=np.int32)]), b.del_from, b.del_to)
    mg = M().HostGraph(g.n, 3, g.row, g.col, g.w)
    mb = M().Batch(3, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to)
    u, mask = M().apply_batch(mg, mb)
    ou, om = O.apply_batch(g, b)
    assert np.array_equal(u.row, ou.row) and np.array_equal(u.col, ou.col)
    assert np.array_equal(u.w, ou.w) and np.array_equal(mask, om)


def test_road_grid_generator_shapes_and_connectivity():
    g1 = M().gen_road_grid(30, 30, 3, seed=1, drop=0.0, diag=0.0)
    assert (g1.n, g1.m) == (900, 3480)
    g = M().gen_road_grid(120, 90, 3, seed=5)
    again = M().gen_road_grid(120, 90, 3, seed=5)
    assert np.array_equal(g.col, again.col) and np.array_equal(g.w, again.w)
    og = O.Graph(g.n, 3, g.row, g.col, g.w)
    assert gen.reachable(og, 0).all()
    grid_edges = 120 * 89 + 90 * 119
    undirected = grid_edges - round(0.3 * grid_edges) + round(0.02 * g.n)
    assert g.m == 2 * undirected
    w = g.w.reshape(-1, 3)
    assert w.min() >= 1 and w.max() <= 100
    # symmetric: (u,v) present iff (v,u) present
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    fwd = set(zip(src.tolist(), g.col.tolist()))
    assert all((v, u) in fwd for u, v in fwd)


def test_rmat_generator_has_no_self_loops_or_duplicates():
    g = M().gen_rmat(10, 8, 2, seed=3)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    assert (src != g.col).all()
    pairs = src.astype(np.int64) * g.n + g.col
    assert len(np.unique(pairs)) == g.m
    assert 0 < g.m <= 8 << 10


def test_batch_generator_safe_keeps_reachability():
    g = M().gen_road_grid(60, 60, 3, seed=11)
    b = M().gen_batch(g, 800, 50.0, seed=12, safe=True)
    assert b.n_ins == 400 and b.n_del == 400
    assert (b.ins_from != b.ins_to).all()
    og = O.Graph(g.n, 3, g.row, g.col, g.w)
    ob = Batch(3, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to)
    upd, _ = O.apply_batch(og, ob)
    assert gen.reachable(upd, 0).all()
    # deletions are distinct existing arcs
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    arcs = set(zip(src.tolist(), g.col.tolist()))
    dels = list(zip(b.del_from.tolist(), b.del_to.tolist()))
    assert len(set(dels)) == len(dels) and all(d in arcs for d in dels)
    unsafe = M().gen_batch(g, 800, 50.0, seed=12, safe=False)
    assert unsafe.n_del == 400


def test_host_helpers_match_oracle():
    for pref, K in (([], 3), ([4, 1, 4], 3), ([1, 2, 3, 4, 5], 5), ([0, 1], 2), ([1, 2], 3),
                    ([1 << 20, 3], 2)):
        assert M().preference_scale(pref, K) == O.preference_scale(pref, K)
    for n, w in ((131071, (1 << 30) + (1 << 14)), (1_960_000, 100), (1 << 22, 1 << 24),
                 (2**31 - 1, 2**31 - 1), (1, 5), (900, 100)):
        assert M().choose_packing(n, w) == O.choose_packing(n, w)
    for m, n, s in ((10, 5, 100), (0, 5, 1), (5_562_480, 1_960_000, 280_000_000)):
        assert M().default_delta(m, n, s) == O.default_delta(m, n, s)


@pytest.mark.parametrize("seed", range(4))
def test_roofline_bytes_match_oracle(seed):
    from mosp_amd.roofline import alg_bytes
    g = gen.grid_graph(25, 20, 2, 30, 0.1, seed)
    b = gen.uniform_batch(g, 60, 50, seed + 3, wmax=30)
    upd, mask = O.apply_batch(g, b)
    rev = O.transpose(upd)
    d0, p0 = gen.initial_trees(g, 2)
    d1, _ = gen.initial_trees(upd, 2)
    changed = [O.changed_lists(b, mask, k) for k in range(2)]
    want = O.alg_bytes(upd, rev, 2, 0, p0, d0, d1, changed, b.ins_to, 123)
    got = alg_bytes(g.n, upd.row, rev.row, 0, p0, d0, d1, changed, b.ins_to, 123)
    assert got == want
    # The parent-recovery term of distance-only objectives: 12 n + 16 m each.
    fb = alg_bytes(g.n, upd.row, rev.row, 0, p0, d0, d1, changed, b.ins_to, 123,
                   distance_only=[True, False])
    assert fb - got == 12 * upd.n + 16 * int(upd.row[-1])


@pytest.mark.skipif(not MOSP_BIN.exists(), reason="bin/mosp not built")
@pytest.mark.parametrize("values,rows", [
    ("0 5\n2 7\n", "0\n1\n2\n2\n"), ("4 5\n-3 7\n", "0\n1\n2\n2\n"),
    ("4294967297 5\n2 7\n", "0\n1\n2\n2\n"), ("4 5\n2147483648 7\n", "0\n1\n2\n2\n"),
    ("4 5\n2 7\n", "0\n1\n4294967298\n2\n"), ("4 5\n2 7 1\n", "0\n1\n2\n2\n"),
])
def test_cli_rejects_malformed_graphs(tmp_path, values, rows):
    g = tmp_path / "g"
    g.mkdir()
    (g / "gRowPtr.txt").write_text(rows)
    (g / "gColInd.txt").write_text("1\n2\n")
    (g / "gValues.txt").write_text(values)
    (tmp_path / "c").mkdir()
    (tmp_path / "c" / "insert.txt").write_text("")
    (tmp_path / "c" / "delete.txt").write_text("")
    r = subprocess.run([str(MOSP_BIN), "--graph", str(g / "g"), "--changes",
                        str(tmp_path / "c"), "--init", str(tmp_path / "i")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Error" in r.stdout


@pytest.mark.skipif(not MOSP_BIN.exists(), reason="bin/mosp not built")
@pytest.mark.parametrize("insert,dist,tree", [
    ("2 0 0 3\n", "0 0\n1 4\n2 INF\n", "0 -1\n1 0\n2 -1\n"),          # weight 0
    ("2 0 -5 3\n", "0 0\n1 4\n2 INF\n", "0 -1\n1 0\n2 -1\n"),
    ("2 0 4294967297 3\n", "0 0\n1 4\n2 INF\n", "0 -1\n1 0\n2 -1\n"),
    ("2 3 1 1\n", "0 0\n1 4\n2 INF\n", "0 -1\n1 0\n2 -1\n"),          # endpoint 3
    ("2 0 1 1\n", "0 0\n1 4\n", "0 -1\n1 0\n2 -1\n"),                # incomplete
    ("2 0 1 1\n", "0 0\n1 4\n1 4\n2 6\n", "0 -1\n1 0\n2 -1\n"),      # duplicate
    ("2 0 1 1\n", "0 0\n1 -4\n2 6\n", "0 -1\n1 0\n2 -1\n"),          # negative
    ("2 0 1 1\n", "0 0\n1 4\n2 INF\n", "0 -1\n1 0\n"),               # tree incomplete
    ("2 0 1 1\n", "0 0\n1 4\n2 INF\n", "0 -1\n1 0\n1 0\n2 1\n"),     # tree duplicate
])
def test_cli_rejects_malformed_batches_and_trees(tmp_path, insert, dist, tree):
    g = tmp_path / "g"
    g.mkdir()
    (g / "gRowPtr.txt").write_text("0\n1\n2\n2\n")
    (g / "gColInd.txt").write_text("1\n2\n")
    (g / "gValues.txt").write_text("4 5\n2147483647 7\n")
    c = tmp_path / "c"
    c.mkdir()
    (c / "insert.txt").write_text(insert)
    (c / "delete.txt").write_text("0 1\n")
    for k in range(2):
        o = tmp_path / "i" / f"obj{k}"
        o.mkdir(parents=True)
        (o / "distancesOriginal.txt").write_text(dist)
        (o / "SSSPTreeOriginal.txt").write_text(tree)
    r = subprocess.run([str(MOSP_BIN), "--graph", str(g / "g"), "--changes", str(c), "--init",
                        str(tmp_path / "i"), "--no-output"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Error" in r.stdout


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_reference_readers_agree_on_the_same_literals(tmp_path):
    """The same malformed literals are rejected by the reference's own readers
    (oracle/_ref), so the CLI tests above pin the reference's behaviour."""
    g = tmp_path / "g"
    g.mkdir()
    (g / "gRowPtr.txt").write_text("0\n1\n2\n2\n")
    (g / "gColInd.txt").write_text("1\n2\n")
    (g / "gValues.txt").write_text("4 5\n2147483648 7\n")
    assert O.ref().ref_read_csr_ok(str(g / "g").encode()) == 0
    (g / "gValues.txt").write_text("4 5\n2147483647 7\n")
    assert O.ref().ref_read_csr_ok(str(g / "g").encode()) == 1
    p = tmp_path / "d.txt"
    p.write_text("0 0\n1 4\n1 4\n2 6\n")
    assert O.ref().ref_read_distances_ok(str(p).encode(), 3) == 0


def _weighted_sum(a) -> int:
    a = np.asarray(a, dtype=np.int64).astype(np.uint64)
    pos = np.arange(1, a.size + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int((a * pos).sum(dtype=np.uint64))


@pytest.mark.skipif(not PREP_BIN.exists(), reason="bin/mosp_prep not built")
def test_large_csr_text_is_read_in_parallel_parts(tmp_path):
    """Text files above 4 MiB are cut at line boundaries and parsed by one
    thread per part: the arrays equal a sequential read (order-sensitive
    checksums), and an error deep in a file reports its line, as one pass."""
    rng = np.random.default_rng(3)
    n, K = 200_000, 3
    deg = rng.integers(0, 8, n)
    row = np.concatenate([[0], np.cumsum(deg)])
    m = int(row[-1])
    col = rng.integers(0, n, m)
    w = rng.integers(1, 2**31 - 1, (m, K))
    pre = tmp_path / "g"
    rows_txt = "\n".join(map(str, row)) + "\n"
    cols_txt = "\n".join(map(str, col)) + "\n"
    vals = [" ".join(map(str, r)) for r in w]
    (tmp_path / "gRowPtr.txt").write_text(rows_txt)
    (tmp_path / "gColInd.txt").write_text(cols_txt)
    (tmp_path / "gValues.txt").write_text("\n".join(vals) + "\n")
    assert (tmp_path / "gValues.txt").stat().st_size > 3 * (4 << 20)
    r = subprocess.run([str(PREP_BIN), "check", str(pre)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    want = (f"n={n} m={m} K={K} rows={_weighted_sum(row)} cols={_weighted_sum(col)} "
            f"weights={_weighted_sum(w.reshape(-1))}")
    assert r.stdout.strip() == want
    # errors deep in the files: the first one, with its line number
    bad = list(vals)
    bad[m - 10] = "1 2"                      # too few objectives on line m - 9
    bad[m - 5] = "1 x 3"                     # a later invalid token
    (tmp_path / "gValues.txt").write_text("\n".join(bad) + "\n")
    r = subprocess.run([str(PREP_BIN), "check", str(pre)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert f"Inconsistent number of objectives at {pre}Values.txt:{m - 9}" in r.stdout
    bad = list(vals)
    bad[m // 2] = "1 2 0"                    # weight 0, in a middle part
    (tmp_path / "gValues.txt").write_text("\n".join(bad) + "\n")
    r = subprocess.run([str(PREP_BIN), "check", str(pre)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and f"{pre}Values.txt:{m // 2 + 1}" in r.stdout
    (tmp_path / "gValues.txt").write_text("\n".join(vals) + "\n")
    cols = cols_txt.split("\n")
    cols[m - 3] = "-4"
    (tmp_path / "gColInd.txt").write_text("\n".join(cols))
    r = subprocess.run([str(PREP_BIN), "check", str(pre)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and f"{pre}ColInd.txt:{m - 2}" in r.stdout


@pytest.mark.skipif(not LEGACY_BIN.exists(), reason="bin/mosp_legacy not built")
def test_file_wrappers_reject_bad_arguments_before_the_gpu(tmp_path):
    """parallelSOSPUpdate / parallelCombinedGraph argument checks (the
    reference's messages), all before any GPU work."""
    g = tmp_path / "g"
    g.mkdir()
    (g / "gRowPtr.txt").write_text("0\n1\n2\n2\n")
    (g / "gColInd.txt").write_text("1\n2\n")
    (g / "gValues.txt").write_text("4 5\n3 7\n")
    pre = str(g / "g")

    def run(*args):
        return subprocess.run([str(LEGACY_BIN), *map(str, args)], capture_output=True,
                              text=True, timeout=60)
    r = run("sosp", pre, "d", "t", "i", "x", 2, 0, tmp_path / "o1", tmp_path / "o2")
    assert r.returncode == 1 and "objectiveIndex out of range" in r.stdout
    r = run("sosp", pre, "d", "t", "i", "x", 0, 7, tmp_path / "o1", tmp_path / "o2")
    assert r.returncode == 1 and "source vertex out of range" in r.stdout
    r = run("sosp", str(tmp_path / "nope"), "d", "t", "i", "x", 0, 0, "a", "b")
    assert r.returncode == 1 and "Could not read original CSR graph" in r.stdout
    r = run("combined", pre, 2, 0, "a", "b", "1,0", "t0", "t1")
    assert r.returncode == 1 and "invalid preference vector" in r.stdout
    r = run("combined", pre, 2, 5, "a", "b", "-", "t0", "t1")
    assert r.returncode == 1 and "out of range" in r.stdout
    r = run("combined", pre, 2, 0, "a", "b", "-", "t0")
    assert r.returncode == 2
