"""End-to-end GPU test of the CLI tools, after the reference's
scripts/endToEndTest.sh:31-238: inputs prepared by mosp_prep (mtx2csr,
init, changes), `mosp --validate` on connectivity-safe, disconnecting and
tree-targeted batches, outputs byte-equal to `mosp_prep expected` and equal
to the oracle's pipeline, the mospCosts.txt layout and values, --canonicalize
of highest-id tie trees, -k 2 --pref 2,1 with the binary cache written then
read (and the cache of `mosp_prep cache`), the --timing CSV (stages and
counters), --serial, and rejection of trees of another source."""
from __future__ import annotations

import pathlib
import re
import subprocess

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import Batch, i32

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
BIN = ROOT / "mosp-cuda_amd" / "bin"
RESULT = re.compile(r"^RESULT gpu_compute_ms=\d+\.\d{3} end_to_end_ms=\d+\.\d{3}$", re.M)


def run(*args, ok=True, timeout=120):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=timeout)
    if ok:
        assert r.returncode == 0, f"{args}\n{r.stdout}\n{r.stderr}"
    return r


def mosp(*args, validate=True, ok=True):
    r = run(BIN / "mosp", *args, *(["--validate"] if validate else []), ok=ok)
    if ok:
        assert RESULT.search(r.stdout), r.stdout
        if validate:
            assert re.search(r"^VALIDATE combined PASS", r.stdout, re.M), r.stdout
            assert not re.search(r"^VALIDATE .* FAIL", r.stdout, re.M), r.stdout
    return r


def write_mtx(path: pathlib.Path, seed: int) -> int:
    """A 30 x 30 grid, both directions, 5% of the links missing, plus 300
    random arcs (the script's awk generator, with numpy's stream)."""
    rng = np.random.default_rng(seed)
    W = H = 30
    arcs = []
    for y in range(H):
        for x in range(W):
            u = y * W + x + 1
            if x + 1 < W and rng.random() >= 0.05:
                arcs += [(u, u + 1), (u + 1, u)]
            if y + 1 < H and rng.random() >= 0.05:
                arcs += [(u, u + W), (u + W, u)]
    for _ in range(300):
        a, b = (int(v) + 1 for v in rng.integers(0, W * H, 2))
        if a != b:
            arcs.append((a, b))
    path.write_text("%%MatrixMarket matrix coordinate pattern general\n"
                    f"{W * H} {W * H} {len(arcs)}\n" + "".join(f"{a} {b}\n" for a, b in arcs))
    return W * H


def read_graph(prefix: pathlib.Path) -> O.Graph:
    d = prefix.parent
    row = i32([int(x) for x in (d / (prefix.name + "RowPtr.txt")).read_text().split()])
    col = i32([int(x) for x in (d / (prefix.name + "ColInd.txt")).read_text().split()])
    lines = (d / (prefix.name + "Values.txt")).read_text().splitlines()
    K = len(lines[0].split())
    w = i32([int(x) for ln in lines for x in ln.split()])
    return O.Graph(len(row) - 1, K, row, col, w)


def read_pairs(path: pathlib.Path, dtype):
    out = {}
    for ln in path.read_text().splitlines():
        v, x = ln.split()
        out[int(v)] = O.INF if x == "INF" else int(x)
    return np.array([out[v] for v in range(len(out))], dtype=dtype)


def write_tree(d: pathlib.Path, k: int, dist, parent):
    o = d / f"obj{k}"
    o.mkdir(parents=True, exist_ok=True)
    (o / "distancesOriginal.txt").write_text(
        "".join(f"{v} {'INF' if x >= O.INF // 2 else x}\n" for v, x in enumerate(dist)))
    (o / "SSSPTreeOriginal.txt").write_text("".join(f"{v} {p}\n" for v, p in enumerate(parent)))


def write_batch(d: pathlib.Path, b: Batch):
    d.mkdir(parents=True, exist_ok=True)
    w = np.asarray(b.ins_w).reshape(-1, b.K) if b.n_ins else np.zeros((0, b.K), int)
    (d / "insert.txt").write_text("".join(
        f"{u} {v} " + " ".join(map(str, ws)) + "\n" for u, v, ws in zip(b.ins_from, b.ins_to, w)))
    (d / "delete.txt").write_text("".join(f"{u} {v}\n" for u, v in zip(b.del_from, b.del_to)))


def read_batch(d: pathlib.Path, K: int) -> Batch:
    ins = [list(map(int, ln.split())) for ln in (d / "insert.txt").read_text().splitlines() if ln]
    dels = [list(map(int, ln.split())) for ln in (d / "delete.txt").read_text().splitlines() if ln]
    return Batch(K, i32([x[0] for x in ins]), i32([x[1] for x in ins]),
                 i32([w for x in ins for w in x[2:2 + K]]), i32([x[0] for x in dels]),
                 i32([x[1] for x in dels]))


def check_against_oracle(g: O.Graph, b: Batch, init: pathlib.Path, out: pathlib.Path, K: int,
                         pref, source: int = 0):
    d0 = np.stack([read_pairs(init / f"obj{k}" / "distancesOriginal.txt", np.int64)
                   for k in range(K)])
    p0 = np.stack([read_pairs(init / f"obj{k}" / "SSSPTreeOriginal.txt", np.int32)
                   for k in range(K)])
    want = O.mosp_update(g, b, d0, p0, K, source, pref)
    for k in range(K):
        dist = read_pairs(out / f"obj{k}" / "distancesUpdated.txt", np.int64)
        par = read_pairs(out / f"obj{k}" / "SSSPTreeUpdated.txt", np.int32)
        assert np.array_equal(np.minimum(dist, O.INF), np.minimum(want["dist"][k], O.INF))
        assert np.array_equal(par, want["parents"][k])
    cd = read_pairs(out / "combinedGraph" / "distancesCsr.txt", np.int64)
    cp = read_pairs(out / "combinedGraph" / "SSSPTreeCsr.txt", np.int32)
    assert np.array_equal(np.minimum(cd, O.INF), np.minimum(want["comb_dist"], O.INF))
    assert np.array_equal(cp, want["comb_parent"])
    # mospCosts.txt: "v c1 .. cKG" (KG = all objectives of the graph), the
    # first-occurrence edge costs along the MOSP tree.
    lines = (out / "combinedGraph" / "mospCosts.txt").read_text().splitlines()
    assert len(lines) == g.n and all(len(ln.split()) == 1 + g.K for ln in lines)
    costs = O.path_costs(want["updated"], want["comb_parent"], source)
    got = np.array([[O.INF if x == "INF" else int(x) for x in ln.split()[1:]] for ln in lines])
    assert np.array_equal(np.minimum(got, O.INF), np.minimum(costs, O.INF))


@pytest.fixture(scope="module")
def work(tmp_path_factory):
    t = tmp_path_factory.mktemp("e2e")
    mtx = t / "graph.mtx"
    write_mtx(mtx, 1)
    pre = t / "g" / "csr" / "graphCsr"
    run(BIN / "mosp_prep", "mtx2csr", mtx, pre, 3, 1, 20, 1)
    run(BIN / "mosp_prep", "init", pre, t / "g" / "init")
    run(BIN / "mosp_prep", "changes", pre, t / "g" / "safe", 400, 50, 1, 1)
    run(BIN / "mosp_prep", "changes", pre, t / "g" / "unsafe", 1500, 20, 0, 1)
    g = read_graph(pre)
    # Targeted: deletions of tree edges of objective 0 (whole subtrees).
    p0 = read_pairs(t / "g" / "init" / "obj0" / "SSSPTreeOriginal.txt", np.int32)
    rng = np.random.default_rng(2)
    vs = [int(v) for v in rng.choice(np.nonzero(p0 >= 0)[0], 200, replace=False)]
    z = np.zeros(0, np.int32)
    write_batch(t / "g" / "targeted", Batch(3, z, z, z, i32([p0[v] for v in vs]), i32(vs)))
    return t, pre, g


@pytest.mark.parametrize("batch", ["safe", "unsafe", "targeted"])
def test_validate_and_expected(work, batch):
    t, pre, g = work
    out, exp = t / f"out-{batch}", t / f"exp-{batch}"
    mosp("--graph", pre, "--changes", t / "g" / batch, "--init", t / "g" / "init", "--quiet",
         "--out", out)
    run(BIN / "mosp_prep", "expected", pre, t / "g" / batch, exp)
    for k in range(3):
        for f in ("distancesUpdated.txt", "SSSPTreeUpdated.txt"):
            assert (out / f"obj{k}" / f).read_bytes() == (exp / f"obj{k}" / f).read_bytes(), \
                f"{batch} obj{k}/{f} differs from mosp_prep expected"
    check_against_oracle(g, read_batch(t / "g" / batch, 3), t / "g" / "init", out, 3, [])


def test_expected_equals_oracle_dijkstra(work):
    t, pre, g = work
    b = read_batch(t / "g" / "unsafe", 3)
    upd, _ = O.apply_batch(g, b)
    exp = t / "exp-unsafe"
    if not exp.exists():
        run(BIN / "mosp_prep", "expected", pre, t / "g" / "unsafe", exp)
    for k in range(3):
        d, p = O.dijkstra(upd, k, 0)
        assert np.array_equal(read_pairs(exp / f"obj{k}" / "SSSPTreeUpdated.txt", np.int32), p)
        dist = read_pairs(exp / f"obj{k}" / "distancesUpdated.txt", np.int64)
        assert np.array_equal(np.minimum(dist, O.INF), np.minimum(d, O.INF))
        assert (d >= O.INF // 2).any(), "the unsafe batch should disconnect some vertices"


def test_canonicalize_highest_id_ties(work):
    """Initial trees from another tool that break ties the other way: with
    --canonicalize the update equals `mosp_prep expected` and the oracle."""
    t, _, _ = work
    pre = t / "ties" / "csr" / "graphCsr"
    run(BIN / "mosp_prep", "mtx2csr", t / "graph.mtx", pre, 2, 1, 2, 1)
    run(BIN / "mosp_prep", "changes", pre, t / "ties" / "changes", 400, 50, 1, 1, 1, 2)
    g = read_graph(pre)
    init = t / "ties" / "init-high"
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    changed = False
    for k in range(2):
        d, p = O.dijkstra(g, k, 0)
        high = np.full(g.n, -1, np.int32)
        w = g.w.reshape(-1, 2)[:, k]
        tight = (d[src] < O.INF // 2) & (d[src] + w == d[g.col]) & (g.col != 0)
        for u, v in zip(src[tight], g.col[tight]):
            high[v] = max(high[v], u)
        changed = changed or bool((high != p).any())
        write_tree(init, k, d, high)
    assert changed, "the weights in [1, 2] should give ties"
    out, exp = t / "out-ties", t / "exp-ties"
    mosp("--graph", pre, "--changes", t / "ties" / "changes", "--init", init, "--out", out,
         "--canonicalize", "--quiet")
    run(BIN / "mosp_prep", "expected", pre, t / "ties" / "changes", exp)
    for k in range(2):
        for f in ("distancesUpdated.txt", "SSSPTreeUpdated.txt"):
            assert (out / f"obj{k}" / f).read_bytes() == (exp / f"obj{k}" / f).read_bytes()


def test_objectives_prefs_and_binary_cache(work):
    """-k 2 --pref 2,1 with --cache: the first run writes the cache, the
    second reads it (not rebuilt); both equal the text graph's outputs, and
    so does the cache written by `mosp_prep cache`."""
    t, pre, g = work
    base = ["--graph", pre, "--init", t / "g" / "init", "--quiet", "--changes",
            t / "g" / "unsafe", "-k", 2, "--pref", "2,1"]
    cache = t / "graph.bin"
    r = mosp(*base, "--cache", cache, "--out", t / "out-cache-write")
    assert cache.exists()
    stamp = cache.stat().st_mtime_ns
    r = mosp(*base, "--cache", cache, "--out", t / "out-cache-read")
    assert "Note" not in r.stdout and cache.stat().st_mtime_ns == stamp
    mosp(*base, "--out", t / "out-text")
    run(BIN / "mosp_prep", "cache", pre, t / "prep.bin")
    r = mosp(*base, "--cache", t / "prep.bin", "--out", t / "out-prep-cache")
    assert "Note" not in r.stdout
    files = sorted(p.relative_to(t / "out-text") for p in (t / "out-text").rglob("*.txt"))
    assert len(files) == 2 * 2 + 3
    for d in ("out-cache-write", "out-cache-read", "out-prep-cache"):
        for f in files:
            assert (t / d / f).read_bytes() == (t / "out-text" / f).read_bytes(), f"{d}/{f}"
    check_against_oracle(g, read_batch(t / "g" / "unsafe", 3), t / "g" / "init",
                         t / "out-text", 2, [2, 1])


def test_serial_schedule_and_timing_csv(work):
    """--serial (K launches + 1) writes the same outputs as the fused launch;
    --timing writes the stage rows and the counter rows."""
    t, pre, _ = work
    base = ["--graph", pre, "--init", t / "g" / "init", "--quiet", "--changes", t / "g" / "safe"]
    mosp(*base, "--out", t / "out-fused", "--timing", t / "fused.csv")
    mosp(*base, "--out", t / "out-serial", "--serial", "--timing", t / "serial.csv")
    for p in (t / "out-fused").rglob("*.txt"):
        assert p.read_bytes() == (t / "out-serial" / p.relative_to(t / "out-fused")).read_bytes()
    for csv in ("fused.csv", "serial.csv"):
        rows = [ln.split(",") for ln in (t / csv).read_text().splitlines()]
        assert rows[0] == ["kind", "name", "value"]
        stages = {r[1] for r in rows[1:] if r[0] == "stage"}
        counters = {r[1]: float(r[2]) for r in rows[1:] if r[0] == "counter"}
        for s in ("device_context_init", "read_graph", "read_changes_and_trees", "upload",
                  "apply_batch", "obj0/sosp_update_gpu", "obj2/sosp_update_gpu",
                  "combined_graph_gpu", "download", "write_outputs", "TOTAL_end_to_end"):
            assert s in stages, (csv, s)
        for c in ("obj0/invalidated", "obj1/iterations", "obj2/pushes", "obj0/jump_rounds",
                  "comb/edges", "comb/L", "comb/delta", "comb/iterations", "comb/pushes"):
            assert c in counters, (csv, c)
        assert counters["comb/L"] == 1 and counters["comb/edges"] > 0


def test_trees_of_another_source_are_rejected(work):
    t, pre, _ = work
    r = mosp("--graph", pre, "--init", t / "g" / "init", "--changes", t / "g" / "safe",
             "--source", 5, "--no-output", validate=False, ok=False)
    assert r.returncode != 0 and "not rooted at the source 5" in r.stderr
