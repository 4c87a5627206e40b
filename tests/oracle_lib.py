"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference's
own host code built into oracle/_ref/libmospref.so.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
from dataclasses import dataclass

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libmospref.so"
INF = (2**63 - 1) // 4

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I, _L, _D = C.c_int, C.c_int64, C.c_double


def _lib() -> C.CDLL:
    if not ORACLE_SO.exists():
        raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
    lib = C.CDLL(str(ORACLE_SO))
    lib.or_dijkstra.argtypes = [_I, _i32p, _i32p, _i32p, _I, _I, _I, _i64p, _i32p]
    lib.or_apply_batch.restype = _I
    lib.or_apply_batch.argtypes = [_I, _I, _i32p, _i32p, _i32p, _I, _i32p, _i32p,
                                   _i32p, _I, _i32p, _i32p, _i32p, _i32p, _i32p,
                                   _u32p]
    lib.or_transpose.argtypes = [_I, _I, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
    lib.or_sequential_update.restype = _I
    lib.or_sequential_update.argtypes = [_I, _i32p, _i32p, _i32p, _i32p, _i32p,
                                         _i32p, _I, _I, _I, _I, _i32p, _i32p, _I,
                                         _i32p, _i64p, _i32p]
    lib.or_preference_scale.restype = _L
    lib.or_preference_scale.argtypes = [_i32p, _I, _I]
    lib.or_combined_edge_weight.restype = _L
    lib.or_combined_edge_weight.argtypes = [C.c_uint, _i32p, _I, _I, _L]
    lib.or_combined_graph.restype = _I
    lib.or_combined_graph.argtypes = [_i32p, _I, _I, _I, _i32p, _I, _i32p, _i32p,
                                      _i32p]
    lib.or_path_costs.restype = _I
    lib.or_path_costs.argtypes = [_I, _I, _i32p, _i32p, _i32p, _i32p, _I, _i64p]
    lib.or_canonicalize_tree.argtypes = [_I, _i32p, _i32p, _i32p, _I, _I, _I,
                                         _i64p, _i32p]
    lib.or_check_tree.argtypes = [_I, _i32p, _i32p, _i32p, _I, _I, _I, _i64p,
                                  _i32p, _i64p, C.c_void_p, _i64p]
    lib.or_default_delta.restype = _L
    lib.or_default_delta.argtypes = [_L, _I, _L]
    lib.or_choose_packing.restype = _I
    lib.or_choose_packing.argtypes = [_I, _L, C.POINTER(_I), C.POINTER(C.c_uint64)]
    lib.or_mosp_update.argtypes = [_I, _I, _i32p, _i32p, _i32p, _I, _i32p, _i32p,
                                   _i32p, _I, _i32p, _i32p, _I, _I, _i32p, _I,
                                   _i64p, _i32p, _i32p, _i32p, _i32p, _i64p,
                                   _i32p, _f64p, _i64p]
    lib.or_alg_bytes.restype = _L
    lib.or_alg_bytes.argtypes = [_I, _i32p, _i32p, _I, _I, _i32p, _i64p, _i64p,
                                 _i32p, _i32p, _i32p, _I, _i32p, _L]
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


def i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).reshape(-1))


@dataclass
class Graph:
    """Reference CsrGraph layout: edge-major weights w[e*K + k]."""
    n: int
    K: int
    row: np.ndarray
    col: np.ndarray
    w: np.ndarray

    @property
    def m(self) -> int:
        return int(self.row[-1])

    def weight(self, e: int, k: int) -> int:
        return int(self.w[e * self.K + k])


@dataclass
class Batch:
    K: int
    ins_from: np.ndarray
    ins_to: np.ndarray
    ins_w: np.ndarray  # insertion-major, K per insertion
    del_from: np.ndarray
    del_to: np.ndarray

    @property
    def n_ins(self) -> int:
        return len(self.ins_from)

    @property
    def n_del(self) -> int:
        return len(self.del_from)


def graph_from_edges(n: int, K: int, edges) -> Graph:
    """edges: iterable of (u, v, [w_0..w_K-1]); row order = stable by u."""
    edges = sorted(list(edges), key=lambda e: e[0])
    row = np.zeros(n + 1, dtype=np.int32)
    for u, _, _ in edges:
        row[u + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    col = i32([e[1] for e in edges])
    w = i32([x for e in edges for x in e[2]]) if edges else np.zeros(0, np.int32)
    return Graph(n, K, row, col, w)


def dijkstra(g: Graph, k: int, source: int):
    d = np.zeros(g.n, np.int64)
    p = np.zeros(g.n, np.int32)
    lib().or_dijkstra(g.n, g.row, g.col, g.w, g.K, k, source, d, p)
    return d, p


def apply_batch(g: Graph, b: Batch):
    cap = g.m + b.n_ins
    row = np.zeros(g.n + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * g.K, np.int32)
    mask = np.zeros(max(b.n_ins, 1), np.uint32)
    m = lib().or_apply_batch(g.n, g.K, g.row, g.col, g.w, b.n_ins, i32(b.ins_from),
                             i32(b.ins_to), i32(b.ins_w), b.n_del, i32(b.del_from),
                             i32(b.del_to), row, col, w, mask)
    assert m >= 0
    return Graph(g.n, g.K, row, col[:m].copy(), w[: m * g.K].copy()), mask[: b.n_ins].copy()


def transpose(g: Graph) -> Graph:
    row = np.zeros(g.n + 1, np.int32)
    col = np.zeros(max(g.m, 1), np.int32)
    w = np.zeros(max(g.m, 1) * g.K, np.int32)
    lib().or_transpose(g.n, g.K, g.row, g.col, g.w, row, col, w)
    return Graph(g.n, g.K, row, col[: g.m].copy(), w[: g.m * g.K].copy())


def changed_lists(b: Batch, mask: np.ndarray, k: int):
    """changed_k = all deletions + inserts whose increase bit k is set."""
    sel = (mask.astype(np.int64) >> k) & 1 == 1
    fr = np.concatenate([i32(b.del_from), i32(b.ins_from)[sel]]).astype(np.int32)
    to = np.concatenate([i32(b.del_to), i32(b.ins_to)[sel]]).astype(np.int32)
    return fr, to


def sequential_update(upd: Graph, rev: Graph, k: int, source: int, ch_from, ch_to,
                      heads, dist, parent):
    d = i64(dist).copy()
    p = i32(parent).copy()
    it = lib().or_sequential_update(upd.n, upd.row, upd.col, upd.w, rev.row, rev.col,
                                    rev.w, upd.K, k, source, len(ch_from), i32(ch_from),
                                    i32(ch_to), len(heads), i32(heads), d, p)
    assert it >= 0
    return d, p


def preference_scale(pref, K: int) -> int:
    a = i32(pref) if len(pref) else np.zeros(1, np.int32)
    return lib().or_preference_scale(a, len(pref), K)


def combined_graph(parents: np.ndarray, n: int, K: int, source: int, pref) -> Graph | None:
    a = i32(pref) if len(pref) else np.zeros(1, np.int32)
    row = np.zeros(n + 1, np.int32)
    col = np.zeros(max(K * n, 1), np.int32)
    w = np.zeros(max(K * n, 1), np.int32)
    ec = lib().or_combined_graph(i32(parents), n, K, source, a, len(pref), row, col, w)
    if ec < 0:
        return None
    return Graph(n, 1, row, col[:ec].copy(), w[:ec].copy())


def path_costs(g: Graph, parent, source: int):
    costs = np.zeros(g.n * g.K, np.int64)
    ok = lib().or_path_costs(g.n, g.K, g.row, g.col, g.w, i32(parent), source, costs)
    return (costs.reshape(g.n, g.K) if ok else None)


def canonicalize(g: Graph, k: int, source: int, dist, parent):
    p = i32(parent).copy()
    lib().or_canonicalize_tree(g.n, g.row, g.col, g.w, g.K, k, source, i64(dist), p)
    return p


def check_tree(rev: Graph, k: int, source: int, dist, parent, ref_dist, ref_parent=None):
    out = np.zeros(4, np.int64)
    rp = i32(ref_parent) if ref_parent is not None else None
    lib().or_check_tree(rev.n, rev.row, rev.col, rev.w, rev.K, k, source, i64(dist),
                        i32(parent), i64(ref_dist),
                        rp.ctypes.data if rp is not None else None, out)
    return out


def default_delta(m: int, n: int, wsum: int) -> int:
    return lib().or_default_delta(m, n, wsum)


def choose_packing(n: int, max_w: int):
    bits = _I(0)
    bound = C.c_uint64(0)
    ok = lib().or_choose_packing(n, max_w, C.byref(bits), C.byref(bound))
    return (bits.value, bound.value) if ok else None


def mosp_update(g: Graph, b: Batch, dist, parents, Kuse: int, source: int, pref):
    """Whole CPU pipeline.  Returns dict with updated graph, trees, combined."""
    n = g.n
    cap = g.m + b.n_ins
    row = np.zeros(n + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * g.K, np.int32)
    d = i64(dist).copy()
    p = i32(parents).copy()
    cd = np.zeros(n, np.int64)
    cp = np.zeros(n, np.int32)
    times = np.zeros(3, np.float64)
    info = np.zeros(4, np.int64)
    a = i32(pref) if len(pref) else np.zeros(1, np.int32)
    lib().or_mosp_update(n, g.K, g.row, g.col, g.w, b.n_ins, i32(b.ins_from),
                         i32(b.ins_to), i32(b.ins_w), b.n_del, i32(b.del_from),
                         i32(b.del_to), Kuse, source, a, len(pref), d, p, row, col, w,
                         cd, cp, times, info)
    m = int(info[0])
    return {
        "ok": bool(info[3]),
        "updated": Graph(n, g.K, row, col[:m].copy(), w[: m * g.K].copy()),
        "dist": d.reshape(Kuse, n), "parents": p.reshape(Kuse, n),
        "comb_dist": cd, "comb_parent": cp, "combined_edges": int(info[1]),
        "scale": int(info[2]), "times_ms": times.copy(),
    }


def alg_bytes(upd: Graph, rev: Graph, Kuse: int, source: int, old_parents, old_dist,
              new_dist, changed: list, heads, combined_edges: int) -> int:
    offs = [0]
    fr, to = [], []
    for cf, ct in changed:
        fr.extend(list(cf))
        to.extend(list(ct))
        offs.append(len(fr))
    return lib().or_alg_bytes(upd.n, upd.row, rev.row, Kuse, source, i32(old_parents),
                              i64(old_dist), i64(new_dist), i32(offs),
                              i32(fr) if fr else np.zeros(1, np.int32),
                              i32(to) if to else np.zeros(1, np.int32),
                              len(heads), i32(heads) if len(heads) else np.zeros(1, np.int32),
                              combined_edges)


# --------------------------------------------------------------------------
# The reference's own host code (oracle/_ref), when it was built here.
# --------------------------------------------------------------------------

def ref_available() -> bool:
    return REF_SO.exists()


_REF: C.CDLL | None = None


def ref() -> C.CDLL:
    global _REF
    if _REF is None:
        r = C.CDLL(str(REF_SO))
        r.ref_dijkstra.argtypes = [_I, _i32p, _i32p, _i32p, _I, _I, _I, _i64p, _i32p]
        r.ref_apply_batch.restype = _I
        r.ref_apply_batch.argtypes = lib().or_apply_batch.argtypes
        s = C.c_char_p
        r.ref_sequential_update_files.restype = _I
        r.ref_sequential_update_files.argtypes = [s, s, s, s, s, _I, _I, s, s]
        r.ref_generate_change_batch_files.restype = _I
        r.ref_generate_change_batch_files.argtypes = [s, _I, _D, _I, _I, _I, C.c_uint,
                                                      _I, _I, _I, s, s]
        r.ref_generate_graph_csr.restype = _I
        r.ref_generate_graph_csr.argtypes = [_I, _I, _I, s, _I, _I, _I, C.c_uint]
        r.ref_generate_changed_edges.restype = _I
        r.ref_generate_changed_edges.argtypes = [_I, _I, _I, _I, _I, _D, _D, _I, _I, _I,
                                                 _I, s, s, s, C.c_uint]
        for name in ("ref_read_csr_ok", "ref_read_distances_ok", "ref_read_parents_ok"):
            getattr(r, name).restype = _I
        r.ref_read_csr_ok.argtypes = [s]
        r.ref_read_distances_ok.argtypes = [s, _I]
        r.ref_read_parents_ok.argtypes = [s, _I]
        r.ref_save_cache.restype = _I
        r.ref_save_cache.argtypes = [s, s]
        r.ref_load_cache_ok.restype = _I
        r.ref_load_cache_ok.argtypes = [s, s]
        r.ref_read_batch_ok.restype = _I
        r.ref_read_batch_ok.argtypes = [s, s, _I, _I]
        _REF = r
    return _REF


def ref_dijkstra(g: Graph, k: int, source: int):
    d = np.zeros(g.n, np.int64)
    p = np.zeros(g.n, np.int32)
    ref().ref_dijkstra(g.n, g.row, g.col, g.w, g.K, k, source, d, p)
    return d, p


def ref_apply_batch(g: Graph, b: Batch):
    cap = g.m + b.n_ins
    row = np.zeros(g.n + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    w = np.zeros(max(cap, 1) * g.K, np.int32)
    mask = np.zeros(max(b.n_ins, 1), np.uint32)
    m = ref().ref_apply_batch(g.n, g.K, g.row, g.col, g.w, b.n_ins, i32(b.ins_from),
                              i32(b.ins_to), i32(b.ins_w), b.n_del, i32(b.del_from),
                              i32(b.del_to), row, col, w, mask)
    assert m >= 0
    return Graph(g.n, g.K, row, col[:m].copy(), w[: m * g.K].copy()), mask[: b.n_ins].copy()


# --------------------------------------------------------------------------
# Text files in the reference's layouts (for the file-based reference calls)
# --------------------------------------------------------------------------

def write_csr(prefix: str, g: Graph) -> None:
    os.makedirs(os.path.dirname(prefix), exist_ok=True)
    with open(prefix + "RowPtr.txt", "w") as f:
        f.write("".join(f"{x}\n" for x in g.row))
    with open(prefix + "ColInd.txt", "w") as f:
        f.write("".join(f"{x}\n" for x in g.col))
    w = g.w.reshape(-1, g.K) if g.m else np.zeros((0, g.K), np.int32)
    with open(prefix + "Values.txt", "w") as f:
        f.write("".join(" ".join(str(x) for x in r) + "\n" for r in w))


def write_batch(ins_path: str, del_path: str, b: Batch) -> None:
    os.makedirs(os.path.dirname(ins_path), exist_ok=True)
    w = i32(b.ins_w).reshape(-1, b.K) if b.n_ins else np.zeros((0, b.K), np.int32)
    with open(ins_path, "w") as f:
        for u, v, ws in zip(b.ins_from, b.ins_to, w):
            f.write(f"{u} {v} " + " ".join(str(x) for x in ws) + "\n")
    with open(del_path, "w") as f:
        for u, v in zip(b.del_from, b.del_to):
            f.write(f"{u} {v}\n")


def read_batch(ins_path: str, del_path: str, K: int) -> Batch:
    fr, to, ws, df, dt = [], [], [], [], []
    with open(ins_path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 2 + K:
                fr.append(int(t[0]))
                to.append(int(t[1]))
                ws.extend(int(x) for x in t[2:2 + K])
    with open(del_path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 2:
                df.append(int(t[0]))
                dt.append(int(t[1]))
    return Batch(K, i32(fr), i32(to), i32(ws), i32(df), i32(dt))


def read_csr(prefix: str) -> Graph:
    row = np.loadtxt(prefix + "RowPtr.txt", dtype=np.int64).astype(np.int32).reshape(-1)
    col = np.loadtxt(prefix + "ColInd.txt", dtype=np.int64).astype(np.int32).reshape(-1)
    with open(prefix + "Values.txt") as f:
        rows = [list(map(int, line.split())) for line in f if line.strip()]
    K = len(rows[0]) if rows else 1
    return Graph(len(row) - 1, K, row, col, i32([x for r in rows for x in r]))


def write_tree(dist_path: str, tree_path: str, dist, parent) -> None:
    os.makedirs(os.path.dirname(dist_path), exist_ok=True)
    with open(dist_path, "w") as f:
        f.write("".join(f"{v} {'INF' if d >= INF // 2 else d}\n" for v, d in enumerate(dist)))
    with open(tree_path, "w") as f:
        f.write("".join(f"{v} {p}\n" for v, p in enumerate(parent)))


def read_tree(dist_path: str, tree_path: str, n: int):
    d = np.zeros(n, np.int64)
    p = np.zeros(n, np.int32)
    with open(dist_path) as f:
        for line in f:
            t = line.split()
            if t:
                d[int(t[0])] = INF if t[1] == "INF" else int(t[1])
    with open(tree_path) as f:
        for line in f:
            t = line.split()
            if t:
                p[int(t[0])] = int(t[1])
    return d, p
