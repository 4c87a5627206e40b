"""Seeded test-case generators (numpy) for the parity tests.

Graph families and batch kinds follow the reference's oracle suite
(src/mospTest.cu:64-116 graphs, 300-356 change sets; src/changeGenerator.cu
modes Uniform / Reweight / Increase and the connectivity-safe filter of
changeGenerator.cu:78-116).  The byte streams are this module's own
(numpy PCG64), so the instances differ from the reference's; the properties
checked on them do not depend on the instance.
"""
from __future__ import annotations

import numpy as np

from oracle_lib import Batch, Graph, apply_batch, dijkstra, i32


def _csr(n: int, K: int, us, vs, ws) -> Graph:
    us = np.asarray(us, dtype=np.int64)
    order = np.argsort(us, kind="stable")
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, us + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    col = np.asarray(vs, dtype=np.int32)[order]
    w = np.asarray(ws, dtype=np.int32).reshape(-1, K)[order].reshape(-1)
    return Graph(n, K, row, np.ascontiguousarray(col), np.ascontiguousarray(w))


def random_graph(n: int, extra: int, K: int, wmax: int, seed: int,
                 wmin: int = 1, duplicates: bool = True) -> Graph:
    """Chain 0 -> 1 -> ... -> n-1 plus `extra` random arcs u != v."""
    rng = np.random.default_rng(seed)
    us = list(range(n - 1))
    vs = list(range(1, n))
    seen = set(zip(us, vs))
    tries = 0
    while len(us) < n - 1 + extra and tries < 50 * (extra + 1):
        tries += 1
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u == v or (not duplicates and (u, v) in seen):
            continue
        seen.add((u, v))
        us.append(u)
        vs.append(v)
    ws = rng.integers(wmin, wmax + 1, size=(len(us), K))
    return _csr(n, K, us, vs, ws)


def grid_graph(width: int, height: int, K: int, wmax: int, drop: float, seed: int,
               wmin: int = 1) -> Graph:
    """Road-like grid, both directions, a fraction of grid edges dropped."""
    rng = np.random.default_rng(seed)
    us, vs = [], []
    for y in range(height):
        for x in range(width):
            for dx, dy in ((1, 0), (0, 1)):
                nx, ny = x + dx, y + dy
                if nx >= width or ny >= height or rng.random() < drop:
                    continue
                a, b = y * width + x, ny * width + nx
                us += [a, b]
                vs += [b, a]
    ws = rng.integers(wmin, wmax + 1, size=(len(us), K))
    return _csr(width * height, K, us, vs, ws)


def reachable(g: Graph, source: int) -> np.ndarray:
    seen = np.zeros(g.n, bool)
    seen[source] = True
    stack = [source]
    while stack:
        u = stack.pop()
        for v in g.col[g.row[u]:g.row[u + 1]]:
            if not seen[v]:
                seen[v] = True
                stack.append(int(v))
    return seen


def _empty(K: int) -> Batch:
    z = np.zeros(0, np.int32)
    return Batch(K, z, z, z, z, z)


def uniform_batch(g: Graph, changes: int, ins_pct: float, seed: int, wmax: int = 50,
                  safe: bool = False, source: int = 0, wmin: int = 1) -> Batch:
    """Inserts join random pairs u != v; deletions pick existing arcs."""
    rng = np.random.default_rng(seed)
    n_ins = int(round(changes * ins_pct / 100.0))
    n_del = changes - n_ins
    fr, to = [], []
    while len(fr) < n_ins:
        u, v = (int(x) for x in rng.integers(0, g.n, 2))
        if u != v:
            fr.append(u)
            to.append(v)
    ws = rng.integers(wmin, wmax + 1, size=(n_ins, g.K)).reshape(-1)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    pick = rng.integers(0, max(g.m, 1), size=n_del) if g.m else np.zeros(0, int)
    b = Batch(g.K, i32(fr), i32(to), i32(ws), i32(src[pick]), i32(g.col[pick]))
    if safe:
        b = filter_unsafe(g, b, source)
    return b


def filter_unsafe(g: Graph, b: Batch, source: int) -> Batch:
    """Drop deletions touching vertices that lose reachability, repeated
    (the rule of changeGenerator.cu:78-116)."""
    before = reachable(g, source)
    while b.n_del > 0:
        upd, _ = apply_batch(g, b)
        lost = before & ~reachable(upd, source)
        if not lost.any():
            break
        keep = ~(lost[b.del_from] | lost[b.del_to])
        if keep.all():
            break
        b = Batch(b.K, b.ins_from, b.ins_to, b.ins_w, i32(b.del_from[keep]),
                  i32(b.del_to[keep]))
    return b


def increase_batch(g: Graph, changes: int, seed: int, source: int = 0,
                   wmax: int = 50) -> Batch:
    """Raise every weight of distinct tree edges (of random objectives)."""
    rng = np.random.default_rng(seed)
    trees = [dijkstra(g, k, source)[1] for k in range(g.K)]
    picked = set()
    fr, to, ws = [], [], []
    for _ in range(changes * 4):
        if len(fr) >= changes:
            break
        k = int(rng.integers(0, g.K))
        v = int(rng.integers(0, g.n))
        p = int(trees[k][v])
        if p < 0 or (p, v) in picked:
            continue
        e = next(e for e in range(g.row[p], g.row[p + 1]) if g.col[e] == v)
        picked.add((p, v))
        fr.append(p)
        to.append(v)
        ws.extend(int(g.w[e * g.K + j]) + int(rng.integers(1, wmax + 1))
                  for j in range(g.K))
    z = np.zeros(0, np.int32)
    return Batch(g.K, i32(fr), i32(to), i32(ws), z, z)


def reweight_batch(g: Graph, changes: int, seed: int, wmax: int = 50) -> Batch:
    """Overwrite existing arcs with new random weights (up and down)."""
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    pick = rng.integers(0, g.m, size=changes)
    ws = rng.integers(1, wmax + 1, size=(changes, g.K)).reshape(-1)
    z = np.zeros(0, np.int32)
    return Batch(g.K, i32(src[pick]), i32(g.col[pick]), i32(ws), z, z)


def disconnecting_batch(g: Graph, seed: int, source: int = 0) -> Batch:
    """Delete every in-edge of a few random vertices (src/mospTest.cu:309-325)."""
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    fr, to = [], []
    for _ in range(max(1, g.n // 50)):
        v = int(rng.integers(1, g.n))
        sel = np.nonzero(g.col == v)[0]
        fr.extend(int(x) for x in src[sel])
        to.extend([v] * len(sel))
    z = np.zeros(0, np.int32)
    return Batch(g.K, z, z, z, i32(fr), i32(to))


BATCH_KINDS = ("uniform-safe", "uniform-unsafe", "deletions-only", "insertions-only",
               "tree-weight-increases", "reweight", "disconnecting")


def make_batch(kind: str, g: Graph, changes: int, seed: int, source: int = 0) -> Batch:
    if kind == "uniform-safe":
        return uniform_batch(g, changes, 50, seed, safe=True, source=source)
    if kind == "uniform-unsafe":
        return uniform_batch(g, changes, 50, seed)
    if kind == "deletions-only":
        return uniform_batch(g, changes, 0, seed)
    if kind == "insertions-only":
        return uniform_batch(g, changes, 100, seed)
    if kind == "tree-weight-increases":
        return increase_batch(g, changes, seed, source)
    if kind == "reweight":
        return reweight_batch(g, changes, seed)
    if kind == "disconnecting":
        return disconnecting_batch(g, seed, source)
    raise ValueError(kind)


def initial_trees(g: Graph, K: int, source: int = 0):
    d = np.zeros((K, g.n), np.int64)
    p = np.zeros((K, g.n), np.int32)
    for k in range(K):
        d[k], p[k] = dijkstra(g, k, source)
    return d, p
