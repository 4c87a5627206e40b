"""Algorithmic HBM bytes of one MOSP update (the roofline numerator).

Per objective k, with the updated graph's in/out degrees:
  B_k = 28 n                                  old dist+parent read, new written,
                                              parent re-read
      + sum over C_k of (8 + 16 indeg(v))     pull: row-pointer pair + per in-edge
                                              (column 4 + weight 4 + state 8)
      + sum over D_k of (8 + 16 outdeg(u))    push: every vertex whose final label
                                              was (re)produced relaxes its out-edges
  C_k = I_k u (H \\ {s}),  D_k = {u != s : d'(u) < INF and (u in I_k or d'(u) != d(u))}
  I_k = subtrees (in the old tree) of the heads of changed tree edges.
Combined step: 4 K n (read K parents) + 4 (n+1) + 8 E (write CSR)
             + 8 n + 16 E (one relaxation pass) + 12 n (write dist + parent).
An objective with distance-only words (its (n-1) * maxWeight bound does not
fit beside the parent ids and the optimistic packed words were redone; the
reference's fallback, src/sospUpdateGpu.cu:490-511) adds its lowest-id
parent recovery over every vertex and edge:
  F_k = 12 n (each vertex's distance and row pointer) + 16 m (each out-edge's
        column and weight, and its head's distance).

Everything is computed with numpy from the inputs and the outputs of the
update, so it is deterministic and independent of the kernel's schedule.
"""
from __future__ import annotations

import numpy as np

INF = (2**63 - 1) // 4


def invalidated(parent: np.ndarray, ch_from: np.ndarray, ch_to: np.ndarray,
                source: int) -> np.ndarray:
    """Vertices whose old tree path runs through a changed tree edge
    (pointer jumping over the parent array, vectorised)."""
    n = len(parent)
    flag = np.zeros(n, bool)
    if len(ch_to):
        hit = parent[ch_to] == ch_from
        flag[ch_to[hit]] = True
    anc = parent.astype(np.int64).copy()
    anc[source] = -1
    while True:
        live = anc >= 0
        if not live.any():
            break
        idx = np.nonzero(live)[0]
        a = anc[idx]
        flag[idx] |= flag[a]
        anc[idx] = anc[a]
    return flag


def alg_bytes(n: int, out_row: np.ndarray, in_row: np.ndarray, source: int,
              old_parents: np.ndarray, old_dist: np.ndarray, new_dist: np.ndarray,
              changed: list, heads: np.ndarray, combined_edges: int,
              distance_only=None) -> int:
    """distance_only: per objective, whether it ran with distance-only words
    (adds its parent-recovery sweep F_k); None: none did."""
    outdeg = np.diff(out_row).astype(np.int64)
    indeg = np.diff(in_row).astype(np.int64)
    K = len(old_parents)
    total = 0
    h = np.zeros(n, bool)
    if len(heads):
        h[heads] = True
    h[source] = False
    for k in range(K):
        cf, ct = changed[k]
        inv = invalidated(old_parents[k], cf, ct, source)
        cand = inv | h
        nd, od = new_dist[k], old_dist[k]
        d = (nd < INF // 2) & (inv | (nd != od))
        d[source] = False
        total += 28 * n + int((8 + 16 * indeg[cand]).sum()) + int((8 + 16 * outdeg[d]).sum())
    m = int(out_row[-1]) if len(out_row) else 0
    for k in range(K):
        if distance_only is not None and distance_only[k]:
            total += 12 * n + 16 * m
    E = int(combined_edges)
    total += 4 * K * n + 4 * (n + 1) + 8 * E + 8 * n + 16 * E + 12 * n
    return int(total)
