"""mosp_sosp_update_noalloc (the entry point that allocates nothing; its
scratch is one static device array) against mosp_sosp_update, the entry
point it varies, through the C ABI: two calls in a row, the second on a
smaller graph, over both kernel variants (a grid and an R-MAT graph with
hubs) and large weights: packed words bounded by their own limit (the
optimistic packing), and distance-only words (MOSP_PACK_LIMIT=1 makes the
optimistic launch meet a larger distance and be redone with them).  Results
must be bit-identical, and equal canonical Dijkstra on the updated graph.
Graphs above the stated maximum are refused before any device work."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

import gen
import oracle_lib as O
from oracle_lib import i32

pytestmark = pytest.mark.gpu

MAX_VERTICES = 1 << 21


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def M():
    import mosp_amd
    return mosp_amd


def case(kind: str, seed: int):
    Mm = M()
    if kind == "grid-large":
        g = Mm.gen_road_grid(300, 300, 2, seed=seed)
    elif kind == "grid-small":
        g = Mm.gen_road_grid(40, 40, 2, seed=seed)
    elif kind == "rmat":
        g = Mm.gen_rmat(14, 8, 2, seed=seed)
    else:  # large weights: (n - 1) * max weight does not fit beside the parent bits
        g = Mm.gen_road_grid(320, 320, 2, seed=seed, wmin=1 << 29, wmax=(1 << 31) - 1)
    large = kind in ("fallback", "optimistic")
    b = Mm.gen_batch(g, max(50, g.n // 20), 50.0, seed=seed + 1, safe=False,
                     wmin=1 << 29 if large else 1, wmax=(1 << 31) - 1 if large else 100)
    return g, b


def run_both(g, b, k: int):
    Mm = M()
    og = O.Graph(g.n, g.K, g.row, g.col, g.w)
    d0, p0 = O.dijkstra(og, k, 0)
    upd, mask = Mm.apply_batch(g, b)
    cf, ct = Mm.changed_lists(b, mask, k)
    maxw, wsum = Mm.objective_bounds(g, b, g.K)
    delta = Mm.default_delta(max(g.m, 1), g.n, int(wsum[k]))
    dg = Mm.DeviceGraph(upd)
    ws = Mm.Workspace(g.n)
    want_d, want_p, st_ws = Mm.sosp_update(dg, k, cf, ct, b.ins_to, 0, delta, int(maxw[k]), ws,
                                           d0, p0)
    got_d, got_p, st = Mm.sosp_update_noalloc(dg, k, cf, ct, b.ins_to, 0, delta, int(maxw[k]),
                                              d0, p0)
    ou = O.Graph(upd.n, upd.K, upd.row, upd.col, upd.w)
    ref_d, ref_p = O.dijkstra(ou, k, 0)
    ws.close()
    dg.close()
    return (got_d, got_p, st), (want_d, want_p, st_ws), (ref_d, ref_p)


@pytest.mark.parametrize("kinds", [("grid-large", "grid-small"), ("rmat", "grid-small"),
                                   ("fallback", "grid-small"), ("optimistic", "grid-small")])
def test_noalloc_equals_workspace_entry_point(kinds, monkeypatch):
    for i, kind in enumerate(kinds):
        if kind == "fallback":
            monkeypatch.setenv("MOSP_PACK_LIMIT", "1")
        else:
            monkeypatch.delenv("MOSP_PACK_LIMIT", raising=False)
        g, b = case(kind, 11 + 7 * i)
        for k in range(2):
            (gd, gp, st), (wd, wp, sw), (rd, rp) = run_both(g, b, k)
            say(f"{kind} n={g.n} obj{k}: invalidated {st['invalidated']} iterations "
                f"{st['iterations']} packed {st['packed_parents']} kernel {st['kernel_ms']:.3f} ms")
            assert np.array_equal(gd, wd) and np.array_equal(gp, wp), f"{kind} obj{k} vs workspace"
            assert np.array_equal(gd, rd) and np.array_equal(gp, rp), f"{kind} obj{k} vs Dijkstra"
            for key in ("invalidated", "jump_rounds", "packed_parents"):
                assert st[key] == sw[key], (kind, k, key)
            if kind in ("fallback", "optimistic"):
                assert st["packed_parents"] == (0 if kind == "fallback" else 1)


def test_noalloc_refuses_graphs_above_its_maximum():
    Mm = M()
    g = Mm.gen_road_grid(20, 20, 1, seed=3)
    dg = Mm.DeviceGraph(g)
    out, inn = dg.view(0)
    big_out = Mm.Csr(MAX_VERTICES + 1, out.m, out.row_ptr, out.col_ind, out.weights)
    big_in = Mm.Csr(MAX_VERTICES + 1, inn.m, inn.row_ptr, inn.col_ind, inn.weights)
    ch = Mm.Changes(None, None, 0, None, 0)
    st = Mm.SospStats()
    ok = Mm.lib().mosp_sosp_update_noalloc(C.byref(big_out), C.byref(big_in), C.byref(ch), 0, 10,
                                           100, None, None, C.byref(st), None)
    assert ok == 0 and "at most 2097152 vertices" in Mm.last_error()
    dg.close()
