"""Bounded grid-barrier waits (no counterpart in the reference, which relies
on its cooperative launch).  The host makes every wait give up at once
(limit 0 ms): the launch must end -- every workgroup leaves, the call returns
an error, the device stays usable -- and the next launch on the same
workspace or batch, with the default limit, must be exact (the scratch the
aborted launch left is cleared first)."""
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


def instance(width, height, K, changes, seed, pref=()):
    Mm = M()
    g = Mm.gen_road_grid(width, height, K, seed=seed)
    b = Mm.gen_batch(g, changes, 50.0, seed=seed + 1, safe=False)
    og = O.Graph(g.n, g.K, g.row, g.col, g.w)
    d0, p0 = gen.initial_trees(og, K)
    want = O.mosp_update(og, O.Batch(K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to),
                         d0, p0, K, 0, list(pref))
    return g, b, d0, p0, want


def same(got, want, tag):
    for key in ("dist", "parents", "comb_dist", "comb_parent"):
        assert np.array_equal(got[key], want[key]), f"{tag}: {key}"


def test_single_launch_gives_up_then_recovers():
    g, b, d0, p0, want = instance(80, 80, 3, 600, 71)
    s = M().Session(g, b, d0, p0, K=3)
    s.ws.set_barrier_timeout(0.0)
    for attempt in range(2):
        with pytest.raises(RuntimeError, match="gave up"):
            s.run()
        say(f"attempt {attempt}: the launch gave up and returned")
    assert M().lib().mosp_device_synchronize() == 1
    s.ws.set_barrier_timeout(2000.0)
    for rep in range(2):
        s.run()
        same(s.results(), want, f"after the aborted launches, run {rep}")
    s.close()


def test_batched_launch_gives_up_then_recovers():
    from mosp_amd import batched as Bm
    cases = [instance(70, 60, 2, 400, 81, (1, 2)), instance(50, 90, 4, 500, 83, (2, 1, 1, 3))]
    sessions = [M().Session(g, b, d0, p0, K=len(w["dist"]), pref=pr)
                for (g, b, d0, p0, w), pr in zip(cases, [(1, 2), (2, 1, 1, 3)])]
    bu = Bm.BatchedUpdate(sessions)
    bu.run(use_graph=True)
    for i, c in enumerate(cases):
        same(bu.results(i), c[4], f"instance {i} before")
    sizes, _ = bu.schedule()
    assert min(sizes) > 1, sizes  # groups of several workgroups: their barriers wait
    bu.set_barrier_timeout(0.0)
    for graph in (True, False):
        with pytest.raises(RuntimeError, match="gave up"):
            bu.run(use_graph=graph)
        say(f"batched (graph={graph}) gave up and returned")
    bu.set_barrier_timeout(2000.0)
    for graph in (True, False, True):
        bu.run(use_graph=graph)
        bu.check()
        for i, c in enumerate(cases):
            same(bu.results(i), c[4], f"instance {i} after the aborted runs (graph={graph})")
    bu.close()
    for s in sessions:
        s.close()
