"""Batched launches (mosp_batch_*): many instances in ONE persistent launch.

Every instance's outputs -- K updated trees, combined distances and MOSP tree
-- must equal the oracle's pipeline on that instance, and equal what the
plain fused launch of the same instance writes, whatever the schedule:
graph replay or direct launch, groups sized to the instances or several
instances per single-workgroup group, XCD placement on or off, any lane
width or near-far bucket width.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
from oracle_lib import i32

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def B():
    from mosp_amd import batched
    return batched


# Uneven small instances: sizes, K, batches and preferences all differ.
SMALL = [  # width, height, K, changes, pref, seed
    (12, 12, 2, 40, (1, 2), 11),
    (20, 17, 3, 120, (1, 1, 1), 12),
    (31, 29, 5, 300, (1, 2, 3, 4, 1), 13),
    (40, 40, 8, 500, (4, 3, 2, 1, 4, 3, 2, 1), 14),
    (9, 50, 4, 90, (2, 2, 3, 3), 15),
    (25, 25, 2, 200, (3, 1), 16),
    (3, 3, 3, 6, (), 17),
]


@pytest.fixture(scope="module")
def small():
    specs = [B().InstanceSpec(w, h, K, ch, pref, seed) for (w, h, K, ch, pref, seed) in SMALL]
    built = [B().build_instance(s) for s in specs]
    yield specs, built
    for _, s in built:
        s.close()


def oracle_of(spec, host):
    g, b, dist, par = host
    og = O.Graph(g.n, g.K, g.row, g.col, g.w)
    ob = O.Batch(b.K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to)
    want = O.mosp_update(og, ob, dist, par, spec.K, 0, list(spec.pref))
    assert want["ok"]
    return want


def assert_same(got, want, tag):
    K = len(got["dist"])
    for k in range(K):
        assert np.array_equal(got["dist"][k], want["dist"][k]), f"{tag} obj{k} distances"
        assert np.array_equal(got["parents"][k], want["parents"][k]), f"{tag} obj{k} parents"
    assert np.array_equal(got["comb_dist"], want["comb_dist"]), f"{tag} combined distances"
    assert np.array_equal(got["comb_parent"], want["comb_parent"]), f"{tag} MOSP tree"


def test_batched_matches_oracle_and_fused(small):
    specs, built = small
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions)
    say("batched: graph replay of", len(sessions), "instances; schedule", bu.schedule())
    bu.run(use_graph=True)
    bu.check()
    batched = [bu.results(i) for i in range(len(sessions))]
    for i, (spec, (host, s)) in enumerate(zip(specs, built)):
        want = oracle_of(spec, host)
        assert_same(batched[i], want, f"instance {i} (batched vs oracle)")
        assert batched[i]["combined_edges"] == want["combined_edges"]
        s.run()
        assert_same(s.results(), batched[i], f"instance {i} (fused vs batched)")
    bu.close()


def test_batched_full_lists_fall_back_to_sweeps(small, monkeypatch):
    """Near / far lists capped far below n (MOSP_LIST_CAP) in the batched
    kernel: the iterations after a full list sweep every vertex; results
    equal the oracle's."""
    specs, built = small
    monkeypatch.setenv("MOSP_LIST_CAP", "8")
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions)
    say("batched with capped lists:", len(sessions), "instances")
    bu.run(use_graph=False)
    bu.check()
    for i, (spec, (host, _)) in enumerate(zip(specs, built)):
        assert_same(bu.results(i), oracle_of(spec, host), f"instance {i} (capped lists)")
    bu.close()


@pytest.mark.parametrize("mode", ["1", "2"])
def test_batched_step1_pull_and_push(small, monkeypatch, mode):
    """Step 1's candidates for invalidated vertices always pulled (1) or
    always pushed from the valid vertices (2) in the batched kernel."""
    specs, built = small
    monkeypatch.setenv("MOSP_INIT_PUSH", mode)
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions)
    bu.run(use_graph=True)
    bu.check()
    for i, (spec, (host, _)) in enumerate(zip(specs, built)):
        assert_same(bu.results(i), oracle_of(spec, host), f"instance {i} (Step 1 mode {mode})")
    bu.close()


def test_replays_and_direct_launch_agree(small):
    specs, built = small
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions)
    ref = None
    for rep, graph in enumerate([True, False, True, True]):
        say(f"replay {rep} graph={graph}")
        bu.run(use_graph=graph)
        res = [bu.results(i) for i in range(len(sessions))]
        for s in sessions:  # eager work between the replays
            s.run()
        if ref is None:
            ref = res
        for i in range(len(sessions)):
            assert_same(res[i], ref[i], f"replay {rep} instance {i}")
    sizes, mode = bu.schedule()
    say("group sizes", sizes, "graph mode", mode)
    assert mode in (1, 2) and sum(sizes) <= 256 and min(sizes) >= 1
    bu.close()


@pytest.mark.parametrize("workgroups,lane_bits,delta_scale,xcd", [
    (4, 2, 1.0, True),     # more instances than workgroups: groups run several in turn
    (3, 1, 3.0, False),
    (64, 3, 0.5, True),
    (0, 0, 6.0, False),
])
def test_schedule_does_not_change_results(small, workgroups, lane_bits, delta_scale, xcd):
    specs, built = small
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions, workgroups=workgroups, lane_bits=lane_bits,
                           delta_scale=delta_scale, xcd_aware=xcd)
    say(f"schedule wg={workgroups} lanes={lane_bits} delta x{delta_scale} xcd={xcd}",
        bu.schedule())
    bu.run(use_graph=True)
    for i, (spec, (host, s)) in enumerate(zip(specs, built)):
        assert_same(bu.results(i), oracle_of(spec, host), f"instance {i}")
    bu.close()


@pytest.mark.parametrize("threads,queue_cap,workgroups", [
    (512, 0, 0),       # two 512-thread workgroups per CU
    (512, 256, 7),     # a queue of 256 entries: most appends go straight to the lists
    (1024, 1024, 0),
    (512, 2048, 300),
])
def test_workgroup_shapes_do_not_change_results(small, threads, queue_cap, workgroups):
    """The batched kernel's two workgroup shapes and smaller LDS push queues
    (the autotuner's second stage) give the oracle's results."""
    specs, built = small
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions)
    bu.configure(workgroups=workgroups, threads=threads, queue_cap=queue_cap)
    sizes, mode = bu.schedule()
    say(f"shape {threads} threads, queue {queue_cap}, wg {workgroups}: groups {sizes}")
    bu.run(use_graph=True)
    bu.run(use_graph=True)
    for i, (spec, (host, s)) in enumerate(zip(specs, built)):
        assert_same(bu.results(i), oracle_of(spec, host), f"instance {i}")
    bu.close()


def test_uneven_mix_subset_against_fused():
    """Eight instances drawn from the batched mix (10k-120k vertices)."""
    specs = B().road_specs(count=8, n_max=120_000, ch_max=20_000, seed=9)
    built = [B().build_instance(s) for s in specs]
    sessions = [s for _, s in built]
    bu = B().BatchedUpdate(sessions)
    bu.run(use_graph=True)
    bu.check()
    for i, (spec, (host, s)) in enumerate(zip(specs, built)):
        say(f"mix instance {i}: n={spec.n} K={spec.K} changes={spec.changes}")
        got = bu.results(i)
        s.run()
        assert_same(got, s.results(), f"mix instance {i} (batched vs fused)")
        if spec.n <= 40_000:
            assert_same(got, oracle_of(spec, host), f"mix instance {i} (oracle)")
    bu.close()
    for s in sessions:
        s.close()


def test_invalid_batches_are_rejected():
    from mosp_amd import BatchInstance, lib
    import ctypes as C
    assert not lib().mosp_batch_create(None, 0)
    arr = (BatchInstance * 1)()
    assert not lib().mosp_batch_create(arr, 1)  # no graph
