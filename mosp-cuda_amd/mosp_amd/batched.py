"""Batched updates: many independent update instances in ONE persistent launch.

`BatchedUpdate` wraps the C ABI mosp_batch_* over a list of `Session`s (each
owns its graph, batch, old trees and output buffers in HBM).  The launch
sequence (control-block reset + persistent kernel) is captured as a HIP graph
and replayed; `run()` returns the device time of one pass.  Each instance's
outputs are exactly what `Session.run()` (one fused launch per instance)
writes -- the batched schedule shares only the GPU.

`road_instances()` builds the uneven mix of the batched workload: road-like
grids of log-uniform size, K uniform in {2..8}, batches of log-uniform size
with 40% insertions and random integer preferences in [1, 4].
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import (BatchInstance, Changes, CombineStats, Session, SospStats, _check, _track,
               gen_batch, gen_road_grid, initial_trees, last_error, lib)


class BatchedUpdate:
    """mosp_batch: the sessions' updates in one launch (optionally a HIP graph)."""

    def __init__(self, sessions, workgroups: int = 0, lane_bits: int = -1,
                 delta_scale: float = 1.0, xcd_aware: bool = True):
        self.sessions = list(sessions)
        arr = (BatchInstance * len(self.sessions))()
        self._keep = []
        for i, s in enumerate(self.sessions):
            dl = (C.c_int64 * s.K)(*[int(x) for x in s.delta])
            mw = (C.c_int64 * s.K)(*[int(x) for x in s.maxw])
            pr = (C.c_int32 * max(1, len(s.pref)))(*[int(x) for x in s.pref])
            self._keep += [dl, mw, pr]
            arr[i] = BatchInstance(s.dg.handle, s.K, C.cast(s.changes, C.POINTER(Changes)),
                                   s.source, dl, mw, pr, len(s.pref), s.init_dist.ptr,
                                   s.init_par.ptr, s.dist.ptr, s.par.ptr, s.cdist.ptr, s.cpar.ptr)
        self._arr = arr
        self.handle = lib().mosp_batch_create(arr, len(self.sessions))
        if not self.handle:
            raise RuntimeError(f"mosp_batch_create failed: {last_error()}")
        _track(self)
        self.configure(workgroups, lane_bits, delta_scale, xcd_aware)

    def configure(self, workgroups: int = 0, lane_bits: int = -1, delta_scale: float = 1.0,
                  xcd_aware: bool = True, latency_items: float | None = None,
                  threads: int = 1024, queue_cap: int = 0) -> None:
        """Launch shape: workgroups (0: every co-resident one of the shape),
        lanes per frontier vertex (2^lane_bits), width scale, XCD placement,
        group-sizing model, threads per workgroup (1024: one per CU; 512: two)
        and the push loop's LDS queue entries (256..4096, 0: all)."""
        self.config = {"workgroups": workgroups, "lane_bits": lane_bits,
                       "delta_scale": delta_scale, "xcd_aware": bool(xcd_aware),
                       "threads": threads, "queue_cap": queue_cap}
        _check(lib().mosp_batch_set_shape(self.handle, int(threads), int(queue_cap)),
               "mosp_batch_set_shape")
        _check(lib().mosp_batch_configure(self.handle, workgroups, lane_bits, delta_scale,
                                          int(bool(xcd_aware))), "mosp_batch_configure")
        if latency_items is not None:
            self.config["latency_items"] = latency_items
            _check(lib().mosp_batch_set_cost_model(self.handle, float(latency_items)),
                   "mosp_batch_set_cost_model")

    def set_barrier_timeout(self, ms: float) -> None:
        """Longest grid-barrier wait (ms) before a run gives up with an error."""
        _check(lib().mosp_batch_set_barrier_timeout(self.handle, float(ms)), "barrier timeout")

    def run(self, use_graph: bool = True) -> float:
        """One pass over every instance; returns its device time in ms."""
        ms = C.c_float(0)
        _check(lib().mosp_batch_run(self.handle, int(use_graph), C.byref(ms)), "mosp_batch_run")
        return ms.value

    def check(self) -> None:
        _check(lib().mosp_batch_check(self.handle), "mosp_batch_check")

    def schedule(self) -> tuple[list[int], int]:
        sizes = (C.c_int32 * len(self.sessions))()
        mode = C.c_int32(0)
        _check(lib().mosp_batch_schedule(self.handle, sizes, C.byref(mode)), "schedule")
        return list(sizes), mode.value

    def results(self, i: int) -> dict:
        s = self.sessions[i]
        st = (SospStats * s.K)()
        cs = CombineStats()
        _check(lib().mosp_batch_stats(self.handle, i, st, C.byref(cs)), "mosp_batch_stats")
        n, K = s.n, s.K
        return {"dist": s.dist.to_numpy(np.int64, K * n).reshape(K, n),
                "parents": s.par.to_numpy(np.int32, K * n).reshape(K, n),
                "comb_dist": s.cdist.to_numpy(np.int64, n),
                "comb_parent": s.cpar.to_numpy(np.int32, n),
                "stats": [st[k].as_dict() for k in range(K)],
                "combined_edges": cs.n_edges,
                "sosp_phase_ms": cs.sosp_phase_ms, "combined_phase_ms": cs.combined_phase_ms,
                "combined_iterations": cs.search.iterations}

    def close(self) -> None:
        if self.handle:
            lib().mosp_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class InstanceSpec:
    width: int
    height: int
    K: int
    changes: int
    pref: tuple
    seed: int

    @property
    def n(self) -> int:
        return self.width * self.height


def road_specs(count: int = 64, n_min: int = 10_000, n_max: int = 500_000, k_min: int = 2,
               k_max: int = 8, ch_min: int = 1_000, ch_max: int = 50_000,
               seed: int = 4) -> list[InstanceSpec]:
    """The uneven batched mix: n log-uniform in [n_min, n_max] (square road
    grids), K uniform in {k_min..k_max}, changes log-uniform in
    [ch_min, ch_max], Pref random ints in [1, 4]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        side = max(2, int(round(math.sqrt(n))))
        K = int(rng.integers(k_min, k_max + 1))
        ch = int(round(math.exp(rng.uniform(math.log(ch_min), math.log(ch_max)))))
        pref = tuple(int(x) for x in rng.integers(1, 5, size=K))
        out.append(InstanceSpec(side, side, K, ch, pref, 1000 + 7 * i))
    return out


def build_instance(spec: InstanceSpec, ins_pct: float = 40.0, safe: bool = True,
                   delta_scale: float = 1.0, grid: int = 0):
    """Graph, batch and initial trees of one instance -> (original, batch,
    dist, parents) on the host, plus its resident Session."""
    g = gen_road_grid(spec.width, spec.height, spec.K, spec.seed)
    b = gen_batch(g, spec.changes, ins_pct, spec.seed + 1, safe=safe)
    dist, par = initial_trees(g, spec.K)
    s = Session(g, b, dist, par, K=spec.K, source=0, pref=spec.pref, grid=grid,
                delta_scale=delta_scale)
    return (g, b, dist, par), s
