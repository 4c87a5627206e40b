"""Served updates: resident graphs, HBM-budget admission, continuous batching.

A `Server` keeps every graph and its K trees resident in HBM and takes update
requests (a change batch against one graph) as they arrive:

* **Resident state** (`ResidentGraph`): the graph in HBM (a current and a
  spare device graph), its current trees and the buffers for the next ones,
  and the weight statistics the reference takes from the original graph of
  each update (max weight for the packing bound, m and the weight sum for the
  near-far width; src/mospUpdate.cu:104-122, 177-181).
* **Requests on one graph keep their order**: one is on the GPU at a time.
  That is what makes every result independent of the mix: a request sees
  exactly the state the previous request on its graph left.
* **Slots**: `slots` workers, each with its own stream, a reusable batched
  workspace (`mosp_batch_create_capacity`, no allocation per request) and a
  device arena for change lists.  A free slot takes the oldest waiting
  requests of idle graphs, as many as fit its scratch (the HBM budget is
  split between the resident state and the slots' scratch) and
  `max_coalesce`, and runs them in ONE batched launch.  Slots run
  concurrently, so requests that arrive while others are on the GPU start as
  soon as a slot, the memory and CUs are free.
* **CUs** (`cu_policy`): "static" gives every slot a fixed 1/slots of the
  CUs; "greedy" lets a launch take the idle CUs, split evenly over the
  launches that could start at that moment (at least `min_workgroups`, else
  the requests wait), so a lone request runs on the whole GPU.  Either way
  the workgroups of concurrent launches sum to at most one per CU, which
  keeps every persistent workgroup co-resident.
* **Per request**: the batch is applied to the graph in HBM
  (`mosp_device_graph_apply`, into the spare graph), the changed lists are
  built from the returned weight-increase mask, the update (K objectives +
  combined graph + its search) runs with the other admitted requests, then
  the graph's current/spare graphs and trees swap.

Latency = completion - arrival; throughput = changes / wall time.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import (BatchInstance, Changes, DeviceBuffer, DeviceGraph, HostGraph, Batch, _check, _i32,
               default_delta, lib, last_error)


@dataclass
class Request:
    rid: int
    gid: int
    batch: Batch
    arrival: float = 0.0     # seconds after the server clock started
    admitted: float = math.nan
    done: float = math.nan
    slot: int = -1
    coalesced: int = 0       # requests in the same launch
    result: dict | None = None

    @property
    def changes(self) -> int:
        return self.batch.n_ins + self.batch.n_del

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3


class ResidentGraph:
    """A graph and its K trees kept in HBM between requests."""

    def __init__(self, gid: int, graph: HostGraph, dist: np.ndarray, parents: np.ndarray,
                 K: int = 0, pref=(), source: int = 0):
        self.gid, self.n, self.K = gid, graph.n, K or graph.K
        self.pref = tuple(int(x) for x in pref)
        self.source = source
        self.graphs = [DeviceGraph(graph), DeviceGraph()]
        self.cur = 0
        kn = self.K * self.n
        self.dist = [DeviceBuffer(8 * kn), DeviceBuffer(8 * kn)]
        self.par = [DeviceBuffer(4 * kn), DeviceBuffer(4 * kn)]
        d = np.ascontiguousarray(np.asarray(dist)[: self.K].reshape(-1), np.int64)
        p = np.ascontiguousarray(np.asarray(parents)[: self.K].reshape(-1), np.int32)
        _check(lib().mosp_memcpy(self.dist[0].ptr, d.ctypes.data, d.nbytes, 1), "trees")
        _check(lib().mosp_memcpy(self.par[0].ptr, p.ctypes.data, p.nbytes, 1), "trees")
        self.cdist = DeviceBuffer(8 * self.n)
        self.cpar = DeviceBuffer(4 * self.n)
        self.m = graph.m
        self.maxw, self.wsum = self.graphs[0].weight_stats()
        self.queue: deque[Request] = deque()
        self.busy = False
        self.updates = 0

    def footprint(self) -> int:
        """Scratch bytes one update of this graph needs in a slot."""
        return int(lib().mosp_batch_instance_bytes(self.n, self.K))

    def resident_bytes(self) -> int:
        g = self.graphs[0]
        edges = int(g.m * 1.125) + 4096
        graph = 2 * (4 * (self.n + 1) + 4 * edges * (1 + g.K))
        trees = 2 * 12 * self.K * self.n + 12 * self.n
        return 2 * graph + trees

    def trees(self) -> dict:
        n, K = self.n, self.K
        return {"dist": self.dist[self.cur].to_numpy(np.int64, K * n).reshape(K, n),
                "parents": self.par[self.cur].to_numpy(np.int32, K * n).reshape(K, n),
                "comb_dist": self.cdist.to_numpy(np.int64, n),
                "comb_parent": self.cpar.to_numpy(np.int32, n)}

    def close(self) -> None:
        for g in self.graphs:
            g.close()


class Slot:
    """A worker: its own stream, reusable batched workspace and change arena."""

    def __init__(self, index: int, workgroups: int, slab_bytes: int, max_instances: int,
                 max_objectives: int, arena_ints: int, delta_scale: float, launch: int):
        self.index = index
        self.handle = lib().mosp_batch_create_capacity(max_instances, max_objectives, slab_bytes)
        if not self.handle:
            raise RuntimeError(f"slot {index}: {last_error()}")
        _check(lib().mosp_batch_configure(self.handle, workgroups, -1, 1.0, 1), "configure")
        self.workgroups = workgroups
        self.stream = lib().mosp_batch_stream(self.handle)
        self.arena = DeviceBuffer(4 * max(arena_ints, 1))
        self.arena_ints = arena_ints
        self.delta_scale = delta_scale
        self.launch = launch
        self.kernel_ms = 0.0
        # Wall seconds per phase: host preparation + graph application,
        # the batched launch (reset, run, check), statistics after it.
        self.phase_s = {"apply": 0.0, "launch": 0.0, "post": 0.0}
        self.launches = 0
        self.busy = False

    def run(self, reqs: list[Request], graphs: list[ResidentGraph], keep_results: bool,
            workgroups: int = 0) -> None:
        if workgroups and workgroups != self.workgroups:
            _check(lib().mosp_batch_configure(self.handle, workgroups, -1, 1.0, 1), "configure")
            self.workgroups = workgroups
        keep = []
        specs = (BatchInstance * len(reqs))()
        offset = 0
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            rg = graphs[r.gid]
            cur, spare = rg.graphs[rg.cur], rg.graphs[1 - rg.cur]
            b = r.batch
            mask = cur.apply(b, spare, stream=self.stream)
            K, n = rg.K, rg.n
            ins_w = _i32(b.ins_w).reshape(-1, cur.K)[:, :K] if b.n_ins else np.zeros((0, K), np.int32)
            maxw = [max(int(rg.maxw[k]), int(ins_w[:, k].max()) if b.n_ins else 1, 1)
                    for k in range(K)]
            delta = [max(1, int(default_delta(max(rg.m, 1), n, int(rg.wsum[k])) * self.delta_scale))
                     for k in range(K)]
            # Changed lists: every deletion + insertions that raised objective k.
            heads = _i32(b.ins_to)
            lists = []
            for k in range(K):
                sel = ((mask.astype(np.int64) >> k) & 1) == 1
                lists.append((np.concatenate([_i32(b.del_from), _i32(b.ins_from)[sel]]),
                              np.concatenate([_i32(b.del_to), _i32(b.ins_to)[sel]])))
            host = np.concatenate([heads] + [np.concatenate(x) for x in lists]).astype(np.int32)
            if offset + host.size > self.arena_ints:
                raise RuntimeError("change arena too small for the admitted requests")
            base = self.arena.ptr + 4 * offset
            if host.size:
                _check(lib().mosp_memcpy(base, host.ctypes.data, host.nbytes, 1), "changes")
            changes = (Changes * K)()
            at = base + 4 * heads.size
            for k, (f, t) in enumerate(lists):
                changes[k] = Changes(at, at + 4 * f.size, f.size, base, heads.size)
                at += 4 * (f.size + t.size)
            offset += host.size
            dl = (C.c_int64 * K)(*delta)
            mw = (C.c_int64 * K)(*maxw)
            pr = (C.c_int32 * max(1, len(rg.pref)))(*rg.pref)
            keep += [changes, dl, mw, pr]
            nxt = 1 - rg.cur
            specs[i] = BatchInstance(spare.handle, K, C.cast(changes, C.POINTER(Changes)),
                                     rg.source, dl, mw, pr, len(rg.pref), rg.dist[rg.cur].ptr,
                                     rg.par[rg.cur].ptr, rg.dist[nxt].ptr, rg.par[nxt].ptr,
                                     rg.cdist.ptr, rg.cpar.ptr)
        t1 = time.perf_counter()
        _check(lib().mosp_batch_reset(self.handle, specs, len(reqs)), "mosp_batch_reset")
        ms = C.c_float(0)
        _check(lib().mosp_batch_run(self.handle, self.launch, C.byref(ms)), "mosp_batch_run")
        _check(lib().mosp_batch_check(self.handle), "mosp_batch_check")
        self.kernel_ms += ms.value
        self.launches += 1
        t2 = time.perf_counter()
        for r in reqs:
            rg = graphs[r.gid]
            rg.cur = 1 - rg.cur
            new = rg.graphs[rg.cur]
            rg.m = int(lib().mosp_device_graph_edges(new.handle))
            rg.maxw, rg.wsum = new.weight_stats(self.stream)
            rg.updates += 1
            r.coalesced = len(reqs)
            if keep_results:
                r.result = rg.trees()
        del keep
        t3 = time.perf_counter()
        self.phase_s["apply"] += t1 - t0
        self.phase_s["launch"] += t2 - t1
        self.phase_s["post"] += t3 - t2

    def close(self) -> None:
        if self.handle:
            lib().mosp_batch_destroy(self.handle)
            self.handle = None


@dataclass
class ServeStats:
    requests: int
    changes: int
    wall_s: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    max_ms: float
    changes_per_s: float
    mean_coalesced: float
    service_p50_ms: float = math.nan   # admission to completion
    service_p99_ms: float = math.nan
    slot_kernel_ms: list = field(default_factory=list)
    phase_s: dict = field(default_factory=dict)   # summed over slots
    launches: int = 0


class Server:
    """Continuous batching of update requests over resident graphs."""

    def __init__(self, graphs: list[ResidentGraph], slots: int = 4, slot_workgroups: int = 0,
                 hbm_budget: int | None = None, max_coalesce: int = 4, max_batch: int = 100_000,
                 delta_scale: float = 1.0, launch: int = 2, keep_results: bool = False,
                 cu_policy: str = "static", min_workgroups: int = 0):
        if cu_policy not in ("static", "greedy"):
            raise ValueError(f"unknown CU policy {cu_policy!r}")
        self.graphs = graphs
        cus = int(lib().mosp_device_cus()) or 256
        self.cus = cus
        self.cu_policy = cu_policy
        self.slot_workgroups = slot_workgroups or max(1, cus // slots)
        self.min_workgroups = min_workgroups or max(1, cus // (2 * slots))
        self.free_cus = cus
        foot = max(g.footprint() for g in graphs)
        resident = sum(g.resident_bytes() for g in graphs)
        if hbm_budget is None:
            hbm_budget = resident + slots * foot * max_coalesce
        self.hbm_budget = hbm_budget
        self.resident = resident
        per_slot = (hbm_budget - resident) // slots
        if per_slot < foot:
            raise RuntimeError(f"HBM budget {hbm_budget} leaves {per_slot} B per slot; the largest "
                               f"update needs {foot} B")
        self.slab_bytes = per_slot
        self.max_coalesce = max_coalesce
        kmax = max(g.K for g in graphs)
        arena = max_coalesce * (kmax * 2 * 2 * max_batch + max_batch)
        self.slots = [Slot(i, self.slot_workgroups, per_slot, max_coalesce, kmax * max_coalesce,
                           arena, delta_scale, launch) for i in range(slots)]
        self.keep_results = keep_results
        self.cv = threading.Condition()
        self.pending = 0
        self.stop = False
        self.t0 = time.perf_counter()
        self.errors: list[BaseException] = []
        self.threads = [threading.Thread(target=self._worker, args=(s,), daemon=True)
                        for s in self.slots]

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def start(self) -> None:
        self.t0 = time.perf_counter()
        for t in self.threads:
            t.start()

    def submit(self, req: Request) -> None:
        with self.cv:
            if math.isnan(req.arrival) or req.arrival <= 0:
                req.arrival = self.now()
            self.graphs[req.gid].queue.append(req)
            self.pending += 1
            self.cv.notify_all()

    def _admit(self) -> tuple[list[Request], int]:
        """Oldest waiting requests of idle graphs that fit one slot, and the
        workgroups their launch gets (0 = the slot's fixed share)."""
        heads = [g.queue[0] for g in self.graphs if g.queue and not g.busy]
        if not heads:
            return [], 0
        workgroups = 0
        if self.cu_policy == "greedy":
            idle_slots = sum(1 for s in self.slots if not s.busy)
            launches = max(1, min(idle_slots, -(-len(heads) // self.max_coalesce)))
            workgroups = self.free_cus // launches
            if workgroups < self.min_workgroups:
                return [], 0
        heads.sort(key=lambda r: (r.arrival, r.rid))
        chosen, used = [], 0
        for r in heads:
            need = self.graphs[r.gid].footprint()
            if used + need > self.slab_bytes or len(chosen) >= self.max_coalesce:
                continue
            chosen.append(r)
            used += need
        for r in chosen:
            g = self.graphs[r.gid]
            g.queue.popleft()
            g.busy = True
        self.free_cus -= workgroups
        return chosen, workgroups

    def _worker(self, slot: Slot) -> None:
        while True:
            with self.cv:
                while not self.stop:
                    reqs, workgroups = self._admit()
                    if reqs:
                        break
                    self.cv.wait()
                else:
                    return
                slot.busy = True
                t = self.now()
                for r in reqs:
                    r.admitted = t
                    r.slot = slot.index
            try:
                slot.run(reqs, self.graphs, self.keep_results, workgroups)
            except BaseException as exc:  # surfaced by drain()
                with self.cv:
                    self.errors.append(exc)
                    self.stop = True
                    self.cv.notify_all()
                return
            with self.cv:
                t = self.now()
                for r in reqs:
                    r.done = t
                    self.graphs[r.gid].busy = False
                self.pending -= len(reqs)
                self.free_cus += workgroups
                slot.busy = False
                self.cv.notify_all()

    def drain(self, timeout: float = 600.0) -> None:
        """Wait until every submitted request is done, then stop the workers."""
        end = time.perf_counter() + timeout
        with self.cv:
            while self.pending > 0 and not self.errors:
                left = end - time.perf_counter()
                if left <= 0:
                    raise TimeoutError(f"{self.pending} requests still pending")
                self.cv.wait(timeout=min(left, 1.0))
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join(timeout=60)
        if self.errors:
            raise RuntimeError(f"server worker failed: {self.errors[0]}") from self.errors[0]

    def close(self) -> None:
        for s in self.slots:
            s.close()

    def play(self, requests: list[Request], timeout: float = 600.0) -> ServeStats:
        """Submit `requests` at their arrival times (seconds from start) and
        serve them; returns latency and throughput figures."""
        self.start()
        for r in sorted(requests, key=lambda r: r.arrival):
            wait = r.arrival - self.now()
            if wait > 0:
                time.sleep(wait)
            self.submit(r)
        self.drain(timeout)
        return summarize(requests, self.slots)


def summarize(requests: list[Request], slots=()) -> ServeStats:
    lat = np.array([r.latency_ms for r in requests])
    service = np.array([(r.done - r.admitted) * 1e3 for r in requests])
    wall = max(r.done for r in requests) - min(r.arrival for r in requests)
    changes = sum(r.changes for r in requests)
    phases = {}
    for s in slots:
        for k, v in s.phase_s.items():
            phases[k] = phases.get(k, 0.0) + v
    return ServeStats(len(requests), changes, wall, float(np.percentile(lat, 50)),
                      float(np.percentile(lat, 99)), float(lat.mean()), float(lat.max()),
                      changes / wall if wall > 0 else float("nan"),
                      float(np.mean([r.coalesced for r in requests])),
                      float(np.percentile(service, 50)), float(np.percentile(service, 99)),
                      [s.kernel_ms for s in slots], phases, sum(s.launches for s in slots))


def poisson_requests(graphs: list[ResidentGraph], count: int, rate: float, seed: int,
                     size_min: int = 100, size_max: int = 100_000, ins_pct: float = 50.0,
                     disconnecting_pct: float = 10.0, originals: list[HostGraph] = ()
                     ) -> list[Request]:
    """Poisson arrivals at `rate` per second over the graphs (uniformly), batch
    sizes log-uniform in [size_min, size_max]: ins_pct insertions of random
    pairs u != v with weights U[1, 100]; the rest deletions of arcs of the
    graph as loaded (`originals`), disconnecting_pct of all changes drawn
    anywhere in the graph (they may disconnect vertices), the others from
    arcs whose head has another in-arc (so they do not cut that head off
    directly).  A deletion of an arc an earlier request already removed is a
    no-op, as in applyChangeBatch."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    arcs = {}
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        gid = int(rng.integers(0, len(graphs)))
        g = originals[gid]
        size = int(round(math.exp(rng.uniform(math.log(size_min), math.log(size_max)))))
        n_ins = int(round(size * ins_pct / 100.0))
        n_cut = int(round(size * disconnecting_pct / 100.0))
        n_del = max(size - n_ins - n_cut, 0)
        fr = rng.integers(0, g.n, n_ins)
        to = rng.integers(0, g.n, n_ins)
        to = np.where(to == fr, (to + 1) % g.n, to)
        w = rng.integers(1, 101, size=(n_ins, g.K))
        if gid not in arcs:
            indeg = np.bincount(g.col, minlength=g.n)
            arcs[gid] = (np.repeat(np.arange(g.n, dtype=np.int32), np.diff(g.row)),
                         np.nonzero(indeg[g.col] > 1)[0])
        src, safe_arcs = arcs[gid]
        picks = [rng.integers(0, g.m, n_cut)]
        if safe_arcs.size:
            picks.append(safe_arcs[rng.integers(0, safe_arcs.size, n_del)])
        pick = np.concatenate(picks).astype(np.int64)
        b = Batch(g.K, _i32(fr), _i32(to), _i32(w.reshape(-1)), _i32(src[pick]), _i32(g.col[pick]))
        out.append(Request(rid, gid, b, arrival=t))
    return out
