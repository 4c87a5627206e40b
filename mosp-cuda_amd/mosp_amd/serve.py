"""Served updates: resident graphs, HBM-budget admission, continuous batching.

A thin wrapper over the library's native server (C ABI `mosp_server_*`,
csrc/host/serve.cpp), which does the scheduling in C++ worker threads:

* **Resident graphs** (`Server.add_graph`): the graph in HBM (a current and a
  spare device graph), its K current trees and buffers for the next ones, and
  the weight statistics the reference takes from the original graph of each
  update (max weight for the packing bound, m and the weight sum for the
  near-far width; src/mospUpdate.cu:104-122, 177-181), recomputed on the
  device after every apply.
* **Requests on one graph keep their order**: one is on the GPU at a time,
  so a request sees exactly the state the previous request on its graph left
  and every result is independent of the mix.
* **Slots**: `slots` worker threads, each with a reusable batched workspace
  (no allocation per request), its stream and a device arena for change
  lists.  A free slot takes the oldest waiting requests of idle graphs, as
  many as fit its share of the HBM budget and `max_coalesce`, applies each
  batch to its graph in HBM, and runs them in ONE batched launch.
* **CUs** (`cu_policy`): "static" gives every slot a fixed 1/slots of the
  CUs; "greedy" lets a launch take the idle CUs, split evenly over the
  launches that could start at that moment (at least `min_workgroups`, else
  the requests wait), so a lone request runs on the whole GPU.

Latency = completion - submission; throughput = changes / wall time.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import Batch, HostGraph, _check, _i32, _nz, _track, lib, last_error

_POLICIES = {"static": 0, "greedy": 1}


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_coalesce", C.c_int32), ("max_batch", C.c_int32),
                ("cu_policy", C.c_int32), ("min_workgroups", C.c_int32),
                ("keep_results", C.c_int32), ("delta_scale", C.c_double),
                ("hbm_budget", C.c_int64)]


class RequestInfo(C.Structure):
    _fields_ = [("graph", C.c_int32), ("slot", C.c_int32), ("coalesced", C.c_int32),
                ("status", C.c_int32), ("changes", C.c_int64), ("arrival", C.c_double),
                ("admitted", C.c_double), ("done", C.c_double)]


class ServerCounters(C.Structure):
    _fields_ = [("slots", C.c_int32), ("cus", C.c_int32), ("launches", C.c_int64),
                ("kernel_ms", C.c_double), ("apply_s", C.c_double), ("launch_s", C.c_double),
                ("post_s", C.c_double), ("hbm_budget", C.c_int64),
                ("resident_bytes", C.c_int64), ("slot_bytes", C.c_int64)]


@dataclass
class Request:
    rid: int
    gid: int
    batch: Batch
    arrival: float = 0.0     # target: seconds after the stream starts (play)
    admitted: float = math.nan
    done: float = math.nan
    slot: int = -1
    coalesced: int = 0       # requests in the same launch
    result: dict | None = None
    handle: int = -1         # the server's request id
    submitted: float = math.nan  # server clock at submission

    @property
    def changes(self) -> int:
        return self.batch.n_ins + self.batch.n_del

    @property
    def latency_ms(self) -> float:
        return (self.done - self.submitted) * 1e3


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


def footprint(n: int, K: int) -> int:
    """Slot scratch bytes one update of a graph of n vertices and K objectives needs."""
    return int(lib().mosp_batch_instance_bytes(n, K))


class Server:
    """Continuous batching of update requests over resident graphs."""

    def __init__(self, slots: int = 4, max_coalesce: int = 4, max_batch: int = 100_000,
                 cu_policy: str = "static", min_workgroups: int = 0,
                 keep_results: bool = False, delta_scale: float = 1.0,
                 hbm_budget: int | None = None):
        if cu_policy not in _POLICIES:
            raise ValueError(f"unknown CU policy {cu_policy!r}")
        cfg = ServerConfig(slots, max_coalesce, max_batch, _POLICIES[cu_policy], min_workgroups,
                           int(keep_results), float(delta_scale), int(hbm_budget or 0))
        self.handle = lib().mosp_server_create(C.byref(cfg))
        if not self.handle:
            raise RuntimeError(f"mosp_server_create: {last_error()}")
        _track(self)
        self.slots = slots
        self.keep_results = keep_results
        self.graphs: list[tuple[int, int]] = []   # (n, K) per graph id
        self.started = False

    # -------------------------------------------------------------- setup --
    def add_graph(self, graph: HostGraph, dist: np.ndarray, parents: np.ndarray, K: int = 0,
                  pref=(), source: int = 0) -> int:
        K = K or graph.K
        d = np.ascontiguousarray(np.asarray(dist)[:K].reshape(-1), np.int64)
        p = np.ascontiguousarray(np.asarray(parents)[:K].reshape(-1), np.int32)
        pr = _i32(pref)
        gid = lib().mosp_server_add_graph(self.handle, graph.n, graph.K, graph.row, _nz(graph.col),
                                          _nz(graph.w), K, d, p, _nz(pr), len(pr), source)
        if gid < 0:
            raise RuntimeError(f"mosp_server_add_graph: {last_error()}")
        self.graphs.append((graph.n, K))
        return gid

    def start(self) -> None:
        _check(lib().mosp_server_start(self.handle), "mosp_server_start")
        self.started = True

    def now(self) -> float:
        return float(lib().mosp_server_clock(self.handle))

    # ------------------------------------------------------------ requests --
    def submit(self, req: Request) -> None:
        b = req.batch
        h = lib().mosp_server_submit(self.handle, req.gid, b.n_ins, _nz(_i32(b.ins_from)),
                                     _nz(_i32(b.ins_to)), _nz(_i32(b.ins_w)), b.n_del,
                                     _nz(_i32(b.del_from)), _nz(_i32(b.del_to)))
        if h < 0:
            raise RuntimeError(f"mosp_server_submit: {last_error()}")
        req.handle = int(h)

    def wait(self, req: Request, timeout: float = 600.0) -> None:
        _check(lib().mosp_server_wait(self.handle, req.handle, timeout), "mosp_server_wait")
        self.collect(req)

    def drain(self, timeout: float = 600.0) -> None:
        _check(lib().mosp_server_drain(self.handle, timeout), "mosp_server_drain")

    def collect(self, req: Request) -> None:
        """Times, slot and coalescing of a finished request (and its trees on
        keep_results servers); then the server forgets it."""
        info = RequestInfo()
        _check(lib().mosp_server_request(self.handle, req.handle, C.byref(info)), "request")
        req.submitted, req.admitted, req.done = info.arrival, info.admitted, info.done
        req.slot, req.coalesced = info.slot, info.coalesced
        if self.keep_results and info.status == 1:
            n, K = self.graphs[req.gid]
            r = {"dist": np.zeros(K * n, np.int64), "parents": np.zeros(K * n, np.int32),
                 "comb_dist": np.zeros(n, np.int64), "comb_parent": np.zeros(n, np.int32)}
            _check(lib().mosp_server_request_trees(self.handle, req.handle, r["dist"],
                                                   r["parents"], r["comb_dist"],
                                                   r["comb_parent"]), "request trees")
            r["dist"] = r["dist"].reshape(K, n)
            r["parents"] = r["parents"].reshape(K, n)
            req.result = r
        _check(lib().mosp_server_release(self.handle, req.handle), "release")

    def counters(self) -> tuple[ServerCounters, list]:
        st = ServerCounters()
        per = (C.c_double * self.slots)()
        _check(lib().mosp_server_stats(self.handle, C.byref(st), per, self.slots), "stats")
        return st, list(per)

    def reset_stats(self) -> None:
        _check(lib().mosp_server_reset_stats(self.handle), "reset stats")

    def play(self, requests: list[Request], timeout: float = 600.0) -> ServeStats:
        """Submit `requests` at their arrival times (seconds from now) and
        serve them; returns latency and throughput figures over them."""
        if not self.started:
            self.start()
        self.reset_stats()
        t0 = time.perf_counter()
        for r in sorted(requests, key=lambda r: (r.arrival, r.rid)):
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            self.submit(r)
        self.drain(timeout)
        for r in requests:
            self.collect(r)
        st, per = self.counters()
        return summarize(requests, st, per)

    def close(self) -> None:
        if self.handle:
            lib().mosp_server_destroy(self.handle)
            self.handle = None


def summarize(requests: list[Request], st: ServerCounters | None = None,
              slot_kernel_ms=()) -> ServeStats:
    lat = np.array([r.latency_ms for r in requests])
    service = np.array([(r.done - r.admitted) * 1e3 for r in requests])
    wall = max(r.done for r in requests) - min(r.submitted for r in requests)
    changes = sum(r.changes for r in requests)
    phases = {}
    if st is not None:
        phases = {"apply": st.apply_s, "launch": st.launch_s, "post": st.post_s}
    return ServeStats(len(requests), changes, wall, float(np.percentile(lat, 50)),
                      float(np.percentile(lat, 99)), float(lat.mean()), float(lat.max()),
                      changes / wall if wall > 0 else float("nan"),
                      float(np.mean([r.coalesced for r in requests])),
                      float(np.percentile(service, 50)), float(np.percentile(service, 99)),
                      list(slot_kernel_ms), phases, int(st.launches) if st is not None else 0)


def poisson_requests(originals: list[HostGraph], count: int, rate: float, seed: int,
                     size_min: int = 100, size_max: int = 100_000, ins_pct: float = 50.0,
                     disconnecting_pct: float = 10.0, first_rid: int = 0) -> list[Request]:
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
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        gid = int(rng.integers(0, len(originals)))
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
        out.append(Request(first_rid + i, gid, b, arrival=t))
    return out
