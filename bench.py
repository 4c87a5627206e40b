"""Benchmark of the DynaMOSP update hot path on one MI355X.

Workload (default): the road-like 1400x1400 grid (1.96M vertices, ~5.56M
arcs: a random spanning tree kept, 30% of the other grid edges dropped, 2%
diagonal shortcuts, both directions), K=3 int32 weights U[1,100], a batch of
50,000 connectivity-safe changes (50% insertions), Pref=(1,1,1), source 0,
all seeded.  The graph, the old trees, the batch and the outputs are resident
in HBM before timing; one step = one MOSP update (K SOSP updates + combined
graph + its SOSP search) in one persistent cooperative launch, host wall
clock around the synchronous call (the reference's gpu_compute_ms scope).
The timed region is bracketed by device synchronisations through the
library (hipDeviceSynchronize); torch is not used, so the process holds one
HIP runtime.

Prints ONE JSON line.  `roofline` uses the algorithmic bytes of one update
(mosp_amd/roofline.py) over the HIP-event duration of the persistent kernel
(recorded by the library on the launch stream).  `cpu_baseline` times the
oracle's restatement of the reference's CPU path (sequential SOSP update per
objective + host combined graph + Dijkstra) on one host core over the same
workload, and its outputs are compared with the GPU's.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import platform
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mosp-cuda_amd"))

PEAK_HBM_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cpus() -> dict:
    """The host's CPU as the process sees it before any pinning."""
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    sockets, cores = set(), set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":", 1)[1].strip()
            elif line.startswith("core id"):
                core = line.split(":", 1)[1].strip()
                sockets.add(phys)
                cores.add((phys, core))
    except OSError:
        pass
    return {"cpu": cpu_model(), "logical_cpus": os.cpu_count(), "cpus_allowed": allowed,
            "physical_cores": len(cores) or None, "sockets": len(sockets) or None}


def copy_bandwidth(M, gib: int = 1, reps: int = 10) -> float:
    """Device-to-device copy of `gib` GiB (hipMemcpy), GB/s of HBM traffic
    (bytes read + bytes written over the wall time of `reps` copies)."""
    nbytes = gib << 30
    a, b = M.DeviceBuffer(nbytes), M.DeviceBuffer(nbytes)
    L = M.lib()
    L.mosp_memcpy(b.ptr, a.ptr, nbytes, 3)
    M.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.mosp_memcpy(b.ptr, a.ptr, nbytes, 3)
    M.synchronize()
    dt = time.perf_counter() - t0
    a.free()
    b.free()
    return 2 * nbytes * reps / dt / 1e9


def cpu_baseline(g, b, d0, p0, gpu_result) -> dict:
    """The oracle (CPU restatement of the reference's sequential path) on one
    pinned core, over the same workload; also checks the GPU outputs."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O

    host = host_cpus()
    saved = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    og = O.Graph(g.n, g.K, g.row, g.col, g.w)
    ob = O.Batch(g.K, b.ins_from, b.ins_to, b.ins_w, b.del_from, b.del_to)
    t0 = time.perf_counter()
    r = O.mosp_update(og, ob, d0, p0, g.K, 0, [])
    wall = (time.perf_counter() - t0) * 1e3
    compute_ms = float(r["times_ms"][1] + r["times_ms"][2])  # updates + combined
    same = all(np.array_equal(gpu_result[k], r[k]) for k in
               ("dist", "parents", "comb_dist", "comb_parent"))
    if saved is not None:
        os.sched_setaffinity(0, saved)
    return {"value": (b.n_ins + b.n_del) / (compute_ms / 1e3),
            "unit": "changes/s", "cores": 1, "kind": "port",
            "sample": ("full workload once: K=3 sequential SOSP updates (Phases 1-2) + host "
                       "combined graph + Dijkstra, apply excluded "
                       f"(apply {r['times_ms'][0]:.1f} ms, updates {r['times_ms'][1]:.1f} ms, "
                       f"combined {r['times_ms'][2]:.1f} ms, wall {wall:.0f} ms)"),
            "compute_ms": compute_ms, **host,
            "pinned_to": "one core (sched_setaffinity) for the single-threaded reference path",
            "outputs_equal_gpu": bool(same)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1400)
    ap.add_argument("--height", type=int, default=1400)
    ap.add_argument("--changes", type=int, default=50_000)
    ap.add_argument("--grid", type=int, default=0, help="workgroups of the persistent launch")
    ap.add_argument("--delta-scale", type=float, default=1.0,
                    help="multiply the default near-far width (tuning; the asynchronous "
                         "loop measured best at 1, the iteration-synchronous one at 3)")
    ap.add_argument("--comb-delta-scale", type=int, default=10,
                    help="multiply the combined search's default near-far width (tuning)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "traffic_road1400_k3.json"),
                    help="measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--mode", choices=("plain", "rmat", "batched", "served"), default="plain",
                    help="plain: the 1.96M-vertex update (the headline line); rmat: the "
                         "power-law stress update (R-MAT scale 22, K=5, weights up to 2^24); "
                         "batched: 64 uneven instances in one launch vs 64 serial launches; "
                         "served: resident graphs, Poisson requests, continuous batching")
    ap.add_argument("--scale", type=int, default=22, help="rmat: R-MAT scale")
    ap.add_argument("--requests", type=int, default=240)
    ap.add_argument("--rate", type=float, default=0.0,
                    help="served: Poisson requests per second for the latency pass (0: "
                         "--load times the capacity measured by the saturation pass)")
    ap.add_argument("--load", type=float, default=0.6,
                    help="served: offered load of the latency pass as a fraction of capacity")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--cu-policy", choices=("static", "greedy"), default="greedy",
                    help="served: fixed CU share per slot, or launches take the idle CUs")
    ap.add_argument("--coalesce", type=int, default=4)
    ap.add_argument("--hbm-budget-gb", type=float, default=0.0,
                    help="served: HBM budget (0 = resident state + room for every slot)")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--tune-cache", default=str(ROOT / "profiles" / "autotune_batched.json"),
                    help="autotuned batched launch choices (read; a workload it lacks, or --retune, is timed and its choice stored there)")
    ap.add_argument("--retune", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="plain mode: the 1.96M-vertex update alone, without the nested "
                         "batched / served / rmat summaries")
    ap.add_argument("--equal-requests", type=int, default=16,
                    help="served: requests replayed in the mix and alone to compare outputs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the update on a single MI355X "
              "(multi-GPU sharding is described, not built); use --gpus 1", file=sys.stderr)
        return 2
    if args.mode != "plain":
        fn = {"batched": batched_line, "rmat": rmat_line, "served": served_line}[args.mode]
        print(json.dumps(fn(args)), flush=True)
        return 0
    line, host = plain_line(args)
    if not args.headline_only:
        # The rest of the metric, each on its own workload: the batched and
        # served forms (speed-up over serial launches, p50 / p99 latency) and
        # the power-law stress update.  Compact passes, same code as the
        # dedicated modes.
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = 3, 1
        b = batched_line(sub)
        line["batched"] = {
            "workload": b["config"]["workload"], "instances": b["config"]["instances"],
            "changes": b["config"]["changes"], "ms_per_pass": b["ms_per_step"],
            "changes_per_s": b["value"], "serial_ms_per_pass": b["serial"]["ms_per_step"],
            "speedup_vs_serial": b["speedup_vs_serial"],
            "outputs_equal_serial": b["outputs_equal_serial"],
            "graph_mode": b["config"]["graph_mode"]}
        sub.requests = 96
        v = served_line(sub)
        line["served"] = {
            "workload": v["config"]["workload"], "p50_ms": v["p50_ms"], "p99_ms": v["p99_ms"],
            "changes_per_s": v["value"],
            "poisson_changes_per_s": v["config"]["poisson_changes_per_s"],
            "rate_per_s": v["config"]["rate_per_s"], "requests": v["config"]["requests"],
            "mean_coalesced": v["config"]["mean_coalesced"],
            "outputs_equal_alone": v["outputs_equal_alone"],
            "equal_requests": v["equal_requests"]}
        r = rmat_line(sub)
        line["rmat"] = {
            "workload": r["config"]["workload"], "ms_per_step": r["ms_per_step"],
            "changes_per_s": r["value"], "kernel_ms": r["kernel_ms"],
            "roofline_frac": r["roofline"]["frac"], "alg_bytes": r["roofline"]["alg_bytes"],
            "packed_parents": r["config"]["packed_parents"]}
    if host is not None:
        log("cpu baseline: oracle CPU path on one core (about 20 s) ...")
        line["cpu_baseline"] = cpu_baseline(*host)
    print(json.dumps(line), flush=True)
    return 0


def plain_line(args):
    """The headline: one MOSP update of the 1.96M-vertex road grid, resident."""
    import mosp_amd as M
    from mosp_amd.roofline import alg_bytes

    M.set_device(0)
    t = time.perf_counter()
    g = M.gen_road_grid(args.width, args.height, 3, seed=7)
    b = M.gen_batch(g, args.changes, 50.0, seed=8, safe=True)
    log(f"workload: n={g.n} m={g.m} K=3 inserts={b.n_ins} deletes={b.n_del} "
        f"({time.perf_counter() - t:.1f} s to generate)")
    d0, p0 = M.initial_trees(g, 3)
    sess = M.Session(g, b, d0, p0, grid=args.grid, delta_scale=args.delta_scale,
                     comb_delta_scale=args.comb_delta_scale)
    log(f"resident in HBM; persistent grid = {sess.ws.grid} workgroups x 1024")

    for _ in range(args.warmup):
        sess.run()
    M.synchronize()
    kernel_ms, sosp_ms, comb_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sess.run()
        kernel_ms.append(sess.stats[0].kernel_ms)
        sosp_ms.append(sess.cstats.sosp_phase_ms)
        comb_ms.append(sess.cstats.combined_phase_ms)
    M.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    changes = b.n_ins + b.n_del
    res = sess.results()
    log(f"{ms_per_step:.3f} ms/step; kernel {np.mean(kernel_ms):.3f} ms "
        f"(K updates {np.mean(sosp_ms):.3f} ms, combined {np.mean(comb_ms):.3f} ms); "
        f"stats {res['stats']}")
    log(f"phase profile (last step, device clock): {sess.ws.last_profile()}")
    log(f"delta {sess.delta.tolist()}; combined: edges {res['combined_edges']}, "
        f"iterations {sess.cstats.search.iterations}, delta {sess.cstats.delta}")

    # Algorithmic bytes of one update.
    changed = [M.changed_lists(b, sess.mask, k) for k in range(3)]
    in_row = np.zeros(g.n + 1, np.int64)
    np.add.at(in_row, sess.updated.col.astype(np.int64) + 1, 1)
    in_row = np.cumsum(in_row)
    ab = alg_bytes(g.n, sess.updated.row, in_row, 0, p0, d0, res["dist"], changed,
                   b.ins_to, res["combined_edges"])
    k_ms = float(np.mean(kernel_ms))
    achieved = ab / (k_ms * 1e-3) / 1e9
    traffic = None
    tp = pathlib.Path(args.traffic)
    if tp.exists():
        traffic = json.loads(tp.read_text()).get("bytes_per_launch")

    line = {
        "metric": "changes/s + gpu_compute_ms, 2M-vertex grid K=3 50k batch; %HBM peak; p99 ms",
        "value": changes / (ms_per_step / 1e3),
        "unit": "changes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "gpu_compute_ms": ms_per_step,
        "kernel_ms": k_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded road-like grid + batch, GPU-computed "
                                  "canonical initial trees)",
        "config": {"workload": f"road-like grid {args.width}x{args.height} "
                               f"(n={g.n}, m={g.m}), K=3 w~U[1,100], {changes} safe changes "
                               "(50% inserts), Pref=(1,1,1), source 0",
                   "n": g.n, "m": g.m, "K": 3, "batch": changes,
                   "grid_workgroups": sess.ws.grid,
                   "delta": [int(x) for x in sess.delta],
                   "combined_delta": int(sess.cstats.delta)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "traffic_source": f"rocprofv3 TCC_EA0_RDREQ_64B/128B + WRITE_SIZE per "
                                       f"update dispatch, {tp.name}",
                     "alg_bytes": ab, "kernel": "sosp_persistent (one launch per update)"},
    }
    sess.close()
    line["roofline"]["copy_gbs"] = copy_bandwidth(M)
    line["roofline"]["copy_frac"] = line["roofline"]["copy_gbs"] / PEAK_HBM_GBS
    log(f"device copy: {line['roofline']['copy_gbs']:.0f} GB/s")
    host = None if args.no_cpu_baseline else (g, b, d0, p0, res)
    return line, host


def rmat_line(args) -> dict:
    """Power-law stress: R-MAT scale 22 (a=0.57, b=c=0.19, edge factor 16,
    self-loops and duplicates removed), K=5 weights U[1, 2^24] -- (n - 1) *
    maxW does not fit next to a 23-bit parent id, so every objective runs the
    distance-only words and the lowest-id parent recovery -- 100,000 changes
    (50% insertions, deletions drawn from all arcs, so some disconnect
    vertices), Pref=(1,2,3,4,5).  Same step and timing as the plain mode."""
    import mosp_amd as M
    from mosp_amd.roofline import alg_bytes

    M.set_device(0)
    K, pref = 5, (1, 2, 3, 4, 5)
    t = time.perf_counter()
    g = M.gen_rmat(args.scale, 16, K, seed=31, wmin=1, wmax=2**24)
    b = M.gen_batch(g, 100_000, 50.0, seed=32, safe=False, wmax=2**24)
    deg = np.diff(g.row)
    log(f"workload: R-MAT scale {args.scale}: n={g.n} m={g.m} max out-degree {deg.max()} "
        f"K={K} inserts={b.n_ins} deletes={b.n_del} ({time.perf_counter() - t:.1f} s)")
    d0, p0 = M.initial_trees(g, K)
    log(f"initial trees ({time.perf_counter() - t:.1f} s)")
    sess = M.Session(g, b, d0, p0, pref=pref, grid=args.grid, delta_scale=args.delta_scale,
                     comb_delta_scale=args.comb_delta_scale)
    for _ in range(args.warmup):
        sess.run()
    M.synchronize()
    kernel_ms, sosp_ms, comb_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sess.run()
        kernel_ms.append(sess.stats[0].kernel_ms)
        sosp_ms.append(sess.cstats.sosp_phase_ms)
        comb_ms.append(sess.cstats.combined_phase_ms)
    M.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    res = sess.results()
    changes = b.n_ins + b.n_del
    log(f"{ms_per_step:.3f} ms/step; kernel {np.mean(kernel_ms):.3f} ms (K updates "
        f"{np.mean(sosp_ms):.3f} ms, combined {np.mean(comb_ms):.3f} ms); stats {res['stats']}")
    log(f"phase profile (last step, device clock): {sess.ws.last_profile()}")
    changed = [M.changed_lists(b, sess.mask, k) for k in range(K)]
    in_row = np.zeros(g.n + 1, np.int64)
    np.add.at(in_row, sess.updated.col.astype(np.int64) + 1, 1)
    in_row = np.cumsum(in_row)
    ab = alg_bytes(g.n, sess.updated.row, in_row, 0, p0, d0, res["dist"], changed, b.ins_to,
                   res["combined_edges"],
                   distance_only=[s["packed_parents"] == 0 for s in res["stats"]])
    k_ms = float(np.mean(kernel_ms))
    achieved = ab / (k_ms * 1e-3) / 1e9
    # Measured HBM bytes per launch (rocprofv3 read requests by size + WRITE_SIZE, scale 22).
    traffic = None
    tp = ROOT / "profiles" / "traffic_rmat22_k5.json"
    if args.scale == 22 and tp.exists():
        traffic = json.loads(tp.read_text()).get("bytes_per_launch")
    line = {
        "metric": "changes/s, power-law stress update (R-MAT scale 22, K=5, weights up to "
                  "2^24)",
        "value": changes / (ms_per_step / 1e3), "unit": "changes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "gpu_compute_ms": ms_per_step, "kernel_ms": k_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded R-MAT + batch, GPU-computed canonical "
                                  "initial trees)",
        "config": {"workload": f"R-MAT scale {args.scale} (n={g.n}, m={g.m}), K=5 "
                               f"w~U[1,2^24], {changes} changes (50% inserts, unsafe "
                               "deletions), Pref=(1,2,3,4,5), source 0",
                   "n": g.n, "m": g.m, "K": K, "batch": changes,
                   "packed_parents": [s["packed_parents"] for s in res["stats"]],
                   "max_out_degree": int(deg.max())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "alg_bytes": ab,
                     "kernel": "sosp_persistent (hub variant, one launch per update)"},
    }
    sess.close()
    return line


def batched_line(args) -> dict:
    """64 independent update instances of uneven size (road-like grids with
    n log-uniform in [10k, 500k], K uniform in {2..8}, batches log-uniform in
    [1k, 50k] with 40% insertions, Pref random in [1, 4]) in ONE persistent
    launch replayed as a HIP graph, against 64 serial plain launches of the
    same instances; outputs compared bit for bit."""
    import mosp_amd as M
    from mosp_amd import autotune
    from mosp_amd.batched import BatchedUpdate, build_instance, road_specs

    M.set_device(0)
    specs = road_specs(count=args.instances)
    t = time.perf_counter()
    built = []
    for i, spec in enumerate(specs):
        built.append(build_instance(spec))
        if i % 8 == 7:
            log(f"built {i + 1}/{len(specs)} instances ({time.perf_counter() - t:.0f} s)")
    sessions = [s for _, s in built]
    changes = sum(s.batch.n_ins + s.batch.n_del for s in sessions)
    total_n = sum(s.n for s in sessions)
    log(f"workload: {len(sessions)} instances, total n={total_n}, "
        f"K={[s.K for s in sessions]}, changes={changes}")

    bu = BatchedUpdate(sessions)
    key = autotune.signature(sessions)
    choice = None if args.retune else autotune.load(args.tune_cache, key)
    tuned = None
    if choice is None:
        tuned = autotune.tune(bu, log=log)
        choice = tuned["best"]
        autotune.save(args.tune_cache, key, choice)
    bu.configure(**choice)
    log(f"batched launch configuration: {choice}")

    for _ in range(max(args.warmup, 1)):
        bu.run(use_graph=True)
    bu.check()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kms.append(bu.run(use_graph=True))
    batched_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    sizes, mode = bu.schedule()
    batched = [bu.results(i) for i in range(len(sessions))]
    log(f"batched: {batched_ms:.3f} ms per pass (device {np.mean(kms):.3f} ms), "
        f"graph mode {mode}, group sizes {sizes}")
    slow = sorted(range(len(sessions)), key=lambda i: -(batched[i]["sosp_phase_ms"] +
                                                        batched[i]["combined_phase_ms"]))
    for i in slow[:6]:
        r, s = batched[i], sessions[i]
        log(f"  instance {i}: n={s.n} K={s.K} changes={s.batch.n_ins + s.batch.n_del} "
            f"group={sizes[i]} updates {r['sosp_phase_ms']:.2f} ms combined "
            f"{r['combined_phase_ms']:.2f} ms iterations "
            f"{max(x['iterations'] for x in r['stats'])}+{r['combined_iterations']}")

    # Serial: one plain fused launch per instance, at the batched launch's
    # width and at the plain path's tuned width (x3); the faster one counts.
    serial_runs = {}
    for scale in sorted({float(choice["delta_scale"]), 3.0}):
        for s in sessions:
            s.set_delta_scale(scale)
        for s in sessions:
            s.run()
        M.synchronize()
        kms_s = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            per = 0.0
            for s in sessions:
                s.run()
                per += s.stats[0].kernel_ms
            kms_s.append(per)
        M.synchronize()
        serial_runs[scale] = ((time.perf_counter() - t0) * 1e3 / args.steps, float(np.mean(kms_s)))
        log(f"serial (delta x{scale}): {serial_runs[scale][0]:.3f} ms per pass "
            f"(device {serial_runs[scale][1]:.3f} ms)")
    serial_scale = min(serial_runs, key=lambda k: serial_runs[k][0])
    serial_ms, serial_kms = serial_runs[serial_scale]
    same = True
    for i, s in enumerate(sessions):
        r = s.results()
        same &= all(np.array_equal(r[k], batched[i][k])
                    for k in ("dist", "parents", "comb_dist", "comb_parent"))
    log(f"serial best: delta x{serial_scale} {serial_ms:.3f} ms per pass; "
        f"outputs identical: {same}")
    line = {
        "metric": "changes/s, 64 uneven update instances in one launch (vs 64 serial launches)",
        "value": changes / (batched_ms / 1e3), "unit": "changes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": batched_ms, "kernel_ms": float(np.mean(kms)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded road-like grids + batches)",
        "config": {"workload": "64 road-like grids, n log-uniform [10k, 500k], K U{2..8}, "
                               "w~U[1,100], batches log-uniform [1k, 50k] (40% inserts, safe), "
                               "Pref random ints [1, 4]",
                   "instances": len(sessions), "total_n": total_n, "changes": changes,
                   "launch": choice, "graph_mode": mode},
        "serial": {"value": changes / (serial_ms / 1e3), "ms_per_step": serial_ms,
                   "kernel_ms": serial_kms, "delta_scale": serial_scale,
                   "all": {str(k): v[0] for k, v in serial_runs.items()}},
        "speedup_vs_serial": serial_ms / batched_ms,
        "outputs_equal_serial": bool(same),
    }
    if tuned is not None:
        line["autotune"] = {"seconds": tuned["seconds"], "trials": len(tuned["trials"])}
    bu.close()
    for s in sessions:
        s.close()
    return line


def served_line(args) -> dict:
    """8 resident graphs (4 road-like 1400x1400 grids K=3, 4 R-MAT scale 20
    K=4, weights U[1,100]); requests with batch sizes log-uniform in [100,
    100k] (50% insertions, 10% disconnecting deletions), served by the
    library's native server (continuous batching over `slots` concurrent
    launches under an HBM budget).  Two passes on the same resident graphs:
    saturation (every request queued at t=0) gives changes/s; open-loop
    Poisson arrivals at `--rate` (or `--load` x the measured capacity) give
    p50/p99 latency (submission to completion)."""
    import mosp_amd as M
    from mosp_amd import serve as S

    M.set_device(0)
    t = time.perf_counter()
    budget = int(args.hbm_budget_gb * 2**30) if args.hbm_budget_gb > 0 else None
    originals, resident = [], []
    for gid in range(8):
        if gid < 4:
            g = M.gen_road_grid(1400, 1400, 3, seed=100 + gid)
            K, pref = 3, (1, 1, 1)
        else:
            g = M.gen_rmat(20, 16, 4, seed=200 + gid)
            K, pref = 4, (1, 2, 3, 4)
        d, p = M.initial_trees(g, K)
        originals.append(g)
        resident.append((g, d, p, K, pref))
        log(f"graph {gid}: n={g.n} m={g.m} K={K} ({time.perf_counter() - t:.0f} s)")

    def server(**kw):
        srv = S.Server(max_batch=100_000, delta_scale=args.delta_scale, **kw)
        for g, d, p, K, pref in resident:
            srv.add_graph(g, d, p, K=K, pref=pref)
        srv.start()
        return srv

    srv = server(slots=args.slots, max_coalesce=args.coalesce, cu_policy=args.cu_policy,
                 hbm_budget=budget)
    st, _ = srv.counters()
    log(f"server: {args.slots} slots, {st.cus} CUs, policy {args.cu_policy}, coalesce <= "
        f"{args.coalesce}, HBM budget {st.hbm_budget / 2**30:.2f} GiB (resident "
        f"{st.resident_bytes / 2**30:.2f} GiB, {st.slot_bytes / 2**30:.2f} GiB scratch per slot)")
    # Warm-up: one small request per graph (first use of spare graphs and
    # scratch), outside the measured passes.
    warm = [S.Request(-1 - i, i, S.poisson_requests([originals[i]], 1, 1.0, seed=600 + i,
                                                     size_min=100, size_max=100)[0].batch)
            for i in range(len(originals))]
    srv.play(warm)

    def report(tag, st):
        ph = {k: round(v, 2) for k, v in st.phase_s.items()}
        log(f"{tag}: {st.requests} requests, {st.changes} changes in {st.wall_s:.2f} s = "
            f"{st.changes_per_s:.0f} changes/s; latency p50 {st.p50_ms:.1f} p99 {st.p99_ms:.1f} "
            f"ms; service p50 {st.service_p50_ms:.1f} p99 {st.service_p99_ms:.1f} ms; "
            f"{st.launches} launches, mean coalesced {st.mean_coalesced:.2f}; slot kernel ms "
            f"{[round(x) for x in st.slot_kernel_ms]}; slot phase s {ph}")

    # Saturation: every request queued at t=0 -> sustainable changes/s.
    sat_reqs = S.poisson_requests(originals, args.requests, 1.0, seed=11)
    for r in sat_reqs:
        r.arrival = 0.0
    sat = srv.play(sat_reqs, timeout=900)
    report("saturation", sat)
    capacity = sat.requests / sat.wall_s
    rate = args.rate if args.rate > 0 else args.load * capacity
    # Latency: open-loop Poisson arrivals at `rate` (fresh seeds: new batches).
    reqs = S.poisson_requests(originals, args.requests, rate, seed=12)
    stats = srv.play(reqs, timeout=900)
    report(f"poisson {rate:.1f}/s", stats)
    srv.close()
    # Outputs: a sample of requests served in the mix (all queued at once, so
    # launches overlap and coalesce) and one at a time must give identical
    # trees, combined distances and MOSP trees.
    sample = S.poisson_requests(originals, args.equal_requests, 1.0, seed=13)
    runs = []
    for kw in ({"slots": args.slots, "max_coalesce": args.coalesce, "cu_policy": args.cu_policy},
               {"slots": 1, "max_coalesce": 1}):
        eq = server(keep_results=True, **kw)
        rs = [S.Request(r.rid, r.gid, r.batch) for r in sample]
        if kw["slots"] > 1:
            eq.play(rs, timeout=900)
        else:
            for r in rs:
                eq.submit(r)
                eq.wait(r)
        eq.close()
        runs.append(rs)
    equal = all(np.array_equal(a.result[k], b.result[k]) for a, b in zip(*runs)
                for k in ("dist", "parents", "comb_dist", "comb_parent"))
    log(f"outputs of {len(sample)} requests in the mix and alone identical: {equal}")
    line = {
        "metric": "changes/s and p50/p99 latency, served: 8 resident graphs, Poisson requests",
        "value": sat.changes_per_s, "unit": "changes/s",
        "n_gpus": 1, "steps": stats.requests, "warmup": len(warm),
        "ms_per_step": float(sat.wall_s * 1e3 / sat.requests),
        "p50_ms": stats.p50_ms, "p99_ms": stats.p99_ms, "max_ms": stats.max_ms,
        "service_p50_ms": stats.service_p50_ms, "service_p99_ms": stats.service_p99_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded road-like grids, R-MAT, Poisson requests)",
        "config": {"workload": "4 road-like 1400x1400 grids K=3 + 4 R-MAT scale 20 K=4, "
                               "w~U[1,100]; batches log-uniform [100, 100k], 50% inserts, "
                               "10% disconnecting deletions",
                   "value_from": "saturation pass (all requests queued at t=0)",
                   "latency_from": f"open-loop Poisson pass at {rate:.2f} requests/s",
                   "capacity_requests_per_s": capacity, "rate_per_s": rate,
                   "requests": stats.requests, "slots": args.slots, "cu_policy": args.cu_policy,
                   "max_coalesce": args.coalesce, "hbm_budget_bytes": st.hbm_budget,
                   "resident_bytes": st.resident_bytes,
                   "mean_coalesced": stats.mean_coalesced,
                   "poisson_changes_per_s": stats.changes_per_s},
        "outputs_equal_alone": bool(equal), "equal_requests": len(sample),
    }
    return line


if __name__ == "__main__":
    sys.exit(main())
