"""Autotuner of the batched launch: near-far bucket width (as a multiple of
the reference's default rule), lanes per frontier vertex (queue tile size:
1024 >> lane_bits vertices per workgroup pass), workgroups per launch, XCD
placement and the group-sizing cost model (workgroups per instance); then,
from the best of those, the workgroup shape (1024 threads, one workgroup per
CU, or 512 threads, two per CU: the library holds both builds), the LDS
queue entries of the push loop and the sizing model again for that shape.
Timed on the workload itself (HIP-graph replays, device time).  Outputs
never depend on these choices, only the time does.

Choices persist in a JSON file keyed by a workload signature, so a later run
of the same workload reuses them without re-timing.
"""
from __future__ import annotations

import itertools
import json
import pathlib
import time

DEFAULT_SPACE = {
    "workgroups": (0,),
    "lane_bits": (1, 2, 3),
    "delta_scale": (1.0, 2.0, 3.0, 4.0),
    "xcd_aware": (True,),
    "latency_items": (100.0, 400.0, 1600.0),
}
# Second stage, around the first stage's best.
SHAPE_SPACE = {
    "threads": (1024, 512),
    "queue_cap": (1024, 2048, 0),
    "latency_items": (100.0, 400.0, 1600.0),
}


def signature(sessions) -> str:
    ns = sorted((s.n, s.K, s.batch.n_ins + s.batch.n_del) for s in sessions)
    total_n = sum(n for n, _, _ in ns)
    total_k = sum(k for _, k, _ in ns)
    total_c = sum(c for _, _, c in ns)
    return f"batched:{len(ns)}:n{total_n}:k{total_k}:c{total_c}"


def load(path, key: str):
    p = pathlib.Path(path)
    if not p.exists():
        return None
    return json.loads(p.read_text()).get(key)


def save(path, key: str, choice: dict) -> None:
    p = pathlib.Path(path)
    data = json.loads(p.read_text()) if p.exists() else {}
    data[key] = choice
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


def _sweep(bu, base: dict, space: dict, reps: int, log, trials: list) -> dict:
    """Times `base` updated with every combination of `space`; returns the best."""
    keys = list(space)
    best = None
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = {**base, **dict(zip(keys, values))}
        bu.configure(**cfg)
        bu.run(use_graph=True)  # capture + warm
        ms = min(bu.run(use_graph=True) for _ in range(reps))
        trials.append({**cfg, "ms": round(ms, 4)})
        log(f"autotune {cfg}: {ms:.3f} ms")
        if best is None or ms < best[1]:
            best = (cfg, ms)
    return best


def tune(bu, space: dict | None = None, reps: int = 3, log=print,
         shape_space: dict | None = None) -> dict:
    """Times every configuration of `space` on `bu` (a BatchedUpdate), then
    every shape of `shape_space` around the best; leaves the fastest
    configured and returns {"best": config, "ms": ..., "trials": [...]}."""
    space = {**DEFAULT_SPACE, **(space or {})}
    shape_space = SHAPE_SPACE if shape_space is None else shape_space
    trials = []
    t0 = time.perf_counter()
    cfg, ms = _sweep(bu, {"threads": 1024, "queue_cap": 0}, space, reps, log, trials)
    if shape_space:
        cfg2, ms2 = _sweep(bu, cfg, shape_space, reps, log, trials)
        if ms2 < ms:
            cfg, ms = cfg2, ms2
    bu.configure(**cfg)
    return {"best": cfg, "ms": round(ms, 4), "trials": trials,
            "seconds": round(time.perf_counter() - t0, 2)}
