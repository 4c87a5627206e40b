"""Autotuner of the batched launch: near-far bucket width (as a multiple of
the reference's default rule), lanes per frontier vertex (queue tile size:
1024 >> lane_bits vertices per workgroup pass), workgroups per launch, XCD
placement and the group-sizing cost model (workgroups per instance), timed on the workload itself (HIP-graph replays, device
time).  Outputs never depend on these choices, only the time does.

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


def tune(bu, space: dict | None = None, reps: int = 3, log=print) -> dict:
    """Times every configuration of `space` on `bu` (a BatchedUpdate), leaves
    the fastest configured and returns {"best": config, "ms": ..., "trials": [...]}."""
    space = {**DEFAULT_SPACE, **(space or {})}
    keys = list(space)
    trials = []
    t0 = time.perf_counter()
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, values))
        bu.configure(**cfg)
        bu.run(use_graph=True)  # capture + warm
        ms = min(bu.run(use_graph=True) for _ in range(reps))
        trials.append({**cfg, "ms": round(ms, 4)})
        log(f"autotune {cfg}: {ms:.3f} ms")
    best = min(trials, key=lambda r: r["ms"])
    cfg = {k: best[k] for k in keys}
    bu.configure(**cfg)
    return {"best": cfg, "ms": best["ms"], "trials": trials,
            "seconds": round(time.perf_counter() - t0, 2)}
