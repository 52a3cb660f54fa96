#!/usr/bin/env python3
"""Throughput of the batched db-A* node-expansion path on one MI355X.

One step = one dp_expand() over a batch of frontier nodes already resident
in HBM: applicability query, translation, bounds, per-state collision with
first-collision index, check_goal, g/h, duplicate test against the visited
set and the ordered in-batch duplicate / branching-cap resolution.  The unit
is one (frontier node, applicable primitive) pair that received a verdict.

Printed: ONE JSON line with the batched throughput (``value``), the roofline
of the dominant kernel (HIP events on the expansion stream), the plain GPU
path (one node per call, as the reference's loop pops one node at a time)
and the CPU FP64 oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public spec


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workloads():
    from dynoplan_amd import synth
    return {
        "unicycle1_scale": synth.unicycle1_scale,
        "unicycle1_small": synth.unicycle1_small,
        "car_trailer": synth.car_trailer,
        "quadrotor": synth.quadrotor,
    }


def algorithmic_bytes(wl, prim):
    """Bytes one pair must stream.  Translate systems: the primitive's states
    (nx FP64 per state) plus an 8-byte verdict.  Quadrotor: the FP32 actions
    (4 x 4 B per step) plus the FP32 start state (13 x 4 B) plus the verdict."""
    from dynoplan_amd import _native as N
    lens = np.diff(wl.offsets.astype(np.int64))[prim]
    if wl.desc.system == N.SYS_QUADROTOR:
        return float(np.sum((lens - 1) * 16 + 52 + 8))
    return float(np.sum(lens * wl.desc.nx * 8 + 8))


def cpu_baseline(wl, n_nodes):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    nodes = wl.nodes[:n_nodes]
    O.expand(wl, nodes=nodes[:4], g_parent=wl.g_parent[:4])  # warm-up (load, page in)
    times, pairs = [], 0
    for _ in range(3):
        t0 = time.perf_counter()
        _, c = O.expand(wl, nodes=nodes, g_parent=wl.g_parent[:n_nodes])
        times.append(time.perf_counter() - t0)
        pairs = c["pairs"]
    t = float(np.median(times))
    return {"value": pairs / t, "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": f"first {n_nodes} of {wl.nodes.shape[0]} frontier nodes ({pairs} pairs), "
                      f"single-threaded FP64 oracle (grid NN, swept-box cull), median of 3"}


def plain_gpu(ex, wl, n_nodes):
    """One node per dp_expand_batch call (upload + expand + synchronise)."""
    import ctypes as C
    from dynoplan_amd import _native as N
    from dynoplan_amd.expander import _dptr
    lib = ex.lib
    nodes = np.ascontiguousarray(wl.nodes[:n_nodes])
    g = np.ascontiguousarray(wl.g_parent[:n_nodes])
    for i in range(min(4, n_nodes)):  # warm-up
        N.check(lib.dp_expand_batch(ex.ctx, _dptr(nodes[i:i + 1]), _dptr(g[i:i + 1]), 1, C.byref(wl.params), None))
    pairs = 0
    cnt = N.Counts()
    t0 = time.perf_counter()
    for i in range(n_nodes):
        N.check(lib.dp_expand_batch(ex.ctx, _dptr(nodes[i:i + 1]), _dptr(g[i:i + 1]), 1, C.byref(wl.params), None))
        N.check(lib.dp_get_counts(ex.ctx, C.byref(cnt)))
        pairs += cnt.pairs
    t = time.perf_counter() - t0
    return {"value": pairs / t, "unit": "pairs/s", "nodes": n_nodes, "pairs": int(pairs),
            "ms_per_node": 1e3 * t / n_nodes}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="unicycle1_scale", choices=sorted(workloads()))
    ap.add_argument("--cpu-nodes", type=int, default=4096, help="frontier nodes in the CPU baseline sample")
    ap.add_argument("--plain-nodes", type=int, default=256, help="nodes expanded one per call (plain path)")
    ap.add_argument("--no-baselines", action="store_true", help="skip CPU and plain-GPU legs (profiling runs)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the expansion path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration).", file=sys.stderr)
        sys.exit(2)

    import torch
    from dynoplan_amd.expander import expander_for

    wl = workloads()[args.workload]()
    from dynoplan_amd import _native as N
    is_quad = wl.desc.system == N.SYS_QUADROTOR
    torch.cuda.init()
    ex = expander_for(wl, device=0)
    ex.upload_nodes(wl.nodes, wl.g_parent)
    ex.reserve(wl.nodes.shape[0], wl.nodes.shape[0] * (wl.offsets.size - 1))

    progress(f"workload {wl.name}: {wl.nodes.shape[0]} nodes, {wl.offsets.size - 1} primitives")
    for _ in range(args.warmup):
        ex.expand_async(wl.params)
        torch.cuda.synchronize()
        progress(f"warm-up step done: {ex.timing()}")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ex.expand_async(wl.params)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    counts = ex.counts()
    progress(f"timed {args.steps} steps in {elapsed:.3f} s: {counts['pairs']} pairs per step")
    pairs = counts["pairs"]
    value = pairs * args.steps / elapsed

    # Per-stage device time from HIP events recorded on the expansion stream.
    stage = {"applicability": [], "check": [], "resolve": [], "total": []}
    for _ in range(args.steps):
        ex.expand_async(wl.params)
        t = ex.timing()
        for k in stage:
            stage[k].append(t[k])
    stage_ms = {k: float(np.mean(v)) for k, v in stage.items()}
    out = ex.fetch()
    abytes = algorithmic_bytes(wl, out["prim"])
    check_s = stage_ms["check"] * 1e-3
    achieved = abytes / check_s / 1e9

    line = {
        "metric": "checked primitive expansions/s (batched GPU path; plain GPU path and CPU oracle alongside)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 rollout + f64 re-check" if is_quad else "f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "nodes": int(wl.nodes.shape[0]), "primitives": int(wl.offsets.size - 1),
                   "obstacles": int(wl.boxes.shape[0] + wl.spheres.shape[0]), "visited": int(wl.visited.shape[0]),
                   "delta": float(wl.params.delta), "alpha": float(wl.params.alpha),
                   "limit_branching_factor": int(wl.params.limit_branching_factor), "seed": wl.seed,
                   "pairs_per_step": int(pairs)},
        "counts": counts,
        "stage_ms": stage_ms,
        "roofline": {"bound": "hbm", "kernel": "k_quad32+k_quad64" if is_quad else "k_check_planar", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": abytes},
    }
    progress(f"stage ms {stage_ms}")
    if not args.no_baselines:
        pg = plain_gpu(ex, wl, args.plain_nodes)
        progress(f"plain GPU path: {pg}")
        line["plain_gpu"] = pg
        line["speedup_vs_plain_gpu"] = value / pg["value"]
        cb = cpu_baseline(wl, args.cpu_nodes)
        progress(f"cpu baseline: {cb}")
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = value / cb["value"]
    ex.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
