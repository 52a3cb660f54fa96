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
# HBM bytes per launch of the roofline kernel, measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB copy (profiles/ holds the runs):
# workload -> (kernel group, bytes, source).
MEASURED_TRAFFIC = {
    "quadrotor": ("check_main", 37.9e6, "profiles/k_quad32_quadrotor_hbm_traffic.txt"),
    "unicycle1_scale": ("check_main", 1.315e9, "profiles/k_check_planar_unicycle1_scale_hbm_traffic.txt"),
}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workloads():
    from dynoplan_amd import synth
    return {
        "quadrotor": synth.quadrotor,  # thrusts U[0, 1.4] x hover per rotor and piece
        "quadrotor_balanced": lambda: synth.quadrotor(thrust="balanced"),
        "unicycle1_scale": synth.unicycle1_scale,
        "unicycle1_small": synth.unicycle1_small,
        "car_trailer": synth.car_trailer,
        "multi": synth.multi_instance,
    }


# Kernel groups timed by dp_last_timing_detail (HIP events on the expansion
# stream), per system family.
KERNEL_GROUPS = {
    "check_main": {"quad": "k_quad32", "planar": "k_check_planar"},
    "check_post": {"quad": "k_quad_window + k_quad64 (with the visited duplicate test)", "planar": "(none)"},
    "lists": {"quad": "k_cont_count + k_cont_scatter + k_cell_list + k_conf_waves (+ fills, scans)",
              "planar": "k_cont_count + k_cont_scatter + k_cell_list + k_conf_waves (+ fills, scans)"},
    "walk": {"quad": "k_pre_nodes + k_resolve_ticketed",
             "planar": "k_pre_nodes + k_resolve_ticketed"},
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpus():
    """CPUs this process may use: the affinity mask, capped by the cgroup
    CPU quota when one is set; with os.cpu_count() for the record."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    return {"os_cpu_count": os.cpu_count(), "affinity": affinity, "cgroup_quota": quota, "usable": usable,
            "model": cpu_model()}


def cpu_baseline(wl, n_nodes):
    """The FP64 oracle on the host cores: single-threaded (as the reference's
    loop) and with OpenMP over frontier nodes; ``value`` is the threaded one."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cpus = host_cpus()
    threads = cpus["usable"]
    nodes = np.ascontiguousarray(wl.nodes[:n_nodes])
    gp = np.ascontiguousarray(wl.g_parent[:n_nodes])
    O.expand(wl, nodes=nodes[:4], g_parent=gp[:4], want_out=False)  # warm-up (load, page in)

    def timed(reps, thr):
        times, pairs = [], 0
        for _ in range(reps):
            t0 = time.perf_counter()
            _, c = O.expand(wl, nodes=nodes, g_parent=gp, threads=thr, want_out=False)
            times.append(time.perf_counter() - t0)
            pairs = c["pairs"]
        return pairs, float(np.median(times))

    pairs, t1 = timed(3, 0)
    progress(f"cpu oracle single-threaded: {pairs} pairs in {t1:.3f} s")
    _, tn = timed(5, threads)
    progress(f"cpu oracle {threads} threads: {pairs} pairs in {tn:.3f} s")
    return {"value": pairs / tn, "unit": "pairs/s", "cores": threads, "kind": "port",
            "single_thread_value": pairs / t1, "cpu_model": cpus["model"], "host_cpus": cpus,
            "sample": f"first {n_nodes} of {wl.nodes.shape[0]} frontier nodes ({pairs} pairs); FP64 oracle "
                      f"(grid NN, swept-box cull) with OpenMP over nodes on all {threads} usable host CPUs, "
                      f"median of 5; single-threaded median of 3 in single_thread_value"}


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


METRIC = "checked primitive expansions/s, batched+surrogate vs plain GPU at 100% recall"


def train_surrogate(ex, wl, label_pairs):
    """Untimed: exact-checker labels of fresh frontier nodes (same
    distribution, other seed), minibatch Adam on the GPU, then the
    conservative threshold from held-out nodes."""
    import copy
    from dynoplan_amd import synth
    rng = np.random.default_rng(wl.seed + 100)
    lo, hi = wl.lo + 0.5, wl.hi - 0.5

    def extra(st):
        st[:, 3] = synth.wrap_angle(st[:, 2] + rng.uniform(-0.5, 0.5, size=st.shape[0]))

    probe = synth.free_planar_nodes(rng, wl.desc, wl.boxes, lo, hi, 1024, extra)
    p0 = copy.copy(wl.params)
    p0.use_surrogate = 0
    per_node = ex.expand(probe, np.zeros(probe.shape[0]), p0)["pairs"] / probe.shape[0]
    n_train = int(np.ceil(label_pairs / max(per_node, 1e-9)))
    train_nodes = synth.free_planar_nodes(rng, wl.desc, wl.boxes, lo, hi, n_train, extra)
    cal_nodes = synth.free_planar_nodes(rng, wl.desc, wl.boxes, lo, hi, 8192, extra)
    n_labels = ex.expand(train_nodes, np.zeros(n_train), p0)["pairs"]
    progress(f"surrogate: {n_labels} exact-checker labels from {n_train} nodes")
    ex.surrogate_init(seed=wl.seed)
    t0 = time.perf_counter()
    losses = ex.surrogate_train(train_nodes, p0, epochs=8, batch=4096, lr=2e-3, seed=wl.seed)
    t_train = time.perf_counter() - t0
    stats = ex.surrogate_calibrate(cal_nodes, p0, margin=0.5, tail=4.0)
    progress(f"surrogate: trained in {t_train:.1f} s, epoch losses {[round(float(x), 4) for x in losses]}, "
             f"calibration {stats}")
    return {"labels": int(n_labels), "train_nodes": n_train, "epochs": 8, "train_s": t_train,
            "epoch_losses": [float(x) for x in losses], "calibration": stats}


def time_steps(ex, params, steps):
    import torch
    t0 = time.perf_counter()
    for _ in range(steps):
        ex.expand_async(params)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def bench_multi(args):
    """128 independent instances (unicycle1, unicycle2, quadrotor) expanded as
    one hipGraph launch; throughput, per-instance completion latency p50/p99,
    and the same instances launched one stream each and one after another."""
    import copy
    import torch
    from dynoplan_amd import _native as N
    from dynoplan_amd.expander import GpuBatch, expander_for

    t0 = time.perf_counter()
    wls = workloads()["multi"]()
    progress(f"generated {len(wls)} instances in {time.perf_counter() - t0:.1f} s")
    torch.cuda.init()
    exs = []
    for w in wls:
        e = expander_for(w, device=0)
        e.upload_nodes(w.nodes, w.g_parent)
        exs.append(e)
    params = []
    for w in wls:
        p = copy.copy(w.params)
        p.fp32_rollout = 1 if w.desc.system == N.SYS_QUADROTOR else 0
        params.append(p)
    progress(f"uploaded {len(exs)} contexts")
    # launch shapes: the untimed autotuner on the largest instance of each
    # system, applied to every instance of that system
    tiles_by_system = {}
    if not args.no_autotune:
        for sysid in sorted({w.desc.system for w in wls}):
            ids = [i for i, w in enumerate(wls) if w.desc.system == sysid]
            k = max(ids, key=lambda i: wls[i].nodes.shape[0] * (wls[i].offsets.size - 1))
            tiles, _ = exs[k].autotune(params[k])
            tiles_by_system[int(sysid)] = tiles
            for i in ids:
                exs[i].set_tiles(**tiles)
        progress(f"autotune per system: {tiles_by_system}")

    def health_sum():
        hs = [e.health() for e in exs]
        return {k: sum(h[k] for h in hs) for k in ("batches", "wave_giveups", "window_overflows")}

    def run(mode, steps, warmup):
        b = GpuBatch(exs, mode)
        for _ in range(warmup):
            b.expand_async(params)
            torch.cuda.synchronize()
            for e in exs:
                e.counts()  # finishes a batch whose FP64 window overflowed (later batches use a wider one)
        h0 = health_sum()
        t = time.perf_counter()
        for _ in range(steps):
            b.expand_async(params)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        h1 = health_sum()
        integ = {k: h1[k] - h0[k] for k in h0}
        integ["device_complete"] = integ["wave_giveups"] == 0 and integ["window_overflows"] == 0
        run.integrity = integ
        lat = []
        for _ in range(steps):
            b.expand_async(params)
            lat.append(b.latency_ms())
        st = b.stats()
        b.close()
        return el, np.array(lat), st

    warm = max(args.warmup, 3)
    el_g, lat_g, st_g = run(N.BATCH_GRAPH, args.steps, warm)
    integrity = run.integrity
    progress(f"graph integrity {integrity}")
    counts = [e.counts() for e in exs]
    pairs = sum(c["pairs"] for c in counts)
    value = pairs * args.steps / el_g
    progress(f"graph: {args.steps} steps in {el_g:.3f} s, {pairs} pairs per step, stats {st_g}")
    el_s, lat_s, _ = run(N.BATCH_STREAMS, args.steps, 1)
    progress(f"streams: {args.steps} steps in {el_s:.3f} s")
    # one instance after another on one stream (no batching across instances)
    one = torch.cuda.current_stream().cuda_stream
    for e, p in zip(exs, params):
        e.expand_async(p, one)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        for e, p in zip(exs, params):
            e.expand_async(p, one)
    torch.cuda.synchronize()
    el_q = time.perf_counter() - t
    progress(f"sequential: {args.steps} steps in {el_q:.3f} s")
    # one instance at a time with its counts read before the next, as
    # separate planner processes expanding one batch each would do
    t = time.perf_counter()
    for _ in range(args.steps):
        for e, p in zip(exs, params):
            e.expand_async(p, one)
            e.counts()
    torch.cuda.synchronize()
    el_r = time.perf_counter() - t
    progress(f"one at a time with counts: {args.steps} steps in {el_r:.3f} s")
    abytes = 0.0
    for e, w in zip(exs, wls):
        abytes += algorithmic_bytes(w, e.fetch()["prim"])
    achieved = abytes / (el_g / args.steps) / 1e9
    batch_lat = lat_g.max(axis=1)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": 1e3 * el_g / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (planar), f32 rollout + f64 re-check (quadrotor)",
        "data": "synthetic",
        "config": {"workload": f"{len(wls)} mixed instances (unicycle1, unicycle2, quadrotor), O U{{10..300}}, "
                               f"P U{{1000..10000}}, N U{{64..4096}}, V 10000, one hipGraph launch",
                   "instances": len(wls), "nodes": int(sum(w.nodes.shape[0] for w in wls)),
                   "pairs_per_step": int(pairs), "surrogate": False, "tiles_by_system": tiles_by_system},
        "latency_ms": {"instance_p50": float(np.percentile(lat_g, 50)), "instance_p99": float(np.percentile(lat_g, 99)),
                       "batch_p50": float(np.percentile(batch_lat, 50)), "batch_p99": float(np.percentile(batch_lat, 99))},
        "streams_value": pairs * args.steps / el_s,
        "streams_latency_ms": {"instance_p50": float(np.percentile(lat_s, 50)),
                               "instance_p99": float(np.percentile(lat_s, 99))},
        "sequential_value": pairs * args.steps / el_q,
        "one_at_a_time_value": pairs * args.steps / el_r,
        "graph_stats": st_g,
        "integrity": integrity,
        "roofline": {"bound": "hbm", "kernel": "whole batch (all kernels)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": abytes},
    }
    if not args.no_baselines:
        sample = [0, 1, 2, 3]
        cbs = [cpu_baseline(wls[i], 128) for i in sample]
        line["cpu_baseline"] = {"value": float(np.mean([cb["value"] for cb in cbs])), "unit": "pairs/s",
                                "cores": cbs[0]["cores"], "kind": "port",
                                "sample": f"first 128 nodes of instances {sample}, mean of their rates",
                                "per_instance": [cb["value"] for cb in cbs]}
    for e in exs:
        e.close()
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="quadrotor", choices=sorted(workloads()))
    ap.add_argument("--cpu-nodes", type=int, default=0,
                    help="frontier nodes in the CPU baseline sample (0: 4096 planar, 2048 quadrotor)")
    ap.add_argument("--plain-nodes", type=int, default=256, help="nodes expanded one per call (plain path)")
    ap.add_argument("--label-pairs", type=int, default=2_000_000, help="surrogate training labels (car_trailer)")
    ap.add_argument("--no-surrogate", action="store_true", help="car_trailer without the learned pre-filter")
    ap.add_argument("--no-baselines", action="store_true", help="skip CPU and plain-GPU legs (profiling runs)")
    ap.add_argument("--no-variants", action="store_true", help="skip the balanced-thrust quadrotor figure")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (scale reference for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the expansion path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "multi":
        bench_multi(args)
        return

    import copy
    import torch
    from dynoplan_amd import _native as N
    from dynoplan_amd.expander import expander_for

    wl = workloads()[args.workload]()
    is_quad = wl.desc.system == N.SYS_QUADROTOR
    use_sur = args.workload == "car_trailer" and not args.no_surrogate
    torch.cuda.init()
    ex = expander_for(wl, device=0)
    progress(f"workload {wl.name}: {wl.nodes.shape[0]} nodes, {wl.offsets.size - 1} primitives")
    surrogate = train_surrogate(ex, wl, args.label_pairs) if use_sur else None
    params = copy.copy(wl.params)
    params.use_surrogate = 1 if use_sur else 0

    ex.upload_nodes(wl.nodes, wl.g_parent)
    ex.reserve(wl.nodes.shape[0], wl.nodes.shape[0] * (wl.offsets.size - 1))
    if args.copy_mib > 0:
        # a copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel give
        # the scale of the counters read for the expansion kernels
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        progress(f"copied {args.copy_mib} MiB three times")
        del src, dst
    tiles = None
    if not args.no_autotune:
        # untimed: pick the check stage's launch shape for this batch
        tiles, trials = ex.autotune(params)
        progress(f"autotune: {tiles}")
    for _ in range(args.warmup):
        ex.expand_async(params)
        torch.cuda.synchronize()
        progress(f"warm-up step done: {ex.timing()}")
    health0 = ex.health()
    elapsed = time_steps(ex, params, args.steps)
    health1 = ex.health()
    # every timed batch must have finished on the device: no resolution wave
    # gave up waiting and no FP64 window overflowed (either leaves work to a
    # later dp_get_counts, outside the timed region)
    integrity = {"batches": health1["batches"] - health0["batches"],
                 "wave_giveups": health1["wave_giveups"] - health0["wave_giveups"],
                 "window_overflows": health1["window_overflows"] - health0["window_overflows"]}
    integrity["device_complete"] = integrity["wave_giveups"] == 0 and integrity["window_overflows"] == 0
    progress(f"timed {args.steps} steps in {elapsed:.3f} s; integrity {integrity}")
    counts = ex.counts()
    pairs = counts["pairs"]
    value = pairs * args.steps / elapsed

    # Per-stage and per-kernel-group device time from HIP events recorded on
    # the expansion stream, over as many further steps.
    stage = {"applicability": [], "surrogate": [], "check": [], "resolve": [], "total": []}
    detail = {k: [] for k in ("applicability", "surrogate", "check_pre", "check_main", "check_post", "lists", "walk")}
    for _ in range(args.steps):
        ex.expand_async(params)
        t = ex.timing()
        for k in stage:
            stage[k].append(t[k])
        d = ex.timing_detail()
        for k in detail:
            detail[k].append(d[k])
    stage_ms = {k: float(np.mean(v)) for k, v in stage.items()}
    detail_ms = {k: float(np.mean(v)) for k, v in detail.items()}
    progress(f"stage ms {stage_ms}; kernel groups {detail_ms}")
    out = ex.fetch()
    checked = out["verdict"] != N.V_SURROGATE_REJECT
    abytes = algorithmic_bytes(wl, out["prim"][checked])
    fam = "quad" if is_quad else "planar"
    groups = {k: {"kernels": KERNEL_GROUPS[k][fam], "ms": detail_ms[k],
                  "hbm_frac": abytes / (detail_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS if detail_ms[k] > 0 else None}
              for k in KERNEL_GROUPS}
    # The roofline kernel is the main check kernel: it streams the pairs'
    # algorithmic bytes, and it is the longest single kernel of the quadrotor
    # step (rocprof).  The longest kernel group is reported beside it.
    longest = max(groups, key=lambda k: groups[k]["ms"])
    dominant = "check_main"
    dom_ms = groups[dominant]["ms"]
    achieved = abytes / (dom_ms * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 rollout + f64 re-check" if is_quad else ("f64 + bf16 surrogate" if use_sur else "f64"),
        "data": "synthetic",
        "config": {"workload": wl.name, "nodes": int(wl.nodes.shape[0]), "primitives": int(wl.offsets.size - 1),
                   "obstacles": int(wl.boxes.shape[0] + wl.spheres.shape[0]), "visited": int(wl.visited.shape[0]),
                   "delta": float(wl.params.delta), "alpha": float(wl.params.alpha),
                   "limit_branching_factor": int(wl.params.limit_branching_factor), "seed": wl.seed,
                   "surrogate": use_sur, "pairs_per_step": int(pairs)},
        "counts": counts,
        "integrity": integrity,
        "stage_ms": stage_ms,
        "kernel_groups": groups,
        "dominant_kernel": KERNEL_GROUPS[dominant][fam],
        "longest_group": {"group": longest, "kernels": KERNEL_GROUPS[longest][fam], "ms": groups[longest]["ms"]},
        "tiles": tiles,
        "roofline": {"bound": "hbm", "kernel": KERNEL_GROUPS[dominant][fam], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": abytes, "launch_ms": dom_ms},
    }
    key = args.workload
    if key in MEASURED_TRAFFIC and not use_sur and MEASURED_TRAFFIC[key][0] == dominant:
        _, line["roofline"]["traffic"], line["roofline"]["traffic_source"] = MEASURED_TRAFFIC[key]
    if is_quad:
        line["fp32_vs_fp64"] = rollout_error(ex, out)
        progress(f"FP32 vs FP64 rollout error: {line['fp32_vs_fp64']}")
    if use_sur:
        # Untimed validation: the exact path on the same batch decides every
        # pair; the surrogate run must keep every feasible edge and give the
        # same edge set.
        exact = copy.copy(wl.params)
        exact.use_surrogate = 0
        ex.expand_async(exact)
        ref = ex.fetch()
        free = ref["verdict"] == N.V_FREE
        kept = out["verdict"][free] == N.V_FREE
        rej = out["verdict"] == N.V_SURROGATE_REJECT
        recall = float(kept.mean()) if free.any() else 1.0
        same_edges = bool(np.array_equal(out["resolution"], ref["resolution"]))
        t_exact = time_steps(ex, exact, args.steps)
        surrogate.update({"recall": recall, "edge_set_identical": same_edges,
                          "rejected_pairs": int(rej.sum()), "infeasible_pairs": int((~free).sum()),
                          "rejected_share_of_infeasible": float(rej.sum() / max(1, (~free).sum())),
                          "batched_exact_value": pairs * args.steps / t_exact,
                          "speedup_vs_batched_exact": t_exact / elapsed})
        progress(f"surrogate validation: {surrogate}")
        line["surrogate"] = surrogate
    if args.workload == "quadrotor" and not args.no_variants:
        # the same path on primitives that stay near hover (few pairs fail the bounds)
        line["variants"] = {"quadrotor_balanced": variant_line(workloads()["quadrotor_balanced"](), args)}
    if not args.no_baselines:
        pg = plain_gpu(ex, wl, args.plain_nodes)
        progress(f"plain GPU path: {pg}")
        line["plain_gpu"] = pg
        line["speedup_vs_plain_gpu"] = value / pg["value"]
        cpu_nodes = args.cpu_nodes or (2048 if is_quad else 4096)
        cb = cpu_baseline(wl, cpu_nodes)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = value / cb["value"]
    ex.close()
    print(json.dumps(line))
    if not integrity["device_complete"]:
        print("bench: a timed batch needed host-launched completion; the line is not a device-only measurement",
              file=sys.stderr)
        sys.exit(3)


def rollout_error(ex, out, n_sample=2000, seed=0):
    """FP32 rollout against the library's FP64 rollout of the same pairs,
    every state (up to the first out-of-bounds state) and every component
    group: max and mean of the per-pair maxima."""
    from dynoplan_amd import _native as N
    rng = np.random.default_rng(seed)
    n = out["prim"].size
    pick = np.sort(rng.choice(n, min(n_sample, n), replace=False)).astype(np.int32)
    s32 = ex.rollout_states(pick, 32)
    s64 = ex.rollout_states(pick, 64)
    upto = np.where(out["verdict"][pick] == N.V_BOUNDS, out["first_k"][pick] + 1, s32.shape[1])
    mask = np.arange(s32.shape[1])[None, :] < upto[:, None]
    err = np.abs(s32 - s64)
    err = np.where(mask[:, :, None] & np.isfinite(err), err, 0.0)
    res = {"pairs": int(pick.size)}
    for name, sl in (("position_m", slice(0, 3)), ("quaternion", slice(3, 7)), ("velocity_mps", slice(7, 10)),
                     ("rate_radps", slice(10, 13))):
        per_pair = err[:, :, sl].max(axis=(1, 2))
        res[name] = {"max": float(per_pair.max()), "mean": float(per_pair.mean())}
    return res


def variant_line(wl, args):
    """Batched throughput of another workload in the same run (no baselines)."""
    import copy
    import torch
    from dynoplan_amd.expander import expander_for
    with expander_for(wl, device=0) as ex:
        params = copy.copy(wl.params)
        ex.upload_nodes(wl.nodes, wl.g_parent)
        ex.reserve(wl.nodes.shape[0], wl.nodes.shape[0] * (wl.offsets.size - 1))
        tiles, _ = ex.autotune(params)
        for _ in range(args.warmup):
            ex.expand_async(params)
            torch.cuda.synchronize()
        h0 = ex.health()
        el = time_steps(ex, params, args.steps)
        h1 = ex.health()
        c = ex.counts()
        ex.expand_async(params)
        d = ex.timing_detail()
        res = {"workload": wl.name, "value": c["pairs"] * args.steps / el, "ms_per_step": 1e3 * el / args.steps,
               "pairs_per_step": c["pairs"], "bounds": c["bounds"], "collide": c["collide"], "free_": c["free_"],
               "capped": c["capped"], "fp64_deferred": c["fp64_deferred"], "kernel_ms": d, "tiles": tiles,
               "integrity": {k: h1[k] - h0[k] for k in ("batches", "wave_giveups", "window_overflows")}}
        res["integrity"]["device_complete"] = (res["integrity"]["wave_giveups"] == 0
                                               and res["integrity"]["window_overflows"] == 0)
    progress(f"variant {wl.name}: {res}")
    return res


if __name__ == "__main__":
    main()
