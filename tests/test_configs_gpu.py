"""The two largest workloads at their full sizes, against the oracle.

- car + 1 trailer: P=5000 primitives of 10-40 steps, N=4096 frontier nodes,
  300 boxes on a 30 x 30 m map, 100k visited states.  Every pair of the
  exact batch matches the oracle; the learned pre-filter, trained on at
  least 2M exact-checker labels and calibrated on held-out nodes, keeps
  every feasible pair (recall 1) and the identical edge set.
- 128 independent instances (unicycle1, unicycle2, quadrotor; O ~ U{10..300},
  P ~ U{1k..10k}, N ~ U{64..4096}) expanded as one hipGraph launch: each
  instance's first nodes match the oracle run on them (the canonical order
  makes a node prefix's results independent of later nodes), every
  instance's counts are consistent, and after warm-up launches every batch
  finishes on the device.
"""
import copy
import time

import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import GpuBatch, expander_for

import oracle_lib as O
from parity_util import assert_parity, counts_consistent, prefix

pytestmark = pytest.mark.gpu


def test_car_trailer_full_size_with_surrogate(need_gpu):
    wl = synth.car_trailer()
    assert wl.offsets.size - 1 == 5000 and wl.nodes.shape[0] == 4096 and wl.boxes.shape[0] == 300
    p0 = copy.copy(wl.params)
    p0.use_surrogate = 0
    rng = np.random.default_rng(wl.seed + 100)
    lo, hi = wl.lo + 0.5, wl.hi - 0.5

    def extra(st):
        st[:, 3] = synth.wrap_angle(st[:, 2] + rng.uniform(-0.5, 0.5, size=st.shape[0]))

    with expander_for(wl) as ex:
        c_plain = ex.expand(wl.nodes, wl.g_parent, p0)
        plain = ex.fetch()
        # labels: fresh frontier-like nodes of the same map, >= 2M pairs
        probe = synth.free_planar_nodes(rng, wl.desc, wl.boxes, lo, hi, 1024, extra)
        per_node = ex.expand(probe, np.zeros(probe.shape[0]), p0)["pairs"] / probe.shape[0]
        n_train = int(np.ceil(2_100_000 / per_node))
        train_nodes = synth.free_planar_nodes(rng, wl.desc, wl.boxes, lo, hi, n_train, extra)
        cal_nodes = synth.free_planar_nodes(rng, wl.desc, wl.boxes, lo, hi, 8192, extra)
        n_labels = ex.expand(train_nodes, np.zeros(n_train), p0)["pairs"]
        ex.surrogate_init(seed=wl.seed)
        t0 = time.perf_counter()
        losses = ex.surrogate_train(train_nodes, p0, epochs=8, batch=4096, lr=2e-3, seed=wl.seed)
        t_train = time.perf_counter() - t0
        stats = ex.surrogate_calibrate(cal_nodes, p0, margin=0.5, tail=4.0)
        p1 = copy.copy(p0)
        p1.use_surrogate = 1
        c_sur = ex.expand(wl.nodes, wl.g_parent, p1)
        sur = ex.fetch()
        h = ex.health()
    print(f"labels {n_labels} from {n_train} nodes, trained in {t_train:.1f} s, losses {losses}, calibration {stats}")
    print("exact", c_plain, "\nsurrogate", c_sur, "\nhealth", h)
    assert n_labels >= 2_000_000
    assert h["wave_giveups"] == 0 and h["host_completions"] == 0

    # the exact batch, every pair
    ora, co = O.expand(wl, params=p0, threads=16)
    assert_parity(plain, ora, "car+trailer full size")
    counts_consistent(c_plain)

    # the pre-filter: recall 1 and the identical edge set on this unseen batch
    free = plain["verdict"] == N.V_FREE
    rej = sur["verdict"] == N.V_SURROGATE_REJECT
    assert np.all(sur["verdict"][free] == N.V_FREE), "a feasible edge was rejected (recall < 1)"
    assert np.array_equal(sur["resolution"], plain["resolution"])
    keep = ~rej
    for k in ("verdict", "first_k", "goal_index"):
        assert np.array_equal(sur[k][keep], plain[k][keep]), k
    for k in ("final_state", "g", "h"):
        assert np.array_equal(sur[k][free], plain[k][free], equal_nan=True), k
    assert c_sur["surrogate_reject"] == rej.sum()
    print(f"pre-filter rejected {rej.sum()} of {(~free).sum()} infeasible pairs; recall 1.0, edges identical")


def test_many_instances_against_oracle(need_gpu):
    t0 = time.perf_counter()
    wls = synth.multi_instance()
    print(f"generated {len(wls)} instances in {time.perf_counter() - t0:.1f} s", flush=True)
    params = []
    for w in wls:
        p = copy.copy(w.params)
        p.fp32_rollout = 1 if w.desc.system == N.SYS_QUADROTOR else 0
        params.append(p)
    exs = [expander_for(w) for w in wls]
    try:
        for e, w in zip(exs, wls):
            e.upload_nodes(w.nodes, w.g_parent)
        batch = GpuBatch(exs, N.BATCH_GRAPH)
        # warm-up launches: a quadrotor node whose cap falls past its FP64
        # window is finished by dp_get_counts, and the context's later
        # batches use a wider window, on the device
        for _ in range(3):
            batch.expand_async(params)
            for e in exs:
                e.counts()
        h0 = [e.health() for e in exs]
        batch.expand_async(params)
        batch.expand_async(params)
        h1 = [e.health() for e in exs]
        st = batch.stats()
        assert st["captures"] >= 1 and st["launches"] >= 5
        for w, a, b in zip(wls, h0, h1):
            assert b["wave_giveups"] == a["wave_giveups"] and b["window_overflows"] == a["window_overflows"], \
                (w.name, a, b)
        print("FP64 window overflows during warm-up:", {w.name: a["window_overflows"] for w, a in zip(wls, h0)
                                                        if a["window_overflows"]})
        total = 0
        for k, (e, w, p) in enumerate(zip(exs, wls, params)):
            c = e.counts()
            counts_consistent(c)
            total += c["pairs"]
            out = e.fetch()
            n = min(w.nodes.shape[0], 24)
            ora, co = O.expand(w, nodes=w.nodes[:n], g_parent=w.g_parent[:n], params=p, threads=16)
            assert_parity(prefix(out, co["pairs"]), ora, f"instance {w.name}")
        print(f"{len(wls)} instances, {total} pairs per launch; every instance's first nodes match the oracle")
        batch.close()
    finally:
        for e in exs:
            e.close()
