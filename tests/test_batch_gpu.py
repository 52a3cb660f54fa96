"""Many planning instances in one launch (graph and stream modes): every
instance's results equal its own single-context expansion, bit for bit; a
replay with unchanged inputs reuses the captured graph; changing one
instance's frontier size re-captures it."""
import copy

import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import GpuBatch, expander_for

pytestmark = pytest.mark.gpu

KEYS = ("node", "prim", "verdict", "first_k", "goal_index", "resolution", "final_state", "g", "h")


def instances():
    wls = [
        synth.unicycle1_small(seed=21, N_nodes=64, P=400),
        synth.unicycle2_small(seed=22, N_nodes=300, P=700),
        synth.quadrotor(seed=23, P=600, N_nodes=128, n_visited=2000, thrust="balanced"),
        synth.car_trailer(seed=24, P=500, N_nodes=200, n_boxes=80, extent=12.0, n_visited=3000),
        synth.unicycle1_scale(seed=25, N_nodes=512, P=1500, n_visited=20000),
        synth.quadrotor(seed=26, P=300, N_nodes=64, n_visited=500, n_boxes=20, n_spheres=10, thrust="balanced"),
    ]
    for w in wls:
        if w.desc.system == N.SYS_QUADROTOR:
            w.params.fp32_rollout = 1
    return wls


def single(ex, wl, nodes=None):
    nodes = wl.nodes if nodes is None else nodes
    c = ex.expand(nodes, wl.g_parent[: nodes.shape[0]], wl.params)
    return c, ex.fetch()


def same(a, b):
    ca, fa = a
    cb, fb = b
    assert {k: v for k, v in ca.items() if k != "resolve_rounds"} == \
        {k: v for k, v in cb.items() if k != "resolve_rounds"}
    for k in KEYS:
        assert np.array_equal(fa[k], fb[k], equal_nan=fa[k].dtype == np.float64), k


@pytest.mark.parametrize("mode", [N.BATCH_GRAPH, N.BATCH_STREAMS])
def test_batch_equals_single_instances(need_gpu, mode):
    wls = instances()
    exs = [expander_for(w) for w in wls]
    try:
        ref = [single(e, w) for e, w in zip(exs, wls)]
        for e, w in zip(exs, wls):
            e.upload_nodes(w.nodes, w.g_parent)
        batch = GpuBatch(exs, mode)
        params = [w.params for w in wls]
        for rep in range(3):
            batch.expand_async(params)
            lat = batch.latency_ms()
            print(f"mode {mode} replay {rep}: latency ms {np.round(lat, 3).tolist()}")
            assert np.all(np.isfinite(lat)) and np.all(lat > 0)
            for e, r in zip(exs, ref):
                same((e.counts(), e.fetch()), r)
        st = batch.stats()
        assert st["launches"] == 3
        if mode == N.BATCH_GRAPH:
            assert st["captures"] == 1
            # a different frontier size for one instance: captured again
            w = wls[1]
            sub = np.ascontiguousarray(w.nodes[:177])
            r1 = single(exs[1], w, sub)
            exs[1].upload_nodes(sub, w.g_parent[:177])
            batch.expand_async(params)
            batch.latency_ms()
            same((exs[1].counts(), exs[1].fetch()), r1)
            for e, r in zip(exs[2:], ref[2:]):
                same((e.counts(), e.fetch()), r)
            assert batch.stats()["captures"] == 2
        batch.close()
    finally:
        for e in exs:
            e.close()
