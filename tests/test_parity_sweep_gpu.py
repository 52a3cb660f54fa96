"""Seeded parity sweep of the expansion against the CPU FP64 oracle.

Every case runs the whole batch through the C-ABI (dynoplan_amd.expander)
and the oracle on the same inputs, for several seeds of each system and
several branching caps (dbastar.cpp:916-918: -1 never caps, 0 decides only
each node's first FREE pair) and duplicate radii: verdicts, first offending
indices, resolutions and FP64 final states / g / h bit-exact, and the batch
counts equal.  The quadrotor cases use the FP32 rollout with its FP64
re-check.
"""
import pytest

from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

import oracle_lib as O
from parity_util import assert_parity, counts_consistent

pytestmark = pytest.mark.gpu

COUNTS = ("pairs", "bounds", "collide", "free_", "novel", "dup_visited", "dup_batch", "capped", "goal_hits")


def check(wl, p, label):
    with expander_for(wl) as ex:
        cg = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    ora, co = O.expand(wl, params=p)
    assert_parity(gpu, ora, label)
    counts_consistent(cg)
    for k in COUNTS:
        assert cg[k] == co[k], (label, k, cg[k], co[k])
    return co


@pytest.mark.parametrize("seed", range(100, 106))
def test_unicycle1_seeds(need_gpu, seed):
    wl = synth.unicycle1_small(seed=seed, N_nodes=256, n_visited=2000)
    dup_batch = 0
    for lim in (-1, 0, 3, 20):
        for delta in (0.3, 0.6):
            p = synth.make_params(goal=wl.params.goal[:3], delta=delta, limit_branching_factor=lim)
            dup_batch += check(wl, p, f"unicycle1 seed={seed} limit={lim} delta={delta}")["dup_batch"]
    assert dup_batch > 0


@pytest.mark.parametrize("seed", range(200, 204))
@pytest.mark.parametrize("system", ["car", "unicycle2"])
def test_car_unicycle2_seeds(need_gpu, system, seed):
    if system == "car":
        wl = synth.car_trailer(seed=seed, P=800, N_nodes=200, n_boxes=60, n_visited=3000, extent=10.0)
    else:
        wl = synth.unicycle2_small(seed=seed, N_nodes=256)
    for lim in (1, 8):
        p = synth.make_params(goal=wl.params.goal[:wl.desc.nx], delta=0.6, limit_branching_factor=lim)
        check(wl, p, f"{system} seed={seed} limit={lim}")


@pytest.mark.parametrize("seed", range(300, 303))
def test_quadrotor_seeds(need_gpu, seed):
    wl = synth.quadrotor(seed=seed, N_nodes=96, n_visited=5000, thrust="balanced")
    for lim in (-1, 3, 20):
        p = synth.make_params(goal=wl.params.goal, fp32_rollout=1, limit_branching_factor=lim)
        co = check(wl, p, f"quadrotor seed={seed} limit={lim}")
        if lim == -1:
            assert co["dup_batch"] > 0 and co["capped"] == 0
