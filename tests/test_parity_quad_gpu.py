"""GPU parity of the quadrotor rollout path against the CPU FP64 oracle.

Exact mode (fp32_rollout=0) runs the FP64 rollout for every pair; the fast
mode rolls out in FP32 with FP64 quaternion renormalisation and re-checks in
FP64 every pair whose verdict a state within near_eps of contact or of a
bound could change.  Both must reproduce the oracle bit-exactly.  The FP32
rollout error against the FP64 states is reported and bounded by near_eps/10.
"""
import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

import oracle_lib as O
from parity_util import assert_parity, counts_consistent, prefix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    return synth.quadrotor(N_nodes=512, n_visited=20000)


def params(wl, fp32, **kw):
    p = synth.make_params(goal=wl.params.goal, fp32_rollout=fp32, **kw)
    return p


@pytest.mark.parametrize("fp32", [0, 1])
def test_quadrotor_parity(need_gpu, wl, fp32):
    n = 96
    p = params(wl, fp32)
    with expander_for(wl) as ex:
        c = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    counts_consistent(c)
    ora, co = O.expand(wl, nodes=wl.nodes[:n], g_parent=wl.g_parent[:n], params=p)
    assert_parity(prefix(gpu, co["pairs"]), ora, f"quad fp32={fp32}")
    assert co["collide"] > 0 and co["free_"] > 0
    print("counts", c)


def test_quadrotor_fp32_error_and_rechecks(need_gpu, wl):
    """FP32 rollout vs the FP64 states of the same pairs."""
    import ctypes as C
    p = params(wl, 1, check_intermediate_goal=False)
    with expander_for(wl) as ex:
        c = ex.expand(wl.nodes, wl.g_parent, p)
        out = ex.fetch()
        n = c["pairs"]
        dbg = np.zeros((3, n), np.float32)
        lib = ex.lib
        assert lib.dp_debug_fp32_positions(ex.ctx, dbg.ctypes.data_as(C.POINTER(C.c_float)), n) == 0
    free = (out["verdict"] == N.V_FREE) & np.isfinite(dbg).all(axis=0)
    err = np.abs(dbg.T[free].astype(np.float64) - out["final_state"][free, :3]).max(axis=1)
    print(f"FP32 vs FP64 final position error over {free.sum()} FREE pairs: max {err.max():.3e} m, "
          f"mean {err.mean():.3e} m; re-checked in FP64: {c['recheck_fp64']} of {n}")
    assert err.max() < wl.desc.near_eps / 10
    assert c["recheck_fp64"] < 0.2 * n


def test_quadrotor_goal_and_cap(need_gpu, wl):
    goal = wl.nodes[5].copy()
    for fp32 in (0, 1):
        p = params(wl, fp32, delta=0.8, limit_branching_factor=3)
        with expander_for(wl) as ex:
            ex.expand(wl.nodes[:64], wl.g_parent[:64], p)
            gpu = ex.fetch()
        p.goal[:13] = goal.tolist()
        with expander_for(wl) as ex:
            ex.expand(wl.nodes[:64], wl.g_parent[:64], p)
            gpu = ex.fetch()
        ora, co = O.expand(wl, nodes=wl.nodes[:64], g_parent=wl.g_parent[:64], params=p)
        assert co["goal_hits"] > 0
        assert_parity(gpu, ora, f"quad goal fp32={fp32}")
