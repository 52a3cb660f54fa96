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
    return synth.quadrotor(N_nodes=512, n_visited=20000, thrust="balanced")


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


@pytest.mark.parametrize("thrust", ["balanced", "uniform"])
def test_quadrotor_fp32_error_and_rechecks(need_gpu, wl, thrust):
    """FP32 rollout against the FP64 oracle states of the same pairs, every
    state and every component.  The FP32 verdicts are exact only while the
    error stays below the re-check margins (near_eps for positions, 10x for
    velocities, 100x for angular rates): the test bounds it by a tenth of
    each."""
    import ctypes as C
    if thrust == "uniform":
        wl = synth.quadrotor(seed=42, N_nodes=256, n_visited=5000)
    p = params(wl, 1, check_intermediate_goal=False)
    rng = np.random.default_rng(0)
    with expander_for(wl) as ex:
        c = ex.expand(wl.nodes, wl.g_parent, p)
        out = ex.fetch()
        n = c["pairs"]
        pick = np.sort(rng.choice(n, 3000, replace=False)).astype(np.int32)
        L = int(np.diff(wl.offsets).max())
        st = ex.rollout_states(pick, 32)
        st64 = ex.rollout_states(pick, 64)
    lib = O.load()
    pr = O.prims_struct(np.ascontiguousarray(wl.offsets, np.int32), np.ascontiguousarray(wl.states),
                        np.ascontiguousarray(wl.actions))
    ref = np.zeros((L, 13))
    groups = {"position": slice(0, 3), "quaternion": slice(3, 7), "velocity": slice(7, 10), "rate": slice(10, 13)}
    worst = {k: 0.0 for k in groups}
    means = {k: [] for k in groups}
    for t, q in enumerate(pick):
        node = np.ascontiguousarray(wl.nodes[out["node"][q]])
        n1 = lib.oracle_transform_primitive(C.byref(wl.desc), C.byref(pr), int(out["prim"][q]), O.dptr(node),
                                            O.dptr(ref))
        # the library's FP64 rollout is the oracle's, bit for bit
        assert np.array_equal(st64[t, :n1].view(np.int64), ref[:n1].view(np.int64)), t
        upto = n1 if out["verdict"][q] != N.V_BOUNDS else out["first_k"][q] + 1
        for k, sl in groups.items():
            e = np.abs(st[t, :upto, sl] - ref[:upto, sl]).max()
            worst[k] = max(worst[k], float(e))
            means[k].append(float(e))
    print("FP32 vs FP64 over every state of", pick.size, "pairs: max", {k: f"{v:.2e}" for k, v in worst.items()},
          "mean of per-pair max", {k: f"{np.mean(v):.2e}" for k, v in means.items()},
          f"; re-checked in FP64: {c['recheck_fp64']} of {n}")
    margin = {"position": 1.0, "velocity": 10.0, "rate": 100.0}
    for k, m in margin.items():
        assert worst[k] < m * wl.desc.near_eps / 10, k
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


@pytest.mark.parametrize("window,limit,overflow", [(0, 20, False), (2, 20, True), (3, 5, True), (1, 0, None)])
def test_quadrotor_fp64_window(need_gpu, wl, window, limit, overflow):
    """FREE pairs past a node's FP64 window skip the FP64 rollout: the
    default window (3 * limit + 32 FREE pairs) holds every cap of
    this batch; tiny forced windows overflow, and the batch is completed by a
    second resolution with every deferred pair rolled out.  Either way the
    results are the oracle's."""
    n = 128
    p = params(wl, 1, limit_branching_factor=limit)
    with expander_for(wl) as ex:
        ex.debug_set_limits(fp64_window=window)
        c = ex.expand(wl.nodes[:n], wl.g_parent[:n], p)
        gpu = ex.fetch()
        h = ex.health()
    ora, co = O.expand(wl, nodes=wl.nodes[:n], g_parent=wl.g_parent[:n], params=p)
    print("window", window, "limit", limit, "deferred", c["fp64_deferred"], "health", h)
    assert_parity(gpu, ora, f"quad window={window} limit={limit}")
    counts_consistent(c)
    assert c["fp64_deferred"] > 0
    if overflow is False:
        assert h["window_overflows"] == 0 and h["host_completions"] == 0
    elif overflow:
        assert h["window_overflows"] > 0 and h["host_completions"] > 0


def test_quadrotor_uniform_thrust_parity(need_gpu):
    """The stated thrust distribution (each rotor U[0, 1.4] x hover per
    piece): most pairs leave the angular-rate bounds early; BOUNDS indices,
    collisions, the few FREE pairs and their resolution match the oracle."""
    wl = synth.quadrotor(seed=41, P=3000, N_nodes=256, n_visited=5000)
    p = params(wl, 1)
    with expander_for(wl) as ex:
        c = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    ora, co = O.expand(wl, params=p)
    print("counts", c)
    assert_parity(gpu, ora, "quad uniform thrust")
    counts_consistent(c)
    assert co["bounds"] > co["pairs"] // 2 and co["free_"] > 0 and co["collide"] > 0


@pytest.mark.parametrize("thrust", ["uniform", "balanced"])
def test_quadrotor_full_size_batch(need_gpu, thrust):
    """The quadrotor batch at full size (N=2048, P=10k, 100 boxes + 50
    spheres, V=20k): every pair against the oracle, and the device pipeline
    finishes every batch by itself (no FP64 window overflow, no resolution
    wave giving up) over repeated launches."""
    wl = synth.quadrotor(thrust=thrust)
    with expander_for(wl) as ex:
        ex.upload_nodes(wl.nodes, wl.g_parent)
        h0 = ex.health()
        for _ in range(5):
            ex.expand_async(wl.params)
        h1 = ex.health()
        c = ex.counts()
        gpu = ex.fetch()
    print(thrust, "counts", c, "health", h0, h1)
    assert h1["window_overflows"] == h0["window_overflows"] and h1["wave_giveups"] == h0["wave_giveups"]
    ora, co = O.expand(wl, threads=16)
    assert_parity(gpu, ora, f"quad full {thrust}")
    counts_consistent(c)
