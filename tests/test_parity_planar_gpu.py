"""GPU parity of the planar expansion path against the CPU FP64 oracle.

Every test calls the HIP library through its C-ABI (dynoplan_amd.expander)
and compares with oracle/ on the same seeded inputs: verdicts, first
offending indices, goal indices, duplicate/cap resolutions bit-exact, and
FP64 final state / g / h bit patterns on FREE pairs.
"""
import math

import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

import oracle_lib as O
from parity_util import assert_parity, counts_consistent, prefix

pytestmark = pytest.mark.gpu


def run_both(wl, params=None, nodes=None, g_parent=None):
    params = params or wl.params
    nodes = wl.nodes if nodes is None else nodes
    g_parent = wl.g_parent if g_parent is None else g_parent
    with expander_for(wl) as ex:
        cg = ex.expand(nodes, g_parent, params)
        gpu = ex.fetch()
    ora, co = O.expand(wl, nodes=nodes, g_parent=g_parent, params=params)
    return gpu, cg, ora, co


def test_unicycle1_small_full(need_gpu):
    wl = synth.unicycle1_small()
    gpu, cg, ora, co = run_both(wl)
    print("counts", cg)
    assert_parity(gpu, ora, "unicycle1_small")
    counts_consistent(cg)
    for k in ("pairs", "bounds", "collide", "free_", "novel", "dup_visited", "dup_batch", "capped", "goal_hits",
              "checked"):
        assert cg[k] == co[k], k


@pytest.mark.parametrize("limit", [-1, 0, 1, 5])
def test_unicycle1_branching_cap(need_gpu, limit):
    """The reference's break (dbastar.cpp:916-918): a size_t NOVEL count >=
    the int limit after each FREE pair.  -1 never caps; 0 decides only each
    node's first FREE pair."""
    wl = synth.unicycle1_small(seed=11)
    p = synth.make_params(goal=wl.params.goal[:3], limit_branching_factor=limit)
    gpu, cg, ora, co = run_both(wl, params=p)
    assert_parity(gpu, ora, f"limit={limit}")
    counts_consistent(cg)
    assert cg["capped"] == co["capped"] and cg["goal_hits"] == co["goal_hits"]
    if limit < 0:
        assert cg["capped"] == 0
    if limit == 0:
        dec = (gpu["verdict"] == N.V_FREE) & (gpu["resolution"] != N.R_CAPPED)
        with_free = np.unique(gpu["node"][gpu["verdict"] == N.V_FREE])
        assert dec.sum() == with_free.size and np.array_equal(np.unique(gpu["node"][dec]), with_free)


def test_unicycle1_rare_list_paths(need_gpu):
    """Conflict lists forced out of their inline slots: 2 inline entries per
    pair, then a pool too small for every long list, so the walk reads pool
    entries and scans the grid for lists that did not fit.  Results are the
    oracle's."""
    wl = synth.unicycle1_small(seed=13, N_nodes=256, n_visited=2000)
    p = synth.make_params(goal=wl.params.goal[:3], delta=0.6, limit_branching_factor=-1)
    ora, co = O.expand(wl, params=p)
    for slots, pool in ((2, 0), (2, 300), (1, 1)):
        with expander_for(wl) as ex:
            ex.debug_set_limits(conf_slots=slots, pool_entries=pool)
            cg = ex.expand(wl.nodes, wl.g_parent, p)
            gpu = ex.fetch()
        print("slots", slots, "pool", pool, "pool entries used", cg["conf_pool_used"], "dup_batch", cg["dup_batch"])
        assert cg["conf_pool_used"] > 0
        assert_parity(gpu, ora, f"slots={slots} pool={pool}")


@pytest.mark.parametrize("name,window", [("unicycle1", 1), ("unicycle1", 6), ("car", 3), ("unicycle2", 0)])
def test_list_window(need_gpu, name, window):
    """Stored conflict lists only for the first `window` FREE pairs of each
    node (0: the default 3 * limit + 32): with tiny windows most walked pairs
    lie past it and are resolved by scanning the contender grid.  Results
    are the oracle's either way."""
    wl = {"unicycle1": lambda: synth.unicycle1_small(seed=15, N_nodes=256, n_visited=2000),
          "car": lambda: synth.car_trailer(seed=16, P=800, N_nodes=200, n_boxes=60, n_visited=3000, extent=10.0),
          "unicycle2": lambda: synth.unicycle2_small(seed=17, N_nodes=256)}[name]()
    p = synth.make_params(goal=wl.params.goal[:wl.desc.nx], delta=0.6, limit_branching_factor=8)
    ora, co = O.expand(wl, params=p)
    with expander_for(wl) as ex:
        ex.debug_set_limits(fp64_window=window)
        cg = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    print(name, "window", window, "counts", cg)
    assert co["dup_batch"] > 0 and co["capped"] > 0
    assert_parity(gpu, ora, f"{name} list window {window}")
    counts_consistent(cg)


def test_unicycle1_goal_hits(need_gpu):
    """Goal placed on a frontier node's neighbourhood: intermediate-goal
    checks fire and forced-novel pairs appear."""
    wl = synth.unicycle1_small(seed=12)
    goal = wl.nodes[7].copy()
    p = synth.make_params(goal=goal, delta=0.6, limit_branching_factor=-1)
    gpu, cg, ora, co = run_both(wl, params=p)
    assert co["goal_hits"] > 0
    assert_parity(gpu, ora, "goal")


@pytest.mark.parametrize("lpp", [2, 4, 8, 16, 64])
def test_goal_hits_every_pair_shape(need_gpu, lpp):
    """Eight intermediate-goal indices (the maximum) with 2 to 64 lanes per
    pair: narrow shapes test the indices in several rounds and must still
    pick the first hit in index order."""
    wl = synth.unicycle1_small(seed=14)
    goal = wl.nodes[3].copy()
    p = synth.make_params(goal=goal, delta=0.6, limit_branching_factor=5, num_check_goal=8)
    with expander_for(wl) as ex:
        ex.set_tiles(lanes_per_pair=lpp, check_blocks=64)
        cg = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    ora, co = O.expand(wl, params=p)
    assert co["goal_hits"] > 0
    assert_parity(gpu, ora, f"goal-lpp{lpp}")
    counts_consistent(cg)


def test_unicycle1_no_goal_check_alpha(need_gpu):
    wl = synth.unicycle1_small(seed=13)
    p = synth.make_params(goal=wl.params.goal[:3], alpha=0.3, delta=0.5, check_intermediate_goal=False)
    gpu, cg, ora, co = run_both(wl, params=p)
    assert_parity(gpu, ora, "alpha0.3")


def test_unicycle1_scale_prefix(need_gpu):
    """Full 4096-node batch on the GPU; the canonical order makes the first
    256 nodes' results independent of later nodes, so they must equal an
    oracle run on those 256 nodes alone."""
    wl = synth.unicycle1_scale()
    with expander_for(wl) as ex:
        cg = ex.expand(wl.nodes, wl.g_parent, wl.params)
        gpu = ex.fetch()
    counts_consistent(cg)
    n_nodes = 256
    ora, co = O.expand(wl, nodes=wl.nodes[:n_nodes], g_parent=wl.g_parent[:n_nodes])
    assert_parity(prefix(gpu, co["pairs"]), ora, "scale-prefix")
    # size-independent properties on the full batch
    assert np.all(np.diff(gpu["node"]) >= 0)
    same = gpu["node"][1:] == gpu["node"][:-1]
    assert np.all(np.diff(gpu["prim"])[same] > 0)
    free = gpu["verdict"] == N.V_FREE
    assert np.all((gpu["resolution"] != N.R_NONE) == free)
    assert np.all(gpu["first_k"][~free] >= 0)
    assert np.all(gpu["first_k"][free] == -1)


def test_unicycle2_parity(need_gpu):
    wl = synth.unicycle2_small()
    gpu, cg, ora, co = run_both(wl)
    assert co["bounds"] > 0
    assert_parity(gpu, ora, "unicycle2")


def test_car_trailer_parity(need_gpu):
    wl = synth.car_trailer(P=1500, N_nodes=512, n_boxes=300, n_visited=20_000)
    gpu, cg, ora, co = run_both(wl)
    assert co["collide"] > 0
    assert_parity(gpu, ora, "car_trailer")


def test_edge_cases(need_gpu):
    wl = synth.unicycle1_small(seed=21, N_nodes=64)
    # single node (the reference's one-node-per-iteration loop)
    gpu, cg, ora, co = run_both(wl, nodes=wl.nodes[:1], g_parent=wl.g_parent[:1])
    assert_parity(gpu, ora, "N=1")
    # empty visited set, no obstacles
    wl2 = synth.Workload(dict(wl))
    wl2["visited"] = np.zeros((0, 3))
    wl2["boxes"] = np.zeros((0, 6))
    gpu, cg, ora, co = run_both(wl2)
    assert co["collide"] == 0
    assert_parity(gpu, ora, "empty world")
    # a box covering the whole map: every in-bounds pair collides at state 0
    wl3 = synth.Workload(dict(wl))
    wl3["boxes"] = np.array([[5.0, 5.0, 0.0, 6.0, 6.0, 0.0]])
    gpu, cg, ora, co = run_both(wl3)
    assert co["free_"] == 0
    assert np.all(ora["first_k"][ora["verdict"] == N.V_COLLIDE] == 0)
    assert_parity(gpu, ora, "full box")
    # nodes whose heading has no applicable primitive (single primitive set)
    wl4 = synth.Workload(dict(wl))
    wl4["offsets"] = wl.offsets[:2].copy()
    gpu, cg, ora, co = run_both(wl4)
    assert_parity(gpu, ora, "P=1")


def test_deterministic_and_append_novel(need_gpu):
    wl = synth.unicycle1_small(seed=31)
    with expander_for(wl) as ex:
        c1 = ex.expand(wl.nodes, wl.g_parent, wl.params)
        a = ex.fetch()
        c2 = ex.expand(wl.nodes, wl.g_parent, wl.params)
        b = ex.fetch()
        assert {k: v for k, v in c1.items() if k != "resolve_rounds"} == \
            {k: v for k, v in c2.items() if k != "resolve_rounds"}
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=a[k].dtype == np.float64), k
        nv0 = ex.visited_size()
        ex.visited_append_novel()
        assert ex.visited_size() == nv0 + c1["novel"]
        ex.expand(wl.nodes, wl.g_parent, wl.params)
        r = ex.fetch()
    # T_n grown by the novel states in canonical order (dbastar.cpp:875-876)
    vis = np.concatenate([wl.visited, a["final_state"][a["resolution"] == N.R_NOVEL]])
    ora, co = O.expand(wl, visited=vis)
    assert_parity(r, ora, "after append")
