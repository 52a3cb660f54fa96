"""db-A* search on the GPU expansion path (dp_search) against the CPU oracle
search, on a bugtrap map with the reference's in-tree unicycle1 primitives.

- one node per iteration: the whole search (expansions, nodes, rewires,
  solution cost as float, trajectory) equals the oracle search bit for bit;
- solved with cost < 100, as test_dbastar_0.cpp:118-163 asks of its bugtrap;
- the trajectory starts within alpha*delta of the start, ends within the goal
  tolerance, stays in bounds and collision-free, and follows the unicycle
  dynamics inside every primitive (jumps between primitives bounded by delta);
- many nodes per iteration (batched frontier): still solved and valid.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
import oracle_search as OS
from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
START = np.array([3.8, 3.0, 0.0])
GOAL = np.array([5.2, 3.0, 0.0])


def bugtrap():
    d = np.load(os.path.join(GOLDEN, "unicycle1_ref_prims.npz"))
    desc = synth.default_desc_static(N.SYS_UNICYCLE1)
    # a U-shaped trap open to the left around the start, goal outside on the right
    boxes = np.array([[3.0, 4.2, 0, 1.2, 0.1, 0], [3.0, 1.8, 0, 1.2, 0.1, 0], [4.2, 3.0, 0, 0.1, 1.3, 0]])
    return synth.Workload(desc=desc, offsets=d["offsets"].astype(np.int32), states=d["states"], actions=d["actions"],
                          lo=np.array([0.0, 0, 0]), hi=np.array([6.0, 6, 0]), boxes=boxes, spheres=np.zeros((0, 4)),
                          nodes=START[None, :], g_parent=np.zeros(1), visited=np.zeros((0, 3)),
                          params=synth.make_params(goal=GOAL), seed=0)


def check_trajectory(wl, states, actions, delta, alpha=0.5):
    lib = O.load()
    d = wl.desc
    w = O.world_struct(wl.lo, wl.hi, wl.boxes, wl.spheres)
    dist = lambda a, b: lib.oracle_distance(C.byref(d), O.dptr(np.ascontiguousarray(a)), O.dptr(np.ascontiguousarray(b)))
    assert states.shape[0] == actions.shape[0] + 1
    assert dist(states[0], START) <= alpha * delta + 1e-12
    assert dist(states[-1], GOAL) < delta
    for s in states:
        s = np.ascontiguousarray(s)
        assert lib.oracle_state_valid(C.byref(d), C.byref(w), O.dptr(s))
        assert not lib.oracle_state_collides(C.byref(d), C.byref(w), O.dptr(s))
    out = np.zeros(3)
    jumps = 0
    for k in range(actions.shape[0]):
        lib.oracle_unicycle1_step(C.byref(d), O.dptr(np.ascontiguousarray(states[k])),
                                  O.dptr(np.ascontiguousarray(actions[k])), O.dptr(out))
        e = dist(out, states[k + 1])
        if e > 1e-9:
            jumps += 1
            assert e <= delta + 1e-9
    return jumps


@pytest.mark.parametrize("delta", [0.8, 0.5])
def test_search_equals_oracle_search(need_gpu, delta):
    wl = bugtrap()
    ref, rs, ra = OS.search(wl, START, GOAL, delta=delta, max_expands=10000)
    with expander_for(wl) as ex:
        info, st, ac = ex.search(START, GOAL, delta=delta, max_expands=10000, batch_nodes=1)
    print(f"delta {delta}: gpu {info}\n oracle {ref}")
    assert info["solved"] == 1 and info["cost"] < 100.0  # test_dbastar_0.cpp:160-161
    for k in ("solved", "expands", "all_nodes", "rewires", "cost"):
        assert info[k] == ref[k], k
    assert info["status"] == ref["status"] == N.SEARCH_SOLVED
    assert np.array_equal(st, rs) and np.array_equal(ac, ra)
    jumps = check_trajectory(wl, st, ac, delta)
    print(f"trajectory: {st.shape[0]} states, {jumps} primitive junctions with a jump")


def test_search_batched_frontier(need_gpu):
    wl = bugtrap()
    with expander_for(wl) as ex:
        info, st, ac = ex.search(START, GOAL, delta=0.5, max_expands=20000, batch_nodes=64)
        print(f"64 nodes per expansion: {info}")
        assert info["solved"] == 1 and info["cost"] < 100.0
        check_trajectory(wl, st, ac, 0.5)
        # a limit the search cannot meet ends with MAX_EXPANDS and the closest node
        info2, st2, _ = ex.search(START, GOAL, delta=0.5, max_expands=5, batch_nodes=1)
        assert info2["status"] == N.SEARCH_MAX_EXPANDS and info2["solved"] == 0 and info2["expands"] == 5
        assert st2.shape[0] >= 1


def test_idbastar_equals_oracle_over_iterations(need_gpu):
    """Three iterations of iDb-A* (idbastar.cpp:128-312), each solved, with
    the solution pieces inserted before the library: every iteration's
    search runs over the primitives the oracle restatement builds, so the
    summed expansions, the schedule's last delta and primitive count, the
    best cost and the best trajectory all match bit for bit."""
    wl = bugtrap()
    kw = dict(delta_0=0.8, num_primitives_0=60, num_primitives_rate=1.5, delta_rate=0.9, max_it=4, max_num_sol=3)
    ref, rs, ra = OS.idbastar(wl, START, GOAL, max_expands=10000, **kw)
    with expander_for(wl) as ex:
        info, st, ac = ex.idbastar(wl.offsets, wl.states, wl.actions, START, GOAL, search={"max_expands": 10000},
                                   timelimit_s=600.0, **kw)
    print(f"iDb-A*: gpu {info}\n oracle {ref}")
    assert ref["iterations"] >= 2 and ref["primitives_added"] > 0
    for k in ("solved", "iterations", "solutions", "cost", "last_delta", "last_primitives", "primitives_added",
              "expands", "exit_criteria"):
        assert info[k] == ref[k], k
    assert np.array_equal(st, rs) and np.array_equal(ac, ra)


def test_idbastar_schedule(need_gpu):
    """iDb-A* (idbastar.cpp:128-312): the first iteration is the plain search
    at delta_0 with the first num_primitives_0 primitives; later iterations
    follow the schedule and reuse solution pieces as primitives."""
    wl = bugtrap()
    with expander_for(wl) as ex:
        one, st1, ac1 = ex.idbastar(wl.offsets, wl.states, wl.actions, START, GOAL, search={"max_expands": 10000},
                                    delta_0=0.8, num_primitives_0=100, max_it=1, timelimit_s=120.0)
        ref, rs, ra = ex.search(START, GOAL, delta=0.8, max_expands=10000)
        assert one["solved"] == ref["solved"] == 1 and one["iterations"] == 1
        assert one["cost"] == ref["cost"] and one["exit_criteria"] == N.IDB_MAX_IT
        assert np.array_equal(st1, rs) and np.array_equal(ac1, ra)
        info, st, ac = ex.idbastar(wl.offsets, wl.states, wl.actions, START, GOAL, search={"max_expands": 10000},
                                   delta_0=0.8, num_primitives_0=60, num_primitives_rate=1.5, delta_rate=0.9,
                                   max_it=4, max_num_sol=3, timelimit_s=120.0)
    print(f"iDb-A*: {info}")
    assert info["solved"] == 1 and info["cost"] < 100.0
    k = info["iterations"]
    assert 1 <= k <= 4
    assert info["exit_criteria"] in (N.IDB_MAX_IT, N.IDB_MAX_SOLUTION)
    if info["exit_criteria"] == N.IDB_MAX_SOLUTION:
        assert info["solutions"] == 3
    assert info["primitives_added"] == 5 * info["solutions"]
    assert info["last_primitives"] <= min(int(60 * 1.5 ** (k - 1)), 100 + info["primitives_added"])
    assert info["last_delta"] <= 0.8
    check_trajectory(wl, st, ac, 0.8)  # found at some delta <= delta_0
