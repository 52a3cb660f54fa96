"""GPU parity of the other callers of the expansion kernels against the
oracle: a caller-given pair order (the shuffled applicable list,
dbastar.hpp:302-303), tdbastar's reverse search and constraints
(tdbastar.cpp:313-447), and dbrrt's choice of one primitive per node, forward
and over the reversed primitive set (dbrrt.cpp:252-339, 580-660).  Integer
fields and states bit-exact."""
import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import primitives as PR
from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

import oracle_lib as O
from parity_util import assert_parity, bits, counts_consistent

pytestmark = pytest.mark.gpu


def consistent(c):
    counts_consistent(c)


def workloads():
    return {
        "unicycle1": synth.unicycle1_small(P=1000, N_nodes=256),
        "car": synth.car_trailer(P=1500, N_nodes=256, n_boxes=120, n_visited=5000, extent=16.0),
        "unicycle2": synth.unicycle2_small(P=800, N_nodes=200),
    }


@pytest.fixture(scope="module")
def wls():
    return workloads()


@pytest.mark.parametrize("name", ["unicycle1", "car"])
@pytest.mark.parametrize("rows", [1, 7, 256])
def test_pair_order(need_gpu, wls, name, rows):
    wl = wls[name]
    P = wl.offsets.size - 1
    rng = np.random.default_rng(rows)
    order = np.stack([rng.permutation(P) for _ in range(rows)]).astype(np.int32)
    p = synth.make_params(goal=wl.params.goal[:wl.desc.nx], limit_branching_factor=3)
    with expander_for(wl) as ex:
        ex.set_order(order)
        c = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
        ex.set_order(None)
        c0 = ex.expand(wl.nodes, wl.g_parent, p)
        gpu0 = ex.fetch()
    ora, co = O.expand(wl, params=p, order=order)
    assert_parity(gpu, ora, f"{name} order rows={rows}")
    consistent(c)
    assert c["capped"] == co["capped"] and c["novel"] == co["novel"]
    ora0, _ = O.expand(wl, params=p)
    assert_parity(gpu0, ora0, f"{name} ascending after clearing the order")
    assert not np.array_equal(gpu["prim"], gpu0["prim"])


def test_pair_order_quadrotor(need_gpu):
    wl = synth.quadrotor(seed=43, P=2000, N_nodes=128, n_visited=3000, thrust="balanced")
    P = wl.offsets.size - 1
    order = np.random.default_rng(1).permutation(P).astype(np.int32)
    p = synth.make_params(goal=wl.params.goal, fp32_rollout=1, limit_branching_factor=5)
    with expander_for(wl) as ex:
        ex.set_order(order)
        c = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    ora, co = O.expand(wl, params=p, order=order)
    assert_parity(gpu, ora, "quadrotor order")
    counts_consistent(c)


def test_pair_order_rejects_non_permutations(need_gpu, wls):
    wl = wls["unicycle1"]
    P = wl.offsets.size - 1
    bad = np.arange(P, dtype=np.int32)
    bad[5] = 4
    with expander_for(wl) as ex:
        with pytest.raises(N.DynoplanError):
            ex.set_order(bad)
        with pytest.raises(N.DynoplanError):
            ex.set_order(np.arange(P - 1, dtype=np.int32))


@pytest.mark.parametrize("name", ["unicycle1", "car", "unicycle2"])
def test_reverse_search(need_gpu, wls, name):
    """Applicability on last states (raw), the primitive placed to end at the
    node, the new node at the first placed state.  A larger delta gives the
    reverse keys (last states with their positions) enough hits."""
    wl = wls[name]
    p = synth.make_params(delta=2.0, alpha=0.5, goal=wl.params.goal[:wl.desc.nx], limit_branching_factor=4)
    with expander_for(wl) as ex:
        ex.set_direction(False)
        c = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
    ora, co = O.expand(wl, params=p, backward=True)
    print(name, "reverse search counts", c)
    assert_parity(gpu, ora, f"{name} reverse search")
    consistent(c)
    assert co["free_"] > 0 and co["collide"] > 0


def constraints_from(wl, p, n_hit=40, n_rand=200, seed=0, backward=False):
    """Constraints on the placed states of FREE pairs at their arrival times
    (violated by construction), plus random ones."""
    ref, _ = O.expand(wl, params=p, backward=backward)
    rng = np.random.default_rng(seed)
    free = np.nonzero(ref["verdict"] == N.V_FREE)[0]
    pick = rng.choice(free, size=min(n_hit, free.size), replace=False)
    nx = wl.desc.nx
    times, states = [], []
    for q in pick:
        i, j = ref["node"][q], ref["prim"][q]
        a, b = wl.offsets[j], wl.offsets[j + 1]
        k = int(rng.integers(0, b - a - 1))
        st = wl.states[a + k].copy()
        last = wl.states[b - 1]
        st[:2] = st[:2] + (wl.nodes[i][:2] - last[:2] if backward else wl.nodes[i][:2])
        times.append(float(np.float32(wl.g_parent[i])) + k * wl.desc.dt)
        states.append(st)
    for _ in range(n_rand):
        s = wl.nodes[rng.integers(0, wl.nodes.shape[0])].copy()
        s[:2] += rng.normal(0, 0.5, 2)
        times.append(float(rng.uniform(0, 14)))
        states.append(s)
    return np.array(times), np.array(states).reshape(-1, nx)


@pytest.mark.parametrize("name,backward", [("unicycle1", False), ("car", False), ("unicycle1", True),
                                           ("unicycle2", True)])
def test_constraints(need_gpu, wls, name, backward):
    wl = wls[name]
    delta = 2.0 if backward else 0.3
    p = synth.make_params(delta=delta, alpha=0.5, goal=wl.params.goal[:wl.desc.nx], limit_branching_factor=5)
    times, states = constraints_from(wl, p, backward=backward)
    with expander_for(wl) as ex:
        ex.set_direction(not backward)
        ex.set_constraints(times, states)
        c = ex.expand(wl.nodes, wl.g_parent, p)
        gpu = ex.fetch()
        ex.set_constraints([], [])
        c_clear = ex.expand(wl.nodes, wl.g_parent, p)
    ora, co = O.expand(wl, params=p, backward=backward, constraints=(times, states))
    print(name, "backward" if backward else "forward", "constraints", len(times), "counts", c)
    assert_parity(gpu, ora, f"{name} constraints backward={backward}")
    consistent(c)
    assert co["constraint"] >= 20 and c["constraint"] == co["constraint"]
    assert c_clear["constraint"] == 0


def test_constraints_and_reverse_refused_for_quadrotor(need_gpu):
    wl = synth.quadrotor(seed=44, P=200, N_nodes=8, n_visited=10)
    with expander_for(wl) as ex:
        with pytest.raises(N.DynoplanError):
            ex.set_direction(False)
        with pytest.raises(N.DynoplanError):
            ex.set_constraints([1.0], [wl.nodes[0]])


def assert_select(gpu, ora, label):
    for k in ("prim", "pair"):
        bad = np.nonzero(gpu[k] != ora[k])[0]
        assert bad.size == 0, f"{label}: {k} differs at {bad.size} nodes, first {bad[:1]}"
    ok = ora["prim"] >= 0
    for k in ("distance", "g", "h", "state", "start"):
        a, b = gpu[k][ok], ora[k][ok]
        assert np.array_equal(bits(a), bits(b)), f"{label}: {k} differs"
        assert np.isnan(gpu[k][~ok]).all()


# The select radius equal to the float product alpha * delta of the
# comparison run below (0.5f * 0.6f), so both see the same applicable sets.
R = float(np.float32(0.3))


@pytest.mark.parametrize("name", ["unicycle1", "car", "unicycle2"])
@pytest.mark.parametrize("choose_first", [False, True])
def test_dbrrt_select(need_gpu, wls, name, choose_first):
    wl = wls[name]
    rng = np.random.default_rng(11)
    targets = wl.nodes + rng.normal(0, 1.5, wl.nodes.shape)
    P = wl.offsets.size - 1
    order = np.stack([rng.permutation(P) for _ in range(3)]).astype(np.int32)
    goal = wl.params.goal[:wl.desc.nx]
    with expander_for(wl) as ex:
        ex.set_order(order)
        got = ex.expand_select(wl.nodes, wl.g_parent, targets, R, goal, cost_jump=1.0, choose_first=choose_first)
        c = ex.counts()
        pairs = ex.fetch()
    ora, n = O.select(wl, wl.nodes, wl.g_parent, targets, R, goal, choose_first=choose_first, order=order)
    assert c["pairs"] == n
    assert_select(got, ora, f"{name} select first={choose_first}")
    assert (ora["prim"] >= 0).sum() > wl.nodes.shape[0] // 2
    # every pair's verdict is the oracle's too (no duplicate test ran)
    ref, _ = O.expand(wl, params=synth.make_params(delta=0.6, alpha=0.5, goal=goal, limit_branching_factor=-1,
                                                   check_intermediate_goal=False),
                      visited=np.zeros((0, wl.desc.nx)), order=order)
    assert np.array_equal(pairs["verdict"], ref["verdict"]) and np.array_equal(pairs["prim"], ref["prim"])
    assert (pairs["resolution"] == N.R_NONE).all()


@pytest.mark.parametrize("noise", [0.0, 1e-7])
def test_dbrrt_backward_select(need_gpu, wls, noise):
    """dbrrt's backward tree: the reversed primitive set (reverse_motions) in
    a second context, the same selection step."""
    wl = wls["unicycle1"]
    o, st, ac, cost = PR.reverse_primitives(wl.offsets, wl.states, wl.actions, N.SYS_UNICYCLE1, noise=noise, seed=3)
    rng = np.random.default_rng(12)
    targets = wl.nodes + rng.normal(0, 1.5, wl.nodes.shape)
    goal = wl.params.goal[:3]
    with expander_for(wl) as ex:
        ex.upload_primitives(o, st, ac, cost=cost)
        got = ex.expand_select(wl.nodes, wl.g_parent, targets, R, goal, cost_jump=0.5)
    ora, n = O.select(wl, wl.nodes, wl.g_parent, targets, R, goal, cost_jump=0.5, cost=cost, offsets=o, states=st,
                      actions=ac)
    assert_select(got, ora, f"dbrrt backward noise={noise}")
    assert (ora["prim"] >= 0).sum() > wl.nodes.shape[0] // 2


def test_select_then_expand_and_state_errors(need_gpu, wls):
    """dp_expand after dp_expand_select runs the full dbastar pipeline again;
    dp_fetch_select needs a fresh select; select refuses backward mode."""
    wl = wls["unicycle1"]
    with expander_for(wl) as ex:
        ex.expand_select(wl.nodes, wl.g_parent, wl.nodes, 0.3, wl.params.goal[:3])
        c = ex.expand(wl.nodes, wl.g_parent, wl.params)
        gpu = ex.fetch()
        with pytest.raises(N.DynoplanError):
            N.check(ex.lib.dp_fetch_select(ex.ctx, N.SelectOut()))
        ex.set_direction(False)
        with pytest.raises(N.DynoplanError):
            ex.expand_select(wl.nodes, wl.g_parent, wl.nodes, 0.3, wl.params.goal[:3])
    ora, _ = O.expand(wl)
    assert_parity(gpu, ora, "expand after select")
    counts_consistent(c)
