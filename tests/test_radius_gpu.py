"""GPU radius query over the visited set (dp_visited_radius, the rewiring
query T_n->nearestR of dbastar.cpp:883-912) against the oracle's brute
force: identical CSR lists, ascending, for every system, at the duplicate
radius, a wider one and radius 0, with queries outside the world (border
cells) and the capacity error path."""
import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

import oracle_lib as O

pytestmark = pytest.mark.gpu


def workloads():
    return {
        "unicycle1": synth.unicycle1_small(seed=21, n_visited=20000),
        "unicycle2": synth.unicycle2_small(seed=22),
        "car": synth.car_trailer(seed=23, P=200, N_nodes=64, n_visited=20000),
        "quadrotor": synth.quadrotor(seed=24, P=200, N_nodes=64, n_visited=20000),
    }


def queries_for(wl, rng, n=1500):
    vis = wl.visited
    q = np.concatenate([vis[rng.integers(0, vis.shape[0], n // 3)],
                        vis[rng.integers(0, vis.shape[0], n // 3)] + rng.normal(0, 0.05, (n // 3, vis.shape[1])),
                        wl.nodes[rng.integers(0, wl.nodes.shape[0], n - 2 * (n // 3))]])
    if wl.desc.pos_dim == 2:
        q[: n // 10, 0] = rng.uniform(wl.lo[0] - 3, wl.hi[0] + 3, n // 10)  # outside the world
        for c in ([2] if wl.desc.system != N.SYS_CAR_TRAILER else [2, 3]):
            q[:, c] = np.mod(q[:, c] + np.pi, 2 * np.pi) - np.pi
    else:
        q[: n // 10, 2] = rng.uniform(-2, 7, n // 10)
        qn = q[:, 3:7]
        q[:, 3:7] = qn / np.linalg.norm(qn, axis=1, keepdims=True)
    return np.ascontiguousarray(q)


@pytest.mark.parametrize("name", ["unicycle1", "unicycle2", "car", "quadrotor"])
def test_visited_radius_matches_oracle(need_gpu, name):
    wl = workloads()[name]
    rng = np.random.default_rng(7)
    q = queries_for(wl, rng)
    r_dup = (1.0 - float(np.float32(0.5))) * float(np.float32(0.3))
    with expander_for(wl) as ex:
        for r in (r_dup, 2.5 * r_dup, 0.0):
            off_g, idx_g = ex.visited_radius(q, r)
            off_o, idx_o = O.radius(wl.desc, wl.visited, q, r)
            print(name, r, "hits", off_o[-1])
            assert np.array_equal(off_g, off_o), (name, r)
            assert np.array_equal(idx_g, idx_o), (name, r)
        # at radius 0 every exact copy of a visited state finds itself
        copies = len({tuple(x) for x in q} & {tuple(v) for v in wl.visited})
        assert copies > 0 and off_o[-1] >= copies
        # capacity error: offsets complete, the first idx_cap entries written
        off_g, _ = ex.visited_radius(q, r_dup)
        total = int(off_g[-1])
        if total > 2:
            with pytest.raises(N.DynoplanError) as e:
                ex.visited_radius(q, r_dup, idx_cap=total - 1)
            assert e.value.code == N.DP_ERR_CAPACITY


def test_visited_radius_after_expansion_and_append(need_gpu):
    """The grid of the duplicate test serves the query after dp_expand and
    after appending the batch's NOVEL states."""
    wl = synth.unicycle1_small(seed=25)
    rng = np.random.default_rng(8)
    q = queries_for(wl, rng, 600)
    r = (1.0 - float(np.float32(0.5))) * float(np.float32(0.3))
    with expander_for(wl) as ex:
        ex.expand(wl.nodes, wl.g_parent, wl.params)
        off_g, idx_g = ex.visited_radius(q, r)
        off_o, idx_o = O.radius(wl.desc, wl.visited, q, r)
        assert np.array_equal(off_g, off_o) and np.array_equal(idx_g, idx_o)
        out = ex.fetch()
        ex.visited_append_novel()
        novel = out["final_state"][out["resolution"] == N.R_NOVEL]
        vis2 = np.concatenate([wl.visited, novel])
        assert ex.visited_size() == vis2.shape[0]
        off_g, idx_g = ex.visited_radius(q, r)
        off_o, idx_o = O.radius(wl.desc, vis2, q, r)
        assert np.array_equal(off_g, off_o) and np.array_equal(idx_g, idx_o)
