"""Launch shapes of the check stage: every shape gives the same pairs bit for
bit, and the autotuner keeps the fastest one it timed."""
import numpy as np
import pytest

from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

pytestmark = pytest.mark.gpu

KEYS = ("node", "prim", "verdict", "first_k", "goal_index", "resolution", "final_state", "g", "h")


def same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k], equal_nan=a[k].dtype == np.float64), k


def test_planar_shapes_and_autotune(need_gpu):
    wl = synth.car_trailer(seed=31, P=600, N_nodes=300, n_boxes=120, extent=12.0, n_visited=4000)
    with expander_for(wl) as ex:
        c0 = ex.expand(wl.nodes, wl.g_parent, wl.params)
        ref = ex.fetch()
        for lpp in (2, 4, 8, 16, 32, 64):
            for blocks in (7, 4096):
                ex.set_tiles(lanes_per_pair=lpp, check_blocks=blocks)
                c = ex.expand(wl.nodes, wl.g_parent, wl.params)
                assert {k: v for k, v in c.items() if k != "resolve_rounds"} == \
                    {k: v for k, v in c0.items() if k != "resolve_rounds"}
                same(ex.fetch(), ref)
        ex.upload_nodes(wl.nodes, wl.g_parent)
        best, trials = ex.autotune(wl.params)
        print("autotune", best, [(t["lanes_per_pair"], t["check_blocks"], round(t["check_ms"], 4)) for t in trials])
        assert len(trials) == 24 and best["check_ms"] == min(t["check_ms"] for t in trials)
        assert ex.tiles() == best
        ex.expand(wl.nodes, wl.g_parent, wl.params)
        same(ex.fetch(), ref)


def test_quadrotor_shapes_and_autotune(need_gpu):
    wl = synth.quadrotor(seed=32, P=800, N_nodes=128, n_visited=2000, thrust="balanced")
    wl.params.fp32_rollout = 1
    with expander_for(wl) as ex:
        ex.expand(wl.nodes, wl.g_parent, wl.params)
        ref = ex.fetch()
        for q32, t64, b64 in ((3, 64, 5), (8192, 128, 16384), (4096, 256, 1)):
            ex.set_tiles(quad32_blocks=q32, quad64_threads=t64, quad64_blocks=b64)
            ex.expand(wl.nodes, wl.g_parent, wl.params)
            same(ex.fetch(), ref)
        ex.upload_nodes(wl.nodes, wl.g_parent)
        best, trials = ex.autotune(wl.params)
        print("autotune", best)
        assert len(trials) == 36 and best["check_ms"] == min(t["check_ms"] for t in trials)
        ex.expand(wl.nodes, wl.g_parent, wl.params)
        same(ex.fetch(), ref)
