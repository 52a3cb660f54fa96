"""ctypes binding of oracle/liboracle.so (the CPU FP64 checker).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from dynoplan_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

NN_BRUTE, NN_GRID = 0, 1
COL_BRUTE, COL_SWEPT = 0, 1

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    lib = C.CDLL(ORACLE_SO)
    P = C.POINTER
    lib.oracle_expand.restype = C.c_int64
    lib.oracle_expand.argtypes = [P(N.SystemDesc), P(N.Prims), C.c_int32, P(N.World), P(C.c_double), C.c_int64,
                                  P(C.c_double), P(C.c_double), C.c_int32, P(N.ExpandParams), C.c_int32,
                                  C.c_int32, P(N.PairsOut), P(N.Counts)]
    lib.oracle_expand_mt.restype = C.c_int64
    lib.oracle_expand_mt.argtypes = lib.oracle_expand.argtypes + [C.c_int32]
    lib.oracle_radius.restype = C.c_int64
    lib.oracle_radius.argtypes = [P(N.SystemDesc), P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, C.c_double,
                                  P(C.c_int64), P(C.c_int64), C.c_int64]
    lib.oracle_distance.restype = C.c_double
    lib.oracle_distance.argtypes = [P(N.SystemDesc), P(C.c_double), P(C.c_double)]
    lib.oracle_lower_bound_time.restype = C.c_double
    lib.oracle_lower_bound_time.argtypes = [P(N.SystemDesc), P(C.c_double), P(C.c_double)]
    lib.oracle_state_valid.restype = C.c_int
    lib.oracle_state_valid.argtypes = [P(N.SystemDesc), P(N.World), P(C.c_double)]
    lib.oracle_state_collides.restype = C.c_int
    lib.oracle_state_collides.argtypes = [P(N.SystemDesc), P(N.World), P(C.c_double)]
    lib.oracle_obb_aabb.restype = C.c_int
    lib.oracle_obb_aabb.argtypes = [C.c_double] * 10
    lib.oracle_quad_step.restype = None
    lib.oracle_quad_step.argtypes = [P(N.SystemDesc), P(C.c_double), P(C.c_double), P(C.c_double)]
    lib.oracle_unicycle1_step.restype = None
    lib.oracle_unicycle1_step.argtypes = [P(N.SystemDesc), P(C.c_double), P(C.c_double), P(C.c_double)]
    lib.oracle_acos01.restype = C.c_double
    lib.oracle_acos01.argtypes = [C.c_double]
    lib.oracle_cos_sinc.restype = None
    lib.oracle_cos_sinc.argtypes = [C.c_double, P(C.c_double), P(C.c_double)]
    lib.oracle_goal_index.restype = C.c_int32
    lib.oracle_goal_index.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    lib.oracle_transform_primitive.restype = C.c_int32
    lib.oracle_transform_primitive.argtypes = [P(N.SystemDesc), P(N.Prims), C.c_int32, P(C.c_double),
                                               P(C.c_double)]
    _lib = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def prims_struct(offsets, states, actions):
    p = N.Prims()
    p.P = offsets.size - 1
    p.nx = states.shape[1]
    p.nu = actions.shape[1] if actions is not None else 0
    p.offsets = offsets.ctypes.data_as(C.POINTER(C.c_int32))
    p.states = dptr(states)
    p.actions = dptr(actions)
    return p


def world_struct(lo, hi, boxes, spheres):
    w = N.World()
    w.n_boxes = boxes.shape[0]
    w.n_spheres = spheres.shape[0]
    for k in range(3):
        w.lo[k] = float(lo[k])
        w.hi[k] = float(hi[k])
    w.boxes = dptr(boxes)
    w.spheres = dptr(spheres)
    return w


def expand(wl, nodes=None, g_parent=None, visited=None, params=None, nn_mode=NN_GRID, col_mode=COL_SWEPT,
           max_motions=0, threads=0, want_out=True, cap=None):
    """Run the oracle on a synth.Workload (optionally overriding parts).
    threads > 0 runs oracle_expand_mt on that many OpenMP threads; with
    want_out False only the counts are produced (timing runs)."""
    lib = load()
    nodes = np.ascontiguousarray(wl.nodes if nodes is None else nodes, np.float64)
    g_parent = np.ascontiguousarray(wl.g_parent if g_parent is None else g_parent, np.float64)
    visited = np.ascontiguousarray(wl.visited if visited is None else visited, np.float64).reshape(-1, wl.desc.nx)
    params = wl.params if params is None else params
    offsets = np.ascontiguousarray(wl.offsets, np.int32)
    states = np.ascontiguousarray(wl.states, np.float64)
    actions = np.ascontiguousarray(wl.actions, np.float64) if wl.actions is not None else None
    boxes = np.ascontiguousarray(wl.boxes, np.float64)
    spheres = np.ascontiguousarray(wl.spheres, np.float64)
    pr = prims_struct(offsets, states, actions)
    w = world_struct(wl.lo, wl.hi, boxes, spheres)
    nx = wl.desc.nx
    # capacity: at most N * P pairs
    P = offsets.size - 1 if max_motions <= 0 else min(max_motions, offsets.size - 1)
    if cap is None:
        cap = max(1, nodes.shape[0] * P) if want_out else 1
    out = {
        "node": np.zeros(cap, np.int32), "prim": np.zeros(cap, np.int32), "verdict": np.zeros(cap, np.int8),
        "first_k": np.zeros(cap, np.int32), "goal_index": np.zeros(cap, np.int32),
        "resolution": np.zeros(cap, np.int8), "final_state": np.zeros((cap, nx)),
        "g": np.zeros(cap), "h": np.zeros(cap),
    }
    from dynoplan_amd.expander import pairs_out
    po = pairs_out(out, cap)
    cnt = N.Counts()
    args = (C.byref(wl.desc), C.byref(pr), max_motions, C.byref(w), dptr(visited), visited.shape[0],
            dptr(nodes), dptr(g_parent), nodes.shape[0], C.byref(params), nn_mode, col_mode,
            C.byref(po) if want_out else None, C.byref(cnt))
    n = lib.oracle_expand_mt(*args, threads) if threads > 0 else lib.oracle_expand(*args)
    if n < 0:
        raise RuntimeError(f"oracle_expand failed: {n}")
    return {k: v[:n] for k, v in out.items()}, cnt.as_dict()


def radius(desc, visited, queries, r):
    """Brute-force T_n->nearestR of every query: (offsets, indices), ascending."""
    lib = load()
    visited = np.ascontiguousarray(visited, np.float64).reshape(-1, desc.nx)
    queries = np.ascontiguousarray(queries, np.float64).reshape(-1, desc.nx)
    off = np.zeros(queries.shape[0] + 1, np.int64)
    total = lib.oracle_radius(C.byref(desc), dptr(visited), visited.shape[0], dptr(queries), queries.shape[0],
                              float(r), off.ctypes.data_as(C.POINTER(C.c_int64)), None, 0)
    idx = np.zeros(max(total, 1), np.int64)
    lib.oracle_radius(C.byref(desc), dptr(visited), visited.shape[0], dptr(queries), queries.shape[0], float(r),
                      off.ctypes.data_as(C.POINTER(C.c_int64)), idx.ctypes.data_as(C.POINTER(C.c_int64)), total)
    return off, idx[:total]
