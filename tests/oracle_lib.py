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


class OracleExt(C.Structure):
    _fields_ = [("order", C.POINTER(C.c_int32)), ("order_rows", C.c_int32), ("backward", C.c_int32),
                ("n_constraints", C.c_int32), ("reserved", C.c_int32), ("con_times", C.POINTER(C.c_double)),
                ("con_states", C.POINTER(C.c_double))]


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
    lib.oracle_expand_ext.restype = C.c_int64
    lib.oracle_expand_ext.argtypes = lib.oracle_expand.argtypes[:-2] + [P(OracleExt), P(N.PairsOut), P(N.Counts)]
    lib.oracle_constraint_index.restype = C.c_int32
    lib.oracle_constraint_index.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32]
    lib.oracle_select.restype = C.c_int64
    lib.oracle_select.argtypes = [P(N.SystemDesc), P(N.Prims), C.c_int32, P(N.World), P(C.c_double), P(C.c_double),
                                  C.c_int32, C.c_double, P(C.c_double), C.c_int32, C.c_double, P(C.c_double),
                                  P(C.c_int32), C.c_int32, P(C.c_int32), P(C.c_int64), P(C.c_double),
                                  P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]
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
           max_motions=0, threads=0, want_out=True, cap=None, order=None, backward=False, constraints=None):
    """Run the oracle on a synth.Workload (optionally overriding parts).
    threads > 0 runs oracle_expand_mt on that many OpenMP threads; with
    want_out False only the counts are produced (timing runs).  order
    ([rows][P]), backward and constraints ((times, states)) select the
    settings of the other callers (oracle_expand_ext, single-threaded)."""
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
    if order is not None or backward or constraints is not None:
        ext = OracleExt()
        keep = []
        if order is not None:
            order = np.ascontiguousarray(np.atleast_2d(order), np.int32)
            keep.append(order)
            ext.order = order.ctypes.data_as(C.POINTER(C.c_int32))
            ext.order_rows = order.shape[0]
        ext.backward = 1 if backward else 0
        if constraints is not None:
            ct = np.ascontiguousarray(constraints[0], np.float64).reshape(-1)
            cs = np.ascontiguousarray(constraints[1], np.float64).reshape(-1, nx)
            keep += [ct, cs]
            ext.n_constraints = ct.size
            ext.con_times = dptr(ct)
            ext.con_states = dptr(cs)
        n = lib.oracle_expand_ext(*args[:-2], C.byref(ext), *args[-2:])
    else:
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


def select(wl, nodes, g_parent, targets, radius, goal, cost_jump=1.0, choose_first=False, order=None, cost=None,
           offsets=None, states=None, actions=None):
    """dbrrt's expansion step per node (oracle_select); the primitive arrays
    default to the workload's."""
    lib = load()
    nx = wl.desc.nx
    nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, nx)
    g_parent = np.ascontiguousarray(g_parent, np.float64)
    targets = np.ascontiguousarray(targets, np.float64).reshape(-1, nx)
    offsets = np.ascontiguousarray(wl.offsets if offsets is None else offsets, np.int32)
    states = np.ascontiguousarray(wl.states if states is None else states, np.float64)
    actions = wl.actions if actions is None else actions
    actions = np.ascontiguousarray(actions, np.float64) if actions is not None else None
    boxes = np.ascontiguousarray(wl.boxes, np.float64)
    spheres = np.ascontiguousarray(wl.spheres, np.float64)
    pr = prims_struct(offsets, states, actions)
    if cost is not None:
        cost = np.ascontiguousarray(cost, np.float64)
        pr.cost = dptr(cost)
    w = world_struct(wl.lo, wl.hi, boxes, spheres)
    goal = np.ascontiguousarray(np.resize(np.asarray(goal, np.float64), nx))
    n = nodes.shape[0]
    out = {"prim": np.zeros(n, np.int32), "pair": np.zeros(n, np.int64), "distance": np.zeros(n),
           "state": np.zeros((n, nx)), "start": np.zeros((n, nx)), "g": np.zeros(n), "h": np.zeros(n)}
    rows = 0
    optr = None
    if order is not None:
        order = np.ascontiguousarray(np.atleast_2d(order), np.int32)
        rows = order.shape[0]
        optr = order.ctypes.data_as(C.POINTER(C.c_int32))
    npairs = lib.oracle_select(C.byref(wl.desc), C.byref(pr), 0, C.byref(w), dptr(nodes), dptr(g_parent), n,
                               float(radius), dptr(targets), 1 if choose_first else 0, float(cost_jump), dptr(goal),
                               optr, rows, out["prim"].ctypes.data_as(C.POINTER(C.c_int32)),
                               out["pair"].ctypes.data_as(C.POINTER(C.c_int64)), dptr(out["distance"]),
                               dptr(out["state"]), dptr(out["start"]), dptr(out["g"]), dptr(out["h"]))
    if npairs < 0:
        raise RuntimeError(f"oracle_select failed: {npairs}")
    return out, int(npairs)
