"""Primitive and problem files through the C-ABI (dp_load_primitives /
dp_save_primitives, dp_load_problem / dp_save_problem): plain numpy arrays in
the layouts of include/dynoplan_amd.h."""
import ctypes as C

import numpy as np

from . import _native as N


def load_primitives(path, system, max_motions=0, noise=0.0, seed=0, fmt=N.FMT_AUTO):
    lib = N.load()
    ps = N.PrimSet()
    N.check(lib.dp_load_primitives(str(path).encode(), fmt, system, max_motions, noise, seed, C.byref(ps)))
    try:
        P, nx, nu = ps.P, ps.nx, ps.nu
        offsets = np.ctypeslib.as_array(ps.offsets, (P + 1,)).copy()
        S = int(offsets[-1])
        states = np.ctypeslib.as_array(ps.states, (S, nx)).copy()
        actions = np.ctypeslib.as_array(ps.actions, (max(S - P, 1), nu)).copy()[: S - P]
        cost = np.ctypeslib.as_array(ps.cost, (P,)).copy()
    finally:
        lib.dp_free_primitives(C.byref(ps))
    return offsets, states, actions, cost


def reverse_primitives(offsets, states, actions, system, noise=0.0, seed=0):
    """reverse_motions (dbrrt.cpp:252-339) through dp_reverse_primitives:
    (offsets, states, actions, cost) of the reversed set."""
    lib = N.load()
    offsets = np.ascontiguousarray(offsets, np.int32)
    states = np.ascontiguousarray(states, np.float64)
    actions = np.ascontiguousarray(actions, np.float64)
    p = N.Prims()
    p.P = offsets.size - 1
    p.nx = states.shape[1]
    p.nu = actions.shape[1]
    p.offsets = offsets.ctypes.data_as(C.POINTER(C.c_int32))
    p.states = states.ctypes.data_as(C.POINTER(C.c_double))
    p.actions = actions.ctypes.data_as(C.POINTER(C.c_double))
    ps = N.PrimSet()
    N.check(lib.dp_reverse_primitives(C.byref(p), system, float(noise), int(seed), C.byref(ps)))
    try:
        P, nx, nu = ps.P, ps.nx, ps.nu
        o = np.ctypeslib.as_array(ps.offsets, (P + 1,)).copy()
        S = int(o[-1])
        st = np.ctypeslib.as_array(ps.states, (S, nx)).copy()
        ac = np.ctypeslib.as_array(ps.actions, (max(S - P, 1), nu)).copy()[: S - P]
        cost = np.ctypeslib.as_array(ps.cost, (P,)).copy()
    finally:
        lib.dp_free_primitives(C.byref(ps))
    return o, st, ac, cost


def save_primitives(path, offsets, states, actions, cost=None, fmt=N.FMT_AUTO):
    lib = N.load()
    offsets = np.ascontiguousarray(offsets, np.int32)
    states = np.ascontiguousarray(states, np.float64)
    actions = np.ascontiguousarray(actions, np.float64)
    ps = N.PrimSet()
    ps.P = offsets.size - 1
    ps.nx = states.shape[1]
    ps.nu = actions.shape[1]
    ps.offsets = offsets.ctypes.data_as(C.POINTER(C.c_int32))
    ps.states = states.ctypes.data_as(C.POINTER(C.c_double))
    ps.actions = actions.ctypes.data_as(C.POINTER(C.c_double))
    if cost is not None:
        cost = np.ascontiguousarray(cost, np.float64)
        ps.cost = cost.ctypes.data_as(C.POINTER(C.c_double))
    N.check(lib.dp_save_primitives(str(path).encode(), fmt, C.byref(ps)))


def load_problem(path):
    """dynobench problem YAML -> dict(system, robot_type, start, goal, lo, hi,
    boxes [n][6] centre + half sizes, spheres [n][4])."""
    lib = N.load()
    p = N.Problem()
    N.check(lib.dp_load_problem(str(path).encode(), C.byref(p)))
    try:
        nx = p.nx
        out = {
            "system": p.system, "robot_type": p.robot_type.decode(), "start": np.array(p.start[:nx]),
            "goal": np.array(p.goal[:nx]), "lo": np.array(p.lo[:]), "hi": np.array(p.hi[:]),
            "boxes": np.ctypeslib.as_array(p.boxes, (max(p.n_boxes, 1) * 6,))[: p.n_boxes * 6].reshape(-1, 6).copy(),
            "spheres": np.ctypeslib.as_array(p.spheres, (max(p.n_spheres, 1) * 4,))[: p.n_spheres * 4]
            .reshape(-1, 4).copy(),
        }
    finally:
        lib.dp_free_problem(C.byref(p))
    return out


def save_problem(path, system, robot_type, start, goal, lo, hi, boxes=None, spheres=None):
    lib = N.load()
    p = N.Problem()
    p.system = system
    start = np.asarray(start, np.float64)
    p.nx = start.size
    p.robot_type = robot_type.encode()
    for k in range(start.size):
        p.start[k] = float(start[k])
        p.goal[k] = float(goal[k])
    for k in range(3):
        p.lo[k] = float(lo[k]) if k < len(lo) else 0.0
        p.hi[k] = float(hi[k]) if k < len(hi) else 0.0
    boxes = np.ascontiguousarray(boxes if boxes is not None else np.zeros((0, 6)), np.float64).reshape(-1, 6)
    spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros((0, 4)), np.float64).reshape(-1, 4)
    p.n_boxes, p.n_spheres = boxes.shape[0], spheres.shape[0]
    p.boxes = boxes.ctypes.data_as(C.POINTER(C.c_double))
    p.spheres = spheres.ctypes.data_as(C.POINTER(C.c_double))
    N.check(lib.dp_save_problem(str(path).encode(), C.byref(p)))
