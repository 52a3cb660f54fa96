"""Primitive files through the C-ABI loader (dp_load_primitives /
dp_save_primitives): returns plain numpy arrays in the dp_prims layout."""
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
