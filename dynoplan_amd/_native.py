"""ctypes view of include/dynoplan_amd.h.

The structures below mirror the C declarations field by field; ctypes applies
the same alignment rules as the C compiler, so their sizes match the header
(checked by tests/test_native_abi.py).  Loading fails loudly when the HIP
library is missing: there is no CPU fallback for the expansion path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdynoplan_amd.so")

DP_MAX_NX = 16

DP_OK = 0
DP_ERR_ARG = -1
DP_ERR_HIP = -2
DP_ERR_STATE = -3
DP_ERR_CAPACITY = -4
DP_ERR_IO = -5
DP_ERR_FORMAT = -6
DP_ERR_NOMEM = -7
DP_ERR_NOTCONV = -8
DP_NOALLOC_MAX_QUERIES = 65536

SYS_UNICYCLE1 = 0
SYS_UNICYCLE2 = 1
SYS_QUADROTOR = 2
SYS_CAR_TRAILER = 3

V_FREE, V_BOUNDS, V_COLLIDE, V_SURROGATE_REJECT, V_CONSTRAINT = 0, 1, 2, 3, 4
BATCH_GRAPH, BATCH_STREAMS = 0, 1
R_NONE, R_NOVEL, R_DUP_VISITED, R_DUP_BATCH, R_CAPPED = 0, 1, 2, 3, 4


class SystemDesc(C.Structure):
    _fields_ = [
        ("system", C.c_int32), ("nx", C.c_int32), ("nu", C.c_int32), ("pos_dim", C.c_int32),
        ("dt", C.c_double),
        ("x_lb", C.c_double * DP_MAX_NX), ("x_ub", C.c_double * DP_MAX_NX),
        ("dist_w", C.c_double * 4),
        ("max_vel", C.c_double), ("max_angular_vel", C.c_double),
        ("max_acc", C.c_double), ("max_angular_acc", C.c_double),
        ("body_half", C.c_double * 2), ("trailer_half", C.c_double * 2),
        ("hitch", C.c_double), ("wheelbase", C.c_double),
        ("quad", C.c_double * 8),
        ("near_eps", C.c_double),
    ]


class Prims(C.Structure):
    _fields_ = [
        ("P", C.c_int32), ("nx", C.c_int32), ("nu", C.c_int32), ("reserved", C.c_int32),
        ("offsets", C.POINTER(C.c_int32)),
        ("states", C.POINTER(C.c_double)),
        ("actions", C.POINTER(C.c_double)),
        ("cost", C.POINTER(C.c_double)),
    ]


class World(C.Structure):
    _fields_ = [
        ("n_boxes", C.c_int32), ("n_spheres", C.c_int32),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("boxes", C.POINTER(C.c_double)),
        ("spheres", C.POINTER(C.c_double)),
    ]


class ExpandParams(C.Structure):
    _fields_ = [
        ("delta", C.c_float), ("alpha", C.c_float),
        ("delta_factor_goal", C.c_double), ("cost_delta_factor", C.c_double),
        ("limit_branching_factor", C.c_int32), ("num_check_goal", C.c_int32),
        ("check_intermediate_goal", C.c_int32), ("use_surrogate", C.c_int32),
        ("fp32_rollout", C.c_int32), ("reserved", C.c_int32),
        ("goal", C.c_double * DP_MAX_NX),
    ]


class Counts(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "nodes", "pairs", "bounds", "collide", "free_", "surrogate_reject", "novel",
        "dup_visited", "dup_batch", "capped", "goal_hits", "checked", "recheck_fp64",
        "resolve_rounds", "conf_pool_used", "fp64_deferred", "constraint")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class SelectParams(C.Structure):
    _fields_ = [("radius", C.c_double), ("cost_jump", C.c_double), ("choose_first", C.c_int32),
                ("reserved", C.c_int32), ("goal", C.c_double * DP_MAX_NX)]


class SelectOut(C.Structure):
    _fields_ = [("prim", C.POINTER(C.c_int32)), ("pair", C.POINTER(C.c_int64)), ("distance", C.POINTER(C.c_double)),
                ("state", C.POINTER(C.c_double)), ("start", C.POINTER(C.c_double)), ("g", C.POINTER(C.c_double)),
                ("h", C.POINTER(C.c_double))]


class HealthInfo(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("batches", "wave_giveups", "window_overflows", "host_completions")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class PairsOut(C.Structure):
    _fields_ = [
        ("capacity", C.c_int64),
        ("node", C.POINTER(C.c_int32)), ("prim", C.POINTER(C.c_int32)),
        ("verdict", C.POINTER(C.c_int8)), ("first_k", C.POINTER(C.c_int32)),
        ("goal_index", C.POINTER(C.c_int32)), ("resolution", C.POINTER(C.c_int8)),
        ("final_state", C.POINTER(C.c_double)),
        ("g", C.POINTER(C.c_double)), ("h", C.POINTER(C.c_double)),
    ]


class PrimSet(C.Structure):
    _fields_ = [
        ("P", C.c_int32), ("nx", C.c_int32), ("nu", C.c_int32), ("reserved", C.c_int32),
        ("offsets", C.POINTER(C.c_int32)), ("states", C.POINTER(C.c_double)),
        ("actions", C.POINTER(C.c_double)), ("cost", C.POINTER(C.c_double)),
    ]


FMT_AUTO, FMT_YAML, FMT_JSON, FMT_MSGPACK, FMT_BOOST = 0, 1, 2, 3, 4


class SurrogateStats(C.Structure):
    _fields_ = [("threshold", C.c_float), ("reserved", C.c_float)] + [
        (n, C.c_int64) for n in ("pairs", "feasible", "infeasible", "rejectable", "colliding", "rejectable_colliding")]

    def as_dict(self):
        return {n: (float(getattr(self, n)) if n == "threshold" else int(getattr(self, n)))
                for n, _ in self._fields_ if n != "reserved"}

SEARCH_SOLVED, SEARCH_MAX_EXPANDS, SEARCH_EMPTY_QUEUE, SEARCH_MAX_TIME = 0, 1, 2, 3


class Problem(C.Structure):
    _fields_ = [("system", C.c_int32), ("nx", C.c_int32), ("robot_type", C.c_char * 64),
                ("start", C.c_double * DP_MAX_NX), ("goal", C.c_double * DP_MAX_NX),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("n_boxes", C.c_int32), ("n_spheres", C.c_int32),
                ("boxes", C.POINTER(C.c_double)), ("spheres", C.POINTER(C.c_double))]


class SearchOptions(C.Structure):
    _fields_ = [("delta", C.c_float), ("alpha", C.c_float), ("delta_factor_goal", C.c_double),
                ("cost_delta_factor", C.c_double), ("max_expands", C.c_int64), ("search_timelimit_ms", C.c_double),
                ("limit_branching_factor", C.c_int32), ("heuristic", C.c_int32), ("batch_nodes", C.c_int32),
                ("use_surrogate", C.c_int32), ("fp32_rollout", C.c_int32), ("reserved", C.c_int32)]


class SearchInfo(C.Structure):
    _fields_ = [("solved", C.c_int32), ("status", C.c_int32), ("cost", C.c_double), ("distance_to_goal", C.c_double)] + [
        (n, C.c_int64) for n in ("expands", "all_nodes", "closed", "rewires", "gpu_batches", "pairs", "checked",
                                 "traj_states", "traj_actions")] + [
        ("time_search_ms", C.c_double), ("time_gpu_ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


IDB_MAX_IT, IDB_TIME_LIMIT, IDB_MAX_SOLUTION = 0, 1, 2


class IdbOptions(C.Structure):
    _fields_ = [("delta_0", C.c_double), ("num_primitives_0", C.c_int64), ("delta_rate", C.c_double),
                ("num_primitives_rate", C.c_double), ("timelimit_s", C.c_double), ("max_it", C.c_int32),
                ("max_num_sol", C.c_int32), ("add_primitives", C.c_int32), ("cuts", C.c_int32), ("seed", C.c_uint64)]


class IdbInfo(C.Structure):
    _fields_ = [("solved", C.c_int32), ("exit_criteria", C.c_int32), ("iterations", C.c_int32),
                ("solutions", C.c_int32), ("cost", C.c_double), ("last_delta", C.c_double),
                ("last_primitives", C.c_int64), ("primitives_added", C.c_int64), ("expands", C.c_int64),
                ("time_ms", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class TileConfig(C.Structure):
    _fields_ = [("lanes_per_pair", C.c_int32), ("check_blocks", C.c_int32), ("quad32_blocks", C.c_int32),
                ("quad64_threads", C.c_int32), ("quad64_blocks", C.c_int32), ("check_ms", C.c_float)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# name -> (restype, argtypes); every symbol include/dynoplan_amd.h declares.
SIGNATURES = {
    "dp_last_error": (C.c_char_p, []),
    "dp_system_default": (C.c_int, [C.c_int32, C.POINTER(SystemDesc)]),
    "dp_create": (C.c_void_p, [C.POINTER(SystemDesc), C.c_int32]),
    "dp_destroy": (None, [C.c_void_p]),
    "dp_upload_primitives": (C.c_int, [C.c_void_p, C.POINTER(Prims), C.c_int32]),
    "dp_upload_world": (C.c_int, [C.c_void_p, C.POINTER(World)]),
    "dp_visited_reset": (C.c_int, [C.c_void_p]),
    "dp_visited_append": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int64]),
    "dp_visited_append_novel": (C.c_int, [C.c_void_p]),
    "dp_visited_size": (C.c_int64, [C.c_void_p]),
    "dp_visited_radius": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.c_double,
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int64]),
    "dp_visited_radius_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p,
                                            C.c_int64, C.c_void_p]),
    "dp_reserve": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "dp_upload_nodes": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32]),
    "dp_expand": (C.c_int, [C.c_void_p, C.POINTER(ExpandParams), C.c_void_p]),
    "dp_expand_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                                  C.POINTER(ExpandParams), C.c_void_p]),
    "dp_get_counts": (C.c_int, [C.c_void_p, C.POINTER(Counts)]),
    "dp_health": (C.c_int, [C.c_void_p, C.POINTER(HealthInfo)]),
    "dp_debug_set_limits": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int32]),
    "dp_fetch_pairs": (C.c_int, [C.c_void_p, C.POINTER(PairsOut)]),
    "dp_stream": (C.c_void_p, [C.c_void_p]),
    "dp_last_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "dp_last_timing_detail": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "dp_debug_rollout_states": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                          C.POINTER(C.c_double)]),
    "dp_surrogate_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "dp_set_order": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]),
    "dp_set_direction": (C.c_int, [C.c_void_p, C.c_int32]),
    "dp_set_constraints": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32]),
    "dp_expand_select": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(SelectParams), C.c_void_p]),
    "dp_fetch_select": (C.c_int, [C.c_void_p, C.POINTER(SelectOut)]),
    "dp_reverse_primitives": (C.c_int, [C.POINTER(Prims), C.c_int32, C.c_double, C.c_uint64, C.POINTER(PrimSet)]),
    "dp_search_options_default": (None, [C.POINTER(SearchOptions)]),
    "dp_search": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(SearchOptions),
                            C.POINTER(SearchInfo)]),
    "dp_search_trajectory": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "dp_search_trajectory_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "dp_load_problem": (C.c_int, [C.c_char_p, C.POINTER(Problem)]),
    "dp_free_problem": (None, [C.POINTER(Problem)]),
    "dp_save_problem": (C.c_int, [C.c_char_p, C.POINTER(Problem)]),
    "dp_idbastar_options_default": (None, [C.POINTER(IdbOptions)]),
    "dp_idbastar": (C.c_int, [C.c_void_p, C.POINTER(Prims), C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(SearchOptions), C.POINTER(IdbOptions), C.POINTER(IdbInfo)]),
    "dp_set_tiles": (C.c_int, [C.c_void_p, C.POINTER(TileConfig)]),
    "dp_get_tiles": (C.c_int, [C.c_void_p, C.POINTER(TileConfig)]),
    "dp_autotune": (C.c_int, [C.c_void_p, C.POINTER(ExpandParams), C.POINTER(TileConfig), C.POINTER(TileConfig),
                              C.POINTER(C.c_int32)]),
    "dp_debug_resolve_trace": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.c_int64]),
    "dp_batch_create": (C.c_void_p, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32]),
    "dp_batch_destroy": (None, [C.c_void_p]),
    "dp_batch_expand": (C.c_int, [C.c_void_p, C.POINTER(ExpandParams), C.c_void_p]),
    "dp_batch_latency": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "dp_batch_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "dp_load_primitives": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_uint64,
                                     C.POINTER(PrimSet)]),
    "dp_free_primitives": (None, [C.POINTER(PrimSet)]),
    "dp_save_primitives": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(PrimSet)]),
    "dp_surrogate_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "dp_surrogate_train": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.POINTER(ExpandParams),
                                     C.c_int32, C.c_int32, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "dp_surrogate_calibrate": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.POINTER(ExpandParams),
                                         C.c_float, C.c_float, C.POINTER(SurrogateStats)]),
    "dp_surrogate_set_threshold": (C.c_int, [C.c_void_p, C.c_float]),
    "dp_surrogate_scores": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    "dp_surrogate_param_count": (C.c_int64, []),
    "dp_surrogate_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "dp_debug_gemm_bf16": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32, C.c_int32,
                                     C.c_int32, C.POINTER(C.c_float)]),
}

_lib = None


class DynoplanError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dynoplan_amd error {code}: {msg}")
        self.code = code


def load():
    """Load the HIP library (build it with ``python -m dynoplan_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m dynoplan_amd.build` "
                          "(the expansion path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != DP_OK:
        raise DynoplanError(rc, load().dp_last_error().decode())
    return rc
