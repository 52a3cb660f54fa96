"""Python host view of the C-ABI: one GPU expansion context.

``GpuExpander`` mirrors the objects the reference's expansion loop holds
(dbastar.cpp:523-919): the primitive store behind T_m (``upload_primitives``),
the environment (``upload_world``), the visited tree T_n
(``visited_append`` / ``visited_append_novel``) and one batched pass of the
inner loop (``expand``).  Arrays cross the boundary as plain pointers.
"""
import ctypes as C

import numpy as np

from . import _native as N


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class GpuExpander:
    def __init__(self, desc, device=0):
        self.lib = N.load()
        self.desc = desc
        self.nx = desc.nx
        ctx = self.lib.dp_create(C.byref(desc), device)
        if not ctx:
            raise N.DynoplanError(N.DP_ERR_HIP, self.lib.dp_last_error().decode())
        self.ctx = C.c_void_p(ctx)
        self._keep = []

    def close(self):
        if self.ctx:
            self.lib.dp_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- inputs -------------------------------------------------------------
    def upload_primitives(self, offsets, states, actions=None, max_motions=0, cost=None):
        offsets = np.ascontiguousarray(offsets, np.int32)
        states = np.ascontiguousarray(states, np.float64)
        p = N.Prims()
        p.P = offsets.size - 1
        p.nx = states.shape[1]
        p.nu = actions.shape[1] if actions is not None and actions.ndim == 2 else 0
        p.offsets = _iptr(offsets)
        p.states = _dptr(states)
        P = offsets.size - 1 if max_motions <= 0 else min(int(max_motions), offsets.size - 1)
        self._maxlen = int(np.diff(offsets[: P + 1]).max()) if P > 0 else 0
        if actions is not None:
            actions = np.ascontiguousarray(actions, np.float64)
            p.actions = _dptr(actions)
        if cost is not None:
            cost = np.ascontiguousarray(cost, np.float64)
            p.cost = _dptr(cost)
        N.check(self.lib.dp_upload_primitives(self.ctx, C.byref(p), int(max_motions)))

    # ---- settings of the other callers (tdbastar, dbrrt) ----------------------
    def set_order(self, order):
        """Pair order per node: [rows][P] permutations (node i uses row i % rows);
        None restores ascending order."""
        if order is None:
            N.check(self.lib.dp_set_order(self.ctx, None, 0))
            return
        order = np.ascontiguousarray(np.atleast_2d(order), np.int32)
        N.check(self.lib.dp_set_order(self.ctx, _iptr(order), order.shape[0]))

    def set_direction(self, forward):
        N.check(self.lib.dp_set_direction(self.ctx, 1 if forward else 0))

    def set_constraints(self, times, states):
        times = np.ascontiguousarray(times, np.float64).reshape(-1)
        states = np.ascontiguousarray(states, np.float64).reshape(-1, self.nx)
        n = times.size
        N.check(self.lib.dp_set_constraints(self.ctx, _dptr(times) if n else None, _dptr(states) if n else None, n))

    def expand_select(self, nodes, g_parent, targets, radius, goal, cost_jump=1.0, choose_first=False):
        """dbrrt's expansion step for each node: the chosen primitive per node."""
        self.upload_nodes(nodes, g_parent)
        targets = np.ascontiguousarray(targets, np.float64).reshape(-1, self.nx)
        sp = N.SelectParams()
        sp.radius = float(radius)
        sp.cost_jump = float(cost_jump)
        sp.choose_first = 1 if choose_first else 0
        for k, v in enumerate(goal):
            sp.goal[k] = float(v)
        N.check(self.lib.dp_expand_select(self.ctx, _dptr(targets), C.byref(sp), None))
        n = targets.shape[0]
        out = {"prim": np.zeros(n, np.int32), "pair": np.zeros(n, np.int64), "distance": np.zeros(n),
               "state": np.zeros((n, self.nx)), "start": np.zeros((n, self.nx)), "g": np.zeros(n), "h": np.zeros(n)}
        so = N.SelectOut()
        so.prim = _iptr(out["prim"])
        so.pair = out["pair"].ctypes.data_as(C.POINTER(C.c_int64))
        for k in ("distance", "state", "start", "g", "h"):
            setattr(so, k, _dptr(out[k]))
        N.check(self.lib.dp_fetch_select(self.ctx, C.byref(so)))
        return out

    def upload_world(self, lo, hi, boxes, spheres=None):
        boxes = np.ascontiguousarray(boxes, np.float64).reshape(-1, 6)
        spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros((0, 4)), np.float64).reshape(-1, 4)
        w = N.World()
        w.n_boxes = boxes.shape[0]
        w.n_spheres = spheres.shape[0]
        for k in range(3):
            w.lo[k] = float(lo[k])
            w.hi[k] = float(hi[k])
        w.boxes = _dptr(boxes)
        w.spheres = _dptr(spheres)
        N.check(self.lib.dp_upload_world(self.ctx, C.byref(w)))

    def visited_reset(self):
        N.check(self.lib.dp_visited_reset(self.ctx))

    def visited_append(self, states):
        states = np.ascontiguousarray(states, np.float64).reshape(-1, self.nx)
        N.check(self.lib.dp_visited_append(self.ctx, _dptr(states), states.shape[0]))

    def visited_append_novel(self):
        N.check(self.lib.dp_visited_append_novel(self.ctx))

    def visited_size(self):
        return int(self.lib.dp_visited_size(self.ctx))

    def visited_radius(self, queries, radius, idx_cap=None):
        """T_n->nearestR for each query: (offsets[n+1], indices) in CSR form,
        each list ascending (append order of the visited set)."""
        q = np.ascontiguousarray(queries, np.float64).reshape(-1, self.nx)
        off = np.zeros(q.shape[0] + 1, np.int64)
        if idx_cap is None:
            # counts first: with no room for indices the call fills the offsets
            rc = self.lib.dp_visited_radius(self.ctx, _dptr(q), q.shape[0], float(radius),
                                            off.ctypes.data_as(C.POINTER(C.c_int64)), None, 0)
            if rc not in (N.DP_OK, N.DP_ERR_CAPACITY):
                N.check(rc)
            idx_cap = int(off[-1])
        idx = np.zeros(max(idx_cap, 1), np.int64)
        N.check(self.lib.dp_visited_radius(self.ctx, _dptr(q), q.shape[0], float(radius),
                                           off.ctypes.data_as(C.POINTER(C.c_int64)),
                                           idx.ctypes.data_as(C.POINTER(C.c_int64)), idx_cap))
        return off, idx[: off[-1]]

    def reserve(self, max_nodes, max_pairs):
        N.check(self.lib.dp_reserve(self.ctx, int(max_nodes), int(max_pairs)))

    def upload_nodes(self, nodes, g_parent=None):
        nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, self.nx)
        g = np.ascontiguousarray(g_parent if g_parent is not None else np.zeros(nodes.shape[0]), np.float64)
        N.check(self.lib.dp_upload_nodes(self.ctx, _dptr(nodes), _dptr(g), nodes.shape[0]))

    # ---- expansion ------------------------------------------------------------
    def expand_async(self, params, stream=None):
        """Enqueue one expansion of the uploaded nodes on ``stream`` (a
        hipStream_t as int, or None for the context stream)."""
        N.check(self.lib.dp_expand(self.ctx, C.byref(params), C.c_void_p(stream) if stream else None))

    def expand(self, nodes, g_parent, params):
        nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, self.nx)
        g = np.ascontiguousarray(g_parent, np.float64)
        N.check(self.lib.dp_expand_batch(self.ctx, _dptr(nodes), _dptr(g), nodes.shape[0], C.byref(params), None))
        return self.counts()

    def counts(self):
        c = N.Counts()
        N.check(self.lib.dp_get_counts(self.ctx, C.byref(c)))
        return c.as_dict()

    def health(self):
        """Never-reset totals: batches, resolution waves that gave up, FP64
        window overflows, batches completed by host-launched work."""
        h = N.HealthInfo()
        N.check(self.lib.dp_health(self.ctx, C.byref(h)))
        return h.as_dict()

    def debug_set_limits(self, conf_slots=0, pool_entries=0, fp64_window=0):
        N.check(self.lib.dp_debug_set_limits(self.ctx, int(conf_slots), int(pool_entries), int(fp64_window)))

    def timing(self):
        ms = (C.c_float * 5)()
        N.check(self.lib.dp_last_timing(self.ctx, ms))
        return {"applicability": ms[0], "check": ms[1], "resolve": ms[2], "total": ms[3], "surrogate": ms[4]}

    def timing_detail(self):
        """Device ms of the last expansion per kernel group (HIP events)."""
        ms = (C.c_float * 7)()
        N.check(self.lib.dp_last_timing_detail(self.ctx, ms))
        keys = ("applicability", "surrogate", "check_pre", "check_main", "check_post", "lists", "walk")
        return dict(zip(keys, [float(x) for x in ms]))

    def rollout_states(self, pairs, bits):
        """Every rollout state of the listed pairs of the last quadrotor batch,
        FP32 (bits=32) or FP64 (bits=64): [n][L][13], NaN past each length."""
        pairs = np.ascontiguousarray(pairs, np.int32)
        L = int(self._maxlen)
        out = np.full((pairs.size, L, 13), np.nan)
        N.check(self.lib.dp_debug_rollout_states(self.ctx, _iptr(pairs), pairs.size, int(bits), _dptr(out)))
        return out

    def tiles(self):
        t = N.TileConfig()
        N.check(self.lib.dp_get_tiles(self.ctx, C.byref(t)))
        return t.as_dict()

    def set_tiles(self, **kw):
        t = N.TileConfig()
        N.check(self.lib.dp_get_tiles(self.ctx, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        N.check(self.lib.dp_set_tiles(self.ctx, C.byref(t)))

    def autotune(self, params):
        """Pick the fastest launch shape for the uploaded nodes; returns
        (best, all trials) as dicts."""
        best = N.TileConfig()
        trials = (N.TileConfig * 64)()
        n = C.c_int32()
        N.check(self.lib.dp_autotune(self.ctx, C.byref(params), C.byref(best), trials, C.byref(n)))
        return best.as_dict(), [trials[i].as_dict() for i in range(n.value)]

    def stream(self):
        return self.lib.dp_stream(self.ctx)

    # ---- learned surrogate -------------------------------------------------------
    def surrogate_init(self, seed=1):
        N.check(self.lib.dp_surrogate_init(self.ctx, seed))

    def surrogate_train(self, nodes, params, epochs=3, batch=4096, lr=1e-3, seed=7):
        nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, self.nx)
        losses = np.zeros(epochs, np.float32)
        N.check(self.lib.dp_surrogate_train(self.ctx, _dptr(nodes), nodes.shape[0], C.byref(params), epochs, batch,
                                            lr, seed, losses.ctypes.data_as(C.POINTER(C.c_float))))
        return losses

    def surrogate_calibrate(self, nodes, params, margin=0.5, tail=3.0):
        nodes = np.ascontiguousarray(nodes, np.float64).reshape(-1, self.nx)
        st = N.SurrogateStats()
        N.check(self.lib.dp_surrogate_calibrate(self.ctx, _dptr(nodes), nodes.shape[0], C.byref(params), margin,
                                                tail, C.byref(st)))
        return st.as_dict()

    def surrogate_scores(self, n):
        out = np.zeros(n, np.float32)
        N.check(self.lib.dp_surrogate_scores(self.ctx, out.ctypes.data_as(C.POINTER(C.c_float)), n))
        return out

    # ---- db-A* search (dbastar.cpp:523-1010) ------------------------------------
    def search(self, start, goal, **options):
        """Run the host search loop; returns (info dict, states [n][nx],
        actions [m][nu]).  Options are dp_search_options fields."""
        opt = N.SearchOptions()
        self.lib.dp_search_options_default(C.byref(opt))
        for k, v in options.items():
            setattr(opt, k, v)
        start = np.ascontiguousarray(start, np.float64)
        goal = np.ascontiguousarray(goal, np.float64)
        info = N.SearchInfo()
        N.check(self.lib.dp_search(self.ctx, _dptr(start), _dptr(goal), C.byref(opt), C.byref(info)))
        d = info.as_dict()
        states = np.zeros((d["traj_states"], self.nx))
        nu = self.desc.nu
        actions = np.zeros((d["traj_actions"], nu))
        N.check(self.lib.dp_search_trajectory(self.ctx, _dptr(states) if states.size else None,
                                              _dptr(actions) if actions.size else None))
        return d, states, actions

    def idbastar(self, offsets, states, actions, start, goal, search=None, **options):
        """The iDb-A* schedule over a primitive library; returns (info dict,
        best states, best actions)."""
        offsets = np.ascontiguousarray(offsets, np.int32)
        states = np.ascontiguousarray(states, np.float64)
        actions = np.ascontiguousarray(actions, np.float64)
        pr = N.Prims()
        pr.P = offsets.size - 1
        pr.nx = states.shape[1]
        pr.nu = actions.shape[1]
        pr.offsets = _iptr(offsets)
        pr.states = _dptr(states)
        pr.actions = _dptr(actions)
        so = N.SearchOptions()
        self.lib.dp_search_options_default(C.byref(so))
        for k, v in (search or {}).items():
            setattr(so, k, v)
        io = N.IdbOptions()
        self.lib.dp_idbastar_options_default(C.byref(io))
        for k, v in options.items():
            setattr(io, k, v)
        start = np.ascontiguousarray(start, np.float64)
        goal = np.ascontiguousarray(goal, np.float64)
        info = N.IdbInfo()
        N.check(self.lib.dp_idbastar(self.ctx, C.byref(pr), _dptr(start), _dptr(goal), C.byref(so), C.byref(io),
                                     C.byref(info)))
        d = info.as_dict()
        # the best trajectory is the context's last trajectory
        st = np.zeros((0, self.nx))
        ac = np.zeros((0, self.desc.nu))
        if d["solved"]:
            n_s, n_a = self._traj_sizes()
            st = np.zeros((n_s, self.nx))
            ac = np.zeros((n_a, self.desc.nu))
            N.check(self.lib.dp_search_trajectory(self.ctx, _dptr(st) if st.size else None,
                                                  _dptr(ac) if ac.size else None))
        return d, st, ac

    def _traj_sizes(self):
        a, b = C.c_int64(), C.c_int64()
        N.check(self.lib.dp_search_trajectory_size(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def fetch(self):
        n = self.counts()["pairs"]
        out = {
            "node": np.zeros(n, np.int32), "prim": np.zeros(n, np.int32), "verdict": np.zeros(n, np.int8),
            "first_k": np.zeros(n, np.int32), "goal_index": np.zeros(n, np.int32),
            "resolution": np.zeros(n, np.int8), "final_state": np.zeros((n, self.nx)),
            "g": np.zeros(n), "h": np.zeros(n),
        }
        po = pairs_out(out, n)
        N.check(self.lib.dp_fetch_pairs(self.ctx, C.byref(po)))
        return out


class GpuBatch:
    """Several GpuExpander contexts (independent planning instances) expanded
    together: one hipGraph launch (mode N.BATCH_GRAPH) or one stream per
    instance (N.BATCH_STREAMS)."""

    def __init__(self, expanders, mode=N.BATCH_GRAPH):
        self.lib = N.load()
        self.expanders = list(expanders)
        arr = (C.c_void_p * len(self.expanders))(*[e.ctx.value for e in self.expanders])
        h = self.lib.dp_batch_create(arr, len(self.expanders), mode)
        if not h:
            raise N.DynoplanError(N.DP_ERR_ARG, self.lib.dp_last_error().decode())
        self.h = C.c_void_p(h)

    def expand_async(self, params, stream=None):
        arr = (N.ExpandParams * len(self.expanders))(*params)
        N.check(self.lib.dp_batch_expand(self.h, arr, C.c_void_p(stream) if stream else None))

    def latency_ms(self):
        ms = np.zeros(len(self.expanders), np.float32)
        N.check(self.lib.dp_batch_latency(self.h, ms.ctypes.data_as(C.POINTER(C.c_float))))
        return ms

    def stats(self):
        a, b = C.c_int64(), C.c_int64()
        N.check(self.lib.dp_batch_stats(self.h, C.byref(a), C.byref(b)))
        return {"captures": a.value, "launches": b.value}

    def close(self):
        if self.h:
            self.lib.dp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pairs_out(out, n):
    po = N.PairsOut()
    po.capacity = n
    po.node = _iptr(out["node"])
    po.prim = _iptr(out["prim"])
    po.verdict = out["verdict"].ctypes.data_as(C.POINTER(C.c_int8))
    po.first_k = _iptr(out["first_k"])
    po.goal_index = _iptr(out["goal_index"])
    po.resolution = out["resolution"].ctypes.data_as(C.POINTER(C.c_int8))
    po.final_state = _dptr(out["final_state"])
    po.g = _dptr(out["g"])
    po.h = _dptr(out["h"])
    return po


def expander_for(workload, device=0):
    """Context loaded with a synth.Workload's primitives, world and visited set."""
    ex = GpuExpander(workload.desc, device)
    ex.upload_primitives(workload.offsets, workload.states, workload.actions)
    ex.upload_world(workload.lo, workload.hi, workload.boxes, workload.spheres)
    if workload.visited.shape[0]:
        ex.visited_append(workload.visited)
    return ex
