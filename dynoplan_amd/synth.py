"""Seed-deterministic synthetic workloads of the expansion path.

Each generator returns plain numpy arrays in the layouts of
include/dynoplan_amd.h (row-major states, int32 offsets).  The workloads
follow the configuration list the benchmark is quoted on:

  unicycle1_small   unicycle1_v0, 1000 primitives of 5-20 steps, 10x10 m map,
                    20 boxes (sides U[0.5,1.5]), 256 frontier nodes in free
                    space, delta 0.3, FP64
  unicycle1_scale   5000 primitives of 5-30 steps, 4096 nodes, 200 boxes on
                    20x20 m, 100k random visited states, delta 0.3, FP64
  quadrotor         quadrotor_v0, 10k primitives of 10-50 steps, 10x10x5 m
                    world, 100 boxes + 50 spheres (r U[0.2,0.8]), 0.1 m robot,
                    2048 nodes
  car_trailer       car + 1 trailer, 5000 primitives of 10-40 steps, 300
                    boxes on 30x30 m, 4096 nodes

Primitives are canonical (start position at the origin, as every primitive
of the reference's data/ fixtures) and are rolled out with the system
dynamics in FP64; controls are piecewise constant.
"""
import math

import numpy as np

from . import _native as N

TWO_PI = 2.0 * math.pi


def wrap_angle(a):
    """Wrap to (-pi, pi]."""
    a = np.where(a > math.pi, a - TWO_PI, a)
    return np.where(a <= -math.pi, a + TWO_PI, a)


def _piecewise_controls(rng, P, L, lo, hi, max_pieces=3):
    """P x L x nu piecewise-constant controls with 1..max_pieces pieces."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    nu = lo.size
    u = np.empty((P, L, nu))
    pieces = rng.integers(1, max_pieces + 1, size=P)
    for j in range(P):
        cuts = np.sort(rng.integers(1, L, size=pieces[j] - 1)) if L > 1 else np.array([], int)
        bounds = np.concatenate([[0], cuts, [L]])
        for a, b in zip(bounds[:-1], bounds[1:]):
            u[j, a:b] = rng.uniform(lo, hi)
    return u


def _pack(states_list, actions_list):
    lens = np.array([s.shape[0] for s in states_list], np.int32)
    offsets = np.zeros(len(lens) + 1, np.int32)
    np.cumsum(lens, out=offsets[1:])
    states = np.ascontiguousarray(np.concatenate(states_list, axis=0))
    actions = np.ascontiguousarray(np.concatenate(actions_list, axis=0))
    return offsets, states, actions


def unicycle1_primitives(rng, P, lmin, lmax, dt=0.1, vmax=0.5, wmax=0.5):
    """x+ = x + dt v cos(th), y+ = y + dt v sin(th), th+ = wrap(th + dt w)."""
    L = rng.integers(lmin, lmax + 1, size=P)
    Lmax = int(L.max())
    u = _piecewise_controls(rng, P, Lmax, [-vmax, -wmax], [vmax, wmax])
    x = np.zeros((P, Lmax + 1, 3))
    x[:, 0, 2] = wrap_angle(rng.uniform(-math.pi, math.pi, size=P))
    for k in range(Lmax):
        th = x[:, k, 2]
        x[:, k + 1, 0] = x[:, k, 0] + dt * (u[:, k, 0] * np.cos(th))
        x[:, k + 1, 1] = x[:, k, 1] + dt * (u[:, k, 0] * np.sin(th))
        x[:, k + 1, 2] = wrap_angle(th + dt * u[:, k, 1])
    return _pack([x[j, : L[j] + 1] for j in range(P)], [u[j, : L[j]] for j in range(P)])


def unicycle2_primitives(rng, P, lmin, lmax, dt=0.1, amax=0.25, vmax=0.5):
    """unicycle2_v0: (x y th v w), controls (a, alpha)."""
    L = rng.integers(lmin, lmax + 1, size=P)
    Lmax = int(L.max())
    u = _piecewise_controls(rng, P, Lmax, [-amax, -amax], [amax, amax])
    x = np.zeros((P, Lmax + 1, 5))
    x[:, 0, 2] = wrap_angle(rng.uniform(-math.pi, math.pi, size=P))
    x[:, 0, 3] = rng.uniform(-0.3, 0.3, size=P)
    x[:, 0, 4] = rng.uniform(-0.3, 0.3, size=P)
    for k in range(Lmax):
        th, v, w = x[:, k, 2], x[:, k, 3], x[:, k, 4]
        x[:, k + 1, 0] = x[:, k, 0] + dt * (v * np.cos(th))
        x[:, k + 1, 1] = x[:, k, 1] + dt * (v * np.sin(th))
        x[:, k + 1, 2] = wrap_angle(th + dt * w)
        x[:, k + 1, 3] = np.clip(v + dt * u[:, k, 0], -vmax, vmax)
        x[:, k + 1, 4] = np.clip(w + dt * u[:, k, 1], -vmax, vmax)
    return _pack([x[j, : L[j] + 1] for j in range(P)], [u[j, : L[j]] for j in range(P)])


def car_trailer_primitives(rng, P, lmin, lmax, dt=0.1, wheelbase=0.25, hitch=0.5, vmax=0.5, phimax=math.pi / 3):
    """car_first_order_with_1_trailers: (x y th0 th1), controls (v, phi)."""
    L = rng.integers(lmin, lmax + 1, size=P)
    Lmax = int(L.max())
    u = _piecewise_controls(rng, P, Lmax, [-vmax, -phimax], [vmax, phimax])
    x = np.zeros((P, Lmax + 1, 4))
    x[:, 0, 2] = wrap_angle(rng.uniform(-math.pi, math.pi, size=P))
    x[:, 0, 3] = wrap_angle(x[:, 0, 2] + rng.uniform(-0.5, 0.5, size=P))
    for k in range(Lmax):
        th0, th1 = x[:, k, 2], x[:, k, 3]
        v, phi = u[:, k, 0], u[:, k, 1]
        x[:, k + 1, 0] = x[:, k, 0] + dt * (v * np.cos(th0))
        x[:, k + 1, 1] = x[:, k, 1] + dt * (v * np.sin(th0))
        x[:, k + 1, 2] = wrap_angle(th0 + dt * (v / wheelbase * np.tan(phi)))
        x[:, k + 1, 3] = wrap_angle(th1 + dt * (v / hitch * np.sin(th0 - th1)))
    return _pack([x[j, : L[j] + 1] for j in range(P)], [u[j, : L[j]] for j in range(P)])


def random_boxes(rng, n, lo, hi, side_lo=0.5, side_hi=1.5, dim=2):
    b = np.zeros((n, 6))
    for a in range(dim):
        b[:, a] = rng.uniform(lo[a], hi[a], size=n)
        b[:, 3 + a] = 0.5 * rng.uniform(side_lo, side_hi, size=n)
    return b


def planar_footprint_collides(desc, boxes, states):
    """Vectorised separating-axis test of planar footprints against AABBs
    (used only to place frontier nodes in free space)."""
    s = np.asarray(states, float)
    hit = np.zeros(s.shape[0], bool)

    def obb(cx, cy, th, hl, hw):
        c, sn = np.cos(th)[:, None], np.sin(th)[:, None]
        tx = boxes[None, :, 0] - cx[:, None]
        ty = boxes[None, :, 1] - cy[:, None]
        hx, hy = boxes[None, :, 3], boxes[None, :, 4]
        ac, asn = np.abs(c), np.abs(sn)
        sep = np.abs(tx) > hx + (hl * ac + hw * asn)
        sep |= np.abs(ty) > hy + (hl * asn + hw * ac)
        sep |= np.abs(tx * c + ty * sn) > hl + (hx * ac + hy * asn)
        sep |= np.abs(ty * c - tx * sn) > hw + (hx * asn + hy * ac)
        return (~sep).any(axis=1)

    hit |= obb(s[:, 0], s[:, 1], s[:, 2], desc.body_half[0], desc.body_half[1])
    if desc.system == N.SYS_CAR_TRAILER:
        th1 = s[:, 3]
        hit |= obb(s[:, 0] - desc.hitch * np.cos(th1), s[:, 1] - desc.hitch * np.sin(th1), th1,
                   desc.trailer_half[0], desc.trailer_half[1])
    return hit


def free_planar_nodes(rng, desc, boxes, lo, hi, n, extra=None):
    out = []
    while sum(len(o) for o in out) < n:
        m = 4 * n
        s = np.zeros((m, desc.nx))
        s[:, 0] = rng.uniform(lo[0], hi[0], size=m)
        s[:, 1] = rng.uniform(lo[1], hi[1], size=m)
        s[:, 2] = rng.uniform(-math.pi, math.pi, size=m)
        if extra is not None:
            extra(s)
        ok = ~planar_footprint_collides(desc, boxes, s)
        out.append(s[ok])
    return np.ascontiguousarray(np.concatenate(out)[:n])


def system_desc(system):
    lib = N.load()
    d = N.SystemDesc()
    N.check(lib.dp_system_default(system, d))
    return d


def default_desc_static(system):
    """The same defaults as dp_system_default, without loading the library
    (the CPU tests build descriptions for the oracle this way)."""
    d = N.SystemDesc()
    for k in range(N.DP_MAX_NX):
        d.x_lb[k] = -math.inf
        d.x_ub[k] = math.inf
    d.system = system
    d.near_eps = 1e-4
    if system == N.SYS_UNICYCLE1:
        d.nx, d.nu, d.pos_dim, d.dt = 3, 2, 2, 0.1
        d.dist_w[0], d.dist_w[1] = 1.0, 0.5
        d.max_vel, d.max_angular_vel = 0.5, 0.5
        d.body_half[0], d.body_half[1] = 0.25, 0.125
    elif system == N.SYS_UNICYCLE2:
        d.nx, d.nu, d.pos_dim, d.dt = 5, 2, 2, 0.1
        d.x_lb[3], d.x_ub[3], d.x_lb[4], d.x_ub[4] = -0.5, 0.5, -0.5, 0.5
        for k, w in enumerate((1.0, 0.5, 0.25, 0.25)):
            d.dist_w[k] = w
        d.max_vel, d.max_angular_vel, d.max_acc, d.max_angular_acc = 0.5, 0.5, 0.25, 0.25
        d.body_half[0], d.body_half[1] = 0.25, 0.125
    elif system == N.SYS_CAR_TRAILER:
        d.nx, d.nu, d.pos_dim, d.dt = 4, 2, 2, 0.1
        for k, w in enumerate((1.0, 0.5, 0.5)):
            d.dist_w[k] = w
        d.max_vel, d.max_angular_vel = 0.5, 0.5
        d.body_half[0], d.body_half[1] = 0.25, 0.125
        d.trailer_half[0], d.trailer_half[1] = 0.15, 0.125
        d.hitch, d.wheelbase = 0.5, 0.25
    elif system == N.SYS_QUADROTOR:
        d.nx, d.nu, d.pos_dim, d.dt = 13, 4, 3, 0.01
        for k in range(7, 10):
            d.x_lb[k], d.x_ub[k] = -4.0, 4.0
        for k in range(10, 13):
            d.x_lb[k], d.x_ub[k] = -8.0, 8.0
        for k, w in enumerate((1.0, 0.5, 0.1, 0.05)):
            d.dist_w[k] = w
        d.max_vel, d.max_acc, d.max_angular_vel = 4.0, 8.0, 8.0
        d.body_half[0] = 0.1
        for k, v in enumerate((0.034, 9.81, 0.046 * 0.7071067811865476, 0.006,
                               16.571710e-6, 16.655602e-6, 29.261652e-6)):
            d.quad[k] = v
    else:
        raise ValueError(system)
    return d


def make_params(delta=0.3, alpha=0.5, goal=None, limit_branching_factor=20, num_check_goal=4,
                check_intermediate_goal=True, delta_factor_goal=1.0, cost_delta_factor=1.0,
                use_surrogate=False, fp32_rollout=False):
    """Options_dbastar defaults (dbastar/options.hpp:105-156)."""
    p = N.ExpandParams()
    p.delta = delta
    p.alpha = alpha
    p.delta_factor_goal = delta_factor_goal
    p.cost_delta_factor = cost_delta_factor
    p.limit_branching_factor = limit_branching_factor
    p.num_check_goal = num_check_goal
    p.check_intermediate_goal = int(check_intermediate_goal)
    p.use_surrogate = int(use_surrogate)
    p.fp32_rollout = int(fp32_rollout)
    if goal is not None:
        for k, v in enumerate(goal):
            p.goal[k] = float(v)
    return p


class Workload(dict):
    """Inputs of one expansion batch: desc, offsets, states, actions, lo, hi,
    boxes, spheres, nodes, g_parent, visited, params, name."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e


def _planar_workload(name, system, seed, P, lmin, lmax, extent, n_boxes, N_nodes, n_visited,
                     visited_from_nodes=False, limit=20, prim_fn=None, desc=None):
    rng = np.random.default_rng(seed)
    desc = desc or default_desc_static(system)
    offsets, states, actions = prim_fn(rng, P, lmin, lmax)
    lo = np.array([0.0, 0.0, 0.0])
    hi = np.array([extent, extent, 0.0])
    boxes = random_boxes(rng, n_boxes, lo, hi)
    extra = None
    if system == N.SYS_CAR_TRAILER:
        def extra(s):
            s[:, 3] = wrap_angle(s[:, 2] + rng.uniform(-0.5, 0.5, size=s.shape[0]))
    elif system == N.SYS_UNICYCLE2:
        def extra(s):
            s[:, 3] = rng.uniform(-0.3, 0.3, size=s.shape[0])
            s[:, 4] = rng.uniform(-0.3, 0.3, size=s.shape[0])
    nodes = free_planar_nodes(rng, desc, boxes, lo + 0.5, hi - 0.5, N_nodes, extra)
    g_parent = rng.uniform(0.0, 10.0, size=N_nodes)
    if visited_from_nodes:
        visited = nodes.copy()
    else:
        visited = np.zeros((n_visited, desc.nx))
        visited[:, 0] = rng.uniform(lo[0], hi[0], size=n_visited)
        visited[:, 1] = rng.uniform(lo[1], hi[1], size=n_visited)
        visited[:, 2] = rng.uniform(-math.pi, math.pi, size=n_visited)
        if extra is not None:
            extra(visited)
    goal = free_planar_nodes(rng, desc, boxes, lo + 0.5, hi - 0.5, 1, extra)[0]
    params = make_params(goal=goal, limit_branching_factor=limit)
    return Workload(name=name, desc=desc, offsets=offsets, states=states, actions=actions, lo=lo, hi=hi,
                    boxes=boxes, spheres=np.zeros((0, 4)), nodes=nodes, g_parent=g_parent,
                    visited=np.ascontiguousarray(visited), params=params, seed=seed)


def unicycle1_small(seed=1, **kw):
    args = dict(P=1000, lmin=5, lmax=20, extent=10.0, n_boxes=20, N_nodes=256, n_visited=0,
                visited_from_nodes=True)
    args.update(kw)
    return _planar_workload("unicycle1_v0 plain expansion N=256 P=1000 O=20", N.SYS_UNICYCLE1, seed,
                            prim_fn=unicycle1_primitives, **args)


def unicycle1_scale(seed=2, **kw):
    args = dict(P=5000, lmin=5, lmax=30, extent=20.0, n_boxes=200, N_nodes=4096, n_visited=100_000)
    args.update(kw)
    return _planar_workload("unicycle1_v0 at scale N=4096 P=5000 O=200 V=100k", N.SYS_UNICYCLE1, seed,
                            prim_fn=unicycle1_primitives, **args)


def unicycle2_small(seed=5, **kw):
    args = dict(P=1000, lmin=5, lmax=20, extent=10.0, n_boxes=20, N_nodes=256, n_visited=0,
                visited_from_nodes=True)
    args.update(kw)
    return _planar_workload("unicycle2_v0 N=256 P=1000 O=20", N.SYS_UNICYCLE2, seed,
                            prim_fn=unicycle2_primitives, **args)


def car_trailer(seed=4, **kw):
    args = dict(P=5000, lmin=10, lmax=40, extent=30.0, n_boxes=300, N_nodes=4096, n_visited=100_000)
    args.update(kw)
    return _planar_workload("car+1 trailer N=4096 P=5000 O=300", N.SYS_CAR_TRAILER, seed,
                            prim_fn=car_trailer_primitives, **args)


# ---- quadrotor ------------------------------------------------------------------
def _quat_mul(a, b):
    x1, y1, z1, w1 = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    x2, y2, z2, w2 = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
                     w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], axis=-1)


def random_attitudes(rng, n, tilt_max=0.25):
    """Uniform yaw times a tilt of at most tilt_max about a random horizontal axis."""
    yaw = rng.uniform(-math.pi, math.pi, n)
    ax = rng.normal(size=(n, 2))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    tilt = tilt_max * np.sqrt(rng.uniform(0, 1, n))
    qy = np.stack([np.zeros(n), np.zeros(n), np.sin(yaw / 2), np.cos(yaw / 2)], 1)
    qt = np.stack([ax[:, 0] * np.sin(tilt / 2), ax[:, 1] * np.sin(tilt / 2), np.zeros(n), np.cos(tilt / 2)], 1)
    return _quat_mul(qy, qt)


def quad_step_np(desc, s, u):
    """Vectorised FP64 step of the quadrotor model (same equations as the
    library's rollout; used to generate primitive data only)."""
    m, g, arm, t2t, Jx, Jy, Jz = (desc.quad[k] for k in range(7))
    dt = desc.dt
    q = s[:, 3:7]
    v = s[:, 7:10]
    w = s[:, 10:13]
    f = (g / 4) * u.sum(1)
    F = m * g / 4
    tau = np.stack([arm * F * (u[:, 2] + u[:, 3] - u[:, 0] - u[:, 1]),
                    arm * F * (u[:, 1] + u[:, 2] - u[:, 0] - u[:, 3]),
                    t2t * F * (u[:, 1] + u[:, 3] - u[:, 0] - u[:, 2])], 1)
    qx, qy, qz, qw = q.T
    rz = np.stack([2 * (qx * qz + qw * qy), 2 * (qy * qz - qw * qx), 1 - 2 * (qx * qx + qy * qy)], 1)
    acc = f[:, None] * rz - np.array([0, 0, g])
    J = np.array([Jx, Jy, Jz])
    dw = (tau - np.cross(w, J * w)) / J
    out = np.empty_like(s)
    out[:, 0:3] = s[:, 0:3] + dt * v
    out[:, 7:10] = v + dt * acc
    out[:, 10:13] = w + dt * dw
    th = np.linalg.norm(w, axis=1) * dt / 2
    sinc = np.where(th > 1e-12, np.sin(th) / np.where(th > 0, th, 1), 1.0)
    e = np.concatenate([(dt / 2) * sinc[:, None] * w, np.cos(th)[:, None]], 1)
    nq = _quat_mul(q, e)
    out[:, 3:7] = nq / np.linalg.norm(nq, axis=1, keepdims=True)
    return out


def quadrotor_primitives(rng, P, lmin, lmax, desc=None, tilt_max=0.25, w_sd=0.7, thrust="uniform", base=(0.9, 1.1),
                         diff=0.02):
    """P primitives of L ~ U{lmin..lmax} steps: 1 to 3 pieces of constant
    rotor thrusts (hover = 1), rolled out in FP64 from a random attitude and
    angular rate at the origin.  thrust="uniform": each rotor U[0, 1.4] per
    piece (the thrust range of quadrotor_v0); most such primitives leave the
    angular-rate bounds within a few steps.  thrust="balanced": U[base] plus
    U[-diff, diff] per rotor, primitives that stay near hover."""
    desc = desc or default_desc_static(N.SYS_QUADROTOR)
    L = rng.integers(lmin, lmax + 1, size=P)
    Lmax = int(L.max())
    u = np.empty((P, Lmax, 4))
    pieces = rng.integers(1, 4, size=P)
    for j in range(P):
        cuts = np.sort(rng.integers(1, Lmax, size=pieces[j] - 1))
        bnd = np.concatenate([[0], cuts, [Lmax]])
        for a, b in zip(bnd[:-1], bnd[1:]):
            if thrust == "uniform":
                u[j, a:b] = rng.uniform(0.0, 1.4, 4)
            else:
                u[j, a:b] = np.clip(rng.uniform(*base) + rng.uniform(-diff, diff, 4), 0.0, 1.4)
    x = np.zeros((P, Lmax + 1, 13))
    x[:, 0, 3:7] = random_attitudes(rng, P, tilt_max)
    x[:, 0, 10:13] = rng.normal(0, w_sd, (P, 3))
    for k in range(Lmax):
        x[:, k + 1] = quad_step_np(desc, x[:, k], u[:, k])
    return _pack([x[j, : L[j] + 1] for j in range(P)], [u[j, : L[j]] for j in range(P)])


def sphere_collides_np(desc, boxes, spheres, p):
    r = desc.body_half[0]
    c = np.clip(p[:, None, :], boxes[None, :, :3] - boxes[None, :, 3:], boxes[None, :, :3] + boxes[None, :, 3:])
    hit = (((c - p[:, None, :]) ** 2).sum(2) <= r * r).any(1)
    if spheres.shape[0]:
        d2 = ((spheres[None, :, :3] - p[:, None, :]) ** 2).sum(2)
        hit |= (d2 <= (r + spheres[None, :, 3]) ** 2).any(1)
    return hit


def quadrotor(seed=3, P=10000, lmin=10, lmax=50, n_boxes=100, n_spheres=50, N_nodes=2048, n_visited=20000,
              limit=20, tilt_max=0.25, w_sd=0.7, thrust="uniform"):
    """quadrotor_v0 batch: P primitives of 10-50 steps (dt 0.01), a 10x10x5 m
    world with boxes (sides U[0.5, 1.5]) and spheres (r U[0.2, 0.8]), a 0.1 m
    sphere robot, N frontier nodes and V visited states in free space."""
    rng = np.random.default_rng(seed)
    desc = default_desc_static(N.SYS_QUADROTOR)
    offsets, states, actions = quadrotor_primitives(rng, P, lmin, lmax, desc, tilt_max, w_sd, thrust)
    lo = np.array([0.0, 0.0, 0.0])
    hi = np.array([10.0, 10.0, 5.0])
    boxes = random_boxes(rng, n_boxes, lo, hi, 0.5, 1.5, dim=3)
    spheres = np.c_[rng.uniform(lo, hi, (n_spheres, 3)), rng.uniform(0.2, 0.8, n_spheres)]

    def states_in_free_space(n):
        out = []
        while sum(len(o) for o in out) < n:
            m = 4 * n
            s = np.zeros((m, 13))
            s[:, :3] = rng.uniform(lo + 0.5, hi - 0.5, (m, 3))
            s[:, 3:7] = random_attitudes(rng, m, tilt_max)
            s[:, 7:10] = rng.uniform(-0.5, 0.5, (m, 3))
            s[:, 10:13] = rng.normal(0, w_sd, (m, 3))
            out.append(s[~sphere_collides_np(desc, boxes, spheres, s[:, :3])])
        return np.ascontiguousarray(np.concatenate(out)[:n])

    nodes = states_in_free_space(N_nodes)
    visited = states_in_free_space(n_visited)
    goal = states_in_free_space(1)[0]
    params = make_params(goal=goal, limit_branching_factor=limit, fp32_rollout=True)
    tag = "" if thrust == "uniform" else f" thrust={thrust}"
    return Workload(name=f"quadrotor_v0 N={N_nodes} P={P} O={n_boxes}+{n_spheres} V={n_visited}{tag}", desc=desc,
                    offsets=offsets, states=states, actions=actions, lo=lo, hi=hi, boxes=boxes, spheres=spheres,
                    nodes=nodes, g_parent=rng.uniform(0, 10, N_nodes), visited=visited, params=params, seed=seed)


# ---- many instances --------------------------------------------------------------
def _prefix(offsets, states, actions, P):
    """The first P primitives of a library (the reference's max_motions)."""
    off = np.ascontiguousarray(offsets[: P + 1])
    return off, np.ascontiguousarray(states[: off[-1]]), np.ascontiguousarray(actions[: off[-1] - P])


def multi_instance(seed=6, n_instances=128, p_range=(1000, 10000), o_range=(10, 300), n_range=(64, 4096),
                   n_visited=10_000, systems=(N.SYS_UNICYCLE1, N.SYS_UNICYCLE2, N.SYS_QUADROTOR)):
    """Independent planning problems for one batched launch: system drawn
    uniformly from `systems`, obstacles O ~ U{o_range}, primitives P ~
    U{p_range} (the first P of one 10k library per system), frontier N ~
    U{n_range}.  Planar maps are 20x20 m with O boxes; quadrotor worlds are
    10x10x5 m with 2/3 of O boxes and 1/3 spheres."""
    rng = np.random.default_rng(seed)
    pmax = p_range[1]
    libs = {}
    for s in systems:
        lr = np.random.default_rng(seed * 1000 + s)
        if s == N.SYS_UNICYCLE1:
            libs[s] = unicycle1_primitives(lr, pmax, 5, 30)
        elif s == N.SYS_UNICYCLE2:
            libs[s] = unicycle2_primitives(lr, pmax, 5, 20)
        elif s == N.SYS_QUADROTOR:
            libs[s] = quadrotor_primitives(lr, pmax, 10, 50, thrust="balanced")
        else:
            libs[s] = car_trailer_primitives(lr, pmax, 10, 40)
    out = []
    for i in range(n_instances):
        s = systems[int(rng.integers(len(systems)))]
        P = int(rng.integers(p_range[0], p_range[1] + 1))
        O = int(rng.integers(o_range[0], o_range[1] + 1))
        Nn = int(rng.integers(n_range[0], n_range[1] + 1))
        sub_seed = int(rng.integers(1 << 31))
        off, st, ac = _prefix(*libs[s], P)
        if s == N.SYS_QUADROTOR:
            nb = (2 * O + 2) // 3
            w = quadrotor(seed=sub_seed, P=1, n_boxes=nb, n_spheres=O - nb, N_nodes=Nn, n_visited=n_visited)
        else:
            prim_fn = {N.SYS_UNICYCLE1: unicycle1_primitives, N.SYS_UNICYCLE2: unicycle2_primitives,
                       N.SYS_CAR_TRAILER: car_trailer_primitives}[s]
            w = _planar_workload("", s, sub_seed, 1, 5, 6, 20.0, O, Nn, n_visited, prim_fn=prim_fn)
        w["offsets"], w["states"], w["actions"] = off, st, ac
        w["name"] = f"#{i} system {s} N={Nn} P={P} O={O} V={n_visited}"
        out.append(w)
    return out
