"""The learned collision-feasibility surrogate on the GPU.

- the training GEMM (BF16 MFMA, FP32 accumulation) against numpy on the same
  BF16-rounded inputs;
- forward + backward of the MLP against an FP64 numpy restatement that rounds
  to BF16 at the same points (the CPU gradient oracle), with a planted-mistake
  control;
- train on exact-checker labels, calibrate the conservative threshold on
  held-out nodes, then on unseen nodes: 100% feasible recall, an identical
  edge set (every resolution unchanged), and a reported share of exact checks
  skipped -- what the trained network gains over the method without it.
"""
import ctypes as C
import math

import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.expander import expander_for

pytestmark = pytest.mark.gpu


def bf16(x):
    """Round-to-nearest-even to BF16, returned as float32 values."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@pytest.fixture(scope="module")
def ctx():
    wl = synth.car_trailer(P=400, N_nodes=64, n_boxes=60, extent=10.0, n_visited=500)
    ex = expander_for(wl)
    yield ex, wl
    ex.close()


def test_gemm_bf16_numerics(need_gpu, ctx):
    ex, _ = ctx
    rng = np.random.default_rng(0)
    for (M, Nn, K) in [(64, 64, 64), (100, 70, 37), (3, 5, 9), (256, 1088, 256), (4096, 256, 256)]:
        A = rng.standard_normal((M, K)).astype(np.float32)
        B = rng.standard_normal((K, Nn)).astype(np.float32)
        Cg = np.zeros((M, Nn), np.float32)
        N.check(ex.lib.dp_debug_gemm_bf16(ex.ctx, fptr(A), fptr(B), M, Nn, K, fptr(Cg)))
        ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
        err = np.abs(Cg - ref).max() / max(1.0, np.abs(ref).max())
        assert err < 1e-5, (M, Nn, K, err)


def unpack(theta):
    o = 0
    out = {}
    for name, shape in (("W1", (256, 1088)), ("b1", (256,)), ("W2", (256, 256)), ("b2", (256,)),
                        ("W3", (256, 256)), ("b3", (256,)), ("w4", (256,)), ("b4", (1,))):
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def mlp_ref(theta, X, y, relu2_mask=True):
    """FP64 forward/backward rounding activations and weights to BF16 where
    the GPU path does (inputs of every GEMM)."""
    P = {k: v.astype(np.float64) for k, v in unpack(theta).items()}
    W1b, W2b, W3b = (bf16(P[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    Xb = bf16(X).astype(np.float64)
    h1 = np.maximum(Xb @ W1b.T + P["b1"], 0)
    h1b = bf16(h1).astype(np.float64)
    h2 = np.maximum(h1b @ W2b.T + P["b2"], 0)
    h2b = bf16(h2).astype(np.float64)
    h3 = np.maximum(h2b @ W3b.T + P["b3"], 0)
    z = h3 @ P["w4"] + P["b4"][0]
    B = X.shape[0]
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (1 / (1 + np.exp(-z)) - y) / B
    g = {"w4": h3.T @ dz, "b4": np.array([dz.sum()])}
    dz3 = dz[:, None] * P["w4"][None, :] * (h3 > 0)
    g["b3"] = dz3.sum(0)
    dz3b = bf16(dz3).astype(np.float64)
    g["W3"] = dz3b.T @ h2b
    dz2 = (dz3b @ W3b) * ((h2 > 0) if relu2_mask else 1.0)
    g["b2"] = dz2.sum(0)
    dz2b = bf16(dz2).astype(np.float64)
    g["W2"] = dz2b.T @ h1b
    dz1 = (dz2b @ W2b) * (h1 > 0)
    g["b1"] = dz1.sum(0)
    dz1b = bf16(dz1).astype(np.float64)
    g["W1"] = dz1b.T @ Xb
    flat = np.concatenate([g[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4")])
    return loss, flat


def test_gradients_against_fp64_oracle(need_gpu, ctx):
    ex, _ = ctx
    ex.surrogate_init(seed=3)
    n = int(ex.lib.dp_surrogate_param_count())
    rng = np.random.default_rng(1)
    theta = (rng.standard_normal(n) * 0.05).astype(np.float32)
    B = 96
    X = np.zeros((B, 1088), np.float32)
    X[:, :1024] = rng.random((B, 1024)) < 0.3          # occupancy bits
    X[:, 1024:] = rng.standard_normal((B, 64)) * 0.5    # embedding inputs
    y = (rng.random(B) < 0.5).astype(np.float32)
    loss = np.zeros(1, np.float32)
    grad = np.zeros(n, np.float32)
    N.check(ex.lib.dp_surrogate_grad(ex.ctx, fptr(theta), fptr(X), fptr(y), B, fptr(loss), fptr(grad)))
    ref_loss, ref = mlp_ref(theta, X, y)
    assert abs(loss[0] - ref_loss) < 1e-5 * max(1, abs(ref_loss))
    gb, rb = unpack(grad), unpack(ref.astype(np.float32))
    for k in gb:
        rel = np.linalg.norm(gb[k] - rb[k]) / max(np.linalg.norm(rb[k]), 1e-12)
        assert rel < 5e-3, (k, rel)
    # planted mistake in the oracle (dropped ReLU mask of layer 2) is detected
    _, bad = mlp_ref(theta, X, y, relu2_mask=False)
    rel_bad = np.linalg.norm(unpack(bad)["W2"] - gb["W2"]) / np.linalg.norm(gb["W2"])
    assert rel_bad > 50 * 5e-3


def patches_np(nodes, boxes):
    """32x32 occupancy of 0.1 m cells centred on the node (cell centres at
    +-15.5 cells), planar AABB obstacles."""
    c = (np.arange(32) - 15.5) * 0.1
    px = nodes[:, 0, None, None] + c[None, None, :]
    py = nodes[:, 1, None, None] + c[None, :, None]
    occ = np.zeros((nodes.shape[0], 32, 32), bool)
    for b in boxes:
        occ |= (np.abs(px - b[0]) <= b[3]) & (np.abs(py - b[1]) <= b[4])
    return occ.reshape(-1, 1024).astype(np.float32)


def infer_np(theta, emb, patch, prim):
    """The inference path in numpy: A = W1p patch, B = W1e emb + b1 (BF16
    inputs, wide accumulation), h1 = relu(A + B), two BF16 layers, FP32 head
    on BF16 activations."""
    P = unpack(theta)
    W1b = bf16(P["W1"]).astype(np.float64)
    A = patch.astype(np.float64) @ W1b[:, :1024].T
    B = bf16(emb[prim]).astype(np.float64) @ W1b[:, 1024:].T + P["b1"]
    h = bf16(np.maximum(A + B, 0)).astype(np.float64)
    h = bf16(np.maximum(h @ bf16(P["W2"]).astype(np.float64).T + P["b2"], 0)).astype(np.float64)
    h = bf16(np.maximum(h @ bf16(P["W3"]).astype(np.float64).T + P["b3"], 0)).astype(np.float64)
    return h @ P["w4"].astype(np.float64) + float(P["b4"][0])


def test_train_calibrate_recall_and_gain(need_gpu):
    wl = synth.car_trailer(P=1500, N_nodes=512, n_boxes=150, extent=14.0, n_visited=5000)
    rng = np.random.default_rng(99)
    extra = lambda s: s.__setitem__((slice(None), 3), synth.wrap_angle(s[:, 2] + rng.uniform(-0.5, 0.5, s.shape[0])))
    train_nodes = synth.free_planar_nodes(rng, wl.desc, wl.boxes, wl.lo + 0.5, wl.hi - 0.5, 16384, extra)
    cal_nodes = synth.free_planar_nodes(rng, wl.desc, wl.boxes, wl.lo + 0.5, wl.hi - 0.5, 8192, extra)
    p0 = synth.make_params(goal=wl.params.goal[:4], limit_branching_factor=20)
    n = int(N.load().dp_surrogate_param_count())
    with expander_for(wl) as ex:
        ex.surrogate_init(seed=5)
        losses = ex.surrogate_train(train_nodes, p0, epochs=8, batch=4096, lr=2e-3)
        stats = ex.surrogate_calibrate(cal_nodes, p0, margin=0.5, tail=4.0)
        cal = ex.fetch()
        scores = ex.surrogate_scores(cal["prim"].size)
        theta = np.zeros(n, np.float32)
        emb = np.zeros((wl.offsets.size - 1, 64), np.float32)
        N.check(ex.lib.dp_surrogate_get(ex.ctx, fptr(theta), fptr(emb)))
        c_plain = ex.expand(wl.nodes, wl.g_parent, p0)
        plain = ex.fetch()
        p1 = synth.make_params(goal=wl.params.goal[:4], limit_branching_factor=20, use_surrogate=True)
        c_sur = ex.expand(wl.nodes, wl.g_parent, p1)
        sur = ex.fetch()
    print("epoch losses", losses, "calibration", stats)
    # the inference kernel computes the network the numpy restatement computes
    pick = np.random.default_rng(0).choice(cal["prim"].size, 3000, replace=False)
    z = infer_np(theta, emb, patches_np(cal_nodes, wl.boxes)[cal["node"][pick]], cal["prim"][pick])
    err = np.abs(scores[pick] - z) / (1 + np.abs(z))
    assert np.percentile(err, 99) < 2e-2 and err.max() < 1e-1, (np.percentile(err, 99), err.max())
    # trained: held-out BCE of the scores well below the constant predictor's
    y = (cal["verdict"] == N.V_FREE).astype(np.float64)
    zc = scores.astype(np.float64)
    bce = np.mean(np.maximum(zc, 0) - zc * y + np.log1p(np.exp(-np.abs(zc))))
    p = y.mean()
    prior = -(p * math.log(p) + (1 - p) * math.log(1 - p))
    print(f"held-out BCE {bce:.4f} vs label prior {prior:.4f}")
    assert bce < 0.8 * prior and losses[-1] < 0.8 * prior
    # calibration: below every feasible score of the calibration batch by at
    # least the margin, widened by the lower tail's spread
    fz = np.sort(zc[y > 0])[:10]
    assert stats["threshold"] == pytest.approx(fz[0] - max(0.5, 4.0 * (fz[-1] - fz[0])), abs=1e-4)
    # unseen batch: recall 1, same edge set, every kept pair decided exactly
    free = plain["verdict"] == N.V_FREE
    rej = sur["verdict"] == N.V_SURROGATE_REJECT
    assert np.all(sur["verdict"][free] == N.V_FREE), "a feasible edge was rejected (recall < 1)"
    assert np.all(~free[rej])
    assert np.array_equal(sur["resolution"], plain["resolution"])  # identical edge set
    keep = ~rej
    assert np.array_equal(sur["verdict"][keep], plain["verdict"][keep])
    assert np.array_equal(sur["first_k"][keep], plain["first_k"][keep])
    for k in ("final_state", "g", "h"):
        assert np.array_equal(sur[k][free], plain[k][free], equal_nan=True)  # NaN: CAPPED
    print(f"exact checks skipped by the surrogate: {rej.sum()} of {rej.size} pairs ({100 * rej.mean():.2f}%), "
          f"{100 * rej.sum() / max(1, (~free).sum()):.2f}% of infeasible pairs; recall 1.0")
    assert c_sur["surrogate_reject"] == rej.sum()
    assert c_sur["checked"] == c_plain["checked"] - rej.sum() + np.count_nonzero(plain["verdict"][rej] == N.V_BOUNDS)
