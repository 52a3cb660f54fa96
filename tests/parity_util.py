"""Comparison helpers shared by the GPU parity tests."""
import numpy as np

from dynoplan_amd import _native as N


def bits(a):
    a = np.ascontiguousarray(a, np.float64)
    return a.view(np.int64)


def assert_parity(gpu, ora, label=""):
    """Bit-exact agreement of every per-pair field (FP64 fields compared as
    bit patterns on FREE pairs, where they are defined)."""
    n_g, n_o = gpu["node"].size, ora["node"].size
    assert n_g == n_o, f"{label}: pair count {n_g} != oracle {n_o}"
    for k in ("node", "prim", "verdict", "first_k", "goal_index", "resolution"):
        a, b = gpu[k], ora[k]
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, (f"{label}: field {k} differs at {bad.size} pairs, first at {bad[0]}: "
                               f"gpu={a[bad[0]]} oracle={b[bad[0]]} (node {ora['node'][bad[0]]}, "
                               f"prim {ora['prim'][bad[0]]}, verdict {ora['verdict'][bad[0]]})")
    # CAPPED pairs come after their node's loop ended: no final state, g, h
    capped = ora["resolution"] == N.R_CAPPED
    for k in ("final_state", "g", "h"):
        assert np.isnan(gpu[k][capped]).all(), f"{label}: {k} of a CAPPED pair is not NaN"
    free = (ora["verdict"] == N.V_FREE) & ~capped
    if not free.any():
        return
    for k in ("final_state", "g", "h"):
        a, b = bits(gpu[k][free]), bits(ora[k][free])
        bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
        assert bad.size == 0, f"{label}: {k} differs (bitwise) at {bad.size} FREE pairs, first {bad[0]}"


def prefix(out, n_pairs):
    return {k: v[:n_pairs] for k, v in out.items()}


def counts_consistent(c):
    assert c["pairs"] == c["bounds"] + c["collide"] + c["free_"] + c["surrogate_reject"] + c.get("constraint", 0)
    assert c["free_"] == c["novel"] + c["dup_visited"] + c["dup_batch"] + c["capped"]
    assert c["checked"] == c["pairs"] - c["bounds"] - c["surrogate_reject"]
