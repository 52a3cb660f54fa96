"""Primitive file formats (the on-disk boundary; no GPU).

Pins: the reference's own twins -- data/...less.bin (Boost) against its
msgpack twin (bit-exact) and data/motions_...bin against its YAML twin (to
the YAML's 6 printed digits) -- plus writer/reader round trips and the
load_motion_primitives_new semantics (robots.cpp:3061-3174)."""
import os

import msgpack
import numpy as np
import pytest

from dynoplan_amd import _native as N
from dynoplan_amd import synth
from dynoplan_amd.primitives import load_primitives, save_primitives

G = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ref100():
    return np.load(os.path.join(G, "unicycle1_ref_prims.npz"))


def test_boost_archive_matches_msgpack_twin(ref100):
    off, st, ac, cost = load_primitives(os.path.join(G, "unicycle1_ref_prims.bin"), N.SYS_UNICYCLE1)
    assert np.array_equal(off, ref100["offsets"])
    assert np.array_equal(st.view(np.int64), ref100["states"].view(np.int64))
    assert np.array_equal(ac.view(np.int64), ref100["actions"].view(np.int64))
    # archive cost 1e8 > 1e5: ref_dt * #actions (robots.cpp:3153-3159)
    assert np.allclose(cost, 0.1 * (np.diff(off) - 1))


def test_msgpack_reader(tmp_path, ref100):
    off, st, ac = ref100["offsets"], ref100["states"], ref100["actions"]
    data = {"data": [{"states": st[off[j]:off[j + 1]].tolist(), "actions": ac[off[j] - j:off[j + 1] - j - 1].tolist()}
                     for j in range(off.size - 1)]}
    p = tmp_path / "prims.msgpack"
    p.write_bytes(msgpack.packb(data, use_single_float=False))
    o2, s2, a2, _ = load_primitives(p, N.SYS_UNICYCLE1)
    assert np.array_equal(o2, off) and np.array_equal(s2, st) and np.array_equal(a2, ac)


def test_json_reader(tmp_path, ref100):
    import json
    off, st, ac = ref100["offsets"], ref100["states"], ref100["actions"]
    data = {"data": [{"states": st[off[j]:off[j + 1]].tolist(), "actions": ac[off[j] - j:off[j + 1] - j - 1].tolist()}
                     for j in range(5)]}
    p = tmp_path / "prims.json"
    p.write_text(json.dumps(data))
    o2, s2, a2, _ = load_primitives(p, N.SYS_UNICYCLE1)
    assert np.array_equal(s2, st[:off[5]]) and np.array_equal(a2, ac[:off[5] - 5])


def test_yaml_and_boost_twins_of_the_reference():
    oy, sy, ay, cy = load_primitives(os.path.join(G, "unicycle1_ref10.yaml"), N.SYS_UNICYCLE1)
    ob, sb, ab, cb = load_primitives(os.path.join(G, "unicycle1_ref10.bin"), N.SYS_UNICYCLE1)
    ref = np.load(os.path.join(G, "unicycle1_ref10.npz"))
    assert np.array_equal(oy, ob) and np.array_equal(oy, ref["offsets"])
    assert np.array_equal(sy, ref["states"])  # the YAML reader parses the printed digits exactly
    assert np.abs(sy - sb).max() < 1e-5 and np.abs(ay - ab).max() < 1e-5
    assert np.array_equal(cy, cb) and set(cy.tolist()) == {2.5, 5.0, 10.0}  # costs kept (< 1e5)


@pytest.mark.parametrize("ext", ["bin", "msgpack", "yaml"])
def test_writer_round_trip(tmp_path, ext):
    rng = np.random.default_rng(0)
    off, st, ac = synth.quadrotor_primitives(rng, 30, 3, 9)
    cost = rng.uniform(0, 1, 30)
    p = tmp_path / f"q.{ext}"
    save_primitives(p, off, st, ac, cost)
    o2, s2, a2, c2 = load_primitives(p, N.SYS_QUADROTOR)
    assert np.array_equal(o2, off)
    q = st[:, 3:7] / np.linalg.norm(st[:, 3:7], axis=1, keepdims=True)
    assert np.array_equal(np.delete(s2, np.s_[3:7], 1), np.delete(st, np.s_[3:7], 1))
    assert np.abs(s2[:, 3:7] - q).max() < 1e-15  # quaternions normalised on load
    assert np.array_equal(a2, ac) and np.array_equal(c2, cost)


def test_load_semantics(tmp_path, ref100):
    path = os.path.join(G, "unicycle1_ref_prims.bin")
    off, st, ac, _ = load_primitives(path, N.SYS_UNICYCLE1, max_motions=7)
    assert off.size == 8 and np.array_equal(st, ref100["states"][:off[-1]])
    o1, s1, _, _ = load_primitives(path, N.SYS_UNICYCLE1, noise=1e-7, seed=5)
    o2, s2, _, _ = load_primitives(path, N.SYS_UNICYCLE1, noise=1e-7, seed=5)
    o3, s3, _, _ = load_primitives(path, N.SYS_UNICYCLE1, noise=1e-7, seed=6)
    base = ref100["states"]
    ends = np.zeros(base.shape[0], bool)
    ends[o1[:-1]] = True
    ends[o1[1:] - 1] = True
    d = s1 - base
    assert np.array_equal(s1, s2) and not np.array_equal(s1, s3)
    assert np.all(d[~ends] == 0) and np.abs(d[ends]).max() <= 1e-7 and np.abs(d[ends]).max() > 1e-8


def test_loader_errors(tmp_path):
    lib = N.load()
    with pytest.raises(N.DynoplanError) as e:
        load_primitives(tmp_path / "missing.bin", N.SYS_UNICYCLE1)
    assert e.value.code == N.DP_ERR_IO
    with pytest.raises(N.DynoplanError) as e:
        load_primitives(tmp_path / "prims.txt", N.SYS_UNICYCLE1)
    assert e.value.code == N.DP_ERR_ARG
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"\x16" + b"\x00" * 7 + b"not an archive at all!" + b"\x00" * 40)
    with pytest.raises(N.DynoplanError) as e:
        load_primitives(bad, N.SYS_UNICYCLE1)
    assert e.value.code == N.DP_ERR_FORMAT
    # unicycle1 file read as a quadrotor: state size mismatch
    with pytest.raises(N.DynoplanError) as e:
        load_primitives(os.path.join(G, "unicycle1_ref_prims.bin"), N.SYS_QUADROTOR)
    assert e.value.code == N.DP_ERR_FORMAT
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(open(os.path.join(G, "unicycle1_ref_prims.bin"), "rb").read()[:5000])
    with pytest.raises(N.DynoplanError) as e:
        load_primitives(trunc, N.SYS_UNICYCLE1)
    assert e.value.code == N.DP_ERR_FORMAT


# ---- problem files ---------------------------------------------------------------------
PROBLEM = """# a unicycle problem in dynobench's layout (benchmark.py:330-350)
environment:
  min: [0, 0]
  max: [6, 5.5]
  obstacles:
    - type: box
      center: [3.0, 4.2]
      size: [2.4, 0.2]
    - type: box
      center: [1.5, 1.0]
      size: [0.5, 0.25]
robots:
  - type: unicycle1_v0
    start: [3.8, 3, 0]
    goal: [5.2, 3, -1.5707963267948966]
"""


def test_problem_reader(tmp_path):
    from dynoplan_amd.primitives import load_problem
    f = tmp_path / "p.yaml"
    f.write_text(PROBLEM)
    p = load_problem(f)
    assert p["system"] == N.SYS_UNICYCLE1 and p["robot_type"] == "unicycle1_v0"
    assert np.array_equal(p["start"], [3.8, 3, 0]) and p["goal"][2] == -1.5707963267948966
    assert np.array_equal(p["lo"], [0, 0, 0]) and np.array_equal(p["hi"], [6, 5.5, 0])
    # full side lengths in the file, half sizes for dp_world
    assert np.array_equal(p["boxes"], [[3.0, 4.2, 0, 1.2, 0.1, 0], [1.5, 1.0, 0, 0.25, 0.125, 0]])
    assert p["spheres"].shape == (0, 4)


@pytest.mark.parametrize("system,model,nx", [(N.SYS_UNICYCLE1, "unicycle1_v0", 3), (N.SYS_UNICYCLE2, "unicycle2_v0", 5),
                                             (N.SYS_CAR_TRAILER, "car_first_order_with_1_trailers_0", 4),
                                             (N.SYS_QUADROTOR, "quadrotor_v0", 13)])
def test_problem_round_trip(tmp_path, system, model, nx):
    from dynoplan_amd.primitives import load_problem, save_problem
    rng = np.random.default_rng(system)
    dims = 3 if system == N.SYS_QUADROTOR else 2
    boxes = np.zeros((7, 6))
    boxes[:, :dims] = rng.uniform(0, 5, (7, dims))
    boxes[:, 3:3 + dims] = rng.uniform(0.1, 1, (7, dims))
    spheres = np.c_[rng.uniform(0, 5, (3, 3)), rng.uniform(0.2, 0.8, 3)] if dims == 3 else None
    start, goal = rng.normal(size=nx), rng.normal(size=nx)
    f = tmp_path / "q.yaml"
    save_problem(f, system, model, start, goal, [0.0] * dims, [5.0] * dims, boxes, spheres)
    p = load_problem(f)
    assert p["system"] == system and p["robot_type"] == model
    assert np.array_equal(p["start"], start) and np.array_equal(p["goal"], goal)
    assert np.array_equal(p["boxes"], boxes)
    if spheres is not None:
        assert np.array_equal(p["spheres"], spheres)


def test_problem_errors(tmp_path):
    from dynoplan_amd.primitives import load_problem
    for text in ("robots: []\n", PROBLEM.replace("unicycle1_v0", "acrobot_v0"),
                 PROBLEM.replace("start: [3.8, 3, 0]", "start: [3.8, 3]"),
                 PROBLEM.replace("type: box", "type: cylinder", 1)):
        f = tmp_path / "bad.yaml"
        f.write_text(text)
        with pytest.raises(N.DynoplanError) as e:
            load_problem(f)
        assert e.value.code == N.DP_ERR_FORMAT
    with pytest.raises(N.DynoplanError) as e:
        load_problem(tmp_path / "missing.yaml")
    assert e.value.code == N.DP_ERR_IO
