"""Write the golden fixtures of tests/ from the reference's own data files.

Run in a container where the reference tree is mounted (default
/root/reference); the outputs are small data files committed next to this
script, so the tests never read the reference at run time.

  unicycle1_ref_prims.npz   data/motion_primitives/unicycle1_v0/*.less.bin.msgpack
                            (100 primitives: offsets, states, actions)
  unicycle1_ref_prims.bin   the Boost binary twin of the same 100 primitives
  unicycle1_ref10.npz       data/motions__unicycle1_v0__20-07-2023--16-04-22.yaml
  unicycle1_ref10.bin       its Boost binary twin
  unicycle1_ref10.yaml      the YAML file itself (format pin of the YAML reader)
  quadrotor_initguess.npz   benchmark_initguess/quadrotor_v0/*/*.yaml solution
                            trajectories (states, actions, primitive junctions)
"""
import glob
import os
import shutil
import sys

import msgpack
import numpy as np
import yaml

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def pack(trajs):
    lens = [len(t["states"]) for t in trajs]
    offsets = np.zeros(len(lens) + 1, np.int32)
    np.cumsum(lens, out=offsets[1:])
    states = np.array([s for t in trajs for s in t["states"]], np.float64)
    actions = np.array([a for t in trajs for a in t["actions"]], np.float64)
    return offsets, states, actions


def main():
    d = os.path.join(REF, "data/motion_primitives/unicycle1_v0")
    mp = glob.glob(os.path.join(d, "*.msgpack"))[0]
    with open(mp, "rb") as f:
        data = msgpack.unpackb(f.read(), raw=False)["data"]
    offsets, states, actions = pack(data)
    np.savez_compressed(os.path.join(OUT, "unicycle1_ref_prims.npz"), offsets=offsets, states=states, actions=actions)
    shutil.copyfile(mp[: -len(".msgpack")], os.path.join(OUT, "unicycle1_ref_prims.bin"))

    y = os.path.join(REF, "data/motions__unicycle1_v0__20-07-2023--16-04-22.yaml")
    with open(y) as f:
        trajs = yaml.safe_load(f)
    offsets, states, actions = pack(trajs)
    np.savez_compressed(os.path.join(OUT, "unicycle1_ref10.npz"), offsets=offsets, states=states, actions=actions)
    shutil.copyfile(y[: -len(".yaml")] + ".bin", os.path.join(OUT, "unicycle1_ref10.bin"))
    shutil.copyfile(y, os.path.join(OUT, "unicycle1_ref10.yaml"))

    names, st, ac, splits, offs = [], [], [], [], [0]
    for p in sorted(glob.glob(os.path.join(REF, "benchmark_initguess/quadrotor_v0/*/*.yaml"))):
        with open(p) as f:
            doc = yaml.safe_load(f)
        for r in doc.get("result") or []:
            s = np.array(r["states"], np.float64)
            a = np.array(r["actions"], np.float64)
            names.append(os.path.relpath(p, REF))
            st.append(s)
            ac.append(np.vstack([a, np.full((1, 4), np.nan)]))  # pad to one row per state
            sp = np.zeros(len(s), np.int8)
            for k in r.get("splits") or []:
                if 0 <= int(k) < len(s):
                    sp[int(k)] = 1
            splits.append(sp)
            offs.append(offs[-1] + len(s))
    np.savez_compressed(os.path.join(OUT, "quadrotor_initguess.npz"), names=np.array(names), offsets=np.array(offs, np.int32),
                        states=np.concatenate(st), actions=np.concatenate(ac), splits=np.concatenate(splits))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
