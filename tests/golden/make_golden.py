"""Regenerate the committed parity fixtures under tests/golden/.

Inputs are the reference's own example particle files (data, read as text):
  configurations/soft_on_plane/plane.dat   3183 disks, box periodic 100x100x10
  configurations/soft_on_sphere/sphere.dat 316 disks on a sphere r = 10, box fixed 30^3
  configurations/lj_langevin/particles.dat 300 LJ particles, box periodic 100x100x10
  configurations/relax/out.input           68 disks (types 1, 2), box fixed 100x100x10
Each input is stored as inputs_<name>.npz.  Expected outputs come from the CPU
oracle (oracle/samos_oracle.cpp) with the conf parameters of the same example:
half pair list, forces, torques and energies after one loop-head evaluation,
stored as golden_<name>.npz.  MANIFEST.json records parameters and SHA-256.

Usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OLD_KEYS = ["id", "type", "radius", "x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "omega", "length",
            "ix", "iy", "iz"]

CASES = {
    "plane": dict(path="configurations/soft_on_plane/plane.dat", box=(100.0, 100.0, 10.0), periodic=True,
                  nlist=(2.5, 0.5), constraint=("plane", 0.0, 0), soft=(1.0, 2.0, False), polar=1.0),
    "sphere": dict(path="configurations/soft_on_sphere/sphere.dat", box=(30.0, 30.0, 30.0), periodic=False,
                   nlist=(2.4, 0.5), constraint=("sphere", 10.0, 0), soft=(10.0, 1.0, False), nematic=0.5),
    "lj": dict(path="configurations/lj_langevin/particles.dat", box=(100.0, 100.0, 10.0), periodic=True,
               nlist=(3.0, 0.5), constraint=("plane", 0.0, 0), lj=(1.0, 1.0, 2.5, True)),
    "relax": dict(path="configurations/relax/out.input", box=(100.0, 100.0, 10.0), periodic=False,
                  nlist=(2.5, 0.5), constraint=("plane", 0.0, 1), soft=(10.0, 2.0, False), polar=0.3),
}


def parse(path):
    """Reader for the reference input format (src/system/system.cpp:64-452):
    '#' comments, optional 'keys:' header naming the columns, else the old
    fixed column order."""
    keys = None
    rows = []
    with open(path) as fh:
        for line in fh:
            t = line.strip().lower()
            if not t or t.startswith("#"):
                continue
            if t.startswith("keys:"):
                keys = t[5:].split()
                continue
            rows.append([float(v) for v in t.split()])
    width = len(rows[0])
    if keys is None:
        keys = OLD_KEYS[:width]
    a = np.array(rows)
    cols = {k: a[:, i] for i, k in enumerate(keys)}
    n = len(rows)
    out = {}
    for k in ["x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz"]:
        out[k] = cols.get(k, np.zeros(n)).astype(np.float64)
    out["radius"] = cols.get("radius", np.ones(n)).astype(np.float64)
    out["type"] = cols.get("type", np.ones(n)).astype(np.int32)
    return out


def build_oracle(inp, case):
    n = len(inp["x"])
    s = oracle.OracleSystem(n, case["box"], case["periodic"])
    for k, v in inp.items():
        s.set(k, v)
    s.nlist(*case["nlist"])
    if "soft" in case:
        s.pair_soft(*case["soft"])
    if "lj" in case:
        eps, sig, rc, shifted = case["lj"]
        s.pair_lj(eps, sig, rc, shifted)
    if "polar" in case:
        s.align_polar(case["polar"])
    if "nematic" in case:
        s.align_nematic(case["nematic"])
    kind, param, flag = case["constraint"]
    s.constraint(kind, param, flag)
    return s


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def main(ref):
    manifest = {}
    for name, case in CASES.items():
        inp = parse(os.path.join(ref, case["path"]))
        ip = os.path.join(HERE, f"inputs_{name}.npz")
        np.savez_compressed(ip, **inp)
        s = build_oracle(inp, case)
        # state after the constraint command, then one loop-head evaluation
        e = s.compute_forces()
        out = {"pairs": s.pairs(), "energy_pair": np.float64(e["pair"]), "energy_align": np.float64(e["align"])}
        for k in ["x", "y", "z", "nx", "ny", "nz", "fx", "fy", "fz", "tau_x", "tau_y", "tau_z"]:
            out[k] = s.get(k)
        gp = os.path.join(HERE, f"golden_{name}.npz")
        np.savez_compressed(gp, **out)
        man = {k: v for k, v in case.items()}
        man.update(n=len(inp["x"]), npairs=int(len(out["pairs"])), inputs_sha256=sha(ip), golden_sha256=sha(gp))
        manifest[name] = man
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, default=list)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
