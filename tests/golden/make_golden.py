"""Writes tests/golden/reference_golden.json from the reference's own code.

Runs only where the reference tree is present (it builds oracle/_ref from
the reference's octree.cpp / particle.cpp / vector.cpp with
oracle/build_ref.sh, flags -O2 -ffp-contract=off, no -march=native).  For
each case the fixture holds SHA-256 digests of the reference's outputs
(tree in pre-order: depth, leaf, count, mass, COM, DFS leaf order;
accelerations and interaction counts of frame 0; state after `frames`
frames) plus the first rows in clear for diagnosis, and the raw bytes of
three reference `particle` records (the .dat format).

Initial conditions come from the oracle's generator (the reference's
generate_particle is in main.cpp, which does not compile here); their
digest is stored too, so a change in the generator is caught.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

CASES = [
    # name, gen_type, n, prec, theta, damping, frames
    ("cube5000_f64", "cube", 5000, "f64", 0.5, False, 10),
    ("cube5000_f32", "cube", 5000, "f32", 0.5, False, 10),
    ("sphere4096_f64", "sphere", 4096, "f64", 0.5, False, 2),
    ("shell4096_f64", "shell", 4096, "f64", 0.5, False, 2),
    ("cube1024_theta0_f64", "cube", 1024, "f64", 0.0, False, 1),
    ("cube2000_damped_f64", "cube", 2000, "f64", 0.5, True, 1),
]
SEED = 0xDEADBEEF
DT = 0.0333


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tree_digest(d):
    return {
        "n_nodes": int(len(d["depth"])),
        "depth_hist": np.bincount(d["depth"].astype(np.int64)).tolist(),
        "depth": sha(d["depth"].astype(np.int32)),
        "leaf": sha(d["leaf"].astype(np.int32)),
        "count": sha(d["count"].astype(np.int64)),
        "mass": sha(d["mass"].astype(np.float64)),
        "com": sha(d["com"].astype(np.float64)),
        "order": sha(d["order"].astype(np.int32)),
        "head": {k: d[k][:32].tolist() for k in ("depth", "leaf", "count", "mass", "com")},
    }


def case(name, gen, n, prec, theta, damping, frames):
    pos, vel, mass = O.generate(n, gen, seed=SEED, prec=prec)
    ic = {"pos": sha(pos), "vel": sha(vel), "mass": sha(mass)}
    tree = O.ref_tree_dump(pos, mass, prec=prec)
    acc, cnt = O.ref_forces(pos, mass, theta=theta, damping=damping, prec=prec)
    p, v, a = O.ref_run(pos, vel, mass, theta=theta, dt=DT, frames=frames,
                        damping=damping, prec=prec)
    return {
        "gen_type": gen, "n": n, "prec": prec, "theta": theta, "damping": damping,
        "dt": DT, "seed": SEED, "frames": frames, "ic": ic,
        "tree": tree_digest(tree),
        "acc0": {"sha": sha(acc), "head": acc[:16].tolist()},
        "interactions": {"sha": sha(cnt.astype(np.int64)), "sum": int(cnt.sum())},
        "final": {"pos": sha(p), "vel": sha(v), "acc": sha(a), "pos_head": p[:8].tolist(),
                  "vel_head": v[:8].tolist()},
    }


def main():
    subprocess.check_call(["sh", os.path.join(ROOT, "oracle", "build_ref.sh")])
    out = {
        "generated_by": "tests/golden/make_golden.py",
        "reference_build": "g++ -std=c++11 -O2 -ffp-contract=off -D{DOUBLE,FLOAT} "
                           "src/{octree,particle,vector,thread_functions}.cpp "
                           "+ oracle/ref_harness.cpp",
        "gxx": subprocess.check_output(["g++", "--version"], text=True).splitlines()[0],
        "libc": " ".join(platform.libc_ver()),
        "cases": {},
    }
    for c in CASES:
        out["cases"][c[0]] = case(*c)
        print("case", c[0], "nodes", out["cases"][c[0]]["tree"]["n_nodes"])
    recs = np.array([[1.5, -2.25, 3.0, 0.1, 0.2, 0.3, -1e-3, 2e-3, 4e3, 5e10],
                     [-100.0, 1e-300, 7.0, 0, 0, 0, 0, 0, 0, 1.0],
                     [np.pi, np.e, -0.0, 1, 2, 3, 4, 5, 6, 2.5e11]])
    out["dat_records"] = {
        "values": recs.tolist(),
        "f64_hex": O.ref_particle_raw(recs[:, 0:3], recs[:, 3:6], recs[:, 6:9], recs[:, 9],
                                      "f64").hex(),
        "f32_hex": O.ref_particle_raw(recs[:, 0:3], recs[:, 3:6], recs[:, 6:9], recs[:, 9],
                                      "f32").hex(),
    }
    with open(os.path.join(HERE, "reference_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_golden.json"))


if __name__ == "__main__":
    main()
