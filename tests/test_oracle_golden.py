"""The oracle against golden vectors made by the reference's own compiled
octree/particle/vector code (tests/golden/make_golden.py): trees, forces,
interaction counts and 1-10 frame trajectories, bit for bit."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_golden.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", params=sorted(GOLD["cases"]))
def case(request):
    c = GOLD["cases"][request.param]
    pos, vel, mass = O.generate(c["n"], c["gen_type"], seed=c["seed"], prec=c["prec"])
    return c, pos, vel, mass


def test_initial_conditions_digest(case):
    c, pos, vel, mass = case
    assert sha(pos) == c["ic"]["pos"]
    assert sha(vel) == c["ic"]["vel"]
    assert sha(mass) == c["ic"]["mass"]


def test_tree_bit_exact(case):
    c, pos, vel, mass = case
    d = O.Tree(pos, mass, prec=c["prec"]).dump()
    t = c["tree"]
    assert len(d["depth"]) == t["n_nodes"]
    assert np.bincount(d["depth"]).tolist() == t["depth_hist"]
    assert sha(d["depth"].astype(np.int32)) == t["depth"]
    assert sha(d["leaf"].astype(np.int32)) == t["leaf"]
    assert sha(d["count"].astype(np.int64)) == t["count"]
    assert sha(d["mass"]) == t["mass"]
    assert sha(d["com"]) == t["com"]
    assert sha(d["order"].astype(np.int32)) == t["order"]


def test_forces_bit_exact(case):
    c, pos, vel, mass = case
    acc, cnt = O.Tree(pos, mass, prec=c["prec"]).forces(theta=c["theta"],
                                                        damping=c["damping"])
    assert sha(cnt.astype(np.int64)) == c["interactions"]["sha"]
    np.testing.assert_array_equal(acc[:16], np.array(c["acc0"]["head"]))
    assert sha(acc) == c["acc0"]["sha"]


def test_frames_bit_exact(case):
    c, pos, vel, mass = case
    p, v, a = O.run(pos, vel, mass, theta=c["theta"], dt=c["dt"], frames=c["frames"],
                    damping=c["damping"], prec=c["prec"])
    np.testing.assert_array_equal(p[:8], np.array(c["final"]["pos_head"]))
    assert sha(p) == c["final"]["pos"]
    assert sha(v) == c["final"]["vel"]
    assert sha(a) == c["final"]["acc"]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
def test_oracle_against_live_reference_build():
    pos, vel, mass = O.generate(3000, "shell", seed=12345)
    r = O.ref_tree_dump(pos, mass)
    d = O.Tree(pos, mass).dump()
    for k in ("depth", "leaf", "count", "mass", "com", "order"):
        np.testing.assert_array_equal(d[k].astype(r[k].dtype), r[k], err_msg=k)
    ra, rc = O.ref_forces(pos, mass, theta=0.7)
    a, c = O.Tree(pos, mass).forces(theta=0.7)
    np.testing.assert_array_equal(a, ra)
    np.testing.assert_array_equal(c, rc)
