"""The tree the reference builds when threads > 1 (gen_root_threaded,
src/main.cpp:817-886; nbs_config.threaded_tree) on the GPU against the
oracle's restatement (itself pinned to the reference's compiled octree code
by tests/test_oracle_threaded.py): eight root children always (an empty
octant is a cell of mass 0 at the origin), particles on the root's centre
planes in every child whose closed cell holds them.

Bars: tree (depth, leaf, DFS leaf order and ranges, node keys, side) and
FP64 mass / COM bit-exact; interaction counts equal; FP64 forces within
1e-12 relative (summation order only); FP32 forces within 1e-5 rms and 1e-3
max; frames within 1e-9 of the coordinate scale."""
import numpy as np
import pytest

import nbodysim_amd
import oracle as O

pytestmark = pytest.mark.gpu


def sim(**kw):
    return nbodysim_amd.Simulation(threaded_tree=True, **kw)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def plane_lattice():
    g = np.array([-1.5, -0.5, 0.0, 0.5, 1.5])
    pos = np.array([[x, y, z] for x in g for y in g for z in g
                    if not (x > 0 and y > 0 and z > 0)])
    return pos, np.linspace(1e10, 2e10, len(pos))


def cases():
    rng = np.random.default_rng(4)
    pos, mass = plane_lattice()
    yield "plane lattice, empty octant", pos, mass
    cube = rng.uniform(-100, 100, size=(3000, 3))
    yield "random cube", cube, rng.uniform(5e10, 5e11, 3000)
    c = np.concatenate([cube[:500], [[0.0, 0.0, 0.0]], [[-100.0, -100.0, -100.0]],
                        [[100.0, 100.0, 100.0]]])
    yield "centre particle", c, rng.uniform(5e10, 5e11, len(c))
    two = 1000.0 + np.concatenate([rng.uniform(-50, -10, size=(200, 3)),
                                   rng.uniform(10, 50, size=(200, 3))])
    yield "six empty octants", two, rng.uniform(5e10, 5e11, 400)
    # a bigger plane case: a 2000-particle lattice of spacing 1 on the
    # planes plus a random cloud, several waves and many duplicates
    # (the eight corners fix the bbox, so the root's centre is the origin)
    g = np.arange(-6.0, 7.0)
    lat = np.array([[x, y, 0.0] for x in g for y in g])
    corners = np.array([[x, y, z] for x in (-6.5, 6.5) for y in (-6.5, 6.5)
                        for z in (-6.5, 6.5)])
    cloud = rng.uniform(-6, 6, size=(5000, 3))
    pts = np.concatenate([lat, corners, cloud])
    yield "plane sheet and cloud", pts, rng.uniform(1e9, 1e10, len(pts))


def assert_tree_equal(gpu, orc, name):
    assert len(gpu["depth"]) == len(orc["depth"]), name
    for k in ("depth", "leaf", "first", "count", "khi", "klo"):
        np.testing.assert_array_equal(gpu[k].astype(np.int64), orc[k].astype(np.int64),
                                      err_msg="%s: %s" % (name, k))
    for k in ("side", "mass", "com"):
        np.testing.assert_array_equal(gpu[k], orc[k], err_msg="%s: %s" % (name, k))
    np.testing.assert_array_equal(gpu["order"], orc["order"], err_msg=name)


def test_threaded_tree_and_forces_match_oracle():
    for name, pos, mass in cases():
        print("threaded tree:", name, len(pos), flush=True)
        tree = O.Tree(pos, mass, threaded=True)
        oacc, ocnt = tree.forces(theta=0.5)
        with sim(theta=0.5) as s:
            s.set_state(pos, None, mass)
            acc = s.forces()
            _, cnt = s.interactions(per_particle=True)
            assert_tree_equal(s.tree_dump(), tree.dump(), name)
            s.configure(precision="f32")
            acc32 = s.forces()
            s.set_walk_code(records=32)  # the walk's 32-byte records, patched too
            np.testing.assert_array_equal(s.forces(), acc32, err_msg=name)
        np.testing.assert_array_equal(cnt, ocnt, err_msg=name)
        assert rel_err(acc, oacc).max() < 1e-12, name
        e = rel_err(acc32, oacc)
        assert np.sqrt(np.mean(e ** 2)) < 1e-5 and e.max() < 1e-3, name


@pytest.mark.parametrize("damping", [False, True])
def test_threaded_frames_follow_the_oracle(damping):
    rng = np.random.default_rng(9)
    pos, mass = plane_lattice()
    vel = rng.normal(size=pos.shape) * 0.01
    opos, ovel, oacc = O.run(pos, vel, mass, theta=0.5, dt=0.0333, frames=4, damping=damping,
                             threaded=True)
    with sim(theta=0.5, dt=0.0333, damping=damping) as s:
        s.set_state(pos, vel, mass)
        s.step(4)
        p, v, a = s.state()
        with sim(theta=0.5, dt=0.0333, damping=damping) as g:  # graph calls run eagerly
            g.set_state(pos, vel, mass)
            g.step(4, graph=True)
            for x, y in zip(g.state(), (p, v, a)):
                np.testing.assert_array_equal(x, y)
    assert np.abs(p - opos).max() < 1e-9 * np.abs(opos).max()
    assert np.abs(v - ovel).max() < 1e-9 * np.abs(ovel).max()


def test_threaded_tree_differs_from_serial_on_planes():
    """On the plane lattice the threaded tree has more leaves than
    particles and every interaction count exceeds the serial tree's where a
    duplicated particle's second leaf is reached."""
    pos, mass = plane_lattice()
    with sim(theta=0.5) as s, nbodysim_amd.Simulation(theta=0.5) as t:
        s.set_state(pos, None, mass)
        t.set_state(pos, None, mass)
        s.forces()
        t.forces()
        info_s, info_t = s.tree_info(), t.tree_info()
        assert info_s["n_nodes"] - info_s["n_internal"] > len(pos)
        assert info_t["n_nodes"] - info_t["n_internal"] == len(pos)
        assert s.interactions() > t.interactions()


def test_threaded_collisions_and_limits():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-2, 2, size=(2000, 3))
    mass = rng.uniform(1e9, 1e10, 2000)
    with sim(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, None, mass)
        s.set_collisions(0.05)
        s.step(2)
        assert s.n < 2000
        with pytest.raises(nbodysim_amd.NbsError):
            s.keys()
    with sim(theta=0.5) as s:
        s.set_state(pos, None, mass)
        s.set_systems([0, 1000, 2000])
        with pytest.raises(nbodysim_amd.NbsError, match="threaded"):
            s.step(1)
    with sim(theta=0.5) as s:
        s.set_state(pos, None, mass)
        s.set_min_node_size(0.5)
        with pytest.raises(nbodysim_amd.NbsError, match="threaded"):
            s.step(1)
