"""The tree the reference builds when threads > 1 (gen_root_threaded,
src/main.cpp:817-886): eight root children allocated up front, each
filled with the particles its inclusive inside() accepts
(src/octree.cpp:320-336), mass and COM by calc_mass_threaded /
calc_com_threaded (src/octree.cpp:230-313).  The oracle's restatement is
checked against the reference's own compiled octree code driven by the
_ref harness (tree, mass, COM, leaf order, forces, interaction counts and
frames, bit for bit) and against what the construction fixes: particles on
the root's centre planes have a leaf in every child whose closed cell
holds them, an empty octant is a childless cell of mass 0 at the origin,
and data off the planes with every octant occupied gives the serial tree."""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


def plane_lattice():
    # 5^3 lattice at -1.5 .. 1.5 with the 0 planes (the root's centre
    # planes: bbox -1.5 .. 1.5), minus the octant x, y, z > 0
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
    # a particle exactly at the root's centre goes into all eight children
    c = np.concatenate([cube[:500], [[0.0, 0.0, 0.0]], [[-100.0, -100.0, -100.0]],
                        [[100.0, 100.0, 100.0]]])
    yield "centre particle", c, rng.uniform(5e10, 5e11, len(c))
    # two clusters at opposite corners of the bbox: only octants 0 and 7
    # of the bbox-centred root hold particles; far from the origin, where
    # an empty child's COM (0, 0, 0) lies, so the walk accepts it
    two = 1000.0 + np.concatenate([rng.uniform(-50, -10, size=(200, 3)),
                                   rng.uniform(10, 50, size=(200, 3))])
    yield "six empty octants", two, rng.uniform(5e10, 5e11, 400)


@needs_ref
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_threaded_tree_is_the_reference_code(prec):
    for name, pos, mass in cases():
        d = O.Tree(pos, mass, prec=prec, threaded=True).dump()
        r = O.ref_tree_dump(pos, mass, prec=prec, threaded=True)
        for k in ("depth", "leaf", "first", "count", "side", "mass", "com", "order"):
            np.testing.assert_array_equal(d[k], r[k], err_msg="%s: %s" % (name, k))


@needs_ref
@pytest.mark.parametrize("damping", [False, True])
def test_threaded_forces_are_the_reference_walk(damping):
    for name, pos, mass in cases():
        a, c = O.Tree(pos, mass, threaded=True).forces(theta=0.5, damping=damping)
        ra, rc = O.ref_forces(pos, mass, 0.5, damping=damping, threaded=True)
        np.testing.assert_array_equal(c, rc, err_msg=name)
        np.testing.assert_array_equal(a, ra, err_msg=name)


@needs_ref
def test_threaded_frames_are_the_reference_loop():
    rng = np.random.default_rng(9)
    pos, mass = plane_lattice()
    vel = rng.normal(size=pos.shape) * 0.01
    for prec in ("f64", "f32"):
        got = O.run(pos, vel, mass, frames=3, prec=prec, threaded=True)
        want = O.ref_run(pos, vel, mass, frames=3, prec=prec, threaded=True)
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)


def test_plane_particles_are_duplicated_and_empty_octant_kept():
    pos, mass = plane_lattice()
    d = O.Tree(pos, mass, threaded=True).dump()
    # the root has exactly eight children (depth 1), in child order
    kids = np.flatnonzero(d["depth"] == 1)
    assert len(kids) == 8
    # leaves per particle: 2^(number of coordinates on a centre plane),
    # counting only the children that exist -- all eight do
    on = (pos == 0.0).sum(axis=1)
    leaves = np.bincount(d["leaf"][d["leaf"] >= 0], minlength=len(pos))
    np.testing.assert_array_equal(leaves, 2 ** on)
    # the octant x, y, z > 0 (child 7) holds only plane particles' copies;
    # the root's mass counts every copy
    np.testing.assert_allclose(d["mass"][0], (mass * leaves).sum(), rtol=1e-14)


def test_empty_octants_are_massless_childless_cells():
    _, pos, mass = list(cases())[3]
    d = O.Tree(pos, mass, threaded=True).dump()
    kids = np.flatnonzero(d["depth"] == 1)
    assert len(kids) == 8
    empty = kids[d["count"][kids] == 0]
    assert len(empty) == 6
    for k in empty:
        assert d["leaf"][k] == -1 and d["mass"][k] == 0.0
        np.testing.assert_array_equal(d["com"][k], 0.0)
        assert k + 1 == len(d["depth"]) or d["depth"][k + 1] <= 1  # no children
    # an empty child the walk accepts is an interaction of mass 0: the
    # counts exceed the serial tree's, the forces do not change
    a, c = O.Tree(pos, mass, threaded=True).forces(theta=0.5)
    sa, sc = O.Tree(pos, mass).forces(theta=0.5)
    assert (c >= sc).all() and (c > sc).any()
    rel = np.linalg.norm(a - sa, axis=1) / np.linalg.norm(sa, axis=1)
    assert rel.max() < 1e-12


def test_off_plane_data_gives_the_serial_tree():
    # every octant occupied and no particle within rounding of a centre
    # plane: the same cells, masses and COMs as gen_root's tree
    rng = np.random.default_rng(2)
    pos = rng.uniform(-100, 100, size=(2000, 3))
    mass = rng.uniform(5e10, 5e11, 2000)
    t = O.Tree(pos, mass, threaded=True).dump()
    s = O.Tree(pos, mass).dump()
    for k in ("depth", "leaf", "first", "count", "side", "mass", "com", "order"):
        np.testing.assert_array_equal(t[k], s[k], err_msg=k)
