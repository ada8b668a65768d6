"""Pins of the oracle's collision frames (orc_run_collide): what the
reference's rules fix without its main.cpp, which does not compile here.

- Range 0 detects nothing and reduces to the plain frames (orc_run).
- Detection equals brute force: pairs closer than the range, found from
  both ends when each reaches the other as a leaf (src/main.cpp:386-396).
- Merging is transitive (the update table of src/main.cpp:472-490): a chain
  a-b, b-c merges into one particle even though a and c are far apart.
- A merge conserves mass exactly for two bodies, momentum and the centre of
  mass to rounding, and leaves the merged velocity unchanged by the update
  (its acceleration is zero).
- A planted mistake in the weighting (masses swapped) breaks the centre of
  mass, and one in the order of the survivors breaks their identities."""
import numpy as np
import pytest

import oracle as O


def planted(seed=5, n=400, rng_range=0.01):
    """A cube of n particles plus close pairs, a triple and a chain."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-50, 50, size=(n, 3))
    extra = []
    for k in range(6):  # pairs at 0.3 range
        c = rng.uniform(-40, 40, size=3)
        extra += [c, c + np.array([0.3 * rng_range, 0, 0])]
    c = np.array([10.0, 20.0, 30.0])  # a triple, all within range
    extra += [c, c + [0, 0.2 * rng_range, 0], c + [0, 0, 0.25 * rng_range]]
    c = np.array([-25.0, 5.0, 12.0])  # a chain: ends 1.4 range apart
    extra += [c, c + [0.7 * rng_range, 0, 0], c + [1.4 * rng_range, 0, 0]]
    pos = np.concatenate([pos, np.array(extra)])
    perm = rng.permutation(len(pos))
    pos = pos[perm]
    vel = rng.uniform(-1, 1, size=pos.shape)
    mass = rng.uniform(1e6, 1e7, size=len(pos))
    return pos, vel, mass


def components(pos, rng_range):
    """Brute force: groups of the graph of pairs closer than the range."""
    n = len(pos)
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    adj = (d < rng_range) & ~np.eye(n, dtype=bool)
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x
    for i, j in zip(*np.nonzero(adj)):
        a, b = find(i), find(j)
        if a != b:
            parent[max(a, b)] = min(a, b)
    groups = {}
    for i in range(n):
        groups.setdefault(find(i), []).append(i)
    return [g for g in groups.values() if len(g) > 1], int(adj.sum())


def test_range_zero_is_the_plain_frame():
    pos, vel, mass = planted()
    p, v, a, m, npairs, nrem = O.run_collide(pos, vel, mass, frames=3, range=0.0)
    p2, v2, a2 = O.run(pos, vel, mass, frames=3)
    assert npairs == 0 and nrem == 0
    np.testing.assert_array_equal(p, p2)
    np.testing.assert_array_equal(v, v2)
    np.testing.assert_array_equal(a, a2)
    np.testing.assert_array_equal(m, mass)


def test_detection_matches_brute_force_and_chains_merge():
    r = 0.01
    pos, vel, mass = planted(rng_range=r)
    groups, ndirected = components(pos, r)
    sizes = sorted(len(g) for g in groups)
    assert sizes == [2] * 6 + [3, 3]  # the chain merges whole
    p, v, a, m, npairs, nrem = O.run_collide(pos, vel, mass, frames=1, range=r)
    assert npairs == ndirected  # each close pair found from both ends
    assert nrem == sum(len(g) - 1 for g in groups)
    assert len(m) == len(mass) - nrem
    # survivors in order; each merged particle in its lowest member's place
    keep = np.ones(len(mass), dtype=bool)
    for g in groups:
        keep[sorted(g)[1:]] = False
    idx = np.nonzero(keep)[0]
    solo = np.array([all(i not in g for g in groups) for i in idx])
    np.testing.assert_array_equal(m[solo], mass[idx[solo]])
    for g in groups:
        j = int(np.searchsorted(idx, min(g)))
        gm = mass[g]
        if len(g) == 2:
            assert m[j] == gm[0] + gm[1]  # exact for two bodies
        assert abs(m[j] - gm.sum()) <= 1e-15 * gm.sum()
        mom = (gm[:, None] * vel[g]).sum(0)
        np.testing.assert_allclose(m[j] * v[j], mom, rtol=1e-14, atol=1e-9)
        # the merged particle drifts from the centre of mass with zero
        # acceleration
        com = (gm[:, None] * pos[g]).sum(0) / gm.sum()
        np.testing.assert_allclose(p[j], com + v[j] * 0.0333, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(a[j], 0.0)


def test_planted_mistakes_are_caught():
    r = 0.01
    pos, vel, mass = planted(rng_range=r)
    groups, _ = components(pos, r)
    g = sorted(groups[0])
    p, v, a, m, _, _ = O.run_collide(pos, vel, mass, frames=1, range=r)
    keep = np.ones(len(mass), dtype=bool)
    for gg in groups:
        keep[sorted(gg)[1:]] = False
    idx = np.nonzero(keep)[0]
    j = int(np.searchsorted(idx, g[0]))
    gm = mass[g]
    com = (gm[:, None] * pos[g]).sum(0) / gm.sum()
    # weights swapped: the centre-of-mass check fails
    wrong = (pos[g[0]] * gm[1] + pos[g[1]] * gm[0]) / gm.sum() + v[j] * 0.0333
    assert np.abs(wrong - (com + v[j] * 0.0333)).max() > 1e-9
    assert np.abs(p[j] - wrong).max() > 1e-9
    # survivors out of order: identities (masses) no longer match
    solo = [i for i in range(len(idx)) if all(idx[i] not in gg for gg in groups)]
    assert not np.array_equal(m[solo][::-1], mass[idx[solo]])


def test_whole_cluster_merges_into_one():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 0.004, size=(60, 3))
    vel = rng.uniform(-1, 1, size=(60, 3))
    mass = rng.uniform(1e3, 1e4, size=60)
    p, v, a, m, npairs, nrem = O.run_collide(pos, vel, mass, frames=1, range=0.01)
    assert len(m) == 1 and nrem == 59
    # pairs are detected only where a leaf is reached: cells of the blob
    # far enough (side / distance <= theta) are accepted whole, so fewer
    # than all 60 * 59 ordered pairs are seen, and transitivity merges all
    assert 0 < npairs < 60 * 59
    assert abs(m[0] - mass.sum()) <= 1e-15 * mass.sum()
    np.testing.assert_allclose(m[0] * v[0], (mass[:, None] * vel).sum(0), rtol=1e-13)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("frames", [1, 3])
def test_against_the_reference_classes(frames):
    """The reference's own distance(), weighted_average(), particle and
    octree code (oracle/_ref): the same particles absorbed, and everything
    bit-identical outside groups of three or more (whose merge order the
    reference leaves to its hash set), those within rounding."""
    r = 0.01
    pos, vel, mass = planted(rng_range=r)
    groups, ndirected = components(pos, r)
    p, v, a, m, npairs, nrem = O.run_collide(pos, vel, mass, frames=frames, range=r)
    rp, rv, ra, rm, rpairs = O.ref_run_collide(pos, vel, mass, frames=frames, range=r)
    assert len(m) == len(rm) == len(mass) - nrem
    # the oracle counts detections from both ends; the reference's set keeps
    # a close pair once, lower x first -- both orientations when the x are
    # equal (the planted triple)
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    i, j = np.nonzero(np.triu(d < r, 1))
    assert npairs == ndirected == 2 * len(i)
    if frames == 1:
        assert rpairs == len(i) + int((pos[i, 0] == pos[j, 0]).sum())
    keep = np.ones(len(mass), dtype=bool)
    for g in groups:
        keep[sorted(g)[1:]] = False
    idx = np.nonzero(keep)[0]
    big = np.array([any(i == min(g) and len(g) > 2 for g in groups) for i in idx])
    for x, y in ((p, rp), (v, rv), (a, ra), (m, rm)):
        np.testing.assert_array_equal(x[~big], y[~big])
        np.testing.assert_allclose(x[big], y[big], rtol=1e-14, atol=1e-12)


# ---------------------------------------------------------- minimum size --
def collide_np(a, b):
    """collide() of src/main.cpp:423-430 on (m, x[3], v[3]) tuples."""
    (am, ax, av), (bm, bx, bv) = a, b
    return (am + bm, (ax * am + bx * bm) / (am + bm), (av * am + bv * bm) / (am + bm))


def test_min_size_below_every_cell_changes_nothing():
    pos, vel, mass = planted()
    p, v, a, m, _, nrem = O.run_collide(pos, vel, mass, frames=2, range=-1.0, min_size=1e-9)
    p2, v2, a2 = O.run(pos, vel, mass, frames=2)
    assert nrem == 0
    np.testing.assert_array_equal(p, p2)
    np.testing.assert_array_equal(v, v2)


def test_min_size_groups_are_the_occupied_cells():
    rng = np.random.default_rng(8)
    # points on a lattice of spacing 0.25 (plus jitter) inside a cube whose
    # bbox root has side 8 + 2 = 10: with min_size 0.6 the largest cells
    # below it have side 10 / 32 = 0.3125
    pos = rng.integers(0, 32, size=(3000, 3)) * 0.25 + rng.uniform(0.02, 0.06, size=(3000, 3))
    pos[0] = [0.0, 0.0, 0.0]
    pos[1] = [8.0, 8.0, 8.0]
    vel = np.zeros_like(pos)
    mass = rng.uniform(1e3, 1e4, size=len(pos))
    _, _, _, m, _, nrem = O.run_collide(pos, vel, mass, frames=1, range=-1.0, min_size=0.6)
    c, side = O.root(pos)
    lo = np.asarray(c) - side / 2
    cells = np.floor((pos - lo) / (side / 32)).astype(int)
    occupied = len({tuple(r) for r in cells})
    assert len(m) == occupied and nrem == len(pos) - occupied
    assert abs(m.sum() - mass.sum()) <= 1e-12 * mass.sum()


def test_min_size_above_the_root_merges_all():
    rng = np.random.default_rng(9)
    pos = rng.uniform(-1, 1, size=(50, 3))
    vel = rng.uniform(-1, 1, size=(50, 3))
    mass = rng.uniform(1.0, 2.0, size=50)
    p, v, a, m, _, nrem = O.run_collide(pos, vel, mass, frames=1, range=-1.0, min_size=100.0)
    assert len(m) == 1 and nrem == 49
    assert abs(m[0] - mass.sum()) <= 1e-14 * mass.sum()
    np.testing.assert_allclose(m[0] * v[0], (mass[:, None] * vel).sum(0), rtol=1e-13)
    com = (mass[:, None] * pos).sum(0) / mass.sum()
    np.testing.assert_allclose(p[0], com + v[0] * 0.0333, atol=1e-13)


def test_min_size_folds_along_the_tree():
    # root: bbox [0, 8]^3 + 2 -> side 10, centre 4.  C sits in octant 0 of
    # the root; A and B in octant 7, where they part one level down (B in
    # child 1, A in child 7).  With min_size above the root side everything
    # merges as collide(C, collide(B, A)), not as the flat fold in tree
    # order collide(collide(C, B), A) (these values tell the two apart)
    A = (6.733, np.array([8.0, 8.0, 8.0]), np.array([-0.967, 0.627, 0.826]))
    B = (3.428, np.array([7.9, 7.7, 7.3]), np.array([0.213, 0.459, 0.087]))
    C = (1.369, np.array([0.0, 0.0, 0.0]), np.array([0.87, 0.632, -0.995]))
    pos = np.array([A[1], B[1], C[1]])
    vel = np.array([A[2], B[2], C[2]])
    mass = np.array([A[0], B[0], C[0]])
    p, v, a, m, _, nrem = O.run_collide(pos, vel, mass, frames=1, dt=0.0, range=-1.0,
                                        min_size=100.0)
    nested = collide_np(C, collide_np(B, A))
    flat = collide_np(collide_np(C, B), A)
    assert m[0] == nested[0]
    np.testing.assert_array_equal(p[0], nested[1])
    np.testing.assert_array_equal(v[0], nested[2])
    assert not (np.array_equal(nested[1], flat[1]) and np.array_equal(nested[2], flat[2]))
