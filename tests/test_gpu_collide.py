"""Collision merging on the GPU (nbs_set_collisions) against the oracle's
collision frames (oracle.run_collide, itself pinned by
tests/test_oracle_collide.py).  Bar: the same pairs detected (count), the
same particles absorbed, masses bit-identical (both fold collide over a
group's members in increasing index), positions and velocities within the
FP64 walk's tolerance (the forces are summed in another order)."""
import numpy as np
import pytest

import nbodysim_amd
from nbodysim_amd import NbsError
import oracle as O

from test_oracle_collide import planted

pytestmark = pytest.mark.gpu

R = 0.01


def gpu_collide(pos, vel, mass, steps, precision="f64", offsets=None, r=R):
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision=precision) as s:
        s.set_state(pos, vel, mass)
        if offsets is not None:
            s.set_systems(offsets)
        s.set_collisions(r)
        s.step(steps)
        p, v, a = s.state()
        m = s.masses()
        stats = s.collision_stats()
        off = s.offsets(len(offsets) - 1) if offsets is not None else None
    return p, v, a, m, stats, off


def close(a, b, scale):
    assert a.shape == b.shape
    err = np.abs(a - b).max() if a.size else 0.0
    assert err <= 1e-10 * scale, err


@pytest.mark.parametrize("steps", [1, 4])
def test_planted_groups_match_oracle(steps):
    pos, vel, mass = planted(rng_range=R)
    p, v, a, m, (npairs, nrem), _ = gpu_collide(pos, vel, mass, steps)
    op, ov, oa, om, onpairs, onrem = O.run_collide(pos, vel, mass, frames=steps, range=R)
    print("steps %d: %d pairs, %d absorbed, %d left" % (steps, npairs, nrem, len(m)),
          flush=True)
    assert (npairs, nrem) == (onpairs, onrem)
    assert nrem == 10  # six pairs, a triple and a three-particle chain
    np.testing.assert_array_equal(m, om)
    close(p, op, 50.0)
    close(v, ov, 1.0)
    if steps == 1:
        # merged particles: zero acceleration, velocity as merged (exact)
        merged = ~np.isin(om, mass)  # merged masses are sums, not input masses
        assert merged.sum() == 8
        np.testing.assert_array_equal(a[merged], 0.0)
        np.testing.assert_array_equal(v[merged], ov[merged])


def test_fp32_forces_detect_the_same_pairs():
    pos, vel, mass = planted(rng_range=R)
    _, _, _, m, (npairs, nrem), _ = gpu_collide(pos, vel, mass, 2, precision="f32")
    _, _, _, om, onpairs, onrem = O.run_collide(pos, vel, mass, frames=2, range=R)
    assert (npairs, nrem) == (onpairs, onrem)
    np.testing.assert_array_equal(m, om)


def test_dense_cluster_grows_the_pair_buffer():
    # 700 particles within 0.004 of each other: far more detections than
    # the initial pair buffer (4096) holds; the step is redone with a larger
    # one and the whole cluster becomes one particle
    rng = np.random.default_rng(3)
    blob = rng.uniform(0, 0.004, size=(700, 3))
    far = rng.uniform(50, 100, size=(300, 3))
    pos = np.concatenate([blob, far])
    vel = rng.uniform(-1, 1, size=pos.shape)
    mass = rng.uniform(1e3, 1e4, size=len(pos))
    p, v, a, m, (npairs, nrem), _ = gpu_collide(pos, vel, mass, 1)
    op, ov, oa, om, onpairs, onrem = O.run_collide(pos, vel, mass, frames=1, range=R)
    assert npairs > 4096
    assert (npairs, nrem) == (onpairs, onrem) and nrem == 699
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(p[0], op[0])  # merged: drift only, bit-exact
    np.testing.assert_array_equal(v[0], ov[0])
    close(p, op, 100.0)


def test_batched_systems_shrink_their_offsets():
    parts = [planted(seed=5 + k, n=300 + 50 * k, rng_range=R) for k in range(3)]
    pos = np.concatenate([q[0] for q in parts])
    vel = np.concatenate([q[1] for q in parts])
    mass = np.concatenate([q[2] for q in parts])
    off = np.cumsum([0] + [len(q[2]) for q in parts])
    p, v, a, m, (npairs, nrem), noff = gpu_collide(pos, vel, mass, 2, offsets=off)
    want_off = [0]
    tot_pairs = tot_rem = 0
    for k, q in enumerate(parts):
        op, ov, oa, om, onp, onr = O.run_collide(*q, frames=2, range=R)
        want_off.append(want_off[-1] + len(om))
        tot_pairs += onp
        tot_rem += onr
        a0, a1 = want_off[k], want_off[k + 1]
        np.testing.assert_array_equal(m[a0:a1], om)
        close(p[a0:a1], op, 50.0)
        close(v[a0:a1], ov, 1.0)
    assert list(noff) == want_off
    assert (npairs, nrem) == (tot_pairs, tot_rem)


def test_collisions_off_and_rejections():
    pos, vel, mass = planted(rng_range=R)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        s.set_collisions(R)
        with pytest.raises(NbsError):
            s.step(2, graph=True)  # eager steps only
        s.step(1)
        n1 = s.n
        assert n1 == len(mass) - 10
        s.set_collisions(enabled=False)
        s.step(2, graph=True)
        assert s.n == n1
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, adaptive=True) as s:
        s.set_state(pos, vel, mass)
        s.set_collisions(R)
        with pytest.raises(NbsError):
            s.step(1)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        with pytest.raises(NbsError):
            s.set_collisions(-1.0)


def test_fixed_root_merges_the_same_groups():
    # a fixed root prunes the pair search by the MAC alone; the planted
    # groups are far from every other particle, so they merge as with the
    # bbox root of the oracle
    pos, vel, mass = planted(rng_range=R)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, root_center=(0.0, 0.0, 0.0),
                                 root_size=512.0) as s:
        s.set_state(pos, vel, mass)
        s.set_collisions(R)
        s.step(1)
        m = s.masses()
        _, nrem = s.collision_stats()
    _, _, _, om, _, onrem = O.run_collide(pos, vel, mass, frames=1, range=R)
    assert nrem == onrem == 10
    np.testing.assert_array_equal(m, om)


# ---------------------------------------------------------- minimum size --
def gpu_min_size(pos, vel, mass, steps, min_size, dt=0.0333, r=None):
    with nbodysim_amd.Simulation(theta=0.5, dt=dt) as s:
        s.set_state(pos, vel, mass)
        s.set_min_node_size(min_size)
        if r is not None:
            s.set_collisions(r)
        s.step(steps)
        p, v, a = s.state()
        m = s.masses()
        merged = s.set_min_node_size(min_size)
    return p, v, m, merged


def test_min_size_lattice_matches_oracle():
    rng = np.random.default_rng(8)
    pos = rng.integers(0, 32, size=(3000, 3)) * 0.25 + rng.uniform(0.02, 0.06, size=(3000, 3))
    pos[0] = [0.0, 0.0, 0.0]
    pos[1] = [8.0, 8.0, 8.0]
    vel = rng.uniform(-0.01, 0.01, size=pos.shape)
    mass = rng.uniform(1e3, 1e4, size=len(pos))
    p, v, m, merged = gpu_min_size(pos, vel, mass, 2, 0.6)
    op, ov, oa, om, _, onrem = O.run_collide(pos, vel, mass, frames=2, range=-1.0, min_size=0.6)
    print("min size: %d absorbed, %d left" % (merged, len(m)), flush=True)
    assert merged == onrem > 0
    np.testing.assert_array_equal(m, om)
    close(p, op, 10.0)
    close(v, ov, 1.0)


def test_min_size_folds_along_the_tree():
    pos = np.array([[8.0, 8.0, 8.0], [7.9, 7.7, 7.3], [0.0, 0.0, 0.0]])
    vel = np.array([[-0.967, 0.627, 0.826], [0.213, 0.459, 0.087], [0.87, 0.632, -0.995]])
    mass = np.array([6.733, 3.428, 1.369])
    p, v, m, merged = gpu_min_size(pos, vel, mass, 1, 100.0, dt=0.0)
    op, ov, _, om, _, _ = O.run_collide(pos, vel, mass, frames=1, dt=0.0, range=-1.0,
                                        min_size=100.0)
    assert merged == 2
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(v, ov)


def test_min_size_with_collisions():
    pos, vel, mass = planted(rng_range=R)
    p, v, m, merged = gpu_min_size(pos, vel, mass, 2, 0.6, r=R)
    op, ov, _, om, _, onrem = O.run_collide(pos, vel, mass, frames=2, range=R, min_size=0.6)
    assert len(m) == len(om) == len(mass) - 10
    np.testing.assert_array_equal(m, om)
    close(p, op, 50.0)


def test_min_size_then_collisions_that_grow_the_pair_buffer():
    # the dense cluster overflows the initial pair buffer: the step is redone
    # keeping its minimum-size merge (and root) as they were
    rng = np.random.default_rng(3)
    blob = rng.uniform(0, 0.004, size=(700, 3))
    far = rng.uniform(50, 100, size=(300, 3))
    pos = np.concatenate([blob, far])
    vel = rng.uniform(-1, 1, size=pos.shape)
    mass = rng.uniform(1e3, 1e4, size=len(pos))
    p, v, m, merged = gpu_min_size(pos, vel, mass, 1, 3e-4, r=R)
    op, ov, _, om, _, onrem = O.run_collide(pos, vel, mass, frames=1, range=R, min_size=3e-4)
    assert len(m) == len(om) == 301
    np.testing.assert_array_equal(m, om)
    close(p, op, 100.0)


def test_merging_edge_cases():
    # two particles in range merge into one, which then steps on alone
    pos = np.array([[0.0, 0.0, 0.0], [0.004, 0.0, 0.0]])
    vel = np.array([[1.0, 0.0, 0.0], [-1.0, 0.5, 0.0]])
    mass = np.array([2.0, 6.0])
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        s.set_collisions(R)
        s.step(3)
        p, v, a = s.state()
        m = s.masses()
        assert s.n == 1 and s.collision_stats()[1] == 1
        img = s.render_image(64, 48, projection="front", brightness=10.0)
    op, ov, oa, om, _, _ = O.run_collide(pos, vel, mass, frames=3, range=R)
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(a, 0.0)
    np.testing.assert_array_equal(img, O.render_image(p, 64, 48, projection="front",
                                                      brightness=10.0))
    # an empty state: steps with merging on do nothing; its image is black
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        s.set_state(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
        s.set_collisions(R)
        s.set_min_node_size(1.0)
        s.step(2)
        assert s.n == 0
        assert not s.render_image(32, 16, adaptive=True, color="heat").any()
