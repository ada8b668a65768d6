"""Initial conditions generated on the device (nbs_generate_state) against
the host generator (nbs_generate, itself pinned to the reference's draws by
tests/test_host_contract.py).  Bar: masses and cubes bit-identical (plain
arithmetic and sqrt, each rounded once); sphere and shell coordinates within
1e-14 of the radius (they pass through cbrt, acos, sin and cos, where the
device math library and glibc may differ in the last bits)."""
import numpy as np
import pytest

from nbodysim_amd import NbsError, host

pytestmark = pytest.mark.gpu

CUBE = dict(size=512.0, min_mass=5e10, max_mass=5e11, min_vel=0.0, max_vel=1.0)
SPHERE = dict(r_sphere=100.0, rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0),
              min_mass=5e10, max_mass=5e11)


def sim():
    import nbodysim_amd
    return nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")


def compare(dev, ref, exact, r):
    if exact:
        np.testing.assert_array_equal(dev, ref)
        return 1.0
    err = np.abs(dev - ref).max()
    assert err <= 1e-14 * r, (err, r)
    return float(np.mean(dev == ref))


@pytest.mark.parametrize("seed", [0xDEADBEEF, 0, 2147483647, 12345])
def test_cube_bit_identical(seed):
    n = 20000
    with sim() as s:
        s.generate_state(n, gen_type="cube", seed=seed, **CUBE)
        pos, vel, _ = s.state()
        m = s.masses()
    p, v, mm = host.generate(n, gen_type="cube", seed=seed, **CUBE)
    np.testing.assert_array_equal(pos, p)
    np.testing.assert_array_equal(vel, v)
    np.testing.assert_array_equal(m, mm)


@pytest.mark.parametrize("gen,extra", [("sphere", {}), ("shell", {}),
                                       ("shell", {"shell_jitter": 0.01})])
def test_sphere_and_shell_within_last_bits(gen, extra):
    n = 20000
    kw = dict(SPHERE, **extra)
    with sim() as s:
        s.generate_state(n, gen_type=gen, seed=0xDEADBEEF, **kw)
        pos, vel, _ = s.state()
        m = s.masses()
    p, v, mm = host.generate(n, gen_type=gen, seed=0xDEADBEEF, **kw)
    np.testing.assert_array_equal(m, mm)
    same_p = compare(pos, p, False, 100.0 * 1.01)
    same_v = compare(vel, v, False, 0.1 * 100.0 * 1.01)
    print("%s%s: %.4f of positions, %.4f of velocities bit-identical"
          % (gen, extra, same_p, same_v), flush=True)


def test_batched_systems_match_host_generation_and_step():
    systems = [(3000, "cube", 7), (100, "sphere", 8), (5000, "sphere", 9), (777, "shell", 10),
               (64, "cube", 11), (2048, "shell", 0xDEADBEEF)]
    kw = dict(SPHERE, size=512.0, min_vel=0.0, max_vel=1.0)
    with sim() as s:
        off = s.generate_state(systems=systems, **kw)
        pos, vel, _ = s.state()
        m = s.masses()
        s.step(3)
        got = s.state()
    assert list(off) == list(np.cumsum([0] + [k[0] for k in systems]))
    for k, (n_s, g, seed) in enumerate(systems):
        a, b = off[k], off[k + 1]
        p, v, mm = host.generate(n_s, gen_type=g, seed=seed, **kw)
        np.testing.assert_array_equal(m[a:b], mm)
        compare(pos[a:b], p, g == "cube", 100.0)
        compare(vel[a:b], v, g == "cube", 10.0)
    # the same particles uploaded from the host step to the same state
    with sim() as t:
        t.set_state(pos, vel, m)
        t.set_systems(off)
        t.step(3)
        want = t.state()
    for g_, w in zip(got, want):
        np.testing.assert_array_equal(g_, w)


def test_rejects_variable_draw_distributions():
    with sim() as s:
        with pytest.raises(NbsError):
            s.generate_state(100, gen_type="cube", seed=1, mass_dist="exp", **CUBE)
        with pytest.raises(NbsError):
            s.generate_state(100, gen_type="cube", seed=1, vel_dist="normal", **CUBE)
        # spheres draw no speed: vel_dist does not matter there
        s.generate_state(100, gen_type="sphere", seed=1, vel_dist="normal", **SPHERE)


def test_systems_must_hold_particles():
    # an empty system is rejected on the host, before any device work
    with sim() as s:
        with pytest.raises(NbsError):
            s.generate_state(systems=[(100, "cube", 1), (0, "sphere", 2)], **CUBE)
        s.generate_state(200, gen_type="cube", seed=3, **CUBE)
        with pytest.raises(NbsError):
            s.set_systems(np.array([0, 100, 100, 200]))
        s.set_systems(np.array([0, 100, 200]))
        s.step(1)
