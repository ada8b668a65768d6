"""Interaction counting modes of the walk (nbs_set_counting): the per-wave
total counted on the scalar unit, and no counting, change no force,
position or velocity bit, in every walk loop (the hand-written FP32 loops
plain and damped, the hand-written FP64 loop with its exact-tie visits, the
compiler-generated loop); the running total equals the per-particle counts'
sum; per-particle counts of a walk that did not keep them are refused."""
import numpy as np
import pytest

from nbodysim_amd import host

pytestmark = pytest.mark.gpu

SPHERE = dict(r_sphere=100.0, rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0))


def sim(**kw):
    import nbodysim_amd
    return nbodysim_amd.Simulation(**kw)


def run(mode, pos, vel, mass, steps, graph=False, generic=False, records="auto", **kw):
    with sim(**kw) as s:
        s.set_counting(mode)
        s.set_walk_code(generic, records)
        s.set_state(pos, vel, mass)
        acc = s.forces()
        total_forces = s.interactions_total(reset=True)
        per = s.interactions(per_particle=True)[1] if mode == "particles" else None
        s.step(steps, graph=graph)
        p, v, a = s.state()
        total_steps = s.interactions_total(reset=True)
    return acc, per, total_forces, (p, v, a), total_steps


CASES = [dict(precision="f32"), dict(precision="f32", records=32),
         dict(precision="f32", damping=True), dict(precision="f32", damping=True, records=32),
         dict(precision="f64"), dict(precision="f32", generic=True)]


@pytest.mark.parametrize("case", CASES, ids=["f32", "f32-rec32", "f32-damped",
                                             "f32-damped-rec32", "f64", "f32-generic"])
def test_counting_modes_change_no_result_bit(case):
    pos, vel, mass = host.generate(20000, gen_type="sphere", seed=11, **SPHERE)
    kw = dict(theta=0.5, dt=0.0333, **case)
    print("counting modes:", case, flush=True)
    ref = run("particles", pos, vel, mass, 3, **kw)
    assert ref[2] == ref[1].sum()
    for mode in ("total", "none"):
        got = run(mode, pos, vel, mass, 3, **kw)
        np.testing.assert_array_equal(got[0], ref[0], err_msg=mode)
        for k in range(3):
            np.testing.assert_array_equal(got[3][k], ref[3][k], err_msg=mode)
        if mode == "total":
            assert got[2] == ref[2]
            assert got[4] == ref[4]
        else:
            assert got[2] == 0 and got[4] == 0


def test_fp64_lattice_total_equals_the_particle_counts():
    # on a lattice, cells whose opening test is within rounding of the
    # boundary go to the compiled exact visit, whose per-lane counts join
    # the scalar total
    g = np.arange(8, dtype=np.float64)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    mass = np.full(len(pos), 1e10)
    per = run("particles", pos, None, mass, 0, theta=0.5, precision="f64")
    tot = run("total", pos, None, mass, 0, theta=0.5, precision="f64")
    assert tot[2] == per[1].sum() == per[2]
    np.testing.assert_array_equal(tot[0], per[0])


def test_graph_replay_follows_the_counting_mode():
    pos, vel, mass = host.generate(20000, gen_type="sphere", seed=12, **SPHERE)
    kw = dict(theta=0.5, dt=0.0333, precision="f32")
    ref = run("particles", pos, vel, mass, 4, graph=True, **kw)
    got = run("total", pos, vel, mass, 4, graph=True, **kw)
    for k in range(3):
        np.testing.assert_array_equal(got[3][k], ref[3][k])
    assert got[4] == ref[4] > 0
    # a mode change drops the captured steps: the last two add nothing
    first_two = run("total", pos, vel, mass, 2, graph=True, **kw)[4]
    with sim(**kw) as s:
        s.set_counting("total")
        s.set_state(pos, vel, mass)
        s.forces()
        s.interactions_total(reset=True)
        s.step(2, graph=True)
        s.set_counting("none")
        s.step(2, graph=True)
        p, v, _ = s.state()
        assert s.interactions_total(reset=True) == first_two
    np.testing.assert_array_equal(p, ref[3][0])
    np.testing.assert_array_equal(v, ref[3][1])


def test_per_particle_counts_refused_after_a_walk_without_them():
    import nbodysim_amd
    pos, vel, mass = host.generate(5000, gen_type="sphere", seed=13, **SPHERE)
    with sim(theta=0.5, precision="f32") as s:
        s.set_state(pos, vel, mass)
        s.set_counting("total")
        s.forces()
        with pytest.raises(nbodysim_amd.NbsError, match="did not count per particle"):
            s.interactions(per_particle=True)
        s.set_counting("particles")
        s.forces()
        tot, per = s.interactions(per_particle=True)
        assert tot == per.sum() > 0
        with pytest.raises(ValueError):
            s.set_counting("sometimes")
