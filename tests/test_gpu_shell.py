"""The shell workload (gen_type shell, r = 100, rotation 0.1 about z, FP32
forces, adaptive dt with min_adaptive_dt 0.00333, max_pos_change 3,
max_vel_change 3) against the oracle: FP32 forces at theta 0.3 / 0.5 / 0.7
on the reference's shell (src/main.cpp:253-256) and on the 1 %-jittered one,
the adaptive step length and update restated exactly on the host, and the
tree and forces of a shell collapsed through its centre.

Bars: FP32 forces within 1e-5 rms and 1e-3 max relative of the FP64 oracle
tree's; FP64 interaction counts equal to the oracle's; the adaptive step
length, the simulated time and the kick-drift update bit-exact; the tree
(keys, order, membership, mass and COM) bit-exact."""
import numpy as np
import pytest

import oracle as O
from nbodysim_amd import host

pytestmark = pytest.mark.gpu

ADAPT = dict(adaptive=True, min_adaptive_dt=0.00333, max_pos_change=3.0, max_vel_change=3.0)


def sim(**kw):
    import nbodysim_amd
    return nbodysim_amd.Simulation(**kw)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def shell(n, jitter, seed=0xDEADBEEF, **kw):
    return host.generate(n, gen_type="shell", seed=seed, r_sphere=100.0,
                         rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0),
                         shell_jitter=jitter, **kw)


@pytest.mark.parametrize("jitter", [0.0, 0.01])
@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_shell_f32_forces_match_oracle(theta, jitter):
    pos, vel, mass = shell(12000, jitter)
    oacc, ocnt = O.Tree(pos, mass).forces(theta=theta)
    with sim(theta=theta, precision="f32") as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        s.configure(precision="f64")
        acc64 = s.forces()
        _, cnt = s.interactions(per_particle=True)
    e = rel_err(acc, oacc)
    print("shell theta %.1f jitter %.2f: f32 rms %.2e max %.2e" %
          (theta, jitter, np.sqrt(np.mean(e ** 2)), e.max()), flush=True)
    assert np.sqrt(np.mean(e ** 2)) <= 1e-5
    assert e.max() <= 1e-3
    np.testing.assert_array_equal(cnt, ocnt)
    assert rel_err(acc64, oacc).max() < 1e-12


@pytest.mark.parametrize("jitter", [0.0, 0.01])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_shell_adaptive_steps_follow_the_rule(precision, jitter):
    """Five single adaptive steps: each step length equals the rule applied
    to the step's own forces (returned by nbs_get_state) and the velocities
    before its kick; the time is their running sum; the update is
    t = a h; v += t; t = v h; x += t (src/particle.cpp:66-74) with that h."""
    pos, vel, mass = shell(4096, jitter)
    # masses 100x the configured range, so that the velocity limit acts
    # (max |a| ~ 800 > max_vel_change / dt) as it does at 4M particles
    mass = mass * 100.0
    with sim(theta=0.5, dt=0.0333, precision=precision, **ADAPT) as s:
        s.set_state(pos, vel, mass)
        t = 0.0
        p0, v0 = pos, vel
        hs = []
        for k in range(5):
            s.step(1)
            p1, v1, a = s.state()
            h = O.adaptive_dt(a, v0, 0.0333, 0.00333, 3.0, 3.0)
            t = t + h
            hs.append(h)
            assert s.time()[0] == t, k
            v_want = v0 + a * h
            np.testing.assert_array_equal(v1, v_want)
            np.testing.assert_array_equal(p1, p0 + v_want * h)
            p0, v0 = p1, v1
    print("adaptive steps", precision, jitter, hs, flush=True)
    assert min(hs) < 0.0333  # the limits acted


def test_collapsed_shell_tree_and_forces():
    """A cold heavy shell falls through its centre (masses 1000x, so the
    free fall takes ~34 steps of min_adaptive_dt): the state at the smallest
    median radius over 45 FP64 adaptive steps (the oracle's own frames reach
    ~18 at step 34) gives the oracle's tree bit for bit, its interaction
    counts and its forces to 1e-12, and FP32 forces within the FP32 bars."""
    pos, vel, mass = shell(16384, 0.0)
    mass = mass * 1000.0
    best = None
    with sim(theta=0.5, dt=0.0333, precision="f64", **ADAPT) as s:
        s.set_state(pos, vel, mass)
        for k in range(45):
            s.step(1)
            p, v, _ = s.state()
            r = np.median(np.linalg.norm(p, axis=1))
            if best is None or r < best[0]:
                best = (r, p, v, k + 1)
    r, p, v, steps = best
    print("collapsed shell: median radius %.3g after %d steps" % (r, steps), flush=True)
    assert r < 30.0
    tree = O.Tree(p, mass)
    oacc, ocnt = tree.forces(theta=0.5)
    with sim(theta=0.5, precision="f64") as s:
        s.set_state(p, v, mass)
        acc = s.forces()
        _, cnt = s.interactions(per_particle=True)
        d = s.tree_dump()
        info = s.tree_info()
        s.configure(precision="f32")
        acc32 = s.forces()
    od = tree.dump()
    assert info["n_buckets"] == 0
    assert len(d["depth"]) == len(od["depth"])
    for k in ("depth", "leaf", "first", "count", "khi", "klo"):
        np.testing.assert_array_equal(d[k].astype(np.int64), od[k].astype(np.int64),
                                      err_msg=k)
    for k in ("side", "mass", "com"):
        np.testing.assert_array_equal(d[k], od[k], err_msg=k)
    np.testing.assert_array_equal(d["order"], od["order"])
    np.testing.assert_array_equal(cnt, ocnt)
    assert rel_err(acc, oacc).max() < 1e-12
    e = rel_err(acc32, oacc)
    assert np.sqrt(np.mean(e ** 2)) <= 1e-5 and e.max() <= 1e-3
