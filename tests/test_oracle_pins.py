"""Pins of the oracle that do not come from the reference build: known
answers of the random engine (C++ standard), an exact-integer restatement
of libstdc++'s generate_canonical, g++'s argument evaluation order, and the
physics closed forms the walk must reduce to."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

G = 6.67384e-11


def test_minstd_rand0_known_answer():
    # [rand.predef]: the 10000th invocation of a default-constructed
    # minstd_rand0 produces 1043618065.
    assert int(O.minstd_raw(10000)[-1]) == 1043618065


def test_minstd_seed_is_taken_mod_m():
    # seed(s) stores s mod (2^31 - 1); 0xDEADBEEF -> 1588444912
    a = O.minstd_raw(5, seed=0xDEADBEEF)
    b = O.minstd_raw(5, seed=1588444912)
    np.testing.assert_array_equal(a, b)
    x = 1588444912
    want = []
    for _ in range(5):
        x = (x * 16807) % 2147483647
        want.append(x)
    assert a.tolist() == want


def _canonical_double(g1, g2):
    # libstdc++ generate_canonical<double, 53> with minstd_rand0: two draws,
    # r = 2^31 - 2, sum = (g1 - 1) + (g2 - 1) * r, divided by r^2 (rounded
    # to double after the long-double product).
    r = 2147483646.0
    s = float(g1 - 1) + float(g2 - 1) * r
    tmp = float(np.longdouble(r) * np.longdouble(r))
    v = s / tmp
    return v if v < 1.0 else np.nextafter(1.0, 0.0)


def test_uniform_double_restated_from_engine():
    raw = O.minstd_raw(40, seed=0xDEADBEEF).astype(np.int64)
    u = O.uniform_raw(20, -256.0, 256.0, 0xDEADBEEF)
    for k in range(20):
        c = _canonical_double(raw[2 * k], raw[2 * k + 1])
        assert u[k] == c * 512.0 + (-256.0)


def test_uniform_float_uses_one_draw():
    raw = O.minstd_raw(10, seed=99).astype(np.int64)
    u = O.uniform_raw(10, 0.0, 1.0, 99, prec="f32")
    for k in range(10):
        c = np.float32(raw[k] - 1) / np.float32(2147483646.0)
        c = c if c < 1 else np.nextafter(np.float32(1), np.float32(0))
        assert np.float32(u[k]) == np.float32(c * np.float32(1.0) + np.float32(0.0))


def test_gxx_evaluates_constructor_arguments_right_to_left():
    # src/main.cpp:222 builds vector(draw(), draw(), draw()); with g++ the
    # last argument is evaluated first, so z gets the first draw.
    src = r"""
    #include <cstdio>
    static int k = 0;
    int draw() { return ++k; }
    struct V { int x, y, z; V(int a, int b, int c) : x(a), y(b), z(c) {} };
    int main() { V v(draw(), draw(), draw()); printf("%d %d %d\n", v.x, v.y, v.z); }
    """
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        open(cpp, "w").write(src)
        try:
            subprocess.check_call(["g++", "-std=c++11", "-O2", cpp, "-o", exe])
        except (OSError, subprocess.CalledProcessError):
            pytest.skip("g++ not available")
        out = subprocess.check_output([exe], text=True).split()
    assert out == ["3", "2", "1"]


def test_cube_draw_order():
    # first particle: mass, then z, y, x of the position
    raw = O.minstd_raw(8, seed=0xDEADBEEF).astype(np.int64)
    pos, vel, mass = O.generate(1, "cube")
    c = [_canonical_double(raw[2 * k], raw[2 * k + 1]) for k in range(4)]
    assert mass[0] == c[0] * (5e11 - 5e10) + 5e10
    assert pos[0, 2] == c[1] * 512.0 - 256.0
    assert pos[0, 1] == c[2] * 512.0 - 256.0
    assert pos[0, 0] == c[3] * 512.0 - 256.0


def test_theta_zero_is_direct_sum():
    pos, vel, mass = O.generate(600, "sphere")
    acc, cnt = O.Tree(pos, mass).forces(theta=0.0)
    assert (cnt == 599).all()
    d = O.direct(pos, mass)
    rel = np.linalg.norm(acc - d, axis=1) / np.linalg.norm(d, axis=1)
    assert rel.max() < 1e-12


def test_two_body_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    mass = np.array([2e10, 5e10])
    acc, cnt = O.Tree(pos, mass).forces(theta=0.5)
    r = 5.0
    np.testing.assert_allclose(np.linalg.norm(acc[0]), G * 5e10 / r ** 2, rtol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(acc[1]), G * 2e10 / r ** 2, rtol=1e-15)
    np.testing.assert_allclose(acc[0] / np.linalg.norm(acc[0]), [0.6, 0.8, 0], atol=1e-15)


def test_damped_law_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.0]])
    mass = np.array([1e10, 1e10])
    acc, _ = O.Tree(pos, mass).forces(theta=0.5, damping=True)
    want = G * 1e10 / (4.0 + np.exp(-2.0))
    np.testing.assert_allclose(acc[0, 2], want, rtol=1e-15)


def test_lattice_mirror_symmetry_and_momentum():
    g = np.array([-1.5, -0.5, 0.5, 1.5])
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    mass = np.full(len(pos), 1e10)
    acc, _ = O.Tree(pos, mass).forces(theta=0.0)
    idx = {tuple(p): i for i, p in enumerate(pos)}
    for i, p in enumerate(pos):
        j = idx[(-p[0], p[1], p[2])]
        assert abs(acc[i, 0] + acc[j, 0]) <= 1e-13 * abs(acc[i, 0])
        assert abs(acc[i, 1] - acc[j, 1]) <= 1e-13 * max(abs(acc[i, 1]), 1e-30)
    f = mass[:, None] * acc
    assert np.abs(f.sum(0)).max() < 1e-12 * np.abs(f).max()


def test_constant_acceleration_update():
    # one far-away heavy body gives an almost constant field; the kick-drift
    # update is x1 = x0 + (v0 + a dt) dt
    pos = np.array([[0.0, 0.0, 0.0], [1e6, 0.0, 0.0]])
    vel = np.array([[1.0, 2.0, 3.0], [0.0, 0.0, 0.0]])
    mass = np.array([1.0, 1e20])
    p, v, a = O.run(pos, vel, mass, theta=0.5, dt=0.5, frames=1)
    np.testing.assert_allclose(v[0], vel[0] + a[0] * 0.5, rtol=0, atol=1e-15)
    np.testing.assert_allclose(p[0], pos[0] + v[0] * 0.5, rtol=0, atol=1e-15)


def test_tree_membership_invariants():
    pos, vel, mass = O.generate(2000, "shell")
    d = O.Tree(pos, mass).dump()
    leaf = d["leaf"] >= 0
    assert leaf.sum() == 2000 and (d["count"][leaf] == 1).all()
    assert (d["count"][~leaf] >= 2).all()  # a cell splits iff it holds >= 2
    # each node's particles form the contiguous DFS-leaf range it reports
    assert d["count"][0] == 2000
    np.testing.assert_allclose(d["mass"][0], mass.sum(), rtol=1e-13)


def test_planted_mistakes_are_caught():
    """A wrong sign, a dropped term or a swapped axis in the force breaks a
    closed-form check: the checks above are sensitive to each."""
    pos = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    mass = np.array([2e10, 5e10])
    acc, _ = O.Tree(pos, mass).forces(theta=0.5)
    good = G * 5e10 / 25.0 * np.array([0.6, 0.8, 0.0])
    for wrong in (-acc[0], acc[0][[1, 0, 2]], acc[0] * 25.0 / 5.0):
        assert not np.allclose(wrong, good, rtol=1e-12, atol=0)
    np.testing.assert_allclose(acc[0], good, rtol=1e-15, atol=0)
