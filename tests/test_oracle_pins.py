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


# ---------------------------------------------------------- sphere / shell --
# generate_particle (src/main.cpp:224-266) draws, per particle: the mass;
# for a sphere the radius fraction u_r (radius = cbrt(u_r) R); the angle
# theta = U(0, 6.28318530718); the polar cosine u_c = U(-1, 1) (azimuth =
# acos(u_c)); then point = (r sin(az) cos(theta), r sin(az) sin(theta),
# r cos(az)) and velocity = cross(omega, point) with omega the normalised
# rotation vector times its magnitude (src/vector.cpp:162-165).  The checks
# below restate each draw from the raw engine words with the independent
# generate_canonical above and test the geometry each draw fixes, so that
# swapping the theta and u_c draws, a sin/cos mix-up, a wrong cross product
# or a wrong draw count fails.

def _canonical_float(g):
    c = np.float32(g - 1) / np.float32(2147483646.0)
    return c if c < 1 else np.nextafter(np.float32(1), np.float32(0))


def _draws(gen, n, prec, seed):
    """(mass, u_r, theta, u_c) restated from the engine words per particle,
    and the number of words a particle consumes."""
    per_value = 2 if prec == "f64" else 1
    values = 4 if gen == "sphere" else 3
    raw = O.minstd_raw(per_value * values * n, seed=seed).astype(np.int64)
    out = []
    for k in range(n):
        w = raw[per_value * values * k:per_value * values * (k + 1)]
        if prec == "f64":
            c = [_canonical_double(w[2 * j], w[2 * j + 1]) for j in range(values)]
        else:
            c = [_canonical_float(w[j]) for j in range(values)]
        if gen == "shell":
            c = [c[0], None, c[1], c[2]]
        out.append(c)
    return out, per_value * values


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("gen", ["sphere", "shell"])
def test_sphere_shell_draw_accounting(gen, prec):
    # n particles use exactly 8n (sphere) / 6n (shell) engine words in FP64
    # (2 per value) and 4n / 3n in FP32: particle n's mass is the value
    # drawn from the words right after them
    n = 37
    pos, vel, mass = O.generate(n + 1, gen, prec=prec, seed=12345)
    words = {"sphere": 4, "shell": 3}[gen] * (2 if prec == "f64" else 1)
    raw = O.minstd_raw(words * (n + 1), seed=12345).astype(np.int64)
    if prec == "f64":
        c = _canonical_double(raw[words * n], raw[words * n + 1])
        assert mass[n] == c * (5e11 - 5e10) + 5e10
    else:
        c = _canonical_float(raw[words * n])
        f = np.float32
        assert mass[n] == f(c * (f(5e11) - f(5e10)) + f(5e10))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("gen", ["sphere", "shell"])
def test_sphere_shell_geometry_follows_each_draw(gen, prec):
    n, R = 400, 100.0
    pos, vel, mass = O.generate(n, gen, prec=prec, seed=0xDEADBEEF, r_sphere=R)
    draws, _ = _draws(gen, n, prec, 0xDEADBEEF)
    tol = 1e-12 if prec == "f64" else 2e-6
    two_pi = 6.28318530718
    checked_theta = 0
    for k, (m, u_r, u_t, u_c) in enumerate(draws):
        p = pos[k]
        r = np.linalg.norm(p)
        if prec == "f64":
            assert mass[k] == m * (5e11 - 5e10) + 5e10
        want_r = R * np.cbrt(float(u_r)) if gen == "sphere" else R
        assert abs(r - want_r) <= tol * R
        # polar cosine: z / |p| is the fourth (sphere) / third (shell) draw
        assert abs(p[2] / r - (float(u_c) * 2.0 - 1.0)) <= tol * 10
        # theta: the angle of (x, y), where the polar sine leaves it defined
        sin_az = np.hypot(p[0], p[1]) / r
        if sin_az > 1e-3:
            theta = np.arctan2(p[1], p[0]) % (2 * np.pi)
            want = float(u_t) * two_pi
            d = abs(theta - want)
            assert min(d, 2 * np.pi - d) <= tol * 100 / sin_az
            checked_theta += 1
    assert checked_theta > n * 0.9


@pytest.mark.parametrize("gen", ["sphere", "shell"])
@pytest.mark.parametrize("rot", [(0.0, 0.0, 1.0), (1.0, 2.0, 3.0)])
def test_rotation_velocity_is_omega_cross_point(gen, rot):
    # velocity = cross(omega, point), omega = rot / |rot| * 0.1: so v is
    # perpendicular to p and to omega, |v| = |omega| |p_perp|, and for a
    # rotation about +z the motion is counter-clockwise (v = (-w y, w x, 0))
    pos, vel, mass = O.generate(300, gen, seed=7, rotation_vector=rot, rotation_magnitude=0.1)
    w = np.array(rot) / np.linalg.norm(rot) * 0.1
    want = np.cross(w, pos)
    np.testing.assert_allclose(vel, want, rtol=0, atol=1e-13 * np.abs(pos).max())
    pn = np.linalg.norm(pos, axis=1)
    vn = np.linalg.norm(vel, axis=1)
    assert np.abs(np.einsum("ij,ij->i", vel, pos)).max() <= 1e-12 * (vn * pn).max()
    assert np.abs(vel @ w).max() <= 1e-12 * vn.max()
    perp = pos - np.outer(pos @ (w / 0.1), w / 0.1)
    np.testing.assert_allclose(vn, 0.1 * np.linalg.norm(perp, axis=1), rtol=1e-12)
    # the opposite product, cross(point, omega), is caught
    assert not np.allclose(vel, np.cross(pos, w), rtol=0, atol=1e-6)


def test_planted_seeding_mistakes_are_caught():
    """The geometry checks above fail for the plausible mistakes: theta and
    the polar cosine drawn in the other order, sin and cos exchanged."""
    n, R = 50, 100.0
    draws, _ = _draws("sphere", n, "f64", 0xDEADBEEF)
    pos, _, _ = O.generate(n, "sphere", seed=0xDEADBEEF, r_sphere=R)
    def point(u_r, t, c, swap=False, sincos=False):
        if swap:
            t, c = c, t
        th, az = float(t) * 6.28318530718, np.arccos(float(c) * 2 - 1)
        r = R * np.cbrt(float(u_r))
        s, co = (np.cos, np.sin) if sincos else (np.sin, np.cos)
        return np.array([r * s(az) * co(th), r * s(az) * s(th), r * co(az)])
    good = np.array([point(d[1], d[2], d[3]) for d in draws])
    np.testing.assert_allclose(pos, good, rtol=0, atol=1e-12 * R)
    for kw in (dict(swap=True), dict(sincos=True)):
        bad = np.array([point(d[1], d[2], d[3], **kw) for d in draws])
        assert np.abs(bad - pos).max() > 1.0


# ------------------------------------------------------------ adaptive dt --
@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
def test_magnitudes_are_the_reference_hook():
    # |a| and |v| as particle::update(dv, dp) returns them (the reference's
    # compiled vector::magnitude), bit for bit
    rng = np.random.default_rng(3)
    v = rng.normal(size=(2000, 3)) * 10.0 ** rng.integers(-12, 12, size=(2000, 1))
    a = rng.normal(size=(2000, 3)) * 10.0 ** rng.integers(-12, 12, size=(2000, 1))
    dv, dp = O.ref_update_mags(v, a)
    np.testing.assert_array_equal(O.magnitudes(a), dv)
    np.testing.assert_array_equal(O.magnitudes(v), dp)


def test_adaptive_dt_rule_closed_forms():
    z = np.zeros((3, 3))
    acc = np.array([[0.0, 0.0, 0.5], [2.0, 0.0, 0.0], [0.0, -1.0, 0.0]])  # max |a| 2
    vel = np.array([[0.0, 6.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -3.0]])  # max |v| 6
    # at rest with no forces: the configured dt
    assert O.adaptive_dt(z, z, 0.0333, 0.00333, 3.0, 3.0) == 0.0333
    # the velocity limit: 1 / 2 against the position limit 3 / 6
    assert O.adaptive_dt(acc, vel, 1.0, 0.0, 3.0, 1.0) == 0.5
    assert O.adaptive_dt(acc, vel, 1.0, 0.0, 1.5, 3.0) == 0.25
    assert O.adaptive_dt(acc, z, 1.0, 0.0, 1.2, 0.5) == 0.25
    # bounded above by dt, below by min_adaptive_dt
    assert O.adaptive_dt(acc, vel, 0.1, 0.0, 3.0, 3.0) == 0.1
    assert O.adaptive_dt(acc * 1e6, vel, 0.1, 0.00333, 3.0, 3.0) == 0.00333
    # exchanging the two limits, or taking the sum of the maxima instead of
    # the largest magnitude, gives another step: the rule is pinned by these
    assert O.adaptive_dt(acc, vel, 1.0, 0.0, 1.0, 3.0) != O.adaptive_dt(acc, vel, 1.0, 0.0,
                                                                         3.0, 1.0)
    assert O.adaptive_dt(acc, vel, 1.0, 0.0, 3.0, 1.0) != 1.0 / np.abs(acc).sum()
