"""GPU parity: libnbs.so against the oracle (CPU restatement of the
reference step, itself pinned to the reference's compiled octree code).

Bars: integer/index work (keys, sort order, node membership, depths) and the
FP64 centre-of-mass reduction are bit-exact; FP64 accelerations agree to
1e-12 relative per particle (only the summation order differs: depth-first
here, breadth-first in the reference); FP32 forces within 1e-5 rms relative.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def sim(**kw):
    import nbodysim_amd
    return nbodysim_amd.Simulation(**kw)


def rel_err(a, b):
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return num / den


def cube(n, seed=0xDEADBEEF):
    return O.generate(n, "cube", seed=seed)


def assert_tree_equal(gpu, orc):
    assert len(gpu["depth"]) == len(orc["depth"])
    for k in ("depth", "leaf", "first", "count", "khi", "klo"):
        np.testing.assert_array_equal(gpu[k].astype(np.int64), orc[k].astype(np.int64), err_msg=k)
    for k in ("side", "mass", "com"):
        np.testing.assert_array_equal(gpu[k], orc[k], err_msg=k)  # bit-exact FP64
    np.testing.assert_array_equal(gpu["order"], orc["order"])


@pytest.mark.parametrize("gen,n", [("cube", 5000), ("sphere", 4096), ("shell", 4096)])
def test_keys_root_tree_bit_exact(gen, n):
    pos, vel, mass = O.generate(n, gen)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        s.forces()
        c, side = s.roots()
        oc, oside = O.root(pos)
        np.testing.assert_array_equal(c[0], oc)
        assert side[0] == oside
        hi, lo = s.keys()
        ohi, olo = O.keys(pos, oc, oside)
        np.testing.assert_array_equal(hi, ohi)
        np.testing.assert_array_equal(lo, olo)
        info = s.tree_info()
        assert info["n_buckets"] == 0
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())


@pytest.mark.parametrize("damping", [False, True])
def test_forces_f64_match_reference_walk(damping):
    pos, vel, mass = cube(5000)
    oacc, ocnt = O.Tree(pos, mass).forces(theta=0.5, damping=damping)
    with sim(theta=0.5, damping=damping) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        tot, cnt = s.interactions(per_particle=True)
    # same accepted set for every particle (bit-exact MAC) ...
    np.testing.assert_array_equal(cnt, ocnt)
    # ... so FP64 results differ only by summation order
    assert rel_err(acc, oacc).max() < 1e-12


def test_forces_f32_precision():
    pos, vel, mass = O.generate(8192, "sphere")
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    with sim(theta=0.5, precision="f32") as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
    e = rel_err(acc, oacc)
    assert np.sqrt(np.mean(e ** 2)) < 1e-5
    assert e.max() < 1e-3


def test_forces_f32_far_from_the_origin():
    """FP32 differences are formed in FP64 and then rounded, so a system far
    from the origin keeps the precision of one at the origin (in absolute
    f32 coordinates a 3e4 offset would leave ~2e-3 of rounding in every
    position, ~1e-3 relative in the near field)."""
    pos, vel, mass = O.generate(8192, "sphere")
    errs = []
    for shift in ((0.0, 0.0, 0.0), (3e4, -2e4, 1e4)):
        p = pos + np.array(shift)
        oacc, _ = O.Tree(p, mass).forces(theta=0.5)
        with sim(theta=0.5, precision="f32") as s:
            s.set_state(p, vel, mass)
            acc = s.forces()
        e = rel_err(acc, oacc)
        errs.append(np.sqrt(np.mean(e ** 2)))
        assert errs[-1] < 1e-5 and e.max() < 1e-3, shift
    assert errs[1] < 3 * errs[0], errs


@pytest.mark.parametrize("theta", [0.0, 0.3, 0.7, 1.0])
def test_theta_sweep_f64(theta):
    pos, vel, mass = O.generate(3000, "shell")
    oacc, ocnt = O.Tree(pos, mass).forces(theta=theta)
    with sim(theta=theta) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        _, cnt = s.interactions(per_particle=True)
    np.testing.assert_array_equal(cnt, ocnt)
    assert rel_err(acc, oacc).max() < 1e-12
    if theta == 0.0:  # every cell opens: the walk is the direct sum
        assert (cnt == len(pos) - 1).all()
        d = O.direct(pos, mass)
        assert rel_err(acc, d).max() < 1e-12


def test_direct_sum_matches_oracle():
    pos, vel, mass = O.generate(6000, "sphere")
    ids = np.random.default_rng(1).choice(len(pos), 300, replace=False)
    with sim() as s:
        s.set_state(pos, vel, mass)
        s.forces()  # permutes the internal order; ids stay caller indices
        acc = s.direct(ids)
    assert rel_err(acc, O.direct(pos, mass, ids)).max() < 1e-13


def test_ten_frames_f64_trajectory():
    pos, vel, mass = cube(5000)
    opos, ovel, oacc = O.run(pos, vel, mass, theta=0.5, dt=0.0333, frames=10)
    with sim(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        s.step(10)
        p, v, a = s.state()
        t = s.time()[0]
    assert abs(t - 10 * 0.0333) < 1e-12
    assert np.abs(p - opos).max() < 1e-9 * np.abs(opos).max()
    assert rel_err(a, oacc).max() < 1e-8


def test_one_frame_update_bit_exact_given_same_forces():
    # With theta = 0 both sides sum the same terms; the kick-drift update
    # (src/particle.cpp:66-74) is then exact up to the force rounding only.
    pos, vel, mass = O.generate(512, "sphere")
    with sim(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        s.step(1)
        p, v, a = s.state()
    # reproduce the update in numpy from the GPU's own forces
    t1 = a * 0.0333
    v1 = vel + t1
    t2 = v1 * 0.0333
    p1 = pos + t2
    np.testing.assert_array_equal(v, v1)
    np.testing.assert_array_equal(p, p1)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 129, 1000])
def test_small_sizes(n):
    pos, vel, mass = cube(n, seed=7)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        if n == 1:
            assert (acc == 0).all()
            return
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12


def test_lattice_symmetry():
    # 4x4x4 lattice at +-0.5, +-1.5 with masses 1e10 (the reference's
    # src/cuda_test.cpp:18-31 case): mirror symmetry fixes each acceleration
    g = np.array([-1.5, -0.5, 0.5, 1.5])
    pos = np.array([[x, y, z] for x in g for y in g for z in g])
    mass = np.full(len(pos), 1e10)
    with sim(theta=0.0) as s:
        s.set_state(pos, None, mass)
        acc = s.forces()
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
    # a_x(x, y, z) = -a_x(-x, y, z) etc.
    idx = {tuple(p): i for i, p in enumerate(pos)}
    for i, p in enumerate(pos):
        j = idx[(-p[0], p[1], p[2])]
        assert abs(acc[i, 0] + acc[j, 0]) <= 1e-12 * abs(acc[i, 0])
    assert np.abs((mass[:, None] * acc).sum(0)).max() < 1e-12 * np.abs(mass[:, None] * acc).max()


def test_batched_systems_equal_separate_runs():
    parts = [O.generate(n, g, seed=0xDEADBEEF + k)
             for k, (n, g) in enumerate([(1500, "cube"), (700, "sphere"), (2100, "shell")])]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.cumsum([0] + [len(p[2]) for p in parts])
    with sim(theta=0.5, precision="f64") as s:
        s.set_state(pos, vel, mass)
        s.set_systems(off)
        s.step(3)
        p_b, v_b, a_b = s.state()
    for k, (p0, v0, m0) in enumerate(parts):
        with sim(theta=0.5, precision="f64") as s:
            s.set_state(p0, v0, m0)
            s.step(3)
            p1, v1, a1 = s.state()
        np.testing.assert_array_equal(p_b[off[k]:off[k + 1]], p1)
        np.testing.assert_array_equal(v_b[off[k]:off[k + 1]], v1)


def test_adaptive_dt_bounds():
    pos, vel, mass = O.generate(4096, "shell")
    with sim(theta=0.5, dt=0.0333, adaptive=True, min_adaptive_dt=0.00333,
             max_pos_change=3.0, max_vel_change=3.0) as s:
        s.set_state(pos, vel, mass)
        s.step(5)
        t = s.time()[0]
    assert 5 * 0.00333 - 1e-12 <= t <= 5 * 0.0333 + 1e-12


def test_large_sphere_invariants_and_precision():
    n = 1 << 18
    pos, vel, mass = O.generate(n, "sphere")
    rng = np.random.default_rng(0xDEADBEEF)
    ids = rng.choice(n, 1024, replace=False)
    with sim(theta=0.5, precision="f64") as s:
        s.set_state(pos, vel, mass)
        a64 = s.forces()
        d = s.tree_dump()
        ref = s.direct(ids)
        s.configure(precision="f32")
        a32 = s.forces()
    # pre-order invariants: each internal node's count equals the sum of its
    # children's counts, and the root holds everything
    assert d["count"][0] == n and d["depth"][0] == 0
    assert np.isclose(d["mass"][0], mass.sum(), rtol=1e-12)
    leaves = d["leaf"] >= 0
    assert leaves.sum() == n
    np.testing.assert_array_equal(np.sort(d["leaf"][leaves]), np.arange(n))
    e_prec = np.sqrt(np.mean(rel_err(a32[ids], a64[ids]) ** 2))
    e_trunc = np.sqrt(np.mean(rel_err(a64[ids], ref) ** 2))
    assert e_prec < 1e-5
    assert e_trunc < 1e-2


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_graph_replays_equal_eager_steps(precision):
    pos, vel, mass = O.generate(20000, "sphere")
    with sim(theta=0.5, precision=precision) as s:
        s.set_state(pos, vel, mass)
        s.step(7)
        want = s.state()
    with sim(theta=0.5, precision=precision) as s:
        s.set_state(pos, vel, mass)
        # several replays with eager work and synchronisation in between
        s.step(3, graph=True)
        mid = s.state()
        s.step(2, graph=True)
        s.synchronize()
        s.step(2, graph=True)
        got = s.state()
    with sim(theta=0.5, precision=precision) as s:
        s.set_state(pos, vel, mass)
        s.step(3)
        np.testing.assert_array_equal(s.state()[0], mid[0])
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_graph_batched_adaptive():
    parts = [O.generate(n, g, seed=1 + k) for k, (n, g) in
             enumerate([(3000, "cube"), (1025, "sphere"), (64, "shell"), (4000, "sphere")])]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.cumsum([0] + [len(p[2]) for p in parts])
    res = []
    for graph in (False, True):
        with sim(theta=0.5, precision="f32", adaptive=True) as s:
            s.set_state(pos, vel, mass)
            s.set_systems(off)
            s.step(6, graph=graph)
            res.append((s.state(), s.time(len(parts))))
    for a, b in zip(res[0][0], res[1][0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(res[0][1], res[1][1])
    assert (res[0][1] > 0).all()


WALK_SHAPES = [(64, 1, 1), (128, 1, 1), (256, 1, 1), (512, 1, 1), (1024, 1, 1), (64, 8, 1),
               (128, 32, 1), (256, 16, 1), (512, 64, 1), (1024, 8, 1), (64, 1, 2), (256, 8, 2),
               (1024, 64, 2), (128, 16, 2)]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_walk_shapes_give_identical_results(precision):
    """Each particle's accepted cells and its summation order are fixed by
    the traversal order of the pre-order array, whatever set of lanes shares
    the traversal; the XCD interleave only permutes which workgroup takes
    which waves.  So no launch shape the autotuner may pick changes a single
    bit of the forces or steps.  30001 particles: the last 128-particle
    wavefront is partly empty."""
    pos, vel, mass = O.generate(30001, "sphere", seed=7)
    ref = None
    for k, shape in enumerate(WALK_SHAPES):
        with sim(theta=0.5, precision=precision) as s:
            s.set_state(pos, vel, mass)
            s.set_walk_shape(*shape)
            s.set_walk_order(k % 2 == 0)  # longest-first dispatch on / off
            acc = s.forces()
            s.step(3)
            s.step(2, graph=True)
            out = (acc, s.interactions(per_particle=True)[1]) + s.state()
        if ref is None:
            ref = out
        for a, b in zip(out, ref):
            np.testing.assert_array_equal(a, b, err_msg=str(shape))


def test_walk_shapes_batched():
    parts = [O.generate(n, g, seed=3 + k) for k, (n, g) in
             enumerate([(5000, "cube"), (77, "sphere"), (9000, "shell"), (640, "sphere")])]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.cumsum([0] + [len(p[2]) for p in parts])
    ref = None
    for shape in [(256, 1, 1), (64, 64, 1), (128, 8, 1), (1024, 16, 1), (256, 1, 2),
                  (64, 8, 2), (512, 1, 2)]:
        with sim(theta=0.6, precision="f32", adaptive=True) as s:
            s.set_state(pos, vel, mass)
            s.set_systems(off)
            s.set_walk_shape(*shape)
            s.step(3, graph=True)
            out = s.state() + (s.time(len(parts)),)
        with sim(theta=0.6, precision="f32") as s:
            s.set_state(pos, vel, mass)
            s.set_systems(off)
            s.set_walk_shape(*shape)
            s.step(2)
            out = out + s.state() + (s.time(len(parts)),)
        if ref is None:
            ref = out
        for a, b in zip(out, ref):
            np.testing.assert_array_equal(a, b, err_msg=str(shape))


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_autotune_restores_state(precision):
    pos, vel, mass = O.generate(30000, "sphere", seed=7)
    with sim(theta=0.5, precision=precision) as s:
        s.set_state(pos, vel, mass)
        s.step(4)
        want = s.state()
        want_t = s.time()
    with sim(theta=0.5, precision=precision) as s:
        s.set_state(pos, vel, mass)
        s.step(1)
        s.interactions_total(reset=True)
        s.step(1)
        before = s.state()
        t_before = s.time()
        tot = s.interactions_total(reset=False)
        best, times = s.autotune(reps=2)
        assert best in times and len(times) == 50
        assert all(t > 0 for t in times.values())
        assert times[best] == min(times.values())
        # tuning moves neither the state nor time nor the running total
        assert s.interactions_total(reset=False) == tot
        for a, b in zip(s.state(), before):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(s.time(), t_before)
        s.step(2, graph=True)
        got = s.state()
        got_t = s.time()
    for a, b in zip(got[:2], want[:2]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got_t, want_t)


@pytest.mark.parametrize("graph", [False, True])
def test_dense_cluster(graph):
    """9000 particles in a 0.01-wide cube inside a 1000-wide cloud: 21 tree
    levels below the 12 the cloud needs, the reference's tree and forces,
    graph replays equal to eager steps."""
    rng = np.random.default_rng(5)
    far = rng.uniform(-500, 500, size=(11000, 3))
    near = 3.0 + rng.uniform(0, 0.01, size=(9000, 3))
    pos = np.concatenate([far, near])
    mass = rng.uniform(1e9, 1e10, size=len(pos))
    vel = np.zeros_like(pos)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        if graph:
            s.step(2, graph=True)
            p, v, _ = s.state()
            with sim(theta=0.5) as e:
                e.set_state(pos, vel, mass)
                e.step(2)
                for a, b in zip(e.state()[:2], (p, v)):
                    np.testing.assert_array_equal(a, b)
            return
        acc = s.forces()
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12


def test_long_runs_of_equal_primary_keys():
    """15000 particles inside one 21-level cell (a 1e-5 cube in a 1000-wide
    cloud) and 40 inside another: runs of equal primary sort keys longer
    than the in-place insertion sort takes, sorted on the full key by the
    workgroup radix sort; tree and forces as the reference's."""
    rng = np.random.default_rng(6)
    far = rng.uniform(-500, 500, size=(5000, 3))
    big = -7.0 + rng.uniform(0, 1e-5, size=(15000, 3))
    small = 11.0 + rng.uniform(0, 1e-6, size=(40, 3))
    pos = np.concatenate([far, big, small])
    perm = rng.permutation(len(pos))
    pos = pos[perm]
    mass = rng.uniform(1e9, 1e10, size=len(pos))
    vel = np.zeros_like(pos)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
        s.step(2, graph=True)
        got = s.state()
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        s.step(2)
        for a, b in zip(s.state()[:2], got[:2]):
            np.testing.assert_array_equal(a, b)


def _walk_code_outputs(pos, vel, mass, generic, offsets=None, shape=None, theta=0.5,
                       precision="f32", damping=False, records="auto"):
    with sim(theta=theta, precision=precision, damping=damping) as s:
        s.set_state(pos, vel, mass)
        if offsets is not None:
            s.set_systems(offsets)
        if shape is not None:
            s.set_walk_shape(*shape)
        s.set_walk_code(generic, records)
        acc = s.forces()
        cnt = s.interactions(per_particle=True)[1]
        info = s.tree_info()
        s.step(2)
        s.step(2, graph=True)
        return (acc, cnt) + s.state(), info


@pytest.mark.parametrize("damping", [False, True])
def test_hand_written_walk_equals_generic_on_32_byte_records(damping):
    """From 2^21 particles on, the FP32 loop reads the 32-byte records by
    default: a sphere just above that size gives forces, counts and four
    steps (eager and graph-replayed) bit-identical to the generic loop,
    which reads the 64-byte records."""
    pos, vel, mass = O.generate((1 << 21) + 77, "sphere", seed=5)
    want, _ = _walk_code_outputs(pos, vel, mass, True, damping=damping)
    got, _ = _walk_code_outputs(pos, vel, mass, False, damping=damping)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["f32", "f64", "f32-damped"])
@pytest.mark.parametrize("case", ["sphere", "cluster", "buckets", "batched", "theta0"])
def test_hand_written_walk_equals_generic(case, precision):
    """The hand-written gfx950 walk loops (FP32 on 64- and on 32-byte node
    records, FP32 with damping, FP64) make the generic loop's lane
    decisions and arithmetic in the same order: forces,
    interaction counts and four steps (eager and graph-replayed) agree bit
    for bit.  The cases
    reach every branch of the loop: a partly empty last wave (30001
    particles), deep stacks of waiting lanes (a dense cluster), depth-42
    buckets (near-coincident particles), several systems, and theta = 0
    (every cell opened, so no lane ever waits)."""
    rng = np.random.default_rng(11)
    offsets, shape, theta = None, None, 0.5
    if case == "sphere":
        pos, vel, mass = O.generate(30001, "sphere", seed=7)
    elif case == "cluster":
        pos = np.concatenate([rng.uniform(-500, 500, size=(11000, 3)),
                              3.0 + rng.uniform(0, 0.01, size=(9000, 3))])
        mass = rng.uniform(1e9, 1e10, size=len(pos))
        vel = np.zeros_like(pos)
    elif case == "buckets":
        dup = 5.0 + np.arange(40)[:, None] * np.array([1e-13, 2e-13, 3e-13])
        pos = np.concatenate([rng.uniform(-500, 500, size=(3000, 3)), dup])
        pos = pos[rng.permutation(len(pos))]
        mass = rng.uniform(1e9, 1e10, size=len(pos))
        vel = np.zeros_like(pos)
    elif case == "batched":
        parts = [O.generate(n, g, seed=3 + k) for k, (n, g) in
                 enumerate([(5000, "cube"), (77, "sphere"), (9000, "shell"), (640, "sphere")])]
        pos = np.concatenate([p[0] for p in parts])
        vel = np.concatenate([p[1] for p in parts])
        mass = np.concatenate([p[2] for p in parts])
        offsets = np.cumsum([0] + [len(p[2]) for p in parts])
        shape = (128, 8, 1)
    else:
        pos, vel, mass = O.generate(3000, "sphere", seed=9)
        theta = 0.0
    damping = precision.endswith("-damped")
    precision = precision.split("-")[0]
    want, _ = _walk_code_outputs(pos, vel, mass, True, offsets, shape, theta, precision,
                                 damping)
    # the FP32 loop on both record formats (the FP64 loop reads 64 bytes)
    for records in ((64, 32) if precision == "f32" else ("auto",)):
        got, info = _walk_code_outputs(pos, vel, mass, False, offsets, shape, theta,
                                       precision, damping, records)
        if case == "buckets":
            assert info["n_buckets"] > 0
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b, err_msg=str(records))
        if case == "theta0":
            assert (got[1] == len(pos) - 1).all()


@pytest.mark.parametrize("scale", [1e-21, 1e-25, 1e21])
def test_fp64_decisions_at_extreme_scales(scale):
    """The FP64 loop's margin test compares f32(d^2) with the record's f32
    opening radius.  On a fixed root of side 512 * scale the radii are in
    f32: at 1e-21 mostly subnormal, at 1e-25 all but the top levels 0, at
    1e21 all infinite (every cell then goes to the exact test).  Every
    particle's accepted set still equals the oracle's, and the forces stay
    within the FP64 bar."""
    pos, vel, mass = cube(3000, seed=17)
    pos = pos * scale
    tree = O.Tree(pos, mass, center=np.zeros(3), side=512.0 * scale)
    oacc, ocnt = tree.forces(theta=0.5)
    with sim(theta=0.5, root_center=(0.0, 0.0, 0.0), root_size=512.0 * scale) as s:
        s.set_state(pos, None, mass)
        acc = s.forces()
        _, cnt = s.interactions(per_particle=True)
        assert_tree_equal(s.tree_dump(), tree.dump())
    np.testing.assert_array_equal(cnt, ocnt)
    assert rel_err(acc, oacc).max() < 1e-12


def test_fp64_walk_exact_opening_tie():
    """A cell whose opening test is an exact tie: the FP64 walk's margin test
    cannot decide it, so the hand-written loop hands the cell to the exact
    test (side / |d| > theta, the reference's expression).  Five particles,
    bbox [0, 30]^3 so the root side is 32: the cell [-1, 7)^3 holds two equal
    masses with COM (3, 3, 3), and the particle at (19, 3, 3) sits exactly
    side / theta = 16 away from it (8 / 16 = 0.5, not > 0.5: accepted)."""
    pos = np.array([[2.0, 3.0, 3.0], [4.0, 3.0, 3.0], [19.0, 3.0, 3.0],
                    [0.0, 30.0, 0.0], [30.0, 0.0, 30.0]])
    mass = np.full(len(pos), 1e10)
    vel = np.zeros_like(pos)
    got, _ = _walk_code_outputs(pos, vel, mass, False, None, None, 0.5, "f64")
    want, _ = _walk_code_outputs(pos, vel, mass, True, None, None, 0.5, "f64")
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    oacc, ocnt = O.Tree(pos, mass).forces(theta=0.5)
    np.testing.assert_array_equal(got[1], ocnt)
    assert rel_err(got[0], oacc).max() < 1e-12
    # the particle at (19, 3, 3) takes the pair as one cell: 3 interactions
    assert got[1][2] == 3


def test_many_long_runs_sorted_concurrently():
    """Many runs longer than the LDS bitonic sort takes (1025-4000 entries,
    the workgroup radix sort), so many workgroups sort long runs at once,
    adjacent in memory to shorter ones: 24 clusters inside 2e-6 cubes, 40
    clusters of 500-1024 (the bitonic path) and a background cloud (~80000
    particles).  Tree as the reference's, forces within 1e-12."""
    rng = np.random.default_rng(21)
    parts = [rng.uniform(-500, 500, size=(3000, 3))]
    sizes = list(rng.integers(1025, 4000, size=24)) + list(rng.integers(500, 1025, size=40))
    for m in sizes:
        c = rng.uniform(-400, 400, size=3)
        parts.append(c + rng.uniform(0, 2e-6, size=(int(m), 3)))
    pos = np.concatenate(parts)
    pos = pos[rng.permutation(len(pos))]
    mass = rng.uniform(1e9, 1e10, size=len(pos))
    vel = np.zeros_like(pos)
    print("many long runs: n =", len(pos), flush=True)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12


def test_graph_steps_as_the_tree_deepens():
    """Steps captured while the tree is shallow keep building correct trees
    when it deepens: a pair closing in on itself adds ~9 levels within the
    replayed steps.  The end state equals eager steps bit for bit."""
    rng = np.random.default_rng(44)
    cloud = rng.uniform(-500, 500, size=(20000, 3))
    pair = np.array([[0.3, 0.2, 0.1], [1.3, 0.2, 0.1]])
    pos = np.concatenate([cloud, pair])
    vel = np.zeros_like(pos)
    # 2.1407 * 0.0333 per step each: 1.0 apart at first, 0.002 at the
    # eighth build (after 7 steps), past each other after the eighth step
    vel[-2] = [2.1407, 0.0, 0.0]
    vel[-1] = [-2.1407, 0.0, 0.0]
    mass = np.full(len(pos), 1e3)
    with sim(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        s.step(2, graph=True)
        depth0 = s.tree_info()["max_depth"]
        s.step(6, graph=True)
        depth1 = s.tree_info()["max_depth"]
        got = s.state()
    print("levels: depth %d -> %d" % (depth0, depth1), flush=True)
    assert depth1 > depth0 + 5
    with sim(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        s.step(8)
        want = s.state()
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


def test_runs_at_the_sort_path_boundaries():
    """Runs of equal primary keys whose lengths straddle every sort path:
    in place (up to 8), one wavefront (9-64, 65-128, 129-256 entries: one,
    two and four entries per lane), a workgroup's LDS bitonic sort
    (257-1024) and its radix sort (1025 and more).  Tree as the
    reference's, forces within 1e-12.  (Equal full keys, where the caller
    index decides, make depth-42 buckets, which the reference's tree does
    not have: the buckets case of test_hand_written_walk_equals_generic
    sorts such a run.)"""
    rng = np.random.default_rng(33)
    parts = [rng.uniform(-500, 500, size=(3000, 3))]
    sizes = [8, 9, 10, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1024, 1025]
    for m in sizes:
        c = rng.uniform(-400, 400, size=3)
        parts.append(c + rng.uniform(0, 2e-6, size=(m, 3)))
    pos = np.concatenate(parts)
    pos = pos[rng.permutation(len(pos))]
    mass = rng.uniform(1e9, 1e10, size=len(pos))
    print("sort path boundaries: n =", len(pos), flush=True)
    with sim(theta=0.5) as s:
        s.set_state(pos, None, mass)
        acc = s.forces()
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12


def test_many_runs_of_equal_primary_keys():
    """Many moderate runs of equal primary sort keys, as a cold shell makes
    while it passes through its centre: 150 clusters of 9-400 particles,
    each inside a 2e-6 cube (below the 21-level primary key's resolution of
    a 1000-wide root), 120 pairs and triples (the in-place path), and a
    background cloud (~34000 particles, 37 levels).  Tree (order, membership, mass/COM) as the
    reference's, and graph replays equal eager steps."""
    rng = np.random.default_rng(12)
    parts = [rng.uniform(-500, 500, size=(4000, 3))]
    sizes = list(rng.integers(9, 400, size=150)) + list(rng.integers(2, 4, size=120))
    for k, m in enumerate(sizes):
        c = rng.uniform(-400, 400, size=3)
        parts.append(c + rng.uniform(0, 2e-6, size=(int(m), 3)))
    pos = np.concatenate(parts)
    pos = pos[rng.permutation(len(pos))]
    mass = rng.uniform(1e9, 1e10, size=len(pos))
    vel = np.zeros_like(pos)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        assert_tree_equal(s.tree_dump(), O.Tree(pos, mass).dump())
        s.step(2, graph=True)
        got = s.state()
    oacc, _ = O.Tree(pos, mass).forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        s.step(2)
        for a, b in zip(s.state()[:2], got[:2]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("gen", ["cube", "sphere"])
def test_fixed_root_512(gen):
    """NBS_ROOT_FIXED: the root is the configured cell (centre 0, side 512,
    the workloads' root size) instead of the bbox + 2 root; tree bit-exact
    and forces against the oracle's tree on that root."""
    pos, vel, mass = O.generate(6000, gen, seed=21)
    tree = O.Tree(pos, mass, center=np.zeros(3), side=512.0)
    oacc, ocnt = tree.forces(theta=0.5)
    with sim(theta=0.5, root_center=(0.0, 0.0, 0.0), root_size=512.0) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        _, cnt = s.interactions(per_particle=True)
        c, side = s.roots()
        assert side[0] == 512.0 and (c[0] == 0).all()
        assert_tree_equal(s.tree_dump(), tree.dump())
        s.configure(precision="f32")
        acc32 = s.forces()
    np.testing.assert_array_equal(cnt, ocnt)
    assert rel_err(acc, oacc).max() < 1e-12
    e = rel_err(acc32, oacc)
    assert np.sqrt(np.mean(e ** 2)) < 1e-5 and e.max() < 1e-3


def test_long_runs_with_skipped_digits():
    """Runs of 1100-6000 entries (the workgroup radix sort of the run-sort
    kernel) inside 1e-6 cubes of a 1000-wide cloud: every key_hi byte, the
    upper three key_lo bytes and the upper two caller-index bytes are equal
    over a run, so 13 of the twenty digit passes find a single digit value
    and are skipped.  Each wave must read the skip decision before any wave
    starts the next pass, which clears it.  Tree bit-exact, forces within
    1e-12 of the oracle."""
    rng = np.random.default_rng(33)
    parts = [rng.uniform(-500, 500, size=(2000, 3))]
    for m in [1100, 2048, 3000, 4500, 6000, 1500, 2500, 1025]:
        c = rng.uniform(-400, 400, size=3)
        parts.append(c + rng.uniform(0, 1e-6, size=(m, 3)))
    pos = np.concatenate(parts)
    pos = pos[rng.permutation(len(pos))]
    mass = rng.uniform(1e9, 1e10, size=len(pos))
    vel = np.zeros_like(pos)
    print("long runs with skipped digits: n =", len(pos), flush=True)
    tree = O.Tree(pos, mass)
    with sim(theta=0.5) as s:
        s.set_state(pos, vel, mass)
        acc = s.forces()
        assert s.tree_info()["n_buckets"] == 0
        assert_tree_equal(s.tree_dump(), tree.dump())
        s.step(2)
    oacc, _ = tree.forces(theta=0.5)
    assert rel_err(acc, oacc).max() < 1e-12
