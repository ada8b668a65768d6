"""Walks split into pieces over equal ranges of the node array
(nbs_set_walk_split): every particle's accepted cells are the whole
traversal's, so interaction counts are identical for any number of
pieces; forces differ only by the order in which the pieces' partial sums
are added (FP64: 1e-13 relative; FP32: the FP32 sums' rounding, 2e-6 rms).
The hand-written loops equal the generic loop bit for bit under the same
split, including the 32-byte records above 2^21 particles; the library
picks the pieces by wavefront count; graph replays equal eager steps."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def sim(**kw):
    import nbodysim_amd
    return nbodysim_amd.Simulation(**kw)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def forces(pos, vel, mass, pieces, generic=False, **kw):
    with sim(**kw) as s:
        s.set_walk_split(pieces)
        s.set_walk_code(generic)
        s.set_state(pos, vel, mass)
        acc = s.forces()
        cnt = s.interactions(per_particle=True)[1]
        used = s.walk_split
    return acc, cnt, used


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("damping", [False, True])
def test_pieces_keep_counts_and_forces(precision, damping):
    pos, vel, mass = O.generate(20000, "sphere", seed=31)
    kw = dict(theta=0.5, precision=precision, damping=damping)
    ref, rcnt, used = forces(pos, vel, mass, 1, **kw)
    assert used == 1
    for k in (2, 8, 32):
        print("pieces", k, precision, damping, flush=True)
        acc, cnt, used = forces(pos, vel, mass, k, **kw)
        assert used == k
        np.testing.assert_array_equal(cnt, rcnt)
        e = rel_err(acc, ref)
        if precision == "f64":
            assert e.max() < 1e-13, (k, e.max())
        else:
            assert np.sqrt(np.mean(e ** 2)) < 2e-6 and e.max() < 1e-4, (k, e.max())


def test_pieces_match_the_oracle():
    pos, vel, mass = O.generate(5000, "cube", seed=7)
    oacc, ocnt = O.Tree(pos, mass).forces(theta=0.5)
    acc, cnt, used = forces(pos, vel, mass, 0, theta=0.5)
    assert used == 8  # 79 wavefronts: eight pieces each
    np.testing.assert_array_equal(cnt, ocnt)
    assert rel_err(acc, oacc).max() < 1e-12


def test_library_picks_pieces_by_wavefront_count():
    for n, want in ((5000, 8), (131072, 4), (262144, 2), (524288, 1)):
        pos, vel, mass = O.generate(n, "sphere", seed=3)
        with sim(theta=0.5, precision="f32") as s:
            s.set_state(pos, vel, mass)
            assert s.walk_split == want, n


@pytest.mark.parametrize("n", [30001, (1 << 21) + 77])
def test_hand_written_equals_generic_in_pieces(n):
    pos, vel, mass = O.generate(n, "sphere", seed=9)
    for precision in ("f32", "f64"):
        kw = dict(theta=0.5, precision=precision)
        a = forces(pos, vel, mass, 4, generic=True, **kw)
        b = forces(pos, vel, mass, 4, generic=False, **kw)
        np.testing.assert_array_equal(a[0], b[0], err_msg=precision)
        np.testing.assert_array_equal(a[1], b[1], err_msg=precision)


def test_graph_steps_in_pieces_equal_eager_steps():
    pos, vel, mass = O.generate(12000, "sphere", seed=5)
    out = []
    for graph in (False, True):
        with sim(theta=0.5, dt=0.0333, precision="f32") as s:
            s.set_counting("total")
            s.set_state(pos, vel, mass)
            assert s.walk_split == 8
            s.step(6, graph=graph)
            out.append(s.state() + (s.interactions_total(),))
    for k in range(3):
        np.testing.assert_array_equal(out[0][k], out[1][k])
    assert out[0][3] == out[1][3] > 0


def test_batched_systems_in_pieces():
    parts = [O.generate(n, g, seed=40 + k) for k, (n, g) in
             enumerate([(3000, "cube"), (50, "sphere"), (7000, "shell"), (1, "cube")])]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.cumsum([0] + [len(p[2]) for p in parts])
    res = []
    for pieces in (1, 8):
        with sim(theta=0.5, dt=0.0333) as s:
            s.set_walk_split(pieces)
            s.set_state(pos, vel, mass)
            s.set_systems(off)
            s.step(2)
            p, v, a = s.state()
            res.append((p, a, s.interactions(per_particle=True)[1]))
    np.testing.assert_array_equal(res[0][2], res[1][2])
    assert np.abs(res[0][0] - res[1][0]).max() < 1e-12 * np.abs(res[0][0]).max()


def test_walk_split_refuses_other_counts():
    import nbodysim_amd
    with sim(theta=0.5) as s:
        for bad in (3, 64, -1):
            with pytest.raises(nbodysim_amd.NbsError, match="walk pieces"):
                s.set_walk_split(bad)
