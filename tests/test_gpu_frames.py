"""Frame output: asynchronous snapshots (nbs_snapshot_begin / _end) against
nbs_get_state, and the frame loop of the nbodysim command (nbs_run_frames,
src/main.cpp:1359-1504) against the same steps taken through the library.
Bar: bit-identical arrays (the same kernels produce both)."""
import os

import numpy as np
import pytest

import oracle as O
from nbodysim_amd import NbsError, host

pytestmark = pytest.mark.gpu


def sim(**kw):
    import nbodysim_amd
    return nbodysim_amd.Simulation(**kw)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_snapshot_equals_get_state(precision):
    pos, vel, mass = O.generate(5000, "sphere", seed=4)
    with sim(theta=0.5, dt=0.0333, precision=precision) as s:
        s.set_state(pos, vel, mass)
        s.snapshot_begin()
        p, v, a = s.snapshot_end()
        np.testing.assert_array_equal(p, pos)
        np.testing.assert_array_equal(v, vel)
        assert not a.any()  # no force evaluation yet
        s.step(3)
        want = s.state()
        s.snapshot_begin()
        s.step(2)  # submitted while the copy of step 3's state is pending
        got = s.snapshot_end()
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)
        s.step(4, graph=True)
        want = s.state()
        s.snapshot_begin()
        got = s.snapshot_end()
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)


def test_snapshot_buffers_alternate():
    pos, vel, mass = O.generate(3000, "cube", seed=2)
    with sim(theta=0.5, dt=0.0333, precision="f64") as s:
        s.set_state(pos, vel, mass)
        s.step(1)
        s.snapshot_begin()
        first = s.snapshot_end(copy=False)
        keep = [x.copy() for x in first]
        s.step(1)
        s.snapshot_begin()  # fills the other buffer: `first` stays intact
        second = s.snapshot_end(copy=False)
        for x, k in zip(first, keep):
            np.testing.assert_array_equal(x, k)
        assert not np.array_equal(second[0], keep[0])
        np.testing.assert_array_equal(second[0], s.state()[0])
        # one snapshot at a time; end without begin is an error
        s.snapshot_begin()
        with pytest.raises(NbsError):
            s.snapshot_begin()
        s.snapshot_end()
        with pytest.raises(NbsError):
            s.snapshot_end()


def test_snapshot_batched_systems_caller_order():
    parts = [O.generate(n, g, seed=10 + k) for k, (n, g) in
             enumerate([(700, "cube"), (1500, "sphere"), (64, "shell")])]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.cumsum([0] + [len(p[2]) for p in parts])
    with sim(theta=0.5, dt=0.0333, precision="f32") as s:
        s.set_state(pos, vel, mass)
        s.set_systems(off)
        s.step(3)
        want = s.state()
        s.snapshot_begin()
        got = s.snapshot_end()
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)


SETTINGS = """use_seed true
seed 3735928559
num_particles 3000
num_frames 4
gen_type cube
size 512
min_mass 5e10
max_mass 5e11
min_vel 0
max_vel 1
theta 0.5
dt 0.0333
dump_binary true
dump_text false
dump_image false
overwrite_data true
force_precision {prec}
keep_previous_binary {keep}
"""


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_frame_loop_writes_the_library_steps(tmp_path, monkeypatch, prec):
    monkeypatch.chdir(tmp_path)
    with open("settings.cfg", "w") as f:
        f.write(SETTINGS.format(prec=prec, keep="true"))
    print("frame loop, force_precision", prec, flush=True)
    assert host.run_frames("settings.cfg") == 0
    settings = host.read_settings("settings.cfg")[0]
    pos, vel, mass = host.generate(3000, gen_type="cube", seed=0xDEADBEEF, settings=settings)
    with sim(theta=0.5, dt=0.0333, precision=prec) as s:
        s.set_state(pos, vel, mass)
        for frame in range(4):
            s.step(1)
            want = s.state()
            got = host.read_binary(host.frame_filename("./data", frame, True))
            for g, w in zip(got[:3], want):
                np.testing.assert_array_equal(g, w)
            np.testing.assert_array_equal(got[3], mass)
    assert not [f for f in os.listdir("data") if f.endswith(".tmp")]


def test_frame_loop_keeps_only_the_newest_frame(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with open("settings.cfg", "w") as f:
        f.write(SETTINGS.format(prec="f32", keep="false"))
    assert host.run_frames("settings.cfg") == 0
    assert sorted(os.listdir("data")) == ["0003.dat"]
    # resume: the newest frame is read back and the loop continues from
    # frame 4 up to num_frames (nothing left to do)
    with open("settings.cfg", "a") as f:
        f.write("read_existing true\n")
    assert host.run_frames("settings.cfg") == 0
    assert sorted(os.listdir("data")) == ["0003.dat"]


def test_snapshot_of_an_empty_state():
    with sim(theta=0.5, dt=0.0333, precision="f32") as s:
        s.set_state(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
        s.snapshot_begin()
        p, v, a = s.snapshot_end()
        assert p.shape == v.shape == a.shape == (0, 3)
        s.step(1)  # nothing to do, not an error


def test_frame_loop_with_collisions(tmp_path, monkeypatch):
    """collide true in the settings: each frame file holds the particles left
    after that frame's merges, with their merged masses, as the library's
    collision steps give them."""
    monkeypatch.chdir(tmp_path)
    text = SETTINGS.format(prec="f64", keep="true").replace("size 512", "size 4")
    text += "collide true\ncollision_range 0.05\n"
    with open("settings.cfg", "w") as f:
        f.write(text)
    assert host.run_frames("settings.cfg") == 0
    settings = host.read_settings("settings.cfg")[0]
    pos, vel, mass = host.generate(3000, gen_type="cube", seed=0xDEADBEEF, settings=settings)
    with sim(theta=0.5, dt=0.0333, precision="f64") as s:
        s.set_state(pos, vel, mass)
        s.set_collisions(0.05)
        counts = []
        for frame in range(4):
            s.step(1)
            want = s.state()
            got = host.read_binary(host.frame_filename("./data", frame, True))
            counts.append(len(got[3]))
            for g, w in zip(got[:3], want):
                np.testing.assert_array_equal(g, w)
            np.testing.assert_array_equal(got[3], s.masses())
    print("particles per frame:", counts, flush=True)
    assert counts[0] < 3000 and counts == sorted(counts, reverse=True)


def test_frame_loop_with_min_node_size(tmp_path, monkeypatch):
    """min_node_size applies only with verbose set (src/main.cpp:1425): with
    both, each frame file holds the merged particles the library's
    minimum-size steps give; without verbose nothing is merged."""
    monkeypatch.chdir(tmp_path)
    base = SETTINGS.format(prec="f64", keep="true") + "min_node_size 20\n"
    with open("settings.cfg", "w") as f:
        f.write(base + "verbose true\n")
    assert host.run_frames("settings.cfg") == 0
    settings = host.read_settings("settings.cfg")[0]
    pos, vel, mass = host.generate(3000, gen_type="cube", seed=0xDEADBEEF, settings=settings)
    with sim(theta=0.5, dt=0.0333, precision="f64") as s:
        s.set_state(pos, vel, mass)
        s.set_min_node_size(20.0)
        for frame in range(4):
            s.step(1)
            want = s.state()
            got = host.read_binary(host.frame_filename("./data", frame, True))
            for g, w in zip(got[:3], want):
                np.testing.assert_array_equal(g, w)
            np.testing.assert_array_equal(got[3], s.masses())
        assert s.n < 3000
    with open("settings.cfg", "w") as f:
        f.write(base + "verbose false\n")
    assert host.run_frames("settings.cfg") == 0
    assert len(host.read_binary(host.frame_filename("./data", 3, True))[3]) == 3000


def test_frame_files_follow_the_oracle_trajectory(tmp_path, monkeypatch):
    """The FP64 frame files of nbs_run_frames against the oracle's frames of
    the reference loop from the same seeded particles: positions within
    1e-9 of the coordinate scale, velocities within 1e-9 of theirs (only
    the summation order of the accelerations differs)."""
    monkeypatch.chdir(tmp_path)
    with open("settings.cfg", "w") as f:
        f.write(SETTINGS.format(prec="f64", keep="true"))
    assert host.run_frames("settings.cfg") == 0
    pos, vel, mass = O.generate(3000, "cube", seed=0xDEADBEEF, min_vel=0.0, max_vel=1.0)
    for frame in range(4):
        opos, ovel, _ = O.run(pos, vel, mass, theta=0.5, dt=0.0333, frames=frame + 1)
        p, v, a, m = host.read_binary(host.frame_filename("./data", frame, True))
        assert np.abs(p - opos).max() < 1e-9 * np.abs(opos).max(), frame
        assert np.abs(v - ovel).max() < 1e-9 * np.abs(ovel).max(), frame
        np.testing.assert_array_equal(m, mass)


RESUME = """use_seed true
seed 3735928559
num_particles 3000
num_frames {frames}
gen_type sphere
min_mass 5e10
max_mass 5e11
theta 0.5
dt 0.0333
dump_binary true
dump_text false
dump_image false
overwrite_data true
keep_previous_binary true
datatype {datatype}
read_existing {resume}
"""


def test_resume_continues_an_uninterrupted_run(tmp_path, monkeypatch):
    """Frames 0-1, then a resumed run to num_frames 5 (src/main.cpp:790-811:
    the newest frame is read back and the loop goes on from the frame after
    it): frames 2-4 equal an uninterrupted 5-frame run's bit for bit (80-B
    records hold the FP64 state exactly, and caller order is kept)."""
    for d in ("a", "b"):
        (tmp_path / d).mkdir()
    monkeypatch.chdir(tmp_path / "a")
    with open("settings.cfg", "w") as f:
        f.write(RESUME.format(frames=5, datatype="double", resume="false"))
    assert host.run_frames("settings.cfg") == 0
    monkeypatch.chdir(tmp_path / "b")
    with open("settings.cfg", "w") as f:
        f.write(RESUME.format(frames=2, datatype="double", resume="false"))
    assert host.run_frames("settings.cfg") == 0
    assert sorted(os.listdir("data")) == ["0000.dat", "0001.dat"]
    with open("settings.cfg", "w") as f:
        f.write(RESUME.format(frames=5, datatype="double", resume="true"))
    assert host.run_frames("settings.cfg") == 0
    for frame in range(5):
        want = open(tmp_path / "a" / "data" / ("%04d.dat" % frame), "rb").read()
        got = open(tmp_path / "b" / "data" / ("%04d.dat" % frame), "rb").read()
        assert got == want, frame


def test_float_records_write_and_resume(tmp_path, monkeypatch):
    """datatype float (the reference's shipped -DFLOAT build, makefile:3):
    particles generated as the FLOAT build draws them, 40-byte records
    holding the library's FP64 state rounded to float, and a resume that
    reads them back: its frames equal library steps taken from the rounded
    state of the frame it resumed from."""
    monkeypatch.chdir(tmp_path)
    with open("settings.cfg", "w") as f:
        f.write(RESUME.format(frames=2, datatype="float", resume="false"))
    assert host.run_frames("settings.cfg") == 0
    settings = host.read_settings("settings.cfg")[0]
    assert settings["datatype_float"]
    pos, vel, mass = host.generate(3000, gen_type="sphere", seed=0xDEADBEEF, prec="f32",
                                   settings=settings)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    with sim(theta=0.5, dt=0.0333, precision="f64") as s:
        s.set_state(pos, vel, mass)
        for frame in range(2):
            s.step(1)
            path = host.frame_filename("./data", frame, True)
            assert os.path.getsize(path) == 40 * 3000
            got = host.read_binary(path, prec="f32")
            for g, w in zip(got[:3], s.state()):
                np.testing.assert_array_equal(g, f32(w))
            np.testing.assert_array_equal(got[3], f32(mass))
    with open("settings.cfg", "w") as f:
        f.write(RESUME.format(frames=4, datatype="float", resume="true"))
    assert host.run_frames("settings.cfg") == 0
    p1, v1, _, m1 = host.read_binary(host.frame_filename("./data", 1, True), prec="f32")
    with sim(theta=0.5, dt=0.0333, precision="f64") as s:
        s.set_state(p1, v1, m1)
        for frame in (2, 3):
            s.step(1)
            got = host.read_binary(host.frame_filename("./data", frame, True), prec="f32")
            for g, w in zip(got[:3], s.state()):
                np.testing.assert_array_equal(g, f32(w))


def test_frame_loop_threads_selects_the_threaded_tree(tmp_path, monkeypatch):
    """threads 8 (the reference's sample settings.cfg:21) selects
    gen_root_threaded's tree in the frame loop (src/main.cpp:1423): the
    frame files equal the library's steps with threaded_tree set."""
    monkeypatch.chdir(tmp_path)
    with open("settings.cfg", "w") as f:
        f.write(SETTINGS.format(prec="f64", keep="true") + "threads 8\n")
    assert host.run_frames("settings.cfg") == 0
    settings = host.read_settings("settings.cfg")[0]
    pos, vel, mass = host.generate(3000, gen_type="cube", seed=0xDEADBEEF, settings=settings)
    import nbodysim_amd
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, threaded_tree=True) as s:
        s.set_state(pos, vel, mass)
        for frame in range(4):
            s.step(1)
            got = host.read_binary(host.frame_filename("./data", frame, True))
            for g, w in zip(got[:3], s.state()):
                np.testing.assert_array_equal(g, w)
