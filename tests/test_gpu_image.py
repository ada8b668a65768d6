"""The frame image rendered on the device (nbs_render_image) against the
oracle's write_image restatement (tests/test_oracle_image.py pins it), byte
for byte, and the frame loop's ./img/NNNN.png files."""
import os

import numpy as np
import pytest

import nbodysim_amd
from nbodysim_amd import host
import oracle as O

pytestmark = pytest.mark.gpu

SPHERE = dict(r_sphere=100.0, rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0),
              min_mass=5e10, max_mass=5e11)


@pytest.fixture(scope="module")
def sphere():
    pos, vel, mass = host.generate(20000, gen_type="sphere", seed=0xDEADBEEF, **SPHERE)
    return pos, vel, mass


@pytest.mark.parametrize("projection", ["front", "side", "top", "iso"])
@pytest.mark.parametrize("color,adaptive,nonlinear,scale,brightness", [
    ("bw", False, False, 1.0, 255.0),
    ("bw", True, False, 1.5, 3.0),
    ("heat", False, False, 0.75, 37.5),
    ("heat", True, True, 1.0, 0.1),
])
def test_render_matches_oracle(sphere, projection, color, adaptive, nonlinear, scale,
                               brightness):
    pos, vel, mass = sphere
    kw = dict(projection=projection, color=color, adaptive=adaptive, nonlinear=nonlinear,
              scale=scale, brightness=brightness)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        got = s.render_image(320, 200, **kw)
        s.step(3)
        got3 = s.render_image(320, 200, **kw)
        p3 = s.state()[0]
    np.testing.assert_array_equal(got, O.render_image(pos, 320, 200, **kw))
    np.testing.assert_array_equal(got3, O.render_image(p3, 320, 200, **kw))
    assert got.any()


def test_full_hd_default_settings():
    pos, vel, mass = host.generate(200000, gen_type="sphere", seed=7, **SPHERE)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32") as s:
        s.set_state(pos, vel, mass)
        got = s.render_image()  # 1920 x 1080 iso, bw, brightness 255
        with pytest.raises(nbodysim_amd.NbsError):
            s.render_image(0, 10)
    np.testing.assert_array_equal(got, O.render_image(pos))


SETTINGS = """use_seed true
seed 3735928559
num_particles 3000
num_frames 3
gen_type sphere
r_sphere 100
rotation_magnitude 0.1
min_mass 5e10
max_mass 5e11
theta 0.5
dt 0.0333
dump_binary false
dump_text false
dump_image true
img_w 400
img_h 300
projection top
color heat
brightness 60
overwrite_data true
"""


def test_frame_loop_writes_images(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    with open("settings.cfg", "w") as f:
        f.write(SETTINGS)
    assert host.run_frames("settings.cfg") == 0
    assert sorted(os.listdir("img")) == ["0000.png", "0001.png", "0002.png"]
    settings = host.read_settings("settings.cfg")[0]
    pos, vel, mass = host.generate(3000, gen_type="sphere", seed=0xDEADBEEF, settings=settings)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, vel, mass)
        for frame in range(3):
            s.step(1)
            want = s.render_image(400, 300, projection="top", color="heat", brightness=60.0)
            np.testing.assert_array_equal(host.read_png("img/%04d.png" % frame), want)
