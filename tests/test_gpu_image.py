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


@pytest.mark.parametrize("color,counts", [("bw", (5120, 3072, 4096, 1)),
                                          ("heat", (10241, 4389, 5852, 2))])
def test_dense_pixels_non_dyadic_brightness(color, counts):
    """Pixels holding thousands of particles at brightness 0.1, adaptive:
    0.1's odd mantissa (about 2^52) times such a count passes 2^64, so an
    exactness test that forms the product wraps and takes the product where
    the reference's repeated addition differs in the last bits.  With these
    counts that last bit moves a byte of another pixel (152 against 153 for
    3072 particles under a 5120-particle maximum; 218 against 219 in the heat
    map), so the bytes equal the oracle's only with the repeated addition."""
    cols = [10 + 7 * k for k in range(len(counts))]
    pos = np.concatenate([np.tile([c - 160 + 0.5, 0.0, 0.25], (m, 1))
                          for c, m in zip(cols, counts)])
    mass = np.full(len(pos), 1e10)
    kw = dict(projection="front", color=color, adaptive=True, brightness=0.1)
    with nbodysim_amd.Simulation(theta=0.5, dt=0.0333) as s:
        s.set_state(pos, None, mass)
        got = s.render_image(320, 200, **kw)
    want = O.render_image(pos, 320, 200, **kw)
    np.testing.assert_array_equal(got, want)
    assert want[200 - 1 - 100, cols[1]].any()
