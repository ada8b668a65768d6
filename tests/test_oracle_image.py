"""Pins of the oracle's frame image (orc_render_image), the reference's
write_image (src/main.cpp:556-678), which needs FreeImage and does not
build here: closed forms of each projection, the scale law, the bounds
(including the y bound taken from the width), accumulation, adaptive and
nonlinear brightness, the heat map, and the PNG file of nbs_write_png read
back by an independent decoder."""
import numpy as np
import pytest

import oracle as O
from nbodysim_amd import host

W, H = 64, 48


def lit(img):
    """(x, y) of the non-black pixels in the reference's coordinates."""
    rows, cols = np.nonzero(img.any(axis=2))
    return sorted(zip(cols.tolist(), (H - 1 - rows).tolist()))


def render(pos, **kw):
    kw.setdefault("projection", "front")
    return O.render_image(np.atleast_2d(np.asarray(pos, float)), W, H, **kw)


def test_projections_place_the_axes():
    p = [5.0, -3.0, 7.0]
    assert lit(render(p, projection="front")) == [(37, 31)]  # (x + 32, z + 24)
    assert lit(render(p, projection="side")) == [(29, 31)]   # (y + 32, z + 24)
    assert lit(render(p, projection="top")) == [(37, 21)]    # (x + 32, y + 24)
    # iso: ((sqrt(3)/2 (x - y) + w) / 2, (-(x + y + 2z)/2 + h) / 2), scale 2,
    # then flipped: the origin lands on (w/2, h - 1 - h/2)
    assert lit(render([0.0, 0.0, 0.0], projection="iso")) == [(32, 23)]
    x = (np.sqrt(3) / 2 * 8 + W) / 2
    x += (x - 32) * 1
    assert lit(render([4.0, -4.0, 0.0], projection="iso")) == [(int(x), 23)]


def test_scale_and_bounds():
    assert lit(render([5.0, 0.0, 0.0], scale=2.0)) == [(42, 24)]  # 37 + (37 - 32)
    assert lit(render([-40.0, 0.0, 0.0])) == []                    # x < 0: dropped
    # y is bounded by the width (63), then clamped to the height: a particle
    # at y = 50 lands on the top row, one at y = 70 is dropped
    assert lit(render([0.0, 0.0, 26.0])) == [(32, 47)]
    assert lit(render([0.0, 0.0, 46.0])) == []
    assert lit(render([np.nan, 0.0, 0.0])) == []


def test_accumulation_and_brightness():
    for k, want in ((1, 100), (2, 200), (3, 255)):
        img = render([[1.0, 0.0, 1.0]] * k, brightness=100.0)
        assert img[H - 1 - 25, 33].tolist() == [want] * 3
    pos = [[1.0, 0.0, 1.0]] * 4 + [[3.0, 0.0, 1.0]] * 2
    img = render(pos, brightness=1.0, adaptive=True)
    assert img[H - 1 - 25, 33, 0] == 255 and img[H - 1 - 25, 35, 0] == 127
    img = render(pos, brightness=1.0, adaptive=True, nonlinear=True)
    assert img[H - 1 - 25, 35, 0] == int(np.sqrt(2.0) * 255.0 / 2.0)
    # adaptive scaling of an empty image divides 0 by 0: black
    assert not render([-40.0, 0.0, 0.0], adaptive=True).any()
    assert not render([-40.0, 0.0, 0.0], adaptive=True, color="heat").any()


def test_heat_map():
    img = render([[1.0, 0.0, 1.0]], brightness=100.0, color="heat")
    assert img[H - 1 - 25, 33].tolist() == [0, 100, 155]
    img = render([[1.0, 0.0, 1.0]], brightness=300.0, color="heat")
    assert img[H - 1 - 25, 33].tolist() == [44, 211, 0]
    img = render([[1.0, 0.0, 1.0]], brightness=600.0, color="heat")
    assert img[H - 1 - 25, 33].tolist() == [255, 0, 0]
    assert img[0, 0].tolist() == [0, 0, 255]  # an empty pixel is blue


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    img = O.render_image(rng.normal(0, 20, size=(5000, 3)), 320, 200, projection="top",
                         color="heat", brightness=40.0)
    path = str(tmp_path / "0000.png")
    host.write_png(path, img)
    np.testing.assert_array_equal(host.read_png(path), img)
    with pytest.raises(ValueError):
        data = bytearray(open(path, "rb").read())
        data[40] ^= 1
        open(path, "wb").write(bytes(data))
        host.read_png(path)
