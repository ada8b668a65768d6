"""Python view of the host contract in libnbs.so (include/nbs_host.h):
settings.cfg reader, seeded particle generator, .dat/.txt frame files."""
import ctypes

import numpy as np

from . import _native
from ._native import Settings, lib

GEN = {"cube": 0, "sphere": 1, "shell": 2}
DIST = {"linear": 0, "exp": 1, "normal": 2}
PROJ = {"front": 0, "side": 1, "top": 2, "iso": 3}
COLOR = {"bw": 0, "heat": 1}
_ENUMS = {"gen_type": GEN, "mass_dist": DIST, "vel_dist": DIST, "projection": PROJ,
          "color": COLOR}


def _dp(a):
    return a.ctypes.data_as(_native.c_dp) if a is not None else None


def _to_dict(s):
    d = {}
    for name, _ in Settings._fields_:
        v = getattr(s, name)
        if name in _ENUMS:
            v = {k: kk for kk, k in _ENUMS[name].items()}[v]
        elif name == "rotation_vector":
            v = tuple(v)
        elif name == "force_precision":
            v = "f32" if v == 1 else "f64"
        d[name] = v
    for name in ("read_existing", "use_seed", "display_progress", "dump_binary",
                 "dump_text", "dump_image", "overwrite_data", "keep_previous_binary",
                 "keep_previous_text", "verbose", "damping", "adaptive_brightness",
                 "nonlinear_brightness", "collide", "cuda", "adaptive", "datatype_float"):
        d[name] = bool(d[name])
    return d


def _from_dict(d):
    s = Settings()
    lib().nbs_settings_default(ctypes.byref(s))
    for k, v in d.items():
        if k in _ENUMS:
            v = _ENUMS[k][v]
        elif k == "rotation_vector":
            arr = (ctypes.c_double * 3)(*v)
            setattr(s, k, arr)
            continue
        elif k == "force_precision":
            v = 1 if v == "f32" else 0
        setattr(s, k, v)
    return s


def default_settings():
    s = Settings()
    lib().nbs_settings_default(ctypes.byref(s))
    return _to_dict(s)


def read_settings(path):
    """Returns (settings dict, messages) like src/main.cpp:1123-1310."""
    s = Settings()
    lib().nbs_settings_default(ctypes.byref(s))
    buf = ctypes.create_string_buffer(1 << 16)
    rc = lib().nbs_settings_read(ctypes.byref(s), path.encode(), buf, len(buf))
    msg = buf.value.decode()
    if rc == _native.NBS_ERR_IO:
        raise FileNotFoundError(msg.strip())
    if rc != 0:
        raise ValueError(msg.strip())
    return _to_dict(s), msg


def generate(n, gen_type="cube", seed=0xDEADBEEF, prec="f64", settings=None, scale=None,
             **kw):
    """Seeded particles of src/main.cpp:184-275 (float64 arrays)."""
    d = dict(settings or default_settings())
    d["gen_type"] = gen_type
    if scale is not None:
        d["scale_x"], d["scale_y"], d["scale_z"] = scale
    d.update(kw)
    s = _from_dict(d)
    pos, vel, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    rc = lib().nbs_generate(ctypes.byref(s), seed, n, 1 if prec == "f32" else 0,
                            _dp(pos), _dp(vel), _dp(mass))
    if rc:
        raise ValueError("nbs_generate failed (%d)" % rc)
    return pos, vel, mass


def frame_filename(directory, frame, binary):
    buf = ctypes.create_string_buffer(4096)
    rc = lib().nbs_frame_filename(directory.encode(), frame, int(binary), buf, len(buf))
    if rc:
        raise ValueError("name too long")
    return buf.value.decode()


def write_binary(path, pos, vel, acc, mass, prec="f64"):
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    rc = lib().nbs_write_binary(path.encode(), len(mass), _dp(f(pos)), _dp(f(vel)),
                                _dp(f(acc)), _dp(mass), 1 if prec == "f32" else 0)
    if rc:
        raise OSError("cannot write %s" % path)


def write_png(path, rgb):
    """8-bit RGB PNG from [height][width][3] uint8, top row first."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w = rgb.shape[:2]
    if lib().nbs_write_png(path.encode(), w, h,
                           rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))):
        raise OSError("cannot write %s" % path)


def read_png(path):
    """Pixels of an 8-bit RGB, non-interlaced PNG ([h][w][3] uint8)."""
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        ln, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + ln]
        crc = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])[0]
        if zlib.crc32(data[pos + 4:pos + 8 + ln]) != crc:
            raise ValueError("bad CRC in chunk %r" % kind)
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2), "only 8-bit RGB"
        elif kind == b"IDAT":
            idat += body
        pos += 12 + ln
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 3 * w + 1)
    assert (raw[:, 0] == 0).all(), "only filter type 0"
    return raw[:, 1:].reshape(h, w, 3).copy()


def write_text(path, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    if lib().nbs_write_text(path.encode(), len(pos), _dp(pos)):
        raise OSError("cannot write %s" % path)


def binary_count(path, prec="f64"):
    return lib().nbs_binary_count(path.encode(), 1 if prec == "f32" else 0)


def read_binary(path, prec="f64"):
    n = binary_count(path, prec)
    if n < 0:
        raise FileNotFoundError(path)
    pos, vel, acc, mass = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    got = lib().nbs_read_binary(path.encode(), 1 if prec == "f32" else 0, n, _dp(pos),
                                _dp(vel), _dp(acc), _dp(mass))
    if got != n:
        raise OSError("short read of %s" % path)
    return pos, vel, acc, mass


def resume_frame(directory, num_frames):
    return lib().nbs_resume_frame(directory.encode(), num_frames)


def run_frames(cfgfile="settings.cfg"):
    """The nbodysim frame loop in the current directory (nbs_run_frames);
    returns its status (0 = success)."""
    return lib().nbs_run_frames(cfgfile.encode())
