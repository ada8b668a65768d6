"""ctypes front-end of the oracle (CPU restatement of nbodysim's step).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.  Each wrapper
names the reference code it restates (see nbs_oracle.cpp for file:line).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "_build", "libnbs_oracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

GEN_TYPES = {"cube": 0, "sphere": 1, "shell": 2}
DISTS = {"linear": 0, "exp": 1, "normal": 2}

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


class GenSettings(ctypes.Structure):
    _fields_ = [("gen_type", ctypes.c_int), ("mass_dist", ctypes.c_int),
                ("vel_dist", ctypes.c_int), ("pad", ctypes.c_int),
                ("size", ctypes.c_double), ("min_mass", ctypes.c_double),
                ("max_mass", ctypes.c_double), ("min_vel", ctypes.c_double),
                ("max_vel", ctypes.c_double), ("r_sphere", ctypes.c_double),
                ("rotation_magnitude", ctypes.c_double),
                ("rot_x", ctypes.c_double), ("rot_y", ctypes.c_double),
                ("rot_z", ctypes.c_double), ("scale_x", ctypes.c_double),
                ("scale_y", ctypes.c_double), ("scale_z", ctypes.c_double)]


def _p(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


_lib = None


def lib():
    """Loads the oracle library, building it first if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["sh", os.path.join(_HERE, "build_oracle.sh")])
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_generate.restype = ctypes.c_int
        L.orc_generate.argtypes = [ctypes.c_int, ctypes.POINTER(GenSettings),
                                   ctypes.c_ulong, ctypes.c_size_t, _dp, _dp, _dp]
        L.orc_minstd_raw.argtypes = [ctypes.c_ulong, ctypes.c_int, ctypes.c_size_t, _u64p]
        L.orc_uniform_raw.argtypes = [ctypes.c_int, ctypes.c_ulong, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_size_t, _dp]
        L.orc_root.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp, _dp]
        L.orc_keys.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp, ctypes.c_double,
                               _u64p, _u64p]
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_tree_build.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp, ctypes.c_int,
                                     _dp, ctypes.c_double]
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_size.restype = ctypes.c_size_t
        L.orc_tree_size.argtypes = [ctypes.c_void_p]
        L.orc_tree_max_depth.restype = ctypes.c_int
        L.orc_tree_max_depth.argtypes = [ctypes.c_void_p]
        L.orc_tree_dump.argtypes = [ctypes.c_void_p, _ip, _u64p, _u64p, _ip, _i64p, _i64p,
                                    _dp, _dp, _dp, _dp, _ip]
        L.orc_tree_forces.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, _dp, _i64p]
        L.orc_tree_forces_subset.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_size_t, _i64p, _dp, _i64p]
        L.orc_run.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp, _dp, _dp,
                              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
        L.orc_tree_leaves.restype = ctypes.c_size_t
        L.orc_tree_leaves.argtypes = [ctypes.c_void_p]
        L.orc_run_collide.restype = ctypes.c_int64
        L.orc_run_collide.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, _i64p, _i64p,
                                      ctypes.c_double]
        L.orc_direct.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_int, ctypes.c_size_t,
                                 _i64p, ctypes.c_int, _dp]
        _lib = L
    return _lib


def prec_code(prec):
    return {"f64": 0, "double": 0, "f32": 1, "float": 1}[prec]


def generate(n, gen_type="cube", seed=0xDEADBEEF, prec="f64", size=512.0,
             min_mass=5e10, max_mass=5e11, min_vel=0.0, max_vel=0.0, r_sphere=100.0,
             rotation_magnitude=0.1, rotation_vector=(0.0, 0.0, 1.0),
             scale=(1.0, 1.0, 1.0), mass_dist="linear", vel_dist="linear"):
    """IC generation of src/main.cpp:184-275 (engine seeded with `seed`).

    Returns (pos[n,3], vel[n,3], mass[n]) as float64 (float values widened
    exactly when prec is f32).
    """
    s = GenSettings(GEN_TYPES[gen_type], DISTS[mass_dist], DISTS[vel_dist], 0,
                    size, min_mass, max_mass, min_vel, max_vel, r_sphere,
                    rotation_magnitude, *rotation_vector, *scale)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    mass = np.zeros(n)
    rc = lib().orc_generate(prec_code(prec), ctypes.byref(s), seed, n, _p(pos), _p(vel),
                            _p(mass))
    if rc != 0:
        raise ValueError("orc_generate failed: %d" % rc)
    return pos, vel, mass


def minstd_raw(count, seed=None):
    out = np.zeros(count, dtype=np.uint64)
    lib().orc_minstd_raw(seed or 0, 1 if seed is None else 0, count, _p(out, _u64p))
    return out


def uniform_raw(count, a, b, seed, prec="f64"):
    out = np.zeros(count)
    lib().orc_uniform_raw(prec_code(prec), seed, a, b, count, _p(out))
    return out


def root(pos, prec="f64"):
    """gen_root's centre and side (src/main.cpp:887-921)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    c = np.zeros(3)
    s = np.zeros(1)
    lib().orc_root(prec_code(prec), len(pos), _p(pos), _p(c), _p(s))
    return c, float(s[0])


def keys(pos, center, side, prec="f64"):
    """42-level child-index paths by the reference's centre chain."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    c = np.ascontiguousarray(center, dtype=np.float64)
    hi = np.zeros(len(pos), dtype=np.uint64)
    lo = np.zeros(len(pos), dtype=np.uint64)
    lib().orc_keys(prec_code(prec), len(pos), _p(pos), _p(c), side, _p(hi, _u64p),
                   _p(lo, _u64p))
    return hi, lo


class Tree:
    """The reference's one-particle-per-leaf octree with mass and COM.

    threaded=True: the tree gen_root_threaded builds when threads > 1
    (src/main.cpp:817-886): eight eager root children filled by the
    inclusive inside() test (src/octree.cpp:320-336), so a particle on the
    root's centre planes has a leaf in every child holding it."""

    def __init__(self, pos, mass, prec="f64", center=None, side=None, threaded=False):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        mode = 0 if center is None else 1
        if threaded:
            if center is not None:
                raise ValueError("the threaded tree has the bbox + 2 root")
            mode = 2
        c = np.ascontiguousarray(center if center is not None else np.zeros(3),
                                 dtype=np.float64)
        h = lib().orc_tree_build(prec_code(prec), len(self.pos), _p(self.pos),
                                 _p(self.mass), mode, _p(c), side or 0.0)
        if not h:
            raise ValueError("tree overflow: coincident particles")
        self.h = ctypes.c_void_p(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def max_depth(self):
        return lib().orc_tree_max_depth(self.h)

    def dump(self):
        """Pre-order node list (children 0..7) as a dict of arrays."""
        m = lib().orc_tree_size(self.h)
        n = lib().orc_tree_leaves(self.h)  # leaves in DFS order
        d = dict(depth=np.zeros(m, np.int32), khi=np.zeros(m, np.uint64),
                 klo=np.zeros(m, np.uint64), leaf=np.zeros(m, np.int32),
                 first=np.zeros(m, np.int64), count=np.zeros(m, np.int64),
                 side=np.zeros(m), center=np.zeros((m, 3)), mass=np.zeros(m),
                 com=np.zeros((m, 3)), order=np.zeros(n, np.int32))
        lib().orc_tree_dump(self.h, _p(d["depth"], _ip), _p(d["khi"], _u64p),
                            _p(d["klo"], _u64p), _p(d["leaf"], _ip), _p(d["first"], _i64p),
                            _p(d["count"], _i64p), _p(d["side"]), _p(d["center"]),
                            _p(d["mass"]), _p(d["com"]), _p(d["order"], _ip))
        return d

    def forces(self, theta=0.5, damping=False, threads=1):
        n = len(self.pos)
        acc = np.zeros((n, 3))
        cnt = np.zeros(n, np.int64)
        lib().orc_tree_forces(self.h, theta, int(damping), threads, _p(acc), _p(cnt, _i64p))
        return acc, cnt

    def forces_subset(self, idx, theta=0.5, damping=False, threads=1):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        acc = np.zeros((len(idx), 3))
        cnt = np.zeros(len(idx), np.int64)
        lib().orc_tree_forces_subset(self.h, theta, int(damping), threads, len(idx),
                                     _p(idx, _i64p), _p(acc), _p(cnt, _i64p))
        return acc, cnt


def run(pos, vel, mass, theta=0.5, dt=0.0333, frames=1, damping=False, prec="f64",
        threads=1, threaded=False):
    """`frames` frames of the reference loop; returns (pos, vel, acc).
    threaded=True builds gen_root_threaded's tree (threads > 1)."""
    pos = np.array(pos, dtype=np.float64, order="C")
    vel = np.array(vel, dtype=np.float64, order="C")
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    acc = np.zeros_like(pos)
    rc = lib().orc_run(prec_code(prec), len(pos), _p(pos), _p(vel), _p(acc), _p(mass),
                       theta, dt, int(damping), frames, threads, int(threaded))
    if rc:
        raise ValueError("orc_run failed (coincident particles)")
    return pos, vel, acc


def run_collide(pos, vel, mass, theta=0.5, dt=0.0333, frames=1, damping=False, range=0.01,
                threads=1, min_size=0.0):
    """FP64 frames with collision merging (range < 0: none) and, for
    min_size > 0, apply_minimum_size; returns (pos, vel, acc, mass, pairs
    detected, particles absorbed), compacted to the survivors."""
    pos = np.array(pos, dtype=np.float64, order="C")
    vel = np.array(vel, dtype=np.float64, order="C")
    mass = np.array(mass, dtype=np.float64, order="C")
    acc = np.zeros_like(pos)
    npairs, nrem = ctypes.c_int64(), ctypes.c_int64()
    n = lib().orc_run_collide(len(pos), _p(pos), _p(vel), _p(acc), _p(mass), theta, dt,
                              int(damping), range, frames, threads, ctypes.byref(npairs),
                              ctypes.byref(nrem), float(min_size))
    if n < 0:
        raise ValueError("orc_run_collide failed (coincident particles)")
    return pos[:n], vel[:n], acc[:n], mass[:n], npairs.value, nrem.value


def direct(pos, mass, idx=None, damping=False, threads=8):
    """FP64 direct sum (long-double accumulation) on the particles in idx."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    if idx is None:
        idx = np.arange(len(pos))
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    acc = np.zeros((len(idx), 3))
    lib().orc_direct(len(pos), _p(pos), _p(mass), int(damping), len(idx), _p(idx, _i64p),
                     threads, _p(acc))
    return acc


def magnitudes(v):
    """|v| per row as vector::magnitude forms it (src/vector.cpp:104-107):
    sqrt of (x*x + y*y) + z*z, every operation rounded (numpy does not
    contract into FMAs)."""
    v = np.asarray(v, dtype=np.float64)
    return np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])


def adaptive_dt(acc, vel, dt, min_dt, max_pos_change, max_vel_change):
    """Step length of the adaptive-dt rule for one system.

    The reference documents the keys (README.md:54-60: min_adaptive_dt,
    max_pos_change, max_vel_change) but its loop never uses them; its only
    hook is particle::update(dv, dp) (src/particle.cpp:76-80), which returns
    |a| and |v|.  The rule, as this engine defines it from those: with
    A = max |a| (forces of the step) and V = max |v| (before the kick),
        h = dt;  h = min(h, max_vel_change / A) if A > 0;
        h = min(h, max_pos_change / V) if V > 0;  h = max(h, min_dt).
    """
    a = magnitudes(acc).max(initial=0.0)
    v = magnitudes(vel).max(initial=0.0)
    h = dt
    if a > 0:
        h = min(h, max_vel_change / a)
    if v > 0:
        h = min(h, max_pos_change / v)
    return max(h, min_dt)


# ----------------------------------------------------------------- _ref --
_ref = {}


def ref_available(prec="f64"):
    return os.path.exists(os.path.join(REF_DIR, "libnbsref_%s.so" % prec))


def ref_lib(prec="f64"):
    """The reference's own octree/particle/vector code + harness (oracle/_ref)."""
    if prec not in _ref:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libnbsref_%s.so" % prec))
        L.ref_tree_dump.restype = ctypes.c_long
        L.ref_tree_dump.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_long, _ip, _ip, _lp,
                                    _lp, _dp, _dp, _dp, _ip]
        L.ref_forces.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                 _dp, _lp]
        L.ref_run.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp, ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, ctypes.c_int]
        _ref[prec] = L
    return _ref[prec]


def ref_update_mags(vel, acc, prec="f64"):
    """(|a|, |v|) per particle from the reference's particle::update(dv, dp)."""
    L = ref_lib(prec)
    L.ref_update_mags.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp]
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    acc = np.ascontiguousarray(acc, dtype=np.float64)
    dv = np.zeros(len(vel))
    dp = np.zeros(len(vel))
    L.ref_update_mags(len(vel), _p(vel), _p(acc), _p(dv), _p(dp))
    return dv, dp


def ref_particle_raw(pos, vel, acc, mass, prec="f64"):
    """Raw bytes of the reference's particle objects (what .dat holds)."""
    L = ref_lib(prec)
    L.ref_particle_bytes.restype = ctypes.c_int
    rec = L.ref_particle_bytes()
    pos, vel, acc, mass = (np.ascontiguousarray(a, dtype=np.float64)
                           for a in (pos, vel, acc, mass))
    out = np.zeros(len(mass) * rec, np.uint8)
    L.ref_particle_raw.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp,
                                   ctypes.POINTER(ctypes.c_uint8)]
    L.ref_particle_raw(len(mass), _p(pos), _p(vel), _p(acc), _p(mass),
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out.tobytes()


def ref_set_threaded(on, prec="f64"):
    """The _ref harness builds gen_root_threaded's tree (threads > 1)."""
    L = ref_lib(prec)
    L.ref_set_threaded.argtypes = [ctypes.c_int]
    L.ref_set_threaded(int(on))


def ref_tree_dump(pos, mass, prec="f64", threaded=False):
    L = ref_lib(prec)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    n = len(pos)
    cap = 64 * n + 64
    d = dict(depth=np.zeros(cap, np.int32), leaf=np.zeros(cap, np.int32),
             first=np.zeros(cap, np.int64), count=np.zeros(cap, np.int64),
             side=np.zeros(cap), mass=np.zeros(cap), com=np.zeros((cap, 3)),
             order=np.zeros(cap, np.int32))
    ref_set_threaded(threaded, prec)
    try:
        m = L.ref_tree_dump(n, _p(pos), _p(mass), cap, _p(d["depth"], _ip),
                            _p(d["leaf"], _ip), _p(d["first"], _lp), _p(d["count"], _lp),
                            _p(d["side"]), _p(d["mass"]), _p(d["com"]), _p(d["order"], _ip))
    finally:
        ref_set_threaded(False, prec)
    assert m <= cap
    for k in list(d):
        if k != "order":
            d[k] = d[k][:m].copy()
    d["order"] = d["order"][:int((d["leaf"] >= 0).sum())].copy()
    return d


def ref_forces(pos, mass, theta=0.5, damping=False, prec="f64", threaded=False):
    ref_set_threaded(threaded, prec)
    try:
        return _ref_forces(pos, mass, theta, damping, prec)
    finally:
        ref_set_threaded(False, prec)


def _ref_forces(pos, mass, theta=0.5, damping=False, prec="f64"):
    L = ref_lib(prec)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    acc = np.zeros_like(pos)
    cnt = np.zeros(len(pos), np.int64)
    L.ref_forces(len(pos), _p(pos), _p(mass), theta, int(damping), _p(acc), _p(cnt, _lp))
    return acc, cnt


def ref_run(pos, vel, mass, theta=0.5, dt=0.0333, frames=1, damping=False, prec="f64",
            threaded=False):
    L = ref_lib(prec)
    pos = np.array(pos, dtype=np.float64, order="C")
    vel = np.array(vel, dtype=np.float64, order="C")
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    acc = np.zeros_like(pos)
    ref_set_threaded(threaded, prec)
    try:
        L.ref_run(len(pos), _p(pos), _p(vel), _p(acc), _p(mass), theta, dt, int(damping),
                  frames)
    finally:
        ref_set_threaded(False, prec)
    return pos, vel, acc


def ref_run_collide(pos, vel, mass, theta=0.5, dt=0.0333, frames=1, damping=False, range=0.01):
    """Collision frames over the reference's classes (FP64); returns (pos,
    vel, acc, mass, distinct pairs recorded), compacted to the survivors."""
    L = ref_lib("f64")
    L.ref_run_collide.restype = ctypes.c_long
    L.ref_run_collide.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  _lp]
    pos = np.array(pos, dtype=np.float64, order="C")
    vel = np.array(vel, dtype=np.float64, order="C")
    mass = np.array(mass, dtype=np.float64, order="C")
    acc = np.zeros_like(pos)
    npairs = ctypes.c_long()
    n = L.ref_run_collide(len(pos), _p(pos), _p(vel), _p(acc), _p(mass), theta, dt,
                          int(damping), range, frames, ctypes.byref(npairs))
    return pos[:n], vel[:n], acc[:n], mass[:n], npairs.value


def render_image(pos, width=1920, height=1080, projection="iso", color="bw", adaptive=False,
                 nonlinear=False, scale=1.0, brightness=255.0):
    """write_image's pixels for FP64 positions: [height][width][3] uint8,
    top row first."""
    proj = {"front": 0, "side": 1, "top": 2, "iso": 3}[projection]
    col = {"bw": 0, "heat": 1}[color]
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros((height, width, 3), np.uint8)
    L = lib()
    L.orc_render_image.argtypes = [ctypes.c_size_t, _dp, ctypes.c_uint, ctypes.c_uint,
                                   ctypes.c_uint, ctypes.c_uint, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_uint8)]
    L.orc_render_image(len(pos), _p(pos), width, height, proj, col, int(adaptive),
                       int(nonlinear), scale, brightness,
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out
