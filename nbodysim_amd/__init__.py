"""nbodysim_amd -- MI355X Barnes-Hut step engine, a drop-in for
fosted3/nbodysim's per-frame step.

Python front-end over the C ABI of libnbs.so (include/nbs.h):

    sim = nbodysim_amd.Simulation(theta=0.5, dt=0.0333, precision="f32")
    sim.set_state(pos, vel, mass)      # float64 [n,3], [n,3], [n]
    acc = sim.forces()                 # barnes_hut_cuda's role
    sim.step(10)                       # whole frames on the GPU
    pos, vel, acc = sim.state()

plus the host contract of the reference (settings.cfg reader, seeded
generator, .dat/.txt frame files) in nbodysim_amd.host.
"""
import ctypes

import numpy as np

from . import _native
from ._native import (NBS_FORCE_F32, NBS_FORCE_F64, NBS_ROOT_BBOX, NBS_ROOT_FIXED,
                      Config, TreeInfo, lib)

__all__ = ["Simulation", "NbsError", "lib", "NBS_FORCE_F32", "NBS_FORCE_F64"]


class NbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nbs error %d: %s" % (code, msg))
        self.code = code


def _dp(a):
    return a.ctypes.data_as(_native.c_dp) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


class Simulation:
    """One GPU context: particle state resident in HBM, one stream."""

    def __init__(self, theta=0.5, dt=0.0333, damping=False, precision="f64",
                 root_center=None, root_size=None, adaptive=False, min_adaptive_dt=0.00333,
                 max_pos_change=3.0, max_vel_change=3.0, device=0, node_factor=4,
                 threaded_tree=False):
        self._L = lib()
        self.cfg = Config()
        self._L.nbs_config_default(ctypes.byref(self.cfg))
        self._fill_config(theta, dt, damping, precision, root_center, root_size, adaptive,
                          min_adaptive_dt, max_pos_change, max_vel_change, device,
                          node_factor, threaded_tree)
        h = ctypes.c_void_p()
        rc = self._L.nbs_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if rc != 0:
            raise NbsError(rc, "nbs_create failed")
        self._h = h
        self.n = 0
        self._collide = False

    def _fill_config(self, theta, dt, damping, precision, root_center, root_size, adaptive,
                     min_adaptive_dt, max_pos_change, max_vel_change, device, node_factor,
                     threaded_tree=False):
        c = self.cfg
        c.theta = theta
        c.dt = dt
        c.damping = int(bool(damping))
        c.force_precision = {"f64": NBS_FORCE_F64, "f32": NBS_FORCE_F32}[precision]
        if root_center is not None:
            c.root_mode = NBS_ROOT_FIXED
            for k in range(3):
                c.root_center[k] = root_center[k]
            c.root_size = root_size
        else:
            c.root_mode = NBS_ROOT_BBOX
        c.adaptive = int(bool(adaptive))
        c.min_adaptive_dt = min_adaptive_dt
        c.max_pos_change = max_pos_change
        c.max_vel_change = max_vel_change
        c.device = device
        c.node_factor = node_factor
        c.threaded_tree = int(bool(threaded_tree))

    def configure(self, **kw):
        """Changes theta / dt / precision / ... of this context."""
        c = self.cfg
        args = dict(theta=c.theta, dt=c.dt, damping=bool(c.damping),
                    precision="f32" if c.force_precision == NBS_FORCE_F32 else "f64",
                    root_center=list(c.root_center) if c.root_mode == NBS_ROOT_FIXED else None,
                    root_size=c.root_size, adaptive=bool(c.adaptive),
                    min_adaptive_dt=c.min_adaptive_dt, max_pos_change=c.max_pos_change,
                    max_vel_change=c.max_vel_change, device=c.device,
                    node_factor=c.node_factor, threaded_tree=bool(c.threaded_tree))
        args.update(kw)
        self._fill_config(**args)
        self._check(self._L.nbs_configure(self._h, ctypes.byref(self.cfg)))

    def _check(self, rc):
        if rc != 0:
            raise NbsError(rc, self._L.nbs_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.nbs_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- state
    def set_state(self, pos, vel, mass):
        mass = _f64(mass)
        n = len(mass)
        pos = _f64(pos, (n, 3))
        vel = None if vel is None else _f64(vel, (n, 3))
        self._check(self._L.nbs_set_state(self._h, n, _dp(pos), _dp(vel), _dp(mass)))
        self.n = n

    def set_state_device(self, n, pos_ptr, vel_ptr, mass_ptr):
        """nbs_set_state_device: n particles from device arrays (addresses
        of float64 [n][3], [n][3] or 0 for zero velocities, [n]) of the HIP
        runtime this library is bound to; they must be ready (their streams
        synchronized) when called."""
        self._check(self._L.nbs_set_state_device(self._h, int(n), int(pos_ptr),
                                                 int(vel_ptr) or None, int(mass_ptr)))
        self.n = int(n)

    def state_device(self, pos_ptr, vel_ptr=0, acc_ptr=0):
        """nbs_get_state_device: the state in caller order into device
        arrays (addresses, 0 to skip); returns once they are written."""
        self._check(self._L.nbs_get_state_device(self._h, int(pos_ptr) or None,
                                                 int(vel_ptr) or None, int(acc_ptr) or None))

    def generate_state(self, n=0, gen_type="cube", seed=0xDEADBEEF, settings=None,
                       systems=None, **kw):
        """Seeded particles generated on the device straight into the state
        (nbs_generate_state): n particles of gen_type from `seed`, or with
        systems = [(n_s, gen_type, seed), ...] one batched system each.
        Other settings as in host.generate (mass_dist linear)."""
        from . import host as _host
        d = dict(settings or _host.default_settings())
        d["gen_type"] = gen_type
        d.update(kw)
        st = _host._from_dict(d)
        if systems is None:
            systems = [(n, gen_type, seed)]
        off = np.ascontiguousarray(np.cumsum([0] + [int(x[0]) for x in systems]),
                                   dtype=np.int64)
        seeds = np.ascontiguousarray([int(x[2]) for x in systems], dtype=np.uint64)
        kinds = np.ascontiguousarray([_host._ENUMS["gen_type"][x[1]] for x in systems],
                                     dtype=np.uint32)
        self._check(self._L.nbs_generate_state(
            self._h, ctypes.byref(st), len(systems), off.ctypes.data_as(_native.c_i64p),
            seeds.ctypes.data_as(_native.c_u64p),
            kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        self.n = int(off[-1])
        return off

    def set_systems(self, offsets):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self._check(self._L.nbs_set_systems(self._h, len(off) - 1,
                                            off.ctypes.data_as(_native.c_i64p)))

    def state(self):
        n = self.n
        pos, vel, acc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.nbs_get_state(self._h, _dp(pos), _dp(vel), _dp(acc)))
        return pos, vel, acc

    def masses(self):
        """Masses in the caller's order."""
        m = np.zeros(self.n)
        self._check(self._L.nbs_get_mass(self._h, _dp(m)))
        return m

    def snapshot_begin(self):
        """Starts an asynchronous copy of pos, vel, acc (caller order) to a
        pinned host buffer behind the work already submitted; the steps
        submitted next overlap it."""
        self._check(self._L.nbs_snapshot_begin(self._h))

    def snapshot_end(self, copy=True):
        """Waits for the pending snapshot; returns [n][3] pos, vel, acc.
        With copy=False the arrays view the context's pinned buffer, valid
        until the second following snapshot_begin."""
        p, v, a = _native.c_dp(), _native.c_dp(), _native.c_dp()
        n = ctypes.c_size_t()
        self._check(self._L.nbs_snapshot_end(self._h, ctypes.byref(p), ctypes.byref(v),
                                             ctypes.byref(a), ctypes.byref(n)))
        n = n.value
        if n == 0:
            return np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3))
        out = tuple(np.ctypeslib.as_array(x, shape=(n, 3)) for x in (p, v, a))
        return tuple(x.copy() for x in out) if copy else out

    # ----------------------------------------------------------------- step
    def forces(self, download=True):
        if not download:
            self._check(self._L.nbs_forces(self._h, None))
            return None
        acc = np.zeros((self.n, 3))
        self._check(self._L.nbs_forces(self._h, _dp(acc)))
        return acc

    def step(self, nsteps=1, graph=False):
        """nsteps frames; graph=True replays a captured HIP graph."""
        f = self._L.nbs_step_graph if graph else self._L.nbs_step
        self._check(f(self._h, int(nsteps)))
        if self._collide:
            n = ctypes.c_int64()
            self._check(self._L.nbs_particle_count(self._h, ctypes.byref(n), None, 0))
            self.n = n.value

    def set_collisions(self, range=0.01, enabled=True):
        """Collision merging in step() (nbs_set_collisions): pairs closer
        than `range` found by the walk merge; the particle count shrinks."""
        self._check(self._L.nbs_set_collisions(self._h, int(bool(enabled)), float(range)))
        self._collide = bool(enabled)

    def set_min_node_size(self, min_size):
        """apply_minimum_size in step() (nbs_set_min_node_size); 0 turns it
        off.  Returns the particles it absorbed before this call."""
        m = ctypes.c_int64()
        self._check(self._L.nbs_set_min_node_size(self._h, float(min_size), ctypes.byref(m)))
        self._collide = self._collide or min_size > 0
        return m.value

    def collision_stats(self):
        """(pairs detected, particles absorbed) since set_collisions."""
        p, r = ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.nbs_collision_stats(self._h, ctypes.byref(p), ctypes.byref(r)))
        return p.value, r.value

    def render_image(self, width=1920, height=1080, projection="iso", color="bw",
                     adaptive=False, nonlinear=False, scale=1.0, brightness=255.0):
        """The frame image of the current positions (nbs_render_image):
        [height][width][3] uint8, top row first."""
        ip = _native.ImageParams(width, height,
                                 {"front": 0, "side": 1, "top": 2, "iso": 3}[projection],
                                 {"bw": 0, "heat": 1}[color], int(bool(adaptive)),
                                 int(bool(nonlinear)), float(scale), float(brightness))
        out = np.zeros((height, width, 3), np.uint8)
        self._check(self._L.nbs_render_image(self._h, ctypes.byref(ip),
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def offsets(self, nsys):
        """Current system offsets (nsys + 1 entries)."""
        off = np.zeros(nsys + 1, dtype=np.int64)
        n = ctypes.c_int64()
        self._check(self._L.nbs_particle_count(self._h, ctypes.byref(n),
                                               off.ctypes.data_as(_native.c_i64p), nsys + 1))
        return off

    WALK_BLOCKS = (64, 128, 256, 512, 1024)
    WALK_GROUPS = (1, 8, 16, 32, 64)
    WALK_PACKS = (1, 2)

    def autotune(self, reps=3):
        """Times `reps` steps from the current state for every walk shape
        (workgroup size x XCD interleave x particles per lane), keeps the
        fastest and restores the state; returns ((block, groups, pack),
        {(block, groups, pack): ms/step})."""
        bb, bg, bp = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        shapes = [(b, g, p) for p in self.WALK_PACKS for b in self.WALK_BLOCKS
                  for g in self.WALK_GROUPS]
        ms = np.zeros(len(shapes))
        self._check(self._L.nbs_autotune(self._h, int(reps), ctypes.byref(bb), ctypes.byref(bg),
                                         ctypes.byref(bp), _dp(ms)))
        return (bb.value, bg.value, bp.value), dict(zip(shapes, ms.tolist()))

    def set_walk_shape(self, block, groups=1, pack=1):
        self._check(self._L.nbs_set_walk_shape(self._h, int(block), int(groups), int(pack)))

    def set_walk_order(self, on=True):
        """Longest-first workgroup dispatch of the walk (default off)."""
        self._check(self._L.nbs_set_walk_order(self._h, int(bool(on))))

    def set_walk_code(self, generic=False, records="auto"):
        """FP32 walk loop: hand-written gfx950 code (default) or the
        compiler-generated loop (generic=True); the hand-written loop reads
        32- or 64-byte node records ("auto": by particle count).  Results
        are bit-identical."""
        codes = {"auto": 0, 64: 2, 32: 3}
        if records not in codes:
            raise ValueError("records must be 'auto', 32 or 64")
        self._check(self._L.nbs_set_walk_code(self._h, 1 if generic else codes[records]))

    def set_walk_split(self, pieces=0):
        """Pieces per walk traversal: 0 (default) by the wavefront count
        (small systems split each traversal to fill the GPU), or a power of
        two up to 32.
        Interaction counts never change; forces differ only by the order of
        the pieces' sums."""
        self._check(self._L.nbs_set_walk_split(self._h, int(pieces)))

    @property
    def walk_split(self):
        """Pieces per walk traversal in use."""
        return self._L.nbs_walk_split(self._h)

    def set_counting(self, mode="particles"):
        """Interaction counting of the walks: "particles" (default; per
        particle and the running total), "total" (the running total only,
        no per-lane work in the walk) or "none".  Changes no result bit."""
        modes = {"none": _native.NBS_COUNT_NONE, "total": _native.NBS_COUNT_TOTAL,
                 "particles": _native.NBS_COUNT_PARTICLES}
        if mode not in modes:
            raise ValueError("counting mode must be one of %s" % sorted(modes))
        self._check(self._L.nbs_set_counting(self._h, modes[mode]))

    def synchronize(self):
        self._check(self._L.nbs_synchronize(self._h))

    def direct(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        acc = np.zeros((len(ids), 3))
        self._check(self._L.nbs_direct(self._h, len(ids), ids.ctypes.data_as(_native.c_i64p),
                                       _dp(acc)))
        return acc

    def interactions(self, per_particle=False):
        tot = ctypes.c_int64()
        per = np.zeros(self.n, np.int64) if per_particle else None
        self._check(self._L.nbs_interactions(
            self._h, ctypes.byref(tot),
            per.ctypes.data_as(_native.c_i64p) if per is not None else None))
        return (tot.value, per) if per_particle else tot.value

    def interactions_total(self, reset=True):
        """Interactions summed over every walk since the last reset."""
        tot = ctypes.c_int64()
        self._check(self._L.nbs_interactions_total(self._h, ctypes.byref(tot), int(reset)))
        return tot.value

    def time(self, nsys=1):
        t = np.zeros(nsys)
        self._check(self._L.nbs_time(self._h, _dp(t), nsys))
        return t

    # -------------------------------------------------------------- adjoint
    def adjoint_forward(self, nsteps, every=20, record_all=False):
        """nsteps frames with checkpoints every `every` steps (reverse mode)."""
        self._check(self._L.nbs_adjoint_forward(self._h, int(nsteps), int(every),
                                                int(bool(record_all))))

    def adjoint_backward(self, gx_T=None, gv_T=None):
        """Returns (dL/dx0, dL/dv0, dL/dm) for the given dL/dx_T, dL/dv_T."""
        n = self.n
        gx = None if gx_T is None else _f64(gx_T, (n, 3))
        gv = None if gv_T is None else _f64(gv_T, (n, 3))
        gx0, gv0, gm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        self._check(self._L.nbs_adjoint_backward(self._h, _dp(gx), _dp(gv), _dp(gx0),
                                                 _dp(gv0), _dp(gm)))
        return gx0, gv0, gm

    def adjoint_backward_device(self, gx_T_ptr, gv_T_ptr, gx0_ptr, gv0_ptr, gm_ptr):
        """nbs_adjoint_backward_device: seeds and gradients as device array
        addresses (0 = zero seed / not wanted)."""
        p = lambda a: int(a) or None
        self._check(self._L.nbs_adjoint_backward_device(self._h, p(gx_T_ptr), p(gv_T_ptr),
                                                        p(gx0_ptr), p(gv0_ptr), p(gm_ptr)))

    def frozen_forward(self, pos0, vel0, mass):
        """Replays the recorded forward at perturbed inputs (FP64, fixed
        cell decisions); returns (x_T, v_T)."""
        n = self.n
        pos0, mass = _f64(pos0, (n, 3)), _f64(mass, (n,))
        vel0 = None if vel0 is None else _f64(vel0, (n, 3))
        xT, vT = np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.nbs_frozen_forward(self._h, _dp(pos0), _dp(vel0), _dp(mass),
                                               _dp(xT), _dp(vT)))
        return xT, vT

    # ----------------------------------------------------------------- tree
    def tree_info(self):
        info = TreeInfo()
        self._check(self._L.nbs_tree_info_get(self._h, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in TreeInfo._fields_}

    def tree_dump(self):
        info = self.tree_info()
        m, n = info["n_nodes"], info["n_nodes"] - info["n_internal"]  # n: leaves
        d = dict(depth=np.zeros(m, np.int32), leaf=np.zeros(m, np.int32),
                 first=np.zeros(m, np.int64), count=np.zeros(m, np.int64),
                 side=np.zeros(m), mass=np.zeros(m), com=np.zeros((m, 3)),
                 khi=np.zeros(m, np.uint64), klo=np.zeros(m, np.uint64),
                 order=np.zeros(n, np.int32))
        P = lambda a, t: a.ctypes.data_as(t)
        self._check(self._L.nbs_tree_dump(
            self._h, P(d["depth"], _native.c_i32p), P(d["leaf"], _native.c_i32p),
            P(d["first"], _native.c_i64p), P(d["count"], _native.c_i64p), _dp(d["side"]),
            _dp(d["mass"]), _dp(d["com"]), P(d["khi"], _native.c_u64p),
            P(d["klo"], _native.c_u64p), P(d["order"], _native.c_i32p)))
        return d

    def keys(self):
        hi = np.zeros(self.n, np.uint64)
        lo = np.zeros(self.n, np.uint64)
        self._check(self._L.nbs_keys(self._h, hi.ctypes.data_as(_native.c_u64p),
                                     lo.ctypes.data_as(_native.c_u64p)))
        return hi, lo

    def roots(self, nsys=1):
        c = np.zeros((nsys, 3))
        s = np.zeros(nsys)
        self._check(self._L.nbs_roots(self._h, _dp(c), _dp(s), nsys))
        return c, s

    # ------------------------------------------------------------ profiling
    def profile(self, on=True):
        self._check(self._L.nbs_profile_enable(self._h, int(on)))

    def profile_read(self, reset=True):
        ms = np.zeros(4)
        self._check(self._L.nbs_profile_read(self._h, _dp(ms), int(reset)))
        return {"build_ms": ms[0], "walk_ms": ms[1], "update_ms": ms[2],
                "walk_launches": int(ms[3])}
