"""nbs_forces_noalloc (no device allocation: every buffer is carved from
one fixed library-owned array per GPU) against nbs_forces on a context,
through the C ABI: accelerations and per-particle interaction counts equal
bit for bit (the same kernels run on the same inputs), on two calls in a
row with the second smaller (the shared array keeps the first call's
contents), in FP64 and FP32, at the maximum size, and sizes beyond it
rejected."""
import ctypes

import numpy as np
import pytest

import nbodysim_amd
from nbodysim_amd import _native
import oracle as O

pytestmark = pytest.mark.gpu


class HipRuntime:
    """hipMalloc / hipMemcpy / hipFree / streams of the HIP runtime libnbs.so
    is linked against (the caller's side of the boundary: the caller owns
    the input and output device arrays).  Loaded by soname after libnbs.so,
    so the process holds one runtime; torch ships another, which must not be
    started after this one."""

    def __init__(self):
        nbodysim_amd.lib()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]

    def check(self, rc, what):
        assert rc == 0, "%s failed: %d" % (what, rc)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 1)), "hipMalloc")
        self.check(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.check(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self, p):
        self.check(self.L.hipFree(p), "hipFree")


def noalloc(pos, mass, precision="f64", theta=0.5, stream=None):
    hip = HipRuntime()
    L = nbodysim_amd.lib()
    cfg = _native.Config()
    L.nbs_config_default(ctypes.byref(cfg))
    cfg.theta = theta
    cfg.force_precision = {"f64": _native.NBS_FORCE_F64, "f32": _native.NBS_FORCE_F32}[precision]
    n = len(mass)
    acc0 = np.full((n, 3), np.nan)
    cnt0 = np.full(n, -1, dtype=np.int64)
    bufs = [hip.upload(np.asarray(pos, dtype=np.float64)),
            hip.upload(np.asarray(mass, dtype=np.float64)), hip.upload(acc0), hip.upload(cnt0)]
    try:
        msg = ctypes.create_string_buffer(512)
        rc = L.nbs_forces_noalloc(ctypes.byref(cfg), n, bufs[0], bufs[1], bufs[2], bufs[3],
                                  stream, msg, len(msg))
        if rc != 0:
            raise nbodysim_amd.NbsError(rc, msg.value.decode())
        return hip.download(bufs[2], acc0), hip.download(bufs[3], cnt0)
    finally:
        for b in bufs:
            hip.free(b)


def with_context(pos, mass, precision="f64", theta=0.5):
    with nbodysim_amd.Simulation(theta=theta, precision=precision) as s:
        s.set_state(pos, None, mass)
        acc = s.forces()
        _, cnt = s.interactions(per_particle=True)
    return acc, cnt


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_two_calls_second_smaller(precision):
    big = O.generate(30000, "sphere", seed=5)
    small = O.generate(7001, "cube", seed=6)
    print("noalloc", precision, "n = 30000 then 7001", flush=True)
    for pos, _, mass in (big, small):
        got = noalloc(pos, mass, precision)
        want = with_context(pos, mass, precision)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
    # and the FP64 result is the oracle's tree walk
    if precision == "f64":
        oacc, ocnt = O.Tree(small[0], small[2]).forces(theta=0.5)
        np.testing.assert_array_equal(got[1], ocnt)
        rel = np.linalg.norm(got[0] - oacc, axis=1) / np.linalg.norm(oacc, axis=1)
        assert rel.max() < 1e-12


def test_maximum_size_and_beyond():
    n = _native.NBS_NOALLOC_MAX_N
    pos, _, mass = O.generate(n, "shell", seed=8)
    print("noalloc at the maximum, n =", n, flush=True)
    hip = HipRuntime()
    stream = ctypes.c_void_p()
    hip.check(hip.L.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    try:
        got = noalloc(pos, mass, "f32", theta=0.7, stream=stream)
    finally:
        hip.check(hip.L.hipStreamDestroy(stream), "hipStreamDestroy")
    want = with_context(pos, mass, "f32", theta=0.7)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    with pytest.raises(nbodysim_amd.NbsError, match="exceeds the no-allocation maximum"):
        noalloc(np.zeros((n + 1, 3)), np.ones(n + 1))


def test_tiny_inputs():
    for n in (1, 2, 65):
        pos, _, mass = O.generate(n, "cube", seed=9)
        got = noalloc(pos, mass)
        want = with_context(pos, mass)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
