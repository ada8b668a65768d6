"""Builds libnbs.so (HIP kernels for gfx950 + C ABI + host contract code)
and the nbodysim command-line driver, in-tree under nbodysim_amd/_lib/.

    python -m nbodysim_amd.build            # rebuild everything
    python -m nbodysim_amd.build --lazy     # only what changed
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "_lib")
LIBNBS = os.path.join(LIB, "libnbs.so")
CLI = os.path.join(LIB, "nbodysim")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NBS_OFFLOAD_ARCH", "gfx950")
# Packed FP32 (v_pk_*_f32) is switched off for all device code.
HIP_FLAGS = ["-std=c++17", "--offload-arch=" + ARCH, "-O3", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result"]
# Host contract code (settings, IC generation, dumps): the reference's
# arithmetic without FMA contraction, so seeding is bit-exact.
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-Wall",
             "-I" + os.path.join(ROOT, "include")]

HIP_SOURCES = ["scan.hip", "radix_sort.hip", "tree_build.hip", "tree_threaded.hip", "walk.hip",
               "integrate.hip", "adjoint.hip", "ic_device.hip", "collide.hip",
               "image.hip", "api.hip"]
# Device initial conditions repeat the host generator's roundings: no
# contraction into FMAs anywhere in that file (the device default would fuse
# a multiply with the add that follows it, even across inlined helpers).
FILE_FLAGS = {"ic_device.hip": ["-ffp-contract=off"], "collide.hip": ["-ffp-contract=off"],
              "tree_threaded.hip": ["-ffp-contract=off"],
              "image.hip": ["-ffp-contract=off"]}
HOST_SOURCES = ["host/settings.cpp", "host/ic.cpp", "host/dump.cpp", "host/frames.cpp",
                "host/png.cpp"]
CLI_SOURCES = ["host/nbodysim_main.cpp"]


def _deps(src):
    d = [src]
    for f in os.listdir(CSRC):
        if f.endswith(".h"):
            d.append(os.path.join(CSRC, f))
    hd = os.path.join(CSRC, "host")
    if os.path.isdir(hd):
        d += [os.path.join(hd, f) for f in os.listdir(hd) if f.endswith(".h")]
    d.append(os.path.join(ROOT, "include", "nbs.h"))
    return d


def _stale(out, deps):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.exists(d) and os.path.getmtime(d) > t for d in deps)


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "feature flag" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _compile(src, lazy):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src.replace("/", "_") + ".o")
    if lazy and not _stale(obj, _deps(path)):
        return obj
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + ["-c", path, "-o", obj + ".tmp"]
    else:
        cmd = [CXX] + CXX_FLAGS + ["-c", path, "-o", obj + ".tmp"]
    _run(cmd)
    os.replace(obj + ".tmp", obj)
    return obj


def build(lazy=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = HIP_SOURCES + HOST_SOURCES
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, lazy), srcs))
        cli_objs = list(ex.map(lambda s: _compile(s, lazy), CLI_SOURCES))
    if not lazy or _stale(LIBNBS, objs):
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs +
             ["-pthread", "-lz", "-o", LIBNBS + ".tmp"])
        os.replace(LIBNBS + ".tmp", LIBNBS)
    if not lazy or _stale(CLI, cli_objs + [LIBNBS]):
        _run([CXX, "-o", CLI + ".tmp"] + cli_objs +
             ["-L" + LIB, "-lnbs", "-Wl,-rpath,$ORIGIN", "-pthread"])
        os.replace(CLI + ".tmp", CLI)
    return LIBNBS


if __name__ == "__main__":
    print(build(lazy="--lazy" in sys.argv))
