"""libnbs.so loads and exports every function include/*.h declares, with
the signature the Python bindings use.  No compute call is made (CPU only);
argument checking that returns before any device work is exercised."""
import ctypes
import os
import re

import pytest

import nbodysim_amd
from nbodysim_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("nbs.h", "nbs_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(nbs_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("nbs_create", "nbs_destroy", "nbs_set_state", "nbs_forces", "nbs_step",
                 "nbs_get_state", "nbs_tree_dump", "nbs_last_error", "nbs_set_systems",
                 "nbs_direct", "nbs_settings_read", "nbs_generate", "nbs_write_binary"):
        assert must in names


def test_every_declared_symbol_is_exported():
    L = nbodysim_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    bound = {s[0] for s in _native.SIGNATURES}
    assert set(declared_functions()) == bound


def test_config_struct_layout():
    # nbs_config: 2 doubles, 4 ints, 3+1 doubles, 3 doubles, 2 ints, 6 ints
    assert ctypes.sizeof(_native.Config) == 2 * 8 + 4 * 4 + 4 * 8 + 3 * 8 + 8 * 4
    cfg = _native.Config()
    nbodysim_amd.lib().nbs_config_default(ctypes.byref(cfg))
    assert cfg.theta == 0.5 and cfg.dt == 0.0333 and cfg.node_factor == 4


def test_create_rejects_bad_config_without_touching_the_gpu():
    L = nbodysim_amd.lib()
    cfg = _native.Config()
    L.nbs_config_default(ctypes.byref(cfg))
    cfg.force_precision = 7
    h = ctypes.c_void_p()
    assert L.nbs_create(ctypes.byref(cfg), ctypes.byref(h)) == _native.NBS_ERR_ARG
    cfg.force_precision = 0
    cfg.theta = -1.0
    assert L.nbs_create(ctypes.byref(cfg), ctypes.byref(h)) == _native.NBS_ERR_ARG


def test_null_context_is_an_error_not_a_crash():
    L = nbodysim_amd.lib()
    assert L.nbs_step(None, 1) == _native.NBS_ERR_ARG
    assert L.nbs_forces(None, None) == _native.NBS_ERR_ARG
    assert L.nbs_last_error(None) == b"null context"


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", "/nonexistent/libnbs.so")
    with pytest.raises(ImportError):
        _native.lib()


def test_noalloc_rejects_sizes_beyond_its_maximum_without_touching_the_gpu():
    # nbs_forces_noalloc checks the size before any HIP call
    L = nbodysim_amd.lib()
    cfg = _native.Config()
    L.nbs_config_default(ctypes.byref(cfg))
    msg = ctypes.create_string_buffer(256)
    rc = L.nbs_forces_noalloc(ctypes.byref(cfg), _native.NBS_NOALLOC_MAX_N + 1, 8, 8, 8,
                              None, None, msg, len(msg))
    assert rc == _native.NBS_ERR_ARG
    assert b"exceeds the no-allocation maximum of 65536" in msg.value
    assert L.nbs_forces_noalloc(None, 10, 8, 8, 8, None, None, msg, len(msg)) == \
        _native.NBS_ERR_ARG
    assert L.nbs_forces_noalloc(ctypes.byref(cfg), 0, None, None, None, None, None, None,
                                0) == _native.NBS_OK
