"""ctypes bindings of libnbs.so (include/nbs.h, include/nbs_host.h).

The library is built in-tree (python -m nbodysim_amd.build); loading fails
with an ImportError when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libnbs.so")

c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_i32p = ctypes.POINTER(ctypes.c_int32)

NBS_OK, NBS_ERR_ARG, NBS_ERR_HIP, NBS_ERR_STATE, NBS_ERR_OVERFLOW, NBS_ERR_NOMEM, \
    NBS_ERR_IO = range(7)
NBS_FORCE_F64, NBS_FORCE_F32 = 0, 1
NBS_ROOT_BBOX, NBS_ROOT_FIXED = 0, 1
NBS_COUNT_NONE, NBS_COUNT_TOTAL, NBS_COUNT_PARTICLES = 0, 1, 2


class Config(ctypes.Structure):
    _fields_ = [("theta", ctypes.c_double), ("dt", ctypes.c_double),
                ("damping", ctypes.c_int), ("force_precision", ctypes.c_int),
                ("root_mode", ctypes.c_int), ("adaptive", ctypes.c_int),
                ("root_center", ctypes.c_double * 3), ("root_size", ctypes.c_double),
                ("min_adaptive_dt", ctypes.c_double), ("max_pos_change", ctypes.c_double),
                ("max_vel_change", ctypes.c_double), ("device", ctypes.c_int),
                ("node_factor", ctypes.c_int), ("threaded_tree", ctypes.c_int),
                ("reserved", ctypes.c_int * 5)]


class ImageParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("projection", ctypes.c_uint32), ("color", ctypes.c_uint32),
                ("adaptive", ctypes.c_int32), ("nonlinear", ctypes.c_int32),
                ("scale", ctypes.c_double), ("brightness", ctypes.c_double)]


class TreeInfo(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int64), ("n_nodes", ctypes.c_int64),
                ("n_internal", ctypes.c_int64), ("max_depth", ctypes.c_int32),
                ("n_buckets", ctypes.c_int32), ("node_capacity", ctypes.c_int64)]


class Settings(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_int) for k in (
        "read_existing", "use_seed", "display_progress", "dump_binary", "dump_text",
        "dump_image", "overwrite_data", "keep_previous_binary", "keep_previous_text",
        "verbose", "damping", "adaptive_brightness", "nonlinear_brightness", "collide",
        "cuda")] +
        [(k, ctypes.c_uint32) for k in (
            "num_particles", "num_frames", "threads", "img_w", "img_h", "projection")] +
        [(k, ctypes.c_double) for k in ("size", "theta", "dt")] +
        [("seed", ctypes.c_ulong)] +
        [(k, ctypes.c_double) for k in (
            "min_mass", "max_mass", "min_vel", "max_vel", "r_sphere", "rotation_magnitude",
            "scale_x", "scale_y", "scale_z", "brightness", "scale", "collision_range")] +
        [("rotation_vector", ctypes.c_double * 3)] +
        [(k, ctypes.c_uint32) for k in ("gen_type", "mass_dist", "vel_dist", "color")] +
        [("min_node_size", ctypes.c_double), ("adaptive", ctypes.c_int),
         ("min_adaptive_dt", ctypes.c_double), ("max_pos_change", ctypes.c_double),
         ("max_vel_change", ctypes.c_double), ("force_precision", ctypes.c_int),
         ("shell_jitter", ctypes.c_double), ("datatype_float", ctypes.c_int)])


# (name, restype, argtypes) of every exported function, in header order.
SIGNATURES = [
    ("nbs_config_default", None, [ctypes.POINTER(Config)]),
    ("nbs_create", ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]),
    ("nbs_destroy", None, [ctypes.c_void_p]),
    ("nbs_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("nbs_configure", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Config)]),
    ("nbs_set_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, c_dp, c_dp, c_dp]),
    ("nbs_set_systems", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_i64p]),
    ("nbs_get_state", ctypes.c_int, [ctypes.c_void_p, c_dp, c_dp, c_dp]),
    ("nbs_set_state_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("nbs_get_state_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("nbs_get_mass", ctypes.c_int, [ctypes.c_void_p, c_dp]),
    ("nbs_snapshot_begin", ctypes.c_int, [ctypes.c_void_p]),
    ("nbs_snapshot_end", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(c_dp),
                                        ctypes.POINTER(c_dp), ctypes.POINTER(c_dp),
                                        ctypes.POINTER(ctypes.c_size_t)]),
    ("nbs_forces", ctypes.c_int, [ctypes.c_void_p, c_dp]),
    ("nbs_forces_noalloc", ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_size_t,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                          ctypes.c_size_t]),
    ("nbs_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_step_graph", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_set_collisions", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]),
    ("nbs_set_min_node_size", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, c_i64p]),
    ("nbs_collision_stats", ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p]),
    ("nbs_particle_count", ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p, ctypes.c_int]),
    ("nbs_render_image", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ImageParams),
                                        ctypes.POINTER(ctypes.c_uint8)]),
    ("nbs_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_i32p, c_i32p, c_i32p,
                                    c_dp]),
    ("nbs_set_walk_shape", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int]),
    ("nbs_set_walk_order", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_set_walk_code", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_set_walk_split", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_walk_split", ctypes.c_int, [ctypes.c_void_p]),
    ("nbs_set_counting", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("nbs_direct", ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, c_i64p, c_dp]),
    ("nbs_interactions", ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p]),
    ("nbs_interactions_total", ctypes.c_int, [ctypes.c_void_p, c_i64p, ctypes.c_int]),
    ("nbs_time", ctypes.c_int, [ctypes.c_void_p, c_dp, ctypes.c_int]),
    ("nbs_tree_info_get", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TreeInfo)]),
    ("nbs_tree_dump", ctypes.c_int, [ctypes.c_void_p, c_i32p, c_i32p, c_i64p, c_i64p, c_dp,
                                     c_dp, c_dp, c_u64p, c_u64p, c_i32p]),
    ("nbs_keys", ctypes.c_int, [ctypes.c_void_p, c_u64p, c_u64p]),
    ("nbs_roots", ctypes.c_int, [ctypes.c_void_p, c_dp, c_dp, ctypes.c_int]),
    ("nbs_adjoint_forward", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]),
    ("nbs_adjoint_backward", ctypes.c_int, [ctypes.c_void_p, c_dp, c_dp, c_dp, c_dp, c_dp]),
    ("nbs_adjoint_backward_device", ctypes.c_int, [ctypes.c_void_p] * 6),
    ("nbs_frozen_forward", ctypes.c_int, [ctypes.c_void_p, c_dp, c_dp, c_dp, c_dp, c_dp]),
    ("nbs_profile_enable", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("nbs_profile_read", ctypes.c_int, [ctypes.c_void_p, c_dp, ctypes.c_int]),
    ("nbs_stream", ctypes.c_void_p, [ctypes.c_void_p]),
    # nbs_host.h
    ("nbs_settings_default", None, [ctypes.POINTER(Settings)]),
    ("nbs_settings_read", ctypes.c_int, [ctypes.POINTER(Settings), ctypes.c_char_p,
                                         ctypes.c_char_p, ctypes.c_size_t]),
    ("nbs_generate", ctypes.c_int, [ctypes.POINTER(Settings), ctypes.c_ulong, ctypes.c_size_t,
                                    ctypes.c_int, c_dp, c_dp, c_dp]),
    ("nbs_generate_state", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Settings),
                                          ctypes.c_int, c_i64p, c_u64p,
                                          ctypes.POINTER(ctypes.c_uint32)]),
    ("nbs_frame_filename", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint, ctypes.c_int,
                                          ctypes.c_char_p, ctypes.c_size_t]),
    ("nbs_write_binary", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, c_dp, c_dp, c_dp,
                                        c_dp, ctypes.c_int]),
    ("nbs_write_text", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, c_dp]),
    ("nbs_binary_count", ctypes.c_int64, [ctypes.c_char_p, ctypes.c_int]),
    ("nbs_read_binary", ctypes.c_int64, [ctypes.c_char_p, ctypes.c_int, ctypes.c_size_t, c_dp,
                                         c_dp, c_dp, c_dp]),
    ("nbs_resume_frame", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint]),
    ("nbs_run_frames", ctypes.c_int, [ctypes.c_char_p]),
    ("nbs_write_png", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_uint8)]),
]

NBS_NOALLOC_MAX_N = 65536  # include/nbs.h

_lib = None


def lib():
    """Loads libnbs.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libnbs.so not built: run `python -m nbodysim_amd.build` "
                              "(expected at %s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
