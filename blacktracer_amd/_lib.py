"""ctypes bindings of libbtamd.so (the C ABI declared in include/bt_abi.h).

The library is built in-tree by ``__graft_entry__.build()`` (make -C
blacktracer_amd/csrc).  There is no fallback: if the shared object is missing
or a HIP device is absent, the calls raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbtamd.so")

DP = C.POINTER(C.c_double)
FP = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int32)
I8P = C.POINTER(C.c_int8)
U8P = C.POINTER(C.c_uint8)
I64P = C.POINTER(C.c_int64)

BT_OK = 0
BT_ERR_INVALID = -1
BT_ERR_DEVICE = -2
BT_ERR_HIP = -3
BT_ERR_NOMEM = -4
BT_ERR_BUDGET = -5
BT_ERR_NOTFOUND = -6
BT_ERR_STATE = -7

BT_PREC_FP64 = 0
BT_PREC_MIXED = 1


class BtPose(C.Structure):
    _fields_ = [("a", C.c_double), ("theta", C.c_double), ("phi", C.c_double), ("r", C.c_double),
                ("br", C.c_double), ("btheta", C.c_double), ("bphi", C.c_double),
                ("speed", C.c_double), ("user_speed", C.c_int)]


class BtIntegParams(C.Structure):
    _fields_ = [("eps", C.c_double), ("h1", C.c_double), ("z_end", C.c_double),
                ("max_steps", C.c_int), ("apriori_capture", C.c_int),
                ("r_escape", C.c_double), ("r_capture_factor", C.c_double),
                ("newton_cap", C.c_int), ("precision", C.c_int),
                ("crit_eps", C.c_double), ("hmin_fp64", C.c_double),
                ("noise_budget", C.c_double), ("decision_margin", C.c_double),
                ("drift_budget", C.c_double), ("polar_b", C.c_double)]


class BtTraceStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("rays", "integrated", "attempts", "attempts_fp64",
                                         "newton", "golden", "fallback_rays")]


class BtGridParams(C.Structure):
    _fields_ = [("refine_thresh", C.c_double), ("fix_tvertices", C.c_int)]


class BtRenderFlags(C.Structure):
    _fields_ = [("star", C.c_int), ("map", C.c_int), ("lr", C.c_int), ("hor", C.c_float),
                ("ver", C.c_float), ("offset", C.c_float)]


class BtServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("slots", C.c_int), ("max_level", C.c_int),
                ("start_level", C.c_int), ("integ", BtIntegParams), ("step", C.c_int),
                ("view_angle", C.c_double), ("off_up", C.c_double), ("off_right", C.c_double)]


class BtServerBytes(C.Structure):
    _fields_ = [("slot", C.c_size_t), ("view", C.c_size_t), ("grid", C.c_size_t), ("build", C.c_size_t)]


class BtRequest(C.Structure):
    _fields_ = [("a", C.c_double), ("r", C.c_double), ("theta", C.c_double), ("offset", C.c_float),
                ("width", C.c_int), ("height", C.c_int), ("star", C.c_int), ("map", C.c_int),
                ("lr", C.c_int), ("out_bgra", C.c_void_p), ("tag", C.c_uint64)]


class BtCompletion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int), ("grid_hit", C.c_int),
                ("latency_ms", C.c_double), ("queue_ms", C.c_double), ("frame_ms", C.c_double)]


class BtServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("grid_builds", C.c_int64), ("grid_hits", C.c_int64), ("evictions", C.c_int64),
                ("peak_inflight", C.c_int), ("peak_bytes", C.c_size_t), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("frames_per_s", C.c_double),
                ("wall_s", C.c_double)]


class BtTune(C.Structure):
    _fields_ = [("chunk", C.c_int), ("block", C.c_int), ("waves", C.c_int),
                ("rays_per_lane", C.c_double)]


# name -> (restype, argtypes); every symbol here is declared in include/bt_abi.h
SIGNATURES = {
    "bt_frame_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "bt_scene_create_tree": (C.c_int, [C.c_void_p, I32P, C.c_int, FP, FP, C.c_int, U8P, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    "bt_star_catalog_parse": (C.c_int, [C.c_char_p, C.c_size_t, C.c_float, FP, C.c_int, C.POINTER(C.c_int)]),
    "bt_star_file_write": (C.c_int, [C.c_char_p, I32P, C.c_int, FP, FP, C.c_int]),
    "bt_star_file_read": (C.c_int, [C.c_char_p, I32P, FP, FP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bt_grid_file_write": (C.c_int, [C.c_char_p, C.c_void_p]),
    "bt_grid_file_read": (C.c_int, [C.c_char_p] + [C.POINTER(C.c_int)] * 5 + [FP, I32P, I32P]),
    "bt_selfcheck_division": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.POINTER(C.c_int64)]),
    "bt_tune_get": (C.c_int, [C.c_void_p, C.POINTER(BtTune)]),
    "bt_tune_set": (C.c_int, [C.c_void_p, C.POINTER(BtTune)]),
    "bt_autotune": (C.c_int, [C.c_void_p, C.POINTER(BtPose), C.POINTER(BtIntegParams), C.c_int64,
                              C.POINTER(BtTune), DP]),
    "bt_server_config_default": (None, [C.POINTER(BtServerConfig)]),
    "bt_server_footprint": (C.c_int, [C.POINTER(BtServerConfig), C.c_int, C.c_int, C.POINTER(BtServerBytes)]),
    "bt_server_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BtServerConfig), C.POINTER(C.c_void_p)]),
    "bt_server_submit": (C.c_int, [C.c_void_p, C.POINTER(BtRequest)]),
    "bt_server_poll": (C.c_int, [C.c_void_p, C.POINTER(BtCompletion), C.c_int, C.POINTER(C.c_int)]),
    "bt_server_drain": (C.c_int, [C.c_void_p, C.c_double]),
    "bt_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(BtServerStats)]),
    "bt_server_destroy": (None, [C.c_void_p]),
    "bt_last_error": (C.c_char_p, []),
    "bt_abi_version": (C.c_int, []),
    "bt_context_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "bt_context_destroy": (None, [C.c_void_p]),
    "bt_context_sync": (C.c_int, [C.c_void_p]),
    "bt_timing_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "bt_timer_get": (C.c_int, [C.c_void_p, C.c_char_p, I64P, DP]),
    "bt_timer_reset": (C.c_int, [C.c_void_p]),
    "bt_device_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "bt_device_free": (None, [C.c_void_p, C.c_void_p]),
    "bt_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "bt_camera_init": (C.c_int, [C.POINTER(BtPose), DP, FP]),
    "bt_integ_params_default": (None, [C.POINTER(BtIntegParams)]),
    "bt_trace_rays": (C.c_int, [C.c_void_p, C.POINTER(BtPose), C.POINTER(BtIntegParams), C.c_int64,
                                DP, DP, I32P, C.POINTER(BtTraceStats)]),
    "bt_trace_rays_device": (C.c_int, [C.c_void_p, C.POINTER(BtPose), C.POINTER(BtIntegParams),
                                       C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(BtTraceStats)]),
    "bt_trace_rays_device_noalloc": (C.c_int, [C.c_void_p, C.POINTER(BtPose), C.POINTER(BtIntegParams),
                                               C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bt_grid_params_default": (None, [C.POINTER(BtGridParams)]),
    "bt_grid_build": (C.c_int, [C.c_void_p, C.POINTER(BtPose), C.POINTER(BtIntegParams), C.c_int,
                                C.c_int, C.c_int, C.POINTER(BtGridParams), C.POINTER(C.c_void_p)]),
    "bt_grid_build_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BtPose), C.POINTER(BtIntegParams),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                      C.POINTER(BtGridParams), C.POINTER(C.c_void_p),
                                      C.POINTER(BtTraceStats)]),
    "bt_grid_info": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 5 + [I64P]),
    "bt_grid_stats": (C.c_int, [C.c_void_p, C.POINTER(BtTraceStats)]),
    "bt_grid_export": (C.c_int, [C.c_void_p, DP, I32P, I8P]),
    "bt_grid_export_psh": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), FP, I32P,
                                     I32P]),
    "bt_grid_free": (None, [C.c_void_p]),
    "bt_render_flags_default": (None, [C.POINTER(BtRenderFlags)]),
    "bt_make_image": (C.c_int, [C.c_void_p, FP, I32P, C.c_int, FP, FP, C.c_int, C.c_int, C.c_int,
                                C.c_int, U8P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, FP,
                                FP, C.c_float, C.c_int, C.c_int, C.c_int, I32P, FP, I32P, I32P,
                                C.c_int, C.c_int, C.POINTER(BtRenderFlags), U8P]),
    "bt_viewer_matrix": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, FP]),
    "bt_star_tree": (C.c_int, [FP, C.c_int, C.c_int, I32P, FP, FP]),
    "bt_sky_add_low_stars": (C.c_int, [U8P, C.c_int, C.c_int, FP, C.c_int, C.c_float]),
    "bt_scene_create": (C.c_int, [C.c_void_p, FP, C.c_int, C.c_int, U8P, C.c_int, C.c_int,
                                  C.POINTER(C.c_void_p)]),
    "bt_scene_export_table": (C.c_int, [C.c_void_p, FP]),
    "bt_scene_free": (None, [C.c_void_p]),
    "bt_view_create": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_int, C.POINTER(C.c_void_p)]),
    "bt_view_free": (None, [C.c_void_p]),
    "bt_render_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, FP,
                                  C.POINTER(BtRenderFlags), C.c_int, U8P, C.c_void_p]),
    "bt_render_frame_stages": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, FP,
                                         C.POINTER(BtRenderFlags), C.c_int, U8P, C.c_void_p, FP,
                                         I32P, U8P, FP, FP]),
    "bt_trails_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, U8P, C.POINTER(C.c_void_p)]),
    "bt_trails_reset": (None, [C.c_void_p]),
    "bt_trails_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bt_trails_free": (None, [C.c_void_p]),
    "bt_render_frame_trails": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, FP,
                                         C.POINTER(BtRenderFlags), C.c_int, C.c_void_p, U8P,
                                         C.c_void_p, FP, FP]),
    "bt_render_frame_blend": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p,
                                        C.c_void_p, FP, C.POINTER(BtRenderFlags), C.c_int, U8P,
                                        C.c_void_p, FP, FP]),
}

_lib = None


def load():
    """Load libbtamd.so and bind every declared symbol; raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"blacktracer_amd native library missing: {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class BtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"blacktracer_amd error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != BT_OK:
        msg = load().bt_last_error()
        raise BtError(rc, msg.decode() if msg else "")
    return rc
