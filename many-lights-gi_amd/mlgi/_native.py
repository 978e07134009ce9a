"""ctypes binding of libmlgi.so (include/mlgi.h, include/mlgi_scene.h).

The library is loaded from this directory; if it is missing, importing this
module raises, so no caller can silently fall back to another path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmlgi.so"

MLGI_OK = 0
MLGI_ERR_INVALID_ARGUMENT = 1
MLGI_ERR_OUT_OF_MEMORY = 2
MLGI_ERR_HIP = 3
MLGI_ERR_UNSUPPORTED = 4
MLGI_ERR_NO_DEVICE = 5
MLGI_ERR_OUT_OF_BUDGET = 6

PRECISION_FP32 = 0
PRECISION_MIXED = 1

VPL_RECORD_FLOATS = 32
NOALLOC_MAX_POINTS = 8388608


class Params(C.Structure):
    _fields_ = [
        ("light_intensity", C.c_float),
        ("shuffle_lights", C.c_int32),
        ("gi_intensity_factor", C.c_float),
        ("vpl_clamping_value", C.c_float),
        ("vpl_start_index", C.c_int32),
        ("vpl_end_index", C.c_int32),
        ("scale_isms", C.c_int32),
        ("points_only_into_scaled_isms", C.c_int32),
        ("tess_level_factor", C.c_float),
        ("use_push_pull", C.c_int32),
        ("enable_shadowing", C.c_int32),
        ("use_interleaving", C.c_int32),
        ("interleave", C.c_int32),
        ("z_far", C.c_float),
        ("atlas_side", C.c_int32),
        ("pull_push_levels", C.c_int32),
        ("precision", C.c_int32),
        ("tile_w", C.c_int32),
        ("tile_h", C.c_int32),
        ("gather_fetch", C.c_int32),
        ("splat_block", C.c_int32),
        ("disable_culling", C.c_int32),
    ]


class Rsm(C.Structure):
    _fields_ = [
        ("diffuse", C.c_void_p),
        ("face_normal", C.c_void_p),
        ("depth", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("inv_biased_view_projection", C.c_float * 16),
    ]


class Vpls(C.Structure):
    _fields_ = [("records", C.c_void_p), ("count", C.c_int32)]


class Points(C.Structure):
    _fields_ = [
        ("records", C.c_void_p),
        ("bin_offsets", C.c_void_p),
        ("count", C.c_int32),
        ("n_bins", C.c_int32),
    ]


class Atlas(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("side", C.c_int32), ("ism_side", C.c_int32)]


class IsmDepth(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("side", C.c_int32), ("ism_side", C.c_int32)]


class GBuffer(C.Structure):
    _fields_ = [
        ("depth", C.c_void_p),
        ("face_normal", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("view_projection_inverse", C.c_float * 16),
    ]


class Image(C.Structure):
    _fields_ = [("rgb", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32)]


class CompositeInputs(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("normal", C.c_void_p), ("diffuse", C.c_void_p), ("specular", C.c_void_p),
                ("gi", C.c_void_p), ("occlusion", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("vsm", C.c_void_p), ("vsm_width", C.c_int32), ("vsm_height", C.c_int32)]


class CompositeParams(C.Structure):
    _fields_ = [("view_projection_inverse", C.c_float * 16), ("biased_light_view_projection", C.c_float * 16),
                ("camera_eye", C.c_float * 3), ("light_position", C.c_float * 3),
                ("inverse_light_direction", C.c_float * 3), ("light_intensity", C.c_float),
                ("exposure", C.c_float)]


class FrameDesc(C.Structure):
    _fields_ = [("rsm", Rsm), ("points", Points), ("gbuffer", GBuffer), ("params", Params),
                ("vpls", Vpls), ("atlas", Atlas), ("ism", IsmDepth), ("image", Image)]


class TuneResult(C.Structure):
    _fields_ = [("interleave", C.c_int32), ("tile_w", C.c_int32), ("tile_h", C.c_int32),
                ("gather_ms", C.c_float * 3), ("cached", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_inflight", C.c_int32), ("max_batch", C.c_int32),
                ("max_queued_batches", C.c_int32), ("rsm_size", C.c_int32), ("precision", C.c_int32),
                ("min_vpls_per_pixel", C.c_int32), ("vpl_counts", C.POINTER(C.c_int32)),
                ("n_vpl_counts", C.c_int32), ("resolutions", C.POINTER(C.c_int32)), ("n_resolutions", C.c_int32),
                ("streams", C.c_int32)]


class ServerScene(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("n_points", C.c_int64), ("point_seed", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("n_vpl", C.c_int32),
                ("light_seed", C.c_uint64), ("arrival_s", C.c_double), ("host_out", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("rejected", C.c_int64), ("wall_s", C.c_double),
                ("req_per_s", C.c_double), ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_mean_ms", C.c_double), ("latency_max_ms", C.c_double), ("batches", C.c_int32),
                ("max_inflight_seen", C.c_int32), ("mean_batch", C.c_double), ("peak_bytes", C.c_size_t),
                ("overlapped_batches", C.c_int32), ("streams", C.c_int32)]


class StageTimes(C.Structure):
    _fields_ = [
        ("vpl_ms", C.c_float),
        ("ism_clear_ms", C.c_float),
        ("ism_splat_ms", C.c_float),
        ("pullpush_ms", C.c_float),
        ("final_gather_ms", C.c_float),
        ("frame_ms", C.c_float),
        ("frames", C.c_int32),
        ("blur_ms", C.c_float),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("eye", C.c_float * 3),
        ("center", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fovy_deg", C.c_float),
        ("aspect", C.c_float),
        ("z_near", C.c_float),
        ("z_far", C.c_float),
    ]


P = C.POINTER
vp = C.c_void_p
i32 = C.c_int32

# name -> (restype, argtypes)
SIGNATURES = {
    "mlgi_version": (i32, []),
    "mlgi_device_count": (C.c_int, [P(i32)]),
    "mlgi_context_create": (C.c_int, [i32, P(vp)]),
    "mlgi_context_destroy": (None, [vp]),
    "mlgi_last_error": (C.c_char_p, [vp]),
    "mlgi_sync": (C.c_int, [vp]),
    "mlgi_stream": (vp, [vp]),
    "mlgi_context_wait": (C.c_int, [vp, vp]),
    "mlgi_params_default": (None, [P(Params)]),
    "mlgi_malloc": (C.c_int, [vp, C.c_size_t, P(vp)]),
    "mlgi_free": (C.c_int, [vp, vp]),
    "mlgi_memcpy_htod": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "mlgi_memcpy_dtoh": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "mlgi_memcpy_dtod": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "mlgi_vpl_extract": (C.c_int, [vp, P(Rsm), P(Params), P(Vpls)]),
    "mlgi_ism_splat": (C.c_int, [vp, P(Points), P(Vpls), P(Params), P(Atlas)]),
    "mlgi_ism_splat_noalloc": (C.c_int, [vp, P(Points), P(Vpls), P(Params), P(Atlas)]),
    "mlgi_splat_spill_stats": (C.c_int, [vp, P(C.c_uint32), P(C.c_int32)]),
    "mlgi_splat_count": (C.c_int, [vp, P(C.c_uint64)]),
    "mlgi_probe_valu": (C.c_int, [vp, P(C.c_double)]),
    "mlgi_check_division": (C.c_int, [vp, C.c_uint32, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]),
    "mlgi_deferred_composite": (C.c_int, [vp, P(CompositeInputs), P(CompositeParams), P(Image)]),
    "mlgi_ism_pullpush": (C.c_int, [vp, P(Atlas), P(Params), P(IsmDepth)]),
    "mlgi_ism_process": (C.c_int, [vp, P(Points), P(Vpls), P(Params), P(Atlas), P(IsmDepth)]),
    "mlgi_final_gather": (C.c_int, [vp, P(GBuffer), P(Vpls), P(IsmDepth), P(Params), P(Image)]),
    "mlgi_frame": (C.c_int, [vp, P(Rsm), P(Points), P(GBuffer), P(Params), P(Vpls), P(Atlas),
                             P(IsmDepth), P(Image)]),
    "mlgi_count_lit_pairs": (C.c_int, [vp, P(GBuffer), P(Vpls), P(Params), P(C.c_uint64)]),
    "mlgi_graph_capture_frame": (C.c_int, [vp, P(Rsm), P(Points), P(GBuffer), P(Params), P(Vpls),
                                           P(Atlas), P(IsmDepth), P(Image), P(vp)]),
    "mlgi_graph_launch": (C.c_int, [vp, vp]),
    "mlgi_gi_blur": (C.c_int, [vp, P(GBuffer), vp, P(Image), P(Image)]),
    "mlgi_batch_create": (C.c_int, [vp, P(FrameDesc), C.c_int32, P(vp)]),
    "mlgi_batch_run": (C.c_int, [vp, vp]),
    "mlgi_batch_capture": (C.c_int, [vp, vp, P(vp)]),
    "mlgi_batch_destroy": (None, [vp]),
    "mlgi_batch_update_light": (C.c_int, [vp, vp, C.c_int32, vp]),
    "mlgi_autotune": (C.c_int, [vp, P(FrameDesc), C.c_int32, P(TuneResult)]),
    "mlgi_server_create": (C.c_int, [vp, P(ServerConfig), P(ServerScene), C.c_int32, P(vp)]),
    "mlgi_server_run": (C.c_int, [vp, P(Request), C.c_int32, P(ServerStats), P(C.c_double)]),
    "mlgi_server_destroy": (None, [vp]),
    "mlgi_graph_destroy": (None, [vp]),
    "mlgi_profile_enable": (C.c_int, [vp, i32]),
    "mlgi_profile_reset": (C.c_int, [vp]),
    "mlgi_profile_read": (C.c_int, [vp, P(StageTimes)]),
    "mlgi_shuffle_indices": (C.c_int, [i32, C.c_uint32, vp]),
    "mlgi_ism_indices_1d": (i32, [i32]),
    # scene tooling
    "mlgi_scene_create": (C.c_int, [C.c_uint64, i32, P(vp)]),
    "mlgi_scene_destroy": (None, [vp]),
    "mlgi_scene_triangle_count": (i32, [vp]),
    "mlgi_scene_total_area": (C.c_double, [vp]),
    "mlgi_scene_triangles": (C.c_int, [vp, vp, vp, vp]),
    "mlgi_scene_default_camera": (None, [C.c_float, P(Camera)]),
    "mlgi_scene_light_pose": (None, [C.c_uint64, P(Camera)]),
    "mlgi_camera_matrices": (C.c_int, [P(Camera), vp, vp, vp]),
    "mlgi_scene_render_gbuffer": (C.c_int, [vp, P(Camera), i32, i32, i32, vp, vp]),
    "mlgi_scene_render_rsm": (C.c_int, [vp, P(Camera), i32, i32, i32, vp, vp, vp, vp]),
    "mlgi_scene_render_surface": (C.c_int, [vp, P(Camera), i32, i32, i32, vp, vp]),
    "mlgi_scene_device_render_surface": (C.c_int, [vp, vp, P(Camera), i32, i32, vp, vp]),
    "mlgi_scene_upload": (C.c_int, [vp, vp, P(vp)]),
    "mlgi_scene_device_destroy": (None, [vp]),
    "mlgi_scene_device_render_gbuffer": (C.c_int, [vp, vp, P(Camera), i32, i32, vp, vp, vp]),
    "mlgi_scene_device_render_rsm": (C.c_int, [vp, vp, P(Camera), i32, i32, vp, vp, vp, vp]),
    "mlgi_scene_device_sample_points": (C.c_int, [vp, vp, C.c_int64, i32, i32, C.c_uint64, vp, vp]),
    "mlgi_scene_sample_points": (C.c_int, [vp, C.c_int64, i32, i32, C.c_uint64, i32, vp, vp]),
}


def load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python many-lights-gi_amd/build.py` "
            "(there is no non-HIP fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


class MlgiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mlgi error {code}: {msg}")
        self.code = code


def exported_symbols() -> list[str]:
    return list(SIGNATURES)
