"""ctypes binding of libpsdr_hip.so (the C ABI declared in include/psdr_hip.h).

The library is built in-tree by torch_psdr_amd.build; importing this module
never builds it.  Loading fails loudly if the library is missing: there is no
CPU fallback for the render path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpsdr_hip.so"

PSDR_OK = 0
PSDR_ERR_ARG = -1
PSDR_ERR_HIP = -2
PSDR_ERR_STATE = -3
PSDR_ERR_NOMEM = -4
PSDR_ERR_CAPTURE = -5
PSDR_MAT_DIFFUSE = 0
PSDR_MAT_DIFFUSE_BSDF = 1
PSDR_MAT_LIGHT = 2
PSDR_MAT_ENV_MAP = 3
PSDR_MAT_ENV_SH9 = 4

PSDR_GRAD_VERTICES = 1
PSDR_GRAD_ALBEDO = 2
PSDR_GRAD_ROUGHNESS = 4
PSDR_GRAD_EMIT = 8
PSDR_GRAD_NORMAL = 16

PSDR_TERM_INTERIOR = 1
PSDR_TERM_PRIMARY = 2
PSDR_TERM_DIRECT = 4
PSDR_TERM_ALL = 7
PSDR_CAMERA_PERSPECTIVE = 0
PSDR_CAMERA_ORTHOGRAPHIC = 1
PSDR_CAMERA_GRAD = 10
PSDR_NOALLOC_MAX_RAYS = 715827882
PSDR_MAX_TRIS = 268435456
PSDR_TRACE_COUNTERS = 8
PSDR_MAX_CAMERAS = 4
PSDR_MIN_TILE_LANES = 65536

ABI_VERSION = 7


class PsdrObject(C.Structure):
    _fields_ = [
        ("material", C.c_int32),
        ("light", C.c_int32),
        ("albedo_offset", C.c_int32),
        ("albedo_h", C.c_int32),
        ("albedo_w", C.c_int32),
        ("rough_offset", C.c_int32),
        ("rough_h", C.c_int32),
        ("rough_w", C.c_int32),
        ("tri_begin", C.c_int32),
        ("tri_count", C.c_int32),
        ("grad_mask", C.c_int32),
        ("normal_offset", C.c_int32),
        ("normal_h", C.c_int32),
        ("normal_w", C.c_int32),
    ]


class PsdrSceneDesc(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int32),
        ("vertices", C.c_void_p),
        ("n_tris", C.c_int32),
        ("faces", C.c_void_p),
        ("tri_uv", C.c_void_p),
        ("tri_obj", C.c_void_p),
        ("n_objects", C.c_int32),
        ("objects", C.POINTER(PsdrObject)),
        ("n_texels", C.c_int32),
        ("texels", C.c_void_p),
        ("n_lights", C.c_int32),
        ("light_emit", C.c_void_p),
        ("light_obj", C.POINTER(C.c_int32)),
    ]


class PsdrConfig(C.Structure):
    _fields_ = [
        ("n_pass", C.c_uint32),
        ("spp_interior", C.c_uint32),
        ("enable_light_visable", C.c_uint32),
        ("spp_primary_edge", C.c_uint32),
        ("spp_secondary_edge", C.c_uint32),
        ("max_bounce", C.c_uint32),
        ("mis_light_samples", C.c_uint32),
        ("mis_bsdf_samples", C.c_uint32),
        ("primary_edge_preprocess_rounds", C.c_uint32),
        ("secondary_edge_preprocess_rounds", C.c_uint32),
        ("primary_edge_hypercube_resolution", C.c_uint32),
        ("secondary_edge_hypercube_resolution", C.c_uint32 * 3),
    ]


class PsdrCamera(C.Structure):
    _fields_ = [
        ("look_from", C.c_float * 3),
        ("look_at", C.c_float * 3),
        ("up", C.c_float * 3),
        ("vfov", C.c_float),
        ("height", C.c_int32),
        ("width", C.c_int32),
        ("kind", C.c_int32),
        ("view_width", C.c_float),
    ]


class PsdrTuning(C.Structure):
    """Launch shapes (include/psdr_hip.h psdr_tuning)."""
    _fields_ = [
        ("trace_blocks_per_cu", C.c_int32),
        ("trace_stack_lds", C.c_int32),
        ("shade_blocks_per_cu", C.c_int32),
        ("adjoint_blocks_per_cu", C.c_int32),
        ("direct_blocks_per_cu", C.c_int32),
        ("tile_lanes", C.c_int32),
    ]


_VP = C.c_void_p
_I32 = C.c_int32
_SIGS = {
    "psdr_last_error": (C.c_char_p, []),
    "psdr_abi_version": (C.c_int, []),
    "psdr_scene_create": (C.c_int, [C.POINTER(PsdrSceneDesc), C.POINTER(_VP), _VP]),
    "psdr_scene_update": (C.c_int, [_VP, C.POINTER(PsdrSceneDesc), _VP]),
    "psdr_scene_destroy": (C.c_int, [_VP]),
    "psdr_scene_join": (C.c_int, [_VP, _VP]),
    "psdr_scene_share_workspace": (C.c_int, [_VP, _VP, _VP]),
    "psdr_scene_info": (C.c_int, [_VP, C.POINTER(_I32), C.POINTER(_I32), C.POINTER(_I32)]),
    "psdr_render_forward": (C.c_int, [_VP, C.POINTER(PsdrConfig), C.POINTER(PsdrCamera), _VP, _I32, _VP]),
    "psdr_render_forward_multi": (C.c_int, [_VP, C.POINTER(PsdrConfig), C.POINTER(PsdrCamera), _I32,
                                            C.POINTER(C.c_void_p), _I32, _VP]),
    "psdr_render_backward": (C.c_int, [_VP, C.POINTER(PsdrConfig), C.POINTER(PsdrCamera), _VP, _VP, _VP, _VP, _VP,
                                       _I32, _VP]),
    "psdr_render_forward_plain": (C.c_int, [_VP, C.POINTER(PsdrConfig), C.POINTER(PsdrCamera), _VP, _I32, _VP]),
    "psdr_render_backward_plain": (C.c_int, [_VP, C.POINTER(PsdrConfig), C.POINTER(PsdrCamera), _VP, _VP, _VP, _VP,
                                             _VP, _I32, _VP]),
    "psdr_pcg_lane_draws": (C.c_int, [C.c_uint32, _I32, _I32, _VP, _VP]),
    "psdr_camera_rays": (C.c_int, [C.POINTER(PsdrCamera), C.c_uint32, _VP, _VP, _VP, _VP]),
    "psdr_camera_matrices": (C.c_int, [C.POINTER(PsdrCamera), _VP, _VP, _VP, _VP, _VP]),
    "psdr_camera_param_vjp": (C.c_int, [C.POINTER(PsdrCamera), _VP, _VP]),
    "psdr_scene_trace": (C.c_int, [_VP, _VP, _VP, _I32, _VP, _VP, _VP, _VP, _VP]),
    "psdr_scene_trace_noalloc": (C.c_int, [_VP, _VP, _VP, _I32, _VP, _VP, _VP, _VP, _VP]),
    "psdr_scene_derived": (C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "psdr_scene_samplers": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "psdr_scene_bvh": (C.c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "psdr_last_lane_radiance": (C.c_int, [_VP, _VP, _VP]),
    "psdr_last_render_stats": (C.c_int, [_VP, C.POINTER(C.c_int64), _I32, C.POINTER(C.c_int64)]),
    "psdr_trace_timing": (C.c_int, [_I32, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_double)]),
    "psdr_trace_counters": (C.c_int, [_I32, C.POINTER(C.c_int64), _I32]),
    "psdr_get_tuning": (C.c_int, [C.POINTER(PsdrTuning)]),
    "psdr_set_tuning": (C.c_int, [C.POINTER(PsdrTuning)]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load() -> C.CDLL:
    """Load libpsdr_hip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m torch_psdr_amd.build` "
            "(the render path has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.psdr_abi_version() != ABI_VERSION:
        raise RuntimeError("libpsdr_hip.so ABI version mismatch; rebuild it")
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().psdr_last_error().decode(errors="replace")
        raise RuntimeError(f"psdr error {rc}: {msg}")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    return None if t is None else t.data_ptr()


def current_stream() -> int:
    import torch
    return torch.cuda.current_stream().cuda_stream
