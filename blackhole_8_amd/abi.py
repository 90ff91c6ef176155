"""ctypes mirror of include/bh_api.h and the loader of the HIP library.

The structures below must match the C layout byte for byte; the ABI test
checks their sizes against the library's own view via offsets of known fields.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libbh8.so"

BH_MAX_SPHERES = 32
BH_MAX_TEXTURES = 64
BH_MAX_RECTS = 8
BH_MAX_PLANES = 4

BH_OK = 0
BH_ERR_INVALID_ARG = -1
BH_ERR_NO_DEVICE = -2
BH_ERR_HIP = -3
BH_ERR_OUT_OF_MEMORY = -4
BH_ERR_EMPTY_TEXTURE = -5
BH_ERR_UNSUPPORTED = -6
BH_ERR_BAD_HANDLE = -7

HIT_NONE, HIT_HORIZON, HIT_DISC, HIT_OBJECT = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1
FILTER_NEAREST, FILTER_BILINEAR = 0, 1
BATCH_GRID, BATCH_TILES, BATCH_LANES = 1, 2, 3
ROOT_BISECT20, ROOT_EXACT = 0, 1

d3 = C.c_double * 3


class Image(C.Structure):
    _fields_ = [("data", C.c_void_p), ("rows", C.c_int32), ("cols", C.c_int32),
                ("channels", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("position", d3), ("axis_x", d3), ("axis_y", d3), ("axis_z", d3),
                ("width", C.c_int32), ("height", C.c_int32), ("fov", C.c_double)]


class Hole(C.Structure):
    _fields_ = [("position", d3), ("mass", C.c_double)]


class Annulus(C.Structure):
    _fields_ = [("vertex", d3 * 5), ("normal", d3), ("r_outer", C.c_double),
                ("r_inner", C.c_double), ("texture", C.c_int32), ("pad_", C.c_int32)]


class Sphere(C.Structure):
    _fields_ = [("center", d3), ("radius", C.c_double), ("texture", C.c_int32),
                ("pad_", C.c_int32)]


class Scene(C.Structure):
    _fields_ = [("camera", Camera), ("hole", Hole), ("has_disc", C.c_int32),
                ("n_spheres", C.c_int32), ("disc", Annulus),
                ("spheres", Sphere * BH_MAX_SPHERES), ("background", C.c_int32),
                ("pad_", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("nstep", C.c_int32), ("escape", C.c_int32),
                ("step_cap", C.c_int32), ("r_escape", C.c_double), ("filter", C.c_int32),
                ("root", C.c_int32), ("band", C.c_double), ("edge_width", C.c_double),
                ("edge_guard", C.c_double)]


class Outputs(C.Structure):
    _fields_ = [("rgb", C.c_void_p), ("hit", C.c_void_p), ("phi_end", C.c_void_p),
                ("hit_point", C.c_void_p), ("steps", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block", C.c_int32), ("waves_per_simd", C.c_int32),
                ("refill_steps", C.c_int32), ("refill_idle", C.c_int32), ("chunk", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("march_ms", C.c_double), ("band_ms", C.c_double),
                ("rays", C.c_int64), ("band_rays", C.c_int64), ("solo_scenes", C.c_int32),
                ("pad_", C.c_int32)]


class DiffParams(C.Structure):
    _fields_ = [("r_s", C.c_double), ("incl", C.c_double), ("dist", C.c_double),
                ("r_inner", C.c_double), ("r_outer", C.c_double), ("azim", C.c_double),
                ("fov", C.c_double), ("edge_width", C.c_double), ("width", C.c_int32),
                ("height", C.c_int32), ("texture", C.c_int32), ("pad_", C.c_int32)]


class Rect(C.Structure):
    _fields_ = [("vertex", d3 * 5), ("texture", C.c_int32), ("pad_", C.c_int32)]


class Plane(C.Structure):
    _fields_ = [("position", d3), ("axis_x", d3), ("axis_y", d3), ("axis_z", d3),
                ("pattern_size", C.c_double)]


class FlatScene(C.Structure):
    _fields_ = [("camera", Camera), ("n_rects", C.c_int32), ("n_planes", C.c_int32),
                ("rects", Rect * BH_MAX_RECTS), ("planes", Plane * BH_MAX_PLANES),
                ("steps", C.c_int32), ("pad_", C.c_int32)]


# Every symbol include/bh_api.h declares, with its ctypes signature.
P = C.c_void_p
SIGNATURES = {
    "bh_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "bh_ctx_destroy": (None, [P]),
    "bh_last_error": (C.c_char_p, [P]),
    "bh_version": (C.c_int, []),
    "bh_device_info": (C.c_int, [P, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "bh_ctx_set_stream": (C.c_int, [P, P]),
    "bh_synchronize": (C.c_int, [P]),
    "bh_texture_set": (C.c_int, [P, C.c_int, C.POINTER(Image)]),
    "bh_texture_clear": (C.c_int, [P, C.c_int]),
    "bh_scene_check": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options)]),
    "bh_render": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options), C.POINTER(Outputs)]),
    "bh_render_device": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options), C.POINTER(Outputs)]),
    "bh_render_device_noalloc": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options),
                                           C.POINTER(Outputs)]),
    "bh_fma_issue_rate": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bh_chord_sincos": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64]),
    "bh_turning_points": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options),
                                    C.POINTER(C.c_double)]),
    "bh_path_last_ms": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bh_path_create": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options), C.POINTER(Camera),
                                 C.c_int, C.POINTER(P)]),
    "bh_path_launch": (C.c_int, [P, C.POINTER(C.c_float)]),
    "bh_path_read": (C.c_int, [P, C.c_int, P, P]),
    "bh_path_time": (C.c_int, [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bh_path_destroy": (None, [P]),
    "bh_time_render": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options), C.c_int,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int64)]),
    "bh_render_batch_device": (C.c_int, [P, C.c_int, C.POINTER(Scene), C.POINTER(Options),
                                         C.POINTER(Outputs), C.POINTER(BatchTuning),
                                         C.POINTER(C.c_double), C.POINTER(BatchStats)]),
    "bh_batch_autotune": (C.c_int, [P, C.c_int, C.POINTER(Scene), C.POINTER(Options),
                                    C.POINTER(Outputs), C.POINTER(BatchTuning), C.c_int, C.c_int,
                                    C.POINTER(BatchTuning), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "bh_diff_render_device": (C.c_int, [P, C.POINTER(DiffParams), C.c_int, P, P]),
    "bh_diff_adjoint_device": (C.c_int, [P, C.POINTER(DiffParams), C.c_int, P, P, P, P, P]),
    "bh_diff_time": (C.c_int, [P, C.POINTER(DiffParams), C.c_int, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bh_render_flat": (C.c_int, [P, C.POINTER(FlatScene), C.POINTER(Outputs)]),
    "bh_render_flat_device": (C.c_int, [P, C.POINTER(FlatScene), C.POINTER(Outputs)]),
    "bh_time_flat": (C.c_int, [P, C.POINTER(FlatScene), C.c_int, C.POINTER(C.c_double)]),
    "bh_ring_create": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(P)]),
    "bh_ring_render": (C.c_int, [P, C.POINTER(Scene), C.POINTER(Options), C.POINTER(C.c_int)]),
    "bh_ring_render_flat": (C.c_int, [P, C.POINTER(FlatScene), C.POINTER(C.c_int)]),
    "bh_ring_wait": (C.c_int, [P, C.c_int, C.POINTER(P), C.POINTER(P), C.POINTER(C.c_float)]),
    "bh_ring_release": (C.c_int, [P, C.c_int]),
    "bh_ring_destroy": (None, [P]),
    "bh_png_decode": (C.c_int, [P, C.c_size_t, P, C.c_size_t, C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
    "bh_image_load": (C.c_int, [C.c_char_p, P, C.c_size_t, C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
    "bh_image_write": (C.c_int, [C.c_char_p, P, C.c_int, C.c_int]),
    "bh_io_last_error": (C.c_char_p, []),
    "bh_jpeg_decode": (C.c_int, [P, C.c_size_t, P, C.c_size_t, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]),
    "bh_jpeg_last_error": (C.c_char_p, []),
    "bh_jpeg_encode": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_uint8)),
                                 C.POINTER(C.c_size_t)]),
    "bh_free": (None, [P]),
    "bh_video_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                C.POINTER(P)]),
    "bh_video_write": (C.c_int, [P, P]),
    "bh_video_close": (C.c_int, [P]),
    "bh_focal_length": (C.c_double, [C.c_int, C.c_double]),
    "bh_object_rotate": (None, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                C.c_int, C.c_double]),
    "bh_camera_look_at": (None, [C.POINTER(Camera), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "bh_annulus_init": (None, [C.POINTER(Annulus), C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                               C.c_double, C.c_int]),
    "bh_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
    "bh_uniform": (C.c_double, [C.POINTER(C.c_uint64), C.c_double, C.c_double]),
    "bh_gen_background": (None, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "bh_gen_disc_texture": (None, [P, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "bh_gen_noise": (None, [P, C.c_size_t, C.c_uint64]),
}

_lib = None


def load(path: os.PathLike | None = None):
    """Load libbh8.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    # One HIP runtime per process: torch ships its own libamdhip64 under the
    # same soname (libamdhip64.so.7) but asks for it by another file name, so
    # loading this library first would bring a second runtime in when torch
    # is imported later.  Importing torch first makes the library bind to
    # torch's runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def image(arr: np.ndarray) -> Image:
    arr = np.ascontiguousarray(arr, dtype=np.uint8)
    assert arr.ndim == 3 and arr.shape[2] in (3, 4)
    return Image(arr.ctypes.data_as(C.c_void_p), arr.shape[0], arr.shape[1], arr.shape[2])
