"""ctypes binding of ``libjaxsplat_amd.so`` (the C ABI in ``include/jaxsplat_amd.h``).

The shared library is built in-tree by ``jaxsplat_amd._build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

# JAXSPLAT_AMD_LIB selects another build of the library (e.g. for an A/B
# comparison of two kernel variants); by default the in-tree build.
LIB_PATH = Path(os.environ.get("JAXSPLAT_AMD_LIB") or
                Path(__file__).resolve().parent / "libjaxsplat_amd.so")

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_CAPACITY = 3
ERR_WORKSPACE = 4
ADJOINT = {"reference": 0, "consistent": 1}
SCHEDULE = {"grid": 0, "queue": 1}

# Every extern "C" symbol declared in include/jaxsplat_amd.h.
EXPORTS = (
    "jaxsplat_version", "jaxsplat_last_error", "jaxsplat_set_stage_marks",
    "jaxsplat_make_descriptor",
    "jaxsplat_project_fwd", "jaxsplat_project_bwd", "jaxsplat_rasterize_fwd",
    "jaxsplat_rasterize_bwd", "jaxsplat_project_workspace_size",
    "jaxsplat_project_forward", "jaxsplat_project_forward_noalloc", "jaxsplat_project_backward",
    "jaxsplat_bin_workspace_size", "jaxsplat_bin_and_sort", "jaxsplat_bin_check",
    "jaxsplat_bin_depth_workspace_size", "jaxsplat_bin_depth_sort",
    "jaxsplat_bin_emit_workspace_size", "jaxsplat_bin_emit_sort", "jaxsplat_bin_rects_sort",
    "jaxsplat_rasterize_backward_staged_workspace_size", "jaxsplat_rasterize_backward_staged",
    "jaxsplat_tile_rects_workspace_size", "jaxsplat_tile_rects", "jaxsplat_bin_depth_sort_rects",
    "jaxsplat_project_forward_rects",
    "jaxsplat_rasterize_backward_staged_rects", "jaxsplat_render_backward",
    "jaxsplat_project_forward_splats", "jaxsplat_rasterize_forward_splats",
    "jaxsplat_rasterize_forward_binned", "jaxsplat_rasterize_backward_binned",
    "jaxsplat_rasterize_forward_nd", "jaxsplat_rasterize_backward_nd",
    "jaxsplat_batch_create", "jaxsplat_batch_destroy", "jaxsplat_batch_workspace_size",
    "jaxsplat_batch_set_depths", "jaxsplat_batch_set_schedule",
    "jaxsplat_batch_num_pixels", "jaxsplat_batch_num_gaussians", "jaxsplat_batch_pixel_offset",
    "jaxsplat_batch_forward", "jaxsplat_batch_backward", "jaxsplat_batch_check",
    "jaxsplat_batch_status_word", "jaxsplat_fit_activate", "jaxsplat_fit_l2",
    "jaxsplat_fit_adam",
)


class Dim3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint32), ("y", ctypes.c_uint32), ("z", ctypes.c_uint32)]


class Descriptor(ctypes.Structure):
    """Mirror of ``jaxsplat_descriptor`` (80 bytes, reference ``lib/ops.h:9-22``)."""
    _fields_ = [
        ("num_points", ctypes.c_uint32),
        ("img_shape", Dim3),
        ("intrins", ctypes.c_float * 4),      # offset 16 (16-byte aligned in C)
        ("glob_scale", ctypes.c_float),
        ("clip_thresh", ctypes.c_float),
        ("block_width", ctypes.c_uint32),
        ("grid_dim_1d", ctypes.c_uint32),
        ("block_dim_1d", ctypes.c_uint32),
        ("grid_dim_2d", Dim3),
        ("block_dim_2d", Dim3),
        ("_pad", ctypes.c_uint32),
    ]


assert ctypes.sizeof(Descriptor) == 80


class Scene(ctypes.Structure):
    """Mirror of ``jaxsplat_scene``."""
    _fields_ = [("first_gaussian", ctypes.c_int32), ("num_gaussians", ctypes.c_int32)]


class View(ctypes.Structure):
    """Mirror of ``jaxsplat_view``."""
    _fields_ = [("scene", ctypes.c_int32), ("img_h", ctypes.c_int32), ("img_w", ctypes.c_int32),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float),
                ("cy", ctypes.c_float)]


class JaxsplatError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"jaxsplat_amd error {code}: {message}")
        self.code = code


_lock = threading.Lock()
_lib = None

P = ctypes.c_void_p
U32 = ctypes.c_uint32
I32 = ctypes.c_int
F32 = ctypes.c_float
SZ = ctypes.c_size_t
DP = ctypes.POINTER(Descriptor)

_SIGS = {
    "jaxsplat_version": (ctypes.c_char_p, []),
    "jaxsplat_last_error": (ctypes.c_char_p, []),
    "jaxsplat_set_stage_marks": (None, [ctypes.POINTER(P), I32]),
    "jaxsplat_make_descriptor": (I32, [DP, U32, U32, U32, F32, F32, F32, F32, F32, F32, U32]),
    "jaxsplat_project_fwd": (None, [P, ctypes.POINTER(P), ctypes.c_char_p, SZ]),
    "jaxsplat_project_bwd": (None, [P, ctypes.POINTER(P), ctypes.c_char_p, SZ]),
    "jaxsplat_rasterize_fwd": (None, [P, ctypes.POINTER(P), ctypes.c_char_p, SZ]),
    "jaxsplat_rasterize_bwd": (None, [P, ctypes.POINTER(P), ctypes.c_char_p, SZ]),
    "jaxsplat_project_workspace_size": (SZ, [DP]),
    "jaxsplat_project_forward": (I32, [DP] + [P] * 12 + [P, SZ, P]),
    "jaxsplat_project_forward_noalloc": (I32, [DP] + [P] * 12 + [P]),
    "jaxsplat_project_backward": (I32, [DP] + [P] * 16 + [I32, P]),
    "jaxsplat_bin_workspace_size": (SZ, [DP, U32]),
    "jaxsplat_bin_and_sort": (I32, [DP, P, P, P, U32, P, P, P, P, SZ, P]),
    "jaxsplat_bin_check": (I32, [DP, U32, P, P]),
    "jaxsplat_bin_depth_workspace_size": (SZ, [DP]),
    "jaxsplat_bin_depth_sort": (I32, [DP, P, P, P, P, SZ, P]),
    "jaxsplat_bin_emit_workspace_size": (SZ, [DP, U32]),
    "jaxsplat_bin_emit_sort": (I32, [DP, U32, P, P, P, P, P, SZ, P]),
    "jaxsplat_bin_rects_sort": (I32, [DP, P, P, P, U32, P, P, P, P, SZ, P]),
    "jaxsplat_rasterize_backward_staged_workspace_size": (SZ, [DP, U32]),
    "jaxsplat_tile_rects_workspace_size": (SZ, [DP]),
    "jaxsplat_tile_rects": (I32, [DP] + [P] * 7 + [P, SZ, P]),
    "jaxsplat_project_forward_rects": (I32, [DP] + [P] * 14 + [P, SZ, P]),
    "jaxsplat_bin_depth_sort_rects": (I32, [DP, P, P, P, P, SZ, P]),
    "jaxsplat_rasterize_backward_staged_rects": (I32, [DP, P, P, U32, P, P, P, P, P, U32, P, P,
                                                        P, P, P, P, P, P, P, P, P, P, I32, P, SZ,
                                                        P]),
    "jaxsplat_render_backward": (I32, [DP, P, P, U32] + [P] * 7 + [U32] + [P] * 11
                                 + [I32, P, SZ, P]),
    "jaxsplat_project_forward_splats": (I32, [DP] + [P] * 6 + [U32] + [P] * 9 + [P, SZ, P]),
    "jaxsplat_rasterize_forward_splats": (I32, [DP, P, P, U32, P, P, U32, P, P, P, P]),
    "jaxsplat_rasterize_backward_staged": (I32, [DP, P, P, U32, P, P, P, P, P, U32, P, P, P, P, P,
                                                  P, P, P, P, P, P, I32, P, SZ, P]),
    "jaxsplat_rasterize_forward_binned": (I32, [DP, P, P, U32] + [P] * 8 + [P]),
    "jaxsplat_rasterize_backward_binned": (I32, [DP, P, P, U32] + [P] * 14 + [I32, P]),
    "jaxsplat_rasterize_forward_nd": (I32, [DP, P, P, U32] + [P] * 5 + [U32] + [P] * 3 + [P]),
    "jaxsplat_rasterize_backward_nd": (I32, [DP, P, P, U32] + [P] * 5 + [U32] + [P] * 9
                                       + [I32, P]),
    "jaxsplat_batch_create": (I32, [ctypes.POINTER(P), ctypes.POINTER(Scene), I32,
                                    ctypes.POINTER(View), I32, F32, F32, U32, ctypes.c_int64]),
    "jaxsplat_batch_destroy": (None, [P]),
    "jaxsplat_batch_workspace_size": (SZ, [P]),
    "jaxsplat_batch_set_depths": (I32, [P, I32, I32]),
    "jaxsplat_batch_set_schedule": (I32, [P, I32]),
    "jaxsplat_batch_num_pixels": (ctypes.c_int64, [P]),
    "jaxsplat_batch_num_gaussians": (ctypes.c_int64, [P]),
    "jaxsplat_batch_pixel_offset": (ctypes.c_int64, [P, I32]),
    "jaxsplat_batch_forward": (I32, [P] + [P] * 10 + [P, SZ, P]),
    "jaxsplat_batch_backward": (I32, [P] + [P] * 15 + [I32, P, SZ, P]),
    "jaxsplat_batch_check": (I32, [P, P, ctypes.POINTER(ctypes.c_int64), P]),
    "jaxsplat_batch_status_word": (P, [P, P]),
    "jaxsplat_fit_activate": (I32, [ctypes.c_int64] + [P] * 6 + [P]),
    "jaxsplat_fit_l2": (I32, [ctypes.c_int64] + [P] * 4 + [P]),
    "jaxsplat_fit_adam": (I32, [ctypes.c_int64] + [P] * 14 + [F32, F32, F32, F32, P]),
}


def lib():
    """Load (once) and return the ctypes library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m jaxsplat_amd._build` "
                    "(hipcc, gfx950).  jaxsplat_amd has no CPU fallback.")
            L = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(code):
    if code != OK:
        raise JaxsplatError(code, lib().jaxsplat_last_error().decode())


def make_descriptor(num_points, img_shape, f, c, glob_scale, clip_thresh, block_width):
    d = Descriptor()
    H, W = img_shape
    code = lib().jaxsplat_make_descriptor(ctypes.byref(d), int(num_points), int(H), int(W),
                                          float(f[0]), float(f[1]), float(c[0]), float(c[1]),
                                          float(glob_scale), float(clip_thresh),
                                          int(block_width))
    if code != OK:
        raise ValueError(lib().jaxsplat_last_error().decode())
    return d
