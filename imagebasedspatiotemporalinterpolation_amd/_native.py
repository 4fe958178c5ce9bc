"""ctypes binding of the C ABI in include/isti.h (libisti.so).

The structures below mirror include/isti.h field for field. Loading fails
loudly when the HIP library has not been built: there is no CPU fallback on
the product path.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int32, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libisti.so"

ISTI_OK = 0
ISTI_ERR_INVALID_ARG = 1
ISTI_ERR_OOM = 2
ISTI_ERR_DEVICE = 3
ISTI_ERR_UNSUPPORTED_SHAPE = 4
ISTI_MAX_AUX = 4
ISTI_MAX_FRAMES = 64
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "DEVICE", 4: "UNSUPPORTED_SHAPE"}

EXPORTED_SYMBOLS = (
    "isti_create",
    "isti_destroy",
    "isti_keyframe_set",
    "isti_keyframe_swap",
    "isti_render",
    "isti_render_frames",
    "isti_render_noalloc",
    "isti_scratch",
    "isti_set_stage_timing",
    "isti_stage_times",
    "isti_stage_totals",
    "isti_last_error",
    "isti_version",
    "isti_batch_create",
    "isti_batch_destroy",
    "isti_batch_set_graph",
    "isti_render_batch",
    "isti_batch_last_error",
    "isti_autotune",
    "isti_set_tile_shape",
    "isti_set_splat_binning",
    "isti_selftest_division",
)
ISTI_N_TILE_SHAPES = 4
ISTI_AUTOTUNE_TIMES = ISTI_N_TILE_SHAPES + 2
TILE_SHAPES = ((16, 16), (32, 8), (8, 8), (8, 32))


class ViewDesc(ctypes.Structure):
    """isti_view_desc."""

    _fields_ = [
        ("width", c_uint32),
        ("height", c_uint32),
        ("color", c_void_p),
        ("z", c_void_p),
        ("motion_x", c_void_p),
        ("motion_y", c_void_p),
        ("motion_z", c_void_p),
        ("shadow_info", c_void_p),
        ("shadow_color_a", c_void_p),
        ("shadow_color_b", c_void_p),
        ("viewproj", c_float * 16),
        ("inv_viewproj", c_float * 16),
        ("priority", c_int32),
    ]


class RenderInfoC(ctypes.Structure):
    """isti_render_info."""

    _fields_ = [
        ("width", c_uint32),
        ("height", c_uint32),
        ("alpha", c_float),
        ("viewproj", c_float * 16),
        ("inv_viewproj", c_float * 16),
        ("frames_per_anchor", c_int32),
        ("shadow_frames_per_anchor", c_int32),
        ("no_shadows", c_int32),
        ("mode", c_int32),
        ("max_search_iterations", c_int32),
    ]


class Config(ctypes.Structure):
    """isti_config."""

    _fields_ = [("device", c_int32), ("n_aux", c_int32)]


class ScratchView(ctypes.Structure):
    """isti_scratch_view."""

    _fields_ = [
        ("width", c_uint32),
        ("height", c_uint32),
        ("splat", c_void_p),
        ("cam_ff", c_void_p),
        ("full_ff", c_void_p),
    ]


class IstiError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None


def lib() -> ctypes.CDLL:
    """Load libisti.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m imagebasedspatiotemporalinterpolation_amd.build`"
        )
    # torch's HIP runtime must be in the process first: libisti.so then binds
    # to it (same soname) instead of loading a second runtime beside it.
    import torch  # noqa: F401

    L = ctypes.CDLL(str(LIB_PATH))
    L.isti_create.argtypes = [POINTER(Config), POINTER(c_void_p)]
    L.isti_create.restype = c_int32
    L.isti_destroy.argtypes = [c_void_p]
    L.isti_destroy.restype = c_int32
    L.isti_keyframe_set.argtypes = [c_void_p, c_int32, c_int32, POINTER(ViewDesc)]
    L.isti_keyframe_set.restype = c_int32
    L.isti_keyframe_swap.argtypes = [c_void_p]
    L.isti_keyframe_swap.restype = c_int32
    L.isti_render.argtypes = [c_void_p, POINTER(RenderInfoC), c_void_p, c_size_t, c_void_p]
    L.isti_render.restype = c_int32
    L.isti_render_frames.argtypes = [c_void_p, POINTER(RenderInfoC), c_uint32, POINTER(c_void_p), POINTER(c_size_t),
                                     c_void_p]
    L.isti_render_frames.restype = c_int32
    L.isti_render_noalloc.argtypes = [c_void_p, POINTER(RenderInfoC), c_void_p, c_size_t, c_void_p]
    L.isti_render_noalloc.restype = c_int32
    L.isti_scratch.argtypes = [c_void_p, c_int32, c_int32, POINTER(ScratchView)]
    L.isti_scratch.restype = c_int32
    L.isti_set_stage_timing.argtypes = [c_void_p, c_int32]
    L.isti_set_stage_timing.restype = c_int32
    L.isti_stage_times.argtypes = [c_void_p, POINTER(c_float)]
    L.isti_stage_times.restype = c_int32
    L.isti_stage_totals.argtypes = [c_void_p, POINTER(ctypes.c_double), POINTER(ctypes.c_int64)]
    L.isti_stage_totals.restype = c_int32
    L.isti_last_error.argtypes = [c_void_p]
    L.isti_last_error.restype = c_char_p
    L.isti_version.argtypes = []
    L.isti_version.restype = c_char_p
    L.isti_selftest_division.argtypes = [c_int32, ctypes.c_uint64, ctypes.c_uint64, c_void_p, c_void_p]
    L.isti_selftest_division.restype = c_int32
    L.isti_batch_create.argtypes = [c_int32, c_int32, POINTER(c_void_p)]
    L.isti_batch_create.restype = c_int32
    L.isti_batch_destroy.argtypes = [c_void_p]
    L.isti_batch_destroy.restype = c_int32
    L.isti_batch_set_graph.argtypes = [c_void_p, c_int32]
    L.isti_batch_set_graph.restype = c_int32
    L.isti_render_batch.argtypes = [c_void_p, POINTER(c_void_p), POINTER(RenderInfoC), POINTER(c_void_p),
                                    POINTER(c_size_t), c_int32, c_void_p]
    L.isti_render_batch.restype = c_int32
    L.isti_batch_last_error.argtypes = [c_void_p]
    L.isti_batch_last_error.restype = c_char_p
    L.isti_autotune.argtypes = [c_void_p, POINTER(RenderInfoC), c_void_p, c_size_t, c_int32, POINTER(c_float),
                                POINTER(c_int32), POINTER(c_int32), c_void_p]
    L.isti_set_splat_binning.argtypes = [c_void_p, c_int32]
    L.isti_set_splat_binning.restype = c_int32
    L.isti_autotune.restype = c_int32
    L.isti_set_tile_shape.argtypes = [c_void_p, c_int32]
    L.isti_set_tile_shape.restype = c_int32
    _lib = L
    return L


def check(status: int, handle) -> None:
    if status != ISTI_OK:
        msg = lib().isti_last_error(handle)
        raise IstiError(status, msg.decode() if msg else "")


def mat16(m) -> "ctypes.Array":
    """A 4x4 (row, col) matrix or a flat column-major 16-vector -> c_float*16."""
    a = np.asarray(m, dtype=np.float32)
    if a.shape == (4, 4):
        a = a.T.reshape(16)  # column-major
    a = np.ascontiguousarray(a.reshape(16), dtype=np.float32)
    return (c_float * 16)(*a.tolist())


# --- device copies through the HIP runtime (for inspecting scratch) ---
_hip = None


def _hip_lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    return _hip


def device_to_host(ptr: int, nbytes: int, dtype, shape) -> np.ndarray:
    """Synchronous device->host copy of `nbytes` at device address `ptr`."""
    out = np.empty(shape, dtype=dtype)
    assert out.nbytes == nbytes
    err = _hip_lib().hipMemcpy(out.ctypes.data, ptr, nbytes, 2)  # hipMemcpyDeviceToHost
    if err != 0:
        raise RuntimeError(f"hipMemcpy failed with {err}")
    return out


__all__ = [
    "Config",
    "EXPORTED_SYMBOLS",
    "IstiError",
    "RenderInfoC",
    "ScratchView",
    "ViewDesc",
    "check",
    "device_to_host",
    "lib",
    "mat16",
    "c_uint64",
]
