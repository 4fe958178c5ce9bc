"""Seeded synthetic keyframes (ctypes binding of include/isti_scene.h).

Replaces the reference's on-disk keyframe reader (src/file_reader.cu) as the
input of the interpolation path; outputs are numpy arrays in the
reference's formats (RGBA8 colour, NDC z, clip-space motion, soft-shadow
triple).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libisti_scene.so"


class SceneConfig(ctypes.Structure):
    """isti_scene_config."""

    _fields_ = [
        ("seed", c_uint64),
        ("width", c_uint32),
        ("height", c_uint32),
        ("vfov_deg", c_float),
        ("near_m", c_float),
        ("far_m", c_float),
        ("n_aux", c_int32),
        ("aux_width", c_uint32),
        ("aux_height", c_uint32),
        ("aux_offset_m", c_float),
        ("aux_vfov_deg", c_float),
        ("n_eyes", c_int32),
        ("ipd_m", c_float),
        ("n_spheres", c_int32),
        ("radius_min", c_float),
        ("radius_max", c_float),
        ("speed_max", c_float),
        ("sinusoidal", c_int32),
        ("cam_speed", c_float),
        ("cam_yaw_deg", c_float),
        ("keyframe_spacing", c_int32),
        ("shadows", c_int32),
        ("shadow_samples", c_int32),
        ("light_radius_m", c_float),
        ("shadow_frames", c_int32),
    ]

    def copy(self, **changes) -> "SceneConfig":
        c = SceneConfig()
        ctypes.pointer(c)[0] = self
        for k, v in changes.items():
            setattr(c, k, v)
        return c


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m imagebasedspatiotemporalinterpolation_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        L.isti_scene_defaults.argtypes = [c_int, c_uint64, POINTER(SceneConfig)]
        L.isti_scene_defaults.restype = c_int
        L.isti_scene_camera.argtypes = [POINTER(SceneConfig), c_double, c_int, c_void_p, c_void_p,
                                        POINTER(c_uint32), POINTER(c_uint32)]
        L.isti_scene_camera.restype = c_int
        L.isti_scene_camera_parts.argtypes = [POINTER(SceneConfig), c_double, c_int, c_void_p, c_void_p]
        L.isti_scene_camera_parts.restype = c_int
        L.isti_scene_keyframe.argtypes = [POINTER(SceneConfig), c_int, c_int, c_int] + [c_void_p] * 10 + [c_int]
        L.isti_scene_keyframe.restype = c_int
        L.isti_scene_ground_truth.argtypes = [POINTER(SceneConfig), c_double, c_int, c_void_p, c_int]
        L.isti_scene_ground_truth.restype = c_int
        _lib = L
    return _lib


def defaults(workload: int, seed: int = 1, **overrides) -> SceneConfig:
    """Defaults of workload 1 (640x360), 2 (1080p + 2 aux + shadows) or 3 (stereo)."""
    c = SceneConfig()
    if lib().isti_scene_defaults(workload, seed, ctypes.byref(c)) != 0:
        raise ValueError(f"unknown workload {workload}")
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def view_size(cfg: SceneConfig, view: int) -> tuple[int, int]:
    if cfg.n_eyes == 1 and view >= 1:
        return cfg.aux_width, cfg.aux_height
    return cfg.width, cfg.height


def camera(cfg: SceneConfig, t: float, view: int = 0):
    """(viewproj[16], inv_viewproj[16], width, height) of `view` at time t (frames)."""
    vp = np.zeros(16, np.float32)
    ivp = np.zeros(16, np.float32)
    w, h = c_uint32(), c_uint32()
    if lib().isti_scene_camera(ctypes.byref(cfg), float(t), view, vp.ctypes.data, ivp.ctypes.data,
                               ctypes.byref(w), ctypes.byref(h)) != 0:
        raise ValueError("invalid view")
    return vp, ivp, w.value, h.value


def camera_parts(cfg: SceneConfig, t: float, view: int = 0):
    """(view[16], proj[16]) of `view` at time t, column-major; viewproj = proj * view."""
    vm = np.zeros(16, np.float32)
    pm = np.zeros(16, np.float32)
    if lib().isti_scene_camera_parts(ctypes.byref(cfg), float(t), view, vm.ctypes.data, pm.ctypes.data) != 0:
        raise ValueError("invalid view")
    return vm, pm


def keyframe(cfg: SceneConfig, k: int, view: int = 0, direction: int = 1, threads: int = 0) -> dict:
    """Keyframe k of `view`; direction +1 = motion toward keyframe k+1, -1 toward k-1."""
    w, h = view_size(cfg, view)
    out = {
        "color": np.zeros((h, w, 4), np.uint8),
        "z": np.zeros((h, w), np.float32),
        "motion": np.zeros((3, h, w), np.float32),
        "shadow_info": np.zeros((h, w, 4), np.uint8),
        "shadow_color_a": np.zeros((h, w, 4), np.uint8),
        "shadow_color_b": np.zeros((h, w, 4), np.uint8),
        "viewproj": np.zeros(16, np.float32),
        "inv_viewproj": np.zeros(16, np.float32),
    }
    m = out["motion"]
    rc = lib().isti_scene_keyframe(
        ctypes.byref(cfg), k, view, direction,
        out["color"].ctypes.data, out["z"].ctypes.data,
        m[0].ctypes.data, m[1].ctypes.data, m[2].ctypes.data,
        out["shadow_info"].ctypes.data, out["shadow_color_a"].ctypes.data, out["shadow_color_b"].ctypes.data,
        out["viewproj"].ctypes.data, out["inv_viewproj"].ctypes.data, threads)
    if rc != 0:
        raise ValueError("keyframe generation failed")
    if not cfg.shadows:
        out["shadow_info"] = out["shadow_color_a"] = out["shadow_color_b"] = None
    out["width"], out["height"] = w, h
    return out


def ground_truth(cfg: SceneConfig, t: float, view: int = 0, threads: int = 0) -> np.ndarray:
    w, h = view_size(cfg, view)
    img = np.zeros((h, w, 4), np.uint8)
    if lib().isti_scene_ground_truth(ctypes.byref(cfg), float(t), view, img.ctypes.data, threads) != 0:
        raise ValueError("ground truth failed")
    return img


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR over RGB on [0,1] as -10 log10(MSE) (the reference's scripts/eval.py:57,68)."""
    x = a[..., :3].astype(np.float64) / 255.0
    y = b[..., :3].astype(np.float64) / 255.0
    mse = float(np.mean((x - y) ** 2))
    return float("inf") if mse == 0.0 else -10.0 * np.log10(mse)
