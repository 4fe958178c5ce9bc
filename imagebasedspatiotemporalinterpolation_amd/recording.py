"""Keyframe recordings in the reference's on-disk format.

Reading goes through libisti_io.so (include/isti_io.h; the reference's
FrameFileReader, src/file_reader.cu:17-230). Writing is done here: PNG and
OpenEXR encoders plus cam.json, so a seeded synthetic scene can be exported
as a recording that the reference's own splinter reads (and as test input
for the reader). Layout per camera directory:

    cam.json, color/NNNN.png, depth/NNNN.exr,
    forwardFlowField/NNNN.exr, backwardFlowField/NNNN.exr,
    forward|backwardShadowInfo|ShadowColorA|ShadowColorB/NNNN.png
"""
from __future__ import annotations

import ctypes
import json
import struct
import zlib
from ctypes import POINTER, c_char_p, c_float, c_int32, c_size_t, c_uint8, c_uint32, c_void_p
from pathlib import Path

import numpy as np

from . import scene as S

PKG = Path(__file__).resolve().parent
IO_LIB_PATH = PKG / "libisti_io.so"


class Camera(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("client_fps", c_int32), ("server_fps", c_int32),
                ("shadow_fps", c_int32), ("priority", c_int32), ("n_frames", c_int32)]


class Frame(ctypes.Structure):
    _fields_ = [("idx", c_int32), ("w2c", c_float * 16), ("proj", c_float * 16)]


class View(ctypes.Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("color", POINTER(c_uint8)), ("z", POINTER(c_float)),
                ("motion_x", POINTER(c_float)), ("motion_y", POINTER(c_float)), ("motion_z", POINTER(c_float)),
                ("shadow_info", POINTER(c_uint8)), ("shadow_color_a", POINTER(c_uint8)),
                ("shadow_color_b", POINTER(c_uint8)), ("view", c_float * 16), ("proj", c_float * 16),
                ("viewproj", c_float * 16), ("inv_viewproj", c_float * 16)]


_io = None


def io_lib() -> ctypes.CDLL:
    global _io
    if _io is None:
        if not IO_LIB_PATH.exists():
            raise RuntimeError(f"{IO_LIB_PATH} is missing: run `python -m imagebasedspatiotemporalinterpolation_amd.build`")
        L = ctypes.CDLL(str(IO_LIB_PATH))
        L.isti_io_read_camera.argtypes = [c_char_p, POINTER(Camera), POINTER(Frame), c_int32]
        L.isti_io_read_view.argtypes = [c_char_p, c_int32, c_int32, POINTER(View)]
        L.isti_io_free_view.argtypes = [POINTER(View)]
        L.isti_io_free_view.restype = None
        L.isti_io_decode_png.argtypes = [c_char_p, POINTER(c_uint32), POINTER(c_uint32), c_void_p, c_size_t]
        L.isti_io_decode_exr.argtypes = [c_char_p, c_char_p, POINTER(c_uint32), POINTER(c_uint32), c_void_p,
                                         c_size_t]
        for f in ("isti_io_read_camera", "isti_io_read_view", "isti_io_decode_png", "isti_io_decode_exr"):
            getattr(L, f).restype = c_int32
        L.isti_io_camera_at.argtypes = [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_void_p]
        L.isti_io_camera_at.restype = c_int32
        L.isti_io_client_camera.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
        L.isti_io_client_camera.restype = c_int32
        L.isti_io_last_error.argtypes = []
        L.isti_io_last_error.restype = c_char_p
        _io = L
    return _io


def _check(st: int) -> None:
    if st != 0:
        raise IOError(io_lib().isti_io_last_error().decode())


# ---------------------------------------------------------------------------
# reading
# ---------------------------------------------------------------------------
def read_camera(cam_dir) -> tuple:
    """(camera info dict, [(idx, w2c[16], proj[16])]) of <cam_dir>/cam.json."""
    L = io_lib()
    cam = Camera()
    _check(L.isti_io_read_camera(str(cam_dir).encode(), ctypes.byref(cam), None, 0))
    frames = (Frame * cam.n_frames)()
    _check(L.isti_io_read_camera(str(cam_dir).encode(), ctypes.byref(cam), frames, cam.n_frames))
    info = {k: getattr(cam, k) for k, _ in Camera._fields_}
    return info, [(f.idx, np.array(f.w2c, np.float32), np.array(f.proj, np.float32)) for f in frames]


def read_view(cam_dir, pair: int, which: int) -> dict:
    """Keyframe view `which` (0: frame pair, forward data; 1: frame pair+1,
    backward data) as the dict KeyframeView.from_numpy takes."""
    L = io_lib()
    v = View()
    _check(L.isti_io_read_view(str(cam_dir).encode(), pair, which, ctypes.byref(v)))
    try:
        w, h = v.width, v.height

        def arr(p, shape, dtype):
            if not p:
                return None
            n = int(np.prod(shape))
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype).reshape(shape).copy()

        mx = arr(v.motion_x, (h, w), np.float32)
        out = {
            "color": arr(v.color, (h, w, 4), np.uint8),
            "z": arr(v.z, (h, w), np.float32),
            "motion": None if mx is None else np.stack([mx, arr(v.motion_y, (h, w), np.float32),
                                                        arr(v.motion_z, (h, w), np.float32)]),
            "shadow_info": arr(v.shadow_info, (h, w, 4), np.uint8),
            "shadow_color_a": arr(v.shadow_color_a, (h, w, 4), np.uint8),
            "shadow_color_b": arr(v.shadow_color_b, (h, w, 4), np.uint8),
            "view": np.array(v.view, np.float32),
            "proj": np.array(v.proj, np.float32),
            "viewproj": np.array(v.viewproj, np.float32),
            "inv_viewproj": np.array(v.inv_viewproj, np.float32),
            "width": w,
            "height": h,
        }
    finally:
        L.isti_io_free_view(ctypes.byref(v))
    return out


def camera_at(view_fwd, view_bwd, proj_fwd, t: float) -> tuple:
    """(view_t, viewproj, inv_viewproj) of the client camera at t between two
    keyframes (glm::interpolate of the view matrices, src/main.cu:26-29)."""
    f = lambda m: np.ascontiguousarray(m, np.float32).reshape(16)
    a, b, p = f(view_fwd), f(view_bwd), f(proj_fwd)
    v, vp, ivp = (np.zeros(16, np.float32) for _ in range(3))
    _check(io_lib().isti_io_camera_at(a.ctypes.data, b.ctypes.data, p.ctypes.data, float(t), v.ctypes.data,
                                      vp.ctypes.data, ivp.ctypes.data))
    return v, vp, ivp


def client_camera(w2c, proj) -> tuple:
    """(viewproj, inv_viewproj) of a camera-path frame: normalizeProjMat(proj)
    * w2c (src/main.cu:158-161, src/file_reader.cu:219-230)."""
    f = lambda m: np.ascontiguousarray(m, np.float32).reshape(16)
    a, p = f(w2c), f(proj)
    vp, ivp = np.zeros(16, np.float32), np.zeros(16, np.float32)
    _check(io_lib().isti_io_client_camera(a.ctypes.data, p.ctypes.data, vp.ctypes.data, ivp.ctypes.data))
    return vp, ivp


def decode_png(path) -> np.ndarray:
    L = io_lib()
    w, h = c_uint32(), c_uint32()
    _check(L.isti_io_decode_png(str(path).encode(), ctypes.byref(w), ctypes.byref(h), None, 0))
    out = np.zeros((h.value, w.value, 4), np.uint8)
    _check(L.isti_io_decode_png(str(path).encode(), ctypes.byref(w), ctypes.byref(h), out.ctypes.data, out.nbytes))
    return out


def decode_exr(path, channels=("R", "G", "B", "A")) -> np.ndarray:
    L = io_lib()
    w, h = c_uint32(), c_uint32()
    names = ",".join(channels).encode()
    _check(L.isti_io_decode_exr(str(path).encode(), names, ctypes.byref(w), ctypes.byref(h), None, 0))
    out = np.zeros((h.value, w.value, len(channels)), np.float32)
    _check(L.isti_io_decode_exr(str(path).encode(), names, ctypes.byref(w), ctypes.byref(h), out.ctypes.data,
                                out.size))
    return out


# ---------------------------------------------------------------------------
# writing
# ---------------------------------------------------------------------------
def _png_chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def encode_png(img: np.ndarray, filters=None) -> bytes:
    """8-bit PNG of an [H, W] (grey) or [H, W, C] (C = 2, 3, 4) image. Row r
    uses filter filters[r % len(filters)] (default: all five in turn)."""
    img = np.ascontiguousarray(img, np.uint8)
    if img.ndim == 2:
        img = img[..., None]
    h, w, c = img.shape
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[c]
    filters = filters or (0, 1, 2, 3, 4)
    prev = np.zeros(w * c, np.int32)
    rows = []
    for y in range(h):
        cur = img[y].reshape(-1).astype(np.int32)
        f = filters[y % len(filters)]
        a = np.concatenate([np.zeros(c, np.int32), cur[:-c]])
        b = prev
        cc = np.concatenate([np.zeros(c, np.int32), prev[:-c]])
        if f == 0:
            pred = np.zeros_like(cur)
        elif f == 1:
            pred = a
        elif f == 2:
            pred = b
        elif f == 3:
            pred = (a + b) >> 1
        else:
            p = a + b - cc
            pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - cc)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, cc))
        rows.append(bytes([f]) + ((cur - pred) & 0xFF).astype(np.uint8).tobytes())
        prev = cur
    ihdr = struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(b"".join(rows), 6))
            + _png_chunk(b"IEND", b""))


def _exr_attr(name: str, kind: str, data: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(data)) + data


def encode_exr(channels: dict, compression: str = "zip", pixel_type: str = "float") -> bytes:
    """Single-part scanline OpenEXR of {name: [H, W] float32 array}; compression
    "none", "zips" or "zip", pixel type "half" or "float"."""
    names = sorted(channels)
    h, w = channels[names[0]].shape
    pt = {"half": 1, "float": 2}[pixel_type]
    comp = {"none": 0, "zips": 2, "zip": 3}[compression]
    lines = {0: 1, 2: 1, 3: 16}[comp]
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", pt, 0, 1, 1) for n in names) + b"\0"
    header = (b"\x76\x2f\x31\x01" + struct.pack("<I", 2)
              + _exr_attr("channels", "chlist", chlist)
              + _exr_attr("compression", "compression", bytes([comp]))
              + _exr_attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
              + _exr_attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
              + _exr_attr("lineOrder", "lineOrder", b"\0")
              + _exr_attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _exr_attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
              + _exr_attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    dt = np.float16 if pt == 1 else np.float32
    planes = [np.ascontiguousarray(channels[n], np.float32).astype(dt) for n in names]
    n_chunks = (h + lines - 1) // lines
    chunks = []
    for k in range(n_chunks):
        y0 = k * lines
        raw = b"".join(p[y].tobytes() for y in range(y0, min(h, y0 + lines)) for p in planes)
        data = raw
        if comp != 0:
            b = np.frombuffer(raw, np.uint8)
            inter = np.concatenate([b[0::2], b[1::2]]).astype(np.int32)  # inverse of the de-interleave
            pred = inter.copy()
            pred[1:] = (inter[1:] - inter[:-1] + 128) & 0xFF  # inverse of the predictor
            z = zlib.compress(pred.astype(np.uint8).tobytes(), 6)
            if len(z) < len(raw):
                data = z
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    offsets, pos = [], len(header) + 8 * n_chunks
    for c in chunks:
        offsets.append(pos)
        pos += len(c)
    return header + struct.pack(f"<{n_chunks}Q", *offsets) + b"".join(chunks)


# normalizeProjMat's matrices (include/file_reader.h:64-67), column-major
_REVERSE_Z = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 1, 1], np.float64).reshape(4, 4).T
_DEPTH_01_TO_11 = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 2, 0, 0, 0, -1, 1], np.float64).reshape(4, 4).T
_UNITY = np.array([1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], np.float64).reshape(4, 4).T
NORMALIZE = _UNITY @ _DEPTH_01_TO_11 @ _REVERSE_Z


def raw_projection(proj_gl: np.ndarray) -> np.ndarray:
    """The recording's projection whose normalisation is proj_gl (column-major [16])."""
    P = np.asarray(proj_gl, np.float64).reshape(4, 4).T
    return (np.linalg.inv(NORMALIZE) @ P).T.reshape(16).astype(np.float32)


def _mat_json(m16) -> dict:
    m = np.asarray(m16, np.float32).reshape(4, 4)  # m[col][row]
    return {f"e{r}{c}": float(m[c, r]) for r in range(4) for c in range(4)}


def z_to_file_depth(z: np.ndarray) -> np.ndarray:
    """NDC z -> reversed-Z [0, 1] depth as stored in the recordings."""
    return (np.float32(1.0) - (z.astype(np.float32) + np.float32(1.0)) * np.float32(0.5)).astype(np.float32)


def write_recording(out_dir, cfg, n_keyframes: int, view: int = 0, cam_name: str = "main",
                    exr_compression: str = "zip", client_fps: int = 60, shadows: bool = True) -> Path:
    """Exports keyframes 0..n_keyframes-1 of one view of a seeded synthetic
    scene in the reference's recording format (server frame k = scene time
    k * keyframe_spacing client frames). Returns the camera directory."""
    d = Path(out_dir) / cam_name
    for sub in ("color", "depth", "forwardFlowField", "backwardFlowField"):
        (d / sub).mkdir(parents=True, exist_ok=True)
    if shadows and cfg.shadows:
        for sub in ("ShadowInfo", "ShadowColorA", "ShadowColorB"):
            (d / ("forward" + sub)).mkdir(exist_ok=True)
            (d / ("backward" + sub)).mkdir(exist_ok=True)
    w, h = S.view_size(cfg, view)
    frames = []
    for k in range(n_keyframes):
        fwd = S.keyframe(cfg, k, view, +1)
        (d / "color" / f"{k:04d}.png").write_bytes(encode_png(fwd["color"]))
        (d / "depth" / f"{k:04d}.exr").write_bytes(
            encode_exr({"R": z_to_file_depth(fwd["z"])}, exr_compression))
        vm, pm = S.camera_parts(cfg, k * cfg.keyframe_spacing, view)
        frames.append({"idx": k, "w2c": _mat_json(vm), "proj": _mat_json(raw_projection(pm))})
        if k + 1 < n_keyframes:
            bwd = S.keyframe(cfg, k + 1, view, -1)
            for name, kf in (("forward", fwd), ("backward", bwd)):
                m = kf["motion"]
                (d / f"{name}FlowField" / f"{k:04d}.exr").write_bytes(encode_exr(
                    {"R": m[0], "G": m[1], "B": m[2], "A": np.zeros_like(m[0])}, exr_compression))
                if shadows and cfg.shadows:
                    for sub, key in (("ShadowInfo", "shadow_info"), ("ShadowColorA", "shadow_color_a"),
                                     ("ShadowColorB", "shadow_color_b")):
                        (d / f"{name}{sub}" / f"{k:04d}.png").write_bytes(encode_png(kf[key]))
    spacing = cfg.keyframe_spacing
    cam = {"resolution": {"x": w, "y": h}, "clientFps": client_fps, "serverFps": client_fps // spacing,
           "shadowFps": client_fps // spacing * max(cfg.shadow_frames, 1) if cfg.shadows else client_fps,
           "objectMotion": True, "shadowGradients": bool(shadows and cfg.shadows), "frames": frames}
    (d / "cam.json").write_text(json.dumps(cam, indent=1))
    return d


def write_camera_path(cam_dir, cfg, n_frames: int, client_fps: int, frame_time: float = 1.0, view: int = 0) -> Path:
    """Writes a client camera path (<cam_dir>/cam.json, the file splinter -c
    reads; src/file_reader.cu:78-107): frame n is the scene's camera at scene
    time n * frame_time (the scene's own client-frame units)."""
    d = Path(cam_dir)
    d.mkdir(parents=True, exist_ok=True)
    w, h = S.view_size(cfg, view)
    frames = []
    for n in range(n_frames):
        vm, pm = S.camera_parts(cfg, n * frame_time, view)
        frames.append({"idx": n, "w2c": _mat_json(vm), "proj": _mat_json(raw_projection(pm))})
    cam = {"resolution": {"x": w, "y": h}, "clientFps": client_fps, "frames": frames}
    (d / "cam.json").write_text(json.dumps(cam, indent=1))
    return d
