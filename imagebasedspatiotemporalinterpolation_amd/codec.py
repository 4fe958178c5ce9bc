"""Keyframe wire codec of the reference's compression round trip
(compression.py), over the GPU kernels of include/isti_codec.h.

  - ShadowInfo, lossless (compression.py:223-325): per frame the third colour
    channel, optionally XOR-ed with the anchor frame taken every server_fps
    frames, optionally run-length coded, framed as
    [u16 name length][name][u32 width][u32 height][u32 length][bytes]
    (little endian) and compressed as one zstd frame (the system's libzstd;
    the reference uses the zstandard module at level 10). The reference's
    pipeline calls it with both options off (:526-528); they are kept as
    switches.
  - depth and flow before the video encoder (normalize_exr :84-91,
    denormalize_exr :93-98): 16-bit normalisation between the image's min
    and max and back to half precision. The HEVC stage that follows in the
    reference needs ffmpeg, which is not part of this build.

Tensors are torch GPU tensors (device memory only; the work runs in the HIP
kernels of libisti.so).
"""
from __future__ import annotations

import ctypes
import struct
from ctypes import POINTER, c_char_p, c_float, c_int32, c_size_t, c_uint64, c_void_p

from . import _native as N

_bound = False


def _lib():
    global _bound
    L = N.lib()
    if not _bound:
        L.isti_codec_create.argtypes = [c_int32, c_uint64, POINTER(c_void_p)]
        L.isti_codec_destroy.argtypes = [c_void_p]
        for f in ("isti_codec_rle_encode", "isti_codec_rle_decode"):
            getattr(L, f).argtypes = [c_void_p, c_void_p, c_uint64, c_void_p, c_uint64, POINTER(c_uint64), c_void_p]
        L.isti_codec_xor.argtypes = [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p]
        L.isti_codec_extract_channel.argtypes = [c_void_p, c_uint64, c_int32, c_void_p, c_void_p]
        L.isti_codec_expand_channel.argtypes = [c_void_p, c_uint64, c_void_p, c_void_p]
        L.isti_codec_normalize.argtypes = [c_void_p, c_void_p, c_uint64, c_int32, c_void_p, POINTER(c_float), c_void_p]
        L.isti_codec_denormalize.argtypes = [c_void_p, c_uint64, c_float, c_float, c_void_p, c_void_p]
        L.isti_codec_last_error.argtypes = [c_void_p]
        L.isti_codec_last_error.restype = c_char_p
        for f in ("isti_codec_create", "isti_codec_destroy", "isti_codec_rle_encode", "isti_codec_rle_decode",
                  "isti_codec_xor", "isti_codec_extract_channel", "isti_codec_expand_channel",
                  "isti_codec_normalize", "isti_codec_denormalize"):
            getattr(L, f).restype = c_int32
        _bound = True
    return L


def _stream(stream):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


class Codec:
    """GPU byte and quantisation kernels with working memory for inputs of up
    to max_bytes bytes (or floats) per call."""

    def __init__(self, max_bytes: int, device: int = 0):
        h = c_void_p()
        st = _lib().isti_codec_create(device, max_bytes, ctypes.byref(h))
        if st != N.ISTI_OK:
            raise N.IstiError(st, "isti_codec_create")
        self._h, self.max_bytes, self.device = h, max_bytes, device

    def _check(self, st, what):
        if st != N.ISTI_OK:
            raise N.IstiError(st, f"{what}: {_lib().isti_codec_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().isti_codec_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rle_encode(self, data, stream=None):
        """uint8 GPU tensor -> uint8 GPU tensor of (count, value) pairs."""
        import torch

        n = data.numel()
        out = torch.empty(max(2 * n, 2), dtype=torch.uint8, device=data.device)
        ln = c_uint64()
        self._check(_lib().isti_codec_rle_encode(self._h, data.data_ptr(), n, out.data_ptr(), out.numel(),
                                                 ctypes.byref(ln), _stream(stream)), "rle_encode")
        return out[:ln.value]

    def rle_decode(self, data, n_out: int, stream=None):
        """(count, value) pairs -> the uint8 bytes (at most n_out of them)."""
        import torch

        out = torch.empty(max(n_out, 1), dtype=torch.uint8, device=data.device)
        ln = c_uint64()
        self._check(_lib().isti_codec_rle_decode(self._h, data.data_ptr(), data.numel(), out.data_ptr(), n_out,
                                                 ctypes.byref(ln), _stream(stream)), "rle_decode")
        return out[:ln.value]

    def normalize(self, plane, zero_small: bool = True, stream=None):
        """float32 GPU tensor -> (uint16 codes as int16 tensor, min, max)."""
        import torch

        out = torch.empty(plane.shape, dtype=torch.int16, device=plane.device)
        mm = (c_float * 2)()
        self._check(_lib().isti_codec_normalize(self._h, plane.data_ptr(), plane.numel(), int(bool(zero_small)),
                                                out.data_ptr(), mm, _stream(stream)), "normalize")
        return out, float(mm[0]), float(mm[1])


def xor(a, b, stream=None):
    import torch

    out = torch.empty_like(a)
    st = _lib().isti_codec_xor(a.data_ptr(), b.data_ptr(), a.numel(), out.data_ptr(), _stream(stream))
    if st != N.ISTI_OK:
        raise N.IstiError(st, "isti_codec_xor")
    return out


def extract_channel(rgba, channel: int = 2, stream=None):
    import torch

    h, w = rgba.shape[:2]
    out = torch.empty((h, w), dtype=torch.uint8, device=rgba.device)
    st = _lib().isti_codec_extract_channel(rgba.data_ptr(), h * w, channel, out.data_ptr(), _stream(stream))
    if st != N.ISTI_OK:
        raise N.IstiError(st, "isti_codec_extract_channel")
    return out


def expand_channel(plane, stream=None):
    import torch

    h, w = plane.shape
    out = torch.empty((h, w, 4), dtype=torch.uint8, device=plane.device)
    st = _lib().isti_codec_expand_channel(plane.data_ptr(), h * w, out.data_ptr(), _stream(stream))
    if st != N.ISTI_OK:
        raise N.IstiError(st, "isti_codec_expand_channel")
    return out


def denormalize(codes, mn: float, mx: float, stream=None):
    """uint16 codes (int16 tensor) -> float32 tensor of half-precision values."""
    import torch

    out = torch.empty(codes.shape, dtype=torch.float32, device=codes.device)
    st = _lib().isti_codec_denormalize(codes.data_ptr(), codes.numel(), mn, mx, out.data_ptr(), _stream(stream))
    if st != N.ISTI_OK:
        raise N.IstiError(st, "isti_codec_denormalize")
    return out


# ---------------------------------------------------------------------------
# zstd through the system library (compression.py uses the zstandard module)
# ---------------------------------------------------------------------------
class _InBuf(ctypes.Structure):
    _fields_ = [("src", c_void_p), ("size", c_size_t), ("pos", c_size_t)]


class _OutBuf(ctypes.Structure):
    _fields_ = [("dst", c_void_p), ("size", c_size_t), ("pos", c_size_t)]


_zstd = None


def zstd_lib():
    global _zstd
    if _zstd is None:
        Z = ctypes.CDLL("libzstd.so.1")
        Z.ZSTD_compressBound.argtypes = [c_size_t]
        Z.ZSTD_compressBound.restype = c_size_t
        Z.ZSTD_compress.argtypes = [c_void_p, c_size_t, c_void_p, c_size_t, ctypes.c_int]
        Z.ZSTD_compress.restype = c_size_t
        Z.ZSTD_isError.argtypes = [c_size_t]
        Z.ZSTD_isError.restype = ctypes.c_uint
        Z.ZSTD_getErrorName.argtypes = [c_size_t]
        Z.ZSTD_getErrorName.restype = c_char_p
        Z.ZSTD_createDStream.restype = c_void_p
        Z.ZSTD_freeDStream.argtypes = [c_void_p]
        Z.ZSTD_initDStream.argtypes = [c_void_p]
        Z.ZSTD_initDStream.restype = c_size_t
        Z.ZSTD_decompressStream.argtypes = [c_void_p, POINTER(_OutBuf), POINTER(_InBuf)]
        Z.ZSTD_decompressStream.restype = c_size_t
        _zstd = Z
    return _zstd


def zstd_compress(data: bytes, level: int = 10) -> bytes:
    Z = zstd_lib()
    cap = Z.ZSTD_compressBound(len(data))
    dst = ctypes.create_string_buffer(cap)
    r = Z.ZSTD_compress(dst, cap, data, len(data), level)
    if Z.ZSTD_isError(r):
        raise IOError(Z.ZSTD_getErrorName(r).decode())
    return dst.raw[:r]


def zstd_decompress(data: bytes) -> bytes:
    """Streaming decompression: frames written without a content size (the
    reference's stream writer) decode as well."""
    Z = zstd_lib()
    ds = Z.ZSTD_createDStream()
    try:
        Z.ZSTD_initDStream(ds)
        src = ctypes.create_string_buffer(data, len(data))
        inb = _InBuf(ctypes.cast(src, c_void_p), len(data), 0)
        chunk = 1 << 20
        dst = ctypes.create_string_buffer(chunk)
        parts = []
        while True:
            outb = _OutBuf(ctypes.cast(dst, c_void_p), chunk, 0)
            r = Z.ZSTD_decompressStream(ds, ctypes.byref(outb), ctypes.byref(inb))
            if Z.ZSTD_isError(r):
                raise IOError(Z.ZSTD_getErrorName(r).decode())
            parts.append(dst.raw[:outb.pos])
            if inb.pos >= inb.size and outb.pos < chunk:
                break
        return b"".join(parts)
    finally:
        Z.ZSTD_freeDStream(ds)


# ---------------------------------------------------------------------------
# record framing (compression.py:262-268, 296-309)
# ---------------------------------------------------------------------------
def pack_records(records) -> bytes:
    """[(name, width, height, payload bytes)] -> the uncompressed record stream."""
    out = bytearray()
    for name, w, h, payload in records:
        nb = name.encode("utf-8")
        out += struct.pack("<H", len(nb)) + nb + struct.pack("<III", w, h, len(payload)) + payload
    return bytes(out)


def unpack_records(data: bytes):
    recs, p = [], 0
    while p < len(data):
        if p + 2 > len(data):
            raise IOError("truncated record header")
        (nl,) = struct.unpack_from("<H", data, p)
        p += 2
        if p + nl + 12 > len(data):
            raise IOError("truncated record header")
        name = data[p:p + nl].decode("utf-8")
        p += nl
        w, h, ln = struct.unpack_from("<III", data, p)
        p += 12
        if p + ln > len(data):
            raise IOError("truncated record payload")
        recs.append((name, w, h, data[p:p + ln]))
        p += ln
    return recs


# ---------------------------------------------------------------------------
# the ShadowInfo stream
# ---------------------------------------------------------------------------
def compress_shadow_info(codec: Codec, frames, names, server_fps: int, delta: bool = False, rle: bool = False,
                         level: int = 10) -> bytes:
    """frames: RGBA8 GPU tensors [H, W, 4] in order; returns the zstd stream.
    compress_raw_images_lossless (compression.py:223-279)."""
    recs, anchor = [], None
    for k, (f, name) in enumerate(zip(frames, names)):
        plane = extract_channel(f, 2).reshape(-1)
        data = plane
        if delta:
            if k % server_fps == 0:
                anchor = plane.clone()
            elif anchor is not None:
                data = xor(plane, anchor)
        if rle:
            data = codec.rle_encode(data)
        h, w = f.shape[:2]
        recs.append((name, w, h, data.cpu().numpy().tobytes()))
    return zstd_compress(pack_records(recs), level)


def decompress_shadow_info(codec: Codec, blob: bytes, server_fps: int, delta: bool = False, rle: bool = False):
    """The inverse (decompress_raw_images_lossless, compression.py:281-325):
    [(name, RGBA8 GPU tensor with the plane in channel 2, alpha 255)]."""
    import numpy as np
    import torch

    out, anchor = [], None
    for k, (name, w, h, payload) in enumerate(unpack_records(zstd_decompress(blob))):
        data = torch.from_numpy(np.frombuffer(payload, np.uint8).copy()).cuda()
        if rle:
            data = codec.rle_decode(data, w * h)
        if data.numel() != w * h:
            raise IOError(f"{name}: {data.numel()} bytes for a {w}x{h} plane")
        if delta:
            if k % server_fps == 0:
                anchor = data.clone()
            else:
                data = xor(data, anchor)
        out.append((name, expand_channel(data.reshape(h, w))))
    return out


def quantize_round_trip(codec: Codec, plane, zero_small: bool = True):
    """normalize_exr then denormalize_exr: the depth / flow values a receiver
    gets back before any video coding loss (float32 GPU tensor)."""
    codes, mn, mx = codec.normalize(plane, zero_small)
    return denormalize(codes, mn, mx), mn, mx
