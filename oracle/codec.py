"""CPU restatement of the keyframe wire codec (the reference's
compression.py) - the checker for include/isti_codec.h's GPU kernels.

TEST INFRASTRUCTURE ONLY: imported by the test suite, never by the product
path. Each function names the lines it restates; the arithmetic is the
reference's (numpy, binary32 where the reference computes in float32).
"""
from __future__ import annotations

import numpy as np


def rle_encode(data: bytes) -> bytes:
    """compression.py:176-204: runs of equal bytes as (count, value) pairs,
    pieces of at most 255, a count of 1 written as 0."""
    arr = np.frombuffer(data, dtype=np.uint8)
    n = len(arr)
    if n == 0:
        return b""
    change = np.nonzero(np.diff(arr) != 0)[0] + 1
    starts = np.concatenate(([0], change, [n]))
    out = bytearray()
    for i in range(len(starts) - 1):
        value = int(arr[starts[i]])
        run = int(starts[i + 1] - starts[i])
        while run > 0:
            count = min(run, 255)
            out.append(count if count > 1 else 0)
            out.append(value)
            run -= count
    return bytes(out)


def rle_decode(data: bytes) -> bytes:
    """compression.py:206-218."""
    out = bytearray()
    for i in range(0, len(data), 2):
        count, value = data[i], data[i + 1]
        out.extend([value] * (count if count else 1))
    return bytes(out)


def xor_delta(frames, server_fps: int):
    """compression.py:249-252: frame k is kept when k % server_fps == 0 (the
    anchor) and XOR-ed with the latest anchor otherwise."""
    out, anchor = [], None
    for k, f in enumerate(frames):
        f = np.asarray(f, np.uint8).ravel()
        if k % server_fps == 0:
            anchor = f.copy()
            out.append(f.copy())
        else:
            out.append(np.bitwise_xor(f, anchor))
    return out


def load_exr_clean(img: np.ndarray) -> np.ndarray:
    """compression.py:63: |v| < 1e-4 reads as 0 (binary32 data)."""
    img = np.asarray(img, np.float32)
    return np.where(np.abs(img) < np.float32(0.0001), np.float32(0.0), img).astype(np.float32)


def normalize_exr(img: np.ndarray):
    """compression.py:84-91 for binary32 images: (codes uint16, min, max)."""
    img = np.asarray(img, np.float32)
    mn, mx = np.min(img), np.max(img)
    if abs(mx - mn) <= 1e-6:
        return np.zeros_like(img, dtype=np.uint16), np.float32(0), np.float32(0)
    t = np.clip((img - mn) / (mx - mn), np.float32(0.0), np.float32(1.0)) * np.float32(65535)
    return np.round(t).astype(np.uint16), mn, mx


def denormalize_exr(codes: np.ndarray, mn, mx) -> np.ndarray:
    """compression.py:93-98: to half precision (returned as the float32 values)."""
    mn, mx = np.float32(mn), np.float32(mx)
    if abs(mx - mn) <= 1e-6:
        return np.zeros(codes.shape, np.float32)
    v = (codes.astype(np.float32) / np.float32(65535)) * np.float32(mx - mn) + mn
    return v.astype(np.float16).astype(np.float32)
