"""Exactness of the rewritten arithmetic the HIP kernels use in place of
IEEE binary32 divisions (checked on the CPU; the GPU parity tests then check
the kernels end to end).

pixel2CS (include/util/helper_common.h:66-69) computes
    RN32(RN32(2*(x+0.5)) / w) - 1 = RN32((2x+1)/w) - 1.
The kernels evaluate RN32((2x+1)/w) as RN32(RN64((2x+1) * RN64(1/w))).
For integers 2x+1 < 2^17 and w < 2^16 the exact quotient is never a binary32
rounding midpoint, and it lies at least 2^-41 (relative) away from every
midpoint, while the double product is within 2^-52 of it; so both roundings
agree. The tests below check that claim exhaustively for every w <= 4096 and
every pixel of the workload sizes, and on random samples up to 65535.
"""
import numpy as np


def _ref(x, w):
    return ((np.float32(2.0) * (x.astype(np.float32) + np.float32(0.5))) / np.float32(w)).astype(np.float32)


def _fast(x, w):
    inv = np.float64(1.0) / np.float64(w)
    return ((2 * x.astype(np.float64) + 1.0) * inv).astype(np.float32)


def test_pixel_division_all_small_sizes():
    for w in range(1, 4097):
        x = np.arange(w, dtype=np.int64)
        assert np.array_equal(_ref(x, w), _fast(x, w)), w


def test_pixel_division_workload_sizes():
    for w in (1920, 1080, 960, 540, 640, 360, 2064, 2208, 1280, 720, 4096, 8192, 16384, 65535):
        x = np.arange(w, dtype=np.int64)
        assert np.array_equal(_ref(x, w), _fast(x, w)), w


def test_pixel_division_random_large():
    rng = np.random.default_rng(0)
    for w in rng.integers(4097, 65536, 400):
        x = rng.integers(0, w, 4000)
        assert np.array_equal(_ref(x, int(w)), _fast(x, int(w))), w
