"""Closed-form and exhaustive pins of the oracle's math helpers.

Reference lines: include/util/helper_common.h:8-95 (pixel/NDC maps, colour
quantisation), src/renderer.cu:59-74 (splat key codec), :204-215 (reprojection
error), :359-364 (__float2uint_rn/rd/ru), src/util/image_buffer.cu:80-101
(mean z). Each check below is fixed by mathematics, not by re-running the
oracle's own formula.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

RES = [(640, 360), (1920, 1080), (960, 540), (2064, 2208), (7, 5), (1, 1)]


@pytest.mark.parametrize("w,h", RES)
def test_pixel_roundtrip_all_pixels(oracle_lib, w, h):
    """CS2pixel(pixel2CS(p)) == p for every pixel (helper_common.h:66-75)."""
    L = oracle_lib.lib()
    out2 = (ctypes.c_float * 2)()
    outi = (ctypes.c_int32 * 2)()
    xs = np.unique(np.r_[np.arange(min(w, 64)), np.linspace(0, w - 1, 97).astype(int)])
    ys = np.unique(np.r_[np.arange(min(h, 64)), np.linspace(0, h - 1, 97).astype(int)])
    for y in ys:
        for x in xs:
            L.oracle_pixel2cs(int(x), int(y), w, h, out2)
            # pixel centre in NDC is within one float ulp of the exact value
            exact_x = 2.0 * (x + 0.5) / w - 1.0
            assert abs(out2[0] - exact_x) <= 2e-7 * max(1.0, abs(exact_x)) + 1e-7
            L.oracle_cs2pixel(out2[0], out2[1], w, h, outi)
            assert (outi[0], outi[1]) == (x, y)


def test_cs2pixel_edges(oracle_lib):
    L = oracle_lib.lib()
    outi = (ctypes.c_int32 * 2)()
    L.oracle_cs2pixel(-1.0, -1.0, 640, 360, outi)
    assert tuple(outi) == (0, 0)
    L.oracle_cs2pixel(1.0, 1.0, 640, 360, outi)  # right/bottom edge maps one past the end
    assert tuple(outi) == (640, 360)
    L.oracle_cs2pixel(-1.0001, 0.0, 640, 360, outi)  # floor, not truncation, below zero
    assert outi[0] == -1
    L.oracle_cs2pixel(float("nan"), 0.0, 640, 360, outi)  # GPU conversion: NaN -> 0
    assert outi[0] == 0


def test_mat_vec_exact_on_integers(oracle_lib):
    """Column-major product; small integers make every association exact."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(0)
    for _ in range(50):
        M = rng.integers(-20, 20, (4, 4)).astype(np.float32)
        v = rng.integers(-20, 20, 4).astype(np.float32)
        m16 = np.ascontiguousarray(M.T.reshape(16))  # column-major
        out = np.zeros(4, np.float32)
        L.oracle_mat_vec(m16.ctypes.data, v.ctypes.data, out.ctypes.data)
        np.testing.assert_array_equal(out, M @ v)


def test_cs2ws_ws2cs_inverse(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(1)
    A = rng.normal(size=(4, 4)) + 4 * np.eye(4)
    Ai = np.linalg.inv(A)
    a16 = np.ascontiguousarray(A.T.reshape(16), np.float32)
    ai16 = np.ascontiguousarray(Ai.T.reshape(16), np.float32)
    for _ in range(20):
        p = np.r_[rng.uniform(-1, 1, 3), 1.0].astype(np.float32)
        w = np.zeros(4, np.float32)
        back = np.zeros(4, np.float32)
        L.oracle_cs2ws(p.ctypes.data, ai16.ctypes.data, w.ctypes.data)
        assert w[3] == 1.0  # divided by w
        L.oracle_ws2cs(w.ctypes.data, a16.ctypes.data, back.ctypes.data)
        np.testing.assert_allclose(back[:3], p[:3], atol=2e-5)


def test_key_codec_roundtrip_and_order(oracle_lib):
    """Key = depth bits << 32 | x << 16 | y; u64 order == depth order for depth >= 0."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(2)
    d = ctypes.c_float()
    x = ctypes.c_uint32()
    y = ctypes.c_uint32()
    keys, depths = [], []
    for _ in range(300):
        depth = float(np.float32(rng.uniform(0, 1) ** 3))
        sx, sy = int(rng.integers(0, 65536)), int(rng.integers(0, 65536))
        k = L.oracle_encode_key(depth, sx, sy)
        assert k >> 32 == np.float32(depth).view(np.uint32)
        assert (k >> 16) & 0xFFFF == sx and k & 0xFFFF == sy
        L.oracle_decode_key(k, ctypes.byref(d), ctypes.byref(x), ctypes.byref(y))
        assert (np.float32(d.value), x.value, y.value) == (np.float32(depth), sx, sy)
        keys.append(k)
        depths.append(np.float32(depth))
    order_k = np.argsort(np.array(keys, dtype=np.uint64), kind="stable")
    assert np.all(np.diff(np.array(depths)[order_k]) >= 0)
    # a NaN depth (either sign) takes the canonical NaN 0x7fffffff of the
    # reference's GPUs: after every finite or infinite depth, before the clear value
    for nan_bits in (0x7FC00000, 0xFFC00000, 0x7F800001):
        nan = float(np.array([nan_bits], np.uint32).view(np.float32)[0])
        k = L.oracle_encode_key(nan, 5, 7)
        assert k == (0x7FFFFFFF << 32) | (5 << 16) | 7
        assert L.oracle_encode_key(float("inf"), 0xFFFF, 0xFFFF) < k < 0xFEFEFEFEFEFEFEFE
    # the 0xfe.. clear value decodes to a negative depth (an empty cell)
    L.oracle_decode_key(0xFEFEFEFEFEFEFEFE, ctypes.byref(d), ctypes.byref(x), ctypes.byref(y))
    assert d.value < 0 and math.isfinite(d.value) and x.value == 0xFEFE and y.value == 0xFEFE


def test_reproj_error_hand_values(oracle_lib):
    L = oracle_lib.lib()
    # q = target pixel, zero motion -> zero error
    out2 = (ctypes.c_float * 2)()
    L.oracle_pixel2cs(10, 20, 640, 360, out2)
    assert L.oracle_reproj_error(10, 20, 640, 360, out2[0], out2[1], 0.0, 0.0) == 0.0
    # one pixel to the right in a 640x360 view: NDC 2/640, scaled by 640/360 -> 2/360
    L.oracle_pixel2cs(11, 20, 640, 360, out2)
    e = L.oracle_reproj_error(10, 20, 640, 360, out2[0], out2[1], 0.0, 0.0)
    assert e == pytest.approx(2.0 / 360.0, rel=1e-5)
    # motion that lands exactly on the target cancels
    e = L.oracle_reproj_error(10, 20, 640, 360, out2[0], out2[1], 2.0 / 640.0, 0.0)
    assert e == pytest.approx(0.0, abs=1e-6)
    # 3-4-5 triangle in a square view
    L.oracle_pixel2cs(0, 0, 100, 100, out2)
    e = L.oracle_reproj_error(0, 0, 100, 100, out2[0] - 0.03, out2[1] - 0.04, 0.0, 0.0)
    assert e == pytest.approx(0.05, rel=1e-5)


def test_float2uint_rounding(oracle_lib):
    L = oracle_lib.lib()
    rn, rd, ru = L.oracle_float2uint_rn, L.oracle_float2uint_rd, L.oracle_float2uint_ru
    assert [rn(v) for v in (0.4, 0.5, 1.5, 2.5, 3.5, 2.6)] == [0, 0, 2, 2, 4, 3]  # ties to even
    assert [rd(v) for v in (0.0, 2.7, 3.0, 1e-9)] == [0, 2, 3, 0]
    assert [ru(v) for v in (0.0, 2.1, 3.0, -1e-9)] == [0, 3, 3, 0]
    assert rn(-3.0) == 0 and rn(float("nan")) == 0 and rd(5e9) == 0xFFFFFFFF


def test_to_8bit_truncates(oracle_lib):
    L = oracle_lib.lib()
    assert L.oracle_to_8bit(1.0) == 255
    assert L.oracle_to_8bit(0.5) == 127  # 127.5 truncates
    assert L.oracle_to_8bit(-0.2) == 0 and L.oracle_to_8bit(3.0) == 255
    assert L.oracle_to_8bit(float("nan")) == 0


def test_lerp_color_truncation_exhaustive(oracle_lib):
    """lerp_color + color_to_uchar4 (helper_common.h:22-38) over all channel pairs.

    The exact blend v = a + t(b-a) is a multiple of 1/4 for these t. In IEEE
    binary32, (v/255)*255 never lands below such a v (every quarter-integer
    in [0, 255] round-trips), so after TRUNCATION the result is exactly
    floor(v). Rounding instead of truncating fails at t=0.5 and t=0.25.
    """
    L = oracle_lib.lib()
    a4 = (ctypes.c_uint8 * 4)()
    b4 = (ctypes.c_uint8 * 4)()
    o4 = (ctypes.c_uint8 * 4)()
    for t in (0.0, 0.25, 0.5, 1.0):
        for a in range(0, 256, 3):
            for b in range(0, 256, 5):
                a4[:] = [a, b, 255, 0]
                b4[:] = [b, a, 0, 255]
                L.oracle_lerp_color(a4, b4, t, o4)
                for ch, (ua, ub) in enumerate(((a, b), (b, a), (255, 0), (0, 255))):
                    v = Fraction(ua) + Fraction(t) * (ub - ua)
                    fl = math.floor(v)
                    assert o4[ch] == fl, (a, b, t, ch, o4[ch], v)


def test_mean_z_exact(oracle_lib):
    """Deterministic double reduction; on dyadic data every sum is exact."""
    rng = np.random.default_rng(3)
    z = (rng.integers(-512, 512, (37, 129)) / 1024.0).astype(np.float32)
    expect = math.fsum(z.astype(np.float64).ravel()) / z.size
    assert oracle_lib.mean_z(z) == np.float32(expect)


def test_search_field_and_window_starts_compose_the_pixel_routines(oracle_lib):
    """The whole-render helpers (used to check learned starts) agree with the
    per-pixel init_search / search / reproj_error entry points, and a search
    started at its own result stays there (fixed point)."""
    from tests import _frames as F

    O = oracle_lib
    L = O.lib()
    cfg = F.scaled(1, 3, 4)
    fwd, bwd = F.pair(cfg, 0)
    vp, ivp, w, h = F.target(cfg, 0, 2)
    _, sc = F.oracle_frame(O, fwd, bwd, vp, ivp, 0.5, 4, 0, max_iter=8, no_shadows=True)
    cam, splat = sc[(1, 0)]["cam_ff"], sc[(1, 0)]["splat"]
    ws = O.window_starts(splat, 1)
    q, err = O.search_field(cam, ws, 8)
    rng = np.random.default_rng(0)
    d, sx, sy, qx, qy, pcs = (ctypes.c_float(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(),
                              ctypes.c_uint32(), (ctypes.c_float * 2)())
    cam_c = np.ascontiguousarray(cam, np.float32)
    for y, x in zip(rng.integers(0, h, 400), rng.integers(0, w, 400)):
        L.oracle_init_search(splat.ctypes.data, w, h, int(x), int(y), 1, ctypes.byref(d), ctypes.byref(sx),
                             ctypes.byref(sy))
        if not d.value > 0:
            assert ws[y, x] == 0xFFFFFFFF and q[y, x] == 0xFFFFFFFF and err[y, x] == np.finfo(np.float32).max
            continue
        assert ws[y, x] == (sy.value << 16 | sx.value)
        L.oracle_search(cam_c.ctypes.data, w, h, sx.value, sy.value, int(x), int(y), w, h, 8, ctypes.byref(qx),
                        ctypes.byref(qy))
        assert q[y, x] == (qy.value << 16 | qx.value)
        L.oracle_pixel2cs(int(x), int(y), w, h, pcs)
        m = cam_c[qy.value, qx.value]
        e = L.oracle_reproj_error(qx.value, qy.value, w, h, pcs[0], pcs[1], float(m[0]), float(m[1]))
        assert err[y, x] == np.float32(e)
    # a converged search restarted at its result does not move
    conv = (q != 0xFFFFFFFF)
    q2, err2 = O.search_field(cam, np.where(conv, q, np.uint32(0xFFFFFFFF)), 8)
    q1, _ = O.search_field(cam, np.where(conv, q, np.uint32(0xFFFFFFFF)), 1)
    fixed = conv & (q1 == q)
    assert fixed.mean() > 0.5
    assert np.array_equal(q2[fixed], q[fixed]) and np.array_equal(err2[fixed], err[fixed])
