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


# ---------------------------------------------------------------------------
# Further rewrites in the resolve kernels (each checked against the literal
# operation of the reference's code).
# ---------------------------------------------------------------------------
def test_div_255_as_double_reciprocal():
    """lerp_color's v / 255 (helper_common.h:22-29) as RN32(RN64(v) * RN64(1/255))
    over random binary32 values of every magnitude (subnormal quotients
    included), every quarter-integer in [-512, 512] and every integer."""
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2**31 - 2**23, 4_000_000, dtype=np.int64).astype(np.uint32)  # finite, positive
    v = bits.view(np.float32)
    v = np.concatenate([v, -v, np.arange(-2048, 2049, dtype=np.float32) / np.float32(4),
                        rng.uniform(-300, 300, 2_000_000).astype(np.float32)])
    ref = v / np.float32(255)
    fast = (v.astype(np.float64) * (1.0 / 255.0)).astype(np.float32)
    same = (ref == fast) | (np.isnan(ref) & np.isnan(fast))
    assert same.all(), v[~same][:5]


def test_div_small_magic_multiply():
    """floor(x / r) = umulhi(x, floor(2^32 / r) + 1) for x, r < 2^16 (r > 1)."""
    def fast(x, r):
        m = (2**32 // r + 1) & 0xFFFFFFFF
        return np.where(r == 1, x, (x.astype(np.uint64) * np.uint64(m)) >> np.uint64(32)).astype(np.int64)

    x = np.arange(65536, dtype=np.int64)
    for r in list(range(1, 130)) + [255, 256, 257, 1023, 4095, 32767, 32768, 65534, 65535]:
        assert np.array_equal(fast(x, r), x // r), r
    rng = np.random.default_rng(2)
    for r in rng.integers(2, 65536, 200):
        assert np.array_equal(fast(x, int(r)), x // int(r)), r


def test_staged_row_split_24_bit():
    """The LDS staging splits cell index i into (i / SPAN, i % SPAN) as
    (i * ceil(2^20 / SPAN)) >> 20 on the 24-bit multiplier: exact for every
    i < 2^10 and every row stride 1 < SPAN < 2^10 (the product stays below
    2^32 and both operands below 2^24)."""
    i = np.arange(1024, dtype=np.int64)
    for span in range(2, 1024):
        m = ((1 << 20) + span - 1) // span
        assert m < 2**24 and int(i[-1]) * m < 2**32
        assert np.array_equal((i * m) >> 20, i // span), span


def test_pixel_index_24_bit_products():
    """y * w + x on the 24-bit multiplier is the 32-bit index for views up to
    65535 pixels per axis: the product of two 16-bit values is below 2^32."""
    rng = np.random.default_rng(5)
    w = np.concatenate([rng.integers(1, 65536, 100_000), [1, 65535]]).astype(np.uint64)
    y = (rng.integers(0, 2**16, w.shape).astype(np.uint64) % w)
    low24 = np.uint64(0xFFFFFF)
    prod24 = ((y & low24) * (w & low24)) & np.uint64(0xFFFFFFFF)  # v_mul_u32_u24: low 32 bits
    assert np.array_equal(prod24, y * w)
    assert int((y * w).max()) < 2**32


def test_shadow_code_min_max_bit_parallel():
    """getMinMaxShadowValues (renderer.cu:217-225): the loop over the 2-bit
    codes of sub-frames 0..sfpa-1 against the kernels' bit-parallel form, for
    random 24-bit fields and every sfpa 0..12, plus the quotient
    (my - min) / (max - min) as RN64 reciprocal products (0/0 = NaN)."""
    rng = np.random.default_rng(3)
    bits = np.concatenate([rng.integers(0, 2**24, 200_000), np.array([0, 0xFFFFFF, 0x555555, 0xAAAAAA])])
    for sfpa in range(13):
        mn = np.full(bits.shape, 3)
        mx = np.zeros(bits.shape, dtype=np.int64)
        for i in range(sfpa):
            c = (bits >> (2 * i)) & 3
            mn = np.minimum(mn, c)
            mx = np.maximum(mx, c)
        used = ((1 << (2 * sfpa)) - 1) & 0x555555
        lo, hi = bits & used, (bits >> 1) & used
        has0 = (~lo & ~hi & used) != 0
        has1, has2, has3 = (lo & ~hi) != 0, (hi & ~lo) != 0, (lo & hi) != 0
        fmn = np.where(has0, 0, np.where(has1, 1, np.where(has2, 2, 3)))
        fmx = np.where(has3, 3, np.where(has2, 2, np.where(has1, 1, 0)))
        assert np.array_equal(fmn, mn) and np.array_equal(fmx, mx), sfpa
    with np.errstate(invalid="ignore", divide="ignore"):
        for a in range(4):
            for b in range(4):
                ref = np.float32(a) / np.float32(b)
                inv = {1: 1.0, 2: 0.5, 3: 1.0 / 3.0}.get(b, np.nan)
                fast = np.float32(np.float64(a) * inv)
                assert ref == fast or (np.isnan(ref) and np.isnan(fast)) or (b == 0 and a > 0), (a, b)


def test_pixel_row_from_float_reciprocal():
    """K1's row of a linear pixel, p // w for p < w h (w, h <= 65535), as
    trunc(RN32(RN32(p) * RN32(1/w))) with one remainder correction:
    exhaustive over p for small and awkward widths, sampled for the rest."""
    f32 = np.float32

    def fast(p, w):
        rcp = f32(1.0 / w)
        q = (p.astype(f32) * rcp).astype(np.float64).astype(np.int64)  # v_cvt_u32_f32 truncates
        r = p - q * w
        return np.where(r < 0, q - 1, np.where(r >= w, q + 1, q))

    rng = np.random.default_rng(9)
    widths = list(range(1, 65)) + [97, 127, 255, 511, 640, 960, 1023, 1280, 1920, 2064, 4095, 4096, 8191, 65535] \
        + [int(v) for v in rng.integers(65, 65536, 30)]
    for w in widths:
        h = min(65535, (2**28 - 1) // w)
        n = w * h
        p = np.arange(n, dtype=np.int64) if n <= 1_000_000 else np.unique(np.concatenate([
            rng.integers(0, n, 300_000), np.arange(w) * 0 + n - 1 - np.arange(w), (np.arange(1, h) * w)[-5000:],
            (np.arange(1, h) * w)[-5000:] - 1]))
        assert np.array_equal(fast(p, w), p // w), w


def _visible_fast(c, w, bound_factor):
    return np.abs(c.astype(np.float64)) <= bound_factor * np.abs(w.astype(np.float64))


def test_visibility_bound_without_division():
    """K1's main-view visibility -lim < RN32(c / w) < lim (renderer.cu:105-107,
    lim = RN32(1 - RN32(1e-3))) is evaluated as |c| <= LIM_MID |w| in binary64
    for finite non-zero w, LIM_MID = 33520877 / 2^25 the midpoint below lim.
    Checked on random operands over wide exponent ranges (subnormal w
    included) and on c within two ulps of the bound, where a wrong bound or an
    inexact product shows (using lim itself as the bound fails here)."""
    lim = np.float32(1.0) - np.float32(1e-3)
    lim_mid = 33520877.0 / 33554432.0
    assert (float(lim) + float(np.nextafter(lim, np.float32(0)))) / 2 == lim_mid
    rng = np.random.default_rng(12)
    n = 400_000
    w = (rng.uniform(1.0, 2.0, n) * 2.0 ** rng.integers(-140, 120, n)).astype(np.float32)
    w *= np.where(rng.random(n) < 0.5, -1, 1).astype(np.float32)
    w = w[(w != 0) & np.isfinite(w)]
    up, dn = np.float32(np.inf), np.float32(-np.inf)
    near = (lim_mid * w.astype(np.float64)).astype(np.float32)  # nearest binary32 to the bound
    cands = {"near": near, "up1": np.nextafter(near, up), "dn1": np.nextafter(near, dn)}
    cands["up2"] = np.nextafter(cands["up1"], up)
    cands["dn2"] = np.nextafter(cands["dn1"], dn)
    cands["random"] = (w.astype(np.float64) * rng.uniform(-2.0, 2.0, w.size)).astype(np.float32)
    mixed = 0
    with np.errstate(over="ignore", under="ignore", invalid="ignore", divide="ignore"):
        for sign in (np.float32(1), np.float32(-1)):
            for name, c in cands.items():
                c = c * sign
                ok = np.isfinite(c)
                cc, ww = c[ok], w[ok]
                q = cc / ww  # binary32 IEEE division
                ref = (q > -lim) & (q < lim)
                assert np.array_equal(ref, _visible_fast(cc, ww, lim_mid)), name
                mixed += bool(ref.any() and not ref.all())
                if name == "near":  # the wrong bound (lim itself) differs here
                    assert not np.array_equal(ref, _visible_fast(cc, ww, float(lim)))
    assert mixed >= 4  # the near-bound sets hold both outcomes


# ---------------------------------------------------------------------------
# Reprojection errors and the second-chance flow distance compared without
# their square roots (resolve_device.hpp: reproj_error2, sqrt_less, decide,
# second_chance; renderer.cu:204-215, 315-324, 366-396, 441-447).
# ---------------------------------------------------------------------------
EPS_SQ = {  # isti_device.hpp: bit patterns of EPS_1_FULL_RES_SQ, EPS_1_LOW_RES_SQ, EPS_3_SQ
    0.002: 0x368637BD,
    0.008: 0x388637BD,
    0.01: 0x38D1B716,
}


def test_squared_thresholds_are_the_exact_boundaries():
    """RN32(sqrt(a)) < eps exactly when a < T (RN32 o sqrt is monotone, so the
    boundary pair decides every a), for the three thresholds of the resolve."""
    for eps, bits in EPS_SQ.items():
        e = np.float32(eps)
        t = np.array([bits], np.uint32).view(np.float32)[0]
        below = np.nextafter(t, np.float32(0))
        assert np.sqrt(t) >= e and np.sqrt(below) < e, eps
        # and around it, every binary32 within 2^16 ulps either side
        a = (np.arange(-65536, 65537, dtype=np.int64) + bits).astype(np.uint32).view(np.float32)
        assert np.array_equal(np.sqrt(a) < e, a < t), eps
    assert np.float32(0.002) * np.float32(4.0) == np.float32(0.008)  # EPS_1_LOW_RES = 4 * EPS_1_FULL_RES


def _sqrt_less(a, b):
    """The kernels' sqrt_less, vectorised: a < b, then the 2^-18 margin, then the square roots."""
    lt = a < b
    with np.errstate(over="ignore", invalid="ignore"):
        clear = (a >= np.float32(2.0 ** -100)) & (b > a * np.float32(1 + 2.0 ** -18))
        exact = np.sqrt(a) < np.sqrt(b)
    return lt & (clear | exact)


def test_sqrt_less_equals_comparing_square_roots():
    """sqrt_less(a, b) == (RN32(sqrt(a)) < RN32(sqrt(b))) on near-ties (b a few
    ulps above a, where the roots often round together), on pairs just above and
    below the margin, on every magnitude (subnormals included) and on 0, +inf and
    NaN. Both sides run in binary32 with IEEE square roots."""
    rng = np.random.default_rng(7)
    bits = rng.integers(1, 0x7F800000, 2_000_000, dtype=np.int64)
    a = bits.astype(np.uint32).view(np.float32)
    pairs = [(a, (bits + rng.integers(-8, 9, bits.size)).clip(0, 0x7F800000).astype(np.uint32).view(np.float32))]
    for k in (-2, -1, 0, 1, 2):  # around the margin b = RN32(a (1 + 2^-18))
        m = (a * np.float32(1 + 2.0 ** -18)).view(np.uint32).astype(np.int64) + k
        pairs.append((a, m.clip(0, 0x7F800000).astype(np.uint32).view(np.float32)))
    pairs.append((a, rng.permutation(a)))
    sub = rng.integers(0, 1 << 23, 200_000, dtype=np.int64)
    sa = sub.astype(np.uint32).view(np.float32)
    pairs.append((sa, (sub + rng.integers(0, 4, sub.size)).astype(np.uint32).view(np.float32)))
    special = np.array([0.0, np.inf, np.nan, 1e-45, 4e-6, 6.4e-5, 3.4e38], np.float32)
    x, y = np.meshgrid(special, special)
    pairs.append((x.ravel(), y.ravel()))
    for p, q in pairs:
        with np.errstate(invalid="ignore"):
            ref = np.sqrt(p) < np.sqrt(q)
        assert np.array_equal(_sqrt_less(p, q), ref)
        with np.errstate(invalid="ignore"):
            assert np.array_equal(_sqrt_less(q, p), np.sqrt(q) < np.sqrt(p))


def test_one_valid_error_decides_the_comparison():
    """decide(): with exactly one of fe, be below EPS_1 the reference's fe < be is
    "fe is the valid one and be is not NaN"; the kernel reads it off the flags."""
    e1 = np.float32(0.002)
    vals = np.array([0.0, 1e-4, 0.0019, 0.002, 0.0021, 1.0, np.float32(3.4e38), np.inf, np.nan], np.float32)
    fe, be = np.meshgrid(vals, vals)
    fe, be = fe.ravel(), be.ravel()
    with np.errstate(invalid="ignore"):
        fok, bok = fe < e1, be < e1
        one = fok ^ bok
        assert np.array_equal((fe < be)[one], (fok & ~np.isnan(be))[one])
