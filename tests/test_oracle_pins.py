"""Known-answer pins of the oracle's aux-camera selection, second-chance blend
weight and timewarp (mode 4).

Every scene uses affine view-projection matrices with dyadic entries, so the
reprojection arithmetic is exact and each expected frame follows from the
reference's definition alone (src/renderer.cu):

  - aux selection (:429-464): a main-camera hole takes the best aux colour;
    a valid main pixel is replaced only by an aux result that is not
    visible in the main keyframe and is closer at non-similar depth; among
    aux cameras the first valid one is replaced by a closer one, or, at
    similar depth, by one of higher-or-equal priority and lower error;
  - blend weight (:323-324, :381, :386): the chosen keyframe's colour c_from
    is blended with the matched colour c_to of the other keyframe as
    lerp(c_to, c_from, w), w = 1 - alpha when fwd is chosen and alpha when
    bwd is chosen;
  - timewarp (:467-522): the ray through each target pixel hits the plane
    through the fwd camera's column-3 point with the fwd view direction
    (toward the mean-z point) as normal; the hit is projected into fwd and
    fetched clamped, or into aux[0] when it leaves fwd; |denom| < 1e-6
    writes nothing.

Expected colours are exact rationals floored (for quarter-integer blends the
binary32 quantisation (v/255)*255 truncates to floor(v), pinned exhaustively
in test_oracle_math.py). The plane normal enters timewarp only through
t = dot(p - o, n) / dot(d, n), which is invariant under n -> -n; the cases
below instead pin the plane's position (through the fwd camera point, not the
mean-z point and not the ray origin), the sign of t and the ray direction.
"""
from fractions import Fraction
import math

import numpy as np
import pytest

W, H = 64, 32
AW, AH = 32, 16
EYE4 = np.eye(4)
EMPTY = 0xFEFEFEFEFEFEFEFE


def colmajor(m):
    """4x4 row-major math matrix (M @ v) -> glm column-major float[16]."""
    return np.asarray(m, np.float32).T.reshape(16).copy()


def view(color, z, motion=None, vp=EYE4, ivp=EYE4):
    return {"color": np.ascontiguousarray(color, np.uint8), "z": np.ascontiguousarray(z, np.float32),
            "motion": None if motion is None else np.ascontiguousarray(motion, np.float32),
            "viewproj": colmajor(vp), "inv_viewproj": colmajor(ivp),
            "shadow_info": None, "shadow_color_a": None, "shadow_color_b": None}


def pattern(seed, h=H, w=W, hi=200):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, hi, (h, w, 4), dtype=np.uint8)
    c[..., 3] = 255
    return c


def render(O, fwd, bwd, alpha, prev, priorities=None, mode=0, w=W, h=H, vp=EYE4, ivp=EYE4, max_iter=3):
    info = O.render_info_c(w, h, alpha, colmajor(vp), colmajor(ivp), fpa=4, sfpa=0, mode=mode,
                           max_iter=max_iter)
    out = prev.copy()
    sc = O.render(fwd, bwd, info, out, priorities=priorities, want_scratch=(mode != 4))
    return out, sc


# ---------------------------------------------------------------------------
# Aux-camera selection (src/renderer.cu:429-464)
# ---------------------------------------------------------------------------
Z_BEHIND = 1.0 / 64  # bwd keyframes sit 1/64 NDC behind fwd: not similar at either EPS_2


def main_views(valid, z_main=0.0, seed=10):
    """Main fwd/bwd at W x H. valid: static, fwd closer by Z_BEHIND (so fwd is
    chosen and its second chance fails on depth: the output is exactly the fwd
    colour). invalid: every source leaves the view (a hole everywhere)."""
    zf = np.full((H, W), z_main, np.float32)
    zb = zf + np.float32(Z_BEHIND)
    if valid:
        return view(pattern(seed), zf), view(pattern(seed + 1), zb)
    om = np.zeros((3, H, W), np.float32)
    om[0] = 10.0
    return view(pattern(seed), zf, om), view(pattern(seed + 1), zb, om)


def aux_views(z_aux, seed, flow_y=0.0, valid=True):
    """Aux fwd/bwd at AW x AH, identity cameras. flow_y: a constant clip-space
    offset delta of cam_ff.y (motion 2*delta at weight 1/2), giving the
    reprojection error |delta| (the y axis is the short side) while every
    search still lands on the pixel's own aux cell."""
    zf = np.full((AH, AW), z_aux, np.float32)
    om = np.zeros((3, AH, AW), np.float32)
    om[1] = 2.0 * flow_y
    if not valid:
        om[0] = 10.0
    return view(pattern(seed, AH, AW), zf, om), view(pattern(seed + 1, AH, AW), zf + np.float32(Z_BEHIND), om)


def upsample(aux_color):
    """Render pixel (x, y) reads aux pixel (x // 2, y // 2)."""
    return np.repeat(np.repeat(aux_color, 2, axis=0), 2, axis=1)


def run_aux(O, main, auxes, priorities):
    prev = np.full((H, W, 4), 7, np.uint8)
    fwd = [main[0]] + [a[0] for a in auxes]
    bwd = [main[1]] + [a[1] for a in auxes]
    out, sc = render(O, fwd, bwd, 0.5, prev, priorities=[0] + list(priorities))
    return out, sc, prev


NOT_VISIBLE_NEAR = -1.0 + 2.0 ** -10  # |z| > 1 - 1e-3: outside the main camera; depth 2^-11 > 0


def test_aux_fills_main_hole(oracle_lib):
    main = main_views(valid=False)
    aux = aux_views(0.25, seed=20)
    out, sc, _ = run_aux(oracle_lib, main, [aux], [3])
    assert np.all(sc[(0, 0)]["splat"] == EMPTY) and np.all(sc[(1, 0)]["splat"] == EMPTY)
    np.testing.assert_array_equal(out, upsample(aux[0]["color"]))


def test_aux_not_visible_in_main_and_closer_replaces_main(oracle_lib):
    main = main_views(valid=True, z_main=0.0)
    aux = aux_views(NOT_VISIBLE_NEAR, seed=20)
    out, sc, _ = run_aux(oracle_lib, main, [aux], [3])
    assert np.all(sc[(0, 1)]["cam_ff"][..., 3] == 0.0)  # aux fwd is outside the main fwd camera
    np.testing.assert_array_equal(out, upsample(aux[0]["color"]))


def test_aux_visible_in_main_never_replaces_main(oracle_lib):
    """Same depths ordering (aux closer, not similar) but the aux surface lies
    inside the main keyframe's frustum: the main colour stays."""
    main = main_views(valid=True, z_main=0.0)
    aux = aux_views(-0.5, seed=20)
    out, sc, _ = run_aux(oracle_lib, main, [aux], [3])
    assert np.all(sc[(0, 1)]["cam_ff"][..., 3] == 1.0)
    np.testing.assert_array_equal(out, main[0]["color"])


def test_aux_similar_depth_keeps_main(oracle_lib):
    """Aux not visible in main and nearer, but within EPS_2_LOW_RES of the main
    depth: similar, so the main colour stays."""
    main = main_views(valid=True, z_main=-1.0 + 2.0 ** -9)
    aux = aux_views(NOT_VISIBLE_NEAR, seed=20)
    out, _, _ = run_aux(oracle_lib, main, [aux], [3])
    np.testing.assert_array_equal(out, main[0]["color"])


def test_aux_farther_keeps_main(oracle_lib):
    """Aux not visible in main (beyond the far side, |z| > 1 - 1e-3) and not
    similar, but farther than the main surface: the main colour stays."""
    main = main_views(valid=True, z_main=0.0)
    aux_far = aux_views(1.0 - 2.0 ** -10, seed=20)
    out, sc, _ = run_aux(oracle_lib, main, [aux_far], [3])
    assert np.all(sc[(0, 1)]["cam_ff"][..., 3] == 0.0)
    np.testing.assert_array_equal(out, main[0]["color"])


def test_first_valid_aux_then_closer_wins(oracle_lib):
    main = main_views(valid=False)
    a0 = aux_views(0.25, seed=20)
    a1 = aux_views(0.0, seed=30)  # closer, not similar
    out, _, _ = run_aux(oracle_lib, main, [a0, a1], [3, 3])
    np.testing.assert_array_equal(out, upsample(a1[0]["color"]))
    # order reversed: the closer one is first and a farther one never replaces it
    out, _, _ = run_aux(oracle_lib, main, [a1, a0], [3, 3])
    np.testing.assert_array_equal(out, upsample(a1[0]["color"]))
    # an invalid first aux camera leaves the choice to the second
    bad = aux_views(0.0, seed=40, valid=False)
    out, _, _ = run_aux(oracle_lib, main, [bad, a0], [3, 3])
    np.testing.assert_array_equal(out, upsample(a0[0]["color"]))


@pytest.mark.parametrize("p0,p1,e0,e1,winner", [
    (3, 3, 2.0 ** -8, 2.0 ** -9, 1),   # equal priority, lower error -> new
    (3, 5, 2.0 ** -8, 2.0 ** -9, 1),   # higher priority, lower error -> new
    (5, 3, 2.0 ** -8, 2.0 ** -9, 0),   # lower priority -> keep the first
    (3, 5, 2.0 ** -9, 2.0 ** -8, 0),   # higher error -> keep the first
    (3, 3, 2.0 ** -9, 2.0 ** -9, 0),   # equal error: strict '<' keeps the first
])
def test_similar_depth_aux_priority_and_error(oracle_lib, p0, p1, e0, e1, winner):
    main = main_views(valid=False)
    a0 = aux_views(0.25, seed=20, flow_y=e0)
    a1 = aux_views(0.25, seed=30, flow_y=e1)
    out, sc, _ = run_aux(oracle_lib, main, [a0, a1], [p0, p1])
    # both cameras correspond everywhere with the stated error (searches stay on the own cell)
    for v, e in ((1, e0), (2, e1)):
        np.testing.assert_array_equal(sc[(0, v)]["cam_ff"][..., 1], np.float32(e))
    np.testing.assert_array_equal(out, upsample((a0, a1)[winner][0]["color"]))


# ---------------------------------------------------------------------------
# Second-chance blend weight (src/renderer.cu:297-327, 381, 386)
# ---------------------------------------------------------------------------
def floor_lerp(a, b, t):
    """floor(a + t (b - a)) per channel, exact rationals."""
    a = a.astype(np.int64)
    b = b.astype(np.int64)
    tf = Fraction(t)
    f = np.vectorize(lambda x, y: math.floor(Fraction(int(x)) + tf * (int(y) - int(x))))
    return f(a, b).astype(np.uint8)


def shifted(img, k):
    out = np.zeros_like(img)
    if k >= 0:
        out[:, k:] = img[:, : img.shape[1] - k]
    else:
        out[:, :k] = img[:, -k:]
    return out


@pytest.mark.parametrize("alpha,bwd_err", [(0.25, 0.0), (0.75, 0.0), (0.25, 2.0 ** -10), (0.75, 2.0 ** -10)])
def test_second_chance_blend_weight(oracle_lib, alpha, bwd_err):
    """A 4-pixel translation between the keyframes: at alpha the target pixel p
    sees fwd source p - 4 alpha and bwd source p + 4 (1 - alpha), and the
    second chance matches them (full flows cancel, equal z). The bwd image is
    the fwd image moved by 4 with +40 on RGB, so c_bwd = c_fwd + 40 at the
    matched pair.

    bwd_err = 0: equal errors -> bwd chosen -> lerp(c_fwd, c_bwd, alpha).
    bwd_err > 0: the bwd flow carries a sub-pixel y offset (error 2^-10 <
    EPS_1), so fwd has the lower error -> fwd chosen -> lerp(c_bwd, c_fwd,
    1 - alpha). Either way the result is floor(c_fwd + 40 alpha): the two
    weights swapped would give c_fwd + 40 (1 - alpha)."""
    O = oracle_lib
    c = pattern(7)
    z = np.full((H, W), 0.25, np.float32)
    om_f = np.zeros((3, H, W), np.float32)
    om_f[0] = 4 * (2.0 / W)
    om_b = np.zeros((3, H, W), np.float32)
    om_b[0] = -4 * (2.0 / W)
    # cam_ff.y = (1 - alpha) * om_b.y for the bwd view
    om_b[1] = bwd_err / (1.0 - alpha)
    cb = shifted(c, 4).astype(np.int32)
    cb[..., :3] += 40
    fwd = view(c, z, om_f)
    bwd = view(np.clip(cb, 0, 255).astype(np.uint8), z, om_b)
    prev = np.full((H, W, 4), 3, np.uint8)
    out, sc = render(O, [fwd], [bwd], alpha, prev)
    if bwd_err:
        assert np.all(sc[(1, 0)]["cam_ff"][..., 1] == np.float32(bwd_err))
    s = int(round(4 * alpha))
    c_fwd = shifted(c, s)           # fwd colour seen at p: c[p - 4 alpha]
    c_bwd = c_fwd.copy()
    c_bwd[..., :3] += 40
    expect = floor_lerp(c_fwd, c_bwd, Fraction(alpha)) if not bwd_err else \
        floor_lerp(c_bwd, c_fwd, 1 - Fraction(alpha))
    expect[..., 3] = 255
    lo, hi = 4, W - 4  # interior: both keyframes cover the pixel
    np.testing.assert_array_equal(out[:, lo:hi], expect[:, lo:hi])
    direct = c_fwd.copy()
    direct[..., :3] += np.uint8(int(40 * alpha))
    np.testing.assert_array_equal(out[:, lo:hi, :3], direct[:, lo:hi, :3])


# ---------------------------------------------------------------------------
# Timewarp (src/renderer.cu:467-522, 615-629)
# ---------------------------------------------------------------------------
def ray_camera(c=0.0, sx=0.0, sy=0.0, z0=0.0):
    """Inverse view-projection of the target camera: clip (x, y, z) -> world
    (x + c + sx z, y + sy z, z + z0). Its rays (clip z 0 -> 1) start at
    (x + c, y, z0) and run along (sx, sy, 1)."""
    return np.array([[1, 0, sx, c], [0, 1, sy, 0], [0, 0, 1, z0], [0, 0, 0, 1]], np.float64)


def tw_views(z_mean=0.5, aux=None):
    """fwd keyframe with identity camera: its column-3 point is the world
    origin and its mean-z point (0, 0, z_mean) lies along +z, so the plane is
    world z = 0."""
    f = view(pattern(50), np.full((H, W), z_mean, np.float32))
    fwd = [f]
    if aux is not None:
        fwd.append(aux)
    return fwd


def timewarp(O, fwd, ivp_from, prev):
    out, _ = render(O, fwd, fwd, 0.0, prev, mode=4, ivp=ivp_from)
    return out


def expected_fwd_fetch(img, dx, dy):
    """out[y, x] = img[clamp(y + dy), clamp(x + dx)] for integer pixel shifts."""
    ys = np.clip(np.arange(H) + dy, 0, H - 1)
    xs = np.clip(np.arange(W) + dx, 0, W - 1)
    return img[ys][:, xs]


@pytest.mark.parametrize("c_px", [0, 3, -5])
def test_timewarp_translation_parallel_to_image_plane(oracle_lib, c_px):
    """A target camera translated by c along x (parallel to the fwd image
    plane): every pixel shows the fwd pixel c_px columns over; pixels beyond
    the border fetch the clamped border column (no aux camera)."""
    fwd = tw_views()
    prev = np.full((H, W, 4), 9, np.uint8)
    out = timewarp(oracle_lib, fwd, ray_camera(c=c_px * 2.0 / W, z0=-0.25), prev)
    np.testing.assert_array_equal(out, expected_fwd_fetch(fwd[0]["color"], c_px, 0))


@pytest.mark.parametrize("z0,sx,sy", [(-0.25, 0.5, 0.25), (0.5, 0.25, -0.125), (-0.125, -1.0, 0.5)])
def test_timewarp_ray_meets_plane_through_camera_point(oracle_lib, z0, sx, sy):
    """Sheared rays starting off the plane: the hit point is the ray origin
    moved by -z0 along (sx, sy, 1), i.e. a shift of (c - z0 sx, -z0 sy) in
    clip space. A plane through the ray origin, a plane through the mean-z
    point or a flipped t would each give another shift."""
    c_px = 2
    fwd = tw_views(z_mean=0.5)
    prev = np.full((H, W, 4), 9, np.uint8)
    out = timewarp(oracle_lib, fwd, ray_camera(c=c_px * 2.0 / W, sx=sx, sy=sy, z0=z0), prev)
    dx = Fraction(c_px) + Fraction(-z0) * Fraction(sx) * W / 2
    dy = Fraction(-z0) * Fraction(sy) * H / 2
    assert dx.denominator == 1 and dy.denominator == 1
    np.testing.assert_array_equal(out, expected_fwd_fetch(fwd[0]["color"], int(dx), int(dy)))


def test_timewarp_parallel_ray_writes_nothing(oracle_lib):
    """Rays along (1, 0, eps): denom = eps. eps = 2^-21 < 1e-6 leaves every
    pixel untouched; eps = 2^-19 writes every pixel (the hit is far right, so
    each pixel fetches the clamped last column of its row)."""
    fwd = tw_views()
    prev = pattern(60)
    m = np.array([[1, 0, 1, 0], [0, 1, 0, 0], [0, 0, 2.0 ** -21, -0.25], [0, 0, 0, 1]], np.float64)
    np.testing.assert_array_equal(timewarp(oracle_lib, fwd, m, prev), prev)
    m[2, 2] = 2.0 ** -19
    out = timewarp(oracle_lib, fwd, m, prev)
    np.testing.assert_array_equal(out, np.repeat(fwd[0]["color"][:, W - 1:W], W, axis=1))


@pytest.mark.parametrize("c_px", [6, 40, -12])
def test_timewarp_falls_back_to_aux_zero(oracle_lib, c_px):
    """aux[0]: a half-resolution camera with twice the field of view (clip =
    world / 2). Hits inside fwd fetch fwd; hits outside fwd fetch aux[0] at
    floor((hit / 2 + 1) * res_aux / 2), clamped to the aux border."""
    aux = view(pattern(70, AH, AW), np.full((AH, AW), 0.5, np.float32),
               vp=np.diag([0.5, 0.5, 1.0, 1.0]), ivp=np.diag([2.0, 2.0, 1.0, 1.0]))
    fwd = tw_views(aux=aux)
    prev = np.full((H, W, 4), 9, np.uint8)
    out = timewarp(oracle_lib, fwd, ray_camera(c=c_px * 2.0 / W, z0=-0.25), prev)
    expect = np.zeros_like(prev)
    for y in range(H):
        for x in range(W):
            tx = x + c_px
            if 0 <= tx < W:
                expect[y, x] = fwd[0]["color"][y, tx]
                continue
            hx = Fraction(2 * x + 1, W) - 1 + Fraction(2 * c_px, W)   # hit in clip = world
            hy = Fraction(2 * y + 1, H) - 1
            ax = math.floor((hx / 2 + 1) * AW / 2)
            ay = math.floor((hy / 2 + 1) * AH / 2)
            expect[y, x] = aux["color"][min(max(ay, 0), AH - 1), min(max(ax, 0), AW - 1)]
    np.testing.assert_array_equal(out, expect)
    if c_px == 40:  # some hits also leave the aux view: clamped to its last column
        assert np.any([(Fraction(2 * x + 1, W) - 1 + Fraction(80, W)) / 2 + 1 >= 2 for x in range(W)])
