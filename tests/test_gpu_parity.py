"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Bar (per case): splat winner buffers bit-exact; cam_ff / full_ff within
1e-3 px; final RGBA within 1/255 max-abs and >= 50 dB PSNR (RGB, as
scripts/eval.py:57,68 of the reference). Every case starts from a random
previous output, so hole retention (write-only-if-valid, renderer.cu:416-419,
453-464) is checked too.
"""
import numpy as np
import pytest

from tests import _frames as F

pytestmark = pytest.mark.gpu


def _prev(h, w, seed=11):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4), dtype=np.uint8)


def _check(O, cfg, fwd, bwd, b, mode=0, max_iter=-1, no_shadows=False, alpha=None, k=0):
    fpa = cfg.keyframe_spacing
    sfpa = cfg.shadow_frames if cfg.shadows else 0
    vp, ivp, w, h = F.target(cfg, k, b)
    a = b / fpa if alpha is None else alpha
    prev = _prev(h, w)
    print(f"case: {w}x{h} n_aux={len(fwd)-1} mode={mode} alpha={a} iters={max_iter}", flush=True)
    ref, sref = F.oracle_frame(O, fwd, bwd, vp, ivp, a, fpa, sfpa, mode, max_iter, no_shadows, prev)
    got, sgot = F.gpu_frame(fwd, bwd, vp, ivp, a, fpa, sfpa, mode, max_iter, no_shadows, prev)
    if mode != 4:
        for key, s in sref.items():
            g = sgot[key]
            np.testing.assert_array_equal(g["splat"], s["splat"], err_msg=f"splat {key}")
            assert F.flow_err_px(g["cam_ff"], s["cam_ff"]) <= 1e-3, f"cam_ff {key}"
            assert F.flow_err_px(g["full_ff"], s["full_ff"]) <= 1e-3, f"full_ff {key}"
            F.assert_visible_flags(g, s, key)
            # the device record holds the moved z, RN32(z + flow z), where the oracle keeps the flow z
            z = (fwd if key[0] == 0 else bwd)[key[1]]["z"]
            np.testing.assert_array_equal(g["cam_ff"][..., 2], z + s["cam_ff"][..., 2], err_msg=f"cam_ff z {key}")
    F.assert_rgb_parity(got, ref)
    return got, ref


@pytest.fixture(scope="module")
def c1():
    cfg = F.scaled(1, 5)
    return cfg, F.pair(cfg, 0)


@pytest.mark.parametrize("b", [1, 2, 3])
def test_c1_mode0(oracle_lib, c1, b):
    cfg, (fwd, bwd) = c1
    _check(oracle_lib, cfg, fwd, bwd, b)


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_c1_other_modes(oracle_lib, c1, mode):
    cfg, (fwd, bwd) = c1
    _check(oracle_lib, cfg, fwd, bwd, 2, mode=mode)


def test_c1_late_keyframe_extrapolation(oracle_lib, c1):
    """alpha > 1 (latency, main.cu:157): fwd search skipped, no blending."""
    cfg, (fwd, bwd) = c1
    _check(oracle_lib, cfg, fwd, bwd, 1, alpha=1.25)


def test_c1_no_object_motion(oracle_lib, c1):
    cfg, (fwd, bwd) = c1
    fwd2 = [F.strip(v, motion=False) for v in fwd]
    bwd2 = [F.strip(v, motion=False) for v in bwd]
    _check(oracle_lib, cfg, fwd2, bwd2, 2)


def test_c1_search_cap_32(oracle_lib, c1):
    cfg, (fwd, bwd) = c1
    _check(oracle_lib, cfg, fwd, bwd, 3, max_iter=32)


@pytest.fixture(scope="module")
def c2_small():
    cfg = F.scaled(2, 9, div=4)  # 480x270, aux 240x135, shadows on
    return cfg, F.pair(cfg, 0)


@pytest.mark.parametrize("b", [1, 3, 4, 7])
def test_c2_small_shadows_aux(oracle_lib, c2_small, b):
    cfg, (fwd, bwd) = c2_small
    _check(oracle_lib, cfg, fwd, bwd, b, max_iter=32)


@pytest.mark.parametrize("iters", [1, 2, 5, 6, 7, 31, 33])
def test_c2_small_search_cycles(oracle_lib, c2_small, iters):
    """The aux searches fall into short cycles; the kernel reads q_N off the
    cycle instead of stepping, which must agree with N plain steps for every
    N (odd and even, shorter and longer than the cycle entry)."""
    cfg, (fwd, bwd) = c2_small
    _check(oracle_lib, cfg, fwd, bwd, 6, max_iter=iters)


def test_c2_small_no_shadows_flag(oracle_lib, c2_small):
    cfg, (fwd, bwd) = c2_small
    _check(oracle_lib, cfg, fwd, bwd, 5, max_iter=32, no_shadows=True)


def test_c2_small_missing_shadow_planes(oracle_lib, c2_small):
    """Absent planes read as the reference's texture handle -1 (renderer.cu:277)."""
    cfg, (fwd, bwd) = c2_small
    fwd2 = [F.strip(v, shadows=False) for v in fwd]
    bwd2 = [dict(v, shadow_color_b=None) for v in bwd]
    _check(oracle_lib, cfg, fwd2, bwd2, 3, max_iter=32)


@pytest.mark.parametrize("mode", [1, 3, 4])
def test_c2_small_modes(oracle_lib, c2_small, mode):
    cfg, (fwd, bwd) = c2_small
    _check(oracle_lib, cfg, fwd, bwd, 2, mode=mode)


def test_c3_small_stereo_eye(oracle_lib):
    cfg = F.scaled(3, 4, div=4)  # 516x552 per eye, 64 spheres on sinusoidal paths
    for eye in (0, 1):
        fwd, bwd = F.pair(cfg, 0, eye)
        fpa = cfg.keyframe_spacing
        vp, ivp, w, h = F.target(cfg, 0, 3, eye)
        prev = _prev(h, w)
        ref, _ = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, 0.5, fpa, 0, prev=prev)
        got, _ = F.gpu_frame(fwd, bwd, vp, ivp, 0.5, fpa, 0, prev=prev, want_scratch=False)
        F.assert_rgb_parity(got, ref)


def test_c3_full_stereo_eyes(oracle_lib):
    """Config C3 at its real shape: both 2064x2208 eyes of a fast-camera frame
    (0.15 m + 2 degrees per frame, 64 moving spheres), splat buffers, flows
    and pixels."""
    cfg = F.scaled(3, 4)
    fpa = cfg.keyframe_spacing
    for eye in (0, 1):
        fwd, bwd = F.pair(cfg, 0, eye)
        vp, ivp, w, h = F.target(cfg, 0, 3, eye)
        prev = _prev(h, w)
        print(f"case: C3 eye {eye} {w}x{h}", flush=True)
        ref, sref = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, 3 / fpa, fpa, 0, prev=prev)
        got, sgot = F.gpu_frame(fwd, bwd, vp, ivp, 3 / fpa, fpa, 0, prev=prev)
        for key, s in sref.items():
            g = sgot[key]
            np.testing.assert_array_equal(g["splat"], s["splat"], err_msg=f"splat {key}")
            assert F.flow_err_px(g["cam_ff"], s["cam_ff"]) <= 1e-3, f"cam_ff {key}"
            assert F.flow_err_px(g["full_ff"], s["full_ff"]) <= 1e-3, f"full_ff {key}"
        F.assert_rgb_parity(got, ref)


def test_c2_full_1080p(oracle_lib):
    """Full-size workload: 1920x1080, 2 aux 960x540, shadows, search cap 32."""
    cfg = F.scaled(2, 3)
    fwd, bwd = F.pair(cfg, 0)
    _check(oracle_lib, cfg, fwd, bwd, 4, max_iter=32)


def test_degenerate_unprojection_nan_column(oracle_lib, c1):
    """The fwd keyframe's unprojection has w == 0 on one pixel column (the
    projective division w / w of CS2WS, helper_common.h:77-87, is then NaN):
    its flows are NaN on both sides and every other value matches. NaN depths
    reach the splat as keys with the canonical NaN 0x7fffffff (the reference's
    GPUs return it), which win cells no number reaches: the splat buffers
    are compared byte for byte; flows match NaN for NaN."""
    cfg, (fwd, bwd) = c1
    w = cfg.width
    x0 = w // 3
    cx0 = np.float32(2.0) * (np.float32(x0) + np.float32(0.5)) / np.float32(w) - np.float32(1.0)
    ivp = fwd[0]["inv_viewproj"].copy()
    ivp[3], ivp[7], ivp[11], ivp[15] = 1.0, 0.0, 0.0, -cx0  # w row: cs.x - cs.x(x0)
    fwd2 = [dict(fwd[0], inv_viewproj=ivp)] + list(fwd[1:])
    fpa = cfg.keyframe_spacing
    vp, tivp, tw, th = F.target(cfg, 0, 2)
    prev = _prev(th, tw)
    print(f"case: {tw}x{th} degenerate unprojection column {x0}", flush=True)
    ref, sref = F.oracle_frame(oracle_lib, fwd2, bwd, vp, tivp, 2 / fpa, fpa, 0, prev=prev)
    got, sgot = F.gpu_frame(fwd2, bwd, vp, tivp, 2 / fpa, fpa, 0, prev=prev)
    cf = sgot[(0, 0)]["cam_ff"]  # full_ff is NaN there only where the motion is zero
    assert np.isnan(cf[:, x0, 0]).all() and not np.isnan(cf[:, x0 + 1, 0]).any()
    nan_keys = (sgot[(0, 0)]["splat"] >> np.uint64(32)) == np.uint64(0x7FFFFFFF)
    assert nan_keys.any()  # the case reaches the NaN-key path
    for key, s in sref.items():
        g = sgot[key]
        np.testing.assert_array_equal(g["splat"], s["splat"], err_msg=f"splat {key}")
        assert F.flow_err_px(g["cam_ff"], s["cam_ff"]) <= 1e-3, f"cam_ff {key}"
        assert F.flow_err_px(g["full_ff"], s["full_ff"]) <= 1e-3, f"full_ff {key}"
        F.assert_visible_flags(g, s, key)
    F.assert_rgb_parity(got, ref)


def test_repeat_run_bit_identical(c1):
    """K1's atomicMin is order-independent: two renders give identical bytes."""
    cfg, (fwd, bwd) = c1
    vp, ivp, w, h = F.target(cfg, 0, 2)
    a, sa = F.gpu_frame(fwd, bwd, vp, ivp, 0.5, 4, 0)
    b, sb = F.gpu_frame(fwd, bwd, vp, ivp, 0.5, 4, 0)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sa[(0, 0)]["splat"], sb[(0, 0)]["splat"])


def test_projective_division_is_ieee():
    """The reprojection divides by w through one refined reciprocal per
    projection; its quotients must be the IEEE quotients bit for bit, checked
    on the device over 2^28 operand sets (random bit patterns, moderate
    exponents, quotients next to rounding midpoints, all exponents with
    subnormals)."""
    import ctypes

    import torch

    from imagebasedspatiotemporalinterpolation_amd import _native as N

    counts = torch.zeros(8, dtype=torch.int32, device="cuda")
    print("case: division self-test, 2^28 operand sets", flush=True)
    st = N.lib().isti_selftest_division(0, 1 << 28, 12345, ctypes.c_void_p(counts.data_ptr()),
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == N.ISTI_OK
    torch.cuda.synchronize()
    c = counts.tolist()
    print("mismatches", c[:4], "reciprocal-path sets", c[4:])
    assert c[:4] == [0, 0, 0, 0]
    per_family = (1 << 28) // 4
    # the reciprocal path answered the moderate and near-midpoint families
    # entirely, and part of the other two (the rest is the fallback)
    assert c[5] == per_family and c[6] == per_family
    assert 0 < c[4] < per_family and 0 < c[7] < per_family


def test_view_size_limits_are_rejected():
    """Keyframe views hold at most 65535 pixels per axis and 2^28 - 1 pixels
    in all (32-bit byte offsets into the float4 flow planes); larger views
    are refused before any device work."""
    import ctypes

    from imagebasedspatiotemporalinterpolation_amd import _native as N
    from imagebasedspatiotemporalinterpolation_amd import renderer as R

    r = R.Renderer(n_aux=0)
    L = N.lib()
    dummy = 16  # never dereferenced: the size checks come first
    for w, h in ((65535, 4097), (65536, 1), (0, 5)):
        d = N.ViewDesc()
        d.width, d.height = w, h
        d.color, d.z = dummy, dummy
        assert L.isti_keyframe_set(r._h, 0, 0, ctypes.byref(d)) == N.ISTI_ERR_UNSUPPORTED_SHAPE, (w, h)
    d = N.ViewDesc()
    d.width, d.height = 65535, 4096  # 268,431,360 pixels: allowed
    d.color, d.z = dummy, dummy
    assert L.isti_keyframe_set(r._h, 0, 0, ctypes.byref(d)) == N.ISTI_OK
    r.close()
