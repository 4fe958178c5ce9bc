"""Seed sweep of the parity bar: the 2-aux soft-shadow scene at 480x270
over several seeds, client frames and modes (0: plain search, 1: window-3
backward bidirectional with always-valid blends, 3: splatting only), plus
the fast-camera stereo eye. Splat buffers bit-exact, flows within 1e-3 px,
RGBA identical to the oracle's bytes (stricter than the 1/255 bar), from a
random previous frame so hole retention is covered."""
import numpy as np
import pytest

from tests import _frames as F

pytestmark = pytest.mark.gpu

CASES = [(seed, b, mode) for seed in (21, 22, 23, 24) for b, mode in ((1, 0), (4, 1), (7, 3), (5, 0))]


@pytest.mark.parametrize("seed,b,mode", CASES)
def test_two_aux_scene_sweep(oracle_lib, seed, b, mode):
    cfg = F.scaled(2, seed, 4)
    fwd, bwd = F.pair(cfg, 0)
    vp, ivp, w, h = F.target(cfg, 0, b)
    fpa, sfpa = cfg.keyframe_spacing, cfg.shadow_frames
    prev = np.random.default_rng(seed * 10 + b).integers(0, 256, (h, w, 4), dtype=np.uint8)
    print(f"case: seed={seed} b={b} mode={mode} {w}x{h}", flush=True)
    ref, sref = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, b / fpa, fpa, sfpa, mode, 32, False, prev)
    got, sgot = F.gpu_frame(fwd, bwd, vp, ivp, b / fpa, fpa, sfpa, mode, 32, False, prev)
    for key, s in sref.items():
        g = sgot[key]
        np.testing.assert_array_equal(g["splat"], s["splat"], err_msg=f"splat {key}")
        assert F.flow_err_px(g["cam_ff"], s["cam_ff"]) <= 1e-3, f"cam_ff {key}"
        assert F.flow_err_px(g["full_ff"], s["full_ff"]) <= 1e-3, f"full_ff {key}"
    assert np.array_equal(got, ref), int((got != ref).any(-1).sum())


@pytest.mark.parametrize("seed", (31, 32))
def test_fast_camera_eye_sweep(oracle_lib, seed):
    cfg = F.scaled(3, seed, 4)
    fwd, bwd = F.pair(cfg, 0, eye=1)
    vp, ivp, w, h = F.target(cfg, 0, 3, eye=1)
    fpa = cfg.keyframe_spacing
    print(f"case: stereo eye seed={seed} {w}x{h}", flush=True)
    ref, _ = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, 3 / fpa, fpa, 0, 0, 32, False)
    got, _ = F.gpu_frame(fwd, bwd, vp, ivp, 3 / fpa, fpa, 0, 0, 32, False, want_scratch=False)
    assert np.array_equal(got, ref), int((got != ref).any(-1).sum())
