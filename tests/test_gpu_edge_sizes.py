"""Edge sizes and aux counts against the oracle: a 1x1 view, views smaller
than one tile, odd sizes whose tiles are partial on both edges, render/aux
ratios of 1, 2 and 3 (different per axis in one case: the ratio is taken
per axis, renderer.cu:432-434), and 3 and 4 aux cameras (the scene's two aux
views repeated; 4 is the largest count the boundary takes). Same bar as the
other parity cases: splat buffers byte for byte, flows within 1e-3 px, the
frame within 1/255 and >= 50 dB. Every resolve tile shape renders each case
(the one-wave 8x8 tile stages its halo and aux windows with 64 threads), and
every other one splats through the LDS bin (views narrower or shorter than
its 32x8 source tile included)."""
import numpy as np
import pytest

from imagebasedspatiotemporalinterpolation_amd import _native as N
from imagebasedspatiotemporalinterpolation_amd import renderer as R
from tests import _frames as F

pytestmark = pytest.mark.gpu

# (width, height, aux width, aux height, aux cameras, client frame, mode)
CASES = [
    (1, 1, 1, 1, 2, 2, 0),
    (7, 5, 7, 5, 2, 1, 0),
    (7, 5, 7, 5, 2, 1, 4),
    (14, 10, 7, 5, 2, 3, 0),
    (99, 63, 33, 21, 2, 3, 0),
    (98, 62, 49, 31, 2, 3, 1),
    (96, 60, 32, 30, 2, 4, 0),
    (130, 70, 65, 35, 3, 2, 0),
    (130, 70, 65, 35, 4, 5, 3),
]


@pytest.mark.parametrize("w,h,aw,ah,n_aux,b,mode", CASES)
def test_edge_sizes_and_aux_counts(oracle_lib, w, h, aw, ah, n_aux, b, mode):
    cfg = F.scaled(2, 7, width=w, height=h, aux_width=aw, aux_height=ah)
    fwd, bwd = F.pair(cfg, 0)
    views = [0] + [1 + (i % 2) for i in range(n_aux)]
    fwd, bwd = [fwd[i] for i in views], [bwd[i] for i in views]
    vp, ivp, tw, th = F.target(cfg, 0, b)
    fpa, sfpa = cfg.keyframe_spacing, cfg.shadow_frames
    prev = np.random.default_rng(11).integers(0, 256, (th, tw, 4), dtype=np.uint8)
    print(f"case: {w}x{h} aux {aw}x{ah} x{n_aux} b={b} mode={mode}", flush=True)
    ref, sref = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, b / fpa, fpa, sfpa, mode, 32, False, prev)
    for shape in range(N.ISTI_N_TILE_SHAPES):
        binned = shape % 2  # the LDS-binned splat on every other shape
        print(f"  tile shape {N.TILE_SHAPES[shape]} splat {'binned' if binned else 'direct'}", flush=True)
        r = R.Renderer(n_aux=n_aux)
        r.set_tile_shape(shape)
        r.set_splat_binning(binned)
        got, sgot = F.gpu_frame(fwd, bwd, vp, ivp, b / fpa, fpa, sfpa, mode, 32, False, prev, renderer=r)
        if mode != 4:
            for key, s in sref.items():
                g = sgot[key]
                np.testing.assert_array_equal(g["splat"], s["splat"], err_msg=f"splat {key}")
                assert F.flow_err_px(g["cam_ff"], s["cam_ff"]) <= 1e-3, f"cam_ff {key}"
                assert F.flow_err_px(g["full_ff"], s["full_ff"]) <= 1e-3, f"full_ff {key}"
                F.assert_visible_flags(g, s, key)
        F.assert_rgb_parity(got, ref)
        r.close()
