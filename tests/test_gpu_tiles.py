"""Resolve tile shapes and the autotuner: every tile shape renders the same
bytes as the oracle (hence as each other), including ragged image edges,
aux cameras and the window-3 mode; the autotuner times every shape and its
choice is used by later renders."""
import numpy as np
import pytest
import torch

from imagebasedspatiotemporalinterpolation_amd import _native as N
from imagebasedspatiotemporalinterpolation_amd import renderer as R
from tests import _frames as F

pytestmark = pytest.mark.gpu


def _setup(cfg, b, mode, n_views):
    fwd, bwd = F.pair(cfg, 0)
    fwd, bwd = fwd[:n_views], bwd[:n_views]
    vp, ivp, w, h = F.target(cfg, 0, b)
    fpa = cfg.keyframe_spacing
    sfpa = cfg.shadow_frames if cfg.shadows else 0
    r = R.Renderer(n_aux=n_views - 1)
    kf = [R.KeyframeView.from_numpy(v) for v in fwd]
    kb = [R.KeyframeView.from_numpy(v) for v in bwd]
    fd = R.FrameData(kf[0], kb[0], [R.AuxCamData(kf[i], kb[i], F.AUX_PRIORITY) for i in range(1, n_views)])
    r.set_frame(fd)
    info = R.RenderInfo((w, h), b / fpa, vp, ivp, R.AdditionalRenderInfo(fpa, sfpa, False, R.RenderMode(mode)), 8)
    return r, fd, info, fwd, bwd, vp, ivp, (w, h, fpa, sfpa)


# (workload, size divisor, client frame, mode, views): 1080p/5 = 384x216 and
# the 2-aux shadow scene at /4 = 480x270 leave partial tiles at the bottom
# edge for every shape; mode 1 uses the 7x7 window (its own LDS halo).
CASES = [(1, 1, 2, 0, 1), (2, 4, 3, 0, 3), (2, 4, 5, 1, 3), (2, 5, 2, 3, 1)]


@pytest.mark.parametrize("wl,div,b,mode,views", CASES)
def test_every_tile_shape_matches_the_oracle(oracle_lib, wl, div, b, mode, views):
    cfg = F.scaled(wl, 5, div)
    r, fd, info, fwd, bwd, vp, ivp, (w, h, fpa, sfpa) = _setup(cfg, b, mode, views)
    prev = np.random.default_rng(3).integers(0, 256, (h, w, 4), dtype=np.uint8)
    ref, _ = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, b / fpa, fpa, sfpa, mode, 8, False, prev)
    for shape in range(N.ISTI_N_TILE_SHAPES):
        print(f"case: {w}x{h} aux={views - 1} mode={mode} shape={N.TILE_SHAPES[shape]}", flush=True)
        r.set_tile_shape(shape)
        out = torch.from_numpy(prev.copy()).cuda()
        r.render(out, info)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (shape, int(np.abs(got.astype(int) - ref).max()))
    r.close()


def test_autotune_times_every_shape_and_later_renders_use_it():
    cfg = F.scaled(2, 6, 2)
    r, fd, info, *_ = _setup(cfg, 3, 0, 3)
    w, h = info.resolution
    base = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    r.render(base, info)
    out = torch.zeros_like(base)
    at = r.autotune(out, info, reps=5)
    torch.cuda.synchronize()
    best, ms = at["tile"], at["resolve_ms"]
    print("tile shape times (ms):", dict(zip(N.TILE_SHAPES, ms)), "best", N.TILE_SHAPES[best],
          "splat (direct, binned):", at["splat_ms"])
    assert 0 <= best < N.ISTI_N_TILE_SHAPES
    assert all(t > 0 for t in ms) and ms[best] == min(ms)
    assert at["binned"] in (0, 1) and at["splat_ms"][at["binned"]] == min(at["splat_ms"])
    assert torch.equal(out, base)  # the autotuner's frame is the frame
    # a fresh renderer of the same size, aux count and window uses the tuned shape
    r2, fd2, info2, *_ = _setup(cfg, 3, 0, 3)
    out2 = torch.zeros_like(base)
    r2.render(out2, info2)
    torch.cuda.synchronize()
    assert torch.equal(out2, base)
    with pytest.raises(N.IstiError):
        r2.set_tile_shape(N.ISTI_N_TILE_SHAPES)
    r.close()
    r2.close()


# C1 at its own size, the 2-aux shadow scene at /4, and the fast-camera stereo
# eye at /4 (large motions: many targets leave the +-16-pixel bins)
@pytest.mark.parametrize("wl,div,b,mode,views", [(1, 1, 2, 0, 1), (2, 4, 3, 0, 3), (3, 4, 2, 0, 1), (2, 4, 5, 1, 3)])
def test_lds_binned_splat_matches_the_oracle(oracle_lib, wl, div, b, mode, views):
    """The LDS-binned splat leaves the oracle's splat buffers bit for bit
    (and so the same frame) in every mode, including mode 1's linear splat."""
    cfg = F.scaled(wl, 8, div)
    r, fd, info, fwd, bwd, vp, ivp, (w, h, fpa, sfpa) = _setup(cfg, b, mode, views)
    _, sref = F.oracle_frame(oracle_lib, fwd, bwd, vp, ivp, b / fpa, fpa, sfpa, mode, 8, False)
    outs = []
    for binned in (0, 1):
        print(f"case: {w}x{h} aux={views - 1} mode={mode} binned={binned}", flush=True)
        r.set_splat_binning(binned)
        out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
        r.render(out, info)
        torch.cuda.synchronize()
        for key, sc in sref.items():
            got = r.scratch(*key)
            assert np.array_equal(got["splat"], sc["splat"]), (binned, key)
        outs.append(out.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    with pytest.raises(N.IstiError):
        r.set_splat_binning(2)
    r.close()
