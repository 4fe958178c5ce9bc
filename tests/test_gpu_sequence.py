"""Successive frames on one renderer: the scratch (splat buffers, flows)
carried from frame to frame must leave every frame with the oracle's splat
buffers, flows toward the other keyframe and pixels, also when the path
changes between frames (plain render, autotune, batch, learned-start
resolve, window-3 mode), when the keyframe size changes (the scratch is
reallocated), and when keyframes change at the same size (the flow toward
the other keyframe is kept per pair and must be recomputed)."""
import ctypes

import numpy as np
import pytest
import torch

from imagebasedspatiotemporalinterpolation_amd import _native as N
from imagebasedspatiotemporalinterpolation_amd import learned as Ln
from imagebasedspatiotemporalinterpolation_amd import renderer as R
from tests import _frames as F

pytestmark = pytest.mark.gpu


def _frame_data(fwd, bwd):
    kf = [R.KeyframeView.from_numpy(v) for v in fwd]
    kb = [R.KeyframeView.from_numpy(v) for v in bwd]
    return R.FrameData(kf[0], kb[0], [R.AuxCamData(kf[i], kb[i], F.AUX_PRIORITY) for i in range(1, len(fwd))])


class _Seq:
    def __init__(self, O, cfg, r):
        self.O, self.cfg, self.r = O, cfg, r
        self.fwd, self.bwd = F.pair(cfg, 0)
        self.fd = _frame_data(self.fwd, self.bwd)
        r.set_frame(self.fd)
        self.fpa = cfg.keyframe_spacing
        self.sfpa = cfg.shadow_frames if cfg.shadows else 0
        self.k = 0

    def use_pair(self, k):
        """Keyframe pair k, set on the renderer (all views)."""
        self.fwd, self.bwd = F.pair(self.cfg, k)
        self.fd = _frame_data(self.fwd, self.bwd)
        self.r.set_frame(self.fd)
        self.k = k

    def info(self, b, mode=0, iters=8):
        vp, ivp, w, h = F.target(self.cfg, self.k, b)
        ri = R.RenderInfo((w, h), b / self.fpa, vp, ivp,
                          R.AdditionalRenderInfo(self.fpa, self.sfpa, False, R.RenderMode(mode)), iters)
        return ri, vp, ivp, w, h

    def render(self, b, mode=0):
        ri, _, _, w, h = self.info(b, mode)
        out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
        self.r.render(out, ri)
        return out

    def check(self, out, b, mode=0, iters=8, prev=None, what=""):
        """Compares the frame just rendered (and the renderer's splat
        buffers) with the oracle's frame."""
        _, vp, ivp, w, h = self.info(b, mode, iters)
        torch.cuda.synchronize()
        ref, sref = F.oracle_frame(self.O, self.fwd, self.bwd, vp, ivp, b / self.fpa, self.fpa, self.sfpa, mode,
                                   iters, False, prev)
        for key, sc in sref.items():
            g = self.r.scratch(*key)
            got = g["splat"]
            assert np.array_equal(got, sc["splat"]), (what, key, int((got != sc["splat"]).sum()))
            assert F.flow_err_px(g["full_ff"], sc["full_ff"]) <= 1e-3, (what, key, "full_ff")
        F.assert_rgb_parity(out.cpu().numpy(), ref)


def _prev(w, h, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4), dtype=np.uint8)


def test_sequence_of_plain_renders(oracle_lib):
    cfg = F.scaled(2, 7, 4)
    r = R.Renderer(n_aux=2)
    S = _Seq(oracle_lib, cfg, r)
    for i, (b, mode) in enumerate([(1, 0), (2, 0), (3, 0), (4, 1), (5, 3), (2, 0)]):
        ri, vp, ivp, w, h = S.info(b, mode)
        prev = _prev(w, h, i)
        print(f"case: frame {i} b={b} mode={mode} {w}x{h}", flush=True)
        out = torch.from_numpy(prev.copy()).cuda()
        r.render(out, ri)
        S.check(out, b, mode, prev=prev, what=f"frame {i}")
    r.close()


def _splat_ptrs(r, n_aux):
    ptrs = []
    for sl in (0, 1):
        for v in range(n_aux + 1):
            sv = N.ScratchView()
            N.check(N.lib().isti_scratch(r._h, sl, v, ctypes.byref(sv)), r._h)
            ptrs.append(sv.splat)
    return ptrs


def test_splat_buffers_cleared_ahead_by_the_resolve(oracle_lib):
    """A plain resolve clears each view's spare splat buffer beside its own
    work and the next frame splats into it without a clear of its own: the
    buffers alternate from frame to frame, and every frame's splat buffers
    still hold exactly the oracle's keys (a stale key left by an incomplete
    clear would show there), also around isti_render_frames, which clears
    ahead on its second buffer set as well, batched frames (the first of which
    takes the cleared spare; a batch clears its other buffers itself) and a
    keyframe size change."""
    cfg = F.scaled(2, 11, 4)
    r = R.Renderer(n_aux=2)
    S = _Seq(oracle_lib, cfg, r)
    seen = []
    for i, b in enumerate((1, 2, 3, 4)):
        print(f"case: plain frame {i} b={b}", flush=True)
        S.check(S.render(b), b, what=f"plain {i}")
        seen.append(_splat_ptrs(r, 2))
    assert all(p != q for p, q in zip(seen[0], seen[1])), "second frame reused the first frame's buffers"
    assert seen[2] == seen[0] and seen[3] == seen[1], "buffers do not alternate"

    print("case: render_frames, then plain", flush=True)
    infos = [S.info(b)[0] for b in (1, 2, 3, 4, 5)]
    w, h = S.info(1)[3:5]
    outs = [torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda") for _ in infos]
    r.render_frames(outs, infos)
    torch.cuda.synchronize()
    for b, out in zip((1, 2, 3, 4, 5), outs):
        ref = torch.zeros_like(out)
        r.render(ref, S.info(b)[0])
        S.check(ref, b, what=f"plain after render_frames, b={b}")
        assert torch.equal(out, ref), b

    print("case: batched frames", flush=True)
    batch = R.Batch(max_requests=2)
    for b in (2, 3, 4):
        ri, _, _, w, h = S.info(b)
        out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
        batch.render([(r, ri, out)])
        S.check(out, b, what=f"batch {b}")
    S.check(S.render(5), 5, what="plain after batch")
    S.check(S.render(1), 1, what="plain after plain after batch")
    batch.close()

    print("case: keyframe size change", flush=True)
    S2 = _Seq(oracle_lib, F.scaled(2, 12, 5), r)
    for b in (1, 2, 3):
        S2.check(S2.render(b), b, what=f"new size {b}")
    r.close()


def test_sequence_across_paths_and_sizes(oracle_lib):
    cfg = F.scaled(2, 8, 4)
    r = R.Renderer(n_aux=2)
    S = _Seq(oracle_lib, cfg, r)
    batch = R.Batch(max_requests=2)
    print("case: plain, plain", flush=True)
    S.check(S.render(1), 1, what="plain 1")
    S.check(S.render(2), 2, what="plain 2")

    print("case: autotune", flush=True)
    ri, _, _, w, h = S.info(3)
    out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    r.autotune(out, ri, reps=2)
    S.check(out, 3, what="autotune")
    S.check(S.render(4), 4, what="plain after autotune")

    print("case: batch", flush=True)
    ri, _, _, w, h = S.info(5)
    out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    batch.render([(r, ri, out)])
    S.check(out, 5, what="batch")
    S.check(S.render(2), 2, what="plain after batch")

    print("case: learned-start resolve with learned starts off", flush=True)
    ri, _, _, w, h = S.info(3)
    li = Ln.LearnedInit(w, h, 2)
    out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    li.render(r, out, ri, conf_threshold=1.0)
    S.check(out, 3, what="learned")
    # the prepared frame stays resolvable after its resolve
    out2 = torch.zeros_like(out)
    Ln.render_learned(r, out2, ri, li.out, 1.0)
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    S.check(S.render(4, 1), 4, 1, what="window 3 after learned")

    print("case: keyframe size change", flush=True)
    cfg2 = F.scaled(2, 9, 5)
    S2 = _Seq(oracle_lib, cfg2, r)
    S2.check(S2.render(2), 2, what="new size 1")
    S2.check(S2.render(3), 3, what="new size 2")
    li.close()
    batch.close()
    r.close()



def _set_view(r, slot, view, kv, priority):
    d = kv.desc(priority)
    N.check(N.lib().isti_keyframe_set(r._h, slot, view, ctypes.byref(d)), r._h)


def test_pair_flow_reused_until_a_keyframe_changes(oracle_lib):
    """The flow toward the other keyframe depends on the keyframe pair alone:
    the first render after a keyframe is set computes it, later renders of the
    pair reuse it. Every frame is checked against the oracle (splat buffers,
    that flow, pixels) across frames of one pair, a new pair of the same size,
    one view's partner set alone, a plane rewritten in place and set again, a
    swap, the batch path and the linear-initialisation mode."""
    cfg = F.scaled(2, 7, 4)
    r = R.Renderer(n_aux=2)
    S = _Seq(oracle_lib, cfg, r)
    batch = R.Batch(max_requests=2)
    for b in (1, 2, 3):
        print(f"case: pair 0 frame {b}", flush=True)
        S.check(S.render(b), b, what=f"pair 0 frame {b}")

    S.use_pair(1)
    for b in (1, 2):
        print(f"case: pair 1 frame {b}", flush=True)
        S.check(S.render(b), b, what=f"pair 1 frame {b}")

    # the bwd main view alone (the partner of the fwd main view's flow)
    print("case: partner set alone", flush=True)
    other = F.pair(cfg, 2)[1][0]
    kv = R.KeyframeView.from_numpy(other)
    _set_view(r, 1, 0, kv, 0)
    S.bwd = [other] + S.bwd[1:]
    S.check(S.render(3), 3, what="partner set alone")
    S.check(S.render(1), 1, what="partner set alone, next frame")

    # a plane rewritten in place, then the view set again
    print("case: plane rewritten in place", flush=True)
    z2 = F.pair(cfg, 3)[1][0]["z"]
    kv.z.copy_(torch.from_numpy(z2).cuda())
    _set_view(r, 1, 0, kv, 0)
    S.bwd = [dict(other, z=z2)] + S.bwd[1:]
    S.check(S.render(2), 2, what="rewritten in place")

    print("case: swap", flush=True)
    r.swap()
    S.fwd, S.bwd = S.bwd, S.fwd
    S.check(S.render(2), 2, what="after swap")
    S.check(S.render(4), 4, what="after swap, next frame")

    print("case: batch", flush=True)
    S.use_pair(0)
    for b in (3, 4):
        ri, _, _, w, h = S.info(b)
        out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
        batch.render([(r, ri, out)])
        S.check(out, b, what=f"batch frame {b}")

    # replayed stage graphs across a pair change: the per-job flow flag
    # travels in the uploaded work table, not in the graph
    print("case: batch with stage graphs", flush=True)
    gbatch = R.Batch(max_requests=2, graph=True)
    for k, b in ((0, 5), (0, 6), (1, 1), (1, 2)):
        if k != S.k:
            S.use_pair(k)
        ri, _, _, w, h = S.info(b)
        out = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
        gbatch.render([(r, ri, out)])
        S.check(out, b, what=f"graph batch pair {k} frame {b}")
    gbatch.close()

    print("case: linear initialisation after reuse", flush=True)
    mode = int(R.RenderMode.BACKWARD_BIDIRECTIONAL)
    S.check(S.render(5, mode), 5, mode, what="linear init")
    S.check(S.render(2), 2, what="plain after linear init")
    del kv
    batch.close()
    r.close()
