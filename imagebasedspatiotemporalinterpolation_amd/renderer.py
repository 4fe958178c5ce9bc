"""Python mirror of the reference's Renderer interface over the C ABI.

Reference: include/renderer.h:9-56 (RenderMode, AdditionalRenderInfo,
RenderInfo, Renderer::render) and include/data.h (FrameData with fwd/bwd,
aux cameras with curr/next and swap()). The call shape is the reference's:

    renderer.render(render_buffer, render_info, frame_data)

Keyframe planes and the render buffer are torch tensors on the GPU (device
memory and stream plumbing only; the work runs in the HIP kernels of
libisti.so). Only valid pixels are written, so the render buffer keeps the
previous frame in holes, as in the reference.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional

import numpy as np

from . import _native as N


class RenderMode(IntEnum):  # include/renderer.h:9-16
    OURS = 0
    BACKWARD_BIDIRECTIONAL = 1
    BACKWARD_UNIDIRECTIONAL = 2
    SPLATTING = 3
    TIMEWARP = 4


@dataclass
class AdditionalRenderInfo:  # include/renderer.h:18-24
    frames_per_anchor: int = 1
    shadow_frames_per_anchor: int = 0
    no_shadows: bool = False
    mode: RenderMode = RenderMode.OURS


@dataclass
class RenderInfo:  # include/renderer.h:26-35
    resolution: tuple
    alpha: float
    viewproj_mat: np.ndarray      # 4x4 (row, col) or flat column-major [16]
    inv_viewproj_mat: np.ndarray
    additional_info: AdditionalRenderInfo = field(default_factory=AdditionalRenderInfo)
    max_search_iterations: int = -1  # < 0: the mode's own count

    def to_c(self) -> N.RenderInfoC:
        c = N.RenderInfoC()
        c.width, c.height = int(self.resolution[0]), int(self.resolution[1])
        c.alpha = float(self.alpha)
        c.viewproj = N.mat16(self.viewproj_mat)
        c.inv_viewproj = N.mat16(self.inv_viewproj_mat)
        a = self.additional_info
        c.frames_per_anchor = int(a.frames_per_anchor)
        c.shadow_frames_per_anchor = int(a.shadow_frames_per_anchor)
        c.no_shadows = 1 if a.no_shadows else 0
        c.mode = int(a.mode)
        c.max_search_iterations = int(self.max_search_iterations)
        return c


@dataclass
class KeyframeView:
    """MotionZBufferImage (include/data.h:12-52) as device tensors.

    color: uint8 [H, W, 4]; z: float32 [H, W] (NDC); motion: float32 [3, H, W]
    or None; shadow_*: uint8 [H, W, 4] or None; matrices column-major [16].
    """

    color: object
    z: object
    viewproj: np.ndarray
    inv_viewproj: np.ndarray
    motion: Optional[object] = None
    shadow_info: Optional[object] = None
    shadow_color_a: Optional[object] = None
    shadow_color_b: Optional[object] = None

    @property
    def dims(self):
        return int(self.z.shape[1]), int(self.z.shape[0])

    def desc(self, priority: int = 0) -> N.ViewDesc:
        for name in ("color", "z", "motion", "shadow_info", "shadow_color_a", "shadow_color_b"):
            t = getattr(self, name)
            if t is not None and (not t.is_cuda or not t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous GPU tensor")
        d = N.ViewDesc()
        d.width, d.height = self.dims
        d.color = self.color.data_ptr()
        d.z = self.z.data_ptr()
        if self.motion is not None:
            d.motion_x = self.motion[0].data_ptr()
            d.motion_y = self.motion[1].data_ptr()
            d.motion_z = self.motion[2].data_ptr()
        for name in ("shadow_info", "shadow_color_a", "shadow_color_b"):
            t = getattr(self, name)
            setattr(d, name, t.data_ptr() if t is not None else None)
        d.viewproj = N.mat16(np.asarray(self.viewproj, np.float32).reshape(16))
        d.inv_viewproj = N.mat16(np.asarray(self.inv_viewproj, np.float32).reshape(16))
        d.priority = int(priority)
        return d

    @staticmethod
    def from_numpy(kf: dict, device="cuda") -> "KeyframeView":
        import torch

        def up(a):
            return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(device)

        return KeyframeView(color=up(kf["color"]), z=up(kf["z"]), viewproj=kf["viewproj"],
                            inv_viewproj=kf["inv_viewproj"], motion=up(kf.get("motion")),
                            shadow_info=up(kf.get("shadow_info")), shadow_color_a=up(kf.get("shadow_color_a")),
                            shadow_color_b=up(kf.get("shadow_color_b")))


@dataclass
class AuxCamData:  # include/data.h:64-77
    curr: KeyframeView
    next: KeyframeView
    priority: int = 0

    def swap(self):
        self.curr, self.next = self.next, self.curr


@dataclass
class FrameData:  # include/data.h:79-116
    fwd: KeyframeView
    bwd: KeyframeView
    aux_cam_images: List[AuxCamData] = field(default_factory=list)

    def swap(self):
        self.fwd, self.bwd = self.bwd, self.fwd
        for a in self.aux_cam_images:
            a.swap()


class Renderer:
    """Renderer (include/renderer.h:52-93) bound to one GPU."""

    def __init__(self, n_aux: int = 0, device: int = 0):
        L = N.lib()
        h = ctypes.c_void_p()
        cfg = N.Config(int(device), int(n_aux))
        N.check(L.isti_create(ctypes.byref(cfg), ctypes.byref(h)), None)
        self._h = h
        self.n_aux = n_aux
        self.device = device

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().isti_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_frame(self, frame_data: FrameData) -> None:
        """Hands the keyframe pair to the renderer for the renders that follow.
        They read the planes afresh, except the pair's flow toward the other
        keyframe, which the first of them computes and the others reuse; call
        again when the pair or a plane's contents change."""
        L = N.lib()
        if len(frame_data.aux_cam_images) != self.n_aux:
            raise ValueError(f"renderer expects {self.n_aux} aux cameras")
        views = [(0, 0, frame_data.fwd, 0), (1, 0, frame_data.bwd, 0)]
        for i, a in enumerate(frame_data.aux_cam_images):
            views += [(0, i + 1, a.curr, a.priority), (1, i + 1, a.next, a.priority)]
        for slot, view, kv, prio in views:
            self.set_view(slot, view, kv, prio)

    def set_view(self, slot: int, view: int, kv: KeyframeView, priority: int = 0) -> None:
        """Sets one view (0 main, 1.. aux) of one keyframe slot (0 fwd, 1 bwd)."""
        d = kv.desc(priority)
        N.check(N.lib().isti_keyframe_set(self._h, int(slot), int(view), ctypes.byref(d)), self._h)

    def render(self, render_buffer, render_info: RenderInfo, frame_data: Optional[FrameData] = None,
               stream=None) -> None:
        """Renderer::render (src/renderer.cu:607-694); stream-ordered, no host sync.
        With frame_data the keyframes are set first (everything recomputed, as
        the reference); without it the pair last set is rendered."""
        import torch

        if frame_data is not None:
            self.set_frame(frame_data)
        if not (render_buffer.is_cuda and render_buffer.dtype == torch.uint8 and render_buffer.is_contiguous()):
            raise ValueError("render_buffer must be a contiguous uint8 GPU tensor [H, W, 4]")
        h, w = render_buffer.shape[:2]
        if (w, h) != (int(render_info.resolution[0]), int(render_info.resolution[1])):
            raise ValueError("render_buffer shape does not match render_info.resolution")
        s = stream if stream is not None else torch.cuda.current_stream()
        ci = render_info.to_c()
        N.check(N.lib().isti_render(self._h, ctypes.byref(ci), render_buffer.data_ptr(), w * 4,
                                    ctypes.c_void_p(s.cuda_stream)), self._h)

    def render_frames(self, render_buffers, render_infos, stream=None) -> None:
        """The frames of the current keyframe pair (isti_render_frames): the
        bytes of render() called for each (buffer, info) in order, with
        consecutive frames overlapped on two internal streams. Buffers must be
        distinct contiguous uint8 GPU tensors [H, W, 4]."""
        import torch

        bufs, infos = list(render_buffers), list(render_infos)
        if len(bufs) != len(infos) or not bufs:
            raise ValueError("one render buffer per render info, at least one")
        n = len(bufs)
        cinfos = (N.RenderInfoC * n)()
        outs = (ctypes.c_void_p * n)()
        pitches = (ctypes.c_size_t * n)()
        for i, (b, info) in enumerate(zip(bufs, infos)):
            if not (b.is_cuda and b.dtype == torch.uint8 and b.is_contiguous()):
                raise ValueError("render buffers must be contiguous uint8 GPU tensors [H, W, 4]")
            if (int(b.shape[1]), int(b.shape[0])) != (int(info.resolution[0]), int(info.resolution[1])):
                raise ValueError("a render buffer's shape does not match its render_info.resolution")
            cinfos[i] = info.to_c()
            outs[i] = b.data_ptr()
            pitches[i] = int(b.shape[1]) * 4
        s = stream if stream is not None else torch.cuda.current_stream()
        N.check(N.lib().isti_render_frames(self._h, cinfos, n, outs, pitches, ctypes.c_void_p(s.cuda_stream)), self._h)

    def autotune(self, render_buffer, render_info: RenderInfo, reps: int = 20, stream=None) -> dict:
        """Times the splat variants and every resolve tile shape on this frame
        and keeps the fastest for renders of its size, aux count and window.
        Returns {"tile": shape index, "resolve_ms": [per shape], "binned": 0/1,
        "splat_ms": [direct, binned]}. Synchronises the stream."""
        import torch

        h, w = render_buffer.shape[:2]
        s = stream if stream is not None else torch.cuda.current_stream()
        ci = render_info.to_c()
        ms = (ctypes.c_float * N.ISTI_AUTOTUNE_TIMES)()
        best, best_splat = ctypes.c_int32(-1), ctypes.c_int32(-1)
        N.check(N.lib().isti_autotune(self._h, ctypes.byref(ci), render_buffer.data_ptr(), w * 4, int(reps), ms,
                                      ctypes.byref(best), ctypes.byref(best_splat), ctypes.c_void_p(s.cuda_stream)),
                self._h)
        t = [float(v) for v in ms]
        return {"tile": best.value, "resolve_ms": t[:N.ISTI_N_TILE_SHAPES], "binned": best_splat.value,
                "splat_ms": t[N.ISTI_N_TILE_SHAPES:]}

    def set_splat_binning(self, mode: int) -> None:
        """Force the splat variant: 0 direct, 1 LDS-binned, -1 tuned."""
        N.check(N.lib().isti_set_splat_binning(self._h, int(mode)), self._h)

    def set_tile_shape(self, shape: int) -> None:
        """Force a resolve tile shape (index into _native.TILE_SHAPES), or -1 for the tuned one."""
        N.check(N.lib().isti_set_tile_shape(self._h, int(shape)), self._h)

    def swap(self) -> None:
        """FrameData::swap on the renderer's own keyframe descriptors."""
        N.check(N.lib().isti_keyframe_swap(self._h), self._h)

    def set_stage_timing(self, enable: bool = True) -> None:
        """Record HIP events between the stages of every render (off by default)."""
        N.check(N.lib().isti_set_stage_timing(self._h, 1 if enable else 0), self._h)

    def stage_times(self) -> tuple:
        """(clear, reproject+splat, resolve) GPU milliseconds of the last timed render."""
        ms = (ctypes.c_float * 3)()
        N.check(N.lib().isti_stage_times(self._h, ms), self._h)
        return tuple(ms)

    def stage_totals(self):
        """((clear, splat, resolve) summed GPU ms, render count) since timing was enabled."""
        tot = (ctypes.c_double * 3)()
        n = ctypes.c_int64()
        N.check(N.lib().isti_stage_totals(self._h, tot, ctypes.byref(n)), self._h)
        return tuple(tot), n.value

    def scratch(self, slot: int, view: int) -> dict:
        """Host copies of (slot, view)'s splat buffer and flow planes after the last render."""
        sv = N.ScratchView()
        N.check(N.lib().isti_scratch(self._h, slot, view, ctypes.byref(sv)), self._h)
        n = sv.width * sv.height
        if n == 0 or not sv.splat:
            return {}
        return {
            "splat": N.device_to_host(sv.splat, n * 8, np.uint64, (sv.height, sv.width)),
            "cam_ff": N.device_to_host(sv.cam_ff, n * 16, np.float32, (sv.height, sv.width, 4)),
            "full_ff": N.device_to_host(sv.full_ff, n * 12, np.float32, (sv.height, sv.width, 3)),
        }


class Batch:
    """Many renders in one persistent launch per stage (isti_render_batch).

    Each request is (renderer, RenderInfo, render_buffer); a renderer may
    appear once per call and must have its keyframes set. Results equal
    separate Renderer.render calls. graph=True replays the stage sequence
    from captured hipGraphs.
    """

    def __init__(self, max_requests: int = 64, device: int = 0, graph: bool = False):
        h = ctypes.c_void_p()
        st = N.lib().isti_batch_create(int(device), int(max_requests), ctypes.byref(h))
        if st != N.ISTI_OK:
            raise N.IstiError(st, N.lib().isti_last_error(None).decode())
        self._h = h
        self.max_requests = max_requests
        self.set_graph(graph)

    def set_graph(self, enable: bool) -> None:
        self._check(N.lib().isti_batch_set_graph(self._h, 1 if enable else 0))

    def _check(self, st):
        if st != N.ISTI_OK:
            raise N.IstiError(st, N.lib().isti_batch_last_error(self._h).decode())

    def render(self, requests, stream=None) -> None:
        """requests: iterable of (Renderer, RenderInfo, render_buffer tensor)."""
        import torch

        reqs = list(requests)
        n = len(reqs)
        rs = (ctypes.c_void_p * max(n, 1))()
        infos = (N.RenderInfoC * max(n, 1))()
        outs = (ctypes.c_void_p * max(n, 1))()
        pitches = (ctypes.c_size_t * max(n, 1))()
        for i, (r, info, buf) in enumerate(reqs):
            rs[i] = r._h.value
            infos[i] = info.to_c()
            outs[i] = buf.data_ptr()
            pitches[i] = int(buf.shape[1]) * 4
        s = stream if stream is not None else torch.cuda.current_stream()
        self._check(N.lib().isti_render_batch(self._h, rs, infos, outs, pitches, n, ctypes.c_void_p(s.cuda_stream)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().isti_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
