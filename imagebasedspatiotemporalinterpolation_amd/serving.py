"""Continuous-batching request loop for interpolation requests.

The reference renders one client frame per call with a host sync inside
(src/main.cu:149-191, src/renderer.cu:650); a server for many clients needs
more: this module keeps every session's keyframes and render state resident
on the GPU, admits sessions by their HBM footprint, and serves interpolation
requests with continuous batching:

  - each of `n_streams` HIP streams runs one batch at a time
    (isti_render_batch: one persistent launch per stage for the whole batch,
    replayed from a hipGraph);
  - whenever a stream is idle, the oldest pending requests whose session is
    not already on the GPU form its next batch, so requests that arrive while
    others run join as soon as a stream frees up, and batches on different
    streams overlap on the device;
  - a session's requests run in arrival order (one in flight at a time,
    since they share the session's render state);
  - latency is measured from arrival to the host observing completion.

Results do not depend on the mix: a request renders exactly what it would
render alone (tests/test_gpu_serving.py).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import renderer as R
from . import scene

# Bytes per pixel and keyframe slot of a session's device state: keyframe
# planes (colour 4 + z 4 + motion 12) and renderer scratch (splat 8 + cam_ff
# 16 + full_ff 16), plus the 4-byte output per pixel.
BYTES_PER_PX_SLOT = 20 + 40
BYTES_PER_PX_OUT = 4


def session_footprint(width: int, height: int, eyes: int = 1) -> int:
    return eyes * width * height * (2 * BYTES_PER_PX_SLOT + BYTES_PER_PX_OUT)


@dataclass
class SessionClass:
    """A client kind: resolution and eye count, with the scene it is fed from."""

    name: str
    workload: int
    width: int
    height: int
    eyes: int = 1


SERVING_MIX = [  # resolutions and weights of the serving workload
    (SessionClass("640x360", 1, 640, 360), 0.30),
    (SessionClass("1280x720", 1, 1280, 720), 0.30),
    (SessionClass("1920x1080", 1, 1920, 1080), 0.25),
    (SessionClass("stereo 2x2064x2208", 3, 2064, 2208, eyes=2), 0.15),
]


class KeyframeCache:
    """One generated keyframe pair per session class (host arrays); sessions
    get their own device copies."""

    def __init__(self, seed: int = 1):
        self.seed = seed
        self._pairs: Dict[str, tuple] = {}

    def config(self, cls: SessionClass):
        return scene.defaults(cls.workload, self.seed, width=cls.width, height=cls.height, n_aux=0,
                              shadows=0, shadow_frames=0)

    def pair(self, cls: SessionClass):
        if cls.name not in self._pairs:
            cfg = self.config(cls)
            eyes = []
            for e in range(cls.eyes):
                eyes.append((scene.keyframe(cfg, 0, e, +1), scene.keyframe(cfg, 1, e, -1)))
            self._pairs[cls.name] = (cfg, eyes)
        return self._pairs[cls.name]


class Session:
    """Device-resident keyframes, renderers (one per eye) and outputs of one client."""

    def __init__(self, sid: int, cls: SessionClass, cache: KeyframeCache):
        import torch

        self.sid, self.cls = sid, cls
        self.cfg, eyes = cache.pair(cls)
        self.fpa = self.cfg.keyframe_spacing
        self.renderers: List[R.Renderer] = []
        self.frames: List[R.FrameData] = []
        self.outs = []
        for fwd, bwd in eyes:
            r = R.Renderer(n_aux=0)
            fd = R.FrameData(R.KeyframeView.from_numpy(fwd), R.KeyframeView.from_numpy(bwd))
            r.set_frame(fd)
            self.renderers.append(r)
            self.frames.append(fd)
            self.outs.append(torch.zeros((cls.height, cls.width, 4), dtype=torch.uint8, device="cuda"))
        self.footprint = session_footprint(cls.width, cls.height, cls.eyes)
        self.cameras = {}

    def info(self, b: int, eye: int, max_iter: int) -> R.RenderInfo:
        key = (b, eye)
        if key not in self.cameras:
            vp, ivp, _, _ = scene.camera(self.cfg, b, eye)
            self.cameras[key] = (vp, ivp)
        vp, ivp = self.cameras[key]
        return R.RenderInfo((self.cls.width, self.cls.height), b / self.fpa, vp, ivp,
                            R.AdditionalRenderInfo(self.fpa, 0, False, R.RenderMode.OURS), max_iter)

    def close(self):
        for r in self.renderers:
            r.close()
        self.renderers, self.frames, self.outs = [], [], []


@dataclass
class Request:
    sid: int
    b: int                     # client frame within the keyframe pair (alpha = b / fpa)
    arrival: float
    launched: float = 0.0
    done: float = 0.0
    outs: Optional[list] = None  # per-request output buffers (default: the session's)


@dataclass
class Stats:
    completed: int = 0
    pixels: int = 0
    latencies_ms: List[float] = field(default_factory=list)
    batches: int = 0
    batch_sizes: List[int] = field(default_factory=list)
    first_arrival: float = float("inf")
    last_done: float = 0.0

    def summary(self) -> dict:
        lat = np.array(self.latencies_ms) if self.latencies_ms else np.zeros(1)
        span = max(self.last_done - self.first_arrival, 1e-9)
        return {
            "requests": self.completed,
            "mpix_per_s": round(self.pixels / span / 1e6, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "batches": self.batches,
            "mean_batch": round(float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0, 2),
        }


class Server:
    """Admission by HBM footprint + continuous batching over several streams."""

    def __init__(self, hbm_budget: int, max_batch: int = 32, n_streams: int = 2, graph: bool = True,
                 max_iter: int = -1, cache: Optional[KeyframeCache] = None):
        import torch

        self.budget = hbm_budget
        self.used = 0
        self.max_batch = max_batch
        self.max_iter = max_iter
        self.cache = cache or KeyframeCache()
        self.sessions: Dict[int, Session] = {}
        self.waiting: collections.deque = collections.deque()  # sessions waiting for HBM
        self.pending: collections.deque = collections.deque()
        self.busy: set = set()  # session ids with work on the GPU
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        # each batch request is one eye: a stereo session counts twice
        self.batches = [R.Batch(max_requests=2 * max_batch, graph=graph) for _ in range(n_streams)]
        self.inflight: List[Optional[tuple]] = [None] * n_streams
        self.stats = Stats()
        self._next_sid = 0

    # -- sessions ---------------------------------------------------------
    def open_session(self, cls: SessionClass) -> int:
        """Queues a session; it becomes active once its footprint fits the budget."""
        fp = session_footprint(cls.width, cls.height, cls.eyes)
        if fp > self.budget:
            raise ValueError(f"session {cls.name} needs {fp} B, more than the whole budget")
        sid = self._next_sid
        self._next_sid += 1
        self.waiting.append((sid, cls))
        self._admit()
        return sid

    def _admit(self):
        while self.waiting:
            sid, cls = self.waiting[0]
            fp = session_footprint(cls.width, cls.height, cls.eyes)
            if self.used + fp > self.budget:
                return
            self.waiting.popleft()
            self.sessions[sid] = Session(sid, cls, self.cache)
            self.used += fp

    def active(self, sid: int) -> bool:
        return sid in self.sessions

    def close_session(self, sid: int):
        self.drain()
        s = self.sessions.pop(sid)
        self.used -= s.footprint
        self.pending = collections.deque(r for r in self.pending if r.sid != sid)
        s.close()
        self._admit()

    # -- requests ---------------------------------------------------------
    def submit(self, req: Request):
        if req.sid not in self.sessions:
            raise KeyError(f"session {req.sid} is not active")
        self.pending.append(req)
        self.stats.first_arrival = min(self.stats.first_arrival, req.arrival)

    def _retire(self, i: int, now: float):
        ev, reqs = self.inflight[i]
        for r in reqs:
            r.done = now
            self.busy.discard(r.sid)
            s = self.sessions.get(r.sid)
            if s is not None:
                self.stats.pixels += s.cls.width * s.cls.height * s.cls.eyes
            self.stats.latencies_ms.append((now - r.arrival) * 1e3)
            self.stats.completed += 1
            self.stats.last_done = max(self.stats.last_done, now)
        self.inflight[i] = None

    def _launch(self, i: int, now: float):
        chosen, work, seen = [], [], set()
        keep = collections.deque()
        while self.pending and len(chosen) < self.max_batch:
            r = self.pending.popleft()
            if r.sid in self.busy or r.sid in seen:
                keep.append(r)  # the session is on the GPU, or earlier in this batch
                continue
            seen.add(r.sid)
            chosen.append(r)
        keep.extend(self.pending)
        self.pending = keep
        if not chosen:
            return
        for r in chosen:
            s = self.sessions[r.sid]
            if r.b == 1:  # the first frame of a keyframe pair: the pair's flows are computed once more
                for rend, fd in zip(s.renderers, s.frames):
                    rend.set_frame(fd)
            outs = r.outs if r.outs is not None else s.outs
            for e, rend in enumerate(s.renderers):
                work.append((rend, s.info(r.b, e, self.max_iter), outs[e]))
            r.launched = now
            self.busy.add(r.sid)
        import torch

        with torch.cuda.stream(self.streams[i]):
            self.batches[i].render(work, stream=self.streams[i])
            ev = torch.cuda.Event()
            ev.record(self.streams[i])
        self.inflight[i] = (ev, chosen)
        self.stats.batches += 1
        self.stats.batch_sizes.append(len(chosen))

    def step(self, now: Optional[float] = None):
        """Retires finished batches and starts new ones on idle streams."""
        now = time.perf_counter() if now is None else now
        for i, f in enumerate(self.inflight):
            if f is not None and f[0].query():
                self._retire(i, time.perf_counter())
        for i, f in enumerate(self.inflight):
            if f is None and self.pending:
                self._launch(i, now)

    def drain(self):
        while self.pending or any(f is not None for f in self.inflight):
            self.step()

    def close(self):
        self.drain()
        for sid in list(self.sessions):
            self.close_session(sid)
        for b in self.batches:
            b.close()


def poisson_arrivals(rate_hz: float, duration_s: float, rng) -> np.ndarray:
    n = rng.poisson(rate_hz * duration_s * 1.2) + 8
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    return t[t < duration_s]


def run_mix(n_sessions: int = 48, rate_hz: float = 90.0, duration_s: float = 2.0, seed: int = 1,
            hbm_budget: int = 64 << 30, max_batch: int = 32, n_streams: int = 2, graph: bool = True,
            max_iter: int = -1, cache: Optional[KeyframeCache] = None) -> dict:
    """Serving mix: sessions drawn from SERVING_MIX, Poisson arrivals per session."""
    rng = np.random.default_rng(seed)
    classes = [c for c, _ in SERVING_MIX]
    weights = np.array([w for _, w in SERVING_MIX])
    picks = rng.choice(len(classes), size=n_sessions, p=weights / weights.sum())
    srv = Server(hbm_budget, max_batch=max_batch, n_streams=n_streams, graph=graph, max_iter=max_iter, cache=cache)
    sids = [srv.open_session(classes[k]) for k in picks]
    events = []
    for sid in sids:
        for t in poisson_arrivals(rate_hz, duration_s, rng):
            events.append((float(t), sid))
    events.sort()
    # warm up every session once (graph capture, scratch allocation)
    for sid in sids:
        if srv.active(sid):
            srv.submit(Request(sid, 1, time.perf_counter()))
    srv.drain()
    srv.stats = Stats()
    t0 = time.perf_counter()
    k = 0
    frame = collections.Counter()
    while k < len(events) or srv.pending or any(f is not None for f in srv.inflight):
        now = time.perf_counter() - t0
        while k < len(events) and events[k][0] <= now:
            t, sid = events[k]
            k += 1
            if srv.active(sid):
                frame[sid] += 1
                srv.submit(Request(sid, 1 + frame[sid] % (srv.sessions[sid].fpa - 1), t0 + t))
        srv.step()
    out = srv.stats.summary()
    out.update({"sessions": n_sessions, "rate_hz_per_session": rate_hz, "duration_s": duration_s,
                "offered_mpix_per_s": round(sum(srv.sessions[s].cls.width * srv.sessions[s].cls.height *
                                                srv.sessions[s].cls.eyes for s in sids if srv.active(s))
                                            * rate_hz / 1e6, 1),
                "streams": n_streams, "graph": graph})
    srv.close()
    return out
