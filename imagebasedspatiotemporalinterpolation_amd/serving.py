"""Continuous-batching request loop for interpolation requests.

The reference renders one client frame per call with a host sync inside
(src/main.cu:149-191, src/renderer.cu:650); a server for many clients needs
more: this module keeps keyframes and every session's render state resident
on the GPU, admits sessions by their HBM footprint, and serves interpolation
requests with continuous batching:

  - keyframes live in a device-resident store per session class (colour, z
    and the motion toward both neighbouring keyframes of every keyframe);
    sessions point their renderers at a pair of it, and advance to the next
    pair as the reference's FrameData::swap + readNextFrame do
    (include/data.h:110-116, src/file_reader.cu:167-217): the renderer's
    descriptors are swapped, then both slots are set to the new pair (no
    plane is copied);
  - each of `n_streams` HIP streams runs one batch at a time
    (isti_render_batch: one persistent launch per stage for the whole batch,
    replayed from a hipGraph);
  - whenever a stream is idle, the oldest pending requests whose session is
    not already on the GPU form its next batch, so requests that arrive while
    others run join as soon as a stream frees up, and batches on different
    streams overlap on the device;
  - a session's requests run in arrival order (one in flight at a time,
    since they share the session's render state);
  - with search="learned" every request takes the learned-start path
    (splat + features, the network, the learned resolve; learned.py) on its
    stream instead of the batched plain launch, so the two methods can be
    served side by side on the same mix; the network runs on the first
    request of each keyframe pair of a session eye and its output (kept per
    session eye, counted in the admission footprint) serves the pair's later
    requests (learned_per_pair=False runs it on every request);
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

# Device bytes per pixel and eye. A session's renderer holds, per keyframe
# slot, the splat cells (8), the cam_ff records (16) and the full_ff records
# (12), and the session owns its 4-byte output. The class's keyframe store
# holds per keyframe colour (4), z (4) and the motion toward the next and the
# previous keyframe (12 each).
BYTES_PER_PX_SLOT = 8 + 16 + 12
BYTES_PER_PX_OUT = 4
BYTES_PER_PX_KEYFRAME = 4 + 4 + 12 + 12


def session_footprint(width: int, height: int, eyes: int = 1) -> int:
    return eyes * width * height * (2 * BYTES_PER_PX_SLOT + BYTES_PER_PX_OUT)


def store_footprint(width: int, height: int, eyes: int, n_keyframes: int) -> int:
    return eyes * width * height * BYTES_PER_PX_KEYFRAME * n_keyframes


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


class KeyframeStore:
    """Device-resident keyframes 0..n_keyframes-1 of every eye of each session
    class, generated on first use. Keyframe k carries its colour and z once
    and two motion planes: toward k+1 (its role as the fwd keyframe of pair k)
    and toward k-1 (as the bwd keyframe of pair k-1). Sessions of a class read
    them without copying."""

    def __init__(self, seed: int = 1, n_keyframes: int = 3):
        if n_keyframes < 2:
            raise ValueError("a store needs at least one keyframe pair")
        self.seed = seed
        self.n_keyframes = n_keyframes
        self._kf: Dict[tuple, dict] = {}
        self._views: Dict[tuple, R.KeyframeView] = {}

    @property
    def n_pairs(self) -> int:
        return self.n_keyframes - 1

    def config(self, cls: SessionClass):
        return scene.defaults(cls.workload, self.seed, width=cls.width, height=cls.height, n_aux=0,
                              shadows=0, shadow_frames=0)

    def footprint(self, cls: SessionClass) -> int:
        return store_footprint(cls.width, cls.height, cls.eyes, self.n_keyframes)

    def _view(self, cls: SessionClass, k: int, eye: int, direction: int) -> R.KeyframeView:
        import torch

        key = (cls.name, k, eye, direction)
        if key not in self._views:
            cfg = self.config(cls)
            host = scene.keyframe(cfg, k, eye, direction)
            base = self._kf.get((cls.name, k, eye))
            if base is None:
                base = {"color": torch.from_numpy(host["color"]).cuda(), "z": torch.from_numpy(host["z"]).cuda()}
                self._kf[(cls.name, k, eye)] = base
            motion = torch.from_numpy(np.ascontiguousarray(host["motion"])).cuda()
            self._views[key] = R.KeyframeView(color=base["color"], z=base["z"], viewproj=host["viewproj"],
                                              inv_viewproj=host["inv_viewproj"], motion=motion)
        return self._views[key]

    def pair(self, cls: SessionClass, pair: int, eye: int) -> R.FrameData:
        """FrameData of keyframe pair (pair, pair + 1) of one eye."""
        if not 0 <= pair < self.n_pairs:
            raise IndexError(f"pair {pair} outside the store's {self.n_pairs} pairs")
        return R.FrameData(self._view(cls, pair, eye, +1), self._view(cls, pair + 1, eye, -1))

    def preload(self, cls: SessionClass) -> None:
        for p in range(self.n_pairs):
            for e in range(cls.eyes):
                self.pair(cls, p, e)


class Session:
    """Renderers (one per eye), the current keyframe pair and outputs of one client."""

    def __init__(self, sid: int, cls: SessionClass, store: KeyframeStore, pair: int = 0):
        import torch

        self.sid, self.cls, self.store = sid, cls, store
        self.cfg = store.config(cls)
        self.fpa = self.cfg.keyframe_spacing
        self.pair = pair
        self.renderers: List[R.Renderer] = []
        self.outs = []
        for e in range(cls.eyes):
            r = R.Renderer(n_aux=0)
            r.set_frame(store.pair(cls, pair, e))
            self.renderers.append(r)
            self.outs.append(torch.zeros((cls.height, cls.width, 4), dtype=torch.uint8, device="cuda"))
        self.footprint = session_footprint(cls.width, cls.height, cls.eyes)
        self.cameras = {}
        self.advances = 0

    def advance(self, pair: int) -> None:
        """Moves every eye to keyframe pair `pair`. The next pair is reached as
        the reference does: swap the keyframes (the bwd one becomes fwd), then
        set the fwd slot's forward motion and the new bwd keyframe; any other
        pair is set directly. Either way the pair's flows are computed by the
        first render that follows."""
        if pair == self.pair:
            return
        for e, r in enumerate(self.renderers):
            fd = self.store.pair(self.cls, pair, e)
            if pair == self.pair + 1:
                r.swap()
            r.set_view(0, 0, fd.fwd)
            r.set_view(1, 0, fd.bwd)
        self.pair = pair
        self.advances += 1

    def info(self, pair: int, b: int, eye: int, max_iter: int) -> R.RenderInfo:
        """Client frame b of keyframe pair `pair` (alpha = b / fpa)."""
        key = (pair, b, eye)
        if key not in self.cameras:
            vp, ivp, _, _ = scene.camera(self.cfg, pair * self.fpa + b, eye)
            self.cameras[key] = (vp, ivp)
        vp, ivp = self.cameras[key]
        return R.RenderInfo((self.cls.width, self.cls.height), b / self.fpa, vp, ivp,
                            R.AdditionalRenderInfo(self.fpa, 0, False, R.RenderMode.OURS), max_iter)

    def close(self):
        for r in self.renderers:
            r.close()
        self.renderers, self.outs = [], []


@dataclass
class Request:
    sid: int
    b: int                     # client frame within the keyframe pair (alpha = b / fpa)
    arrival: float
    pair: int = 0              # keyframe pair of the store it interpolates
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


def learned_out_footprint(width: int, height: int, eyes: int = 1) -> int:
    """Bytes of the network outputs a learned-start session keeps (both
    directions, 16 fp32 channels per cell, per eye)."""
    from . import learned as Ln

    k = learned_shift(width, height)
    nw, nh = Ln.net_grid(width, height, k)
    return eyes * 2 * nw * nh * 16 * 4


def learned_shift(width: int, height: int) -> int:
    """Reduction (as a power of two, at least 2x) from a session's size to
    the learned path's network grid: the smallest that keeps the grid within
    512 cells a side (1920x1080 runs the network at 480 x 272)."""
    k = 1
    while -(-width // (1 << k)) > 512 or -(-height // (1 << k)) > 512:
        k += 1
    return k


class Server:
    """Admission by HBM footprint + continuous batching over several streams.

    The budget covers the sessions' renderer state and outputs and the
    keyframe store of every class with an admitted session (counted once per
    class, when its first session is admitted)."""

    def __init__(self, hbm_budget: int, max_batch: int = 32, n_streams: int = 2, graph: bool = True,
                 max_iter: int = -1, store: Optional[KeyframeStore] = None, search: str = "plain",
                 learned_params=None, learned_per_pair: bool = True):
        import torch

        if search not in ("plain", "learned"):
            raise ValueError("search is 'plain' or 'learned'")
        self.search = search
        self.learned_per_pair = learned_per_pair
        self._net_out: Dict[tuple, tuple] = {}  # (session id, eye) -> (pair, network output)
        self._params = learned_params
        if search == "learned" and learned_params is None:
            from . import learned as Ln

            self._params, _ = Ln.load_params(Ln.WEIGHTS)
        self._learned: Dict[tuple, object] = {}  # (class name, stream) -> LearnedInit

        self.budget = hbm_budget
        self.used = 0
        self.max_batch = max_batch
        self.max_iter = max_iter
        self.store = store or KeyframeStore()
        self.sessions: Dict[int, Session] = {}
        self.resident: set = set()  # classes whose keyframe store is counted in `used`
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
    def session_bytes(self, cls: SessionClass) -> int:
        """A session's own device bytes: renderer state, outputs and, for
        learned starts kept per pair, its network outputs."""
        fp = session_footprint(cls.width, cls.height, cls.eyes)
        if self.search == "learned" and self.learned_per_pair:
            fp += learned_out_footprint(cls.width, cls.height, cls.eyes)
        return fp

    def footprint(self, cls: SessionClass) -> int:
        """Bytes admitting a session of `cls` would add now."""
        fp = self.session_bytes(cls)
        if cls.name not in self.resident:
            fp += self.store.footprint(cls)
        return fp

    def open_session(self, cls: SessionClass, pair: int = 0) -> int:
        """Queues a session starting at keyframe pair `pair`; it becomes active
        once its footprint fits the budget."""
        need = self.session_bytes(cls) + self.store.footprint(cls)
        if need > self.budget:
            raise ValueError(f"session {cls.name} needs {need} B, more than the whole budget")
        sid = self._next_sid
        self._next_sid += 1
        self.waiting.append((sid, cls, pair))
        self._admit()
        return sid

    def _admit(self):
        while self.waiting:
            sid, cls, pair = self.waiting[0]
            fp = self.footprint(cls)
            if self.used + fp > self.budget:
                return
            self.waiting.popleft()
            self.sessions[sid] = Session(sid, cls, self.store, pair)
            self.sessions[sid].footprint = self.session_bytes(cls)
            self.resident.add(cls.name)
            self.used += fp

    def active(self, sid: int) -> bool:
        return sid in self.sessions

    def close_session(self, sid: int):
        self.drain()
        s = self.sessions.pop(sid)
        self.used -= s.footprint
        for e in range(s.cls.eyes):
            self._net_out.pop((sid, e), None)
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
        chosen, work, owners, seen = [], [], [], set()
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
            s.advance(r.pair)
            outs = r.outs if r.outs is not None else s.outs
            for e, rend in enumerate(s.renderers):
                work.append((rend, s.info(r.pair, r.b, e, self.max_iter), outs[e]))
                owners.append((s.cls, r.sid, e, r.pair))
            r.launched = now
            self.busy.add(r.sid)
        import torch

        with torch.cuda.stream(self.streams[i]):
            if self.search == "learned":
                for (cls, sid, e, pair), (rend, info, out) in zip(owners, work):
                    self._render_learned(i, cls, sid, e, pair, rend, info, out)
            else:
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
        for li in self._learned.values():
            li.close()
        self._learned = {}

    def _render_learned(self, i, cls, sid, eye, pair, rend, info, out):
        """One eye of a request on stream i through the learned path; with
        per-pair outputs, the network runs only when the eye's stored output
        belongs to another pair (a session's requests never overlap, so its
        output is not in use by another stream)."""
        import torch

        from . import learned as Ln

        li = self.learned_init(cls, i)
        st = self.streams[i]
        if not self.learned_per_pair:
            li.render(rend, out, info, stream=st)
            return
        cached = self._net_out.get((sid, eye))
        if cached is not None and cached[0] == pair:
            Ln.features(rend, info, None, st, net_size=(li.net_w, li.net_h))
            Ln.render_learned(rend, out, info, cached[1], stream=st)
            return
        buf = cached[1] if cached is not None else torch.empty_like(li.out)
        Ln.features(rend, info, li.x, st)
        li.net.forward(li.x, buf, st)
        Ln.render_learned(rend, out, info, buf, stream=st)
        self._net_out[(sid, eye)] = (pair, buf)

    def learned_init(self, cls: SessionClass, i: int):
        """The learned-start network state of stream i for sessions of cls."""
        from . import learned as Ln

        key = (cls.name, i)
        if key not in self._learned:
            self._learned[key] = Ln.LearnedInit(cls.width, cls.height, learned_shift(cls.width, cls.height),
                                                self._params)
        return self._learned[key]


def poisson_arrivals(rate_hz: float, duration_s: float, rng) -> np.ndarray:
    n = rng.poisson(rate_hz * duration_s * 1.2) + 8
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    return t[t < duration_s]


def frame_schedule(n: int, fpa: int, phase: int, n_pairs: int) -> tuple:
    """(pair, b) of a session's n-th request: client frames 1..fpa-1 of a
    pair, then the next pair (cycling through the store's pairs)."""
    per_pair = fpa - 1
    return (phase + n // per_pair) % n_pairs, 1 + n % per_pair


def run_mix(n_sessions: int = 48, rate_hz: float = 90.0, duration_s: float = 2.0, seed: int = 1,
            hbm_budget: int = 64 << 30, max_batch: int = 32, n_streams: int = 2, graph: bool = True,
            max_iter: int = -1, store: Optional[KeyframeStore] = None, search: str = "plain") -> dict:
    """Serving mix: sessions drawn from SERVING_MIX, Poisson arrivals per
    session; every session walks through the store's keyframe pairs."""
    rng = np.random.default_rng(seed)
    classes = [c for c, _ in SERVING_MIX]
    weights = np.array([w for _, w in SERVING_MIX])
    picks = rng.choice(len(classes), size=n_sessions, p=weights / weights.sum())
    store = store or KeyframeStore(seed)
    for k in sorted(set(picks)):
        store.preload(classes[k])
    srv = Server(hbm_budget, max_batch=max_batch, n_streams=n_streams, graph=graph, max_iter=max_iter, store=store,
                 search=search)
    phases = rng.integers(0, store.n_pairs, n_sessions)
    sids = [srv.open_session(classes[k], int(p)) for k, p in zip(picks, phases)]
    events = []
    for sid in sids:
        for t in poisson_arrivals(rate_hz, duration_s, rng):
            events.append((float(t), sid))
    events.sort()
    # warm up every session once (graph capture, scratch allocation)
    for sid, p in zip(sids, phases):
        if srv.active(sid):
            srv.submit(Request(sid, 1, time.perf_counter(), pair=int(p)))
    srv.drain()
    srv.stats = Stats()
    t0 = time.perf_counter()
    k = 0
    count = collections.Counter()
    while k < len(events) or srv.pending or any(f is not None for f in srv.inflight):
        now = time.perf_counter() - t0
        while k < len(events) and events[k][0] <= now:
            t, sid = events[k]
            k += 1
            if srv.active(sid):
                s = srv.sessions[sid]
                count[sid] += 1
                pair, b = frame_schedule(count[sid], s.fpa, int(phases[sid]), store.n_pairs)
                srv.submit(Request(sid, b, t0 + t, pair=pair))
        srv.step()
    out = srv.stats.summary()
    active = [srv.sessions[s] for s in sids if srv.active(s)]
    out.update({"sessions": n_sessions, "rate_hz_per_session": rate_hz, "duration_s": duration_s,
                "offered_mpix_per_s": round(sum(s.cls.width * s.cls.height * s.cls.eyes for s in active)
                                            * rate_hz / 1e6, 1),
                "pair_advances": int(sum(s.advances for s in active)),
                "hbm_used_bytes": int(srv.used),
                "streams": n_streams, "graph": graph and search == "plain", "search": search})
    srv.close()
    return out
