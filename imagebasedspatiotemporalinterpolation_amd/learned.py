"""Learned search initialiser: ctypes binding of include/isti_learned.h,
training on the GPU, checkpoints and evaluation against the plain search.

UNet wraps the network (hand-written HIP forward/backward kernels on MFMA);
the module-level helpers connect it to a Renderer: capture of the plain
search's results (training targets), feature building, the training loss and
the learned-start render. Tensors are torch GPU tensors (device memory only).

    python -m imagebasedspatiotemporalinterpolation_amd.learned train --out unet.bin
    python -m imagebasedspatiotemporalinterpolation_amd.learned eval --weights unet.bin
"""
from __future__ import annotations

import ctypes
import json
import struct
import time
from concurrent.futures import ThreadPoolExecutor
from ctypes import POINTER, c_float, c_int32, c_size_t, c_uint32, c_uint64, c_void_p
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

from . import _native as N
from . import scene as S

CH = 16
# (channels from the first input, channels from the skip input, outputs, kernel,
#  first input upsampled 2x, ReLU): e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b head
LAYERS = [
    (16, 0, 16, 3, 0, 1), (16, 0, 16, 3, 0, 1), (16, 0, 32, 3, 0, 1), (32, 0, 32, 3, 0, 1),
    (32, 0, 64, 3, 0, 1), (64, 0, 64, 3, 0, 1), (64, 0, 128, 3, 0, 1), (128, 0, 128, 3, 0, 1),
    (128, 64, 64, 3, 1, 1), (64, 0, 64, 3, 0, 1), (64, 32, 32, 3, 1, 1), (32, 0, 32, 3, 0, 1),
    (32, 16, 16, 3, 1, 1), (16, 0, 16, 3, 0, 1), (16, 0, 16, 1, 0, 0),
]

LEVEL = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0]  # resolution level of each layer (size / 2^level)


def flop_per_cell() -> float:
    """Forward FLOPs per network cell (one multiply-add = 2), from LAYERS:
    each layer's 2 * cout * k^2 * cin at its level's share of the cells."""
    return float(sum(2 * co * ks * ks * (ca + cb) / 4 ** lv for (ca, cb, co, ks, _, _), lv in zip(LAYERS, LEVEL)))


_bound = False


def _lib():
    global _bound
    L = N.lib()
    if not _bound:
        L.isti_unet_create.argtypes = [c_int32, c_uint32, c_uint32, c_int32, c_uint64, POINTER(c_void_p)]
        L.isti_unet_destroy.argtypes = [c_void_p]
        L.isti_unet_param_count.argtypes = []
        L.isti_unet_param_count.restype = c_size_t
        for f in ("isti_unet_get_params", "isti_unet_set_params", "isti_unet_get_grads"):
            getattr(L, f).argtypes = [c_void_p, c_void_p]
        L.isti_unet_forward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
        L.isti_unet_backward.argtypes = [c_void_p, c_void_p, c_void_p]
        L.isti_unet_zero_grads.argtypes = [c_void_p, c_void_p]
        L.isti_unet_adam.argtypes = [c_void_p, c_float, c_float, c_float, c_float, c_void_p]
        L.isti_unet_autotune.argtypes = [c_void_p, c_int32, c_void_p]
        L.isti_set_search_capture.argtypes = [c_void_p, c_void_p]
        L.isti_learned_features.argtypes = [c_void_p, POINTER(N.RenderInfoC), c_void_p, c_uint32, c_uint32, c_void_p]
        L.isti_learned_loss.argtypes = [c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_float, c_void_p,
                                        c_void_p, c_void_p]
        L.isti_render_learned.argtypes = [c_void_p, POINTER(N.RenderInfoC), c_void_p, c_uint32, c_uint32, c_float,
                                          c_void_p, c_void_p, c_size_t, c_void_p]
        for f in ("isti_unet_create", "isti_unet_destroy", "isti_unet_get_params", "isti_unet_set_params",
                  "isti_unet_get_grads", "isti_unet_forward", "isti_unet_backward", "isti_unet_zero_grads",
                  "isti_unet_adam", "isti_unet_autotune", "isti_set_search_capture", "isti_learned_features",
                  "isti_learned_loss",
                  "isti_render_learned"):
            getattr(L, f).restype = c_int32
        _bound = True
    return L


def _stream(stream):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def _ok(st, what):
    if st != N.ISTI_OK:
        raise N.IstiError(st, what)


def param_layout():
    """[(w_offset, (cout, ks, ks, cin), b_offset, cout)] per layer, in parameter order."""
    out, off = [], 0
    for ca, cb, co, ks, _, _ in LAYERS:
        shape = (co, ks, ks, ca + cb)
        n = co * ks * ks * (ca + cb)
        out.append((off, shape, off + n, co))
        off += n + co
    return out


class UNet:
    def __init__(self, width: int, height: int, batch: int = 2, seed: int = 0, device: int = 0):
        h = c_void_p()
        _ok(_lib().isti_unet_create(device, width, height, batch, seed, ctypes.byref(h)), "isti_unet_create")
        self._h, self.width, self.height, self.batch = h, width, height, batch
        self.n_params = int(_lib().isti_unet_param_count())

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _ok(_lib().isti_unet_get_params(self._h, p.ctypes.data), "get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _ok(_lib().isti_unet_set_params(self._h, p.ctypes.data), "set_params")

    def get_grads(self) -> np.ndarray:
        g = np.zeros(self.n_params, np.float32)
        _ok(_lib().isti_unet_get_grads(self._h, g.ctypes.data), "get_grads")
        return g

    def forward(self, x, out, stream=None) -> None:
        """x: int16 (bf16 bits) [batch, height, width, 16]; out: float32 same shape."""
        _ok(_lib().isti_unet_forward(self._h, x.data_ptr(), out.data_ptr(), _stream(stream)), "forward")

    def autotune(self, reps: int = 10, stream=None) -> None:
        """Picks the fastest wave shape of every forward convolution on this
        size (same bits for every shape); synchronises, clobbers activations."""
        _ok(_lib().isti_unet_autotune(self._h, reps, _stream(stream)), "autotune")

    def backward(self, dout, stream=None) -> None:
        _ok(_lib().isti_unet_backward(self._h, dout.data_ptr(), _stream(stream)), "backward")

    def zero_grads(self, stream=None) -> None:
        _ok(_lib().isti_unet_zero_grads(self._h, _stream(stream)), "zero_grads")

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None) -> None:
        _ok(_lib().isti_unet_adam(self._h, lr, beta1, beta2, eps, _stream(stream)), "adam")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().isti_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def net_grid(width: int, height: int, shift: int = 1):
    """Network grid for a render size: size / 2^shift rounded up to multiples of 8."""
    cw, ch = -(-width // (1 << shift)), -(-height // (1 << shift))
    return -(-cw // 8) * 8, -(-ch // 8) * 8


def set_search_capture(renderer, q_out) -> None:
    """q_out: int32/uint32 GPU tensor [H, W, 2], or None to disable."""
    _ok(_lib().isti_set_search_capture(renderer._h, q_out.data_ptr() if q_out is not None else None), "capture")


def features(renderer, info, x, stream=None, net_size=None) -> None:
    """Splat stage of frame `info` plus the network input x (int16 bf16 bits,
    [2, net_h, net_w, 16]); prepares the frame for loss() and render_learned().
    x None runs the splat stage alone, for a frame that reuses the network
    output of an earlier frame of its keyframe pair (net_size = (net_w, net_h))."""
    ci = info.to_c()
    if x is None:
        nw, nh = net_size
        ptr = None
    else:
        n_, nh, nw, _ = x.shape
        ptr = x.data_ptr()
    N.check(_lib().isti_learned_features(renderer._h, ctypes.byref(ci), ptr, nw, nh, _stream(stream)), renderer._h)


def loss(renderer, out, q_target, dout, loss_acc, conf_weight=0.1, stream=None) -> None:
    """Training loss of `out` for the frame the last features() call prepared."""
    _, nh, nw, _ = out.shape
    st = _lib().isti_learned_loss(renderer._h, out.data_ptr(), q_target.data_ptr(), nw, nh, conf_weight,
                                  dout.data_ptr(), loss_acc.data_ptr(), _stream(stream))
    N.check(st, renderer._h)


def render_learned(renderer, render_buffer, info, net_out, conf_threshold=0.5, stats=None, stream=None) -> None:
    ci = info.to_c()
    _, nh, nw, _ = net_out.shape
    h, w = render_buffer.shape[:2]
    st = _lib().isti_render_learned(renderer._h, ctypes.byref(ci), net_out.data_ptr(), nw, nh, conf_threshold,
                                    stats.data_ptr() if stats is not None else None, render_buffer.data_ptr(), w * 4,
                                    _stream(stream))
    N.check(st, renderer._h)


# ---------------------------------------------------------------------------
# Checkpoints: "ISTIUNET", u32 format version, u32 parameter count, u32 JSON
# metadata length, the metadata, then the float32 parameters (little endian).
# ---------------------------------------------------------------------------
MAGIC = b"ISTIUNET"
WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_c1_960x540.bin"


def save_params(path, params: np.ndarray, meta: dict) -> None:
    params = np.ascontiguousarray(params, "<f4")
    blob = json.dumps(meta, sort_keys=True).encode()
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<III", 1, params.size, len(blob)) + blob)
        f.write(params.tobytes())


def load_params(path) -> tuple:
    data = Path(path).read_bytes()
    if data[:8] != MAGIC:
        raise ValueError(f"{path}: not a U-Net checkpoint")
    ver, n, m = struct.unpack_from("<III", data, 8)
    if ver != 1:
        raise ValueError(f"{path}: unknown checkpoint version {ver}")
    meta = json.loads(data[20:20 + m].decode())
    params = np.frombuffer(data, "<f4", count=n, offset=20 + m).astype(np.float32)
    if n != int(_lib().isti_unet_param_count()):
        raise ValueError(f"{path}: {n} parameters, the network has {int(_lib().isti_unet_param_count())}")
    return params, meta


# ---------------------------------------------------------------------------
# Inference: features -> network -> learned-start resolve, per frame
# ---------------------------------------------------------------------------
class LearnedInit:
    """Learned search starts for renders of one size (network at size / 2^shift)."""

    def __init__(self, width: int, height: int, shift: int, params: np.ndarray = None, device: int = 0):
        import torch

        self.width, self.height, self.shift = width, height, shift
        self.net_w, self.net_h = net_grid(width, height, shift)
        self.net = UNet(self.net_w, self.net_h, batch=2, device=device)
        if params is not None:
            self.net.set_params(params)
        self.net.autotune()
        dev = f"cuda:{device}"
        self.x = torch.zeros((2, self.net_h, self.net_w, CH), dtype=torch.int16, device=dev)
        self.out = torch.zeros((2, self.net_h, self.net_w, CH), dtype=torch.float32, device=dev)
        self.stats = torch.zeros(8, dtype=torch.int32, device=dev)

    def render(self, renderer, render_buffer, info, conf_threshold: float = 0.5, with_stats: bool = False,
               stream=None, reuse: bool = False) -> None:
        """Splat stage + features, network forward, learned-start resolve (stream-ordered).

        reuse=True skips the features and the network and resolves with the
        network output of the last frame rendered without reuse: the network
        runs once per keyframe pair, on its first frame (render_pair_frame)."""
        if reuse:
            features(renderer, info, None, stream, net_size=(self.net_w, self.net_h))
        else:
            features(renderer, info, self.x, stream)
            self.net.forward(self.x, self.out, stream)
        render_learned(renderer, render_buffer, info, self.out, conf_threshold,
                       self.stats if with_stats else None, stream)

    def render_pair_frame(self, renderer, render_buffer, info, first_of_pair: bool, conf_threshold: float = 0.5,
                          with_stats: bool = False, stream=None) -> None:
        """The per-pair schedule: the network runs on the first frame of a
        keyframe pair and its output serves the pair's later frames."""
        self.render(renderer, render_buffer, info, conf_threshold, with_stats, stream, reuse=not first_of_pair)

    def read_stats(self, reset: bool = True) -> dict:
        v = self.stats.cpu().numpy().astype(np.int64)
        if reset:
            self.stats.zero_()
        return {"kept": int(v[0]), "rejected": int(v[1]), "low_confidence": int(v[2]), "searches": int(v[3]),
                "steps": int(v[4])}

    def close(self):
        self.net.close()


# ---------------------------------------------------------------------------
# Training on plain-search results of seeded synthetic keyframe pairs
# ---------------------------------------------------------------------------
@dataclass
class TrainConfig:
    pairs: int = 4000          # keyframe pairs (one training step each)
    sequences: int = 40        # seeded scenes; pairs are consecutive keyframes of each
    workload: int = 1          # scene generator workload (1: checker plane + 16 spheres)
    width: int = 960
    height: int = 540
    shift: int = 1             # network at render size / 2
    max_iter: int = 32         # plain search cap of the target renders
    lr: float = 1e-3
    lr_final: float = 1e-4     # cosine decay to this
    conf_weight: float = 0.1
    seed: int = 0
    seed_base: int = 100000    # scene seeds seed_base + sequence


def _pair_data(cfg: TrainConfig, i: int):
    """Keyframes and target camera of training pair i (deterministic)."""
    per = -(-cfg.pairs // cfg.sequences)
    seq, k = divmod(i, per)
    sc = S.defaults(cfg.workload, cfg.seed_base + seq, width=cfg.width, height=cfg.height)
    fwd = S.keyframe(sc, k, 0, +1)
    bwd = S.keyframe(sc, k + 1, 0, -1)
    b = 1 + (i % (sc.keyframe_spacing - 1))  # client frame inside the pair
    vp, ivp, _, _ = S.camera(sc, k * sc.keyframe_spacing + b)
    return fwd, bwd, vp, ivp, b / sc.keyframe_spacing, sc.keyframe_spacing


def _frame_objects(fwd, bwd, vp, ivp, alpha, fpa, w, h, max_iter):
    from .renderer import AdditionalRenderInfo, FrameData, KeyframeView, RenderInfo

    fd = FrameData(KeyframeView.from_numpy(fwd), KeyframeView.from_numpy(bwd))
    info = RenderInfo((w, h), alpha, vp, ivp, AdditionalRenderInfo(fpa, 0, True, 0), max_search_iterations=max_iter)
    return fd, info


def train(cfg: TrainConfig = TrainConfig(), init_params: np.ndarray = None, log=print, log_every: int = 100):
    """Trains the network on the GPU; returns (params, per-step losses)."""
    import torch

    from .renderer import Renderer

    w, h = cfg.width, cfg.height
    nw, nh = net_grid(w, h, cfg.shift)
    net = UNet(nw, nh, batch=2, seed=cfg.seed)
    if init_params is not None:
        net.set_params(init_params)
    r = Renderer(0)
    buf = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    q = torch.zeros((h, w, 2), dtype=torch.int32, device="cuda")
    x = torch.zeros((2, nh, nw, CH), dtype=torch.int16, device="cuda")
    out = torch.zeros((2, nh, nw, CH), dtype=torch.float32, device="cuda")
    dout = torch.zeros_like(out)
    loss_acc = torch.zeros(cfg.pairs, dtype=torch.float32, device="cuda")
    set_search_capture(r, q)
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=2) as pool:
        ahead = [pool.submit(_pair_data, cfg, i) for i in range(min(4, cfg.pairs))]
        for i in range(cfg.pairs):
            fwd, bwd, vp, ivp, alpha, fpa = ahead[i].result()
            ahead[i] = None
            if i + 4 < cfg.pairs:
                ahead.append(pool.submit(_pair_data, cfg, i + 4))
            fd, info = _frame_objects(fwd, bwd, vp, ivp, alpha, fpa, w, h, cfg.max_iter)
            r.render(buf, info, fd)                 # plain search -> captured targets
            features(r, info, x)
            net.forward(x, out)
            loss(r, out, q, dout, loss_acc[i:i + 1], cfg.conf_weight)
            net.zero_grads()
            net.backward(dout)
            frac = i / max(cfg.pairs - 1, 1)
            lr = cfg.lr_final + 0.5 * (cfg.lr - cfg.lr_final) * (1.0 + np.cos(np.pi * frac))
            net.adam(lr)
            if log and ((i + 1) % log_every == 0 or i + 1 == cfg.pairs):
                recent = loss_acc[max(0, i + 1 - log_every):i + 1].mean().item()
                log(f"step {i + 1}/{cfg.pairs} loss {recent:.4f} lr {lr:.2e} {time.time() - t0:.1f}s")
    set_search_capture(r, None)
    torch.cuda.synchronize()
    params = net.get_params()
    losses = loss_acc.cpu().numpy()
    net.close()
    r.close()
    return params, losses


# ---------------------------------------------------------------------------
# Evaluation against the plain search (same kernel, learned starts disabled)
# ---------------------------------------------------------------------------
def evaluate(params: np.ndarray, width: int = 1920, height: int = 1080, shift: int = 2, workload: int = 1,
             seeds=(900001, 900002), pairs_per_seed: int = 2, conf_threshold: float = 0.5, max_iter: int = 32,
             per_pair: bool = False):
    """Iterations saved, fallback rate and PSNR of learned starts vs the plain search.

    PSNR (RGB, the reference's scripts/eval.py:57,68) is measured against the
    generator's ground-truth render of each target frame, over pixels both
    methods wrote; `psnr_learned_vs_plain` compares the two outputs directly.
    per_pair: the network runs on the first frame of each keyframe pair only
    and the pair's later frames reuse its output.
    """
    import torch

    from .renderer import Renderer

    r = Renderer(0)
    li = LearnedInit(width, height, shift, params)
    plain_buf = torch.zeros((height, width, 4), dtype=torch.uint8, device="cuda")
    learn_buf = torch.zeros_like(plain_buf)
    tot_p = {"steps": 0, "searches": 0}
    tot_l = {"kept": 0, "rejected": 0, "low_confidence": 0, "searches": 0, "steps": 0}
    mse_p = mse_l = mse_pl = 0.0
    n_px = 0
    for sd in seeds:
        sc = S.defaults(workload, sd, width=width, height=height)
        for k in range(pairs_per_seed):
            fwd, bwd = S.keyframe(sc, k, 0, +1), S.keyframe(sc, k + 1, 0, -1)
            for b in range(1, sc.keyframe_spacing):
                vp, ivp, _, _ = S.camera(sc, k * sc.keyframe_spacing + b)
                fd, info = _frame_objects(fwd, bwd, vp, ivp, b / sc.keyframe_spacing, sc.keyframe_spacing, width,
                                          height, max_iter)
                r.set_frame(fd)
                plain_buf.zero_()
                learn_buf.zero_()
                # plain starts, counted (threshold 1 reads no network output)
                li.render(r, plain_buf, info, conf_threshold=1.0, with_stats=True, reuse=True)
                sp = li.read_stats()
                li.render(r, learn_buf, info, conf_threshold=conf_threshold, with_stats=True,
                          reuse=per_pair and b > 1)
                sl = li.read_stats()
                for key in tot_p:
                    tot_p[key] += sp[key]
                for key in tot_l:
                    tot_l[key] += sl[key]
                gt = S.ground_truth(sc, k * sc.keyframe_spacing + b)
                a, c = plain_buf.cpu().numpy(), learn_buf.cpu().numpy()
                both = (a[..., 3] > 0) & (c[..., 3] > 0)
                g = gt[..., :3][both].astype(np.float64) / 255.0
                pa = a[..., :3][both].astype(np.float64) / 255.0
                pc = c[..., :3][both].astype(np.float64) / 255.0
                mse_p += float(((pa - g) ** 2).sum())
                mse_l += float(((pc - g) ** 2).sum())
                mse_pl += float(((pa - pc) ** 2).sum())
                n_px += int(both.sum()) * 3
    li.close()
    r.close()

    def db(m):
        return float("inf") if m == 0 else -10.0 * np.log10(m / max(n_px, 1))

    attempts = tot_l["kept"] + tot_l["rejected"]
    res = {
        "size": [width, height], "net_shift": shift, "frames": len(seeds) * pairs_per_seed * (sc.keyframe_spacing - 1),
        "searches": tot_l["searches"],
        "steps_per_search_plain": tot_p["steps"] / max(tot_p["searches"], 1),
        "steps_per_search_learned": tot_l["steps"] / max(tot_l["searches"], 1),
        "learned_start_rate": attempts / max(tot_l["searches"], 1),
        "kept_rate": tot_l["kept"] / max(tot_l["searches"], 1),
        "fallback_rate": tot_l["rejected"] / max(attempts, 1),
        "psnr_plain_vs_gt": db(mse_p), "psnr_learned_vs_gt": db(mse_l),
        "psnr_learned_vs_plain": db(mse_pl),
    }
    res["iterations_saved_per_search"] = res["steps_per_search_plain"] - res["steps_per_search_learned"]
    res["delta_psnr_db"] = res["psnr_learned_vs_gt"] - res["psnr_plain_vs_gt"]
    return res


def time_forward(width: int = 480, height: int = 272, reps: int = 100) -> dict:
    """Average time of the network's forward pass (both directions, batch 2)
    at width x height cells, on seeded inputs, after 5 untimed passes."""
    import time

    import torch

    params = load_params(WEIGHTS)[0] if WEIGHTS.exists() else None
    net = UNet(width, height, batch=2)
    if params is not None:
        net.set_params(params)
    net.autotune()
    g = torch.Generator().manual_seed(0)
    x = torch.randn((2, height, width, CH), generator=g).to(torch.bfloat16).view(torch.int16).cuda()
    out = torch.zeros((2, height, width, CH), dtype=torch.float32, device="cuda")
    for _ in range(5):
        net.forward(x, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(x, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    net.close()
    gflop = flop_per_cell() * 2 * width * height / 1e9
    return {"width": width, "height": height, "forward_ms": round(ms, 4), "gflop": round(gflop, 2),
            "tflop_s": round(gflop / ms, 1)}


def _main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    for f, v in asdict(TrainConfig()).items():
        t.add_argument("--" + f.replace("_", "-"), type=type(v), default=v)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=str(WEIGHTS))
    e.add_argument("--width", type=int, default=1920)
    e.add_argument("--height", type=int, default=1080)
    e.add_argument("--shift", type=int, default=2)
    e.add_argument("--conf", type=float, default=0.5)
    e.add_argument("--json", default=None)
    e.add_argument("--per-pair", action="store_true", help="network once per keyframe pair")
    f = sub.add_parser("forward", help="time the network's forward pass (shipped weights, seeded input)")
    f.add_argument("--width", type=int, default=480)
    f.add_argument("--height", type=int, default=272)
    f.add_argument("--reps", type=int, default=100)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        cfg = TrainConfig(**{f: getattr(a, f) for f in asdict(TrainConfig())})
        params, losses = train(cfg, log=lambda m: print(m, flush=True))
        meta = {"train": asdict(cfg), "loss_first100": float(losses[:100].mean()),
                "loss_last100": float(losses[-100:].mean())}
        save_params(a.out, params, meta)
        print(json.dumps(meta), flush=True)
    elif a.cmd == "forward":
        print(json.dumps(time_forward(a.width, a.height, a.reps)), flush=True)
    else:
        params, meta = load_params(a.weights)
        res = evaluate(params, a.width, a.height, a.shift, conf_threshold=a.conf, per_pair=a.per_pair)
        print(json.dumps(res), flush=True)
        if a.json:
            Path(a.json).write_text(json.dumps(res, indent=1) + "\n")


if __name__ == "__main__":
    _main()
