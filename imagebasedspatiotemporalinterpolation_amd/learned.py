"""Learned search initialiser: ctypes binding of include/isti_learned.h.

UNet wraps the network (hand-written HIP forward/backward kernels on MFMA);
the module-level helpers connect it to a Renderer: capture of the plain
search's results (training targets), feature building, the training loss and
the learned-start render. Tensors are torch GPU tensors (device memory only).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int32, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

from . import _native as N

CH = 16
# (channels from the first input, channels from the skip input, outputs, kernel,
#  first input upsampled 2x, ReLU): e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b head
LAYERS = [
    (16, 0, 16, 3, 0, 1), (16, 0, 16, 3, 0, 1), (16, 0, 32, 3, 0, 1), (32, 0, 32, 3, 0, 1),
    (32, 0, 64, 3, 0, 1), (64, 0, 64, 3, 0, 1), (64, 0, 128, 3, 0, 1), (128, 0, 128, 3, 0, 1),
    (128, 64, 64, 3, 1, 1), (64, 0, 64, 3, 0, 1), (64, 32, 32, 3, 1, 1), (32, 0, 32, 3, 0, 1),
    (32, 16, 16, 3, 1, 1), (16, 0, 16, 3, 0, 1), (16, 0, 16, 1, 0, 0),
]

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
        L.isti_set_search_capture.argtypes = [c_void_p, c_void_p]
        L.isti_learned_features.argtypes = [c_void_p, POINTER(N.RenderInfoC), c_void_p, c_uint32, c_uint32, c_void_p]
        L.isti_learned_loss.argtypes = [c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_uint32, c_float, c_void_p,
                                        c_void_p, c_void_p]
        L.isti_render_learned.argtypes = [c_void_p, POINTER(N.RenderInfoC), c_void_p, c_uint32, c_uint32, c_float,
                                          c_void_p, c_void_p, c_size_t, c_void_p]
        for f in ("isti_unet_create", "isti_unet_destroy", "isti_unet_get_params", "isti_unet_set_params",
                  "isti_unet_get_grads", "isti_unet_forward", "isti_unet_backward", "isti_unet_zero_grads",
                  "isti_unet_adam", "isti_set_search_capture", "isti_learned_features", "isti_learned_loss",
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


def features(renderer, info, x, stream=None) -> None:
    """x: int16 [2, net_h, net_w, 16] for the frame `info` the renderer has just rendered."""
    ci = info.to_c()
    n_, nh, nw, _ = x.shape
    _ok(_lib().isti_learned_features(renderer._h, ctypes.byref(ci), x.data_ptr(), nw, nh, _stream(stream)),
        "features")


def loss(out, q_target, width, height, dout, loss_acc, conf_weight=0.1, stream=None) -> None:
    _, nh, nw, _ = out.shape
    _ok(_lib().isti_learned_loss(out.data_ptr(), q_target.data_ptr(), width, height, nw, nh, conf_weight,
                                 dout.data_ptr(), loss_acc.data_ptr(), _stream(stream)), "loss")


def render_learned(renderer, render_buffer, info, net_out, conf_threshold=0.5, stats=None, stream=None) -> None:
    ci = info.to_c()
    _, nh, nw, _ = net_out.shape
    h, w = render_buffer.shape[:2]
    st = _lib().isti_render_learned(renderer._h, ctypes.byref(ci), net_out.data_ptr(), nw, nh, conf_threshold,
                                    stats.data_ptr() if stats is not None else None, render_buffer.data_ptr(), w * 4,
                                    _stream(stream))
    N.check(st, renderer._h)
