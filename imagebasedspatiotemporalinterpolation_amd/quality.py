"""PSNR and SSIM of RGBA8 GPU images (binding of include/isti_quality.h).

Definitions of the reference's evaluation: PSNR = -10 log10(MSE) over RGB
in [0, 1] (scripts/eval.py:57); SSIM with an 11x11 Gaussian window,
sigma 1.5 (scripts/loss_utils.py:17-57). Images are torch uint8 [H, W, 4]
CUDA tensors; only RGB is scored.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import c_size_t, c_uint32, c_void_p

from . import _native as N

_bound = False


def _lib():
    global _bound
    L = N.lib()
    if not _bound:
        L.isti_image_quality.argtypes = [c_void_p, c_size_t, c_void_p, c_size_t, c_uint32, c_uint32, c_void_p,
                                         c_void_p]
        L.isti_image_quality.restype = ctypes.c_int32
        _bound = True
    return L


def accumulate(a, b, acc, stream=None) -> None:
    """acc (float64 CUDA tensor [2]) <- (sum of squared RGB differences, sum of the SSIM map)."""
    import torch

    for t in (a, b):
        if not (t.is_cuda and t.dtype == torch.uint8 and t.dim() == 3 and t.shape[2] == 4 and t.is_contiguous()):
            raise ValueError("images must be contiguous uint8 CUDA tensors [H, W, 4]")
    if a.shape != b.shape:
        raise ValueError("images differ in size")
    if not (acc.is_cuda and acc.dtype == torch.float64 and acc.numel() >= 2):
        raise ValueError("acc must be a float64 CUDA tensor of 2 elements")
    h, w = a.shape[:2]
    s = stream if stream is not None else torch.cuda.current_stream()
    st = _lib().isti_image_quality(a.data_ptr(), w * 4, b.data_ptr(), w * 4, w, h, acc.data_ptr(),
                                   c_void_p(s.cuda_stream))
    if st != N.ISTI_OK:
        raise N.IstiError(st, "isti_image_quality")


def psnr_ssim(a, b) -> tuple:
    """(PSNR dB, SSIM) of two images; synchronises."""
    import torch

    acc = torch.zeros(2, dtype=torch.float64, device=a.device)
    accumulate(a, b, acc)
    sse, ssim_sum = acc.cpu().tolist()
    n = 3.0 * a.shape[0] * a.shape[1]
    mse = sse / n
    return (float("inf") if mse == 0.0 else -10.0 * math.log10(mse)), ssim_sum / n
