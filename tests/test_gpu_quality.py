"""GPU PSNR / SSIM against a PyTorch CPU float64 restatement of the
reference's definitions (scripts/eval.py:57, scripts/loss_utils.py:17-57),
plus closed forms: identical images, constant images."""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as Fn

from imagebasedspatiotemporalinterpolation_amd import quality as Q

pytestmark = pytest.mark.gpu


def _window():
    g = torch.tensor([math.exp(-(x - 5) ** 2 / float(2 * 1.5 ** 2)) for x in range(11)], dtype=torch.float32)
    g = g / g.sum()
    return (g[:, None] @ g[None, :]).double()


def _ref(a: np.ndarray, b: np.ndarray, want_map=False):
    x = torch.from_numpy(a[..., :3]).permute(2, 0, 1)[None].double() / 255.0
    y = torch.from_numpy(b[..., :3]).permute(2, 0, 1)[None].double() / 255.0
    psnr = float(-10.0 * torch.log10(Fn.mse_loss(x, y)))
    win = _window().expand(3, 1, 11, 11).contiguous()
    mu1, mu2 = Fn.conv2d(x, win, padding=5, groups=3), Fn.conv2d(y, win, padding=5, groups=3)
    s11 = Fn.conv2d(x * x, win, padding=5, groups=3) - mu1 * mu1
    s22 = Fn.conv2d(y * y, win, padding=5, groups=3) - mu2 * mu2
    s12 = Fn.conv2d(x * y, win, padding=5, groups=3) - mu1 * mu2
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu1 * mu2 + c1) * (2 * s12 + c2)) / ((mu1 * mu1 + mu2 * mu2 + c1) * (s11 + s22 + c2))
    return (psnr, float(m.mean()), m[0]) if want_map else (psnr, float(m.mean()))


@pytest.mark.parametrize("h,w", [(45, 67), (360, 640), (1080, 1920)])
def test_psnr_ssim_match_reference(h, w):
    rng = np.random.default_rng(h)
    a = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    # a correlated second image: a plus noise, clipped
    b = np.clip(a.astype(np.int16) + rng.integers(-40, 41, (h, w, 4)), 0, 255).astype(np.uint8)
    p, s = Q.psnr_ssim(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    rp, rs = _ref(a, b)
    print(f"{w}x{h}: psnr {p:.6f} vs {rp:.6f}, ssim {s:.7f} vs {rs:.7f}")
    assert abs(p - rp) < 1e-4
    assert abs(s - rs) < 2e-5


def test_identical_images():
    a = torch.from_numpy(np.random.default_rng(1).integers(0, 256, (70, 90, 4), dtype=np.uint8)).cuda()
    p, s = Q.psnr_ssim(a, a.clone())
    assert p == float("inf") and s == 1.0


def test_constant_images_closed_form():
    h, w, c1v, c2v = 64, 96, 200, 120
    a = torch.full((h, w, 4), c1v, dtype=torch.uint8, device="cuda")
    b = torch.full((h, w, 4), c2v, dtype=torch.uint8, device="cuda")
    p, s = Q.psnr_ssim(a, b)
    assert abs(p - (-10 * math.log10(((c1v - c2v) / 255) ** 2))) < 1e-4
    # away from the zero-padded border, sigma = 0 and SSIM = (2 m1 m2 + C1) / (m1^2 + m2^2 + C1)
    m1, m2 = c1v / 255, c2v / 255
    interior = (2 * m1 * m2 + 1e-4) / (m1 * m1 + m2 * m2 + 1e-4)
    _, rs, rmap = _ref(a.cpu().numpy(), b.cpu().numpy(), want_map=True)
    assert torch.allclose(rmap[:, 5:-5, 5:-5], torch.full_like(rmap[:, 5:-5, 5:-5], interior), atol=1e-12)
    # single precision (as in the reference) cancels in E[x^2] - mu^2 ~ 0 and
    # divides the residue by C2 = 9e-4: a few 1e-5 of absolute error
    assert abs(s - rs) < 1e-4
