"""HIP forward/backward of the learned initialiser's network against a
PyTorch CPU float64 reference of the same op graph."""
import numpy as np
import pytest
import torch

from imagebasedspatiotemporalinterpolation_amd import learned as Ln
from tests import _unet_ref as ref

pytestmark = pytest.mark.gpu


def _bf16_bits(a: np.ndarray) -> np.ndarray:
    return torch.from_numpy(a.astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy()


def _run(n, h, w, seed):
    """GPU forward/backward on seeded parameters and inputs; returns what the
    reference needs and the GPU's outputs."""
    net = Ln.UNet(w, h, batch=n, seed=5)
    rng = np.random.default_rng(seed)
    p = net.get_params()
    p += rng.normal(0, 0.02, p.shape).astype(np.float32) * (p != 0)  # non-zero biases would be zero after init
    for _, _, b_off, co in Ln.param_layout()[:-1]:
        p[b_off:b_off + co] = rng.normal(0, 0.05, co)
    net.set_params(p)
    x = rng.uniform(-1, 1, (n, h, w, 16)).astype(np.float32)
    x[..., 15] = 0
    xb = torch.from_numpy(_bf16_bits(x)).cuda()  # alive until backward has read it
    out = torch.zeros((n, h, w, 16), dtype=torch.float32, device="cuda")
    net.forward(xb, out)
    dout = rng.normal(0, 1, (n, h, w, 16)).astype(np.float32)
    dout[..., 3:] = 0
    net.zero_grads()
    net.backward(torch.from_numpy(dout).cuda())
    torch.cuda.synchronize()
    got_out, got_g = out.cpu().numpy(), net.get_grads()
    net.close()
    return p, x, dout, got_out, got_g


def _reference(p, x, dout, bf16_grads):
    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    ro = ref.forward(P, torch.from_numpy(x).to(torch.bfloat16).to(torch.float64), bf16_grads=bf16_grads)
    (ro * torch.from_numpy(dout).double()).sum().backward()
    return ro.detach().numpy(), P.grad.numpy()


# Activations are rounded to bf16 after every layer. Where an exact
# pre-activation sits within float rounding of a bf16 rounding boundary, the
# GPU's fp32 sum and the reference's float64 sum round to neighbouring bf16
# values, and the difference grows through the later layers (observed: a
# forward error of 2e-4..6e-3 and gradient errors up to 15 % on such inputs,
# varying from seed to seed). Such ties show in the forward output, so each
# case takes the first seed whose forward pass matches the reference to float
# rounding (a tie-free input) and checks its gradients tightly there: every
# layer's weight and bias gradient and every kernel tap's slice within 1 %
# of the bf16-emulating reference (a tap left out of a gradient would be off
# by 1/9 of the layer or 100 % of the tap) and within 2 % of plain float64.
@pytest.mark.parametrize("n,h,w", [(1, 16, 48), (2, 16, 48), (2, 32, 32), (1, 64, 32)])
def test_forward_backward_match_reference(n, h, w):
    lay = Ln.param_layout()
    for seed in (1, 100, 101, 102, 103, 104, 105):
        p, x, dout, got_out, got_g = _run(n, h, w, seed)
        exp_out, _ = _reference(p, x, dout, True)
        e = np.linalg.norm(got_out[..., :3] - exp_out[..., :3]) / np.linalg.norm(exp_out[..., :3])
        print(f"seed {seed}: forward rel error {e:.2e}")
        assert e < 1e-2, f"forward rel error {e}"
        assert np.all(got_out[..., 3:] == 0)
        if e < 1e-5:
            break
    else:
        pytest.fail("no tie-free input among the seeds")
    for bf16_grads, tol in ((True, 1e-2), (False, 2e-2)):
        _, exp_g = _reference(p, x, dout, bf16_grads)
        assert np.all(np.isfinite(got_g))
        for l, (w_off, shape, b_off, co) in enumerate(lay):
            nw = int(np.prod(shape))
            gw, ew = got_g[w_off:w_off + nw], exp_g[w_off:w_off + nw]
            gb, eb = got_g[b_off:b_off + co], exp_g[b_off:b_off + co]
            if l == len(lay) - 1:  # head: only outputs 0..2 live
                gw, ew = gw.reshape(shape)[:3].ravel(), ew.reshape(shape)[:3].ravel()
                gb, eb = gb[:3], eb[:3]
            if np.linalg.norm(ew) == 0.0:  # a layer whose ReLUs are all off here
                assert np.all(gw == 0) and np.all(gb == 0), l
                continue
            rw = np.linalg.norm(gw - ew) / np.linalg.norm(ew)
            rb = np.linalg.norm(gb - eb) / max(np.linalg.norm(eb), 1e-12)
            # per kernel tap: the gradient slice of each (ky, kx) over all (co, ci)
            ks = shape[1]
            gt, et = gw.reshape(-1, ks * ks, shape[3]), ew.reshape(-1, ks * ks, shape[3])
            rt = [np.linalg.norm(gt[:, k] - et[:, k]) / max(np.linalg.norm(et[:, k]), 1e-12) for k in range(ks * ks)]
            print(f"bf16 grads {bf16_grads} layer {l}: weight {rw:.4f}, bias {rb:.4f}, worst tap {max(rt):.4f}")
            assert rw < tol and rb < tol and max(rt) < 2 * tol, (bf16_grads, l, rw, rb, rt)


def test_adam_reduces_a_simple_loss():
    """A few Adam steps on a fixed target lower the loss (training loop plumbing)."""
    n, h, w = 1, 16, 16
    net = Ln.UNet(w, h, batch=n, seed=2)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(_bf16_bits(rng.uniform(-1, 1, (n, h, w, 16)) * (np.arange(16) < 15))).cuda()
    target = torch.zeros((n, h, w, 16), device="cuda")
    target[..., 0] = 1.5
    out = torch.zeros((n, h, w, 16), device="cuda")
    losses = []
    for it in range(30):
        net.forward(x, out)
        d = (out - target)
        d[..., 1:] = 0
        losses.append(float((d ** 2).sum().item()))
        net.zero_grads()
        net.backward(2 * d / d.numel())
        net.adam(lr=3e-3)
    assert losses[-1] < 0.5 * losses[0], losses
    net.close()


def test_autotuned_shapes_give_the_same_bits():
    """isti_unet_autotune only picks among wave shapes that sum every output
    in the same order: the forward pass after tuning equals the one before,
    bit for bit, at the 1080p network grid and at a small odd one."""
    for w, h in ((480, 272), (72, 40)):
        net = Ln.UNet(w, h, batch=2, seed=3)
        rng = np.random.default_rng(w)
        x = torch.from_numpy(_bf16_bits(rng.uniform(-1, 1, (2, h, w, 16)))).cuda()
        a = torch.zeros((2, h, w, 16), dtype=torch.float32, device="cuda")
        b = torch.zeros_like(a)
        net.forward(x, a)
        net.autotune(reps=2)
        net.forward(x, b)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (w, h)
        net.close()
