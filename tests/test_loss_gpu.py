"""GPU checks of the fused square-root-intensity loss (``pty_sqrt_intensity_loss``) behind
``serving.loss_and_grad_intensity``: the loss sum (sqrt(I + eps) - sqrt(I_meas))^2 and dL/dI =
(sqrt(I + eps) - sqrt(I_meas)) / sqrt(I + eps) against a float64 numpy evaluation of the same formula,
bitwise reproducibility of the FP64 loss, and the autograd path."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _case(shape, seed):
    rng = np.random.default_rng(seed)
    pat = rng.gamma(0.5, 2.0, shape).astype(np.float32)
    meas = rng.poisson(pat * 50.0).astype(np.float32) / 50.0  # counts: many exact zeros
    pat.flat[::97] = 0.0  # dark pixels: the loss takes sqrt(eps) there
    return pat, meas


@pytest.mark.parametrize("shape", [(3, 17, 19), (64, 128, 128), (1,)])
def test_sqrt_loss_matches_fp64_formula(shape):
    from ptyrodactyl_amd import serving
    pat, meas = _case(shape, seed=sum(shape))
    print(f"sqrt loss {shape}", flush=True)
    loss, grad = serving.loss_and_grad_intensity(torch.tensor(pat, device="cuda"), torch.tensor(meas, device="cuda"))
    torch.cuda.synchronize()
    eps = np.float32(serving.EPS_SQRT)
    root = np.sqrt(pat.astype(np.float64) + np.float64(eps))
    diff = root - np.sqrt(meas.astype(np.float64))
    ref_loss = float(np.sum(diff ** 2))
    ref_grad = diff / root
    g = grad.cpu().numpy()
    assert g.dtype == np.float32 and g.shape == pat.shape
    # float32 arithmetic against the float64 formula: a few ulp of the operands
    assert np.max(np.abs(g - ref_grad)) <= 1e-6 * max(1.0, float(np.max(np.abs(ref_grad))))
    assert abs(float(loss) - ref_loss) <= 1e-6 * ref_loss + 1e-12


def test_sqrt_loss_is_bitwise_reproducible_and_differentiable():
    from ptyrodactyl_amd import serving
    pat, meas = _case((32, 64, 64), seed=5)
    p = torch.tensor(pat, device="cuda")
    m = torch.tensor(meas, device="cuda")
    l1, g1 = serving.loss_and_grad_intensity(p, m)
    l2, g2 = serving.loss_and_grad_intensity(p, m)
    assert torch.equal(l1, l2) and torch.equal(g1, g2)
    leaf = p.clone().requires_grad_(True)
    loss, grad = serving.loss_and_grad_intensity(leaf, m)
    (3.0 * loss).backward()
    assert torch.equal(loss.detach(), l1)
    assert torch.allclose(leaf.grad, 3.0 * g1, rtol=1e-6, atol=0)
