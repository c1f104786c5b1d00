"""The warm-start network's GPU kernels (BF16 MFMA convolutions, pooling, fused upsampling and
concatenation, head, loss, backward pass, Adam) against the CPU FP64 restatement (oracle/unet_ref.py,
BF16 rounding at the same points): forward outputs and the gradient of every parameter tensor.

Activations are rounded to BF16 after every layer, so a different accumulation order flips some of
them by one BF16 unit (2^-8 relative) and the flips propagate. The noise floor of the comparison is
therefore the difference between two valid CPU evaluations, FP64 and FP32 accumulation of the same
BF16 operands; the GPU (FP32 MFMA accumulation) must stay within 4 times that floor, and within 1e-3
(forward) and 1e-2 (gradients) where the floor is lower."""
import numpy as np
import pytest

import unet_ref as U

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from msamd import gpu as G
    G.lib()
    return G


def _inputs(B, H, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((B, 2, H, H)).astype(np.float32)
    t = (2.0 * rng.standard_normal((B, 2, H, H))).astype(np.float32)
    return x, t


@pytest.mark.parametrize("scales,base,B,H", [(2, 32, 2, 64), (3, 64, 1, 128), (4, 32, 1, 256)])
def test_unet_forward_and_gradient_match_cpu(gpu, scales, base, B, H):
    print(f"unet scales {scales} base {base} batch {B} {H}^2", flush=True)
    net = gpu.UNet(scales, base, seed=11)
    p = net.get_params().astype(np.float64)
    x, t = _inputs(B, H, scales)
    xd = torch.tensor(x, device="cuda")
    td = torch.tensor(t, device="cuda")
    out = net.forward(xd).cpu().numpy().astype(np.float64)
    loss, g = net.gradient(xd, td, want_grad=True)
    net.close()
    ref_out, _ = U.forward(p, x.astype(np.float64), scales, base)
    alt_out, _ = U.forward(p, x.astype(np.float64), scales, base, acc32=True)
    floor = np.linalg.norm(alt_out - ref_out) / np.linalg.norm(ref_out)
    rl = np.linalg.norm(out - ref_out) / np.linalg.norm(ref_out)
    print(f"forward relative L2 difference {rl:.2e} (FP32 vs FP64 accumulation on the CPU: {floor:.2e})", flush=True)
    assert rl < max(1e-3, 4 * floor)
    ref_loss, ref_g, _ = U.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), scales, base)
    _, alt_g, _ = U.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), scales, base, acc32=True)
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    acts, convs, pools, head, npar = U.graph(scales, base)
    assert g.size == npar
    worst = 0.0
    for k, c in enumerate(convs):
        cin = c["cin1"] + c["cin2"]
        for nm, a, n in (("w", c["woff"], 9 * c["cout"] * cin), ("b", c["boff"], c["cout"])):
            ref = ref_g[a:a + n]
            e = np.linalg.norm(g[a:a + n] - ref) / max(np.linalg.norm(ref), 1e-30)
            fl = np.linalg.norm(alt_g[a:a + n] - ref) / max(np.linalg.norm(ref), 1e-30)
            worst = max(worst, e)
            assert e < max(1e-2, 4 * fl), (k, nm, e, fl)
    for a, n in ((head["woff"], 2 * base), (head["boff"], 2)):
        e = np.linalg.norm(g[a:a + n] - ref_g[a:a + n]) / np.linalg.norm(ref_g[a:a + n])
        fl = np.linalg.norm(alt_g[a:a + n] - ref_g[a:a + n]) / np.linalg.norm(ref_g[a:a + n])
        assert e < max(1e-2, 4 * fl), ("head", e, fl)
    print(f"largest per-tensor relative L2 gradient difference {worst:.2e}", flush=True)


def test_unet_training_lowers_the_loss(gpu):
    """Adam on one fixed batch: a smooth displacement predicted from a pair lowers the loss."""
    H = 128
    rng = np.random.default_rng(3)
    x = rng.random((2, 2, H, H)).astype(np.float32)
    yy, xx = np.mgrid[0:H, 0:H] / H
    t = np.stack([np.stack([2 * np.sin(2 * np.pi * yy + k), np.cos(2 * np.pi * xx - k)]) for k in range(2)])
    xd = torch.tensor(x, device="cuda")
    td = torch.tensor(t.astype(np.float32), device="cuda")
    net = gpu.UNet(3, 32, seed=5)
    losses = [net.train_step(xd, td, 2e-3) for _ in range(150)]
    net.close()
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert np.isfinite(losses).all()
    assert np.mean(losses[-10:]) < 0.75 * np.mean(losses[:10])
