"""Learned warm start of the gradient flow (no counterpart in the reference).

The network (msg_unet_*, a BF16 U-Net on MFMA) is trained on (template, reference, converged
deformation) triples that the GPU method itself produces, predicts a deformation for a new pair, and
the plain gradient flow refines from it on the finest level. The start falls back to zero when the
warm-started solve ends above the plain method's energy.

Pipeline (all device-resident; torch only allocates and synthesises the training frames):
  * stem_pairs_gpu: (reference, template) pairs of STEM frames with independent scan distortions,
    the model of msamd.synth (lattice of Gaussian columns, Poisson noise, row-wise random walk plus
    sine modes), drawn with torch's generator on the GPU; normalised like the series
    (enhanceContrastSaturationPercentage 0.15, normalizeMinToZero) by msg_normalize;
  * plain_solve: the plain multilevel method (msg_solve from start_level to the finest level, fixed
    iteration count, lambda 10) on a batch of pairs; its deformations (in pixels) are the targets;
  * train: Adam on the mean squared error of the predicted deformation;
  * compare: plain method vs warm start on fresh pairs: iterations and wall clock to reach the plain
    method's final energy, fallback count, and the network's deformation error.
"""
import math
import time

import numpy as np

from . import gpu


def unet_macs(scales, base, B, H, W):
    """Multiply-adds of one forward pass of the network (its 3x3 convolutions; the 1x1 head, pooling
    and the elementwise work are left out): per scale s two convolutions at (H >> s) x (W >> s), the
    encoder's from 16 (the padded input) or base * 2^(s-1) channels to base * 2^s, the decoder's from
    the upsampled deeper tensor plus the skip tensor."""
    total, first = 0, None
    ch = [base << s for s in range(scales)]
    for s in range(scales):
        px = B * (H >> s) * (W >> s)
        cin = 16 if s == 0 else ch[s - 1]
        m1 = px * 9 * cin * ch[s]
        first = m1 if s == 0 else first
        total += m1 + px * 9 * ch[s] * ch[s]
    for s in range(scales - 2, -1, -1):
        px = B * (H >> s) * (W >> s)
        total += px * 9 * (ch[s + 1] + ch[s]) * ch[s] + px * 9 * ch[s] * ch[s]
    return total, first


def _torch():
    import torch
    return torch


def _profile(t, spacing, sigma, offset):
    torch = _torch()
    k0 = torch.floor((t - offset) / spacing)
    g = torch.zeros_like(t)
    for dk in range(-3, 4):
        c = offset + (k0 + dk) * spacing
        g = g + torch.exp(-((t - c) ** 2) / (2.0 * sigma * sigma))
    return g


def stem_frames_gpu(N, count, gen, spacing=12.0, sigma=2.5, jitter_step=0.3, max_shift=4.0, modes=4, mode_amp=2.0,
                    counts=200.0, background=0.1, amplitude=1.0):
    """(count, N, N) float32 frames on the GPU, each with its own scan distortion."""
    torch = _torch()
    dev = "cuda"
    yy, xx = torch.meshgrid(torch.arange(N, device=dev, dtype=torch.float32),
                            torch.arange(N, device=dev, dtype=torch.float32), indexing="ij")
    steps = torch.randn((count, N), generator=gen, device=dev) * jitter_step
    walk = torch.clamp(torch.cumsum(steps, dim=1), -max_shift, max_shift)
    ux = walk[:, :, None].expand(count, N, N).clone()
    uy = torch.zeros((count, N, N), device=dev)
    for _ in range(modes):
        kxy = torch.randint(1, 3, (count, 2), generator=gen, device=dev).float()
        phase = torch.rand((count,), generator=gen, device=dev) * (2 * math.pi)
        comp = torch.randint(0, 2, (count,), generator=gen, device=dev)
        s = mode_amp / modes * 2 * torch.sin(2 * math.pi * (kxy[:, 0, None, None] * xx + kxy[:, 1, None, None] * yy) / N
                                             + phase[:, None, None])
        ux = ux + s * (comp == 0)[:, None, None]
        uy = uy + s * (comp == 1)[:, None, None]
    off = 0.5 * spacing
    img = background + amplitude * _profile(xx + ux, spacing, sigma, off) * _profile(yy + uy, spacing, sigma, off)
    return torch.poisson(counts * img, generator=gen) / counts


def stem_pairs_gpu(N, count, seed, chunk=64, **kw):
    """(count, 2, N, N) float32 normalised (reference, template) pairs on the GPU."""
    torch = _torch()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    out = torch.empty((count, 2, N, N), dtype=torch.float32, device="cuda")
    for i in range(0, count, chunk):
        k = min(chunk, count - i)
        f = stem_frames_gpu(N, 2 * k, gen, **kw).reshape(k, 2, N, N).contiguous()
        gpu.normalize(f, 0.15, True)
        out[i:i + k] = f
    torch.cuda.synchronize()
    return out


def pairs_from_frames(frames_np, reference_index=0):
    """(count, 2, N, N) normalised device pairs (frame 0 as the reference, each other frame as the
    template) from a host series of float64 frames (msamd.synth)."""
    torch = _torch()
    f = torch.tensor(np.asarray(frames_np), dtype=torch.float32, device="cuda").contiguous()
    gpu.normalize(f, 0.15, True)
    ref = f[reference_index]
    tpl = torch.cat([f[:reference_index], f[reference_index + 1:]])
    return torch.stack([ref.expand_as(tpl), tpl], dim=1).contiguous()


def plain_solve(pairs, start_level, iters, lam=10.0, batch=32):
    """The plain multilevel method on every pair: (deformations in pixels (count, 2, N, N) float32,
    final energies, gradient-flow iterations per pair (mean), seconds)."""
    torch = _torch()
    count, _, N, _ = pairs.shape
    L = int(round(math.log2(N)))
    phi = torch.empty((count, 2, N, N), dtype=torch.float32, device="cuda")
    energies = np.zeros(count)
    its = 0
    secs = 0.0
    for i in range(0, count, batch):
        k = min(batch, count - i)
        ctx = gpu.Context(L, k, gpu.F32)
        try:
            ctx.set_mode(2)
            ctx.set_images_device(pairs[i:i + k, 0].contiguous(), pairs[i:i + k, 1].contiguous())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            energies[i:i + k] = ctx.solve(start_level, L, lam, iters, -1.0, resample_prolong=1)
            secs += time.perf_counter() - t0
            its += ctx.counts()[0]
            out = torch.empty((k, 2, N, N), dtype=torch.float32, device="cuda")
            ctx.get_deformations_device(out)
            phi[i:i + k] = out * (N - 1)
        finally:
            ctx.close()
    torch.cuda.synchronize()
    return phi, energies, its / count, secs


def net_input(pairs):
    """Network input channels (template, reference) from (reference, template) pairs."""
    torch = _torch()
    return torch.stack([pairs[:, 1], pairs[:, 0]], dim=1).contiguous()


def train(net, pairs, targets, epochs, batch=8, lr=1e-3, seed=0, log=None):
    """Adam on the mean squared deformation error; returns the mean loss per epoch."""
    torch = _torch()
    x_all = net_input(pairs)
    rng = np.random.default_rng(seed)
    count = pairs.shape[0]
    hist = []
    for ep in range(epochs):
        order = rng.permutation(count)
        tot, nb = 0.0, 0
        for i in range(0, count - batch + 1, batch):
            idx = torch.as_tensor(order[i:i + batch], device="cuda")
            x = x_all.index_select(0, idx).contiguous()
            y = targets.index_select(0, idx).contiguous()
            tot += net.train_step(x, y, lr)
            nb += 1
        hist.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f}")
    return hist


def compare(net, pairs, start_level, iters, lam=10.0, cap=None, batch=32):
    """Plain method vs warm start on `pairs`. The warm start solves the finest level from the network's
    prediction until it reaches the plain method's final energy (at most `cap` iterations); a pair
    that does not, or whose predicted start has a higher energy than zero, falls back to the plain
    result. Returns a dict of per-pair and aggregate figures."""
    torch = _torch()
    count, _, N, _ = pairs.shape
    L = int(round(math.log2(N)))
    cap = cap or 2 * iters * (L - start_level + 1)
    phi_plain, e_plain, it_plain, t_plain = plain_solve(pairs, start_level, iters, lam, batch)
    x = net_input(pairs)
    need = np.full(count, -1)
    e_warm = np.zeros(count)
    e_start = np.zeros(count)
    e_zero = np.zeros(count)
    t_net, t_warm = 0.0, 0.0
    err = []
    for i in range(0, count, batch):
        k = min(batch, count - i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pred = net.forward(x[i:i + k].contiguous())
        torch.cuda.synchronize()
        t_net += time.perf_counter() - t0
        err.append((pred - phi_plain[i:i + k]).abs().amax(dim=(1, 2, 3)).cpu().numpy())
        ctx = gpu.Context(L, k, gpu.F32)
        try:
            ctx.set_mode(2)
            ctx.set_images_device(pairs[i:i + k, 0].contiguous(), pairs[i:i + k, 1].contiguous())
            e_zero[i:i + k] = ctx.solve_level(L, lam, 0, -1.0)[0]  # energy of the zero start
            ctx.set_deformations_device((pred / (N - 1)).contiguous())
            e_start[i:i + k] = ctx.solve_level(L, lam, 0, -1.0)[0]  # energy of the predicted start
            e, it, tr = ctx.solve_level(L, lam, cap, -1.0, trace_cap=cap)
            for j in range(k):
                hit = np.nonzero(tr[j][:, 0] <= e_plain[i + j])[0]
                need[i + j] = hit[0] + 1 if hit.size else -1
                e_warm[i + j] = tr[j][hit[0], 0] if hit.size else e[j]
            # wall clock to the plain energy for the whole batch: the largest count needed
            kmax = int(max([n for n in need[i:i + k] if n > 0], default=0))
            if kmax > 0:
                ctx.set_deformations_device((pred / (N - 1)).contiguous())
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ctx.solve_level(L, lam, kmax, -1.0)
                t_warm += time.perf_counter() - t0
        finally:
            ctx.close()
    ok = need > 0
    it_warm = np.where(ok, need, np.nan)
    final = np.where(ok, e_warm, e_plain)  # the result kept per pair
    # a fallback pair costs the failed attempt plus the plain solve
    it_eff = np.where(ok, need, cap + it_plain)
    # cost in finest-level iterations (a level-l iteration costs 4^(l - L) of one): the plain schedule
    # and the warm start's finest-level iterations (a fallback: the attempt plus the plain schedule)
    plain_fine = float(sum(iters * 4.0 ** (l - L) for l in range(start_level, L + 1)))
    cost_eff = np.where(ok, need, cap + plain_fine)
    return dict(pairs=int(count), size=int(N), start_level=int(start_level), iterations_per_level=int(iters),
                start_energy_below_zero_start=int((e_start < e_zero).sum()),
                plain_iterations=float(it_plain), plain_seconds=t_plain,
                warm_iterations_mean=float(np.nanmean(it_warm)) if ok.any() else None,
                warm_iterations_effective_mean=float(np.mean(it_eff)), fallbacks=int((~ok).sum()),
                iteration_ratio=float(it_plain / np.mean(it_eff)),
                plain_fine_level_equivalent_iterations=plain_fine,
                cost_ratio=float(plain_fine / np.mean(cost_eff)), net_seconds=t_net, warm_seconds=t_warm,
                net_error_px_mean_of_max=float(np.mean(np.concatenate(err))),
                plain_energy_mean=float(np.mean(e_plain)), warm_energy_mean=float(np.mean(e_warm[ok])) if ok.any() else None,
                plain_energies=e_plain, final_energies=final)
