"""Seeded synthetic STEM series: Gaussian atomic columns on a square lattice, Poisson noise, and a
per-frame scan distortion (row-wise random-walk shift plus low-frequency sine modes).

Frame f of workload k uses seed (20261202 + k) * 1000 + f. Images are float64, row-major
(node (x, y) at x + y*N), matching the reference's image layout.
"""
import numpy as np

BASE_SEED = 20261202


def frame_seed(workload, frame):
    return (BASE_SEED + workload) * 1000 + frame


def _lattice_profile(t, spacing, sigma, offset):
    """sum_k exp(-(t - offset - k*spacing)^2 / (2 sigma^2)) over the columns near t."""
    k0 = np.floor((t - offset) / spacing)
    g = np.zeros_like(t)
    for dk in range(-3, 4):
        c = offset + (k0 + dk) * spacing
        g += np.exp(-((t - c) ** 2) / (2.0 * sigma * sigma))
    return g


def distortion(N, rng, jitter_step, max_shift=4.0, modes=4, mode_amp=2.0):
    """Pixel displacement (ux, uy) of one frame: a horizontal per-row random walk (step
    N(0, jitter_step^2) px/row, clipped to +-max_shift) plus `modes` sine modes of amplitude
    mode_amp px with wave numbers 1-2 per axis and random phase."""
    steps = rng.normal(0.0, jitter_step, size=N)
    walk = np.clip(np.cumsum(steps), -max_shift, max_shift)
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    ux = np.repeat(walk[:, None], N, axis=1)
    uy = np.zeros((N, N))
    for _ in range(modes):
        kx, ky = rng.integers(1, 3, size=2)
        phase = rng.uniform(0, 2 * np.pi)
        comp = rng.integers(0, 2)
        s = mode_amp / modes * 2 * np.sin(2 * np.pi * (kx * x + ky * y) / N + phase)
        if comp == 0:
            ux += s
        else:
            uy += s
    return ux, uy


def stem_frame(N, seed, spacing=12.0, sigma=2.5, jitter_step=0.3, counts=200.0, background=0.1,
               amplitude=1.0, noise=True):
    rng = np.random.default_rng(seed)
    ux, uy = distortion(N, rng, jitter_step)
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    # the lattice origin is shared by every frame of a series (the object does not move)
    off = 0.5 * spacing
    img = background + amplitude * _lattice_profile(x + ux, spacing, sigma, off) * _lattice_profile(y + uy, spacing, sigma, off)
    if noise:
        img = rng.poisson(counts * img) / counts
    return img.astype(np.float64)


def stem_series(N, frames, workload=0, **kw):
    """(frames, N, N) float64 series; frame f uses frame_seed(workload, f)."""
    return np.stack([stem_frame(N, frame_seed(workload, f), **kw) for f in range(frames)])


# The two image workloads of the benchmark suite.
CPU_SERIES = dict(N=256, frames=8, workload=0, spacing=12.0, sigma=2.5, jitter_step=0.3)
GPU_SERIES = dict(N=1024, frames=32, workload=1, spacing=16.0, sigma=3.0, jitter_step=0.4)
