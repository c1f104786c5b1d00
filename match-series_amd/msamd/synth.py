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


def ricker(t, f):
    """Ricker wavelet of peak frequency f (Hz) at times t (s)."""
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def seismic_gather(seed, samples=2048, traces=512, dt=0.5e-3, interfaces=60, f=40.0, vmin=1500.0, vmax=2500.0,
                   max_offset=1000.0, max_static=3.0):
    """One common-midpoint gather as an image: row j is trace j (offset j * max_offset / (traces - 1)),
    column i time sample i (dt seconds apart). Reflectivity: `interfaces` random layer times with
    coefficients uniform in +-0.2, each moved out hyperbolically, t(x) = sqrt(t0^2 + x^2 / v^2) with
    one velocity v uniform in [vmin, vmax] per gather, plus a random static shift per trace (uniform
    in +-max_static samples), convolved with a Ricker wavelet of peak frequency f. Values in float64,
    offset by 1 so that no row is empty."""
    rng = np.random.default_rng(seed)
    t0 = np.sort(rng.uniform(0.05, 0.95, interfaces)) * samples * dt
    r = rng.uniform(-0.2, 0.2, interfaces)
    v = rng.uniform(vmin, vmax)
    x = np.linspace(0.0, max_offset, traces)
    static = rng.uniform(-max_static, max_static, traces) * dt
    tx = np.sqrt(t0[None, :] ** 2 + (x[:, None] / v) ** 2) + static[:, None]  # (traces, interfaces)
    # the wavelet is below 1e-6 of its peak beyond 1.2 / f: evaluate it on that window only
    half = int(np.ceil(1.2 / f / dt))
    c = np.rint(tx / dt).astype(np.int64)  # nearest sample of each arrival
    idx = c[:, :, None] + np.arange(-half, half + 1)[None, None, :]  # (traces, interfaces, window)
    ok = (idx >= 0) & (idx < samples)
    val = r[None, :, None] * ricker(idx * dt - tx[:, :, None], f) * ok
    flat = np.arange(traces)[:, None, None] * samples + idx
    img = np.bincount(flat[ok], weights=val[ok], minlength=traces * samples).reshape(traces, samples)
    return img + 1.0


def seismic_series(count, workload=4, **kw):
    return np.stack([seismic_gather(frame_seed(workload, g), **kw) for g in range(count)])


SEISMIC_SERIES = dict(count=24, samples=2048, traces=512, workload=4)
