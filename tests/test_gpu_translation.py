"""Phase-correlation translation (msg_register_translation) and the series option
useCorrelationToInitTranslation on the GPU, against the oracle's restatement
(PhaseCorrelationRegistration::registerImages, quoc/paramReg.h:2057-2096; Convolution::
phaseCorrelation, quoc/convolution.h:1049-1114; matchSeries.h:145-153). The oracle itself is
pinned on the CPU against numpy's FFT and brute-force cross-correlation
(tests/test_oracle_properties.py)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from msamd import gpu as G
    G.lib()
    return G


def shifted_pair(N, bx, by, seed):
    """Two N x N crops of one STEM-like field (not periodic), the second moved by (bx, by) pixels:
    a(x + b) = b_img(x) wherever both are defined."""
    from msamd import synth
    big = synth.stem_frame(N + 64, synth.frame_seed(seed, 0), spacing=max(12 * N / 256, 4.0),
                           sigma=max(2.5 * N / 256, 1.0), jitter_step=0.2)
    a = big[32:32 + N, 32:32 + N]
    b = big[32 + by:32 + by + N, 32 + bx:32 + bx + N]
    return np.ascontiguousarray(a), np.ascontiguousarray(b)


@pytest.mark.parametrize("N,bx,by", [(64, 3, -5), (256, -17, 9), (1024, 25, 31), (32, 0, 0)])
def test_phase_correlation_matches_oracle(gpu, N, bx, by):
    a, b = shifted_pair(N, bx, by, N)
    ta = torch.tensor(a, dtype=torch.float64, device="cuda")
    tb = torch.tensor(b, dtype=torch.float64, device="cuda")
    got, corr = gpu.register_translation(ta, tb, want_corr=True)
    ref_corr = O.phase_correlation(a, b, True)
    assert np.abs(corr.cpu().numpy() - ref_corr).max() < 1e-12
    assert got == O.register_translation(a, b) == (bx, by)


@pytest.mark.parametrize("N,bx,by,m", [(64, 3, -5, 6), (128, 20, -3, 4), (64, 1, 1, 1)])
def test_windowed_correlation_matches_oracle(gpu, N, bx, by, m):
    a, b = shifted_pair(N, bx, by, 7)
    ta = torch.tensor(a, dtype=torch.float64, device="cuda")
    tb = torch.tensor(b, dtype=torch.float64, device="cuda")
    got, corr = gpu.register_translation(ta, tb, max_shift=m, want_corr=True)
    ref_corr = O.phase_correlation(a, b, False)
    assert np.abs(corr.cpu().numpy() - ref_corr).max() < 1e-9 * np.abs(ref_corr).max()
    assert got == O.register_translation(a, b, m)
    if max(abs(bx), abs(by)) <= m:
        assert got == (bx, by)


@pytest.mark.parametrize("N,m", [(64, 63), (64, 46), (8, 7), (16, 15)])
def test_window_wider_than_half_the_frame(gpu, N, m):
    """maxCorrShift up to n - 1: the (2m+1)^2 window wraps around the periodic correlation surface
    and holds up to (2n-1)^2 values, more than any n x n work buffer; the scan order (x fastest from
    (-m, -m), strict maximum, start at (0, 0)) decides between the repeated values as the oracle does."""
    for seed, (bx, by) in enumerate(((3, -5), (N // 2 - 1, 2), (-N // 2 + 1, N // 2 - 1))):
        if N < 16:
            rng = np.random.default_rng(seed)
            a = rng.random((N, N))
            b = np.roll(a, (-by, -bx), axis=(0, 1))
        else:
            a, b = shifted_pair(N, bx % 8, by % 8, 11 + seed)
        ta = torch.tensor(a, dtype=torch.float64, device="cuda")
        tb = torch.tensor(b, dtype=torch.float64, device="cuda")
        assert gpu.register_translation(ta, tb, max_shift=m) == O.register_translation(a, b, m)


def test_float_frames_and_periodic_shifts(gpu):
    rng = np.random.default_rng(5)
    N = 128
    a = rng.random((N, N))
    for bx, by in ((0, 0), (5, -7), (N // 2, -N // 2 + 1), (-1, N // 2)):
        b = np.roll(a, (-by, -bx), axis=(0, 1))  # a(x + b) = b_img(x), periodically
        for dt in (torch.float32, torch.float64):
            got = gpu.register_translation(torch.tensor(a, dtype=dt, device="cuda"),
                                           torch.tensor(b, dtype=dt, device="cuda"))
            assert got == (bx, by), (bx, by, dt)


def test_rejects_sizes_beyond_the_transform(gpu):
    t = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="maxCorrShift"):
        gpu.register_translation(t, t, max_shift=8)


def series_params(mod, **kw):
    base = dict(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
                lambda_=10.0, lambda_factor=-1, max_iter=15, stop_eps=-1.0, saturation_pct=0.15,
                normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1, dont_accumulate=0, data_term=0,
                corr_init=1)
    base.update(kw)
    if mod is not O:
        base["spec"] = 0
    return mod.SeriesParams(**base)


@pytest.mark.parametrize("max_shift", [0, 8])
def test_match_series_with_translation_init(gpu, max_shift):
    """Frames drifting by several pixels per frame: every pairwise solve starts from the translation
    its frames' phase correlation gives, as in the oracle's (and the reference's) series loop."""
    from msamd import synth
    N = 64
    big = synth.stem_frame(N + 64, synth.frame_seed(3, 0), spacing=6.0, sigma=1.5, jitter_step=0.1)
    frames = np.stack([big[32 + 4 * f:32 + 4 * f + N, 32 - 5 * f:32 - 5 * f + N] for f in range(4)])
    p = series_params(O, max_corr_shift=max_shift)
    ref = O.match_series(frames[1:], frames[0], p)
    t = torch.tensor(frames, dtype=torch.float64, device="cuda")
    out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), series_params(gpu, max_corr_shift=max_shift),
                           want_pairwise=True)
    d = out["deformations"].cpu().numpy()
    assert np.abs(d - ref["deformations"]).max() * (N - 1) < 1e-6
    np.testing.assert_allclose(out["energies"], ref["energies"], rtol=1e-9)
    np.testing.assert_allclose(out["mean"].cpu().numpy(), ref["mean"], rtol=1e-9, atol=1e-12)
    # phase correlation finds the drift: frame f is frame f-1 seen (-5, 4) px further on, so each
    # pairwise deformation (T(x + phi) ~ R) is about (5, -4) px. (The plain correlation of the
    # windowed search peaks near the origin on these frames, as the reference's would.)
    if max_shift == 0:
        pw = out["pairwise"].cpu().numpy() * (N - 1)
        assert np.abs(np.median(pw[:, 0]) - 5) < 0.5 and np.abs(np.median(pw[:, 1]) + 4) < 0.5
