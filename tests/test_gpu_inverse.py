"""Approximate inverse deformation (msg_approx_inverse) and calcInverseDeformation on the GPU against
the oracle, which reproduces the reference's qc::approxInverseDeformation bit for bit and its
program's <i>-r-inv/ solves (tests/test_oracle_golden.py)."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from msamd import gpu as G
    G.lib()
    return G


@pytest.mark.parametrize("N", [8, 32, 64])
def test_approx_inverse_golden(gpu, N):
    with np.load(os.path.join(GOLD, "inverse.npz")) as z:
        for k in range(3):
            phi = torch.tensor(z[f"phi{N}_{k}"], dtype=torch.float64, device="cuda")
            got = gpu.approx_inverse(phi).cpu().numpy()
            assert np.abs(got - z[f"inv{N}_{k}"]).max() * (N - 1) < 1e-12, k


@pytest.mark.parametrize("N,amp", [(256, 3.0), (1024, 6.0)])
def test_approx_inverse_large(gpu, N, amp):
    rng = np.random.default_rng(N)
    h = 1.0 / (N - 1)
    y, x = np.mgrid[0:N, 0:N] * h
    phi = np.stack([amp * h * np.sin(6 * y + 1) + 0.5 * h * rng.standard_normal((N, N)),
                    amp * h * np.cos(5 * x) - 3 * h])
    ref = O.approx_inverse(phi)
    got = gpu.approx_inverse(torch.tensor(phi, dtype=torch.float64, device="cuda")).cpu().numpy()
    assert np.abs(got - ref).max() * (N - 1) < 1e-12
    g32 = gpu.approx_inverse(torch.tensor(phi, dtype=torch.float32, device="cuda")).cpu().numpy()
    # FP32 fields: the corners move by float-rounded displacements; a node near a triangle edge may
    # take the neighbouring triangle's value, within one interpolation step of the deformation
    assert np.median(np.abs(g32 - ref)) * (N - 1) < 1e-5


def test_match_series_inverse(gpu):
    from msamd import synth
    N = 64
    frames = synth.stem_series(N, 4, workload=0, spacing=6.0, sigma=1.5, jitter_step=0.15)
    kw = dict(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
              lambda_=10.0, lambda_factor=-1, max_iter=15, stop_eps=-1.0, saturation_pct=0.15,
              normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1, dont_accumulate=0, data_term=0,
              calc_inverse=1)
    ref = O.match_series(frames[1:], frames[0], O.SeriesParams(**kw))
    t = torch.tensor(frames, dtype=torch.float64, device="cuda")
    gp = gpu.SeriesParams(spec=0, **kw)
    for mode in (2, 0):  # captured graphs (the swap must not meet stale graphs), then host rounds
        gpu.set_default_mode(mode)
        out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gp, want_levels=True)
        assert np.abs(out["deformations"].cpu().numpy() - ref["deformations"]).max() * (N - 1) < 1e-6
        assert np.abs(out["inverses"].cpu().numpy() - ref["inverses"]).max() * (N - 1) < 1e-6
        kinds = [lv["kind"] for lv in out["levels"]]
        assert kinds.count("inverse") == 3 * 2
        last = [lv for lv in out["levels"] if lv["kind"] == "inverse" and lv["level"] == 6]
        for lv in last:
            assert np.abs(lv["field"] - ref["inverses"][lv["frame"]]).max() * (N - 1) < 1e-6
