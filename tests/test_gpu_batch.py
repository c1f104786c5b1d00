"""msg_match_series_batch: several independent series of different frame sizes and lengths in one
call (pairwise solves batched across series, chains in lockstep per size, size groups concurrent).
Every series' outputs must be bitwise those of msg_match_series on it alone: each pair's arithmetic
does not depend on the batch it is solved in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from msamd import gpu as G
    G.lib()
    return G


def series(N, frames, workload):
    from msamd import synth
    sc = N / 256.0
    return synth.stem_series(N, frames, workload=workload, spacing=max(12 * sc, 4.0), sigma=max(2.5 * sc, 1.0),
                             jitter_step=0.3 * sc)


def params(gpu, N, data_term=0, dont_accumulate=0, its=20):
    L = int(np.log2(N))
    return gpu.SeriesParams(start_level=L - 2, stop_level=L, refine_start_level=L - 1, refine_stop_level=L,
                            precision_level=L, lambda_=10.0, lambda_factor=-1.0, max_iter=its, stop_eps=-1.0,
                            saturation_pct=0.15, normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1,
                            dont_accumulate=dont_accumulate, spec=0, data_term=data_term)


# (frame size, frames, data term (0 SSD, 1 NCC), dont_accumulate)
LAYOUT = [(64, 4, 0, 0), (128, 6, 0, 0), (64, 7, 0, 0), (256, 3, 0, 0), (128, 3, 0, 0), (64, 5, 1, 0),
          (64, 4, 0, 1)]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_equals_single_series(gpu, dtype):
    jobs = []
    for k, (N, nf, dtm, dacc) in enumerate(LAYOUT):
        fr = torch.tensor(series(N, nf + 1, 10 + k), dtype=dtype, device="cuda")
        p = params(gpu, N, gpu.NCC if dtm else gpu.SSD, dacc)
        jobs.append((fr[1:].contiguous(), fr[0].contiguous(), p))
    print(f"batch of {len(jobs)} series, {dtype}", flush=True)
    outs = gpu.match_series_batch(jobs)
    for k, (tpl, ref, p) in enumerate(jobs):
        one = gpu.match_series(tpl, ref, params(gpu, tpl.shape[-1], p.data_term, p.dont_accumulate))
        b = outs[k]
        assert torch.equal(b["deformations"], one["deformations"]), k
        np.testing.assert_array_equal(b["energies"], one["energies"])
        for key in ("mean", "count", "median"):
            assert torch.equal(b[key], one[key]), (k, key)
        assert np.isfinite(b["energies"]).all()


def test_batch_refuses_what_it_does_not_take(gpu):
    fr = torch.tensor(series(64, 3, 1), dtype=torch.float64, device="cuda")
    p = params(gpu, 64)
    p.calc_inverse = 1
    with pytest.raises(RuntimeError, match="calcInverseDeformation"):
        gpu.match_series_batch([(fr[1:].contiguous(), fr[0].contiguous(), p)])
