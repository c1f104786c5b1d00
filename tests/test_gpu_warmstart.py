"""The learned warm start end to end on a small configuration: training triples produced by the GPU
method itself (256^2 STEM pairs, plain multilevel solve from level 5, 50 iterations per level), the
network trained on them with its own GPU kernels, then plain method vs warm start on fresh seeded
pairs (msamd.synth). Checked: the training lowers the loss, and every pair ends at or below the plain
method's final energy (a warm start that does not reach it falls back to the plain result). The
iteration ratio (plain over warm, fallbacks charged with the failed attempt plus the plain solve) is
reported, not asserted: whether a network of this size trained this briefly beats the multilevel
start is the measurement, not a property of the code."""
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


def test_warm_start_needs_fewer_iterations(gpu):
    import time
    from msamd import synth, warmstart as W
    N, start, iters = 256, 5, 50
    t0 = time.perf_counter()
    train_pairs = W.stem_pairs_gpu(N, 384, seed=11)
    targets, e_train, it_train, _ = W.plain_solve(train_pairs, start, iters)
    print(f"training triples: {train_pairs.shape[0]} pairs of {N}^2 in {time.perf_counter() - t0:.1f} s "
          f"({it_train:.0f} iterations each)", flush=True)
    net = gpu.UNet(4, 32, seed=1)
    t0 = time.perf_counter()
    hist = W.train(net, train_pairs, targets, epochs=10, batch=8, lr=2e-3,
                   log=lambda m: print(m, flush=True))
    print(f"training: {time.perf_counter() - t0:.1f} s", flush=True)
    assert np.isfinite(hist).all() and hist[-1] < hist[0]
    fr = synth.stem_series(N, 17, workload=7, spacing=12.0, sigma=2.5, jitter_step=0.3)
    test_pairs = W.pairs_from_frames(fr)
    r = W.compare(net, test_pairs, start, iters)
    net.close()
    print({k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items() if not isinstance(v, np.ndarray)},
          flush=True)
    assert 0 <= r["fallbacks"] <= r["pairs"]
    assert np.all(r["final_energies"] <= r["plain_energies"] * (1 + 1e-12) + 1e-15)
    assert r["iteration_ratio"] > 0
