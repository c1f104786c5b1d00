"""msg_energy_noalloc (the energy pass with its scratch in a fixed device array, no allocation)
against msg_energy through the C ABI: bit-identical per-pair energies on two calls in a row, the
second with a smaller input (the array keeps the first call's contents), in FP32 and FP64; sizes
beyond the array are refused with a message and leave the next call intact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(B, n, dtype, seed):
    import torch
    g = np.random.default_rng(seed)
    R = g.random((B, n, n))
    T = np.roll(R, 1, axis=2) + 0.05 * g.standard_normal((B, n, n))
    phi = 3.0 / (n - 1) * g.standard_normal((B, 2, n, n))
    dev = torch.device("cuda:0")
    return [torch.tensor(a, dtype=dtype, device=dev) for a in (R, T, phi)]


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_noalloc_energy_matches_msg_energy(precision):
    import torch
    from msamd import gpu
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    dt = torch.float32 if precision == "f32" else torch.float64
    for (B, n, seed) in [(6, 256, 1), (3, 64, 2)]:  # second call smaller than the first
        print(f"noalloc {precision}: batch {B} n {n} seed {seed}", flush=True)
        R, T, phi = _inputs(B, n, dt, seed)
        torch.cuda.synchronize()
        es, er = gpu.energy(R, T, phi)
        ns, nr = gpu.energy_noalloc(R, T, phi)
        assert np.all(np.isfinite(es)) and np.all(es > 0)
        np.testing.assert_array_equal(ns, es)
        np.testing.assert_array_equal(nr, er)


def test_noalloc_energy_refuses_oversized_batches():
    import ctypes
    import torch
    from msamd import gpu
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    n = 2048
    R = torch.zeros((1, n, n), dtype=torch.float32, device="cuda:0")
    phi = torch.zeros((1, 2, n, n), dtype=torch.float32, device="cuda:0")
    es = np.zeros(256)
    # 256 pairs of 2048^2 exceed the 32 MiB array (their partial sums alone are 64 MiB); nothing is
    # launched (the pointers are only checked for NULL, and the size check comes first)
    rc = gpu.lib().msg_energy_noalloc(gpu.F32, gpu._ptr(R), gpu._ptr(R), gpu._ptr(phi), n, 256,
                                      es.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None)
    assert rc == 1
    assert b"scratch" in gpu.lib().msg_last_error()
    R2, T2, phi2 = _inputs(2, 128, torch.float32, 3)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(gpu.energy_noalloc(R2, T2, phi2)[0], gpu.energy(R2, T2, phi2)[0])
