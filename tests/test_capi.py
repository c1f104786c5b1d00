"""The C ABI library loads without a GPU and exports every function include/ms_gpu.h declares; the
ctypes binding declares exactly those functions. No compute calls here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ms_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(msg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("msg_ctx_create", "msg_solve", "msg_solve_level", "msg_energy", "msg_gradient", "msg_h1_smooth",
                 "msg_warp", "msg_compose", "msg_series_stats", "msg_match_series", "msg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from msamd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libmsgpu.so not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(gpu.SIGNATURES) == declared()


def test_error_reporting_without_device_work():
    from msamd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libmsgpu.so not built")
    L = gpu.lib()
    # argument validation fails before any HIP call
    assert L.msg_energy(7, None, None, None, 64, 1, None, None, None) == 1
    assert b"dtype" in L.msg_last_error()
    assert L.msg_h1_smooth(0, None, None, 48, 1, 1.0, None) == 1
    assert b"power of two" in L.msg_last_error()
    assert L.msg_version().startswith(b"msgpu")


def test_wrappers_reject_mismatched_tensors_before_the_library():
    """The ctypes wrappers check dtype, shape, contiguity and device of every tensor argument with
    explicit raises (the library trusts the sizes it is given). CPU tensors reach the device check
    last, so every mismatch below is reported without a GPU and without a library call."""
    torch = pytest.importorskip("torch")
    from msamd import gpu
    n = 8
    R = torch.zeros((2, n, n), dtype=torch.float64)
    T = torch.zeros((2, n, n), dtype=torch.float64)
    phi = torch.zeros((2, 2, n, n), dtype=torch.float64)
    with pytest.raises(TypeError, match="phi: dtype"):
        gpu.energy(R, T, phi.float())
    with pytest.raises(TypeError, match="T: dtype"):
        gpu.gradient(R, T.float(), phi, 1.0)
    with pytest.raises(ValueError, match="phi: shape"):
        gpu.energy(R, T, phi[:1].contiguous())
    with pytest.raises(ValueError, match="T: shape"):
        gpu.ncc_energy(R, T[:, :4, :4].contiguous(), phi)
    with pytest.raises(ValueError, match="contiguous"):
        gpu.energy(R, T.transpose(1, 2), phi)
    with pytest.raises(ValueError, match="device tensor"):
        gpu.energy(R, T, phi)
    with pytest.raises(TypeError, match="float32 or float64"):
        gpu.h1_smooth(torch.zeros((1, n, n), dtype=torch.float16), 1.0)
    with pytest.raises(ValueError, match="phi: shape"):
        gpu.warp(R, torch.zeros((3, 2, n, n), dtype=torch.float64))
    with pytest.raises(TypeError, match="phi2: dtype"):
        gpu.compose(phi[0], phi[1].float())
    with pytest.raises(ValueError, match="templ: shape"):
        gpu.register_translation(R[0], torch.zeros((n, 2 * n), dtype=torch.float64))
    with pytest.raises(ValueError, match="reference: shape"):
        gpu.match_series(R, torch.zeros((n + 1, n + 1), dtype=torch.float64), gpu.SeriesParams())
    with pytest.raises(ValueError, match="expected shape"):
        gpu.series_stats(torch.zeros((3, n, n + 1), dtype=torch.float64), 0.0)
