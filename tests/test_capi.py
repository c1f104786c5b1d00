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
