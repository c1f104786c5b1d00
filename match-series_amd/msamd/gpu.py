"""ctypes binding of libmsgpu.so (include/ms_gpu.h) for tests and benchmarks.

Device arrays are torch tensors on a ROCm device (torch is only the allocator here); every call goes
through the library's C ABI. Importing this module loads the library and fails loudly if it is
missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmsgpu.so")

F32, F64 = 0, 1
SSD, NCC, XSHIFT = 0, 1, 3  # data terms: deformationModel 22 / deformationModel 0 / seismic row x-shift

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class SolveParams(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("lambda_factor", ctypes.c_double),
                ("lambda_ref_level", ctypes.c_int), ("max_iter", ctypes.c_int), ("stop_eps", ctypes.c_double),
                ("resample_prolong", ctypes.c_int), ("spec", ctypes.c_int)]


class IterTrace(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("energy", "ereg", "tau", "tau_first", "sigma", "evals", "passes", "beta")]


class SeriesParams(ctypes.Structure):
    _fields_ = [
        ("start_level", ctypes.c_int), ("stop_level", ctypes.c_int),
        ("refine_start_level", ctypes.c_int), ("refine_stop_level", ctypes.c_int),
        ("precision_level", ctypes.c_int),
        ("lambda_", ctypes.c_double), ("lambda_factor", ctypes.c_double),
        ("max_iter", ctypes.c_int), ("stop_eps", ctypes.c_double),
        ("saturation_pct", ctypes.c_double), ("normalize_min_to_zero", ctypes.c_int),
        ("dont_normalize", ctypes.c_int), ("resample_prolong", ctypes.c_int),
        ("dont_accumulate", ctypes.c_int), ("spec", ctypes.c_int), ("pre_smooth_sigma", ctypes.c_double),
        ("pairwise_batch", ctypes.c_int), ("data_term", ctypes.c_int),
        ("average_ext_mean", ctypes.c_int), ("average_nearest", ctypes.c_int),
        ("reduce_deformations", ctypes.c_int), ("reduce_max_iter", ctypes.c_int), ("reduce_stop_eps", ctypes.c_double),
        ("pair0_init", ctypes.c_void_p), ("corr_init", ctypes.c_int), ("max_corr_shift", ctypes.c_int),
        ("calc_inverse", ctypes.c_int), ("inverse_out", ctypes.c_void_p),
        ("alt_start_level", ctypes.c_int), ("energy_as_alt_criterion", ctypes.c_int), ("first_try_out", ctypes.c_void_p),
        ("raw_templates", ctypes.c_void_p), ("y_reg_factor", ctypes.c_double), ("keep_start_transformation", ctypes.c_int),
    ]


class SeriesJob(ctypes.Structure):
    _fields_ = [("params", ctypes.POINTER(SeriesParams)), ("templates", ctypes.c_void_p), ("n", ctypes.c_int),
                ("reference", ctypes.c_void_p), ("deformations", ctypes.c_void_p), ("energies", _dp),
                ("mean", ctypes.c_void_p), ("count", ctypes.c_void_p), ("median", ctypes.c_void_p)]


LEVEL_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, _dp, ctypes.c_void_p,
                                  ctypes.c_int, _dp)


class SeriesLevels(ctypes.Structure):
    _fields_ = [("fields", _dp), ("iters", ctypes.POINTER(IterTrace)), ("start", _dp),
                ("iterations", ctypes.POINTER(ctypes.c_int)), ("on_level", LEVEL_CALLBACK), ("user", ctypes.c_void_p)]


# exported symbols with their argument types (also the list the C-ABI test checks)
SIGNATURES = {
    "msg_last_error": ([], ctypes.c_char_p),
    "msg_version": ([], ctypes.c_char_p),
    "msg_ctx_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "msg_ctx_destroy": ([_vp], ctypes.c_int),
    "msg_set_reference": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "msg_set_template": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "msg_set_images_device": ([_vp, _vp, _vp, ctypes.c_int], ctypes.c_int),
    "msg_set_deformation": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "msg_get_deformation": ([_vp, ctypes.c_int, _dp], ctypes.c_int),
    "msg_set_deformation_zero": ([_vp, ctypes.c_int], ctypes.c_int),
    "msg_set_deformation_level": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "msg_get_deformation_level": ([_vp, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
    "msg_set_deformations_device": ([_vp, _vp], ctypes.c_int),
    "msg_get_deformations_device": ([_vp, _vp], ctypes.c_int),
    "msg_solve_level": ([_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _ip,
                         _vp, ctypes.c_int], ctypes.c_int),
    "msg_solve": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SolveParams), _dp], ctypes.c_int),
    "msg_last_counts": ([_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                         ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "msg_ctx_buffer": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "msg_energy": ([ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _vp], ctypes.c_int),
    "msg_energy_noalloc": ([ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _vp], ctypes.c_int),
    "msg_gradient": ([ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp], ctypes.c_int),
    "msg_h1_smooth": ([ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp], ctypes.c_int),
    "msg_warp": ([ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                  ctypes.c_int, _vp], ctypes.c_int),
    "msg_restrict": ([ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "msg_prolong": ([ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "msg_resample": ([ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "msg_compose": ([ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "msg_series_stats": ([ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp, _vp, _vp],
                         ctypes.c_int),
    "msg_normalize": ([ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp], ctypes.c_int),
    "msg_release_cache": ([], ctypes.c_int),
    "msg_ctx_set_mode": ([_vp, ctypes.c_int], ctypes.c_int),
    "msg_ctx_set_data_term": ([_vp, ctypes.c_int], ctypes.c_int),
    "msg_ctx_set_y_reg_factor": ([_vp, ctypes.c_double], ctypes.c_int),
    "msg_ncc_energy": ([ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _vp], ctypes.c_int),
    "msg_ncc_gradient": ([ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp],
                         ctypes.c_int),
    "msg_xshift_energy": ([ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _vp],
                          ctypes.c_int),
    "msg_xshift_gradient": ([ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            ctypes.c_double, _vp], ctypes.c_int),
    "msg_match_series_batch": ([ctypes.c_int, ctypes.c_int, ctypes.POINTER(SeriesJob)], ctypes.c_int),
    "msg_series_last_counts": ([ctypes.POINTER(ctypes.c_longlong)] * 3 + [_dp, _dp], ctypes.c_int),
    "msg_set_default_mode": ([ctypes.c_int], ctypes.c_int),
    "msg_set_default_exact_decisions": ([ctypes.c_int], ctypes.c_int),
    "msg_profile_enable": ([ctypes.c_uint], ctypes.c_int),
    "msg_profile_reset": ([], ctypes.c_int),
    "msg_profile_read": ([ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_longlong), _dp], ctypes.c_int),
    "msg_time_energy_pass": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "msg_match_series": ([ctypes.POINTER(SeriesParams), ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _dp, _vp,
                          _vp, _vp, ctypes.POINTER(SeriesLevels), _vp], ctypes.c_int),
    "msg_resample_rect": ([ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "msg_approx_inverse": ([ctypes.c_int, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "msg_register_translation": ([ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _ip, _vp, _vp], ctypes.c_int),
    "msg_unet_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong], ctypes.c_int),
    "msg_unet_destroy": ([_vp], ctypes.c_int),
    "msg_unet_param_count": ([_vp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "msg_unet_get_params": ([_vp, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "msg_unet_set_params": ([_vp, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "msg_unet_forward": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "msg_unet_gradient": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_float),
                           _vp], ctypes.c_int),
    "msg_unet_adam_step": ([_vp, ctypes.c_float], ctypes.c_int),
    "msg_reduce_deformations": ([ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _dp,
                                 ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
}

_lib = None


def lib():
    """Load libmsgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError(f"libmsgpu error {rc}: {lib().msg_last_error().decode()}")


def _torch():
    import torch
    return torch


def dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise TypeError("tensors must be float32 or float64")


def _ptr(t):
    if not (t.is_cuda and t.is_contiguous()):
        raise ValueError("expected a contiguous device tensor")
    return ctypes.c_void_p(t.data_ptr())


def _arg(t, name, code=None, shape=None):
    """Checks one tensor argument before it reaches the library (which trusts sizes and types):
    float32/float64, the dtype `code` and `shape` when given, contiguous. Returns its dtype code.
    (The device is checked afterwards for all arguments together, _on_device.)"""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"{name}: tensors must be float32 or float64, got {t.dtype}")
    c = dtype_code(t)
    if code is not None and c != code:
        raise TypeError(f"{name}: dtype {t.dtype} differs from the other arguments' {torch_dtype(code)}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: expected a contiguous tensor")
    return c


def _on_device(**named):
    """After the type and shape checks of _arg: every tensor on one ROCm device."""
    first = None
    for name, t in named.items():
        if t is None:
            continue
        if not t.is_cuda:
            raise ValueError(f"{name}: expected a device tensor, got one on {t.device}")
        if first is not None and t.device != first[1]:
            raise ValueError(f"{name}: on {t.device}, {first[0]} on {first[1]}")
        first = first or (name, t.device)


def _square(t, name, lead):
    """(lead..., n, n) with n >= 2: returns n."""
    if t.dim() != len(lead) + 2 or t.shape[-1] != t.shape[-2] or t.shape[-1] < 2:
        raise ValueError(f"{name}: expected shape {tuple(lead) + ('n', 'n')}, got {tuple(t.shape)}")
    for k, (got, want) in enumerate(zip(t.shape, lead)):
        if want is not None and got != want:
            raise ValueError(f"{name}: dimension {k} is {got}, expected {want}")
    return t.shape[-1]


def _pair_args(R, T, phi):
    """R, T: (B, n, n); phi: (B, 2, n, n); one dtype and device. Returns (code, B, n)."""
    n = _square(R, "R", (None,))
    B = R.shape[0]
    code = _arg(R, "R")
    _arg(T, "T", code, (B, n, n))
    _arg(phi, "phi", code, (B, 2, n, n))
    _on_device(R=R, T=T, phi=phi)
    return code, B, n


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _stream():
    """torch's current stream: the operator entry points run on it, ordered with torch's own work."""
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ready():
    """Context calls and msg_match_series run on the library's own streams: wait for torch's pending
    work on their device inputs (and on freshly allocated outputs) first."""
    _torch().cuda.current_stream().synchronize()


def torch_dtype(code):
    torch = _torch()
    return torch.float64 if code == F64 else torch.float32


# ------------------------------------------------------------------------------------------------
# Operator surface
# ------------------------------------------------------------------------------------------------
def energy(R, T, phi):
    """Per-pair (E_SSD, 1/2 sum_c phi_c K phi_c). R, T: (B, n, n); phi: (B, 2, n, n)."""
    code, B, n = _pair_args(R, T, phi)
    es = np.zeros(B)
    er = np.zeros(B)
    check(lib().msg_energy(code, _ptr(R), _ptr(T), _ptr(phi), n, B, _dptr(es), _dptr(er), _stream()))
    return es, er


def energy_noalloc(R, T, phi):
    """energy() through msg_energy_noalloc: no device allocation, fixed-size scratch array."""
    code, B, n = _pair_args(R, T, phi)
    es = np.zeros(B)
    er = np.zeros(B)
    check(lib().msg_energy_noalloc(code, _ptr(R), _ptr(T), _ptr(phi), n, B, _dptr(es), _dptr(er), _stream()))
    return es, er


def ncc_energy(R, T, phi):
    """Per-pair (E_NCC, 1/2 sum_c phi_c K phi_c) of the deformationModel 0 data term."""
    code, B, n = _pair_args(R, T, phi)
    es = np.zeros(B)
    er = np.zeros(B)
    check(lib().msg_ncc_energy(code, _ptr(R), _ptr(T), _ptr(phi), n, B, _dptr(es), _dptr(er), _stream()))
    return es, er


def ncc_gradient(R, T, phi, lam):
    code, B, n = _pair_args(R, T, phi)
    out = _torch().empty_like(phi)
    check(lib().msg_ncc_gradient(code, _ptr(R), _ptr(T), _ptr(phi), _ptr(out), n, B, lam, _stream()))
    return out


def xshift_energy(R, T, phi, y_reg_factor):
    """Per-pair (row-averaged 1-D NCC, 1/2 phi_x (K_xx + y K_yy) phi_x) of the x-shift registration;
    phi: (B, 2, n, n), its x component read."""
    code, B, n = _pair_args(R, T, phi)
    es = np.zeros(B)
    er = np.zeros(B)
    check(lib().msg_xshift_energy(code, _ptr(R), _ptr(T), _ptr(phi), n, B, float(y_reg_factor), _dptr(es), _dptr(er),
                                  _stream()))
    return es, er


def xshift_gradient(R, T, phi, lam, y_reg_factor):
    code, B, n = _pair_args(R, T, phi)
    out = _torch().empty_like(phi)
    check(lib().msg_xshift_gradient(code, _ptr(R), _ptr(T), _ptr(phi), _ptr(out), n, B, float(lam),
                                    float(y_reg_factor), _stream()))
    return out


def gradient(R, T, phi, lam):
    code, B, n = _pair_args(R, T, phi)
    out = _torch().empty_like(phi)
    check(lib().msg_gradient(code, _ptr(R), _ptr(T), _ptr(phi), _ptr(out), n, B, lam, _stream()))
    return out


def h1_smooth(d, sigma):
    """d: (..., n, n) fields smoothed independently."""
    if d.dim() < 2:
        raise ValueError("d: expected (..., n, n)")
    n = _square(d, "d", (None,) * (d.dim() - 2))
    code = _arg(d, "d")
    _on_device(d=d)
    k = d.numel() // (n * n)
    out = _torch().empty_like(d)
    check(lib().msg_h1_smooth(code, _ptr(d), _ptr(out), n, k, sigma, _stream()))
    return out


def warp(img, phi, mode=0, ext=float("inf"), nearest=False):
    """img: (k, n, n); phi: (k, 2, n, n) or (2, n, n) shared. mode 0: constant outside, 1: clamp, 2: NaN."""
    if img.dim() < 2:
        raise ValueError("img: expected (..., n, n)")
    n = _square(img, "img", (None,) * (img.dim() - 2))
    code = _arg(img, "img")
    k = img.numel() // (n * n)
    shared = 1 if phi.dim() == 3 else 0
    _arg(phi, "phi", code, (2, n, n) if shared else (k, 2, n, n))
    _on_device(img=img, phi=phi)
    out = _torch().empty_like(img)
    check(lib().msg_warp(code, _ptr(img), _ptr(phi), _ptr(out), n, k, shared, mode, ext, int(nearest), _stream()))
    return out


def restrict(fine):
    if fine.dim() < 2:
        raise ValueError("fine: expected (..., n, n)")
    n = _square(fine, "fine", (None,) * (fine.dim() - 2))
    code = _arg(fine, "fine")
    _on_device(fine=fine)
    k = fine.numel() // (n * n)
    out = _torch().empty(fine.shape[:-2] + (n // 2, n // 2), dtype=fine.dtype, device=fine.device)
    check(lib().msg_restrict(code, _ptr(fine), _ptr(out), n, k, _stream()))
    return out


def prolong(coarse):
    if coarse.dim() < 2:
        raise ValueError("coarse: expected (..., n, n)")
    nc = _square(coarse, "coarse", (None,) * (coarse.dim() - 2))
    code = _arg(coarse, "coarse")
    _on_device(coarse=coarse)
    k = coarse.numel() // (nc * nc)
    out = _torch().empty(coarse.shape[:-2] + (2 * nc, 2 * nc), dtype=coarse.dtype, device=coarse.device)
    check(lib().msg_prolong(code, _ptr(coarse), _ptr(out), nc, k, _stream()))
    return out


def resample(src, n):
    if src.dim() < 2:
        raise ValueError("src: expected (..., n, n)")
    nc = _square(src, "src", (None,) * (src.dim() - 2))
    code = _arg(src, "src")
    _on_device(src=src)
    k = src.numel() // (nc * nc)
    out = _torch().empty(src.shape[:-2] + (n, n), dtype=src.dtype, device=src.device)
    check(lib().msg_resample(code, _ptr(src), _ptr(out), nc, n, k, _stream()))
    return out


def compose(phi1, phi2, start=None):
    n = _square(phi1, "phi1", (2,))
    code = _arg(phi1, "phi1")
    _arg(phi2, "phi2", code, (2, n, n))
    if start is not None:
        _arg(start, "start", code, (2, n, n))
    _on_device(phi1=phi1, phi2=phi2, start=start)
    out = _torch().zeros_like(phi1) if start is None else start.clone()
    check(lib().msg_compose(code, _ptr(phi1), _ptr(phi2), _ptr(out), phi1.shape[-1], _stream()))
    return out


def series_stats(stack, fill):
    n = _square(stack, "stack", (None,))
    code = _arg(stack, "stack")
    _on_device(stack=stack)
    k = stack.shape[0]
    torch = _torch()
    mean = torch.empty((n, n), dtype=stack.dtype, device=stack.device)
    cnt = torch.empty_like(mean)
    med = torch.empty_like(mean)
    check(lib().msg_series_stats(code, _ptr(stack), k, n, fill, _ptr(mean), _ptr(cnt), _ptr(med), _stream()))
    return mean, cnt, med


def normalize(img, pct=-1.0, min_to_zero=True):
    """In place on (k, n, n) images."""
    if img.dim() < 2:
        raise ValueError("img: expected (..., n, n)")
    n = _square(img, "img", (None,) * (img.dim() - 2))
    code = _arg(img, "img")
    _on_device(img=img)
    k = img.numel() // (n * n)
    check(lib().msg_normalize(code, _ptr(img), n, k, pct, int(min_to_zero), _stream()))
    return img


def register_translation(reference, templ, max_shift=0, want_corr=False):
    """Integer translation (bx, by) in pixels with reference(. + b) ~ templ, by phase correlation
    (max_shift > 0: plain correlation within |b|_inf <= max_shift); with want_corr also the
    correlation surface as an (n, n) float64 device tensor."""
    n = _square(reference, "reference", ())
    code = _arg(reference, "reference")
    _arg(templ, "templ", code, (n, n))
    _on_device(reference=reference, templ=templ)
    sh = (ctypes.c_int * 2)()
    corr = _torch().empty((n, n), dtype=_torch().float64, device=reference.device) if want_corr else None
    check(lib().msg_register_translation(code, _ptr(reference), _ptr(templ), n, int(max_shift), sh,
                                         _ptr(corr) if want_corr else None, _stream()))
    b = (int(sh[0]), int(sh[1]))
    return (b, corr) if want_corr else b


# ------------------------------------------------------------------------------------------------
# Registration context
# ------------------------------------------------------------------------------------------------
class Context:
    """A batch of registration pairs with image pyramids up to `level` (msg_ctx)."""

    def __init__(self, level, npairs, dtype=F32):
        self.level, self.npairs, self.dtype = level, npairs, dtype
        self.n = 1 << level
        h = ctypes.c_void_p()
        check(lib().msg_ctx_create(ctypes.byref(h), level, npairs, dtype))
        self.h = h

    def set_data_term(self, data_term):
        check(lib().msg_ctx_set_data_term(self.h, data_term))

    def set_mode(self, mode):
        check(lib().msg_ctx_set_mode(self.h, mode))

    def set_y_reg_factor(self, y):
        check(lib().msg_ctx_set_y_reg_factor(self.h, y))

    def close(self):
        if self.h:
            lib().msg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reference(self, pair, img):
        a = np.ascontiguousarray(img, dtype=np.float64)
        check(lib().msg_set_reference(self.h, pair, _dptr(a)))

    def set_template(self, pair, img):
        a = np.ascontiguousarray(img, dtype=np.float64)
        check(lib().msg_set_template(self.h, pair, _dptr(a)))

    def set_images_device(self, ref, tpl):
        shape = (self.npairs, self.n, self.n)
        if ref is None and tpl is None:
            raise ValueError("set_images_device: no image given")
        code = _arg(ref, "ref", None, shape) if ref is not None else _arg(tpl, "tpl", None, shape)
        if ref is not None and tpl is not None:
            _arg(tpl, "tpl", code, shape)
        _on_device(ref=ref, tpl=tpl)
        _ready()
        check(lib().msg_set_images_device(self.h, _ptr(ref) if ref is not None else None,
                                          _ptr(tpl) if tpl is not None else None,
                                          dtype_code(ref if ref is not None else tpl)))

    def set_deformation(self, pair, phi):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        check(lib().msg_set_deformation(self.h, pair, _dptr(a)))

    def get_deformation(self, pair):
        out = np.zeros((2, self.n, self.n))
        check(lib().msg_get_deformation(self.h, pair, _dptr(out)))
        return out

    def set_deformation_zero(self, pair):
        check(lib().msg_set_deformation_zero(self.h, pair))

    def set_deformations_device(self, phi):
        _arg(phi, "phi", self.dtype, (self.npairs, 2, self.n, self.n))
        _on_device(phi=phi)
        _ready()
        check(lib().msg_set_deformations_device(self.h, _ptr(phi)))

    def get_deformations_device(self, phi):
        _arg(phi, "phi", self.dtype, (self.npairs, 2, self.n, self.n))
        _on_device(phi=phi)
        _ready()
        check(lib().msg_get_deformations_device(self.h, _ptr(phi)))

    def solve_level(self, level, lam, max_iter, stop_eps, spec=0, trace_cap=0):
        e = np.zeros(self.npairs)
        it = np.zeros(self.npairs, dtype=np.int32)
        tr = None
        if trace_cap:
            tr = (IterTrace * (trace_cap * self.npairs))()
        check(lib().msg_solve_level(self.h, level, lam, max_iter, stop_eps, spec, _dptr(e), it.ctypes.data_as(_ip),
                                    ctypes.cast(tr, _vp) if tr is not None else None, trace_cap))
        traces = None
        if tr is not None:
            raw = np.ctypeslib.as_array(ctypes.cast(tr, _dp), shape=(self.npairs, trace_cap, 8)).copy()
            traces = [raw[p, : it[p]] for p in range(self.npairs)]
        return e, it, traces

    def solve(self, start, stop, lam, max_iter, stop_eps, lambda_factor=-1.0, lambda_ref_level=None,
              resample_prolong=0, spec=0):
        p = SolveParams(lam, lambda_factor, stop if lambda_ref_level is None else lambda_ref_level, max_iter, stop_eps,
                        resample_prolong, spec)
        e = np.zeros(self.npairs)
        check(lib().msg_solve(self.h, start, stop, ctypes.byref(p), _dptr(e)))
        return e

    def counts(self):
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().msg_last_counts(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def time_energy_pass(self, level, ncand=3, reps=50):
        """Mean time (ms) and algorithmic bytes of one energy pass over all pairs of the context, with
        `ncand` step sizes along the direction of the last solve at this level."""
        ms, b = ctypes.c_double(), ctypes.c_double()
        check(lib().msg_time_energy_pass(self.h, level, ncand, reps, ctypes.byref(ms), ctypes.byref(b)))
        return ms.value, b.value


def series_level_layout(params, n, reuse=False):
    """(solve kind, frame index, level) of every per-level record of match_series, in order."""
    npw = 1 if reuse else n
    recs = [("pairwise", i, l) for i in range(npw) for l in range(params.start_level, params.stop_level + 1)]
    if not params.dont_accumulate:
        recs += [("refine", i, l) for i in range(1, n)
                 for l in range(params.refine_start_level, params.refine_stop_level + 1)]
        if params.calc_inverse:
            recs += [("inverse", i, l) for i in range(n)
                     for l in range(params.refine_start_level, params.refine_stop_level + 1)]
    if params.alt_start_level > 0 and params.alt_start_level != params.start_level:
        recs += [("alt", i, l) for i in range(npw) for l in range(params.alt_start_level, params.stop_level + 1)]
    return recs


def reduce_deformations(defs, max_iter=1000, stop_eps=5e-7):
    """reduceDeformations on device fields defs (n, 2, N, N): (phi, energy, iterations)."""
    N = _square(defs, "defs", (None, 2))
    code = _arg(defs, "defs")
    _on_device(defs=defs)
    if N & (N - 1):
        raise ValueError("defs: the grid side must be a power of two")
    n = defs.shape[0]
    phi = _torch().zeros_like(defs[0])
    e = np.zeros(1)
    it = ctypes.c_int(0)
    _ready()
    check(lib().msg_reduce_deformations(code, _ptr(defs), n, int(np.log2(N)), max_iter, stop_eps, _ptr(phi),
                                        _dptr(e), ctypes.byref(it)))
    return phi, float(e[0]), it.value


def resample_rect(src, n):
    """ScalarArray<QC_2D>::resampleFrom of an (h, w) device image onto (n, n)."""
    if src.dim() != 2:
        raise ValueError(f"src: expected (h, w), got {tuple(src.shape)}")
    code = _arg(src, "src")
    _on_device(src=src)
    h, w = src.shape[-2], src.shape[-1]
    out = _torch().empty((n, n), dtype=src.dtype, device=src.device)
    check(lib().msg_resample_rect(code, _ptr(src), w, h, _ptr(out), n, _stream()))
    return out


def approx_inverse(phi):
    """qc::approxInverseDeformation of a (2, n, n) device displacement field."""
    _square(phi, "phi", (2,))
    code = _arg(phi, "phi")
    _on_device(phi=phi)
    out = _torch().empty_like(phi)
    check(lib().msg_approx_inverse(code, _ptr(phi), _ptr(out), phi.shape[-1], _stream()))
    return out


def match_series(templates, reference, params, pairwise_in=None, want_pairwise=False, want_levels=False):
    """One stage of series matching on device tensors. templates: (n, N, N); reference: (N, N).
    want_levels: also return the per-level deformations and iteration records of every solve; with
    params.calc_inverse the inverse solves' deformations come back as "inverses"."""
    torch = _torch()
    N = _square(templates, "templates", (None,))
    code = _arg(templates, "templates")
    n = templates.shape[0]
    _arg(reference, "reference", code, (N, N))
    if pairwise_in is not None:
        _arg(pairwise_in, "pairwise_in", code, (n, 2, N, N))
    _on_device(templates=templates, reference=reference, pairwise_in=pairwise_in)
    lv = None
    if want_levels:
        recs = series_level_layout(params, n, pairwise_in is not None)
        nf = sum(2 << (2 * l) for _, _, l in recs)
        mi = max(params.max_iter, 0)
        fields = np.zeros(nf)
        iters = (IterTrace * max(len(recs) * mi, 1))()
        start = np.zeros(2 * len(recs))
        its = np.zeros(len(recs), dtype=np.int32)
        lv = SeriesLevels(_dptr(fields), iters, _dptr(start), its.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    defs = torch.zeros((n, 2, N, N), dtype=templates.dtype, device=templates.device)
    pw = torch.zeros_like(defs) if want_pairwise else None
    en = np.zeros(n)
    mean = torch.zeros((N, N), dtype=templates.dtype, device=templates.device)
    cnt = torch.zeros_like(mean)
    med = torch.zeros_like(mean)
    red = torch.zeros((2, N, N), dtype=templates.dtype, device=templates.device) if params.reduce_deformations else None
    inv = torch.zeros_like(defs) if params.calc_inverse and not params.dont_accumulate else None
    first = torch.zeros_like(defs) if params.alt_start_level > 0 else None
    # every output's zero fill is on torch's stream: finished before the library's streams write
    _ready()
    params.inverse_out = ctypes.c_void_p(inv.data_ptr()) if inv is not None else None
    params.first_try_out = ctypes.c_void_p(first.data_ptr()) if first is not None else None
    check(lib().msg_match_series(ctypes.byref(params), code, _ptr(templates), n, _ptr(reference),
                                 _ptr(pairwise_in) if pairwise_in is not None else None, _ptr(defs),
                                 _ptr(pw) if pw is not None else None, _dptr(en), _ptr(mean), _ptr(cnt), _ptr(med),
                                 ctypes.byref(lv) if lv is not None else None, _ptr(red) if red is not None else None))
    params.inverse_out = None
    params.first_try_out = None
    out = dict(deformations=defs, pairwise=pw, energies=en, mean=mean, count=cnt, median=med, reduced=red,
               inverses=inv, first_try=first)
    if lv is not None:
        levels, off = [], 0
        tr = np.ctypeslib.as_array(iters).view(np.float64).reshape(-1, 8) if mi > 0 else np.zeros((0, 8))
        for r, (kind, i, l) in enumerate(recs):
            m = 1 << l
            levels.append(dict(kind=kind, frame=i, level=l, field=fields[off:off + 2 * m * m].reshape(2, m, m),
                               start=(start[2 * r], start[2 * r + 1]), iterations=int(its[r]),
                               trace=tr[r * mi:r * mi + int(its[r])]))
            off += 2 * m * m
        out["levels"] = levels
    return out


def match_series_batch(series):
    """Several series in one call (msg_match_series_batch). series: list of (templates (n, N, N),
    reference (N, N), SeriesParams), device tensors of one dtype. Returns one dict per series as
    match_series does (deformations, energies, mean, count, median)."""
    torch = _torch()
    code = None
    jobs = (SeriesJob * len(series))()
    outs, keep = [], []
    for k, (templates, reference, params) in enumerate(series):
        N = _square(templates, "templates", (None,))
        c = _arg(templates, "templates")
        code = c if code is None else code
        if c != code:
            raise ValueError("all series must have the same dtype")
        _arg(reference, "reference", code, (N, N))
        _on_device(templates=templates, reference=reference)
        n = templates.shape[0]
        o = dict(deformations=torch.zeros((n, 2, N, N), dtype=templates.dtype, device=templates.device),
                 energies=np.zeros(n), mean=torch.zeros((N, N), dtype=templates.dtype, device=templates.device))
        o["count"] = torch.zeros_like(o["mean"])
        o["median"] = torch.zeros_like(o["mean"])
        outs.append(o)
        keep.append(params)
        jobs[k] = SeriesJob(ctypes.pointer(params), _ptr(templates), n, _ptr(reference), _ptr(o["deformations"]),
                            _dptr(o["energies"]), _ptr(o["mean"]), _ptr(o["count"]), _ptr(o["median"]))
    _ready()
    check(lib().msg_match_series_batch(code, len(series), jobs))
    return outs


def series_last_counts():
    """Counts of the last match_series call: dict of iterations, evals, passes (totals) and
    node_iterations, node_evals (weighted by each level's node count)."""
    v = [ctypes.c_longlong() for _ in range(3)]
    w = np.zeros(2)
    check(lib().msg_series_last_counts(*[ctypes.byref(x) for x in v], w[0:1].ctypes.data_as(_dp),
                                       w[1:2].ctypes.data_as(_dp)))
    return dict(iterations=v[0].value, evals=v[1].value, passes=v[2].value, node_iterations=float(w[0]),
                node_evals=float(w[1]))


# ------------------------------------------------------------------------------------------------
# Kernel timing
# ------------------------------------------------------------------------------------------------
KERNEL_CLASSES = ("energy", "variation", "h1_vcycle", "cg_direction", "update", "control", "composition")


def profile_enable(classes=("energy",)):
    mask = 0
    for c in classes:
        mask |= 1 << KERNEL_CLASSES.index(c)
    check(lib().msg_profile_enable(mask))


def profile_reset():
    check(lib().msg_profile_reset())


def profile_read(cls):
    ms, n, b = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
    check(lib().msg_profile_read(KERNEL_CLASSES.index(cls), ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
    return dict(ms=ms.value, launches=n.value, bytes=b.value)


def set_default_mode(mode):
    check(lib().msg_set_default_mode(mode))


def release_cache():
    """Frees the contexts msg_match_series keeps for reuse."""
    check(lib().msg_release_cache())


# ------------------------------------------------------------------------------------------------
# Learned warm start (msg_unet_*)
# ------------------------------------------------------------------------------------------------
class UNet:
    """The warm-start network on the GPU: forward, gradient and Adam steps through the C ABI.
    Inputs (B, 2, H, W) float32 device tensors (template, reference); outputs and targets
    (B, 2, H, W) float32 displacements in pixels."""

    def __init__(self, scales=4, base=32, seed=0):
        self.h = ctypes.c_void_p()
        check(lib().msg_unet_create(ctypes.byref(self.h), scales, base, seed))
        self.scales, self.base = scales, base
        n = ctypes.c_longlong()
        check(lib().msg_unet_param_count(self.h, ctypes.byref(n)))
        self.nparams = n.value

    def close(self):
        if self.h:
            check(lib().msg_unet_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        a = np.zeros(self.nparams, dtype=np.float32)
        check(lib().msg_unet_get_params(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {a.size}")
        check(lib().msg_unet_set_params(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def _io(self, x, name):
        torch = _torch()
        if not isinstance(x, torch.Tensor) or x.dtype != torch.float32 or x.dim() != 4 or x.shape[1] != 2:
            raise ValueError(f"{name}: expected a float32 tensor of shape (B, 2, H, W)")
        if not (x.is_cuda and x.is_contiguous()):
            raise ValueError(f"{name}: expected a contiguous device tensor")
        return x.shape[0], x.shape[2], x.shape[3]

    def forward(self, x):
        B, H, W = self._io(x, "x")
        out = _torch().empty_like(x)
        check(lib().msg_unet_forward(self.h, _ptr(x), _ptr(out), B, H, W, _stream()))
        return out

    def gradient(self, x, target, want_grad=False):
        B, H, W = self._io(x, "x")
        if tuple(target.shape) != tuple(x.shape):
            raise ValueError("target: shape must match x")
        self._io(target, "target")
        loss = ctypes.c_double()
        g = np.zeros(self.nparams, dtype=np.float32) if want_grad else None
        check(lib().msg_unet_gradient(self.h, _ptr(x), _ptr(target), B, H, W, ctypes.byref(loss),
                                      g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if want_grad else None,
                                      _stream()))
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr):
        check(lib().msg_unet_adam_step(self.h, float(lr)))

    def train_step(self, x, target, lr):
        loss = self.gradient(x, target)
        self.adam_step(lr)
        return loss

