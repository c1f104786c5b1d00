"""ctypes binding of liboracle.so, the FP64 CPU restatement of the registration hot path.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() (as the checker) and
bench.py's cpu_baseline leg. Arrays are float64 numpy arrays in the reference's layout
(node (x, y) at x + y*N; a displacement field is shape (2, N, N), x component first).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


SSD, NCC, XSHIFT = 0, 1, 2  # data terms: deformationModel 22 / deformationModel 0 / seismic line shift


class LevelParams(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("max_iter", ctypes.c_int), ("stop_eps", ctypes.c_double),
                ("data_term", ctypes.c_int), ("y_reg", ctypes.c_double)]


class SeriesParams(ctypes.Structure):
    _fields_ = [
        ("start_level", ctypes.c_int), ("stop_level", ctypes.c_int),
        ("refine_start_level", ctypes.c_int), ("refine_stop_level", ctypes.c_int),
        ("precision_level", ctypes.c_int),
        ("lambda_", ctypes.c_double), ("lambda_factor", ctypes.c_double),
        ("max_iter", ctypes.c_int), ("stop_eps", ctypes.c_double),
        ("saturation_pct", ctypes.c_double), ("normalize_min_to_zero", ctypes.c_int),
        ("dont_normalize", ctypes.c_int), ("resample_prolong", ctypes.c_int),
        ("dont_accumulate", ctypes.c_int), ("data_term", ctypes.c_int),
        ("reduce_deformations", ctypes.c_int), ("reduce_max_iter", ctypes.c_int), ("reduce_stop_eps", ctypes.c_double),
        ("pair0_init", ctypes.POINTER(ctypes.c_double)), ("corr_init", ctypes.c_int), ("max_corr_shift", ctypes.c_int),
        ("calc_inverse", ctypes.c_int), ("inverse_out", ctypes.POINTER(ctypes.c_double)),
        ("alt_start_level", ctypes.c_int), ("energy_as_alt_criterion", ctypes.c_int),
        ("first_try_out", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle lib)")
        L = ctypes.CDLL(_LIB_PATH)
        L.msor_interpolate.restype = ctypes.c_double
        L.msor_ssd_energy.restype = ctypes.c_double
        L.msor_reg_energy.restype = ctypes.c_double
        L.msor_ncc_normalize.restype = ctypes.c_double
        L.msor_ncc_energy.restype = ctypes.c_double
        L.msor_consistency_energy.restype = ctypes.c_double
        L.msor_level_solve.restype = ctypes.c_int
        L.msor_match_series.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def restrict(fine):
    fine = _f64(fine)
    N = fine.shape[-1]
    out = np.empty((N // 2, N // 2))
    lib().msor_restrict(_p(fine), N, _p(out))
    return out


def prolong(coarse):
    coarse = _f64(coarse)
    n = coarse.shape[-1]
    out = np.empty((2 * n, 2 * n))
    lib().msor_prolong(_p(coarse), n, _p(out))
    return out


def resample(src, N):
    src = _f64(src)
    out = np.empty((N, N))
    lib().msor_resample(_p(src), src.shape[-1], _p(out), N)
    return out


def deform_image(img, phi, ext=np.inf, nearest=False):
    img, phi = _f64(img), _f64(phi)
    out = np.empty_like(img)
    lib().msor_deform_image(_p(img), _p(phi), img.shape[-1], ctypes.c_double(ext), int(nearest), _p(out))
    return out


def fast_deform_image(img, phi):
    img, phi = _f64(img), _f64(phi)
    out = np.empty_like(img)
    lib().msor_fast_deform_image(_p(img), _p(phi), img.shape[-1], _p(out))
    return out


def ssd_energy(R, T, phi):
    R, T, phi = _f64(R), _f64(T), _f64(phi)
    return lib().msor_ssd_energy(_p(R), _p(T), _p(phi), R.shape[-1])


def ssd_force(R, T, phi):
    R, T, phi = _f64(R), _f64(T), _f64(phi)
    out = np.empty_like(phi)
    lib().msor_ssd_force(_p(R), _p(T), _p(phi), R.shape[-1], _p(out))
    return out


def stiff_apply(u):
    u = _f64(u)
    out = np.empty_like(u)
    lib().msor_stiff_apply(_p(u), u.shape[-1], _p(out))
    return out


def reg_energy(phi):
    phi = _f64(phi)
    return lib().msor_reg_energy(_p(phi), phi.shape[-1])


def mass_apply(u):
    u = _f64(u)
    out = np.empty_like(u)
    lib().msor_mass_apply(_p(u), u.shape[-1], _p(out))
    return out


def ncc_normalize(img):
    """(normalised image, sqrt(u.Mu)) as normalizeImageForNCC computes them."""
    img = _f64(img)
    out = np.empty_like(img)
    var = lib().msor_ncc_normalize(_p(img), img.shape[-1], _p(out))
    return out, var


def ncc_energy(R, T, phi):
    R, T, phi = _f64(R), _f64(T), _f64(phi)
    return lib().msor_ncc_energy(_p(R), _p(T), _p(phi), R.shape[-1])


def ncc_force(R, T, phi):
    R, T, phi = _f64(R), _f64(T), _f64(phi)
    out = np.empty_like(phi)
    lib().msor_ncc_force(_p(R), _p(T), _p(phi), R.shape[-1], _p(out))
    return out


def heat_apply(u, tau):
    u = _f64(u)
    out = np.empty_like(u)
    lib().msor_heat_apply(_p(u), u.shape[-1], ctypes.c_double(tau), _p(out))
    return out


def vcycle(d, sigma):
    d = _f64(d)
    out = np.empty_like(d)
    lib().msor_vcycle(_p(d), d.shape[-1], ctypes.c_double(sigma), _p(out))
    return out


def level_solve(R, T, phi0, lam, max_iter, stop_eps, trace_cap=None, data_term=SSD, y_reg=0.0):
    """Returns (phi, info) with info: iterations, final energy, energy trace per iteration and
    tau / sigma / evaluations per Armijo call."""
    R, T = _f64(R), _f64(T)
    phi = _f64(phi0).copy()
    N = R.shape[-1]
    cap = trace_cap or (2 * max_iter + 2)
    et = np.zeros(cap)
    tt = np.zeros(cap)
    st = np.zeros(cap)
    ev = np.zeros(cap, dtype=np.int32)
    calls = ctypes.c_int(0)
    fe = ctypes.c_double(0)
    p = LevelParams(lam, max_iter, stop_eps, data_term, y_reg)
    it = lib().msor_level_solve(_p(R), _p(T), _p(phi), N, ctypes.byref(p), _p(et), _p(tt), _p(st),
                                ev.ctypes.data_as(_ip), cap, ctypes.byref(calls), ctypes.byref(fe))
    nc = calls.value
    return phi, dict(iterations=it, energy=fe.value, energies=et[: it + 1].copy(), taus=tt[:nc].copy(),
                     sigmas=st[:nc].copy(), evals=ev[:nc].copy())


def xshift_energy(R, T, phix, y_reg):
    """(E_data, E_reg) of the seismic x-shift registration (the mean row NCC and 1/2 phix.A phix)."""
    R, T, phix = _f64(R), _f64(T), _f64(phix)
    ed, er = ctypes.c_double(), ctypes.c_double()
    lib().msor_xshift_energy(_p(R), _p(T), _p(phix), R.shape[-1], ctypes.c_double(y_reg), ctypes.byref(ed),
                             ctypes.byref(er))
    return ed.value, er.value


def xshift_gradient(R, T, phix, y_reg, lam):
    R, T, phix = _f64(R), _f64(T), _f64(phix)
    out = np.zeros_like(phix)
    lib().msor_xshift_gradient(_p(R), _p(T), _p(phix), R.shape[-1], ctypes.c_double(y_reg), ctypes.c_double(lam),
                               _p(out))
    return out


def normalize(img, pct=-1.0, min_to_zero=True):
    out = _f64(img).copy()
    lib().msor_normalize(_p(out), out.size, ctypes.c_double(pct), int(min_to_zero))
    return out


def compose(phi1, phi2, start=None):
    phi1, phi2 = _f64(phi1), _f64(phi2)
    out = np.zeros_like(phi1) if start is None else _f64(start).copy()
    lib().msor_compose(_p(phi1), _p(phi2), phi1.shape[-1], _p(out))
    return out


def series_stats(stack, fill):
    stack = _f64(stack)
    F = stack.shape[0]
    NN = stack[0].size
    mean = np.empty(stack.shape[1:])
    cnt = np.empty(stack.shape[1:])
    med = np.empty(stack.shape[1:])
    lib().msor_series_stats(_p(stack), F, NN, ctypes.c_double(fill), _p(mean), _p(cnt), _p(med))
    return mean, cnt, med


def consistency_energy(psi, phi):
    psi, phi = _f64(psi), _f64(phi)
    return lib().msor_consistency_energy(_p(psi), _p(phi), psi.shape[-1])


def consistency_force(psi, phi):
    psi, phi = _f64(psi), _f64(phi)
    out = np.empty_like(phi)
    lib().msor_consistency_force(_p(psi), _p(phi), psi.shape[-1], _p(out))
    return out


def reduce_deformations(psis, max_iter=1000, stop_eps=5e-7):
    """(phi, energy, iterations) of SeriesMatching::reduceDeformations over psis (n, 2, N, N)."""
    psis = _f64(psis)
    n, N = psis.shape[0], psis.shape[-1]
    phi = np.zeros((2, N, N))
    e = ctypes.c_double(0)
    it = lib().msor_reduce_deformations(_p(psis), n, N, max_iter, ctypes.c_double(stop_eps), _p(phi), ctypes.byref(e))
    return phi, e.value, it


def phase_correlation(a, b, normalise=True):
    """Convolution::phaseCorrelation (quoc/convolution.h:1049-1114): real inverse FFT of
    FFT(a) * conj(FFT(b)) (unit modulus per frequency when normalise) / N^2, shape (N, N)."""
    a, b = _f64(a), _f64(b)
    out = np.zeros_like(a)
    lib().msor_phase_correlation(_p(a), _p(b), a.shape[-1], int(bool(normalise)), _p(out))
    return out


def register_translation(reference, templ, max_shift=0):
    """PhaseCorrelationRegistration::registerImages (quoc/paramReg.h:2057-2096): (bx, by) in
    pixels with reference(. + b) ~ templ."""
    a, b = _f64(reference), _f64(templ)
    sh = (ctypes.c_int * 2)()
    lib().msor_register_translation(_p(a), _p(b), a.shape[-1], int(max_shift), sh)
    return int(sh[0]), int(sh[1])


def resample_rect(src, N):
    """ScalarArray<QC_2D>::resampleFrom of an (h, w) image onto (N, N)."""
    src = _f64(src)
    out = np.zeros((N, N))
    lib().msor_resample_rect(_p(src), src.shape[1], src.shape[0], _p(out), N)
    return out


def approx_inverse(phi):
    """qc::approxInverseDeformation (quoc/deformations.h:341-355) of a (2, N, N) displacement."""
    phi = _f64(phi)
    out = np.zeros_like(phi)
    lib().msor_approx_inverse(_p(phi), phi.shape[-1], _p(out))
    return out


def match_series(templates, reference, params, pairwise_in=None, pair0_init=None):
    """One stage of series matching; returns dict of deformations, pairwise deformations,
    energies, mean, count, median (and the reduced deformation with reduce_deformations)."""
    T = _f64(templates)
    n, N = T.shape[0], T.shape[-1]
    ref = _f64(reference)
    defs = np.zeros((n, 2, N, N))
    pw = np.zeros((n, 2, N, N))
    en = np.zeros(n)
    mean = np.zeros((N, N))
    cnt = np.zeros((N, N))
    med = np.zeros((N, N))
    pin = None if pairwise_in is None else _p(_f64(pairwise_in))
    init = None if pair0_init is None else _f64(pair0_init).copy()
    params.pair0_init = _p(init) if init is not None else None
    red = np.zeros((2, N, N))
    inv = np.zeros((n, 2, N, N)) if params.calc_inverse and not params.dont_accumulate else None
    params.inverse_out = _p(inv) if inv is not None else None
    first = np.zeros((n, 2, N, N))
    params.first_try_out = _p(first)
    lib().msor_match_series(_p(T), n, _p(ref), N, ctypes.byref(params), pin, _p(defs), _p(pw), _p(en),
                            _p(mean), _p(cnt), _p(med), _p(red))
    params.pair0_init = None
    params.inverse_out = None
    params.first_try_out = None
    return dict(deformations=defs, pairwise=pw, energies=en, mean=mean, count=cnt, median=med,
                reduced=red if params.reduce_deformations else None, inverses=inv, first_try=first)
