"""The oracle's x-shift registration (matchSeismicSeries' findXShift) against the reference's own
operators and gradient flow (tests/golden/seismic.npz, written by tests/golden/make_seismic.py from
oracle/_ref/ref_harness), and against closed forms that the fixtures leave open: identical rows give
NCC -1, the stiffness form of fields linear in x or y, and the variation as the energy's derivative.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def g():
    with np.load(os.path.join(GOLD, "seismic.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("N", [8, 32, 64])
def test_xshift_energy_and_variation_golden(g, N):
    R, T, phix = g[f"xs{N}_R"], g[f"xs{N}_T"], g[f"xs{N}_phix"]
    lam, yreg = g[f"xs{N}_lam_yreg"]
    ed, er = O.xshift_energy(R, T, phix, yreg)
    assert rel(ed, g[f"xs{N}_e_ncc"]) < 1e-13
    assert rel(er, g[f"xs{N}_e_reg"]) < 1e-13
    assert rel(O.xshift_gradient(R, T, phix, yreg, lam), g[f"xs{N}_DE"]) < 1e-12


def test_xshift_level_solve_golden(g):
    R, T = g["solve_R"], g["solve_T"]
    lam, yreg, iters = g["solve_params"]
    N = R.shape[0]
    phi, info = O.level_solve(R, T, np.zeros((2, N, N)), lam, int(iters), 0.0, data_term=O.XSHIFT, y_reg=yreg)
    assert info["iterations"] == int(g["solve_iterations"])
    np.testing.assert_array_equal(info["taus"], g["solve_taus"])
    np.testing.assert_allclose(info["energy"], g["solve_energy"], rtol=1e-12)
    assert np.abs(phi[0] - g["solve_phix"]).max() * (N - 1) < 1e-9
    assert np.abs(phi[1]).max() == 0  # the y component never moves


def test_xshift_closed_forms():
    N = 33
    h = 1.0 / (N - 1)
    rng = np.random.default_rng(0)
    R = rng.uniform(0.2, 1.0, (N, N))
    # identical rows and no shift: every row's NCC is -1, so their mean is -1
    ed, er = O.xshift_energy(R, R, np.zeros((N, N)), 0.7)
    assert abs(ed + 1) < 1e-13 and er == 0
    # stiffness form: phi = a x gives 1/2 a^2 (K_xx), phi = a y gives 1/2 y a^2 (K_yy) on the unit square
    y, x = np.meshgrid(np.arange(N) * h, np.arange(N) * h, indexing="ij")
    a, yreg = 0.3, 0.7
    assert abs(O.xshift_energy(R, R, a * x, yreg)[1] - 0.5 * a * a) < 1e-13
    assert abs(O.xshift_energy(R, R, a * y, yreg)[1] - 0.5 * yreg * a * a) < 1e-13
    # a constant shift carries no regularisation
    assert abs(O.xshift_energy(R, R, np.full((N, N), 2 * h), yreg)[1]) < 1e-13


def test_xshift_regulariser_variation_is_the_derivative():
    """The regulariser's variation is exactly the derivative of its quadratic form (central differences
    agree to rounding). The data term's variation integrates the template's derivative at quadrature
    points while its energy warps nodally (FastDeformImage), so it is not the derivative of the
    discrete energy; the reference's own vectors above pin it."""
    with np.load(os.path.join(GOLD, "seismic.npz"), allow_pickle=False) as z:
        R, T, phix = z["xs32_R"], z["xs32_T"], z["xs32_phix"]
    N = R.shape[0]
    lam, yreg = 5.0, 0.7
    dreg = (O.xshift_gradient(R, T, phix, yreg, lam) - O.xshift_gradient(R, T, phix, yreg, 0.0)) / lam
    rng = np.random.default_rng(1)
    eps = 1e-6
    for _ in range(24):
        i, j = rng.integers(0, N, 2)
        p, m = phix.copy(), phix.copy()
        p[i, j] += eps
        m[i, j] -= eps
        rp, rm = O.xshift_energy(R, T, p, yreg)[1], O.xshift_energy(R, T, m, yreg)[1]
        assert abs((rp - rm) / (2 * eps) - dreg[i, j]) < 1e-7 * np.abs(dreg).max()
