"""The oracle against what the mathematics fixes, independently of the reference's code: closed
forms, invariants, finite differences and brute force. Each check is chosen so that a plausible
slip (a dropped quadrature weight, a transposed index, a sign, a wrong boundary mirror) fails it.
"""
import numpy as np
import pytest

import oracle as O


def grid(N):
    y, x = np.mgrid[0:N, 0:N].astype(float)
    return x, y, 1.0 / (N - 1)


@pytest.mark.parametrize("N", [4, 16, 64])
def test_ssd_constant_offset_closed_form(N):
    # T = R + c, phi = 0: E = 1/2 c^2 |[0,1]^2| and the variation is c * grad T weighted by the basis
    rng = np.random.default_rng(N)
    R = rng.random((N, N))
    c = 0.3
    assert abs(O.ssd_energy(R, R + c, np.zeros((2, N, N))) - 0.5 * c * c) < 1e-14
    # constant images: no force
    assert np.abs(O.ssd_force(np.zeros((N, N)), np.full((N, N), c), np.zeros((2, N, N)))).max() < 1e-15


def test_ssd_force_is_the_derivative():
    # central differences of E along random directions (FirstDerivativeValidator's check)
    N = 32
    x, y, h = grid(N)
    R = np.sin(2 * np.pi * x / N) * np.cos(2 * np.pi * y / N)
    T = np.sin(2 * np.pi * (x + 0.7) / N) * np.cos(2 * np.pi * (y - 0.4) / N)
    rng = np.random.default_rng(0)
    phi = 0.3 * h * rng.standard_normal((2, N, N))
    F = O.ssd_force(R, T, phi)
    for k in range(3):
        v = rng.standard_normal((2, N, N)) * h
        eps = 1e-6
        fd = (O.ssd_energy(R, T, phi + eps * v) - O.ssd_energy(R, T, phi - eps * v)) / (2 * eps)
        assert abs(fd - np.sum(F * v)) <= 1e-5 * abs(fd)


@pytest.mark.parametrize("N", [2, 8, 32])
def test_mass_closed_forms(N):
    x, y, h = grid(N)
    # interior stencil h^2/36 [1 4 1; 4 16 4; 1 4 1], total mass 1, exact second moment of x
    if N >= 4:
        u = np.zeros((N, N))
        u[N // 2, N // 2] = 1.0
        blk = O.mass_apply(u)[N // 2 - 1:N // 2 + 2, N // 2 - 1:N // 2 + 2]
        np.testing.assert_allclose(blk, h * h / 36 * np.array([[1, 4, 1], [4, 16, 4], [1, 4, 1]]), rtol=1e-14)
    assert abs(O.mass_apply(np.ones((N, N))).sum() - 1.0) < 1e-14
    assert abs(np.sum(O.mass_apply(x * h) * (x * h)) - 1.0 / 3.0) < 1e-14
    rng = np.random.default_rng(N)
    a, b = rng.random((N, N)), rng.random((N, N))
    assert abs(np.sum(O.mass_apply(a) * b) - np.sum(a * O.mass_apply(b))) < 1e-14


@pytest.mark.parametrize("N", [8, 64])
def test_ncc_closed_forms(N):
    # phi = 0: T an increasing affine function of R gives -1, a decreasing one +1; the normalised
    # reference has unit M-norm and zero M-mean
    rng = np.random.default_rng(N + 1)
    R = rng.random((N, N))
    z = np.zeros((2, N, N))
    assert abs(O.ncc_energy(R, 3.0 * R + 2.0, z) + 1.0) < 1e-12
    assert abs(O.ncc_energy(R, 0.5 - 2.0 * R, z) - 1.0) < 1e-12
    T = rng.random((N, N))
    assert abs(O.ncc_energy(R, 7.0 * T - 1.0, z) - O.ncc_energy(R, T, z)) < 1e-12
    rn, var = O.ncc_normalize(R)
    assert abs(np.sum(O.mass_apply(rn) * rn) - 1.0) < 1e-12 and abs(O.mass_apply(rn).sum()) < 1e-12
    # a constant template has no variance: no division, energy 0
    assert abs(O.ncc_energy(R, np.full((N, N), 0.7), z)) < 1e-12


def test_ncc_force_along_a_translation():
    # the variation is the continuous one (quadrature-point gradients), the energy warps nodally:
    # they agree to O(h) along smooth directions, and a wrong sign or scale fails this
    N = 64
    x, y, h = grid(N)
    R = np.sin(4 * np.pi * x / N) * np.cos(4 * np.pi * y / N) + 0.5
    T = np.sin(4 * np.pi * (x + 0.7) / N) * np.cos(4 * np.pi * (y - 0.4) / N) + 0.5
    phi = 0.3 * h * np.random.default_rng(0).standard_normal((2, N, N))
    F = O.ncc_force(R, T, phi)
    for c in range(2):
        v = np.zeros((2, N, N))
        v[c] = 0.5 * h
        fd = (O.ncc_energy(R, T, phi + 1e-6 * v) - O.ncc_energy(R, T, phi - 1e-6 * v)) / 2e-6
        assert abs(fd - np.sum(F * v)) <= 0.05 * abs(fd)


def test_consistency_closed_forms_and_derivative():
    # psi = 0, phi = c: 1/2 |c|^2 over the unit square; psi = -c, phi = c: 0
    N = 16
    c = np.array([0.013, -0.021])
    phi = np.ones((2, N, N)) * c[:, None, None]
    assert abs(O.consistency_energy(np.zeros((2, N, N)), phi) - 0.5 * np.sum(c * c)) < 1e-15
    assert abs(O.consistency_energy(-phi, phi)) < 1e-30
    # the variation is the derivative of the quadrature energy (central differences, no clipping)
    N = 32
    x, y, h = grid(N)
    psi = 0.02 * np.stack([np.sin(2 * np.pi * y * h), np.cos(2 * np.pi * x * h)])
    rng = np.random.default_rng(1)
    phi = 0.2 * h * rng.standard_normal((2, N, N)) + 0.01
    phi[:, :2, :], phi[:, -2:, :], phi[:, :, :2], phi[:, :, -2:] = 0, 0, 0, 0  # stay inside the domain
    F = O.consistency_force(psi, phi)
    for k in range(3):
        v = rng.standard_normal((2, N, N)) * h
        v[:, :2, :], v[:, -2:, :], v[:, :, :2], v[:, :, -2:] = 0, 0, 0, 0
        eps = 1e-7
        fd = (O.consistency_energy(psi, phi + eps * v) - O.consistency_energy(psi, phi - eps * v)) / (2 * eps)
        assert abs(fd - np.sum(F * v)) <= 1e-5 * abs(fd)


@pytest.mark.parametrize("N", [2, 8, 33 - 1])
def test_stiffness_closed_forms(N):
    x, y, h = grid(N)
    assert np.abs(O.stiff_apply(np.ones((N, N)))).max() < 1e-14
    # interior 9-point stencil: centre 8/3, the eight neighbours -1/3
    u = np.zeros((N, N))
    if N >= 4:
        u[N // 2, N // 2] = 1.0
        k = O.stiff_apply(u)
        blk = k[N // 2 - 1:N // 2 + 2, N // 2 - 1:N // 2 + 2]
        np.testing.assert_allclose(blk, np.array([[-1, -1, -1], [-1, 8, -1], [-1, -1, -1]]) / 3.0, atol=1e-15)
    # 1/2 int |grad phi|^2 of the bilinear interpolant of phi_x = x (domain units): 1/2 per component
    phi = np.stack([x * h, y * h])
    assert abs(O.reg_energy(phi) - 1.0) < 1e-12


@pytest.mark.parametrize("N", [2, 8, 64])
def test_vcycle_invariants(N):
    rng = np.random.default_rng(N)
    u, v = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    # sigma = 0: the heat operator is the identity and so is the cycle
    assert np.array_equal(O.vcycle(u, 0.0), u)
    # linear
    np.testing.assert_allclose(O.vcycle(2 * u - 3 * v, 0.5), 2 * O.vcycle(u, 0.5) - 3 * O.vcycle(v, 0.5), atol=1e-12)
    # constants are kernel of the Laplacian (mirror boundary): reproduced
    np.testing.assert_allclose(O.vcycle(np.full((N, N), 1.5), 1.0), 1.5, rtol=1e-13)
    # smoothing damps: the cycle never increases the l2 norm of a random field
    assert np.linalg.norm(O.vcycle(u, 1.0)) <= np.linalg.norm(u)


def test_heat_operator_mirror_boundary():
    # A u = u + hs (4u - neighbours) with the missing neighbour mirrored: for u = x^2 on an 8-grid
    N = 8
    x, y, _ = grid(N)
    u = x ** 2
    tau = 0.1
    hs = tau * (N - 1) ** 2
    Au = O.heat_apply(u, tau)
    lap = np.zeros_like(u)  # 4u - sum of neighbours with mirrored ghost nodes
    up = np.pad(u, 1, mode="reflect")
    lap = 4 * u - up[1:-1, 2:] - up[1:-1, :-2] - up[2:, 1:-1] - up[:-2, 1:-1]
    np.testing.assert_allclose(Au, u + hs * lap, atol=1e-12)


def test_transfer_invariants():
    rng = np.random.default_rng(3)
    c = rng.random((8, 8))
    assert np.array_equal(O.restrict(O.prolong(c)), c)
    f = rng.random((16, 16))
    assert abs(O.restrict(f).mean() - f.mean()) < 1e-15
    # resampling reproduces bilinear functions exactly (affine here)
    x, y, _ = grid(8)
    a = 0.2 + 0.5 * x - 0.3 * y
    X, Y, _ = grid(15 + 1)
    s = 7.0 / 15.0
    np.testing.assert_allclose(O.resample(a, 16), 0.2 + 0.5 * X * s - 0.3 * Y * s, atol=1e-13)


def test_warp_invariants():
    N = 16
    x, y, h = grid(N)
    img = np.random.default_rng(1).random((N, N))
    assert np.array_equal(O.deform_image(img, np.zeros((2, N, N))), img)
    # integer translation by (2, -1) px: shifted image inside, Inf outside
    phi = np.stack([np.full((N, N), 2 * h), np.full((N, N), -1 * h)])
    w = O.deform_image(img, phi, np.inf)
    np.testing.assert_allclose(w[1:, : N - 2], img[: N - 1, 2:], atol=1e-14)
    assert np.all(np.isinf(w[0, :])) and np.all(np.isinf(w[:, N - 1]))
    # bilinear interpolation is exact on affine images
    aff = 1.0 + 0.25 * x + 0.5 * y
    phi = np.stack([0.37 * h * np.ones((N, N)), 0.61 * h * np.ones((N, N))])
    w = O.fast_deform_image(aff, phi)
    exp = 1.0 + 0.25 * np.minimum(x + 0.37, N - 1) + 0.5 * np.minimum(y + 0.61, N - 1)
    np.testing.assert_allclose(w, exp, atol=1e-13)


def test_compose_identities():
    N = 16
    x, y, h = grid(N)
    p1 = np.stack([0.5 * h * np.sin(x), 0.4 * h * np.cos(y)])
    # composing with zero displacement returns phi1 exactly
    assert np.array_equal(O.compose(p1, np.zeros((2, N, N))), p1)
    # composing zero with a translation keeps it where defined and extends it harmonically (a constant)
    t = np.stack([np.full((N, N), 1.5 * h), np.full((N, N), -2.0 * h)])
    np.testing.assert_allclose(O.compose(np.zeros((2, N, N)), t), t, atol=1e-9)


def test_series_stats_brute_force():
    rng = np.random.default_rng(9)
    F, N = 7, 12
    st = rng.random((F, N, N))
    st[rng.random((F, N, N)) < 0.4] = np.inf
    m, c, med = O.series_stats(st, 0.5)
    for j in range(N * N):
        vals = st.reshape(F, -1)[:, j]
        vals = vals[np.isfinite(vals)]
        if len(vals) == 0:
            assert m.ravel()[j] == 0.5 and med.ravel()[j] == 0.5 and c.ravel()[j] == 0
        else:
            assert c.ravel()[j] == len(vals)
            assert abs(m.ravel()[j] - vals.mean()) < 1e-15
            assert med.ravel()[j] == np.median(vals)


def test_normalize_formula():
    rng = np.random.default_rng(4)
    img = rng.random((32, 32)) * 5 - 2
    n = img.size
    srt = np.sort(img.ravel())
    lo, hi = srt[int(0.15 / 200 * (n - 1))], srt[int((1 - 0.15 / 200) * (n - 1))]
    exp = (np.clip(img, lo, hi) - lo) / (hi - lo)
    np.testing.assert_allclose(O.normalize(img, 0.15, True), exp, atol=1e-15)
    exp2 = (img - img.min()) / (img.max() - img.min())
    np.testing.assert_allclose(O.normalize(img, -1, True), exp2, atol=1e-15)


def test_level_solve_identical_images_filter_schedule():
    # identical images: zero gradient, every step is rejected (tau = 0); the filter width halves each
    # iteration from 1 until 0.5 * sigma <= 1e-4, i.e. 13 halvings, and the 14th iteration stops.
    N = 16
    x, y, _ = grid(N)
    R = np.sin(x) + np.cos(y)
    phi, info = O.level_solve(R, R, np.zeros((2, N, N)), 10.0, 1000, 0.0)
    assert info["iterations"] == 14
    assert np.all(phi == 0)
    assert np.all(info["taus"] == 0)
    np.testing.assert_array_equal(info["sigmas"][::2], 0.5 ** np.arange(14))


def test_level_solve_decreases_energy_monotonically():
    N = 32
    x, y, h = grid(N)
    R = np.sin(2 * np.pi * x / N) * np.sin(2 * np.pi * y / N)
    T = np.sin(2 * np.pi * (x - 1.3) / N) * np.sin(2 * np.pi * (y + 0.8) / N)
    _, info = O.level_solve(R, T, np.zeros((2, N, N)), 1.0, 60, -1.0)
    e = info["energies"]
    assert np.all(np.diff(e) <= 1e-15)   # Armijo accepts only descent steps
    assert e[-1] < 0.2 * e[0]


@pytest.mark.parametrize("N", [8, 16])
def test_phase_correlation_is_the_circular_cross_correlation(N):
    # un-normalised: out(s) = sum_x a(x + s) b(x) (circular), by brute force; normalised: the same
    # formula on unit-modulus spectra, by numpy's independent FFT
    rng = np.random.default_rng(N)
    a, b = rng.random((N, N)), rng.random((N, N))
    brute = np.array([[np.sum(np.roll(a, (-sy, -sx), axis=(0, 1)) * b) for sx in range(N)] for sy in range(N)])
    np.testing.assert_allclose(O.phase_correlation(a, b, False), brute, rtol=1e-13, atol=1e-12)
    g = np.fft.fft2(a) * np.conj(np.fft.fft2(b))
    np.testing.assert_allclose(O.phase_correlation(a, b, True), np.real(np.fft.ifft2(g / np.abs(g))), atol=1e-14)
    # a pure circular shift gives a unit delta at the shift
    pc = O.phase_correlation(a, np.roll(a, (-2, 3), axis=(0, 1)), True)
    assert abs(pc[2, N - 3] - 1.0) < 1e-12 and np.abs(np.delete(pc.ravel(), 2 * N + N - 3)).max() < 1e-12


def test_register_translation_conventions():
    # reference(x + b) = templ(x): b is returned signed in (-N/2, N/2], x first
    rng = np.random.default_rng(1)
    N = 32
    a = rng.random((N, N))
    for bx, by in ((0, 0), (3, -5), (-7, 2), (N // 2, 1), (1 - N // 2, -N // 2 + 1)):
        assert O.register_translation(a, np.roll(a, (-by, -bx), axis=(0, 1))) == (bx, by)
    # a window that holds the shift finds it on the plain correlation; one that does not returns the
    # best position inside it (brute force over the window, first strict maximum in x-fastest order)
    b = np.roll(a, (-2, -3), axis=(0, 1))
    assert O.register_translation(a, b, 4) == (3, 2)
    m = 1
    cc = O.phase_correlation(a, b, False)
    best, at = cc[0, 0], (0, 0)
    for dy in range(-m, m + 1):
        for dx in range(-m, m + 1):
            if cc[dy % N, dx % N] > best:
                best, at = cc[dy % N, dx % N], (dx, dy)
    assert O.register_translation(a, b, m) == at


def test_approx_inverse_of_translations():
    # phi = t (whole pixels): every node x - t inside the grid is covered, with value -t; the
    # others take the midpoint of the covered range, also -t. The zero field inverts to zero.
    N = 16
    h = 1.0 / (N - 1)
    assert np.abs(O.approx_inverse(np.zeros((2, N, N)))).max() < 1e-15
    for tx, ty in ((2, 0), (-3, 1), (1, -4)):
        phi = np.stack([np.full((N, N), tx * h), np.full((N, N), ty * h)])
        inv = O.approx_inverse(phi)
        np.testing.assert_allclose(inv[0], -tx * h, atol=1e-14)
        np.testing.assert_allclose(inv[1], -ty * h, atol=1e-14)


def test_approx_inverse_composes_to_identity():
    # a small smooth deformation: phi(x + inv(x)) + inv(x) ~ 0 (the inverse is exact at the nodes up
    # to the piecewise-linear interpolation of phi), checked on the interior
    N = 64
    h = 1.0 / (N - 1)
    y, x = np.mgrid[0:N, 0:N] * h
    phi = np.stack([1.5 * h * np.sin(2 * np.pi * y), 1.5 * h * np.cos(2 * np.pi * x)])
    inv = O.approx_inverse(phi)
    px, py = x + inv[0], y + inv[1]
    i, j = np.clip((px / h).astype(int), 0, N - 2), np.clip((py / h).astype(int), 0, N - 2)
    fx, fy = px / h - i, py / h - j

    def bil(f):
        return ((1 - fx) * (1 - fy) * f[j, i] + fx * (1 - fy) * f[j, i + 1] + (1 - fx) * fy * f[j + 1, i]
                + fx * fy * f[j + 1, i + 1])
    r = np.stack([bil(phi[0]) + inv[0], bil(phi[1]) + inv[1]])[:, 4:-4, 4:-4]
    assert np.abs(r).max() < 0.05 * h
