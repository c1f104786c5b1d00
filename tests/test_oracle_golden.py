"""The oracle (oracle/liboracle.so) against golden vectors written by the reference itself
(tests/golden/make_golden.py: the reference's operators and the reference matchSeries program).

Bit-exact where the oracle follows the reference's operation order (SSD energy and variation,
V-cycle, warps, restriction, injection, normalisation, statistics); rounding-level where the
summation order differs (assembled stiffness rows, long-double resampling steps, the PCG of the
composition); exact step-size traces for complete level solves.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("N", [8, 32, 64])
def test_ssd_energy_force_stiffness(N):
    g = load(f"ssd{N}")
    assert O.ssd_energy(g["R"], g["T"], g["phi"]) == g["e_ssd"]
    assert np.array_equal(O.ssd_force(g["R"], g["T"], g["phi"]), g["de_ssd"])
    assert abs(O.reg_energy(g["phi"]) - g["e_reg"]) <= 1e-14 * g["e_reg"]
    k = np.stack([O.stiff_apply(g["phi"][0]), O.stiff_apply(g["phi"][1])])
    assert rel(k, g["kphi"]) < 1e-14


@pytest.mark.parametrize("name", ["smooth8_0", "smooth8_1", "smooth64_0", "smooth64_1"])
def test_vcycle(name):
    g = load(name)
    assert np.array_equal(O.vcycle(g["d"], float(g["sigma"])), g["out"])


def test_warps():
    g = load("warp32")
    a = O.deform_image(g["img"], g["phi"], np.inf)
    assert np.array_equal(np.isinf(a), np.isinf(g["deform_inf"]))
    assert np.array_equal(a[np.isfinite(a)], g["deform_inf"][np.isfinite(a)])
    assert np.array_equal(O.fast_deform_image(g["img"], g["phi"]), g["fast"])
    g = load("warp32_nn")
    assert np.array_equal(O.deform_image(g["img"], g["phi"], 0.25, nearest=True), g["deform_nn"])


def test_transfer():
    g = load("transfer32")
    assert np.array_equal(O.restrict(g["fine"]), g["restrict"])
    assert np.array_equal(O.prolong(g["coarse"]), g["prolong"])
    np.testing.assert_array_equal(O.resample(g["coarse"], 32), g["resample"])


def test_compose():
    g = load("compose32")
    assert rel(O.compose(g["phi1"], g["phi2"]), g["out"]) < 1e-13


@pytest.mark.parametrize("tag", ["sat", "nosat"])
def test_normalize_and_pyramid(tag):
    g = load(f"normalize32_{tag}")
    out = O.normalize(g["img"], float(g["pct"]), True)
    assert np.array_equal(out, g["out"])
    assert np.array_equal(O.restrict(out), g["level4"])  # the 16x16 level of the pyramid


def test_series_stats():
    g = load("stats16")
    m, c, med = O.series_stats(g["stack"], float(g["fill"]))
    assert np.array_equal(c, g["count"])
    assert np.array_equal(m, g["mean"])
    assert np.array_equal(med, g["median"])


@pytest.mark.parametrize("N", [8, 32, 64])
def test_ncc_energy_variation_mass(N):
    """deformationModel 0: NormalizedCrossCorrelationEnergy / its variation and the assembled mass
    matrix, against the reference's own classes (bit for bit)."""
    g = load(f"ncc{N}")
    assert O.ncc_energy(g["R"], g["T"], g["phi"]) == g["e_ncc"]
    np.testing.assert_array_equal(O.ncc_force(g["R"], g["T"], g["phi"]), g["de_ncc"])
    np.testing.assert_array_equal(O.mass_apply(g["R"]), g["mass_r"])
    rn, var = O.ncc_normalize(g["R"])
    np.testing.assert_array_equal(rn, g["r_norm"])
    assert var == g["var_r"]


@pytest.mark.parametrize("name", ["solve32_lam10", "solve64_lam10", "solve64_lam1", "solve32_lam10_ncc",
                                  "solve64_lam1_ncc"])
def test_level_solve_trace(name):
    g = load(name)
    N = g["R"].shape[0]
    phi, info = O.level_solve(g["R"], g["T"], np.zeros((2, N, N)), float(g["lam"]), int(g["max_iter"]),
                              float(g["stop_eps"]), data_term=O.NCC if int(g["model"]) == 0 else O.SSD)
    assert info["iterations"] == int(g["iterations"])
    np.testing.assert_array_equal(info["taus"], g["taus"])
    np.testing.assert_array_equal(info["sigmas"], g["sigmas"])
    np.testing.assert_array_equal(info["evals"], g["evals"])
    np.testing.assert_allclose(info["energies"], g["energies"], rtol=1e-12)
    assert np.abs(phi - g["phi"]).max() * (N - 1) < 1e-10


def e2e_params(lam, data_term=O.SSD):
    return O.SeriesParams(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
                          lambda_=lam, lambda_factor=1.0, max_iter=15, stop_eps=-1.0, saturation_pct=0.15,
                          normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1, dont_accumulate=0,
                          data_term=data_term)


def test_reference_program_two_stages_ncc():
    """deformationModel 0 end to end against the reference program's output tree."""
    g = load("e2e64_ncc")
    fr = g["frames"]
    N = fr.shape[-1]
    r1 = O.match_series(fr[1:], fr[0], e2e_params(10.0, O.NCC))
    assert np.abs(r1["deformations"] - g["s1_deformations"]).max() * (N - 1) < 1e-10
    assert np.abs(r1["pairwise"] - g["s1_pairwise"]).max() * (N - 1) < 1e-10
    assert np.array_equal(r1["count"], g["s1_numSamples"])
    np.testing.assert_allclose(r1["median"], g["s1_median"], atol=1e-10)
    np.testing.assert_allclose(r1["energies"], g["s1_energies"], rtol=1e-5)
    pw = np.zeros((4, 2, N, N))
    pw[1:] = g["s1_pairwise"]
    r2 = O.match_series(fr, g["s1_median"], e2e_params(1.0, O.NCC), pw)
    assert np.abs(r2["deformations"] - g["s2_deformations"]).max() * (N - 1) < 1e-7
    assert np.array_equal(r2["count"], g["s2_numSamples"])
    np.testing.assert_allclose(r2["median"], g["s2_median"], atol=1e-8)
    np.testing.assert_allclose(r2["energies"], g["s2_energies"], rtol=1e-5)


def test_reference_program_two_stages():
    g = load("e2e64")
    fr = g["frames"]
    N = fr.shape[-1]
    r1 = O.match_series(fr[1:], fr[0], e2e_params(10.0))
    assert np.abs(r1["deformations"] - g["s1_deformations"]).max() * (N - 1) < 1e-10
    assert np.abs(r1["pairwise"] - g["s1_pairwise"]).max() * (N - 1) < 1e-10
    assert np.array_equal(r1["count"], g["s1_numSamples"])
    np.testing.assert_allclose(r1["mean"], g["s1_average"], atol=1e-11)
    np.testing.assert_allclose(r1["median"], g["s1_median"], atol=1e-11)
    np.testing.assert_allclose(r1["energies"], g["s1_energies"], rtol=1e-5)  # energies.txt holds 6 digits
    # stage 2: all frames against the stage-1 median, lambda * 0.1, stage-1 pairwise deformations reused
    pw = np.zeros((4, 2, N, N))
    pw[1:] = g["s1_pairwise"]
    r2 = O.match_series(fr, g["s1_median"], e2e_params(1.0), pw)
    assert np.abs(r2["deformations"] - g["s2_deformations"]).max() * (N - 1) < 1e-4
    assert np.array_equal(r2["count"], g["s2_numSamples"])
    np.testing.assert_allclose(r2["median"], g["s2_median"], atol=1e-4)
    np.testing.assert_allclose(r2["energies"], g["s2_energies"], rtol=1e-5)


@pytest.mark.parametrize("N", [8, 32, 64])
def test_consistency_energy_and_variation(N):
    """reduceDeformations' ConsistencyEnergyOp and its variation, bit for bit."""
    g = load(f"consistency{N}")
    assert O.consistency_energy(g["psi"], g["phi"]) == g["e"]
    np.testing.assert_array_equal(O.consistency_force(g["psi"], g["phi"]), g["de"])


def test_reduce_deformations_solve():
    g = load("reduce32")
    phi, e, it = O.reduce_deformations(g["psi"], int(g["max_iter"]), float(g["stop_eps"]))
    assert it == int(g["iterations"])
    assert e == g["energy"]
    np.testing.assert_array_equal(phi, g["phi"])


def test_reference_program_two_stages_reduce():
    """reduceDeformations 1 end to end: the reduced deformation of each stage, the averages it enters,
    and stage 2's template 0 starting from stage 1's reduced deformation."""
    g = load("e2e64_reduce")
    fr = g["frames"]
    N = fr.shape[-1]

    def params(lam):
        p = e2e_params(lam)
        p.reduce_deformations, p.reduce_max_iter, p.reduce_stop_eps = 1, 15, -1.0
        return p
    r1 = O.match_series(fr[1:], fr[0], params(10.0))
    assert np.abs(r1["deformations"] - g["s1_deformations"]).max() * (N - 1) < 1e-10
    assert np.abs(r1["reduced"] - g["s1_reduceDef"]).max() * (N - 1) < 1e-10
    assert np.array_equal(r1["count"], g["s1_numSamples"])
    np.testing.assert_allclose(r1["mean"], g["s1_average"], atol=1e-11)
    np.testing.assert_allclose(r1["median"], g["s1_median"], atol=1e-11)
    pw = np.zeros((4, 2, N, N))
    pw[1:] = g["s1_pairwise"]
    r2 = O.match_series(fr, g["s1_median"], params(1.0), pw, pair0_init=g["s1_reduceDef"])
    assert np.abs(r2["deformations"] - g["s2_deformations"]).max() * (N - 1) < 1e-4
    assert np.abs(r2["reduced"] - g["s2_reduceDef"]).max() * (N - 1) < 1e-4
    assert np.array_equal(r2["count"], g["s2_numSamples"])
    np.testing.assert_allclose(r2["median"], g["s2_median"], atol=1e-4)



@pytest.mark.parametrize("N", [8, 32, 64])
def test_approx_inverse(N):
    # qc::approxInverseDeformation through the reference harness: smooth, noisy and folding fields
    g = load("inverse")
    for k in range(3):
        np.testing.assert_array_equal(O.approx_inverse(g[f"phi{N}_{k}"]), g[f"inv{N}_{k}"])


def test_reference_program_inverse_solves():
    # calcInverseDeformation: the reference program's <i>-r-inv/ solves of stage 1
    g = load("e2e64_inv")
    fr = g["frames"]
    N = fr.shape[-1]
    p = e2e_params(10.0)
    p.calc_inverse = 1
    r1 = O.match_series(fr[1:], fr[0], p)
    assert np.abs(r1["deformations"] - g["s1_deformations"]).max() * (N - 1) < 1e-10
    for i in range(3):
        assert np.abs(r1["inverses"][i] - g[f"s1_{i}_r_inv_deformation_06"]).max() * (N - 1) < 1e-10, i


def test_resample_rect():
    # resizeInput's ScalarArray<QC_2D>::resampleFrom of non-square sources, through the harness
    g = load("resize")
    for k in range(3):
        N = g[f"out{k}"].shape[0]
        np.testing.assert_array_equal(O.resample_rect(g[f"src{k}"], N), g[f"out{k}"])
