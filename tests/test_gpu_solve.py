"""GPU level solve and series matching against the oracle's restatement of the reference loop.

MSG_F64 must reproduce the oracle's trajectory: the same step size at every Armijo call, the same
number of iterations, energies to 1e-10 and the deformation to 1e-8 px. MSG_F32 (FP32 fields,
FP64 energies, slopes and decisions) is held to the north-star bar: final energy within 1e-6
relative and final deformation within 1e-3 px after the fixed iteration schedule.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from msamd import gpu as G
    G.lib()
    return G


def stem_pair(N, seed_offset=0):
    from msamd import synth
    sc = N / 256.0
    s = np.stack([synth.stem_frame(N, synth.frame_seed(0, f + seed_offset), spacing=max(12 * sc, 3.0),
                                   sigma=max(2.5 * sc, 0.8), jitter_step=0.3 * sc) for f in range(2)])
    return O.normalize(s[0], 0.15, True), O.normalize(s[1], 0.15, True)


def gpu_calls(trace):
    """Per Armijo call step sizes from the per-iteration trace (tau_first, then tau if it fell back)."""
    out = []
    for row in trace:
        out.append(row[3])
        if row[3] == 0:
            out.append(row[2])
    return np.array(out)


@pytest.mark.parametrize("N,lam,its,eps", [(16, 10.0, 30, -1.0), (64, 10.0, 50, -1.0), (64, 1.0, 200, 1e-6),
                                            (128, 10.0, 25, -1.0)])
def test_level_solve_f64_trajectory(gpu, N, lam, its, eps):
    R, T = stem_pair(N)
    L = int(np.log2(N))
    ctx = gpu.Context(L, 1, gpu.F64)
    ctx.set_reference(0, R)
    ctx.set_template(0, T)
    e, it, tr = ctx.solve_level(L, lam, its, eps, trace_cap=its)
    phi = ctx.get_deformation(0)
    ctx.close()
    phi_o, info = O.level_solve(R, T, np.zeros((2, N, N)), lam, its, eps)
    assert it[0] == info["iterations"]
    np.testing.assert_array_equal(gpu_calls(tr[0]), info["taus"])
    np.testing.assert_allclose(tr[0][:, 0], info["energies"][1:], rtol=1e-10)
    assert abs(e[0] - info["energy"]) <= 1e-10 * abs(info["energy"])
    assert np.abs(phi - phi_o).max() * (N - 1) < 1e-8


@pytest.mark.parametrize("N,lam,its", [(8, 10.0, 15), (16, 10.0, 15), (32, 10.0, 40), (64, 1.0, 60), (128, 10.0, 25)])
def test_level_solve_ncc_f64_trajectory(gpu, N, lam, its):
    """deformationModel 0 (NCC): the same step sizes as the oracle at every Armijo call."""
    R, T = stem_pair(N)
    L = int(np.log2(N))
    ctx = gpu.Context(L, 1, gpu.F64)
    ctx.set_data_term(gpu.NCC)
    ctx.set_reference(0, R)
    ctx.set_template(0, T)
    e, it, tr = ctx.solve_level(L, lam, its, -1.0, trace_cap=its)
    phi = ctx.get_deformation(0)
    ctx.close()
    phi_o, info = O.level_solve(R, T, np.zeros((2, N, N)), lam, its, -1.0, data_term=O.NCC)
    assert it[0] == info["iterations"]
    np.testing.assert_array_equal(gpu_calls(tr[0]), info["taus"])
    np.testing.assert_allclose(tr[0][:, 0], info["energies"][1:], rtol=1e-10)
    assert np.abs(phi - phi_o).max() * (N - 1) < 1e-8


@pytest.mark.parametrize("N,lam,its", [(64, 10.0, 50), (256, 10.0, 30)])
def test_level_solve_f32(gpu, N, lam, its):
    R, T = stem_pair(N)
    L = int(np.log2(N))
    ctx = gpu.Context(L, 1, gpu.F32)
    ctx.set_reference(0, R)
    ctx.set_template(0, T)
    e, it, _ = ctx.solve_level(L, lam, its, -1.0)
    phi = ctx.get_deformation(0)
    ctx.close()
    # the oracle on the FP32-rounded images
    R32, T32 = R.astype(np.float32).astype(np.float64), T.astype(np.float32).astype(np.float64)
    phi_o, info = O.level_solve(R32, T32, np.zeros((2, N, N)), lam, its, -1.0)
    assert it[0] == info["iterations"]
    assert abs(e[0] - info["energy"]) <= 1e-6 * abs(info["energy"])
    assert np.abs(phi - phi_o).max() * (N - 1) < 1e-3


def test_batched_pairs_are_independent(gpu):
    N, L = 32, 5
    pairs = [stem_pair(N, 3 * k) for k in range(3)]
    ctx = gpu.Context(L, 3, gpu.F64)
    for k, (R, T) in enumerate(pairs):
        ctx.set_reference(k, R)
        ctx.set_template(k, T)
    e, it, _ = ctx.solve_level(L, 5.0, 40, -1.0)
    for k, (R, T) in enumerate(pairs):
        phi_o, info = O.level_solve(R, T, np.zeros((2, N, N)), 5.0, 40, -1.0)
        assert it[k] == info["iterations"]
        assert abs(e[k] - info["energy"]) <= 1e-10 * abs(info["energy"])
        assert np.abs(ctx.get_deformation(k) - phi_o).max() * (N - 1) < 1e-8
    ctx.close()


@pytest.mark.parametrize("resample,dterm", [(0, 0), (1, 0), (1, 1)])
def test_match_series_f64(gpu, resample, dterm):
    from msamd import synth
    N, L = 64, 6
    frames = synth.stem_series(N, 4, workload=0, spacing=6.0, sigma=1.5, jitter_step=0.15)
    p = O.SeriesParams(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
                       lambda_=10.0, lambda_factor=-1, max_iter=15, stop_eps=-1.0, saturation_pct=0.15,
                       normalize_min_to_zero=1, dont_normalize=0, resample_prolong=resample, dont_accumulate=0,
                       data_term=dterm)
    ref = O.match_series(frames[1:], frames[0], p)
    gp = gpu.SeriesParams(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
                          lambda_=10.0, lambda_factor=-1, max_iter=15, stop_eps=-1.0, saturation_pct=0.15,
                          normalize_min_to_zero=1, dont_normalize=0, resample_prolong=resample, dont_accumulate=0,
                          spec=0, data_term=dterm)
    t = torch.tensor(frames, dtype=torch.float64, device="cuda")
    out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gp)
    # twice: contexts come from the cache the second time and must not carry state over
    out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gp)
    d = out["deformations"].cpu().numpy()
    assert np.abs(d - ref["deformations"]).max() * (N - 1) < 1e-6
    np.testing.assert_allclose(out["energies"], ref["energies"], rtol=1e-9)
    np.testing.assert_array_equal(out["count"].cpu().numpy(), ref["count"])
    np.testing.assert_allclose(out["mean"].cpu().numpy(), ref["mean"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out["median"].cpu().numpy(), ref["median"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("N,its,B,dterm", [(32, 40, 3, 0), (128, 30, 2, 0), (64, 30, 2, 1)])
def test_device_driven_modes_match_host_mode(gpu, mode, N, its, B, dterm):
    """The device-driven iteration (control in one-workgroup-per-pair kernels, 8 iterations per host check,
    optionally one hipGraph) gives exactly the host-synchronised result."""
    L = int(np.log2(N))
    pairs = [stem_pair(N, 5 * k) for k in range(B)]
    out = {}
    for m in (0, mode):
        ctx = gpu.Context(L, B, gpu.F64)
        ctx.set_mode(m)
        ctx.set_data_term(dterm)
        for k, (R, T) in enumerate(pairs):
            ctx.set_reference(k, R)
            ctx.set_template(k, T)
        e, it, tr = ctx.solve_level(L, 10.0, its, -1.0, trace_cap=its)
        out[m] = (e, it, [ctx.get_deformation(k) for k in range(B)], tr)
        # a second solve reuses the captured graph
        e2, it2, _ = ctx.solve_level(L, 10.0, its, -1.0, trace_cap=its)
        np.testing.assert_array_equal(it2, it)
        ctx.close()
    e0, it0, phi0, tr0 = out[0]
    e1, it1, phi1, tr1 = out[mode]
    np.testing.assert_array_equal(it1, it0)
    np.testing.assert_array_equal(e1, e0)
    for k in range(B):
        np.testing.assert_array_equal(phi1[k], phi0[k])
        np.testing.assert_array_equal(tr1[k], tr0[k])


def test_match_series_independent_of_pairwise_batching(gpu):
    """The chain overlaps the later pairwise batches on a second host thread; the result is the same
    for every batch size (one pair per batch up to all pairs in one)."""
    from msamd import synth
    N = 64
    frames = synth.stem_series(N, 6, workload=0, spacing=6.0, sigma=1.5, jitter_step=0.15)
    t = torch.tensor(frames, dtype=torch.float32, device="cuda")
    outs = []
    for b in (1, 2, 5, 0):
        gp = gpu.SeriesParams(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6,
                              precision_level=6, lambda_=10.0, lambda_factor=-1, max_iter=12, stop_eps=-1.0,
                              saturation_pct=0.15, normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1,
                              dont_accumulate=0, spec=0, pairwise_batch=b)
        o = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gp)
        outs.append({k: (v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)) for k, v in o.items()})
    for o in outs[1:]:
        for k in ("deformations", "energies", "mean", "count", "median"):
            np.testing.assert_array_equal(o[k], outs[0][k], err_msg=k)


@pytest.mark.parametrize("ext_mean,nearest", [(1, 0), (0, 1), (1, 1)])
def test_match_series_averaging_options(gpu, ext_mean, nearest):
    """deformForAverageExtendedWithMean / ...UsingNearestNeighborInterpolation: the reconstruction from
    the returned deformations, restated with the oracle's DeformImage and statistics."""
    from msamd import synth
    N = 64
    frames = synth.stem_series(N, 4, workload=0, spacing=6.0, sigma=1.5, jitter_step=0.15)
    gp = gpu.SeriesParams(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
                          lambda_=10.0, lambda_factor=-1, max_iter=10, stop_eps=-1.0, saturation_pct=0.15,
                          normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1, dont_accumulate=0,
                          average_ext_mean=ext_mean, average_nearest=nearest)
    t = torch.tensor(frames, dtype=torch.float64, device="cuda")
    out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gp)
    d = out["deformations"].cpu().numpy()
    tpl = frames[1:]
    stack = np.stack([O.deform_image(tpl[i], d[i], ext=tpl[i].mean() if ext_mean else np.inf, nearest=bool(nearest))
                      for i in range(len(tpl))])
    mean, cnt, med = O.series_stats(stack, tpl[-1].mean())
    np.testing.assert_array_equal(out["count"].cpu().numpy(), cnt)
    np.testing.assert_allclose(out["mean"].cpu().numpy(), mean, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out["median"].cpu().numpy(), med, rtol=1e-12, atol=1e-12)


def test_reduce_deformations_f64(gpu):
    """reduceDeformations: plain H1 gradient descent (fixed filter width, gradient-norm stopping) on
    the consistency energy of the given deformations, against the oracle (bit-exact with the
    reference, fixture reduce32)."""
    import os
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reduce32.npz")) as z:
        g = {k: z[k] for k in z.files}
    N = g["psi"].shape[-1]
    psi = torch.tensor(g["psi"], dtype=torch.float64, device="cuda")
    for mode in (0, 2):
        gpu.set_default_mode(mode)
        gpu.release_cache()
        phi, e, it = gpu.reduce_deformations(psi, int(g["max_iter"]), float(g["stop_eps"]))
        assert it == int(g["iterations"])
        assert abs(e - g["energy"]) <= 1e-10 * abs(g["energy"])
        assert np.abs(phi.cpu().numpy() - g["phi"]).max() * (N - 1) < 1e-8
    gpu.set_default_mode(0)
    gpu.release_cache()


def test_match_series_reduce_f64(gpu):
    from msamd import synth
    N = 64
    frames = synth.stem_series(N, 4, workload=0, spacing=6.0, sigma=1.5, jitter_step=0.15)
    kw = dict(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6, lambda_=10.0,
              lambda_factor=-1, max_iter=12, stop_eps=-1.0, saturation_pct=0.15, normalize_min_to_zero=1,
              dont_normalize=0, resample_prolong=1, dont_accumulate=0, reduce_deformations=1, reduce_max_iter=12,
              reduce_stop_eps=-1.0)
    t = torch.tensor(frames, dtype=torch.float64, device="cuda")
    init = np.zeros((2, N, N))
    init[0] = 0.01 * np.sin(np.linspace(0, 3, N))[None, :]
    ref = O.match_series(frames[1:], frames[0], O.SeriesParams(**kw), pair0_init=init)
    ti = torch.tensor(init, dtype=torch.float64, device="cuda")
    gp = gpu.SeriesParams(**kw)
    gp.pair0_init = ti.data_ptr()
    out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gp)
    assert np.abs(out["deformations"].cpu().numpy() - ref["deformations"]).max() * (N - 1) < 1e-6
    assert np.abs(out["reduced"].cpu().numpy() - ref["reduced"]).max() * (N - 1) < 1e-6
    np.testing.assert_array_equal(out["count"].cpu().numpy(), ref["count"])
    np.testing.assert_allclose(out["median"].cpu().numpy(), ref["median"], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("energy_criterion", [0, 1])
def test_match_series_alt_start_level(gpu, energy_criterion):
    """altStartLevel: each pairwise solve repeated from zero at level 5 and the better one kept (by the
    template's overlap with the domain, or by energy), against the oracle's series loop."""
    from msamd import synth
    N = 64
    frames = synth.stem_series(N, 4, workload=0, spacing=6.0, sigma=1.5, jitter_step=0.15)
    kw = dict(start_level=4, stop_level=6, refine_start_level=5, refine_stop_level=6, precision_level=6,
              lambda_=10.0, lambda_factor=-1, max_iter=15, stop_eps=-1.0, saturation_pct=0.15,
              normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1, dont_accumulate=0, data_term=0,
              alt_start_level=5, energy_as_alt_criterion=energy_criterion)
    ref = O.match_series(frames[1:], frames[0], O.SeriesParams(**kw))
    t = torch.tensor(frames, dtype=torch.float64, device="cuda")
    out = gpu.match_series(t[1:].contiguous(), t[0].contiguous(), gpu.SeriesParams(spec=0, **kw), want_pairwise=True,
                           want_levels=True)
    assert np.abs(out["pairwise"].cpu().numpy() - ref["pairwise"]).max() * (N - 1) < 1e-6
    assert np.abs(out["first_try"].cpu().numpy() - ref["first_try"]).max() * (N - 1) < 1e-6
    assert np.abs(out["deformations"].cpu().numpy() - ref["deformations"]).max() * (N - 1) < 1e-6
    np.testing.assert_allclose(out["energies"], ref["energies"], rtol=1e-9)
    assert sum(lv["kind"] == "alt" for lv in out["levels"]) == 3 * 2
