"""Write tests/golden/seismic.npz: golden vectors of the x-shift registration (matchSeismicSeries).

  * xs_<N> (N = 8, 32, 64): seeded seismic-like (R, T) and an x-shift field phi_x; the reference's
    FELineShiftRegisEnergy (registration.h:2138-2193) energy and derivative and the stiffness form
    1/2 phi_x^T (K_xx + y K_yy) phi_x of findXShift (matchSeries.h:805-868), from ref_harness "xshift";
  * xs_solve32: findXShift's gradient flow (H1 descent on the dyadic grid, 60 steps, lambda 5,
    yRegFactor 1) from zero, from ref_harness "xshift_solve": the deformation, the step count, the
    final energy, the energy trace and the accepted step sizes;
  * series: the reference program matchSeismicSeries (oracle/_ref/matchSeismicSeries) on a 4-frame
    64x64 series with the dyadic path (dyadicDepth = precisionLevel = stopLevel), two stages,
    initialAverage by stacking: its initial average, every pairwise and refinement deformation
    (level 6), the averages and medians and the energy logs of both stages.

Frames: each row is a trace (a reflectivity series convolved with a Ricker wavelet, plus 0.5) shifted
along x by a smooth per-row amount. Runs in the build container only (needs oracle/_ref).
Usage: python tests/golden/make_seismic.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_golden as G  # noqa: E402

SEISMIC = os.path.join(os.path.dirname(G.MATCHSERIES), "matchSeismicSeries")

SERIES_PAR = """templateNamePattern in/frame_%02d.dat.bz2
templateNumOffset 0
templateNumStep 1
numTemplates 4
saveDirectory out/
lambda 5
lambdaFactor 1
maxGDIterations 40
stopEpsilon 0
startLevel 4
stopLevel 6
precisionLevel 6
dyadicDepth 6
refineStartLevel 5
refineStopLevel 6
numExtraStages 1
extraStagesLambdaFactor 0.1
useMedianAsNewTarget 1
reuseStage1Results 1
skipStage1 0
yRegFactor 1
laplaceWeight 0
refLineDistance 2
justStackIniAverage 1
keepEmptyTopBottomBlocks 0
resampleInsteadOfProlongateDeformation 1
dontAccumulateDeformation 0
onlySaveDisplacement 1
checkboxWidth 8
"""


def trace(N, rng, nrefl):
    """A reflectivity series (nrefl spikes, coefficients in +-0.2) convolved with a Ricker wavelet."""
    refl = np.zeros(N)
    idx = rng.choice(np.arange(2, N - 2), nrefl, replace=False)
    refl[idx] = rng.uniform(-0.2, 0.2, nrefl)
    f = 0.12 * 64 / N
    k = min(10, (N - 1) // 2)
    t = np.arange(-k, k + 1)
    w = (1 - 2 * (np.pi * f * t) ** 2) * np.exp(-(np.pi * f * t) ** 2)
    return np.convolve(refl, w, mode="same")


def seis_pair(N, seed):
    rng = np.random.default_rng(seed)
    xs = np.arange(N)
    tr = trace(N, rng, max(N // 8, 2))
    R = np.stack([np.interp(xs - 0.6 * np.sin(2 * np.pi * y / N), xs, tr) for y in range(N)]) + 0.5
    T = np.stack([np.interp(xs - 0.6 * np.sin(2 * np.pi * y / N) - 1.3 * np.cos(3 * np.pi * y / N), xs, tr)
                  for y in range(N)]) + 0.5
    return R, T


def seis_series(N, count, seed):
    rng = np.random.default_rng(seed)
    xs = np.arange(N)
    tr = trace(N, rng, 12)
    out = []
    for _ in range(count):
        sh = rng.uniform(-1.5, 1.5) + 0.8 * np.sin(2 * np.pi * (np.arange(N) / N + rng.uniform()))
        F = np.stack([np.interp(xs - sh[y] - 0.5 * np.sin(2 * np.pi * y / N), xs, tr) for y in range(N)]) + 0.5
        out.append(F * 100)
    return np.stack(out)


def main():
    if not os.path.exists(SEISMIC):
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for N in (8, 32, 64):
            R, T = seis_pair(N, N)
            h = 1.0 / (N - 1)
            rng = np.random.default_rng(N + 1)
            phix = 1.5 * h * np.sin(np.arange(N * N).reshape(N, N) / N) + 0.3 * h * rng.standard_normal((N, N))
            lam, yreg = 5.0, 0.7
            o = G.harness("xshift", N, [lam, yreg], [R, T, phix], tmp)
            out.update({f"xs{N}_R": R, f"xs{N}_T": T, f"xs{N}_phix": phix, f"xs{N}_lam_yreg": np.array([lam, yreg]),
                        f"xs{N}_e_ncc": o[0], f"xs{N}_e_reg": o[1], f"xs{N}_DE": o[2:].reshape(N, N)})
        N = 32
        R, T = seis_pair(N, 5)
        lam, yreg, iters = 5.0, 1.0, 60
        o = G.harness("xshift_solve", N, [lam, yreg, iters], [R, T, np.zeros((N, N))], tmp)
        NN = N * N
        it = int(o[NN])
        nc = int(o[NN + 3 + it])
        out.update(solve_R=R, solve_T=T, solve_params=np.array([lam, yreg, iters]), solve_phix=o[:NN].reshape(N, N),
                   solve_iterations=np.array(it), solve_energy=o[NN + 1], solve_trace=o[NN + 2:NN + 3 + it],
                   solve_taus=o[NN + 4 + it:NN + 4 + it + nc])
    frames = seis_series(64, 4, 3)
    out["frames"] = frames
    out["par"] = np.frombuffer(SERIES_PAR.encode(), dtype=np.uint8)
    with tempfile.TemporaryDirectory() as wd:
        os.makedirs(os.path.join(wd, "in"))
        for i, f in enumerate(frames):
            G.write_p9(os.path.join(wd, "in", f"frame_{i:02d}.dat.bz2"), f)
        with open(os.path.join(wd, "run.par"), "w") as f:
            f.write(SERIES_PAR)
        subprocess.run([SEISMIC, "run.par"], cwd=wd, env=dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON"),
                       capture_output=True, text=True, check=True)
        od = os.path.join(wd, "out")
        out["initial_average"] = G.read_p9(os.path.join(od, "initialAverage.q2bz"))
        for st, subs in ((1, ["0", "1", "2", "3", "1-r", "2-r", "3-r"]), (2, ["0", "1-r", "2-r", "3-r"])):
            sd = os.path.join(od, f"stage{st}")
            for sub in subs:
                key = f"s{st}_def_{sub.replace('-', '_')}"
                out[key] = np.stack([G.read_p9(os.path.join(sd, sub, f"deformation_06_{c}.dat.bz2")) for c in range(2)])
                if st == 1:
                    with open(os.path.join(sd, sub, "energy_06.txt")) as f:
                        out[f"s1_energy06_{sub.replace('-', '_')}"] = np.loadtxt(f, ndmin=2)
            for name in ("average", "median"):
                out[f"s{st}_{name}"] = G.read_p9(os.path.join(sd, f"{name}.q2bz"))
            out[f"s{st}_energies_txt"] = np.frombuffer(open(os.path.join(sd, "energies.txt"), "rb").read(),
                                                        dtype=np.uint8)
        out["s1_defnorms_txt"] = np.frombuffer(open(os.path.join(od, "stage1", "defNorms.txt"), "rb").read(),
                                               dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "seismic.npz"), **out)
    print("wrote seismic.npz with", len(out), "arrays")


if __name__ == "__main__":
    main()
