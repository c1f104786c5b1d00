"""Write the golden fixtures of the benchmark configurations at sizes the CPU finishes:

  * cfg1_256.npz: 8 STEM frames of 256x256 (msamd.synth.CPU_SERIES), stage 1 of matchSeries with
    deformationModel 22, levels 6->8, refinement 7->8, 50 iterations per level (stopEpsilon -1),
    lambda 10;
  * cfg2_256.npz / cfg2_512.npz: the 1024x1024 generator of the GPU benchmark (GPU_SERIES)
    cropped to its top-left 256x256 (4 frames) / 512x512 (3 frames) blocks, levels L-3 -> L,
    refinement L-1 -> L, 100 iterations per level, lambda 10.

Each holds (deformations in pixels as float32, i.e. within 5e-7 px of the computed values):
  * ref_*: the reference program (oracle/_ref/matchSeries) on the FP64 frames: the deformations
    used for the average, the energies and the median;
  * o32_*: the oracle's series loop (oracle/ms_oracle.c) with dontNormalizeInputImages on the frames
    normalised in FP64 (enhanceContrastSaturationPercentage 0.15, normalizeMinToZero) and rounded
    to FP32 -- the inputs the GPU's FP32 mode computes on when handed those images, so the GPU's
    FP32 result can be held to the oracle without the input rounding in between.
and the oracle's own sensitivity of those results to rounding-level changes of its inputs (its result
on perturbed inputs against its result on the inputs):
  * ulp_spread_px / ulp_spread_energy: the FP64 frames times (1 +- 2^-52), random signs;
  * f32_spread_px / f32_spread_energy: the FP32-rounded normalised frames times (1 +- 2^-24).
A configuration whose results move by far more than the perturbation under the first is decided by
rounding-level step decisions (no implementation that sums in another order can follow it closer than
that spread); the second is the scale of what FP32 storage alone can move.
Frames are regenerated from the seeded generator by the tests (not stored).
Runs in the build container only (needs oracle/_ref and oracle's library).
Usage: python tests/golden/make_configs.py [cfg1] [cfg2_256] [cfg2_512] [--sensitivity-only]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "match-series_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import make_golden as G  # noqa: E402
import oracle as O  # noqa: E402
from msamd import synth  # noqa: E402

PAR = """templateNamePattern in/frame_%02d.dat.bz2
templateNumOffset 0
templateNumStep 1
numTemplates {nf}
deformationModel 22
saveDirectory out/
lambda 10
maxGDIterations {its}
stopEpsilon -1
startLevel {l0}
stopLevel {L}
precisionLevel {L}
refineStartLevel {r0}
refineStopLevel {L}
numExtraStages 0
extraStagesLambdaFactor 0.1
useMedianAsNewTarget 1
reuseStage1Results 1
skipStage1 0
dontNormalizeInputImages 0
enhanceContrastSaturationPercentage 0.15
normalizeMinToZero 1
resampleInsteadOfProlongateDeformation 1
dontAccumulateDeformation 0
reduceDeformations 0
onlySaveDisplacement 1
checkboxWidth 8
"""


def config(name):
    """(frames, levels (l0, L), refine start, iterations)"""
    if name == "cfg1":
        s = synth.CPU_SERIES
        fr = synth.stem_series(s["N"], s["frames"], workload=s["workload"], spacing=s["spacing"], sigma=s["sigma"],
                               jitter_step=s["jitter_step"])
        return fr, (6, 8), 7, 50
    S = int(name.split("_")[1])
    nf = 4 if S == 256 else 3
    g = synth.GPU_SERIES
    fr = synth.stem_series(g["N"], nf, workload=g["workload"], spacing=g["spacing"], sigma=g["sigma"],
                           jitter_step=g["jitter_step"])[:, :S, :S]
    L = int(np.log2(S))
    return np.ascontiguousarray(fr), (L - 3, L), L - 1, 100


def run_reference(frames, l0, L, r0, its, wd):
    os.makedirs(os.path.join(wd, "in"))
    for i, f in enumerate(frames):
        G.write_p9(os.path.join(wd, "in", f"frame_{i:02d}.dat.bz2"), f)
    with open(os.path.join(wd, "run.par"), "w") as f:
        f.write(PAR.format(nf=len(frames), its=its, l0=l0, L=L, r0=r0))
    env = dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON")
    subprocess.check_call([G.MATCHSERIES, "run.par"], cwd=wd, env=env, stdout=subprocess.DEVNULL,
                          stderr=subprocess.DEVNULL)
    sd = os.path.join(wd, "out", "stage1")
    N = 1 << L
    defs = []
    for i in range(len(frames) - 1):
        sub = "0" if i == 0 else f"{i}-r"
        defs.append(np.stack([G.read_p9(os.path.join(sd, sub, f"deformation_{L:02d}_{c}.dat.bz2")) for c in range(2)]))
    return dict(ref_defs_px=(np.stack(defs) * (N - 1)).astype(np.float32),
                ref_energies=np.loadtxt(os.path.join(sd, "energies.txt"))[:, 1],
                ref_median=G.read_p9(os.path.join(sd, "median.q2bz")).astype(np.float32),
                ref_count=G.read_p9(os.path.join(sd, "numSamples.q2bz")).astype(np.float32))


def run_oracle_rounded(frames, l0, L, r0, its):
    norm = np.stack([O.normalize(f, 0.15, True) for f in frames]).astype(np.float32).astype(np.float64)
    p = O.SeriesParams(start_level=l0, stop_level=L, refine_start_level=r0, refine_stop_level=L, precision_level=L,
                       lambda_=10.0, lambda_factor=-1, max_iter=its, stop_eps=-1.0, saturation_pct=-1.0,
                       normalize_min_to_zero=1, dont_normalize=1, resample_prolong=1, dont_accumulate=0, data_term=0)
    r = O.match_series(norm[1:], norm[0], p)
    N = 1 << L
    return dict(o32_defs_px=(r["deformations"] * (N - 1)).astype(np.float32), o32_energies=r["energies"])


def oracle_series(frames, l0, L, r0, its, normalised):
    p = O.SeriesParams(start_level=l0, stop_level=L, refine_start_level=r0, refine_stop_level=L, precision_level=L,
                       lambda_=10.0, lambda_factor=-1, max_iter=its, stop_eps=-1.0,
                       saturation_pct=-1.0 if normalised else 0.15, normalize_min_to_zero=1,
                       dont_normalize=1 if normalised else 0, resample_prolong=1, dont_accumulate=0, data_term=0)
    return O.match_series(frames[1:], frames[0], p)


def sensitivity(frames, l0, L, r0, its):
    N = 1 << L
    rng = np.random.default_rng(2026)
    out = {}
    norm = np.stack([O.normalize(f, 0.15, True) for f in frames]).astype(np.float32).astype(np.float64)
    for tag, base_in, rel, normalised in (("ulp", frames, 2.0 ** -52, False), ("f32", norm, 2.0 ** -24, True)):
        a = oracle_series(base_in, l0, L, r0, its, normalised)
        b = oracle_series(base_in * (1 + rel * rng.choice([-1.0, 1.0], size=base_in.shape)), l0, L, r0, its, normalised)
        out[f"{tag}_spread_px"] = np.abs(a["deformations"] - b["deformations"]).max() * (N - 1)
        out[f"{tag}_spread_energy"] = np.max(np.abs(a["energies"] - b["energies"]) / np.abs(a["energies"]))
    return out


def main():
    names = [a for a in sys.argv[1:] if not a.startswith("--")] or ["cfg1", "cfg2_256", "cfg2_512"]
    for name in names:
        frames, (l0, L), r0, its = config(name)
        fn = os.path.join(HERE, f"{name if name != 'cfg1' else 'cfg1_256'}.npz")
        if "--sensitivity-only" in sys.argv:
            with np.load(fn) as z:
                out = {k: z[k] for k in z.files}
            out.update(sensitivity(frames, l0, L, r0, its))
            np.savez_compressed(fn, **out)
            print("updated", fn, {k: float(out[k]) for k in out if "spread" in k}, flush=True)
            continue
        with tempfile.TemporaryDirectory() as wd:
            out = run_reference(frames, l0, L, r0, its, wd)
        out.update(run_oracle_rounded(frames, l0, L, r0, its))
        out.update(sensitivity(frames, l0, L, r0, its))
        out.update(levels=np.array([l0, L, r0, its]))
        np.savez_compressed(fn, **out)
        print("wrote", fn, os.path.getsize(fn), "bytes", flush=True)


if __name__ == "__main__":
    main()
