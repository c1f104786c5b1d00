"""Write the golden fixtures of matchSeries' optional outputs from the reference itself:
  * inverse.npz: qc::approxInverseDeformation (oracle/_ref/ref_harness "inverse") of smooth, noisy
    and folding displacements on 8x8, 32x32 and 64x64 grids;
  * e2e64_inv.npz: the reference program on the 4-frame 64x64 series of e2e64.npz with
    "calcInverseDeformation 1" (its <i>-r-inv/ solves among the per-level records);
  * e2e64_named_<a|b>.npz: the same series with saveNamedDeformedTemplates (plain; extended with the
    mean and nearest-neighbour with reduceDeformations), saveNamedDeformations and
    averageSaveIncrement 2 / 1: every extra file of both stages;
  * e2e128_to64 / e2e_crop64 / e2e_resize64.npz: the program on frames larger than the grid, cropped
    (cropInput) and resampled (resizeInput); resize.npz: ScalarArray<QC_2D>::resampleFrom alone;
  * e2e64_dm3.npz: the series as DM3 files, deformed templates written back into DM3 containers;
  * e2e64_alt / e2e64_alt_energy.npz: altStartLevel 5 (chosen by overlap / energyAsAltCriterion), with
    the chosen deformation-<iii> files and the <i>-alt/ records.
Runs in the build container only (needs oracle/_ref). Usage: python tests/golden/make_options.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_golden as G  # noqa: E402


def main():
    if not os.path.exists(G.HARNESS):
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for N in (8, 32, 64):
            rng = np.random.default_rng(500 + N)
            h = 1.0 / (N - 1)
            y, x = np.mgrid[0:N, 0:N] / (N - 1.0)
            for k, amp in enumerate((0.5, 3.0, 8.0)):
                phi = np.stack([amp * h * np.sin(2 * np.pi * y + 1) + 0.3 * amp * h * rng.standard_normal((N, N)),
                                amp * h * np.cos(2 * np.pi * x) + 2 * h])
                out[f"phi{N}_{k}"] = phi
                out[f"inv{N}_{k}"] = G.harness("inverse", N, [], [phi], tmp).reshape(2, N, N)
        np.savez_compressed(os.path.join(HERE, "inverse.npz"), **out)
        with np.load(os.path.join(HERE, "e2e64.npz")) as z:
            frames = z["frames"]
        e2e = G.run_e2e(frames, G.E2E_PAR + "calcInverseDeformation 1\n", os.path.join(tmp, "e2e_inv"))
        np.savez_compressed(os.path.join(HERE, "e2e64_inv.npz"), **e2e)
        # frames larger than the grid (128^2 at precisionLevel 6: normalised there, then restricted)
        # and frames cropped to it (80x72, cropInput at (5, 3))
        from msamd import synth
        big = synth.stem_series(128, 4, workload=0, spacing=12.0, sigma=3.0, jitter_step=0.3)
        np.savez_compressed(os.path.join(HERE, "e2e128_to64.npz"),
                            **G.run_e2e(big, G.E2E_PAR, os.path.join(tmp, "e2e128")))
        crop = np.ascontiguousarray(G.synth.stem_series(96, 4, workload=0, spacing=6.0, sigma=1.5,
                                                        jitter_step=0.15)[:, :72, :80])
        np.savez_compressed(os.path.join(HERE, "e2e_crop64.npz"), **G.run_e2e(
            crop, G.E2E_PAR + "cropInput 1\ncropStartX 5\ncropStartY 3\ndontResizeOrCropReference 0\n",
            os.path.join(tmp, "e2e_crop")))
        # resizeInput: ScalarArray<QC_2D>::resampleFrom through the harness, and the program on
        # 150x130 frames (resampled to 128^2, normalised there, restricted to 64^2)
        rz = {}
        rng = np.random.default_rng(77)
        for k, (w, h, N) in enumerate(((80, 72, 64), (100, 60, 128), (33, 47, 32))):
            src = rng.random((h, w))
            rz[f"src{k}"] = src
            rz[f"out{k}"] = G.harness("resize", N, [w, h], [src], tmp).reshape(N, N)
        np.savez_compressed(os.path.join(HERE, "resize.npz"), **rz)
        rs = np.ascontiguousarray(G.synth.stem_series(160, 4, workload=0, spacing=9.0, sigma=2.2,
                                                      jitter_step=0.2)[:, :130, :150])
        np.savez_compressed(os.path.join(HERE, "e2e_resize64.npz"), **G.run_e2e(
            rs, G.E2E_PAR + "resizeInput 1\ndontResizeOrCropReference 0\n", os.path.join(tmp, "e2e_resize")))
        # altStartLevel: a second pairwise solve from level 5, chosen by overlap (and by energy)
        for tag, extra in (("alt", "altStartLevel 5\n"), ("alt_energy", "altStartLevel 5\nenergyAsAltCriterion 1\n")):
            wd = os.path.join(tmp, tag)
            e2e = G.run_e2e(frames, G.E2E_PAR + extra, wd)
            for st in (1, 2):
                sd = os.path.join(wd, "out", f"stage{st}")
                for fn in sorted(os.listdir(sd)):
                    if fn.startswith("deformation-") and fn.endswith("_0.dat.bz2"):
                        e2e[f"s{st}_chosen_{fn[len('deformation-'):len('deformation-') + 3]}"] = np.stack(
                            [G.read_p9(os.path.join(sd, fn.replace("_0.dat", f"_{c}.dat"))) for c in range(2)])
                e2e[f"s{st}_defNorms"] = np.loadtxt(os.path.join(sd, "defNorms.txt"), usecols=(0, 1)) \
                    if os.path.exists(os.path.join(sd, "defNorms.txt")) else np.zeros((0, 2))
            np.savez_compressed(os.path.join(HERE, f"e2e64_{tag}.npz"), **e2e)
        # DM3 frames (float32 data) with the deformed templates written back into DM3 containers
        np.savez_compressed(os.path.join(HERE, "e2e64_dm3.npz"), **run_dm(frames, os.path.join(tmp, "dm3")))
        for tag, make_par in NAMED.items():
            np.savez_compressed(os.path.join(HERE, f"e2e64_named_{tag}.npz"),
                                **run_named(frames, make_par(G.E2E_PAR), os.path.join(tmp, f"named_{tag}")))
    print("wrote inverse.npz, e2e64_inv.npz and e2e64_named_*.npz")


NAMED = {
    "a": lambda par: par + "saveNamedDeformations 1\nsaveNamedDeformedTemplates 1\naverageSaveIncrement 2\n",
    "b": lambda par: par.replace("reduceDeformations 0", "reduceDeformations 1") + (
        "saveNamedDeformedTemplates 1\nsaveNamedDeformedTemplatesExtendedWithMean 1\n"
        "saveNamedDeformedTemplatesUsingNearestNeighborInterpolation 1\ndeformForAverageExtendedWithMean 1\n"
        "averageSaveIncrement 1\n"),
}


def run_dm(frames, wd):
    """The reference program on the frames stored as DM3 files (float32 image data, as
    tests/golden/make_formats.py writes them), with saveNamedDeformedTemplates and
    saveNamedDeformedDMXTemplatesAsDMX: the input files, every DM3 file it writes and the
    stage-1 deformations."""
    import make_formats as F
    os.makedirs(os.path.join(wd, "in"))
    out = {}
    for i, f in enumerate(frames):
        b = F.dm_file(3, [("ImageList", [("", [("ImageData", F.dm_image(f.astype(np.float32), 6))])])])
        with open(os.path.join(wd, "in", f"frame_{i:02d}.dm3"), "wb") as fh:
            fh.write(b)
        out[f"in/frame_{i:02d}.dm3"] = np.frombuffer(b, dtype=np.uint8)
    par = (G.E2E_PAR.replace("in/frame_%02d.dat.bz2", "in/frame_%02d.dm3")
           + "saveNamedDeformedTemplates 1\nsaveNamedDeformedDMXTemplatesAsDMX 1\n")
    with open(os.path.join(wd, "run.par"), "w") as fh:
        fh.write(par)
    env = dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON")
    subprocess.check_call([G.MATCHSERIES, "run.par"], cwd=wd, env=env, stdout=subprocess.DEVNULL,
                          stderr=subprocess.DEVNULL)
    out["par"] = np.frombuffer(par.encode(), dtype=np.uint8)
    for st in (1, 2):
        sd = os.path.join(wd, "out", f"stage{st}")
        for fn in sorted(os.listdir(sd)):
            if fn.endswith(".dm3"):
                out[f"s{st}/{fn}"] = np.fromfile(os.path.join(sd, fn), dtype=np.uint8)
    out["s1_deformations"] = np.stack([np.stack([G.read_p9(os.path.join(
        wd, "out", "stage1", f"{i}" if i == 0 else f"{i}-r", f"deformation_06_{c}.dat.bz2")) for c in range(2)])
        for i in range(3)])
    return out


def run_named(frames, par, wd):
    """The reference program with the extra options; every file of its stage directories that the
    plain run does not write (named deformed templates and deformations, intermediate
    reconstructions), keyed "s<stage>/<file name>"."""
    os.makedirs(os.path.join(wd, "in"))
    for i, f in enumerate(frames):
        G.write_p9(os.path.join(wd, "in", f"frame_{i:02d}.dat.bz2"), f)
    with open(os.path.join(wd, "run.par"), "w") as f:
        f.write(par)
    env = dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON")
    subprocess.check_call([G.MATCHSERIES, "run.par"], cwd=wd, env=env, stdout=subprocess.DEVNULL,
                          stderr=subprocess.DEVNULL)
    out = dict(frames=frames, par=np.frombuffer(par.encode(), dtype=np.uint8))
    for st in (1, 2):
        sd = os.path.join(wd, "out", f"stage{st}")
        for fn in sorted(os.listdir(sd)):
            path = os.path.join(sd, fn)
            if fn.startswith("frame_") or "_00" in fn:
                if fn.endswith(".png"):
                    out[f"s{st}/{fn}"] = G.read_png_gray8(path)
                elif fn.endswith((".q2bz", ".bz2")):
                    out[f"s{st}/{fn}"] = G.read_p9(path)
    return out


if __name__ == "__main__":
    main()
