"""Write tests/golden/actions64.npz: the reference program's ACTION 2 (analyzeDeformations) and ACTION 3
(applyDeformation) on the 4-frame 64x64 series of e2e64.npz.

The reference program (oracle/_ref/matchSeries) first matches the series (ACTION 0, two stages);
its saved deformations are recorded, then ACTION 2 with saveDeformedTemplates and
saveXDerivativeOfDef (its stderr report, defTempl_<iii>.dat.bz2 and .png, def_<iii>.q2bz) and
ACTION 3 on the sum of stage 1's pairwise deformations (numDeformations 3,
deformationBaseNamePattern out/stage1/%d/deformation_06), once extended with the mean and once
with nearest-neighbour interpolation (<basename>.q2bz). Runs in the build container only (needs
oracle/_ref). Usage: python tests/golden/make_actions.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_golden as G  # noqa: E402
from pngread import read_png_gray8  # noqa: E402

A3_PAR = "numDeformations 3\ndeformationBaseNamePattern out/stage1/%d/deformation_06\n"


def run(wd, par, action):
    with open(os.path.join(wd, "run.par"), "w") as f:
        f.write(par)
    env = dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON")
    r = subprocess.run([G.MATCHSERIES, "run.par", str(action)], cwd=wd, env=env, capture_output=True, text=True,
                       check=True)
    return r.stderr


def main():
    if not os.path.exists(G.MATCHSERIES):
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    with np.load(os.path.join(HERE, "e2e64.npz")) as z:
        frames = z["frames"]
        par = bytes(z["par"]).decode()
    out = dict(frames=frames, par=np.frombuffer(par.encode(), dtype=np.uint8))
    with tempfile.TemporaryDirectory() as wd:
        os.makedirs(os.path.join(wd, "in"))
        for i, f in enumerate(frames):
            G.write_p9(os.path.join(wd, "in", f"frame_{i:02d}.dat.bz2"), f)
        run(wd, par, 0)
        od = os.path.join(wd, "out")
        out["s1_median"] = G.read_p9(os.path.join(od, "stage1", "median.q2bz"))
        for st, subs in ((1, ["0", "1", "2", "1-r", "2-r"]), (2, ["0", "1-r", "2-r", "3-r"])):
            for sub in subs:
                out[f"s{st}_def_{sub.replace('-', '_')}"] = np.stack(
                    [G.read_p9(os.path.join(od, f"stage{st}", sub, f"deformation_06_{c}.dat.bz2")) for c in range(2)])
        # ACTION 2
        a2 = "saveDeformedTemplates 1\nsaveXDerivativeOfDef 1\n"
        err = run(wd, par + a2, 2)
        out["a2_par"] = np.frombuffer(a2.encode(), dtype=np.uint8)
        out["a2_report"] = np.frombuffer("\n".join(
            ln for ln in err.splitlines() if ln.startswith(("Translation", "average", "ratio"))).encode(), dtype=np.uint8)
        for st, n in ((1, 3), (2, 4)):
            sd = os.path.join(od, f"stage{st}")
            for i in range(n):
                out[f"a2_s{st}_defTempl_{i}"] = G.read_p9(os.path.join(sd, f"defTempl_{i:03d}.dat.bz2"))
                out[f"a2_s{st}_xderiv_{i}"] = G.read_p9(os.path.join(sd, f"def_{i:03d}.q2bz"))
                png = [f for f in os.listdir(sd) if f.startswith(f"defTempl_{i:03d}-") and f.endswith(".png")]
                assert len(png) == 1, png
                out[f"a2_s{st}_png_name_{i}"] = np.frombuffer(png[0].encode(), dtype=np.uint8)
                out[f"a2_s{st}_png_{i}"] = read_png_gray8(os.path.join(sd, png[0]))
        # ACTION 3, twice
        for tag, extra in (("mean", "saveNamedDeformedTemplatesExtendedWithMean 1\n"),
                           ("nn", "saveNamedDeformedTemplatesUsingNearestNeighborInterpolation 1\n")):
            a3 = A3_PAR + extra
            run(wd, par + a3, 3)
            out[f"a3_{tag}_par"] = np.frombuffer(a3.encode(), dtype=np.uint8)
            for st, n in ((1, 3), (2, 4)):
                sd = os.path.join(od, f"stage{st}")
                names = sorted(f for f in os.listdir(sd) if f.startswith("frame_") and f.endswith(".q2bz"))
                assert len(names) == n, names
                for i, fn in enumerate(names):
                    assert fn == f"frame_{i:02d}.dat.q2bz", fn
                    out[f"a3_{tag}_s{st}_{i}"] = G.read_p9(os.path.join(sd, fn))
                    os.remove(os.path.join(sd, fn))
    np.savez_compressed(os.path.join(HERE, "actions64.npz"), **out)
    print("wrote actions64.npz with", len(out), "arrays")


if __name__ == "__main__":
    main()
