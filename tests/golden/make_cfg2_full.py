"""Write tests/golden/cfg2_full.npz: the reference program (oracle/_ref/matchSeries) on the whole
benchmark series -- 32 STEM frames of 1024x1024 (msamd.synth.GPU_SERIES), stage 1 of matchSeries with
deformationModel 22, levels 7->10, refinement 9->10, 100 iterations per level (stopEpsilon -1),
lambda 10 -- the configuration bench.py times.

Stored (the full fields would be 500 MB): the 31 final energies (energies.txt, 6 significant digits),
every template's deformation used for the average sampled on a 64 x 64 lattice of nodes (every 16th
node from node 8, in pixels, float32), the median sampled alike, and the reference's wall-clock time
on this machine's one core.

The run takes about 2.5 hours on one core. Runs in the build container only (needs oracle/_ref).
Usage: python tests/golden/make_cfg2_full.py [--from-dir DIR]
  --from-dir: take the output tree of an earlier run of the same parameter file in DIR (DIR/out,
              DIR/ref.log) instead of running the reference again.
"""
import argparse
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "match-series_amd"))
import make_configs as MC  # noqa: E402
import make_golden as G  # noqa: E402
from msamd import synth  # noqa: E402

SAMPLE = slice(8, 1024, 16)


def extract(wd, n):
    od = os.path.join(wd, "out", "stage1")
    out = {}
    with open(os.path.join(od, "energies.txt")) as f:
        out["ref_energies"] = np.array([float(ln.split()[1]) for ln in f if ln.strip()])
    defs = []
    for i in range(n):
        sub = "0" if i == 0 else f"{i}-r"
        d = np.stack([G.read_p9(os.path.join(od, sub, f"deformation_10_{c}.dat.bz2")) for c in range(2)])
        defs.append((d[:, SAMPLE, SAMPLE] * 1023).astype(np.float32))
    out["ref_defs_px_sampled"] = np.stack(defs)
    out["ref_median_sampled"] = G.read_p9(os.path.join(od, "median.q2bz"))[SAMPLE, SAMPLE].astype(np.float32)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--from-dir")
    args = ap.parse_args()
    g = synth.GPU_SERIES
    par = MC.PAR.format(nf=g["frames"], its=100, l0=7, L=10, r0=9)
    if args.from_dir:
        wd = args.from_dir
        with open(os.path.join(wd, "run.par")) as f:
            assert f.read() == par, "the directory holds a run of another parameter file"
        secs = None
        log = open(os.path.join(wd, "ref.log")).read()
        m = re.search(r"Wall clock time = (\d+)h *(\d+)m *([\d.]+)s", log) or re.search(r"Wall clock time = (\d+)m *([\d.]+)s", log)
        if m:
            v = [float(x) for x in m.groups()]
            secs = v[0] * 3600 + v[1] * 60 + v[2] if len(v) == 3 else v[0] * 60 + v[1]
        out = extract(wd, g["frames"] - 1)
    else:
        if not os.path.exists(G.MATCHSERIES):
            sys.exit("oracle/_ref is not built (make -C oracle ref)")
        fr = synth.stem_series(g["N"], g["frames"], workload=g["workload"], spacing=g["spacing"], sigma=g["sigma"],
                               jitter_step=g["jitter_step"])
        with tempfile.TemporaryDirectory() as wd:
            os.makedirs(os.path.join(wd, "in"))
            for i, f in enumerate(fr):
                G.write_p9(os.path.join(wd, "in", f"frame_{i:02d}.dat.bz2"), f)
            with open(os.path.join(wd, "run.par"), "w") as f:
                f.write(par)
            t0 = time.perf_counter()
            subprocess.run([G.MATCHSERIES, "run.par"], cwd=wd, env=dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON"),
                           capture_output=True, check=True)
            secs = time.perf_counter() - t0
            out = extract(wd, g["frames"] - 1)
    out["par"] = np.frombuffer(par.encode(), dtype=np.uint8)
    out["ref_seconds_one_core"] = np.array(np.nan if secs is None else secs)
    np.savez_compressed(os.path.join(HERE, "cfg2_full.npz"), **out)
    print("wrote cfg2_full.npz:", {k: v.shape for k, v in out.items()}, "seconds", secs)


if __name__ == "__main__":
    main()
