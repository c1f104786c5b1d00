"""Write the golden fixtures of calcInverseDeformation from the reference itself:
  * inverse.npz: qc::approxInverseDeformation (oracle/_ref/ref_harness "inverse") of smooth, noisy
    and folding displacements on 8x8, 32x32 and 64x64 grids;
  * e2e64_inv.npz: the reference program on the 4-frame 64x64 series of e2e64.npz with
    "calcInverseDeformation 1" (its <i>-r-inv/ solves among the per-level records).
Runs in the build container only (needs oracle/_ref). Usage: python tests/golden/make_inverse.py
"""
import os
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
    print("wrote inverse.npz and e2e64_inv.npz")


if __name__ == "__main__":
    main()
