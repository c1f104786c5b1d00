"""Per-frame-size tuning of the knobs that change speed but not results.

Two knobs of a series solve leave every result bitwise unchanged and only move its speed:
  * spec: the step sizes an Armijo energy pass evaluates at once (1..4; the search replays the
    sequential algorithm over them, so the accepted steps are the same);
  * pairwise_batch: how many pairwise solves share one batched context (each pair's arithmetic is
    independent of its batch).
For each frame size this times one series stage (9 frames, levels L-3..L, refinement L-1..L, a fixed
schedule of `its` iterations per level, lambda 10, SSD, FP32 fields) over a grid of both knobs and
keeps the fastest. The table (tuning.json next to this file) is what msamd.autotune.tuned() returns;
the multi-series benchmark reads it.

    python -m msamd.autotune [--sizes 256 512 1024 2048] [--its 30] [--out path]
"""
import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")


def tuned(N, table=None):
    """(spec, pairwise_batch) for frames of N x N: the table's entry for N, else the nearest size's,
    else the library defaults (0, 0)."""
    path = table or TABLE
    if not os.path.exists(path):
        return 0, 0
    with open(path) as f:
        t = json.load(f).get("sizes", {})
    if not t:
        return 0, 0
    key = min(t, key=lambda k: abs(int(k) - N))
    return int(t[key]["spec"]), int(t[key]["pairwise_batch"])


def tune_size(gpu, torch, N, frames=9, its=30, specs=(1, 2, 3, 4), batches=(0, 2, 4, 8)):
    from . import synth
    L = int(np.log2(N))
    sc = N / 1024.0
    fr = synth.stem_series(N, frames, workload=3, spacing=max(16.0 * sc, 4.0), sigma=max(3.0 * sc, 1.0),
                           jitter_step=0.4 * sc)
    t = torch.tensor(fr, dtype=torch.float32, device="cuda")
    tpl, ref = t[1:].contiguous(), t[0].contiguous()
    results = []
    for spec in specs:
        for b in batches:
            p = gpu.SeriesParams(start_level=L - 3, stop_level=L, refine_start_level=L - 1, refine_stop_level=L,
                                 precision_level=L, lambda_=10.0, lambda_factor=-1.0, max_iter=its, stop_eps=-1.0,
                                 saturation_pct=0.15, normalize_min_to_zero=1, dont_normalize=0, resample_prolong=1,
                                 dont_accumulate=0, spec=spec, pairwise_batch=b)
            gpu.match_series(tpl, ref, p)  # warm-up: contexts and graphs of this shape
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gpu.match_series(tpl, ref, p)
            torch.cuda.synchronize()
            results.append(dict(spec=spec, pairwise_batch=b, ms=1000.0 * (time.perf_counter() - t0)))
    best = min(results, key=lambda r: r["ms"])
    return best, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[256, 512, 1024, 2048])
    ap.add_argument("--its", type=int, default=30)
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args()
    import torch
    from . import gpu
    torch.cuda.set_device(0)
    table = {"workload": f"one stage of 9 frames per size, levels L-3..L, refine L-1..L, {args.its} it/level, FP32",
             "device": torch.cuda.get_device_name(0), "sizes": {}}
    for N in args.sizes:
        best, allr = tune_size(gpu, torch, N, its=args.its)
        table["sizes"][str(N)] = dict(spec=best["spec"], pairwise_batch=best["pairwise_batch"], ms=round(best["ms"], 2),
                                      tried=[{k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.items()}
                                             for r in allr])
        print(f"{N}: spec {best['spec']}, pairwise_batch {best['pairwise_batch']} ({best['ms']:.1f} ms)", flush=True)
        gpu.release_cache()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
