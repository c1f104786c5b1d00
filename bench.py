"""Benchmark: registration + reconstruction of a 32-frame 1024x1024 STEM series on one MI355X.

One step = stage 1 of matchSeries (deformationModel 22, SSD + Dirichlet regulariser) on the whole
series, inputs already resident in HBM: 31 pairwise multilevel solves (levels 7-10, batched), the
accumulate-and-refine chain against the first frame (levels 9-10), and the series mean / median.
100 gradient-flow iterations per level (stopEpsilon -1: fixed schedule), lambda 10. Fields FP32,
energies / slopes / step decisions FP64.

Prints one JSON line with frames/s, the roofline of the dominant kernel and a CPU baseline.

Roofline: the energy pass (k_energy_f32) at the workload's dominant shape, one pairwise batch of 8
pairs at 1024^2 evaluating 3 step sizes along a real CG direction (the state two iterations of a
level solve leave), timed with HIP events around `reps` back-to-back launches on the context's
stream. Algorithmic bytes: an energy evaluation reads the deformation (2 fields), the direction (2),
reference and template, 6 FP32 fields of N^2; one launch performs 3 evaluations of 8 pairs, so
3 * 6 * 4 B * 8 * 1024^2 = 576 MiB per launch. The pass reads its inputs once for all its step sizes,
which the measured `traffic` shows (rocprofv3 FETCH_SIZE + WRITE_SIZE of the same launch, calibrated
on a plain copy, from profiles/k_energy_f32_8pairs_1024_traffic.json). `--roofline-only` runs just this measurement (for rocprofv3).

Whole step: the iterations and energy evaluations every solve of the step executed (counted by the
library, weighted by each level's node count) priced with the path's byte model, per node and
iteration 4 B x (73 1/3 + 6 n_E) (DE 6 field passes, H1 V-cycle 49 1/3, NCG 12, update 6, and 6 per
energy evaluation), plus the series average (6 passes per frame); the fraction of 8 TB/s it reaches.

FP64 line: one more step of the same series with FP64 fields (the reference's arithmetic), its time
and its gap to the FP32 result (largest deformation difference, largest relative energy difference).

Seismic line (matchSeismicSeries, the x-shift registration): 24 synthetic gathers of 2048 time
samples x 512 traces (msamd.synth.seismic_gather: 60 interfaces, 40 Hz Ricker at 0.5 ms, hyperbolic
moveout, static shifts), resampled onto the dyadic 1024^2 grid the reference's registration object
solves on (its default dyadicDepth for 2048 samples), stage 1 at levels 7-10, refinement 9-10, at
most 100 iterations per level, lambda 5, yRegFactor 1, FP32 fields. The reference's
full-resolution step (Cholesky H1 metric) is not part of this build.

Batched uneven series (msg_match_series_batch): 16 independent series, frame sizes drawn from
{256, 512, 1024, 2048}^2, 10-40 templates each, lattice spacing 10-20 px (the STEM model, frames made
on the GPU by msamd.warmstart.stem_frames_gpu), stage 1 at levels L-3..L, refinement L-1..L, 100
iterations per level, lambda 10, FP32 fields; the Armijo candidates per energy pass for each frame
size from msamd/tuning.json (msamd.autotune). One call runs all 16: series of one size share contexts
(pairwise solves batched across series, chains in lockstep), sizes run concurrently.

Warm-start network line: the BF16 U-Net (4 scales, 32 base channels, 1.95 M parameters) of the
learned warm start, training steps (forward, loss, backward, Adam) on 8 pairs of 256^2 and forward
passes, timed over 20 each; FLOPs counted from the architecture (2 per multiply-add of the 3x3
convolutions; a training step is the forward, the data gradient of every layer but the first and the
weight gradient), against MI355X's dense BF16 peak of 2.5 PFLOP/s. `--unet-only` runs just this.

CPU baseline: the reference's own level solve (oracle/_ref/ref_harness, compiled from the reference
sources) timed for a few iterations at 1024^2 on this host, extrapolated to the same schedule (cost
proportional to N^2).
"""
import argparse
import collections
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "match-series_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def series_params(gpu, levels=(7, 10), refine=(9, 10), its=100, lam=10.0, spec=0):
    L = levels[1]
    return gpu.SeriesParams(start_level=levels[0], stop_level=levels[1], refine_start_level=refine[0],
                            refine_stop_level=refine[1], precision_level=L, lambda_=lam, lambda_factor=-1.0,
                            max_iter=its, stop_eps=-1.0, saturation_pct=0.15, normalize_min_to_zero=1,
                            dont_normalize=0, resample_prolong=1, dont_accumulate=0, spec=spec)


def cpu_baseline(N, its_sample, schedule, frames, cores=1):
    """Reference level solve at N x N for its_sample iterations on one host core, extrapolated; with
    cores > 1 also the throughput with that many concurrent copies."""
    from msamd import synth
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    kind = "reference"
    s = synth.stem_series(N, 2, workload=1, spacing=16.0, sigma=3.0, jitter_step=0.4)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    R, T = O.normalize(s[0], 0.15, True), O.normalize(s[1], 0.15, True)
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        np.concatenate([R.ravel(), T.ravel(), np.zeros(2 * N * N)]).astype("<f8").tofile(fin)
        if os.path.exists(harness):
            t0 = time.perf_counter()
            subprocess.check_call([harness, "solve", str(N), "10", str(its_sample), "-1", fin, fout],
                                  stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            t = time.perf_counter() - t0
        else:  # the oracle's restatement of the same loop
            kind = "port"
            t0 = time.perf_counter()
            O.level_solve(R, T, np.zeros((2, N, N)), 10.0, its_sample, -1.0)
            t = time.perf_counter() - t0
        # the same sample on every host core at once (independent processes): the pairwise solves
        # of a series are independent, the accumulate-and-refine chain is not
        par = None
        if os.path.exists(harness) and cores > 1:
            t0 = time.perf_counter()
            procs = [subprocess.Popen([harness, "solve", str(N), "10", str(its_sample), "-1", fin, fout + f".{k}"],
                                      stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL) for k in range(cores)]
            rcs = [q.wait() for q in procs]
            if all(rc == 0 for rc in rcs):
                par = time.perf_counter() - t0
    t_iter = t / its_sample  # includes the initial energy and setup: slightly pessimistic for the CPU
    L = int(np.log2(N))
    (l0, l1), (r0, r1), its = schedule
    per_pair = sum(its * t_iter * 4.0 ** (l - L) for l in range(l0, l1 + 1))
    per_refine = sum(its * t_iter * 4.0 ** (l - L) for l in range(r0, r1 + 1))
    total = (frames - 1) * per_pair + (frames - 2) * per_refine
    what = "oracle/_ref/ref_harness solve" if kind == "reference" else "the oracle"
    one = dict(value=frames / total, unit="frames/s", cores=1, kind=kind,
               sample=f"{its_sample} gradient-flow iterations of one {N}x{N} pair (level {L}) in {what} "
                      f"= {t_iter:.3f} s/iteration, extrapolated to the step's schedule with cost ~ N^2")
    allc = None
    if par is not None:
        t_par = par / its_sample  # wall time per iteration with every core busy
        per_pair_p = sum(its * t_par * 4.0 ** (l - L) for l in range(l0, l1 + 1))
        total_p = (frames - 1) * per_pair_p / cores + (frames - 2) * per_refine
        allc = dict(value=frames / total_p, unit="frames/s", cores=cores, kind=kind,
                    sample=f"{cores} concurrent copies of the same sample = {t_par:.3f} s/iteration each; pairwise "
                           f"solves spread over {cores} cores, the chain on one")
    return one, allc


def unit_range(img):
    lo, hi = float(img.min()), float(img.max())
    return (img - lo) / (hi - lo if hi > lo else 1.0)


def energy_roofline(gpu, frames, defs=None, batch=8, reps=100):
    """Mean launch time and algorithmic bytes of the energy pass over `batch` pairs (frames k, k+1)."""
    N = frames.shape[-1]
    L = int(np.log2(N))
    ctx = gpu.Context(L, batch, gpu.F32)
    try:
        for p in range(batch):
            ctx.set_reference(p, unit_range(frames[p]))
            ctx.set_template(p, unit_range(frames[p + 1]))
            if defs is not None:
                ctx.set_deformation(p, defs[p])
        # a real deformation and CG direction: 2 iterations from the series' deformations, or 30 from 0
        ctx.solve_level(L, 10.0, 2 if defs is not None else 30, -1.0)
        ms, nbytes = ctx.time_energy_pass(L, 3, reps)
    finally:
        ctx.close()
    evals = 3
    algo = evals * nbytes  # nbytes: the 6 fields of one evaluation for every pair
    gbs = algo / (ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "k_energy_f32_8pairs_1024_traffic.json")
    if os.path.exists(tf) and batch == 8 and N == 1024:
        with open(tf) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")
    return {"bound": "hbm", "kernel": "msg::k_energy_f32", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "launches": reps,
            "avg_launch_us": round(1000.0 * ms, 2), "bytes_per_launch": algo, "evals_per_launch": evals,
            "bytes_per_evaluation": nbytes,
            "shape": f"{batch} pairs x {N}^2, {evals} step sizes, FP32 fields"}


def seismic_line(gpu, torch, count=24, size=1024, its=100):
    """Stage 1 of matchSeismicSeries' dyadic registration on `count` gathers resampled to size^2, with
    FP32 fields (timed) and FP64 fields (timed, and the FP32 result's gap to it)."""
    from msamd import synth
    L = int(np.log2(size))
    g = synth.seismic_series(count, workload=synth.SEISMIC_SERIES["workload"])
    p = series_params(gpu, (L - 3, L), (L - 1, L), its, lam=5.0)
    p.data_term = gpu.XSHIFT
    p.y_reg_factor = 1.0
    p.keep_start_transformation = 1
    p.stop_eps = 0.0
    p.saturation_pct = -1.0
    res = {}
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        raw = torch.tensor(g, dtype=dt, device="cuda")
        dev = torch.stack([gpu.resample_rect(raw[k].contiguous(), size) for k in range(count)]).contiguous()
        tpl, ref = dev[1:].contiguous(), dev[0].contiguous()
        gpu.match_series(tpl, ref, p)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = gpu.match_series(tpl, ref, p)
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0, out, gpu.series_last_counts())
        del raw, dev, tpl, ref
    t, out, cnt = res["f32"]
    t64, out64, _ = res["f64"]
    dd = (out["deformations"].double() - out64["deformations"]).abs().reshape(count - 1, -1).amax(dim=1).cpu().numpy()
    de = np.abs(out["energies"] - out64["energies"]) / np.abs(out64["energies"])
    return {"workload": f"matchSeismicSeries stage 1 (dyadic path): {count} gathers of 2048 samples x 512 traces "
                        f"resampled to {size}^2, levels {L - 3}-{L}, refine {L - 1}-{L}, <= {its} it/level, "
                        f"lambda 5, yRegFactor 1",
            "value": round(count / t, 3), "unit": "gathers/s", "ms_per_step": round(1000.0 * t, 2), "dtype": "f32",
            "iterations": cnt["iterations"], "energy_evaluations": cnt["evals"],
            "f64_ms_per_step": round(1000.0 * t64, 2),
            "fp32_vs_fp64_max_deformation_px": round(float(dd.max()) * (size - 1), 6),
            "fp32_vs_fp64_max_rel_energy": float(de.max()),
            "energies_finite": bool(np.isfinite(out["energies"]).all())}


def batch_line(gpu, torch, sequential=False, nseries=16, its=100):
    """One msg_match_series_batch call over nseries synthetic series of uneven sizes and lengths."""
    from msamd import autotune, synth
    from msamd import warmstart as W
    rng = np.random.default_rng(synth.BASE_SEED + 3)
    sizes = rng.choice([256, 512, 1024, 2048], nseries)
    counts = rng.integers(10, 41, nseries)
    spacing = rng.uniform(10.0, 20.0, nseries)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(synth.BASE_SEED + 3)
    jobs = []
    for k in range(nseries):
        N = int(sizes[k])
        L = int(np.log2(N))
        fr = W.stem_frames_gpu(N, int(counts[k]) + 1, gen, spacing=float(spacing[k]), sigma=2.5 * spacing[k] / 12.0,
                               jitter_step=0.3)
        spec, _ = autotune.tuned(N)  # the batch API sizes its own pairwise contexts
        p = series_params(gpu, (L - 3, L), (L - 1, L), its, spec=spec)
        jobs.append((fr[1:].contiguous(), fr[0].contiguous(), p))
    torch.cuda.synchronize()
    gpu.match_series_batch(jobs)  # warm-up: contexts and graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = gpu.match_series_batch(jobs)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    nt = int(counts.sum())
    line = {"workload": f"{nseries} series, sizes {sorted(collections.Counter(int(x) for x in sizes).items())}, "
                        f"{nt} templates in all (10-40 per series), levels L-3..L, refine L-1..L, {its} it/level, "
                        f"lambda 10, SSD, FP32 fields, one msg_match_series_batch call",
            "value": round(nt / t, 3), "unit": "frames/s", "ms_per_step": round(1000.0 * t, 2),
            "pixels_per_s": round(float(sum(int(counts[k]) * int(sizes[k]) ** 2 for k in range(nseries))) / t, 1),
            "tuning": {str(N): autotune.tuned(N) for N in (256, 512, 1024, 2048)},
            "energies_finite": bool(all(np.isfinite(o["energies"]).all() for o in outs))}
    if sequential:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for tpl, ref, p in jobs:
            gpu.match_series(tpl, ref, p)
        torch.cuda.synchronize()
        line["sequential_ms"] = round(1000.0 * (time.perf_counter() - t0), 2)
    del jobs, outs
    gpu.release_cache()
    return line


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def unet_line(gpu, torch, B=8, H=256, steps=20):
    """Training-step and forward throughput of the warm-start network."""
    from msamd import warmstart as W
    net = gpu.UNet(4, 32, seed=1)
    try:
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        x = torch.rand((B, 2, H, H), device="cuda", generator=g)
        t = torch.randn((B, 2, H, H), device="cuda", generator=g)
        net.train_step(x, t, 1e-3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            net.train_step(x, t, 1e-3)
        torch.cuda.synchronize()
        ms_train = 1000.0 * (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for _ in range(steps):
            net.forward(x)
        torch.cuda.synchronize()
        ms_fwd = 1000.0 * (time.perf_counter() - t0) / steps
        params = net.nparams
    finally:
        net.close()
    macs, first = W.unet_macs(4, 32, B, H, H)
    train_flops = 2.0 * (3 * macs - first)
    return {"workload": f"BF16 U-Net, 4 scales, 32 base channels, {params} parameters, batch {B} x 2 x {H}^2",
            "train_ms_per_step": round(ms_train, 3), "forward_ms": round(ms_fwd, 3),
            "train_tflops": round(train_flops / (ms_train * 1e-3) / 1e12, 1),
            "forward_tflops": round(2.0 * macs / (ms_fwd * 1e-3) / 1e12, 1),
            "peak_tflops": BF16_PEAK_TFLOPS,
            "train_frac": round(train_flops / (ms_train * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
            "forward_frac": round(2.0 * macs / (ms_fwd * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=32)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=100)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--spec", type=int, default=0)
    ap.add_argument("--mode", type=int, default=2, help="0 host-synchronised, 1 device-driven, 2 + hipGraph")
    ap.add_argument("--cpu-iters", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=0, help="pairwise solves per batch (0: automatic)")
    ap.add_argument("--classes", action="store_true", help="also time every kernel class in one extra step")
    ap.add_argument("--roofline-only", action="store_true", help="only the energy-pass roofline measurement")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 comparison step")
    ap.add_argument("--chain-shape", action="store_true",
                    help="only the chain's shape: one pair of the series at levels 9 and 10, 100 iterations each, "
                         "three times (for rocprofv3)")
    ap.add_argument("--no-seismic", action="store_true", help="skip the seismic (x-shift) line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-series line")
    ap.add_argument("--no-unet", action="store_true", help="skip the warm-start network line")
    ap.add_argument("--unet-only", action="store_true", help="only the warm-start network line (for rocprofv3)")
    ap.add_argument("--batch-vs-sequential", action="store_true",
                    help="also time the batched series one msg_match_series call at a time")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the series shards by frame across GPUs "
                                   "in a multi-GPU deployment, which is not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from msamd import gpu, synth

    torch.cuda.set_device(0)
    gpu.set_default_mode(args.mode)
    N, F = args.size, args.frames
    L = int(np.log2(N))
    levels, refine = (L - 3, L), (L - 1, L)
    dt = torch.float32 if args.precision == "f32" else torch.float64
    frames = synth.stem_series(N, F, workload=1, spacing=16.0, sigma=3.0, jitter_step=0.4)
    if args.unet_only:
        print(json.dumps({"unet": unet_line(gpu, torch)}), flush=True)
        return
    if args.chain_shape:
        ctx = gpu.Context(L, 1, gpu.F32)
        try:
            ctx.set_reference(0, unit_range(frames[0]))
            ctx.set_template(0, unit_range(frames[1]))
            res = []
            for rep_ in range(3):
                ctx.set_deformation_zero(0)
                for lv in (L - 1, L):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    e, it, _ = ctx.solve_level(lv, 10.0, args.iters, -1.0)
                    torch.cuda.synchronize()
                    res.append({"repetition": rep_, "level": lv, "iterations": int(it[0]),
                                "us_per_iteration": round(1e6 * (time.perf_counter() - t0) / max(int(it[0]), 1), 1)})
        finally:
            ctx.close()
        print(json.dumps({"chain_shape": res}), flush=True)
        return
    if args.roofline_only:
        r = energy_roofline(gpu, frames)
        # calibration for HBM byte counters: plain copies of 256 MiB (read 256 MiB, write 256 MiB each)
        a = torch.rand(1 << 26, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        r["calibration_copy_bytes"] = a.numel() * 4
        print(json.dumps({"roofline": r}), flush=True)
        return
    dev = torch.tensor(frames, dtype=dt, device="cuda")
    tpl, ref = dev[1:].contiguous(), dev[0].contiguous()
    p = series_params(gpu, levels, refine, args.iters, spec=args.spec)
    p.pairwise_batch = args.batch

    out = None
    for _ in range(args.warmup):
        out = gpu.match_series(tpl, ref, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = gpu.match_series(tpl, ref, p)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt_s = t1 - t0
    ms_step = 1000.0 * dt_s / args.steps
    value = F * args.steps / dt_s
    cnt = gpu.series_last_counts()  # the last timed step
    step_bytes = 4.0 * ((73.0 + 1.0 / 3.0) * cnt["node_iterations"] + 6.0 * cnt["node_evals"]) + 4.0 * N * N * 6 * F
    whole_step = {"iterations": cnt["iterations"], "energy_evaluations": cnt["evals"], "energy_passes": cnt["passes"],
                  "n_E": round(cnt["evals"] / max(cnt["iterations"], 1), 3),
                  "node_iterations": cnt["node_iterations"], "algorithmic_bytes": step_bytes,
                  "achieved_gbs": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
                  "frac": round(step_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}

    f64 = None
    if not args.no_f64 and args.precision == "f32":
        d64 = torch.tensor(frames, dtype=torch.float64, device="cuda")
        t64, r64 = d64[1:].contiguous(), d64[0].contiguous()
        gpu.match_series(t64, r64, p)  # warm-up (contexts, graphs)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        o64 = gpu.match_series(t64, r64, p)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        dd = (out["deformations"].double() - o64["deformations"]).abs().reshape(F - 1, -1).amax(dim=1).cpu().numpy()
        de = np.abs(out["energies"] - o64["energies"]) / np.abs(o64["energies"])
        f64 = {"ms_per_step": round(1000.0 * (tb - ta), 2), "value": round(F / (tb - ta), 3), "unit": "frames/s",
               "fp32_vs_fp64_max_deformation_px": round(float(dd.max()) * (N - 1), 6),
               "fp32_vs_fp64_median_deformation_px": round(float(np.median(dd)) * (N - 1), 6),
               "fp32_vs_fp64_max_rel_energy": float(de.max()),
               "frames_over_1e-3_px": int((dd * (N - 1) > 1e-3).sum())}
        del d64, t64, r64, o64

    roofline = energy_roofline(gpu, frames, out["deformations"][1:9].double().cpu().numpy())

    classes = None
    if args.classes:  # every kernel class bracketed by HIP events during one more step (graphs off meanwhile)
        gpu.profile_reset()
        gpu.profile_enable(gpu.KERNEL_CLASSES)
        gpu.match_series(tpl, ref, p)
        torch.cuda.synchronize()
        gpu.profile_enable(())
        classes = {c: gpu.profile_read(c) for c in gpu.KERNEL_CLASSES}

    # measured HBM ceiling: a plain device copy of 1 GiB (read + write), torch's copy kernel
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    roofline["copy_ceiling_gbs"] = round(10 * 2 * a.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
    del a, b

    seismic = None
    if not args.no_seismic:
        seismic = seismic_line(gpu, torch)
    batch = None
    if not args.no_batch:
        batch = batch_line(gpu, torch, args.batch_vs_sequential)
    unet = None
    if not args.no_unet:
        unet = unet_line(gpu, torch)

    cpu = cpu_all = None
    if not args.no_cpu:
        ncores = min(16, len(os.sched_getaffinity(0)))
        cpu, cpu_all = cpu_baseline(N, args.cpu_iters, (levels, refine, args.iters), F, ncores)

    finite = bool(torch.isfinite(out["median"]).all().item()) if out is not None else False
    line = {
        "metric": "frames/s at 1024^2 to fixed energy (plain vs NN warm-start); % HBM roofline",
        "value": round(value, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic (seeded STEM lattice series, Poisson noise, scan distortion)",
        "config": {"workload": f"plain matchSeries stage 1: {F} frames of {N}x{N}, levels {levels[0]}-{levels[1]}, "
                               f"refine {refine[0]}-{refine[1]}, {args.iters} it/level, lambda 10, SSD",
                   "frames": F, "size": N, "iterations_per_level": args.iters, "method": "plain",
                   "iteration_mode": args.mode, "armijo_candidates": args.spec or 3},
        "roofline": roofline,
        "whole_step": whole_step,
        "fp64": f64,
        "seismic": seismic,
        "batched_series": batch,
        "warm_start_network": unet,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "kernel_classes": classes,
        "energies_final_mean": float(np.mean(out["energies"])) if out is not None else None,
        "median_finite": finite,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
