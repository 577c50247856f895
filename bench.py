"""Benchmark: 256^3 RI volumes/s from 60 x 1024^2 hologram pairs on one MI355X.

Workload (one "step"): one cell of 60 illumination angles, each a 1024^2 float32 sample hologram
plus background (seeded synthetic phantom, inputs resident in HBM before timing), sideband retrieval
to 256^2, DCT least-squares unwrapping, Rytov spectra, Ewald-cap gridding into a 256^3 complex64
k-space, inverse 3-D FFT and 10 positivity iterations -> one 256^3 complex RI volume.

value       = batched (fused, hipGraph-replayed) volumes/s
plain_value = the reference-structured plain pipeline on the same GPU (per angle, unfused,
              host-synchronous), reported beside it with the ratio
roofline    = the dominant kernel (largest share of the step) at its algorithmic bytes per launch
cpu_baseline= the FP64 host oracle on one whole volume of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "256³ RI vol/s (60×1024² holos): batched vs plain, same GPU, RI err ≤1e-5"
HBM_PEAK_GBS = 8000.0
A, H, V, ITERS = 60, 1024, 256, 10
# HBM bytes per launch of each volume pass at 256^3 (16 pencils per workgroup), measured with
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (separate runs), fetch scaled by 2.000 and write by
# 1.000 so that the copy yardstick reads as its known 128 MiB in + 128 MiB out. The profiler cannot run
# inside this process, so the figures are carried here; profiles/traffic_config2_pmc.md has the runs.
TRAFFIC_KIB = {0: 131209 + 131072, 1: 131157 + 131072, 2: 170176 + 131072}
PASS_NAMES = {0: "k_vol_x (inverse x-FFT + projection + forward x-FFT)",
              1: "k_vol_y (y-FFT)",
              2: "k_vol_zdc (z-FFT + data consistency + inverse z-FFT)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-config4", action="store_true", help="skip the 8-cell uneven batch")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(cells shard across GPUs without communication; see DESIGN.md).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import ctdm_amd

    dev = torch.device("cuda:0")
    log(f"step: synth {A} x {H}^2 hologram pairs")
    holo, bg, ang = ctdm_amd.synth_cell(2, A)
    hd = torch.from_numpy(holo).to(dev)
    bd = torch.from_numpy(bg).to(dev)
    optics = ctdm_amd.optics_for(V, H)
    ctx = ctdm_amd.Context(0)
    tune_report = "" if args.no_tune else ctx.autotune([V])
    if tune_report:
        log(tune_report.strip())

    ri_b = torch.empty((V, V, V), dtype=torch.complex64, device=dev)
    cell = dict(holo=hd, bg=bd, angles=ang, V=V, iters=ITERS, optics=optics, ri=ri_b)
    log("step: batched first call (eager + graph capture)")
    ctx.reconstruct([cell], mode=1, graph=True)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        ctx.reconstruct([cell], mode=1, graph=True)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"step: timing {args.steps} batched steps")
    e0.record()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reconstruct([cell], mode=1, graph=True)
    e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    ms = e0.elapsed_time(e1) / args.steps
    value = 1000.0 / ms

    # plain, reference-structured pipeline on the same inputs
    ri_p = torch.empty_like(ri_b)
    cellp = dict(holo=hd, bg=bd, angles=ang, V=V, iters=ITERS, optics=optics, ri=ri_p)
    plain_value = dn = None
    if args.plain_steps > 0:  # 0 skips the plain leg (short profiling runs)
        log("step: plain warm-up")
        ctx.reconstruct([cellp], mode=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            ctx.reconstruct([cellp], mode=0)
        torch.cuda.synchronize()
        ms_plain = (time.perf_counter() - t0) * 1e3 / args.plain_steps
        plain_value = 1000.0 / ms_plain
        dn = (ri_b - ri_p).abs().max().item()
        log(f"batched {ms:.3f} ms/vol ({value:.1f} vol/s), plain {ms_plain:.2f} ms/vol, max|dn| {dn:.2e}")

    # roofline of the dominant kernel: live HIP-event timing of each fused volume pass inside the
    # loop's own sequence (x, y, z+dc, y) on the stream the kernels run on
    per_vol = {0: ITERS, 1: 2 * ITERS + 1, 2: ITERS}  # launches per volume of each pass
    lms, lbytes = ctx.time_loop(V, 2 * ITERS)
    kernels = {}
    for k in (0, 1, 2):
        kernels[k] = dict(name=PASS_NAMES[k], ms=lms[k], bytes=lbytes[k], gbs=lbytes[k] / lms[k] / 1e6,
                          share=per_vol[k] * lms[k] / ms)
    cms, cbytes = ctx.time_pass(4, V, 0, 20)  # plain streaming copy of the same volume: the yardstick
    dom = max(kernels, key=lambda k: per_vol[k] * kernels[k]["ms"])
    d = kernels[dom]
    roofline = {"bound": "hbm", "achieved": round(d["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_KIB[dom] * 1024,
                "kernel": d["name"], "bytes_per_launch": int(d["bytes"]),
                "avg_launch_ms": round(d["ms"], 5), "share_of_step": round(d["share"], 3),
                "other_kernels": {PASS_NAMES[k]: {"GB/s": round(v["gbs"], 1), "avg_ms": round(v["ms"], 5)}
                                  for k, v in kernels.items() if k != dom},
                "copy_yardstick": {"GB/s": round(cbytes / cms / 1e6, 1), "avg_ms": round(cms, 5),
                                   "bytes_per_launch": int(cbytes)}}

    cpu = None
    if not args.no_cpu:
        from oracle import binding as orc

        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        threads = orc.threads(threads)
        crop = cell["crop_used"].astuple()
        log(f"step: cpu baseline (FP64 oracle, {threads} threads, whole volumes for about 10 s)")
        t0 = time.perf_counter()
        nvol = 0
        while nvol < 8 and (nvol == 0 or time.perf_counter() - t0 < 10.0):
            ref = orc.reconstruct_cell(holo, bg, ang, V, crop, optics.wavelength, optics.n_med, optics.dx, ITERS)
            nvol += 1
        tc = (time.perf_counter() - t0) / nvol
        err = float(np.abs(ri_b.cpu().numpy() - ref).max())
        cpu = {"value": round(1.0 / tc, 4), "unit": "volumes/s", "cores": threads, "kind": "port",
               "sample": f"{nvol} full volumes of the same workload (60 x 1024^2 pairs -> 256^3, "
                         f"10 iterations) through the FP64 oracle, {tc:.1f} s each",
               "ri_max_err_vs_gpu": err}
        log(f"cpu baseline {tc:.1f} s/volume, GPU-vs-oracle max|dn| {err:.2e}")

    config4 = None if args.no_config4 else run_config4(ctx, args)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "volumes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipsoid cell with 5 nucleoli, 60-angle NA 1.0 cone)",
        "config": {"workload": "one cell: 60 x 1024^2 f32 hologram pairs -> 256^3 complex64 RI, "
                               "10 positivity iterations", "angles": A, "hologram": H, "volume": V,
                   "iters": ITERS, "mode": "batched, hipGraph replay"},
        "plain_value": plain_value and round(plain_value, 3),
        "speedup_vs_plain": plain_value and round(value / plain_value, 2),
        "ri_max_diff_batched_vs_plain": dn, "wall_ms_per_step": round(wall, 4),
        "roofline": roofline, "cpu_baseline": cpu, "config4": config4,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def run_config4(ctx, args):
    """Eight cells, A = {30,45,60,60,90,90,120,120}, V cyclic over {128,192,256}, 1024^2 holograms,
    10 iterations: one captured graph for all cells (two lanes; one lane timed beside it) vs the plain
    per-cell pipeline."""
    import torch

    import ctdm_amd

    angles = [30, 45, 60, 60, 90, 90, 120, 120]
    sides = [128, 192, 256, 128, 192, 256, 128, 192]
    log("step: config4 synth (615 hologram pairs)")
    cells = []
    for i, (A_, V_) in enumerate(zip(angles, sides)):
        holo, bg, ang = ctdm_amd.synth_cell(40 + i, A_)
        cells.append(dict(holo=torch.from_numpy(holo).cuda(), bg=torch.from_numpy(bg).cuda(), angles=ang,
                          V=V_, iters=ITERS, optics=ctdm_amd.optics_for(V_, H),
                          ri=torch.empty((V_, V_, V_), dtype=torch.complex64, device="cuda")))
        del holo, bg
    if not args.no_tune:
        log(ctx.autotune(sorted(set(sides))).strip())
    ev = lambda: torch.cuda.Event(enable_timing=True)
    log("step: config4 batched graph")
    ctx.reconstruct(cells, mode=1, graph=True)
    ctx.reconstruct(cells, mode=1, graph=True)
    torch.cuda.synchronize()
    reps = 5
    e0, e1 = ev(), ev()
    e0.record()
    for _ in range(reps):
        ctx.reconstruct(cells, mode=1, graph=True)
    e1.record()
    torch.cuda.synchronize()
    ms_batch = e0.elapsed_time(e1) / reps
    outs = [c["ri"].clone() for c in cells]
    # the same sequence on one lane (cells strictly one after another)
    ctx.reconstruct(cells, mode=1, graph=True, lanes=1)
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    e0.record()
    for _ in range(reps):
        ctx.reconstruct(cells, mode=1, graph=True, lanes=1)
    e1.record()
    torch.cuda.synchronize()
    ms_one_lane = e0.elapsed_time(e1) / reps
    lane_diff = max((c["ri"] - o).abs().max().item() for c, o in zip(cells, outs))
    per_cell = []
    for c in cells:  # each cell alone, its own captured graph
        ctx.reconstruct([c], mode=1, graph=True)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        for _ in range(reps):
            ctx.reconstruct([c], mode=1, graph=True)
        e1.record()
        torch.cuda.synchronize()
        per_cell.append(round(e0.elapsed_time(e1) / reps, 3))
    log("step: config4 plain")
    plain_ms, dn = 0.0, 0.0
    for c, ob in zip(cells, outs):
        rp = torch.empty_like(ob)
        cp = dict(c, ri=rp)
        t0 = time.perf_counter()
        ctx.reconstruct([cp], mode=0)
        torch.cuda.synchronize()
        plain_ms += (time.perf_counter() - t0) * 1e3
        dn = max(dn, (rp - ob).abs().max().item())
    res = {"cells": [{"A": a, "V": v} for a, v in zip(angles, sides)],
           "batched_ms_all_cells": round(ms_batch, 3), "volumes_per_s": round(8000.0 / ms_batch, 2),
           "lanes": 2, "batched_ms_one_lane": round(ms_one_lane, 3), "ri_max_diff_two_vs_one_lane": lane_diff,
           "per_cell_latency_ms": per_cell, "plain_ms_all_cells": round(plain_ms, 2),
           "plain_volumes_per_s": round(8000.0 / plain_ms, 2), "speedup_vs_plain": round(plain_ms / ms_batch, 2),
           "ri_max_diff_batched_vs_plain": dn}
    log(f"config4: batched {ms_batch:.2f} ms for 8 cells ({res['volumes_per_s']} vol/s; one lane "
        f"{ms_one_lane:.2f} ms), plain {plain_ms:.1f} ms, "
        f"per-cell {per_cell}")
    for c in cells:
        del c["holo"], c["bg"]
    torch.cuda.empty_cache()
    return res


if __name__ == "__main__":
    sys.exit(main())
