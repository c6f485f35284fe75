"""Benchmark: fused depth frames/s of the per-keyframe depth pipeline on one MI355X.

Workload (synthetic, seeded): 752x480 uint8 frames of a scene of 12 random textured planes and
boxes at 0.5-8 m, fx=fy=460, a 20-frame sinusoidal trajectory with 3-15 cm spacing; each
keyframe uses its 5 nearest frames as measurement frames; 64 planes over inverse depth
0.125-2.0; 8-path SGM; WTA + subpixel + filter; Gaussian x Beta fusion over all 20 frames.

One step = one pass of the whole 20-keyframe sequence (cost over 5 sources -> SGM -> filter ->
fusion per keyframe) with every frame already resident in HBM.  value = keyframes / second.

Prints one JSON line (rank 0).  Extra objects: "roofline" (the SGM pass, HBM-bound),
"stages" (per-stage ms and roofline fractions) and "cpu_baseline" (the float/OpenMP oracle on
a bounded sample of the same workload, on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
# HBM bytes of one SGM pass from rocprofv3 FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0) per dispatch,
# calibrated on a 1 GiB copy (profiles/depth_pipeline_752x480x64_s5_p8_f32.md)
SGM_PASS_TRAFFIC = {}
FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cost-type", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--form", default="plain", choices=["plain", "batched", "diff"],
                    help="plain: one keyframe at a time; batched: all keyframes' cost, SGM and "
                         "filter in one autotuned, hipGraph-captured batch, then the fusion steps")
    ap.add_argument("--median", type=int, default=0)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=12.0,
                    help="bound on the CPU baseline sample (0 disables it)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the attainable-bandwidth copy (0 disables it)")
    return ap.parse_args()


def cpu_baseline(scene, seconds: float) -> dict | None:
    """float/OpenMP oracle over as many keyframes of the workload as fit in ~`seconds`."""
    if seconds <= 0:
        return None
    import oracle
    from oracle.sequence import OracleSequence
    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    s = scene
    seq = OracleSequence(s["images"], s["poses"], s["sources"], s["scene"].K, s["planes"],
                         s["idep_min"], s["idep_max"], s["knobs"], s["paths"],
                         precision="f32")
    done = 0
    t0 = time.perf_counter()
    while done < len(s["images"]):
        seq.run_keyframe(done)
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "fused depth frames/s", "cores": threads,
            "kind": "port",
            "sample": f"first {done} of 20 keyframes of the same 752x480x64, 5-source, "
                      f"8-path workload (float instantiation of the oracle, OpenMP)"}


def attainable_copy(torch, mib: int) -> dict | None:
    """Device-to-device copy of `mib` MiB: the attainable streaming bandwidth on this box, and
    the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters of a profile."""
    if mib <= 0:
        return None
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda")
    src.fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 5
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return {"bytes": 2 * n, "ms": ms, "gbs": 2 * n / (ms * 1e-3) / 1e9}


def bench_batched(args, torch, seq, s, copy):
    """Batched form: autotune the batch, capture it into a hipGraph on a side stream, then
    time K steps of (graph replay for all keyframes' cost/SGM/filter + the fusion steps)."""
    nkf = len(s["images"])
    tuning = seq.batch.autotune(3)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        seq.batch.capture()
        torch.cuda.synchronize()

        def step():
            seq.reset()
            seq.batch.launch()
            for i in range(nkf):
                seq.fusion_step(i)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(side)
        for _ in range(args.steps):
            step()
        e1.record(side)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        # per-stage times of the whole batch, from an instrumented pass after the timed steps
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev[0].record(side)
        for st in range(3):
            seq.batch.run_stage(st)
            ev[st + 1].record(side)
        seq.reset()
        for i in range(nkf):
            seq.fusion_step(i)
        ev[4].record(side)
        torch.cuda.synchronize()
    ms = [ev[i].elapsed_time(ev[i + 1]) / nkf for i in range(4)]
    H, W, D, P = seq.H, seq.W, seq.D, seq.paths
    S = len(s["sources"][0])
    N, V = H * W, H * W * D
    bc = {"f32": 4, "f16": 2, "bf16": 2}[args.cost_type]
    sgm_bytes = P * V * (bc + 8)
    sgm_gbs = sgm_bytes / (ms[1] * 1e-3) / 1e9
    value = args.steps * nkf / elapsed
    stages = {
        "cost": {"ms": ms[0], "bound": "valu",
                 "achieved_tflops": V * S * 9 * 20 / (ms[0] * 1e-3) / 1e12,
                 "peak_tflops": FP32_VALU_PEAK_TFLOPS},
        "sgm": {"ms": ms[1], "bound": "hbm", "achieved_gbs": sgm_gbs,
                "frac": sgm_gbs / HBM_PEAK_GBS},
        "filter": {"ms": ms[2], "bound": "hbm"},
        "fusion": {"ms": ms[3], "bound": "hbm"},
    }
    base = cpu_baseline(s, args.cpu_baseline_seconds)
    out = {
        "metric": "fused depth frames/s @752x480x64 planes, plain vs FP16-batched; % HBM peak",
        "value": value,
        "unit": "fused depth frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.cost_type == "f32" else f"{args.cost_type}-storage/f32",
        "data": "synthetic (seeded textured-plane scene, 20-frame trajectory)",
        "config": {"workload": "752x480 frames, 64 planes (idep 0.125-2.0), 5 sources, "
                               "8-path SGM, filter, Gaussian x Beta fusion, 20-keyframe "
                               "sequence per step",
                   "form": "batched (20 keyframes per hipGraph replay)",
                   "cost_storage": args.cost_type, "keyframes_per_step": nkf,
                   "tuning": tuning},
        "roofline": {"bound": "hbm", "kernel": "sgm pass (8 path launches over the batch)",
                     "achieved": sgm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sgm_gbs / HBM_PEAK_GBS, "algorithmic_bytes": sgm_bytes,
                     "traffic": None},
        "stages": stages,
        "attainable_copy": copy,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


def bench_diff(args, torch):
    """Differentiable form: forward (cost, 8-path SGM, soft-argmin readout, L1 loss) + hand-written
    backward (dL/dI_ref, dL/d(H, t)) of one 640x480 keyframe with 4 sources and 64 planes, in
    the store and recompute modes."""
    from probabilistic_mapping_amd import diff, pipeline, synth
    s = synth.diff_scene(args.seed)
    H, W = s["ref"].shape
    D, S = s["planes"], len(s["srcs"])
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    ref = torch.from_numpy(s["ref"]).cuda().float()
    quads = [pipeline.make_quad(torch.from_numpy(im).cuda()) for im in s["srcs"]]
    homs = []
    for R, t in s["rel_poses"]:
        Hm, th = diff.se3_homography(s["K"], R, t, np.zeros(6))
        homs.append((Hm.astype(np.float32), th.astype(np.float32)))
    gt = torch.from_numpy(s["gt_depth"].astype(np.float32)).cuda()
    stream = torch.cuda.current_stream()
    modes = {}
    for mode in ("store", "recompute"):
        op = diff.SoftArgminDepth(W, H, D, S, s["knobs"], paths=8, temperature=s["temperature"],
                                  cost_scale=s["cost_scale"], recompute=(mode == "recompute"))
        for _ in range(args.warmup):
            op.forward(ref, quads, homs, s["idep_min"], ds, gt)
            op.backward(1.0)
        torch.cuda.synchronize()
        evs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record(stream)
            op.forward(ref, quads, homs, s["idep_min"], ds, gt)
            e[1].record(stream)
            op.backward(1.0)
            e[2].record(stream)
            evs.append(e)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        modes[mode] = {"keyframes_per_s": args.steps / elapsed,
                       "ms_per_step": elapsed / args.steps * 1e3,
                       "forward_ms": float(np.mean([a.elapsed_time(b) for a, b, _ in evs])),
                       "backward_ms": float(np.mean([b.elapsed_time(c) for _, b, c in evs])),
                       "device_bytes": op.memory_bytes}
        op.close()
    V = H * W * D
    # backward algorithmic HBM bytes, store mode: readout adjoint V*4 read + V*4 write; 8 SGM
    # adjoint sweeps of (L read, dL/dagg read, dL/dC read+write) = 16 B; cost adjoint reads
    # dL/dC and the counts (5 B)
    bwd_bytes = V * 8 + 8 * V * 16 + V * 5
    st = modes["store"]
    out = {
        "metric": "differentiable soft-argmin depth: forward + backward keyframes/s "
                  "@640x480x64 planes, 4 sources",
        "value": st["keyframes_per_s"],
        "unit": "keyframes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": st["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded textured-plane scene with ground-truth depth)",
        "config": {"workload": "640x480 reference, 4 measurement frames, 64 planes "
                               "(idep 0.2-1.0), 8-path SGM, soft-argmin T=0.05 over agg/255, "
                               "mean L1 loss, backward to I_ref and (H, t) per source",
                   "form": "diff"},
        "modes": modes,
        "backward_algorithmic_bytes_store": bwd_bytes,
        "backward_gbs_store": bwd_bytes / (st["backward_ms"] * 1e-3) / 1e9,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(keyframe streams would shard one stream per GPU)"}))
        sys.exit(2)
    import torch
    from probabilistic_mapping_amd import pipeline as pl
    from probabilistic_mapping_amd import synth
    from probabilistic_mapping_amd.pipeline import KeyframeSequence

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.form == "diff":
        return bench_diff(args, torch)
    s = synth.trajectory_scene(args.seed)
    nkf = len(s["images"])
    Seq = pl.BatchedSequence if args.form == "batched" else KeyframeSequence
    seq = Seq(s["images"], s["poses"], s["sources"], s["scene"].K, s["planes"],
              s["idep_min"], s["idep_max"], s["knobs"], s["paths"],
              cost_type=args.cost_type, median=bool(args.median))
    H, W, D, P = seq.H, seq.W, seq.D, seq.paths
    S = len(s["sources"][0])
    N = H * W
    V = N * D
    bc = {"f32": 4, "f16": 2, "bf16": 2}[args.cost_type]

    copy = attainable_copy(torch, args.copy_mib)

    if args.form == "batched":
        return bench_batched(args, torch, seq, s, copy)

    for _ in range(args.warmup):
        seq.run()
    torch.cuda.synchronize()

    # timed region: K full sequences; HIP events on the launch stream bracket each stage
    stream = torch.cuda.current_stream()
    ev = {k: [] for k in ("cost", "sgm", "median", "fusion")}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seq.reset()
        for i in range(nkf):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
            e[0].record(stream)
            pl.calc_cost(seq.images[i], [seq.quads[j] for j in seq.sources[i]], seq.homs[i],
                         seq.idep_min, seq.dep_sample, seq.D, seq.cost_type, out=seq.cost)
            e[1].record(stream)
            pl.sgm_filter_paths(seq.images[i], seq.cost, seq.knobs, seq.paths, seq.idep_min,
                                seq.dep_sample, workspace=seq.sgm_ws, agg=seq.agg,
                                depth=seq.dep_raw if seq.median else seq.dep)
            e[2].record(stream)
            if seq.median:
                pl.median3x3(seq.dep_raw, out=seq.dep)
            e[3].record(stream)
            seq.fusion_stage(i)
            e[4].record(stream)
            for k, (a, b) in zip(ev, zip(e[:-1], e[1:])):
                ev[k].append((a, b))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}

    frames = args.steps * nkf
    value = frames / elapsed
    # the SGM pass: all paths at once (read cost, write the path's L volume), then one pass
    # that reads the P volumes, writes the aggregate and the filtered depth.  The same bytes
    # as the reference's path-after-path read-modify-write (P V (b_c + 8)) plus its filter
    # pass (V 4 read, N 4 write), minus the filter's re-read of the aggregate.
    sgm_bytes = P * V * (bc + 8) + V * 4 + N * 4
    sgm_gbs = sgm_bytes / (ms["sgm"] * 1e-3) / 1e9
    cost_flop = V * S * 9 * 20                          # ~20 FLOP per warp term
    cost_bytes = V * bc + (S + 1) * N
    fus_bytes = 104 * N
    stages = {
        "cost": {"ms": ms["cost"], "bound": "valu", "achieved_tflops": cost_flop / (ms["cost"] * 1e-3) / 1e12,
                 "peak_tflops": FP32_VALU_PEAK_TFLOPS,
                 "hbm_gbs": cost_bytes / (ms["cost"] * 1e-3) / 1e9},
        "sgm": {"ms": ms["sgm"], "bound": "hbm", "achieved_gbs": sgm_gbs,
                "frac": sgm_gbs / HBM_PEAK_GBS, "includes": "filter (WTA + subpixel)",
                "compulsory_gbs": (V * (bc + 4) + N * 4) / (ms["sgm"] * 1e-3) / 1e9},
        "median": {"ms": ms["median"], "enabled": bool(args.median)},
        "fusion": {"ms": ms["fusion"], "bound": "hbm",
                   "achieved_gbs": fus_bytes / (ms["fusion"] * 1e-3) / 1e9},
    }
    base = cpu_baseline(s, args.cpu_baseline_seconds)
    out = {
        "metric": "fused depth frames/s @752x480x64 planes, plain vs FP16-batched; % HBM peak",
        "value": value,
        "unit": "fused depth frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.cost_type == "f32" else f"{args.cost_type}-storage/f32",
        "data": "synthetic (seeded textured-plane scene, 20-frame trajectory)",
        "config": {"workload": "752x480 frames, 64 planes (idep 0.125-2.0), 5 sources, "
                               "8-path SGM, filter, Gaussian x Beta fusion, 20-keyframe "
                               "sequence per step",
                   "form": "plain", "cost_storage": args.cost_type,
                   "median": bool(args.median), "keyframes_per_step": nkf},
        "roofline": {"bound": "hbm",
                     "kernel": "sgm pass (all paths in one launch + combine/filter)",
                     "achieved": sgm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sgm_gbs / HBM_PEAK_GBS,
                     "algorithmic_bytes": sgm_bytes,
                     "traffic": SGM_PASS_TRAFFIC.get((args.cost_type, P, H, W, D))},
        "stages": stages,
        "attainable_copy": copy,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
