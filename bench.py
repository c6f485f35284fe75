"""Benchmark: fused depth frames/s of the per-keyframe depth pipeline on one MI355X.

Workload (synthetic, seeded): 752x480 uint8 frames of a scene of 12 random textured planes and
boxes at 0.5-8 m, fx=fy=460, a 20-frame sinusoidal trajectory with 3-15 cm spacing; each
keyframe uses its 5 nearest frames as measurement frames; 64 planes over inverse depth
0.125-2.0; 8-path SGM; WTA + subpixel + filter; 3x3 median (the median/consistency filter,
--median 0 drops it); Gaussian x Beta fusion over all 20 frames.

One step = one pass of the whole 20-keyframe sequence (cost over 5 sources -> SGM -> filter ->
fusion per keyframe) with every frame already resident in HBM.  value = keyframes / second.

Forms: "plain" runs one keyframe at a time (FP32 cost, exact SGM storage by default); "batched"
runs the keyframes' cost, SGM and filter as autotuned batches of keyframes (grouped so that a
group's cost volumes stay in the 256 MB Infinity Cache: 4 groups of 5 here) and the fusion
steps, the whole step captured into one hipGraph in which each group's cost launch runs beside
the previous group's SGM pass, filter and fusion steps (--overlap 0: the groups' captured
batches one after another, then the fusion steps) (FP16 cost, compact SGM storage by default).  The default, "both", reports the
plain FP32 form as ``value`` and the FP16 batched form beside it, with their ratio.
``--form uneven`` times 16 independent keyframes of mixed size (320x240-1280x720, 2-8 sources,
32-128 planes) as one FP16 batch, beside the same keyframes one at a time; ``--form mapper``
times the 20-keyframe sequence through the C++ StereoMapper (initReference, update per source,
outputFusion; host images in, host depth out) with a C++ driver; ``--form diff`` times the
differentiable readout's forward + backward; ``--form driver`` times the reference node's
processImg loop, with and without the upload-once frame ring.

Prints one JSON line.  Extra objects: "roofline" (the SGM pass, HBM-bound: the bytes it streams
over its event time, against 8 TB/s, with its counter-measured traffic), "stages" (per-stage ms;
the VALU-bound cost stage as a fraction of the chip's VALU issue slots) and "cpu_baseline" (the
float/OpenMP oracle on a bounded sample of the same workload, built and timed on this host's
physical cores).
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
METRIC = "fused depth frames/s @752x480x64 planes, plain vs FP16-batched; % HBM peak"
WORKLOAD = ("752x480 frames, 64 planes (idep 0.125-2.0), 5 sources, 8-path SGM, filter, "
            "3x3 median (unless --median 0), Gaussian x Beta fusion, 20-keyframe sequence "
            "per step")
COST_BYTES = {"f32": 4, "f16": 2, "bf16": 2}
# Counter-measured figures per keyframe of this workload, from the rocprofv3 --pmc runs in
# COUNTER_PROFILE (keyed by form, cost storage, SGM storage):
#   * SGM_PASS_TRAFFIC: HBM bytes of the SGM pass (paths launch + combine/filter launch),
#     FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 in bytes (scales from a device copy of known size);
COUNTER_PROFILE = "profiles/counters_752x480x64_s5_p8.md"
SGM_PASS_TRAFFIC = {("plain", "f32", "exact"): 2.2033e9,
                    ("batched", "f16", "compact"): 1.1583e9}
#   * COST_VALU: the cost launch's SQ_INSTS_VALU (wave64 instructions) and the fraction of the
#     chip's VALU issue slots they fill over the launch's duration in the profiled run (at the
#     2.4 GHz peak clock).
#   * UNEVEN_SGM_TRAFFIC: HBM bytes of the uneven batch's SGM pass (its paths and combine/filter
#     launches) per replay, default keyframe set (UNEVEN_PROFILE).
UNEVEN_PROFILE = "profiles/uneven_batch_16kf_f16.md"
UNEVEN_SGM_TRAFFIC = 21.210e9
COST_VALU = {("plain", "f32", "exact"): (113.12e6, 0.826),
             ("batched", "f16", "compact"): (112.27e6, 0.936)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cost-type", default="f32", choices=["f32", "f16", "bf16"],
                    help="cost storage of the plain / batched form (both: of the plain form)")
    ap.add_argument("--form", default="both",
                    choices=["both", "plain", "batched", "uneven", "mapper", "diff", "driver"])
    ap.add_argument("--sgm-storage", default=None, choices=["exact", "compact", "sweep"],
                    help="SGM per-path storage (default: exact for f32 cost, compact otherwise; "
                         "sweep exists in the batched form only)")
    ap.add_argument("--median", type=int, default=1,
                    help="3x3 median over the filtered depth before fusion (the workload's "
                         "median/consistency filter; 0 disables it)")
    ap.add_argument("--cpu-baseline-seconds", type=float, default=12.0,
                    help="bound on the CPU baseline sample (0 disables it)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="size of the attainable-bandwidth copy (0 disables it)")
    ap.add_argument("--overlap", type=int, default=1,
                    help="batched form: 1 = each keyframe group's cost launch beside the "
                         "previous group's SGM pass and fusion steps, the whole step one "
                         "hipGraph; 0 = the groups' captured batches one after another")
    ap.add_argument("--inflight", type=int, default=2,
                    help="both/batched forms: also time this many independent 20-keyframe "
                         "sequences in flight on as many streams, each with its own buffers "
                         "and captured graph (reported beside the one-at-a-time figure; 1 "
                         "disables it)")
    ap.add_argument("--tuning", default=None,
                    help="batched form: launch shapes as JSON (pm_batch_tuning fields) instead "
                         "of running the autotuner")
    return ap.parse_args()


def default_storage(cost_type: str, storage: str | None) -> str:
    return storage or ("exact" if cost_type == "f32" else "compact")


def cpu_baseline(seed: int, seconds: float, median: bool) -> dict | None:
    """The float/OpenMP oracle over as many keyframes of the workload as fit in ~`seconds`,
    built with -march=native on this host and timed in a child process with OMP_NUM_THREADS =
    the physical cores this process may use (at most 16) and OMP_PROC_BIND=close."""
    if seconds <= 0:
        return None
    from oracle import cpu_baseline as cb
    return cb.run(seconds, seed, median, max_threads=16)


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


def reference_dataflow_bytes(seq, cost_type: str) -> int:
    """Bytes of the SGM pass with its filter in the reference's dataflow: each path streams the
    cost and read-modify-writes the FP32 aggregate, P V (b_c + 8) (sgm.cu:52,101-108), plus the
    filter (V 4 read, N 4 written).  The passes here move fewer bytes than that, so its rate is
    an effective figure, not a fraction of peak."""
    N = seq.H * seq.W
    V = N * seq.D
    return seq.paths * V * (COST_BYTES[cost_type] + 8) + V * 4 + N * 4


def stage_report(seq, s, cost_type, ms_cost, ms_sgm, ms_fusion, storage, form):
    H, W, D, P = seq.H, seq.W, seq.D, seq.paths
    S = len(s["sources"][0])
    N, V = H * W, H * W * D
    bc = COST_BYTES[cost_type]
    design = sgm_pass_design_bytes(W, H, D, P, bc, storage)
    ref_bytes = reference_dataflow_bytes(seq, cost_type)
    # compulsory bytes of the pass: the cost read once, the aggregate written once, the depth
    compulsory = V * (bc + 4) + N * 4
    key = (form, cost_type, storage)
    valu, valu_frac = COST_VALU.get(key, (None, None))
    traffic = SGM_PASS_TRAFFIC.get(key)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    return {
        "cost": {"ms": ms_cost, "bound": "valu", "valu_insts": valu,
                 "valu_issue_frac": valu_frac, "valu_source": COUNTER_PROFILE,
                 "hbm_gbs": gbs(V * bc + (S + 1) * N, ms_cost)},
        "sgm": {"ms": ms_sgm, "bound": "hbm", "storage": storage,
                "includes": "filter (WTA + subpixel)",
                "design_bytes": design, "achieved_gbs": gbs(design, ms_sgm),
                "frac": gbs(design, ms_sgm) / HBM_PEAK_GBS,
                "traffic": traffic,
                "traffic_frac": gbs(traffic, ms_sgm) / HBM_PEAK_GBS if traffic else None,
                "reference_dataflow_bytes": ref_bytes,
                "reference_dataflow_effective_gbs": gbs(ref_bytes, ms_sgm),
                "compulsory_gbs": gbs(compulsory, ms_sgm)},
        "fusion": {"ms": ms_fusion, "bound": "hbm",
                   "achieved_gbs": gbs(104 * N, ms_fusion)},
    }


def measure_plain(args, torch, s, cost_type, storage):
    """One keyframe at a time; HIP events on the launch stream bracket each stage."""
    from probabilistic_mapping_amd import pipeline as pl
    seq = pl.KeyframeSequence(s["images"], s["poses"], s["sources"], s["scene"].K, s["planes"],
                              s["idep_min"], s["idep_max"], s["knobs"], s["paths"],
                              cost_type=cost_type, median=bool(args.median),
                              sgm_storage=storage, write_agg=False)
    nkf = len(s["images"])
    for _ in range(args.warmup):
        seq.run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = {k: [] for k in ("cost", "sgm", "fusion")}
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    for j in range(args.steps):
        marks[j].record(stream)
        seq.reset()
        for i in range(nkf):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            e[0].record(stream)
            seq.cost_stage(i)
            e[1].record(stream)
            seq.sgm_stage(i, seq.dep_raw if seq.median else seq.dep)
            e[2].record(stream)
            seq.fusion_stage(i)   # the median runs inside the fusion step's launch
            e[3].record(stream)
            for k, (a, b) in zip(ev, zip(e[:-1], e[1:])):
                ev[k].append((a, b))
    marks[-1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_ms = [marks[j].elapsed_time(marks[j + 1]) for j in range(args.steps)]
    ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}
    stages = stage_report(seq, s, cost_type, ms["cost"], ms["sgm"], ms["fusion"], storage,
                          "plain")
    stages["median"] = {"enabled": bool(args.median),
                        "in": "the fusion step's launch (stages.fusion)"}
    return {"value": args.steps * nkf / elapsed, "ms_per_step": elapsed / args.steps * 1e3,
            "ms_per_step_median": float(np.median(step_ms)), "stages": stages,
            "form": "plain", "cost_storage": cost_type, "sgm_storage": storage,
            "keyframes_per_step": nkf}


def measure_batched(args, torch, s, cost_type, storage):
    """Autotune the batch, capture it into a hipGraph on a side stream, then time K steps of
    (graph replay for all keyframes' cost / SGM / filter + the fusion steps)."""
    from probabilistic_mapping_amd import pipeline as pl
    seq = pl.BatchedSequence(s["images"], s["poses"], s["sources"], s["scene"].K, s["planes"],
                             s["idep_min"], s["idep_max"], s["knobs"], s["paths"],
                             cost_type=cost_type, median=bool(args.median),
                             sgm_storage=storage, write_agg=False, lookahead=True)
    nkf = len(s["images"])
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        if args.tuning:
            seq.batch.tuning = json.loads(args.tuning)
            tuning = seq.batch.tuning
        else:
            seq.batch.autotune(3)
            tuning = seq.batch.tuning
        # every group's own launch shapes (groups are tuned separately)
        tunings = getattr(seq.batch, "tunings", [tuning])
        if args.overlap:
            # one graph of the whole step: each group's cost launch beside the previous
            # group's SGM pass, filter and fusion steps (BatchedSequence.run_overlapped)
            seq.capture_overlapped()
        else:
            seq.batch.capture()
        torch.cuda.synchronize()

        def step():
            if args.overlap:
                seq.launch_overlapped()
                return
            seq.reset()
            seq.batch.launch()
            for i in range(nkf):
                seq.fusion_step(i)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        t0 = time.perf_counter()
        for j in range(args.steps):
            marks[j].record(side)
            step()
        marks[-1].record(side)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        step_ms = [marks[j].elapsed_time(marks[j + 1]) for j in range(args.steps)]
        # per-stage times of the whole batch, from instrumented passes after the timed steps
        # (each keyframe group's three stages in turn, as its graph runs them)
        reps = max(1, min(args.steps, 5))
        acc = np.zeros(4)
        groups = getattr(seq.batch, "batches", [seq.batch])
        for _ in range(reps):
            for b in groups:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
                ev[0].record(side)
                for st in range(3):
                    b.run_stage(st)
                    ev[st + 1].record(side)
                torch.cuda.synchronize()
                acc[:3] += [ev[i].elapsed_time(ev[i + 1]) for i in range(3)]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(side)
            seq.reset()
            for i in range(nkf):
                seq.fusion_step(i)
            e1.record(side)
            torch.cuda.synchronize()
            acc[3] += e0.elapsed_time(e1)
    ms = acc / reps / nkf
    inflight = measure_inflight(args, torch, s, seq, cost_type, storage, tunings) \
        if args.inflight > 1 else None
    # the SGM pass is stage 1 (paths) + stage 2 (filter; compact: combine + filter)
    stages = stage_report(seq, s, cost_type, ms[0], ms[1] + ms[2], ms[3], storage, "batched")
    stages["sgm"]["paths_ms"] = float(ms[1])
    stages["sgm"]["filter_ms"] = float(ms[2])
    return {"value": args.steps * nkf / elapsed, "ms_per_step": elapsed / args.steps * 1e3,
            "ms_per_step_median": float(np.median(step_ms)), "stages": stages,
            "form": (f"batched (20 keyframes per step, {len(groups)} cache-sized groups of "
                     f"{len(groups[0].outputs)}, " +
                     ("one hipGraph replay per step; each group's cost launch beside the "
                      "previous group's SGM pass and fusion)" if args.overlap
                      else "one hipGraph replay each)") if len(groups) > 1
                     else "batched (20 keyframes per step in one batch, one hipGraph replay)"),
            "overlap": bool(args.overlap),
            "cost_storage": cost_type,
            "sgm_storage": storage, "keyframes_per_step": nkf, "tuning": tuning,
            "inflight": inflight}


def measure_inflight(args, torch, s, first, cost_type, storage, tunings):
    """Several independent 20-keyframe sequences in flight: sequence j of the timed steps runs
    on stream j % n with its own batch buffers, captured graphs and fusion state, so one
    sequence's fusion steps (small, latency-bound launches) overlap the next one's batch.
    Each sequence runs its groups' captured batches one after another (with two sequences in
    flight that measured faster than each as one overlapped graph: 2,772 against 2,668
    keyframes/s).  Every step still computes a whole sequence; the figure is keyframes/s over
    all of them."""
    from probabilistic_mapping_amd import pipeline as pl
    n = args.inflight
    nkf = len(s["images"])
    seqs, streams = [first], [torch.cuda.Stream()]   # first: already captured
    for _ in range(1, n):
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            q = pl.BatchedSequence(s["images"], s["poses"], s["sources"], s["scene"].K,
                                   s["planes"], s["idep_min"], s["idep_max"], s["knobs"],
                                   s["paths"], cost_type=cost_type, median=bool(args.median),
                                   sgm_storage=storage, write_agg=False, lookahead=True)
            for b, t in zip(getattr(q.batch, "batches", [q.batch]), tunings):
                b.tuning = t
            q.batch.capture()
        seqs.append(q)
        streams.append(st)
    if args.overlap:   # the first sequence was captured as one overlapped graph
        with torch.cuda.stream(streams[0]):
            first.batch.capture()
    torch.cuda.synchronize()

    def step(j):
        q, st = seqs[j % n], streams[j % n]
        with torch.cuda.stream(st):
            q.reset()
            q.batch.launch()
            for i in range(nkf):
                q.fusion_step(i)

    for j in range(max(args.warmup, 1) * n):
        step(j)
    torch.cuda.synchronize()
    steps = max(args.steps, n)
    t0 = time.perf_counter()
    for j in range(steps):
        step(j)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    same = all(torch.equal(q.fused, seqs[0].fused) for q in seqs[1:])
    for q in seqs[1:]:
        q.batch.close()
    return {"sequences_in_flight": n, "value": steps * nkf / elapsed,
            "ms_per_step": elapsed / steps * 1e3, "steps": steps,
            "same_fused_result": bool(same)}


def roofline(sg: dict, kernel: str) -> dict:
    """The SGM pass against HBM peak: achieved = the bytes this pass streams (each launch's
    reads and writes with no cache reuse, sgm_pass_design_bytes) over the event-timed pass;
    traffic = the counter-measured HBM bytes of the same pass (COUNTER_PROFILE)."""
    return {"bound": "hbm", "kernel": kernel, "achieved": sg["achieved_gbs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": sg["frac"],
            "algorithmic_bytes": sg["design_bytes"], "traffic": sg["traffic"],
            "traffic_source": COUNTER_PROFILE,
            "reference_dataflow_effective_gbs": sg["reference_dataflow_effective_gbs"]}


def report(args, main: dict, extra: dict | None, copy, base) -> dict:
    sg = main["stages"]["sgm"]
    out = {
        "metric": METRIC,
        "value": main["value"],
        "unit": "fused depth frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main["ms_per_step"],
        "ms_per_step_median": main["ms_per_step_median"],
        "value_from_median": main["keyframes_per_step"] / (main["ms_per_step_median"] * 1e-3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main["cost_storage"] == "f32" else f"{main['cost_storage']}-storage/f32",
        "data": "synthetic (seeded textured-plane scene, 20-frame trajectory)",
        "config": {"workload": WORKLOAD, "form": main["form"],
                   "cost_storage": main["cost_storage"], "sgm_storage": main["sgm_storage"],
                   "median": bool(args.median), "keyframes_per_step": main["keyframes_per_step"]},
        "roofline": roofline(sg, "sgm pass (all paths in one launch + combine/filter)"),
        "stages": main["stages"],
        "attainable_copy": copy,
        "cpu_baseline": base,
    }
    if "tuning" in main:
        out["config"]["tuning"] = main["tuning"]
    if main.get("inflight"):
        out["inflight"] = main["inflight"]
    if extra is not None:
        out["fp16_batched"] = {k: extra[k] for k in ("value", "ms_per_step",
                                                     "ms_per_step_median", "form",
                                                     "cost_storage", "sgm_storage", "stages",
                                                     "tuning", "inflight")}
        out["fp16_batched"]["roofline"] = roofline(
            extra["stages"]["sgm"], "sgm pass (compact paths launch + combine/filter)")
        out["fp16_batched_vs_plain"] = extra["value"] / main["value"]
        if extra.get("inflight"):
            out["fp16_batched_inflight_vs_plain"] = extra["inflight"]["value"] / main["value"]
    return out


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


def bench_driver(args, torch):
    """The reference node's loop (evaluate.cc processImg through pm::FrameDriver): a 640x480
    stream of 60 frames, 64 planes, aggNum 3 at interval 6, CASE 3 (CPU fusion rules), host
    images in and host depths out.  Timed per processImg step that runs (after the first,
    which allocates the mapper), with the frames reaching the mapper through the upload-once
    frame ring and, beside it, uploaded at every use as the reference does."""
    from scipy.spatial.transform import Rotation
    from probabilistic_mapping_amd import driver, mapper, synth
    n_frames, planes = 60, 64
    s = synth.trajectory_scene(args.seed, width=640, height=480, n_frames=n_frames, n_sources=3,
                               planes=planes, fx=481.2017)
    K = s["scene"].K
    cfg = dict(agg_num=3, interval=6, planes=planes, fx=float(K[0, 0]), fy=float(K[1, 1]),
               cx=float(K[0, 2]), cy=float(K[1, 2]),
               dep_sample=(s["idep_max"] - s["idep_min"]) / (planes - 1))
    poses = []
    for R, T in s["poses"]:
        q = Rotation.from_matrix(np.asarray(R, np.float64)).as_quat()
        poses.append((np.asarray(T, np.float64).reshape(3), q))
    old = mapper.set_knobs(case_mode=3, idep_min=s["idep_min"], sgm_paths=4, **s["knobs"])
    forms = {}
    try:
        for name, ring in (("frame_ring", True), ("per_use_upload", False)):
            runs = []
            for _ in range(max(args.warmup, 0) + 1):
                d = driver.FrameDriver(frame_ring=ring, **cfg)
                times = []
                for f in range(n_frames):
                    d.grab_pose(1000 + f, *poses[f])
                    d.grab_image(1000 + f, s["images"][f])
                    t0 = time.perf_counter()
                    r = d.process()
                    if r is not None:
                        times.append(time.perf_counter() - t0)
                runs.append((times, d.ring_puts))
                d.close()
            times, puts = runs[-1]
            steady = times[1:]
            forms[name] = {"steps": len(times), "ms_per_step": float(np.mean(steady)) * 1e3,
                           "ms_per_step_p50": float(np.median(steady)) * 1e3,
                           "ms_steps": [round(t * 1e3, 3) for t in times],
                           "steps_per_s": 1.0 / float(np.mean(steady)),
                           "frame_uploads": puts if ring else len(times) * (cfg["agg_num"] + 1)}
    finally:
        mapper.set_knobs(**old)
    fr, pu = forms["frame_ring"], forms["per_use_upload"]
    out = {
        "metric": "processImg steps/s (host images in, host depth out) @640x480x64 planes, "
                  "aggNum 3, interval 6, CASE 3",
        "value": fr["steps_per_s"], "unit": "steps/s",
        "n_gpus": 1, "steps": fr["steps"] - 1, "warmup": args.warmup,
        "ms_per_step": fr["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded textured-plane scene, 60-frame trajectory)",
        "config": {"workload": "640x480 frames, 64 planes, 4-path SGM, aggNum 3 at interval 6, "
                               "CPU fusion rules, host-side driver loop", "form": "driver"},
        "forms": forms,
        "ring_speedup": pu["ms_per_step"] / fr["ms_per_step"],
    }
    print(json.dumps(out), flush=True)


def sgm_pass_design_bytes(W: int, H: int, D: int, P: int, bc: int, storage: str,
                          write_agg: bool = False) -> int:
    """Bytes the SGM pass (paths launch + combine/filter launch) streams when nothing is
    re-read from a cache: each path reads the cost volume and the reference row and writes its
    own volume (FP32 L for the exact storage, 16-bit excess for the compact one); the combine
    reads the per-path volumes (compact: and the cost once), writes the depth and, if asked,
    the FP32 aggregate."""
    N = W * H
    V = N * D
    vol = 4 if storage == "exact" else 2
    paths = P * (V * bc + N + V * vol)
    combine = P * V * vol + (V * bc if storage == "compact" else 0) + N * 4
    return paths + combine + (V * 4 if write_agg else 0)


def bench_uneven(args, torch):
    """16 independent keyframes of uneven resolution, source count and plane count
    (synth.uneven_keyframes) as one FP16-cost, compact-SGM batch: autotuned, captured into one
    hipGraph and replayed.  value = keyframes/s of the replay (cost over every source, 8-path
    SGM, filter; independent keyframes, so no fusion).  Beside it: per-stage times of the
    batch, and the same keyframes run one at a time through the single-keyframe entry points
    (FP16 compact, and FP32 exact as the plain form runs them)."""
    from probabilistic_mapping_amd import pipeline as pl, synth
    host = synth.uneven_keyframes(synth.UNEVEN_SEED if args.seed == 0 else args.seed)
    knobs, P = host[0]["knobs"], host[0]["paths"]
    specs = [pl.keyframe_spec(k) for k in host]
    nkf = len(specs)
    shapes = [(k["width"], k["height"], k["planes"], k["num_sources"]) for k in host]
    side = torch.cuda.Stream()
    batch = pl.Batch(specs, "f16", knobs, P, "compact", write_agg=False)
    with torch.cuda.stream(side):
        if args.tuning:
            batch.tuning = json.loads(args.tuning)
        else:
            batch.autotune(3)
        tuning = batch.tuning
        batch.capture()
        for _ in range(args.warmup):
            batch.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.launch()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        reps = max(1, min(args.steps, 5))
        acc = np.zeros(3)
        for _ in range(reps):
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            ev[0].record(side)
            for st in range(3):
                batch.run_stage(st)
                ev[st + 1].record(side)
            torch.cuda.synchronize()
            acc += [ev[i].elapsed_time(ev[i + 1]) for i in range(3)]
        ms = acc / reps
    batch.close()
    torch.cuda.empty_cache()
    design = sum(sgm_pass_design_bytes(W, H, D, P, 2, "compact") for W, H, D, _ in shapes)
    sgm_ms = float(ms[1] + ms[2])
    # counter bytes exist for the default keyframe set only
    traffic = UNEVEN_SGM_TRAFFIC if args.seed == 0 else None

    # one keyframe at a time through the single-keyframe entry points
    def one_at_a_time(cost_type, storage):
        bufs = []
        for kf in specs:
            H, W = kf["ref"].shape
            D = kf["planes"]
            cost = torch.empty((H, W, D), dtype=pl.COST_TYPES[cost_type][1], device="cuda")
            ws = (pl.sgm_excess_workspace if storage == "compact" else pl.sgm_workspace)(
                H, W, D, P)
            bufs.append((cost, ws, torch.empty((H, W), dtype=torch.float32, device="cuda")))

        def run():
            for kf, (cost, ws, dep) in zip(specs, bufs):
                pl.calc_cost(kf["ref"], kf["quads"], kf["homs"], kf["idep_min"],
                             kf["dep_sample"], kf["planes"], cost_type, out=cost)
                fn = pl.sgm_filter_excess if storage == "compact" else pl.sgm_filter_paths
                fn(kf["ref"], cost, knobs, P, kf["idep_min"], kf["dep_sample"], workspace=ws,
                   depth=dep, write_agg=False)
        for _ in range(max(args.warmup, 1)):
            run()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        del bufs
        torch.cuda.empty_cache()
        return {"value": args.steps * nkf / dt, "ms_per_step": dt / args.steps * 1e3}

    seq_f16 = one_at_a_time("f16", "compact")
    seq_f32 = one_at_a_time("f32", "exact")
    V = sum(W * H * D for W, H, D, _ in shapes)
    out = {
        "metric": "uneven keyframe batch: depth keyframes/s (16 keyframes, resolutions "
                  "320x240-1280x720, 2-8 sources, 32-128 planes, 8-path SGM)",
        "value": args.steps * nkf / elapsed, "unit": "keyframes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16-storage/f32", "data": "synthetic (16 seeded scenes and trajectories)",
        "config": {"workload": "16 independent keyframes, one hipGraph replay per step",
                   "form": "uneven batch (FP16 cost, compact SGM storage)",
                   "keyframes": [list(x) for x in shapes], "cost_volume_elements": V,
                   "tuning": tuning},
        "stages": {"cost_ms": float(ms[0]), "sgm_paths_ms": float(ms[1]),
                   "combine_filter_ms": float(ms[2])},
        "roofline": {"bound": "hbm", "kernel": "sgm pass (compact paths launch + "
                                                "combine/filter launch), whole batch",
                     "achieved": design / (sgm_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": design / (sgm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "algorithmic_bytes": design, "traffic": traffic,
                     "traffic_source": UNEVEN_PROFILE if traffic else None},
        "one_at_a_time": {"f16_compact": seq_f16, "f32_exact": seq_f32},
        "batch_vs_one_at_a_time_f16": (args.steps * nkf / elapsed) / seq_f16["value"],
    }
    print(json.dumps(out), flush=True)


def write_mapper_workload(path, s, steps, warmup, median, ring):
    """The workload file pm_bench_mapper reads (layout in csrc/pm_bench_mapper.cpp)."""
    import struct
    H, W = s["images"][0].shape
    kn = s["knobs"]
    with open(path, "wb") as f:
        f.write(struct.pack("<10i", W, H, s["planes"], len(s["images"]), len(s["sources"][0]),
                            steps, warmup, int(median), s["paths"], ring))
        f.write(np.asarray(s["scene"].K, "<f8").tobytes())
        f.write(np.array([s["idep_min"], s["idep_max"], kn["pi1"], kn["pi2"], kn["tau_so"],
                          kn["sgm_q1"], kn["sgm_q2"], kn["var_scale"]], "<f8").tobytes())
        for R, T in s["poses"]:
            f.write(np.concatenate([np.asarray(R, "<f8").reshape(9),
                                    np.asarray(T, "<f8").reshape(3)]).tobytes())
        f.write(np.asarray(s["sources"], "<i4").tobytes())
        for im in s["images"]:
            f.write(np.ascontiguousarray(im, np.uint8).tobytes())


def bench_mapper(args, torch):
    """The drop-in path timed from C++: the 20-keyframe workload through pm::StereoMapper
    (initReference, update per source, outputFusion returning the host depth), run by the
    pm_bench_mapper executable in a child process.  Two forms: host images at every use
    (pinned staging, as the reference uploads them) and the upload-once frame ring.  The last
    step's fused and filtered depth must equal the device-resident sequence's bit for bit."""
    import subprocess
    import tempfile
    from probabilistic_mapping_amd import build as pm_build, pipeline as pl, synth
    s = synth.trajectory_scene(args.seed)
    H, W = s["images"][0].shape
    D, P, S = s["planes"], s["paths"], len(s["sources"][0])
    F = len(s["images"])
    kn = s["knobs"]
    exe = pm_build.BENCH_MAPPER
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} is missing: run __graft_entry__.build()")
    seq = pl.KeyframeSequence(s["images"], s["poses"], s["sources"], s["scene"].K, D,
                              s["idep_min"], s["idep_max"], kn, P, cost_type="f32",
                              median=bool(args.median))
    seq.run()
    torch.cuda.synchronize()
    want_fused, want_dep = seq.fused.cpu().numpy(), seq.dep.cpu().numpy()
    del seq
    torch.cuda.empty_cache()
    forms = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, ring in (("host_images", 0), ("frame_ring", 1)):
            path = os.path.join(tmp, f"workload_{ring}.bin")
            write_mapper_workload(path, s, args.steps, args.warmup, bool(args.median), ring)
            out = subprocess.run([exe, path, tmp], capture_output=True, text=True, check=True)
            r = json.loads(out.stdout.strip().splitlines()[-1])
            fused = np.fromfile(os.path.join(tmp, "fused.f32"), np.float32).reshape(H, W)
            dep = np.fromfile(os.path.join(tmp, "dep.f32"), np.float32).reshape(H, W)
            ms = np.asarray(r["step_ms"])
            forms[name] = {"value": F / (ms.mean() * 1e-3), "ms_per_step": float(ms.mean()),
                           "ms_per_step_median": float(np.median(ms)),
                           "frame_uploads_per_step": r["ring_puts"] if ring else F * (S + 1),
                           "same_result_as_sequence": bool(np.array_equal(fused, want_fused)
                                                           and np.array_equal(dep, want_dep))}
    hi = forms["host_images"]
    out = {
        "metric": "fused depth frames/s through the C++ StereoMapper (drop-in path) "
                  "@752x480x64 planes, 5 sources, 8-path SGM",
        "value": hi["value"], "unit": "fused depth frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": hi["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded textured-plane scene, 20-frame trajectory)",
        "config": {"workload": WORKLOAD, "form": "mapper (host images in, host depth out; "
                                                 "C++ driver, std::chrono)"},
        "forms": forms,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; --gpus must be 1 "
                                   "(keyframe streams would shard one stream per GPU)"}))
        sys.exit(2)
    import torch
    from probabilistic_mapping_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.form == "diff":
        return bench_diff(args, torch)
    if args.form == "driver":
        return bench_driver(args, torch)
    if args.form == "uneven":
        return bench_uneven(args, torch)
    if args.form == "mapper":
        return bench_mapper(args, torch)
    s = synth.trajectory_scene(args.seed)
    copy = attainable_copy(torch, args.copy_mib)
    storage = default_storage(args.cost_type, args.sgm_storage)
    if storage == "sweep" and args.form != "batched":
        print(json.dumps({"error": "the sweep SGM storage runs in the batched form only "
                                   "(--form batched)"}))
        sys.exit(2)
    extra = None
    if args.form == "batched":
        main_r = measure_batched(args, torch, s, args.cost_type, storage)
    else:
        main_r = measure_plain(args, torch, s, args.cost_type, storage)
        if args.form == "both":
            torch.cuda.empty_cache()
            extra = measure_batched(args, torch, s, "f16", "compact")
    base = cpu_baseline(args.seed, args.cpu_baseline_seconds, bool(args.median))
    print(json.dumps(report(args, main_r, extra, copy, base)), flush=True)


if __name__ == "__main__":
    main()
