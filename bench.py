"""Benchmark: 1080p retime output frames/s on one MI355X.

Workload: `vfx apply retime` on a synthetic 1920x1080 RGBA float32 clip of 120 frames with the
keyed speed ramp 100 % @ 0, 25 % @ 60, 100 % @ 120 (defaults otherwise: Best vectors, Blend,
sensitivity 70), 120 output frames. One step = the whole clip rendered through the HIP backend
(frame selection, ingest, pyramids, every flow solve, warp+blend and finish) with the clip and the
output resident in HBM.

Prints ONE JSON line. `roofline` covers the dominant kernel (HIP-event timed around its launches
inside the timed steps); `cpu_baseline` times the oracle (the C++ restatement of the reference
CPU path) on a bounded sample of the same output frames and checks those frames' bytes against
the device output.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SPEED = "speed_percent=100@0;25@60;100@120"

# Measured HBM bytes over algorithmic bytes per kernel family, from the FETCH_SIZE / WRITE_SIZE
# counter runs (copy-calibrated) recorded in profiles/retime_1080p_hbm_traffic_pmc.md.
TRAFFIC_RATIO = {"frames": 0.77, "new_disp_gauss": 1.04, "calc_b_gauss": 1.03, "bm_cols": 1.23, "bm_rows": 1.30,
                 "ingest": 1.00, "pyr_luma_blur5_band_resample": 1.08, "bm_transpose": 1.39, "upsample": 1.27,
                 "tensor_gauss": 1.00}
# The share of the SIMDs' issue slots the family's VALU instructions fill over its time, from the
# SQ_INSTS_VALU counter run recorded in the same profile.
VALU_SHARE = {"calc_b_gauss": 0.73, "frames": 0.61, "bm_rows": 0.55, "bm_cols": 0.50, "new_disp_gauss": 0.58,
              "pyr_luma_blur5_band_resample": 0.46, "tensor_gauss": 0.67, "bm_transpose": 0.19, "upsample": 0.37,
              "ingest": 0.06}
# What bounds a family where neither the HBM fraction nor the VALU share says it: the block-matching
# sweeps are serial chains, each step's decision feeding the next step's texel gathers.
LIMITER = {"bm_rows": "serial chain latency (each step's decision feeds the next step's texel gathers)",
           "bm_cols": "serial chain latency (each step's decision feeds the next step's texel gathers)"}
# The library counts a fused kernel's bytes once (its inputs read once, its outputs written once).
# The step-boundary model of the CPU path counts every step's inputs and outputs, 4 bytes a float:
# grad + tensor 24 and the tensor blur 48 B/px (fused: 24), calc_b 32 and the B blur 32 (fused: 32),
# new_disp 36 and the flow blur 32 (fused: 36); the block-matching transposes are layout work the
# CPU path does not have.
STEP_MODEL_FACTOR = {"tensor_gauss": 72.0 / 24.0, "calc_b_gauss": 64.0 / 32.0, "new_disp_gauss": 68.0 / 36.0,
                     "bm_transpose": 0.0}
FLOW_FAMILIES = ("ingest", "pyr_luma_blur5_band_resample", "tensor_gauss", "calc_b_gauss", "new_disp_gauss", "bm_rows", "bm_cols",
                 "bm_transpose", "range", "upsample", "negate", "solve_init", "coarse_solve")
WARP_FAMILIES = ("frames",)
STAGE_TARGET = 0.60  # of the HBM peak, for the flow and the warp stages


def copy_rate_gbs(torch, mib=1024, reps=10):
    """The practical HBM ceiling of this device: a device-to-device copy of `mib` MiB (torch's copy
    kernel), timed with HIP events over `reps` copies after a warm-up; read plus written bytes over
    the time. Larger than the last-level cache (256 MB), so every byte crosses HBM."""
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * (mib << 20) / (ms / 1000.0) / 1e9


def step_model(params, n_out, select, W, H, bpp, w0, h0, levels, sensitivity):
    """The step-boundary byte model of one render (the CPU path's steps, each step's distinct
    inputs read once and outputs written once): per solved working pixel, level and direction 516
    B (three iterations of calc_b 32 + B blur 32 + new_disp 36 + flow blur 32, four sweeps 96,
    range check and damping or upsample 24), per solved working pixel, level and distinct X frame
    72 B (grad + tensor 24, tensor blur 48: they depend on X alone, and the device computes them
    once per frame where the CPU path repeats them per solve), per pyramid the frame read plus
    every level written, per source frame its ingest (the wire frame read, the engine image
    written), per in-between frame both frames read, the output written and both flows read, per
    exact frame one frame read and one written. Returns (flow bytes, warp bytes, distinct pairs,
    in-between frames)."""
    import math

    pairs, vector_frames, exact_frames, sources = set(), 0, 0, set()
    for n in range(n_out):
        s = select(params, n)
        sources.update((s.a,) if s.exact else (s.a, s.b))
        if s.exact:
            exact_frames += 1
        else:
            vector_frames += 1
            pairs.add((s.a, s.b))
    sizes, w, h = [], w0, h0
    for _ in range(levels):
        sizes.append(w * h)
        w, h = max(w // 2, 2), max(h // 2, 2)
    x = math.log(float(sensitivity) ** 2) / math.log(2.0) + levels  # flow.rs:179-197
    start = min(max(int(math.floor(x)), 0), levels - 1) if x >= 0 else 0
    solved = sum(sizes[:start + 1])
    frames_used = {f for p in pairs for f in p}
    # both directions of every pair are solved, so every frame of a pair is the X of a solve
    flow = (2 * len(pairs) * solved * 516 + len(frames_used) * solved * 72
            + len(frames_used) * (W * H * bpp + 4 * sum(sizes)) + len(sources) * 2 * W * H * bpp)
    warp = vector_frames * (3 * W * H * bpp + 2 * w0 * h0 * 8) + exact_frames * 2 * W * H * bpp
    return flow, warp, len(pairs), vector_frames


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--per-pair", action="store_true", help="also time one launch sequence per output frame")
    ap.add_argument("--no-profile", action="store_true", help="skip the profiled pass (per-launch HIP events)")
    ap.add_argument("--no-graphs", action="store_true", help="time plain launches instead of the captured graph")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (calibrates the HBM traffic counters)")
    ap.add_argument("--batch-clips", type=int, default=64,
                    help="clips of the mixed-size batch leg (batched against one call per output frame); 0 skips it")
    ap.add_argument("--batch-only", action="store_true", help="run only the mixed-size batch leg")
    ap.add_argument("--diff-steps", type=int, default=50,
                    help="refinement steps of the differentiable warp+blend leg (1280x720); 0 skips it")
    ap.add_argument("--diff-only", action="store_true", help="run only the differentiable warp+blend leg")
    ap.add_argument("--vectors-frames", type=int, default=60,
                    help="frames of the 3840x2160 Create Vectors / Vectors In leg; 0 skips it")
    ap.add_argument("--vectors-only", action="store_true", help="run only the Create Vectors / Vectors In leg")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: there is no multi-GPU path (asked for {args.gpus}); a clip would "
              "shard by output-frame range, each GPU rendering its frames from the solves they need", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from twixtor_amd import synth
    from twixtor_amd.backend import HipBackend
    from twixtor_amd.params import Retime

    dev = torch.device("cuda:0")
    if args.batch_only:
        be = HipBackend(0)
        print(json.dumps({"batch": batch_leg(be, args.batch_clips, args.steps, args.warmup)}), flush=True)
        return
    if args.diff_only:
        print(json.dumps({"diff": diff_leg(args.diff_steps)}), flush=True)
        return
    if args.vectors_only:
        print(json.dumps({"vectors": vectors_leg(HipBackend(0), args.vectors_frames, args.steps, args.warmup)}), flush=True)
        return
    spec = synth.spec_1080p(frames=args.frames)
    gen = synth.Clip(spec, device=dev)
    clip = torch.cat([gen.wire(t) for t in range(spec.frames)]).contiguous()
    cfg = Retime(spec.width, spec.height, 32, "24", [SPEED])
    params = cfg.struct()
    fb = cfg.frame_bytes()
    n_out = args.frames
    out = torch.empty(fb * n_out, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    be = HipBackend(0)
    be.graphs(not args.no_graphs)

    def step():
        be.retime_frames_dev(cfg, clip.data_ptr(), spec.frames, 0, n_out, out.data_ptr(), params)

    # timed steps: the whole clip per step; with graphs on, the first call (warm-up) captures the
    # batched render and the timed steps replay it
    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    s0 = be.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s1 = be.stats()
    ms_per_step = elapsed * 1000.0 / args.steps
    fps = n_out * args.steps / elapsed
    # the per-kernel HIP-event timing: the same steps again as plain launches, each launch
    # bracketed by events on the library's stream (kernel durations do not depend on the graph)
    prof, ms_profiled = {}, None
    if not args.no_profile:
        be.profile(True)
        step()
        torch.cuda.synchronize()
        be.profile_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        ms_profiled = (time.perf_counter() - t0) * 1000.0 / args.steps
        prof = be.profile_read()
        be.profile(False)

    copy_gbs = copy_rate_gbs(torch)
    # dominant kernel by HIP-event time over the timed steps
    kern = max(prof, key=lambda k: prof[k]["ms"]) if prof else None
    roofline = None
    if kern:
        k = prof[kern]
        avg_ms = k["ms"] / max(k["launches"], 1)
        achieved = (k["bytes"] / max(k["launches"], 1)) / (avg_ms / 1000.0) / 1e9 if avg_ms > 0 else 0.0
        alg_per_launch = k["bytes"] / max(k["launches"], 1)
        traffic = round(TRAFFIC_RATIO[kern] * alg_per_launch) if kern in TRAFFIC_RATIO else None
        factor = STEP_MODEL_FACTOR.get(kern, 1.0)
        roofline = {"bound": "hbm", "kernel": kern, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algorithmic_bytes_per_launch": round(alg_per_launch),
                    "avg_launch_ms": round(avg_ms, 4), "launches_per_step": k["launches"] // args.steps,
                    "share_of_step": round(k["ms"] / args.steps / ms_per_step, 3),
                    "byte_models": {
                        "fused": {"frac": round(achieved / HBM_PEAK_GBS, 4),
                                  "note": "achieved/frac above: the kernel's own inputs read once, outputs written once"},
                        "step_model": {"achieved": round(achieved * factor, 1),
                                       "frac": round(achieved * factor / HBM_PEAK_GBS, 4),
                                       "note": "the CPU path's step boundaries (each step's inputs and outputs)"},
                        "counter_traffic": None if traffic is None else {
                            "achieved": round(traffic / (avg_ms / 1000.0) / 1e9, 1),
                            "frac": round(traffic / (avg_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                            "note": "FETCH_SIZE + WRITE_SIZE bytes (copy-calibrated) per launch"}},
                    "copy_reference": {"GBs": round(copy_gbs, 1), "frac_of_peak": round(copy_gbs / HBM_PEAK_GBS, 4),
                                       "achieved_over_copy": round(achieved / copy_gbs, 4),
                                       "note": "a 1 GiB device-to-device copy on the same device, HIP events: the "
                                               "bandwidth a plain streaming kernel reaches here"},
                    "valu_issue_share": VALU_SHARE.get(kern),
                    "limiter": LIMITER.get(kern) or ("VALU issue (no FMA, no packed FP32 under the bit-exact contract)"
                                                     if VALU_SHARE.get(kern, 0.0) > achieved / HBM_PEAK_GBS else "HBM")}
    # whole-path algorithmic bytes per step over the step time
    total_bytes = sum(v["bytes"] for v in prof.values()) / args.steps if prof else 0.0
    path = {"algorithmic_GB_per_step": round(total_bytes / 1e9, 3),
            "achieved_GBs": round(total_bytes / (ms_per_step / 1000.0) / 1e9, 1),
            "kernels": {n: {"ms_per_step": round(v["ms"] / args.steps, 3),
                            "algorithmic_bytes_per_step": round(v["bytes"] / args.steps),
                            "launches_per_step": v["launches"] // args.steps,
                            "GBs": round(v["bytes"] / (v["ms"] / 1000.0) / 1e9, 1) if v["ms"] > 0 else 0.0}
                        for n, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])}}
    # the flow and warp stages against the step-boundary byte model of the CPU path
    from twixtor_amd.backend import flow_geometry, flow_settings, select

    fw0, fh0, levels = flow_geometry(flow_settings(quality=4), spec.width, spec.height)
    flow_b, warp_b, n_pairs, n_vec = step_model(params, n_out, select, spec.width, spec.height, 16, fw0, fh0, levels,
                                                0.7)
    stages = {"target_frac": STAGE_TARGET, "copy_GBs": round(copy_gbs, 1)}
    for name, fams, model in (("flow", FLOW_FAMILIES, flow_b), ("warp", WARP_FAMILIES, warp_b)):
        ms = sum(prof[f]["ms"] for f in fams if f in prof) / args.steps if prof else 0.0
        stages[name] = {"ms_per_step": round(ms, 3), "step_model_GB": round(model / 1e9, 3),
                        "GBs": round(model / (ms / 1000.0) / 1e9, 1) if ms > 0 else None,
                        "frac": round(model / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None,
                        "over_copy": round(model / (ms / 1000.0) / 1e9 / copy_gbs, 4) if ms > 0 else None,
                        "families": list(fams)}
    stages["distinct_pairs"] = n_pairs
    stages["in_between_frames"] = n_vec
    path["stages"] = stages
    # the tile shapes the autotuner picked for the fused solve kernels, per level area class
    shapes = {0: "64x32 pu4", 1: "64x32 pu1", 2: "64x16 pu4", 3: "64x16 pu1", 4: "64x64 pu4", 5: "64x64 pu1"}
    path["tile_choices"] = [{"kernel": ("tensor", "calc_b", "new_disp")[k], "log2_area": a, "log2_jobs": j,
                             "shape": shapes[v]} for k, a, j, v in HipBackend.tile_choices()]

    per_pair = None
    if args.per_pair:
        be.graphs(False)  # one plain launch sequence per output frame, as the reference's loop
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for n in range(n_out):
            be.retime_frames_dev(cfg, clip.data_ptr(), spec.frames, n, 1, out.data_ptr() + n * fb, params)
        torch.cuda.synchronize()
        pp = time.perf_counter() - t0
        per_pair = {"ms_per_clip": round(pp * 1000.0, 2), "batched_speedup": round(pp * 1000.0 / ms_per_step, 2)}
        be.graphs(not args.no_graphs)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, clip, out, spec.frames, n_out, fb, args.cpu_seconds)
    del clip, out
    torch.cuda.empty_cache()
    batch = batch_leg(be, args.batch_clips, args.steps, args.warmup) if args.batch_clips > 0 else None
    diff = diff_leg(args.diff_steps) if args.diff_steps > 0 else None
    vectors = vectors_leg(be, args.vectors_frames, args.steps, args.warmup) if args.vectors_frames > 1 else None

    line = {
        "metric": "1080p retime output frames/s",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"retime 1920x1080 RGBA f32, {spec.frames} frames, {SPEED}, Best, Blend -> {n_out} output frames",
                   "width": 1920, "height": 1080, "frames_in": spec.frames, "frames_out": n_out,
                   "unique_solves_per_step": (s1["pair_solves"] - s0["pair_solves"]) // args.steps,
                   "vector_frames_per_step": (s1["composes"] - s0["composes"]) // args.steps,
                   "graph_replays_timed": s1["graph_replays"] - s0["graph_replays"],
                   "ms_per_step_profiled_plain_launches": None if ms_profiled is None else round(ms_profiled, 3)},
        "roofline": roofline,
        "path": path,
        "per_pair": per_pair,
        "batch": batch,
        "diff": diff,
        "vectors": vectors,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def batch_leg(be, n_clips, steps, warmup):
    """The mixed-size batch: n_clips RGBA 8-bit clips of 24 frames (widths 320..2048, heights
    180..1152), each retimed at 33.3333 % into 70 output frames. Batched: one twx_retime_batch call
    over every clip (captured into a graph by the warm-up call and replayed by the timed ones).
    Against: one plain call per output frame (the reference's render loop), which solves the pair
    of every in-between frame. The output bytes of both must be equal."""
    import torch

    from twixtor_amd import synth
    from twixtor_amd.params import Retime

    dev = torch.device("cuda:0")
    specs = synth.spec_batch(clips=n_clips)
    n_in, n_out = 24, 70
    cfgs = [Retime(s.width, s.height, 8, "24", ["speed_percent=33.3333"]) for s in specs]
    structs = [c.struct() for c in cfgs]
    clips = [synth.Clip(s, device=dev).clip(n_in).contiguous() for s in specs]
    outs = [torch.zeros(c.frame_bytes() * n_out, dtype=torch.uint8, device=dev) for c in cfgs]
    n = len(cfgs)
    cptr, optr = [c.data_ptr() for c in clips], [o.data_ptr() for o in outs]
    torch.cuda.synchronize()

    def batched():
        be.retime_batch(structs, cptr, [n_in] * n, [0] * n, [n_out] * n, optr)

    for _ in range(warmup):
        batched()
    torch.cuda.synchronize()
    s0 = be.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        batched()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    s1 = be.stats()
    solves = (s1["pair_solves"] - s0["pair_solves"]) // steps
    outs2 = [torch.zeros_like(o) for o in outs]
    be.graphs(False)  # the per-pair form: one plain launch sequence per output frame
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, cfg in enumerate(cfgs):
        fb = cfg.frame_bytes()
        for i in range(n_out):
            be.retime_frames_dev(cfg, cptr[k], n_in, i, 1, outs2[k].data_ptr() + i * fb, structs[k])
    torch.cuda.synchronize()
    t_pp = time.perf_counter() - t0
    be.graphs(True)
    same = all(torch.equal(a, b) for a, b in zip(outs, outs2))
    px = sum(s.width * s.height for s in specs)
    return {"workload": f"{n} RGBA 8-bit clips, widths 320-2048, heights 180-1152, {n_in} frames each, "
                        f"speed_percent=33.3333 -> {n_out} output frames each",
            "clips": n, "output_frames": n * n_out, "megapixels_per_output_frame_avg": round(px / n / 1e6, 3),
            "unique_solves_per_step": solves, "batched_ms": round(t_batch * 1000.0, 2),
            "batched_frames_per_s": round(n * n_out / t_batch, 1), "per_pair_ms": round(t_pp * 1000.0, 2),
            "per_pair_calls": n * n_out, "speedup": round(t_pp / t_batch, 2), "bytes_equal": bool(same)}


def diff_leg(steps):
    """The differentiable warp+blend: pair 0 of the 1280x720 RGB clip rendered at twice the pair
    rate, solver flows plus N(0, 0.5^2) px noise, `steps` Adam steps of photometric refinement
    (forward, L2 loss against the true middle frame, backward, update). Timed as one
    twx_refine_flows call (every step on the device, fixed-order sums) and, for comparison, as the
    torch loop (autograd over the forward and backward entry points, torch.optim.Adam). Reports the
    time per step, the loss before and after, and the HIP-event time and algorithmic bytes of the
    device loop's kernels."""
    import torch

    from twixtor_amd import diff, synth

    dev = torch.device("cuda:0")
    clip = synth.Clip(synth.spec_720p_2x(frames=3), device="cpu")
    a, b, truth, fab, fba = diff.photometric_case(clip, 0)
    h = diff.handle(dev)
    diff.refine_flows_device(a, b, truth, fab, fba, 0.5, steps=2)  # warm-up
    diff.refine_flows(a, b, truth, fab, fba, 0.5, steps=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, tl = diff.refine_flows(a, b, truth, fab, fba, 0.5, steps=steps)
    torch.cuda.synchronize()
    el_torch = time.perf_counter() - t0
    h.profile(True)
    h.profile_reset()
    t0 = time.perf_counter()
    _, _, losses = diff.refine_flows_device(a, b, truth, fab, fba, 0.5, steps=steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    prof = h.profile_read()
    h.profile(False)
    # the same call without the per-launch events: the time a caller sees
    t0 = time.perf_counter()
    diff.refine_flows_device(a, b, truth, fab, fba, 0.5, steps=steps)
    torch.cuda.synchronize()
    el = min(el, time.perf_counter() - t0)
    kern = {}
    for name, v in prof.items():
        n = max(v["launches"], 1)
        avg = v["ms"] / n
        gbs = (v["bytes"] / n) / (avg / 1000.0) / 1e9 if avg > 0 else 0.0
        kern[name] = {"launches": v["launches"], "avg_launch_ms": round(avg, 4), "ms_per_step": round(v["ms"] / steps, 4),
                      "algorithmic_bytes_per_launch": round(v["bytes"] / n), "GBs": round(gbs, 1),
                      "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return {"workload": "1280x720 RGB f32 pair (frames 0 and 2 of a clip at twice the pair rate, truth frame 1), "
                        "flows 640x360 from the solver + N(0, 0.5^2) px, blend at wa = 0.5, L2 loss, Adam lr 0.05",
            "steps": steps, "ms_per_step": round(el * 1000.0 / steps, 3),
            "ms_per_step_torch_loop": round(el_torch * 1000.0 / steps, 3),
            "loss_first": losses[0], "loss_last": losses[-1], "loss_last_torch_loop": tl[-1],
            "deterministic": "every sum in a fixed order or in fixed point: the same bits on every run",
            "kernels": kern}


def _families(prof, steps=1):
    out = {}
    for name, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"]):
        n = max(v["launches"], 1)
        gbs = v["bytes"] / (v["ms"] / 1000.0) / 1e9 if v["ms"] > 0 else 0.0
        out[name] = {"ms": round(v["ms"] / steps, 3), "launches": v["launches"] // steps, "GBs": round(gbs, 1),
                     "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def vectors_leg(be, frames, steps, warmup):
    """Create Vectors and the Vectors In read-back at 3840x2160 RGBA f32: a clip under a
    global affine motion with 10 discs; Create Vectors with Normal (half res) vectors, track
    Previous, Max Displace 2048: encoded frame n = the flow of frame n towards frame n - 1, for
    n = 1 .. frames - 1, one twx_create_vectors_dev call over the clip (device-resident), written
    as frames 1 .. frames - 1 of a vector layer whose frame 0 is transparent. The read-back is
    `vfx apply vectors-in` at 50 % (30 fps, warping 2 Forward, Blend, Max Displace 2048) with that
    layer as Vectors: Prev: output frames 2 .. 2 + 2 * rb_pairs, every odd one an in-between frame
    (the layer's flow decoded and filled, its inversion through the Gouraud forward warp, the
    forward compositions with coverage and fill), one twx_vectors_in_frames_dev call, device to
    device, after one untimed call."""
    import torch

    from twixtor_amd import synth
    from twixtor_amd.backend import flow_settings
    from twixtor_amd.params import VectorsIn

    dev = torch.device("cuda:0")
    spec = synth.spec_4k_vectors(frames=frames)
    gen = synth.Clip(spec, device=dev)
    W, H, fmt = spec.width, spec.height, 2
    clip = torch.cat([gen.wire(t) for t in range(frames)]).contiguous()
    fb = clip.numel() // frames
    xs, ys = list(range(1, frames)), list(range(0, frames - 1))
    layer = torch.zeros(fb * frames, dtype=torch.uint8, device=dev)  # frame 0 transparent
    fs = flow_settings(quality=4)
    md = 2048.0
    torch.cuda.synchronize()

    def run():
        be.create_vectors_dev(fs, W, H, fmt, clip.data_ptr(), frames, xs, ys, md, layer.data_ptr() + fb)

    for _ in range(warmup):
        run()
    be.profile(True)
    be.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    t_cv = (time.perf_counter() - t0) / steps
    prof_cv = be.profile_read()
    be.profile(False)

    rb_pairs = 4
    cfg = VectorsIn(W, H, 32, "30", ["speed_percent=50", "vectors_max_displace=2048"])
    first, count = 2, 2 * rb_pairs
    out = torch.empty(fb * count, dtype=torch.uint8, device=dev)

    def read_back():
        be.vectors_in_frames_dev(cfg, clip.data_ptr(), frames, None, (layer.data_ptr(), frames), None, first, count,
                                 out.data_ptr())

    read_back()  # warm-up: the first launch of each kernel loads its code
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    read_back()
    torch.cuda.synchronize()
    t_rb = time.perf_counter() - t0
    be.profile(True)
    be.profile_reset()
    read_back()
    torch.cuda.synchronize()
    prof_rb = be.profile_read()
    be.profile(False)
    kern_rb = sum(v["ms"] for v in prof_rb.values())
    finite = bool(torch.isfinite(out.view(torch.float32)).all().item())
    del clip, layer, out
    torch.cuda.empty_cache()
    return {"workload": f"Create Vectors 3840x2160 RGBA f32, {frames} frames (affine background 0.2 deg, x1.002, (3, -1) px per "
                        "frame; 10 discs up to 40 px/frame), Normal vectors (working 1920x1080), track Previous, Max Displace "
                        f"2048 -> {len(xs)} encoded frames; Vectors In read-back of {count} output frames at 50 % "
                        f"({rb_pairs} in-between), warping 2, Blend, device-resident",
            "create_vectors": {"output_frames": len(xs), "ms": round(t_cv * 1000.0, 2),
                               "frames_per_s": round(len(xs) / t_cv, 1), "kernels": _families(prof_cv, steps)},
            "read_back": {"output_frames": count, "in_between_frames": rb_pairs, "wall_ms": round(1000.0 * t_rb, 2),
                          "wall_ms_per_in_between_frame": round(1000.0 * t_rb / rb_pairs, 2),
                          "kernel_ms_per_in_between_frame": round(kern_rb / rb_pairs, 3),
                          "kernels": _families(prof_rb, 1), "finite": finite}}


def cpu_baseline(cfg, clip, out, n_in, n_out, fb, budget_s):
    """The oracle on the host: the CPU path restated in C++ with OpenMP over rows (the reference
    splits rows over rayon's threads the same way), over a bounded, evenly spread sample of the
    output frames, on every core this process may use and on one thread; the sampled frames' bytes
    must equal the device output. Each frame is rendered whole as the CPU path does: an in-between
    frame solves its pair again (the device solves each distinct pair once per clip)."""
    from oracle import cpuref

    nproc = os.cpu_count()
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    host_clip = clip.cpu().numpy().tobytes()
    host_out = out.cpu().numpy()
    params = cfg.struct()
    order = list(range(1, n_out, 7)) + [0]

    def run(n_threads, budget):
        cpuref.set_threads(n_threads)
        done, t_sum, same = 0, 0.0, True
        for n in order:
            t0 = time.perf_counter()
            got = cpuref.retime_frames(params, host_clip, n_in, n, 1)
            t_sum += time.perf_counter() - t0
            done += 1
            same &= got == host_out[n * fb:(n + 1) * fb].tobytes()
            if t_sum >= budget:
                break
        return done, t_sum, same

    done, t_all, same = run(threads, budget_s)
    done1, t_one, same1 = run(1, budget_s / 2)
    cpuref.set_threads(threads)
    vector_frames = sum(1 for n in range(n_out) if not cpuref.select(params, n).exact)
    return {"value": round(done / t_all, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{done} of {n_out} output frames (every 7th from frame 1, then frame 0), the oracle rendering each "
                      "frame whole",
            "host": {"nproc": nproc, "affinity_cpus": affinity, "OMP_NUM_THREADS": omp},
            "one_thread": {"value": round(done1 / t_one, 4), "unit": "frames/s", "sample_frames": done1},
            "solves_per_clip": {"cpu_path": 2 * vector_frames,
                                "note": "the CPU path solves both directions of each in-between frame's pair; the device "
                                        "solves each distinct pair once per clip"},
            "bytes_equal_to_device": bool(same and same1)}


if __name__ == "__main__":
    main()
