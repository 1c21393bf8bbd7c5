#!/usr/bin/env python3
"""Throughput of the MI355X stitch path on one GPU.

A step is one frame of the batched workload of the headline metric: 8 cameras
of three sizes (2 x 2560x1440 rectilinear 90 deg, 4 x 1920x1080 rectilinear
100 deg, 2 x 1280x1280 fisheye 190 deg) remapped and multiband-blended into a
7680x3840 equirect canvas with 9 pyramid levels.  Inputs (camera frames,
maps, seams, weights) are resident in HBM before timing starts; the timed
region is K back-to-back render() calls on one stream, bracketed by device
synchronisation.  value = output Mpix/s.

Prints one JSON line (see the field list in main()).  The CPU baseline is the
oracle's restatement of CPUPanoramaRender::render with row-parallel reprojection
and per-camera pyramid threads, timed on a bounded sample of the same frames.
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

METRIC = "stitched output Mpix/s (8-cam→7680x3840) + p99 frame latency, 16 streams"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(wl, srcs, xs, ys, ms, luts, frames: int):
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    ob = O.Blender()
    ob.prepare(np.stack(ms), wl.levels)
    bgr = [s[..., :3].copy() for s in srcs]
    t0 = time.perf_counter()
    for _ in range(frames):
        ob.render(bgr, xs, ys, luts, nthreads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(wl.out_pixels * frames / dt / 1e6, 3),
        "unit": "Mpix/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} frame(s) of the same workload, oracle CPUPanoramaRender::render "
                  f"restatement, {threads} threads, {dt:.1f} s",
    }


def read_traffic(workload: str):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed8")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the stitch path shards by stream, one rig per GPU, "
              "and this build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import panostitcher_amd as P
    from panostitcher_amd import synth

    P.lib()
    wl = synth.workload(args.workload, args.scale)
    t0 = time.perf_counter()
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = [synth.bgrx(f) for f in synth.frames(wl)]
    luts = synth.luts(wl)
    log(f"inputs {wl.name} {wl.dst_w}x{wl.dst_h} n={wl.n} in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    r = P.PanoramaRender()
    ok = r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys,
                        ms, wl.levels)
    if not ok:
        raise SystemExit("prepare failed")
    log(f"prepare {time.perf_counter() - t0:.1f}s, levels {r.num_levels}, "
        f"device bytes {r.mem_bytes() / 1e9:.2f} GB")
    dsrc = [torch.from_numpy(s).cuda() for s in srcs]
    out = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        r.render(dsrc, out, luts, stream)
    torch.cuda.synchronize()
    log("warmup done")

    r.set_timing(True)
    r.stage_ms()  # reset the period
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render(dsrc, out, luts, stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stages, frames = r.stage_ms()
    r.set_timing(False)
    log(f"timed {args.steps} frames in {elapsed * 1e3:.1f} ms; stages {stages}")

    # per-frame device latency (frames alone on the stream)
    lat = []
    r.set_timing(True)
    for _ in range(min(args.steps, 20)):
        r.render(dsrc, out, luts, stream)
        st, nf = r.stage_ms()
        lat.append(sum(st.values()))
    r.set_timing(False)

    ms_per_step = elapsed * 1e3 / args.steps
    value = wl.out_pixels * args.steps / elapsed / 1e6
    per_stage = {k: v / max(frames, 1) for k, v in stages.items()}
    dom = max(per_stage, key=per_stage.get)
    stage_idx = P.PanoramaRender.STAGES.index(dom)
    dom_bytes = r.model_bytes(stage_idx)
    achieved = dom_bytes / (per_stage[dom] * 1e-3) / 1e9
    traffic = read_traffic(wl.name)
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get(dom) if isinstance(traffic, dict) else None,
        "algorithmic_bytes": dom_bytes,
        "stage_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "frame_model_bytes": r.model_bytes(4),
        "frame_model_gbs": round(r.model_bytes(4) / (ms_per_step * 1e-3) / 1e9, 1),
    }
    cpu = None
    if not args.no_cpu and args.cpu_frames > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(wl, srcs, xs, ys, ms, luts, args.cpu_frames)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: 8 cams (2x2560x1440 rect 90, 4x1920x1080 rect 100, "
                        f"2x1280x1280 fisheye 190) -> {wl.dst_w}x{wl.dst_h} equirect, "
                        f"L={r.num_levels}, one render() per frame",
            "frames": args.steps,
            "streams": 1,
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4),
                       "streams": 1},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
