#!/usr/bin/env python3
"""Throughput of the MI355X stitch path on one GPU.

Workload of the headline metric: 8 cameras of three sizes (2 x 2560x1440
rectilinear 90 deg, 4 x 1920x1080 rectilinear 100 deg, 2 x 1280x1280 fisheye
190 deg) remapped and multiband-blended into a 7680x3840 equirect canvas with
9 pyramid levels.  A step stitches one frame of each of S streams (default
16) with one launch per pyramid stage (pano_batch_render); inputs (camera
frames, maps, seams, weights) are resident in HBM before timing starts and the
timed region is K back-to-back steps on one stream, bracketed by device
synchronisation.  value = output Mpix/s over all streams.  The same line
reports one stream alone (render()) and the plain drop-in (reproject +
blender, the reference's GPU API usage).

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


def timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def plain_dropin(wl, xs, ys, ms, dsrc, steps):
    """The reference's GPU API used as CudaPanoramaRender2 does: cudaTransform +
    cudaReprojectTo16S per camera into full-canvas 16SC4 images, then
    CudaTilingMultibandBlendFast::blend (RicohUtil.cpp:1301-1313)."""
    import torch

    import panostitcher_amd as P

    b = P.CudaTilingMultibandBlendFast()
    assert b.prepare(ms, wl.levels)
    dx = [torch.from_numpy(x).cuda() for x in xs]
    dy = [torch.from_numpy(y).cuda() for y in ys]
    rep = [torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.int16, device="cuda") for _ in xs]
    out = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")

    def frame():
        for i in range(wl.n):
            P.cuda_reproject_to16s(dsrc[i], dx[i], dy[i], rep[i])
        b.blend(rep, out)

    dt = timed(frame, steps, 2)
    return {"value": round(wl.out_pixels * steps / dt / 1e6, 2), "unit": "Mpix/s",
            "ms_per_frame": round(dt * 1e3 / steps, 4),
            "path": "reproject_to16s per camera + blender.blend, full-canvas 16-bit planes"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mixed8")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--streams", type=int, default=16)
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the stitch path shards by stream, one rig per GPU, "
              "and this build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import panostitcher_amd as P
    from panostitcher_amd import synth

    P.lib()
    if args.calib_copy_mib > 0:
        # a copy kernel of known size: reads and writes calib_copy_mib MiB once
        n = args.calib_copy_mib << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy of {args.calib_copy_mib} MiB")
        del a, b
    wl = synth.workload(args.workload, args.scale)
    t0 = time.perf_counter()
    xs, ys, ms = synth.maps_and_masks(wl)
    frames = synth.frames(wl)
    srcs = [synth.bgrx(f) for f in frames]
    luts = synth.luts(wl)
    log(f"inputs {wl.name} {wl.dst_w}x{wl.dst_h} n={wl.n} in {time.perf_counter() - t0:.1f}s")
    S = max(1, args.streams)
    renders = []
    t0 = time.perf_counter()
    for k in range(S):
        r = P.PanoramaRender()
        ok = r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs,
                            ys, ms, wl.levels)
        if not ok or not r.fused:
            raise SystemExit("prepare failed")
        renders.append(r)
    log(f"prepared {S} renders in {time.perf_counter() - t0:.1f}s, levels {renders[0].num_levels}, "
        f"{renders[0].mem_bytes() / 1e9:.2f} GB each")
    dsrc = [torch.from_numpy(s).cuda() for s in srcs]
    # every stream has its own resident frame buffers and canvas
    stream_src = [dsrc] + [[t.clone() for t in dsrc] for _ in range(S - 1)]
    outs = [torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda") for _ in range(S)]
    stream = torch.cuda.current_stream()

    # single stream: render() of one rig per call
    r0 = renders[0]
    dt1 = timed(lambda: r0.render(dsrc, outs[0], luts, stream), args.steps, args.warmup)
    single = {"value": round(wl.out_pixels * args.steps / dt1 / 1e6, 2), "unit": "Mpix/s",
              "ms_per_frame": round(dt1 * 1e3 / args.steps, 4)}
    log(f"single stream {single}")

    # batched: one launch per stage for the frames of all S streams
    batch = P.Batch(renders)
    lut_list = [luts] * S

    def step():
        batch.render(stream_src, outs, lut_list, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    batch.set_timing(True)
    batch.stage_ms()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stages, nframes = batch.stage_ms()
    # per-batch device latency: every stream's frame is ready when its batch ends
    lat = []
    for _ in range(min(args.steps, 20)):
        step()
        st, _ = batch.stage_ms()
        lat.append(sum(st.values()))
    batch.set_timing(False)
    log(f"batched {S} streams: {args.steps} steps in {elapsed * 1e3:.1f} ms; stages {stages}")

    ms_per_step = elapsed * 1e3 / args.steps
    value = S * wl.out_pixels * args.steps / elapsed / 1e6
    per_stage = {k: v / max(nframes, 1) for k, v in stages.items()}
    dom = max(per_stage, key=per_stage.get)
    stage_idx = P.PanoramaRender.STAGES.index(dom)
    dom_bytes = S * r0.model_bytes(stage_idx)
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
        "step_model_bytes": S * r0.model_bytes(4),
        "step_model_gbs": round(S * r0.model_bytes(4) / (ms_per_step * 1e-3) / 1e9, 1),
    }
    plain = None
    if not args.no_plain:
        log("plain drop-in ...")
        plain = plain_dropin(wl, xs, ys, ms, dsrc, max(3, args.steps // 4))
        log(f"plain {plain}")
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
                        f"L={r0.num_levels}; a step stitches one frame of each of {S} streams "
                        f"with one launch per pyramid stage",
            "streams": S,
            "frames_per_step": S,
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4),
                       "streams": S,
                       "note": "device time of a step; each stream's frame completes with it"},
        "single_stream": single,
        "plain_dropin": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
