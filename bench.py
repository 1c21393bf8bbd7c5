#!/usr/bin/env python3
"""Throughput of the MI355X stitch path on one GPU.

Default (--config c3), the headline metric's workload: 8 cameras of three sizes
(2 x 2560x1440 rectilinear 90 deg, 4 x 1920x1080 rectilinear 100 deg, 2 x
1280x1280 fisheye 190 deg) remapped and multiband-blended into a 7680x3840
equirect canvas with 9 pyramid levels.  A step stitches one frame of each of S
streams (default 16) with one launch per pyramid stage (pano_batch_render);
inputs (camera frames, maps, seams, weights) are resident in HBM before timing
starts and the timed region is K back-to-back steps on one stream, bracketed by
device synchronisation.  value = output Mpix/s over all streams, from the faster
of the eager batched step and its replay as a captured HIP graph (both are in
the line).  The same line reports one stream alone (render()) and the plain
drop-in (reproject + blender, the reference's GPU API usage).

Other configurations of the same path (one JSON line each, same fields):
  c1  2 fisheyes 1920x1920 195 deg -> 2048x1024, L=5, one frame per step
  c2  6 rectilinear 1920x1080 100 deg -> 3840x1920, L=7, gains as LUTs: the
      plain drop-in (transform + reproject + blender) and the fused render
  c4  the serving loop (pano_server): 16 concurrent streams of seeded rigs (2-8
      cameras, 1280x720 to 3840x2160, fisheye or rectilinear, outputs 2048x1024 /
      3840x1920 / 7680x3840 with 5 / 7 / 9 levels) admitted by an HBM budget; one
      stream leaves and another joins at Poisson times (mean 0.1 s); each stream
      submits 30 frames/s; then the same 16 streams kept saturated
  c5  4 fisheyes 3840x2160 200 deg -> 7680x3840, L=10, gains re-estimated from
      overlap means every frame under a seeded brightness random walk
  lin the mixed8 rig in the render's linear-blend mode (highQualityBlend = 0):
      Gaussian-blurred seam weights, one weighted-remap launch per frame

The roofline object gives the dominant stage's algorithmic bytes (the 16-bit
stage model below) over its mean device time, measured with HIP events on the
stream the kernels run on.  The CPU baseline is the oracle's restatement of
CPUPanoramaRender::render with row-parallel reprojection and per-camera pyramid
threads, timed on a bounded sample of the same frames.
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
STAGES = ("remap", "down", "blend", "final")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_model_bytes(wl, xs, ys, ms, L):
    """Algorithmic bytes of one frame per stage, the 16-bit stage model: every
    plane written once and read once per consuming stage, 3-channel int16 planes
    on each camera's support, int16 weights.  With S_i source pixels, P_i valid
    map points of camera i, q_j = 4^-j and P canvas pixels:
      remap  3 S_i + P_i (8 map + 6 G0 write + 6 G0 read) + q_1 P_i (6 G1 write + 4 alpha)
      down   P_i sum_{1<=j<L} q_j 6 + P_i sum_{2<=j<=L} q_j (6 + 4)
      blend  P_i sum_{1<=j<=L} q_j (6 G + 2 W) + P_i sum_{1<=j<L} q_{j+1} 6
             + P sum_{1<=j<=L} q_j (6 R write + 6 R read + 6 C write) + P sum_{1<=j<L} q_{j+1} 6
      final  P_i (6 + 2) + q_1 P_i 6 + P (12 + 3) + q_1 P 6
    plus, when the masks do not cover the canvas, 4 q_j P per level for the
    weight sums and 1 P for the outside mask.  The sum is
    3 S + P_i K_cam + P K_out with K_cam ~ 38 B and K_out ~ 23 B."""
    S = sum(c.width * c.height for c in wl.cameras)
    Pi = float(sum(int(((x >= 0) & (y >= 0) & (x < c.width) & (y < c.height)).sum())
                   for x, y, c in zip(xs, ys, wl.cameras)))
    P = float(wl.dst_w * wl.dst_h)
    full = bool(np.logical_or.reduce([m > 0 for m in ms]).all())
    q = [4.0 ** -j for j in range(L + 1)]
    st = {
        "remap": 3 * S + Pi * 20 + q[1] * Pi * 10,
        "down": Pi * sum(q[j] * 6 for j in range(1, L)) + Pi * sum(q[j] * 10 for j in range(2, L + 1)),
        "blend": Pi * sum(q[j] * 8 for j in range(1, L + 1)) + Pi * sum(q[j + 1] * 6 for j in range(1, L))
        + P * sum(q[j] * 18 for j in range(1, L + 1)) + P * sum(q[j + 1] * 6 for j in range(1, L)),
        "final": Pi * 8 + q[1] * Pi * 6 + P * 15 + q[1] * P * 6,
    }
    if not full:
        st["blend"] += P * sum(q[j] * 4 for j in range(1, L + 1))
        st["final"] += P * 5
    return st


def cpu_baseline(wl, srcs, xs, ys, ms, luts, frames: int, seconds: float = 10.0):
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    ob = O.Blender()
    ob.prepare(np.stack(ms), wl.levels)
    bgr = [s[..., :3].copy() for s in srcs]
    # whole frames until about `seconds` of CPU work (at least `frames` frames)
    t0 = time.perf_counter()
    done = 0
    while done < frames or time.perf_counter() - t0 < seconds:
        ob.render(bgr, xs, ys, luts, nthreads=threads)
        done += 1
    dt = time.perf_counter() - t0
    frames = done
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


def plain_dropin(wl, xs, ys, ms, dsrc, steps, luts=None, float_weights=False):
    """The reference's GPU API used as CudaPanoramaRender2 does: cudaTransform
    (when gains are given) + cudaReprojectTo16S per camera into full-canvas 16SC4
    images, then CudaTilingMultibandBlendFast::blend (RicohUtil.cpp:1301-1313).
    float_weights: cudaReproject to 8UC4 images and the float-weight blender
    CudaTilingMultibandBlendFast32F instead."""
    import torch

    import panostitcher_amd as P

    if float_weights:
        b = P.CudaTilingMultibandBlendFast32F()
        rep_type, reproject = torch.uint8, P.cuda_reproject
    else:
        b = P.CudaTilingMultibandBlendFast()
        rep_type, reproject = torch.int16, P.cuda_reproject_to16s
    assert b.prepare(ms, wl.levels)
    dx = [torch.from_numpy(x).cuda() for x in xs]
    dy = [torch.from_numpy(y).cuda() for y in ys]
    rep = [torch.empty((wl.dst_h, wl.dst_w, 4), dtype=rep_type, device="cuda") for _ in xs]
    adj = [torch.empty_like(s) for s in dsrc] if luts is not None else dsrc
    out = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")

    def frame():
        for i in range(wl.n):
            if luts is not None:
                P.cuda_transform(dsrc[i], luts[i], adj[i])
            reproject(adj[i], dx[i], dy[i], rep[i])
        b.blend(rep, out)

    dt = timed(frame, steps, 2)
    path = ("transform + " if luts is not None else "")
    path += ("reproject (8UC4) per camera + blender32f.blend, full-canvas float planes" if float_weights
             else "reproject_to16s per camera + blender.blend, full-canvas 16-bit planes")
    return {"value": round(wl.out_pixels * steps / dt / 1e6, 2), "unit": "Mpix/s",
            "ms_per_frame": round(dt * 1e3 / steps, 4), "path": path}


def roofline(model, per_stage_ms, frames_per_launch, own_bytes, traffic):
    """The dominant stage's launch: algorithmic bytes (16-bit stage model) over
    its device time; also the whole frame, and the same stage under the 8-bit
    planes this build stores (pano_render_model_bytes)."""
    dom = max(per_stage_ms, key=per_stage_ms.get)
    launch_s = per_stage_ms[dom] * frames_per_launch * 1e-3
    algo = model[dom] * frames_per_launch
    achieved = algo / launch_s / 1e9 if launch_s else 0.0
    total_s = sum(per_stage_ms.values()) * 1e-3
    frame_gbs = sum(model.values()) / total_s / 1e9 if total_s else 0.0
    own = own_bytes[STAGES.index(dom)] * frames_per_launch
    return {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get(dom) if isinstance(traffic, dict) else None,
        "algorithmic_bytes": int(algo),
        "model": "16-bit stage model (3-channel int16 planes, int16 weights)",
        "stage_ms_per_frame": {k: round(v, 4) for k, v in per_stage_ms.items()},
        "frame_model_bytes": int(sum(model.values())),
        "frame_achieved": round(frame_gbs, 1),
        "frame_frac": round(frame_gbs / HBM_PEAK_GBS, 4),
        "achieved_8bit_model": round(own / launch_s / 1e9, 1) if launch_s else 0.0,
    }


def base_line(metric, value, steps, warmup, ms_per_step, config):
    return {
        "metric": metric,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": config,
    }


def setup(name, scale):
    import panostitcher_amd as P
    from panostitcher_amd import synth

    P.lib()
    wl = synth.workload(name, scale)
    t0 = time.perf_counter()
    xs, ys, ms = synth.maps_and_masks(wl)
    srcs = [synth.bgrx(f) for f in synth.frames(wl)]
    luts = synth.luts(wl)
    log(f"inputs {wl.name} {wl.dst_w}x{wl.dst_h} n={wl.n} in {time.perf_counter() - t0:.1f}s")
    return wl, xs, ys, ms, srcs, luts


def new_render(wl, xs, ys, ms, tile_shape=0):
    import panostitcher_amd as P

    r = P.PanoramaRender()
    r.set_tile_shape(tile_shape)
    ok = r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys,
                        ms, wl.levels)
    if not ok or not r.fused:
        raise SystemExit("prepare failed")
    return r


def per_frame_stages(obj):
    stages, nf = obj.stage_ms()
    return {k: v / max(nf, 1) for k, v in stages.items()}


def run_batched(args):
    """c3: S streams of the mixed8 rig, one launch per stage per step."""
    import torch

    import panostitcher_amd as P

    wl, xs, ys, ms, srcs, luts = setup(args.workload or "mixed8", args.scale)
    S = max(1, args.streams)
    # down tile shape: the project's autotuner times each shape on one rig
    shape, shape_tune = 0, None
    if not args.no_autotune:
        tr = P.PanoramaRender()
        tsrc = [torch.from_numpy(x).cuda() for x in srcs]
        tout = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
        shape, tms = tr.autotune_tile_shape([c.param() for c in wl.cameras],
                                            [(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h),
                                            wl.levels, tsrc, tout, iters=5)
        shape_tune = {"chosen": shape, "tile": list(P.PanoramaRender.TILE_SHAPES[shape]),
                      "ms_per_frame": dict(zip(["x".join(map(str, t)) for t in P.PanoramaRender.TILE_SHAPES], tms))}
        del tr, tsrc, tout
        log(f"tile shape autotune {shape_tune}")
    renders = []
    t0 = time.perf_counter()
    for _ in range(S):
        renders.append(new_render(wl, xs, ys, ms, shape))
    r0 = renders[0]
    prep_each = (time.perf_counter() - t0) / S
    log(f"prepared {S} renders in {time.perf_counter() - t0:.1f}s, levels {r0.num_levels}, "
        f"{r0.mem_bytes() / 1e9:.2f} GB each; last prepare by phase (ms) {renders[-1].prepare_ms()}")
    dsrc = [torch.from_numpy(s).cuda() for s in srcs]
    stream_src = [dsrc] + [[t.clone() for t in dsrc] for _ in range(S - 1)]
    outs = [torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda") for _ in range(S)]
    stream = torch.cuda.current_stream()

    dt1 = timed(lambda: r0.render(dsrc, outs[0], luts, stream), args.steps, args.warmup)
    single = {"value": round(wl.out_pixels * args.steps / dt1 / 1e6, 2), "unit": "Mpix/s",
              "ms_per_frame": round(dt1 * 1e3 / args.steps, 4)}
    log(f"single stream {single}")

    batch = P.Batch(renders)
    lut_list = [luts] * S
    tune = None
    if not args.no_autotune:
        chosen, tms = batch.autotune(stream_src, outs, lut_list, stream, iters=3)
        tune = {"chosen": dict(zip(STAGES, chosen)),
                "ms_per_step": {f"order{m}": dict(zip(STAGES, [round(float(v), 4) for v in tms[m]]))
                                for m in range(len(tms))},
                "orders": "0: block = job; 1: one contiguous run of jobs per XCD"}
        log(f"autotune {tune}")

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
    per_step = per_frame_stages(batch)
    lat = []
    for _ in range(min(args.steps, 20)):
        step()
        st, _ = batch.stage_ms()
        lat.append(sum(st.values()))
    batch.set_timing(False)
    log(f"batched {S} streams: {args.steps} steps in {elapsed * 1e3:.1f} ms; per step {per_step}")
    eager = {"value": round(S * wl.out_pixels * args.steps / elapsed / 1e6, 2), "unit": "Mpix/s",
             "ms_per_step": round(elapsed * 1e3 / args.steps, 4)}

    # the same step captured as a HIP graph and replayed (frames refreshed in place)
    graph = None
    if not args.no_graph:
        side = torch.cuda.Stream()   # a capturable (non-null) stream, default priority
        torch.cuda.synchronize()
        batch.capture(stream_src, outs, lut_list, side)
        dtg = timed(lambda: batch.replay(lut_list, side), args.steps, args.warmup)
        graph = {"value": round(S * wl.out_pixels * args.steps / dtg / 1e6, 2), "unit": "Mpix/s",
                 "ms_per_step": round(dtg * 1e3 / args.steps, 4)}
        log(f"graph replay {graph}")
        if dtg < elapsed:
            elapsed = dtg

    # the encoder hand-off: NV12 planes written by the level-0 kernel, against
    # the BGRx canvas followed by the standalone conversion (cvtBGR32ToNV12)
    enc = None
    if not args.no_encoder:
        yps = [torch.empty((wl.dst_h, wl.dst_w), dtype=torch.uint8, device="cuda") for _ in range(S)]
        uvs = [torch.empty((wl.dst_h // 2, wl.dst_w), dtype=torch.uint8, device="cuda") for _ in range(S)]
        nv12 = [P.Output.nv12(y, uv) for y, uv in zip(yps, uvs)]
        dtn = timed(lambda: batch.render_to(stream_src, nv12, lut_list, stream), args.steps, args.warmup)

        def then_convert():
            batch.render(stream_src, outs, lut_list, stream)
            for k in range(S):
                P.cvt_bgr32_to_nv12(outs[k], yps[k], uvs[k], stream)

        dtc = timed(then_convert, max(3, args.steps // 2), 1)

        # the planes downloaded to page-locked host buffers after each step (the
        # PCIe-inclusive rate of a task that encodes on the host)
        hy = [torch.empty((wl.dst_h, wl.dst_w), dtype=torch.uint8, pin_memory=True) for _ in range(S)]
        huv = [torch.empty((wl.dst_h // 2, wl.dst_w), dtype=torch.uint8, pin_memory=True) for _ in range(S)]

        def with_d2h():
            batch.render_to(stream_src, nv12, lut_list, stream)
            for k in range(S):
                hy[k].copy_(yps[k], non_blocking=True)
                huv[k].copy_(uvs[k], non_blocking=True)

        dtd = timed(with_d2h, max(3, args.steps // 2), 1)

        # the same with the downloads on a copy stream, overlapping the next
        # step's render (two sets of device planes, events between the streams)
        yb = [torch.empty_like(t) for t in yps]
        uvb = [torch.empty_like(t) for t in uvs]
        sets = [(yps, uvs, nv12), (yb, uvb, [P.Output.nv12(y, uv) for y, uv in zip(yb, uvb)])]
        copy_stream = torch.cuda.Stream()   # default priority
        copied = [None, None]
        state = {"k": 0}

        def overlapped():
            cur = state["k"] & 1
            state["k"] += 1
            ys_, uvs_, outs_ = sets[cur]
            if copied[cur] is not None:
                stream.wait_event(copied[cur])      # the set's previous download is done
            batch.render_to(stream_src, outs_, lut_list, stream)
            rendered = torch.cuda.Event()
            rendered.record(stream)
            copy_stream.wait_event(rendered)
            with torch.cuda.stream(copy_stream):
                for k in range(S):
                    hy[k].copy_(ys_[k], non_blocking=True)
                    huv[k].copy_(uvs_[k], non_blocking=True)
            copied[cur] = torch.cuda.Event()
            copied[cur].record(copy_stream)

        dto = timed(overlapped, max(3, args.steps // 2), 2)
        enc = {"nv12_fused": {"value": round(S * wl.out_pixels * args.steps / dtn / 1e6, 2), "unit": "Mpix/s",
                              "ms_per_step": round(dtn * 1e3 / args.steps, 4)},
               "bgrx_then_nv12": {"value": round(S * wl.out_pixels * max(3, args.steps // 2) / dtc / 1e6, 2),
                                  "unit": "Mpix/s", "ms_per_step": round(dtc * 1e3 / max(3, args.steps // 2), 4)},
               "nv12_fused_then_d2h_pinned": {
                   "value": round(S * wl.out_pixels * max(3, args.steps // 2) / dtd / 1e6, 2), "unit": "Mpix/s",
                   "ms_per_step": round(dtd * 1e3 / max(3, args.steps // 2), 4),
                   "host_GBps": round(S * wl.out_pixels * 1.5 * max(3, args.steps // 2) / dtd / 1e9, 2),
                   "note": "PCIe-inclusive: NV12 planes copied to page-locked host memory each step"},
               "nv12_fused_d2h_overlapped": {
                   "value": round(S * wl.out_pixels * max(3, args.steps // 2) / dto / 1e6, 2), "unit": "Mpix/s",
                   "ms_per_step": round(dto * 1e3 / max(3, args.steps // 2), 4),
                   "host_GBps": round(S * wl.out_pixels * 1.5 * max(3, args.steps // 2) / dto / 1e9, 2),
                   "note": "downloads on a copy stream overlapping the next step's render"}}
        log(f"encoder hand-off {enc}")

    ms_per_step = elapsed * 1e3 / args.steps
    value = S * wl.out_pixels * args.steps / elapsed / 1e6
    per_frame = {k: v / S for k, v in per_step.items()}
    model = stage_model_bytes(wl, xs, ys, ms, r0.num_levels)
    own = [r0.model_bytes(k) for k in range(4)]
    roof = roofline(model, per_frame, S, own, read_traffic(wl.name))
    plain = plain32 = None
    if not args.no_plain:
        log("plain drop-in ...")
        plain = plain_dropin(wl, xs, ys, ms, dsrc, max(3, args.steps // 4), luts)
        log(f"plain {plain}")
        plain32 = plain_dropin(wl, xs, ys, ms, dsrc, max(3, args.steps // 4), luts, float_weights=True)
        log(f"plain float weights {plain32}")
    cpu = None
    if not args.no_cpu and args.cpu_frames > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(wl, srcs, xs, ys, ms, luts, args.cpu_frames, args.cpu_seconds)
    line = base_line(METRIC, value, args.steps, args.warmup, ms_per_step, {
        "workload": f"{wl.name}: 8 cams (2x2560x1440 rect 90, 4x1920x1080 rect 100, "
                    f"2x1280x1280 fisheye 190) -> {wl.dst_w}x{wl.dst_h} equirect, "
                    f"L={r0.num_levels}; a step stitches one frame of each of {S} streams "
                    f"with one launch per pyramid stage",
        "streams": S,
        "frames_per_step": S,
        "parallelism": "single GPU",
    })
    line["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 4),
                          "p99": round(float(np.percentile(lat, 99)), 4),
                          "streams": S,
                          "note": "device time of a step; each stream's frame completes with it"}
    line["prepare"] = {"s_per_rig": round(prep_each, 3), "last_ms_by_phase": renders[-1].prepare_ms(),
                       "note": "host maps uploaded; tables, pyramids and map storage built on the device"}
    # a rig prepared from its camera parameters (maps generated on the device),
    # with the tables on the device and, for comparison, on the host
    params = [c.param() for c in wl.cameras]
    sizes = [(c.width, c.height) for c in wl.cameras]
    for on, key in ((True, "from_params_s"), (False, "from_params_host_tables_s")):
        pr = P.PanoramaRender()
        pr.set_prepare_on_device(on)
        t1 = time.perf_counter()
        if not pr.prepare(params, sizes, (wl.dst_w, wl.dst_h), wl.levels):
            raise SystemExit("prepare from params failed")
        line["prepare"][key] = round(time.perf_counter() - t1, 3)
        line["prepare"][key.replace("_s", "_ms_by_phase")] = pr.prepare_ms()
        del pr
    log(f"prepare {line['prepare']}")
    line["single_stream"] = single
    line["batched_eager"] = eager
    line["batched_graph"] = graph
    line["autotune"] = tune
    line["tile_shape_autotune"] = shape_tune
    line["encoder_handoff"] = enc
    line["plain_dropin"] = plain
    line["plain_dropin_f32"] = plain32
    if plain:
        line["vs_plain_dropin"] = round(value / plain["value"], 3)
        line["vs_plain_dropin_f32"] = round(value / plain32["value"], 3)
    line["roofline"] = roof
    line["cpu_baseline"] = cpu
    return line


def timed_stages(r, fn, steps):
    """Per-frame device latencies (one synchronising frame at a time), then the
    mean per-stage device time over back-to-back frames."""
    r.set_timing(True)
    r.stage_ms()
    lat = []
    for _ in range(min(steps, 50)):
        fn()
        st, _ = r.stage_ms()
        lat.append(sum(st.values()))
    for _ in range(steps):
        fn()
    per_frame = per_frame_stages(r)
    r.set_timing(False)
    return lat, per_frame


def run_single(args, name, label, with_plain):
    """c1 / c2: one rig, one frame per step through render()."""
    import torch

    wl, xs, ys, ms, srcs, luts = setup(name, args.scale)
    r = new_render(wl, xs, ys, ms)
    dsrc = [torch.from_numpy(s).cuda() for s in srcs]
    out = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def fn():
        r.render(dsrc, out, luts, stream)

    dt = timed(fn, args.steps, args.warmup)
    lat, per_frame = timed_stages(r, fn, args.steps)
    eager = {"value": round(wl.out_pixels * args.steps / dt / 1e6, 2), "unit": "Mpix/s",
             "ms_per_frame": round(dt * 1e3 / args.steps, 4)}
    # the frame's launch sequence captured once as a HIP graph and replayed (a
    # single-rig batch): small canvases are launch-bound
    graph = None
    if not args.no_graph:
        import panostitcher_amd as P

        b = P.Batch()
        b.set([r])
        side = torch.cuda.Stream()   # a capturable (non-null) stream, default priority
        torch.cuda.synchronize()
        b.capture([dsrc], [out], [luts], side)
        dtg = timed(lambda: b.replay([luts], side), args.steps, args.warmup)
        graph = {"value": round(wl.out_pixels * args.steps / dtg / 1e6, 2), "unit": "Mpix/s",
                 "ms_per_frame": round(dtg * 1e3 / args.steps, 4)}
        log(f"graph replay {graph}")
        if dtg < dt:
            dt = dtg
    value = wl.out_pixels * args.steps / dt / 1e6
    model = stage_model_bytes(wl, xs, ys, ms, r.num_levels)
    own = [r.model_bytes(k) for k in range(4)]
    line = base_line(f"stitched output Mpix/s ({label})", value, args.steps, args.warmup,
                     dt * 1e3 / args.steps, {"workload": f"{wl.name} -> {wl.dst_w}x{wl.dst_h}, "
                                             f"L={r.num_levels}, {wl.n} cams, one frame per step",
                                             "streams": 1, "parallelism": "single GPU"})
    line["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 4),
                          "p99": round(float(np.percentile(lat, 99)), 4), "streams": 1}
    line["roofline"] = roofline(model, per_frame, 1, own, read_traffic(wl.name))
    line["eager"] = eager
    line["graph"] = graph
    if with_plain:
        line["plain_dropin"] = plain_dropin(wl, xs, ys, ms, dsrc, max(3, args.steps // 2), luts)
        # the same drop-in with the float-weight blender (float32 weights)
        line["plain_dropin_f32"] = plain_dropin(wl, xs, ys, ms, dsrc, max(3, args.steps // 2), luts,
                                                float_weights=True)
    line["cpu_baseline"] = None
    if not args.no_cpu and args.cpu_frames > 0:
        line["cpu_baseline"] = cpu_baseline(wl, srcs, xs, ys, ms, luts, args.cpu_frames, args.cpu_seconds)
    return line


def run_linear(args):
    """lin: the render's linear-blend mode (highQualityBlend = 0) on the mixed8 rig:
    Gaussian-blurred seam weights at prepare, then one weighted-remap launch per
    frame (pano_linear_render)."""
    import torch

    import panostitcher_amd as P

    wl, xs, ys, ms, srcs, luts = setup(args.workload or "mixed8", args.scale)
    radius = args.linear_radius
    lr = P.LinearRender()
    t0 = time.perf_counter()
    assert lr.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs, ys, ms, radius)
    prep_s = time.perf_counter() - t0
    log(f"linear prepare (radius {radius}) in {prep_s:.2f}s")
    wts = [lr.weights(i) for i in range(wl.n)]   # read once, before any frame
    dsrc = [torch.from_numpy(s).cuda() for s in srcs]
    out = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    dt = timed(lambda: lr.render(dsrc, out, luts, stream), args.steps, args.warmup)
    # device time of the frames (one launch each without LUTs) with events on the
    # stream, the queue kept busy by frames issued before the first event
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(4):
        lr.render(dsrc, out, luts, stream)
    ev0.record(stream)
    for _ in range(args.steps):
        lr.render(dsrc, out, luts, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    kms = ev0.elapsed_time(ev1) / args.steps
    # algorithmic bytes per frame: a weight and a map point per listed (camera, pixel),
    # the source pixel's 3 bytes per sampled one, the 4-byte output per canvas pixel
    listed = sampled = 0
    for i in range(wl.n):
        w = wts[i]
        sampled += int(((w > 0) & (xs[i] >= 0) & (ys[i] >= 0) & (xs[i] < wl.cameras[i].width)
                        & (ys[i] < wl.cameras[i].height)).sum())
        m = w > 0                                              # tiles list a camera where its weight is
        TH = 16                                                # non-zero; 64 x 16 canvas tiles
        th, tw = (wl.dst_h + TH - 1) // TH, (wl.dst_w + 63) // 64
        pad = np.zeros((th * TH, tw * 64), bool)
        pad[:wl.dst_h, :wl.dst_w] = m
        listed += int(pad.reshape(th, TH, tw, 64).any(axis=(1, 3)).sum()) * 64 * TH
    src_px = sum(c.width * c.height for c in wl.cameras)
    bytes_frame = 12 * listed + 3 * sampled + 3 * src_px + 4 * wl.out_pixels
    achieved = bytes_frame / (kms * 1e-3) / 1e9
    value = wl.out_pixels * args.steps / dt / 1e6
    line = base_line(f"stitched output Mpix/s (linear blend, {wl.name} -> {wl.dst_w}x{wl.dst_h})", value,
                     args.steps, args.warmup, dt * 1e3 / args.steps,
                     {"workload": f"{wl.name} -> {wl.dst_w}x{wl.dst_h}, linear blend radius {radius}, "
                                  f"{wl.n} cams, one frame per step", "streams": 1,
                      "parallelism": "single GPU"})
    line["prepare_s"] = round(prep_s, 3)
    lp = P.LinearRender()
    t0 = time.perf_counter()
    assert lp.prepare([c.param() for c in wl.cameras], [(c.width, c.height) for c in wl.cameras],
                      (wl.dst_w, wl.dst_h), radius)
    line["prepare_from_params_s"] = round(time.perf_counter() - t0, 3)
    del lp
    line["roofline"] = {"bound": "hbm", "kernel": "linear_render", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": (read_traffic(f"{wl.name}_linear") or {}).get("linear_render"),
                        "algorithmic_bytes": int(bytes_frame),
                        "model": "4 B weight and 8 B map per listed camera pixel + 3 B source per sample "
                                 "+ 3 B per source pixel + 4 B output per canvas pixel",
                        "kernel_ms": round(kms, 4)}
    line["cpu_baseline"] = None
    if not args.no_cpu and args.cpu_frames > 0:
        from oracle import oracle as O

        threads = min(16, os.cpu_count() or 1)
        wts = np.stack(wts)
        bgr = [s[..., :3].copy() for s in srcs]
        t0 = time.perf_counter()
        nf = 0
        while nf < args.cpu_frames or time.perf_counter() - t0 < args.cpu_seconds:
            O.linear_render(bgr, xs, ys, wts, luts, nthreads=threads)
            nf += 1
        cdt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(wl.out_pixels * nf / cdt / 1e6, 3), "unit": "Mpix/s",
                                "cores": threads, "kind": "port",
                                "sample": f"{nf} frame(s), oracle linear-branch restatement "
                                          f"(reprojection row-parallel, {threads} threads), {cdt:.1f} s"}
    return line


def run_autogain(args):
    """c5: fisheye4, gains re-estimated from overlap means every frame."""
    import torch

    wl, xs, ys, ms, srcs, _ = setup("fisheye4", args.scale)
    r = new_render(wl, xs, ys, ms)
    rng = np.random.default_rng(20261203 + 5000)
    F = 8
    k = np.ones(wl.n)
    base = [torch.from_numpy(s).cuda() for s in srcs]
    frames = []
    for _ in range(F):
        # seeded random walk in brightness, N(0, 0.01) per camera per frame
        k = np.clip(k + rng.normal(0, 0.01, size=wl.n), 0.5, 1.5)
        fr = []
        for i, t in enumerate(base):
            v = (t.float() * float(k[i])).round().clamp(0, 255).to(torch.uint8)
            v[..., 3] = 0
            fr.append(v)
        frames.append(fr)
    out = torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    it = [0]

    def fn():
        r.render_autogain(frames[it[0] % F], out, tint=True, stream=stream)
        it[0] += 1

    dt = timed(fn, args.steps, args.warmup)
    lat, per_frame = timed_stages(r, fn, args.steps)
    gains, red, blue, anchor = r.gains()
    value = wl.out_pixels * args.steps / dt / 1e6
    model = stage_model_bytes(wl, xs, ys, ms, r.num_levels)
    own = [r.model_bytes(k) for k in range(4)]
    line = base_line("stitched output Mpix/s (4 fisheye 3840x2160 -> 7680x3840, per-frame gains)",
                     value, args.steps, args.warmup, dt * 1e3 / args.steps,
                     {"workload": f"fisheye4 -> {wl.dst_w}x{wl.dst_h}, L={r.num_levels}, gains and "
                                  f"tint re-estimated every frame (overlap means, on the device), "
                                  f"brightness random walk N(0, 0.01) over {F} resident frames",
                      "streams": 1, "parallelism": "single GPU"})
    line["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 4),
                          "p99": round(float(np.percentile(lat, 99)), 4), "streams": 1,
                          "budget_60fps": 16.667}
    line["roofline"] = roofline(model, per_frame, 1, own, None)
    line["roofline"]["note"] = "remap stage time includes the gain statistics and solve; the remap applies the LUTs"
    line["last_gains"] = {"exposure": [round(float(g), 5) for g in gains], "anchor": anchor}
    line["cpu_baseline"] = None
    return line


def run_served(args):
    """c4: the serving loop under a paced 30 fps load with stream churn, then
    saturated."""
    import concurrent.futures as cf

    import torch

    import panostitcher_amd as P
    from panostitcher_amd import synth

    R, S, fps = 24, 16, 30.0
    budget = int(args.hbm_budget_gb * 2 ** 30)
    t0 = time.perf_counter()

    def prep(k):
        wl = synth.served_rig(k)
        xs, ys, ms = synth.maps_and_masks(wl)
        r = P.PanoramaRender()
        ok = r.prepare_maps([(c.width, c.height) for c in wl.cameras], (wl.dst_w, wl.dst_h), xs,
                            ys, ms, wl.levels)
        if not ok or not r.fused:
            raise SystemExit(f"prepare of rig {k} failed")
        return wl, r

    with cf.ThreadPoolExecutor(4) as ex:
        rigs = list(ex.map(prep, range(R)))
    log(f"prepared {R} rigs in {time.perf_counter() - t0:.1f}s, "
        f"{sum(r.mem_bytes() for _, r in rigs) / 1e9:.1f} GB resident")
    gen = torch.Generator(device="cuda").manual_seed(20261203 + 4000)
    srcs, outs = [], []
    for wl, _ in rigs:
        fr = []
        for c in wl.cameras:
            t = torch.randint(0, 256, (c.height, c.width, 4), dtype=torch.uint8, device="cuda",
                              generator=gen)
            t[..., 3] = 0
            fr.append(t)
        srcs.append(fr)
        outs.append([torch.empty((wl.dst_h, wl.dst_w, 4), dtype=torch.uint8, device="cuda")
                     for _ in range(3)])
    torch.cuda.synchronize()
    srv = P.Server(budget, max_inflight=2)
    rng = np.random.default_rng(20261203 + 4001)
    active, pending_join = {}, []
    rejected = 0
    for k in range(R):
        if len(active) == S:
            break
        sid = srv.admit(rigs[k][1])
        if sid is None:
            rejected += 1
            continue
        active[k] = sid
    sent = {}
    frame_no = {k: 0 for k in range(R)}

    def submit(k, now):
        sid = active[k]
        i = frame_no[k]
        frame_no[k] += 1
        sent[(sid, i)] = (k, now)
        srv.submit(sid, srcs[k], outs[k][i % 3], tag=i)

    def collect(done, lat, pix):
        for sid, tag, ms, st, bf in done:
            if st != 0:
                raise SystemExit(f"frame failed: status {st}")
            k, _ = sent.pop((sid, tag))
            lat.append(ms)
            pix.append(rigs[k][0].out_pixels)
            batch_sizes.append(bf)

    # warm-up: one frame of every rig alone
    for k in list(active):
        submit(k, time.perf_counter())
    srv.drain()
    batch_sizes = []
    srv.poll(1024)
    sent.clear()

    # paced phase: 30 fps per stream, one leave + one join at Poisson times
    lat, pix = [], []
    T = args.serve_seconds
    start = time.perf_counter()
    due = {k: start + rng.uniform(0, 1 / fps) for k in active}
    next_event = start + rng.exponential(0.1)
    joins = leaves = 0
    while True:
        now = time.perf_counter()
        if now - start >= T:
            break
        if now >= next_event:
            k_out = int(rng.choice(list(active)))
            srv.retire(active.pop(k_out))
            leaves += 1
            idle = [k for k in range(R) if k not in active and k != k_out]
            rng.shuffle(idle)
            for k_in in idle:
                sid = srv.admit(rigs[k_in][1])
                if sid is None:
                    rejected += 1
                    continue
                active[k_in] = sid
                due[k_in] = time.perf_counter()
                joins += 1
                break
            next_event += rng.exponential(0.1)
        for k in list(active):
            while due[k] <= now:
                submit(k, now)
                due[k] += 1 / fps
        collect(srv.poll(1024, wait_ms=1), lat, pix)
    srv.drain()
    collect(srv.poll(4096), lat, pix)
    paced_s = time.perf_counter() - start
    paced = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
             "p99_ms": round(float(np.percentile(lat, 99)), 3),
             "frames": len(lat), "seconds": round(paced_s, 2),
             "mpix_s": round(sum(pix) / paced_s / 1e6, 1),
             "joins": joins, "leaves": leaves, "rejected_admissions": rejected,
             "mean_batch_frames": round(float(np.mean(batch_sizes)), 2) if batch_sizes else 0}
    log(f"paced {paced}")

    # saturated phase: the current streams keep two frames queued each
    lat, pix, batch_sizes = [], [], []
    for k in list(active):
        submit(k, 0)
        submit(k, 0)
    start = time.perf_counter()
    while time.perf_counter() - start < args.serve_seconds / 2:
        done = srv.poll(1024, wait_ms=2)
        for sid, tag, ms, st, bf in done:
            k = sent[(sid, tag)][0]
            if k in active:
                submit(k, 0)
        collect(done, lat, pix)
    srv.drain()
    collect(srv.poll(4096), lat, pix)
    sat_s = time.perf_counter() - start
    stats = srv.stats()
    srv.close()
    value = sum(pix) / sat_s / 1e6
    line = base_line("served aggregate output Mpix/s (16 streams, mixed rigs) + p99 frame latency",
                     value, len(lat), 0, sat_s * 1e3 / max(len(lat), 1),
                     {"workload": f"served mix: {R} seeded rigs (2-8 cams, 1280x720..3840x2160, "
                                  f"fisheye/rect) -> 2048x1024 / 3840x1920 / 7680x3840, L 5/7/9; "
                                  f"{S} concurrent streams, HBM budget {args.hbm_budget_gb} GiB",
                      "streams": len(active), "parallelism": "single GPU"})
    line["latency_ms"] = {"p50": paced["p50_ms"], "p99": paced["p99_ms"], "streams": S,
                          "load": "30 frames/s per stream with churn"}
    line["paced"] = paced
    line["saturated"] = {"p50_ms": round(float(np.percentile(lat, 50)), 3),
                         "p99_ms": round(float(np.percentile(lat, 99)), 3),
                         "frames": len(lat), "seconds": round(sat_s, 2),
                         "mean_batch_frames": round(float(np.mean(batch_sizes)), 2)}
    line["server"] = stats
    line["cpu_baseline"] = None
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "lin"])
    ap.add_argument("--linear-radius", type=int, default=96)
    ap.add_argument("--serve-seconds", type=float, default=6.0)
    ap.add_argument("--hbm-budget-gb", type=float, default=10.0)
    ap.add_argument("--workload", default=None)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--streams", type=int, default=16)
    ap.add_argument("--cpu-frames", type=int, default=1, help="at least this many CPU baseline frames")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: whole frames until about this much CPU time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-encoder", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the stitch path shards by stream, one rig per GPU, "
              "and this build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import panostitcher_amd as P

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
    if args.config == "c3":
        line = run_batched(args)
    elif args.config == "c1":
        line = run_single(args, "fisheye2", "2 fisheye 1920x1920 -> 2048x1024", False)
    elif args.config == "c2":
        line = run_single(args, "rect6", "6 rect 1920x1080 -> 3840x1920, gains", True)
    elif args.config == "c4":
        line = run_served(args)
    elif args.config == "lin":
        line = run_linear(args)
    else:
        line = run_autogain(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
