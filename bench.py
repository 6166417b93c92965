#!/usr/bin/env python3
"""BM3Dv2 throughput on one MI355X.

A step is one pass of the filter over the whole synthetic clip resident in
HBM: for the default workload, 32 frames of 1920x1080 GrayS float32 with
sigma 3, block_step 4, bm_range 16, radius 0 (plain FP32 form).  `value` is
frames per second over the timed steps (each step's launch sequence replayed
from the library's HIP graph at radius 0; launched directly at radius > 0,
the library's default there).  Per-kernel durations come from HIP events
the library records around every launch on the launch stream, over further
steps launched directly; the dominant kernel's algorithmic work divided by
its average duration gives the roofline fraction.  The CPU baseline is the
C restatement of the reference shaders (the oracle) timed on a bounded
sample of the same workload on the host cores.

Beside the headline line's `value`, the same JSON line carries the other
forms (`forms`):

* ``fwd_bwd``: the differentiable form, 16 frames of 512x512 (sigma 3,
  block_step 4, bm_range 9, radius 0): forward with saved groups, masks and
  weights, then the backward pass, frames/s of forward+backward;
* ``temporal_r2``: 32 frames of 1920x1080 luma shifted by (1.5t, 0.75t) px,
  radius 2, ps_num 2, ps_range 7, extractor_exp 2;
* ``host_io``: the headline workload through the host-buffer entry point
  (the VapourSynth getFrame path): upload, filter and download per frame
  request, so PCIe-inclusive;
* ``pipe``: the vspipe counterpart (``vsfeel.pipe_bench``) with the
  reference harness's BM3Dv2 defaults (sigma 0.7, radius 2, bm_range 16,
  ps_range 7, block_step 4): 1080p frames requested from 4 threads through
  the host-buffer path;
* ``mfma_fp16`` / ``mfma_bf16``: the precision form (block-matching
  distances of the pixels rounded once to FP16 / BF16 on MFMA, shortlists
  rescored in FP32) on the temporal workload (1080p, radius 2) with its
  ratio to the plain form's frames/s, and on the headline workload
  (``radius0``); rooflines of the matching against the FP16/BF16 matrix peak;
* ``c4``: 16 frames of 3840x2160 (radius 1, ps 2/7, sigma 5) in the FP32
  form and both precision forms, with frame 0 compared against the FP64
  oracle (max-abs, PSNR, denoising PSNR against the clean frame, share of
  groups identical to the plain form's and to the oracle's).

The line also carries ``aggregation``: the tile reduction (the fused
aggregation at radius 0) against the HBM peak.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
                    [--forms plain,mfma,fwd_bwd,temporal_r2,host_io,pipe,c4] [--cpu-seconds S]
                    [--copy-calibration-mb MB]
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vapoursynth-feel_amd"))
sys.path.insert(0, ROOT)

METRIC = "BM3Dv2 frames/sec at 1080p GrayS (plain vs MFMA-distance), fwd+bwd fps, %peak"
PEAK_FP32_TFLOPS = 157.3
PEAK_MFMA16_TFLOPS = 2500.0   # dense FP16 / BF16 matrix peak
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bm3d_match_1080p_hbm_traffic.md")


def ref_positions(w, h, step):
    nx = (w - 8 + step + step - 1) // step
    ny = (h - 8 + step + step - 1) // step
    xs = np.minimum(np.arange(nx) * step, w - 8)
    ys = np.minimum(np.arange(ny) * step, h - 8)
    return xs, ys


def matching_flops(w, h, step, bm_range):
    """Spatial block-matching FLOP per frame, the reference's own count: 24
    FLOP per 8-row column SSD, (rw + 7) column SSDs per scanned row, 7 FLOP per
    candidate for the tree sum (rows x (24 (rw + 7) + 7 rw) per ref block)."""
    xs, ys = ref_positions(w, h, step)
    rw = np.minimum(xs + bm_range, w - 8) - np.maximum(xs - bm_range, 0) + 1
    rh = np.minimum(ys + bm_range, h - 8) - np.maximum(ys - bm_range, 0) + 1
    per_x = 24 * (rw + 7) + 7 * rw
    return float(np.sum(rh) * np.sum(per_x))


TRANSFORM_FLOP_PER_GROUP = 192 * 44 + 192 * 51 + 512 * 3  # fwd + inv DCTs + threshold


def temporal_flops(w, h, step, radius, ps_num, ps_range):
    """Temporal matching FLOP per frame in the reference's count, interior
    windows (every predictor window (2 ps_range + 1)^2 candidates): per
    window row 24 (wh + 7) + 7 wh, 2 radius ps_num windows per ref block."""
    xs, ys = ref_positions(w, h, step)
    wh = 2 * ps_range + 1
    return float(len(xs) * len(ys) * 2 * radius * ps_num * wh * (24 * (wh + 7) + 7 * wh))


def candidates(w, h, step, bm_range, radius=0, ps_num=0, ps_range=0):
    """Candidates per frame: the spatial windows (clamped) plus the temporal
    predictor windows (interior size)."""
    xs, ys = ref_positions(w, h, step)
    rw = np.minimum(xs + bm_range, w - 8) - np.maximum(xs - bm_range, 0) + 1
    rh = np.minimum(ys + bm_range, h - 8) - np.maximum(ys - bm_range, 0) + 1
    wh = 2 * ps_range + 1
    return float(np.sum(rw) * np.sum(rh)) + float(len(xs) * len(ys) * 2 * radius * ps_num * wh * wh)


def mfma_instructions(w, h, step, bm_range):
    """16x16x32 MFMA instructions of the precision form's spatial search per
    frame: per window 2 per 16x16 tile (h hi, lo) and 6 per 16-row band (the
    k < 7 correction of its first column tile)."""
    xs, ys = ref_positions(w, h, step)
    rw = np.minimum(xs + bm_range, w - 8) - np.maximum(xs - bm_range, 0) + 1
    rh = np.minimum(ys + bm_range, h - 8) - np.maximum(ys - bm_range, 0) + 1
    tr, tc = (rh + 15) // 16, (rw + 15) // 16
    return float(np.sum(2 * tr[:, None] * tc[None, :] + 6 * tr[:, None] + 0 * tc[None, :]))


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (its line
    'traffic_bytes_per_launch <kernel> <workload> <bytes>'), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                m = re.match(r"\s*traffic_bytes_per_launch\s+(\S+)\s+(\S+)\s+([0-9.eE+]+)", line)
                if m and m.group(1) == kernel and m.group(2) == workload:
                    return float(m.group(3))
    except OSError:
        pass
    return None


def copy_calibration(mb):
    """A plain device copy of `mb` MiB (read mb, write mb), three launches,
    so that a counter run sees the FETCH_SIZE / WRITE_SIZE of known traffic."""
    import torch
    n = mb * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def timed(inst, run, steps, warmup):
    """Warm up, time `steps` calls of run() (launch sequences replayed from
    HIP graphs where the entry point replays them by default), then run max(2, steps // 2)
    more calls with the library's per-kernel HIP events on (direct launches)
    for the kernel durations; returns (seconds, {kernel: (avg ms, launches)})."""
    import torch
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    inst.set_timing(True)
    for _ in range(max(2, steps // 2)):
        run()
    torch.cuda.synchronize()
    times = inst.kernel_times()
    inst.set_timing(False)
    per_kernel = {k: (ms / cnt, cnt) for k, (ms, cnt) in times.items() if cnt > 0}
    return t1 - t0, per_kernel


def roofline(per_kernel, w, h, kw, workload, kernel=None, precision="fp32"):
    """Roofline object of `kernel` (default: the one with the most time):
    algorithmic work per launch over its average launch duration."""
    n_ref = len(ref_positions(w, h, kw["block_step"])[0]) * len(ref_positions(w, h, kw["block_step"])[1])
    dominant = kernel or max(per_kernel, key=lambda k: per_kernel[k][0] * per_kernel[k][1])
    avg_ms = per_kernel[dominant][0]
    r = kw.get("radius", 0)
    tw = 2 * r + 1
    traffic = measured_traffic(dominant, workload)
    if dominant == "bm3d_match" and precision != "fp32":
        # precision form: the spatial search's dense 8x8 dot products (norms
        # - 2 dot, 128 FLOP per candidate) against the FP16/BF16 matrix peak
        ncand = candidates(w, h, kw["block_step"], kw["bm_range"])
        work = 128.0 * ncand
        issued = 16384.0 * mfma_instructions(w, h, kw["block_step"], kw["bm_range"])
        achieved = work / (avg_ms * 1e-3) / 1e12
        return dict(bound="mfma", achieved=achieved, peak=PEAK_MFMA16_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_MFMA16_TFLOPS, traffic=traffic, kernel=dominant,
                    avg_launch_ms=avg_ms, work_per_launch=work,
                    work_unit=f"FLOP (128 per candidate: the 8x8 dot products of the spatial search's "
                              f"{ncand:.4g} candidates; {precision.upper()} matrix peak)",
                    mfma_flop_issued_per_launch=issued,
                    mfma_issued_tflops=issued / (avg_ms * 1e-3) / 1e12)
    if dominant in ("bm3d_match", "bm3d_match_temporal"):
        spatial = matching_flops(w, h, kw["block_step"], kw["bm_range"])
        temporal = temporal_flops(w, h, kw["block_step"], r, kw["ps_num"], kw["ps_range"]) if r > 0 else 0.0
        if dominant == "bm3d_match_temporal":
            work, what = temporal, "temporal predictive matching, interior predictor windows"
        elif r > 0 and "bm3d_match_temporal" not in per_kernel:
            work, what = spatial + temporal, "spatial + temporal matching (interior predictor windows)"
        else:
            work, what = spatial, "spatial matching"
        achieved = work / (avg_ms * 1e-3) / 1e12
        return dict(bound="valu_fp32", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_TFLOPS, traffic=traffic, kernel=dominant,
                    avg_launch_ms=avg_ms, work_per_launch=work,
                    work_unit=f"FLOP ({what}, reference count; FP32 vector peak)")
    if dominant == "bm3d_estimate":
        work = n_ref * TRANSFORM_FLOP_PER_GROUP
        achieved = work / (avg_ms * 1e-3) / 1e12
        return dict(bound="valu_fp32", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_TFLOPS, traffic=traffic, kernel=dominant,
                    avg_launch_ms=avg_ms, work_per_launch=work,
                    work_unit="FLOP (3-D transforms + threshold; FP32 vector peak)")
    if dominant == "bm3d_tile_reduce":
        # per pixel: the covering tiles' (wdst, weight) pairs read once is
        # not algorithmic; the aggregation's own bytes are the slice planes
        # read and the output written: (2 tw + 1) planes at radius 0 fused
        work = (2 * (tw if r == 0 else 1) + 1) * w * h * 4.0
    else:
        work = (2 * tw + 1) * w * h * 4.0
    achieved = work / (avg_ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                frac=achieved / PEAK_HBM_GBS, traffic=traffic, kernel=dominant,
                avg_launch_ms=avg_ms, work_per_launch=work, work_unit="bytes")


def describe(config, clip_shape, kw, form):
    nf, h, w = clip_shape
    return (f"config {config}: {nf} x {w}x{h} GrayS float32, {form}, "
            + ", ".join(f"{k}={v}" for k, v in kw.items()))


def bench_plain(config, steps, warmup, frames=None, precision="fp32"):
    import torch
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    spec = synth.WORKLOADS[config]
    kw = dict(spec["args"])
    planes = synth.workload(config, frames=frames)
    clip = planes[0] if isinstance(planes, tuple) else planes
    nf, h, w = clip.shape
    x = torch.from_numpy(clip).to("cuda:0")
    inst = Instance(nf, h, w, precision=precision, **kw)
    out = torch.empty_like(x)
    secs, per_kernel = timed(inst, lambda: inst.process(x, out=out), steps, warmup)
    return clip, kw, secs, per_kernel


def bench_fwd_bwd(steps, warmup):
    import torch
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    kw = dict(synth.WORKLOADS[5]["args"])
    clip = synth.workload(5)
    nf, h, w = clip.shape
    x = torch.from_numpy(clip).to("cuda:0")
    g = torch.from_numpy(np.random.default_rng(5).standard_normal(clip.shape).astype(np.float32)).to("cuda:0")
    inst = Instance(nf, h, w, **kw)

    def run():
        _, saved = inst.forward_saved(x)
        inst.backward(g, saved)

    secs, per_kernel = timed(inst, run, steps, warmup)
    return clip, kw, secs, per_kernel


def bench_c4(steps, warmup, oracle_threads):
    """The precision-form workload (16 frames of 3840x2160, shift (1, 0),
    noise 5/255, sigma 5, step 4, range 16, radius 1, ps 2/7) in the plain
    FP32 form and the FP16 / BF16 MFMA forms: frames/s of each; on frame 0
    the share of groups identical to the plain form's and to the FP64
    oracle's, max-abs error and PSNR against the FP64 oracle's output, and
    the denoising PSNR against the clean frame beside the FP64 oracle's."""
    import torch
    from oracle import pyoracle
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    spec = synth.WORKLOADS[4]
    kw = dict(spec["args"])
    clip = synth.workload(4)
    nf, h, w = clip.shape
    clean0 = synth.shifted_clip(4, 1, w, h, 0.0, spec["shift"])[0].astype(np.float64)
    x = torch.from_numpy(clip).to("cuda:0")
    p = pyoracle.make_params(w, h, nthreads=oracle_threads, **kw)
    ref = pyoracle.clip(p, clip, first=0, count=1, f64=True)[0]
    r = kw.get("radius", 0)
    window = [clip[min(max(0 - r + z, 0), nf - 1)] for z in range(2 * r + 1)]
    _, g64 = pyoracle.estimate(p, window, f64=True)

    def psnr(a, b):
        mse = float(np.mean((a - b) ** 2))
        return 10.0 * np.log10(1.0 / mse) if mse > 0 else None

    psnr64 = psnr(ref, clean0)
    out = {}
    g_plain = None
    for prec in ("fp32", "fp16", "bf16"):
        inst = Instance(nf, h, w, precision=prec, **kw)
        o = torch.empty_like(x)
        secs, per_kernel = timed(inst, lambda: inst.process(x, out=o), steps, warmup)
        got = o[0].cpu().numpy().astype(np.float64)
        g, _ = inst.groups(x, 0)
        if g_plain is None:
            g_plain = g
        den = psnr(got, clean0)
        out[prec] = dict(value=nf * steps / secs, unit="frames/s", ms_per_step=secs * 1e3 / steps,
                         kernels_ms_avg={k: v[0] for k, v in per_kernel.items()},
                         roofline=roofline(per_kernel, w, h, kw, "", kernel="bm3d_match", precision=prec),
                         roofline_temporal=roofline(per_kernel, w, h, kw, "", kernel="bm3d_match_temporal"),
                         groups_identical_to_plain=float((g == g_plain).all(axis=(2, 3)).mean()),
                         groups_identical_to_fp64=float((g == g64).all(axis=(2, 3)).mean()),
                         max_abs_vs_fp64=float(np.abs(got - ref).max()),
                         psnr_vs_fp64_db=psnr(got, ref),
                         denoise_psnr_db=den,
                         denoise_psnr_change_vs_fp64_db=(den - psnr64) if den and psnr64 else None)
        del inst
    plain = out.pop("fp32")
    for prec in ("fp16", "bf16"):
        out[prec]["ratio_to_plain"] = out[prec]["value"] / plain["value"]
    return dict(plain, workload=describe(4, clip.shape, kw, "plain FP32 form, then the MFMA forms"),
                frame0_fp64_oracle_denoise_psnr_db=psnr64, oracle_threads=oracle_threads,
                mfma_fp16=out["fp16"], mfma_bf16=out["bf16"])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--frames", type=int, default=None, help="override the clip length")
    ap.add_argument("--forms", default="plain,mfma,fwd_bwd,temporal_r2,host_io,pipe,c4",
                    help="comma list of forms to time (plain is the headline value)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall-time budget of the CPU baseline sample (0: skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: measures one GPU only; frames would shard across GPUs (each GPU "
              "filtering a contiguous range of frames from its own copy of the range and a "
              "2r-frame source halo), but no multi-GPU run is implemented", file=sys.stderr)
        sys.exit(2)

    forms = [f for f in args.forms.split(",") if f]
    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb)

    clip, kw, secs, per_kernel = bench_plain(args.config, args.steps, args.warmup, args.frames)
    nf, h, w = clip.shape
    fps = nf * args.steps / secs
    workload = f"{w}x{h}_step{kw['block_step']}_range{kw['bm_range']}_r{kw.get('radius', 0)}"
    roof = roofline(per_kernel, w, h, kw, workload)
    aggregation = None
    if "bm3d_tile_reduce" in per_kernel:
        # radius 0: the tile reduction is the aggregation (fused division)
        aggregation = roofline(per_kernel, w, h, kw, workload, kernel="bm3d_tile_reduce")

    extra = {}
    if "fwd_bwd" in forms:
        c5, kw5, s5, pk5 = bench_fwd_bwd(args.steps, args.warmup)
        extra["fwd_bwd"] = dict(value=c5.shape[0] * args.steps / s5, unit="frames/s",
                                ms_per_step=s5 * 1e3 / args.steps,
                                workload=describe(5, c5.shape, kw5, "forward with saved state + backward"),
                                kernels_ms_avg={k: v[0] for k, v in pk5.items()},
                                roofline=roofline(pk5, c5.shape[2], c5.shape[1], kw5, ""))
    if "host_io" in forms:
        from vsfeel import BM3Dv2, Clip
        from vsfeel.bm3d import set_graphs
        host = {}
        for graphs in (False, True):
            node = BM3Dv2(Clip.gray(clip), **kw)
            set_graphs(node, requests=graphs)
            # set-up outside the timed region: every output slot's first
            # request (its pinned staging planes; with graphs, its graph)
            for n in range(5):
                node.get_frame(n)
            h0 = time.perf_counter()
            for n in range(nf):
                node.get_frame(n)
            h1 = time.perf_counter()
            host[graphs] = nf / (h1 - h0)
            del node
        extra["host_io"] = dict(value=host[False], unit="frames/s", graph_launches=host[True],
                                workload=describe(args.config, clip.shape, kw,
                                                  "host buffers: upload, filter and download per "
                                                  "frame request, one thread (PCIe-inclusive); value "
                                                  "with direct launches (the default), graph_launches "
                                                  "with per-request graphs"))
    if "pipe" in forms:
        # the vspipe counterpart with the reference harness's BM3Dv2 defaults
        # (benchmark/bench.py:167-180): 4 threads requesting 1080p frames
        from vsfeel import pipe_bench
        frames = 300
        pipe_fps = pipe_bench.run(frames=frames, synthetic=False, threads=4)
        pipe_graphs = pipe_bench.run(frames=frames, synthetic=False, threads=4, graphs=True)
        # radius 0 from 8 threads with 8 output slots: the host path against
        # the device-resident headline rate
        pipe_r0 = pipe_bench.run(frames=2 * frames, synthetic=False, threads=8, radius=0, sigma=3.0,
                                 num_streams=8)
        extra["pipe"] = dict(value=pipe_fps, unit="frames/s", graph_launches=pipe_graphs,
                             workload=f"{frames} frames of 1920x1080 GrayS float32 (16 synthetic frames cycled), "
                                      "host buffers from 4 threads (PCIe-inclusive), sigma=0.7, block_step=4, "
                                      "bm_range=16, radius=2, ps_range=7, num_streams=4; value with direct "
                                      "launches, graph_launches with per-request graphs",
                             radius0_8threads=dict(
                                 value=pipe_r0, unit="frames/s",
                                 workload=f"{2 * frames} frames of 1920x1080 GrayS float32, host buffers from 8 "
                                          "threads (PCIe-inclusive), sigma=3.0, block_step=4, bm_range=16, "
                                          "radius=0, num_streams=8, direct launches"))
    if "c4" in forms:
        extra["c4"] = bench_c4(max(1, args.steps // 2), 1, max(1, min(16, os.cpu_count() or 1)))
    tsteps = max(1, args.steps // 2)
    plain_r2 = None
    if "temporal_r2" in forms or "mfma" in forms:
        c3, kw3, s3, pk3 = bench_plain(3, tsteps, 1, args.frames)
        h3, w3 = c3.shape[1:]
        wl3 = f"{w3}x{h3}_step{kw3['block_step']}_range{kw3['bm_range']}_r{kw3['radius']}"
        plain_r2 = c3.shape[0] * tsteps / s3
        extra["temporal_r2"] = dict(value=plain_r2, unit="frames/s", ms_per_step=s3 * 1e3 / tsteps,
                                    workload=describe(3, c3.shape, kw3, "plain form, luma"),
                                    kernels_ms_avg={k: v[0] for k, v in pk3.items()},
                                    roofline=roofline(pk3, w3, h3, kw3, wl3, kernel="bm3d_match_temporal"),
                                    roofline_spatial=roofline(pk3, w3, h3, kw3, wl3, kernel="bm3d_match"))
    if "mfma" in forms:
        # the precision form on the temporal workload (the north star's
        # 1080p radius 2 target) and on the headline workload (radius 0)
        for prec in ("fp16", "bf16"):
            cm, kwm, sm, pkm = bench_plain(3, tsteps, 1, args.frames, precision=prec)
            hm, wm = cm.shape[1:]
            c0, kw0, s0, pk0 = bench_plain(args.config, args.steps, args.warmup, args.frames, precision=prec)
            h0, w0 = c0.shape[1:]
            fps_m, fps_0 = cm.shape[0] * tsteps / sm, c0.shape[0] * args.steps / s0
            extra["mfma_" + prec] = dict(
                value=fps_m, unit="frames/s", ms_per_step=sm * 1e3 / tsteps, ratio_to_plain=fps_m / plain_r2,
                workload=describe(3, cm.shape, kwm, f"precision form ({prec.upper()} MFMA distances), luma"),
                kernels_ms_avg={k: v[0] for k, v in pkm.items()},
                roofline=roofline(pkm, wm, hm, kwm, "", kernel="bm3d_match", precision=prec),
                radius0=dict(value=fps_0, unit="frames/s", ms_per_step=s0 * 1e3 / args.steps,
                             ratio_to_plain=fps_0 / fps,
                             workload=describe(args.config, c0.shape, kw0, f"precision form ({prec.upper()})"),
                             kernels_ms_avg={k: v[0] for k, v in pk0.items()},
                             roofline=roofline(pk0, w0, h0, kw0, "", kernel="bm3d_match", precision=prec)))

    cpu = None
    if args.cpu_seconds > 0:
        from oracle import pyoracle
        # the cores this process may run on; one GPU's box grants 16
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        cores = max(1, min(16, avail))
        p = pyoracle.make_params(w, h, nthreads=cores, **kw)
        done = 0
        c0 = time.perf_counter()
        while done < nf and (time.perf_counter() - c0) < args.cpu_seconds:
            pyoracle.clip(p, clip, first=done, count=1)
            done += 1
        c1 = time.perf_counter()
        cpu = dict(value=done / (c1 - c0), unit="frames/s", cores=cores, kind="port",
                   cpu_model=cpu_model(), cores_available=avail,
                   sample=f"{done} of the {nf} frames, oracle restatement of the reference "
                          f"shaders (not the reference binary: no Vulkan loader, glslc or lavapipe), "
                          f"{cores} threads on {cpu_model()}")

    line = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": describe(args.config, clip.shape, kw, "plain form"), "frames_per_step": nf},
        "kernels_ms_avg": {k: v[0] for k, v in per_kernel.items()},
        "roofline": roof,
        "aggregation": aggregation,
        "cpu_baseline": cpu,
        "forms": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
