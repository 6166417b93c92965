#!/usr/bin/env python3
"""BM3Dv2 throughput on one MI355X.

A step is one pass of the filter over the whole synthetic clip resident in
HBM: for the default workload, 32 frames of 1920x1080 GrayS float32 with
sigma 3, block_step 4, bm_range 16, radius 0 (plain FP32 form).  `value` is
frames per second over the timed steps (each step's launch sequence replayed
from the library's HIP graph).  Per-kernel durations come from HIP events
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
* ``c4``: 16 frames of 3840x2160 (radius 1, ps 2/7, sigma 5) in the FP32
  form, with frame 0 compared against the FP64 oracle (max-abs, PSNR, share
  of identical groups).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2]
                    [--forms plain,fwd_bwd,temporal_r2,host_io,pipe,c4] [--cpu-seconds S]
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
    HIP graphs where the entry point has them), then run max(2, steps // 2)
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


def roofline(per_kernel, w, h, kw, workload):
    n_ref = len(ref_positions(w, h, kw["block_step"])[0]) * len(ref_positions(w, h, kw["block_step"])[1])
    dominant = max(per_kernel, key=lambda k: per_kernel[k][0] * per_kernel[k][1])
    avg_ms = per_kernel[dominant][0]
    tw = 2 * kw.get("radius", 0) + 1
    traffic = measured_traffic(dominant, workload)
    if dominant == "bm3d_match":
        work = matching_flops(w, h, kw["block_step"], kw["bm_range"])
        achieved = work / (avg_ms * 1e-3) / 1e12
        return dict(bound="mfma", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_TFLOPS, traffic=traffic, kernel=dominant,
                    avg_launch_ms=avg_ms, work_per_launch=work,
                    work_unit="FLOP (spatial matching, reference count; FP32 vector peak)")
    if dominant == "bm3d_estimate":
        work = n_ref * TRANSFORM_FLOP_PER_GROUP
        achieved = work / (avg_ms * 1e-3) / 1e12
        return dict(bound="mfma", achieved=achieved, peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=achieved / PEAK_FP32_TFLOPS, traffic=traffic, kernel=dominant,
                    avg_launch_ms=avg_ms, work_per_launch=work,
                    work_unit="FLOP (3-D transforms + threshold; FP32 vector peak)")
    work = (2 * tw + 1) * w * h * 4.0
    achieved = work / (avg_ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=achieved, peak=PEAK_HBM_GBS, unit="GB/s",
                frac=achieved / PEAK_HBM_GBS, traffic=traffic, kernel=dominant,
                avg_launch_ms=avg_ms, work_per_launch=work, work_unit="bytes")


def describe(config, clip_shape, kw, form):
    nf, h, w = clip_shape
    return (f"config {config}: {nf} x {w}x{h} GrayS float32, {form}, "
            + ", ".join(f"{k}={v}" for k, v in kw.items()))


def bench_plain(config, steps, warmup, frames=None):
    import torch
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    spec = synth.WORKLOADS[config]
    kw = dict(spec["args"])
    planes = synth.workload(config, frames=frames)
    clip = planes[0] if isinstance(planes, tuple) else planes
    nf, h, w = clip.shape
    x = torch.from_numpy(clip).to("cuda:0")
    inst = Instance(nf, h, w, **kw)
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
    noise 5/255, sigma 5, step 4, range 16, radius 1, ps 2/7) in the FP32
    form: frames/s, and on frame 0 the error against the FP64 oracle (max-abs,
    PSNR) and the share of its groups identical to the FP64 oracle's."""
    import torch
    from oracle import pyoracle
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    kw = dict(synth.WORKLOADS[4]["args"])
    clip = synth.workload(4)
    nf, h, w = clip.shape
    x = torch.from_numpy(clip).to("cuda:0")
    inst = Instance(nf, h, w, **kw)
    out = torch.empty_like(x)
    secs, per_kernel = timed(inst, lambda: inst.process(x, out=out), steps, warmup)
    got = out[0].cpu().numpy().astype(np.float64)
    g32, _ = inst.groups(x, 0)
    p = pyoracle.make_params(w, h, nthreads=oracle_threads, **kw)
    ref = pyoracle.clip(p, clip, first=0, count=1, f64=True)[0]
    r = kw.get("radius", 0)
    window = [clip[min(max(0 - r + z, 0), nf - 1)] for z in range(2 * r + 1)]
    _, g64 = pyoracle.estimate(p, window, f64=True)
    same = (g32 == g64).all(axis=(2, 3))
    mse = float(np.mean((got - ref) ** 2))
    return dict(value=nf * steps / secs, unit="frames/s", ms_per_step=secs * 1e3 / steps,
                workload=describe(4, clip.shape, kw, "plain FP32 form"),
                kernels_ms_avg={k: v[0] for k, v in per_kernel.items()},
                vs_fp64_oracle=dict(frame=0, max_abs=float(np.abs(got - ref).max()),
                                    psnr_db=10.0 * np.log10(1.0 / mse) if mse > 0 else None,
                                    groups_identical=float(same.mean()),
                                    oracle_threads=oracle_threads))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--frames", type=int, default=None, help="override the clip length")
    ap.add_argument("--forms", default="plain,fwd_bwd,temporal_r2,host_io,pipe,c4",
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

    extra = {}
    if "fwd_bwd" in forms:
        c5, kw5, s5, pk5 = bench_fwd_bwd(args.steps, args.warmup)
        extra["fwd_bwd"] = dict(value=c5.shape[0] * args.steps / s5, unit="frames/s",
                                ms_per_step=s5 * 1e3 / args.steps,
                                workload=describe(5, c5.shape, kw5, "forward with saved state + backward"),
                                kernels_ms_avg={k: v[0] for k, v in pk5.items()})
    if "host_io" in forms:
        from vsfeel import BM3Dv2, Clip
        node = BM3Dv2(Clip.gray(clip), **kw)
        node.get_frame(0)
        h0 = time.perf_counter()
        for n in range(nf):
            node.get_frame(n)
        h1 = time.perf_counter()
        extra["host_io"] = dict(value=nf / (h1 - h0), unit="frames/s",
                                workload=describe(args.config, clip.shape, kw,
                                                  "host buffers: upload, filter and download per "
                                                  "frame request, one thread (PCIe-inclusive)"))
        del node
    if "pipe" in forms:
        # the vspipe counterpart with the reference harness's BM3Dv2 defaults
        # (benchmark/bench.py:167-180): 4 threads requesting 1080p frames
        from vsfeel import pipe_bench
        frames = 300
        pipe_fps = pipe_bench.run(frames=frames, synthetic=False, threads=4)
        extra["pipe"] = dict(value=pipe_fps, unit="frames/s",
                             workload=f"{frames} frames of 1920x1080 GrayS float32 (16 synthetic frames cycled), "
                                      "host buffers from 4 threads (PCIe-inclusive), sigma=0.7, block_step=4, "
                                      "bm_range=16, radius=2, ps_range=7, num_streams=4")
    if "c4" in forms:
        extra["c4"] = bench_c4(max(1, args.steps // 2), 1, max(1, min(16, os.cpu_count() or 1)))
    if "temporal_r2" in forms:
        c3, kw3, s3, pk3 = bench_plain(3, max(1, args.steps // 2), 1, args.frames)
        extra["temporal_r2"] = dict(value=c3.shape[0] * max(1, args.steps // 2) / s3, unit="frames/s",
                                    workload=describe(3, c3.shape, kw3, "plain form, luma"),
                                    kernels_ms_avg={k: v[0] for k, v in pk3.items()})

    cpu = None
    if args.cpu_seconds > 0:
        from oracle import pyoracle
        cores = max(1, min(16, os.cpu_count() or 1))
        p = pyoracle.make_params(w, h, nthreads=cores, **kw)
        done = 0
        c0 = time.perf_counter()
        while done < nf and (time.perf_counter() - c0) < args.cpu_seconds:
            pyoracle.clip(p, clip, first=done, count=1)
            done += 1
        c1 = time.perf_counter()
        cpu = dict(value=done / (c1 - c0), unit="frames/s", cores=cores, kind="port",
                   sample=f"{done} of the {nf} frames, oracle restatement of the reference "
                          f"shaders (not the reference binary), {cores} threads")

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
        "cpu_baseline": cpu,
        "forms": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
