"""Frame-request throughput runner, the counterpart of the reference's
vspipe benchmark (benchmark/bench.py:32-54 builds the script, :108-121 the
BM3Dv2 call, :167-180 its defaults, :226-235 times vspipe).

It requests frames 0..N-1 of a BM3Dv2 node through the host-buffer path
that the VapourSynth plugin uses (upload, filter, download per request),
from several threads like vspipe, and prints vspipe's summary line.  The
clip is either a blank 1920x1080 GrayS clip (the reference's --synthetic
BlankClip) or a pool of synthetic frames cycled over the clip's length, so
that a long clip needs no more host memory than the pool.

    python -m vsfeel.pipe_bench --synthetic --frames 1000
"""

from __future__ import annotations

import argparse
import threading
import time

import numpy as np

from . import bm3d, synth


class _Frames:
    """A (num_frames, h, w) plane whose frame n is pool[n % len(pool)]."""

    def __init__(self, pool: np.ndarray, num_frames: int):
        self._pool = pool
        self.shape = (num_frames,) + pool.shape[1:]

    def __getitem__(self, n):
        return self._pool[n % self._pool.shape[0]]


def make_clip(frames: int, synthetic: bool, width: int = 1920, height: int = 1080, pool: int = 16) -> bm3d.Clip:
    if synthetic:
        frames_pool = np.zeros((1, height, width), np.float32)  # BlankClip: black
    else:
        frames_pool = synth.workload(2, frames=min(pool, frames), width=width, height=height)
    clip = bm3d.Clip.gray(frames_pool[:1])
    clip._planes = [_Frames(frames_pool, frames)]
    return clip


def run(frames: int = 1000, synthetic: bool = True, threads: int = 4, width: int = 1920, height: int = 1080,
        sigma: float = 0.7, radius: int = 2, bm_range: int = 16, ps_range: int = 7, block_step: int = 4,
        num_streams: int = 4, graphs: bool = False) -> float:
    """Frames per second of requesting every frame once, in order, from
    `threads` threads; graphs=True launches each request's kernels from its
    output slot's HIP graph instead of directly."""
    clip = make_clip(frames, synthetic, width, height)
    node = bm3d.BM3Dv2(clip, sigma=sigma, radius=radius, bm_range=bm_range, ps_range=ps_range,
                       block_step=block_step, num_streams=num_streams)
    if graphs:
        bm3d.set_graphs(node, requests=True)
    # device set-up outside the timed region, as vspipe's first frame: one
    # request per output slot (its pinned staging planes and, with graphs,
    # its first graph)
    for n in range(min(num_streams, frames)):
        node.get_frame(n)
    nxt = [0]
    lock = threading.Lock()
    errors = []

    def worker():
        while True:
            with lock:
                n = nxt[0]
                nxt[0] += 1
            if n >= frames:
                return
            try:
                node.get_frame(n)
            except Exception as e:  # surfaced after the join
                errors.append(e)
                return

    pool = [threading.Thread(target=worker) for _ in range(max(1, threads))]
    t0 = time.perf_counter()
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    dt = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return frames / dt


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description="BM3Dv2 frame-request throughput (vspipe counterpart)")
    ap.add_argument("--frames", type=int, default=1000)
    ap.add_argument("--synthetic", action="store_true", help="blank 1920x1080 GrayS clip (BlankClip)")
    ap.add_argument("--threads", type=int, default=4, help="concurrent frame requests")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bm3d-sigma", type=float, default=0.7)
    ap.add_argument("--bm3d-radius", type=int, default=2)
    ap.add_argument("--bm3d-bm-range", type=int, default=16)
    ap.add_argument("--bm3d-ps-range", type=int, default=7)
    ap.add_argument("--bm3d-block-step", type=int, default=4)
    ap.add_argument("--num-streams", type=int, default=4)
    ap.add_argument("--graphs", action="store_true", help="launch each request's kernels from a graph")
    ns = ap.parse_args(argv)
    fps = run(ns.frames, ns.synthetic, ns.threads, ns.width, ns.height, ns.bm3d_sigma, ns.bm3d_radius,
              ns.bm3d_bm_range, ns.bm3d_ps_range, ns.bm3d_block_step, ns.num_streams, ns.graphs)
    print(f"Output {ns.frames} frames in {ns.frames / fps:.2f} seconds ({fps:.2f} fps)")


if __name__ == "__main__":
    main()
