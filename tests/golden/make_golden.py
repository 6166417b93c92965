"""Writes the golden vectors of tests/golden/*.npz from the oracle.

The reference cannot run here (GLSL compute shaders behind Vulkan and
VapourSynth) and its repository holds no numeric vectors, so these fixtures
freeze the oracle's outputs on small seeded inputs: inputs, parameters, the
group table of every frame (x, y, z of its 8 members) and the denoised
frames.  tests/test_oracle.py checks that the oracle still reproduces them
and the GPU parity tests check the HIP path against them, so a change of
either side shows up against a fixed record.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "vapoursynth-feel_amd"))

from oracle import pyoracle  # noqa: E402
from vsfeel import synth  # noqa: E402

CASES = {
    # spatial search, the reference workload's generator and parameters
    "spatial_96x80": (lambda: synth.workload(1, frames=2, width=96, height=80),
                      dict(sigma=3.0, block_step=8, bm_range=9, radius=0, ps_num=2, ps_range=4)),
    # temporal predictive search with the extractor (sums exact)
    "temporal_72x56_r2": (lambda: synth.shifted_clip(21, 5, 72, 56, 3 / 255, (1.5, 0.75)),
                          dict(sigma=3.0, block_step=4, bm_range=6, radius=2, ps_num=2, ps_range=3,
                                extractor_exp=2)),
    # exact SSD metric (radius 3)
    "exact_48x40_r3": (lambda: synth.shifted_clip(22, 4, 48, 40, 3 / 255, (1.0, 0.5)),
                       dict(sigma=3.0, block_step=5, bm_range=4, radius=3, ps_num=2, ps_range=2)),
}


def golden(name):
    make, kw = CASES[name]
    clip = np.ascontiguousarray(make(), np.float32)
    nf, h, w = clip.shape
    p = pyoracle.make_params(w, h, nthreads=4, **kw)
    out = pyoracle.clip(p, clip)
    r = kw.get("radius", 0)
    groups = []
    for m in range(nf):
        window = [clip[min(max(m - r + z, 0), nf - 1)] for z in range(2 * r + 1)]
        _, g = pyoracle.estimate(p, window)
        groups.append(g)
    return clip, kw, np.stack(groups).astype(np.int16), out.astype(np.float32)


def main():
    for name in CASES:
        clip, kw, groups, out = golden(name)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), clip=clip, groups=groups, out=out,
                            args=np.frombuffer(json.dumps(kw).encode(), np.uint8))
        print(name, clip.shape, groups.shape)


if __name__ == "__main__":
    main()
