"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Block-match groups must be identical entry for entry (order included: the
group axis of the 3-D transform runs over them in list order), and the
denoised frames must agree within 1e-5 max-abs on [0, 1] data (the
reference accumulates with unordered float atomics, so the last bits of the
sums are not defined).  With a non-zero extractor every accumulated value is
a multiple of a power of two and the sums are exact, so the output must be
bit-identical.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = max(1, min(16, os.cpu_count() or 1))
TOL = 1e-5

# name -> (clip factory, BM3Dv2 arguments)
CASES = {
    # the CPU-runnable reference workload at its own size (640x360, step 8, range 9)
    "c1": (lambda s: s.workload(1, frames=2),
           dict(sigma=3.0, block_step=8, bm_range=9, radius=0, ps_num=2, ps_range=4)),
    # the differentiable-form workload shape (512x512, step 4, range 9)
    "c5": (lambda s: s.workload(5, frames=2),
           dict(sigma=3.0, block_step=4, bm_range=9, radius=0)),
    # the 1080p parameters (step 4, range 16) on a smaller frame
    "c2_small": (lambda s: s.workload(2, frames=2, width=480, height=272),
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=0)),
    # the benchmark sizes: 1080p plain (frame 0, and frame 7 whose texture
    # runs along the rows), and 1080p radius 2 over 5 frames
    "c2_full": (lambda s: s.workload(2, frames=8)[[0, 7]],
                dict(sigma=3.0, block_step=4, bm_range=16, radius=0)),
    "c3_full": (lambda s: s.workload(3, frames=5)[0],
                dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7, extractor_exp=2)),
    # the reference harness's case: blank 1080p frames (every distance ties)
    # with its BM3Dv2 defaults
    "blank_full": (lambda s: np.zeros((5, 1080, 1920), np.float32),
                   dict(sigma=0.7, block_step=4, bm_range=16, radius=2, ps_range=7)),
    # the precision-form workload's shape: 3840x2160, radius 1 (3 frames)
    "c4_full": (lambda s: s.workload(4, frames=3),
                dict(sigma=5.0, block_step=4, bm_range=16, radius=1, ps_num=2, ps_range=7)),
    # temporal: radius 2, predictive search 2/7, extractor 2^2, moving content
    "c3_small": (lambda s: s.workload(3, frames=6, width=256, height=144)[0],
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7,
                      extractor_exp=2)),
    "r1": (lambda s: s.workload(4, frames=4, width=200, height=120),
           dict(sigma=5.0, block_step=4, bm_range=16, radius=1, ps_num=2, ps_range=7)),
    # radius >= 3 switches the shader to the exact SSD scan
    "r3_exact": (lambda s: s.shifted_clip(7, 7, 96, 72, 3 / 255, (1.0, 0.5)),
                 dict(sigma=3.0, block_step=5, bm_range=6, radius=3, ps_num=3, ps_range=3)),
    "r4_exact": (lambda s: s.shifted_clip(8, 5, 64, 48, 3 / 255, (0.5, 1.0)),
                 dict(sigma=2.0, block_step=8, bm_range=4, radius=4, ps_num=1, ps_range=2)),
    # a search window wider than one wavefront row (chunked columns)
    "wide_range": (lambda s: s.gray_clip(9, 1, 160, 120, 3 / 255),
                   dict(sigma=3.0, block_step=6, bm_range=40, radius=0)),
    # ragged sizes: width/height not multiples of the step, odd widths
    "ragged": (lambda s: s.gray_clip(10, 2, 77, 45, 5 / 255),
               dict(sigma=4.0, block_step=3, bm_range=5, radius=1, ps_num=8, ps_range=2)),
    # the smallest frame: a single reference block, fewer than 8 candidates
    "tiny": (lambda s: s.gray_clip(11, 3, 9, 8, 5 / 255),
             dict(sigma=3.0, block_step=1, bm_range=1, radius=1, ps_num=2, ps_range=1)),
    # flat content: every distance ties, the seq order decides
    "flat": (lambda s: np.full((2, 40, 48), 0.5, np.float32),
             dict(sigma=3.0, block_step=4, bm_range=7, radius=0)),
    # ... also in the temporal search (overlapping predictor windows, every
    # frame top-8 decided by seq, more than 16 tied candidates)
    "flat_temporal": (lambda s: np.full((5, 40, 48), 0.25, np.float32),
                      dict(sigma=3.0, block_step=4, bm_range=5, radius=2, ps_num=3, ps_range=3)),
    # wide predictor windows (25 x 25, three per frame): one ref block per
    # workgroup of the temporal search
    "wide_predictive": (lambda s: s.shifted_clip(12, 5, 120, 96, 3 / 255, (2.0, -1.0)),
                        dict(sigma=3.0, block_step=6, bm_range=8, radius=2, ps_num=3, ps_range=12)),
    # one predictor per step (one staging half idle), 4 group entries
    "one_predictor": (lambda s: s.shifted_clip(14, 4, 80, 64, 3 / 255, (1.5, -0.5)),
                      dict(sigma=3.0, block_step=4, bm_range=6, radius=2, ps_num=1, ps_range=5)),
    # four predictors per step (two staging passes), 8 group entries
    "four_predictors": (lambda s: s.shifted_clip(15, 4, 88, 72, 3 / 255, (-1.0, 1.0)),
                        dict(sigma=3.0, block_step=5, bm_range=7, radius=1, ps_num=4, ps_range=3)),
    # texture constant along rows: a block's best matches crowd into its own
    # row and the rows a period away, past the first threshold's 16 slots
    # (frame 7 of the 1080p generator, at a smaller size, and pure stripes)
    "row_texture": (lambda s: s.workload(2, frames=8, width=320, height=192)[6:],
                    dict(sigma=3.0, block_step=4, bm_range=16, radius=0)),
    "stripes": (lambda s: np.clip(0.5 + 0.1 * np.sin(2 * np.pi * np.arange(96) / 9.0)[None, :, None]
                                  + np.random.default_rng(16).normal(0.0, 3 / 255, (2, 96, 160)),
                                  0.0, 1.0).astype(np.float32),
                dict(sigma=3.0, block_step=4, bm_range=12, radius=1, ps_num=2, ps_range=4)),
    # one predictor of 27 x 27 (34 staged columns, more than a 32-lane half
    # holds): the per-wavefront search takes over
    "one_wide_predictor_r1": (lambda s: s.shifted_clip(17, 3, 112, 88, 3 / 255, (1.5, -1.0)),
                              dict(sigma=3.0, block_step=5, bm_range=6, radius=1, ps_num=1, ps_range=13)),
    "one_wide_predictor_r2": (lambda s: s.shifted_clip(18, 5, 112, 88, 3 / 255, (-1.0, 1.5)),
                              dict(sigma=3.0, block_step=5, bm_range=6, radius=2, ps_num=1, ps_range=13)),
    # predictor windows too wide for the row-parallel temporal search
    # (61 columns + 7): the per-wavefront search takes over
    "widest_predictive": (lambda s: s.shifted_clip(13, 3, 96, 80, 3 / 255, (1.0, 1.0)),
                          dict(sigma=3.0, block_step=8, bm_range=6, radius=1, ps_num=2, ps_range=30)),
}


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _run_case(name, fp_mode=1):
    torch = _torch()
    from oracle import pyoracle
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    make, kw = CASES[name]
    clip = np.ascontiguousarray(make(synth), np.float32)
    nf, h, w = clip.shape
    print(f"case {name} fp_mode={fp_mode} clip={clip.shape} args={kw}", flush=True)
    inst = Instance(nf, h, w, fp_mode=fp_mode, **kw)
    x = torch.from_numpy(clip).to("cuda:0")
    out = inst.process(x)
    torch.cuda.synchronize()
    got = out.cpu().numpy()

    p = pyoracle.make_params(w, h, fp_mode=fp_mode, nthreads=NT, **kw)
    ref = pyoracle.clip(p, clip)
    r = kw.get("radius", 0)
    for m in range(nf):
        g, _ = inst.groups(x, m)
        window = [clip[min(max(m - r + z, 0), nf - 1)] for z in range(2 * r + 1)]
        _, og = pyoracle.estimate(p, window)
        if not np.array_equal(g, og):
            bad = np.argwhere((g != og).any(axis=(2, 3)))
            by, bx = bad[0]
            raise AssertionError(
                f"{name}: frame {m}: {len(bad)} groups differ; first at ref block ({bx},{by}):\n"
                f"gpu    {g[by, bx].tolist()}\noracle {og[by, bx].tolist()}")
    return got, ref, kw


@pytest.mark.parametrize("name", list(CASES))
def test_groups_and_output_match_oracle(name):
    got, ref, kw = _run_case(name)
    assert np.isfinite(got).all()
    err = float(np.abs(got - ref).max())
    if kw.get("extractor_exp", 0):
        assert np.array_equal(got, ref), f"{name}: extractor sums must be exact, max diff {err}"
    assert err <= TOL, f"{name}: max |gpu - oracle| = {err}"


@pytest.mark.parametrize("name", ["c3_small", "four_predictors", "wide_predictive", "r4_exact"])
def test_binning_overflow_matches_oracle(name, monkeypatch):
    """Where a group takes at least 8 temporal insertions (2 radius ps_num)
    the other frames' blocks are binned by 32x32 cell; a block whose cell
    list is full adds its pixels there atomically.  A list capacity of 3
    sends most blocks down that path; the frames must still match the
    oracle (bit-exact with the extractor)."""
    monkeypatch.setenv("VSFEEL_BM3D_CELL_CAP", "3")
    got, ref, kw = _run_case(name)
    err = float(np.abs(got - ref).max())
    if kw.get("extractor_exp", 0):
        assert np.array_equal(got, ref), f"{name}: extractor sums must be exact, max diff {err}"
    assert err <= TOL, f"{name}: max |gpu - oracle| = {err}"


@pytest.mark.parametrize("fp_mode", [0, 2])
@pytest.mark.parametrize("name", ["c1", "c3_small", "r3_exact"])
def test_fp_modes_match_oracle(name, fp_mode):
    got, ref, _ = _run_case(name, fp_mode)
    err = float(np.abs(got - ref).max())
    assert err <= TOL, f"{name} fp_mode {fp_mode}: max |gpu - oracle| = {err}"


@pytest.mark.parametrize("name", ["spatial_96x80", "temporal_72x56_r2", "exact_48x40_r3"])
def test_matches_golden_vectors(name):
    """The committed fixtures (tests/golden, written by the oracle)."""
    torch = _torch()
    from test_oracle import load_golden
    from vsfeel.torch_bm3d import Instance

    clip, kw, groups, want = load_golden(name)
    nf, h, w = clip.shape
    print(f"golden {name} clip={clip.shape} args={kw}", flush=True)
    inst = Instance(nf, h, w, **kw)
    x = torch.from_numpy(np.ascontiguousarray(clip)).to("cuda:0")
    got = inst.process(x).cpu().numpy()
    for m in range(nf):
        g, _ = inst.groups(x, m)
        assert np.array_equal(g, groups[m]), f"{name}: frame {m} groups differ from the fixture"
    if kw.get("extractor_exp", 0):
        assert np.array_equal(got, want), f"{name}: extractor sums must be exact"
    assert float(np.abs(got - want).max()) <= TOL


def test_final_pass_matches_oracle():
    """The Wiener pass (ref clip) against the oracle, on the oracle's own basic estimate."""
    torch = _torch()
    from oracle import pyoracle
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    clip = synth.shifted_clip(12, 5, 160, 96, 5 / 255, (1.0, 0.5))
    nf, h, w = clip.shape
    kw = dict(sigma=5.0, block_step=4, bm_range=8, radius=2, ps_num=2, ps_range=4)
    basic = pyoracle.clip(pyoracle.make_params(w, h, nthreads=NT, **kw), clip)
    pf = pyoracle.make_params(w, h, final=True, nthreads=NT, **kw)
    ref = pyoracle.clip(pf, clip, ref=basic)
    inst = Instance(nf, h, w, final=True, **kw)
    x = torch.from_numpy(clip).to("cuda:0")
    b = torch.from_numpy(np.ascontiguousarray(basic)).to("cuda:0")
    got = inst.process(x, b).cpu().numpy()
    for m in range(nf):
        g, _ = inst.groups(x, m, b)
        window = [basic[min(max(m - 2 + z, 0), nf - 1)] for z in range(5)]
        inputs = [clip[min(max(m - 2 + z, 0), nf - 1)] for z in range(5)]
        _, og = pyoracle.estimate(pf, window, inputs)
        assert np.array_equal(g, og), f"final pass: frame {m} groups differ"
    err = float(np.abs(got - ref).max())
    assert err <= TOL, f"final pass: max |gpu - oracle| = {err}"
    assert float(np.abs(got - basic).max()) > 1e-4, "the final pass must refine the basic estimate"


def test_estimate_stack_matches_oracle():
    """One estimation dispatch: the accumulated (wdst, weight) planes."""
    torch = _torch()
    from oracle import pyoracle
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    clip = synth.workload(3, frames=5, width=192, height=112)[0]
    kw = dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7)
    nf, h, w = clip.shape
    inst = Instance(nf, h, w, **kw)
    x = torch.from_numpy(clip).to("cuda:0")
    res = inst.estimate(x, 2).cpu().numpy()
    p = pyoracle.make_params(w, h, nthreads=NT, **kw)
    ores, _ = pyoracle.estimate(p, [clip[z] for z in range(5)])
    assert res.shape == ores.shape
    scale = np.abs(ores).max()
    err = float(np.abs(res - ores).max())
    assert err <= 1e-5 * max(1.0, scale), f"estimate stack differs by {err}"


@pytest.mark.parametrize("radius", [1, 2, 3])
def test_sigma_zero_reproduces_the_clip_at_radius(radius):
    """sigma = 0 keeps every coefficient, so every slice aggregated for frame
    n holds frame n's own pixels: the HIP path reproduces a moving clip at
    radius 1-3 (a slice landing on the wrong frame would not)."""
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    clip = synth.shifted_clip(64 + radius, 2 * radius + 2, 40, 32, 5 / 255, (1.0, -0.5))
    nf, h, w = clip.shape
    print(f"sigma 0 radius {radius} clip={clip.shape}", flush=True)
    inst = Instance(nf, h, w, sigma=0.0, block_step=4, bm_range=5, radius=radius, ps_num=2, ps_range=3)
    out = inst.process(torch.from_numpy(clip).to("cuda:0")).cpu().numpy()
    np.testing.assert_allclose(out, clip, atol=2e-6)
