"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors for BM3Dv2 (its tests only compare
runs with each other and with other plugins), so every part of the C
restatement is checked here against an independent formulation: closed
forms of the DCT butterflies, a float64 3-D DCT with scipy, brute-force
distances and a global stable sort for the block matching, the agg_z table
worked out by hand, and invariants (sigma = 0 reproduces the input, a
constant frame stays constant, the extractor quantises to its grid).
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest
from scipy.fft import dct, dctn, idct, idctn

from oracle import pyoracle as po
from vsfeel import synth

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- transforms
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dct8_is_four_times_orthonormal_dct2(mode):
    for _ in range(50):
        v = RNG.uniform(-1, 1, 8)
        fwd = po.dct8(v, fp_mode=mode)
        np.testing.assert_allclose(fwd, 4 * dct(v, norm="ortho"), rtol=0, atol=3e-6)
        inv = po.dct8(v, inverse=True, fp_mode=mode)
        np.testing.assert_allclose(inv, 4 * idct(v, norm="ortho"), rtol=0, atol=3e-6)
        f64 = po.dct8(v, fp_mode=mode, f64=True)
        np.testing.assert_allclose(f64, 4 * dct(v, norm="ortho"), rtol=0, atol=1e-7)


def test_dct8_fused_and_unfused_differ_only_in_rounding():
    v = RNG.uniform(-1, 1, (200, 8)).astype(np.float32)
    a = np.stack([po.dct8(x, fp_mode=0) for x in v])
    b = np.stack([po.dct8(x, fp_mode=1) for x in v])
    assert np.abs(a - b).max() < 1e-5
    assert (a != b).any(), "fp_mode 0 must round the GLSL fma() calls separately"


def test_group_sigma_zero_keeps_everything_and_reproduces_the_input():
    cube = RNG.uniform(0, 1, (8, 8, 8)).astype(np.float32)
    out, w, kept = po.group(cube, 0.0)
    assert kept == 512
    assert w == np.float32(1.0 / 512)
    np.testing.assert_allclose(out, cube, atol=2e-6)


def _np_hard(cube, sigma):
    """Float64 collaborative hard threshold: the coefficients of the shader are
    4^3 = 64 times the orthonormal 3-D DCT-II; kept ones scaled by 1/4096 and
    brought back by 64 x the orthonormal inverse."""
    c = 64.0 * dctn(cube.astype(np.float64), norm="ortho")
    keep = np.abs(c) >= sigma
    keep[0, 0, 0] = True
    out = idctn(np.where(keep, c, 0.0) / 64.0, norm="ortho")
    return out, 1.0 / keep.sum(), keep, c


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_group_hard_threshold_matches_scipy(mode):
    sigma = float(np.float32(3.0) * np.float32(po.SIGMA_FACTOR_BASIC))
    checked = 0
    for _ in range(40):
        cube = (0.5 + 0.02 * RNG.standard_normal((8, 8, 8))).astype(np.float32)
        ref, w, keep, c = _np_hard(cube, sigma)
        margin = np.abs(np.abs(c) - sigma).min()
        if margin < 1e-3:
            continue  # a coefficient within rounding distance of the threshold
        out, wo, kept = po.group(cube, sigma, fp_mode=mode)
        assert kept == keep.sum()
        assert abs(wo - w) <= 1e-7 * w
        np.testing.assert_allclose(out, ref, atol=2e-6)
        checked += 1
    assert checked >= 20


def test_group_wiener_matches_scipy():
    sigma = float(np.float32(3.0) * np.float32(po.SIGMA_FACTOR_FINAL))
    for _ in range(20):
        cube = (0.5 + 0.03 * RNG.standard_normal((8, 8, 8))).astype(np.float32)
        refc = (cube + 0.01 * RNG.standard_normal((8, 8, 8))).astype(np.float32)
        out, w, _ = po.group(cube, sigma, ref_cube=refc)
        R = 64.0 * dctn(refc.astype(np.float64), norm="ortho")
        coeff = R * R / (R * R + sigma * sigma)
        coeff[0, 0, 0] = 1.0
        expect = idctn(coeff * dctn(cube.astype(np.float64), norm="ortho"), norm="ortho")
        np.testing.assert_allclose(out, expect, atol=3e-6)
        assert abs(w - 1.0 / (coeff * coeff).sum()) <= 1e-6 / (coeff * coeff).sum()


# ---------------------------------------------------------- block matching
def _col_ssd64(plane, cur, a, X, cy):
    return float(((cur[:, a].astype(np.float64) - plane[cy:cy + 8, X].astype(np.float64)) ** 2).sum())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_exact_metric_is_the_true_ssd(seed):
    plane = synth.gray_clip(20 + seed, 1, 48, 40, 5 / 255)[0]
    x, y, R = 17, 13, 6
    d, (left, top) = po.window_dist(plane, x, y, R, x, y, exact=True)
    cur = plane[y:y + 8, x:x + 8].astype(np.float64)
    for cy in range(d.shape[0]):
        for cx in range(d.shape[1]):
            blk = plane[top + cy:top + cy + 8, left + cx:left + cx + 8].astype(np.float64)
            ssd = ((cur - blk) ** 2).sum()
            assert abs(d[cy, cx] - ssd) <= 1e-5 * max(ssd, 1e-6) + 1e-9
    assert d[y - top, x - left] == 0.0


@pytest.mark.parametrize("seed", [0, 1])
def test_sliding_metric_closed_form(seed):
    """The radius < 3 scan shifts column sums that were taken against other
    reference columns: candidate k of a row adds S(min(c + k, 7), left+k+c)
    over c, S(a, X) being the 8-row SSD of reference column a and plane column
    X.  For k >= 7 every term uses reference column 7."""
    plane = synth.gray_clip(30 + seed, 1, 56, 44, 5 / 255)[0]
    x, y, R = 20, 15, 9
    d, (left, top) = po.window_dist(plane, x, y, R, x, y, exact=False)
    dx, _ = po.window_dist(plane, x, y, R, x, y, exact=True)
    cur = plane[y:y + 8, x:x + 8]
    for r in range(d.shape[0]):
        cy = top + r
        assert d[r, 0] == dx[r, 0], "the first candidate of a row is an exact SSD"
        for k in range(d.shape[1]):
            closed = sum(_col_ssd64(plane, cur, min(c + k, 7), left + k + c, cy) for c in range(8))
            assert abs(d[r, k] - closed) <= 2e-5 * max(closed, 1e-6) + 1e-9
    own = d[y - top, x - left]
    assert own > 0.0, "the reference block's own distance is not zero under this metric"
    assert np.abs(d - dx).max() > 1e-3


def _brute_groups(plane, p):
    """Top-8 of every window by (distance, seq) with a global stable sort,
    placeholders (0, 0) when a window has fewer than 8 candidates, then the
    self-block rule."""
    n, nx, ny = po.ref_grid(p)
    out = np.zeros((ny, nx, 8, 3), np.int32)
    W, H, s, R = p.width, p.height, p.block_step, p.bm_range
    for by in range(ny):
        for bx in range(nx):
            x, y = min(bx * s, W - 8), min(by * s, H - 8)
            d, (left, top) = po.window_dist(plane, x, y, R, x, y, exact=p.radius >= 3)
            flat = d.ravel()
            order = np.argsort(flat, kind="stable")[:8]
            rw = d.shape[1]
            sel = [(left + int(o) % rw, top + int(o) // rw) for o in order]
            sel += [(0, 0)] * (8 - len(sel))
            if (x, y) not in sel:
                sel = [(x, y)] + sel[:7]
            out[by, bx, :, 0] = [q[0] for q in sel]
            out[by, bx, :, 1] = [q[1] for q in sel]
    return out


@pytest.mark.parametrize("shape,step,rng_", [((40, 32), 4, 5), ((24, 17), 3, 9), ((9, 8), 1, 1),
                                              ((64, 48), 8, 9)])
def test_spatial_groups_equal_a_global_stable_sort(shape, step, rng_):
    w, h = shape
    plane = synth.gray_clip(40 + w, 1, w, h, 5 / 255)[0]
    p = po.make_params(w, h, block_step=step, bm_range=rng_)
    _, g = po.estimate(p, [plane])
    np.testing.assert_array_equal(g, _brute_groups(plane, p))


def test_flat_frame_ties_resolve_by_scan_order():
    plane = np.full((24, 24), 0.25, np.float32)
    p = po.make_params(24, 24, block_step=4, bm_range=3)
    _, g = po.estimate(p, [plane])
    np.testing.assert_array_equal(g, _brute_groups(plane, p))


def test_every_group_holds_its_reference_block():
    clip = synth.shifted_clip(50, 5, 64, 48, 5 / 255, (1.0, 0.5))
    p = po.make_params(64, 48, block_step=4, bm_range=8, radius=2, ps_num=2, ps_range=3)
    _, g = po.estimate(p, [clip[z] for z in range(5)])
    n, nx, ny = po.ref_grid(p)
    for by in range(ny):
        for bx in range(nx):
            x, y = min(bx * 4, 56), min(by * 4, 40)
            ent = [tuple(e) for e in g[by, bx]]
            assert (x, y, 2) in ent
            assert all(0 <= e[2] <= 4 for e in ent)


# ------------------------------------------------------------- host logic
def test_ref_grid_counts():
    assert po.ref_grid(po.make_params(640, 360, block_step=8))[0] == 80 * 45
    assert po.ref_grid(po.make_params(1920, 1080, block_step=4))[0] == 479 * 269
    assert po.ref_grid(po.make_params(3840, 2160, block_step=4))[0] == 959 * 539
    assert po.ref_grid(po.make_params(512, 512, block_step=4))[0] == 127 * 127


def test_agg_z_table_by_hand():
    # n = 0, r = 2: frame 0's own slice three times, then frames 1 and 2
    assert [po.agg_z(i, 0, 10, 2) for i in range(5)] == [2, 2, 2, 1, 0]
    assert [po.agg_z(i, 1, 10, 2) for i in range(5)] == [3, 3, 2, 1, 0]
    assert [po.agg_z(i, 5, 10, 2) for i in range(5)] == [4, 3, 2, 1, 0]
    assert [po.agg_z(i, 9, 10, 2) for i in range(5)] == [4, 3, 2, 2, 2]
    for r in range(1, 5):
        for nf in (1, 2, 5, 12):
            for n in range(nf):
                for i in range(2 * r + 1):
                    m = min(max(n - r + i, 0), nf - 1)
                    z = po.agg_z(i, n, nf, r)
                    assert min(max(m - r + z, 0), nf - 1) == n, "every slice lands on frame n"


def test_sigma_scaling_constants():
    assert po.SIGMA_FACTOR_BASIC == np.float32(0.50823534)
    assert po.SIGMA_FACTOR_FINAL == np.float32(0.1882353)
    p = po.make_params(16, 16, sigma=3.0)
    assert np.float32(p.sigma) == np.float32(3.0) * np.float32(po.SIGMA_FACTOR_BASIC)


# ----------------------------------------------------------- whole filter
def test_sigma_zero_reproduces_the_clip():
    clip = synth.gray_clip(60, 2, 72, 40, 5 / 255)
    out = po.clip(po.make_params(72, 40, sigma=0.0, block_step=4, bm_range=6), clip)
    np.testing.assert_allclose(out, clip, atol=2e-6)


def test_constant_clip_stays_constant():
    clip = np.full((3, 32, 40), 0.375, np.float32)
    out = po.clip(po.make_params(40, 32, block_step=3, bm_range=5, radius=1, ps_range=2), clip)
    np.testing.assert_allclose(out, 0.375, atol=1e-6)


def test_extractor_quantises_to_its_grid():
    clip = synth.gray_clip(61, 1, 48, 32, 3 / 255)
    p = po.make_params(48, 32, block_step=4, bm_range=6, extractor_exp=2)
    res, _ = po.estimate(p, [clip[0]])
    # (w*p + 4) - 4 lies on the float grid of [4, 8): multiples of 2^-21,
    # or 2^-22 for sums that fall below 4
    q = res.astype(np.float64) * 2.0 ** 22
    assert np.array_equal(q, np.round(q))


def test_denoising_reduces_the_error():
    clean = synth.clean_image(np.random.Generator(np.random.PCG64(7)), 96, 64).astype(np.float32)
    noisy = np.clip(clean + np.random.default_rng(7).normal(0, 5 / 255, clean.shape), 0, 1).astype(np.float32)
    out = po.clip(po.make_params(96, 64, sigma=5.0, block_step=4, bm_range=8), noisy[None])[0]
    inner = (slice(4, -4), slice(4, -4))
    err_in = np.abs(noisy - np.clip(clean, 0, 1))[inner].mean()
    err_out = np.abs(out - np.clip(clean, 0, 1))[inner].mean()
    assert err_out < 0.8 * err_in


def test_f64_oracle_close_to_f32():
    clip = synth.gray_clip(62, 2, 64, 48, 3 / 255)
    p = po.make_params(64, 48, block_step=4, bm_range=8)
    a = po.clip(p, clip)
    b = po.clip(p, clip, f64=True)
    assert np.abs(a - b).max() < 5e-3


# ------------------------------------------------------------ golden vectors
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    """(clip, BM3Dv2 arguments, groups (nf, ny, nx, 8, 3), denoised frames)."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    kw = json.loads(bytes(z["args"]).decode())
    return z["clip"], kw, z["groups"].astype(np.int32), z["out"]


GOLDEN_CASES = ["spatial_96x80", "temporal_72x56_r2", "exact_48x40_r3"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_reproduces_the_golden_vectors(name):
    clip, kw, groups, out = load_golden(name)
    nf, h, w = clip.shape
    p = po.make_params(w, h, nthreads=3, **kw)
    got = po.clip(p, clip)
    r = kw.get("radius", 0)
    for m in range(nf):
        window = [clip[min(max(m - r + z, 0), nf - 1)] for z in range(2 * r + 1)]
        _, g = po.estimate(p, window)
        assert np.array_equal(g, groups[m]), f"{name}: frame {m} groups changed"
    if kw.get("extractor_exp", 0):
        assert np.array_equal(got, out)
    else:
        np.testing.assert_allclose(got, out, rtol=0, atol=1e-6)
