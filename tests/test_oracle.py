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


def _d32(cur, plane, left, top, rw, rh):
    """float32 distances of the sliding metric over a window, written out
    independently of the oracle: column SSDs v_j = sum_i (cur[i, min(j, 7)]
    - plane[cy + i, left + j])^2 with even and odd rows in two partial sums
    (each product and sum rounded to float32, none fused), then
    D(k) = ((v_k + v_k+1) + (v_k+2 + v_k+3)) + ((v_k+4 + v_k+5) + (v_k+6 + v_k+7))."""
    ncol = rw + 7
    a = np.minimum(np.arange(ncol), 7)
    ref = cur[:, a].astype(np.float32)                                   # [8, ncol]
    rows = np.lib.stride_tricks.sliding_window_view(plane[top:top + rh + 7, left:left + ncol], 8, axis=0)
    d = ref[None, :, :] - np.transpose(rows, (0, 2, 1))                  # [rh, 8, ncol] float32
    e0 = d[:, 0] * d[:, 0]
    e1 = d[:, 1] * d[:, 1]
    for i in range(2, 8, 2):
        e0 = e0 + d[:, i] * d[:, i]
        e1 = e1 + d[:, i + 1] * d[:, i + 1]
    v = e0 + e1                                                          # [rh, ncol]
    k = lambda j: v[:, j:j + rw]
    return ((k(0) + k(1)) + (k(2) + k(3))) + ((k(4) + k(5)) + (k(6) + k(7)))


def _window(cx, cy, rng_, W, H):
    left, right = max(cx - rng_, 0), min(cx + rng_, W - 8)
    top, bottom = max(cy - rng_, 0), min(cy + rng_, H - 8)
    return left, top, right - left + 1, bottom - top + 1


FLT_MAX = float(np.finfo(np.float32).max)


def _smallest8(cands):
    """The 8 smallest (distance, seq) of [(d, seq, x, y)], placeholders
    (FLT_MAX, 0, 0) past the last; distances at FLT_MAX never count."""
    cands = sorted((c for c in cands if c[0] < FLT_MAX), key=lambda c: (c[0], c[1]))[:8]
    return cands + [(FLT_MAX, 2 ** 31 - 1, 0, 0)] * (8 - len(cands))


def _window_cands(cur, plane, left, top, rw, rh, seq_base):
    d = _d32(cur, plane, left, top, rw, rh)
    return [(float(d[r, k]), seq_base + r * rw + k, left + k, top + r) for r in range(rh) for k in range(rw)]


def _brute_temporal_groups(clip, p, m):
    """The groups of frame m's estimate, restated from the shader's semantics
    (src/bm3d.comp:479-663) with global sorts instead of lane lists: the
    spatial top-8 by (distance, seq); then per direction -1, +1 and step
    t = 1..r the windows +-ps_range around the first ps_num predictors (the
    previous list's entries; the spatial top-8 for t = 1), seq = i (2
    ps_range + 1)^2 + position, the step's top-8 by (distance, seq), its first
    ps_num inserted into the group with a strict '<' (after equal distances);
    finally the self-block rule."""
    W, H, s, R, r = p.width, p.height, p.block_step, p.bm_range, p.radius
    PN, PR = p.ps_num, p.ps_range
    win = (2 * PR + 1) ** 2
    nf = len(clip)
    frame = lambda z: clip[min(max(m - r + z, 0), nf - 1)]
    _, nx, ny = po.ref_grid(p)
    out = np.zeros((ny, nx, 8, 3), np.int32)
    for by in range(ny):
        for bx in range(nx):
            x, y = min(bx * s, W - 8), min(by * s, H - 8)
            cur = frame(r)[y:y + 8, x:x + 8]
            sp = _smallest8(_window_cands(cur, frame(r), *_window(x, y, R, W, H), 0))
            group = [(c[0], c[2], c[3], r) for c in sp]
            for direction in (-1, 1):
                preds = [(c[2], c[3]) for c in sp]
                for t in range(1, r + 1):
                    z = r + direction * t
                    cands = []
                    for i in range(PN):
                        cands += _window_cands(cur, frame(z), *_window(*preds[i], PR, W, H), i * win)
                    step = _smallest8(cands)
                    for i in range(PN):
                        e = step[i][0]
                        if e < group[7][0]:
                            pos = next(q for q in range(8) if e < group[q][0])
                            group = group[:pos] + [(e, step[i][2], step[i][3], z)] + group[pos:7]
                    preds = [(c[2], c[3]) for c in step]
            if not any(g[1] == x and g[2] == y and g[3] == r for g in group):
                group = [(-1.0, x, y, r)] + group[:7]
            out[by, bx] = [[g[1], g[2], g[3]] for g in group]
    return out


def test_d32_restatement_equals_the_oracle_distances_bitwise():
    """The float32 restatement above reproduces the oracle's window
    distances bit for bit (fp_mode 1 rounds these sums unfused)."""
    plane = synth.gray_clip(63, 1, 48, 40, 5 / 255)[0]
    for (x, y, R) in ((17, 13, 6), (0, 0, 5), (40, 32, 4)):
        d, (left, top) = po.window_dist(plane, x, y, R, x, y, exact=False)
        mine = _d32(plane[y:y + 8, x:x + 8], plane, left, top, d.shape[1], d.shape[0])
        assert np.array_equal(d.view(np.uint32), mine.view(np.uint32))


# three shifted clips: radius 1 and 2, one to three predictors, windows
# clamped at the frame edges (small frames, motion up to 2 px per frame)
TEMPORAL_CLIPS = [
    ((70, 4, 40, 32, 3 / 255, (1.5, -1.0)), dict(block_step=4, bm_range=5, radius=1, ps_num=1, ps_range=3)),
    ((71, 5, 44, 36, 3 / 255, (-1.0, 0.5)), dict(block_step=4, bm_range=4, radius=2, ps_num=2, ps_range=2)),
    ((72, 5, 36, 28, 5 / 255, (2.0, 1.0)), dict(block_step=3, bm_range=6, radius=2, ps_num=3, ps_range=4)),
]


def _periodic_clip(nf, w, h):
    """Identical noise-free frames of a pattern periodic in x (3 px) and y
    (2 px), exact in float32: distances tie across positions, windows and
    directions, so the seq order of the windows and the order of the two
    directions decide the groups."""
    y, x = np.mgrid[0:h, 0:w]
    frame = ((x % 3) + 4 * (y % 2) + 0.5 * ((x // 7) % 2)).astype(np.float32) / 16
    return np.repeat(frame[None], nf, axis=0)


@pytest.mark.parametrize("case", range(len(TEMPORAL_CLIPS) + 1))
def test_temporal_groups_equal_a_brute_force(case):
    if case < len(TEMPORAL_CLIPS):
        (seed, nf, w, h, noise, shift), kw = TEMPORAL_CLIPS[case]
        clip = synth.shifted_clip(seed, nf, w, h, noise, shift)
    else:
        nf, w, h = 5, 36, 30
        kw = dict(block_step=4, bm_range=4, radius=2, ps_num=2, ps_range=2)
        clip = _periodic_clip(nf, w, h)
    p = po.make_params(w, h, **kw)
    r = kw["radius"]
    for m in range(nf):
        window = [clip[min(max(m - r + z, 0), nf - 1)] for z in range(2 * r + 1)]
        _, g = po.estimate(p, window)
        want = _brute_temporal_groups(clip, p, m)
        bad = np.argwhere((g != want).any(axis=(2, 3)))
        assert len(bad) == 0, f"frame {m}: {len(bad)} groups differ, first {bad[0]}: {g[tuple(bad[0])].tolist()} " \
                              f"vs {want[tuple(bad[0])].tolist()}"


@pytest.mark.parametrize("radius", [1, 2, 3])
def test_sigma_zero_reproduces_the_clip_at_radius(radius):
    """sigma = 0 keeps every coefficient: every group reproduces its blocks,
    weights are equal, and every slice aggregated for frame n holds frame n's
    own pixels (agg_z), so the output is the input at any radius."""
    clip = synth.shifted_clip(64 + radius, 2 * radius + 2, 40, 32, 5 / 255, (1.0, -0.5))
    p = po.make_params(40, 32, sigma=0.0, block_step=4, bm_range=5, radius=radius, ps_num=2, ps_range=3)
    out = po.clip(p, clip)
    np.testing.assert_allclose(out, clip, atol=2e-6)


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
