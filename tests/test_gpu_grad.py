"""Differentiable form: backward pass against float64 references.

With the block-match groups held fixed, the basic filter is
out = N(x) / D with N = sum_g w_g G_g^T C^T M_g C G_g x: gather the group's
8 blocks, orthonormal 3-D DCT, keep mask, inverse, weight, scatter.  The
float64 model below implements that with scipy and is used two ways:

* reference gradient: with the GPU's own masks and weights, the gradient of
  sum(g * out) is A^T (g / D) = A (g / D), evaluated in float64;
* finite differences: with the masks recomputed from the float64 spectrum at
  every evaluation (central differences, eps = 1e-6) on 64x64 crops, at
  pixels away from threshold boundaries (no coefficient of a group touching
  the pixel within 1e-4 of the threshold, and the GPU and float64 masks
  agreeing on those groups).
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.fft import dctn, idctn

pytestmark = pytest.mark.gpu

KW = dict(sigma=3.0, block_step=4, bm_range=9)


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _unpack(groups_packed):
    g = groups_packed.astype(np.int64) & 0xFFFFFFFF
    return np.stack([g & 0x3FFF, (g >> 14) & 0x3FFF, g >> 28], axis=-1)


def _index(groups):
    ar = np.arange(8)
    ys = groups[..., 1][..., None, None] + ar[:, None]
    xs = groups[..., 0][..., None, None] + ar[None, :]
    return ys, xs  # broadcast to (nref, 8, 8, 8)


def _masks_from_words(words):
    w = words.astype(np.uint64)
    bits = (w[..., None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)  # (nref, u, i)
    m = bits.reshape(w.shape[0], 8, 8, 8)  # (nref, u, w, v)
    return np.transpose(m, (0, 2, 3, 1)).astype(bool)  # (nref, w, v, u)


def _apply(x, groups, keep, weights):
    ys, xs = _index(groups)
    blocks = x[ys, xs]
    out = idctn(np.where(keep, dctn(blocks, axes=(1, 2, 3), norm="ortho"), 0.0), axes=(1, 2, 3), norm="ortho")
    n = np.zeros_like(x)
    d = np.zeros_like(x)
    np.add.at(n, (ys, xs), weights[:, None, None, None] * out)
    np.add.at(d, (ys, xs), np.broadcast_to(weights[:, None, None, None], out.shape))
    return n, d


def _model(x, groups, sigma):
    """float64 forward with fixed groups, masks from x's own spectrum."""
    ys, xs = _index(groups)
    c = 64.0 * dctn(x[ys, xs], axes=(1, 2, 3), norm="ortho")
    keep = np.abs(c) >= sigma
    keep[:, 0, 0, 0] = True
    w = 1.0 / keep.sum(axis=(1, 2, 3))
    n, d = _apply(x, groups, keep, w)
    return n / d, keep, c


def _gpu_forward_backward(clip, g, recompute=False):
    torch = _torch()
    from vsfeel.autograd import bm3d_differentiable
    from vsfeel.torch_bm3d import Instance
    x = torch.from_numpy(clip).to("cuda:0").requires_grad_(True)
    out = bm3d_differentiable(x, recompute=recompute, **KW)
    out.backward(torch.from_numpy(g).to("cuda:0"))
    inst = Instance(*clip.shape, radius=0, **KW)
    _, saved = inst.forward_saved(x.detach())
    torch.cuda.synchronize()
    return (out.detach().cpu().numpy(), x.grad.cpu().numpy(),
            [t.cpu().numpy() for t in saved], inst)


def test_forward_saved_equals_the_plain_forward():
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    clip = synth.workload(5, frames=4, width=128, height=96)
    inst = Instance(*clip.shape, radius=0, **KW)
    x = torch.from_numpy(clip).to("cuda:0")
    plain = inst.process(x).cpu().numpy()
    out, (groups, masks, weights, denom) = inst.forward_saved(x)
    assert np.array_equal(out.cpu().numpy(), plain)
    g, _ = inst.groups(x, 1)
    assert np.array_equal(_unpack(groups[1].cpu().numpy()), g.reshape(-1, 8, 3))


def test_backward_matches_float64_reference_gradient():
    from oracle import pyoracle
    from vsfeel import synth
    clip = synth.workload(5, frames=2, width=160, height=128)
    g = np.random.default_rng(5).standard_normal(clip.shape).astype(np.float32)
    out, grad, (groups, masks, weights, denom), inst = _gpu_forward_backward(clip, g)
    for f in range(clip.shape[0]):
        gr = _unpack(groups[f])
        keep = _masks_from_words(masks[f])
        w = weights[f].astype(np.float64)
        n, d = _apply(clip[f].astype(np.float64), gr, keep, w)
        np.testing.assert_allclose(out[f], n / d, atol=2e-6)
        np.testing.assert_allclose(denom[f], d, rtol=1e-6)
        ref, _ = _apply(g[f].astype(np.float64) / d, gr, keep, w)
        err = np.abs(grad[f] - ref).max()
        assert err <= 1e-3 * np.abs(ref).max(), f"frame {f}: max err {err}"
        assert np.linalg.norm(grad[f] - ref) <= 1e-5 * np.linalg.norm(ref)


def test_backward_matches_finite_differences_on_crops():
    from oracle import pyoracle
    from vsfeel import synth
    sigma = float(np.float32(3.0) * np.float32(pyoracle.SIGMA_FACTOR_BASIC))
    full = synth.workload(5, frames=2, width=512, height=512)
    crops = np.stack([full[0, 100:164, 200:264], full[1, 300:364, 40:104]]).astype(np.float32)
    g = np.random.default_rng(11).standard_normal(crops.shape).astype(np.float32)
    _, grad, (groups, masks, _, _), _ = _gpu_forward_backward(crops, g)
    rng = np.random.default_rng(3)
    eps = 1e-6
    checked = 0
    for f in range(crops.shape[0]):
        x = crops[f].astype(np.float64)
        gg = g[f].astype(np.float64)
        gr = _unpack(groups[f])
        out0, keep0, c0 = _model(x, gr, sigma)
        margin = np.abs(np.abs(c0) - sigma)
        margin[:, 0, 0, 0] = np.inf
        bad_group = (margin.min(axis=(1, 2, 3)) < 1e-4) | (keep0 != _masks_from_words(masks[f])).any(axis=(1, 2, 3))
        ys, xs = _index(gr)
        bad = np.zeros_like(x)
        np.add.at(bad, (ys, xs), np.broadcast_to(bad_group[:, None, None, None].astype(float), (len(gr), 8, 8, 8)))
        ok = np.argwhere(bad == 0)
        assert len(ok) > 200, f"crop {f}: only {len(ok)} pixels away from threshold boundaries"
        for p in ok[rng.choice(len(ok), size=48, replace=False)]:
            xp = x.copy()
            xm = x.copy()
            xp[tuple(p)] += eps
            xm[tuple(p)] -= eps
            op, kp, _ = _model(xp, gr, sigma)
            om, km, _ = _model(xm, gr, sigma)
            if not (np.array_equal(kp, keep0) and np.array_equal(km, keep0)):
                continue
            fd = ((gg * op).sum() - (gg * om).sum()) / (2 * eps)
            an = grad[f][tuple(p)]
            assert abs(an - fd) <= 1e-3 * abs(fd) + 1e-5, f"crop {f} pixel {tuple(p)}: {an} vs FD {fd}"
            checked += 1
    assert checked >= 80


def test_recompute_variant_gives_the_same_gradient():
    from vsfeel import synth
    clip = synth.workload(5, frames=2, width=96, height=80)
    g = np.random.default_rng(9).standard_normal(clip.shape).astype(np.float32)
    _, a, _, _ = _gpu_forward_backward(clip, g, recompute=False)
    _, b, _, _ = _gpu_forward_backward(clip, g, recompute=True)
    assert np.array_equal(a, b)


def test_full_size_batch_runs_and_is_finite():
    """The differentiable workload's shape: 16 frames of 512x512."""
    torch = _torch()
    from vsfeel import synth
    from vsfeel.autograd import bm3d_differentiable
    clip = synth.workload(5)
    x = torch.from_numpy(clip).to("cuda:0").requires_grad_(True)
    out = bm3d_differentiable(x, **KW)
    g = torch.from_numpy(np.random.default_rng(1).standard_normal(clip.shape).astype(np.float32)).to("cuda:0")
    (out * g).sum().backward()
    torch.cuda.synchronize()
    assert torch.isfinite(x.grad).all()
    # linearity of the backward in the upstream gradient
    x2 = x.detach().clone().requires_grad_(True)
    out2 = bm3d_differentiable(x2, **KW)
    (out2 * (2.0 * g)).sum().backward()
    torch.cuda.synchronize()
    assert torch.allclose(x2.grad, 2.0 * x.grad, rtol=1e-5, atol=1e-6)


def test_frame_lanes_leave_saved_state_and_gradient_unchanged(monkeypatch):
    # forward_saved and backward run a batch's frames on up to three lanes
    # (own streams, result slots and buffers); with VSFEEL_BM3D_LANES=1 they
    # run one after another: every output, saved tensor and gradient is
    # bit-identical
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance
    clip = synth.workload(5, frames=7, width=160, height=128)
    g = torch.from_numpy(np.random.default_rng(4).standard_normal(clip.shape).astype(np.float32)).to("cuda:0")
    x = torch.from_numpy(clip).to("cuda:0")
    print(f"frame lanes clip={clip.shape}", flush=True)
    got = {}
    for lanes in ("1", None):
        if lanes is None:
            monkeypatch.delenv("VSFEEL_BM3D_LANES", raising=False)
        else:
            monkeypatch.setenv("VSFEEL_BM3D_LANES", lanes)
        inst = Instance(*clip.shape, radius=0, **KW)
        out, saved = inst.forward_saved(x)
        gin = inst.backward(g, saved)
        torch.cuda.synchronize()
        got[lanes] = [t.cpu().numpy() for t in (out, *saved, gin)]
    for name, a, b in zip(("out", "groups", "masks", "weights", "denom", "grad"), got["1"], got[None]):
        assert np.array_equal(a, b), f"{name} differs between one and three lanes"
