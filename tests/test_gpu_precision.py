"""The precision form: block-matching distances of the pixels rounded once to
FP16 or BF16, computed as norms - 2 dot on MFMA with FP32 accumulation, and
a shortlist of twice the entries a search keeps rescored with the FP32 sums
of the plain form.  Its groups must equal the plain form's on at least
99.9 % of the ref blocks (identical wherever the plain matches lie in the
shortlist), and its output must stay close to the plain form's and to the
FP64 oracle's.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = max(1, min(16, os.cpu_count() or 1))
AGREE = 0.999  # share of identical groups

CASES = {
    # the 1080p workload's parameters on a smaller frame, and frame 7 of the
    # generator (texture along the rows)
    "c2_small": (lambda s: s.workload(2, frames=8, width=480, height=272)[[0, 7]],
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=0)),
    # the temporal workload (radius 2, predictive search 2/7, extractor)
    "c3_small": (lambda s: s.workload(3, frames=6, width=320, height=192)[0],
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7, extractor_exp=2)),
    # the precision-form workload's parameters (radius 1, sigma 5)
    "c4_small": (lambda s: s.workload(4, frames=3, width=480, height=270),
                 dict(sigma=5.0, block_step=4, bm_range=16, radius=1, ps_num=2, ps_range=7)),
    # full-size frames: the 1080p temporal workload (radius 2) and the 4K
    # precision-form workload's own frames (radius 1)
    "c3_1080p": (lambda s: s.workload(3, frames=5)[0],
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7, extractor_exp=2)),
    "c4_4k": (lambda s: s.workload(4, frames=3),
              dict(sigma=5.0, block_step=4, bm_range=16, radius=1, ps_num=2, ps_range=7)),
    # the CPU-runnable workload's parameters (step 8, range 9: 19-wide windows)
    "c1": (lambda s: s.workload(1, frames=2),
           dict(sigma=3.0, block_step=8, bm_range=9, radius=0)),
    # windows wider than 48 columns: a second group of column tiles without
    # the k < 7 correction
    "wide_range": (lambda s: s.gray_clip(12, 2, 200, 120, 4 / 255),
                   dict(sigma=3.0, block_step=4, bm_range=28, radius=0)),
    # ragged sizes and a step that leaves runs of unaligned windows
    "ragged": (lambda s: s.gray_clip(10, 2, 77, 45, 5 / 255),
               dict(sigma=4.0, block_step=3, bm_range=5, radius=0)),
    # a window smaller than the shortlist
    "tiny": (lambda s: s.gray_clip(11, 2, 12, 10, 5 / 255),
             dict(sigma=3.0, block_step=1, bm_range=1, radius=0)),
    # every distance ties: the seq order decides, exactly as in the plain form
    "flat": (lambda s: np.full((2, 40, 48), 0.5, np.float32),
             dict(sigma=3.0, block_step=4, bm_range=7, radius=0)),
    "blank": (lambda s: np.zeros((2, 64, 96), np.float32),
              dict(sigma=0.7, block_step=4, bm_range=16, radius=0)),
}


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _groups(inst, x, nf):
    return np.stack([inst.groups(x, m)[0] for m in range(nf)])


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("name", list(CASES))
def test_groups_agree_with_plain(name, precision):
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    make, kw = CASES[name]
    clip = np.ascontiguousarray(make(synth), np.float32)
    nf, h, w = clip.shape
    print(f"precision case {name} {precision} clip={clip.shape} args={kw}", flush=True)
    x = torch.from_numpy(clip).to("cuda:0")
    plain = Instance(nf, h, w, **kw)
    fast = Instance(nf, h, w, precision=precision, **kw)
    gp, gf = _groups(plain, x, nf), _groups(fast, x, nf)
    same = (gp == gf).all(axis=(3, 4))
    share = float(same.mean())
    print(f"  identical groups {share:.6f} ({int((~same).sum())} of {same.size} differ)", flush=True)
    assert share >= AGREE, f"{name} {precision}: {share:.5f} of the groups identical"
    if name in ("flat", "blank"):
        assert share == 1.0
    op = plain.process(x).cpu().numpy()
    of = fast.process(x).cpu().numpy()
    assert np.isfinite(of).all()
    diff = float(np.abs(op - of).max())
    if share == 1.0:
        # the same groups: the same kernels downstream, the same frames up to
        # the order of float atomics (other frames' blocks at radius 1)
        assert diff <= 1e-5, diff
    else:
        assert diff < 0.05, diff


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_error_against_fp64_oracle(precision):
    """Max-abs error and PSNR of both forms against the FP64 oracle on one
    frame of the precision-form workload's parameters (smaller frame)."""
    torch = _torch()
    from oracle import pyoracle
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    clip = synth.workload(4, frames=3, width=320, height=180)
    nf, h, w = clip.shape
    kw = dict(sigma=5.0, block_step=4, bm_range=16, radius=1, ps_num=2, ps_range=7)
    x = torch.from_numpy(clip).to("cuda:0")
    plain = Instance(nf, h, w, **kw).process(x).cpu().numpy()
    fast = Instance(nf, h, w, precision=precision, **kw).process(x).cpu().numpy()
    p = pyoracle.make_params(w, h, nthreads=NT, **kw)
    ref = pyoracle.clip(p, clip, f64=True)

    def psnr(a, b):
        mse = float(np.mean((a.astype(np.float64) - b) ** 2))
        return 10 * np.log10(1.0 / mse) if mse > 0 else float("inf")

    ep, ef = float(np.abs(plain - ref).max()), float(np.abs(fast - ref).max())
    pp, pf = psnr(plain, ref), psnr(fast, ref)
    print(f"{precision}: max-abs vs FP64 plain {ep:.3g} precision {ef:.3g}; "
          f"PSNR vs FP64 plain {pp:.2f} dB precision {pf:.2f} dB", flush=True)
    assert ef < 0.02
    assert pf > 60.0


def test_precision_arguments_rejected():
    _torch()
    from vsfeel.bm3d import Error
    from vsfeel.torch_bm3d import Instance

    with pytest.raises(Error, match="precision"):
        Instance(5, 64, 64, sigma=3.0, radius=3, precision="fp16")
    with pytest.raises(Error, match="precision"):
        Instance(2, 64, 64, sigma=3.0, precision=3)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_final_pass_groups_agree_with_plain(precision):
    # the Wiener pass matches on the ref clip: the plane set is the ref
    # frame's (src/bm3d.comp:401-443 with the groups of src/bm3d.comp:451-663)
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    clip = synth.workload(4, frames=4, width=320, height=192)
    nf, h, w = clip.shape
    kw = dict(sigma=5.0, block_step=4, bm_range=12, radius=1, ps_num=2, ps_range=5)
    print(f"final pass {precision} clip={clip.shape}", flush=True)
    x = torch.from_numpy(clip).to("cuda:0")
    basic = Instance(nf, h, w, **kw).process(x)
    plain = Instance(nf, h, w, final=True, **kw)
    fast = Instance(nf, h, w, final=True, precision=precision, **kw)
    gp = np.stack([plain.groups(x, m, ref=basic)[0] for m in range(nf)])
    gf = np.stack([fast.groups(x, m, ref=basic)[0] for m in range(nf)])
    same = (gp == gf).all(axis=(3, 4))
    share = float(same.mean())
    print(f"  identical groups {share:.6f} ({int((~same).sum())} of {same.size} differ)", flush=True)
    assert share >= AGREE
    op = plain.process(x, ref=basic).cpu().numpy()
    of = fast.process(x, ref=basic).cpu().numpy()
    assert np.abs(op - of).max() <= (1e-5 if share == 1.0 else 0.05)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_host_requests_equal_the_device_path(precision):
    # BM3Dv2(..., precision=...) through the host-buffer path (getFrame)
    torch = _torch()
    from vsfeel import BM3Dv2, Clip, synth
    from vsfeel.torch_bm3d import Instance

    clip = synth.workload(3, frames=6, width=192, height=128)[0]
    nf, h, w = clip.shape
    kw = dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7, extractor_exp=2)
    print(f"host path {precision} clip={clip.shape}", flush=True)
    want = Instance(nf, h, w, precision=precision, **kw).process(torch.from_numpy(clip).to("cuda:0")).cpu().numpy()
    node = BM3Dv2(Clip.gray(clip), precision=precision, **kw)
    for n in (3, 0, 5, 1):
        assert np.array_equal(node.get_frame(n)[0], want[n]), f"frame {n} differs"
