"""Host logic of the C ABI that needs no GPU: exported symbols and the
argument validation of BM3DCreate (src/bm3d.cpp:1062-1192), which runs
before any device is touched."""

from __future__ import annotations

import os
import re

import numpy as np
import pytest

import vsfeel
from vsfeel import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"^\w[\w\s\*]*?\b(vsfeel_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_entry_point():
    L = _lib.load()
    declared = _declared()
    assert declared, "no declarations found in include/"
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), f"{name} is declared but not exported"


def _gray(n=4, h=32, w=40, dtype=np.float32):
    return vsfeel.Clip.gray(np.zeros((n, h, w), dtype))


def _err(clip, **kw):
    with pytest.raises(vsfeel.Error) as ei:
        vsfeel.BM3Dv2(clip, **kw)
    return str(ei.value)


@pytest.mark.parametrize("kw,msg", [
    (dict(radius=5), 'BM3D: "radius" must be in range [0, 4]'),
    (dict(radius=-1), 'BM3D: "radius" must be in range [0, 4]'),
    (dict(sigma=-0.5), 'BM3D: "sigma" must be non-negative'),
    (dict(sigma=[3.0, -1.0]), 'BM3D: "sigma" must be non-negative'),
    (dict(block_step=0), 'BM3D: "block_step" must be in range [1, 8]'),
    (dict(block_step=[4, 9]), 'BM3D: "block_step" must be in range [1, 8]'),
    (dict(bm_range=0), 'BM3D: "bm_range" must be positive'),
    (dict(ps_num=9), 'BM3D: "ps_num" must be in range [1, 8]'),
    (dict(ps_num=0), 'BM3D: "ps_num" must be in range [1, 8]'),
    (dict(ps_range=0), 'BM3D: "ps_range" must be positive'),
    (dict(num_streams=0), 'BM3D: "num_streams" must be positive'),
])
def test_argument_validation_messages(kw, msg):
    assert _err(_gray(), **kw) == msg


def test_validation_order_follows_the_reference():
    # sigma is checked before block_step, block_step before bm_range, ...
    assert "sigma" in _err(_gray(), sigma=-1, block_step=0, radius=9)
    assert "block_step" in _err(_gray(), block_step=0, bm_range=0)
    assert "bm_range" in _err(_gray(), bm_range=0, radius=9)
    assert "radius" in _err(_gray(), radius=9, ps_num=0)
    assert "ps_num" in _err(_gray(), ps_num=0, ps_range=0)
    assert "ps_range" in _err(_gray(), ps_range=0, num_streams=0)


def test_only_32_bit_float_is_accepted():
    assert _err(_gray(dtype=np.uint8)) == "BM3D: only constant format 32 bit float input supported"
    assert _err(_gray(dtype=np.uint16)) == "BM3D: only constant format 32 bit float input supported"


def test_rgb_is_rejected_with_the_reference_message():
    c = vsfeel.Clip([np.zeros((2, 16, 16), np.float32)] * 3, vsfeel.RGB)
    # the reference passes a pre-prefixed string to its prefixing set_error
    assert _err(c) == "BM3D: BM3D: only Gray and YUV input are currently supported"


def test_ref_clip_must_match():
    clip = _gray(4, 32, 40)
    assert _err(clip, ref=_gray(4, 32, 41)) == 'BM3D: "ref" must be of the same dimensions as "clip"'
    assert _err(clip, ref=_gray(5, 32, 40)) == 'BM3D: "ref" must be of the same number of frames as "clip"'
    assert _err(clip, ref=_gray(4, 32, 40, np.uint8)) == 'BM3D: "ref" must be of the same format as "clip"'
    yuv = vsfeel.Clip.yuv(np.zeros((4, 32, 40), np.float32), np.zeros((4, 16, 20), np.float32),
                          np.zeros((4, 16, 20), np.float32))
    assert _err(clip, ref=yuv) == 'BM3D: "ref" must be of the same format as "clip"'


def test_frames_smaller_than_a_block_are_rejected():
    assert _err(_gray(2, 7, 40)) == "BM3D: the luma plane must be at least 8x8"


def test_valid_arguments_reach_the_device_stage():
    """With valid arguments creation only fails for want of a GPU here."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    msg = _err(_gray(), sigma=[3.0, 2.0, 1.0], block_step=4, bm_range=16, radius=2, ps_num=2,
               ps_range=7, num_streams=2, extractor_exp=2)
    assert msg.startswith("BM3D: ") and "device" in msg.lower()
