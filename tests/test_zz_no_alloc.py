"""The allocation-free entry point (vsfeel_bm3d_process_device_static)
against vsfeel_bm3d_process_device on the same frames, both through the C
ABI: two calls in a row, the second on smaller frames, each bit-identical
with the handle-based entry point (same kernels, order-free fixed-point
tile sums).  Frames too large for its static scratch are rejected.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _args(nf, h, w, keep, **kw):
    from vsfeel import _lib
    from vsfeel.bm3d import FLOAT, GRAY

    a = _lib.Args()
    a.clip = _lib.Format(GRAY, FLOAT, 32, 0, 0, w, h, nf)
    for name, ctype, dt in (("sigma", ctypes.c_double, np.float64), ("block_step", ctypes.c_int64, np.int64),
                            ("bm_range", ctypes.c_int64, np.int64), ("radius", ctypes.c_int64, np.int64)):
        arr = np.ascontiguousarray(np.atleast_1d(np.asarray(kw[name], dt)))
        keep.append(arr)
        setattr(a, name, arr.ctypes.data_as(ctypes.POINTER(ctype)))
        setattr(a, "num_" + name, 1)
    a.fp_mode = 1
    a.device = 0
    return a


def _static(x, out, **kw):
    torch = _torch()
    from vsfeel import _lib

    nf, h, w = x.shape
    keep = []
    a = _args(nf, h, w, keep, **kw)
    c = _lib.DeviceClip()
    c.data = x.data_ptr()
    c.frame_stride = x.stride(0)
    c.pitch = x.stride(1)
    c.ref = None
    err = _lib.err_buf()
    stream = torch.cuda.current_stream(x.device).cuda_stream
    rc = _lib.load().vsfeel_bm3d_process_device_static(ctypes.byref(a), ctypes.byref(c), 0, nf, out.data_ptr(),
                                                      out.stride(1), out.stride(0), ctypes.c_void_p(stream),
                                                      err, len(err))
    return rc, err.value.decode()


CASES = [
    # the largest frame the static scratch holds: 1080p at the benchmark's parameters
    ("1080p", 2, 1080, 1920, dict(sigma=3.0, block_step=4, bm_range=16, radius=0)),
    # then smaller frames at other parameters: the same scratch, other layout
    ("small", 3, 144, 200, dict(sigma=5.0, block_step=3, bm_range=7, radius=0)),
]


def test_static_matches_process_device_twice():
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    for name, nf, h, w, kw in CASES:
        print(f"no-alloc case {name}: {nf}x{h}x{w} {kw}", flush=True)
        clip = synth.gray_clip(21, nf, w, h, 3 / 255)
        x = torch.from_numpy(clip).to("cuda:0")
        want = Instance(nf, h, w, **kw).process(x)
        got = torch.full_like(want, float("nan"))
        rc, msg = _static(x, got, **kw)
        assert rc == 0, msg
        torch.cuda.synchronize()
        g, r = got.cpu().numpy(), want.cpu().numpy()
        assert np.isfinite(g).all()
        assert np.array_equal(g, r), f"{name}: max diff {float(np.abs(g - r).max())}"


def test_static_rejects_what_it_cannot_hold():
    torch = _torch()
    from vsfeel import _lib

    x = torch.zeros((1, 2160, 3840), dtype=torch.float32, device="cuda:0")
    out = torch.empty_like(x)
    rc, msg = _static(x, out, sigma=3.0, block_step=4, bm_range=16, radius=0)
    assert rc == _lib.EINVAL and "scratch" in msg, msg
    y = torch.zeros((3, 64, 64), dtype=torch.float32, device="cuda:0")
    rc, msg = _static(y, torch.empty_like(y), sigma=3.0, block_step=4, bm_range=8, radius=1)
    assert rc == _lib.EINVAL and "radius 0" in msg, msg
