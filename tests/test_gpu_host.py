"""The host-buffer path (the VapourSynth getFrame counterpart) on the GPU.

BM3Dv2(Clip) uploads the frames each request needs, runs the estimates the
frame cache lacks and downloads the result (src/bm3d.cpp:698-1028).  Its
frames must equal the device path's in any request order and from several
threads at once (the reference's fmParallel getFrame,
tests/test_bm3dv2.py:77-103), and chroma passes through bit-exact
(src/bm3d.cpp:715-726).  The frame's own estimate slice is summed in integer
fixed point, so at radius 0 and with the extractor on (every contribution on
the extractor's grid) the frames are bit-identical; at radius > 0 without it
the other slices take float atomics in arrival order, which leaves the last
bits free (1e-5, as in the parity tests).
"""

from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _device_frames(clip, kw):
    torch = _torch()
    from vsfeel.torch_bm3d import Instance
    nf, h, w = clip.shape
    inst = Instance(nf, h, w, **kw)
    return inst.process(torch.from_numpy(np.ascontiguousarray(clip)).to("cuda:0")).cpu().numpy()


@pytest.mark.parametrize("radius,extractor_exp", [(0, 0), (2, 2), (2, 0)])
def test_host_path_equals_device_path_in_any_order(radius, extractor_exp):
    _torch()
    from vsfeel import BM3Dv2, Clip, synth
    clip = synth.workload(3, frames=7, width=160, height=96)[0]
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=radius, ps_num=2, ps_range=4,
              extractor_exp=extractor_exp)
    print(f"host path radius={radius} extractor_exp={extractor_exp} clip={clip.shape}", flush=True)
    want = _device_frames(clip, kw)
    node = BM3Dv2(Clip.gray(clip), **kw)
    order = [3, 0, 6, 1, 5, 2, 4, 4, 0]  # seeks, repeats
    for n in order:
        got = node.get_frame(n)[0]
        if radius == 0 or extractor_exp:
            assert np.array_equal(got, want[n]), f"radius {radius}: frame {n} differs"
        else:
            assert np.abs(got - want[n]).max() <= 1e-5, f"radius {radius}: frame {n} differs"


def test_yuv_chroma_passes_through_and_luma_is_denoised():
    _torch()
    from vsfeel import BM3Dv2, Clip, synth
    y, u, v = synth.workload(3, frames=3, width=128, height=64)
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=1, ps_num=2, ps_range=4, extractor_exp=2)
    print(f"yuv clip={y.shape} chroma={u.shape}", flush=True)
    node = BM3Dv2(Clip.yuv(y, u, v), **kw)
    want = _device_frames(y, kw)
    for n in range(3):
        py, pu, pv = node.get_frame(n)
        assert np.array_equal(py, want[n])
        assert np.array_equal(pu, u[n]) and np.array_equal(pv, v[n])
        assert np.abs(py - y[n]).max() > 1e-4


@pytest.mark.parametrize("radius,num_streams,nthreads", [(1, 4, 4), (0, 2, 6), (2, 3, 6)])
def test_parallel_requests_are_deterministic(radius, num_streams, nthreads):
    # requests overlap (uploads, kernels and downloads of different frames on
    # separate streams); more threads than output slots make them wait for one
    _torch()
    from vsfeel import BM3Dv2, Clip, synth
    clip = synth.workload(3, frames=8, width=128, height=80)[0]
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=radius, ps_num=2, ps_range=4, extractor_exp=2,
              num_streams=num_streams)
    print(f"parallel requests clip={clip.shape} radius={radius} streams={num_streams} threads={nthreads}",
          flush=True)
    want = _device_frames(clip, {k: v for k, v in kw.items() if k != "num_streams"})
    node = BM3Dv2(Clip.gray(clip), **kw)
    got = {}
    errors = []

    def worker(frames):
        try:
            for n in frames:
                got.setdefault(n, []).append(node.get_frame(n)[0])
        except Exception as e:  # surfaced below
            errors.append(e)

    rng = np.random.default_rng(4)
    threads = [threading.Thread(target=worker, args=(list(rng.permutation(8)),)) for _ in range(nthreads)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors, errors
    assert sorted(got) == list(range(8))
    for n, frames in got.items():
        assert len(frames) == nthreads
        for f in frames:
            assert np.array_equal(f, want[n]), f"frame {n} differs under parallel requests"


def _padded(a, fill):
    """a (h, w) float32 in rows of a 64-byte-rounded stride; the padding holds `fill`."""
    h, w = a.shape
    sw = -(-w * 4 // 64) * 16
    buf = np.full((h, sw), fill, np.float32)
    buf[:, :w] = a
    return buf


def _raw_frame(node, n, src, ref, dst):
    """vsfeel_bm3d_get_frame on arbitrary (strided) host planes."""
    import ctypes
    from vsfeel import _lib
    from vsfeel.bm3d import Error
    L = _lib.load()
    first, last = ctypes.c_int32(), ctypes.c_int32()
    assert L.vsfeel_bm3d_frame_window(node._handle, n, ctypes.byref(first), ctypes.byref(last)) == _lib.OK
    VP = ctypes.c_void_p

    def arrays(planes):
        fs = range(first.value, last.value + 1)
        return ((VP * len(fs))(*[planes[f].ctypes.data for f in fs]),
                (ctypes.c_ssize_t * len(fs))(*[planes[f].strides[0] for f in fs]))

    cp, cs = arrays(src)
    rp, rs = arrays(ref) if ref is not None else (None, None)
    err = _lib.err_buf()
    rc = L.vsfeel_bm3d_get_frame(node._handle, n, cp, cs, rp, rs, dst.ctypes.data, dst.strides[0], err, len(err))
    if rc != _lib.OK:
        raise Error(err.value.decode())


@pytest.mark.parametrize("final", [False, True])
def test_padded_strides_leave_destination_padding_untouched(final):
    # VapourSynth frames have row strides rounded up (src/bm3d.cpp:996-1017
    # copies rows of width * 4 bytes out of and into them)
    _torch()
    from vsfeel import BM3Dv2, Clip, synth
    nf, h, w = 5, 70, 157
    clip = synth.gray_clip(21, nf, w, h, 6 / 255)
    ref = synth.gray_clip(22, nf, w, h, 2 / 255) if final else None
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=1, ps_num=2, ps_range=4, extractor_exp=2)
    print(f"padded strides final={final} clip={clip.shape}", flush=True)
    node = BM3Dv2(Clip.gray(clip), Clip.gray(ref) if final else None, **kw)
    want = [node.get_frame(n)[0] for n in range(nf)]
    # a fresh node reads the padded planes (nothing cached)
    node2 = BM3Dv2(Clip.gray(clip), Clip.gray(ref) if final else None, **kw)
    src = [_padded(clip[f], np.nan)[:, :w] for f in range(nf)]
    rsrc = [_padded(ref[f], np.nan)[:, :w] for f in range(nf)] if final else None
    assert src[0].strides[0] % 64 == 0 and src[0].strides[0] > w * 4
    for n in (2, 0, 4, 1, 3):
        dbuf = _padded(np.zeros((h, w), np.float32), -7.0)
        _raw_frame(node2, n, src, rsrc, dbuf[:, :w])
        assert np.array_equal(dbuf[:, :w], want[n]), f"frame {n} differs with padded strides"
        assert (dbuf[:, w:] == -7.0).all(), f"frame {n}: destination padding written"


@pytest.mark.parametrize("stage", [1, 2])
def test_failed_request_leaves_the_handle_usable(stage):
    # a request that fails at its upload (1) or launch (2) returns the
    # "BM3D: ..." error (src/bm3d.cpp:747-764); concurrent requests complete
    # and the next request on the handle succeeds
    _torch()
    from vsfeel import BM3Dv2, Clip, _lib, synth
    from vsfeel.bm3d import Error
    clip = synth.workload(3, frames=8, width=128, height=80)[0]
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=1, ps_num=2, ps_range=4, extractor_exp=2,
              num_streams=3)
    print(f"failure hook stage={stage} clip={clip.shape}", flush=True)
    want = _device_frames(clip, {k: v for k, v in kw.items() if k != "num_streams"})
    node = BM3Dv2(Clip.gray(clip), **kw)
    L = _lib.load()
    assert L.vsfeel_bm3d_fail_next_request(node._handle, 3) == _lib.EINVAL
    assert L.vsfeel_bm3d_fail_next_request(node._handle, stage) == _lib.OK
    got, errors = {}, []

    def worker(frames):
        for n in frames:
            try:
                got.setdefault(n, []).append(node.get_frame(n)[0])
            except Error as e:
                errors.append((n, str(e)))

    rng = np.random.default_rng(stage)
    threads = [threading.Thread(target=worker, args=(list(rng.permutation(8)),)) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert len(errors) == 1, errors
    n_bad, msg = errors[0]
    assert msg.startswith("BM3D: ") and "test hook" in msg, msg
    assert sum(len(v) for v in got.values()) == 4 * 8 - 1
    for n, frames in got.items():
        for f in frames:
            assert np.array_equal(f, want[n]), f"frame {n} differs next to a failed request"
    assert np.array_equal(node.get_frame(n_bad)[0], want[n_bad])


def test_missing_plane_leaves_the_ring_intact():
    # a request whose window lacks a plane it must upload fails before it
    # enqueues any upload ("missing plane of frame f"); the node then serves
    # every frame as before
    import ctypes
    _torch()
    from vsfeel import BM3Dv2, Clip, _lib, synth
    clip = synth.workload(3, frames=6, width=128, height=80)[0]
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=1, ps_num=2, ps_range=4, extractor_exp=2)
    print(f"missing plane clip={clip.shape}", flush=True)
    want = _device_frames(clip, kw)
    node = BM3Dv2(Clip.gray(clip), **kw)
    L = _lib.load()
    planes = [np.ascontiguousarray(clip[f]) for f in range(6)]
    first, last = ctypes.c_int32(), ctypes.c_int32()
    assert L.vsfeel_bm3d_frame_window(node._handle, 2, ctypes.byref(first), ctypes.byref(last)) == _lib.OK
    fs = list(range(first.value, last.value + 1))
    VP = ctypes.c_void_p
    ptrs = (VP * len(fs))(*[planes[f].ctypes.data if f != fs[-1] else None for f in fs])
    strides = (ctypes.c_ssize_t * len(fs))(*[planes[f].strides[0] for f in fs])
    dst = np.zeros((80, 128), np.float32)
    err = _lib.err_buf()
    rc = L.vsfeel_bm3d_get_frame(node._handle, 2, ptrs, strides, None, None, dst.ctypes.data, dst.strides[0],
                                 err, len(err))
    assert rc == _lib.EINVAL and err.value.decode() == f"BM3D: missing plane of frame {fs[-1]}"
    for n in (2, 0, 5, 3):
        assert np.array_equal(node.get_frame(n)[0], want[n]), f"frame {n} differs after the failed request"
