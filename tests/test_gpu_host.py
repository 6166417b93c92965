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
