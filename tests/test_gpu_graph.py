"""Graph replay of the device entry point against direct launches.

process() captures its launch sequence into a HIP graph on the first call
with a given set of arguments and replays it afterwards.  Every replay, with
the program's own eager work and synchronisations between replays, must give
bit-identical frames to the directly launched sequence (the accumulation is
integer fixed point, so the results are deterministic), also after the clip
memory is rewritten in place and after the output moves.  Host-buffer
requests (getFrame) with request graphs on launch from per-slot graphs
updated per request, and must equal direct launches in any order and from
several threads.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    "spatial": (lambda s: s.workload(1, frames=3),
                dict(sigma=3.0, block_step=8, bm_range=9, radius=0)),
    "temporal": (lambda s: s.workload(3, frames=5, width=192, height=112)[0],
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7,
                      extractor_exp=2)),
}


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replays_match_direct_launches(name):
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    make, kw = CASES[name]
    clip = np.ascontiguousarray(make(synth), np.float32)
    nf, h, w = clip.shape
    print(f"graph case {name} clip={clip.shape} args={kw}", flush=True)
    x = torch.from_numpy(clip).to("cuda:0")
    direct = Instance(nf, h, w, **kw)
    direct.set_graphs(False)
    want = direct.process(x).cpu().numpy()

    inst = Instance(nf, h, w, **kw)
    inst.set_graphs(True)  # (the default at radius 0 only)
    out = torch.empty_like(x)
    for rep in range(3):
        got = inst.process(x, out=out)
        scratch = (x * 2.0 + 1.0).sum()  # eager work of the program between replays
        torch.cuda.synchronize()
        assert np.isfinite(float(scratch))
        assert np.array_equal(got.cpu().numpy(), want), f"{name}: replay {rep} differs"

    # same memory, new content: the replay reads the clip at run time
    clip2 = np.ascontiguousarray(clip[::-1])
    x.copy_(torch.from_numpy(clip2))
    want2 = direct.process(x).cpu().numpy()
    got2 = inst.process(x, out=out).cpu().numpy()
    assert np.array_equal(got2, want2), f"{name}: replay after rewriting the clip differs"

    # a new output buffer is a new capture
    out2 = torch.empty_like(x)
    got3 = inst.process(x, out=out2).cpu().numpy()
    assert np.array_equal(got3, want2), f"{name}: capture for a new output differs"

    # a sub-range of frames
    got4 = inst.process(x, first=1, count=nf - 2).cpu().numpy()
    want4 = direct.process(x, first=1, count=nf - 2).cpu().numpy()
    assert np.array_equal(got4, want4), f"{name}: sub-range replay differs"


@pytest.mark.parametrize("radius,extractor_exp", [(0, 0), (2, 2)])
def test_host_requests_from_graphs_match_direct_launches(radius, extractor_exp):
    # vsfeel_bm3d_get_frame launches a request's kernels as the graph of its
    # output slot and recompute pattern, updated per request: frames in any
    # order and from 4 threads equal those of direct launches bit for bit
    import threading

    _torch()
    from vsfeel import BM3Dv2, Clip, _lib, synth
    clip = synth.workload(3, frames=9, width=160, height=96)[0]
    kw = dict(sigma=3.0, block_step=4, bm_range=8, radius=radius, ps_num=2, ps_range=4,
              extractor_exp=extractor_exp, num_streams=3)
    print(f"host graphs radius={radius} clip={clip.shape}", flush=True)
    L = _lib.load()
    direct = BM3Dv2(Clip.gray(clip), **kw)
    assert L.vsfeel_bm3d_set_graphs(direct._handle, 0) == _lib.OK
    want = [direct.get_frame(n)[0] for n in range(9)]

    node = BM3Dv2(Clip.gray(clip), **kw)
    assert L.vsfeel_bm3d_set_graphs(node._handle, 4) == _lib.EINVAL
    assert L.vsfeel_bm3d_set_graphs(node._handle, 3) == _lib.OK  # request graphs on (off by default)
    for n in [4, 0, 8, 5, 5, 1, 7]:  # seeks (whole windows), repeats, neighbours
        assert np.array_equal(node.get_frame(n)[0], want[n]), f"graph frame {n} differs"
    got, errors = {}, []

    def worker(frames):
        try:
            for n in frames:
                got.setdefault(n, []).append(node.get_frame(n)[0])
        except Exception as e:  # surfaced below
            errors.append(e)

    rng = np.random.default_rng(9)
    threads = [threading.Thread(target=worker, args=(list(rng.permutation(9)),)) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors, errors
    for n, frames in got.items():
        assert len(frames) == 4
        for f in frames:
            assert np.array_equal(f, want[n]), f"graph frame {n} differs under parallel requests"
    # switching request graphs off on a live node drops the slots' graphs
    assert L.vsfeel_bm3d_set_graphs(node._handle, 1) == _lib.OK
    assert np.array_equal(node.get_frame(3)[0], want[3])


# more frames than lanes, so that every lane's buffers are reused
LANE_CASES = {
    "spatial": (lambda s: s.workload(1, frames=7),
                dict(sigma=3.0, block_step=8, bm_range=9, radius=0)),
    "temporal": (lambda s: s.workload(3, frames=7, width=192, height=112)[0],
                 dict(sigma=3.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=7,
                      extractor_exp=2)),
}


@pytest.mark.parametrize("name", list(LANE_CASES))
def test_lane_count_leaves_frames_unchanged(name, monkeypatch):
    # process_device runs consecutive estimates on up to three lanes (own
    # streams and buffer sets); VSFEEL_BM3D_LANES=1, read when the instance
    # is created, keeps everything on one stream: the frames are identical
    torch = _torch()
    from vsfeel import synth
    from vsfeel.torch_bm3d import Instance

    make, kw = LANE_CASES[name]
    clip = np.ascontiguousarray(make(synth), np.float32)
    nf, h, w = clip.shape
    print(f"lanes case {name} clip={clip.shape}", flush=True)
    x = torch.from_numpy(clip).to("cuda:0")
    got = {}
    for lanes in ("1", "2", None, "0", "9"):  # (0 and 9 are clamped to 1 and 3)
        if lanes is None:
            monkeypatch.delenv("VSFEEL_BM3D_LANES", raising=False)
        else:
            monkeypatch.setenv("VSFEEL_BM3D_LANES", lanes)
        inst = Instance(nf, h, w, **kw)
        for graphs in (False, True):
            inst.set_graphs(graphs)
            got[(lanes, graphs)] = inst.process(x).cpu().numpy()
            # a sub-range: the first estimates of the call on other lanes
            got[(lanes, graphs, "sub")] = inst.process(x, first=2, count=nf - 3).cpu().numpy()
    want = got[("1", False)]
    want_sub = got[("1", False, "sub")]
    for key, frames in got.items():
        ref = want_sub if len(key) == 3 else want
        assert np.array_equal(frames, ref), f"{name}: lanes={key[0]} graphs={key[1]} {key[2:]} differs"
