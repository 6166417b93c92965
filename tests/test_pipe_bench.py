"""The frame-request throughput runner (the reference's vspipe benchmark,
benchmark/bench.py:32-54, :108-121, :167-180, :226-235)."""

from __future__ import annotations

import numpy as np
import pytest


def test_defaults_are_the_reference_harness_defaults(capsys):
    from vsfeel import pipe_bench
    seen = {}

    def fake_run(frames, synthetic, threads, width, height, sigma, radius, bm_range, ps_range, block_step,
                 num_streams, graphs):
        seen.update(frames=frames, synthetic=synthetic, width=width, height=height, sigma=sigma,
                    radius=radius, bm_range=bm_range, ps_range=ps_range, block_step=block_step,
                    num_streams=num_streams, graphs=graphs)
        return 50.0

    orig = pipe_bench.run
    pipe_bench.run = fake_run
    try:
        pipe_bench.main(["--synthetic"])
    finally:
        pipe_bench.run = orig
    # BM3Dv2 defaults of the reference harness (benchmark/bench.py:167-180)
    assert seen == dict(frames=1000, synthetic=True, width=1920, height=1080, sigma=0.7, radius=2,
                        bm_range=16, ps_range=7, block_step=4, num_streams=4, graphs=False)
    assert capsys.readouterr().out.strip() == "Output 1000 frames in 20.00 seconds (50.00 fps)"


def test_clip_cycles_a_pool_without_materialising_it():
    from vsfeel import pipe_bench
    clip = pipe_bench.make_clip(1000, synthetic=False, width=64, height=48, pool=4)
    assert (clip.num_frames, clip.height, clip.width) == (1000, 48, 64)
    assert np.array_equal(clip.get_frame(5)[0], clip.get_frame(1)[0])
    assert not np.array_equal(clip.get_frame(1)[0], clip.get_frame(2)[0])
    blank = pipe_bench.make_clip(10, synthetic=True, width=32, height=24)
    assert blank.num_frames == 10 and not blank.get_frame(9)[0].any()


@pytest.mark.gpu
def test_runs_on_the_device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from vsfeel import pipe_bench
    for synthetic in (True, False):
        fps = pipe_bench.run(frames=12, synthetic=synthetic, threads=3, width=160, height=96, radius=1,
                             bm_range=8, ps_range=3)
        assert fps > 0
