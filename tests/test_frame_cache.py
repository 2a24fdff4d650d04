"""Per-output-frame rendering over the device cache of anchor stacks
(vnlbhip_frame_cache_*): the frame-request model of the VapourSynth plugins,
where vnlb.Basic / vnlb.Final make one anchor frame's stack per request and
vnlb.Aggregate sums the stacks around an output frame
(src/vs/plugin.cpp:616-758,966-1045).

Each rendered frame must equal that frame of the whole-clip stage (which
sums the same contributions in one fp64 accumulator set; the stacks are
rounded to fp32 per anchor, hence 1e-5), and of the oracle's stage within
the plain-form bound; rendering every frame in order computes each anchor
stack exactly once, also when the cache is only as large as one output
frame needs.
"""
import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, VNLBError

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim == 3:
        a = a[:, None]
    return torch.from_numpy(a).to("cuda:0")


def _render_all(ctx, stage, params, x, ref, order, capacity=None):
    frames, channels, height, width = x.shape
    cache = ctx.frame_cache(stage, params, frames, channels, height, width, capacity)
    try:
        out = {n: cache.render(n, x, ref).cpu().numpy() for n in order}
        return np.stack([out[n] for n in range(frames)]), cache.stats()
    finally:
        cache.close()


@pytest.mark.parametrize("name", ["gray_pt2_step4", "yuv_pt2_tau200"])
def test_rendered_frames_match_the_stage(gpu, oracle_lib, name):
    noisy, _, bp, fp = workloads.make(name)
    x = _dev(noisy)
    frames = x.shape[0]
    anchors = frames - bp.patch_time + 1
    basic = gpu.stage(BASIC, bp, x)
    for stage, params, ref in ((BASIC, bp, None), (FINAL, fp, basic)):
        print(f"[frame_cache] {name} stage={stage}", flush=True)
        want = gpu.stage(stage, params, x, ref).cpu().numpy()
        got, stats = _render_all(gpu, stage, params, x, ref, range(frames))
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
        assert stats["misses"] == anchors and stats["evictions"] == 0, stats
        # Any request order gives the same frames.
        shuffled, _ = _render_all(gpu, stage, params, x, ref, np.random.default_rng(1).permutation(frames))
        np.testing.assert_array_equal(shuffled, got)
    oracle_basic, _ = oracle_lib.stage(BASIC, bp, noisy)
    got_basic, _ = _render_all(gpu, BASIC, bp, x, None, range(frames))
    assert np.abs(got_basic - oracle_basic.reshape(got_basic.shape)).max() <= 1e-3


def test_long_clip_computes_each_stack_once(gpu):
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    clip, _ = synth.make_clip(width=96, height=64, frames=14, channels=1, sigma8=20, seed=21, motion="subpixel")
    x = _dev(clip)
    frames = x.shape[0]
    need = bp.search_bwd + bp.search_fwd + bp.patch_time
    want = gpu.stage(BASIC, bp, x).cpu().numpy()
    got, stats = _render_all(gpu, BASIC, bp, x, None, range(frames), capacity=need)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
    assert stats["misses"] == frames - bp.patch_time + 1 and stats["evictions"] > 0, stats
    # Backwards over the clip with the same small cache: same frames.
    back, _ = _render_all(gpu, BASIC, bp, x, None, range(frames - 1, -1, -1), capacity=need)
    np.testing.assert_array_equal(back, got)


def test_fast_form_and_errors():
    noisy, _, bp, _ = workloads.make("yuv_pt2_tau200")
    x = _dev(noisy)
    frames, channels, height, width = x.shape
    ctx = vnlbhip.Context(0)
    try:
        ctx.set_form("fast")
        print("[frame_cache] fast", flush=True)
        want = ctx.stage(BASIC, bp, x).cpu().numpy()
        got, _ = _render_all(ctx, BASIC, bp, x, None, range(frames))
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-4)
        need = bp.search_bwd + bp.search_fwd + bp.patch_time
        with pytest.raises(VNLBError):
            ctx.frame_cache(BASIC, bp, frames, channels, height, width, need - 1)
        cache = ctx.frame_cache(BASIC, bp, frames, channels, height, width)
        try:
            with pytest.raises(VNLBError):
                cache.render(frames, x)
        finally:
            cache.close()
    finally:
        ctx.close()


def test_closing_the_context_frees_its_caches():
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    x = _dev(noisy)
    frames, channels, height, width = x.shape
    ctx = vnlbhip.Context(0)
    cache = ctx.frame_cache(BASIC, bp, frames, channels, height, width)
    first = cache.render(1, x).cpu().numpy()
    assert np.isfinite(first).all()
    ctx.close()
    assert not cache._handle
    cache.close()  # already freed with its context: a no-op
