"""The boundary as the reference's frame requests use it, on the GPU.

* Windows of a longer clip (vnlbhip_anchor_stack_window,
  vnlbhip_frame_cache_render_window): render_stage_frame gives
  process_*_anchor only the anchor's input frames in a geometry that keeps
  the absolute anchor index and the clip length (src/vs/plugin.cpp:576-700,
  src/core/video.hpp:19-47).  Stacks over windows must equal the whole-clip
  stacks (same groups; fp64 accumulators, so up to the last fp32 rounding)
  and the oracle's window stacks; frames rendered from windows of a 40-frame
  clip must equal the oracle's whole-clip stage.
* vnlbhip_reserve: after it, stage calls of that shape allocate nothing.
* Contexts share nothing: two contexts driven from two host threads give the
  results of one.
* Settings that change a stage (eigen re-solve gap, search offsets) are part
  of the keys of captured graphs and cached stacks.
"""
import threading

import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, VNLBError, parse_stage_parameters

pytestmark = pytest.mark.gpu

COMMON = dict(sigma=20, block_size=8, block_step=4, group_size=8, bm_range=5, patch_time=2, radius=2,
              rank=4, tau=0)


def _dev(a):
    a = np.ascontiguousarray(a, np.float32)
    return torch.from_numpy(a[:, None] if a.ndim == 3 else a).to("cuda:0")


def _clip(frames, width=40, height=32, seed=3):
    noisy, _ = synth.make_clip(width, height, frames, 1, sigma8=20, seed=seed, motion="subpixel")
    noisy = np.ascontiguousarray(noisy, np.float32)
    return noisy[:, None] if noisy.ndim == 3 else noisy


def input_range(params, anchor, total):
    max_origin = total - params.patch_time
    start, end = anchor - params.search_bwd, anchor + params.search_fwd
    shift = min(0, start) + max(0, end - max_origin)
    return max(0, start - shift), min(max_origin, end - shift) + params.patch_time - 1


def test_window_stacks_match_whole_clip_and_oracle(gpu, oracle_lib):
    clip = _clip(12)
    total = clip.shape[0]
    p = parse_stage_parameters(BASIC, COMMON)
    x = _dev(clip)
    for anchor in range(total - p.patch_time + 1):
        print(f"[window] stack anchor {anchor}", flush=True)
        lo, hi = input_range(p, anchor, total)
        whole = gpu.anchor_stack(BASIC, p, x, None, anchor).cpu().numpy()
        got = gpu.anchor_stack_window(BASIC, p, x[lo:hi + 1].contiguous(), None, lo, total, anchor)
        got = got.cpu().numpy()
        np.testing.assert_allclose(got, whole, rtol=1e-6, atol=1e-7, err_msg=f"anchor {anchor}")
        want = oracle_lib.anchor_stack_window(BASIC, p, clip[lo:hi + 1], None, lo, total, anchor)
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6, err_msg=f"anchor {anchor} vs oracle")
    with pytest.raises(VNLBError, match="window"):
        lo, hi = input_range(p, 5, total)
        gpu.anchor_stack_window(BASIC, p, x[lo + 1:hi + 1].contiguous(), None, lo + 1, total, 5)


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_windowed_renders_of_a_40_frame_clip(oracle_lib, form):
    """Each output frame rendered from the smallest window its missing stacks
    need, in request order, against the whole-clip stage."""
    clip = _clip(40, width=48, height=40, seed=9)
    total = clip.shape[0]
    bp = parse_stage_parameters(BASIC, COMMON)
    fp = parse_stage_parameters(FINAL, COMMON)
    x = _dev(clip)
    ctx = vnlbhip.Context(0)
    try:
        ctx.set_form(form)
        basic_whole = ctx.stage(BASIC, bp, x)
        final_whole = ctx.stage(FINAL, fp, x, basic_whole)
        torch.cuda.synchronize()
        for stage, params, ref, whole in ((BASIC, bp, None, basic_whole), (FINAL, fp, basic_whole, final_whole)):
            cache = ctx.frame_cache(stage, params, total, 1, clip.shape[2], clip.shape[3])
            try:
                frames = []
                for n in range(total):
                    anchors = range(max(0, n - params.search_fwd - params.patch_time + 1),
                                    min(total - params.patch_time, n + params.search_bwd) + 1)
                    lo = min([n] + [input_range(params, a, total)[0] for a in anchors])
                    hi = max([n] + [input_range(params, a, total)[1] for a in anchors])
                    if n % 13 == 0:
                        print(f"[window] {form} stage {stage} frame {n} window [{lo}, {hi}]", flush=True)
                    win_ref = None if ref is None else ref[lo:hi + 1].contiguous()
                    frames.append(cache.render_window(n, x[lo:hi + 1].contiguous(), win_ref, lo).cpu().numpy())
                stats = cache.stats()
            finally:
                cache.close()
            got = np.stack(frames)
            if form == "plain":
                np.testing.assert_allclose(got, whole.cpu().numpy(), rtol=0, atol=1e-5)
            else:
                # The fast form centres its bf16 operands on per-channel means of
                # the frames a call receives, so a window rounds differently from
                # the whole clip and near-tie decisions may flip: the bar is the
                # fast form's own (PSNR against the whole-clip result).
                want = whole.cpu().numpy()
                mse = float(np.mean((got.astype(np.float64) - want) ** 2))
                psnr = 10 * np.log10(1.0 / max(mse, 1e-20))
                assert psnr >= 60.0, psnr
                assert np.abs(got - want).max() <= 2e-2
            assert stats["misses"] == total - params.patch_time + 1, stats
        if form == "plain":
            want, _ = oracle_lib.stage(BASIC, bp, clip, flags=oracle_lib.ACCUM_DOUBLE)
            assert np.abs(basic_whole.cpu().numpy() - want.reshape(clip.shape)).max() <= 1e-3
    finally:
        ctx.close()


def test_reserve_then_no_allocation():
    noisy, _, bp, fp = workloads.make("yuv_pt2_tau200")
    x = _dev(noisy)
    frames, channels, height, width = x.shape
    for form in ("plain", "fast"):
        ctx = vnlbhip.Context(0)
        try:
            ctx.set_form(form)
            ctx.reserve(BASIC, bp, frames, channels, height, width)
            ctx.reserve(FINAL, fp, frames, channels, height, width)
            basic = ctx.stage(BASIC, bp, x)
            assert ctx.stats()["reallocations"] == 0, form
            ctx.stage(FINAL, fp, x, basic)
            assert ctx.stats()["reallocations"] == 0, form
            fresh = vnlbhip.Context(0)
            try:
                fresh.set_form(form)
                fresh.stage(BASIC, bp, x)
                assert fresh.stats()["reallocations"] > 0
            finally:
                fresh.close()
        finally:
            ctx.close()


def test_two_contexts_on_two_threads_match_one():
    noisy, _, bp, fp = workloads.make("yuv_pt2_tau200")
    x = _dev(noisy)
    want = {}
    one = vnlbhip.Context(0)
    try:
        for form in ("plain", "fast"):
            one.set_form(form)
            b = one.stage(BASIC, bp, x)
            want[form] = (b.cpu().numpy(), one.stage(FINAL, fp, x, b).cpu().numpy())
    finally:
        one.close()
    results = {}
    errors = []

    def worker(form):
        try:
            ctx = vnlbhip.Context(0)
            stream = torch.cuda.Stream()
            try:
                ctx.set_form(form)
                with torch.cuda.stream(stream):
                    for _ in range(3):
                        b = ctx.stage(BASIC, bp, x)
                        f = ctx.stage(FINAL, fp, x, b)
                    stream.synchronize()
                    results[form] = (b.cpu().numpy(), f.cpu().numpy())
            finally:
                ctx.close()
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    print("[window] two threads", flush=True)
    threads = [threading.Thread(target=worker, args=(form,)) for form in ("plain", "fast")]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for form in ("plain", "fast"):
        tol = 1e-6 if form == "plain" else 1e-4
        np.testing.assert_allclose(results[form][0], want[form][0], rtol=0, atol=tol)
        np.testing.assert_allclose(results[form][1], want[form][1], rtol=0, atol=tol)


def test_graph_replay_follows_a_new_refine_gap():
    """Capture in the fast form, change the fp64 re-solve gap: the next call
    must match an eager run with the new gap (the gap is a kernel argument
    of the captured sequence)."""
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    bp = parse_stage_parameters(BASIC, dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=9,
                                            patch_time=2, radius=2, rank=4, tau=0))
    x = _dev(noisy)
    graph = vnlbhip.Context(0)
    eager = vnlbhip.Context(0)
    try:
        for c in (graph, eager):
            c.set_form("fast")
        graph.set_graph(True)
        out = torch.empty_like(x)
        for gap in (1e-3, 1e-3, 0.5, 0.5, 0.0):
            print(f"[window] graph gap {gap}", flush=True)
            graph.set_eigen_refine(gap)
            eager.set_eigen_refine(gap)
            graph.stage(BASIC, bp, x, out=out)
            got_refined = graph.stats()["refined"]
            want = eager.stage(BASIC, bp, x)
            torch.cuda.synchronize()
            assert got_refined == eager.stats()["refined"], gap
            np.testing.assert_allclose(out.cpu().numpy(), want.cpu().numpy(), rtol=0, atol=1e-4)
        assert eager.stats()["refined"] == 0
    finally:
        graph.close()
        eager.close()


def test_cache_restarts_when_search_offsets_change(gpu):
    clip = _clip(8)
    frames, channels, height, width = clip.shape
    p = parse_stage_parameters(BASIC, COMMON)
    x = _dev(clip)
    temporal = p.search_bwd + p.search_fwd + 1
    field = torch.zeros((frames, temporal, height, width, 2), dtype=torch.int16, device="cuda:0")
    field[..., 0] = 3
    cache = gpu.frame_cache(BASIC, p, frames, channels, height, width, capacity=frames)
    try:
        plain = [cache.render(n, x).cpu().numpy() for n in range(frames)]
        misses = cache.stats()["misses"]
        gpu.set_search_offsets(field)
        try:
            guided = [cache.render(n, x).cpu().numpy() for n in range(frames)]
            assert cache.stats()["misses"] == 2 * misses
            want = gpu.stage(BASIC, p, x).cpu().numpy()
            np.testing.assert_allclose(np.stack(guided), want, rtol=0, atol=1e-5)
            # A field made for another clip shape is refused before any read.
            small = field[:, :, :-1].contiguous()
            gpu.set_search_offsets(small)
            with pytest.raises(VNLBError):
                cache.render(0, x)
        finally:
            gpu.set_search_offsets(None)
        again = [cache.render(n, x).cpu().numpy() for n in range(frames)]
        np.testing.assert_array_equal(np.stack(again), np.stack(plain))
    finally:
        cache.close()
