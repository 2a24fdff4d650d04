"""Contribution stacks of single anchor frames, in the layout a vnlb.Basic /
vnlb.Final output frame holds them (src/vs/plugin.cpp:264-271,625-641;
packed as the CUDA plugin's pack_kernel, src/cuda/aggregation.cu:508-528).

The layout is pinned by the aggregation rule of vnlb.Aggregate
(src/vs/plugin.cpp:966-1045; aggregate_frame,
src/aggregate/aggregate.cpp:95-163): output t sums the stacks of the anchor
frames a in [max(0, t - fwd - (Pt - 1)), min(n - 1, t + bwd)] at slot
t - a + bwd, and divides the numerators by the weight (the source sample
where the weight is 0).  Applied to the oracle's stacks it must give the
oracle's stage, which aggregates its per-anchor stacks internally; the HIP
stacks must match the oracle's and aggregate to the HIP stage.
"""
import numpy as np
import pytest

import workloads
from vnlbhip.params import BASIC, FINAL

CASES = ["gray_pt2_step4", "yuv_pt2_tau200"]


def _clip4(a):
    a = np.ascontiguousarray(a, np.float32)
    return a[:, None] if a.ndim == 3 else a


def aggregate_stacks(stacks, noisy, params):
    """vnlb.Aggregate over packed stacks [C, slots * 2 * H, W], one per frame."""
    noisy = _clip4(noisy)
    frames, channels, height, width = noisy.shape
    slots = params.search_bwd + params.search_fwd + params.patch_time
    out = np.empty_like(noisy)
    for t in range(frames):
        first = max(0, t - (params.search_fwd + params.patch_time - 1))
        last = min(frames - 1, t + params.search_bwd)
        numerator = np.zeros((channels, height, width))
        weight = np.zeros((height, width))
        for a in range(first, last + 1):
            stack = stacks[a].reshape(channels, slots, 2, height, width)
            numerator += stack[:, t - a + params.search_bwd, 0]
            weight += stack[0, t - a + params.search_bwd, 1]
        w32 = weight.astype(np.float32)
        safe = np.where(w32 > 0, w32, np.float32(1))
        out[t] = np.where(w32 > 0, numerator.astype(np.float32) / safe, noisy[t])
    return out


def _check_layout(stack, params, anchor_frame, frames):
    channels = stack.shape[0]
    slots = params.search_bwd + params.search_fwd + params.patch_time
    s = stack.reshape(channels, slots, 2, -1)
    # The weight plane is repeated for every channel.
    for c in range(1, channels):
        np.testing.assert_array_equal(s[c, :, 1], s[0, :, 1])
    # Slots whose output frame lies outside the clip stay empty.
    for slot in range(slots):
        frame = anchor_frame - params.search_bwd + slot
        if frame < 0 or frame >= frames:
            assert not np.any(s[:, slot]), (anchor_frame, slot)
    if anchor_frame > frames - params.patch_time:
        assert not np.any(s)
    else:
        assert np.any(s[0, :, 1] > 0)


@pytest.mark.parametrize("name", CASES)
def test_oracle_stacks_aggregate_to_the_stage(oracle_lib, name):
    noisy, _, bp, fp = workloads.make(name)
    frames = _clip4(noisy).shape[0]
    basic, _ = oracle_lib.stage(BASIC, bp, noisy)
    for stage, params, ref in ((BASIC, bp, None), (FINAL, fp, basic)):
        want, _ = oracle_lib.stage(stage, params, noisy, ref)
        stacks = [oracle_lib.anchor_stack(stage, params, noisy, ref, a) for a in range(frames)]
        for a, stack in enumerate(stacks):
            _check_layout(stack, params, a, frames)
        got = aggregate_stacks(stacks, noisy, params)
        np.testing.assert_allclose(got, _clip4(want), rtol=0, atol=1e-5)


def _dev(a):
    import torch

    return torch.from_numpy(_clip4(a)).to("cuda:0")


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_stacks_match_the_oracle_and_aggregate_to_the_stage(gpu, oracle_lib, name):
    noisy, _, bp, fp = workloads.make(name)
    frames = _clip4(noisy).shape[0]
    x = _dev(noisy)
    basic = gpu.stage(BASIC, bp, x)
    basic_np = basic.cpu().numpy()
    for stage, params, ref, ref_np in ((BASIC, bp, None, None), (FINAL, fp, basic, basic_np)):
        print(f"[stack] {name} stage={stage}", flush=True)
        out = gpu.stage(stage, params, x, ref).cpu().numpy()
        stacks = []
        for a in range(frames):
            got = gpu.anchor_stack(stage, params, x, ref, a).cpu().numpy()
            want = oracle_lib.anchor_stack(stage, params, noisy, ref_np, a)
            _check_layout(got, params, a, frames)
            scale = max(1.0, float(np.abs(want).max()))
            np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5 * scale, err_msg=f"anchor {a}")
            stacks.append(got)
        np.testing.assert_allclose(aggregate_stacks(stacks, noisy, params), out, rtol=0, atol=1e-5)


@pytest.mark.gpu
def test_fast_form_stacks_aggregate_to_the_stage():
    import vnlbhip

    noisy, _, bp, fp = workloads.make("yuv_pt2_tau200")
    frames = _clip4(noisy).shape[0]
    x = _dev(noisy)
    fast = vnlbhip.Context(0)
    try:
        fast.set_form("fast")
        basic = fast.stage(BASIC, bp, x)
        for stage, params, ref in ((BASIC, bp, None), (FINAL, fp, basic)):
            print(f"[stack] fast stage={stage}", flush=True)
            out = fast.stage(stage, params, x, ref).cpu().numpy()
            stacks = [fast.anchor_stack(stage, params, x, ref, a).cpu().numpy() for a in range(frames)]
            for a, stack in enumerate(stacks):
                _check_layout(stack, params, a, frames)
            # fp32 accumulators: the order of the adds differs between runs.
            np.testing.assert_allclose(aggregate_stacks(stacks, noisy, params), out, rtol=0, atol=1e-4)
        with pytest.raises(vnlbhip.VNLBError):
            fast.anchor_stack(BASIC, bp, x, None, frames)
    finally:
        fast.close()
