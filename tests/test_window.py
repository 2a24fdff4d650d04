"""Stages over a window of a longer clip, as the reference's frame requests run
them: render_stage_frame hands process_*_anchor only the frames of
input_frame_range_for_anchor (src/vs/plugin.cpp:576-700,
src/core/core.cpp:2026-2034) in a VideoGeometry whose first_frame /
total_frames keep the absolute anchor index and the clip length
(src/core/video.hpp:19-47).  Grid phases (core.cpp:934-950) and temporal
clamps (core.cpp:873-896) read those, so a window must reproduce the
whole-clip stack of the same anchor bit for bit.

Oracle only (no GPU): the window form of the oracle against its whole-clip
form, which the pariasm fixtures and the closed forms of test_oracle.py pin.
"""
import numpy as np
import pytest

from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

COMMON = dict(sigma=20, block_size=8, block_step=4, group_size=8, bm_range=5, patch_time=2, radius=2,
              rank=4, tau=0)


def input_range(params, anchor, total):
    """input_frame_range_for_anchor with compute_temporal_range, restated
    from the reference's integer rule for the window bounds."""
    max_origin = total - params.patch_time
    start, end = anchor - params.search_bwd, anchor + params.search_fwd
    shift = min(0, start) + max(0, end - max_origin)
    low, high = max(0, start - shift), min(max_origin, end - shift)
    return low, high + params.patch_time - 1


def _clip(frames=11, width=40, height=32, seed=3):
    noisy, _ = synth.make_clip(width, height, frames, 1, sigma8=20, seed=seed, motion="subpixel")
    noisy = np.ascontiguousarray(noisy, np.float32)
    return noisy[:, None] if noisy.ndim == 3 else noisy


def test_input_range_examples():
    # Hand-worked: radius 2, Pt 2, 11 frames (last origin 9).
    p = parse_stage_parameters(BASIC, COMMON)
    assert input_range(p, 0, 11) == (0, 5)    # shifted forward at the start
    assert input_range(p, 5, 11) == (3, 8)
    assert input_range(p, 9, 11) == (5, 10)   # shifted back at the end


@pytest.mark.parametrize("stage", [BASIC, FINAL])
def test_window_stack_equals_whole_clip_stack(oracle_lib, stage):
    clip = _clip()
    total = clip.shape[0]
    p = parse_stage_parameters(stage, COMMON)
    ref = None
    if stage == FINAL:
        ref, _ = oracle_lib.stage(BASIC, parse_stage_parameters(BASIC, COMMON), clip)
        ref = np.ascontiguousarray(ref, np.float32).reshape(clip.shape)
    for anchor in range(total):
        whole = oracle_lib.anchor_stack(stage, p, clip, ref, anchor)
        if anchor > total - p.patch_time:
            assert not whole.any()
            lo, hi = total - 1, total - 1
        else:
            lo, hi = input_range(p, anchor, total)
        window = clip[lo:hi + 1]
        got = oracle_lib.anchor_stack_window(stage, p, window, None if ref is None else ref[lo:hi + 1],
                                             lo, total, anchor)
        np.testing.assert_array_equal(got, whole, err_msg=f"anchor {anchor}")
        # A wider window changes nothing either.
        lo2, hi2 = max(0, lo - 1), min(total - 1, hi + 2)
        got2 = oracle_lib.anchor_stack_window(stage, p, clip[lo2:hi2 + 1],
                                              None if ref is None else ref[lo2:hi2 + 1], lo2, total, anchor)
        np.testing.assert_array_equal(got2, whole, err_msg=f"anchor {anchor} (wide window)")


def test_grid_phase_follows_the_absolute_frame(oracle_lib):
    """The same window content at another absolute position (a longer clip)
    moves the staggered grid: phase_y is the absolute anchor frame, 0 only on
    the clip's last patch origin.  Shifting the clip by `step` frames keeps
    the phase; shifting it by one frame changes the stack."""
    clip = _clip(frames=6)
    p = parse_stage_parameters(BASIC, COMMON)
    base = oracle_lib.anchor_stack_window(BASIC, p, clip, None, 0, 20, 2)
    same_phase = oracle_lib.anchor_stack_window(BASIC, p, clip, None, 4, 24, 6)
    other_phase = oracle_lib.anchor_stack_window(BASIC, p, clip, None, 1, 21, 3)
    np.testing.assert_array_equal(same_phase, base)
    assert not np.array_equal(other_phase, base)


def test_window_must_hold_the_input_frames(oracle_lib):
    clip = _clip()
    p = parse_stage_parameters(BASIC, COMMON)
    lo, hi = input_range(p, 5, clip.shape[0])
    with pytest.raises(ValueError, match="window"):
        oracle_lib.anchor_stack_window(BASIC, p, clip[lo + 1:hi + 1], None, lo + 1, clip.shape[0], 5)
    with pytest.raises(ValueError, match="window"):
        oracle_lib.anchor_stack_window(BASIC, p, clip[lo:hi], None, lo, clip.shape[0], 5)
