"""MVTools search-centre guidance: the oracle's restatement of
MVToolsFlowProvider::center_for (src/flow/flow.hpp:340-390), the MVTools
property parsers of vnlbhip.flow (src/flow/flow.hpp:196-318) and, on the
GPU, the dense offset field of vnlbhip_mvtools_offsets and guided stages.

Pins: the reference's known-answer test (tests/cpp/test_core.cpp:40-65:
grids at frames 1 and 2 with the vector (10, -10), anchor (3, 3) of frame 2,
candidate frame 0, 5 x 5 frame, patch 2 -> centre (4, 0)); the nearest block
by brute force over the block centres; hand-worked 1/pel roundings; chains
whose per-frame vectors are distinct, so a grid taken from the wrong frame
or direction changes the result.
"""
import numpy as np
import pytest

import workloads
from vnlbhip import flow
from vnlbhip.params import BASIC, FINAL, VNLBError, parse_stage_parameters


def _shape(bx=1, by=1, ox=0, oy=0, cx=1, cy=1, pel=1):
    return {"block_size_x": bx, "block_size_y": by, "overlap_x": ox, "overlap_y": oy,
            "block_count_x": cx, "block_count_y": cy, "pel": pel}


def _clip(shape, vectors, valid):
    return shape, np.asarray(vectors, np.int32), np.asarray(valid, np.uint8)


def test_center_for_known_answer(oracle_lib):
    vectors = np.zeros((3, 1, 1, 2), np.int32)
    vectors[1:, 0, 0] = (10, -10)
    fw = _clip(_shape(), vectors, [0, 1, 1])
    assert oracle_lib.center_for(fw, None, 3, 3, 2, 0, 2, 5, 5) == (4, 0)
    # SameLocationProvider (tests/cpp/test_core.cpp:34-38): no clips.
    assert oracle_lib.center_for(None, None, 7, 9, 2, 4, 1, 0, 0, grid_frames=0) == (7, 9)
    # The anchor frame itself keeps the anchor.
    assert oracle_lib.center_for(fw, None, 3, 3, 2, 2, 2, 5, 5) == (3, 3)


@pytest.mark.parametrize("block,overlap,count", [(8, 0, 7), (8, 4, 9), (16, 8, 5), (7, 2, 6), (1, 0, 40)])
def test_nearest_block_by_brute_force(oracle_lib, block, overlap, count):
    """One step from frame 1 to frame 0 moves x by the chosen block's index:
    it must be the block whose centre block/2 + i * step is nearest to
    x + patch/2 (ties to the later block), clamped to the grid."""
    step = max(1, block - overlap)
    width = 200
    patch = 5
    vectors = np.zeros((2, 1, count, 2), np.int32)
    vectors[1, 0, :, 0] = np.arange(count)
    fw = _clip(_shape(bx=block, ox=overlap, cx=count), vectors, [0, 1])
    centers = block // 2 + step * np.arange(count)
    for x in range(width - count):
        c = x + patch // 2
        dist = np.abs(c - centers)
        want = int(np.nonzero(dist == dist.min())[0].max())
        got = oracle_lib.center_for(fw, None, x, 0, 1, 0, patch, width, 4)
        assert got == (x + want, 0), (x, want, got)


@pytest.mark.parametrize("pel,cases", [
    (2, [(3, 2), (-3, -2), (1, 1), (-1, -1), (2, 1), (4, 2)]),
    (4, [(2, 1), (-2, -1), (5, 1), (6, 2), (-6, -2), (1, 0), (-1, 0)]),
])
def test_subpixel_vectors_round_half_away_from_zero(oracle_lib, pel, cases):
    for value, pixels in cases:
        vectors = np.zeros((2, 1, 1, 2), np.int32)
        vectors[1, 0, 0] = (value, -value)
        fw = _clip(_shape(pel=pel), vectors, [0, 1])
        assert oracle_lib.center_for(fw, None, 50, 50, 1, 0, 1, 100, 100) == (50 + pixels, 50 - pixels)


def test_chain_directions_frames_and_validity(oracle_lib):
    frames = 6
    fw_v = np.zeros((frames, 1, 1, 2), np.int32)
    bw_v = np.zeros((frames, 1, 1, 2), np.int32)
    fw_v[:, 0, 0, 0] = np.arange(frames) + 1          # fw grid of frame f: dx = f + 1
    bw_v[:, 0, 0, 1] = 10 * (np.arange(frames) + 1)   # bw grid of frame f: dy = 10 (f + 1)
    fw = _clip(_shape(), fw_v, np.ones(frames))
    bw = _clip(_shape(), bw_v, np.ones(frames))
    # Earlier frames: fw grids of frames 3, 2 (not 1) for 3 -> 1.
    assert oracle_lib.center_for(fw, bw, 0, 0, 3, 1, 1, 100, 100) == (4 + 3, 0)
    # Later frames: bw grids of frames 2, 3 (not 4) for 2 -> 4.
    assert oracle_lib.center_for(fw, bw, 0, 0, 2, 4, 1, 100, 100) == (0, 30 + 40)
    # An unusable grid on the way is skipped.
    bw_gap = _clip(_shape(), bw_v, [1, 1, 1, 0, 1, 1])
    assert oracle_lib.center_for(fw, bw_gap, 0, 0, 2, 4, 1, 100, 100) == (0, 30)
    # Without the clip of that direction the centre stays.
    assert oracle_lib.center_for(fw, None, 5, 5, 2, 4, 1, 100, 100) == (5, 5)


def test_dense_field_follows_the_schedule(oracle_lib):
    """vo_mvtools_offsets entry (f, ti) is center_for towards the ti-th frame
    of the anchor's schedule (anchor, later ascending, earlier descending)."""
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    frames, height, width = np.ascontiguousarray(noisy).shape[0], 16, 20
    rng = np.random.default_rng(4)
    shape = _shape(bx=8, by=8, ox=4, oy=4, cx=4, cy=3, pel=2)
    fw = _clip(shape, rng.integers(-9, 10, (frames, 3, 4, 2)), rng.integers(0, 2, frames))
    bw = _clip(shape, rng.integers(-9, 10, (frames, 3, 4, 2)), rng.integers(0, 2, frames))
    field = oracle_lib.mvtools_offsets(bp, fw, bw, width, height, frames)
    temporal = bp.search_bwd + bp.search_fwd + 1
    max_origin = frames - bp.patch_time
    for f in range(frames):
        start, end = f - bp.search_bwd, f + bp.search_fwd
        shift = min(0, start) + max(0, end - max_origin)
        low, high = max(0, start - shift), min(max_origin, end - shift)
        schedule = [f] + list(range(f + 1, high + 1)) + list(range(f - 1, low - 1, -1))
        for ti in range(temporal):
            for y, x in ((0, 0), (5, 7), (height - 1, width - 1), (9, 2)):
                want = (0, 0)
                if ti < len(schedule):
                    cx, cy = oracle_lib.center_for(fw, bw, x, y, f, schedule[ti], bp.patch_size, width, height)
                    want = (cx - x, cy - y)
                assert tuple(field[f, ti, y, x]) == want, (f, ti, x, y)


def _analysis(**kw):
    base = dict(block_size_x=8, block_size_y=8, pel=2, level_count=3, delta_frame=1, backwards=False,
                width=64, height=48, overlap_x=4, overlap_y=4, block_count_x=15, block_count_y=11,
                bits_per_sample=8, chroma_ratio_x=1, chroma_ratio_y=1, padding_x=16, padding_y=16)
    base.update(kw)
    return flow.MVToolsAnalysis(**base)


def test_mvtools_blobs_round_trip():
    a = _analysis(delta_frame=-1, backwards=True)
    assert flow.parse_mvtools_analysis(flow.analysis_bytes(a)) == a
    rng = np.random.default_rng(0)
    finest = rng.integers(-40, 41, (11 * 15, 2))
    blob = flow.vectors_bytes([rng.integers(-5, 5, (6, 2)), rng.integers(-9, 9, (20, 2)), finest])
    ok, v = flow.parse_mvtools_vectors(blob, a)
    assert ok and v.shape == (11, 15, 2)
    np.testing.assert_array_equal(v.reshape(-1, 2), finest)
    assert flow.parse_mvtools_vectors(flow.vectors_bytes([finest], valid=False), a) == (False, None)


@pytest.mark.parametrize("mutate,message", [
    (lambda b: b[:-4], "21 32-bit fields"),
    (lambda b: b[:8] + b"\0\0\0\0" + b[12:], "horizontal block size"),
    (lambda b: b[:16] + b"\0\0\0\0" + b[20:], "invalid MVTools pel"),
    (lambda b: b[:48] + (8).to_bytes(4, "little") + b[52:], "overlap must be smaller"),
])
def test_analysis_errors(mutate, message):
    with pytest.raises(VNLBError, match=message):
        flow.parse_mvtools_analysis(mutate(flow.analysis_bytes(_analysis())))


def test_vector_errors():
    a = _analysis()
    good = flow.vectors_bytes([np.zeros((165, 2))])
    with pytest.raises(VNLBError, match="too small"):
        flow.parse_mvtools_vectors(good[:4], a)
    with pytest.raises(VNLBError, match="size header"):
        flow.parse_mvtools_vectors(good + b"\0", a)
    with pytest.raises(VNLBError, match="does not match analysis"):
        flow.parse_mvtools_vectors(flow.vectors_bytes([np.zeros((160, 2))]), a)


def test_vector_clip_checks_and_scene_changes():
    a = _analysis()
    blobs = [flow.vectors_bytes([np.full((165, 2), f)]) for f in range(4)]
    props = [{}, {"_SceneChangePrev": 1}, {"Scenechange": 0}, {"_SceneChangeNext": 1}]
    clip = flow.VectorClip.from_mvtools(flow.analysis_bytes(a), blobs, props)
    np.testing.assert_array_equal(clip.valid, [1, 0, 1, 1])
    assert clip.vectors[3, 0, 0, 0] == 3
    clip.check("mvfw", 4, 64, 48, False)
    with pytest.raises(VNLBError, match="frame count"):
        clip.check("mvfw", 5, 64, 48, False)
    with pytest.raises(VNLBError, match="dimensions"):
        clip.check("mvfw", 4, 64, 40, False)
    with pytest.raises(VNLBError, match="isb=True"):
        clip.check("mvbw", 4, 64, 48, True)
    back = flow.VectorClip.from_mvtools(flow.analysis_bytes(_analysis(backwards=True)), blobs, props)
    np.testing.assert_array_equal(back.valid, [1, 1, 1, 0])


def test_parameters_accept_vector_clips():
    a = _analysis()
    clip = flow.VectorClip(a, np.zeros((2, 11, 15, 2)), np.ones(2))
    parse_stage_parameters(BASIC, {"sigma": 10, "mvfw": clip, "mvbw": clip})
    with pytest.raises(VNLBError, match="mvfw/mvbw"):
        parse_stage_parameters(FINAL, {"sigma": 10, "mvbw": np.zeros(3)})


def _random_clips(frames, width, height, seed):
    rng = np.random.default_rng(seed)
    clips = []
    for backwards in (False, True):
        a = _analysis(width=width, height=height, block_count_x=(width - 4) // 4, block_count_y=(height - 4) // 4,
                      backwards=backwards, pel=2)
        vectors = rng.integers(-14, 15, (frames, a.block_count_y, a.block_count_x, 2))
        valid = (rng.random(frames) > 0.2).astype(np.uint8)
        clips.append(flow.VectorClip(a, vectors, valid))
    return clips


def _oracle_clip(c):
    return c.analysis.grid_shape(), c.vectors, c.valid


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["gray_pt2_step4", "yuv_pt2_tau200"])
def test_gpu_mvtools_field_and_guided_stages(gpu, oracle_lib, name):
    import torch

    import vnlbhip

    noisy, clean, bp, fp = workloads.make(name)
    x = np.ascontiguousarray(noisy, np.float32)
    if x.ndim == 3:
        x = x[:, None]
    frames, channels, height, width = x.shape
    mvfw, mvbw = _random_clips(frames, width, height, seed=6)
    print(f"[flow] {name} field", flush=True)
    got = gpu.mvtools_offsets(bp, frames, height, width, mvfw, mvbw).cpu().numpy()
    want = oracle_lib.mvtools_offsets(bp, _oracle_clip(mvfw), _oracle_clip(mvbw), width, height, frames)
    np.testing.assert_array_equal(got, want)
    only_bw = gpu.mvtools_offsets(bp, frames, height, width, None, mvbw).cpu().numpy()
    np.testing.assert_array_equal(
        only_bw, oracle_lib.mvtools_offsets(bp, None, _oracle_clip(mvbw), width, height, frames))
    print(f"[flow] {name} guided stages", flush=True)
    temporal = bp.search_bwd + bp.search_fwd + 1
    try:
        oracle_lib.set_search_offsets(want, temporal)
        want_basic, _ = oracle_lib.stage(BASIC, bp, noisy)
        want_final, _ = oracle_lib.stage(FINAL, fp, noisy, want_basic)
    finally:
        oracle_lib.set_search_offsets(None)
    # The plugin-level entry points with the vector clips.
    args = workloads.WORKLOADS[name][1]
    dev = torch.from_numpy(x).to("cuda:0")
    got_basic = vnlbhip.Basic(dev, mvfw=mvfw, mvbw=mvbw, **args)
    ref = torch.from_numpy(np.ascontiguousarray(want_basic, np.float32).reshape(x.shape)).to("cuda:0")
    got_final = vnlbhip.Final(dev, ref, mvfw=mvfw, mvbw=mvbw, **args)
    basic_err = float(np.abs(got_basic.cpu().numpy() - want_basic.reshape(x.shape)).max())
    final_err = float(np.abs(got_final.cpu().numpy() - want_final.reshape(x.shape)).max())
    print(f"[flow] {name} max-abs Basic {basic_err:.2e} Final {final_err:.2e}", flush=True)
    assert basic_err <= 1e-3 and final_err <= 1e-3
    # Guidance changes the result (the field is not a no-op here).
    unguided = vnlbhip.Basic(dev, **args).cpu().numpy()
    assert np.abs(unguided - got_basic.cpu().numpy()).max() > 1e-3
