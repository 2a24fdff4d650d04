"""MFMA ("fast") form against the plain form, through the C ABI.

The fast form takes candidate distances from bf16x3 products
(E(r) + E(c) - 2 <r, c>) and filters dual-PCA groups through their
cross-Gram (bf16x3 on the matrix cores) instead of explicit basis vectors.
Distances agree with the exact fp32 sums to a small relative error, the
kept candidates can differ where distances are within that error of a
threshold, and the outputs agree in quality: the bound is
|PSNR(fast) - PSNR(plain)| <= 0.05 dB against the clean clip, checked both
on the benchmark's configuration and on one that denoises by several dB.
Shapes the fast kernels do not take (patch sizes other than 8) run the
plain kernels and must give bit-identical results.
"""
import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

pytestmark = pytest.mark.gpu

PSNR_TOL = 0.05
DIST_RTOL = 2e-4  # relative to the distance plus the patch energy scale


@pytest.fixture(scope="module")
def contexts():
    plain = vnlbhip.Context(0)
    fast = vnlbhip.Context(0)
    fast.set_form("fast")
    yield plain, fast
    plain.close()
    fast.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


ARGS = dict(sigma=15, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2,
            radius=2, rank=16, tau=200, cap_factor=4.0, model_cap_factor=1.0, chroma=1)


# 93 x 69: the energy kernels' strips of 8 origins end inside the frame.
@pytest.mark.parametrize("width,height", [(96, 72), (93, 69)])
@pytest.mark.parametrize("stage", [BASIC, FINAL])
def test_fast_distances_track_exact(contexts, stage, width, height):
    plain, fast = contexts
    noisy, _ = synth.make_clip(width=width, height=height, frames=6, channels=3, sigma8=15, seed=11,
                               motion="subpixel")
    params = parse_stage_parameters(stage, ARGS)
    ref = None
    if stage == FINAL:
        ref = plain.stage(BASIC, parse_stage_parameters(BASIC, ARGS), _dev(noisy))
    print(f"[fast] distances stage={stage} {width}x{height}", flush=True)
    want = plain.debug_frame(stage, params, _dev(noisy), ref, 2)
    got = fast.debug_frame(stage, params, _dev(noisy), ref, 2)
    np.testing.assert_array_equal(got["anchor_x"], want["anchor_x"])
    same_sets = 0
    checked = 0
    for a, n in enumerate(want["counts"]):
        m = got["counts"][a]
        assert m >= 1
        # The anchor itself leads at distance 0.
        assert got["distances"][a, 0] == 0.0
        exact = dict(zip(want["positions"][a, :n].tolist(), want["distances"][a, :n].tolist()))
        scale = float(np.max(want["distances"][a, :n])) + 1.0
        for pos, dist in zip(got["positions"][a, :m].tolist(), got["distances"][a, :m].tolist()):
            if pos in exact:
                assert abs(dist - exact[pos]) <= DIST_RTOL * scale, (a, pos, dist, exact[pos])
                checked += 1
        same_sets += set(exact) == set(got["positions"][a, :m].tolist())
    assert checked > 0
    assert same_sets >= 0.9 * len(want["counts"]), (same_sets, len(want["counts"]))


WIDE = dict(ARGS, bm_range=12, radius=3)  # 25 x 25 windows, 7 candidate frames: two-row anchor tiles


@pytest.mark.parametrize("stage,args", [(BASIC, ARGS), (FINAL, ARGS), (BASIC, WIDE), (FINAL, WIDE)])
def test_fused_selection_matches_staged(stage, args):
    """The k-nearest selection inside the MFMA matcher keeps exactly the
    lists that the plain form's selection (select.h) makes from the same
    staged distances: counts, positions and distances are equal on every
    anchor, and the stage outputs agree up to the order of the fp32
    accumulator adds (a few float ulps of each sum)."""
    fused = vnlbhip.Context(0)
    staged = vnlbhip.Context(0)
    try:
        fused.set_form("fast")
        staged.set_form("fast")
        staged.set_fused_selection(False)
        noisy, _ = synth.make_clip(width=128, height=96, frames=8, channels=3, sigma8=15, seed=5,
                                   motion="subpixel")
        params = parse_stage_parameters(stage, args)
        ref = None
        if stage == FINAL:
            ref = fused.stage(BASIC, parse_stage_parameters(BASIC, args), _dev(noisy))
        for frame in (0, 2, 4):
            print(f"[fast] fused vs staged selection stage={stage} frame={frame}", flush=True)
            a = fused.debug_frame(stage, params, _dev(noisy), ref, frame)
            b = staged.debug_frame(stage, params, _dev(noisy), ref, frame)
            np.testing.assert_array_equal(a["counts"], b["counts"])
            assert np.all(a["counts"] >= 1)
            for i, n in enumerate(a["counts"]):
                np.testing.assert_array_equal(a["positions"][i, :n], b["positions"][i, :n], err_msg=str(i))
                np.testing.assert_array_equal(a["distances"][i, :n], b["distances"][i, :n], err_msg=str(i))
        out_a = fused.stage(stage, params, _dev(noisy), ref).cpu().numpy()
        out_b = staged.stage(stage, params, _dev(noisy), ref).cpu().numpy()
        np.testing.assert_allclose(out_a, out_b, rtol=0, atol=1e-4)
        with pytest.raises(vnlbhip.VNLBError):
            staged.set_fused_selection(2)
    finally:
        fused.close()
        staged.close()


@pytest.mark.parametrize("channels,width,height", [(3, 128, 96), (1, 112, 80), (3, 40, 36)])
def test_fast_stage_quality_matches_plain(contexts, channels, width, height):
    plain, fast = contexts
    noisy, clean = synth.make_clip(width=width, height=height, frames=6, channels=channels,
                                   sigma8=15, seed=5, motion="subpixel")
    bp = parse_stage_parameters(BASIC, ARGS)
    fp = parse_stage_parameters(FINAL, ARGS)
    print(f"[fast] stages {channels}x{width}x{height}", flush=True)
    x = _dev(noisy)
    b_plain = plain.stage(BASIC, bp, x)
    f_plain = plain.stage(FINAL, fp, x, b_plain)
    b_fast = fast.stage(BASIC, bp, x)
    f_fast = fast.stage(FINAL, fp, x, b_fast)
    for got, want in ((b_fast, b_plain), (f_fast, f_plain)):
        g = got.cpu().numpy()
        w = want.cpu().numpy()
        assert np.isfinite(g).all()
        assert abs(synth.psnr(g, clean) - synth.psnr(w, clean)) <= PSNR_TOL


def test_fast_form_falls_back_to_plain_matcher(contexts):
    plain, fast = contexts
    noisy, _ = synth.make_clip(width=64, height=48, frames=4, channels=1, sigma8=20, seed=3,
                               motion="integer")
    args = dict(sigma=20, block_size=5, block_step=3, group_size=8, bm_range=7, patch_time=1,
                radius=1, rank=6)
    bp = parse_stage_parameters(BASIC, args)
    x = _dev(noisy)
    a = plain.stage(BASIC, bp, x).cpu().numpy()
    b = fast.stage(BASIC, bp, x).cpu().numpy()
    np.testing.assert_array_equal(a, b)


def test_form_is_validated(contexts):
    _, fast = contexts
    with pytest.raises(vnlbhip.VNLBError):
        fast.set_form("fp8")
    assert fast.form == "fast"


DENOISE = dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2,
               radius=2, rank=4, chroma=1)


@pytest.mark.parametrize("channels", [3, 1])
def test_fast_form_on_a_denoising_config(contexts, channels):
    """rank 4 shrinks the dual-PCA model enough to denoise by several dB:
    the fast matcher and the fast (dual-algebra) filter must keep the plain
    form's quality."""
    plain, fast = contexts
    noisy, clean = synth.make_clip(width=128, height=96, frames=6, channels=channels, sigma8=20,
                                   seed=9, motion="subpixel")
    bp = parse_stage_parameters(BASIC, DENOISE)
    fp = parse_stage_parameters(FINAL, DENOISE)
    print(f"[fast] denoising config channels={channels}", flush=True)
    x = _dev(noisy)
    f_plain = plain.stage(FINAL, fp, x, plain.stage(BASIC, bp, x)).cpu().numpy()
    f_fast = fast.stage(FINAL, fp, x, fast.stage(BASIC, bp, x)).cpu().numpy()
    p_noisy = synth.psnr(noisy, clean)
    p_plain = synth.psnr(f_plain, clean)
    p_fast = synth.psnr(f_fast, clean)
    assert p_plain > p_noisy + 3.0, (p_noisy, p_plain)
    assert abs(p_fast - p_plain) <= PSNR_TOL, (p_fast, p_plain)
    assert float(np.abs(f_fast - f_plain).mean()) < 2e-3


@pytest.mark.parametrize("stage", [BASIC, FINAL])
def test_parallel_jacobi_matches_cyclic(contexts, stage):
    """The FAST form solves the 16 x 16 group Grams with the parallel
    (r, r ^ k) ordering; on groups whose retained lists agree with the plain
    form (same Gram), eigenvalues agree with the cyclic solver's to 1e-4 of
    the group's largest eigenvalue."""
    plain, fast = contexts
    noisy, _ = synth.make_clip(width=96, height=72, frames=6, channels=3, sigma8=15, seed=21,
                               motion="subpixel")
    params = parse_stage_parameters(stage, ARGS)
    ref = None
    if stage == FINAL:
        ref = plain.stage(BASIC, parse_stage_parameters(BASIC, ARGS), _dev(noisy))
    print(f"[fast] jacobi stage={stage}", flush=True)
    want = plain.debug_frame(stage, params, _dev(noisy), ref, 2)
    got = fast.debug_frame(stage, params, _dev(noisy), ref, 2)
    compared = 0
    for a, n in enumerate(want["counts"]):
        if not (want["active"][a] and got["active"][a]) or (want["group_flags"][a] & 4) == 0:
            continue
        if got["counts"][a] != n or not np.array_equal(got["positions"][a, :n], want["positions"][a, :n]):
            continue
        w = want["eigenvalues"][a]
        g = got["eigenvalues"][a]
        assert np.isfinite(g).all()
        scale = float(np.max(np.abs(w))) + 1e-12
        assert np.max(np.abs(g - w)) <= 1e-4 * scale, (a, g, w)
        compared += 1
    assert compared >= 20, compared


def test_fast_form_nan_region(contexts):
    """NaN regions: the fast form forms the same groups as the plain form
    (anchors whose candidates are all NaN have none) and yields NaN exactly
    where the plain form does."""
    plain, fast = contexts
    rng = np.random.default_rng(17)
    noisy = rng.random((4, 1, 48, 64), dtype=np.float32)
    noisy[:, :, :, :24] = np.nan
    args = dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=5, patch_time=2,
                radius=1, rank=4)
    bp = parse_stage_parameters(BASIC, args)
    x = _dev(noisy)
    print("[fast] nan region", flush=True)
    want = plain.stage(BASIC, bp, x).cpu().numpy()
    got = fast.stage(BASIC, bp, x).cpu().numpy()
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    assert plain.stats()["groups"] > 0


def test_eigen_refinement_in_fp64(gpu, oracle_lib):
    """Rank 4 of 16 model rows puts a cutoff inside every spectrum.  With the
    gap threshold at 1e30 every modelled group is re-solved by the fp64
    Jacobi: the refined count equals the modelled groups, eigenvalues match
    the oracle's float64 model to 1e-5 of the largest, and the output stays
    within 0.05 dB of the fp32 solve's; threshold 0 refines none."""
    noisy, clean = synth.make_clip(width=96, height=80, frames=5, channels=3, sigma8=20, seed=41,
                                   motion="subpixel")
    bp = parse_stage_parameters(BASIC, DENOISE)
    ctx = vnlbhip.Context(0)
    ctx.set_form("fast")
    try:
        ctx.set_eigen_refine(0.0)
        base = ctx.stage(BASIC, bp, _dev(noisy)).cpu().numpy()
        assert ctx.stats()["refined"] == 0
        ctx.set_eigen_refine(1e30)
        print("[fast] eigen refinement", flush=True)
        refined = ctx.stage(BASIC, bp, _dev(noisy)).cpu().numpy()
        stats = ctx.stats()
        got = ctx.debug_frame(BASIC, bp, _dev(noisy), None, 2)
        with pytest.raises(vnlbhip.VNLBError):
            ctx.set_eigen_refine(-1.0)
    finally:
        ctx.close()
    assert stats["refined"] == stats["groups"], stats
    want = oracle_lib.frame_debug(BASIC, bp, noisy, None, 2,
                                  flags=oracle_lib.ACCUM_DOUBLE | oracle_lib.MODEL_DOUBLE)
    compared = 0
    for a, n in enumerate(want["counts"]):
        if not (want["active"][a] and got["active"][a]) or (want["group_flags"][a] & 4) == 0:
            continue
        if got["counts"][a] != n or not np.array_equal(got["positions"][a, :n], want["positions"][a, :n]):
            continue
        w = want["eigenvalues"][a].astype(np.float64)
        g = got["eigenvalues"][a].astype(np.float64)
        assert np.max(np.abs(g - w)) <= 1e-5 * np.max(np.abs(w)), (a, g, w)
        compared += 1
    assert compared >= 20, compared
    assert abs(synth.psnr(refined, clean) - synth.psnr(base, clean)) <= PSNR_TOL


def test_autotune_picks_a_part_count_and_keeps_results():
    """The autotuner times the stage at 1..4 pipeline parts on the first call
    and keeps the fastest; the part count changes only the launch sequence,
    so outputs match the default two-part run (plain form: identical up to
    the order of the fp64 accumulator adds)."""
    noisy, _ = synth.make_clip(width=96, height=80, frames=8, channels=3, sigma8=15, seed=21,
                               motion="subpixel")
    params = parse_stage_parameters(BASIC, ARGS)
    x = _dev(noisy)
    tuned = vnlbhip.Context(0)
    default = vnlbhip.Context(0)
    try:
        tuned.set_autotune(True)
        print("[fast] autotune plain", flush=True)
        a = tuned.stage(BASIC, params, x).cpu().numpy()
        parts = tuned.stats()["parts"]
        assert 1 <= parts <= 4
        b = tuned.stage(BASIC, params, x).cpu().numpy()
        assert tuned.stats()["parts"] == parts
        c = default.stage(BASIC, params, x).cpu().numpy()
        assert default.stats()["parts"] == 2
        np.testing.assert_allclose(a, c, rtol=0, atol=1e-6)
        np.testing.assert_allclose(b, c, rtol=0, atol=1e-6)
        with pytest.raises(vnlbhip.VNLBError):
            tuned.set_autotune(2)
    finally:
        tuned.close()
        default.close()


@pytest.mark.parametrize("name", ["gray_pt2_step4", "gray_bm12_radius2"])
def test_tile_shapes_give_the_same_lists(name):
    """The MFMA matcher's anchor tiles only group anchors for the matrix
    products: every shape gives the same distances (the same products in the
    same order per anchor pair) and the same lists; the autotuner keeps one
    of them."""
    noisy, _, bp, _ = workloads.make(name)
    x = _dev(noisy)
    x = x[:, None].contiguous() if x.dim() == 3 else x
    ctx = vnlbhip.Context(0)
    try:
        ctx.set_form("fast")
        want = ctx.match_lists(BASIC, bp, x, None, 2)
        used = set()
        for rows, cols in ((4, 4), (3, 4), (2, 4), (1, 4), (4, 3), (4, 2), (2, 2), (4, 1)):
            print(f"[fast] {name}: tile {rows}x{cols}", flush=True)
            ctx.set_tile_shape(rows, cols)
            got = ctx.match_lists(BASIC, bp, x, None, 2)
            used.add(ctx.stats()["tile_shape"])
            for key in ("anchors", "counts", "positions", "distances"):
                torch.testing.assert_close(got[key], want[key], rtol=0, atol=0)
        assert len(used) > 1
        for rows, cols in ((4, 5), (2, 8), (5, 2)):
            with pytest.raises(vnlbhip.VNLBError):
                ctx.set_tile_shape(rows, cols)
        ctx.set_tile_shape(4, 4)
        ctx.set_autotune(True)
        ctx.stage(BASIC, bp, x)
        shape = ctx.stats()["tile_shape"]
        print(f"[fast] {name}: autotuned tile {shape}, parts {ctx.stats()['parts']}", flush=True)
        assert 1 <= shape[0] <= 4 and 1 <= shape[1] <= 4
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["gray_bm12_radius2", "yuv_bm12_radius3"])
def test_fast_form_on_wide_windows(contexts, name):
    """25 x 25 windows (bm_range 12, radius 2 and 3): the MFMA matcher on
    two-row anchor tiles, the fast model and filter; Basic + Final quality
    within the fast-form bound."""
    import workloads

    plain, fast = contexts
    noisy, clean, bp, fp = workloads.make(name)
    x = _dev(noisy)
    print(f"[fast] wide window {name}", flush=True)
    want = plain.stage(FINAL, fp, x, plain.stage(BASIC, bp, x)).cpu().numpy()
    got = fast.stage(FINAL, fp, x, fast.stage(BASIC, bp, x)).cpu().numpy()
    assert np.isfinite(got).all()
    clean = np.asarray(clean, np.float32).reshape(want.shape)
    assert abs(synth.psnr(got, clean) - synth.psnr(want, clean)) <= PSNR_TOL
