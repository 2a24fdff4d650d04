"""Solver limits and reference-legal shapes at the edges, on the GPU.

* The Jacobi sweep limit (JacobiEigenOptions::max_sweeps,
  src/linalg/linalg.hpp:154-158): with one sweep the problems do not
  converge; the plain form still equals the oracle run with the same limit
  bit for bit on the eigenvalues, unconverged problems are counted, and the
  CUDA plugin's behaviour (src/vs/cuda_plugin.cpp:1831-1846) makes the stage
  fail instead.
"""
import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from test_gpu_parity import _check_output, _compare_frame
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, VNLBError, parse_stage_parameters

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1


def test_more_than_8192_candidates(gpu, oracle_lib):
    """tests/cpp/test_cuda_block_match.cpp:931-945's shape: Final matching
    of a 27x27 YUV clip of 13 frames with 1x1 patches, a 27-wide window and
    radius 6, N = 27 * 27 * 13 = 9477 candidates, K = 60, cap 240."""
    noisy, clean = synth.make_clip(27, 27, 13, 3, sigma8=20, seed=60, motion="subpixel")
    args = dict(sigma=20, block_size=1, block_step=1, group_size=60, bm_range=13, patch_time=1, radius=6,
                rank=3, tau=240, cap_factor=4.0)
    bp = parse_stage_parameters(BASIC, args)
    fp = parse_stage_parameters(FINAL, args)
    basic, _ = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
    for stage, params, ref in ((BASIC, bp, None), (FINAL, fp, basic)):
        print(f"[limits] N = 9477, stage {stage}", flush=True)
        want = _compare_frame(gpu, oracle_lib, stage, params, noisy, ref, 6)
        assert want["counts"].max() > 60
        out, _ = oracle_lib.stage(stage, params, noisy, ref, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
        got = gpu.stage(stage, params, torch.from_numpy(noisy).to("cuda:0"),
                        None if ref is None else torch.from_numpy(ref).to("cuda:0"))
        _check_output(got, out, clean)


def test_search_fwd_int_max_minus_one(gpu, oracle_lib):
    """tests/cpp/test_cuda_block_match.cpp:745-762: a legal but extreme
    search_fwd is clamped to the clip before any addition; the temporal range
    it gives is the whole clip, as search_fwd = frames - patch_time gives."""
    noisy, clean = synth.make_clip(24, 20, 5, 1, sigma8=20, seed=61, motion="subpixel")
    base = dict(sigma=20, block_size=3, block_step=2, group_size=3, bm_range=2, patch_time=1, rank=3, tau=5,
                cap_factor=4.0, search_bwd=0)
    huge = parse_stage_parameters(BASIC, dict(base, search_fwd=INT_MAX - 1))
    clamp = parse_stage_parameters(BASIC, dict(base, search_fwd=4))
    x = torch.from_numpy(noisy).to("cuda:0")
    for anchor in (2, 4):
        print(f"[limits] search_fwd INT_MAX - 1, anchor frame {anchor}", flush=True)
        got = gpu.debug_frame(BASIC, huge, x, None, anchor)
        want = _compare_frame(gpu, oracle_lib, BASIC, clamp, noisy, None, anchor)
        np.testing.assert_array_equal(got["counts"], want["counts"])
        for a, n in enumerate(want["counts"]):
            np.testing.assert_array_equal(got["positions"][a, :n], want["positions"][a, :n])
    out, _ = oracle_lib.stage(BASIC, clamp, noisy, flags=oracle_lib.ACCUM_DOUBLE)
    _check_output(gpu.stage(BASIC, huge, x), out, clean)


def test_tuned_rgb_profile_rank_39(gpu, oracle_lib):
    """tests/cpp/test_cuda_group_filter.cpp:709-733: the tuned reference
    profile's K = 100 (Basic) / 60 (Final) with rank 39 at sample dimension
    128: dual Grams of 100 and 60 model rows, beyond one wavefront of the
    cyclic Jacobi solver (two wavefronts per problem)."""
    noisy, clean = synth.make_clip(64, 56, 5, 1, sigma8=20, seed=70, motion="subpixel")
    common = dict(sigma=20, block_size=8, block_step=4, bm_range=9, patch_time=2, radius=1, rank=39, tau=0)
    bp = parse_stage_parameters(BASIC, dict(common, group_size=100))
    fp = parse_stage_parameters(FINAL, dict(common, group_size=60))
    basic, _ = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
    for stage, params, ref in ((BASIC, bp, None), (FINAL, fp, basic)):
        print(f"[limits] K {params.similar} rank 39, stage {stage}", flush=True)
        want = _compare_frame(gpu, oracle_lib, stage, params, noisy, ref, 2)
        assert ((want["group_flags"] & 4) != 0).any()
        out, _ = oracle_lib.stage(stage, params, noisy, ref, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
        x = torch.from_numpy(noisy).to("cuda:0")
        r = None if ref is None else torch.from_numpy(ref).to("cuda:0")
        _check_output(gpu.stage(stage, params, x, r), out, clean)
    # The fast form keeps its matrix-core matcher and takes the exact model
    # kernels for groups this large.
    fast = vnlbhip.Context(0)
    try:
        fast.set_form("fast")
        x = torch.from_numpy(noisy).to("cuda:0")
        got = fast.stage(FINAL, fp, x, fast.stage(BASIC, bp, x)).cpu().numpy()
        plain = gpu.stage(FINAL, fp, x, gpu.stage(BASIC, bp, x)).cpu().numpy()
        assert abs(synth.psnr(got, clean) - synth.psnr(plain, clean)) <= 0.05
    finally:
        fast.close()


def test_retained_cap_above_256(gpu, oracle_lib):
    """cap_factor 20 with K = 16 (soft cap 320) and a tau that admits most
    of the window: groups retain more than 256 patches."""
    noisy, clean = synth.make_clip(64, 56, 4, 1, sigma8=5, seed=62, motion="integer")
    args = dict(sigma=5, block_size=8, block_step=8, group_size=16, bm_range=9, patch_time=1, radius=1, rank=8,
                tau=40000, cap_factor=20.0)
    bp = parse_stage_parameters(BASIC, args)
    print("[limits] retained cap 320", flush=True)
    want = _compare_frame(gpu, oracle_lib, BASIC, bp, noisy, None, 1)
    assert want["cap"] == 320 and want["counts"].max() > 256
    out, _ = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE)
    _check_output(gpu.stage(BASIC, bp, torch.from_numpy(noisy).to("cuda:0")), out, clean)


def _dev(a):
    a = np.ascontiguousarray(a, np.float32)
    return torch.from_numpy(a[:, None] if a.ndim == 3 else a).to("cuda:0")


def test_jacobi_sweep_limit_and_unconverged_problems(oracle_lib):
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    x = _dev(noisy)
    ctx = vnlbhip.Context(0)
    try:
        for form in ("plain", "fast"):
            print(f"[limits] jacobi sweeps, {form}", flush=True)
            ctx.set_form(form)
            ctx.set_jacobi(64)
            full = ctx.stage(BASIC, bp, x).cpu().numpy()
            assert ctx.stats()["unconverged_groups"] == 0
            ctx.set_jacobi(1)
            capped = ctx.stage(BASIC, bp, x).cpu().numpy()
            stats = ctx.stats()
            assert stats["unconverged_groups"] > 0
            assert np.isfinite(capped).all()
            assert np.abs(capped - full).max() > 0.0
            ctx.set_jacobi(1, fail_unconverged=True)
            with pytest.raises(VNLBError, match="did not converge"):
                ctx.stage(BASIC, bp, x)
            ctx.set_jacobi(64, fail_unconverged=True)
            # (Outputs agree up to the order of the accumulators' atomic sums.)
            np.testing.assert_allclose(ctx.stage(BASIC, bp, x).cpu().numpy(), full, rtol=0, atol=1e-5)
        # Plain form: the one-sweep eigenvalues are the oracle's with the same limit.
        ctx.set_form("plain")
        ctx.set_jacobi(1)
        got = ctx.debug_frame(BASIC, bp, x, None, 2)
        oracle_lib.set_max_sweeps(1)
        try:
            want = oracle_lib.frame_debug(BASIC, bp, noisy, None, 2)
        finally:
            oracle_lib.set_max_sweeps(64)
        modelled = (got["group_flags"] & 4) != 0
        assert modelled.any()
        np.testing.assert_array_equal(got["eigenvalues"][modelled], want["eigenvalues"][modelled])
        with pytest.raises(VNLBError):
            ctx.set_jacobi(0)
    finally:
        ctx.close()
