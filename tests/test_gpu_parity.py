"""HIP path against the CPU oracle, through the C ABI.

The plain form follows the oracle's operation order, so match lists
(positions and order), distances, paste-mask decisions and model
eigenvalues are compared bit for bit; aggregation weights to 1e-6 relative
(exp/log are evaluated in double on both sides and can differ by one float
ulp in rare double-rounding cases); stage outputs within max-abs 1e-3 and
PSNR 0.02 dB (the plain-form tolerance), and in practice far closer.
"""
import numpy as np
import pytest
import torch

import workloads
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, StageParameters

pytestmark = pytest.mark.gpu

MAX_ABS = 1e-3
PSNR_TOL = 0.02


def _dev(a):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim == 3:
        a = a[:, None]
    return torch.from_numpy(a).to("cuda:0")


def _compare_frame(gpu, oracle_lib, stage, params, noisy, ref, anchor_frame):
    want = oracle_lib.frame_debug(stage, params, noisy, ref, anchor_frame)
    got = gpu.debug_frame(stage, params, _dev(noisy), _dev(ref) if ref is not None else None,
                          anchor_frame)
    assert got["cap"] == want["cap"]
    np.testing.assert_array_equal(got["anchor_x"], want["anchor_x"])
    np.testing.assert_array_equal(got["anchor_y"], want["anchor_y"])
    np.testing.assert_array_equal(got["counts"], want["counts"])
    for a, n in enumerate(want["counts"]):
        np.testing.assert_array_equal(got["positions"][a, :n], want["positions"][a, :n])
        np.testing.assert_array_equal(got["distances"][a, :n], want["distances"][a, :n])
    np.testing.assert_array_equal(got["active"], want["active"])
    np.testing.assert_array_equal(got["group_flags"], want["group_flags"])
    act = want["active"] == 1
    modelled = act & ((want["group_flags"] & 4) != 0)
    np.testing.assert_array_equal(got["eigenvalues"][modelled], want["eigenvalues"][modelled])
    for a in np.nonzero(act)[0]:
        n = want["counts"][a]
        np.testing.assert_allclose(got["weights"][a, :n], want["weights"][a, :n], rtol=1e-6, atol=0)
    return want


def _check_output(got, want, clean=None):
    got = got.cpu().numpy() if torch.is_tensor(got) else got
    diff = np.abs(got - want)
    assert diff.max() <= MAX_ABS, f"max abs {diff.max()}"
    if clean is not None:
        assert abs(synth.psnr(got, clean) - synth.psnr(want, clean)) <= PSNR_TOL
    return float(diff.max())


@pytest.mark.parametrize("name", list(workloads.WORKLOADS))
def test_match_paste_model_bitwise(gpu, oracle_lib, name):
    noisy, clean, bp, fp = workloads.make(name)
    print(f"[parity] {name} {noisy.shape}", flush=True)
    frames = noisy.shape[0]
    for af in sorted({0, 1, frames - bp.patch_time}):
        _compare_frame(gpu, oracle_lib, BASIC, bp, noisy, None, af)
    basic, _ = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
    for af in sorted({0, 2, frames - fp.patch_time}):
        _compare_frame(gpu, oracle_lib, FINAL, fp, noisy, basic, af)


@pytest.mark.parametrize("name", list(workloads.WORKLOADS))
def test_stage_outputs(gpu, oracle_lib, name):
    noisy, clean, bp, fp = workloads.make(name)
    print(f"[stage] {name} {noisy.shape}", flush=True)
    want_basic, groups = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
    got_basic = gpu.stage(BASIC, bp, _dev(noisy))
    assert gpu.stats()["groups"] == groups
    _check_output(got_basic, want_basic, clean)
    # Final on the same guide clip, then the chained GPU form.
    want_final, _ = oracle_lib.stage(FINAL, fp, noisy, want_basic, flags=oracle_lib.ACCUM_DOUBLE, threads=4)
    got_final = gpu.stage(FINAL, fp, _dev(noisy), _dev(want_basic))
    _check_output(got_final, want_final, clean)
    chained = gpu.stage(FINAL, fp, _dev(noisy), got_basic)
    _check_output(chained, want_final, clean)
    # The reference's float-order accumulation stays within the tolerance too.
    ref_order_basic, _ = oracle_lib.stage(BASIC, bp, noisy, flags=0, threads=4)
    _check_output(got_basic, ref_order_basic, clean)


@pytest.mark.parametrize("case", range(6))
def test_pariasm_fixture_cases(gpu, oracle_lib, case):
    name, noisy, basic_ref, final_ref, bp, fp, tol = workloads.pariasm_cases()[case]
    print(f"[pariasm] {name}", flush=True)
    want_basic, _ = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE)
    got_basic = gpu.stage(BASIC, bp, _dev(noisy))
    _check_output(got_basic, want_basic)
    got_final = gpu.stage(FINAL, fp, _dev(noisy), got_basic)
    want_final, _ = oracle_lib.stage(FINAL, fp, noisy, want_basic, flags=oracle_lib.ACCUM_DOUBLE)
    _check_output(got_final, want_final)
    for got, ref in ((got_basic, basic_ref), (got_final, final_ref)):
        diff = np.abs(got.cpu().numpy()[:, 0] - ref)
        assert diff.max() <= tol["max_abs"] and diff.mean() <= tol["mean_abs"]
        assert np.sqrt((diff ** 2).mean()) <= tol["rmse"]
    _compare_frame(gpu, oracle_lib, BASIC, bp, noisy, None, 1)


def test_integer_valued_clip_is_exact(gpu, oracle_lib):
    """Samples on a 1/16 lattice make every distance exact in any order;
    ties are everywhere, so this pins the (distance, scan position) order."""
    noisy = workloads.integer_clip()
    p = StageParameters(sigma=10 / 255, patch_size=4, patch_time=1, search_window=7, search_bwd=1,
                        search_fwd=1, similar=8, rank=4, proc_step=2, tau=0.05)
    _compare_frame(gpu, oracle_lib, BASIC, p, noisy, None, 1)
    want, _ = oracle_lib.stage(BASIC, p, noisy, flags=oracle_lib.ACCUM_DOUBLE)
    got = gpu.stage(BASIC, p, _dev(noisy)).cpu().numpy()
    assert np.abs(got - want).max() <= 1e-6


def _nan_clip():
    rng = np.random.default_rng(17)
    noisy = rng.random((4, 1, 48, 64), dtype=np.float32)
    noisy[:, :, :, :24] = np.nan
    return noisy


def test_nan_region_anchors_are_skipped(gpu, oracle_lib):
    """Anchors whose every candidate distance is NaN have no matches: they
    form no group and clear nothing (the oracle skips them), and their
    pixels fall back to the (NaN) source."""
    noisy = _nan_clip()
    p = StageParameters(sigma=20 / 255, patch_size=4, patch_time=1, search_window=7, search_bwd=1,
                        search_fwd=1, similar=8, rank=4, proc_step=2)
    print("[nan] debug frame", flush=True)
    want_dbg = _compare_frame(gpu, oracle_lib, BASIC, p, noisy, None, 1)
    assert (want_dbg["counts"] == 0).any() and (want_dbg["counts"] > 0).any()
    print("[nan] stage", flush=True)
    want, groups = oracle_lib.stage(BASIC, p, noisy, flags=oracle_lib.ACCUM_DOUBLE)
    got = gpu.stage(BASIC, p, _dev(noisy)).cpu().numpy()
    assert gpu.stats()["groups"] == groups
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    finite = ~np.isnan(want)
    assert np.abs(got[finite] - want[finite]).max() <= MAX_ABS


def test_constant_clip(gpu):
    noisy = np.full((3, 1, 24, 20), 0.25, np.float32)
    p = StageParameters(sigma=0.1, patch_size=2, patch_time=1, search_window=3, search_bwd=1,
                        search_fwd=1, similar=4, rank=4, beta=2.0, tau=1000.0, proc_step=2)
    out = gpu.stage(BASIC, p, _dev(noisy)).cpu().numpy()
    np.testing.assert_allclose(out, 0.25, atol=1e-5)
    out = gpu.stage(FINAL, p, _dev(noisy), _dev(noisy)).cpu().numpy()
    np.testing.assert_allclose(out, 0.25, atol=1e-5)


def test_single_patch_groups_and_fallback(gpu):
    noisy, _ = synth.make_clip(40, 32, 3, 1, sigma8=10, seed=2)
    p = StageParameters(sigma=10 / 255, patch_size=4, patch_time=1, search_window=5, similar=1,
                        rank=0, proc_step=9)
    out = gpu.stage(BASIC, p, _dev(noisy)).cpu().numpy()
    np.testing.assert_allclose(out, noisy, atol=1e-6)


def test_host_buffer_entry_point(gpu, oracle_lib):
    noisy, clean, bp, fp = workloads.make("gray_p8_step8")
    noisy = noisy[:, :, :64, :80]
    got = gpu.stage(BASIC, bp, noisy)  # numpy in -> vnlbhip_stage_host
    want, _ = oracle_lib.stage(BASIC, bp, noisy, flags=oracle_lib.ACCUM_DOUBLE)
    _check_output(got, want)


def test_invalid_parameters_raise(gpu):
    import vnlbhip

    noisy = _dev(np.zeros((4, 1, 32, 32), np.float32))
    p = StageParameters(sigma=0.1, rank=32, similar=16, patch_time=1)
    with pytest.raises(vnlbhip.VNLBError, match="group_size must cover the effective VNLB rank"):
        gpu.stage(BASIC, p, noisy)
