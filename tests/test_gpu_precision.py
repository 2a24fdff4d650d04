"""Both arithmetic forms against the oracle with its group model in float64.

The oracle's MODEL_DOUBLE flag runs the Gram, the eigen solve, the
projections, the Wiener shrinkage and the aggregation weights in double
(candidate distances stay fp32, as in the reference), so it measures the
error the fp32 (plain) and bf16x3 / parallel-Jacobi (fast) model paths add.
Workload: a 160 x 128 crop of the benchmark configuration (YUV444, chroma,
Pt = 2, step 4, K = 16, tau = 200, cap 4), both stages chained.  Bounds,
on groups whose match lists agree with the oracle's: Gram entries (the
B x B dual Gram the eigen solver receives) within 2e-6 (plain) / 5e-5
(fast) of the group's largest entry (measured 2.6e-7 / 4.8e-6); eigenvalues within 1e-4 (plain) /
1e-3 (fast) of the group's largest eigenvalue; output pixels within
max-abs 1e-3 (plain) and 0.05 dB of PSNR (fast).
"""
import numpy as np
import pytest
import torch

import vnlbhip
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

pytestmark = pytest.mark.gpu

ARGS = dict(sigma=15, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2,
            radius=2, rank=16, tau=200, cap_factor=4.0, model_cap_factor=1.0, chroma=1)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


@pytest.fixture(scope="module")
def crop():
    noisy, clean = synth.make_clip(width=160, height=128, frames=6, channels=3, sigma8=15, seed=31,
                                   motion="subpixel")
    return noisy, clean, parse_stage_parameters(BASIC, ARGS), parse_stage_parameters(FINAL, ARGS)


def _model_errors(got, want):
    worst = 0.0
    worst_gram = 0.0
    compared = 0
    gram_compared = 0
    for a, n in enumerate(want["counts"]):
        if not (want["active"][a] and got["active"][a]) or (want["group_flags"][a] & 4) == 0:
            continue
        if got["counts"][a] != n or not np.array_equal(got["positions"][a, :n], want["positions"][a, :n]):
            continue
        w = want["eigenvalues"][a].astype(np.float64)
        g = got["eigenvalues"][a].astype(np.float64)
        scale = float(np.max(np.abs(w))) + 1e-30
        worst = max(worst, float(np.max(np.abs(g - w))) / scale)
        compared += 1
        gw = want["grams"][a]
        gg = got["grams"][a].astype(np.float64)
        present = ~np.isnan(gw)
        if present.any():
            np.testing.assert_array_equal(~np.isnan(gg), present)
            gram_scale = float(np.max(np.abs(gw[present]))) + 1e-30
            worst_gram = max(worst_gram, float(np.max(np.abs(gg[present] - gw[present]))) / gram_scale)
            gram_compared += 1
    return worst, compared, worst_gram, gram_compared


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_error_against_float64_model(gpu, oracle_lib, crop, form):
    noisy, clean, bp, fp = crop
    flags = oracle_lib.ACCUM_DOUBLE | oracle_lib.MODEL_DOUBLE
    ctx = vnlbhip.Context(0)
    ctx.set_form(form)
    try:
        print(f"[precision] {form} eigenvalues", flush=True)
        want_dbg = oracle_lib.frame_debug(BASIC, bp, noisy, None, 2, flags=flags)
        got_dbg = ctx.debug_frame(BASIC, bp, _dev(noisy), None, 2)
        eig_err, compared, gram_err, gram_compared = _model_errors(got_dbg, want_dbg)
        assert compared >= 50, compared
        assert gram_compared >= 50, gram_compared
        print(f"[precision] {form} stages", flush=True)
        want_basic, _ = oracle_lib.stage(BASIC, bp, noisy, flags=flags, threads=4)
        want_final, _ = oracle_lib.stage(FINAL, fp, noisy, want_basic, flags=flags, threads=4)
        got_basic = ctx.stage(BASIC, bp, _dev(noisy))
        got_final = ctx.stage(FINAL, fp, _dev(noisy), got_basic).cpu().numpy()
        got_basic = got_basic.cpu().numpy()
    finally:
        ctx.close()
    basic_err = float(np.abs(got_basic - want_basic).max())
    final_err = float(np.abs(got_final - want_final).max())
    dpsnr = synth.psnr(got_final, clean) - synth.psnr(want_final, clean)
    print(f"[precision] {form}: Gram rel. error {gram_err:.2e} over {gram_compared} groups; "
          f"eigenvalue rel. error {eig_err:.2e} over {compared} groups; "
          f"max-abs Basic {basic_err:.2e}, Final {final_err:.2e}; dPSNR Final {dpsnr:+.4f} dB", flush=True)
    if form == "plain":
        assert gram_err <= 2e-6
        assert eig_err <= 1e-4
        assert basic_err <= 1e-3 and final_err <= 1e-3
    else:
        assert gram_err <= 5e-5
        assert eig_err <= 1e-3
        assert abs(dpsnr) <= 0.05
