"""vnlbhip_stage_scratch: the stage without device allocation.

The entry point carves its buffers from a fixed per-GPU scratch array; its
results must equal vnlbhip_stage's (same kernels, same order) on two calls
in a row, the second with a smaller clip, in both arithmetic forms, and a
stage larger than the array must be refused before any kernel runs.
"""
import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from vnlbhip.params import BASIC, FINAL, VNLBError, parse_stage_parameters

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.ascontiguousarray(a, np.float32)
    return torch.from_numpy(a[:, None] if a.ndim == 3 else a).to("cuda:0")


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_scratch_stage_matches_stage(form):
    big, _, bp, fp = workloads.make("yuv_pt2_tau200")
    x = _dev(big)
    # The second call: a smaller clip (fewer frames, a crop).
    small = x[:4, :, :48, :64].contiguous()
    ctx = vnlbhip.Context(0)
    ref_ctx = vnlbhip.Context(0)
    try:
        ctx.set_form(form)
        ref_ctx.set_form(form)
        for name, clip in (("full", x), ("smaller", small)):
            print(f"[scratch] {form} {name} {tuple(clip.shape)}", flush=True)
            want_b = ref_ctx.stage(BASIC, bp, clip)
            want_f = ref_ctx.stage(FINAL, fp, clip, want_b)
            got_b = ctx.stage_scratch(BASIC, bp, clip)
            stats_b = ctx.stats()
            got_f = ctx.stage_scratch(FINAL, fp, clip, want_b)
            stats_f = ctx.stats()
            torch.cuda.synchronize()
            assert stats_b["reallocations"] == 0 and stats_f["reallocations"] == 0
            want = ref_ctx.stats()
            assert stats_f["groups"] == want["groups"] and stats_f["retained"] == want["retained"]
            # Same kernels on the same inputs (Final on the same Basic output):
            # match lists and groups agree exactly, outputs up to the order of
            # the accumulators' atomic additions.
            tol = 1e-6 if form == "plain" else 1e-5
            np.testing.assert_allclose(got_b.cpu().numpy(), want_b.cpu().numpy(), rtol=0, atol=tol)
            np.testing.assert_allclose(got_f.cpu().numpy(), want_f.cpu().numpy(), rtol=0, atol=tol)
        # The context's own buffers were left alone: its ordinary stage still
        # works and allocates once.
        again = ctx.stage(BASIC, bp, x)
        np.testing.assert_allclose(again.cpu().numpy(), ref_ctx.stage(BASIC, bp, x).cpu().numpy(),
                                   rtol=0, atol=1e-6 if form == "plain" else 1e-5)
    finally:
        ctx.close()
        ref_ctx.close()


def test_scratch_stage_refuses_stages_larger_than_the_array():
    p = parse_stage_parameters(BASIC, dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=9,
                                           patch_time=2, radius=2, rank=16, tau=0))
    # 1024x512x20 GrayS: the fp64 accumulators alone take 2 * 8 B per pixel
    # and frame = 168 MB, over the 128 MiB array.
    x = torch.rand((20, 1, 512, 1024), dtype=torch.float32, device="cuda:0")
    out = torch.full_like(x, -1.0)
    ctx = vnlbhip.Context(0)
    try:
        with pytest.raises(VNLBError, match="scratch"):
            ctx.stage_scratch(BASIC, p, x, out=out)
        torch.cuda.synchronize()
        assert bool((out == -1.0).all()), "a refused call must not launch anything"
        # The same context still runs a stage that fits.
        small = x[:4, :, :64, :80].contiguous()
        got = ctx.stage_scratch(BASIC, p, small)
        want = ctx.stage(BASIC, p, small)
        np.testing.assert_allclose(got.cpu().numpy(), want.cpu().numpy(), rtol=0, atol=1e-6)
    finally:
        ctx.close()
