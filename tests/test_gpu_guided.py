"""Search-centre guidance (MVToolsFlowProvider's centres,
src/flow/flow.hpp:321-390, as a dense offset field; the reference centres
each scheduled frame's window on them, src/core/core.cpp:1328-1340) on the
HIP path against the oracle with the same field.

Plain form: match lists, distances, paste decisions and eigenvalues bit for
bit, stage outputs within max-abs 1e-3.  Fast form: guided stages keep the
plain matcher, so match lists are identical to the oracle's; outputs within
0.05 dB of PSNR.
"""
import numpy as np
import pytest
import torch

import vnlbhip
import workloads
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL

from test_gpu_parity import _check_output, _compare_frame, _dev

pytestmark = pytest.mark.gpu

CASES = ["gray_pt2_step4", "yuv_pt2_tau200"]


def _field(noisy, params, seed, span=6):
    clip = np.ascontiguousarray(noisy, np.float32)
    frames = clip.shape[0]
    height, width = clip.shape[-2:]
    temporal = params.search_bwd + params.search_fwd + 1
    rng = np.random.default_rng(seed)
    field = rng.integers(-span, span + 1, size=(frames, temporal, height, width, 2)).astype(np.int16)
    return field, temporal


@pytest.mark.parametrize("name", CASES)
def test_guided_plain_matches_the_oracle(gpu, oracle_lib, name):
    noisy, clean, bp, fp = workloads.make(name)
    field, temporal = _field(noisy, bp, seed=5)
    frames = np.ascontiguousarray(noisy).shape[0]
    dev_field = torch.from_numpy(field).to("cuda:0")
    try:
        oracle_lib.set_search_offsets(field, temporal)
        gpu.set_search_offsets(dev_field)
        print(f"[guided] {name} frames", flush=True)
        for anchor_frame in sorted({0, frames // 2, frames - bp.patch_time}):
            _compare_frame(gpu, oracle_lib, BASIC, bp, noisy, None, anchor_frame)
        print(f"[guided] {name} stages", flush=True)
        want_basic, groups = oracle_lib.stage(BASIC, bp, noisy)
        got_basic = gpu.stage(BASIC, bp, _dev(noisy))
        assert gpu.stats()["groups"] == groups
        _check_output(got_basic, want_basic.reshape(got_basic.shape))
        want_final, _ = oracle_lib.stage(FINAL, fp, noisy, want_basic)
        got_final = gpu.stage(FINAL, fp, _dev(noisy), _dev(want_basic))
        _check_output(got_final, want_final.reshape(got_final.shape))
        _compare_frame(gpu, oracle_lib, FINAL, fp, noisy, want_basic, frames // 2)
    finally:
        gpu.set_search_offsets(None)
        oracle_lib.set_search_offsets(None)


def test_zero_field_is_no_guidance(gpu):
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    field, _ = _field(noisy, bp, seed=0)
    x = _dev(noisy)
    want = gpu.stage(BASIC, bp, x).cpu().numpy()
    try:
        gpu.set_search_offsets(torch.zeros(field.shape, dtype=torch.int16, device="cuda:0"))
        got = gpu.stage(BASIC, bp, x).cpu().numpy()
    finally:
        gpu.set_search_offsets(None)
    np.testing.assert_array_equal(got, want)


def test_guided_fast_form(oracle_lib):
    noisy, clean, bp, fp = workloads.make("yuv_pt2_tau200")
    field, temporal = _field(noisy, bp, seed=9)
    frames = np.ascontiguousarray(noisy).shape[0]
    ctx = vnlbhip.Context(0)
    try:
        ctx.set_form("fast")
        ctx.set_search_offsets(torch.from_numpy(field).to("cuda:0"))
        oracle_lib.set_search_offsets(field, temporal)
        print("[guided] fast frame", flush=True)
        want = oracle_lib.frame_debug(BASIC, bp, noisy, None, frames // 2)
        got = ctx.debug_frame(BASIC, bp, _dev(noisy), None, frames // 2)
        np.testing.assert_array_equal(got["counts"], want["counts"])
        for a, n in enumerate(want["counts"]):
            np.testing.assert_array_equal(got["positions"][a, :n], want["positions"][a, :n])
        print("[guided] fast stages", flush=True)
        want_basic, _ = oracle_lib.stage(BASIC, bp, noisy)
        want_final, _ = oracle_lib.stage(FINAL, fp, noisy, want_basic)
        got_basic = ctx.stage(BASIC, bp, _dev(noisy))
        got_final = ctx.stage(FINAL, fp, _dev(noisy), got_basic).cpu().numpy()
    finally:
        ctx.close()
        oracle_lib.set_search_offsets(None)
    dpsnr = synth.psnr(got_final.reshape(want_final.shape), clean) - synth.psnr(want_final, clean)
    print(f"[guided] fast dPSNR {dpsnr:+.4f} dB", flush=True)
    assert abs(dpsnr) <= 0.05


def test_guided_graph_replays_follow_the_field(gpu):
    """A replayed stage graph holds the field's address: a new field (another
    tensor) is a new graph, and replays match eager runs."""
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    first, _ = _field(noisy, bp, seed=1)
    second, _ = _field(noisy, bp, seed=2)
    x = _dev(noisy)
    fields = [torch.from_numpy(f).to("cuda:0") for f in (first, second)]
    eager = []
    try:
        for f in fields:
            gpu.set_search_offsets(f)
            eager.append(gpu.stage(BASIC, bp, x).cpu().numpy())
        assert np.abs(eager[0] - eager[1]).max() > 0
        ctx = vnlbhip.Context(0)
        try:
            ctx.set_graph(True)
            out = torch.empty_like(x)
            for call in range(4):
                ctx.set_search_offsets(fields[call % 2])
                got = ctx.stage(BASIC, bp, x, out=out).cpu().numpy()
                np.testing.assert_array_equal(got, eager[call % 2], err_msg=f"call {call}")
            assert ctx.stats()["graph_replays"] >= 2
        finally:
            ctx.close()
    finally:
        gpu.set_search_offsets(None)


def test_guided_rejects_a_mismatched_field(gpu):
    noisy, _, bp, _ = workloads.make("gray_pt2_step4")
    field, temporal = _field(noisy, bp, seed=0)
    wrong = torch.zeros((field.shape[0], temporal + 2) + field.shape[2:], dtype=torch.int16, device="cuda:0")
    try:
        gpu.set_search_offsets(wrong)
        with pytest.raises(vnlbhip.VNLBError):
            gpu.stage(BASIC, bp, _dev(noisy))
    finally:
        gpu.set_search_offsets(None)
    with pytest.raises(vnlbhip.VNLBError):
        gpu.set_search_offsets(torch.zeros((2, 2), dtype=torch.int16, device="cuda:0"))
