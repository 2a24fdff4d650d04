"""Seeded test workloads shaped like the benchmark configurations, at sizes
the CPU oracle finishes in seconds."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

GOLDEN = Path(__file__).resolve().parent / "golden"


def params_pair(common: dict, basic_extra: dict | None = None, final_extra: dict | None = None):
    bp = parse_stage_parameters(BASIC, {**common, **(basic_extra or {})})
    fp = parse_stage_parameters(FINAL, {**common, **(final_extra or {})})
    return bp, fp


# name -> (clip kwargs, common plugin args, basic extra, final extra)
WORKLOADS = {
    # Plain VNLB, GrayS, configuration of the reference CPU case (256x256x5).
    "gray_p8_step8": (
        dict(width=256, height=256, frames=5, channels=1, sigma8=10, seed=11, motion="integer"),
        dict(sigma=10, block_size=8, block_step=8, group_size=8, bm_range=9, patch_time=1, radius=1,
             rank=8, tau=0),
        None, None),
    # GrayS, Pt=2 (d=128), step 4, K=16, radius 2, sub-pixel motion (crop).
    "gray_pt2_step4": (
        dict(width=112, height=96, frames=6, channels=1, sigma8=20, seed=12, motion="subpixel"),
        dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2, radius=2,
             rank=16, tau=0),
        None, None),
    # YUV444, coupled chroma (d=384), tau=200 and cap_factor 4: uneven groups.
    "yuv_pt2_tau200": (
        dict(width=96, height=80, frames=6, channels=3, sigma8=15, seed=13, motion="subpixel"),
        dict(sigma=15, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2, radius=2,
             rank=16, tau=200, cap_factor=4.0, model_cap_factor=1.0, chroma=1),
        None, None),
    # GrayS, 25x25 window (bm_range 12, 625 candidates per temporal slot),
    # radius 2: the matching shape of the learned-proposal configuration
    # (crop; the MFMA matcher's tiles take windows up to 19, so both forms
    # match with the plain matcher here).
    "gray_bm12_radius2": (
        dict(width=88, height=72, frames=6, channels=1, sigma8=20, seed=15, motion="subpixel"),
        dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=12, patch_time=2, radius=2,
             rank=16, tau=0),
        None, None),
    # YUV444, bm_range 12, radius 3 (4375 candidates), tau 200, cap 4, sigma 25:
    # the shape of the full 4K configuration (crop; rank clamped to K = 16).
    "yuv_bm12_radius3": (
        dict(width=80, height=64, frames=8, channels=3, sigma8=25, seed=16, motion="subpixel"),
        dict(sigma=25, block_size=8, block_step=4, group_size=16, bm_range=12, patch_time=2, radius=3,
             rank=16, tau=200, cap_factor=4.0, model_cap_factor=1.0, chroma=1),
        None, None),
    # GrayS at low noise with a large tau: groups widen from K = 16 up to the
    # cap of 64, and model_cap_factor 2 makes up to 32 of them model rows
    # (rows beyond K included).
    "gray_model_cap2": (
        dict(width=80, height=64, frames=5, channels=1, sigma8=5, seed=17, motion="subpixel"),
        dict(sigma=5, block_size=8, block_step=4, group_size=16, bm_range=7, patch_time=2, radius=1,
             rank=16, tau=10000, cap_factor=4.0, model_cap_factor=2.0),
        None, None),
    # YUV444 with uncoupled channels (three per-channel models).
    "yuv_uncoupled": (
        dict(width=64, height=56, frames=4, channels=3, sigma8=15, seed=14, motion="integer"),
        dict(sigma=15, block_size=8, block_step=6, group_size=12, bm_range=5, patch_time=1, radius=1,
             rank=8, chroma=0),
        None, None),
}


def make(name: str):
    clip_kw, common, bex, fex = WORKLOADS[name]
    noisy, clean = synth.make_clip(**clip_kw)
    bp, fp = params_pair(common, bex, fex)
    return noisy, clean, bp, fp


def pariasm_cases():
    meta = json.loads((GOLDEN / "pariasm" / "cases.json").read_text())
    out = []
    for case in meta["cases"]:
        d = GOLDEN / "pariasm" / case
        noisy = np.load(d / "input.npz")["noisy"]
        basic_ref = np.load(d / "basic_ref.npz")["basic"]
        final_ref = np.load(d / "final_ref.npz")["final"]
        bp, fp = params_pair(meta["common"], meta["basic"], meta["final"])
        out.append((case, noisy, basic_ref, final_ref, bp, fp, meta["tolerance"]))
    return out


def integer_clip(width=48, height=40, frames=4, seed=5):
    """Samples that are multiples of 1/16 in [0, 1]: every squared distance
    and every partial sum is exact in float32, so any summation order gives
    the same bits."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 17, size=(height + 8, width + 8)).astype(np.float32) / 16.0
    frames_ = [base[t % 4: t % 4 + height, 0:width] for t in range(frames)]
    return np.stack(frames_)[:, None].astype(np.float32)
