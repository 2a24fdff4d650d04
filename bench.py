"""VNLB Basic + Final throughput on one MI355X.

One step = vnlbhip Basic over the whole clip, then Final with ref = that
Basic output, both returning aggregated frames, with the clip already
resident in HBM.  Metric: Mpix/s counted once per spatial position.

Default workload: YUV444PS 1920x1080, 16 frames, chroma=1, patch_time=2,
block_step=4, group_size=16, bm_range=9, radius=2, tau=200, cap_factor=4,
model_cap_factor=1, rank=16 (rank 32 is not valid with group_size 16: the
reference requires group_size >= min(dim, rank)), sigma=15.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "vapoursynth-vnlb_amd"))

MI355X_PEAK = {"hbm_gbs": 8000.0, "fp32_tflops": 157.3, "bf16_tflops": 2500.0}

WORKLOADS = {
    "yuv444ps_1080p_16f": dict(
        clip=dict(width=1920, height=1080, frames=16, channels=3, sigma8=15, seed=2026, motion="subpixel"),
        args=dict(sigma=15, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2, radius=2,
                  rank=16, tau=200, cap_factor=4.0, model_cap_factor=1.0, chroma=1)),
    "grays_1080p_16f": dict(
        clip=dict(width=1920, height=1080, frames=16, channels=1, sigma8=20, seed=2026, motion="subpixel"),
        args=dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2, radius=2,
                  rank=16, tau=0)),
    # The matching shape of the learned-proposal configuration (bm_range 12)
    # without the learned part: exhaustive search.
    "grays_1080p_16f_bm12": dict(
        clip=dict(width=1920, height=1080, frames=16, channels=1, sigma8=20, seed=2026, motion="subpixel"),
        args=dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=12, patch_time=2, radius=2,
                  rank=16, tau=0)),
    # The full 4K configuration (rank clamped to K = 16), exhaustive search.
    "yuv444ps_2160p_8f": dict(
        clip=dict(width=3840, height=2160, frames=8, channels=3, sigma8=25, seed=2026, motion="subpixel"),
        args=dict(sigma=25, block_size=8, block_step=4, group_size=16, bm_range=12, patch_time=2, radius=3,
                  rank=16, tau=200, cap_factor=4.0, model_cap_factor=1.0, chroma=1)),
    "grays_256_5f": dict(
        clip=dict(width=256, height=256, frames=5, channels=1, sigma8=10, seed=2026, motion="integer"),
        args=dict(sigma=10, block_size=8, block_step=8, group_size=8, bm_range=9, patch_time=1, radius=1,
                  rank=8, tau=0)),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_match(params, width, height, frames, channels, anchors, stage):
    """Per launch-set: GEMM-form distance flops 2*N*d per anchor, N the
    candidates of the boundary-shifted window (constant per anchor)."""
    nx = min(params.search_window, width - params.patch_size + 1)
    ny = min(params.search_window, height - params.patch_size + 1)
    nt = min(params.search_bwd + params.search_fwd + 1, frames - params.patch_time + 1)
    d = params.patch_size ** 2 * params.patch_time * (channels if stage == 1 else 1)
    return 2.0 * nx * ny * nt * d * anchors


def measured_traffic(kernel: str, launches_per_step: int):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, copy-calibrated), rescaled
    from the profiled run's launches per step (default pipeline parts) to
    this run's, or None."""
    path = REPO / "profiles" / "hbm_traffic_fast_kernels_yuv444ps_1080p_16f.md"
    if not path.exists():
        return None
    per_launch = None
    profiled = None
    for line in path.read_text().splitlines():
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            per_launch = float(line.split(":", 1)[1])
        if line.startswith(f"traffic_launches_per_step {kernel}:"):
            profiled = int(line.split(":", 1)[1])
    if per_launch is None or profiled is None or launches_per_step <= 0:
        return None
    return per_launch * profiled / launches_per_step


def match_roofline(form, times, bp, fp, clip, sb, sf) -> dict:
    """Roofline of the distance kernel: algorithmic GEMM-form flops (2 N d
    per anchor, the candidates each anchor needs, not the tile unions or the
    three bf16 products) over its average launch time from hipEvents."""
    flops = (algorithmic_match(bp, clip["width"], clip["height"], clip["frames"], clip["channels"], sb["anchors"], 0)
             + algorithmic_match(fp, clip["width"], clip["height"], clip["frames"], clip["channels"], sf["anchors"], 1))
    match = times["match"]
    launches = max(1, match["launches"])
    avg_s = match["ms"] / launches / 1000.0
    achieved = flops / launches / avg_s / 1e12 if avg_s > 0 else 0.0
    if form == "fast":
        kernel, peak = "match_fast_kernel", MI355X_PEAK["bf16_tflops"]
        traffic = measured_traffic("match_fast_kernel", launches)
    else:
        kernel, peak, traffic = "match_kernel", MI355X_PEAK["fp32_tflops"], None
    return {
        "kernel": kernel,
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": "TFLOP/s",
        "frac": round(achieved / peak, 5),
        "traffic": traffic,
        "launch_ms": round(avg_s * 1000.0, 4),
        "algorithmic_per_launch": round(flops / launches, 1),
    }


def filter_roofline(form, times, bp, fp, clip, sb, sf) -> dict:
    """Roofline of the filter + aggregation kernel by the max rule: the
    bound is whichever of its algorithmic flops (at the fp32 peak) and its
    algorithmic HBM bytes (at the HBM peak) takes longer, over the kernel's
    average launch time from hipEvents (a stage launches it once per part of
    its anchor frames, so per launch is the stage total over the launches).
      flops, per group of S retained samples, B model rows, rank R and
      dimension D: projection + reconstruction 4 S R D and the dual map
      2 R B D;
      minimum bytes, per stage: the fp32 accumulators written and read once,
      frames * (C + 1) * pixels * 4 * 2, and one read of the clip (and of the
      Basic estimate in Final), frames * C * pixels * 4;
      scatter-design bytes, reported beside it: S * D * 4 (filtered rows
      read) + S * Pt * P^2 * (C + 1) * 8 (accumulator read-modify-write)."""
    channels = clip["channels"]
    pixels = clip["width"] * clip["height"]
    frames = clip["frames"]
    flops = 0.0
    minimum = 0.0
    scatter = 0.0
    for stage, params, stats in ((0, bp, sb), (1, fp, sf)):
        dim = params.patch_size ** 2 * params.patch_time * channels
        rows = max(1, int(params.similar * params.model_cap_factor))
        rank = min(params.rank, rows)
        flops += 4.0 * rank * dim * stats["retained"] + 2.0 * rank * rows * dim * stats["groups"]
        minimum += 2.0 * frames * (channels + 1) * pixels * 4.0 + frames * channels * pixels * 4.0 * (1 + stage)
        contributions = params.patch_time * params.patch_size ** 2 * (channels + 1)
        scatter += stats["retained"] * (dim * 4.0 + contributions * 8.0)
    filt = times["filter"]
    launches = max(1, filt["launches"])
    avg_s = filt["ms"] / launches / 1000.0
    kernel = "filter_fast_kernel" if form == "fast" else "filter_kernel"

    def gbs(total):
        achieved = total / launches / avg_s / 1e9 if avg_s > 0 else 0.0
        return {"bytes_per_launch": round(total / launches, 1), "achieved_gbs": round(achieved, 2),
                "frac": round(achieved / MI355X_PEAK["hbm_gbs"], 5)}

    flop_time = flops / (MI355X_PEAK["fp32_tflops"] * 1e12)
    byte_time = minimum / (MI355X_PEAK["hbm_gbs"] * 1e9)
    if flop_time >= byte_time:
        achieved = flops / launches / avg_s / 1e12 if avg_s > 0 else 0.0
        head = {"bound": "mfma", "achieved": round(achieved, 3), "peak": MI355X_PEAK["fp32_tflops"],
                "unit": "TFLOP/s", "frac": round(achieved / MI355X_PEAK["fp32_tflops"], 5),
                "algorithmic_per_launch": round(flops / launches, 1)}
    else:
        g = gbs(minimum)
        head = {"bound": "hbm", "achieved": g["achieved_gbs"], "peak": MI355X_PEAK["hbm_gbs"], "unit": "GB/s",
                "frac": g["frac"], "algorithmic_per_launch": g["bytes_per_launch"]}
    return {
        "kernel": kernel,
        **head,
        "traffic": measured_traffic(kernel, launches) if form == "fast" else None,
        "launch_ms": round(avg_s * 1000.0, 4),
        "hbm_minimum": gbs(minimum),
        "scatter_design": gbs(scatter),
    }


def train_proposer(ctx, steps: int, seed: int = 7):
    """The learned candidate proposer, trained on the GPU (its own forward,
    backward and Adam kernels) on the plain matcher's exact 16 nearest over
    fresh seeded clips of the precision configuration's kind (GrayS, sigma
    20, sub-pixel motion, bm_range 12, radius 2): about 200k (anchor, window)
    pairs from a 960x540x6 clip, matched on the noisy clip (Basic) and on its
    Basic estimate (Final).  Returns (proposer, summary)."""
    import torch

    from vnlbhip import proposer as prop
    from vnlbhip import synth
    from vnlbhip.params import BASIC, parse_stage_parameters

    args = dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=12, patch_time=2, radius=2, rank=4,
                tau=0)
    params = parse_stage_parameters(BASIC, args)
    t0 = time.perf_counter()
    noisy, _ = synth.make_clip(960, 540, 6, 1, sigma8=20, seed=seed, motion="subpixel")
    x = torch.from_numpy(noisy).to("cuda:0")
    basic = ctx.stage(BASIC, params, x)
    examples = [prop.make_examples(ctx, x, params, anchor_frames=(1, 2, 3), seed=seed, max_anchors=100_000),
                prop.make_examples(ctx, basic, params, anchor_frames=(1, 2, 3), seed=seed + 1, max_anchors=100_000)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    proposer = prop.Proposer(0, seed=seed)
    history = prop.train(proposer, examples, steps=steps, batch_anchors=64, learning_rate=2e-3,
                         log=lambda s, v: log(f"[bench] proposer step {s}: loss {v:.4f}") if s % 500 == 0 else None)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # FP8 copies calibrated on a sample of the training rows.
    rows, _ = prop.batch_rows(proposer, examples[0], torch.arange(64, device="cuda:0"))
    proposer.calibrate(rows)
    return proposer, {
        "pairs": sum(len(e) for e in examples),
        "rows_per_pair": examples[0].group,
        "steps": steps,
        "labels_s": round(t1 - t0, 2),
        "train_s": round(t2 - t1, 2),
        "loss_first": round(history[0][1], 4),
        "loss_last": round(float(np.mean([v for _, v in history[-4:]])), 4),
    }


def learned_form(ctx, proposer, step, bp, fp, noisy, clean_np, final_plain_np, pixels, steps: int, margin: float,
                 precision: str) -> dict:
    """The combined form (the context's form with learned proposals): step
    time, quality against the plain form, recall of one anchor frame's lists,
    exhaustive-fallback share and the proposing matcher's time."""
    import torch

    from vnlbhip import synth
    from vnlbhip.params import BASIC, FINAL

    exact = {}
    ctx.set_proposer(None)
    form = ctx.form
    ctx.set_form("plain")
    frame = min(2, int(noisy.shape[0]) - 3)
    exact[BASIC] = ctx.match_lists(BASIC, bp, noisy, None, frame)
    ctx.set_form(form)
    ctx.set_proposer(proposer, top_k=48, margin=margin, precision=precision)
    got = ctx.match_lists(BASIC, bp, noisy, None, frame)
    k = bp.similar
    g = got["positions"][:, :k].cpu().numpy()
    w = exact[BASIC]["positions"][:, :k].cpu().numpy()
    recall = float(np.mean([len(set(a) & set(b)) / k for a, b in zip(g, w)]))
    step()  # warm (the first call of a stage captures its graph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, final, sb, sf = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    ctx.set_profiling(True)
    step()
    torch.cuda.synchronize()
    times = ctx.kernel_times()
    ctx.set_profiling(False)
    ctx.set_proposer(None)
    final_np = final.cpu().numpy()
    psnr = synth.psnr(final_np, clean_np)
    share = (sb["exhaustive_fallbacks"] + sf["exhaustive_fallbacks"]) / max(1, sb["proposed"] + sf["proposed"])
    return {
        "precision": precision,
        "top_k": 48,
        "margin": margin,
        "mpix_s": round(pixels / (ms / 1000.0) / 1e6, 4),
        "ms_per_step": round(ms, 3),
        "dpsnr_vs_plain_db": round(psnr - synth.psnr(final_plain_np, clean_np), 4),
        "max_abs_vs_plain": float(np.abs(final_np - final_plain_np).max()),
        "recall_at_16_basic": round(recall, 4),
        "exhaustive_fallback_share": round(share, 4),
        "match_ms": round(times["match"]["ms"], 3),
        "match_launches": times["match"]["launches"],
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_rate(workload: dict, width: int, height: int, threads: int) -> tuple[float, float]:
    """(Mpix/s, seconds) of the oracle's Basic + Final on a width x height
    crop of the workload's clip."""
    from oracle import oracle
    from vnlbhip import synth
    from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

    clip = dict(workload["clip"])
    clip.update(width=width, height=height)
    noisy, _ = synth.make_clip(**clip)
    bp = parse_stage_parameters(BASIC, workload["args"])
    fp = parse_stage_parameters(FINAL, workload["args"])
    t0 = time.perf_counter()
    basic, _ = oracle.stage(BASIC, bp, noisy, flags=oracle.SELECT_FAITHFUL, threads=threads)
    oracle.stage(FINAL, fp, noisy, basic, flags=oracle.SELECT_FAITHFUL, threads=threads)
    elapsed = time.perf_counter() - t0
    return width * height * clip["frames"] / elapsed / 1e6, elapsed


def cpu_baseline(workload: dict) -> dict:
    """The CPU oracle (a C++ restatement of the reference's CPU path, with its
    sequential early-exit selection), frame-parallel over the host cores this
    job has (at most 16) on a 480x270 crop of the same clip, beside a
    1-thread run on a 160x90 crop and the reference CPU configuration
    (GrayS 256x256x5) in full at both thread counts."""
    threads = min(16, os.cpu_count() or 1)
    value, seconds = _oracle_rate(workload, 480, 270, threads)
    one, one_s = _oracle_rate(workload, 160, 90, 1)
    small = WORKLOADS["grays_256_5f"]
    cfg1, cfg1_s = _oracle_rate(small, 256, 256, threads)
    cfg1_one, cfg1_one_s = _oracle_rate(small, 256, 256, 1)
    frames = workload["clip"]["frames"]
    return {
        "value": round(value, 4),
        "unit": "Mpix/s",
        "cores": threads,
        "kind": "port",
        "sample": f"Basic+Final of the same workload on a 480x270x{frames} crop ({seconds:.1f} s, frame-parallel, "
                  f"reference sequential selection with early exit)",
        "nproc": os.cpu_count(),
        "cpu_model": _cpu_model(),
        "one_thread": {"value": round(one, 4), "sample": f"160x90x{frames} crop ({one_s:.1f} s)"},
        "grays_256_5f": {"threads": threads, "value": round(cfg1, 4), "seconds": round(cfg1_s, 2),
                         "one_thread_value": round(cfg1_one, 4), "one_thread_seconds": round(cfg1_one_s, 2)},
    }


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=3)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--workload", default="yuv444ps_1080p_16f", choices=sorted(WORKLOADS))
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-graph", action="store_true",
                        help="launch every kernel eagerly instead of replaying captured stage graphs")
    parser.add_argument("--no-autotune", action="store_true",
                        help="keep the default pipeline part count instead of timing the candidates")
    parser.add_argument("--form", default="fast", choices=("fast", "plain"),
                        help="fast: MFMA candidate distances (the metric's form); plain: exact fp32 order")
    parser.add_argument("--plain-steps", type=int, default=5,
                        help="timed steps of the plain form (after one warm step)")
    parser.add_argument("--no-denoising", action="store_true",
                        help="skip the rank-4 denoising configuration (fast and plain on the same clip)")
    parser.add_argument("--denoising-steps", type=int, default=3,
                        help="timed steps per form of the rank-4 denoising configuration")
    parser.add_argument("--no-learned", action="store_true",
                        help="skip the learned-proposal (combined) form's report")
    parser.add_argument("--proposer-steps", type=int, default=3000,
                        help="Adam steps of the proposer's training (64 anchors x 128 rows each)")
    parser.add_argument("--learned-steps", type=int, default=1,
                        help="timed steps of the combined form (the slowest form; after one warm step)")
    parser.add_argument("--learned-margin", type=float, default=0.5,
                        help="score margin below which an anchor's window is searched exhaustively")
    parser.add_argument("--frame-cache", action="store_true",
                        help="also time per-output-frame rendering over the anchor-stack cache "
                             "(Basic then Final, every frame once, in order)")
    args = parser.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(anchor frames would be sharded across GPUs; see the design notes)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import vnlbhip
    from vnlbhip import synth
    from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

    workload = WORKLOADS[args.workload]
    clip = workload["clip"]
    log(f"[bench] generating {args.workload}")
    noisy_np, clean_np = synth.make_clip(**clip)
    noisy = torch.from_numpy(noisy_np).to("cuda:0")
    bp = parse_stage_parameters(BASIC, workload["args"])
    fp = parse_stage_parameters(FINAL, workload["args"])
    ctx = vnlbhip.Context(0)
    ctx.set_form(args.form)
    # The first warmup step picks each stage's pipeline part count by timing
    # (without warmup steps the default count is used).
    ctx.set_autotune(args.warmup > 0 and not args.no_autotune)
    # Stage graphs: the first call of each stage runs eagerly and captures its
    # launch sequence; later calls with the same tensors replay it (the kernel
    # timing step below runs eagerly).
    ctx.set_graph(not args.no_graph)
    basic_out = torch.empty_like(noisy)
    final_out = torch.empty_like(noisy)

    def step():
        basic = ctx.stage(BASIC, bp, noisy, out=basic_out)
        stats_b = ctx.stats()
        final = ctx.stage(FINAL, fp, noisy, basic, out=final_out)
        stats_f = ctx.stats()
        return basic, final, stats_b, stats_f

    for i in range(args.warmup):
        log(f"[bench] warmup {i + 1}/{args.warmup} ({args.form})")
        step()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} steps ({args.form})")
    t0 = time.perf_counter()
    for i in range(args.steps):
        basic, final, stats_b, stats_f = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000.0 / args.steps
    pixels = clip["width"] * clip["height"] * clip["frames"]
    value = pixels / (ms_per_step / 1000.0) / 1e6

    # Kernel timing (hipEvents on the stage stream) over one extra step.
    ctx.set_profiling(True)
    step_profile = step()
    torch.cuda.synchronize()
    times = ctx.kernel_times()
    ctx.set_profiling(False)
    log("[bench] kernel ms: " + ", ".join(f"{k}={v['ms']:.2f}/{v['launches']}" for k, v in times.items()))
    _, _, sb, sf = step_profile
    # The dominant kernel (largest share of the step) carries the roofline.
    rooflines = {
        "match": match_roofline(args.form, times, bp, fp, clip, sb, sf),
        "filter": filter_roofline(args.form, times, bp, fp, clip, sb, sf),
    }
    dominant = max(rooflines, key=lambda k: times[k]["ms"])
    roofline = rooflines[dominant]

    # The plain form on the same clip: quality reference of the metric and
    # its own throughput (one untimed warm step, then --plain-steps timed).
    quality = {}
    final_np = final.cpu().numpy()
    psnr_final = synth.psnr(final_np, clean_np)
    if args.form != "plain":
        ctx.set_form("plain")
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            _, final_plain, _, _ = step()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1000.0 / args.plain_steps
        plain_np = final_plain.cpu().numpy()
        psnr_plain = synth.psnr(plain_np, clean_np)
        quality = {
            "psnr_final_plain_db": round(psnr_plain, 4),
            "dpsnr_vs_plain_db": round(psnr_final - psnr_plain, 4),
            "max_abs_vs_plain": float(np.abs(final_np - plain_np).max()),
            "plain_mpix_s": round(pixels / (plain_ms / 1000.0) / 1e6, 4),
            "plain_steps": args.plain_steps,
            "fast_over_plain": round(plain_ms / ms_per_step, 4),
        }
        log(f"[bench] plain form: {plain_ms:.1f} ms/step, dPSNR {psnr_final - psnr_plain:+.4f} dB")
        ctx.set_form(args.form)

    # The denoising configuration: the same clip and parameters at rank 4
    # (the headline's rank 16 equals the model rows, which keeps almost all
    # of the noise), fast and plain, for the speed-up at a quality that
    # differs from the noisy input.
    denoising = {}
    if args.form != "plain" and not args.no_denoising:
        den_args = dict(workload["args"], rank=4)
        dbp = parse_stage_parameters(BASIC, den_args)
        dfp = parse_stage_parameters(FINAL, den_args)

        def den_step():
            b = ctx.stage(BASIC, dbp, noisy, out=basic_out)
            return ctx.stage(FINAL, dfp, noisy, b, out=final_out)

        den_ms = {}
        den_out = {}
        for form in (args.form, "plain"):
            ctx.set_form(form)
            den_step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.denoising_steps):
                out = den_step()
            torch.cuda.synchronize()
            den_ms[form] = (time.perf_counter() - t0) * 1000.0 / args.denoising_steps
            den_out[form] = out.cpu().numpy()
        ctx.set_form(args.form)
        p_fast = synth.psnr(den_out[args.form], clean_np)
        p_plain = synth.psnr(den_out["plain"], clean_np)
        denoising = {"rank4_denoising": {
            "rank": 4,
            "steps": args.denoising_steps,
            "mpix_s": round(pixels / (den_ms[args.form] / 1000.0) / 1e6, 4),
            "plain_mpix_s": round(pixels / (den_ms["plain"] / 1000.0) / 1e6, 4),
            "fast_over_plain": round(den_ms["plain"] / den_ms[args.form], 4),
            "psnr_noisy_db": round(synth.psnr(noisy_np, clean_np), 4),
            "psnr_final_db": round(p_fast, 4),
            "psnr_final_plain_db": round(p_plain, 4),
            "dpsnr_vs_plain_db": round(p_fast - p_plain, 4),
        }}
        log(f"[bench] rank-4 denoising: {denoising['rank4_denoising']}")

    learned = {}
    if not args.no_learned and args.form != "plain":
        # The combined form: this form's model and aggregation with candidates
        # proposed by the learned network (bf16, and fp8 after calibration).
        log(f"[bench] training the candidate proposer ({args.proposer_steps} steps)")
        ctx.set_autotune(False)
        proposer, training = train_proposer(ctx, args.proposer_steps)
        log(f"[bench] proposer: {training}")
        learned = {"learned_training": training}
        for precision in ("bf16", "fp8"):
            log(f"[bench] combined form, {precision} proposals")
            entry = learned_form(ctx, proposer, step, bp, fp, noisy, clean_np, plain_np, pixels,
                                 args.learned_steps, args.learned_margin, precision)
            log(f"[bench] combined form {precision}: {entry}")
            learned[f"learned_{precision}"] = entry
        proposer.close()

    frame_cache = {}
    if args.frame_cache:
        # Frame-request rendering: each output frame from the cached anchor
        # stacks around it (every stack computed once over the pass).
        frames, channels, height, width = noisy.shape
        basic_frames = torch.empty_like(noisy)
        final_frames = torch.empty_like(noisy)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        misses = 0
        for stage, params, ref, dst in ((BASIC, bp, None, basic_frames), (FINAL, fp, basic_frames, final_frames)):
            cache = ctx.frame_cache(stage, params, frames, channels, height, width)
            for n in range(frames):
                cache.render(n, noisy, ref, out=dst[n])
            torch.cuda.synchronize()
            misses += cache.stats()["misses"]
            cache.close()
        cache_ms = (time.perf_counter() - t0) * 1000.0
        frame_cache = {
            "frame_cache": {
                "mpix_s": round(pixels / (cache_ms / 1000.0) / 1e6, 4),
                "ms": round(cache_ms, 3),
                "stacks_computed": misses,
                "max_abs_vs_stage": float(np.abs(final_frames.cpu().numpy() - final_np).max()),
            }
        }
        log(f"[bench] frame cache: {cache_ms:.1f} ms for Basic+Final, {misses} stacks")

    result = {
        "metric": "Mpix/s Basic+Final 1080p YUV444PS; dPSNR vs plain GPU form <=0.05 dB",
        "value": round(value, 4),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3/f32" if args.form == "fast" else "f32",
        "data": "synthetic (seeded gratings + rectangles, sub-pixel motion, Gaussian noise)",
        "config": {
            "workload": args.workload,
            "form": args.form,
            "width": clip["width"], "height": clip["height"], "frames": clip["frames"],
            "channels": clip["channels"], **workload["args"],
            "groups_basic": sb["groups"], "groups_final": sf["groups"],
            "psnr_noisy_db": round(synth.psnr(noisy_np, clean_np), 3),
            "psnr_final_db": round(psnr_final, 4),
            **quality,
            **denoising,
            **learned,
            **frame_cache,
            "kernel_ms": {k: round(v["ms"], 3) for k, v in times.items()},
            "retained_basic": sb["retained"], "retained_final": sf["retained"],
            "pipeline_parts": {"basic": sb["parts"], "final": sf["parts"]},
            "stage_graphs": not args.no_graph, "graph_replays": stats_f["graph_replays"],
            "rooflines": rooflines,
        },
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        result["cpu_baseline"] = cpu_baseline(workload)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
