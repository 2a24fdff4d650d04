"""The learned candidate proposer on the GPU, through the C ABI.

* features, scores and gradients of the library's kernels against the CPU
  restatement (oracle/proposer_oracle.py, bf16 operands emulated);
* training on the plain matcher's exact k nearest lowers the loss, and the
  trained network's proposals recall the exact k nearest far better than
  the untrained network's and at least as well as the non-learned proposal
  (box-mean SSD, the proposer's own cheapest feature) at the same budget;
* the proposing matcher: exact distances bit for bit where it finds the
  exhaustive neighbours, the margin fallback gives the exhaustive lists,
  the whole stage stays within 0.05 dB of the plain form;
* fp8 inference after calibration tracks bf16.
"""
import numpy as np
import pytest
import torch

import vnlbhip
from oracle import proposer_oracle as po
from vnlbhip import proposer as prop
from vnlbhip import synth
from vnlbhip.params import BASIC, FINAL, parse_stage_parameters

pytestmark = pytest.mark.gpu

# Config 4's matching shape (GrayS, bm_range 12: 25 x 25 per temporal slot,
# radius 2, step 4, K 16, Pt 2) on small clips of config 2's kind.
ARGS = dict(sigma=20, block_size=8, block_step=4, group_size=16, bm_range=12, patch_time=2, radius=2, rank=4,
            tau=0)
TOP_K = 48
NEVER = -3.0e38  # margin below every gap: never fall back


def _clip(width, height, frames, seed):
    noisy, clean = synth.make_clip(width, height, frames, 1, sigma8=20, seed=seed, motion="subpixel")
    noisy = np.ascontiguousarray(noisy, np.float32)
    noisy = noisy[:, None] if noisy.ndim == 3 else noisy
    clean = clean[:, None] if clean.ndim == 3 else clean
    return noisy, clean


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


@pytest.fixture(scope="module")
def trained():
    """A proposer trained on two fresh clips, its untrained twin's
    parameters, and the loss history."""
    torch.manual_seed(0)
    params = parse_stage_parameters(BASIC, ARGS)
    ctx = vnlbhip.Context(0)
    proposer = prop.Proposer(0, seed=1)
    initial = proposer.get_params()
    examples = []
    for seed in (101, 102, 103):
        clip, _ = _clip(192, 144, 6, seed)
        x = _dev(clip)
        # Basic matches the noisy clip, Final the Basic estimate: examples of both.
        examples.append(prop.make_examples(ctx, x, params, anchor_frames=(1, 2, 3), seed=seed))
        basic = ctx.stage(BASIC, params, x)
        examples.append(prop.make_examples(ctx, basic, params, anchor_frames=(1, 2, 3), seed=seed + 50))
    print(f"[proposer] examples {[len(e) for e in examples]}", flush=True)
    history = prop.train(proposer, examples, steps=1500, batch_anchors=64, learning_rate=2e-3,
                         log=lambda s, v: print(f"[proposer] step {s} loss {v:.4f}", flush=True))
    yield proposer, initial, history, ctx
    ctx.close()
    proposer.close()


def _params_dict(flat):
    return po.unflatten(np.asarray(flat, np.float64))


def test_features_scores_and_gradients_match_the_oracle():
    params = parse_stage_parameters(BASIC, ARGS)
    ctx = vnlbhip.Context(0)
    proposer = prop.Proposer(0, seed=7)
    try:
        clip, _ = _clip(96, 80, 6, 55)
        x_dev = _dev(clip)
        ex = prop.make_examples(ctx, x_dev, params, anchor_frames=(2,), seed=5, max_anchors=8)
        index = torch.arange(8, device="cuda:0")
        x, t = prop.batch_rows(proposer, ex, index)
        group = ex.group
        # Features: the kernel's float arithmetic is the oracle's, so the
        # bf16 rows agree exactly.
        anchors = ex.anchors.cpu().numpy().view(np.uint32)
        cands = ex.candidates.cpu().numpy().view(np.uint32)
        xs = x.float().cpu().numpy()
        for i in range(8):
            ax, ay, af = prop.unpack(anchors[i])
            for r in range(0, group, 7):
                cx, cy, cf = prop.unpack(cands[i, r])
                want = po.bf16(po.features(clip, 2, (int(ax), int(ay), int(af)), (int(cx), int(cy), int(cf))))
                np.testing.assert_array_equal(xs[i * group + r], want, err_msg=f"row {i} {r}")
        flat = proposer.get_params()
        pd = _params_dict(flat)
        # Scores of the inference tile.
        s = proposer.scores(x).cpu().numpy()
        cache = po.forward(pd, xs, emulate_bf16=True)
        scale = np.abs(cache["s"]).max()
        assert np.abs(s - cache["s"]).max() <= 2e-3 * scale + 1e-6
        # Gradients.
        grads, losses = proposer.gradients(x, t, group)
        torch.cuda.synchronize()
        loss, ds = po.listwise_loss(cache["s"], t.cpu().numpy(), group)
        np.testing.assert_allclose(losses.cpu().numpy().mean(), loss, rtol=1e-4)
        want = po.backward(pd, cache, ds, emulate_bf16=True)
        got = po.unflatten(grads.cpu().numpy().astype(np.float64))
        for name, _ in po.SHAPES:
            g, w = got[name], want[name]
            if name == "b3":
                # sum over a group of softmax - targets is 0: only rounding is left.
                assert abs(float(g[0])) <= 1e-6 and abs(float(w[0])) <= 1e-12
                continue
            err = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-30)
            print(f"[proposer] grad {name}: rel err {err:.2e}", flush=True)
            assert err <= 5e-3, name
        # One Adam step moves the parameters as the oracle's step does.
        proposer.train_step(x, t, group, 1e-3)
        new = proposer.get_params().astype(np.float64)
        m0 = np.zeros(po.PARAM_COUNT)
        expect, _, _ = po.adam(flat.astype(np.float64), grads.cpu().numpy().astype(np.float64), m0, m0, 1, 1e-3)
        np.testing.assert_allclose(new, expect, rtol=0, atol=1e-6)
    finally:
        proposer.close()
        ctx.close()


def _recall(got, want, k=16):
    g = got["positions"][:, :k].cpu().numpy()
    w = want["positions"][:, :k].cpu().numpy()
    hits = [(len(set(a) & set(b))) / k for a, b in zip(g, w)]
    return float(np.mean(hits))


def _box_image(clip):
    """[F - 1, H - 7, W - 7, 32] box means of every 8x8 patch (Pt 2)."""
    c = clip[:, 0].astype(np.float32)
    f, h, w = c.shape
    q = ((c[:, :-1, :-1] + c[:, :-1, 1:]) + (c[:, 1:, :-1] + c[:, 1:, 1:])) * np.float32(0.25)
    out = np.empty((f - 1, h - 7, w - 7, 32), np.float32)
    for t in range(2):
        for i in range(4):
            for j in range(4):
                out[..., t * 16 + i * 4 + j] = q[t:f - 1 + t, 2 * i:2 * i + h - 7, 2 * j:2 * j + w - 7]
    return out


def _proxy_recall(clip, params, lists, k=16, budget=TOP_K, sample=300):
    """Recall of the non-learned proposal: the budget candidates of least
    box-mean SSD, then the exact k nearest among them."""
    frames, _, height, width = clip.shape
    boxes = _box_image(clip)
    anchors = lists["anchors"].cpu().numpy().view(np.uint32)
    exact = lists["positions"][:, :k].cpu().numpy().view(np.uint32)
    windows = prop.search_windows(params, width, height, frames, anchors)
    rng = np.random.default_rng(0)
    pick = rng.choice(anchors.size, size=min(sample, anchors.size), replace=False)
    hits = []
    c64 = clip[:, 0].astype(np.float64)
    for i in pick:
        ax, ay, af = (int(v) for v in prop.unpack(anchors[i]))
        x0, y0, nx, ny, t0, t1 = (int(v) for v in windows[i])
        fs, ys, xs = np.meshgrid(np.arange(t0, t1 + 1), np.arange(y0, y0 + ny), np.arange(x0, x0 + nx), indexing="ij")
        fs, ys, xs = fs.ravel(), ys.ravel(), xs.ravel()
        d = ((boxes[fs, ys, xs] - boxes[af, ay, ax]) ** 2).sum(axis=1)
        top = np.argsort(d, kind="stable")[:budget]
        ref = c64[af:af + 2, ay:ay + 8, ax:ax + 8]
        exact_d = [((c64[fs[j]:fs[j] + 2, ys[j]:ys[j] + 8, xs[j]:xs[j] + 8] - ref) ** 2).sum() for j in top]
        best = top[np.argsort(exact_d, kind="stable")[:k]]
        chosen = set(int(v) for v in prop.pack(xs[best], ys[best], fs[best]))
        hits.append(len(chosen & set(int(v) for v in exact[i])) / k)
    return float(np.mean(hits))


def test_training_gains_over_untrained_and_non_learned_proposals(trained):
    proposer, initial, history, ctx = trained
    first = history[0][1]
    last = float(np.mean([v for _, v in history[-5:]]))
    floor = np.log(16.0)  # the loss when the 16 positives share all the mass
    print(f"[proposer] loss {first:.4f} -> {last:.4f} (floor {floor:.4f})", flush=True)
    assert last - floor < 0.5 * (first - floor)
    params = parse_stage_parameters(BASIC, ARGS)
    clip, _ = _clip(128, 96, 6, 909)  # unseen clip
    x = _dev(clip)
    eval_ctx = vnlbhip.Context(0)
    untrained = prop.Proposer(0, seed=1)
    try:
        exact = eval_ctx.match_lists(BASIC, params, x, None, 2)
        eval_ctx.set_proposer(proposer, top_k=TOP_K, margin=NEVER)
        learned = eval_ctx.match_lists(BASIC, params, x, None, 2)
        stats = eval_ctx.stats()
        eval_ctx.set_proposer(untrained, top_k=TOP_K, margin=NEVER)
        blind = eval_ctx.match_lists(BASIC, params, x, None, 2)
        r_trained = _recall(learned, exact)
        r_untrained = _recall(blind, exact)
        r_proxy = _proxy_recall(clip, params, exact)
        print(f"[proposer] recall@16 with top {TOP_K}: trained {r_trained:.4f}, untrained {r_untrained:.4f}, "
              f"box-mean SSD {r_proxy:.4f}; fallbacks {stats['exhaustive_fallbacks']} of {stats['proposed']}",
              flush=True)
        assert stats["proposed"] == exact["anchors"].numel() and stats["exhaustive_fallbacks"] == 0
        assert r_trained >= 0.9
        assert r_trained >= r_untrained + 0.2
        assert r_trained >= r_proxy - 0.01
        # Where the proposals hold the exhaustive neighbours, the lists are
        # the exhaustive lists, distances bit for bit.
        same = (learned["positions"][:, :16] == exact["positions"][:, :16]).all(dim=1)
        assert int(same.sum()) > 0
        torch.testing.assert_close(learned["distances"][same, :16], exact["distances"][same, :16], rtol=0, atol=0)
    finally:
        untrained.close()
        eval_ctx.close()


def test_margin_fallback_is_the_exhaustive_search(trained):
    proposer, _, _, _ = trained
    params = parse_stage_parameters(BASIC, ARGS)
    clip, _ = _clip(96, 80, 6, 77)
    x = _dev(clip)
    ctx = vnlbhip.Context(0)
    try:
        exact = ctx.match_lists(BASIC, params, x, None, 2)
        ctx.set_proposer(proposer, top_k=TOP_K, margin=3.0e38)  # every margin is below: always fall back
        fallback = ctx.match_lists(BASIC, params, x, None, 2)
        stats = ctx.stats()
        assert stats["exhaustive_fallbacks"] == stats["proposed"] == exact["anchors"].numel()
        for key in ("anchors", "counts", "positions", "distances"):
            torch.testing.assert_close(fallback[key], exact[key], rtol=0, atol=0)
        # A finite margin: some anchors fall back, the rest are proposed.
        ctx.set_proposer(proposer, top_k=TOP_K, margin=1.0)
        ctx.match_lists(BASIC, params, x, None, 2)
        mid = ctx.stats()
        print(f"[proposer] margin 1.0: {mid['exhaustive_fallbacks']} of {mid['proposed']} anchors fall back",
              flush=True)
        assert 0 <= mid["exhaustive_fallbacks"] <= mid["proposed"]
    finally:
        ctx.close()


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_learned_stages_stay_within_the_plain_form(trained, form):
    proposer, _, _, _ = trained
    bp = parse_stage_parameters(BASIC, ARGS)
    fp = parse_stage_parameters(FINAL, ARGS)
    clip, clean = _clip(128, 96, 6, 313)
    x = _dev(clip)
    plain = vnlbhip.Context(0)
    learned = vnlbhip.Context(0)
    try:
        b0 = plain.stage(BASIC, bp, x)
        f0 = plain.stage(FINAL, fp, x, b0).cpu().numpy()
        p0 = synth.psnr(f0, clean)
        learned.set_form(form)
        # Recall of each stage's matching.
        for stage, params, ref in ((BASIC, bp, None), (FINAL, fp, b0)):
            exact = plain.match_lists(stage, params, x, ref, 2)
            learned.set_proposer(proposer, top_k=TOP_K, margin=NEVER)
            got = learned.match_lists(stage, params, x, ref, 2)
            print(f"[proposer] {form}: stage {stage} recall {_recall(got, exact):.4f}", flush=True)
        results = {}
        for margin in (NEVER, 0.0, 0.5, 1.0, 2.0):
            learned.set_proposer(proposer, top_k=TOP_K, margin=margin)
            b1 = learned.stage(BASIC, bp, x)
            sb = learned.stats()
            f1 = learned.stage(FINAL, fp, x, b1).cpu().numpy()
            sf = learned.stats()
            p1 = synth.psnr(f1, clean)
            share = (sb["exhaustive_fallbacks"] + sf["exhaustive_fallbacks"]) / max(1, sb["proposed"] + sf["proposed"])
            results[margin] = (p1, share)
            print(f"[proposer] {form}: margin {margin:g}: PSNR plain {p0:.3f}, learned {p1:.3f} dB "
                  f"(noisy {synth.psnr(clip, clean):.3f}); exhaustive fallbacks {share:.3f}", flush=True)
            assert sb["proposed"] > 0 and sf["proposed"] > 0
        assert any(abs(p1 - p0) <= 0.05 and share <= 0.5 for p1, share in results.values())
    finally:
        plain.close()
        learned.close()


def test_fp8_inference_tracks_bf16(trained):
    proposer, _, _, ctx = trained
    params = parse_stage_parameters(BASIC, ARGS)
    clip, _ = _clip(96, 80, 6, 606)
    x = _dev(clip)
    ex = prop.make_examples(ctx, x, params, anchor_frames=(2,), seed=3, max_anchors=64)
    rows, _ = prop.batch_rows(proposer, ex, torch.arange(64, device="cuda:0"))
    with pytest.raises(vnlbhip.VNLBError):
        ctx.set_proposer(proposer, top_k=TOP_K, margin=NEVER, precision="fp8")
    proposer.calibrate(rows)
    s16 = proposer.scores(rows).cpu().numpy()
    s8 = proposer.scores(rows, "fp8").cpu().numpy()
    corr = np.corrcoef(s16, s8)[0, 1]
    print(f"[proposer] fp8 vs bf16 score correlation {corr:.5f}", flush=True)
    assert corr >= 0.98
    eval_ctx = vnlbhip.Context(0)
    try:
        exact = eval_ctx.match_lists(BASIC, params, x, None, 2)
        eval_ctx.set_proposer(proposer, top_k=TOP_K, margin=NEVER)
        r16 = _recall(eval_ctx.match_lists(BASIC, params, x, None, 2), exact)
        eval_ctx.set_proposer(proposer, top_k=TOP_K, margin=NEVER, precision="fp8")
        r8 = _recall(eval_ctx.match_lists(BASIC, params, x, None, 2), exact)
        print(f"[proposer] recall bf16 {r16:.4f} fp8 {r8:.4f}", flush=True)
        assert r8 >= r16 - 0.03
    finally:
        eval_ctx.close()


def test_invalid_proposer_settings():
    ctx = vnlbhip.Context(0)
    proposer = prop.Proposer(0, seed=3)
    try:
        with pytest.raises(vnlbhip.VNLBError):
            ctx.set_proposer(proposer, top_k=0)
        with pytest.raises(vnlbhip.VNLBError):
            ctx.set_proposer(proposer, top_k=300)
        ctx.set_proposer(proposer, top_k=8, margin=NEVER)
        params = parse_stage_parameters(BASIC, ARGS)  # K = 16 > top_k
        clip, _ = _clip(64, 48, 4, 1)
        with pytest.raises(vnlbhip.VNLBError, match="top_k"):
            ctx.stage(BASIC, params, _dev(clip))
        bad = parse_stage_parameters(BASIC, dict(ARGS, block_size=6))
        ctx.set_proposer(proposer, top_k=TOP_K, margin=NEVER)
        with pytest.raises(vnlbhip.VNLBError, match="8x8"):
            ctx.stage(BASIC, bad, _dev(clip))
        with pytest.raises(vnlbhip.VNLBError):
            proposer.set_params(np.zeros(10, np.float32))
    finally:
        proposer.close()
        ctx.close()
