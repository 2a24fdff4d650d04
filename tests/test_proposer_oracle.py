"""Pins of the proposer restatement (oracle/proposer_oracle.py).

The reference has no learned part, so nothing in it pins these functions;
they are pinned by mathematics instead: the analytic gradients against
central finite differences in float64, the softmax-gradient closed form,
the features against hand-built patches, and the bf16 rounding against
hand-worked bit patterns.
"""
import numpy as np
import pytest

from oracle import proposer_oracle as po


def _params(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in po.SHAPES:
        p[name] = rng.normal(0.0, scale, size=shape)
    return p


def _loss(params, x, t, group):
    cache = po.forward(params, x, emulate_bf16=False)
    return po.listwise_loss(cache["s"], t, group)[0]


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    group = 6
    rows = 3 * group
    x = rng.normal(0.0, 0.5, size=(rows, po.INPUT))
    t = np.zeros(rows)
    for g in range(rows // group):
        t[g * group + rng.integers(group)] = 0.75
        t[g * group + (rng.integers(group - 1) + 1) % group] += 0.25
    params = _params(2)
    cache = po.forward(params, x, emulate_bf16=False)
    _, ds = po.listwise_loss(cache["s"], t, group)
    grads = po.backward(params, cache, ds, emulate_bf16=False)
    h = 1e-6
    for name, shape in po.SHAPES:
        flat = params[name].reshape(-1)
        picks = rng.choice(flat.size, size=min(flat.size, 12), replace=False)
        for i in picks:
            keep = flat[i]
            flat[i] = keep + h
            up = _loss(params, x, t, group)
            flat[i] = keep - h
            down = _loss(params, x, t, group)
            flat[i] = keep
            numeric = (up - down) / (2 * h)
            analytic = grads[name].reshape(-1)[i]
            assert abs(numeric - analytic) <= 1e-6 + 1e-5 * abs(numeric), (name, i, numeric, analytic)


def test_softmax_gradient_closed_form():
    s = np.array([0.0, np.log(3.0), 0.0, 0.0])  # probabilities 1/6, 1/2, 1/6, 1/6
    t = np.array([0.0, 1.0, 0.0, 0.0])
    loss, g = po.listwise_loss(s, t, 4)
    assert loss == pytest.approx(np.log(2.0))
    np.testing.assert_allclose(g, [1 / 6, -0.5, 1 / 6, 1 / 6], atol=1e-15)
    # Two groups: each contributes half.
    loss2, g2 = po.listwise_loss(np.concatenate([s, s]), np.concatenate([t, t]), 4)
    assert loss2 == pytest.approx(np.log(2.0))
    np.testing.assert_allclose(g2[:4], g / 2, atol=1e-15)


def test_bf16_rounding_bit_patterns():
    one_plus = np.float32(1.0) + np.float32(2.0 ** -8)       # exactly half way: ties to even (1.0)
    assert po.bf16(np.array([one_plus]))[0] == np.float32(1.0)
    three_q = np.float32(1.0) + np.float32(3 * 2.0 ** -9)     # 1 + 0.75 ulp -> rounds up
    assert po.bf16(np.array([three_q]))[0] == np.float32(1.0 + 2.0 ** -7)
    odd_half = np.float32(1.0 + 2.0 ** -7 + 2.0 ** -8)        # half way from an odd mantissa: up
    assert po.bf16(np.array([odd_half]))[0] == np.float32(1.0 + 2.0 ** -6)
    assert po.bf16(np.array([-2.5]))[0] == np.float32(-2.5)


def test_box_means_and_features_on_hand_built_patches():
    clip = np.zeros((4, 1, 12, 12), np.float32)
    # Anchor patch at (1, 2) of frame 0: value 1 on its top-left 2x2 box,
    # its second frame (frame 1) all 0.25.  Candidate at (3, 1) of frame 2:
    # 0.5 everywhere, its second frame (frame 3) zero.
    clip[0, 0, 2:4, 1:3] = 1.0
    clip[1, 0, 2:10, 1:9] = 0.25
    clip[2, 0, 1:9, 3:11] = 0.5
    a = po.box_means(clip, 1, 2, 0, 2)
    assert a[0] == 1.0 and np.all(a[1:16] == 0.0) and np.all(a[16:] == 0.25)
    row = po.features(clip, 2, (1, 2, 0), (3, 1, 2))
    c = po.box_means(clip, 3, 1, 2, 2)
    assert np.all(c[:16] == 0.5) and np.all(c[16:] == 0.0)
    np.testing.assert_array_equal(row[:32], a - 0.5)
    np.testing.assert_array_equal(row[32:64], c - 0.5)
    np.testing.assert_array_equal(row[64:96], np.abs(a - c) * 4)
    assert row[96] == pytest.approx(2 / 16) and row[97] == pytest.approx(-1 / 16) and row[98] == 0.5
    assert row[99] == pytest.approx(4 * np.sum((a.astype(np.float64) - c) ** 2), rel=1e-6)
    assert np.all(row[100:] == 0.0)
    # Patch time 1: the second patch frame's slots stay zero.
    row1 = po.features(clip, 1, (1, 2, 0), (3, 1, 2))
    assert np.all(row1[16:32] == 0.0) and np.all(row1[48:64] == 0.0) and np.all(row1[80:96] == 0.0)
    # The proxy is the box-mean SSD.
    assert po.proxy_distance(clip, 2, (1, 2, 0), (3, 1, 2)) == pytest.approx(row[99] / 4, rel=1e-6)


def test_bf16_emulation_changes_only_operands():
    rng = np.random.default_rng(3)
    x = rng.normal(0.0, 0.3, size=(8, po.INPUT))
    params = _params(4)
    exact = po.forward(params, x, emulate_bf16=False)["s"]
    emulated = po.forward(params, x, emulate_bf16=True)["s"]
    # bf16 operands carry 8 significant bits: scores move by about 2^-8 of
    # their scale, not more.
    scale = np.abs(exact).max()
    assert np.abs(exact - emulated).max() <= 0.02 * scale
    assert np.abs(exact - emulated).max() > 0.0


def test_adam_first_step_moves_by_lr_times_sign():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -1e-3, 0.0])
    new, m, v = po.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(new[:2], p[:2] - 0.01 * np.sign(g[:2]), rtol=1e-6)
    assert new[2] == p[2]
