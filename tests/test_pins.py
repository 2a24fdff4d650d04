"""Independent pins of the oracle functions no reference fixture exercises:
aggregation weights (a21, src/core/core.cpp:184-294,326-412; the formula in
README.md:122-185), the aggregation window (a22, core.cpp:296-324) and the
staggered anchor grid (a5, core.cpp:934-967).

The pariasm fixtures run with weight_alpha = weight_beta = weight_gamma = 0,
so nothing there reaches these functions.  Each case below is worked out by
hand from the published formulas for inputs chosen so that a plausible
mistake (a swapped noise variance, a dropped volume or residual term, a
missing minimum shift, a wrong clamp, a phase taken from the local frame)
changes the expected number.  The library's own host copies of the window
and the grid (vnlbhip_aggregation_window, vnlbhip_anchor_grid) are held to the
same numbers; the weights computed on the GPU are held to the oracle by the
parity tests.
"""
import math

import numpy as np
import pytest

import vnlbhip
from vnlbhip.params import BASIC, FINAL, StageParameters, parse_stage_parameters

S8 = math.sqrt(8.0)  # sqrt(2 D) for D = 4


def _params(**kw):
    p = StageParameters()
    p.sigma = 1.0
    p.beta = 1.0
    p.weight_epsilon = 1e-6
    p.membership_noise_floor = 0.25
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _weights(oracle_lib, stage, p, rows, eig=None, basis=None, B=None, models=1, precision=1):
    return oracle_lib.aggregation_weights(stage, p, np.asarray(rows, float), eig, basis, B, models, precision)


# ---------------------------------------------------------------- a21 weights

def test_weights_disabled_are_one(oracle_lib):
    p = _params(weight_alpha=0.0, weight_beta=0.0, weight_gamma=0.0)
    rows = np.random.default_rng(0).normal(size=(5, 4))
    w = _weights(oracle_lib, BASIC, p, rows, [3.0], [[1, 0, 0, 0]], 5)
    np.testing.assert_array_equal(w, np.ones(5, np.float32))


def test_rank0_group_weight_is_the_mean_term_power(oracle_lib):
    # tau = eps + D / B with no signal components: w = (eps + 64/8)^-alpha.
    p = _params(weight_alpha=0.75, weight_beta=0.0, weight_gamma=1.0)
    rows = np.random.default_rng(1).normal(size=(12, 64))
    w = _weights(oracle_lib, BASIC, p, rows, B=8)
    np.testing.assert_allclose(w, np.full(12, (1e-6 + 8.0) ** -0.75), rtol=1e-6)
    # The default model of equal groups (no model added) gives the same.
    w0 = _weights(oracle_lib, BASIC, _params(weight_alpha=0.75, weight_beta=0.5, weight_gamma=1.0), rows, B=8,
                  models=0)
    np.testing.assert_allclose(w0, np.full(12, (1e-6 + 8.0) ** -0.75), rtol=1e-6)


def test_rank0_membership_is_the_shifted_norm_score(oracle_lib):
    # Z_i = (|x_i|^2 / sigma_m^2 - D) / sqrt(2D), sigma_m^2 = max(1, 0.25) = 1;
    # log w_i = -beta_w (Z_i - min Z) with alpha = 0.
    p = _params(weight_alpha=0.0, weight_beta=0.5, weight_gamma=1.0)
    rows = [[1, 0, 0, 0], [2, 0, 0, 0], [0, 0, 0, 3]]  # |x|^2 = 1, 4, 9
    want = [1.0, math.exp(-0.5 * 3 / S8), math.exp(-0.5 * 8 / S8)]
    for precision in (0, 1):
        np.testing.assert_allclose(_weights(oracle_lib, BASIC, p, rows, B=3, precision=precision), want,
                                   rtol=2e-6)


def test_rank1_basic_weights_by_hand(oracle_lib):
    # sigma = 1, beta = 1: model noise = estimate noise = 1; lambda = 5 ->
    # lambda^ = 4; sigma_m^2 = 1.  tau = eps + D/B + 4/5 = eps + 2.8.
    # volume = ln(1 + 4/1) / 4.  Row (2,0,0,0): p = 2, distance 4/(4+1) = 0.8;
    # row (0,1,0,0): residual 1, distance 1.  Z = (distance - 4) / sqrt(8).
    p = _params(weight_alpha=1.0, weight_beta=0.5, weight_gamma=1.0)
    rows = [[2, 0, 0, 0], [0, 1, 0, 0]]
    tau = 1e-6 + 2.0 + 0.8
    vol = math.log(5.0) / 4
    z0, z1 = (0.8 - 4) / S8, (1.0 - 4) / S8
    want = [math.exp(-math.log(tau) - 0.5 * vol), math.exp(-math.log(tau) - 0.5 * ((z1 - z0) + vol))]
    for precision in (0, 1):
        got = _weights(oracle_lib, BASIC, p, rows, [5.0], [[1, 0, 0, 0]], B=2, precision=precision)
        np.testing.assert_allclose(got, want, rtol=2e-6)
    # The row off the model direction ranks lower although its norm is smaller.
    assert want[1] < want[0]


def test_rank1_final_noise_terms_by_hand(oracle_lib):
    # Final: model noise beta * sigma_b^2 = 0.25 (lambda^ = 4.75), estimate
    # noise beta * sigma^2 = 1 (trace 4.75/5.75), membership noise
    # max(sigma_b^2, floor * sigma^2) = max(0.25, 0.1) = 0.25.
    p = _params(weight_alpha=1.0, weight_beta=0.5, weight_gamma=1.0, sigma_basic=0.5,
                membership_noise_floor=0.1)
    rows = [[2, 0, 0, 0], [0, 1, 0, 0]]
    tau = 1e-6 + 2.0 + 4.75 / 5.75
    vol = math.log(1 + 4.75 / 0.25) / 4
    z0 = (4 / (4.75 + 0.25) - 4) / S8
    z1 = (1 / 0.25 - 4) / S8
    want = [math.exp(-math.log(tau) - 0.5 * vol), math.exp(-math.log(tau) - 0.5 * ((z1 - z0) + vol))]
    got = _weights(oracle_lib, FINAL, p, rows, [5.0], [[1, 0, 0, 0]], B=2)
    np.testing.assert_allclose(got, want, rtol=2e-6)
    # Basic with the same numbers uses sigma for every noise term: different.
    basic = _weights(oracle_lib, BASIC, p, rows, [5.0], [[1, 0, 0, 0]], B=2)
    assert not np.allclose(basic, want, rtol=1e-3)


def test_log_weights_saturate_at_80(oracle_lib):
    rows = np.random.default_rng(2).normal(size=(6, 64))
    low = _weights(oracle_lib, BASIC, _params(weight_alpha=200.0, weight_beta=0.0, weight_gamma=1.0), rows, B=8)
    np.testing.assert_array_equal(low, np.full(6, np.float32(math.exp(-80.0))))
    # Covariance-path shapes (D = 4 < B = 64) give tau < 1: positive logs.
    rows = np.random.default_rng(3).normal(size=(64, 4))
    high = _weights(oracle_lib, BASIC, _params(weight_alpha=100.0, weight_beta=0.0, weight_gamma=1.0), rows,
                    B=64)
    np.testing.assert_array_equal(high, np.full(64, np.float32(math.exp(80.0))))


def test_weights_invariances(oracle_lib):
    """Scale-free (rows and sigma by c, eigenvalues by c^2), rotation-free
    (rows and basis rotated together), row-permutation equivariant; two equal
    models average to one."""
    rng = np.random.default_rng(4)
    S, D, R = 9, 6, 2
    rows = rng.normal(size=(S, D))
    q, _ = np.linalg.qr(rng.normal(size=(D, D)))
    basis = q[:R]
    eig = np.array([7.0, 3.0])
    p = _params(weight_alpha=0.8, weight_beta=0.6, weight_gamma=1.0, sigma=0.9)
    base = _weights(oracle_lib, BASIC, p, rows, eig, basis, B=S)
    c = 3.0
    scaled = _weights(oracle_lib, BASIC, _params(weight_alpha=0.8, weight_beta=0.6, weight_gamma=1.0, sigma=0.9 * c),
                      rows * c, eig * c * c, basis, B=S)
    np.testing.assert_allclose(scaled, base, rtol=1e-6)
    rot, _ = np.linalg.qr(rng.normal(size=(D, D)))
    rotated = _weights(oracle_lib, BASIC, p, rows @ rot, eig, basis @ rot, B=S)
    np.testing.assert_allclose(rotated, base, rtol=1e-6)
    perm = rng.permutation(S)
    np.testing.assert_allclose(_weights(oracle_lib, BASIC, p, rows[perm], eig, basis, B=S), base[perm], rtol=1e-6)
    np.testing.assert_allclose(_weights(oracle_lib, BASIC, p, rows, eig, basis, B=S, models=2), base, rtol=1e-6)
    # The row with the smallest score has membership = volume only: its weight
    # is the group weight times exp(-beta_w * volume).
    lam = np.maximum(eig - 0.81, 0)
    tau = 1e-6 + D / S + np.sum(lam / (lam + 0.81))
    vol = np.sum(np.log1p(lam / 0.81)) / D
    np.testing.assert_allclose(base.max(), math.exp(-0.8 * math.log(tau) - 0.6 * vol), rtol=1e-6)


# ----------------------------------------------------------------- a22 window

def _window_params(P, gamma):
    return _params(patch_size=P, weight_gamma=gamma)


@pytest.mark.parametrize("source", ["oracle", "library"])
def test_aggregation_window_closed_forms(oracle_lib, source):
    get = oracle_lib.aggregation_window if source == "oracle" else vnlbhip.aggregation_window
    # P = 3, gamma = 1: centre 1, edges exp(-1/2), corners exp(-1) (radius 1).
    w3 = get(_window_params(3, 1.0))
    e, c = math.exp(-0.5), math.exp(-1.0)
    np.testing.assert_allclose(w3, [[c, e, c], [e, 1, e], [c, e, c]], rtol=1e-6)
    # P = 8, gamma = 2: the four central samples are the maximum (1); a
    # corner is at squared distance 2, the centre samples at 2/49 (radius 3.5).
    w8 = get(_window_params(8, 2.0))
    assert w8[3, 3] == w8[3, 4] == w8[4, 3] == w8[4, 4] == 1.0
    np.testing.assert_allclose(w8[0, 0], math.exp(-(2 - 2 / 49)), rtol=1e-6)
    np.testing.assert_allclose(w8[0, 3], math.exp(-((3.5 ** 2 + 0.25) / 3.5 ** 2 - 2 / 49)), rtol=1e-6)
    np.testing.assert_array_equal(w8, w8.T)
    np.testing.assert_array_equal(w8, w8[::-1, ::-1])
    # P = 2: every sample is a corner, all equal after normalisation.
    np.testing.assert_array_equal(get(_window_params(2, 1.0)), np.ones((2, 2), np.float32))
    # gamma = 0 and P = 1: no window.
    np.testing.assert_array_equal(get(_window_params(8, 0.0)), np.ones((8, 8), np.float32))
    np.testing.assert_array_equal(get(_window_params(1, 1.0)), np.ones((1, 1), np.float32))


# ------------------------------------------------------------------- a5 grid

def _grid(source, oracle_lib, p, width, height, total, frame):
    if source == "oracle":
        return oracle_lib.grid_mask(p, width, height, total, frame).astype(bool)
    xs, ys = vnlbhip.anchor_grid(p, width, height, total, frame)
    mask = np.zeros((height, width), bool)
    mask[ys, xs] = True
    # Raster order.
    assert np.all(np.diff(ys.astype(np.int64) * width + xs) > 0)
    return mask


GRID_CASES = [(64, 48, 4, 2, 9), (71, 53, 3, 1, 7), (40, 40, 8, 8, 5), (37, 29, 5, 4, 6)]


@pytest.mark.parametrize("source", ["oracle", "library"])
@pytest.mark.parametrize("width,height,step,P,total", GRID_CASES)
def test_grid_invariants(oracle_lib, source, width, height, step, P, total):
    p = parse_stage_parameters(BASIC, dict(sigma=20, block_size=P, block_step=step, patch_time=1, radius=1))
    max_x, max_y = width - P, height - P
    masks = [_grid(source, oracle_lib, p, width, height, total, f) for f in range(total)]
    for f, m in enumerate(masks):
        assert not m[:, max_x + 1:].any() and not m[max_y + 1:].any()
        rows = np.flatnonzero(m.any(axis=1))
        assert rows[0] == 0 and rows[-1] == max_y
        for y in rows:
            assert m[y, 0] and m[y, max_x], (f, y)
        # Every pixel lies in some anchor's patch when step <= P.
        if step <= P:
            cover = np.zeros_like(m)
            for y, x in zip(*np.nonzero(m)):
                cover[y:y + P, x:x + P] = True
            assert cover.all(), f
    # The phase moves by one per frame: away from the border rows and
    # columns, frame f + 1 is frame f shifted down one row and right one
    # column, or two columns when the row phase wraps (its rows then start a
    # new block of `step` rows, which adds one to the column phase) -- except
    # into the last frame, whose phase is 0 whatever its index.
    y0, y1, x0, x1 = step + 1, max_y - step, 2 * step + 2, max_x - step
    for f in range(total - 2):
        dx = 2 if f % step == step - 1 else 1
        a = masks[f][y0:y1, x0:x1]
        b = masks[f + 1][y0 + 1:y1 + 1, x0 + dx:x1 + dx]
        np.testing.assert_array_equal(a, b, err_msg=f"frame {f}")
    last = masks[total - 1]
    first_row = np.flatnonzero(last[1:max_y].any(axis=1))
    assert (first_row[0] + 1) % step == 0          # phase_y = 0 on the last frame
    # Its row `step` is the second block of rows: column phase 1.
    assert step < 2 or (last[step, 1] and not last[step, 2])


@pytest.mark.parametrize("source", ["oracle", "library"])
def test_grid_counts_at_the_benchmark_shapes(oracle_lib, source):
    # Maximum anchors per frame over the phases (independently derived
    # figures for the 1080p step-4 and 256x256 step-8 configurations).
    p4 = parse_stage_parameters(BASIC, dict(sigma=20, block_size=8, block_step=4, patch_time=2, radius=2))
    counts = [int(_grid(source, oracle_lib, p4, 1920, 1080, 16, f).sum()) for f in (0, 1, 14, 15)]
    assert max(counts) == 129532 and counts[0] < counts[1]
    p8 = parse_stage_parameters(BASIC, dict(sigma=10, block_size=8, block_step=8, patch_time=1, radius=1))
    assert max(int(_grid(source, oracle_lib, p8, 256, 256, 5, f).sum()) for f in range(5)) == 1085


def test_library_grid_equals_oracle_grid(oracle_lib):
    for width, height, step, P, total in GRID_CASES:
        p = parse_stage_parameters(BASIC, dict(sigma=20, block_size=P, block_step=step, patch_time=2, radius=1))
        for f in range(total):
            np.testing.assert_array_equal(_grid("library", oracle_lib, p, width, height, total, f),
                                          _grid("oracle", oracle_lib, p, width, height, total, f))
