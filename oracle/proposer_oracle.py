"""CPU restatement of the learned candidate proposer (test infrastructure only).

Only tests/ and bench.py's checks use this module, as the checker: the
product's features, network, loss, gradients and optimizer run in HIP
kernels (vapoursynth-vnlb_amd/csrc/proposer.hip, match.hip).

The reference has no learned part; its block matching scans every
candidate of the window (src/core/core.cpp:1245-1362).  The proposer
scores the window's candidates with a small MLP so that only the best
scored ones are rescored with exact distances.  What this module fixes:

* features(): the network input of one (anchor, candidate) pair -- 2x2 box
  means of the 8x8 patches (4x4 per patch frame) of channel 0, the
  anchor's and the candidate's means centred on 0.5, their absolute
  differences x4, the offset (dx / 16, dy / 16, dt / 4), the box-mean SSD
  x4, zero padding to 128.  float32 in the kernel's operation order.
* forward() / backward(): s = w3 . relu(W2 relu(W1 x + b1) + b2) + b3 and
  the gradients of the listwise loss L = mean_g -sum_r t_r log softmax(s_g)_r
  over groups g of rows (one anchor's candidates each, targets t summing to
  1 per group).  With emulate_bf16 the operands of every matrix product
  are rounded to bfloat16 (round to nearest even) exactly where the kernels
  round them; products and sums are float64.
* adam(): the optimizer step.

Pins (tests/test_proposer_oracle.py): backward() against central finite
differences of forward() + loss in float64, the softmax gradient's closed
form, features() against a direct numpy restatement on hand-built patches,
and the zero-padding / scaling invariants.
"""
from __future__ import annotations

import numpy as np

INPUT = 128
HIDDEN = 256
PATCH = 8
DS = 4  # box means per patch side

# Parameter layout of the flat vector (kernels use the same order).
SHAPES = (("W1", (HIDDEN, INPUT)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
          ("w3", (HIDDEN,)), ("b3", (1,)))
PARAM_COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(flat: np.ndarray) -> dict:
    out = {}
    at = 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[at:at + n]).reshape(shape)
        at += n
    return out


def flatten(params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x) -> np.ndarray:
    """float32 -> bfloat16 (round to nearest, ties to even) -> float32."""
    a = np.ascontiguousarray(x, np.float32)
    bits = a.view(np.uint32).astype(np.uint64)
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return rounded.view(np.float32).reshape(a.shape)


def box_means(clip: np.ndarray, x: int, y: int, frame: int, patch_time: int) -> np.ndarray:
    """2x2 box means of the 8x8 patch at (x, y) of channel 0, per patch
    frame: float32 [patch_time * 16], index t * 16 + i * 4 + j, each
    ((p00 + p01) + (p10 + p11)) * 0.25 with p_ab = v(x + 2j + b, y + 2i + a)."""
    f32 = np.float32
    out = np.empty(patch_time * DS * DS, np.float32)
    for t in range(patch_time):
        plane = clip[frame + t, 0]
        for i in range(DS):
            for j in range(DS):
                p00 = f32(plane[y + 2 * i, x + 2 * j])
                p01 = f32(plane[y + 2 * i, x + 2 * j + 1])
                p10 = f32(plane[y + 2 * i + 1, x + 2 * j])
                p11 = f32(plane[y + 2 * i + 1, x + 2 * j + 1])
                out[t * 16 + i * 4 + j] = f32(f32(f32(p00 + p01) + f32(p10 + p11)) * f32(0.25))
    return out


def features(clip: np.ndarray, patch_time: int, anchor, candidate) -> np.ndarray:
    """Input row of one (anchor, candidate) pair; anchor / candidate are
    (x, y, frame).  clip: float32 [frames][channels][height][width]."""
    f32 = np.float32
    ax, ay, af = anchor
    cx, cy, cf = candidate
    a = box_means(clip, ax, ay, af, patch_time)
    c = box_means(clip, cx, cy, cf, patch_time)
    n = patch_time * 16
    row = np.zeros(INPUT, np.float32)
    half = f32(0.5)
    ssd = f32(0.0)
    for k in range(n):
        diff = f32(a[k] - c[k])
        row[k] = f32(a[k] - half)
        row[32 + k] = f32(c[k] - half)
        row[64 + k] = f32(abs(diff) * f32(4.0))
        ssd = f32(ssd + f32(diff * diff))
    row[96] = f32(cx - ax) * f32(1.0 / 16.0)
    row[97] = f32(cy - ay) * f32(1.0 / 16.0)
    row[98] = f32(cf - af) * f32(0.25)
    row[99] = f32(ssd * f32(4.0))
    return row


def proxy_distance(clip: np.ndarray, patch_time: int, anchor, candidate) -> float:
    """Box-mean SSD: the non-learned proposal score (lower is better)."""
    a = box_means(clip, anchor[0], anchor[1], anchor[2], patch_time).astype(np.float64)
    c = box_means(clip, candidate[0], candidate[1], candidate[2], patch_time).astype(np.float64)
    return float(np.sum((a - c) ** 2))


def forward(params: dict, x: np.ndarray, emulate_bf16: bool = True) -> dict:
    """Scores and the activations backward() needs; x [rows][128]."""
    rnd = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    x64 = rnd(x).astype(np.float64)
    W1 = rnd(params["W1"]).astype(np.float64)
    W2 = rnd(params["W2"]).astype(np.float64)
    z1 = x64 @ W1.T + np.asarray(params["b1"], np.float64)
    h1 = rnd(np.maximum(z1, 0.0).astype(np.float32 if emulate_bf16 else np.float64)).astype(np.float64)
    z2 = h1 @ W2.T + np.asarray(params["b2"], np.float64)
    h2 = np.maximum(z2, 0.0)
    s = h2 @ np.asarray(params["w3"], np.float64) + float(np.asarray(params["b3"]).ravel()[0])
    return {"x": x64, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "s": s, "W2": W2}


def listwise_loss(s: np.ndarray, targets: np.ndarray, group: int):
    """Mean over groups of -sum t log softmax(s); returns (loss, dL/ds)."""
    s = np.asarray(s, np.float64).reshape(-1, group)
    t = np.asarray(targets, np.float64).reshape(-1, group)
    m = s.max(axis=1, keepdims=True)
    e = np.exp(s - m)
    z = e.sum(axis=1, keepdims=True)
    logp = s - m - np.log(z)
    groups = s.shape[0]
    loss = float(-(t * logp).sum() / groups)
    grad = (e / z - t) / groups
    return loss, grad.ravel()


def backward(params: dict, cache: dict, ds: np.ndarray, emulate_bf16: bool = True) -> dict:
    """Gradients of all parameters given dL/ds."""
    rnd = bf16 if emulate_bf16 else (lambda v: np.asarray(v, np.float64))
    ds = np.asarray(ds, np.float64)
    w3 = np.asarray(params["w3"], np.float64)
    g = {}
    g["w3"] = cache["h2"].T @ ds
    g["b3"] = np.array([ds.sum()])
    dz2 = (ds[:, None] * w3[None, :]) * (cache["z2"] > 0)
    dz2_op = rnd(dz2.astype(np.float32) if emulate_bf16 else dz2).astype(np.float64)
    g["W2"] = dz2_op.T @ cache["h1"]
    g["b2"] = dz2.sum(axis=0)
    dh1 = dz2_op @ cache["W2"]
    dz1 = dh1 * (cache["z1"] > 0)
    dz1_op = rnd(dz1.astype(np.float32) if emulate_bf16 else dz1).astype(np.float64)
    g["W1"] = dz1_op.T @ cache["x"]
    g["b1"] = dz1.sum(axis=0)
    return g


def adam(params: np.ndarray, grads: np.ndarray, m: np.ndarray, v: np.ndarray, step: int, lr: float,
         beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
    """One Adam step (step counts from 1); returns new (params, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grads
    v = beta2 * v + (1.0 - beta2) * grads * grads
    mhat = m / (1.0 - beta1 ** step)
    vhat = v / (1.0 - beta2 ** step)
    return params - lr * mhat / (np.sqrt(vhat) + eps), m, v


def recall(proposed: np.ndarray, exact: np.ndarray) -> float:
    """Mean over anchors of |proposed_i & exact_i| / |exact_i| (rows are
    candidate ids; -1 entries are ignored)."""
    total = 0.0
    for p, e in zip(proposed, exact):
        e = set(int(v) for v in e if v >= 0)
        p = set(int(v) for v in p if v >= 0)
        total += len(e & p) / max(1, len(e))
    return total / max(1, len(exact))
