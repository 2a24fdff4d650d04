"""Learned candidate proposal: training data, training and the proposer handle.

The reference scans every candidate of a search window
(find_similar_patches, src/core/core.cpp:1245-1362).  A proposer is a small
MLP (include/vnlbhip.h, csrc/proposer.h) that scores the window's candidates
so that only the best scored ones are rescored with exact distances.  It is
trained on the GPU by the library's own forward / backward / Adam kernels on
the exact k nearest that the plain (exhaustive fp32) matcher returns:

* positives: the exact K nearest of each anchor (targets 1 / K each);
* hard negatives: the exhaustive ranks skip + 1 .. skip + hard, from a
  plain matching run with group_size skip + hard (ranks K + 1 .. skip,
  skip = the proposal budget, may be proposed at no cost and stay out of
  the loss);
* random negatives: uniform candidates of the anchor's window (its
  temporal range and boundary-shifted spatial range).

Loss: the listwise softmax cross-entropy over each anchor's rows.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _abi
from .params import BASIC, StageParameters, VNLBError


def pack(x, y, frame):
    return (np.asarray(x, np.uint32) | (np.asarray(y, np.uint32) << 12) | (np.asarray(frame, np.uint32) << 24))


def unpack(p):
    p = np.asarray(p, np.uint32)
    return (p & 0xFFF).astype(np.int64), ((p >> 12) & 0xFFF).astype(np.int64), (p >> 24).astype(np.int64)


def search_windows(params: StageParameters, width: int, height: int, total_frames: int, anchors) -> np.ndarray:
    """[n, 6] int32: x_low, y_low, nx, ny, t_low, t_high of each packed anchor
    (the library's compute_temporal_range / compute_spatial_range)."""
    lib = _abi.load()
    a = np.ascontiguousarray(anchors, np.uint32)
    out = np.empty((a.size, 6), np.int32)
    rc = lib.vnlbhip_search_windows(ctypes.byref(_abi.to_struct(params)), int(width), int(height), int(total_frames),
                                    int(a.size), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc != 0:
        raise VNLBError("VNLB HIP: invalid search window arguments")
    return out


class Proposer:
    """The network's parameters, Adam state and inference copies on one GPU."""

    def __init__(self, device: int = 0, seed: int = 0):
        self._lib = _abi.load()
        self._handle = ctypes.c_void_p()
        rc = self._lib.vnlbhip_proposer_create(int(device), ctypes.c_uint64(int(seed)), ctypes.byref(self._handle))
        if rc != 0:
            raise RuntimeError(f"VNLB HIP: cannot create a proposer on device {device}")
        self.device = int(device)
        self.calibrated = False

    def close(self) -> None:
        if self._handle:
            self._lib.vnlbhip_proposer_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != 0:
            message = self._lib.vnlbhip_proposer_last_error(self._handle).decode()
            raise (VNLBError if rc == -1 else RuntimeError)(f"VNLB HIP: {message}")

    def _stream(self):
        import torch

        return ctypes.c_void_p(torch.cuda.current_stream(torch.device("cuda", self.device)).cuda_stream)

    def get_params(self) -> np.ndarray:
        out = np.empty(_abi.PROPOSER_PARAMS, np.float32)
        self._check(self._lib.vnlbhip_proposer_get_params(self._handle, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != _abi.PROPOSER_PARAMS:
            raise VNLBError(f"VNLB HIP: a proposer has {_abi.PROPOSER_PARAMS} parameters")
        self._check(self._lib.vnlbhip_proposer_set_params(self._handle, p.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))
        self.calibrated = False

    def features(self, clip, patch_time: int, anchors, candidates):
        """bf16 input rows [rows, 128] of the (anchor, candidate) pairs
        (packed positions, int32 GPU tensors) of a device clip."""
        import torch

        frames, channels, height, width = (int(s) for s in clip.shape)
        rows = int(anchors.numel())
        out = torch.empty((rows, _abi.PROPOSER_INPUT), dtype=torch.bfloat16, device=clip.device)
        self._check(self._lib.vnlbhip_proposer_features(
            self._handle, ctypes.c_void_p(clip.data_ptr()), width, height, frames, channels, int(patch_time),
            ctypes.c_void_p(anchors.data_ptr()), ctypes.c_void_p(candidates.data_ptr()), rows,
            ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def gradients(self, x, targets, group: int):
        """(flat gradient, per-group losses) of the listwise loss, GPU tensors."""
        import torch

        rows = int(x.shape[0])
        grads = torch.empty(_abi.PROPOSER_PARAMS, dtype=torch.float32, device=x.device)
        losses = torch.empty(rows // group, dtype=torch.float32, device=x.device)
        self._check(self._lib.vnlbhip_proposer_gradients(
            self._handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(targets.data_ptr()), rows, int(group),
            ctypes.c_void_p(grads.data_ptr()), ctypes.c_void_p(losses.data_ptr()), self._stream()))
        return grads, losses

    def train_step(self, x, targets, group: int, learning_rate: float, losses=None) -> None:
        self._check(self._lib.vnlbhip_proposer_train_step(
            self._handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(targets.data_ptr()), int(x.shape[0]),
            int(group), float(learning_rate), ctypes.c_void_p(losses.data_ptr()) if losses is not None else None,
            self._stream()))
        self.calibrated = False

    def scores(self, x, precision: str = "bf16"):
        import torch

        out = torch.empty(int(x.shape[0]), dtype=torch.float32, device=x.device)
        self._check(self._lib.vnlbhip_proposer_scores(
            self._handle, ctypes.c_void_p(x.data_ptr()), int(x.shape[0]), _abi.PRECISIONS[precision],
            ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def calibrate(self, x) -> None:
        """FP8 inference copies and scales from sample input rows."""
        self._check(self._lib.vnlbhip_proposer_calibrate(
            self._handle, ctypes.c_void_p(x.data_ptr()), int(x.shape[0]), self._stream()))
        self.calibrated = True


@dataclasses.dataclass
class Examples:
    """Training rows of a set of anchors: row r of anchor i is
    (anchors[i], candidates[i, r]) with target targets[i, r]."""
    clip: object          # device clip [F, C, H, W]
    patch_time: int
    anchors: object       # int32 GPU [A]
    candidates: object    # int32 GPU [A, group]
    targets: object       # float32 GPU [A, group]

    @property
    def group(self) -> int:
        return int(self.candidates.shape[1])

    def __len__(self) -> int:
        return int(self.anchors.shape[0])


def make_examples(context, clip, params: StageParameters, anchor_frames, skip: int = 48, hard: int = 64,
                  random: int = 48, seed: int = 0, max_anchors: int | None = None) -> Examples:
    """Labelled rows from the plain matcher on a device clip (Basic matching:
    channel 0).  params: the matching parameters (K = params.similar).  Rows
    per anchor: the K nearest (positives), the exhaustive ranks skip + 1 ..
    skip + hard (hard negatives; ranks K + 1 .. skip may sit among the
    top-`skip` proposals at no cost, so they are left out of the loss), and
    `random` uniform candidates of the window."""
    import torch

    rng = np.random.default_rng(seed)
    frames, channels, height, width = (int(s) for s in clip.shape)
    k = int(params.similar)
    exact = dataclasses.replace(params, similar_cap_factor=1.0, tau=0.0)
    wide = dataclasses.replace(exact, similar=skip + hard, rank=min(int(params.rank), skip + hard))
    all_anchors, all_cands = [], []
    saved = context.form
    context.set_form("plain")
    context.set_proposer(None)
    try:
        for f in anchor_frames:
            lists = context.match_lists(BASIC, wide, clip, None, int(f))
            a = lists["anchors"].cpu().numpy().view(np.uint32)
            counts = lists["counts"].cpu().numpy()
            ranked = lists["positions"].cpu().numpy().view(np.uint32)
            keep = counts >= skip + hard
            a = a[keep]
            pos = np.concatenate([ranked[keep, :k], ranked[keep, skip:skip + hard]], axis=1)
            windows = search_windows(params, width, height, frames, a)
            n = a.size
            ti = rng.integers(0, np.iinfo(np.int32).max, size=(n, random)) % (windows[:, 5:6] - windows[:, 4:5] + 1)
            ry = rng.integers(0, np.iinfo(np.int32).max, size=(n, random)) % windows[:, 3:4]
            rx = rng.integers(0, np.iinfo(np.int32).max, size=(n, random)) % windows[:, 2:3]
            rand = pack(windows[:, 0:1] + rx, windows[:, 1:2] + ry, windows[:, 4:5] + ti)
            all_anchors.append(a)
            all_cands.append(np.concatenate([pos, rand], axis=1))
    finally:
        context.set_form(saved)
    anchors = np.concatenate(all_anchors)
    cands = np.concatenate(all_cands)
    if max_anchors is not None and anchors.size > max_anchors:
        pick = rng.choice(anchors.size, size=max_anchors, replace=False)
        anchors, cands = anchors[pick], cands[pick]
    # Random negatives that coincide with one of the K nearest are replaced
    # by the last hard negative (exhaustive rank skip + hard), not a positive.
    positives = cands[:, :k]
    tail = cands[:, k + hard:]
    for j in range(k):
        clash = tail == positives[:, j:j + 1]
        tail[clash] = np.broadcast_to(cands[:, k + hard - 1:k + hard], tail.shape)[clash]
    cands[:, k + hard:] = tail
    targets = np.zeros(cands.shape, np.float32)
    targets[:, :k] = 1.0 / k
    dev = clip.device
    return Examples(clip=clip, patch_time=int(params.patch_time),
                    anchors=torch.from_numpy(anchors.view(np.int32).copy()).to(dev),
                    candidates=torch.from_numpy(cands.view(np.int32).copy()).to(dev),
                    targets=torch.from_numpy(targets).to(dev))


def batch_rows(proposer: Proposer, examples: Examples, index):
    """(inputs, targets) of the anchors `index` (a GPU index tensor)."""
    group = examples.group
    anchors = examples.anchors[index].repeat_interleave(group)
    cands = examples.candidates[index].reshape(-1).contiguous()
    x = proposer.features(examples.clip, examples.patch_time, anchors.contiguous(), cands)
    return x, examples.targets[index].reshape(-1).contiguous()


def train(proposer: Proposer, examples: list, steps: int, batch_anchors: int = 64, learning_rate: float = 1e-3,
          seed: int = 0, log=None) -> list:
    """Adam steps over random batches of anchors drawn from the example
    sets; returns the mean loss of each step (read back every 50 steps)."""
    import torch

    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    sizes = np.array([len(e) for e in examples], np.float64)
    history = []
    group = examples[0].group
    losses = torch.empty(batch_anchors, dtype=torch.float32, device=examples[0].clip.device)
    for step in range(steps):
        which = int(torch.multinomial(torch.from_numpy(sizes / sizes.sum()), 1, generator=gen))
        ex = examples[which]
        index = torch.randint(0, len(ex), (batch_anchors,), generator=gen).to(ex.clip.device)
        x, t = batch_rows(proposer, ex, index)
        lr = learning_rate * (0.1 + 0.9 * 0.5 * (1.0 + np.cos(np.pi * step / max(1, steps))))
        proposer.train_step(x, t, group, lr, losses)
        if step % 50 == 0 or step == steps - 1:
            value = float(losses.mean())
            history.append((step, value))
            if log is not None:
                log(step, value)
    return history
