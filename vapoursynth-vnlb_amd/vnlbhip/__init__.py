"""vnlbhip: the VNLB Basic / Final stages on an MI355X.

Python mirror of the reference's VapourSynth interface for the hot path:
``Basic(clip, sigma, ...)`` and ``Final(clip, ref, sigma, ...)`` accept the
argument names and defaults of ``vnlb.Basic`` / ``vnlb.Final``
(src/vs/plugin.cpp:1162-1185) and, like the CUDA plugin ``vnlbcu``, return
the aggregated frames (``vnlb.Aggregate(vnlb.Basic(...), src)``).

A clip is a float32 array shaped (frames, channels, height, width) with
channels 1 (GrayS) or 3 (YUV444PS), or (frames, height, width) for GrayS.
numpy arrays are processed through host buffers; torch tensors on a ROCm
device stay in device memory and run on the current stream.
"""
from __future__ import annotations

import ctypes
import threading
import weakref

import numpy as np

from . import _abi
from .params import BASIC, FINAL, StageParameters, VNLBError, parse_stage_parameters

__all__ = ["Basic", "Final", "Context", "StageParameters", "VNLBError", "parse_stage_parameters",
           "BASIC", "FINAL", "anchor_grid", "aggregation_window"]

_ERROR_NAMES = {-1: "invalid argument", -2: "device error", -3: "model error"}

# VNLBHIP_SCRATCH_BYTES: the per-GPU scratch array of Context.stage_scratch.
SCRATCH_BYTES = 128 * 1024 * 1024


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _canonical_shape(shape) -> tuple[int, int, int, int]:
    if len(shape) == 3:
        return int(shape[0]), 1, int(shape[1]), int(shape[2])
    if len(shape) == 4:
        return tuple(int(s) for s in shape)  # type: ignore[return-value]
    raise VNLBError("VNLB HIP: only constant GrayS and YUV444PS clips are supported")


class Context:
    """One device context (not thread-safe; see the module-level helpers)."""

    def __init__(self, device: int = 0):
        self._lib = _abi.load()
        self._handle = ctypes.c_void_p()
        rc = self._lib.vnlbhip_create(int(device), ctypes.byref(self._handle))
        if rc != 0:
            raise RuntimeError(f"VNLB HIP: cannot create a context on device {device} ({_ERROR_NAMES.get(rc, rc)})")
        self.device = int(device)
        self.form = "plain"
        self._search_offsets = None
        self._proposer = None
        self._caches = weakref.WeakSet()

    def close(self) -> None:
        if self._handle:
            # Frame caches hold device stacks of this context: free them first.
            for cache in list(getattr(self, "_caches", ())):
                cache.close()
            self._lib.vnlbhip_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def _raise(self, rc: int) -> None:
        message = self._lib.vnlbhip_last_error(self._handle).decode()
        raise VNLBError(f"VNLB HIP: {message}") if rc in (-1, -3) else RuntimeError(f"VNLB HIP: {message}")

    def stats(self) -> dict:
        s = _abi.Stats()
        self._lib.vnlbhip_get_stats(self._handle, ctypes.byref(s))
        return {"anchors": s.anchors, "groups": s.groups,
                "unconverged_groups": s.unconverged_groups, "batches": s.batches,
                "retained": s.retained, "refined": s.refined, "parts": s.parts,
                "graph_replays": s.graph_replays, "reallocations": s.reallocations,
                "proposed": s.proposed, "exhaustive_fallbacks": s.exhaustive_fallbacks,
                "tile_shape": (s.tile_rows, s.tile_cols)}

    def set_proposer(self, proposer, top_k: int = 48, margin: float = 0.0, precision: str = "bf16") -> None:
        """Match through a learned candidate proposer (vnlbhip.proposer.Proposer;
        None restores exhaustive matching): only the top_k scored candidates of
        each window get exact distances, unless the score margin between the
        K-th and the (top_k + 1)-th best is below `margin` (then that window is
        scanned whole).  precision 'bf16' or 'fp8' (after calibration).  The
        proposer is kept alive while it is set."""
        if precision not in _abi.PRECISIONS:
            raise VNLBError(f"VNLB HIP: precision must be one of {sorted(_abi.PRECISIONS)}")
        handle = None if proposer is None else proposer._handle
        rc = self._lib.vnlbhip_set_proposer(self._handle, handle, int(top_k), float(margin),
                                            _abi.PRECISIONS[precision])
        if rc != 0:
            self._raise(rc)
        self._proposer = proposer

    def match_lists(self, stage: int, params: StageParameters, noisy, ref, anchor_frame: int) -> dict:
        """Match lists of one anchor frame (matching only) with the context's
        form and proposer: GPU tensors anchors [A] (packed x | y << 12 |
        frame << 24, int32 bits), counts [A], positions [A, cap] (packed),
        distances [A, cap]."""
        import torch

        frames, channels, height, width = _canonical_shape(tuple(noisy.shape))
        self._check_offsets(frames, height, width)
        if not _is_torch(noisy) or noisy.dtype != torch.float32 or not noisy.is_cuda:
            raise VNLBError("VNLB HIP: device clips must be float32 tensors on the GPU")
        if stage == FINAL and (ref is None or tuple(ref.shape) != tuple(noisy.shape)):
            raise VNLBError("VNLB HIP: ref clip properties must match input clip")
        p = _abi.to_struct(params)
        cap = int(self._lib.vnlbhip_retained_cap(ctypes.byref(p)))
        step = max(1, int(params.proc_step))
        capacity = (-(-width // step) + 2) * (-(-height // step) + 2)
        src = noisy.contiguous()
        guide = ref.contiguous() if ref is not None else None
        dev = src.device
        anchors = torch.empty(capacity, dtype=torch.int32, device=dev)
        counts = torch.empty(capacity, dtype=torch.int32, device=dev)
        positions = torch.empty((capacity, cap), dtype=torch.int32, device=dev)
        distances = torch.empty((capacity, cap), dtype=torch.float32, device=dev)
        count = ctypes.c_int(0)
        stream = torch.cuda.current_stream(dev).cuda_stream
        rc = self._lib.vnlbhip_match_lists(
            self._handle, stage, ctypes.byref(p), ctypes.c_void_p(src.data_ptr()),
            ctypes.c_void_p(guide.data_ptr()) if guide is not None else None, width, height, frames, channels,
            int(anchor_frame), capacity, ctypes.byref(count), ctypes.c_void_p(anchors.data_ptr()),
            ctypes.c_void_p(counts.data_ptr()), ctypes.c_void_p(positions.data_ptr()),
            ctypes.c_void_p(distances.data_ptr()), ctypes.c_void_p(stream))
        if rc != 0:
            self._raise(rc)
        a = count.value
        return {"anchors": anchors[:a], "counts": counts[:a], "positions": positions[:a],
                "distances": distances[:a]}

    def set_form(self, form: str) -> None:
        """'plain' (fp32 in the CPU path's order, the default) or 'fast'
        (candidate distances from bf16x3 products on the matrix cores)."""
        if form not in _abi.FORMS:
            raise VNLBError(f"VNLB HIP: form must be one of {sorted(_abi.FORMS)}")
        rc = self._lib.vnlbhip_set_form(self._handle, _abi.FORMS[form])
        if rc != 0:
            self._raise(rc)
        self.form = form

    def set_tile_shape(self, rows: int, cols: int) -> None:
        """Fast form: anchor tiles of the MFMA matcher (rows x cols anchors,
        1 to 4 each; default 4 x 4) for calls the autotuner does not tune."""
        rc = self._lib.vnlbhip_set_tile_shape(self._handle, int(rows), int(cols))
        if rc != 0:
            raise VNLBError("VNLB HIP: a tile has 1 .. 4 grid rows of 1 .. 4 anchors")

    def set_jacobi(self, max_sweeps: int = 64, fail_unconverged: bool = False) -> None:
        """Jacobi sweep limit of every eigen solve (default 64) and whether a
        problem that does not converge within it fails the stage (the CUDA
        plugin's behaviour) or is used and counted in
        stats()["unconverged_groups"] (the CPU path's, the default)."""
        rc = self._lib.vnlbhip_set_jacobi(self._handle, int(max_sweeps), 1 if fail_unconverged else 0)
        if rc != 0:
            raise VNLBError("VNLB HIP: max_sweeps must be in [1, 10000]")

    def set_eigen_refine(self, relative_gap: float) -> None:
        """Fast form: re-solve in fp64 the group models whose eigenvalue gap at
        the rank cutoff is below relative_gap times the largest eigenvalue
        (default 1e-3, 0 disables)."""
        rc = self._lib.vnlbhip_set_eigen_refine(self._handle, float(relative_gap))
        if rc != 0:
            raise VNLBError("VNLB HIP: relative_gap must be finite and non-negative")

    def set_graph(self, enable: bool) -> None:
        """Capture each stage's launch sequence into a hipGraph after an eager
        run and replay it on calls with the same stage, parameters, geometry
        and tensor addresses (pass out= to keep the output address)."""
        rc = self._lib.vnlbhip_set_graph(self._handle, int(enable))
        if rc != 0:
            self._raise(rc)

    def set_autotune(self, enable: bool) -> None:
        """Time the stage once per pipeline part count (1..4) on the first
        call for a stage / form / geometry / parameter set and keep the
        fastest (stats()["parts"] reports the count used)."""
        rc = self._lib.vnlbhip_set_autotune(self._handle, int(enable))
        if rc != 0:
            self._raise(rc)

    def set_fused_selection(self, enable: bool) -> None:
        """Fast form: select the k nearest inside the MFMA matcher (True,
        the default) or from staged distances in a second kernel (False);
        both give the same match lists."""
        rc = self._lib.vnlbhip_set_fused_selection(self._handle, int(enable))
        if rc != 0:
            self._raise(rc)

    def set_search_offsets(self, offsets) -> None:
        """Centre each scheduled frame's search window on the anchor plus a
        per-anchor offset (MVToolsFlowProvider::center_for,
        src/flow/flow.hpp:340-390, from a dense field): an int16 GPU tensor
        [frames, search_bwd + search_fwd + 1, height, width, 2] of (dx, dy)
        per anchor position and schedule index (the anchor frame, later
        frames ascending, earlier frames descending), or None to clear it.
        The tensor is kept alive while it is set."""
        import torch

        if offsets is None:
            rc = self._lib.vnlbhip_set_search_offsets(self._handle, None, 0, 0, 0, 0)
            self._search_offsets = None
        else:
            if (not offsets.is_cuda or offsets.dtype != torch.int16 or offsets.dim() != 5
                    or offsets.shape[-1] != 2 or not offsets.is_contiguous()):
                raise VNLBError("VNLB HIP: offsets must be a contiguous int16 GPU tensor "
                                "[frames, temporal, height, width, 2]")
            frames, temporal, height, width = (int(s) for s in offsets.shape[:4])
            rc = self._lib.vnlbhip_set_search_offsets(self._handle, ctypes.c_void_p(offsets.data_ptr()),
                                                      temporal, frames, height, width)
            self._search_offsets = offsets
        if rc != 0:
            self._raise(rc)

    def mvtools_offsets(self, params: StageParameters, frames: int, height: int, width: int,
                        mvfw=None, mvbw=None):
        """The search-offset field of set_search_offsets from MVTools vector
        clips (vnlbhip.flow.VectorClip; mvfw is walked towards earlier
        frames, mvbw towards later ones, as MVToolsFlowProvider::center_for
        does, src/flow/flow.hpp:340-390): an int16 GPU tensor [frames,
        search_bwd + search_fwd + 1, height, width, 2], computed on the GPU."""
        import torch

        device = torch.device("cuda", self.device)
        temporal = params.search_bwd + params.search_fwd + 1
        out = torch.empty((frames, temporal, height, width, 2), dtype=torch.int16, device=device)
        keep = []

        def clip_args(clip, name, backwards):
            if clip is None:
                return None, None, None
            clip.check(name, frames, width, height, backwards)
            shape = _abi.MvShape(**clip.analysis.grid_shape())
            vectors = torch.from_numpy(clip.vectors).to(device)
            valid = torch.from_numpy(clip.valid).to(device)
            keep.extend((shape, vectors, valid))
            return (ctypes.byref(shape), ctypes.c_void_p(vectors.data_ptr()),
                    ctypes.c_void_p(valid.data_ptr()))

        fw = clip_args(mvfw, "mvfw", False)
        bw = clip_args(mvbw, "mvbw", True)
        stream = torch.cuda.current_stream(device).cuda_stream
        rc = self._lib.vnlbhip_mvtools_offsets(
            self._handle, ctypes.byref(_abi.to_struct(params)), fw[0], fw[1], fw[2], bw[0], bw[1], bw[2],
            ctypes.c_void_p(out.data_ptr()), int(width), int(height), int(frames), ctypes.c_void_p(stream))
        if rc != 0:
            self._raise(rc)
        return out

    def frame_cache(self, stage: int, params: StageParameters, frames: int, channels: int, height: int,
                    width: int, capacity: int | None = None) -> "FrameCache":
        """A per-output-frame renderer over a device cache of `capacity`
        anchor stacks (default: the search_bwd + search_fwd + patch_time
        stacks one output frame needs)."""
        cache = FrameCache(self, stage, params, frames, channels, height, width, capacity)
        self._caches.add(cache)
        return cache

    def _check_offsets(self, frames: int, height: int, width: int) -> None:
        field = self._search_offsets
        if field is not None and (field.shape[0] != frames or tuple(field.shape[2:4]) != (height, width)):
            raise VNLBError(f"VNLB HIP: search offsets {tuple(field.shape)} do not cover a clip of "
                            f"{frames} frames of {width}x{height}")

    def set_profiling(self, enable: bool) -> None:
        self._lib.vnlbhip_set_profiling(self._handle, 1 if enable else 0)

    def kernel_times(self) -> dict:
        t = _abi.KernelTimes()
        self._lib.vnlbhip_get_kernel_times(self._handle, ctypes.byref(t))
        return {name: {"ms": t.milliseconds[i], "launches": t.launches[i]}
                for i, name in enumerate(_abi.KERNEL_NAMES)}

    def anchor_stack(self, stage: int, params: StageParameters, noisy, ref, anchor_frame: int):
        """Contribution stack of one anchor frame as a vnlb.Basic / vnlb.Final
        output frame holds it: a float32 GPU tensor [channels, slots * 2 *
        height, width], per slot a numerator plane then the weight plane."""
        import torch

        frames, channels, height, width = _canonical_shape(tuple(noisy.shape))
        self._check_offsets(frames, height, width)
        if not _is_torch(noisy) or noisy.dtype != torch.float32 or not noisy.is_cuda:
            raise VNLBError("VNLB HIP: device clips must be float32 tensors on the GPU")
        if stage == FINAL and (ref is None or tuple(ref.shape) != tuple(noisy.shape)):
            raise VNLBError("VNLB HIP: ref clip properties must match input clip")
        p = _abi.to_struct(params)
        slots = int(params.search_bwd) + int(params.search_fwd) + int(params.patch_time)
        src = noisy.contiguous()
        guide = ref.contiguous() if ref is not None else None
        stack = torch.empty((channels, slots * 2 * height, width), dtype=torch.float32, device=src.device)
        stream = torch.cuda.current_stream(src.device).cuda_stream
        rc = self._lib.vnlbhip_anchor_stack(
            self._handle, stage, ctypes.byref(p), ctypes.c_void_p(src.data_ptr()),
            ctypes.c_void_p(guide.data_ptr()) if guide is not None else None, width, height,
            frames, channels, int(anchor_frame), ctypes.c_void_p(stack.data_ptr()),
            ctypes.c_void_p(stream))
        if rc != 0:
            self._raise(rc)
        return stack

    def anchor_stack_window(self, stage: int, params: StageParameters, noisy, ref, first_frame: int,
                            total_frames: int, anchor_frame: int):
        """anchor_stack over a window of a longer clip: noisy / ref hold the
        frames [first_frame, first_frame + len(noisy)) of a clip of
        total_frames; anchor_frame is absolute (the reference's frame
        requests, src/vs/plugin.cpp:616-758)."""
        import torch

        frames, channels, height, width = _canonical_shape(tuple(noisy.shape))
        self._check_offsets(frames, height, width)
        if not _is_torch(noisy) or noisy.dtype != torch.float32 or not noisy.is_cuda:
            raise VNLBError("VNLB HIP: device clips must be float32 tensors on the GPU")
        if stage == FINAL and (ref is None or tuple(ref.shape) != tuple(noisy.shape)):
            raise VNLBError("VNLB HIP: ref clip properties must match input clip")
        p = _abi.to_struct(params)
        slots = int(params.search_bwd) + int(params.search_fwd) + int(params.patch_time)
        src = noisy.contiguous()
        guide = ref.contiguous() if ref is not None else None
        stack = torch.empty((channels, slots * 2 * height, width), dtype=torch.float32, device=src.device)
        stream = torch.cuda.current_stream(src.device).cuda_stream
        rc = self._lib.vnlbhip_anchor_stack_window(
            self._handle, stage, ctypes.byref(p), ctypes.c_void_p(src.data_ptr()),
            ctypes.c_void_p(guide.data_ptr()) if guide is not None else None, width, height,
            frames, channels, int(first_frame), int(total_frames), int(anchor_frame),
            ctypes.c_void_p(stack.data_ptr()), ctypes.c_void_p(stream))
        if rc != 0:
            self._raise(rc)
        return stack

    def reserve(self, stage: int, params: StageParameters, frames: int, channels: int, height: int,
                width: int) -> None:
        """Size every device buffer a stage of this shape needs in the
        current form, so that later calls of that shape allocate nothing
        (StagePipeline::reserve, src/cuda/pipeline.hpp:82-107)."""
        rc = self._lib.vnlbhip_reserve(self._handle, int(stage), ctypes.byref(_abi.to_struct(params)),
                                       int(width), int(height), int(frames), int(channels))
        if rc != 0:
            self._raise(rc)

    def stage_scratch(self, stage: int, params: StageParameters, noisy, ref=None, out=None):
        """stage() on device tensors without any device allocation inside
        the call (vnlbhip_stage_scratch): the library's buffers come from a
        fixed per-GPU scratch array of SCRATCH_BYTES; larger stages raise.
        Calls that overlap on one GPU race."""
        import torch

        frames, channels, height, width = _canonical_shape(tuple(noisy.shape))
        self._check_offsets(frames, height, width)
        if not _is_torch(noisy) or noisy.dtype != torch.float32 or not noisy.is_cuda:
            raise VNLBError("VNLB HIP: device clips must be float32 tensors on the GPU")
        if stage == FINAL and (ref is None or tuple(ref.shape) != tuple(noisy.shape)):
            raise VNLBError("VNLB HIP: ref clip properties must match input clip")
        if out is None:
            out = torch.empty_like(noisy)
        for t in (noisy, ref, out):
            if t is not None and not t.is_contiguous():
                raise VNLBError("VNLB HIP: stage_scratch takes contiguous tensors (no copies are made)")
        if tuple(out.shape) != tuple(noisy.shape) or out.dtype != torch.float32 or out.device != noisy.device:
            raise VNLBError("VNLB HIP: out must be a contiguous float32 tensor like the clip")
        p = _abi.to_struct(params)
        stream = torch.cuda.current_stream(noisy.device).cuda_stream
        rc = self._lib.vnlbhip_stage_scratch(
            self._handle, stage, ctypes.byref(p), ctypes.c_void_p(noisy.data_ptr()),
            ctypes.c_void_p(ref.data_ptr()) if ref is not None else None,
            ctypes.c_void_p(out.data_ptr()), width, height, frames, channels, ctypes.c_void_p(stream))
        if rc != 0:
            self._raise(rc)
        return out

    def stage(self, stage: int, params: StageParameters, noisy, ref=None, out=None):
        """Run one stage over a whole clip and return the aggregated frames
        (device clips: into out when given, a float32 tensor of the clip's
        shape on the same device)."""
        frames, channels, height, width = _canonical_shape(tuple(noisy.shape))
        self._check_offsets(frames, height, width)
        if channels not in (1, 3):
            raise VNLBError("VNLB HIP: only constant GrayS and YUV444PS clips are supported")
        if stage == FINAL:
            if ref is None:
                raise VNLBError("VNLB HIP: ref must be a video node")
            if tuple(ref.shape) != tuple(noisy.shape):
                raise VNLBError("VNLB HIP: ref clip properties must match input clip")
        p = _abi.to_struct(params)
        if _is_torch(noisy):
            import torch

            if noisy.dtype != torch.float32 or not noisy.is_cuda:
                raise VNLBError("VNLB HIP: device clips must be float32 tensors on the GPU")
            src = noisy.contiguous()
            guide = ref.contiguous() if ref is not None else None
            if out is None:
                out = torch.empty_like(src)
            elif (tuple(out.shape) != tuple(src.shape) or out.dtype != torch.float32
                  or out.device != src.device or not out.is_contiguous()):
                raise VNLBError("VNLB HIP: out must be a contiguous float32 tensor like the clip")
            stream = torch.cuda.current_stream(src.device).cuda_stream
            rc = self._lib.vnlbhip_stage(
                self._handle, stage, ctypes.byref(p), ctypes.c_void_p(src.data_ptr()),
                ctypes.c_void_p(guide.data_ptr()) if guide is not None else None,
                ctypes.c_void_p(out.data_ptr()), width, height, frames, channels,
                ctypes.c_void_p(stream))
            if rc != 0:
                self._raise(rc)
            return out
        src = np.ascontiguousarray(noisy, dtype=np.float32)
        guide = np.ascontiguousarray(ref, dtype=np.float32) if ref is not None else None
        out = np.empty_like(src)
        fptr = ctypes.POINTER(ctypes.c_float)
        rc = self._lib.vnlbhip_stage_host(
            self._handle, stage, ctypes.byref(p), src.ctypes.data_as(fptr),
            guide.ctypes.data_as(fptr) if guide is not None else None,
            out.ctypes.data_as(fptr), width, height, frames, channels)
        if rc != 0:
            self._raise(rc)
        return out

    def debug_frame(self, stage: int, params: StageParameters, noisy, ref, anchor_frame: int,
                    gram_dim: int = 16) -> dict:
        """Per-anchor diagnostics of one anchor frame (device tensors in); grams
        holds the dual Gram of each modelled dual group with at most gram_dim
        model rows (NaN elsewhere)."""
        import torch

        frames, channels, height, width = _canonical_shape(tuple(noisy.shape))
        self._check_offsets(frames, height, width)
        p = _abi.to_struct(params)
        cap = self._lib.vnlbhip_retained_cap(ctypes.byref(p))
        if cap <= 0:
            raise VNLBError("VNLB HIP: invalid parameters")
        capacity = width * height
        ints = lambda n: np.zeros(n, dtype=np.int32)  # noqa: E731
        num = ctypes.c_int(0)
        ax, ay, counts = ints(capacity), ints(capacity), ints(capacity)
        positions = np.full(capacity * cap, -1, dtype=np.int32)
        distances = np.full(capacity * cap, np.nan, dtype=np.float32)
        active = np.zeros(capacity, dtype=np.uint8)
        weights = np.full(capacity * cap, np.nan, dtype=np.float32)
        eig = np.full(capacity * max(params.rank, 1), np.nan, dtype=np.float32)
        flags = ints(capacity)
        grams = np.full(capacity * gram_dim * gram_dim, np.nan, dtype=np.float32)
        src = noisy.contiguous()
        guide = ref.contiguous() if ref is not None else None
        ip = ctypes.POINTER(ctypes.c_int)
        fp = ctypes.POINTER(ctypes.c_float)
        rc = self._lib.vnlbhip_debug_frame(
            self._handle, stage, ctypes.byref(p), ctypes.c_void_p(src.data_ptr()),
            ctypes.c_void_p(guide.data_ptr()) if guide is not None else None,
            width, height, frames, channels, anchor_frame, capacity, ctypes.byref(num),
            ax.ctypes.data_as(ip), ay.ctypes.data_as(ip), counts.ctypes.data_as(ip),
            positions.ctypes.data_as(ip), distances.ctypes.data_as(fp),
            active.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), weights.ctypes.data_as(fp),
            eig.ctypes.data_as(fp), flags.ctypes.data_as(ip), gram_dim, grams.ctypes.data_as(fp))
        if rc != 0:
            self._raise(rc)
        torch.cuda.synchronize()
        n = num.value
        return {
            "cap": cap,
            "anchor_x": ax[:n], "anchor_y": ay[:n], "counts": counts[:n],
            "positions": positions[: n * cap].reshape(n, cap),
            "distances": distances[: n * cap].reshape(n, cap),
            "active": active[:n], "weights": weights[: n * cap].reshape(n, cap),
            "eigenvalues": eig[: n * max(params.rank, 1)].reshape(n, max(params.rank, 1)),
            "group_flags": flags[:n],
            "grams": grams[: n * gram_dim * gram_dim].reshape(n, gram_dim, gram_dim),
        }


class FrameCache:
    """Output frames one at a time, as a VapourSynth frame request makes them
    (vnlb.Basic / vnlb.Final stacks aggregated by vnlb.Aggregate), with each
    anchor frame's stack computed once while it stays in the device cache."""

    def __init__(self, context: Context, stage: int, params: StageParameters, frames: int, channels: int,
                 height: int, width: int, capacity: int | None = None):
        self._context = context
        self._lib = context._lib
        self._shape = (frames, channels, height, width)
        if capacity is None:
            capacity = params.search_bwd + params.search_fwd + params.patch_time
        self._handle = ctypes.c_void_p()
        rc = self._lib.vnlbhip_frame_cache_create(context._handle, int(stage), ctypes.byref(_abi.to_struct(params)),
                                                  int(width), int(height), int(frames), int(channels),
                                                  int(capacity), ctypes.byref(self._handle))
        if rc != 0:
            context._raise(rc)

    def render(self, frame: int, noisy, ref=None, out=None):
        """Output frame `frame` [channels, height, width] of the clip
        (a float32 GPU tensor [frames, channels, height, width]; the same
        tensors on every call keep the cache)."""
        import torch

        frames, channels, height, width = self._shape
        self._context._check_offsets(frames, height, width)
        src = noisy if noisy.dim() == 4 else noisy.unsqueeze(1)
        if tuple(src.shape) != self._shape or src.dtype != torch.float32 or not src.is_cuda:
            raise VNLBError(f"VNLB HIP: the clip must be a float32 GPU tensor of shape {self._shape}")
        src = src.contiguous()
        guide = None
        if ref is not None:
            guide = (ref if ref.dim() == 4 else ref.unsqueeze(1)).contiguous()
        if out is None:
            out = torch.empty((channels, height, width), dtype=torch.float32, device=src.device)
        stream = torch.cuda.current_stream(src.device).cuda_stream
        rc = self._lib.vnlbhip_frame_cache_render(
            self._handle, int(frame), ctypes.c_void_p(src.data_ptr()),
            ctypes.c_void_p(guide.data_ptr()) if guide is not None else None,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
        if rc != 0:
            self._context._raise(rc)
        return out

    def render_window(self, frame: int, noisy, ref, first_frame: int, out=None):
        """Output frame `frame` from a window of the clip: noisy / ref hold the
        frames [first_frame, first_frame + len(noisy)) (float32 GPU tensors),
        which must include `frame` and the input frames of every stack the
        call computes.  Stacks are cached by frame number."""
        import torch

        frames, channels, height, width = self._shape
        src = noisy if noisy.dim() == 4 else noisy.unsqueeze(1)
        window = int(src.shape[0])
        self._context._check_offsets(window, height, width)
        if (tuple(src.shape[1:]) != (channels, height, width) or src.dtype != torch.float32
                or not src.is_cuda):
            raise VNLBError("VNLB HIP: the window must be a float32 GPU tensor of the clip's frame shape")
        src = src.contiguous()
        guide = None
        if ref is not None:
            guide = (ref if ref.dim() == 4 else ref.unsqueeze(1)).contiguous()
            if tuple(guide.shape) != tuple(src.shape):
                raise VNLBError("VNLB HIP: ref clip properties must match input clip")
        if out is None:
            out = torch.empty((channels, height, width), dtype=torch.float32, device=src.device)
        stream = torch.cuda.current_stream(src.device).cuda_stream
        rc = self._lib.vnlbhip_frame_cache_render_window(
            self._handle, int(frame), ctypes.c_void_p(src.data_ptr()),
            ctypes.c_void_p(guide.data_ptr()) if guide is not None else None, int(first_frame), window,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
        if rc != 0:
            self._context._raise(rc)
        return out

    def invalidate(self) -> None:
        """Drop every cached stack (the clip's content changed)."""
        self._lib.vnlbhip_frame_cache_invalidate(self._handle)

    def stats(self) -> dict:
        s = _abi.FrameCacheStats()
        self._lib.vnlbhip_frame_cache_get_stats(self._handle, ctypes.byref(s))
        return {"hits": s.hits, "misses": s.misses, "evictions": s.evictions}

    def close(self) -> None:
        if self._handle:
            self._lib.vnlbhip_frame_cache_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


def anchor_grid(params: StageParameters, width: int, height: int, total_frames: int, anchor_frame: int):
    """The staggered anchor grid of one anchor frame (absolute index) of a
    clip of total_frames as (xs, ys) int32 arrays in raster order
    (src/core/core.cpp:934-967), from the library's host code; needs no GPU."""
    lib = _abi.load()
    p = _abi.to_struct(params)
    count = lib.vnlbhip_anchor_grid(ctypes.byref(p), int(width), int(height), int(total_frames),
                                    int(anchor_frame), 0, None, None)
    if count < 0:
        raise VNLBError("VNLB HIP: invalid grid arguments")
    xs = np.zeros(count, np.int32)
    ys = np.zeros(count, np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.vnlbhip_anchor_grid(ctypes.byref(p), int(width), int(height), int(total_frames), int(anchor_frame), count,
                            xs.ctypes.data_as(ip), ys.ctypes.data_as(ip))
    return xs, ys


def aggregation_window(params: StageParameters):
    """The aggregation window a(u) [patch_size, patch_size]
    (src/core/core.cpp:296-324), from the library's host code; needs no GPU."""
    lib = _abi.load()
    out = np.zeros(params.patch_size * params.patch_size, np.float32)
    rc = lib.vnlbhip_aggregation_window(ctypes.byref(_abi.to_struct(params)),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if rc != 0:
        raise VNLBError("VNLB HIP: invalid window arguments")
    return out.reshape(params.patch_size, params.patch_size)


_local = threading.local()


def _context(device: int) -> Context:
    contexts = getattr(_local, "contexts", None)
    if contexts is None:
        contexts = _local.contexts = {}
    if device not in contexts:
        contexts[device] = Context(device)
    return contexts[device]


def _guided(context: Context, stage: int, params: StageParameters, clip, ref, mvfw, mvbw):
    """A stage, with search centres from mvfw / mvbw when either is given
    (process_*_anchor_mvtools, src/vs/plugin.cpp:700-737)."""
    if mvfw is None and mvbw is None:
        return context.stage(stage, params, clip, ref)
    if not _is_torch(clip):
        raise VNLBError("VNLB HIP: mvfw/mvbw guidance needs the clip as a GPU tensor")
    frames, _, height, width = _canonical_shape(tuple(clip.shape))
    field = context.mvtools_offsets(params, frames, height, width, mvfw, mvbw)
    context.set_search_offsets(field)
    try:
        return context.stage(stage, params, clip, ref)
    finally:
        context.set_search_offsets(None)


def Basic(clip, sigma, block_size=None, block_step=None, group_size=None, bm_range=None,
          patch_time=None, radius=None, search_bwd=None, search_fwd=None, rank=None,
          cap_factor=None, model_cap_factor=None, beta=None, tau=None,
          variance_threshold=None, weight_alpha=None, weight_beta=None, weight_gamma=None,
          weight_epsilon=None, membership_noise_floor=None, chroma=None, mvfw=None,
          mvbw=None, device_id=None, num_streams=None, chunk_size=None, form=None):
    """vnlb.Basic followed by vnlb.Aggregate(src=clip), on the GPU.
    form: 'plain' (default) or 'fast' (MFMA candidate distances)."""
    args = dict(locals())
    clip = args.pop("clip")
    form = args.pop("form") or "plain"
    params = parse_stage_parameters(BASIC, args)
    context = _context(int(device_id or 0))
    context.set_form(form)
    return _guided(context, BASIC, params, clip, None, mvfw, mvbw)


def Final(clip, ref, sigma, block_size=None, block_step=None, group_size=None, bm_range=None,
          patch_time=None, radius=None, search_bwd=None, search_fwd=None, rank=None,
          cap_factor=None, model_cap_factor=None, beta=None, tau=None,
          variance_threshold=None, sigma_basic=None, gamma=None, flat_areas=None,
          weight_alpha=None, weight_beta=None, weight_gamma=None, weight_epsilon=None,
          membership_noise_floor=None, chroma=None, mvfw=None, mvbw=None, device_id=None,
          num_streams=None, chunk_size=None, form=None):
    """vnlb.Final(ref=basic) followed by vnlb.Aggregate(src=clip), on the GPU.
    form: 'plain' (default) or 'fast' (MFMA candidate distances)."""
    args = dict(locals())
    clip = args.pop("clip")
    ref = args.pop("ref")
    form = args.pop("form") or "plain"
    params = parse_stage_parameters(FINAL, args)
    context = _context(int(device_id or 0))
    context.set_form(form)
    return _guided(context, FINAL, params, clip, ref, mvfw, mvbw)
