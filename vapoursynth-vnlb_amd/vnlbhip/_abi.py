"""ctypes binding of the vnlbhip C ABI (include/vnlbhip.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when libvnlbhip.so has not been built.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

from .params import StageParameters

LIBRARY_PATH = Path(__file__).resolve().parent.parent / "_build" / "libvnlbhip.so"

EXPORTED_SYMBOLS = (
    "vnlbhip_default_params",
    "vnlbhip_validate",
    "vnlbhip_create",
    "vnlbhip_destroy",
    "vnlbhip_last_error",
    "vnlbhip_reserve",
    "vnlbhip_stage",
    "vnlbhip_stage_scratch",
    "vnlbhip_stage_host",
    "vnlbhip_anchor_stack",
    "vnlbhip_anchor_stack_window",
    "vnlbhip_debug_frame",
    "vnlbhip_retained_cap",
    "vnlbhip_anchor_grid",
    "vnlbhip_aggregation_window",
    "vnlbhip_get_stats",
    "vnlbhip_set_profiling",
    "vnlbhip_get_kernel_times",
    "vnlbhip_set_form",
    "vnlbhip_set_eigen_refine",
    "vnlbhip_set_fused_selection",
    "vnlbhip_set_autotune",
    "vnlbhip_set_graph",
    "vnlbhip_set_search_offsets",
    "vnlbhip_mvtools_offsets",
    "vnlbhip_frame_cache_create",
    "vnlbhip_frame_cache_render",
    "vnlbhip_frame_cache_render_window",
    "vnlbhip_frame_cache_invalidate",
    "vnlbhip_frame_cache_get_stats",
    "vnlbhip_frame_cache_destroy",
    "vnlbhip_set_jacobi",
    "vnlbhip_set_tile_shape",
    "vnlbhip_proposer_create",
    "vnlbhip_proposer_destroy",
    "vnlbhip_proposer_last_error",
    "vnlbhip_proposer_get_params",
    "vnlbhip_proposer_set_params",
    "vnlbhip_proposer_features",
    "vnlbhip_proposer_gradients",
    "vnlbhip_proposer_train_step",
    "vnlbhip_proposer_scores",
    "vnlbhip_proposer_calibrate",
    "vnlbhip_set_proposer",
    "vnlbhip_match_lists",
    "vnlbhip_search_windows",
)

PROPOSER_INPUT = 128
PROPOSER_HIDDEN = 256
PROPOSER_PARAMS = 99073
PRECISIONS = {"bf16": 0, "fp8": 1}

KERNEL_NAMES = ("match", "paste", "prepare", "jacobi", "filter", "normalize", "select", "energy")


class FrameCacheStats(ctypes.Structure):
    _fields_ = [("hits", ctypes.c_int64), ("misses", ctypes.c_int64), ("evictions", ctypes.c_int64)]


class MvShape(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in (
        "block_size_x", "block_size_y", "overlap_x", "overlap_y", "block_count_x", "block_count_y", "pel")]


class Params(ctypes.Structure):
    _fields_ = [
        ("sigma", ctypes.c_float),
        ("patch_size", ctypes.c_int),
        ("patch_time", ctypes.c_int),
        ("search_window", ctypes.c_int),
        ("search_bwd", ctypes.c_int),
        ("search_fwd", ctypes.c_int),
        ("similar", ctypes.c_int),
        ("rank", ctypes.c_int),
        ("similar_cap_factor", ctypes.c_float),
        ("model_cap_factor", ctypes.c_float),
        ("beta", ctypes.c_float),
        ("tau", ctypes.c_float),
        ("variance_threshold", ctypes.c_float),
        ("sigma_basic", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("weight_alpha", ctypes.c_float),
        ("weight_beta", ctypes.c_float),
        ("weight_gamma", ctypes.c_float),
        ("weight_epsilon", ctypes.c_float),
        ("membership_noise_floor", ctypes.c_float),
        ("proc_step", ctypes.c_int),
        ("flat_areas", ctypes.c_int),
        ("couple_channels", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("anchors", ctypes.c_int64),
        ("groups", ctypes.c_int64),
        ("unconverged_groups", ctypes.c_int64),
        ("batches", ctypes.c_int32),
        ("parts", ctypes.c_int32),
        ("retained", ctypes.c_int64),
        ("refined", ctypes.c_int64),
        ("graph_replays", ctypes.c_int64),
        ("reallocations", ctypes.c_int64),
        ("proposed", ctypes.c_int64),
        ("exhaustive_fallbacks", ctypes.c_int64),
        ("tile_rows", ctypes.c_int32),
        ("tile_cols", ctypes.c_int32),
    ]


class KernelTimes(ctypes.Structure):
    _fields_ = [
        ("milliseconds", ctypes.c_double * len(KERNEL_NAMES)),
        ("launches", ctypes.c_int64 * len(KERNEL_NAMES)),
    ]


def to_struct(p: StageParameters) -> Params:
    s = Params()
    for name, _ in Params._fields_:
        value = getattr(p, name)
        setattr(s, name, int(value) if isinstance(value, bool) else value)
    return s


# Arithmetic forms (VNLBHIP_FORM_*).
FORMS = {"plain": 0, "fast": 1}

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY_PATH.exists():
        raise RuntimeError(
            f"vnlbhip: {LIBRARY_PATH} is missing; build it with "
            "`python vapoursynth-vnlb_amd/build.py` (there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIBRARY_PATH))
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    fp = P(ctypes.c_float)
    ip = P(ctypes.c_int)
    lib.vnlbhip_default_params.argtypes = [ctypes.c_int, P(Params)]
    lib.vnlbhip_default_params.restype = ctypes.c_int
    lib.vnlbhip_validate.argtypes = [ctypes.c_int, P(Params), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    lib.vnlbhip_validate.restype = ctypes.c_int
    lib.vnlbhip_create.argtypes = [ctypes.c_int, P(vp)]
    lib.vnlbhip_create.restype = ctypes.c_int
    lib.vnlbhip_destroy.argtypes = [vp]
    lib.vnlbhip_destroy.restype = None
    lib.vnlbhip_last_error.argtypes = [vp]
    lib.vnlbhip_last_error.restype = ctypes.c_char_p
    lib.vnlbhip_reserve.argtypes = [vp, ctypes.c_int, P(Params), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
    lib.vnlbhip_reserve.restype = ctypes.c_int
    lib.vnlbhip_stage.argtypes = [vp, ctypes.c_int, P(Params), vp, vp, vp, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    lib.vnlbhip_stage.restype = ctypes.c_int
    lib.vnlbhip_stage_scratch.argtypes = lib.vnlbhip_stage.argtypes
    lib.vnlbhip_stage_scratch.restype = ctypes.c_int
    lib.vnlbhip_anchor_stack.argtypes = [vp, ctypes.c_int, P(Params), vp, vp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         vp, vp]
    lib.vnlbhip_anchor_stack.restype = ctypes.c_int
    lib.vnlbhip_anchor_stack_window.argtypes = [vp, ctypes.c_int, P(Params), vp, vp] + [ctypes.c_int] * 7 + [vp, vp]
    lib.vnlbhip_anchor_stack_window.restype = ctypes.c_int
    lib.vnlbhip_stage_host.argtypes = [vp, ctypes.c_int, P(Params), fp, fp, fp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.vnlbhip_stage_host.restype = ctypes.c_int
    lib.vnlbhip_debug_frame.argtypes = [vp, ctypes.c_int, P(Params), vp, vp, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ip, ip, ip, ip, ip, fp,
                                        P(ctypes.c_ubyte), fp, fp, ip, ctypes.c_int, fp]
    lib.vnlbhip_debug_frame.restype = ctypes.c_int
    lib.vnlbhip_retained_cap.argtypes = [P(Params)]
    lib.vnlbhip_retained_cap.restype = ctypes.c_int
    lib.vnlbhip_anchor_grid.argtypes = [P(Params)] + [ctypes.c_int] * 5 + [ip, ip]
    lib.vnlbhip_anchor_grid.restype = ctypes.c_int
    lib.vnlbhip_aggregation_window.argtypes = [P(Params), fp]
    lib.vnlbhip_aggregation_window.restype = ctypes.c_int
    lib.vnlbhip_get_stats.argtypes = [vp, P(Stats)]
    lib.vnlbhip_get_stats.restype = ctypes.c_int
    lib.vnlbhip_set_profiling.argtypes = [vp, ctypes.c_int]
    lib.vnlbhip_set_profiling.restype = ctypes.c_int
    lib.vnlbhip_set_form.argtypes = [vp, ctypes.c_int]
    lib.vnlbhip_set_form.restype = ctypes.c_int
    lib.vnlbhip_set_eigen_refine.argtypes = [vp, ctypes.c_float]
    lib.vnlbhip_set_eigen_refine.restype = ctypes.c_int
    lib.vnlbhip_set_fused_selection.argtypes = [vp, ctypes.c_int]
    lib.vnlbhip_set_fused_selection.restype = ctypes.c_int
    lib.vnlbhip_set_autotune.argtypes = [vp, ctypes.c_int]
    lib.vnlbhip_set_autotune.restype = ctypes.c_int
    lib.vnlbhip_set_graph.argtypes = [vp, ctypes.c_int]
    lib.vnlbhip_set_graph.restype = ctypes.c_int
    lib.vnlbhip_set_search_offsets.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.vnlbhip_set_search_offsets.restype = ctypes.c_int
    mv = P(MvShape)
    lib.vnlbhip_mvtools_offsets.argtypes = [vp, P(Params), mv, vp, vp, mv, vp, vp, vp, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, vp]
    lib.vnlbhip_mvtools_offsets.restype = ctypes.c_int
    ci = ctypes.c_int
    lib.vnlbhip_frame_cache_create.argtypes = [vp, ci, P(Params), ci, ci, ci, ci, ci, P(vp)]
    lib.vnlbhip_frame_cache_create.restype = ci
    lib.vnlbhip_frame_cache_render.argtypes = [vp, ci, vp, vp, vp, vp]
    lib.vnlbhip_frame_cache_render.restype = ci
    lib.vnlbhip_frame_cache_render_window.argtypes = [vp, ci, vp, vp, ci, ci, vp, vp]
    lib.vnlbhip_frame_cache_render_window.restype = ci
    lib.vnlbhip_frame_cache_invalidate.argtypes = [vp]
    lib.vnlbhip_frame_cache_invalidate.restype = ci
    lib.vnlbhip_frame_cache_get_stats.argtypes = [vp, P(FrameCacheStats)]
    lib.vnlbhip_frame_cache_get_stats.restype = ci
    lib.vnlbhip_frame_cache_destroy.argtypes = [vp]
    lib.vnlbhip_frame_cache_destroy.restype = None
    lib.vnlbhip_get_kernel_times.argtypes = [vp, P(KernelTimes)]
    lib.vnlbhip_get_kernel_times.restype = ctypes.c_int
    lib.vnlbhip_set_jacobi.argtypes = [vp, ci, ci]
    lib.vnlbhip_set_jacobi.restype = ci
    lib.vnlbhip_set_tile_shape.argtypes = [vp, ci, ci]
    lib.vnlbhip_set_tile_shape.restype = ci
    lib.vnlbhip_proposer_create.argtypes = [ci, ctypes.c_uint64, P(vp)]
    lib.vnlbhip_proposer_create.restype = ci
    lib.vnlbhip_proposer_destroy.argtypes = [vp]
    lib.vnlbhip_proposer_destroy.restype = None
    lib.vnlbhip_proposer_last_error.argtypes = [vp]
    lib.vnlbhip_proposer_last_error.restype = ctypes.c_char_p
    lib.vnlbhip_proposer_get_params.argtypes = [vp, fp]
    lib.vnlbhip_proposer_get_params.restype = ci
    lib.vnlbhip_proposer_set_params.argtypes = [vp, fp]
    lib.vnlbhip_proposer_set_params.restype = ci
    lib.vnlbhip_proposer_features.argtypes = [vp, vp, ci, ci, ci, ci, ci, vp, vp, ci, vp, vp]
    lib.vnlbhip_proposer_features.restype = ci
    lib.vnlbhip_proposer_gradients.argtypes = [vp, vp, vp, ci, ci, vp, vp, vp]
    lib.vnlbhip_proposer_gradients.restype = ci
    lib.vnlbhip_proposer_train_step.argtypes = [vp, vp, vp, ci, ci, ctypes.c_float, vp, vp]
    lib.vnlbhip_proposer_train_step.restype = ci
    lib.vnlbhip_proposer_scores.argtypes = [vp, vp, ci, ci, vp, vp]
    lib.vnlbhip_proposer_scores.restype = ci
    lib.vnlbhip_proposer_calibrate.argtypes = [vp, vp, ci, vp]
    lib.vnlbhip_proposer_calibrate.restype = ci
    lib.vnlbhip_set_proposer.argtypes = [vp, vp, ci, ctypes.c_float, ci]
    lib.vnlbhip_set_proposer.restype = ci
    lib.vnlbhip_match_lists.argtypes = [vp, ci, P(Params), vp, vp, ci, ci, ci, ci, ci, ci, P(ci), vp, vp, vp, vp, vp]
    lib.vnlbhip_match_lists.restype = ci
    lib.vnlbhip_search_windows.argtypes = [P(Params), ci, ci, ci, ci, P(ctypes.c_uint32), P(ctypes.c_int32)]
    lib.vnlbhip_search_windows.restype = ci
    _lib = lib
    return lib
