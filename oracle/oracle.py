"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker or the timed CPU baseline, never as the thing
the product runs.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIBRARY = ORACLE_DIR / "_build" / "libvnlb_oracle.so"

ACCUM_DOUBLE = 1
SELECT_FAITHFUL = 2
MODEL_DOUBLE = 4

_FIELDS = [
    ("sigma", ctypes.c_float), ("patch_size", ctypes.c_int), ("patch_time", ctypes.c_int),
    ("search_window", ctypes.c_int), ("search_bwd", ctypes.c_int), ("search_fwd", ctypes.c_int),
    ("similar", ctypes.c_int), ("rank", ctypes.c_int), ("similar_cap_factor", ctypes.c_float),
    ("model_cap_factor", ctypes.c_float), ("beta", ctypes.c_float), ("tau", ctypes.c_float),
    ("variance_threshold", ctypes.c_float), ("sigma_basic", ctypes.c_float),
    ("gamma", ctypes.c_float), ("weight_alpha", ctypes.c_float), ("weight_beta", ctypes.c_float),
    ("weight_gamma", ctypes.c_float), ("weight_epsilon", ctypes.c_float),
    ("membership_noise_floor", ctypes.c_float), ("proc_step", ctypes.c_int),
    ("flat_areas", ctypes.c_int), ("couple_channels", ctypes.c_int),
]


class VoParams(ctypes.Structure):
    _fields_ = _FIELDS


MV_SHAPE_FIELDS = ("block_size_x", "block_size_y", "overlap_x", "overlap_y", "block_count_x",
                   "block_count_y", "pel")


class MvShape(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int) for name in MV_SHAPE_FIELDS]


def _vector_clip(clip):
    """(shape dict, vectors [frames, by, bx, 2] int32, valid [frames] uint8) or
    None -> ctypes arguments (kept alive by the returned tuple)."""
    P = ctypes.POINTER
    if clip is None:
        return None, None, None, None
    shape, vectors, valid = clip
    s = MvShape(*[int(shape[name]) for name in MV_SHAPE_FIELDS])
    vectors = np.ascontiguousarray(vectors, np.int32)
    valid = np.ascontiguousarray(valid, np.uint8)
    return (ctypes.byref(s), vectors.ctypes.data_as(P(ctypes.c_int32)),
            valid.ctypes.data_as(P(ctypes.c_ubyte)), (s, vectors, valid))


def center_for(fw, bw, anchor_x, anchor_y, anchor_frame, candidate_frame, patch_size, width, height,
               grid_frames=None):
    """MVToolsFlowProvider::center_for over dense vector clips (fw: the
    mvfw clip, walked towards earlier frames; bw: the mvbw clip)."""
    lib = load()
    a = _vector_clip(fw)
    b = _vector_clip(bw)
    if grid_frames is None:
        grid_frames = len((fw or bw)[2]) if (fw or bw) is not None else 0
    out = (ctypes.c_int * 2)()
    lib.vo_center_for(a[0], a[1], a[2], b[0], b[1], b[2], int(grid_frames), int(anchor_x),
                      int(anchor_y), int(anchor_frame), int(candidate_frame), int(patch_size),
                      int(width), int(height), out)
    return int(out[0]), int(out[1])


def mvtools_offsets(params, fw, bw, width, height, frames):
    """Dense search offsets [frames, temporal, height, width, 2] int16 from
    the vector clips, in the layout of set_search_offsets."""
    lib = load()
    struct = to_struct(params)
    temporal = struct.search_bwd + struct.search_fwd + 1
    out = np.zeros((frames, temporal, height, width, 2), np.int16)
    a = _vector_clip(fw)
    b = _vector_clip(bw)
    rc = lib.vo_mvtools_offsets(ctypes.byref(struct), a[0], a[1], a[2], b[0], b[1], b[2], int(width),
                                int(height), int(frames),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)))
    if rc != 0:
        raise ValueError("vo_mvtools_offsets: invalid arguments")
    return out


def to_struct(params) -> VoParams:
    s = VoParams()
    source = params if isinstance(params, dict) else params.as_dict()
    for name, _ in _FIELDS:
        value = source[name]
        setattr(s, name, int(value) if isinstance(value, bool) else value)
    return s


def build() -> Path:
    subprocess.run(["make", "-C", str(ORACLE_DIR), "-s"], check=True)
    return LIBRARY


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            build()
        lib = ctypes.CDLL(str(LIBRARY))
        P = ctypes.POINTER
        fp, ip, dp = P(ctypes.c_float), P(ctypes.c_int), P(ctypes.c_double)
        lib.vo_stage.argtypes = [ctypes.c_int, P(VoParams), fp, fp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, fp, ctypes.c_int, ctypes.c_int,
                                 P(ctypes.c_int64)]
        lib.vo_frame_debug.argtypes = [ctypes.c_int, P(VoParams), fp, fp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ip, ip, ip, ip, ip, fp, P(ctypes.c_ubyte),
                                       fp, fp, ip, ctypes.c_int, ctypes.c_int, dp]
        lib.vo_anchor_stack.argtypes = [ctypes.c_int, P(VoParams), fp, fp, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, ctypes.c_int]
        lib.vo_anchor_stack_window.argtypes = [ctypes.c_int, P(VoParams), fp, fp] + [ctypes.c_int] * 7 + [
            fp, ctypes.c_int]
        lib.vo_aggregation_weights.argtypes = [ctypes.c_int, P(VoParams), dp, ctypes.c_int, ctypes.c_int, dp, dp,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
        lib.vo_aggregation_window.argtypes = [P(VoParams), fp]
        lib.vo_grid_mask.argtypes = [P(VoParams), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     P(ctypes.c_ubyte)]
        lib.vo_set_search_offsets.argtypes = [P(ctypes.c_int16), ctypes.c_int]
        lib.vo_set_search_offsets.restype = None
        lib.vo_set_max_sweeps.argtypes = [ctypes.c_int]
        lib.vo_set_max_sweeps.restype = None
        mv = P(MvShape)
        i32, u8 = P(ctypes.c_int32), P(ctypes.c_ubyte)
        lib.vo_center_for.argtypes = [mv, i32, u8, mv, i32, u8] + [ctypes.c_int] * 8 + [ip]
        lib.vo_center_for.restype = None
        lib.vo_mvtools_offsets.argtypes = [P(VoParams), mv, i32, u8, mv, i32, u8, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, P(ctypes.c_int16)]
        lib.vo_retained_cap.argtypes = [P(VoParams)]
        lib.vo_grid_count.argtypes = [P(VoParams), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
        lib.vo_linalg_mean.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp]
        lib.vo_linalg_gram_centered.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
        lib.vo_linalg_covariance_centered.argtypes = [dp, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_double, dp]
        lib.vo_linalg_topk_eigen.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, dp, ip]
        lib.vo_linalg_topk_eigen_f32.argtypes = [fp, ctypes.c_int, ctypes.c_int, fp, fp, ip]
        lib.vo_linalg_dual_basis.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, dp]
        lib.vo_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _video(a) -> tuple[np.ndarray, tuple[int, int, int, int]]:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim == 3:
        a = a[:, None]
    f, c, h, w = a.shape
    return a, (w, h, f, c)


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def stage(stage_id: int, params, noisy, ref=None, flags: int = ACCUM_DOUBLE, threads: int = 1):
    lib = load()
    src, (w, h, f, c) = _video(noisy)
    guide = _video(ref)[0] if ref is not None else None
    out = np.empty_like(src)
    groups = ctypes.c_int64(0)
    p = to_struct(params)
    rc = lib.vo_stage(stage_id, ctypes.byref(p), _fp(src), _fp(guide), w, h, f, c, _fp(out),
                      flags, threads, ctypes.byref(groups))
    if rc != 0:
        raise ValueError(lib.vo_last_error().decode())
    return out, int(groups.value)


_offsets = None  # keeps the array the oracle points at alive


def set_search_offsets(offsets, temporal: int = 0) -> None:
    """Search-centre offsets [frames, temporal, height, width, 2] int16 (dx, dy)
    per anchor position and scheduled candidate frame, or None to clear."""
    global _offsets
    lib = load()
    if offsets is None:
        _offsets = None
        lib.vo_set_search_offsets(None, 0)
        return
    _offsets = np.ascontiguousarray(offsets, dtype=np.int16)
    lib.vo_set_search_offsets(_offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)), int(temporal))


def set_max_sweeps(max_sweeps: int = 64) -> None:
    """Jacobi sweep limit of every eigen solve (64 restores the default)."""
    load().vo_set_max_sweeps(int(max_sweeps))


def anchor_stack(stage_id: int, params, noisy, ref, anchor_frame: int, flags: int = ACCUM_DOUBLE):
    """Contribution stack of one anchor frame, [channels, slots * 2 * height,
    width]: per slot a numerator plane then the weight plane (the layout of a
    vnlb.Basic / vnlb.Final output frame)."""
    lib = load()
    src, (w, h, f, c) = _video(noisy)
    guide = _video(ref)[0] if ref is not None else None
    p = to_struct(params)
    slots = int(p.search_bwd) + int(p.search_fwd) + int(p.patch_time)
    packed = np.empty((c, slots * 2 * h, w), np.float32)
    rc = lib.vo_anchor_stack(stage_id, ctypes.byref(p), _fp(src), _fp(guide), w, h, f, c,
                             anchor_frame, _fp(packed), flags)
    if rc != 0:
        raise ValueError(lib.vo_last_error().decode())
    return packed


def anchor_stack_window(stage_id: int, params, noisy, ref, first_frame: int, total_frames: int,
                        anchor_frame: int, flags: int = ACCUM_DOUBLE):
    """anchor_stack over a window: noisy / ref hold the frames [first_frame,
    first_frame + len(noisy)) of a clip of total_frames; anchor_frame is
    absolute."""
    lib = load()
    src, (w, h, f, c) = _video(noisy)
    guide = _video(ref)[0] if ref is not None else None
    p = to_struct(params)
    slots = int(p.search_bwd) + int(p.search_fwd) + int(p.patch_time)
    packed = np.empty((c, slots * 2 * h, w), np.float32)
    rc = lib.vo_anchor_stack_window(stage_id, ctypes.byref(p), _fp(src), _fp(guide), w, h, f, c,
                                    int(first_frame), int(total_frames), int(anchor_frame), _fp(packed),
                                    flags)
    if rc != 0:
        raise ValueError(lib.vo_last_error().decode())
    return packed


def frame_debug(stage_id: int, params, noisy, ref, anchor_frame: int, flags: int = ACCUM_DOUBLE,
                gram_dim: int = 16) -> dict:
    lib = load()
    src, (w, h, f, c) = _video(noisy)
    guide = _video(ref)[0] if ref is not None else None
    p = to_struct(params)
    cap = lib.vo_retained_cap(ctypes.byref(p))
    capacity = w * h
    rank = max(int(p.rank), 1)
    num = ctypes.c_int(0)
    ax = np.zeros(capacity, np.int32)
    ay = np.zeros(capacity, np.int32)
    counts = np.zeros(capacity, np.int32)
    positions = np.full(capacity * cap, -1, np.int32)
    distances = np.full(capacity * cap, np.nan, np.float32)
    active = np.zeros(capacity, np.uint8)
    weights = np.full(capacity * cap, np.nan, np.float32)
    eig = np.full(capacity * rank, np.nan, np.float32)
    gflags = np.zeros(capacity, np.int32)
    grams = np.full(capacity * gram_dim * gram_dim, np.nan, np.float64)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib.vo_frame_debug(stage_id, ctypes.byref(p), _fp(src), _fp(guide), w, h, f, c,
                            anchor_frame, capacity, ctypes.byref(num), ax.ctypes.data_as(ip),
                            ay.ctypes.data_as(ip), counts.ctypes.data_as(ip),
                            positions.ctypes.data_as(ip), _fp(distances),
                            active.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _fp(weights),
                            _fp(eig), gflags.ctypes.data_as(ip), flags, gram_dim, _dp(grams))
    if rc != 0:
        raise ValueError(lib.vo_last_error().decode())
    n = num.value
    return {
        "cap": cap, "anchor_x": ax[:n], "anchor_y": ay[:n], "counts": counts[:n],
        "positions": positions[: n * cap].reshape(n, cap),
        "distances": distances[: n * cap].reshape(n, cap), "active": active[:n],
        "weights": weights[: n * cap].reshape(n, cap),
        "eigenvalues": eig[: n * rank].reshape(n, rank), "group_flags": gflags[:n],
        "grams": grams[: n * gram_dim * gram_dim].reshape(n, gram_dim, gram_dim),
    }


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def aggregation_weights(stage_id: int, params, rows, eigenvalues=None, basis=None, basis_similar=None,
                        models: int = 1, precision: int = 1):
    """Weights of one group (reset, `models` x accumulate_aggregation_model,
    finalize) from its centred membership rows [S, D] and rank eigen pairs."""
    lib = load()
    r = np.ascontiguousarray(rows, np.float64)
    similar, dim = r.shape
    e = np.ascontiguousarray(np.zeros(0) if eigenvalues is None else eigenvalues, np.float64)
    b = np.ascontiguousarray(np.zeros((0, dim)) if basis is None else basis, np.float64)
    rank = e.shape[0]
    out = np.zeros(similar, np.float32)
    rc = lib.vo_aggregation_weights(stage_id, ctypes.byref(to_struct(params)), _dp(r), similar, dim, _dp(e),
                                    _dp(b), rank, int(basis_similar or similar), int(models), int(precision),
                                    _fp(out))
    if rc != 0:
        raise ValueError(lib.vo_last_error().decode())
    return out


def aggregation_window(params):
    lib = load()
    p = to_struct(params)
    out = np.zeros(p.patch_size * p.patch_size, np.float32)
    assert lib.vo_aggregation_window(ctypes.byref(p), _fp(out)) == 0
    return out.reshape(p.patch_size, p.patch_size)


def grid_mask(params, width: int, height: int, total_frames: int, anchor_frame: int):
    lib = load()
    mask = np.zeros((height, width), np.uint8)
    count = lib.vo_grid_mask(ctypes.byref(to_struct(params)), width, height, total_frames, anchor_frame,
                             mask.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    assert count >= 0
    return mask


def linalg_mean(samples):
    lib = load()
    s = np.ascontiguousarray(samples, np.float64)
    mean = np.zeros(s.shape[1])
    assert lib.vo_linalg_mean(_dp(s), s.shape[0], s.shape[1], _dp(mean)) == 0
    return mean


def linalg_gram_centered(centered, scale=1.0):
    lib = load()
    s = np.ascontiguousarray(centered, np.float64)
    gram = np.zeros((s.shape[0], s.shape[0]))
    assert lib.vo_linalg_gram_centered(_dp(s), s.shape[0], s.shape[1], scale, _dp(gram)) == 0
    return gram


def linalg_covariance_centered(centered, scale=1.0):
    lib = load()
    s = np.ascontiguousarray(centered, np.float64)
    cov = np.zeros((s.shape[1], s.shape[1]))
    assert lib.vo_linalg_covariance_centered(_dp(s), s.shape[0], s.shape[1], scale, _dp(cov)) == 0
    return cov


def linalg_topk_eigen(matrix, rank, dtype=np.float64):
    lib = load()
    m = np.ascontiguousarray(matrix, dtype)
    n = m.shape[0]
    values = np.zeros(rank, dtype)
    vectors = np.zeros((rank, n), dtype)
    converged = ctypes.c_int(0)
    if dtype == np.float64:
        sweeps = lib.vo_linalg_topk_eigen(_dp(m), n, rank, _dp(values), _dp(vectors),
                                          ctypes.byref(converged))
    else:
        sweeps = lib.vo_linalg_topk_eigen_f32(_fp(m), n, rank, _fp(values), _fp(vectors),
                                              ctypes.byref(converged))
    assert sweeps >= 0
    return values, vectors, bool(converged.value), sweeps


def linalg_dual_basis(centered, gram_values, gram_vectors):
    lib = load()
    s = np.ascontiguousarray(centered, np.float64)
    values = np.ascontiguousarray(gram_values, np.float64)
    vectors = np.ascontiguousarray(gram_vectors, np.float64)
    rank = values.shape[0]
    basis = np.zeros((rank, s.shape[1]))
    assert lib.vo_linalg_dual_basis(_dp(s), s.shape[0], s.shape[1], _dp(values), _dp(vectors),
                                    rank, _dp(basis)) == 0
    return basis
