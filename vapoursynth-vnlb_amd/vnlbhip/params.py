"""Argument parsing of vnlb.Basic / vnlb.Final for vnlbhip.

Mirrors parse_stage_parameters and parse_temporal_parameters of the
reference CPU plugin (src/vs/plugin.cpp:228-397): the StageParameters
defaults (src/core/core.hpp:17-41), the Final stage overrides
(src/vs/plugin.cpp:331-339), 8-bit scaling of sigma / sigma_basic (/255) and
tau (/255^2) carried out in float32 as the plugin does, radius expanding to
search_bwd / search_fwd, bm_range to the odd search window, chroma to
couple_channels.
"""
from __future__ import annotations

from dataclasses import dataclass, fields

import numpy as np

EIGHT_BIT_SCALE = np.float32(255.0)
EIGHT_BIT_DISTANCE_SCALE = np.float32(255.0 * 255.0)

BASIC = 0
FINAL = 1

# Argument strings registered by the reference plugin
# (src/vs/plugin.cpp:1162-1185); vnlbhip accepts the same names.
BASIC_ARGS = (
    "sigma block_size block_step group_size bm_range patch_time radius search_bwd "
    "search_fwd rank cap_factor model_cap_factor beta tau variance_threshold "
    "weight_alpha weight_beta weight_gamma weight_epsilon membership_noise_floor "
    "chroma mvfw mvbw"
).split()
FINAL_ARGS = (
    "sigma block_size block_step group_size bm_range patch_time radius search_bwd "
    "search_fwd rank cap_factor model_cap_factor beta tau variance_threshold "
    "sigma_basic gamma flat_areas weight_alpha weight_beta weight_gamma weight_epsilon "
    "membership_noise_floor chroma mvfw mvbw"
).split()
# Device extras in the style of the CUDA plugin (src/vs/cuda_plugin.cpp:2154-2178).
DEVICE_ARGS = ("device_id", "num_streams", "chunk_size")


class VNLBError(ValueError):
    """Raised with the reference's messages, prefixed "VNLB HIP: "."""


@dataclass
class StageParameters:
    sigma: float = 1.0
    patch_size: int = 8
    patch_time: int = 1
    search_window: int = 19
    search_bwd: int = 1
    search_fwd: int = 1
    similar: int = 8
    rank: int = 8
    similar_cap_factor: float = 4.0
    model_cap_factor: float = 1.0
    beta: float = 1.0
    tau: float = 0.0
    variance_threshold: float = 1.1
    sigma_basic: float = 0.0
    gamma: float = 0.95
    weight_alpha: float = 0.75
    weight_beta: float = 0.35
    weight_gamma: float = 1.0
    weight_epsilon: float = 1.0e-6
    membership_noise_floor: float = 0.25
    proc_step: int = 8
    flat_areas: bool = False
    couple_channels: bool = True

    def as_dict(self) -> dict:
        return {f.name: getattr(self, f.name) for f in fields(self)}


def _f32(value) -> float:
    return float(np.float32(value))


def _int(name: str, value) -> int:
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
        raise VNLBError(f"VNLB HIP: {name} must be an int")
    return int(value)


def _float(name: str, value) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float, np.integer, np.floating)):
        raise VNLBError(f"VNLB HIP: {name} must be a float")
    return _f32(value)


def parse_stage_parameters(stage: int, args: dict) -> StageParameters:
    """StageParameters from vnlb keyword arguments (plugin.cpp:327-397)."""
    allowed = set(BASIC_ARGS if stage == BASIC else FINAL_ARGS) | set(DEVICE_ARGS) | {"clip", "ref"}
    unknown = sorted(k for k, v in args.items() if k not in allowed and v is not None)
    if unknown:
        raise VNLBError(f"VNLB HIP: unknown argument(s): {', '.join(unknown)}")

    def get(name):
        return args.get(name)

    p = StageParameters()
    if get("sigma") is None:
        raise VNLBError("VNLB HIP: missing required argument: sigma")
    p.sigma = float(np.float32(_float("sigma", get("sigma"))) / EIGHT_BIT_SCALE)
    if stage == FINAL:
        p.tau = float(np.float32(400.0) / EIGHT_BIT_DISTANCE_SCALE)
        p.variance_threshold = _f32(1.7)
        p.flat_areas = True
        p.weight_alpha = 1.0
        p.weight_beta = 0.5
        p.weight_gamma = 1.0

    if get("block_size") is not None:
        p.patch_size = _int("block_size", get("block_size"))
    # parse_temporal_parameters (plugin.cpp:228-257).
    if get("patch_time") is not None:
        p.patch_time = _int("patch_time", get("patch_time"))
    if get("radius") is not None:
        radius = _int("radius", get("radius"))
        if radius < 0:
            raise VNLBError("VNLB HIP: radius must be non-negative")
        p.search_bwd = radius
        p.search_fwd = radius
    if get("search_bwd") is not None:
        p.search_bwd = _int("search_bwd", get("search_bwd"))
    if get("search_fwd") is not None:
        p.search_fwd = _int("search_fwd", get("search_fwd"))
    if p.patch_time <= 0 or p.search_bwd < 0 or p.search_fwd < 0:
        raise VNLBError(
            "VNLB HIP: patch_time must be positive and search_bwd/search_fwd must be non-negative"
        )
    if get("bm_range") is not None:
        bm_range = _int("bm_range", get("bm_range"))
        if bm_range < 0 or bm_range > (2**31 - 2) // 2:
            raise VNLBError("VNLB HIP: bm_range must be non-negative")
        p.search_window = bm_range * 2 + 1
    if get("group_size") is not None:
        p.similar = _int("group_size", get("group_size"))
    if get("rank") is not None:
        p.rank = _int("rank", get("rank"))
    if get("cap_factor") is not None:
        p.similar_cap_factor = _float("cap_factor", get("cap_factor"))
    if get("model_cap_factor") is not None:
        p.model_cap_factor = _float("model_cap_factor", get("model_cap_factor"))
    if get("beta") is not None:
        p.beta = _float("beta", get("beta"))
    if get("tau") is not None:
        p.tau = float(np.float32(_float("tau", get("tau"))) / EIGHT_BIT_DISTANCE_SCALE)
    if get("variance_threshold") is not None:
        p.variance_threshold = _float("variance_threshold", get("variance_threshold"))
    if get("block_step") is not None:
        p.proc_step = _int("block_step", get("block_step"))
    if stage == FINAL:
        if get("sigma_basic") is not None:
            p.sigma_basic = float(np.float32(_float("sigma_basic", get("sigma_basic"))) / EIGHT_BIT_SCALE)
        if get("gamma") is not None:
            p.gamma = _float("gamma", get("gamma"))
        if get("flat_areas") is not None:
            p.flat_areas = _int("flat_areas", get("flat_areas")) != 0
    for name in ("weight_alpha", "weight_beta", "weight_gamma", "weight_epsilon", "membership_noise_floor"):
        if get(name) is not None:
            setattr(p, name, _float(name, get(name)))
    if get("chroma") is not None:
        p.couple_channels = _int("chroma", get("chroma")) != 0
    for name in ("mvfw", "mvbw"):
        clip = get(name)
        if clip is not None and type(clip).__name__ != "VectorClip":
            raise VNLBError(f"VNLB HIP: mvfw/mvbw must be vnlbhip.flow.VectorClip objects ({name})")
    return p


def slot_count(p: StageParameters) -> int:
    return p.search_bwd + p.search_fwd + p.patch_time
