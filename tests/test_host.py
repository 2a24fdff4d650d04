"""Host logic without a GPU: argument parsing and validation mirroring the
reference plugin (src/vs/plugin.cpp:228-397, src/core/core.cpp:1948-2024),
and the C ABI library's exported surface."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from vnlbhip import _abi
from vnlbhip.params import BASIC, FINAL, StageParameters, VNLBError, parse_stage_parameters

REPO = Path(__file__).resolve().parent.parent


def test_basic_defaults_follow_stage_parameters():
    p = parse_stage_parameters(BASIC, {"sigma": 10})
    d = StageParameters()
    for name, value in d.as_dict().items():
        if name == "sigma":
            continue
        assert getattr(p, name) == pytest.approx(value), name
    assert p.sigma == float(np.float32(10) / np.float32(255))


def test_final_overrides():
    p = parse_stage_parameters(FINAL, {"sigma": 20})
    assert p.tau == float(np.float32(400) / np.float32(65025))
    assert p.variance_threshold == pytest.approx(1.7)
    assert p.flat_areas and p.weight_alpha == 1.0 and p.weight_beta == 0.5 and p.weight_gamma == 1.0


def test_argument_mapping():
    p = parse_stage_parameters(FINAL, dict(sigma=15, block_size=8, block_step=4, group_size=16,
                                           bm_range=9, patch_time=2, radius=2, search_fwd=3,
                                           rank=16, tau=200, cap_factor=4.0, chroma=0,
                                           sigma_basic=5, flat_areas=0))
    assert p.search_window == 19 and p.search_bwd == 2 and p.search_fwd == 3
    assert p.proc_step == 4 and p.similar == 16 and p.patch_time == 2
    assert p.tau == float(np.float32(200) / np.float32(65025))
    assert p.sigma_basic == float(np.float32(5) / np.float32(255))
    assert not p.couple_channels and not p.flat_areas


@pytest.mark.parametrize("args,message", [
    ({}, "missing required argument: sigma"),
    ({"sigma": 10, "radius": -1}, "radius must be non-negative"),
    ({"sigma": 10, "bm_range": -2}, "bm_range must be non-negative"),
    ({"sigma": 10, "patch_time": 0}, "patch_time must be positive"),
    ({"sigma": 10, "mvfw": object()}, "mvfw/mvbw"),
    ({"sigma": 10, "flat_areas": 1}, "unknown argument"),
])
def test_argument_errors(args, message):
    with pytest.raises(VNLBError, match=message):
        parse_stage_parameters(BASIC, args)


def test_library_exports_every_declared_symbol():
    header = (REPO / "include" / "vnlbhip.h").read_text()
    declared = set(re.findall(r"\b(vnlbhip_[a-z_]+)\s*\(", header))
    assert declared == set(_abi.EXPORTED_SYMBOLS)
    lib = ctypes.CDLL(str(_abi.LIBRARY_PATH))
    for name in declared:
        assert hasattr(lib, name), name


def _validate(stage, params, w, h, f, c):
    lib = _abi.load()
    buf = ctypes.create_string_buffer(256)
    rc = lib.vnlbhip_validate(stage, ctypes.byref(_abi.to_struct(params)), w, h, f, c, buf, 256)
    return rc, buf.value.decode()


def test_validate_accepts_benchmark_shapes():
    for stage in (BASIC, FINAL):
        p = parse_stage_parameters(stage, dict(sigma=15, block_size=8, block_step=4, group_size=16,
                                               bm_range=9, patch_time=2, radius=2, rank=16,
                                               tau=200, cap_factor=4.0, chroma=1))
        assert _validate(stage, p, 1920, 1080, 16, 3) == (0, "")


@pytest.mark.parametrize("change,message", [
    (dict(rank=32), "group_size must cover the effective VNLB rank"),
    (dict(rank=400), "rank must not exceed the effective patch dimension"),
    (dict(block_size=0), "block_size must be positive"),
    (dict(cap_factor=0.5), "cap_factor must be zero or at least 1.0"),
    (dict(block_size=40), "block_size must fit inside the frame"),
])
def test_validate_messages(change, message):
    args = dict(sigma=15, block_size=8, block_step=4, group_size=16, bm_range=9, patch_time=2,
                radius=2, rank=16, chroma=1)
    args.update(change)
    p = parse_stage_parameters(BASIC, args)
    rc, text = _validate(BASIC, p, 32, 32, 8, 3)
    assert rc == -1 and message in text


def test_validate_reports_device_limits():
    p = parse_stage_parameters(BASIC, dict(sigma=10, bm_range=60, radius=1))
    rc, text = _validate(BASIC, p, 256, 256, 4, 1)
    assert rc == -1 and "32768 candidates" in text


def test_default_params_entry_point():
    lib = _abi.load()
    for stage in (BASIC, FINAL):
        s = _abi.Params()
        assert lib.vnlbhip_default_params(stage, ctypes.byref(s)) == 0
        p = parse_stage_parameters(stage, {"sigma": 1})
        for name, _ in _abi.Params._fields_:
            if name == "sigma":
                continue
            assert getattr(s, name) == pytest.approx(float(getattr(p, name))), (stage, name)


def test_retained_cap():
    lib = _abi.load()
    p = parse_stage_parameters(BASIC, dict(sigma=15, group_size=16, cap_factor=4.0, bm_range=9, radius=2))
    assert lib.vnlbhip_retained_cap(ctypes.byref(_abi.to_struct(p))) == 64
    p = parse_stage_parameters(BASIC, dict(sigma=15, group_size=16, cap_factor=4.0, bm_range=1, radius=0))
    assert lib.vnlbhip_retained_cap(ctypes.byref(_abi.to_struct(p))) == 9
