"""Build libvnlbhip.so (gfx950) in-tree with hipcc.

Usage: python vapoursynth-vnlb_amd/build.py [--jobs N]
Output: vapoursynth-vnlb_amd/_build/libvnlbhip.so
"""
from __future__ import annotations

import argparse
import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_DIR = PACKAGE_DIR.parent
CSRC = PACKAGE_DIR / "csrc"
BUILD = PACKAGE_DIR / "_build"
LIBRARY = BUILD / "libvnlbhip.so"

HIP_SOURCES = ["match.hip", "match_fast.hip", "paste.hip", "prepare.hip", "model.hip", "filter.hip", "filter_fast.hip", "aggregate.hip", "flow.hip", "scratch.hip", "proposer.hip"]
HOST_SOURCES = ["stage.cpp"]

# -ffp-contract=off: only explicit fmaf calls fuse, so the summation order
# matches the CPU restatement.  Packed FP32 ops are disabled on purpose.
COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    f"-I{REPO_DIR / 'include'}",
    f"-I{CSRC}",
]
DEVICE_FLAGS = [
    "-munsafe-fp-atomics",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: loop-carried tiles are not copied
    # between AGPRs and VGPRs around every MFMA.
    "-mllvm",
    "-amdgpu-mfma-vgpr-form=1",
]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if candidate and (Path(candidate).exists() or candidate == "hipcc"):
            return candidate
    raise RuntimeError("hipcc not found")


def _compile(source: Path, obj: Path, device: bool) -> None:
    cmd = [hipcc(), *COMMON_FLAGS, *(DEVICE_FLAGS if device else []), "-c", str(source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def build(jobs: int = 8, verbose: bool = False) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    tasks = [(CSRC / s, BUILD / (Path(s).stem + ".o"), True) for s in HIP_SOURCES]
    tasks += [(CSRC / s, BUILD / (Path(s).stem + ".o"), False) for s in HOST_SOURCES]
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        futures = [pool.submit(_compile, *t) for t in tasks]
        for f in futures:
            f.result()
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *[str(t[1]) for t in tasks]]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--jobs", type=int, default=8)
    args = parser.parse_args()
    build(args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
