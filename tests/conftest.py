import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for path in (REPO, REPO / "vapoursynth-vnlb_amd", REPO / "tests"):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0.  Fails (does not skip) without a GPU."""
    import torch

    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    import vnlbhip

    ctx = vnlbhip.Context(0)
    yield ctx
    ctx.close()
