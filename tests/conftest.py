import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and librsvdpack.so")
    config.addinivalue_line("markers", "slow: long-running case")


def _have_gpu() -> bool:
    try:
        import torch  # noqa: F401  (device count only; no HIP context)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu_lib():
    """librsvdpack loaded and initialised on the GPU (fails loudly if missing)."""
    from lowrankmatrixdecompositioncodes_amd import rsvdpack as rp
    lib = rp.load()
    rc = lib.rsvd_init()
    if rc != 0:
        pytest.fail("rsvd_init failed: " + rp.last_error())
    return rp


@pytest.fixture(scope="session")
def oracle():
    from oracle import ref
    ref.load()
    return ref
