from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import hfcpu

    hfcpu.lib()
    return hfcpu


@pytest.fixture(scope="session")
def gpu():
    """The GPU module; skips when no device is visible."""
    import cpp_hf_amd._native as native

    if native.device_count() < 1:
        pytest.skip("no GPU visible")
    return native
