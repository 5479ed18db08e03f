import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _ensure_oracle():
    lib = ROOT / "oracle" / "lib" / "libmcl_oracle.so"
    if not lib.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, stdout=subprocess.DEVNULL)
    return lib


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The device library; GPU tests fail loudly if it is missing or no device is usable."""
    from thoth_amd import mcl
    mcl.lib()
    assert mcl.available(), "no usable gfx950 device: " + mcl.device_info()
    return mcl
