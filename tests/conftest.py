import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


_EMU_LIB = os.environ.get('LENSFDTD_EMU_LIB')
if _EMU_LIB:
    # CPU bounds-checking runs: the GPU tests drive the library built for the CPU under
    # AddressSanitizer (tests/emu/build.py) instead of liblensfdtd.so
    from lens_sensor_amd.engine import hip as _hip
    _hip._LIB_PATH = _EMU_LIB


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device and the built liblensfdtd.so')
    config.addinivalue_line('markers', 'slow: longer CPU runs')


@pytest.fixture
def in_tmp(tmp_path, monkeypatch):
    """Run a solver in a temporary cwd (it writes results/)."""
    monkeypatch.chdir(tmp_path)
    return tmp_path
