import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device and the built liblensfdtd.so')
    config.addinivalue_line('markers', 'slow: longer CPU runs')


@pytest.fixture
def in_tmp(tmp_path, monkeypatch):
    """Run a solver in a temporary cwd (it writes results/)."""
    monkeypatch.chdir(tmp_path)
    return tmp_path
