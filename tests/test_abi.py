"""The C-ABI library loads and exports every entry point include/lensfdtd.h declares
(no compute call: this runs on the CPU-only machine too)."""
import ctypes
import os
import re

import pytest

from lens_sensor_amd.engine import hip

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'include',
                      'lensfdtd.h')


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char \*)\s*(lf_\w+)\s*\(', text, re.M)))


def test_header_declares_the_binding_set():
    assert set(_declared()) == set(hip.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(hip.library_path()):
        pytest.skip('liblensfdtd.so not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(hip.library_path())
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_step_struct_matches_the_header():
    from lens_sensor_amd.engine.spec import STEP_DTYPE
    assert STEP_DTYPE.itemsize == 120
    text = open(HEADER).read()
    body = text[text.index('typedef struct lf_step'):text.index('} lf_step;')]
    doubles = sum(len(m.split(',')) for m in re.findall(r'double ([^;]+);', body))
    floats = sum(len(m.split(',')) for m in re.findall(r'float ([^;]+);', body))
    uints = sum(len(m.split(',')) for m in re.findall(r'unsigned int ([^;]+);', body))
    assert doubles * 8 + floats * 4 + uints * 4 == STEP_DTYPE.itemsize
    names = [n for n in STEP_DTYPE.names]
    assert names[:9] == ['env', 'ssin', 'scos', 'pr', 'pi', 'hr', 'hi', 'dr', 'di']
