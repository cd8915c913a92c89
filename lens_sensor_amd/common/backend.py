"""Backend selection for the solvers (reference: common/backend.py:9-30).

The reference picks CuPy when importable and falls back to NumPy.  Here the compute
backend is the HIP step engine in ``liblensfdtd.so`` (hand-written gfx950 kernels behind
a C ABI, loaded with ctypes).  There is no silent CPU fallback: when the library or a
GPU is missing, ``make_engine`` raises with the reason.

Host-side array work (materials, coefficients, result post-processing) stays NumPy, so
``xp`` is NumPy and ``to_np`` is the identity, as in the reference's CPU mode.

``use_engine(factory)`` installs another engine factory for the current process; the
test-suite uses it to run the same drivers on the CPU oracle.  Product code never
calls it.
"""
import contextlib
import os

import numpy as np

xp = np
cp = None
DEFAULT_MODE = os.environ.get('LENSFDTD_MODE', 'ref32')

_factory_override = None
_hip_error = None


def to_np(a):
    return np.asarray(a)


def _hip_factory():
    from lens_sensor_amd.engine import hip
    return hip.HipEngine


def hip_available():
    """True when liblensfdtd.so loads and reports at least one HIP device."""
    global _hip_error
    try:
        from lens_sensor_amd.engine import hip
        return hip.device_count() > 0
    except Exception as e:                    # library missing or no device
        _hip_error = str(e)
        return False


GPU_AVAILABLE = None     # resolved lazily by gpu_available()


def gpu_available():
    global GPU_AVAILABLE
    if GPU_AVAILABLE is None:
        GPU_AVAILABLE = hip_available()
    return GPU_AVAILABLE


def make_engine(kind, dims, mode=None):
    mode = mode or DEFAULT_MODE
    if _factory_override is not None:
        return _factory_override(kind, dims, mode)
    cls = _hip_factory()
    return cls(kind, dims, mode)


def engine_name():
    if _factory_override is not None:
        return getattr(_factory_override, '__name__', 'override')
    return 'hip'


@contextlib.contextmanager
def use_engine(factory):
    global _factory_override
    prev = _factory_override
    _factory_override = factory
    try:
        yield
    finally:
        _factory_override = prev


# Runs of at least this many steps tune the engine's temporal blocking (depth, x chunk)
# before stepping; LENSFDTD_AUTOTUNE=0 turns it off.  The blocked passes give identical
# results for every choice, so tuning changes only speed.
AUTOTUNE_MIN_STEPS = 2000
AUTOTUNE = os.environ.get('LENSFDTD_AUTOTUNE', '1') != '0'


def tune(eng, row, nsteps, cache=None, key=None):
    """Tune ``eng`` with the step record ``row`` when the run is long enough and the engine
    supports it.  ``cache`` (a dict) with ``key`` reuses one tuning for engines of the same
    shape and setup, e.g. the windows of a stitch run.  Returns (depth, lx, lead) or None."""
    if not AUTOTUNE or nsteps < AUTOTUNE_MIN_STEPS or not hasattr(eng, 'autotune'):
        return None
    if getattr(eng, 'absorbers', False) or getattr(eng, 'td', False):
        return None                          # sponge / Mur / TF/SF runs advance one step per launch
    if cache is not None and key in cache:
        depth, lx, lead = cache[key]
        eng.set_temporal(depth, lx)
        eng.set_tb_lead(lead)
        return depth, lx, lead
    depth, lx, lead, _ms = eng.autotune(row)
    if cache is not None:
        cache[key] = (depth, lx, lead)
    return depth, lx, lead
