"""Operator-level 2D CPML updates on device arrays (reference: planar_beads/fused2d.py).

The reference fuses each CPML update into one CuPy kernel per field when the arrays are
CuPy float32 C-contiguous arrays (fused2d.py:27-34), called from CPML_TE / CPML_TM
.update_H / .update_E (boundaries.py:187-188, 197-198, 241-242, 251-252).  Here the
same functions run hand-written HIP kernels of liblensfdtd.so (lf_op_* in
include/lensfdtd.h) on device arrays: torch tensors on the HIP device, float32,
contiguous, in the reference's shapes

    TE: Ez (Nx,Ny), Hx (Nx,Ny-1), Hy (Nx-1,Ny)
    TM: Hz (Nx-1,Ny-1), Ex (Nx-1,Ny), Ey (Nx,Ny-1)

updated in place on the current stream.  ``p`` is the CPML object (boundaries.CPML_TE /
CPML_TM), which holds the profiles and the psi arrays on the device.  Per element the
float32 operations are the reference's, in its order.

The loop-level engine (HipEngine.run, one call per chunk of steps) is the fast path;
these functions serve callers that keep their own time loop over device arrays.
"""
import ctypes

import numpy as np

from lens_sensor_amd.engine import hip

ENABLED = True       # switch for comparison runs, as in the reference


def _torch():
    import torch
    return torch


def available(*arrays):
    """True when every array is a float32, contiguous device tensor and the library
    loads (fused2d.py:27-34 asks the same of CuPy arrays)."""
    if not ENABLED:
        return False
    try:
        torch = _torch()
    except ImportError:
        return False
    for a in arrays:
        if not (isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == torch.float32
                and a.is_contiguous()):
            return False
    try:
        hip.load()
    except OSError:
        return False
    return True


def _ptr(a):
    return a.data_ptr()


def _stream(a):
    return _torch().cuda.current_stream(a.device).cuda_stream


def _shape(a, want, name):
    if tuple(a.shape) != tuple(want):
        raise ValueError(f'{name} has shape {tuple(a.shape)}, expected {tuple(want)}')


def te_update_H(p, Ez, Hx, Hy, Ch):
    """Hx, Hy += CPML curl of Ez (boundaries.py:186-195)."""
    Nx, Ny = Ez.shape
    _shape(Hx, (Nx, Ny - 1), 'Hx')
    _shape(Hy, (Nx - 1, Ny), 'Hy')
    p._device_state(Ez)
    op = p._op2d(Ez, p.psi_Hxy, p.psi_Hyx)
    hip.op_call('lf_op_te_update_H', op, Nx, Ny, _ptr(Ez), _ptr(Hx), _ptr(Hy),
                float(np.float32(Ch)), _stream(Ez))


def te_update_E(p, Ez, Hx, Hy, Ce_E, Ce_H):
    """Interior Ez <- Ce_E Ez + Ce_H (CPML curl of H) (boundaries.py:197-205)."""
    Nx, Ny = Ez.shape
    _shape(Ce_E, (Nx, Ny), 'Ce_E')
    _shape(Ce_H, (Nx, Ny), 'Ce_H')
    p._device_state(Ez)
    op = p._op2d(Ez, p.psi_Ezx, p.psi_Ezy)
    hip.op_call('lf_op_te_update_E', op, Nx, Ny, _ptr(Ez), _ptr(Hx), _ptr(Hy), _ptr(Ce_E),
                _ptr(Ce_H), _stream(Ez))


def tm_update_H(p, Ex, Ey, Hz, Ch):
    """Hz -= Ch (CPML curl of E) (boundaries.py:240-248)."""
    Nx, Ny = Ey.shape[0], Ex.shape[1]
    _shape(Ex, (Nx - 1, Ny), 'Ex')
    _shape(Hz, (Nx - 1, Ny - 1), 'Hz')
    p._device_state(Ex)
    op = p._op2d(Ex, p.psi_Hzx, p.psi_Hzy)
    hip.op_call('lf_op_tm_update_H', op, Nx, Ny, _ptr(Ex), _ptr(Ey), _ptr(Hz),
                float(np.float32(Ch)), _stream(Ex))


def tm_update_E(p, Ex, Ey, Hz, Ca_ex, Ca_ey):
    """Interior Ex, Ey += Ca (CPML curl of Hz) (boundaries.py:250-258)."""
    Nx, Ny = Ey.shape[0], Ex.shape[1]
    _shape(Ey, (Nx, Ny - 1), 'Ey')
    _shape(Ca_ex, (Nx - 1, Ny), 'Ca_ex')
    _shape(Ca_ey, (Nx, Ny - 1), 'Ca_ey')
    p._device_state(Ex)
    op = p._op2d(Ex, p.psi_Exy, p.psi_Eyx)
    hip.op_call('lf_op_tm_update_E', op, Nx, Ny, _ptr(Ex), _ptr(Ey), _ptr(Hz), _ptr(Ca_ex),
                _ptr(Ca_ey), _stream(Ex))


def te_terminate(p, Ez):
    """PEC on the ends that carry a PML (boundaries.py:207-215)."""
    Nx, Ny = Ez.shape
    pec = _int4(p.pec[0] + p.pec[1])
    hip.op_call('lf_op_te_terminate', Nx, Ny, _ptr(Ez), pec, _stream(Ez))


def tm_terminate(p, Ex, Ey):
    """PEC on the ends that carry a PML (boundaries.py:260-268)."""
    Nx, Ny = Ey.shape[0], Ex.shape[1]
    pec = _int4(p.pec[0] + p.pec[1])
    hip.op_call('lf_op_tm_terminate', Nx, Ny, _ptr(Ex), _ptr(Ey), pec, _stream(Ex))


def _int4(flags):
    return (ctypes.c_int * 4)(*[int(bool(f)) for f in flags])
