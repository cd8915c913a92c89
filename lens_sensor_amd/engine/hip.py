"""ctypes binding of liblensfdtd.so: the HIP step engine behind the drivers.

Implements the step-engine contract of lens_sensor_amd/engine/spec.py on the C ABI of
include/lensfdtd.h.  The library must be built (lens_sensor_amd/engine/build.py,
__graft_entry__.build()); a missing library or device raises, there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from lens_sensor_amd.engine.spec import STEP_DTYPE, comp_shape, resolve_box

_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         'liblensfdtd.so')
_lib = None

KIND = {'te2d': 0, 'tm2d': 1, 'yee3d': 2}
MODE = {'ref32': 0, 'fp64acc': 1}
ARR = {'Ex': 0, 'Ey': 1, 'Ez': 2, 'Hx': 3, 'Hy': 4, 'Hz': 5, 'ceE': 6, 'ceH': 7, 'cax': 8,
       'cay': 9, 'cex': 10, 'cey': 11, 'cez': 12, 'dax': 13, 'day': 14, 'daz': 15}
SQ = {'te': 0, 'tm': 1, 'tm_ex2': 2, '3d_s': 3, '3d_p': 4}
SRC = {'Ex': 0, 'Ey': 1, 'Ez': 2, 'ey_node': 3, 'ex_node': 4}
ACC = {'c64p': 0, 'c64': 1, 'split32': 1, 'c128': 2}

EXPORTS = ('lf_create', 'lf_destroy', 'lf_last_error', 'lf_device_count', 'lf_device_name',
           'lf_upload', 'lf_download', 'lf_set_ch', 'lf_set_cpml', 'lf_finalize',
           'lf_set_source', 'lf_set_drive', 'lf_add_box', 'lf_add_sq', 'lf_add_dft',
           'lf_add_flux', 'lf_monitor_bytes', 'lf_monitor_read', 'lf_monitor_write',
           'lf_monitor_zero', 'lf_run', 'lf_sync', 'lf_set_timing', 'lf_kernel_ms',
           'lf_set_variant', 'lf_set_tile', 'lf_set_temporal', 'lf_tb_stats',
           'lf_batch_create', 'lf_batch_add', 'lf_batch_set_temporal', 'lf_batch_set_timing',
           'lf_batch_run', 'lf_batch_stats', 'lf_batch_destroy', 'lf_batch_set_table',
           'lf_batch_run_at', 'lf_autotune', 'lf_set_temporal_mon', 'lf_set_tb_lead',
           'lf_op_te_update_H', 'lf_op_te_update_E', 'lf_op_tm_update_H', 'lf_op_tm_update_E',
           'lf_op_te_terminate', 'lf_op_tm_terminate', 'lf_op3d_create', 'lf_op3d_set_cpml',
           'lf_op3d_finalize', 'lf_op3d_step_H', 'lf_op3d_step_E', 'lf_op3d_destroy',
           'lf_monitor_change', 'lf_norm_diff', 'lf_norm_diff_noalloc', 'lf_set_sponge',
           'lf_add_mur', 'lf_build_coeff', 'lf_add_box_form', 'lf_add_sx', 'lf_add_rec',
           'lf_add_tfsf', 'lf_load_tfsf', 'lf_field_maxabs')


class OpProfiles2D(ctypes.Structure):
    """lf_op2d: device pointers of the 1-D CPML profiles and of the two psi arrays of an
    operator-level 2D update (include/lensfdtd.h)."""
    _fields_ = [(n, ctypes.c_void_p) for n in
                ('bex', 'cex', 'kex', 'bhx', 'chx', 'khx', 'bey', 'cey', 'key', 'bhy', 'chy', 'khy',
                 'psi0', 'psi1')]


def library_path():
    return _LIB_PATH


def load():
    """Load liblensfdtd.so (raises OSError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise OSError(f'{_LIB_PATH} is not built (run __graft_entry__.build())')
    # One HIP runtime per process.  torch ships its own libamdhip64 with the same SONAME
    # (libamdhip64.so.7) as the ROCm one this library links, but requests it by another
    # file name: loaded after this library it maps a second runtime, which finds no GPU.
    # Loading torch first binds this library to torch's runtime, so the device pointers
    # and stream handles of torch tensors (the operator-level API) are valid here too.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(_LIB_PATH)
    P, I, LL, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        'lf_create': [I, I, I, I, I, ctypes.POINTER(P)],
        'lf_destroy': [P],
        'lf_device_count': [ip],
        'lf_device_name': [ctypes.c_char_p, I],
        'lf_upload': [P, I, fp, LL],
        'lf_download': [P, I, fp, LL],
        'lf_set_ch': [P, F],
        'lf_set_cpml': [P, I, fp, fp, fp, fp, fp, fp, I, I],
        'lf_finalize': [P],
        'lf_set_source': [P, I, I, I, fp, fp, fp],
        'lf_set_drive': [P, I, fp, I, ip, fp, I],
        'lf_add_box': [P, ip, ip],
        'lf_add_sq': [P, I, I, ip],
        'lf_add_dft': [P, I, I, I, I, I, I, ip],
        'lf_add_flux': [P, I, ip],
        'lf_monitor_bytes': [P, I, ctypes.POINTER(LL)],
        'lf_monitor_read': [P, I, P, LL],
        'lf_monitor_write': [P, I, P, LL],
        'lf_monitor_zero': [P, I],
        'lf_run': [P, P, I],
        'lf_sync': [P],
        'lf_set_timing': [P, I],
        'lf_set_variant': [P, I, I],
        'lf_set_tile': [P, I],
        'lf_set_temporal': [P, I, I],
        'lf_set_temporal_mon': [P, I],
        'lf_tb_stats': [P, ctypes.POINTER(ctypes.c_double), ip, ip],
        'lf_autotune': [P, P, I, ip, ip, ip, ctypes.POINTER(ctypes.c_double)],
        'lf_set_tb_lead': [P, I],
        'lf_kernel_ms': [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ip],
        'lf_batch_create': [ctypes.POINTER(P)],
        'lf_batch_add': [P, P],
        'lf_batch_set_temporal': [P, I, I],
        'lf_batch_set_timing': [P, I],
        'lf_batch_run': [P, P, I, P, P],
        'lf_batch_stats': [P, ctypes.POINTER(ctypes.c_double), ip],
        'lf_batch_destroy': [P],
        'lf_batch_set_table': [P, P, I],
        'lf_batch_run_at': [P, P, I, P, P],
        'lf_op_te_update_H': [ctypes.POINTER(OpProfiles2D), I, I, P, P, P, F, P],
        'lf_op_te_update_E': [ctypes.POINTER(OpProfiles2D), I, I, P, P, P, P, P, P],
        'lf_op_tm_update_H': [ctypes.POINTER(OpProfiles2D), I, I, P, P, P, F, P],
        'lf_op_tm_update_E': [ctypes.POINTER(OpProfiles2D), I, I, P, P, P, P, P, P],
        'lf_op_te_terminate': [I, I, P, ip, P],
        'lf_op_tm_terminate': [I, I, P, P, ip, P],
        'lf_op3d_create': [I, I, I, ctypes.POINTER(P)],
        'lf_op3d_set_cpml': [P, I, fp, fp, fp, fp, fp, fp, I, I],
        'lf_op3d_finalize': [P],
        'lf_op3d_step_H': [P, P, P, P, P, P, P, F, P],
        'lf_op3d_step_E': [P, P, P, P, P, P, P, P, P, P, P, P, P, P],
        'lf_op3d_destroy': [P],
        'lf_monitor_change': [P, I, I, ctypes.c_double, ctypes.c_double,
                              ctypes.POINTER(ctypes.c_double)],
        'lf_norm_diff': [P, P, LL, I, ctypes.c_double, ctypes.c_double, P, P],
        'lf_norm_diff_noalloc': [P, P, LL, I, ctypes.c_double, ctypes.c_double, P, P],
        'lf_set_sponge': [P, I, fp, ctypes.c_uint],
        'lf_add_mur': [P, I, I, I, fp, LL, I],
        'lf_build_coeff': [P, I, ctypes.POINTER(ctypes.c_ubyte), I, fp, I, I, fp],
        'lf_add_box_form': [P, ip, I, ip],
        'lf_add_sx': [P, I, ip],
        'lf_add_rec': [P, I, I, I, I, ip],
        'lf_add_tfsf': [P, I, I, I, I, I, I, I, I, I, ip],
        'lf_load_tfsf': [P, I, fp, I],
        'lf_field_maxabs': [P, I, fp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = I
    lib.lf_last_error.restype = ctypes.c_char_p
    lib.lf_last_error.argtypes = []
    _lib = lib
    return lib


def op_call(name, *args):
    """Call an operator-level entry point (lf_op_*, lf_op3d_*) and raise on error."""
    lib = load()
    _check(getattr(lib, name)(*args), name)


def _check(rc, what):
    if rc != 0:
        msg = _lib.lf_last_error().decode(errors='replace')
        raise RuntimeError(f'{what} failed ({rc}): {msg}')


def device_count():
    lib = load()
    n = ctypes.c_int(0)
    rc = lib.lf_device_count(ctypes.byref(n))
    if rc != 0:
        return 0
    return n.value


def device_name():
    lib = load()
    buf = ctypes.create_string_buffer(256)
    _check(lib.lf_device_name(buf, 256), 'lf_device_name')
    return buf.value.decode()


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class HipEngine:
    """One HIP context: the Yee domain of a run (or stitch window) resident in HBM."""

    def __init__(self, kind, dims, mode='ref32'):
        if mode not in MODE:
            raise ValueError(f"HIP engine modes: ref32 | fp64acc (got {mode!r})")
        lib = load()
        if device_count() < 1:
            raise RuntimeError('no HIP device visible: the HIP step engine needs an MI355X')
        self.kind, self.dims, self.mode = kind, tuple(int(d) for d in dims), mode
        nz = self.dims[2] if kind == 'yee3d' else 1
        h = ctypes.c_void_p()
        _check(lib.lf_create(KIND[kind], MODE[mode], self.dims[0], self.dims[1], nz,
                             ctypes.byref(h)), 'lf_create')
        self.h = h
        self._lib = lib
        self.mon_kind = []
        self.absorbers = False          # sponge / Mur faces set: one step per launch
        self.td = False                 # TF/SF corrections or plane records: one step per launch

    @property
    def wide_acc(self):
        return self.mode != 'ref32'

    def close(self):
        if getattr(self, 'h', None):
            self._lib.lf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(self._lib.lf_sync(self.h), 'lf_sync')

    # ---------------------------------------------------------------- setup
    def _upload(self, name, arr):
        shp = comp_shape(self.kind, self.dims, {'ceE': 'node', 'ceH': 'node', 'cax': 'Ex',
                                                 'cay': 'Ey', 'cex': 'Ex', 'cey': 'Ey',
                                                 'cez': 'Ez', 'dax': 'Ex', 'day': 'Ey',
                                                 'daz': 'Ez'}.get(name, name))
        a = np.ascontiguousarray(arr, dtype=np.float32)
        if a.shape != shp:
            raise ValueError(f'{name}: shape {a.shape} != {shp}')
        _check(self._lib.lf_upload(self.h, ARR[name], _fptr(a), a.size), f'lf_upload({name})')

    def set_coeffs(self, ch, **arrays):
        _check(self._lib.lf_set_ch(self.h, float(np.float32(ch))), 'lf_set_ch')
        for k, v in arrays.items():
            if v is not None:
                self._upload(k, v)

    def build_coeffs(self, ch, packed):
        """Coefficient arrays assembled on the device from their compact form
        (fdtd2d_core.window_coeffs: name -> (xclass, table, j0, band))."""
        _check(self._lib.lf_set_ch(self.h, float(np.float32(ch))), 'lf_set_ch')
        for name, (xc, table, j0, band) in packed.items():
            xc = np.ascontiguousarray(xc, dtype=np.uint8)
            t = np.ascontiguousarray(table, dtype=np.float32)
            b = np.ascontiguousarray(band, dtype=np.float32)
            _check(self._lib.lf_build_coeff(self.h, ARR[name],
                                            xc.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                            t.shape[0], _fptr(t), int(j0), b.shape[1], _fptr(b)),
                   f'lf_build_coeff({name})')

    def set_cpml(self, axis, e_prof, h_prof, pec_lo, pec_hi):
        ek, eb, ec = (np.ascontiguousarray(x, dtype=np.float32) for x in e_prof)
        hk, hb, hc = (np.ascontiguousarray(x, dtype=np.float32) for x in h_prof)
        _check(self._lib.lf_set_cpml(self.h, axis, _fptr(ek), _fptr(eb), _fptr(ec), _fptr(hk),
                                     _fptr(hb), _fptr(hc), int(bool(pec_lo)), int(bool(pec_hi))),
               'lf_set_cpml')

    def finalize(self):
        _check(self._lib.lf_finalize(self.h), 'lf_finalize')

    def set_source(self, comp, ix, amp, ph1, ph2):
        a = np.ascontiguousarray(amp, dtype=np.float32).ravel()
        p1 = np.ascontiguousarray(ph1, dtype=np.float32).ravel()
        p2 = np.ascontiguousarray(ph2, dtype=np.float32).ravel()
        _check(self._lib.lf_set_source(self.h, ARR[comp], int(ix), a.size, _fptr(a), _fptr(p1),
                                       _fptr(p2)), 'lf_set_source')

    def set_drive(self, ncells, wtap, drives, math):
        comps = list(drives)
        cid = (ctypes.c_int * len(comps))(*[ARR[c] for c in comps])
        buf = np.concatenate([np.ascontiguousarray(drives[c], dtype=np.complex64).view(np.float32).ravel()
                              for c in comps])
        w = np.ascontiguousarray(wtap, dtype=np.float32)
        _check(self._lib.lf_set_drive(self.h, int(ncells), _fptr(w), len(comps), cid, _fptr(buf),
                                      0 if math == 'c128' else 1), 'lf_set_drive')

    def set_sponge(self, n, g, faces):
        """Graded sponge (3D): n cells, factors g (g[0] outermost), faces bitmask
        (bit 2*axis + hi)."""
        a = np.ascontiguousarray(g, dtype=np.float32).ravel()
        if a.size != int(n):
            raise ValueError('sponge: need n damping factors')
        _check(self._lib.lf_set_sponge(self.h, int(n), _fptr(a), int(faces)), 'lf_set_sponge')
        self.absorbers = True

    def add_mur(self, comp, axis, hi, m, flags):
        """First-order Mur on face (axis, hi) of E component comp; m scalar or one value
        per cell of the boundary plane; faces apply in the order added."""
        a = np.ascontiguousarray(np.atleast_1d(m), dtype=np.float32).ravel()
        _check(self._lib.lf_add_mur(self.h, ARR[comp], int(axis), int(bool(hi)), _fptr(a), a.size,
                                    int(flags)), 'lf_add_mur')
        self.absorbers = True

    def write_field(self, comp, arr):
        self._upload(comp, arr)

    def read_coeff(self, name):
        """A per-cell coefficient array (reference shape) as the engine holds it."""
        shp = comp_shape(self.kind, self.dims, {'ceE': 'node', 'ceH': 'node', 'cax': 'Ex',
                                                 'cay': 'Ey', 'dax': 'Ex', 'day': 'Ey'}[name])
        out = np.empty(shp, dtype=np.float32)
        _check(self._lib.lf_download(self.h, ARR[name], _fptr(out), out.size), f'lf_download({name})')
        return out

    def read_field(self, comp):
        shp = comp_shape(self.kind, self.dims, comp)
        out = np.empty(shp, dtype=np.float32)
        _check(self._lib.lf_download(self.h, ARR[comp], _fptr(out), out.size), 'lf_download')
        return out

    # ---------------------------------------------------------------- monitors
    def _new_id(self, rc, mid, info):
        _check(rc, 'lf_add_monitor')
        self.mon_kind.append(info)
        return mid.value

    def add_box(self, box, form=0):
        """form 1 (2D TM): the node Ey^2 alone instead of the node-grid I_E."""
        mid = ctypes.c_int()
        if self.kind == 'yee3d':
            r = []
            for c in ('Ex', 'Ey', 'Ez'):
                for lo, hi in resolve_box(comp_shape(self.kind, self.dims, c), box):
                    r += [lo, hi]
        else:
            (a0, a1), (b0, b1) = resolve_box(self.dims, box)
            r = [a0, a1, b0, b1] + [0] * 14
        arr = (ctypes.c_int * 18)(*r)
        rc = (self._lib.lf_add_box_form(self.h, arr, int(form), ctypes.byref(mid)) if form
              else self._lib.lf_add_box(self.h, arr, ctypes.byref(mid)))
        return self._new_id(rc, mid, dict(t='box'))

    # -- time-domain stitch
    def add_sx(self, il):
        """FP64 sum over the masked steps of the instantaneous flux in +x through plane il."""
        mid = ctypes.c_int()
        return self._new_id(self._lib.lf_add_sx(self.h, int(il), ctypes.byref(mid)), mid,
                            dict(t='box'))

    def add_rec(self, comp, il, phase, rows):
        """Plane il of comp into row step['row'] of a (rows,) + plane buffer on the masked
        steps; phase 0 at the start of the step, 1 after the H update and its TF/SF
        corrections."""
        mid = ctypes.c_int()
        shp = (int(rows),) + comp_shape(self.kind, self.dims, comp)[1:]
        rc = self._lib.lf_add_rec(self.h, ARR[comp], int(il), int(phase), int(rows),
                                  ctypes.byref(mid))
        self.td = True
        return self._new_id(rc, mid, dict(t='rec', shape=shp))

    def add_tfsf(self, comp, il, sign, coef, rng, rows):
        """TF/SF correction F[il][rng] +-= c inc[row][rng] (c = Ch for coef None, else the
        coefficient array coef at the cell); rng: one (lo, hi) per plane axis.  Returns its
        index for load_tfsf."""
        if self.kind == 'yee3d':
            (j0, j1), (k0, k1) = rng
        else:
            (j0, j1), (k0, k1) = (0, 1), rng[0]
        idx = ctypes.c_int()
        _check(self._lib.lf_add_tfsf(self.h, ARR[comp], int(il), int(sign),
                                     -1 if coef is None else ARR[coef], j0, j1, k0, k1, int(rows),
                                     ctypes.byref(idx)), 'lf_add_tfsf')
        self.td = True
        return idx.value

    def load_tfsf(self, idx, rows):
        """Rows 0..len(rows)-1 of a TF/SF correction's incident series."""
        a = np.ascontiguousarray(rows, dtype=np.float32)
        _check(self._lib.lf_load_tfsf(self.h, int(idx), _fptr(a), len(a)), 'lf_load_tfsf')

    def field_maxabs(self, comp):
        out = ctypes.c_float()
        _check(self._lib.lf_field_maxabs(self.h, ARR[comp], ctypes.byref(out)), 'lf_field_maxabs')
        return float(out.value)

    def add_sq(self, form, nx_acc=None):
        mid = ctypes.c_int()
        if self.kind == 'yee3d':
            nxa = self.dims[0] if nx_acc is None else int(nx_acc)
            tgt = 'Ez' if form == '3d_s' else 'Ey'
            shp = (nxa,) + comp_shape(self.kind, self.dims, tgt)[1:]
        else:
            nxa = self.dims[0]
            shp = tuple(self.dims)
        return self._new_id(self._lib.lf_add_sq(self.h, SQ[form], nxa, ctypes.byref(mid)), mid,
                            dict(t='sq', shape=shp))

    def add_dft(self, src, x0, x1, kplane=None, phasor='p', acc='c64p'):
        mid = ctypes.c_int()
        if self.mode != 'ref32' and acc != 'c128':
            acc = 'c128'
        if src in ('ey_node', 'ex_node'):
            base = self.dims
        else:
            base = comp_shape(self.kind, self.dims, src)
        shp = (x1 - x0,) + tuple(base[1:])
        if kplane is not None:
            shp = shp[:2]
        dt = np.complex128 if acc == 'c128' else np.complex64
        ph = 1 if phasor == 'q' else 0
        rc = self._lib.lf_add_dft(self.h, SRC[src], int(x0), int(x1),
                                  -1 if kplane is None else int(kplane), ph, ACC[acc],
                                  ctypes.byref(mid))
        return self._new_id(rc, mid, dict(t='dft', shape=shp, dtype=dt))

    def add_flux_plane(self, il):
        from lens_sensor_amd.common.flux import PAIRS
        mid = ctypes.c_int()
        dt = np.complex128 if self.wide_acc else np.complex64
        shapes = [comp_shape(self.kind, self.dims, ec)[1:] for _s, ec, _h in PAIRS[self.kind]]
        return self._new_id(self._lib.lf_add_flux(self.h, int(il), ctypes.byref(mid)), mid,
                            dict(t='flux', shapes=shapes, dtype=dt))

    def _bytes(self, mid):
        b = ctypes.c_longlong()
        _check(self._lib.lf_monitor_bytes(self.h, mid, ctypes.byref(b)), 'lf_monitor_bytes')
        return b.value

    def read_monitor(self, mid):
        info = self.mon_kind[mid]
        nb = self._bytes(mid)
        raw = np.empty(nb, dtype=np.uint8)
        _check(self._lib.lf_monitor_read(self.h, mid, raw.ctypes.data, nb), 'lf_monitor_read')
        if info['t'] == 'box':
            return float(np.sum(raw.view(np.float64)))
        if info['t'] in ('sq', 'rec'):
            return raw.view(np.float32).reshape(info['shape']).copy()
        if info['t'] == 'dft':
            return raw.view(info['dtype']).reshape(info['shape']).copy()
        out, off = [], 0
        it = np.dtype(info['dtype']).itemsize
        for shp in info['shapes']:
            n = int(np.prod(shp)) * it
            aE = raw[off:off + n].view(info['dtype']).reshape(shp).copy()
            off += n
            aH = raw[off:off + n].view(info['dtype']).reshape(shp).copy()
            off += n
            out.append((aE, aH))
        return out

    def read_flux(self, mid):
        return self.read_monitor(mid)

    def write_monitor(self, mid, value):
        info = self.mon_kind[mid]
        if info['t'] == 'box':
            raw = np.array([float(value)], dtype=np.float64).view(np.uint8)
            raw = np.concatenate([raw, np.zeros(self._bytes(mid) - 8, np.uint8)])
        elif info['t'] == 'flux':
            parts = []
            for aE, aH in value:
                parts.append(np.ascontiguousarray(aE, dtype=info['dtype']).view(np.uint8).ravel())
                parts.append(np.ascontiguousarray(aH, dtype=info['dtype']).view(np.uint8).ravel())
            raw = np.concatenate(parts)
        elif info['t'] in ('sq', 'rec'):
            raw = np.ascontiguousarray(value, dtype=np.float32).view(np.uint8).ravel()
        else:
            raw = np.ascontiguousarray(value, dtype=info['dtype']).view(np.uint8).ravel()
        raw = np.ascontiguousarray(raw)
        _check(self._lib.lf_monitor_write(self.h, mid, raw.ctypes.data, raw.size), 'lf_monitor_write')

    def zero_monitor(self, mid):
        _check(self._lib.lf_monitor_zero(self.h, mid), 'lf_monitor_zero')

    def monitor_change(self, mid_a, mid_b, da=1.0, db=1.0):
        """(sum |A/da - B/db|^2, sum |A/da|^2) of two DFT monitors, reduced on the device
        in FP64 (two doubles cross to the host)."""
        out = (ctypes.c_double * 2)()
        _check(self._lib.lf_monitor_change(self.h, mid_a, mid_b, float(da), float(db), out),
               'lf_monitor_change')
        return out[0], out[1]

    # ---------------------------------------------------------------- stepping
    def run(self, tab):
        tab = np.ascontiguousarray(tab, dtype=STEP_DTYPE)
        _check(self._lib.lf_run(self.h, tab.ctypes.data, len(tab)), 'lf_run')

    def set_variant(self, variant, lx=32):
        _check(self._lib.lf_set_variant(self.h, int(variant), int(lx)), 'lf_set_variant')

    def set_tile(self, ty):
        _check(self._lib.lf_set_tile(self.h, int(ty)), 'lf_set_tile')

    def set_temporal(self, depth, lx=64):
        _check(self._lib.lf_set_temporal(self.h, int(depth), int(lx)), 'lf_set_temporal')

    def set_temporal_mon(self, depth):
        """Depth of the 2D passes that also sum the CW-window monitors (1 = off)."""
        _check(self._lib.lf_set_temporal_mon(self.h, int(depth)), 'lf_set_temporal_mon')

    def set_tb_lead(self, cls):
        """Pass class dispatched first in the 2D blocked passes (0 slab-free, 1 y-slab rows,
        2 x-slab planes, 3 excitation planes; 4 every class in turn on one stream); a speed
        knob, results are identical."""
        _check(self._lib.lf_set_tb_lead(self.h, int(cls)), 'lf_set_tb_lead')

    def autotune(self, row, reps=None):
        """Time the candidate blocking depths / x chunks (2D: then the lead class) on this
        engine with monitor-free copies of step row `row` (one lf_step record); keeps the
        fastest and restores the fields.  Returns (depth, lx, lead, ms per step)."""
        row = np.ascontiguousarray(np.asarray(row, dtype=STEP_DTYPE).reshape(1))
        d, lx, ld, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        if reps is None:
            reps = 3 if self.kind == 'yee3d' else 8
        _check(self._lib.lf_autotune(self.h, row.ctypes.data, int(reps), ctypes.byref(d),
                                     ctypes.byref(lx), ctypes.byref(ld), ctypes.byref(ms)),
               'lf_autotune')
        return d.value, lx.value, ld.value, ms.value

    def tb_stats(self):
        ms, n, d = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _check(self._lib.lf_tb_stats(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(d)),
               'lf_tb_stats')
        return ms.value, n.value, d.value

    def set_timing(self, on=True):
        _check(self._lib.lf_set_timing(self.h, int(bool(on))), 'lf_set_timing')

    def kernel_ms(self):
        u, t, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(self._lib.lf_kernel_ms(self.h, ctypes.byref(u), ctypes.byref(t), ctypes.byref(n)),
               'lf_kernel_ms')
        return u.value, t.value, n.value


class HipBatch:
    """Many 2D HipEngine instances advanced together through the batch engine
    (lf_batch_*): every temporally blocked pass is one launch per wavefront class over all
    instances that take it.  The engines must stay open while the batch exists."""

    def __init__(self, engines=(), depth=6, lx=64):
        lib = load()
        h = ctypes.c_void_p()
        _check(lib.lf_batch_create(ctypes.byref(h)), 'lf_batch_create')
        self.h, self._lib, self.engines = h, lib, []
        self.set_temporal(depth, lx)
        for e in engines:
            self.add(e)

    def add(self, eng):
        _check(self._lib.lf_batch_add(self.h, eng.h), 'lf_batch_add')
        self.engines.append(eng)

    def set_temporal(self, depth, lx=64):
        _check(self._lib.lf_batch_set_temporal(self.h, int(depth), int(lx)), 'lf_batch_set_temporal')

    def set_timing(self, on=True):
        _check(self._lib.lf_batch_set_timing(self.h, int(bool(on))), 'lf_batch_set_timing')

    def run(self, tab, masks=None, nact=None):
        """Advance every instance over the shared step table `tab`; masks: (ninst, nsteps)
        monitor masks (default tab['mask']); nact: steps per instance (default all)."""
        tab = np.ascontiguousarray(tab, dtype=STEP_DTYPE)
        n = len(tab)
        mp = ap = None
        if masks is not None:
            m = np.ascontiguousarray(masks, dtype=np.uint32)
            if m.shape != (len(self.engines), n):
                raise ValueError(f'masks must be ({len(self.engines)}, {n})')
            mp = m.ctypes.data
        if nact is not None:
            a = np.ascontiguousarray(nact, dtype=np.int32)
            if a.shape != (len(self.engines),):
                raise ValueError('nact needs one entry per instance')
            ap = a.ctypes.data
        _check(self._lib.lf_batch_run(self.h, tab.ctypes.data, n, mp, ap), 'lf_batch_run')

    def set_table(self, tab):
        """Upload the resident step table used by run_at."""
        self._table = np.ascontiguousarray(tab, dtype=STEP_DTYPE)
        _check(self._lib.lf_batch_set_table(self.h, self._table.ctypes.data, len(self._table)),
               'lf_batch_set_table')

    def run_at(self, offs, nsteps, masks=None, nact=None):
        """Instance i takes table[offs[i] + s] for s < nact[i] <= nsteps."""
        ni = len(self.engines)
        o = np.ascontiguousarray(offs, dtype=np.int32)
        if o.shape != (ni,):
            raise ValueError('offs needs one entry per instance')
        mp = ap = None
        if masks is not None:
            m = np.ascontiguousarray(masks, dtype=np.uint32)
            if m.shape != (ni, nsteps):
                raise ValueError(f'masks must be ({ni}, {nsteps})')
            mp = m.ctypes.data
        if nact is not None:
            a = np.ascontiguousarray(nact, dtype=np.int32)
            if a.shape != (ni,):
                raise ValueError('nact needs one entry per instance')
            ap = a.ctypes.data
        _check(self._lib.lf_batch_run_at(self.h, o.ctypes.data, int(nsteps), mp, ap), 'lf_batch_run_at')

    def stats(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(self._lib.lf_batch_stats(self.h, ctypes.byref(ms), ctypes.byref(n)), 'lf_batch_stats')
        return ms.value, n.value

    def close(self):
        if getattr(self, 'h', None):
            self._lib.lf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
