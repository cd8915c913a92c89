"""CPU oracle of the Yee time-stepping path (NumPy).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product path (lens_sensor_amd) never imports it.

It is a restatement of the reference's per-step arithmetic with the same NumPy
expressions (and therefore the same float32 rounding and dtype promotion), behind the
step-engine contract of lens_sensor_amd/engine/spec.py so that the package drivers can
run on it unchanged:

  2D TE update_H / update_E / terminate   planar_beads/boundaries.py:186-215
  2D TM update_H / update_E / terminate   planar_beads/boundaries.py:240-268
  TM loss E <- D E before the E update     planar_beads/fdtd2d_core.py:796-797
  3D step_H / step_E (psi strips, PEC)     planar_3d/boundaries3d.py:240-251, 311-362
  Huygens inject                           common/vcsel.py:151-156
  stitch overlap imposition                fdtd2d_core.py:1426-1427, fdtd3d_core.py:1626-1630
  detector box sums                        fdtd2d_core.py:835-837, fdtd3d_core.py:477-481
  <|E|^2> accumulators                     fdtd2d_core.py:848-850, fdtd3d_core.py:438-474
  CW phasor DFTs                           fdtd2d_core.py:838-847, 1434-1441; fdtd3d_core.py:1150-1160, 1636-1648
  flux running sums                        common/flux.py:30-43
  time-domain stitch: plane records, TF/SF corrections, instantaneous Poynting sums,
  max |F|                                  fdtd2d_core.py:1767-1835, fdtd3d_core.py:1966-2004

Modes: 'ref32' is the reference's float32 path (pinned against the golden master of
tests/test_characterization.py); 'f64' runs every array in float64 (the FP64 oracle,
same coefficients upcast); 'fp64acc' is float32 fields with the FP64 phase source and
complex128 accumulators (the device's long-run mode).

Pinning (see tests/test_oracle_*.py): the golden-master values of the reference at
RTOL 1e-6 (2D TE full, 3D TE full, 3D TE stitch), plus invariants that fail for a
dropped term, a wrong sign or index: discrete div(H) conservation, linearity in the
source amplitude, plane-wave numerical dispersion, energy decay in a closed CPML box,
and the f64 mode agreeing with ref32 to float32 accuracy.
"""
import numpy as np

from lens_sensor_amd.engine.spec import (STEP_DTYPE, comp_shape, resolve_box, E_COMPS, H_COMPS)


def _strip_range(c):
    """(lo_end, hi_start) of the PML slabs of a profile: indices < lo_end or >= hi_start
    have c != 0 (boundaries3d.py:216-224 scans c the same way; 'full' if they meet)."""
    cn = np.abs(np.asarray(c, dtype=np.float64))
    N = len(cn)
    nlo = 0
    while nlo < N and cn[nlo] > 0.0:
        nlo += 1
    nhi = 0
    while nhi < N and cn[N - 1 - nhi] > 0.0:
        nhi += 1
    if nlo + nhi >= N:
        return N, N
    return nlo, N - nhi


class _Strip:
    """psi of one stretch axis, held only in its PML slabs (boundaries3d.py:201-251)."""

    def __init__(self, shape, axis, b, c, k, dtype):
        self.parts = []
        N = shape[axis]
        lo_end, hi_start = _strip_range(c)
        spans = []
        if lo_end > 0:
            spans.append((0, lo_end))
        if hi_start < N:
            spans.append((hi_start, N))

        def bc(v):
            return v.reshape(tuple(-1 if d == axis else 1 for d in range(len(shape))))
        for a0, a1 in spans:
            sl = tuple(slice(a0, a1) if d == axis else slice(None) for d in range(len(shape)))
            shp = tuple((a1 - a0) if d == axis else shape[d] for d in range(len(shape)))
            self.parts.append([sl, bc(b[a0:a1]), bc(c[a0:a1]), np.zeros(shp, dtype=dtype),
                               bc(1.0 / k[a0:a1] - 1.0)])

    def add_into(self, F, dF, coef, sign):
        for sl, b, c, p, km1 in self.parts:
            p *= b
            p += c * dF[sl]
            add = p + km1 * dF[sl]
            if np.ndim(coef) == 0:
                F[sl] += (sign * coef) * add
            else:
                F[sl] += (sign * coef[sl]) * add


class OracleEngine:
    def __init__(self, kind, dims, mode='ref32'):
        assert mode in ('ref32', 'fp64acc', 'f64')
        self.kind, self.dims, self.mode = kind, tuple(int(d) for d in dims), mode
        self.fdt = np.float64 if mode == 'f64' else np.float32
        self.F = {c: np.zeros(comp_shape(kind, self.dims, c), dtype=self.fdt)
                  for c in E_COMPS[kind] + H_COMPS[kind]}
        self.coef = {}
        self.ch = None
        self.prof = {}
        self.pec = {}
        self.src = None
        self.drive = None
        self.mons = []
        self.sponge = None
        self.mur = []
        self.tfsf = []

    @property
    def wide_acc(self):
        return self.mode != 'ref32'

    def sync(self):
        pass

    # ---------------------------------------------------------------- setup
    def set_coeffs(self, ch, **arrays):
        self.ch = self.fdt(ch)
        for k, v in arrays.items():
            self.coef[k] = None if v is None else np.asarray(v).astype(self.fdt)

    def build_coeffs(self, ch, packed):
        """Compact coefficients (name -> (xclass, table, j0, band)) expanded on the host."""
        arrays = {}
        for name, (xc, table, j0, band) in packed.items():
            full = np.asarray(table, dtype=np.float32)[np.asarray(xc).astype(np.intp)]
            full[:, j0:j0 + band.shape[1]] = band
            arrays[name] = full
        self.set_coeffs(ch, **arrays)

    def set_cpml(self, axis, e_prof, h_prof, pec_lo, pec_hi):
        self.prof[axis] = (tuple(np.asarray(a).astype(self.fdt) for a in e_prof),
                           tuple(np.asarray(a).astype(self.fdt) for a in h_prof))
        self.pec[axis] = (bool(pec_lo), bool(pec_hi))

    def finalize(self):
        """Allocate psi once all profiles are known."""
        k, d, dt = self.kind, self.dims, self.fdt
        if k == 'te2d':
            Nx, Ny = d
            self.psi = dict(Hxy=np.zeros((Nx, Ny - 1), dt), Hyx=np.zeros((Nx - 1, Ny), dt),
                            Ezx=np.zeros((Nx - 2, Ny - 2), dt), Ezy=np.zeros((Nx - 2, Ny - 2), dt))
        elif k == 'tm2d':
            Nx, Ny = d
            self.psi = dict(Hzx=np.zeros((Nx - 1, Ny - 1), dt), Hzy=np.zeros((Nx - 1, Ny - 1), dt),
                            Exy=np.zeros((Nx - 1, Ny - 2), dt), Eyx=np.zeros((Nx - 2, Ny - 1), dt))
        else:
            Nx, Ny, Nz = d
            (bxe, cxe, kxe), (bxh, cxh, kxh) = self._bck(0)
            (bye, cye, kye), (byh, cyh, kyh) = self._bck(1)
            (bze, cze, kze), (bzh, czh, kzh) = self._bck(2)
            S = _Strip
            sHx, sHy, sHz = (Nx, Ny - 1, Nz - 1), (Nx - 1, Ny, Nz - 1), (Nx - 1, Ny - 1, Nz)
            sEx, sEy, sEz = (Nx - 1, Ny - 2, Nz - 2), (Nx - 2, Ny - 1, Nz - 2), (Nx - 2, Ny - 2, Nz - 1)
            i = slice(1, -1)
            self.psi = dict(
                Hx_z=S(sHx, 2, bzh, czh, kzh, dt), Hx_y=S(sHx, 1, byh, cyh, kyh, dt),
                Hy_x=S(sHy, 0, bxh, cxh, kxh, dt), Hy_z=S(sHy, 2, bzh, czh, kzh, dt),
                Hz_y=S(sHz, 1, byh, cyh, kyh, dt), Hz_x=S(sHz, 0, bxh, cxh, kxh, dt),
                Ex_y=S(sEx, 1, bye[i], cye[i], kye[i], dt), Ex_z=S(sEx, 2, bze[i], cze[i], kze[i], dt),
                Ey_z=S(sEy, 2, bze[i], cze[i], kze[i], dt), Ey_x=S(sEy, 0, bxe[i], cxe[i], kxe[i], dt),
                Ez_x=S(sEz, 0, bxe[i], cxe[i], kxe[i], dt), Ez_y=S(sEz, 1, bye[i], cye[i], kye[i], dt))

    def _bck(self, ax):
        e, h = self.prof[ax]
        return (e[1], e[2], e[0]), (h[1], h[2], h[0])   # (b, c, kappa)

    def set_source(self, comp, ix, amp, ph1, ph2):
        amp = np.asarray(amp, dtype=np.float32)
        ph1 = np.asarray(ph1, dtype=np.float32)
        ph2 = np.asarray(ph2, dtype=np.float32)
        a64, p1, p2 = (x.astype(np.float64) for x in (amp, ph1, ph2))
        self.src = dict(comp=comp, ix=int(ix), amp=amp, ph1=ph1, ph2=ph2,
                        a1c=a64 * np.cos(p1), a1s=a64 * np.sin(p1),
                        a2c=a64 * np.cos(p2), a2s=a64 * np.sin(p2))

    def set_drive(self, ncells, wtap, drives, math):
        """Overlap imposition F[:O] = w d(t) + (1 - w) F[:O]; drives: comp -> complex64."""
        self.drive = dict(n=int(ncells), wtap=np.asarray(wtap, dtype=np.float32),
                          D={c: np.asarray(v, dtype=np.complex64) for c, v in drives.items()},
                          math=math)

    def set_sponge(self, n, g, faces):
        """Graded sponge of n cells on the faces of bitmask `faces` (bit 2*axis + hi),
        factors g with g[0] at the outermost cell (fdtd3d_core.py:397-414)."""
        self.sponge = (int(n), np.asarray(g, dtype=np.float32).astype(self.fdt), int(faces))

    def add_mur(self, comp, axis, hi, m, flags):
        """First-order Mur face (boundaries3d.py:99-140; 2D inline fdtd2d_core.py:805-833);
        flags bit 0: F0_old is the current value (2D form), bit 1: zero the plane's cells
        on PEC faces where comp is tangential afterwards."""
        m = np.asarray(m)
        m = m.astype(np.float32) if m.dtype != np.float64 or self.fdt != np.float64 else m
        self.mur.append(dict(comp=comp, axis=int(axis), hi=bool(hi), m=m.astype(self.fdt),
                             flags=int(flags), prev=None))

    def write_field(self, comp, arr):
        self.F[comp][...] = np.asarray(arr, dtype=self.fdt)

    def read_field(self, comp):
        return self.F[comp].copy()

    # ---------------------------------------------------------------- monitors
    def _add(self, m):
        if len(self.mons) >= 32:
            raise RuntimeError('at most 32 monitors per engine')
        self.mons.append(m)
        return len(self.mons) - 1

    def add_box(self, box, form=0):
        """form 1 (2D TM): the node Ey^2 alone (fdtd2d_core.py:1811-1815)."""
        if form and self.kind != 'tm2d':
            raise RuntimeError('box form 1 is 2D TM only')
        comps = {'te2d': ('node',), 'tm2d': ('node',), 'yee3d': ('Ex', 'Ey', 'Ez')}[self.kind]
        rng = {c: resolve_box(comp_shape(self.kind, self.dims, c), box) for c in comps}
        return self._add(dict(t='box', rng=rng, val=0.0, form=int(form)))

    # -- time-domain stitch (fdtd2d_core.py:1586-1863, fdtd3d_core.py:1785-2084)
    def add_sx(self, il):
        """Sum over the masked steps of the instantaneous flux in +x through plane il
        (fdtd2d_core.py:489-497, fdtd3d_core.py:762-767)."""
        if not 1 <= il <= self.dims[0] - 2:
            raise RuntimeError('add_sx: plane index out of range')
        return self._add(dict(t='sx', il=int(il), val=0.0))

    def add_rec(self, comp, il, phase, rows):
        """Plane il of comp into row step['row'] of a (rows,) + plane buffer; phase 0 at the
        start of the step, 1 after the H update and its TF/SF corrections."""
        shp = comp_shape(self.kind, self.dims, comp)
        if not 0 <= il < shp[0] or phase not in (0, 1) or rows < 1:
            raise RuntimeError('add_rec: bad plane, phase or rows')
        return self._add(dict(t='rec', comp=comp, il=int(il), phase=int(phase),
                              a=np.zeros((int(rows),) + shp[1:], dtype=self.fdt)))

    def add_tfsf(self, comp, il, sign, coef, rng, rows):
        """F[il][rng] = F[il][rng] +- c inc[row][rng]; c = Ch (coef None, H components)
        or the coefficient array coef at the cell (E components)."""
        shp = comp_shape(self.kind, self.dims, comp)
        sl = tuple(slice(a, b) for a, b in rng)
        self.tfsf.append(dict(comp=comp, il=int(il), sign=1 if sign > 0 else -1, coef=coef,
                              sl=sl, a=np.zeros((int(rows),) + shp[1:], dtype=self.fdt)))
        return len(self.tfsf) - 1

    def load_tfsf(self, idx, rows):
        rows = np.asarray(rows, dtype=np.float32)
        self.tfsf[idx]['a'][:len(rows)] = rows

    def field_maxabs(self, comp):
        return float(np.max(np.abs(self.F[comp])))

    def _record(self, mask, r, phase):
        for i, m in enumerate(self.mons):
            if m['t'] == 'rec' and m['phase'] == phase and (mask >> i) & 1:
                m['a'][r] = self.F[m['comp']][m['il']]

    def _apply_tfsf(self, r, h):
        for t in self.tfsf:
            if (t['comp'][0] == 'H') != h:
                continue
            F = self.F[t['comp']][t['il']]
            inc = t['a'][r][t['sl']]
            c = self.ch if t['coef'] is None else self.coef[t['coef']][t['il']][t['sl']]
            if t['sign'] > 0:
                F[t['sl']] += c * inc
            else:
                F[t['sl']] -= c * inc

    def add_sq(self, form, nx_acc=None):
        if self.kind == 'yee3d':
            tgt = 'Ez' if form == '3d_s' else 'Ey'
            shp = list(comp_shape(self.kind, self.dims, tgt))
            shp[0] = nx_acc if nx_acc is not None else shp[0]
        else:
            shp = list(self.dims)
        acc = np.zeros(shp, dtype=self.fdt)
        return self._add(dict(t='sq', form=form, acc=acc, nx=shp[0]))

    def add_dft(self, src, x0, x1, kplane=None, phasor='p', acc='c64p'):
        base = self._dft_src_shape(src)
        shp = [x1 - x0] + list(base[1:])
        if kplane is not None:
            shp = shp[:2]
        if self.wide_acc or acc == 'c128':            # fp64acc / f64: complex128 sums
            acc, dt = 'c128', np.complex128
        elif acc == 'split32':
            dt = np.complex64
        else:
            dt = np.complex64
        return self._add(dict(t='dft', src=src, x0=x0, x1=x1, k=kplane, ph=phasor, acc=acc,
                              a=np.zeros(shp, dtype=dt)))

    def _dft_src_shape(self, src):
        if src in ('ey_node', 'ex_node'):
            return self.dims
        return comp_shape(self.kind, self.dims, src)

    def add_flux_plane(self, il):
        from lens_sensor_amd.common.flux import PAIRS
        pairs = []
        wide = self.mode != 'ref32'
        for sign, ec, hc in PAIRS[self.kind]:
            shp = self.F[ec].shape[1:]
            dt = np.complex128 if wide else np.complex64
            pairs.append([ec, hc, np.zeros(shp, dt), np.zeros(shp, dt)])
        return self._add(dict(t='flux', il=int(il), pairs=pairs))

    def zero_monitor(self, mid):
        m = self.mons[mid]
        if m['t'] in ('box', 'sx'):
            m['val'] = 0.0
        elif m['t'] == 'rec':
            m['a'][...] = 0
        elif m['t'] == 'sq':
            m['acc'][...] = 0
        elif m['t'] == 'dft':
            m['a'][...] = 0
        else:
            for p in m['pairs']:
                p[2][...] = 0
                p[3][...] = 0

    def monitor_change(self, mid_a, mid_b, da=1.0, db=1.0):
        """(sum |A/da - B/db|^2, sum |A/da|^2) of two DFT monitors in float64 (the engine
        contract of the device's steady-state reduction)."""
        a = self.mons[mid_a]['a'].astype(np.complex128) / da
        b = self.mons[mid_b]['a'].astype(np.complex128) / db
        d = a - b
        return (float(np.sum(d.real * d.real + d.imag * d.imag)),
                float(np.sum(a.real * a.real + a.imag * a.imag)))

    def read_monitor(self, mid):
        m = self.mons[mid]
        if m['t'] in ('box', 'sx'):
            return float(m['val'])
        if m['t'] == 'sq':
            return m['acc'].copy()
        if m['t'] in ('dft', 'rec'):
            return m['a'].copy()
        return [(p[2].copy(), p[3].copy()) for p in m['pairs']]

    def write_monitor(self, mid, value):
        m = self.mons[mid]
        if m['t'] in ('box', 'sx'):
            m['val'] = value
        elif m['t'] == 'sq':
            m['acc'][...] = value
        elif m['t'] == 'dft':
            m['a'][...] = value
        else:
            for p, (aE, aH) in zip(m['pairs'], value):
                p[2][...] = aE
                p[3][...] = aH

    def read_flux(self, mid):
        return self.read_monitor(mid)

    # ---------------------------------------------------------------- stepping
    def run(self, tab):
        tab = np.asarray(tab, dtype=STEP_DTYPE)
        for row in tab:
            self._step(row)
            mask = int(row['mask'])
            if mask:
                for i, m in enumerate(self.mons):
                    if mask >> i & 1:
                        self._monitor(m, row)

    def _step(self, row):
        td = bool(self.tfsf) or any(m['t'] == 'rec' for m in self.mons)
        mask, r = int(row['mask']), int(row['row'])
        if td:
            self._record(mask, r, 0)
        if self.mur:
            self._mur_capture()
        if td:
            # fdtd2d_core.py:1768-1794, fdtd3d_core.py:1966-1990
            self._h_half()
            self._apply_tfsf(r, True)
            self._record(mask, r, 1)
            self._e_half(row)
            self._apply_tfsf(r, False)
        else:
            self._yee_step(row)
        if self.sponge is not None:
            self._apply_sponge()
        for f in self.mur:
            self._mur_apply(f)

    # -- secondary absorbers
    @staticmethod
    def _plane(F, axis, idx):
        return F[idx] if axis == 0 else (F[:, idx] if axis == 1 else F[:, :, idx])

    def _mur_capture(self):
        # the E planes before the step (boundaries3d.py:125-131; fdtd2d_core.py:795, 823-824)
        for f in self.mur:
            F = self.F[f['comp']]
            i0, i1 = (-1, -2) if f['hi'] else (0, 1)
            f['prev'] = (self._plane(F, f['axis'], i1).copy(),
                         self._plane(F, f['axis'], i0).copy())

    def _mur_apply(self, f):
        # F0 <- F1_prev + m (F1_new - F0_old)   (boundaries3d.py:133-140; fdtd2d_core.py:805-810)
        F = self.F[f['comp']]
        ax = f['axis']
        i0, i1 = (-1, -2) if f['hi'] else (0, 1)
        f1_prev, f0_prev = f['prev']
        f0_old = self._plane(F, ax, i0).copy() if f['flags'] & 1 else f0_prev
        val = f1_prev + f['m'] * (self._plane(F, ax, i1) - f0_old)
        if f['flags'] & 2:
            # PEC faces (applied after the Mur faces in the reference): the plane's cells
            # on PEC faces where the component is tangential
            val = val.copy()
            ndim = F.ndim
            cdir = 2 if (ndim == 2 and f['comp'] == 'Ez') else 'xyz'.index(f['comp'][1])
            others = [b for b in range(ndim) if b != ax]
            for pos, b in enumerate(others):
                if b == cdir or b not in self.pec:
                    continue
                lo, hi = self.pec[b]
                sl_lo = [slice(None)] * val.ndim
                sl_hi = [slice(None)] * val.ndim
                sl_lo[pos], sl_hi[pos] = 0, -1
                if lo:
                    val[tuple(sl_lo)] = 0
                if hi:
                    val[tuple(sl_hi)] = 0
        if ax == 0:
            F[i0] = val
        elif ax == 1:
            F[:, i0] = val
        else:
            F[:, :, i0] = val

    def _apply_sponge(self):
        # all six fields, faces in the order x low/high, y low/high, z low/high
        # (fdtd3d_core.py:406-414, applied at :1133-1135)
        n, g, faces = self.sponge
        for c in E_COMPS[self.kind] + H_COMPS[self.kind]:
            F = self.F[c]
            if faces & 1:
                F[:n] *= g[:, None, None]
            if faces & 2:
                F[-n:] *= g[::-1][:, None, None]
            if faces & 4:
                F[:, :n] *= g[None, :, None]
            if faces & 8:
                F[:, -n:] *= g[::-1][None, :, None]
            if faces & 16:
                F[:, :, :n] *= g[None, None, :]
            if faces & 32:
                F[:, :, -n:] *= g[None, None, ::-1]

    def _yee_step(self, row):
        self._h_half()
        self._e_half(row)

    def _h_half(self):
        if self.kind == 'te2d':
            self._te_H()
        elif self.kind == 'tm2d':
            self._tm_H()
        else:
            self._yee_H()

    def _e_half(self, row):
        if self.kind == 'te2d':
            self._te_E()
            self._excite(row)
            self._pec2d()
        elif self.kind == 'tm2d':
            dax = self.coef.get('dax')
            if dax is not None:                        # fdtd2d_core.py:796-797
                self.F['Ex'] *= dax
                self.F['Ey'] *= self.coef['day']
            self._tm_E()
            self._excite(row)
            self._pec2d()
        else:
            self._yee_E()                              # includes loss and PEC
            self._excite(row)

    # -- 2D TE: boundaries.py:189-205
    def _te_H(self):
        Ez, Hx, Hy, Ch = self.F['Ez'], self.F['Hx'], self.F['Hy'], self.ch
        khx, bhx, chx = self.prof[0][1]
        khy, bhy, chy = self.prof[1][1]
        dy = Ez[:, 1:] - Ez[:, :-1]
        self.psi['Hxy'] = bhy[None, :] * self.psi['Hxy'] + chy[None, :] * dy
        Hx -= Ch * (dy / khy[None, :] + self.psi['Hxy'])
        dx_ = Ez[1:, :] - Ez[:-1, :]
        self.psi['Hyx'] = bhx[:, None] * self.psi['Hyx'] + chx[:, None] * dx_
        Hy += Ch * (dx_ / khx[:, None] + self.psi['Hyx'])

    def _te_E(self):
        Ez, Hx, Hy = self.F['Ez'], self.F['Hx'], self.F['Hy']
        kex, bex, cex = self.prof[0][0]
        key, bey, cey = self.prof[1][0]
        ceE, ceH = self.coef['ceE'], self.coef['ceH']
        dHy = Hy[1:, 1:-1] - Hy[:-1, 1:-1]
        dHx = Hx[1:-1, 1:] - Hx[1:-1, :-1]
        self.psi['Ezx'] = bex[1:-1, None] * self.psi['Ezx'] + cex[1:-1, None] * dHy
        self.psi['Ezy'] = bey[None, 1:-1] * self.psi['Ezy'] + cey[None, 1:-1] * dHx
        Ez[1:-1, 1:-1] = (ceE[1:-1, 1:-1] * Ez[1:-1, 1:-1]
                          + ceH[1:-1, 1:-1] * (dHy / kex[1:-1, None] + self.psi['Ezx']
                                               - dHx / key[None, 1:-1] - self.psi['Ezy']))

    # -- 2D TM: boundaries.py:243-258
    def _tm_H(self):
        Ex, Ey, Hz, Ch = self.F['Ex'], self.F['Ey'], self.F['Hz'], self.ch
        khx, bhx, chx = self.prof[0][1]
        khy, bhy, chy = self.prof[1][1]
        dEyx = Ey[1:, :] - Ey[:-1, :]
        dExy = Ex[:, 1:] - Ex[:, :-1]
        self.psi['Hzx'] = bhx[:, None] * self.psi['Hzx'] + chx[:, None] * dEyx
        self.psi['Hzy'] = bhy[None, :] * self.psi['Hzy'] + chy[None, :] * dExy
        Hz -= Ch * ((dEyx / khx[:, None] + self.psi['Hzx'])
                    - (dExy / khy[None, :] + self.psi['Hzy']))

    def _tm_E(self):
        Ex, Ey, Hz = self.F['Ex'], self.F['Ey'], self.F['Hz']
        kex, bex, cex = self.prof[0][0]
        key, bey, cey = self.prof[1][0]
        cax, cay = self.coef['cax'], self.coef['cay']
        dHzy = Hz[:, 1:] - Hz[:, :-1]
        self.psi['Exy'] = bey[None, 1:-1] * self.psi['Exy'] + cey[None, 1:-1] * dHzy
        Ex[:, 1:-1] += cax[:, 1:-1] * (dHzy / key[None, 1:-1] + self.psi['Exy'])
        dHzx = Hz[1:, :] - Hz[:-1, :]
        self.psi['Eyx'] = bex[1:-1, None] * self.psi['Eyx'] + cex[1:-1, None] * dHzx
        Ey[1:-1, :] += -cay[1:-1, :] * (dHzx / kex[1:-1, None] + self.psi['Eyx'])

    def _pec2d(self):
        (xl, xh), (yl, yh) = self.pec[0], self.pec[1]
        if self.kind == 'te2d':
            Ez = self.F['Ez']
            if xl:
                Ez[0, :] = 0.0
            if xh:
                Ez[-1, :] = 0.0
            if yl:
                Ez[:, 0] = 0.0
            if yh:
                Ez[:, -1] = 0.0
        else:
            Ex, Ey = self.F['Ex'], self.F['Ey']
            if xl:
                Ey[0, :] = 0.0
            if xh:
                Ey[-1, :] = 0.0
            if yl:
                Ex[:, 0] = 0.0
            if yh:
                Ex[:, -1] = 0.0

    # -- 3D: boundaries3d.py:311-362
    def _yee_H(self):
        Ex, Ey, Ez = self.F['Ex'], self.F['Ey'], self.F['Ez']
        Hx, Hy, Hz, Ch = self.F['Hx'], self.F['Hy'], self.F['Hz'], self.ch
        P = self.psi
        dEy_dz = Ey[:, :, 1:] - Ey[:, :, :-1]
        dEz_dy = Ez[:, 1:, :] - Ez[:, :-1, :]
        Hx += Ch * (dEy_dz - dEz_dy)
        P['Hx_z'].add_into(Hx, dEy_dz, Ch, +1)
        P['Hx_y'].add_into(Hx, dEz_dy, Ch, -1)
        dEz_dx = Ez[1:, :, :] - Ez[:-1, :, :]
        dEx_dz = Ex[:, :, 1:] - Ex[:, :, :-1]
        Hy += Ch * (dEz_dx - dEx_dz)
        P['Hy_x'].add_into(Hy, dEz_dx, Ch, +1)
        P['Hy_z'].add_into(Hy, dEx_dz, Ch, -1)
        dEx_dy = Ex[:, 1:, :] - Ex[:, :-1, :]
        dEy_dx = Ey[1:, :, :] - Ey[:-1, :, :]
        Hz += Ch * (dEx_dy - dEy_dx)
        P['Hz_y'].add_into(Hz, dEx_dy, Ch, +1)
        P['Hz_x'].add_into(Hz, dEy_dx, Ch, -1)

    def _yee_E(self):
        Ex, Ey, Ez = self.F['Ex'], self.F['Ey'], self.F['Ez']
        Hx, Hy, Hz = self.F['Hx'], self.F['Hy'], self.F['Hz']
        P, c = self.psi, self.coef
        if c.get('dax') is not None:
            Ex *= c['dax']
            Ey *= c['day']
            Ez *= c['daz']
        dHz_dy = Hz[:, 1:, 1:-1] - Hz[:, :-1, 1:-1]
        dHy_dz = Hy[:, 1:-1, 1:] - Hy[:, 1:-1, :-1]
        Exv = Ex[:, 1:-1, 1:-1]
        cev = c['cex'][:, 1:-1, 1:-1]
        Exv += cev * (dHz_dy - dHy_dz)
        P['Ex_y'].add_into(Exv, dHz_dy, cev, +1)
        P['Ex_z'].add_into(Exv, dHy_dz, cev, -1)
        dHx_dz = Hx[1:-1, :, 1:] - Hx[1:-1, :, :-1]
        dHz_dx = Hz[1:, :, 1:-1] - Hz[:-1, :, 1:-1]
        Eyv = Ey[1:-1, :, 1:-1]
        cev = c['cey'][1:-1, :, 1:-1]
        Eyv += cev * (dHx_dz - dHz_dx)
        P['Ey_z'].add_into(Eyv, dHx_dz, cev, +1)
        P['Ey_x'].add_into(Eyv, dHz_dx, cev, -1)
        dHy_dx = Hy[1:, 1:-1, :] - Hy[:-1, 1:-1, :]
        dHx_dy = Hx[1:-1, 1:, :] - Hx[1:-1, :-1, :]
        Ezv = Ez[1:-1, 1:-1, :]
        cev = c['cez'][1:-1, 1:-1, :]
        Ezv += cev * (dHy_dx - dHx_dy)
        P['Ez_x'].add_into(Ezv, dHy_dx, cev, +1)
        P['Ez_y'].add_into(Ezv, dHx_dy, cev, -1)
        (xl, xh), (yl, yh), (zl, zh) = self.pec[0], self.pec[1], self.pec[2]
        if xl:
            Ey[0] = 0
            Ez[0] = 0
        if xh:
            Ey[-1] = 0
            Ez[-1] = 0
        if yl:
            Ex[:, 0] = 0
            Ez[:, 0] = 0
        if yh:
            Ex[:, -1] = 0
            Ez[:, -1] = 0
        if zl:
            Ex[:, :, 0] = 0
            Ey[:, :, 0] = 0
        if zh:
            Ex[:, :, -1] = 0
            Ey[:, :, -1] = 0

    # -- source (vcsel.py:151-156) or stitch drive
    def _excite(self, row):
        if self.drive is not None:
            self._impose(row)
        if self.src is None:
            return
        s = self.src
        F = self.F[s['comp']]
        i = s['ix']
        env = float(row['env'])
        if self.mode == 'ref32':
            wt = np.float32(row['wt32'])
            F[i] += env * np.sin(wt + s['ph1']) * s['amp']
            F[i + 1] -= env * np.sin(wt + s['ph2']) * s['amp']
        else:
            ss, sc = float(row['ssin']), float(row['scos'])
            F[i] += (ss * s['a1c'] + sc * s['a1s']).astype(self.fdt)
            F[i + 1] -= (ss * s['a2c'] + sc * s['a2s']).astype(self.fdt)

    def _impose(self, row):
        d = self.drive
        O = d['n']
        env = float(row['env'])
        shp = (O,) + (1,) * (len(self.dims) - 1)
        w = d['wtap'].reshape(shp)
        for comp, D in d['D'].items():
            F = self.F[comp]
            if d['math'] == 'c128':                    # fdtd2d_core.py:1426-1427
                din = env * np.real(D * np.complex128(complex(row['dr'], row['di'])))
                F[:O] = w * din + (1.0 - w) * F[:O]
            else:                                      # fdtd3d_core.py:1627-1630
                ph = np.complex64(complex(row['dr32'], row['di32']))
                F[:O] = w * (env * np.real(D * ph)) + (1.0 - w) * F[:O]

    # ---------------------------------------------------------------- monitors
    def _node_ey(self):
        Ey = self.F['Ey']                              # fdtd2d_core.py:458-466
        out = np.zeros((Ey.shape[0], Ey.shape[1] + 1), dtype=self.fdt)
        out[:, 1:-1] = 0.5 * (Ey[:, :-1] + Ey[:, 1:])
        out[:, 0] = Ey[:, 0]
        out[:, -1] = Ey[:, -1]
        return out

    def _node_ex(self):
        Ex = self.F['Ex']                              # fdtd2d_core.py:469-475
        out = np.zeros((Ex.shape[0] + 1, Ex.shape[1]), dtype=self.fdt)
        out[1:-1, :] = 0.5 * (Ex[:-1, :] + Ex[1:, :])
        out[0, :] = Ex[0, :]
        out[-1, :] = Ex[-1, :]
        return out

    def _node_ex2(self):
        Ex = self.F['Ex']                              # fdtd2d_core.py:478-486
        s = Ex * Ex
        out = np.zeros((Ex.shape[0] + 1, Ex.shape[1]), dtype=self.fdt)
        out[1:-1, :] = 0.5 * (s[:-1, :] + s[1:, :])
        out[0, :] = s[0, :]
        out[-1, :] = s[-1, :]
        return out

    def _intensity2d(self):
        if self.kind == 'te2d':
            Ez = self.F['Ez']
            return Ez * Ez                             # fdtd2d_core.py:835
        f = self._node_ey()
        return f * f + self._node_ex2()                # fdtd2d_core.py:812

    def _sq_on_grid(self, F, tshape, x0, x1):
        """fdtd3d_core.py:438-462: F^2 on the target grid by averaging squares."""
        nF = F.shape[0]
        if nF == tshape[0]:
            S = F[x0:x1] ** 2
        else:
            i = np.arange(x0, x1)
            S = 0.5 * (F[np.clip(i - 1, 0, nF - 1)] ** 2 + F[np.clip(i, 0, nF - 1)] ** 2)
        for ax in (1, 2):
            n, t = S.shape[ax], tshape[ax]
            if n == t:
                continue
            a = [slice(None)] * 3
            b = [slice(None)] * 3
            a[ax] = slice(None, -1)
            b[ax] = slice(1, None)
            mid = 0.5 * (S[tuple(a)] + S[tuple(b)])
            if n == t + 1:
                S = mid
            else:
                first = [slice(None)] * 3
                last = [slice(None)] * 3
                first[ax] = slice(0, 1)
                last[ax] = slice(-1, None)
                S = np.concatenate([S[tuple(first)], mid, S[tuple(last)]], axis=ax)
        return S

    def _monitor(self, m, row):
        t = m['t']
        if t == 'box':
            if self.kind == 'yee3d':                   # fdtd3d_core.py:477-481
                tot = 0
                for c in ('Ex', 'Ey', 'Ez'):
                    (a0, a1), (b0, b1), (c0, c1) = m['rng'][c]
                    tot = tot + np.sum(self.F[c][a0:a1, b0:b1, c0:c1] ** 2, dtype=np.float64)
                m['val'] += tot
            else:
                (a0, a1), (b0, b1) = m['rng']['node']
                if m.get('form'):                      # fdtd2d_core.py:1814
                    f = self._node_ey()[a0:a1, b0:b1]
                    m['val'] += np.sum(f ** 2, dtype=np.float64)
                else:
                    I = self._intensity2d()
                    m['val'] += np.sum(I[a0:a1, b0:b1], dtype=np.float64)
        elif t == 'sx':                                # fdtd2d_core.py:489-497, fdtd3d_core.py:762-767
            il, F = m['il'], self.F
            if self.kind == 'te2d':
                hy = 0.5 * (F['Hy'][il - 1] + F['Hy'][il])
                m['val'] += -np.sum(F['Ez'][il] * hy, dtype=np.float64)
            elif self.kind == 'tm2d':
                hz = 0.5 * (F['Hz'][il - 1] + F['Hz'][il])
                m['val'] += np.sum(F['Ey'][il] * hz, dtype=np.float64)
            else:
                hz = 0.5 * (F['Hz'][il - 1] + F['Hz'][il])
                hy = 0.5 * (F['Hy'][il - 1] + F['Hy'][il])
                m['val'] += (np.sum(F['Ey'][il] * hz, dtype=np.float64)
                             - np.sum(F['Ez'][il] * hy, dtype=np.float64))
        elif t == 'rec':
            pass                                       # recorded inside the step
        elif t == 'sq':
            if self.kind == 'yee3d':                   # fdtd3d_core.py:465-474
                nx = m['nx']
                Ex, Ey, Ez = self.F['Ex'][:nx - 1], self.F['Ey'][:nx], self.F['Ez'][:nx]
                I = m['acc']
                ch = max(32, nx // 8)
                for x0 in range(0, nx, ch):
                    x1 = min(x0 + ch, nx)
                    I[x0:x1] += (self._sq_on_grid(Ex, I.shape, x0, x1)
                                 + self._sq_on_grid(Ey, I.shape, x0, x1)
                                 + self._sq_on_grid(Ez, I.shape, x0, x1))
            elif m['form'] == 'tm_ex2':
                m['acc'] += self._node_ex2()           # fdtd2d_core.py:1435
            else:
                m['acc'] += self._intensity2d()        # fdtd2d_core.py:848-850
        elif t == 'dft':
            src = m['src']
            if src == 'ey_node':
                F = self._node_ey()
            elif src == 'ex_node':
                F = self._node_ex()
            else:
                F = self.F[src]
            F = F[m['x0']:m['x1']]
            if m['k'] is not None:
                F = F[:, :, m['k']]
            a = m['a']
            if m['acc'] == 'c128':
                a += F.astype(np.complex128) * complex(row['pr'], row['pi'])
            elif m['acc'] == 'c64p':                   # complex128 promotion (NumPy >= 2)
                a += F.astype(np.complex64) * np.complex128(complex(row['pr'], row['pi']))
            elif m['acc'] == 'split32':                # fdtd3d_core.py:1638-1639
                re, im = a.real.copy(), a.imag.copy()
                re += F * np.float32(row['qr32'])
                im -= F * np.float32(-row['qi32'])
                a.real, a.imag = re, im
            else:                                      # complex64 phasor
                key = m['ph']
                ph = np.complex64(complex(row[key[0] + 'r32'], row[key[0] + 'i32']))
                a += F.astype(np.complex64) * ph
        else:                                          # flux.py:40-42
            il = m['il']
            for p in m['pairs']:
                E, H = self.F[p[0]], self.F[p[1]]
                h = 0.5 * (H[il - 1] + H[il])
                if p[2].dtype == np.complex64:
                    pe = np.complex64(complex(row['pr32'], row['pi32']))
                    ph = np.complex64(complex(row['hr32'], row['hi32']))
                    p[2] += E[il].astype(np.complex64) * pe
                    p[3] += h.astype(np.complex64) * ph
                else:
                    p[2] += E[il].astype(np.complex128) * complex(row['pr'], row['pi'])
                    p[3] += h.astype(np.complex128) * complex(row['hr'], row['hi'])

    def close(self):
        pass
