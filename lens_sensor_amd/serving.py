"""Continuous-batching solve server for CW requests (2D windows and 3D windows).

The reference solves one request at a time: run_beads / run_beads_stitched /
run_3d each own the process until their CW criterion is met (fdtd2d_core.py:859-873,
fdtd3d_core.py:1150-1175: relative phasor change between consecutive DFT windows).  Here
a request loop keeps every admitted solve resident on the GPU (fields, CPML psi,
coefficients, monitors) and advances all of them round by round:

* a request is admitted when its HBM footprint fits the budget (FIFO, default 220 GB);
* one round = ``round_steps`` monitor-free steps followed by two one-period CW phasor
  windows (A, then B) with the Poynting flux through the input / output planes; after the
  round the request is done when |A - B| / |A| < ``tol`` (phasor change over one period),
  the two norms reduced on the device (the fields and phasors stay resident, only two
  doubles and the flux plane sums cross to the host per round).
  Rounds that end before one transit of the domain carry no check (no monitors at all);
* all live 2D requests advance together through the batch engine (one launch per
  wavefront class per temporally blocked pass, each request at its own time offset in a
  resident step table); 3D requests are enqueued on their own streams first, so their
  work overlaps the 2D batch on the device;
* requests that arrive during a round join at the next round (continuous batching).

A request's step schedule depends only on the request, and the batch engine is
bit-identical to a lone instance, so every result is the same alone and in the mix.
"""
import time

import numpy as np

from lens_sensor_amd.common import backend
from lens_sensor_amd.common.flux import PlaneFluxBook, guard_in_plane
from lens_sensor_amd.common.physics import C0, MU0, n_at
from lens_sensor_amd.engine.spec import new_steps, fill_phasors, cw_envelope
from lens_sensor_amd.planar_beads import fdtd2d_core as d2
from lens_sensor_amd.planar_beads.boundaries import make_boundary
from lens_sensor_amd.planar_3d import fdtd3d_core as f3d
from lens_sensor_amd.planar_3d.boundaries3d import make_boundary_3d, ALL_FACES

LAMBDA_NM = 850.0


def request_stream(n, rate_hz=20.0, seed=7, frac_3d=0.2, nx2=(4096, 16384), ny2=(1024, 2048),
                   n3=((256, 512), (192, 256), (192, 256))):
    """Seeded Poisson stream of `n` requests: arrival times (s) and window sizes; 80 % 2D
    windows (TE / TM alternating), 20 % 3D windows."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
    out = []
    for i in range(n):
        if rng.uniform() < frac_3d:
            dims = tuple(int(rng.integers(lo, hi + 1)) for lo, hi in n3)
            out.append(dict(id=i, arrival=float(t[i]), kind='3d', dims=dims,
                            bead_x_frac=float(rng.uniform(0.35, 0.65)), bead_r_um=float(rng.uniform(0.2, 0.5))))
        else:
            dims = (int(rng.integers(nx2[0], nx2[1] + 1)), int(rng.integers(ny2[0], ny2[1] + 1)))
            out.append(dict(id=i, arrival=float(t[i]), kind='2d', dims=dims, pol='p' if i % 2 else 's',
                            bead_x_frac=float(rng.uniform(0.3, 0.7)), bead_r_um=float(rng.uniform(0.2, 0.6)),
                            kappa_aq=float(rng.uniform(0.0, 0.003))))
    return out


class _CWSolve:
    """Common part of a served solve: local step counter, round schedule, CW check."""

    round_steps = 1536
    tol = 1e-4

    def _init_monitors(self, acc, dft_src, kplane, planes):
        eng = self.eng
        kw = {} if kplane is None else dict(kplane=kplane)
        self.mon_a = eng.add_dft(dft_src, 0, self.dims[0], acc=acc, **kw)
        self.mon_b = eng.add_dft(dft_src, 0, self.dims[0], acc=acc, **kw)
        self.book = PlaneFluxBook(eng, self.kind)
        for key, il in planes:
            self.book.add_plane(key, il)
        self.flux_bits = sum(1 << self.book.mon[k] for k in self.book.keys())
        self.n = 0
        self.done = False
        self.converged = False
        self.change = float('nan')

    def round_len(self):
        return self.round_steps + 2 * self.spT

    def checking(self):
        """The CW criterion is applied once the field can have crossed the domain (the
        estimated transit time); earlier rounds carry no monitor and block throughout."""
        return self.n + self.round_len() >= self.first_check

    def masks(self):
        m = np.zeros(self.round_len(), np.uint32)
        if self.checking():
            m[self.round_steps:self.round_steps + self.spT] = (1 << self.mon_a) | self.flux_bits
            m[self.round_steps + self.spT:] = (1 << self.mon_b) | self.flux_bits
        return m

    def finish_round(self):
        """Advance the local clock by one round and apply the CW criterion."""
        if not self.checking():
            self.n += self.round_len()
            return False
        self.n += self.round_len()
        # |A - B| / |A| reduced on the device: two doubles cross to the host, not two
        # phasor fields
        num, den = self.eng.monitor_change(self.mon_a, self.mon_b)
        self.change = float(np.sqrt(num) / max(float(np.sqrt(den)), 1e-300))
        if self.change < self.tol or self.n + self.round_len() > self.max_steps:
            self.converged = self.change < self.tol
            self.book.note_samples(self.book.keys(), 2 * self.spT)
            s_in, s_out = self.book.flux('in'), self.book.flux('out')
            self.result = dict(converged=self.converged, steps=self.n, change=self.change,
                               transmission=float(s_out / s_in) if s_in else float('nan'),
                               phasor_norm=float(np.sqrt(den)))
            self.done = True
            return True
        for m in (self.mon_a, self.mon_b):
            self.eng.zero_monitor(m)
        self.book.reset()
        return False

    def table_slice(self, table):
        t = table[self.n:self.n + self.round_len()].copy()
        t['mask'] = self.masks()
        return t


class Solve2D(_CWSolve):
    """A 2D window request (TE or TM): the geometry and host set-up of a full-domain
    run_beads solve (this package's fdtd2d_core), resident on the device until done."""

    dx_nm = 25.0

    def __init__(self, req, mode='fp64acc', npml=24, engine_factory=None):
        self.req = req
        Nx, Ny = self.dims = req['dims']
        self.pol = req['pol']
        self.kind = 'tm2d' if self.pol == 'p' else 'te2d'
        dx = self.dx_nm * 1e-9
        dt = 0.5 * dx / (C0 * np.sqrt(2))                 # fdtd2d_core.py:626
        T_WG, LAM, wg_n = 2.0e-6, LAMBDA_NM * 1e-9, 1.49
        tear = Ny * dx - T_WG - d2.AIR_BUFFER
        r = req['bead_r_um']
        stack = [dict(name='Aqueous', t_um=3.0, n=complex(1.333, req['kappa_aq']), kappa=req['kappa_aq'],
                      mat=None), dict(name='Mucin', t_um=0.5, n=complex(1.34, 0.0), kappa=0.0, mat=None)]
        L_um = Nx * dx * 1e6
        beads = [dict(x_um=req['bead_x_frac'] * L_um, d_um=2 * r, y_um=-r, n=1.59)]
        old_tb, old_lam = d2.TEAR_BUFFER, d2.LAM_LOSS
        d2.TEAR_BUFFER, d2.LAM_LOSS = tear, LAM
        try:
            src_ix = npml + 2
            x_facet_um, _air, zsrc = d2._vcsel_geom_2d(src_ix, dx, 0.0)
            mp = d2._material_parts(
                Nx, Ny, dx, wg_n, 1.59, T_WG, beads[0]['x_um'], 2 * r, 0.0, False, n_aq=1.333,
                n_mucin=1.34, n_cornea=1.376, t_aq_um=3.0, t_mu_um=0.5, t_lip_um=0.0,
                n_lip=n_at('lipid', LAMBDA_NM), input_gap_um=x_facet_um, length_um=L_um,
                curved=False, bead_y_um=None, stack=stack, beads=beads)
            self.eng = eng = (engine_factory or backend.make_engine)(self.kind, (Nx, Ny), mode)
            # coefficients in compact form, assembled by the engine (fdtd2d_core.window_coeffs)
            packed, ch_tm = d2.window_coeffs(self.pol, mp, dt, dx)
            eng.build_coeffs(ch_tm if self.pol == 'p' else np.float32(dt / (MU0 * dx)), packed)
            del packed
            f0 = C0 / LAM
            y0_m = -tear
            n_src = d2._n_source_2d(src_ix * dx * 1e6, int(round((T_WG / 2 - y0_m) / dx)), y0_m, dx,
                                    wg_n, T_WG, 1.333, x_facet_um)
            _iy, vs, vs_ey = d2._vcsel_source_2d(Ny, dx, dt, T_WG, src_ix, 0.0, 3.0, LAMBDA_NM, 0.0,
                                                 n_src, zsrc, f0, y0_m=y0_m)
            make_boundary('cpml', self.pol, Nx, Ny, dx, dt, npml=npml, lam_m=LAM).attach(eng)
            eng.finalize()
            src = vs_ey if self.pol == 'p' else vs
            eng.set_source('Ey' if self.pol == 'p' else 'Ez', src.ix, src.amp, src.ph1, src.ph2)
            L = Nx * dx
            cin = guard_in_plane(int(round(0.1 * L / dx)), max(src_ix, int(round(x_facet_um / (dx * 1e6)))), Nx)
            self.spT = max(4, int(round((1.0 / f0) / dt)))
            steps_ref = int((6 * 4 / (2 * np.pi * f0) + 2 * L * d2.N_G_FACTOR * wg_n / C0) / dt)
            self.max_steps = 3 * steps_ref
            self.first_check = int(L * d2.N_G_FACTOR * wg_n / C0 / dt)        # one transit
            self.omega, self.dt, self.src_omega = 2 * np.pi * f0, dt, src.omega
            self._init_monitors('c128' if eng.wide_acc else 'c64p', 'ey_node' if self.pol == 'p' else 'Ez',
                                None, (('in', max(1, cin)), ('out', int(round(0.9 * L / dx)))))
        finally:
            d2.TEAR_BUFFER, d2.LAM_LOSS = old_tb, old_lam
        self.cells = Nx * Ny
        self.footprint = self.cells * (72 + (8 if self.pol == 'p' else 0))


class Solve3D(_CWSolve):
    """A 3D window request: the host set-up of run_3d (this package's fdtd3d_core:
    build_eps3d, inv_loss, ce_arrays, Gaussian VCSEL source, CPML on all faces); CW
    phasor on the z-centre plane, flux through the 20 % / 80 % planes."""

    dx_nm = 30.0

    def __init__(self, req, mode='fp64acc', npml=12, engine_factory=None):
        self.req = req
        self.kind = 'yee3d'
        Nx, Ny, Nz = self.dims = req['dims']
        dx_um = self.dx_nm / 1000.0
        dx = dx_um * 1e-6
        dt = 0.5 * dx / (C0 * np.sqrt(3))                 # fdtd3d_core.py:838
        LAM = LAMBDA_NM * 1e-9
        wg_n, t_wg, air = 1.49, 1.5, 1.0
        ly, lz, lx = Ny * dx_um, Nz * dx_um, Nx * dx_um
        tear = ly - t_wg - air
        y0_um, z0_um = -tear, -lz / 2.0
        stack = [dict(name='Lipid', t_um=0.1, n=1.46, kappa=0.0, mat=None),
                 dict(name='Aqueous', t_um=1.5, n=1.333, kappa=0.0, mat=None),
                 dict(name='Mucin', t_um=0.3, n=1.34, kappa=0.0, mat=None)]
        r = req['bead_r_um']
        bead = dict(x_um=req['bead_x_frac'] * lx, d_um=2 * r, y_um=-0.1 - r, n=1.59, z_um=0.0)
        old = f3d.LAM_LOSS
        f3d.LAM_LOSS = LAM
        try:
            n_sp = npml + 2                                # source at n_sp + 6, clear of the CPML
            x_src_um, x_wg_start_um, _air, zsrc = f3d._vcsel_geom_3d(n_sp, dx_um, 0.0)
            tp = {}
            eps = f3d.build_eps3d(Nx, Ny, Nz, dx_um, y0_um, z0_um, wg_n, t_wg, 0.0, 1.5, 0.3, 1.46,
                                  1.333, 1.34, 1.376, bead, x_wg_start_um=x_wg_start_um,
                                  tensor_patch=tp, stack=stack)
            eps, eim = f3d.split_eps(eps)
            inv, da = f3d.inv_loss(eps, eim, dt, dx, LAM)
            del eps, eim
            cex, cey, cez = f3d.ce_arrays(inv, tp, dt, dx, dx_um, da=da)
            da = f3d.stagger_da(da)
            del inv
            self.eng = eng = (engine_factory or backend.make_engine)('yee3d', (Nx, Ny, Nz), mode)
            eng.set_coeffs(np.float32(dt / (MU0 * dx)), cex=cex, cey=cey, cez=cez,
                           dax=None if da is None else da[0], day=None if da is None else da[1],
                           daz=None if da is None else da[2])
            del cex, cey, cez
            make_boundary_3d('cpml', dx, dt, ALL_FACES, shape=(Nx, Ny, Nz), npml=npml, lam_m=LAM).attach(eng)
            eng.finalize()
            src_ix, _j, _k, vs = f3d._build_gauss_source(
                n_sp, t_wg, y0_um, z0_um, dx_um, dx, LAM, Ny, Nz - 1, 1.0, 1.0, 0.0, 0.0, 0.0,
                n_src_fn=f3d._n_source_fn_3d(x_src_um, y0_um, dx_um, wg_n, t_wg, 1.333,
                                             x_wg_start_um, 's'),
                pol='s', wg_width_um=None, dt=dt, z_prop_um=zsrc)
            eng.set_source('Ez', src_ix, vs.amp, vs.ph1, vs.ph2)
            f0 = C0 / LAM
            self.spT = max(4, int(round((1.0 / f0) / dt)))
            steps_ref = int((24.0 / (2 * np.pi * f0) + 2.0 * lx * 1e-6 * f3d.N_G_FACTOR * wg_n / C0) / dt)
            self.max_steps = 3 * steps_ref
            self.first_check = int(lx * 1e-6 * f3d.N_G_FACTOR * wg_n / C0 / dt)   # one transit
            self.omega, self.dt, self.src_omega = 2 * np.pi * f0, dt, vs.omega
            cin = guard_in_plane(int(round(0.2 * Nx)), src_ix, Nx)
            self._init_monitors('c128' if eng.wide_acc else 'c64p', 'Ez', (Nz - 1) // 2,
                                (('in', max(1, cin)), ('out', int(round(0.8 * Nx)))))
        finally:
            f3d.LAM_LOSS = old
        self.cells = Nx * Ny * Nz
        self.footprint = self.cells * 76


def step_table(omega, dt, src_omega, n):
    tab = new_steps(n)
    t = np.arange(n) * dt
    fill_phasors(tab, omega, dt, t, cw_envelope(t, 4 / omega), src_omega=src_omega)
    return tab


class Server:
    """Continuous-batching request loop (one process, one GPU)."""

    def __init__(self, mode='fp64acc', hbm_budget=220e9, max_active=64, depth=6, lx=64,
                 engine_factory=None, batched=True, round_steps=1536, log=None):
        self.mode, self.budget, self.max_active = mode, float(hbm_budget), max_active
        self.depth, self.lx = depth, lx
        self.factory, self.batched, self.log = engine_factory, batched, log
        self.round_steps = round_steps
        self._tables = {}

    def _table(self, s):
        key = (s.kind == 'yee3d', s.omega, s.dt)
        if key not in self._tables or len(self._tables[key]) < s.max_steps:
            self._tables[key] = step_table(s.omega, s.dt, s.src_omega, s.max_steps)
        return self._tables[key]

    def _make(self, req):
        cls = Solve3D if req['kind'] == '3d' else Solve2D
        s = cls(req, mode=self.mode, engine_factory=self.factory)
        s.round_steps = self.round_steps
        return s

    def run(self, requests, clock=time.time, realtime=True):
        """Serve `requests` (arrival times in seconds from the start).  realtime: admit by
        the wall clock; otherwise every request is available at once.  Returns the
        per-request records and summary numbers."""
        pending = sorted(requests, key=lambda r: r['arrival'])
        t0 = clock()
        active, records, used = [], [], 0.0
        while pending or active:
            now = clock() - t0
            while (pending and len(active) < self.max_active
                   and (not realtime or pending[0]['arrival'] <= now)):
                req = pending[0]
                est = Solve3D.footprint_of(req) if req['kind'] == '3d' else Solve2D.footprint_of(req)
                if active and used + est > self.budget:
                    break
                pending.pop(0)
                s = self._make(req)
                s.t_admit = clock() - t0
                used += s.footprint
                active.append(s)
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, min(0.05, pending[0]['arrival'] - (clock() - t0))))
                continue
            self._round(active)
            now = clock() - t0
            for s in [a for a in active if a.done]:
                active.remove(s)
                used -= s.footprint
                rec = dict(id=s.req['id'], kind=s.req['kind'], dims=s.req['dims'], arrival=s.req['arrival'],
                           admitted=s.t_admit, finished=now, latency=now - s.req['arrival'], **s.result)
                records.append(rec)
                s.eng.close()
                if self.log:
                    self.log(f"[serve] request {rec['id']} ({rec['kind']} {rec['dims']}) done: "
                             f"{rec['steps']} steps, change {rec['change']:.2e}, latency {rec['latency']:.2f}s")
        lat = np.array([r['latency'] for r in records])
        span = max(r['finished'] for r in records) - min(r['arrival'] for r in records) if records else 0.0
        return records, dict(requests=len(records), p50_s=float(np.percentile(lat, 50)) if len(lat) else None,
                             p99_s=float(np.percentile(lat, 99)) if len(lat) else None,
                             solves_per_s=len(records) / span if span > 0 else None,
                             converged=sum(r['converged'] for r in records))

    def _round(self, active):
        """Advance every active request by one round (2D together, 3D overlapped)."""
        threed = [s for s in active if s.kind == 'yee3d']
        twod = [s for s in active if s.kind != 'yee3d']
        for s in threed:                                  # enqueued first: overlaps the 2D batch
            s.eng.run(s.table_slice(self._table(s)))
        if twod:
            if self.batched and len(twod) > 0 and hasattr(twod[0].eng, 'h'):
                from lens_sensor_amd.engine.hip import HipBatch
                b = HipBatch([s.eng for s in twod], depth=self.depth, lx=self.lx)
                b.set_table(self._table(twod[0]))
                n = twod[0].round_len()
                b.run_at([s.n for s in twod], n, masks=np.stack([s.masks() for s in twod]))
                b.close()
            else:
                for s in twod:
                    s.eng.run(s.table_slice(self._table(s)))
        for s in threed:
            s.eng.sync()
        for s in active:
            s.finish_round()


def _fp2(req):
    Nx, Ny = req['dims']
    return Nx * Ny * (72 + (8 if req.get('pol') == 'p' else 0))


def _fp3(req):
    Nx, Ny, Nz = req['dims']
    return Nx * Ny * Nz * 76


Solve2D.footprint_of = staticmethod(_fp2)
Solve3D.footprint_of = staticmethod(_fp3)
