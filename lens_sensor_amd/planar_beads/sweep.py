"""Batched 2D parameter sweeps: many independent CW solves advanced together.

The reference runs a sweep as one run_beads call after another (run_simulation.py:290-431
loops over the parameter sets; fdtd2d_core.py:531-927 is one solve).  Here every
instance is set up with the same host code as a full-domain run_beads solve (materials
build_materials, coefficients _coeffs / _tm_coeffs / _apply_bead_tm / _tm_loss, CPML
make_boundary, VCSEL Huygens source, source plane npml + 2, flux planes at 10 % / 90 % of
the length, fdtd2d_core.py:445-560 of this package) and the instances are stepped by the
batch engine (lf_batch_*): one launch per wavefront class and temporally blocked pass
covers all instances, instances in their CW window step one step per launch.

Monitor set of a sweep solve: the CW phasor of the driven field over the last 8 periods
and the Poynting flux through the input / output planes, i.e. the CW transmission
S_out / S_in of run_beads (fdtd2d_core.py:897-902 there).  The per-step detector boxes
and the running <|E|^2> of run_beads are not accumulated (they are not part of the sweep
result), which leaves every step before the CW window free of monitors.
"""
import time

import numpy as np

from lens_sensor_amd.common import backend
from lens_sensor_amd.common.flux import PlaneFluxBook, guard_in_plane
from lens_sensor_amd.common.physics import C0, MU0, n_at
from lens_sensor_amd.engine.spec import new_steps, fill_phasors, cw_envelope
from lens_sensor_amd.planar_beads import fdtd2d_core as d2
from lens_sensor_amd.planar_beads.boundaries import make_boundary


def sweep_instances(n=64, seed=2024, dx_nm=25.0, nx_range=(2048, 16384), ny_range=(512, 2048)):
    """Seeded sweep of `n` instances: grid x ~ U[2048, 16384], y ~ U[512, 2048] cells,
    TE / TM alternating, one polystyrene bead (r ~ U(0.2, 0.6) um) at a seeded position in
    the aqueous layer, kappa_aqueous ~ U[0, 0.003], aqueous (3 um) and mucin (0.5 um)
    thicknesses x U[0.8, 1.2]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        r = float(rng.uniform(0.2, 0.6))
        L_um = nx * dx_nm * 1e-3
        t_aq = 3.0 * float(rng.uniform(0.8, 1.2))
        out.append(dict(nx=nx, ny=ny, pol='p' if i % 2 else 's', bead_r_um=r,
                        bead_x_um=float(rng.uniform(0.3, 0.7) * L_um),
                        bead_y_um=-r - float(rng.uniform(0.0, 1.0)) * max(0.0, t_aq - 2 * r),
                        kappa_aq=float(rng.uniform(0.0, 0.003)), t_aq_um=t_aq,
                        t_mu_um=0.5 * float(rng.uniform(0.8, 1.2)), dx_nm=dx_nm))
    return out


class Solve:
    """One CW solve of a sweep: its engine, step count, CW window and flux planes."""

    def __init__(self, p, mode='fp64acc', npml=24, wg_n=1.49, t_wg_um=2.0, lambda_nm=850.0,
                 n_cornea=1.376, steps_cap=None, engine_factory=None):
        dx = p['dx_nm'] * 1e-9
        self.p = p
        self.pol = p['pol']
        self.kind = 'tm2d' if self.pol == 'p' else 'te2d'
        Nx, Ny = p['nx'], p['ny']
        self.cells = Nx * Ny
        dt = 0.5 * dx / (C0 * np.sqrt(2))                 # fdtd2d_core.py:626
        T_WG = t_wg_um * 1e-6
        LAM = lambda_nm * 1e-9
        tear = Ny * dx - T_WG - d2.AIR_BUFFER
        stack = [dict(name='Aqueous', t_um=p['t_aq_um'], n=complex(1.333, p['kappa_aq']),
                      kappa=p['kappa_aq'], mat=None),
                 dict(name='Mucin', t_um=p['t_mu_um'], n=complex(1.34, 0.0), kappa=0.0, mat=None)]
        r = p['bead_r_um']
        beads = [dict(x_um=p['bead_x_um'], d_um=2 * r, y_um=p['bead_y_um'], n=1.59)]
        old_tb, old_lam = d2.TEAR_BUFFER, d2.LAM_LOSS
        d2.TEAR_BUFFER, d2.LAM_LOSS = tear, LAM
        try:
            src_ix = npml + 2
            x_facet_um, _air, zsrc = d2._vcsel_geom_2d(src_ix, dx, 0.0)
            L_um = Nx * dx * 1e6
            mp = d2._material_parts(
                Nx, Ny, dx, wg_n, 1.59, T_WG, p['bead_x_um'], 2 * r, 0.0, False,
                n_aq=1.333, n_mucin=1.34, n_cornea=n_cornea, t_aq_um=p['t_aq_um'],
                t_mu_um=p['t_mu_um'], t_lip_um=0.0, n_lip=n_at('lipid', lambda_nm),
                input_gap_um=x_facet_um, length_um=L_um, curved=False,
                bead_y_um=None, stack=stack, beads=beads)
            factory = engine_factory or backend.make_engine
            self.eng = eng = factory(self.kind, (Nx, Ny), mode)
            # coefficients in compact form, assembled by the engine (fdtd2d_core.window_coeffs)
            packed, ch_tm = d2.window_coeffs(self.pol, mp, dt, dx)
            eng.build_coeffs(ch_tm if self.pol == 'p' else np.float32(dt / (MU0 * dx)), packed)
            del packed
            f0 = C0 / LAM
            sigma_t = 4 / (2 * np.pi * f0)
            y0_m = -tear
            iy0 = int(round((T_WG / 2 - y0_m) / dx))
            n_src = d2._n_source_2d(src_ix * dx * 1e6, iy0, y0_m, dx, wg_n, T_WG,
                                    float(complex(stack[0]['n']).real), x_facet_um)
            _iy, vs, vs_ey = d2._vcsel_source_2d(Ny, dx, dt, T_WG, src_ix, 0.0, 3.0, lambda_nm,
                                                 0.0, n_src, zsrc, f0, y0_m=y0_m)
            bnd = make_boundary('cpml', self.pol, Nx, Ny, dx, dt, npml=npml, lam_m=LAM)
            bnd.attach(eng)
            eng.finalize()
            src = vs_ey if self.pol == 'p' else vs
            self.comp = 'Ey' if self.pol == 'p' else 'Ez'
            eng.set_source(self.comp, src.ix, src.amp, src.ph1, src.ph2)
            L = Nx * dx
            steps = int((6 * sigma_t + 2 * L * d2.N_G_FACTOR * wg_n / C0) / dt)   # :723
            self.steps_full = steps
            self.steps = min(steps, steps_cap) if steps_cap else steps
            spT = max(4, int(round((1.0 / f0) / dt)))
            self.n_cw = min(8 * spT, self.steps)                                 # :757-758
            acc = 'c128' if eng.wide_acc else 'c64p'
            self.mon_cw = eng.add_dft('ey_node' if self.pol == 'p' else 'Ez', 0, Nx, acc=acc)
            hw = max(1, int(round(min(d2.DET_LEN / 2, 0.05 * L) / dx)))
            cin, cout = int(round(0.1 * L / dx)), int(round(0.9 * L / dx))
            cin = guard_in_plane(cin, max(src_ix, int(round(x_facet_um / (dx * 1e6)))), Nx)
            self.book = PlaneFluxBook(eng, self.kind)
            self.book.add_plane('in', max(1, cin))
            self.book.add_plane('out', max(1, min(cout, Nx - 2)))
            self.cw_bits = (1 << self.mon_cw) | sum(1 << self.book.mon[k] for k in self.book.keys())
            self.omega, self.dt, self.sigma_t, self.src_omega = 2 * np.pi * f0, dt, sigma_t, src.omega
            self.hw = hw
        finally:
            d2.TEAR_BUFFER, d2.LAM_LOSS = old_tb, old_lam

    def masks(self, n_total):
        m = np.zeros(n_total, np.uint32)
        m[self.steps - self.n_cw:self.steps] = self.cw_bits
        return m

    def result(self):
        self.book.note_samples(self.book.keys(), self.n_cw)
        Ecw = (2.0 / self.n_cw) * self.eng.read_monitor(self.mon_cw)
        S_in, S_out = self.book.flux('in'), self.book.flux('out')
        Tr = S_out / S_in if abs(S_in) > 0 else float('nan')
        return dict(transmission=float(Tr), S_in=float(S_in), S_out=float(S_out),
                    ecw_norm=float(np.linalg.norm(Ecw)), steps=self.steps, cells=self.cells)


def step_table(omega, dt, sigma_t, src_omega, n):
    tab = new_steps(n)
    t = np.arange(n) * dt
    fill_phasors(tab, omega, dt, t, cw_envelope(t, sigma_t), src_omega=src_omega)
    return tab


def run_sweep(params, batched=True, mode='fp64acc', steps_cap=None, depth=6, lx=64,
              engine_factory=None, chunk=4096, log=None):
    """Solve every instance of `params`; batched = one batch engine for all, otherwise one
    instance after another with the one-step-per-launch path.  Returns (results, seconds of
    stepping, seconds of setup)."""
    t0 = time.time()
    solves = []
    for k, p in enumerate(params):
        solves.append(Solve(p, mode=mode, steps_cap=steps_cap, engine_factory=engine_factory))
        if log:
            log(f'[sweep] set up instance {k + 1}/{len(params)} ({p["nx"]} x {p["ny"]})')
    t_setup = time.time() - t0
    s0 = solves[0]
    for s in solves:
        if s.dt != s0.dt or s.omega != s0.omega:
            raise ValueError('a batched sweep needs one dx and one wavelength')
    nmax = max(s.steps for s in solves)
    tab = step_table(s0.omega, s0.dt, s0.sigma_t, s0.src_omega, nmax)
    for s in solves:
        s.eng.sync()
    t1 = time.time()
    if batched:
        from lens_sensor_amd.engine.hip import HipBatch
        b = HipBatch([s.eng for s in solves], depth=depth, lx=lx)
        masks = np.stack([s.masks(nmax) for s in solves])
        nact = np.array([s.steps for s in solves], np.int32)
        for c0 in range(0, nmax, chunk):
            c1 = min(nmax, c0 + chunk)
            b.run(tab[c0:c1], masks=masks[:, c0:c1], nact=np.clip(nact - c0, 0, c1 - c0))
            if log:
                log(f'[sweep] batched steps {c1}/{nmax}')
        b.close()
    else:
        for k, s in enumerate(solves):
            if hasattr(s.eng, 'set_temporal'):
                s.eng.set_temporal(1, 64)
            tt = tab[:s.steps].copy()
            tt['mask'] = s.masks(s.steps)
            for c0 in range(0, s.steps, chunk):
                s.eng.run(tt[c0:c0 + chunk])
            s.eng.sync()
            if log:
                log(f'[sweep] sequential instance {k + 1}/{len(solves)}')
    t_step = time.time() - t1
    results = [s.result() for s in solves]
    for s in solves:
        s.eng.close()
    return results, t_step, t_setup
