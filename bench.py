"""Benchmark of the Yee time-stepping hot path on one MI355X.

Workload (default, --workload tm2d_stitch): 2D TM (Ey/Hx/Hz) stitch mode, FP32 fields
with FP64 phasor/flux sums: 10 windows of 8192 x 1536 cells at dx = 25 nm, 24-cell CPML
(inner seams end in the non-assembled CPML extension), 20000 steps per window with the
complex-phasor handoff; PMMA core n=1.49 (2 um) over aqueous 1.331+0.001i (3 um),
mucin 1.34 (0.5 um), cornea 1.376; 0-8 polystyrene beads (n=1.59, kappa 0.002,
r ~ U(0.2, 0.6) um) per window drawn from the seed.  A "step" is one Yee time step
of one window (12.6 M cells); value = cell updates / second of device stepping time.

--workload yee3d: 3D FP32 Yee + CPML, 512 x 256 x 256 at dx = 30 nm, 12-cell CPML on
all faces, 4-layer stack + lipid, one bead, Gaussian CW source, 8000 steps, FP64 phasor
sums on one monitor plane; the detector boxes of every step are summed inside the
temporally blocked passes (depth 3), the steps of the final CW window run one per launch.

--workload pulse2d: the two pulse modes of the 2D stitch on a lossy TM geometry (40 um
long, dx = 20 nm, 30 um windows every 18 um, 30-period pulse): the Fourier stitch (one CW
stitch per comb line) and the time-domain TF/SF stitch; value = cell updates / second of
the two whole solves (set-up included).

Prints ONE JSON line: metric, value, roofline of the dominant kernel (the blocked pass;
frac from the counter-measured HBM bytes of the pass), CPU baseline of the NumPy oracle
on a bounded sample of the same workload.  The default run adds, in the same line:
  yee3d        the 3D workload above, with its own roofline and CPU baseline;
  fp64_rel_l2  relative L2 of the device's CW phasors against the FP64 oracle on a
               cropped copy of the stitch workload (2 windows of 1024 x 512, 20000 steps
               each, the same blocking), from tests/golden/fp64_tm_stitch_crop.npz.
"""
import argparse
import contextlib
import io
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell update: every field read+written once, every per-cell
# coefficient read once (psi slabs, sources and monitors not counted)
BYTES_TM2D_LOSSY = 3 * 8 + 4 * 4      # Ex, Ey, Hz r/w + Ca_ex, Ca_ey, D_ex, D_ey = 40 B
BYTES_YEE3D = 6 * 8 + 3 * 4           # 6 fields r/w + ce_x, ce_y, ce_z = 60 B
# HBM bytes per blocked pass (all its launches) from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch), scaled by the same counters on a plain 1 GiB copy
# (fetch x2.0, write x1.0); profiles/*_pmc.md hold the runs.  Keyed by (workload, depth,
# x chunk): a pass of another blocking setting has no entry and reports traffic None.
TRAFFIC_BYTES = {
    # mean over the ten windows: the source window's pass (409.36 MB, four launches) and
    # nine driven windows' passes (526.20 MB, three launches); the slab-free class without
    # excitation reads the coefficients' compact form (layer-column rows from L2, the
    # bead band from HBM)
    ('tm2d_stitch', 8, 128): 514_517_984,
    # yee_tb_kernel<3, false>: FETCH 1,479,486.5 KiB x 2048 + WRITE 1,043,593.9 KiB x 1024
    ('yee3d', 3, 96): 3_029_988_357 + 1_068_640_115,
}


def _timed_factory(records, every=1, fixed=None):
    """Engine class that records HIP-event times of its runs.  every > 1: only one run()
    call in `every` carries the events (the per-launch events of a timed run add to the
    stream's own time, so a sample of the calls is timed and the rest run bare).
    fixed = (depth, lx, lead): blocking setting of every engine (the autotuner is off)."""
    from lens_sensor_amd.engine.hip import HipEngine

    class TimedHip(HipEngine):
        def __init__(self, kind, dims, mode):
            super().__init__(kind, dims, mode)
            self.set_timing(True)
            if fixed is not None:
                self.set_temporal(fixed[0], fixed[1])
                if kind != 'yee3d':
                    self.set_tb_lead(fixed[2])
            self._upd = 0.0
            self._tot = 0.0
            self._n = 0
            self._tb_ms = 0.0
            self._tb_n = 0
            self._tb_depth = 1
            self._tb_units = 0
            self._calls = 0
            if not hasattr(self, '_cfg'):
                self._cfg = (None, None)
            records.append(self)

        def set_temporal(self, depth, lx=64):
            super().set_temporal(depth, lx)
            self._cfg = (depth, lx)

        def autotune(self, row, reps=None):
            d, lx, ld, ms = super().autotune(row, reps)
            self._cfg = (d, lx)
            return d, lx, ld, ms

        def tb_lx(self):
            return self._cfg[1]

        def run(self, tab):
            on = self._calls % every == every // 2
            self._calls += 1
            if every > 1:
                self.set_timing(on)
            super().run(tab)
            if not on:
                return
            u, t, n = self.kernel_ms()
            self._upd += u
            self._tot += t
            self._n += n
            ms, nl, d = self.tb_stats()
            self._tb_ms += ms
            self._tb_n += nl
            self._tb_depth = d
            self._tb_units += nl * d * int(np.prod(self.dims))    # cell updates of the passes

    return TimedHip


def bench_tm2d_stitch(args, warmup=False):
    """The stitch workload (stitch_bench.run) on engines that record HIP-event times.
    warmup: one window of ``args.warmup`` steps, untimed (loads the code objects, runs
    the autotuner's first trials)."""
    from lens_sensor_amd.common import backend
    from lens_sensor_amd.planar_beads import stitch_bench as sb
    nx_win, ny, slide = 8192, 1536, 6144
    n_win = 10 if args.windows is None else args.windows
    steps = 20000 if args.steps_override is None else args.steps_override
    if warmup:
        n_win, steps = 1, args.warmup
    records = []
    info = []
    old_at = backend.AUTOTUNE
    fixed = None
    if args.temporal:
        fixed = tuple(int(v) for v in args.temporal.split(','))
        backend.AUTOTUNE = False
    t0 = time.time()
    try:
        with backend.use_engine(_timed_factory(records, every=5, fixed=fixed)), \
                contextlib.redirect_stdout(io.StringIO()):
            sb.run(nx_win=nx_win, ny=ny, slide=slide, n_windows=n_win, steps=steps,
                   seed=args.seed, engine_mode='fp64acc',
                   window_hook=lambda w, i: (info.append(i),
                                             print(f'[bench] window {w} {i}', file=sys.stderr,
                                                   flush=True)))
    finally:
        backend.AUTOTUNE = old_at
    wall = time.time() - t0
    cells = [i['cells'] for i in info]
    step_s = sum(i['seconds'] for i in info)
    cell_updates = sum(c * i['steps'] for c, i in zip(cells, info))
    upd_ms = sum(e._upd for e in records)
    n_upd = sum(e._n for e in records)
    return dict(cell_updates=cell_updates, step_seconds=step_s, wall=wall,
                steps=sum(i['steps'] for i in info),
                upd_ms_per_step=upd_ms / max(n_upd, 1), cells_per_step=float(np.mean(cells)),
                bytes_per_cell=BYTES_TM2D_LOSSY, windows=len(info), grid=(nx_win, ny),
                steps_per_window=steps, tb_ms=sum(e._tb_ms for e in records),
                tb_launches=sum(e._tb_n for e in records), tb_depth=records[0]._tb_depth,
                tb_lx=records[0].tb_lx(), tb_units=sum(e._tb_units for e in records),
                kernel='tm_tb_kernel')


def fp64_companion():
    """Relative L2 error of the device path against the FP64 oracle on the cropped copy of
    the stitch workload (two 1024 x 512 windows, 20000 steps each, depth 8 and x chunk
    128 as the full run), from the committed FP64 fixture."""
    from lens_sensor_amd.common import backend
    from lens_sensor_amd.engine.hip import HipEngine
    from lens_sensor_amd.planar_beads import stitch_bench as sb
    path = os.path.join(HERE, 'tests', 'golden', 'fp64_tm_stitch_crop.npz')
    ref = np.load(path)

    def blocked(kind, dims, mode):
        e = HipEngine(kind, dims, mode)
        e.set_temporal(8, 128)
        return e
    old_at = backend.AUTOTUNE
    backend.AUTOTUNE = False
    try:
        with backend.use_engine(blocked), contextlib.redirect_stdout(io.StringIO()):
            r = sb.run(engine_mode='fp64acc', return_efull=True, **sb.CROP)
    finally:
        backend.AUTOTUNE = old_at
    sx, sy = sb.CROP_SUB

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return dict(efull=rel(np.asarray(r['Efull'])[::sx, ::sy], ref['efull']),
                ex_phasor=rel(np.asarray(r['Ex_cw_amp'])[::sx, ::sy], ref['ex_amp']),
                transmission=abs(r['transmission'] - float(ref['transmission'])) / abs(float(ref['transmission'])),
                bound=1e-4, case='2D TM stitch, 2 windows of 1024x512, 20000 steps each, depth 8, '
                                 'x chunk 128, fp64acc; vs the f64 oracle (every 4th cell)')


def bench_yee3d(args):
    from lens_sensor_amd.common import backend
    from lens_sensor_amd.planar_3d import fdtd3d_core as f3d
    rng = np.random.default_rng(args.seed)
    steps = 8000 if args.steps_override is None else args.steps_override
    dx_um = 0.03
    nx, ny, nz = 512, 256, 256
    stack = [dict(name='Lipid', t_um=0.1, n=1.46, kappa=0.0, mat=None),
             dict(name='Aqueous', t_um=3.0, n=1.333, kappa=0.0, mat=None),
             dict(name='Mucin', t_um=0.5, n=1.34, kappa=0.0, mat=None)]
    t_wg = 2.0
    air = 0.6
    tear = ny * dx_um - t_wg - air
    bead = dict(x_um=float(rng.uniform(6.0, 12.0)), d_um=1.0, y_um=-0.6, n=1.59,
                z_um=float(rng.uniform(-1.0, 1.0)))
    records = []
    old_at = backend.AUTOTUNE
    fixed = None
    if args.temporal:
        fixed = tuple(int(v) for v in args.temporal.split(','))
        backend.AUTOTUNE = False
    t0 = time.time()
    try:
        with backend.use_engine(_timed_factory(records, fixed=fixed)), \
                contextlib.redirect_stdout(io.StringIO()):
            r = f3d.run_3d('bench3d', 1.49, t_wg_um=t_wg, lam_nm=850.0, lx_um=nx * dx_um, air_um=air,
                           tear_um=tear, lz_um=nz * dx_um, dx_nm=30.0, vcsel_waist_um=1.5,
                           vcsel_waist_z_um=1.5, n_snapshots=1, bead=bead, stack=stack,
                           n_co=1.376, npml=12, engine_mode='fp64acc', steps_override=steps,
                           cw_volume=False)
    finally:
        backend.AUTOTUNE = old_at
    wall = time.time() - t0
    e = records[0]
    cells = nx * ny * nz
    return dict(cell_updates=cells * steps, step_seconds=e._tot / 1e3, wall=wall, steps=e._n,
                upd_ms_per_step=e._upd / max(e._n, 1), cells_per_step=float(cells),
                bytes_per_cell=BYTES_YEE3D, grid=(nx, ny, nz), steps_per_window=steps,
                tb_ms=e._tb_ms, tb_launches=e._tb_n, tb_depth=e._tb_depth,
                tb_lx=e.tb_lx(), tb_units=e._tb_units, kernel='yee_tb_kernel',
                pass_desc='one blocked pass, persistent workgroups')


def bench_sweep(args):
    """64-instance batched 2D sweep vs the same 64 solves one after another with one step
    per launch; both over the same seeded instances and the same steps per solve."""
    from lens_sensor_amd.planar_beads import sweep
    ps = sweep.sweep_instances(64, seed=args.seed)
    cap = args.steps_override or 12288
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    rb, tb, sb = sweep.run_sweep(ps, batched=True, steps_cap=cap, log=log)
    rs, ts, ss = sweep.run_sweep(ps, batched=False, steps_cap=cap, log=log)
    same = all(np.array_equal([a['transmission'], a['S_in'], a['S_out'], a['ecw_norm']],
                              [b['transmission'], b['S_in'], b['S_out'], b['ecw_norm']], equal_nan=True)
               for a, b in zip(rb, rs))
    cell_updates = sum(r['cells'] * r['steps'] for r in rb)
    return dict(solves=len(ps), batched_s=tb, sequential_s=ts, setup_s=(sb + ss) / 2,
                steps_cap=cap, identical=same, cell_updates=cell_updates,
                full_steps=[int(x) for x in [min(r['steps'] for r in rb), max(r['steps'] for r in rb)]])


def bench_serve(args):
    """Seeded Poisson stream (mean 20 req/s) of CW solve requests, 80 % 2D windows
    (4096-16384 x 1024-2048), 20 % 3D windows (256-512 x 192-256 x 192-256), admitted
    against a 220 GB HBM budget, resident until the one-period phasor change < 1e-4."""
    from lens_sensor_amd import serving
    n = args.requests
    reqs = serving.request_stream(n, rate_hz=20.0, seed=args.seed)
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    t0 = time.time()
    recs, summ = serving.Server(log=log).run(reqs, realtime=True)
    wall = time.time() - t0
    cells = sum(int(np.prod(r['dims'])) * r['steps'] for r in recs)
    return recs, summ, wall, cells


PULSE_CASE = dict(wg_mat='pmma', bead_mat='polystyrol', bead_d_um=1.0, dx_nm=20.0,
                  save_frames=False, n_snapshots=24, wg_thickness_um=5.0, lambda_nm=850.0,
                  method='stitch', source_type='pulse', pulse_cycles=30.0, length_um=40.0,
                  bead_x_um=20.0, window_w_um=30.0, slide_um=18.0, polarization='p',
                  engine_mode='fp64acc')


def bench_pulse(args):
    """The two pulse modes of the 2D stitch on one lossy TM geometry (40 um long, dx 20 nm,
    30 um windows every 18 um, 30-period pulse at 850 nm): the Fourier stitch (one CW
    stitch per comb line) and the time-domain TF/SF stitch.  Wall time of each whole
    solve, set-up included."""
    from lens_sensor_amd.planar_beads import fdtd2d_core as d2
    out = {}
    case = dict(PULSE_CASE, pulse_nfreq=args.pulse_lines)
    d2.KAPPA = {'aq': 0.001, 'bead': 0.002}
    try:
        for mode in args.pulse_modes.split(','):
            print(f'[pulse] {mode}', file=sys.stderr, flush=True)
            t0 = time.time()
            r = d2.run_beads(pulse_mode=mode, **case)
            wall = time.time() - t0
            out[mode] = dict(wall_s=wall, cell_updates=int(r['cell_updates']),
                             mcups=r['cell_updates'] / wall / 1e6,
                             transmission=float(r['transmission']),
                             lines=(r['pulse'].get('nfreq') if mode == 'fourier' else None),
                             windows=[dict(steps=w['steps'], cells=w['cells'], s=round(w['seconds'], 2))
                                      for w in r.get('windows') or []] if mode == 'td' else None)
    finally:
        d2.KAPPA = {}
    return out


def cpu_baseline(workload, seconds=12.0):
    """NumPy oracle (one core) on a bounded sample: the same grid and coefficients, as many
    steps as fit in ~``seconds``; Mcell-updates/s."""
    from oracle.engine import OracleEngine
    from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
    from lens_sensor_amd.engine.spec import new_steps, fill_phasors
    if workload == 'tm2d_stitch':
        kind, dims = 'tm2d', (8192 + 28, 1536)
    else:
        kind, dims = 'yee3d', (512, 256, 256)
    eng = OracleEngine(kind, dims, 'ref32')
    dx, dt = 25e-9, 0.5 * 25e-9 / (3e8 * np.sqrt(2))
    if kind == 'tm2d':
        eng.set_coeffs(np.float32(0.4), cax=np.full((dims[0] - 1, dims[1]), 0.1, np.float32),
                       cay=np.full((dims[0], dims[1] - 1), 0.1, np.float32),
                       dax=np.full((dims[0] - 1, dims[1]), 0.999, np.float32),
                       day=np.full((dims[0], dims[1] - 1), 0.999, np.float32))
    else:
        sh = {'cex': (dims[0] - 1, dims[1], dims[2]), 'cey': (dims[0], dims[1] - 1, dims[2]),
              'cez': (dims[0], dims[1], dims[2] - 1)}
        eng.set_coeffs(np.float32(0.3), **{k: np.full(v, 0.1, np.float32) for k, v in sh.items()})
    for ax, N in enumerate(dims):
        e, h = cpml_profile_1d(N, 24 if kind == 'tm2d' else 12, 24 if kind == 'tm2d' else 12,
                               dx, dt, lam_m=850e-9)
        eng.set_cpml(ax, e, h, True, True)
    eng.finalize()
    comp = 'Ey' if kind == 'tm2d' else 'Ez'
    plane = eng.F[comp].shape[1:]
    eng.set_source(comp, 30, np.ones(plane, np.float32), np.zeros(plane, np.float32),
                   np.full(plane, 0.7, np.float32))
    tab = new_steps(1)
    fill_phasors(tab, 2 * np.pi * 3.5e14, dt, np.array([1e-15]), np.array([1.0]))
    eng.run(tab)                                   # warm-up
    n, t0 = 0, time.time()
    while time.time() - t0 < seconds:
        eng.run(tab)
        n += 1
    el = time.time() - t0
    cells = int(np.prod(dims))
    return dict(value=cells * n / el / 1e6, unit='Mcell-updates/s', cores=1, kind='port',
                sample=f'{n} {kind} steps of the NumPy oracle on {"x".join(map(str, dims))} '
                       f'cells (same grid and CPML as the workload), {el:.1f}s, one core; '
                       f'host has {os.cpu_count()} CPUs')


def copy_probe(mib):
    """Plain device copy of known size (calibration of the HBM byte counters): ``mib`` MiB
    read + ``mib`` MiB written per launch, 20 launches; prints one JSON line."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device='cuda:0')
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(20):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 20
    print(json.dumps(dict(probe='copy', bytes_read=n * 4, bytes_written=n * 4, ms_per_copy=ms,
                          gbs=2 * n * 4 / (ms * 1e-3) / 1e9)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='accepted for the driver contract: the workload fixes its steps')
    ap.add_argument('--warmup', type=int, default=16,
                    help='tm2d_stitch: steps of one untimed warm-up window (0: none)')
    ap.add_argument('--workload', default='tm2d_stitch',
                    choices=('tm2d_stitch', 'yee3d', 'sweep', 'serve', 'pulse2d'))
    ap.add_argument('--requests', type=int, default=24, help='serve: requests in the stream')
    ap.add_argument('--pulse-lines', type=int, default=None,
                    help='pulse2d: comb lines of the Fourier stitch (default: the whole band)')
    ap.add_argument('--pulse-modes', default='fourier,td', help='pulse2d: fourier, td or both')
    ap.add_argument('--windows', type=int, default=None, help='tm2d_stitch: windows to run')
    ap.add_argument('--steps-override', type=int, default=None,
                    help='steps per window (reduced runs for profiling only)')
    ap.add_argument('--temporal', default=None, metavar='DEPTH,LX,LEAD',
                    help='tm2d_stitch: fixed blocking setting instead of the autotuner '
                         '(profiling runs of one configuration)')
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-3d', action='store_true', help='tm2d_stitch: skip the 3D record')
    ap.add_argument('--no-fp64', action='store_true',
                    help='tm2d_stitch: skip the FP64-error companion run')
    ap.add_argument('--copy-probe', type=int, default=None, metavar='MIB',
                    help='only run a plain device copy of MIB MiB (byte-counter calibration)')
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print('bench.py: this build measures one GPU only (--gpus 1); multi-GPU runs are not '
              'part of it', file=sys.stderr)
        sys.exit(2)
    from lens_sensor_amd.engine import hip
    dev = hip.device_name()
    if args.workload == 'serve':
        recs, summ, wall, cells = bench_serve(args)
        out = dict(metric='solves/s', value=summ['solves_per_s'], unit='solves/s', n_gpus=1,
                   steps=int(sum(r['steps'] for r in recs)), warmup=0,
                   ms_per_step=None, higher_is_better=True, scaling='weak', vs_baseline=None,
                   dtype='fp32', data='synthetic',
                   config=dict(workload='serve_poisson20_2d80_3d20', requests=summ['requests'],
                               hbm_budget_gb=220, tol=1e-4, fields='fp32', phasors='fp64'),
                   device=dev, latency_p50_s=summ['p50_s'], latency_p99_s=summ['p99_s'],
                   converged=summ['converged'], wall_s=wall, mcups=cells / wall / 1e6,
                   per_request=[dict(id=r['id'], kind=r['kind'], dims=list(r['dims']), steps=r['steps'],
                                     latency_s=round(r['latency'], 3), converged=bool(r['converged']))
                                for r in recs])
        print(json.dumps(out), flush=True)
        return
    if args.workload == 'pulse2d':
        r = bench_pulse(args)
        tot = sum(v['cell_updates'] for v in r.values())
        wall = sum(v['wall_s'] for v in r.values())
        r.setdefault('fourier', None)
        r.setdefault('td', None)
        out = dict(metric='Mcell-updates/s', value=tot / wall / 1e6, unit='Mcell-updates/s',
                   n_gpus=1, steps=None, warmup=0, ms_per_step=None, higher_is_better=True,
                   scaling='weak', vs_baseline=None, dtype='fp32', data='synthetic',
                   config=dict(workload='pulse2d_tm_40um_fourier_and_td', dx_nm=20.0,
                               window_um=30.0, slide_um=18.0, pulse_cycles=30.0,
                               comb_lines=args.pulse_lines or 'band', modes=args.pulse_modes,
                               fields='fp32', phasors='fp64'),
                   device=dev, fourier=r['fourier'], td=r['td'])
        print(json.dumps(out), flush=True)
        return
    if args.workload == 'sweep':
        r = bench_sweep(args)
        out = dict(metric='solves/s', value=r['solves'] / r['batched_s'], unit='solves/s',
                   n_gpus=1, steps=r['steps_cap'], warmup=0,
                   ms_per_step=r['batched_s'] * 1e3 / r['steps_cap'], higher_is_better=True,
                   scaling='weak', vs_baseline=None, dtype='fp32', data='synthetic',
                   config=dict(workload='sweep_64x2d_te_tm', instances=r['solves'],
                               steps_per_solve=r['steps_cap'], dx_nm=25.0, npml=24,
                               fields='fp32', phasors='fp64'),
                   device=dev,
                   sequential=dict(value=r['solves'] / r['sequential_s'], unit='solves/s',
                                   note='same solves, one instance after another, one step per launch'),
                   speedup=r['sequential_s'] / r['batched_s'], identical_results=r['identical'],
                   mcups_batched=r['cell_updates'] / r['batched_s'] / 1e6,
                   mcups_sequential=r['cell_updates'] / r['sequential_s'] / 1e6,
                   setup_s_per_pass=r['setup_s'])
        print(json.dumps(out), flush=True)
        return
    warm = 0
    if args.warmup > 0 and args.workload == 'tm2d_stitch':
        bench_tm2d_stitch(args, warmup=True)
        warm = args.warmup
    if args.workload == 'tm2d_stitch':
        res = bench_tm2d_stitch(args)
        cfg = dict(workload='tm2d_stitch_10x8192x1536', windows=res['windows'],
                   grid=list(res['grid']), steps_per_window=res['steps_per_window'],
                   dx_nm=25.0, npml=24, fields='fp32', phasors='fp64')
    else:
        res = bench_yee3d(args)
        cfg = dict(workload='yee3d_512x256x256', grid=list(res['grid']),
                   steps=res['steps_per_window'], dx_nm=30.0, npml=12, fields='fp32',
                   phasors='fp64')
    out = dict(metric='Mcell-updates/s', value=None, unit='Mcell-updates/s', n_gpus=1,
               steps=res['steps'], warmup=warm, higher_is_better=True, scaling='weak',
               vs_baseline=None, dtype='fp32', data='synthetic', config=cfg, device=dev)
    out.update(_line(res, args.workload))
    if args.workload == 'tm2d_stitch' and not args.no_3d and args.windows is None \
            and args.steps_override is None and not args.temporal:
        # the 3D FP32 Yee + CPML step (512 x 256 x 256, 8000 steps) measured in the same run
        r3 = bench_yee3d(args)
        rec = dict(config=dict(workload='yee3d_512x256x256', grid=list(r3['grid']),
                               steps=r3['steps_per_window'], dx_nm=30.0, npml=12, fields='fp32',
                               phasors='fp64'))
        rec.update(_line(r3, 'yee3d'))
        if not args.no_cpu:
            rec['cpu_baseline'] = cpu_baseline('yee3d')
        out['yee3d'] = rec
    if args.workload == 'tm2d_stitch' and not args.no_fp64:
        out['fp64_rel_l2'] = fp64_companion()
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(args.workload)
    print(json.dumps(out), flush=True)


def _line(res, workload):
    """value, ms_per_step and the roofline record of one stepping workload.

    Dominant kernel: the temporally blocked pass (all its launches; HIP events on the
    context stream bracket each pass).  One pass advances every cell by tb_depth steps.
      traffic         HBM bytes per pass from rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE
                      (copy-calibrated, TRAFFIC_BYTES), None where no profile matches
      achieved / frac traffic / pass time, / 8 TB/s
      algorithmic_*   the pass's minimum: every field read and written once and every
                      coefficient read once per pass (bytes_per_cell_update x cells)
      effective_gbs   bytes_per_cell_update x cell updates per pass / pass time: what a
                      code stepping once per launch would have to move (exceeds the peak
                      under temporal blocking)
    """
    value = res['cell_updates'] / res['step_seconds'] / 1e6
    ms_step = res['step_seconds'] * 1e3 / max(1, res['steps'])
    B = res['bytes_per_cell']
    if res['tb_launches']:
        launch_ms = res['tb_ms'] / res['tb_launches']
        units = res['tb_units'] / res['tb_launches']             # cell updates per pass
        cells = units / res['tb_depth']
        kname = (f"{res['kernel']}<{res['tb_depth']}> "
                 f"({res.get('pass_desc', 'one blocked pass: slab + interior launches')})")
    else:
        launch_ms = res['upd_ms_per_step']
        units = cells = res['cells_per_step']
        kname = f"{res['kernel']} (one step)"
    sec = launch_ms * 1e-3
    traffic = TRAFFIC_BYTES.get((workload, res['tb_depth'], res.get('tb_lx')))
    alg_bytes = B * cells
    hbm = traffic / sec / 1e9 if traffic else None
    roof = dict(bound='hbm', achieved=hbm, peak=HBM_PEAK_GBS, unit='GB/s',
                frac=hbm / HBM_PEAK_GBS if hbm else None, traffic=traffic, kernel=kname,
                depth=res['tb_depth'], x_chunk=res.get('tb_lx'), launch_ms=launch_ms,
                cell_updates_per_launch=units, bytes_per_cell_update=B,
                algorithmic_bytes_per_launch=alg_bytes,
                algorithmic_gbs=alg_bytes / sec / 1e9,
                algorithmic_frac=alg_bytes / sec / 1e9 / HBM_PEAK_GBS,
                effective_gbs=B * units / sec / 1e9,
                traffic_per_algorithmic=(traffic / alg_bytes) if traffic else None,
                update_ms_per_step=res['upd_ms_per_step'])
    return dict(value=value, ms_per_step=ms_step, roofline=roof, wall_s=res['wall'])


if __name__ == '__main__':
    main()
