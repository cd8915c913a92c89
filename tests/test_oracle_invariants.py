"""Pinning of the CPU oracle by what the mathematics fixes (CPU only).

The oracle's whole-run output is pinned against fixtures the reference itself wrote
(test_characterization.py, test_setup_vectors.py).  These tests pin its parts
independently, each chosen so that a plausible slip fails it:

* exact Yee eigenmode recurrence in a PEC cavity (2D TE, 2D TM): a wrong sign, index or
  coefficient in any curl term breaks the discrete Laplacian and the recurrence
  p[n+1] + p[n-1] = (2 - Ch Ce mu) p[n] with mu = 4 (sin^2(kx/2) + sin^2(ky/2));
* discrete div H and (uniform-coefficient) div E conservation in 3D: div curl = 0 holds
  on the Yee grid only if all six curl components carry the right stencils and signs;
* linearity in the source amplitude and the float32 path agreeing with the FP64 oracle;
* CPML absorption: a pulse leaving a CPML-terminated box must not come back;
* sub-pixel fractions summing to the disc / ball volume; the CW phasor of a pure tone
  recovering its complex amplitude exactly.
"""
import numpy as np
import pytest

from lens_sensor_amd.common import subpixel
from lens_sensor_amd.engine.spec import new_steps, fill_phasors
from lens_sensor_amd.planar_beads.boundaries import cpml_profile_1d
from oracle.engine import OracleEngine


def _no_pml(N):
    return ((np.ones(N), np.ones(N), np.zeros(N)), (np.ones(N - 1), np.ones(N - 1), np.zeros(N - 1)))


def _quiet(n=1):
    return new_steps(n)


def test_te_cavity_eigenmode_recurrence():
    Nx, Ny, m, n = 31, 23, 3, 2
    o = OracleEngine('te2d', (Nx, Ny), 'f64')
    ch, ceh = 0.31, 0.27
    o.set_coeffs(ch, ceE=np.ones((Nx, Ny)), ceH=np.full((Nx, Ny), ceh))
    o.set_cpml(0, *_no_pml(Nx), True, True)
    o.set_cpml(1, *_no_pml(Ny), True, True)
    o.finalize()
    i, j = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing='ij')
    kx, ky = np.pi * m / (Nx - 1), np.pi * n / (Ny - 1)
    mode = np.sin(kx * i) * np.sin(ky * j)
    o.write_field('Ez', mode)
    mu = 4.0 * (np.sin(kx / 2) ** 2 + np.sin(ky / 2) ** 2)
    p = [1.0]
    for _ in range(60):
        o.run(_quiet())
        Ez = o.read_field('Ez')
        c = np.sum(Ez * mode) / np.sum(mode * mode)
        assert np.max(np.abs(Ez - c * mode)) <= 1e-12          # stays in the mode
        p.append(c)
    p = np.array(p)
    a = ch * ceh * mu
    assert abs(p[1] - (1.0 - a)) <= 1e-12
    res = p[2:] + p[:-2] - (2.0 - a) * p[1:-1]
    assert np.max(np.abs(res)) <= 1e-12


def test_tm_cavity_eigenmode_recurrence():
    Nx, Ny, m, n = 27, 33, 2, 3
    o = OracleEngine('tm2d', (Nx, Ny), 'f64')
    ch, ca = 0.29, 0.33
    o.set_coeffs(ch, cax=np.full((Nx - 1, Ny), ca), cay=np.full((Nx, Ny - 1), ca))
    o.set_cpml(0, *_no_pml(Nx), True, True)
    o.set_cpml(1, *_no_pml(Ny), True, True)
    o.finalize()
    i, j = np.meshgrid(np.arange(Nx - 1) + 0.5, np.arange(Ny - 1) + 0.5, indexing='ij')
    kx, ky = np.pi * m / (Nx - 1), np.pi * n / (Ny - 1)
    mode = np.cos(kx * i) * np.cos(ky * j)                   # Neumann modes of Hz
    o.write_field('Hz', mode)
    mu = 4.0 * (np.sin(kx / 2) ** 2 + np.sin(ky / 2) ** 2)
    q = [1.0]
    for _ in range(60):
        o.run(_quiet())
        Hz = o.read_field('Hz')
        c = np.sum(Hz * mode) / np.sum(mode * mode)
        assert np.max(np.abs(Hz - c * mode)) <= 1e-12
        q.append(c)
    q = np.array(q)
    res = q[2:] + q[:-2] - (2.0 - ch * ca * mu) * q[1:-1]
    assert np.max(np.abs(res)) <= 1e-12


def _div_h(o):
    Hx, Hy, Hz = (o.read_field(c) for c in ('Hx', 'Hy', 'Hz'))
    return (Hx[1:] - Hx[:-1]) + (Hy[:, 1:] - Hy[:, :-1]) + (Hz[:, :, 1:] - Hz[:, :, :-1])


def _div_e_interior(o):
    Ex, Ey, Ez = (o.read_field(c) for c in ('Ex', 'Ey', 'Ez'))
    return ((Ex[1:, 1:-1, 1:-1] - Ex[:-1, 1:-1, 1:-1]) + (Ey[1:-1, 1:, 1:-1] - Ey[1:-1, :-1, 1:-1])
            + (Ez[1:-1, 1:-1, 1:] - Ez[1:-1, 1:-1, :-1]))


def test_3d_divergence_conservation():
    dims = (14, 12, 16)
    o = OracleEngine('yee3d', dims, 'f64')
    rng = np.random.default_rng(3)
    from lens_sensor_amd.engine.spec import comp_shape
    o.set_coeffs(0.3, cex=np.full(comp_shape('yee3d', dims, 'Ex'), 0.25),
                 cey=np.full(comp_shape('yee3d', dims, 'Ey'), 0.25),
                 cez=np.full(comp_shape('yee3d', dims, 'Ez'), 0.25))
    for ax, N in enumerate(dims):
        o.set_cpml(ax, *_no_pml(N), True, True)
    o.finalize()
    for c in ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'):
        o.write_field(c, rng.standard_normal(comp_shape('yee3d', dims, c)))
    dh0, de0 = _div_h(o), _div_e_interior(o)
    o.run(_quiet(25))
    scale = max(np.abs(o.read_field(c)).max() for c in ('Hx', 'Hy', 'Hz', 'Ex', 'Ey', 'Ez'))
    assert np.max(np.abs(_div_h(o) - dh0)) <= 1e-11 * scale
    assert np.max(np.abs(_div_e_interior(o) - de0)) <= 1e-11 * scale


def _te_with_source(mode, amp, steps=80, Nx=60, Ny=44, npml=8):
    o = OracleEngine('te2d', (Nx, Ny), mode)
    dx, dt = 25e-9, 0.5 * 25e-9 / (3e8 * np.sqrt(2))
    o.set_coeffs(np.float32(dt / (4e-7 * np.pi * dx)), ceE=np.ones((Nx, Ny), np.float32),
                 ceH=np.full((Nx, Ny), dt / (8.8541878128e-12 * 2.25 * dx), np.float32))
    for ax, N in enumerate((Nx, Ny)):
        e, h = cpml_profile_1d(N, npml, npml, dx, dt, lam_m=850e-9)
        o.set_cpml(ax, e, h, True, True)
    o.finalize()
    y = np.arange(Ny)
    prof = np.exp(-((y - Ny / 2) / 6.0) ** 2).astype(np.float32)
    o.set_source('Ez', npml + 2, amp * prof, np.zeros(Ny, np.float32), np.full(Ny, 0.4, np.float32))
    tab = new_steps(steps)
    t = np.arange(steps) * dt
    fill_phasors(tab, 2 * np.pi * 3.527e14, dt, t, 1 - np.exp(-(t / (20 * dt)) ** 2))
    o.run(tab)
    return o


def test_linearity_in_source_amplitude():
    a = _te_with_source('f64', 1.0).read_field('Ez')
    b = _te_with_source('f64', 2.0).read_field('Ez')
    assert np.max(np.abs(a)) > 1e-3
    np.testing.assert_allclose(b, 2.0 * a, rtol=0, atol=1e-12 * np.max(np.abs(a)))


def test_float32_path_matches_fp64_oracle():
    a = _te_with_source('ref32', 1.0).read_field('Ez').astype(np.float64)
    b = _te_with_source('f64', 1.0).read_field('Ez')
    rel = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel <= 1e-4, rel


def test_cpml_absorbs_outgoing_pulse():
    """A pulse started in a CPML box of 40 x 40 cells (8-cell CPML): after it has left, the
    field left inside is tiny; the same box with the psi correction reversed in sign
    (a plausible slip) or without CPML keeps far more."""
    Nx = Ny = 40

    def run(flip):
        o = OracleEngine('te2d', (Nx, Ny), 'f64')
        dx, dt = 25e-9, 0.5 * 25e-9 / (3e8 * np.sqrt(2))
        o.set_coeffs(dt / (4e-7 * np.pi * dx), ceE=np.ones((Nx, Ny)),
                     ceH=np.full((Nx, Ny), dt / (8.8541878128e-12 * dx)))
        for ax in (0, 1):
            (ke, be, ce), (kh, bh, chh) = cpml_profile_1d(Nx, 8, 8, dx, dt, lam_m=850e-9)
            s = -1.0 if flip else 1.0
            o.set_cpml(ax, (ke, be, s * ce.astype(np.float64)), (kh, bh, s * chh.astype(np.float64)),
                       True, True)
        o.finalize()
        i, j = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing='ij')
        E0 = np.exp(-((i - 20) ** 2 + (j - 20) ** 2) / 8.0)
        o.write_field('Ez', E0)
        o.run(_quiet(400))
        return np.abs(o.read_field('Ez')[8:-8, 8:-8]).max()

    good, bad = run(False), run(True)
    assert good < 2e-3, good
    assert bad > 10 * good


def test_subpixel_fractions_sum_to_disc_and_ball():
    r, dx = 0.43, 0.05
    n = 24
    xs = (np.arange(n) - n / 2 + 0.5) * dx
    f2 = subpixel.circle_fraction(xs, xs, 0.013, -0.021, r, dx)
    assert f2.min() >= 0.0 and f2.max() <= 1.0
    assert abs(f2.sum() * dx * dx - np.pi * r * r) <= 0.01 * np.pi * r * r
    f3 = subpixel.sphere_fraction(xs, xs, xs, 0.011, -0.017, 0.007, r, dx)
    assert f3.min() >= 0.0 and f3.max() <= 1.0
    assert abs(f3.sum() * dx ** 3 - 4.0 / 3.0 * np.pi * r ** 3) <= 0.02 * 4.0 / 3.0 * np.pi * r ** 3


def test_cw_phasor_of_pure_tone_is_exact():
    """A field that is A cos(w t + phi) for an integer number of periods accumulates
    sum F e^{-i w t} = (N/2) A e^{i phi} (complex128 accumulator)."""
    Nx, Ny = 5, 4
    o = OracleEngine('te2d', (Nx, Ny), 'fp64acc')
    o.set_coeffs(0.0, ceE=np.ones((Nx, Ny), np.float32), ceH=np.zeros((Nx, Ny), np.float32))
    o.set_cpml(0, *_no_pml(Nx), False, False)
    o.set_cpml(1, *_no_pml(Ny), False, False)
    o.finalize()
    mid = o.add_dft('Ez', 0, Nx, acc='c128')
    N, periods, dt = 96, 4, 1e-16
    w = 2 * np.pi * periods / (N * dt)
    A = np.linspace(0.5, 1.5, Nx * Ny).reshape(Nx, Ny)
    phi = np.linspace(-2.0, 2.5, Nx * Ny).reshape(Nx, Ny)
    t = np.arange(N) * dt
    tab = new_steps(N)
    fill_phasors(tab, w, dt, t, np.ones(N))
    tab['mask'] = 1
    for n in range(N):
        o.write_field('Ez', (A * np.cos(w * t[n] + phi)).astype(np.float32))
        o.run(tab[n:n + 1])
    got = o.read_monitor(mid) * (2.0 / N)
    want = A * np.exp(1j * phi)
    assert np.max(np.abs(got - want)) <= 1e-6          # float32 samples of the tone
