"""The oracle's conductive-media coefficients and its Ricker and cw waveforms against what pins them: golden vectors
from the reference's own numpy path with sigma_e > 0 and those waveforms (tests/golden/make_lossy_golden.py; 2D FP64,
3D float32), and closed forms: a uniform field in a periodic conductive box has no curl, so each step multiplies it
by ca = (1 - s) / (1 + s), s = sigma dt / (2 eps0 eps_r) (the semi-implicit conduction term); the Ricker wavelet's
peak, zeros and spectral peak; the cw ramp's value at one ramp time."""
import math
import os

import numpy as np
import pytest

from oracle.yee2d import EPS0, FACES, Boundaries, Yee2D, make_temporal
from oracle.yee3d import Yee3D
from tests.cases import LOSSY2D, LOSSY3D, build_lossy2d, build_lossy3d

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _boundaries(faces, npml):
    return Boundaries(dict(zip(FACES, faces)), npml=npml)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_oracle_reproduces_reference_lossy_case(polarization):
    """Conductive rectangles (one around the Ricker source line), a cw line source over part of a line: the FP64
    oracle reproduces the reference's FP64 run (sigma_e raster, both calibrated profiles, both waveforms, m = 1 with a
    cw source, fields at two steps to 1e-12 of the peak, the flux spectrum to 1e-12)."""
    golden = np.load(os.path.join(GOLDEN, f'lossy2d_{polarization}.npz'))
    sim, sources, monitor = build_lossy2d(Yee2D, polarization, _boundaries)
    assert (sim.nx, sim.ny) == (int(golden['nx']), int(golden['ny']))
    np.testing.assert_array_equal(sim.materials['sigma_e'].astype(np.float64), golden['sigma_e'])
    assert golden['sigma_e'].max() > 1e5
    done = 0
    for step in LOSSY2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        for name, values in sim.fields.items():
            want = golden[f'{name}_{step}']
            assert np.abs(want).max() > 0
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert float(sim.dt) == float(golden['dt'])
    times = (np.arange(LOSSY2D['waveform_samples'], dtype=np.float64) + 0.5) * float(sim.dt)
    for index, source in enumerate(sources):
        np.testing.assert_array_equal(source.profile.astype(np.float64), golden[f'source{index}_profile'])
        np.testing.assert_array_equal([source.temporal(float(t)) for t in times], golden[f'source{index}_waveform'])
    assert monitor.sample_interval == int(golden['sample_interval']) == 1
    np.testing.assert_allclose(monitor.flux(), golden['flux'], rtol=1e-12)


def test_lossy_calibration_depends_on_sigma():
    """The Ricker line sits in sigma_e = 5e3 S/m: its calibrated profile differs from the same line without the
    conduction term (|eta| of the lossy medium), so the golden profile pins the conductive branch of the calibration."""
    golden = np.load(os.path.join(GOLDEN, 'lossy2d_ez.npz'))
    spec = dict(LOSSY2D, rectangles=tuple((*r[:5], 0.0) for r in LOSSY2D['rectangles']))
    sim = Yee2D(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'],
                boundaries=_boundaries(spec['faces'], spec['npml']))
    for x0, y0, x1, y1, eps, sigma in spec['rectangles']:
        sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps, sigma_e=sigma)
    source = sim.add_line_source(**spec['sources'][0])
    sim.finalize()
    lossless = source.profile.astype(np.float64)
    assert np.max(np.abs(lossless / golden['source0_profile'] - 1.0)) > 1e-4


def test_golden_3d_lossy_setup():
    golden = np.load(os.path.join(GOLDEN, 'lossy3d.npz'))
    sim, dipoles, plane = build_lossy3d(Yee3D)
    sim.finalize()
    assert sim.shape == tuple(golden['shape'])
    assert float(sim.dt) == float(golden['dt'])
    np.testing.assert_array_equal(sim.materials['eps_r'], golden['eps_r'])
    np.testing.assert_array_equal(sim.materials['sigma_e'], golden['sigma_e'])
    for index, dipole in enumerate(dipoles):
        assert dipole.cell == tuple(golden[f'dipole{index}_cell'])
        # the reference rounds cb (with sigma_e in it) and moment / volume to float32 before the product
        assert abs(dipole.weight / float(golden[f'dipole{index}_weight'][0]) - 1.0) < 3e-7
    assert sim.sample_interval == int(golden['sample_interval']) == 1
    assert plane.edge == int(golden['plane_edge'])


def test_oracle_3d_lossy_matches_golden():
    """FP64 oracle vs the reference's float32 run with conductive boxes, a Ricker and a cw dipole: fields within
    float32 accumulation (5e-4 relative L2 after 250 steps), plane flux within 2e-5."""
    golden = np.load(os.path.join(GOLDEN, 'lossy3d.npz'))
    sim, _, plane = build_lossy3d(Yee3D)
    sim.run(n_steps=LOSSY3D['n_steps'])
    for name in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
        assert _rel(sim.fields[name], golden[name]) < 5e-4, name
    np.testing.assert_allclose(plane.flux(), golden['flux'], rtol=2e-5)


def _uniform_decay_2d(polarization, eps_r, sigma, steps=40):
    dx = 25e-9
    sim = Yee2D(8 * dx, 6 * dx, dx, npml=1, boundaries='periodic', polarization=polarization)
    sim.add_dielectric_rectangle(-1.0, -1.0, 1.0, 1.0, eps_r=eps_r, sigma_e=sigma)
    sim.finalize()
    names = ('ez',) if polarization == 'ez' else ('ex', 'ey')
    for name in names:
        sim.fields[name][...] = 1.0
    for _ in range(steps):
        sim.step()
    s = sigma * float(sim.dt) / (2.0 * EPS0 * eps_r)
    ca = (1.0 - s) / (1.0 + s)
    return sim, names, ca ** steps


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
@pytest.mark.parametrize('sigma', [2e4, 3e5])
def test_uniform_field_decays_by_ca(polarization, sigma):
    """No curl in a uniform field of a periodic box: E^{n+1} = ca E^n, so E after N steps is ca^N (1e-12), and H stays
    exactly 0. A conduction term dropped, doubled or applied explicitly fails it (ca differs at O(s^2))."""
    sim, names, want = _uniform_decay_2d(polarization, 2.5, sigma)
    assert want < 0.99
    for name in names:
        np.testing.assert_allclose(sim.fields[name], want, rtol=1e-12, atol=0)
    for name, values in sim.fields.items():
        if name not in names:
            assert not np.any(values), name


def test_uniform_field_decays_by_ca_3d():
    dx, eps_r, sigma, steps = 25e-9, 3.0, 5e4, 30
    faces = {face: 'periodic' for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    sim = Yee3D(6 * dx, 5 * dx, 4 * dx, dx, npml=1, boundaries=faces)
    sim.add_dielectric_box((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), eps_r, sigma_e=sigma)
    sim.finalize()
    for name in ('ex', 'ey', 'ez'):
        sim.fields[name][...] = 1.0
    for _ in range(steps):
        sim.step()
    s = sigma * float(sim.dt) / (2.0 * EPS0 * eps_r)
    want = ((1.0 - s) / (1.0 + s)) ** steps
    for name in ('ex', 'ey', 'ez'):
        np.testing.assert_allclose(sim.fields[name], want, rtol=1e-12, atol=0)
    for name in ('hx', 'hy', 'hz'):
        assert not np.any(sim.fields[name])


def test_ricker_closed_form():
    """The Ricker wavelet (1 - 2 a^2) exp(-a^2), a = pi f (t - t0), t0 = 4 / f: 1 at t0, zeros at t0 +- 1 / (pi f
    sqrt 2), minima -2 exp(-3/2) at t0 +- sqrt(3/2) / (pi f), and its amplitude spectrum (the second derivative of a
    Gaussian: f^2 exp(-f^2 / f_peak^2)) peaks at f_peak."""
    f = 5e14
    ricker = make_temporal('ricker', f, 0.5)
    t0 = 4.0 / f
    assert ricker(t0) == 1.0
    for sign in (-1.0, 1.0):
        assert abs(ricker(t0 + sign / (math.pi * f * math.sqrt(2.0)))) < 1e-15
        assert abs(ricker(t0 + sign * math.sqrt(1.5) / (math.pi * f)) + 2.0 * math.exp(-1.5)) < 1e-15
    dt = 1e-18
    times = np.arange(0, int(8.0 / f / dt)) * dt
    values = np.array([ricker(float(t)) for t in times])
    frequencies = np.linspace(0.2 * f, 2.0 * f, 361)
    spectrum = np.abs(np.exp(-2j * np.pi * frequencies[:, None] * times[None, :]) @ values) * dt
    assert abs(frequencies[np.argmax(spectrum)] / f - 1.0) < 0.006
    assert ricker.spectral_limit() == 4.2 * f


def test_cw_ramp_closed_form():
    """cw: (1 - exp(-t / tau)) sin(2 pi f t), tau = 3 / f: at t = tau (a whole number of periods plus a quarter, sin = 1)
    the value is 1 - 1/e; the ramp is 0 at t = 0; no spectral limit (DFT monitors sample every step)."""
    f = 4.3e14
    cw = make_temporal('cw', f, 0.5)
    quarter = 3.0 / f + 0.25 / f
    want = (1.0 - math.exp(-quarter * f / 3.0))
    assert abs(cw(quarter) - want) < 1e-12
    assert abs(cw(3.0 / f + 0.25 / f) - (1.0 - math.exp(-1.0 - 1.0 / 12.0))) < 1e-12
    assert cw(0.0) == 0.0
    assert cw.spectral_limit() is None
