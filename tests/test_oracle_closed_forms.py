"""Closed-form and invariant checks that pin the oracle independently of the reference's golden vectors:
the discrete Yee dispersion of a periodic-box eigenmode (bulk update and coefficients), the calibrated sheet amplitude
(physics/em/selftests_hz.py plane-wave case), CPML reflection against a longer domain, energy balance R + T = 1 of a
lossless film, and the DFT sample interval."""
import numpy as np
import pytest

from oracle.yee2d import C0, EPS0, MU0, Yee2D, dft_sample_interval, GaussianPulse


def _mode(sim, p, q):
    geo = sim.geo
    x, y = geo.centres
    kx = 2 * np.pi * p / (geo.nx * geo.dx)
    ky = 2 * np.pi * q / (geo.ny * geo.dx)
    return kx, ky, np.cos(kx * x)[:, None] * np.cos(ky * y)[None, :]


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
@pytest.mark.parametrize('p,q', [(1, 0), (2, 3), (5, 1)])
def test_periodic_eigenmode_follows_discrete_dispersion(polarization, p, q):
    """F^{n+1} + F^{n-1} = 2 cos(w dt) F^n with cos(w dt) = 1 - 2 dt^2/(eps0 mu0) sum sin^2(k d/2)/d^2 exactly for a
    single eigenmode of the periodic Yee grid (a sign or spacing error changes the recurrence)."""
    sim = Yee2D(24 * 25e-9, 16 * 25e-9, 25e-9, boundaries='periodic', polarization=polarization)
    _, _, mode = _mode(sim, p, q)
    sim.fields[polarization][...] = mode
    d = float(sim.geo.dx32)   # kernel cell width
    dt = float(sim.dt)
    theta_x, theta_y = 2 * np.pi * p / sim.nx, 2 * np.pi * q / sim.ny   # phase advance per cell
    cos_wdt = 1.0 - 2.0 * dt ** 2 / (EPS0 * MU0) * (np.sin(theta_x / 2) ** 2 + np.sin(theta_y / 2) ** 2) / d ** 2
    series = []
    for _ in range(60):
        sim.step()
        series.append(sim.fields[polarization].copy())
    series = np.array(series)
    lhs = series[2:] + series[:-2]
    rhs = 2.0 * cos_wdt * series[1:-1]
    assert np.max(np.abs(lhs - rhs)) < 1e-11 * np.max(np.abs(series))
    # and the field stays in the mode (no leakage into other shapes)
    last = series[-1]
    coefficient = np.sum(last * mode) / np.sum(mode * mode)
    assert np.max(np.abs(last - coefficient * mode)) < 1e-11


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
@pytest.mark.parametrize('eps_r', [1.0, 4.0])
def test_cw_sheet_radiates_its_amplitude(polarization, eps_r):
    """A calibrated cw sheet in a periodic strip launches amplitude A per side in the local medium (+-2 %,
    physics/em/selftests_hz.py:617-639)."""
    dx = 25e-9
    sim = Yee2D(4e-6, 4 * dx, dx, npml=15, boundaries='periodic_y', polarization=polarization)
    if eps_r != 1.0:
        sim.add_dielectric_rectangle(-1e-6, -1e-6, 5e-6, 1e-6, eps_r=eps_r)
    sim.add_line_source(axis='y', position=1.0e-6, wavelength=800e-9, waveform='cw', amplitude=1.0)
    component = 'ez' if polarization == 'ez' else 'ey'
    period = 800e-9 / C0
    sim.run(duration=40 * period)
    probe = sim.geo.cell(0, 3.0e-6)
    peak = 0.0
    for _ in range(int(np.ceil(period / float(sim.dt))) + 2):
        sim.step()
        peak = max(peak, float(np.max(np.abs(sim.fields[component][probe]))))
    assert abs(peak - 1.0) < 0.02


def _pulse_probe(window_x, steps, polarization):
    dx = 25e-9
    sim = Yee2D(window_x, 4 * dx, dx, npml=15, boundaries='periodic_y', polarization=polarization)
    sim.add_line_source(axis='y', position=1.0e-6, wavelength=600e-9, bandwidth=0.8)
    probe = sim.geo.cell(0, 3.0e-6)
    component = 'ez' if polarization == 'ez' else 'ey'
    trace = []
    for _ in range(steps):
        sim.step()
        trace.append(float(sim.fields[component][probe, 0]))
    return np.array(trace)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_cpml_reflection_is_small(polarization):
    """Probe 1 um before the hi CPML of a 4 um window vs the same probe in a 16 um window (whose far wall is never
    reached in the run): the difference is the CPML reflection."""
    steps = 900
    near = _pulse_probe(4e-6, steps, polarization)
    far = _pulse_probe(16e-6, steps, polarization)
    reflection = np.max(np.abs(near - far)) / np.max(np.abs(far))
    assert reflection < 2e-3


def test_lossless_film_conserves_energy():
    """R + T = 1 for a lossless film (examples/em/thin_film.py:88-96), with R from a monitor behind the source."""
    def run(with_film):
        sim = Yee2D(6e-6, 3e-6, 25e-9, npml=15, boundaries='periodic_y')
        if with_film:
            sim.add_dielectric_rectangle(2.75e-6, 0, 3.25e-6, 3e-6, eps_r=2.25)
        sim.add_line_source(axis='y', position=1.2e-6, wavelength=600e-9, bandwidth=0.8)
        trans = sim.add_flux_monitor(axis='y', position=4.5e-6, span=(0, 3e-6), wavelengths=[500e-9, 550e-9, 600e-9])
        refl = sim.add_flux_monitor(axis='y', position=1.8e-6, span=(0, 3e-6), wavelengths=[500e-9, 550e-9, 600e-9])
        sim.run(duration=4 * 6e-6 / C0)
        return trans.flux(), refl.flux()
    t_film, r_film = run(True)
    t_empty, r_empty = run(False)
    transmission = t_film / t_empty
    reflection = 1.0 - r_film / r_empty
    assert np.max(np.abs(transmission + reflection - 1.0)) < 2e-3
    assert np.all(transmission > 0.8) and np.all(transmission <= 1.0 + 1e-3)


def test_dft_sample_interval_closed_form():
    """m = floor(1 / ((f0 (1 + 3.72 bw) + f_max) dt)): 6 for the thin-film case."""
    sim = Yee2D(6e-6, 3e-6, 25e-9, npml=15, boundaries='periodic_y')
    src = sim.add_line_source(axis='y', position=1.2e-6, wavelength=600e-9, bandwidth=0.8)
    f0 = C0 / 600e-9
    assert src.temporal.spectral_limit() == pytest.approx(f0 * (1 + 3.72 * 0.8), rel=1e-15)
    freqs = [np.array([C0 / 500e-9, C0 / 600e-9])]
    expected = int(np.floor(1.0 / ((f0 * (1 + 3.72 * 0.8) + C0 / 500e-9) * float(sim.dt))))
    assert dft_sample_interval(sim.dt, sim.sources, freqs) == expected == 6


def test_gaussian_pulse_closed_form():
    pulse = GaussianPulse(5e14, 0.5)
    tau = 1.0 / (0.5 * 5e14 * np.pi)
    assert pulse.tau == pytest.approx(tau, rel=1e-15)
    assert pulse.t0 == pytest.approx(4.5 * tau, rel=1e-15)
    assert pulse(pulse.t0) == 0.0
    t = pulse.t0 + 0.25 / 5e14
    assert pulse(t) == pytest.approx(np.exp(-(0.25 / 5e14 / tau) ** 2), rel=1e-14)
