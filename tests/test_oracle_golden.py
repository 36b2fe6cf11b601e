"""The oracle against the golden vectors written by the reference's own numpy path (tests/golden/, precision
'double'): fields at fixed steps, the calibrated profile, clock, DFT interval, flux spectra and transmission."""
import os

import numpy as np
import pytest

from oracle.yee2d import Yee2D
from tests.cases import SNAPSHOT_STEPS, THIN_FILM, add_thin_film_probes, build

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'fixtures')


def run_oracle(polarization, with_slab):
    sim, source, monitor = build(Yee2D, THIN_FILM, polarization, with_slab)
    probes = add_thin_film_probes(sim, polarization)
    snapshots = {}
    done = 0
    for step in SNAPSHOT_STEPS:
        sim.run(n_steps=step - done)
        done = step
        for name, values in sim.fields.items():
            snapshots[f'{name}_{step}'] = values.copy()
    sim.run(duration=THIN_FILM['duration'] - sim.time)
    return sim, source, monitor, snapshots, probes


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_oracle_reproduces_reference_thin_film(polarization):
    golden = np.load(os.path.join(GOLDEN, f'thin_film_{polarization}.npz'))
    flux = {}
    for tag, with_slab in (('film', True), ('empty', False)):
        sim, source, monitor, snapshots, probes = run_oracle(polarization, with_slab)
        assert (sim.nx, sim.ny) == (int(golden[f'{tag}_nx']), int(golden[f'{tag}_ny']))
        assert float(sim.dt) == float(golden[f'{tag}_dt'])
        assert sim.n_step == int(golden[f'{tag}_n_steps'])
        assert sim.time == float(golden[f'{tag}_time'])
        assert monitor.sample_interval == int(golden[f'{tag}_sample_interval'])
        assert source.position == int(golden[f'{tag}_source_position'])
        assert monitor.position == int(golden[f'{tag}_monitor_position'])
        assert tuple(monitor.span) == tuple(golden[f'{tag}_monitor_span'])
        np.testing.assert_array_equal(source.profile.astype(np.float64), golden[f'{tag}_source_profile'])
        for key, values in snapshots.items():
            np.testing.assert_allclose(values, golden[f'{tag}_{key}'], rtol=0, atol=1e-12 * np.abs(
                golden[f'{tag}_{key}']).max() + 1e-300)
        np.testing.assert_allclose(monitor.e_spectrum, golden[f'{tag}_e_spectrum'], rtol=1e-12, atol=0)
        flux[tag] = monitor.flux()
        np.testing.assert_allclose(flux[tag], golden[f'{tag}_flux'], rtol=1e-12)
        # field probes: one sample per step at the time after it; the sample after step s is the reference's
        # field at the probe's cell after step s
        for probe in probes:
            times, values = probe.get_data()
            assert len(times) == len(values) == sim.n_step
            assert times[-1] == sim.time
            np.testing.assert_allclose(np.diff(times), sim.dt, rtol=1e-9)
            for step in SNAPSHOT_STEPS:
                want = golden[f'{tag}_{probe.component}_{step}']
                assert abs(values[step - 1] - want[probe.cell]) <= 1e-12 * np.abs(want).max() + 1e-300
    np.testing.assert_allclose(flux['film'] / flux['empty'], golden['transmission'], rtol=1e-12)


def test_golden_transmission_is_physical():
    """The golden transmission of the 500 nm-thick n = 1.5 film sits near the Airy curve (examples/em/thin_film.py:
    20-26): T(500 nm) ~ 1 (round trip 6 pi), T(600 nm) ~ (1 - R)^2 / (1 + R)^2 with R = 0.04."""
    golden = np.load(os.path.join(GOLDEN, 'thin_film_ez.npz'))
    wavelengths = np.array([500e-9, 600e-9])
    n, d = 1.5, 500e-9
    delta = 2 * np.pi * n * d / wavelengths
    r = ((1 - n) / (1 + n)) ** 2
    analytic = (1 - r) ** 2 / (1 - 2 * r * np.cos(2 * delta) + r ** 2)
    assert np.max(np.abs(golden['transmission'] - analytic)) < 1.5e-2


def _oracle_boundaries(faces, npml):
    from oracle.yee2d import FACES, Boundaries
    return Boundaries(dict(zip(FACES, faces)), npml=npml)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_oracle_reproduces_reference_pec_case(polarization):
    """PEC region under a later dielectric, a source line crossing it and (E_z) a 'pec' x_lo face: the FP64 oracle
    reproduces the reference's FP64 run (fields at two steps, the source profile, the flux spectrum)."""
    from tests.cases import PEC2D, add_pec2d_dft_boxes, build_pec2d
    golden = np.load(os.path.join(GOLDEN, f'pec2d_{polarization}.npz'))
    sim, source, monitor = build_pec2d(Yee2D, polarization, _oracle_boundaries, phase_dtype=np.complex64)
    boxes = add_pec2d_dft_boxes(sim)
    assert (sim.nx, sim.ny) == (int(golden['nx']), int(golden['ny']))
    np.testing.assert_array_equal(sim.pec_regions[0], golden['pec_mask'])
    done = 0
    for step in PEC2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        if step == PEC2D['steps'][0]:
            np.testing.assert_array_equal(source.profile.astype(np.float64), golden['source_profile'])
        for name, values in sim.fields.items():
            want = golden[f'{name}_{step}']
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max() + 1e-300)
    assert monitor.sample_interval == int(golden['sample_interval'])
    np.testing.assert_allclose(monitor.flux(), golden['flux'], rtol=1e-12)
    # DFT boxes: sampled cells, and the complex64 spectra of the same complex64 arithmetic within 1e-6 of the peak
    for index, box in enumerate(boxes):
        samples = golden[f'dft{index}_samples']
        assert [[s.start, s.stop, s.step] for s in (box.x_samples, box.y_samples)] == samples.tolist()
        for name, values in box.spectra.items():
            want = golden[f'dft{index}_{name}']
            assert values.shape == want.shape
            error = np.max(np.abs(values - want)) / np.max(np.abs(want))
            assert error <= 1e-6, (index, name, error)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_oracle_reproduces_reference_point_currents(polarization):
    """Point currents (E_z filament currents; H_z magnetic line current and in-plane dipoles, with the periodic face
    wrap, the 1e-12 share cut and a PEC face) and a Gaussian beam: the FP64 oracle reproduces the reference's FP64 run
    (cells, faces, float32 weights and the beam profile exactly; fields at two steps to 1e-12 of the peak)."""
    from tests.cases import POINT2D, build_point2d
    golden = np.load(os.path.join(FIXTURES, f'point2d_{polarization}.npz'))
    sim, sources = build_point2d(Yee2D, polarization, _oracle_boundaries)
    assert (sim.nx, sim.ny) == (int(golden['nx']), int(golden['ny']))
    assert float(sim.dt) == float(golden['dt'])
    done = 0
    for step in POINT2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        for name, values in sim.fields.items():
            want = golden[f'{name}_{step}']
            assert np.abs(want).max() > 0
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())
    for index, source in enumerate(sources):
        faces = source if isinstance(source, list) else [source]
        want = golden[f'source{index}_weights']
        np.testing.assert_array_equal(np.concatenate([face.weight for face in faces]).astype(np.float64), want)
        if isinstance(source, list):
            cells = [[0 if face.component == 'ex' else 1, *face.cell] for face in faces]
            assert cells == golden[f'source{index}_faces'].tolist()
        else:
            assert list(source.cell) == golden[f'source{index}_cell'].tolist()
    weights = [w for k, w in golden.items() if k.endswith('_weights')]
    assert any(np.any(w == 0) for w in weights)   # a PEC cell or face injects nothing
    beam = sim.sources[-1]   # the Gaussian beam: span and calibrated float32 profile
    assert list(beam.span) == golden['beam_span'].tolist()
    np.testing.assert_array_equal(beam.profile.astype(np.float64), golden['beam_profile'])
    assert golden['beam_profile'].max() > 4 * golden['beam_profile'][[0, -1]].max()   # a bell over the span


def test_oracle_reproduces_reference_walls_and_mirrors():
    """E_z with a symmetric, a pmc and an antisymmetric face (ghost cells, even / odd copies, the corner, the re-pin
    after a source injects into a pinned row): the FP64 oracle reproduces the reference's FP64 run to 1e-12."""
    from tests.cases import WALLS2D, build_walls2d
    golden = np.load(os.path.join(FIXTURES, 'walls2d_ez.npz'))
    sim = build_walls2d(Yee2D, _oracle_boundaries)
    assert (sim.nx, sim.ny) == (int(golden['nx']), int(golden['ny']))
    done = 0
    for step in WALLS2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        for name, values in sim.fields.items():
            want = golden[f'{name}_{step}']
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())
    ez = golden['ez_400']
    assert np.array_equal(ez[0], ez[1]) and np.array_equal(ez[:, 0], -ez[:, 1]) and np.abs(ez[0]).max() > 0
    hz_case = build_walls2d(Yee2D, _oracle_boundaries, 'hz')   # H_z pmc walls, a magnetic current beside them
    done = 0
    for step in WALLS2D['steps']:
        hz_case.run(n_steps=step - done)
        done = step
        for name, values in hz_case.fields.items():
            want = golden[f'hzcase_{name}_{step}']
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_oracle_reproduces_reference_pole_case(polarization):
    """Drude rectangle (eps_inf background) under a PEC corner and a Lorentz disc on a dielectric: the FP64 oracle
    reproduces the reference's FP64 run (masks, eps_r, fields at two steps, flux spectrum); H_z: poles on the E faces."""
    from tests.cases import POLES2D, build_poles2d
    golden = np.load(os.path.join(GOLDEN, f'poles2d_{polarization}.npz'))
    sim, monitor = build_poles2d(Yee2D, polarization, phase_dtype=np.complex64)
    assert (sim.nx, sim.ny) == (int(golden['nx']), int(golden['ny']))
    np.testing.assert_array_equal(sim.poles[0]['mask'], golden['drude_mask'])
    np.testing.assert_array_equal(sim.poles[1]['mask'], golden['lorentz_mask'])
    np.testing.assert_array_equal(sim.pec_regions[0], golden['pec_mask'])
    eps_name = 'eps_r' if polarization == 'ez' else 'eps_xx'
    np.testing.assert_array_equal(sim.materials[eps_name].astype(np.float64), golden['eps_r'])
    done = 0
    for step in POLES2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        for name, values in sim.fields.items():
            want = golden[f'{name}_{step}']
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert monitor.sample_interval == int(golden['sample_interval'])
    np.testing.assert_allclose(monitor.flux(), golden['flux'], rtol=1e-12)


def test_pole_oracle_bulk_plasma_oscillation():
    """A periodic box filled with one Drude pole, uniform initial E_z (k = 0, no curl): E_z rings at the plasma
    frequency omega_p / sqrt(eps_inf) (1 % for omega dt ~ 0.2) and decays at gamma / 2 (5 %), the closed form of
    eps_inf eps0 dE/dt = -J, dJ/dt + gamma J = eps0 omega_p^2 E; a dropped drive or a sign error fails both."""
    dx = 25e-9
    sim = Yee2D(10 * dx, 8 * dx, dx, npml=1, boundaries='periodic')
    omega_p, gamma, eps_inf = 3e15, 4e13, 2.0
    sim.add_drude_rectangle(-1.0, -1.0, 1.0, 1.0, omega_p, gamma, eps_inf)
    sim.finalize()
    sim.fields['ez'][...] = 1.0
    series = []
    for _ in range(4000):
        sim.step()
        series.append(float(sim.fields['ez'][3, 4]))
    series = np.array(series)
    assert np.allclose(sim.fields['ez'], sim.fields['ez'][0, 0])   # stays uniform
    crossings = np.nonzero(np.diff(np.sign(series)) != 0)[0]
    period = 2.0 * float(sim.dt) * np.mean(np.diff(crossings))
    omega = 2.0 * np.pi / period
    assert abs(omega / (omega_p / np.sqrt(eps_inf)) - 1.0) < 1e-2
    peaks = [np.max(np.abs(series[a:b])) for a, b in zip(crossings[:-2:2], crossings[2::2])]
    t_span = (crossings[-3] - crossings[0]) * float(sim.dt)
    decay = -np.log(peaks[-1] / peaks[0]) / t_span
    assert abs(decay / (gamma / 2.0) - 1.0) < 5e-2


@pytest.mark.parametrize('case', ['ez', 'hz', 'ez_mirror'])
def test_oracle_reproduces_reference_tfsf(case):
    """TFSF plane-wave boxes around a dielectric (E_z, H_z, and E_z opening onto a symmetric y_lo mirror): the FP64
    oracle reproduces the reference's FP64 run (box cells, fields at two steps to 1e-12 of the peak, the scattered-
    and total-field flux to 1e-12). The scattered-field region holds only the scatterer's field (no incident leak)."""
    from tests.cases import TFSF2D, build_tfsf2d
    golden = np.load(os.path.join(GOLDEN, f'tfsf2d_{case}.npz'))
    sim, source, monitors = build_tfsf2d(Yee2D, case, _oracle_boundaries)
    assert list(source.cells) == golden['cells'].tolist()
    assert (sim.nx, sim.ny) == (int(golden['nx']), int(golden['ny']))
    done = 0
    for step in TFSF2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        for name, values in sim.fields.items():
            want = golden[f'{name}_{step}']
            assert np.abs(want).max() > 0
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert monitors[0].sample_interval == int(golden['sample_interval'])
    for index, monitor in enumerate(monitors):
        np.testing.assert_allclose(monitor.flux(), golden[f'flux{index}'], rtol=1e-12)


def test_tfsf_without_scatterer_leaks_nothing():
    """An empty TFSF box in vacuum: the scattered-field region stays below 1e-6 of the incident peak (the exact 1D /
    2D consistency of the incident line), while the total-field region carries the full wave; a wrong side
    coefficient, sign or timing leaks at the 1e-2 level."""
    for polarization, name in (('ez', 'ez'), ('hz', 'ey')):
        sim = Yee2D(2e-6, 1.4e-6, 25e-9, npml=10, polarization=polarization)
        source = sim.add_tfsf(0.4e-6, 0.2e-6, 1.6e-6, 1.2e-6, wavelength=600e-9, bandwidth=0.5)
        x0, x1, y0, y1 = source.cells
        sim.run(n_steps=150)
        field = sim.fields[name]
        inside = np.abs(field[x0 + 2:x1 - 2, y0 + 2:y1 - 2]).max()
        outside = max(np.abs(field[:x0 - 2]).max(), np.abs(field[x1 + 2:]).max(), np.abs(field[:, :y0 - 2]).max(),
                      np.abs(field[:, y1 + 2:]).max())
        assert inside > 0.5 and outside < 1e-6 * inside, (polarization, inside, outside)


@pytest.mark.parametrize('name', ['a', 'b'])
def test_oracle_reproduces_reference_hz_walls(name):
    """H_z pec and mirror faces (even copies on pec / antisymmetric faces with the tangential E pinned on their plane,
    odd copies on symmetric faces, ghost cells on mirrors, corners with both factors; 'b' with the y_hi wall on column
    128 of 129): the FP64 oracle reproduces the reference's FP64 run to 1e-12 of the peak; the golden rows are exact
    copies."""
    from tests.cases import HZWALLS2D, build_hzwalls2d
    golden = np.load(os.path.join(GOLDEN, 'hzwalls2d.npz'))
    sim = build_hzwalls2d(Yee2D, name, _oracle_boundaries)
    assert (sim.nx, sim.ny) == tuple(golden[f'{name}_shape'])
    done = 0
    for step in HZWALLS2D['steps']:
        sim.run(n_steps=step - done)
        done = step
        for field, values in sim.fields.items():
            want = golden[f'{name}_{field}_{step}']
            assert np.abs(want).max() > 0
            np.testing.assert_allclose(values, want, rtol=0, atol=1e-12 * np.abs(want).max())
    hz = golden[f'{name}_hz_400']
    faces = HZWALLS2D['cases'][name]['faces']
    if faces[3] == 'pec':
        assert np.array_equal(hz[:, -1], hz[:, -2]) and np.abs(hz[:, -1]).max() > 0
    if faces[1] == 'symmetric':
        assert np.array_equal(hz[-1], -hz[-2])
