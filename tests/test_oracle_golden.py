"""The oracle against the golden vectors written by the reference's own numpy path (tests/golden/, precision
'double'): fields at fixed steps, the calibrated profile, clock, DFT interval, flux spectra and transmission."""
import os

import numpy as np
import pytest

from oracle.yee2d import Yee2D
from tests.cases import SNAPSHOT_STEPS, THIN_FILM, add_thin_film_probes, build

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


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
