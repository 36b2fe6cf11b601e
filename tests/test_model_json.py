"""Model-JSON front end (opensol_amd.model): parameter expressions, the build of the reference's thin-film example
model (examples/em/thin_film.json, held as text in the golden fixture) against the hand-built thin-film case, the
study's step count, and the results-directory writer. GPU runs of the same model are in test_gpu_model_json.py."""
import json
import os

import numpy as np
import pytest

from opensol_amd import model as M
from opensol_amd.fdtd2d import Model2D, steps_for_duration

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'fixtures')


def golden_spec(polarization):
    golden = np.load(os.path.join(FIXTURES, f'model_thin_film_{polarization}.npz'))
    return golden, json.loads(str(golden['model_json']))


def test_expressions_follow_python_arithmetic_and_the_builtin_units():
    names = M.evaluate_parameters([{'name': 'b', 'expression': 'a*2'}, {'name': 'a', 'expression': '600*nm'},
                                   {'name': 'c', 'expression': 'sqrt(b**2) / c_const + 0*pi'}])
    assert list(names) == ['b', 'a', 'c']
    assert names['a'] == 600 * 1e-9 and names['b'] == 600 * 1e-9 * 2
    assert names['c'] == (names['b'] ** 2) ** 0.5 / 299792458.0
    assert M.evaluate('4*Lx/c_const', {'Lx': 6e-6}) == 4 * 6e-6 / 299792458.0
    for bad in ('__import__("os")', 'a.b', '[1]', 'lambda: 1', 'x if 1 else 2', 'unknown*2', '1/0'):
        with pytest.raises((ValueError, ZeroDivisionError)):
            M.evaluate(bad, {'a': 1.0})
    with pytest.raises(ValueError):
        M.evaluate_parameters([{'name': 'a', 'expression': 'b'}, {'name': 'b', 'expression': 'a'}])
    with pytest.raises(ValueError):
        M.evaluate_parameters([{'name': 'pi', 'expression': 3}])
    with pytest.raises(KeyError):
        M.evaluate_parameters([{'name': 'a', 'expression': 1}], {'z': 2})


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_thin_film_model_builds_the_hand_built_case(polarization):
    """The example model's build equals the same case built through the scripting API: grid, dt, materials, the
    source line and profile, the monitor lines and wavelengths; the study runs the reference's step count."""
    golden, spec = golden_spec(polarization)
    names = M.evaluate_parameters(spec['parameters'])
    built = M._Built(spec, dict(names))
    model = built.sim.model
    ref = Model2D(6e-6, 3e-6, 25e-9, npml=15, boundaries='periodic_y', polarization=polarization)
    ref.add_dielectric_rectangle(3e-6 - 250e-9, 0.0, 3e-6 + 250e-9, 3e-6, eps_r=1.5 ** 2)
    ref.add_line_source('y', 1.2e-6, wavelength=600e-9, bandwidth=0.8, span=(0.0, 3e-6))
    wavelengths = np.linspace(400e-9, 800e-9, 50)
    for x in (1.8e-6, 4.5e-6):
        ref.add_flux_monitor('y', x, (0.0, 3e-6), wavelengths=wavelengths)
    model.finalize()
    ref.finalize()
    assert model.shape == ref.shape and model.dt == ref.dt
    np.testing.assert_array_equal(model.materials.ids, ref.materials.ids)
    np.testing.assert_array_equal(model.materials.values(), ref.materials.values())
    (source,), (ref_source,) = model.sources, ref.sources
    assert (source.axis, source.position, source.span) == (ref_source.axis, ref_source.position, ref_source.span)
    np.testing.assert_array_equal(source.profile, ref_source.profile)
    for monitor, ref_monitor in zip(model.monitors, ref.monitors):
        assert (monitor.normal, monitor.position, monitor.span) == (ref_monitor.normal, ref_monitor.position,
                                                                    ref_monitor.span)
        np.testing.assert_array_equal(monitor.frequencies, ref_monitor.frequencies)
        np.testing.assert_allclose(M.C0 / monitor.frequencies, golden['transmitted_flux/0/wavelength'], rtol=1e-15)
    assert model.sample_interval == ref.sample_interval
    duration = M.evaluate(spec['study']['settings']['duration'], names)
    # the reference's summary counts the ticks of both swept runs
    assert 2 * steps_for_duration(0.0, model.dt, duration) == int(golden['summary_ticks'])
    sweep = json.loads(str(golden['summary_sweep']))
    assert [M.evaluate(v, names) for v in spec['study']['sweep']['values']] == sweep['values']


def test_unsupported_models_are_refused():
    _, spec = golden_spec('ez')
    names = M.evaluate_parameters(spec['parameters'])
    for change in ({'type': 'em_fdfd_ez'},):
        other = json.loads(json.dumps(spec))
        other['physics'][0].update(change)
        with pytest.raises(ValueError):
            M._Built(other, names)
    other = json.loads(json.dumps(spec))
    other['physics'][0]['features'].append({'name': 'box', 'type': 'plane_wave_tfsf', 'selection': [],
                                            'params': {}, 'enabled': True})
    with pytest.raises(ValueError):
        M._Built(other, names)
    other = json.loads(json.dumps(spec))
    other['study']['type'] = 'frequency_domain'
    with pytest.raises(ValueError):
        M.run(other)


def test_results_directory_layout(tmp_path):
    """results.npz holds 'result/<key>/<array>' and 'sweep/<i>/<key>/<array>' arrays with a 'metadata' JSON entry of
    format 'open_sol.results' version 1 (io/results_io.py:399-433, 517-532); summary.json the per-result digests."""
    golden, spec = golden_spec('ez')
    result = {'kind': 'flux_spectrum', 'frequency': np.array([1.0, 2.0]), 'wavelength': np.array([2.0, 1.0]),
              'power': np.array([3.0, 4.0]), 'unit': 'a.u.', 'mirror_factor': 1.0}
    runs = [{'em.transmitted_flux': result, 'em.reflection_side_flux': result}] * 2
    fields = {'em.ez': (np.zeros((3, 2), np.float32), np.arange(4.0), np.arange(3.0))}
    summary = {'status': 'completed', 't_end': 1.0, 'results': {k: M.result_digest(v) for k, v in runs[-1].items()}}
    out = M.Results('completed', spec, runs[-1], runs, {'parameter': 'n_film', 'values': [1.0, 1.5]}, fields, summary)
    files = out.save(tmp_path)
    with np.load(tmp_path / files['arrays'], allow_pickle=False) as npz:
        names = set(npz.files)
        metadata = json.loads(str(npz['metadata']))
    reference_names = {str(n) for n in golden['npz_names'] if str(n).split('/')[0] in ('result', 'sweep')}
    ours = {n for n in names if n.split('/')[0] in ('result', 'sweep')}
    assert ours == {n for n in reference_names if not n.endswith('power_per_cw')}
    reference_metadata = json.loads(str(golden['npz_metadata']))
    assert metadata['format'] == reference_metadata['format'] and metadata['version'] == reference_metadata['version']
    assert metadata['results'] == reference_metadata['results']
    assert metadata['sweep_results'] == reference_metadata['sweep_results']
    assert set(metadata) == set(reference_metadata)
    saved = json.loads((tmp_path / 'summary.json').read_text())
    reference_digest = json.loads(str(golden['summary_sweep']))['results'][1]['em.transmitted_flux']
    assert set(saved['results']['em.transmitted_flux']) == set(reference_digest)
    assert json.loads((tmp_path / 'model.json').read_text()) == spec


def test_3d_model_builds_the_scripted_case():
    """The 3D model's build equals the same case through Simulation3D: materials (glass block, PEC box), the dipole,
    the two flux planes, the DFT layer and the probe; the study runs the reference's step count."""
    from opensol_amd.fdtd3d import Model3D
    golden = np.load(os.path.join(FIXTURES, 'model_dipole_3d.npz'))
    spec = json.loads(str(golden['model_json']))
    names = M.evaluate_parameters(spec['parameters'])
    model = M._Built(spec, names).sim.model
    ref = Model3D(0.72e-6, 0.6e-6, 0.55e-6, 25e-9, npml=8)
    ref.add_dielectric_box((0.05e-6, 0.05e-6, 0.0), (0.3e-6, 0.35e-6, 0.25e-6), 2.25)
    ref.add_pec_box((0.45e-6, 0.1e-6, 0.3e-6), (0.6e-6, 0.35e-6, 0.45e-6))
    ref.add_dipole(0.36e-6, 0.29e-6, 0.27e-6, orientation='z', dipole_moment=1e-12, wavelength=600e-9,
                   bandwidth=0.5)
    wavelengths = np.linspace(500e-9, 700e-9, 5)
    ref.add_flux_plane('x', 0.55e-6, wavelengths=wavelengths)
    ref.add_flux_plane('z', 0.1e-6, wavelengths=wavelengths)
    ref.add_flux_box((0.2e-6, 0.15e-6, 0.12e-6), (0.42e-6, 0.45e-6, 0.42e-6), wavelengths=wavelengths)
    ref.add_dft_plane(0.27e-6, wavelengths=np.linspace(500e-9, 700e-9, 3))
    ref.add_field_probe(0.2e-6, 0.45e-6, 0.3e-6, 'ez')
    model.finalize()
    ref.finalize()
    assert model.shape == ref.shape and model.dt == ref.dt
    np.testing.assert_array_equal(model.materials.ids, ref.materials.ids)
    np.testing.assert_array_equal(model.materials.values(), ref.materials.values())
    assert [(s.cell, s.component) for s in model.sources] == [(s.cell, s.component) for s in ref.sources]
    assert [(p.normal, p.edge) for p in model.monitors] == [(p.normal, p.edge) for p in ref.monitors]
    assert [d.k_grid for d in model.dft_monitors] == [d.k_grid for d in ref.dft_monitors]
    assert [p.cell for p in model.probes] == [p.cell for p in ref.probes]
    duration = M.evaluate(spec['study']['settings']['duration'], names)
    assert steps_for_duration(0.0, model.dt, duration) == int(golden['summary_ticks'])
    np.testing.assert_allclose(M.C0 / model.monitors[0].frequencies, golden['flux_x/wavelength'], rtol=1e-15)


def test_dispersive_material_becomes_a_pole_on_its_cells():
    """A material 'dispersion' list (definition_common.py:1009-1111): one DrudePole on the film's cells, the film's
    permittivity its eps_inf; H_z takes it too; unknown pole keys are refused."""
    from opensol_amd import DrudePole
    golden, spec = golden_spec('drude')
    names = M.evaluate_parameters(spec['parameters'])
    model = M._Built(spec, names).sim.model
    (pole, mask), = model.poles
    assert isinstance(pole, DrudePole) and float(pole.omega_p) == np.float32(2e15) and pole.eps_inf == np.float32(2.25)
    film = np.zeros(model.shape, dtype=bool)
    film[tuple(slice(lo, hi) for lo, hi in model.domain.box_ranges((2.75e-6, 0.0), (3.25e-6, 3e-6)))] = True
    np.testing.assert_array_equal(mask, film)
    np.testing.assert_array_equal(model.materials.property_array(0)[film], np.float32(2.25))
    other = json.loads(json.dumps(spec))
    other['physics'][0]['type'] = 'em_fdtd_hz'
    assert len(M._Built(other, names).sim.model.poles) == 1   # H_z: the pole on the film's E faces
    other = json.loads(json.dumps(spec))
    other['materials'][0]['dispersion'][0]['tau'] = 1.0
    with pytest.raises(ValueError):
        M._Built(other, names)


def test_point_source_feature_places_the_point_currents():
    """A 'point_source' feature on a point geometry (definition_common.py:1199-1217): E_z takes a filament current of
    `current`; H_z a magnetic line current (the default excitation) or an in-plane dipole along `orientation`."""
    golden, spec = golden_spec('ez')
    spec['geometry'].append({'name': 'emitter', 'kind': 'point', 'params': {'x': '1.5*um', 'y': '1.2*um'}})
    source = next(f for f in spec['physics'][0]['features'] if f['type'] == 'line_source')
    source.update(type='point_source', selection=['emitter'])
    source['params'] = {'waveform': 'gaussian', 'wavelength': '600*nm', 'bandwidth': 0.5, 'current': 2e-3}
    names = M.evaluate_parameters(spec['parameters'])
    model = M._Built(spec, names).sim.model
    (point,) = model.sources
    cell = model.domain.pos_to_cell(1.5e-6, 1.2e-6)
    area = float(model.grid.dx) ** 2
    assert (point.field, point.component, point.span[0], point.position) == ('e', 'ez', *cell)
    assert point.current_density == np.float32(2e-3 / area)
    spec['physics'][0]['type'] = 'em_fdtd_hz'
    (magnetic,) = M._Built(spec, names).sim.model.sources
    assert (magnetic.field, magnetic.component) == ('h', 'hz')
    assert magnetic.current_density == np.float32(1.0 / area)   # magnetic_current defaults to 1 V
    source['params'].update(excitation='electric_dipole', dipole_moment=1e-3, orientation=90.0)
    faces = M._Built(spec, names).sim.model.sources
    assert [f.component for f in faces] == ['ey', 'ey'] and all(f.field == 'e' for f in faces)


def test_gaussian_beam_feature_is_a_shaped_line_source():
    """A 'gaussian_beam' feature (definition_common.py:1227-1232): the line source of its line with the transverse
    factor exp(-((s - centre) / waist)^2), centre defaulting to the middle of the line."""
    golden, spec = golden_spec('ez')
    source = next(f for f in spec['physics'][0]['features'] if f['type'] == 'line_source')
    source['type'] = 'gaussian_beam'
    source['params'] = dict(source['params'], waist='0.8*um')
    names = M.evaluate_parameters(spec['parameters'])
    (beam,) = M._Built(spec, names).sim.model.sources
    line = beam.weight.astype(np.float64)
    centres = M._Built(spec, names).sim.model.domain.cell_centers()[1][beam.span[0]:beam.span[1]]
    expected = np.exp(-((centres - 1.5e-6) / 0.8e-6) ** 2)   # the line runs y = 0 .. Ly = 3 um: centre 1.5 um
    np.testing.assert_allclose(line / line.max(), expected / expected.max(), rtol=1e-6, atol=0)


def test_3d_dispersive_material_becomes_a_pole_on_its_box():
    """em_fdtd_3d: a material's Drude dispersion becomes a pole on exactly the cells its box painted, with eps_inf
    as their permittivity."""
    golden = np.load(os.path.join(FIXTURES, 'model_dipole_3d.npz'))
    spec = json.loads(str(golden['model_json']))
    spec['materials'][0]['properties'] = {}
    spec['materials'][0]['dispersion'] = [{'model': 'drude', 'omega_p': '2e15', 'gamma': '1e14', 'eps_inf': 2.25}]
    model = M._Built(spec, M.evaluate_parameters(spec['parameters'])).sim.model
    (pole, mask), = model.poles
    painted = model.materials.property_array(0) == np.float32(2.25)
    np.testing.assert_array_equal(mask, painted)
    assert mask.sum() > 0 and float(pole.omega_p) == np.float32(2e15)
