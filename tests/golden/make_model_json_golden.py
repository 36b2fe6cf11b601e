"""Write tests/fixtures/model_thin_film_{ez,hz,drude}.npz and model_dipole_3d.npz (beside the GPU tests that read them): what the reference's headless model API (open_sol.api.run, CPU
backend) returns and saves for its thin-film example model (examples/em/thin_film.json: a line source, two flux
monitors of 50 wavelengths, a film swept over n_film = 1 then 1.5), once as written (E_z), once with the physics
type switched to em_fdtd_hz and once with the film a Drude medium (E_z), and for a small 3D model written here (MODEL_3D). They pin the model-JSON caller end to end: parameter evaluation, geometry and material
selection, feature placement, the study duration and sweep, and the results.npz layout.

Saved per case: the model JSON text (input), each swept run's flux wavelengths, power and power_per_cw per monitor
feature, the results.npz array names and its metadata JSON text (the file layout), the summary's sweep record, tick
count and end time, and a sample of the results.npz values: the stored fields of one snapshot and of the final
field group (3D: their mid-height z cuts), their cell centres and edges, and (E_z, H_z) the empty run's primary field
at that snapshot.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_model_json_golden.py [3d] [ez] [hz] [drude] [big]   (no argument: every case)
The fixture holds inputs and outputs only, no source text of the reference.
"""
import json
import os
import sys
import tempfile

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(os.path.dirname(HERE), 'fixtures')
EXAMPLE = os.path.join(REFERENCE, 'examples', 'em', 'thin_film.json')
MONITORS = ('reflection_side_flux', 'transmitted_flux')
# results.npz values kept (besides the names and metadata): the stored fields of snapshot KEPT_SNAPSHOT and of the
# final field group, with their centres and edges, and the first swept run's primary field at that snapshot
KEPT_SNAPSHOT = 20
KEPT_FIELDS = {'ez': ('ez', 'hx', 'hy', 'e_abs', 'e_intensity', 'h_abs', 'eps_r', 'sigma_e'),
               'hz': ('hz', 'ex', 'ey', 'e_abs', 'e_intensity', 'h_abs', 'eps_xx', 'eps_yy', 'sigma_e')}
# 3D: the stored fields' mid-height z cut in snapshot KEPT_SNAPSHOT_3D and in the final field group, and that
# snapshot's two stored cuts (MODEL_3D's study stores an x and a y plane besides the z cut)
KEPT_SNAPSHOT_3D = 2
KEPT_FIELDS_3D = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz', 'e_abs', 'e_intensity', 'h_abs', 'eps_r', 'absorbed_power')


# A small 3D model of the em_fdtd_3d vocabulary (no example of the reference uses only features of this path): a
# z dipole over a glass block, a PEC box, flux planes normal to x and z, a flux box around the dipole, a DFT layer
# and an E_z probe; 8-cell CPML.
MODEL_3D = {
    'title': 'Dipole over a glass block (3D)', 'version': 1,
    'parameters': [
        {'name': 'L', 'expression': '0.6*um', 'description': 'domain width'},
        {'name': 'H', 'expression': '0.55*um', 'description': 'domain height'},
        {'name': 'lambda0', 'expression': '600*nm', 'description': 'pulse centre wavelength'},
        {'name': 'n_glass', 'expression': 1.5, 'description': 'block index'},
    ],
    'domain': {'size_x': '1.2*L', 'size_y': 'L', 'size_z': 'H'},
    'geometry': [
        {'name': 'dipole_point', 'kind': 'point', 'params': {'x': '0.36*um', 'y': '0.29*um', 'z': '0.27*um'}},
        {'name': 'probe_point', 'kind': 'point', 'params': {'x': '0.2*um', 'y': '0.45*um', 'z': '0.3*um'}},
        {'name': 'block', 'kind': 'box', 'params': {'x0': '0.05*um', 'y0': '0.05*um', 'z0': 0.0, 'x1': '0.3*um',
                                                    'y1': '0.35*um', 'z1': '0.25*um'}},
        {'name': 'metal', 'kind': 'box', 'params': {'x0': '0.45*um', 'y0': '0.1*um', 'z0': '0.3*um', 'x1': '0.6*um',
                                                    'y1': '0.35*um', 'z1': '0.45*um'}},
        {'name': 'surface', 'kind': 'box', 'params': {'x0': '0.2*um', 'y0': '0.15*um', 'z0': '0.12*um',
                                                      'x1': '0.42*um', 'y1': '0.45*um', 'z1': '0.42*um'}},
    ],
    'selections': [],
    'materials': [{'name': 'glass', 'library': None, 'properties': {'eps_r': 'n_glass**2'}, 'dispersion': None,
                   'selection': ['block']}],
    'physics': [{
        'name': 'em', 'type': 'em_fdtd_3d', 'settings': {'pml_cells': 8},
        'features': [
            {'name': 'dipole', 'type': 'dipole_source', 'selection': ['dipole_point'],
             'params': {'dipole_axis': 'z', 'waveform': 'gaussian', 'wavelength': 'lambda0', 'bandwidth': 0.5,
                        'current_moment': 1e-12}, 'enabled': True},
            {'name': 'conductor', 'type': 'pec_domain', 'selection': ['metal'], 'params': {}, 'enabled': True},
            {'name': 'flux_x', 'type': 'flux_plane', 'selection': [],
             'params': {'normal': 'x', 'position': '0.55*um', 'wavelength_start': 5e-07, 'wavelength_stop': 7e-07,
                        'n_wavelengths': 5}, 'enabled': True},
            {'name': 'flux_z', 'type': 'flux_plane', 'selection': [],
             'params': {'normal': 'z', 'position': '0.1*um', 'wavelength_start': 5e-07, 'wavelength_stop': 7e-07,
                        'n_wavelengths': 5}, 'enabled': True},
            {'name': 'box', 'type': 'flux_box', 'selection': ['surface'],
             'params': {'wavelength_start': 5e-07, 'wavelength_stop': 7e-07, 'n_wavelengths': 5}, 'enabled': True},
            {'name': 'layer', 'type': 'dft_plane', 'selection': [],
             'params': {'position': '0.27*um', 'wavelength_start': 5e-07, 'wavelength_stop': 7e-07,
                        'n_wavelengths': 3}, 'enabled': True},
            {'name': 'probe', 'type': 'field_probe', 'selection': ['probe_point'], 'params': {'component': 'ez'},
             'enabled': True},
        ], 'every': 1}],
    'couplings': [],
    'mesh': {'mode': 'uniform', 'dx': '25*nm', 'max_ratio': 1.15, 'wavelength_min': None,
             'points_per_wavelength': 20.0, 'refinements': []},
    'study': {'type': 'time_dependent', 'settings': {'duration': '15*fs', 'output_interval': '5*fs'},
              'stored_fields': [], 'stored_cuts': [['x', '0.31*um'], ['y', 'L/2']], 'sweep': None},
    'derived_values': [],
}


def run_3d():
    """The 3D model through the reference's api.run: flux, per-cw power, DFT layer intensity, the probe series and the
    results.npz layout with a sample of its snapshot and field values (mid-height z cuts)."""
    sys.path.insert(0, REFERENCE)
    from open_sol import api

    text = json.dumps(MODEL_3D)
    out = {'model_json': np.array(text)}
    with tempfile.TemporaryDirectory() as directory:
        path = os.path.join(directory, 'model.json')
        with open(path, 'w', encoding='utf-8') as handle:
            handle.write(text)
        results = api.run(api.load(path), out_dir=os.path.join(directory, 'out'), backend='cpu')
        if results.status != 'completed':
            raise RuntimeError(f'reference run ended {results.status}')
        with np.load(os.path.join(directory, 'out', 'results.npz'), allow_pickle=False) as npz:
            out['npz_names'] = np.array(sorted(npz.files))
            out['npz_metadata'] = np.array(str(npz['metadata']))
            for name in KEPT_FIELDS_3D:   # the mid-height cut of snapshot KEPT_SNAPSHOT_3D and of the final fields
                for group in (f'snapshot/{KEPT_SNAPSHOT_3D}', 'field'):
                    out[f'npz/{group}/em.{name}/values'] = npz[f'{group}/em.{name}/values']
            out['npz/result/em.layer/intensity_per_cw_factor'] = npz['result/em.layer/intensity_per_cw_factor']
            for cut in (0, 1):   # the stored cuts (x, then y) of that snapshot
                for name in KEPT_FIELDS_3D:
                    key = f'snapshot/{KEPT_SNAPSHOT_3D}/cut{cut}/em.{name}/values'
                    out[f'npz/{key}'] = npz[key]
                for axis in 'xy':
                    key = f'snapshot/{KEPT_SNAPSHOT_3D}/cut{cut}/em.ez/{axis}_centers'
                    out[f'npz/{key}'] = npz[key]
            for axis in 'xy':
                out[f'npz/snapshot/{KEPT_SNAPSHOT_3D}/{axis}_centers'] = npz[
                    f'snapshot/{KEPT_SNAPSHOT_3D}/em.ez/{axis}_centers']
                out[f'npz/field/{axis}_edges'] = npz[f'field/em.ez/{axis}_edges']
        with open(os.path.join(directory, 'out', 'summary.json'), encoding='utf-8') as handle:
            summary = json.load(handle)
        out['summary_snapshot_times'] = np.array(summary['snapshot_times'], dtype=np.float64)
        out['summary_stored_cuts'] = np.array(json.dumps(summary['stored_cuts']))
        out['summary_t_end'] = np.float64(summary['t_end'])
        for name in ('flux_x', 'flux_z', 'box'):
            result = results.result('em', name)
            for key in ('wavelength', 'power', 'power_per_cw'):
                out[f'{name}/{key}'] = np.asarray(result[key], dtype=np.float64)
        for index in range(3):
            layer = results.result('em', 'layer', freq_index=index)
            out[f'layer/intensity/{index}'] = np.asarray(layer['intensity'], dtype=np.float64)
        out['layer/x_centers'] = np.asarray(layer['x_centers'], dtype=np.float64)
        out['layer/z'] = np.float64(layer['z'])
        probe = results.result('em', 'probe')
        out['probe/t'] = np.asarray(probe['t'], dtype=np.float64)
        out['probe/values'] = np.asarray(probe['values'], dtype=np.float64)
        out['summary_ticks'] = np.int64(results.summary()['ticks'])
        results.close()
    return out


def model_text(polarization):
    """The example as written (E_z), its H_z variant, or 'drude': E_z with the film a Drude medium (eps_inf 2.25,
    omega_p 2e15 rad/s, gamma 1e14 1/s) in place of its eps_r."""
    with open(EXAMPLE, encoding='utf-8') as handle:
        spec = json.load(handle)
    if polarization == 'drude':
        film = spec['materials'][0]
        film['properties'] = {}
        film['dispersion'] = [{'model': 'drude', 'omega_p': '2e15', 'gamma': '1e14', 'eps_inf': 2.25}]
        return json.dumps(spec)
    spec['physics'][0]['type'] = f'em_fdtd_{polarization}'
    return json.dumps(spec)


def run_case(polarization):
    sys.path.insert(0, REFERENCE)
    from open_sol import api

    text = model_text(polarization)
    out = {'model_json': np.array(text)}
    with tempfile.TemporaryDirectory() as directory:
        path = os.path.join(directory, 'model.json')
        with open(path, 'w', encoding='utf-8') as handle:
            handle.write(text)
        results = api.run(api.load(path), out_dir=os.path.join(directory, 'out'), backend='cpu')
        if results.status != 'completed':
            raise RuntimeError(f'reference run ended {results.status}')
        for name in MONITORS:
            for index, result in enumerate(results.sweep_result('em', name)):
                out[f'{name}/{index}/wavelength'] = np.asarray(result['wavelength'], dtype=np.float64)
                out[f'{name}/{index}/power'] = np.asarray(result['power'], dtype=np.float64)
                out[f'{name}/{index}/power_per_cw'] = np.asarray(result['power_per_cw'], dtype=np.float64)
        with np.load(os.path.join(directory, 'out', 'results.npz'), allow_pickle=False) as npz:
            out['npz_names'] = np.array(sorted(npz.files))
            out['npz_metadata'] = np.array(str(npz['metadata']))
            if polarization in KEPT_FIELDS:   # a sample of the field and snapshot values
                primary = KEPT_FIELDS[polarization][0]
                for name in KEPT_FIELDS[polarization]:
                    out[f'npz/snapshot/{KEPT_SNAPSHOT}/em.{name}/values'] = npz[
                        f'snapshot/{KEPT_SNAPSHOT}/em.{name}/values']
                    out[f'npz/field/em.{name}/values'] = npz[f'field/em.{name}/values']
                for axis in 'xy':
                    out[f'npz/snapshot/{KEPT_SNAPSHOT}/{axis}_centers'] = npz[
                        f'snapshot/{KEPT_SNAPSHOT}/em.{primary}/{axis}_centers']
                    out[f'npz/field/{axis}_edges'] = npz[f'field/em.{primary}/{axis}_edges']
                out[f'npz/sweep_snapshot/0/{KEPT_SNAPSHOT}/em.{primary}/values'] = npz[
                    f'sweep_snapshot/0/{KEPT_SNAPSHOT}/em.{primary}/values']
        with open(os.path.join(directory, 'out', 'summary.json'), encoding='utf-8') as handle:
            summary = json.load(handle)
        out['summary_sweep'] = np.array(json.dumps(summary.get('sweep')))
        out['summary_ticks'] = np.int64(summary['ticks'])
        out['summary_t_end'] = np.float64(summary['t_end'])
        out['summary_keys'] = np.array(sorted(summary))
        results.close()
    return out


def big_model_text():
    """The thin-film example widened to a 553 x 459 cell window (253 827 cells, past the reference's 250 000-cell
    snapshot limit, so its snapshots hold nearest-cell resampled fields), the film over the middle 40 % of y, a 3 fs
    run with 2 snapshots and no sweep."""
    with open(EXAMPLE, encoding='utf-8') as handle:
        spec = json.load(handle)
    for parameter in spec['parameters']:
        if parameter['name'] == 'Lx':
            parameter['expression'] = '13.825*um'
        if parameter['name'] == 'Ly':
            parameter['expression'] = '11.475*um'
    for item in spec['geometry']:
        if item['name'] == 'slab':
            item['params']['y0'], item['params']['y1'] = '0.3*Ly', '0.7*Ly'
    spec['study'] = {'type': 'time_dependent', 'settings': {'duration': '3*fs', 'n_snapshots': 2},
                     'stored_fields': [], 'sweep': None}
    return json.dumps(spec)


def run_big():
    """The widened example through api.run: results.npz names and metadata, and the resampled snapshot fields E_z
    and eps_r of the final snapshot."""
    sys.path.insert(0, REFERENCE)
    from open_sol import api

    text = big_model_text()
    out = {'model_json': np.array(text)}
    with tempfile.TemporaryDirectory() as directory:
        path = os.path.join(directory, 'model.json')
        with open(path, 'w', encoding='utf-8') as handle:
            handle.write(text)
        results = api.run(api.load(path), out_dir=os.path.join(directory, 'out'), backend='cpu')
        if results.status != 'completed':
            raise RuntimeError(f'reference run ended {results.status}')
        with np.load(os.path.join(directory, 'out', 'results.npz'), allow_pickle=False) as npz:
            out['npz_names'] = np.array(sorted(npz.files))
            out['npz_metadata'] = np.array(str(npz['metadata']))
            for name in ('ez', 'eps_r'):
                out[f'npz/snapshot/1/em.{name}/values'] = npz[f'snapshot/1/em.{name}/values']
        results.close()
    return out


def main(cases=('3d', 'ez', 'hz', 'drude', 'big')):
    if '3d' in cases:
        out = run_3d()
        path = os.path.join(FIXTURES, 'model_dipole_3d.npz')
        np.savez_compressed(path, **out)
        print(path, 'ticks', int(out['summary_ticks']), 'flux_x', out['flux_x/power'])
    if 'big' in cases:
        out = run_big()
        path = os.path.join(FIXTURES, 'model_big_window_ez.npz')
        np.savez_compressed(path, **out)
        print(path, 'snapshot shape', out['npz/snapshot/1/em.ez/values'].shape)
    for polarization in [case for case in ('ez', 'hz', 'drude') if case in cases]:
        out = run_case(polarization)
        path = os.path.join(FIXTURES, f'model_thin_film_{polarization}.npz')
        np.savez_compressed(path, **out)
        print(path, 'T(600 nm) runs', [float(out[f'transmitted_flux/{i}/power'][25]) for i in range(2)])


if __name__ == '__main__':
    main(tuple(sys.argv[1:]) or ('3d', 'ez', 'hz', 'drude', 'big'))
