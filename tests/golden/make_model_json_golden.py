"""Write tests/fixtures/model_thin_film_{ez,hz}.npz (beside the GPU tests that read them): what the reference's headless model API (open_sol.api.run, CPU
backend) returns and saves for its thin-film example model (examples/em/thin_film.json: a line source, two flux
monitors of 50 wavelengths, a film swept over n_film = 1 then 1.5), once as written (E_z) and once with the physics
type switched to em_fdtd_hz. They pin the model-JSON caller end to end: parameter evaluation, geometry and material
selection, feature placement, the study duration and sweep, and the results.npz layout.

Saved per case: the model JSON text (input), each swept run's flux wavelengths, power and power_per_cw per monitor
feature, the results.npz array names and its metadata JSON text (the file layout), and the summary's sweep record,
tick count and end time.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_model_json_golden.py
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


def model_text(polarization):
    with open(EXAMPLE, encoding='utf-8') as handle:
        spec = json.load(handle)
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
        with open(os.path.join(directory, 'out', 'summary.json'), encoding='utf-8') as handle:
            summary = json.load(handle)
        out['summary_sweep'] = np.array(json.dumps(summary.get('sweep')))
        out['summary_ticks'] = np.int64(summary['ticks'])
        out['summary_t_end'] = np.float64(summary['t_end'])
        out['summary_keys'] = np.array(sorted(summary))
        results.close()
    return out


def main():
    for polarization in ('ez', 'hz'):
        out = run_case(polarization)
        path = os.path.join(FIXTURES, f'model_thin_film_{polarization}.npz')
        np.savez_compressed(path, **out)
        print(path, 'T(600 nm) runs', [float(out[f'transmitted_flux/{i}/power'][25]) for i in range(2)])


if __name__ == '__main__':
    main()
