"""Snapshot side of the model-JSON front end, on the CPU: the output schedule, the stored-field table, the face-to-centre
averaging and the results.npz snapshot / field layout, against what the reference's own run of its thin-film example
saved (tests/fixtures/model_thin_film_{ez,hz}.npz, written by tests/golden/make_model_json_golden.py).

  * snapshot_schedule() reproduces the reference run's 42 snapshot times bit for bit (output_interval 2 fs: ticks of
    58.4 as, so the times are 0, 1.985, 4.028, ... 79.976 fs and the final 80.034 fs) and its tick count, and with
    output_interval 0 spaces n_snapshots over the duration (first at 0, last the final state, one per grid point);
  * STORED_FIELDS equals the FieldInfo records, in order, that the reference stored for E_z and H_z;
  * _to_centres() is checked against a per-cell loop (periodic and not), so a swapped neighbour or a dropped wrap
    fails;
  * Results.save() of hand-made snapshots writes the reference's array names and metadata entries for a 2D run."""
import json
import math
import os

import numpy as np
import pytest

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'fixtures')


def _golden(polarization):
    return np.load(os.path.join(FIXTURES, f'model_thin_film_{polarization}.npz'))


def _dt(spec):
    from opensol_amd import model as M
    names = M.evaluate_parameters(spec['parameters'])
    built = M._Built(spec, names)
    return built, names


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_schedule_matches_the_reference_snapshot_times(polarization):
    from opensol_amd import model as M
    from opensol_amd.fdtd2d import steps_for_duration
    golden = _golden(polarization)
    spec = json.loads(str(golden['model_json']))
    built, names = _dt(spec)
    settings = spec['study']['settings']
    duration = M.evaluate(settings['duration'], names)
    schedule = M.snapshot_schedule(built.sim.model.dt, duration, M.evaluate(settings['output_interval'], names))
    want = json.loads(str(golden['summary_sweep']))['snapshot_times']
    assert [t for _, t in schedule] == want[0] == want[1]
    assert schedule[-1][0] == steps_for_duration(0.0, built.sim.model.dt, duration)
    assert 2 * schedule[-1][0] == int(golden['summary_ticks'])   # the summary counts both swept runs' ticks
    metadata = json.loads(str(golden['npz_metadata']))
    assert [s['time'] for s in metadata['snapshots']] == want[1]


def test_even_schedule_closed_form():
    """output_interval 0: n_snapshots grid points k * duration / (n - 1), each taken at the first tick within half a
    tick of it, the last one the final state."""
    from opensol_amd import model as M
    dt, duration, n = 1.0e-3, 1.0, 11
    schedule = M.snapshot_schedule(dt, duration, 0.0, n)
    assert len(schedule) == n
    ticks = [tick for tick, _ in schedule]
    assert ticks == [round(k * 100) for k in range(n)]
    assert schedule[-1][0] == 1000
    # one snapshot: only the final state
    assert M.snapshot_schedule(dt, duration, 0.0, 1) == [(1000, schedule[-1][1])]
    # an interval longer than the run: the start and the final state
    assert [tick for tick, _ in M.snapshot_schedule(dt, duration, 5.0)] == [0, 1000]
    # intervals shorter than a tick record every tick
    assert [tick for tick, _ in M.snapshot_schedule(dt, 10 * dt, dt / 4)] == list(range(11))


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_stored_field_table_matches_the_reference(polarization):
    from opensol_amd import model as M
    metadata = json.loads(str(_golden(polarization)['npz_metadata']))
    stored = metadata['snapshots'][0]['fields']
    assert [key.split('.', 1)[1] for key in stored] == [info['name'] for info in M.STORED_FIELDS[polarization]]
    for info in M.STORED_FIELDS[polarization]:
        assert stored[f'em.{info["name"]}']['info'] == info
        assert metadata['fields'][f'em.{info["name"]}']['info'] == info


@pytest.mark.parametrize('periodic', [False, True])
@pytest.mark.parametrize('axis', [0, 1])
def test_face_to_centre_average_against_a_loop(axis, periodic):
    from opensol_amd import model as M
    rng = np.random.default_rng(7 + axis + 2 * periodic)
    face = rng.standard_normal((7, 5)).astype(np.float32)
    got = M._to_centres(face, axis, periodic)
    want = np.empty_like(face)
    for i in range(face.shape[0]):
        for j in range(face.shape[1]):
            index = (i, j)[axis]
            lower = list((i, j))
            lower[axis] = index - 1 if index > 0 else (face.shape[axis] - 1 if periodic else 0)
            want[i, j] = np.float32(np.float32(face[i, j] + face[tuple(lower)]) * np.float32(0.5))
    assert got.dtype == np.float32
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_saved_layout_matches_the_reference_names_and_entries(polarization, tmp_path):
    """The recorder's window layout (built on the CPU, no engine) and Results.save() of snapshots at the reference
    run's times: the window edges and centres, the field / snapshot / sweep_snapshot array names and those groups'
    metadata entries equal the reference's (the values are hand-made here; the GPU test compares a real run's)."""
    from opensol_amd import model as M
    golden = _golden(polarization)
    metadata = json.loads(str(golden['npz_metadata']))
    spec = json.loads(str(golden['model_json']))
    built, _ = _dt(spec)
    recorder = M._Recorder(built)
    np.testing.assert_array_equal(recorder.edges[0], golden['npz/field/x_edges'])
    np.testing.assert_array_equal(recorder.edges[1], golden['npz/field/y_edges'])
    np.testing.assert_array_equal(recorder.centres[0], golden['npz/snapshot/20/x_centers'])
    np.testing.assert_array_equal(recorder.centres[1], golden['npz/snapshot/20/y_centers'])
    shape = (recorder.edges[0].size - 1, recorder.edges[1].size - 1)
    rng = np.random.default_rng(3)
    for snapshot in metadata['snapshots']:
        t = snapshot['time']
        recorder.snapshots.append({'time': t, 'label': None, 'time_label': f't = {t * 1e15:.1f} fs',
                                   'fields': {k: rng.standard_normal(shape).astype(np.float32)
                                              for k in recorder.keys}})
    final = recorder.snapshots[-1]['fields']
    fields = {key: (values, recorder.edges[0], recorder.edges[1]) for key, values in final.items()}
    kept = [(recorder.layout(), recorder.snapshots)] * 2
    out = M.Results('completed', {}, {}, [{}, {}], {'parameter': 'n_film', 'values': [1.0, 1.5]}, fields,
                    {'t_end': recorder.snapshots[-1]['time']}, recorder.snapshots, recorder.layout(), kept)
    out.save(tmp_path)
    groups = ('field', 'snapshot', 'sweep_snapshot')
    with np.load(tmp_path / 'results.npz', allow_pickle=False) as npz:
        ours = {n for n in npz.files if n.split('/')[0] in groups}
        saved = json.loads(str(npz['metadata']))
    assert ours == {str(n) for n in golden['npz_names'] if str(n).split('/')[0] in groups}
    assert saved['fields'] == metadata['fields']
    assert saved['snapshots'] == metadata['snapshots']
    assert saved['sweep_snapshots'] == metadata['sweep_snapshots']


def test_stored_fields_request_is_checked():
    from opensol_amd import model as M
    golden = _golden('ez')
    spec = json.loads(str(golden['model_json']))
    built, _ = _dt(spec)
    recorder = M._Recorder(built, ['em.h_abs', 'em.ez'])
    assert recorder.keys == ['em.h_abs', 'em.ez']
    with pytest.raises(ValueError, match="stored field 'em.hz' is not an output field"):
        M._Recorder(built, ['em.hz'])


def test_absorption_tracking_is_refused():
    from opensol_amd import model as M
    spec = json.loads(str(_golden('ez')['model_json']))
    spec['physics'][0].setdefault('settings', {})['track_absorption'] = True
    with pytest.raises(ValueError, match='absorption tracking'):
        M._Built(spec, M.evaluate_parameters(spec['parameters']))


def test_snapshot_byte_limit_keeps_the_final_state(monkeypatch):
    """Past the byte limit intermediate records stop and the final one is still taken (model/study.py:2121-2129);
    a keep_fields sweep past it keeps only each later value's final snapshot (:683-693)."""
    from opensol_amd import model as M

    class Fake:
        def __init__(self):
            self.keys = ['em.ez']
            self.snapshots = []
            self.full = False
            self.stored_bytes = 0

        def layout(self):
            return {}

    monkeypatch.setattr(M, '_SNAPSHOT_BYTE_LIMIT', 100)
    recorder = Fake()
    recorder._values = lambda: {'ez': np.zeros(10, np.float32)}   # 40 bytes a record
    recorder._cut_values = lambda: []
    recorder._cut_views = lambda: []
    recorder.view = None
    recorder.sim = type('S', (), {'time': 0.0})()
    recorder.time_label = lambda: 't = 0.0 fs'
    for index in range(6):
        M._Recorder.record(recorder, final=index == 5)
    assert len(recorder.snapshots) == 4 and recorder.full   # 3 x 40 passes 100 after the third; final appended
    kept = M._KeptSweepSnapshots()
    kept.take(recorder)
    kept.take(recorder)
    assert [len(s) for _, s in kept.kept] == [1, 1] and math.isclose(kept.stored_bytes, 80)


# ---- 3D: the mid-height z cut ----------------------------------------------------------------------------------

def _golden_3d():
    return np.load(os.path.join(FIXTURES, 'model_dipole_3d.npz'))


def test_3d_schedule_and_field_table_match_the_reference():
    """The 3D model's snapshot times (output_interval 5 fs over 15 fs) and the stored-field records, primary first."""
    from opensol_amd import model as M
    golden = _golden_3d()
    spec = json.loads(str(golden['model_json']))
    built, names = _dt(spec)
    settings = spec['study']['settings']
    schedule = M.snapshot_schedule(built.sim.model.dt, M.evaluate(settings['duration'], names),
                                   M.evaluate(settings['output_interval'], names))
    assert [t for _, t in schedule] == list(golden['summary_snapshot_times'])
    assert schedule[-1][0] == int(golden['summary_ticks'])
    metadata = json.loads(str(golden['npz_metadata']))
    stored = metadata['snapshots'][0]['fields']
    assert list(stored) == [f'em.{info["name"]}' for info in M.STORED_FIELDS['3d']]
    for info in M.STORED_FIELDS['3d']:
        assert stored[f'em.{info["name"]}']['info'] == info


def test_3d_saved_layout_matches_the_reference(tmp_path):
    """A 3D recorder built without an engine: the cut plane (the window cell holding mid-height, by its centre), the
    stored x and y cuts (their planes and in-plane centres), the x / y edges and centres, and Results.save()'s snapshot,
    cut and field names and metadata equal the reference's."""
    from opensol_amd import model as M
    golden = _golden_3d()
    metadata = json.loads(str(golden['npz_metadata']))
    spec = json.loads(str(golden['model_json']))
    built, names = _dt(spec)
    recorder = M._Recorder3D(built, (), M._stored_cuts(spec['study'], names))
    assert recorder.plane == metadata['fields']['em.ez']['plane']
    assert recorder.cut_records() == json.loads(str(golden['summary_stored_cuts']))
    for c, plane in enumerate(recorder.planes[1:]):
        for a, axis in enumerate('xy'):
            np.testing.assert_array_equal(plane['centres'][a], golden[f'npz/snapshot/2/cut{c}/em.ez/{axis}_centers'])
    np.testing.assert_array_equal(recorder.edges[0], golden['npz/field/x_edges'])
    np.testing.assert_array_equal(recorder.edges[1], golden['npz/field/y_edges'])
    np.testing.assert_array_equal(recorder.centres[0], golden['npz/snapshot/2/x_centers'])
    np.testing.assert_array_equal(recorder.centres[1], golden['npz/snapshot/2/y_centers'])
    shape = (recorder.edges[0].size - 1, recorder.edges[1].size - 1)
    cut_shapes = [tuple(c.size for c in plane['centres']) for plane in recorder.planes[1:]]
    for snapshot in metadata['snapshots']:
        t = snapshot['time']
        recorder.snapshots.append({'time': t, 'label': None, 'time_label': f't = {t * 1e15:.1f} fs',
                                   'fields': {k: np.zeros(shape, np.float32) for k in recorder.keys},
                                   'cuts': [{k: np.zeros(cs, np.float32) for k in recorder.keys} for cs in cut_shapes]})
    fields = {key: (values, recorder.edges[0], recorder.edges[1])
              for key, values in recorder.snapshots[-1]['fields'].items()}
    out = M.Results('completed', {}, {}, [{}], None, fields, {'t_end': recorder.snapshots[-1]['time']},
                    recorder.snapshots, recorder.layout())
    out.save(tmp_path)
    groups = ('field', 'snapshot', 'sweep_snapshot')
    with np.load(tmp_path / 'results.npz', allow_pickle=False) as npz:
        ours = {n for n in npz.files if n.split('/')[0] in groups}
        saved = json.loads(str(npz['metadata']))
    assert ours == {str(n) for n in golden['npz_names'] if str(n).split('/')[0] in groups}
    for group in ('fields', 'snapshots', 'sweep_snapshots'):
        assert saved[group] == metadata[group], group


class _CutModel:
    def __init__(self, shape, periodic):
        self.shape, self.periodic = shape, periodic


@pytest.mark.parametrize('cut_axis', [0, 1, 2])
@pytest.mark.parametrize('periodic', [(False, False, False), (True, True, True)])
@pytest.mark.parametrize('window_start', [0, 2])
@pytest.mark.parametrize('k', [0, 3])
def test_3d_centre_average_against_a_loop(cut_axis, periodic, window_start, k):
    """_Recorder3D._centre against a per-cell loop of the reference's rule: pair sums along each averaged axis in
    axis order (the lower neighbour of window cell 0 wraps on a periodic axis, else is the cell itself), times
    1 / 2^axes, in float32."""
    from opensol_amd import model as M
    shape = (7, 6, 5)
    rng = np.random.default_rng(11 + window_start + k)
    volume = {name: rng.standard_normal(shape).astype(np.float32) for name in M.CENTRE_AXES_3D}
    recorder = M._Recorder3D.__new__(M._Recorder3D)
    recorder.sim = type('S', (), {'model': _CutModel(shape, periodic)})()
    recorder.window = (slice(window_start, shape[0] - 1), slice(window_start, shape[1]), slice(window_start, shape[2]))
    k_lower = k - 1 if k > 0 else (shape[cut_axis] - 1 if periodic[cut_axis] else k)
    pick = [slice(None)] * 3
    pick[cut_axis] = [k_lower, k]
    layers = {name: v[tuple(pick)] for name, v in volume.items()}
    in_plane = [b for b in range(3) if b != cut_axis]

    def lower(axis, index):   # grid row 0 pairs with the last row on a periodic axis, else with itself
        return index - 1 if index > 0 else (shape[axis] - 1 if periodic[axis] else index)

    for name, axes in M.CENTRE_AXES_3D.items():
        got = recorder._centre(layers, name, cut_axis)
        want = np.empty_like(got)
        u, v = in_plane
        for a, i in enumerate(range(recorder.window[u].start, recorder.window[u].stop)):
            for b, j in enumerate(range(recorder.window[v].start, recorder.window[v].stop)):
                def value(p):
                    return volume[name][p]

                def summed(p, remaining):   # pair sums along the axes, the first axis innermost
                    if not remaining:
                        return value(p)
                    axis = remaining[-1]
                    q = list(p)
                    q[axis] = lower(axis, p[axis])
                    return np.float32(summed(p, remaining[:-1]) + summed(tuple(q), remaining[:-1]))

                point = [0, 0, 0]
                point[u], point[v], point[cut_axis] = i, j, k
                want[a, b] = np.float32(summed(tuple(point), list(axes)) * np.float32(1.0 / 2 ** len(axes)))
        np.testing.assert_array_equal(got, want, err_msg=name)


def test_a_sweep_without_values_is_refused():
    from opensol_amd import model as M
    spec = json.loads(str(_golden('ez')['model_json']))
    spec['study']['sweep']['values'] = []
    with pytest.raises(ValueError, match="sweep of 'n_film' has no values"):
        M.run(spec)


def test_3d_material_cuts_show_the_dielectric_under_pec():
    """The displayed eps_r of the z cut and both stored cuts equals the reference's on the CPU (no fields needed),
    including the y cut through the PEC box, where the reference shows the dielectric painted under the PEC (its PEC
    lives in the engine's index only, physics/em3d/engine3d.py:351-362)."""
    from opensol_amd import model as M
    golden = _golden_3d()
    spec = json.loads(str(golden['model_json']))
    built, names = _dt(spec)
    built.sim.model.finalize()   # paints the PEC box
    recorder = M._Recorder3D(built, (), M._stored_cuts(spec['study'], names))
    materials = built.sim.model.materials
    crossed = 0
    for c, plane in enumerate(recorder.planes):
        cells = list(recorder.window)
        cells[plane['axis']] = plane['k']
        shown = materials.display_array(2, tuple(cells))
        want = golden['npz/field/em.eps_r/values' if c == 0 else f'npz/snapshot/2/cut{c - 1}/em.eps_r/values']
        np.testing.assert_array_equal(shown, want, err_msg=str(plane['plane']))
        crossed += int(np.count_nonzero(materials.pec_flags()[materials.ids[tuple(cells)]]))
    assert crossed > 0   # a cut does cross the PEC box


def test_pec_keeps_the_dielectric_under_it_for_display():
    from opensol_amd.materials import MaterialMap
    m = MaterialMap((6, 5))
    glass = m.material_id(2.25)
    m.paint_box(((0, 4), (0, 5)), glass)
    mask = np.zeros((6, 5), dtype=bool)
    mask[2:5, 1:3] = True
    assert m.paint_pec([mask])
    assert np.all(m.property_array(0)[mask] == -1.0)                  # the kernels' ids: PEC
    shown = m.display_array(0)
    assert np.all(shown[2:4, 1:3] == np.float32(2.25)) and np.all(shown[4, 1:3] == 1.0)   # glass, vacuum under PEC
    assert not m.paint_pec([mask])                                    # nothing changes on a second pass
    m.paint_box(((0, 6), (2, 3)), m.material_id(4.0))                 # a later dielectric over part of the PEC
    assert m.paint_pec([mask])                                        # PEC again on top of it
    assert np.all(m.display_array(0)[2:5, 2] == 4.0) and np.all(m.display_array(0)[2:4, 1] == np.float32(2.25))
    m.set_property((3,), np.full((6, 5), 7.0, np.float32))            # a write into every cell reaches the dielectric
    assert np.all(m.display_array(3) == 7.0) and np.all(m.property_array(0)[mask] == -1.0)



def test_big_window_snapshots_are_resampled_like_the_reference(tmp_path):
    """A 553 x 459 cell window (past the 250 000-cell snapshot limit): the snapshots hold the nearest-cell resampling
    on a 548 x 455 display grid with native_shape [553, 459] and no centre arrays, the field group keeps the native
    cells. The displayed eps_r of the resampled snapshot equals the reference's (the material needs no GPU), and the
    saved names and snapshot / field metadata equal the reference run's."""
    from opensol_amd import model as M
    golden = np.load(os.path.join(FIXTURES, 'model_big_window_ez.npz'))
    metadata = json.loads(str(golden['npz_metadata']))
    spec = json.loads(str(golden['model_json']))
    built, names = _dt(spec)
    built.sim.model.finalize()
    recorder = M._Recorder(built)
    assert recorder.view is not None and recorder.view[2] == [553, 459]
    shown = M._shown(built.sim.model.materials.display_array(0, recorder.window), recorder.view)
    np.testing.assert_array_equal(shown, golden['npz/snapshot/1/em.eps_r/values'])
    native = tuple(e.size - 1 for e in recorder.edges)
    for snapshot in metadata['snapshots']:
        t = snapshot['time']
        values = {k: np.zeros(native, np.float32) for k in recorder.keys}
        recorder.snapshots.append({'time': t, 'label': None, 'time_label': f't = {t * 1e15:.1f} fs',
                                   'fields': {k: M._shown(v, recorder.view) for k, v in values.items()},
                                   'cuts': []})
        recorder.final_fields = values
    fields = {key: (v, recorder.edges[0], recorder.edges[1]) for key, v in recorder.final_fields.items()}
    out = M.Results('completed', {}, {}, [{}], None, fields, {'t_end': recorder.snapshots[-1]['time']},
                    recorder.snapshots, recorder.layout())
    out.save(tmp_path)
    groups = ('field', 'snapshot', 'sweep_snapshot')
    with np.load(tmp_path / 'results.npz', allow_pickle=False) as npz:
        ours = {n for n in npz.files if n.split('/')[0] in groups}
        saved = json.loads(str(npz['metadata']))
        assert npz['snapshot/1/em.ez/values'].shape == golden['npz/snapshot/1/em.ez/values'].shape == (548, 455)
        assert npz['field/em.ez/values'].shape == (553, 459)
    assert ours == {str(n) for n in golden['npz_names'] if str(n).split('/')[0] in groups}
    for group in ('fields', 'snapshots', 'sweep_snapshots'):
        assert saved[group] == metadata[group], group
