"""Batched 2D instances (E_z and H_z of uneven sizes, each with its own sources, monitors and step count) stepped by
one launch per step give every instance the same fields and flux spectra, bit for bit, as running it alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (window cells x, y, npml, polarization, boundaries, steps)
INSTANCES = [((40, 30), 6, 'ez', None, 150), ((25, 61), 8, 'hz', None, 97), ((70, 18), 5, 'ez', 'periodic_y', 203),
             ((33, 33), 6, 'hz', 'periodic_x', 120), ((90, 41), 10, 'ez', None, 64), ((17, 23), 4, 'hz', None, 180)]


def _make(spec, seed):
    from opensol_amd import Simulation
    (cx, cy), npml, pol, bnd, _ = spec
    dx = 25e-9
    sx, sy = cx * dx, cy * dx
    sim = Simulation(sx, sy, dx, npml=npml, boundaries=bnd, polarization=pol)
    rng = np.random.default_rng(seed)
    for _ in range(5):
        x0, y0 = rng.uniform(0, sx), rng.uniform(0, sy)
        sim.add_dielectric_rectangle(x0, y0, x0 + rng.uniform(0.1, 0.5) * sx, y0 + rng.uniform(0.1, 0.5) * sy,
                                     eps_r=rng.uniform(1.0, 9.0))
    sim.add_line_source(axis='y', position=0.3 * sx, wavelength=rng.uniform(0.4e-6, 0.8e-6), bandwidth=0.6)
    monitor = sim.add_flux_monitor(axis='y', position=0.7 * sx, span=(0.0, sy), wavelengths=[0.45e-6, 0.6e-6])
    return sim, monitor


def _snapshot(sim, monitor):
    sim.engine.synchronize()
    return {name: sim.engine.field(name) for name in sim.model.components}, monitor.flux()


def test_batch_matches_solo_runs_bit_for_bit(gpu):
    from opensol_amd.batch import Batch2D
    solo = []
    for k, spec in enumerate(INSTANCES):
        sim, monitor = _make(spec, k)
        sim.run(n_steps=spec[4], print_every=0)
        first = _snapshot(sim, monitor)
        sim.run(n_steps=31, print_every=0)
        solo.append((first, _snapshot(sim, monitor)))
        sim.close()
    made = [_make(spec, k) for k, spec in enumerate(INSTANCES)]
    batch = Batch2D([sim for sim, _ in made], graph_steps=16)
    batch.run([spec[4] for spec in INSTANCES])
    for (sim, monitor), (first, _) in zip(made, solo):
        fields, flux = _snapshot(sim, monitor)
        for name, values in first[0].items():
            assert np.array_equal(fields[name], values), name
        assert np.array_equal(flux, first[1])
    batch.run(31)   # continue every instance together
    for (sim, monitor), (_, second) in zip(made, solo):
        fields, flux = _snapshot(sim, monitor)
        for name, values in second[0].items():
            assert np.array_equal(fields[name], values), name
        assert np.array_equal(flux, second[1])
    batch.close()


def test_autotuned_shape_gives_the_same_bits(gpu):
    from opensol_amd.autotune import tune
    spec = INSTANCES[0]
    tuned, monitor_t = _make(spec, 0)
    options, timings = tune(tuned, steps=4, use_cache=False)
    assert set(options) == {'lane_columns', 'rows_per_tile'} and len(timings) >= 2
    plain, monitor_p = _make(spec, 0)
    tuned.run(n_steps=40, print_every=0)
    plain.run(n_steps=40, print_every=0)
    a, b = _snapshot(tuned, monitor_t), _snapshot(plain, monitor_p)
    for name in a[0]:
        assert np.array_equal(a[0][name], b[0][name])
    assert np.array_equal(a[1], b[1])


# 3D: (window cells x, y, z, npml, boundaries, steps, planes per tile in the batch)
PERIODIC_Y = {'y_lo': 'periodic', 'y_hi': 'periodic'}
PERIODIC_X = {'x_lo': 'periodic', 'x_hi': 'periodic'}
INSTANCES_3D = [((20, 14, 18), 4, None, 60), ((9, 22, 12), 3, PERIODIC_Y, 41), ((16, 16, 70), 4, None, 75),
                ((30, 8, 11), 3, PERIODIC_X, 33), ((12, 12, 12), 2, None, 90)]


def _make_3d(spec, seed):
    from opensol_amd import Simulation3D
    (cx, cy, cz), npml, bnd, _ = spec
    dx = 25e-9
    size = np.array([cx, cy, cz]) * dx
    sim = Simulation3D(*size, dx, npml=npml, boundaries=bnd)
    rng = np.random.default_rng(100 + seed)
    for _ in range(4):
        lo = rng.uniform(0, 1, 3) * size
        sim.add_dielectric_box(tuple(lo), tuple(lo + rng.uniform(0.2, 0.6, 3) * size), eps_r=rng.uniform(1.0, 9.0))
    centre = size / 2
    sim.add_dipole(*centre, orientation='xyz'[seed % 3], wavelength=0.5e-6, bandwidth=0.6)
    plane = sim.add_flux_plane('x', 0.75 * size[0], wavelengths=[0.45e-6, 0.55e-6])
    return sim, plane


def _snapshot_3d(sim, plane):
    sim.engine.synchronize()
    return {name: sim.engine.field(name) for name in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')}, plane.flux()


@pytest.mark.parametrize('planes', [0, 3])
def test_batch3d_matches_solo_runs_bit_for_bit(gpu, planes):
    from opensol_amd.batch import Batch3D
    solo = []
    for k, spec in enumerate(INSTANCES_3D):
        sim, plane = _make_3d(spec, k)
        sim.run(n_steps=spec[3])
        first = _snapshot_3d(sim, plane)
        sim.run(n_steps=17)
        solo.append((first, _snapshot_3d(sim, plane)))
        sim.close()
    made = [_make_3d(spec, k) for k, spec in enumerate(INSTANCES_3D)]
    batch = Batch3D([sim for sim, _ in made], graph_steps=8, planes_per_tile=planes)
    batch.run([spec[3] for spec in INSTANCES_3D])
    for (sim, plane), (first, _) in zip(made, solo):
        fields, flux = _snapshot_3d(sim, plane)
        for name, values in first[0].items():
            assert np.array_equal(fields[name], values), name
        assert np.array_equal(flux, first[1])
    batch.run(17)
    for (sim, plane), (_, second) in zip(made, solo):
        fields, flux = _snapshot_3d(sim, plane)
        for name, values in second[0].items():
            assert np.array_equal(fields[name], values), name
        assert np.array_equal(flux, second[1])
    batch.close()
    for sim, _ in made:
        sim.close()


def test_autotuned_batches_give_the_same_bits_and_fill_the_cache(gpu, tune_cache):
    """Batch2D / Batch3D with autotune=True: every member's launch shape comes from its size class (tuned as a
    one-member batch, then cached in the session's JSON file), and the fields and flux equal an untuned batch's bit
    for bit; a second member of a tuned class takes the cached shape without timing."""
    import json
    from opensol_amd.autotune import tune_batch_member
    from opensol_amd.batch import Batch2D, Batch3D
    runs = []
    for tuned in (True, False):
        made = [_make(spec, k) for k, spec in enumerate(INSTANCES)]
        batch = Batch2D([sim for sim, _ in made], autotune=tuned)
        batch.run([spec[4] for spec in INSTANCES])
        runs.append([_snapshot(sim, monitor) for sim, monitor in made])
        if tuned:
            assert all(sim._engine_options.get('rows_per_tile') for sim, _ in made)
        batch.close()
        for sim, _ in made:
            sim.close()
    for (fa, xa), (fb, xb) in zip(*runs):
        for name in fa:
            assert np.array_equal(fa[name], fb[name]), name
        assert np.array_equal(xa, xb)
    runs = []
    for tuned in (True, False):
        made = [_make_3d(spec, k) for k, spec in enumerate(INSTANCES_3D)]
        batch = Batch3D([sim for sim, _ in made], autotune=tuned)
        batch.run([spec[3] for spec in INSTANCES_3D])
        runs.append([_snapshot_3d(sim, plane) for sim, plane in made])
        batch.close()
        for sim, _ in made:
            sim.close()
    for (fa, xa), (fb, xb) in zip(*runs):
        for name in fa:
            assert np.array_equal(fa[name], fb[name]), name
        assert np.array_equal(xa, xb)
    with open(tune_cache) as f:
        entries = next(iter(json.load(f).values()))
    assert any(key.startswith('batch:ez:') for key in entries) and any(key.startswith('batch:3d:') for key in entries)
    again, _ = _make(INSTANCES[0], 0)
    options, timings = tune_batch_member(again)
    assert timings == {} and options == {'rows_per_tile': again._engine_options['rows_per_tile']}


def test_tall_grid_single_and_batched_kernels_agree(gpu):
    """A tall grid (77 y tiles, 2 z tiles) stepped by the single-instance kernel and by the batched kernel gives the
    same bits."""
    from opensol_amd.batch import Batch3D
    spec = ((10, 300, 70), 4, None, 45)
    solo, solo_plane = _make_3d(spec, 7)
    solo.run(n_steps=spec[3])
    expected = _snapshot_3d(solo, solo_plane)
    solo.close()
    sim, plane = _make_3d(spec, 7)
    batch = Batch3D([sim])
    batch.run(spec[3])
    fields, flux = _snapshot_3d(sim, plane)
    for name, values in expected[0].items():
        assert np.array_equal(fields[name], values), name
    assert np.array_equal(flux, expected[1])
    batch.close()
    sim.close()
