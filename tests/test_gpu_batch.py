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
