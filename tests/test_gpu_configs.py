"""The benchmark workloads at their stated sizes, on the GPU (the generators are bench.py's own):

  * the 8192^2 E_z roofline grid (20-cell CPML, 400 rectangles): a few steps from random fields against the FP64
    oracle over the whole grid, so every bulk and edge tile of the real launch is checked;
  * the 512^3 generator at 128^3 (its boxes, dipole and flux planes): a dipole run against the FP64 oracle, and
    random fields for a few steps;
  * the batched config (sides 128-1024 cells, E_z and H_z mixed, 12-cell CPML, 20-100 rectangles each): the first 32
    instances of the 256-instance stream in one batch, every instance bit-identical to its solo run, the four smallest
    against the FP64 oracle;
  * the served config (Poisson stream of 2D 128-1024 and 3D 64^3-192^3 jobs, 500-5000 steps): 20 jobs served under
    the 200 GB budget and under a budget that holds jobs back, every result equal to the solo and the plain runs.

Tolerances: fields <= 1e-5 relative L2 and flux <= 1e-5 against the FP64 oracle after a pulse run; random-field
steps <= 5e-6 of each field's peak; batched / served / solo / plain bit for bit.
"""
import numpy as np
import pytest

from tests.cases import rel_l2, scaled_error, update_vector_scales

pytestmark = pytest.mark.gpu


def _bench():
    import bench
    return bench


def _random_fields(sim, ref, seed, names):
    rng = np.random.default_rng(seed)
    for name in names:
        values = rng.standard_normal(ref.fields[name].shape).astype(np.float32)
        ref.fields[name][...] = values
        sim.engine.set_field(name, values)


def _compare_peak(sim, ref, names, tol, label):
    for name in names:
        got, want = sim.engine.field(name), ref.fields[name]
        error = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
        print(f'{label} {name}: max err / peak {error:.3e}', flush=True)
        assert error <= tol, (label, name, error)


def test_ez8192_roofline_grid_matches_oracle(gpu):
    """The exact bench grid: 4 steps from random float32 fields on all 8192^2 cells (bulk tiles, edge tiles, the CPML
    layers and their corners, the source line) against the FP64 oracle of the same model."""
    from oracle.yee2d import Yee2D
    bench = _bench()
    print('building the 8192^2 E_z workload (product and oracle)', flush=True)
    sim = bench.build_ez_workload()
    ref = bench.build_ez_workload(target=Yee2D)
    sim.finalize()
    ref.finalize()
    np.testing.assert_array_equal(sim.model.materials.property_array(0), ref.materials['eps_r'])
    _random_fields(sim, ref, 8192, ('ez', 'hx', 'hy'))
    for _ in range(4):
        ref.step()
    sim.run(n_steps=4, print_every=0)
    _compare_peak(sim, ref, ('ez', 'hx', 'hy'), 5e-6, 'ez8192 random 4 steps')
    sim.close()


def test_yee3d_generator_at_128_matches_oracle(gpu):
    """The 512^3 workload's generator at 128^3 (boxes, the E_z dipole at the centre, two flux planes at 4 wavelengths,
    m = 17): 220 steps of the pulse against the FP64 oracle (each component relative to the largest norm of its field
    vector, E or H: H_z of the centred E_z dipole is near zero by symmetry; flux), then 3
    steps from random fields (every tile of the 3D step, the CPML on all six faces)."""
    from oracle.yee3d import Yee3D
    bench = _bench()
    sim = bench.build_3d_workload(n=128)
    ref = bench.build_3d_workload(n=128, target=Yee3D)
    sim.finalize()
    ref.finalize()
    assert sim.model.sample_interval == ref.sample_interval == 17
    names = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')
    scales = {}
    for chunk in (110, 110):
        print(f'yee3d 128: {chunk} more steps', flush=True)
        sim.run(n_steps=chunk)
        ref.run(n_steps=chunk)
        update_vector_scales({name: ref.fields[name] for name in names}, scales)
        for name in names:
            error = scaled_error(sim.engine.field(name), ref.fields[name], scales, name)
            print(f'yee3d 128 step {sim.engine.n_step} {name}: {error:.3e}', flush=True)
            assert error <= 1e-5, (name, error)
    for plane, ref_plane in zip(sim.model.monitors, ref.monitors):
        flux, want = np.asarray(plane.flux()), np.asarray(ref_plane.flux())
        error = float(np.max(np.abs(flux - want)) / np.max(np.abs(want)))
        print(f'yee3d 128 flux error / peak {error:.3e}', flush=True)
        assert error <= 1e-5
    _random_fields(sim, ref, 128, names)
    for _ in range(3):
        ref.step()
    sim.run(n_steps=3)
    _compare_peak(sim, ref, names, 5e-6, 'yee3d 128 random 3 steps')
    sim.close()


def test_batched_config_instances(gpu):
    """The first 32 instances of the batched config's stream (sides 128-1024, E_z and H_z), 300-800 steps each, in one
    Batch2D with graphs of 50 steps: every instance's fields and flux equal its solo run bit for bit; the four smallest
    match the FP64 oracle (fields <= 1e-5 relative L2, flux <= 1e-5)."""
    from oracle.yee2d import Yee2D
    from opensol_amd.batch import Batch2D
    bench = _bench()
    count = 32
    steps = [int(n) for n in np.random.default_rng(4).integers(300, 801, count)]
    batched, _ = bench.build_batch_instances(count=count, graph_steps=50)
    sizes = [s.model.shape for s in batched]
    assert min(min(s) for s in sizes) >= 128 and max(max(s) for s in sizes) <= 1024
    assert {s.model.polarization for s in batched} == {'ez', 'hz'}
    batch = Batch2D(batched, graph_steps=50)
    print(f'batch of {count}: {sum(nx * ny for nx, ny in sizes)} cells', flush=True)
    batch.run(steps)
    batch.synchronize()
    solo, _ = bench.build_batch_instances(count=count)
    for k, (one, many, n) in enumerate(zip(solo, batched, steps)):
        one.run(n_steps=n, print_every=0)
        for name in one.model.components:
            assert np.array_equal(one.engine.field(name), many.engine.field(name)), (k, name)
        for a, b in zip(one.model.monitors, many.model.monitors):
            assert np.array_equal(a.flux(), b.flux()), k
        one.close()
    oracle, _ = bench.build_batch_instances(count=count, target=Yee2D)
    smallest = sorted(range(count), key=lambda k: sizes[k][0] * sizes[k][1] * steps[k])[:4]
    for k in smallest:
        ref = oracle[k]
        ref.run(n_steps=steps[k])
        sim = batched[k]
        for name in sim.model.components:
            error = rel_l2(sim.engine.field(name), ref.fields[name])
            print(f'batch instance {k} {sizes[k]} {sim.model.polarization} {steps[k]} steps {name}: {error:.3e}',
                  flush=True)
            assert error <= 1e-5, (k, name, error)
        flux, want = np.asarray(sim.model.monitors[0].flux()), np.asarray(ref.monitors[0].flux())
        error = float(np.max(np.abs(flux / want - 1.0)))
        print(f'batch instance {k} flux {error:.3e}', flush=True)
        assert error <= 1e-5
    batch.close()


def test_served_config_stream(gpu):
    """20 jobs of the served config's stream at 400 jobs/s (2D 128-1024 and 3D 64^3-192^3, 500-5000 steps): served
    under the 200 GB budget and under a budget of twice the largest job (which holds jobs back), every job's flux
    spectra equal its solo run and the plain FIFO run bit for bit; the resident jobs never exceed the budget."""
    from opensol_amd.serve import FDTDServer, state_bytes
    bench = _bench()
    jobs = bench.serve_jobs(400, 20)
    kinds = {job.name.split('-')[0] for job in jobs}
    assert kinds == {'2d', '3d'}, kinds
    largest = max(state_bytes(job.build()) for job in jobs)
    results = {}
    for budget in (200e9, 2.0 * largest):
        print(f'serving 20 jobs under {budget:.3e} B', flush=True)
        stream = bench.serve_jobs(400, 20)
        stats = FDTDServer(hbm_budget=budget, chunk_steps=100).run(stream)
        assert stats['jobs'] == 20
        for job in stream:
            resident = [o for o in stream if o.admitted <= job.admitted < o.finished]
            assert sum(o.bytes for o in resident) <= budget
        results[budget] = stats['results']
        print(f'  p50 {stats["p50_latency_s"]:.3f} s p99 {stats["p99_latency_s"]:.3f} s '
              f'{stats["cell_updates_per_s"]:.3e} cell-updates/s', flush=True)
    plain = FDTDServer(hbm_budget=200e9).run_plain(bench.serve_jobs(400, 20))['results']
    for job in bench.serve_jobs(400, 20):
        sim = job.build()
        sim.run(n_steps=job.steps, print_every=0)
        solo = [np.asarray(m.flux()) for m in sim.model.monitors]
        sim.close()
        for other in (*results.values(), plain):
            for a, b in zip(solo, other[job.name]):
                assert np.array_equal(a, b), job.name
