"""GPU parity of the fused 2D HIP step (libyee.so through the Simulation API) against the FP64 oracle.

Tolerances (FP32 fields against the FP64 oracle): relative L2 <= 1e-5 for field snapshots, <= 1e-6 absolute on the
normalised transmission; random-input kernel checks <= 5e-6 of each field's peak after a few steps; graph replay and
batching are bit-identical to eager solo stepping.
"""
import numpy as np
import pytest

from oracle.yee2d import Yee2D
from tests.cases import SNAPSHOT_STEPS, THIN_FILM, build, random_rectangles, rel_l2, scaled_error, update_vector_scales

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
TRANSMISSION_TOL = 1e-6


def simulation_class():
    from opensol_amd import Simulation
    return Simulation


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_thin_film_matches_oracle(gpu, polarization):
    Simulation = simulation_class()
    flux = {}
    for tag, with_slab in (('film', True), ('empty', False)):
        ref, _, ref_monitor = build(Yee2D, THIN_FILM, polarization, with_slab, phase_dtype=np.complex128)
        sim, _, monitor = build(Simulation, THIN_FILM, polarization, with_slab)
        done = 0
        pulse_norm = {}
        for step in SNAPSHOT_STEPS:
            ref.run(n_steps=step - done)
            sim.run(n_steps=step - done, print_every=0)
            done = step
            for name in sim.model.components:
                # relative to the field norm while the pulse is inside the window (steps 60 and 165); by step 400
                # the empty-domain pulse has left, and what remains is a tail far below it
                pulse_norm[name] = max(pulse_norm.get(name, 0.0), float(np.linalg.norm(ref.fields[name])))
                if pulse_norm[name] == 0.0:
                    # the y-invariant case keeps H_x (E_z) / E_x (H_z) exactly 0: every copy of a cell computed by
                    # two tiles rounds identically, so no seam appears
                    assert not np.any(sim.engine.field(name)), (tag, step, name)
                    continue
                scale = pulse_norm[name]
                error = float(np.linalg.norm(sim.engine.field(name) - ref.fields[name])) / scale
                print(f'{polarization} {tag} step {step} {name}: rel L2 {error:.3e}')
                assert error <= FIELD_TOL, (tag, step, name, error)
        ref.run(duration=THIN_FILM['duration'] - ref.time)
        sim.run(duration=THIN_FILM['duration'] - sim.time, print_every=0)
        assert sim.n_step == ref.n_step == 1370
        assert sim.time == ref.time
        flux[tag] = (monitor.flux(), ref_monitor.flux())
        flux_error = np.max(np.abs(flux[tag][0] / flux[tag][1] - 1.0))
        print(f'{polarization} {tag} flux rel error {flux_error:.3e}')
        assert flux_error <= 1e-5
    t_gpu = flux['film'][0] / flux['empty'][0]
    t_ref = flux['film'][1] / flux['empty'][1]
    print(f'{polarization} T gpu {t_gpu} oracle {t_ref} |diff| {np.abs(t_gpu - t_ref)}')
    assert np.max(np.abs(t_gpu - t_ref)) <= TRANSMISSION_TOL


# (window cells x, y, npml, boundaries, rows_per_tile): tile edges inside CPML layers, ragged last tiles, wraps
CASES = [
    (40, 30, 8, None, 64),
    (70, 130, 6, None, 16),
    (33, 127, 5, 'periodic_y', 7),
    (90, 64, 10, 'periodic_x', 13),
    (48, 126, 0, 'periodic', 5),
    (25, 200, 12, None, 3),
    (40, 300, 6, None, 17),          # two wavefront tiles across y (256 columns each)
    (30, 520, 0, 'periodic', 11),    # three tiles, wraps across tile edges
    (24, 517, 7, 'periodic_x', 9),   # odd width: padded pitch, ragged last lane
    (20, 301, 5, 'periodic_y', 8),   # periodic y wrap from the ragged last lane to column 0
]


def _same_model(target, polarization, case, rng_seed, **options):
    wx, wy, npml, boundaries, _ = case
    dx = 25e-9
    sim = target(wx * dx, wy * dx, dx, npml=max(npml, 1), boundaries=boundaries, polarization=polarization,
                 **options)
    rng = np.random.default_rng(rng_seed)
    # every other rectangle conductive (sigma_e up to 6e4 S/m): the semi-implicit E coefficients and, for H_z, the
    # float32 face means of eps and sigma across two ids
    sigmas = np.random.default_rng(rng_seed + 7).uniform(0.0, 6e4, 6)
    for k, (x0, y0, x1, y1, eps) in enumerate(random_rectangles(rng, wx * dx, wy * dx, 6)):
        sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps, sigma_e=float(sigmas[k]) if k % 2 else 0.0)
    sim.add_line_source(axis='x', position=0.4 * wy * dx, wavelength=700e-9, bandwidth=0.6)
    sim.add_line_source(axis='y', position=0.6 * wx * dx, wavelength=500e-9, bandwidth=0.4, amplitude=0.5)
    sim.add_flux_monitor(axis='y', position=0.7 * wx * dx, span=(0.1 * wy * dx, 0.9 * wy * dx),
                         wavelengths=[500e-9, 650e-9])
    sim.add_flux_monitor(axis='x', position=0.3 * wy * dx, span=(0.0, wx * dx), wavelengths=[600e-9])
    return sim


@pytest.mark.parametrize('lane_columns', [2, 4])
@pytest.mark.parametrize('polarization', ['ez', 'hz'])
@pytest.mark.parametrize('case', CASES, ids=[f'{c[0]}x{c[1]}-pml{c[2]}-{c[3]}-r{c[4]}' for c in CASES])
def test_random_fields_match_oracle(gpu, polarization, case, lane_columns):
    """Random float32 initial fields, random materials, both source and monitor orientations: every branch of the
    fused kernel (tile halos, CPML layers and corners, periodic wraps, ragged tiles) against the oracle."""
    Simulation = simulation_class()
    seed = 1234 + CASES.index(case)
    ref = _same_model(Yee2D, polarization, case, seed)
    sim = _same_model(Simulation, polarization, case, seed, rows_per_tile=case[4], lane_columns=lane_columns)
    ref.finalize()
    sim.finalize()
    rng = np.random.default_rng(seed + 99)
    for name in sim.model.components:
        values = rng.standard_normal(sim.model.shape).astype(np.float32)
        ref.fields[name][...] = values
        sim.engine.set_field(name, values)
    for _ in range(9):
        ref.step()
    sim.run(n_steps=9, print_every=0)
    for name in sim.model.components:
        got, want = sim.engine.field(name), ref.fields[name]
        error = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f'{polarization} {case} {name}: max err / peak {error:.3e}')
        assert error <= 5e-6, (name, error)
    for mon_gpu, mon_ref in zip(sim.monitors, ref.monitors):
        e_gpu, h_gpu = mon_gpu.spectra()
        scale = np.max(np.abs(mon_ref.e_spectrum)) + 1e-300
        assert np.max(np.abs(e_gpu - mon_ref.e_spectrum)) <= 5e-6 * scale
        scale = np.max(np.abs(mon_ref.h_spectrum)) + 1e-300
        assert np.max(np.abs(h_gpu - mon_ref.h_spectrum)) <= 5e-6 * scale


def test_graph_replay_is_bit_identical(gpu):
    Simulation = simulation_class()
    eager, _, eager_monitor = build(Simulation, THIN_FILM, 'ez', True)
    graphed, _, graphed_monitor = build(Simulation, THIN_FILM, 'ez', True, graph_steps=16)
    eager.run(n_steps=333, print_every=0)
    graphed.run(n_steps=333, print_every=0)
    for name in eager.model.components:
        np.testing.assert_array_equal(graphed.engine.field(name), eager.engine.field(name))
    np.testing.assert_array_equal(graphed_monitor.flux(), eager_monitor.flux())


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_launch_geometry_does_not_change_results(gpu, polarization):
    """Tile length, lanes' column count and the automatic geometry give bit-identical fields and spectra."""
    Simulation = simulation_class()
    results = []
    first = 'hz' if polarization == 'hz' else 'ez'
    second = 'ey' if polarization == 'hz' else 'hy'
    for rows, lanes in ((1, 4), (9, 2), (64, 4), (300, 2), (0, 4), (0, 2)):
        sim, _, monitor = build(Simulation, THIN_FILM, polarization, True, rows_per_tile=rows, lane_columns=lanes)
        sim.run(n_steps=200, print_every=0)
        results.append((sim.engine.field(first), sim.engine.field(second), monitor.flux()))
    for other in results[1:]:
        for a, b in zip(results[0], other):
            np.testing.assert_array_equal(a, b)


def _product_boundaries(faces, npml):
    from opensol_amd.boundaries import BoundaryConfig
    bc = BoundaryConfig(*faces)
    bc.cpml_npml = npml
    return bc


def _oracle_boundaries(faces, npml):
    from oracle.yee2d import FACES, Boundaries
    return Boundaries(dict(zip(FACES, faces)), npml=npml)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_pec_case_matches_oracle(gpu, polarization):
    """PEC region under a later dielectric, a source line crossing it and (E_z) a 'pec' x_lo face against the FP64
    oracle (which reproduces the reference's own run to 1e-12): fields <= 1e-5 relative L2, flux <= 1e-5; E_z stays
    exactly 0 in PEC cells (E_z) and the E faces touching PEC stay exactly 0 (H_z)."""
    from tests.cases import PEC2D, build_pec2d
    Simulation = simulation_class()
    ref, _, ref_monitor = build_pec2d(Yee2D, polarization, _oracle_boundaries, phase_dtype=np.complex128)
    sim, _, monitor = build_pec2d(Simulation, polarization, _product_boundaries)
    done = 0
    for step in PEC2D['steps']:
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        done = step
        for name in sim.model.components:
            error = rel_l2(sim.engine.field(name), ref.fields[name])
            print(f'{polarization} step {step} {name}: rel L2 {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
    pec = ref.pec_mask()
    if polarization == 'ez':
        assert not np.any(sim.engine.field('ez')[pec])
    else:
        assert not np.any(sim.engine.field('ex')[~ref.keep_x]) and not np.any(sim.engine.field('ey')[~ref.keep_y])
    flux_error = np.max(np.abs(monitor.flux() / ref_monitor.flux() - 1.0))
    print(f'{polarization} PEC flux rel error {flux_error:.3e}')
    assert flux_error <= 1e-5


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_lossy_case_matches_oracle(gpu, polarization):
    """Conductive rectangles (sigma_e 5e3 to 3e5 S/m; the Ricker source line inside one), a cw line source over part of
    a line and a flux line sampled every step, against the FP64 oracle (which reproduces the reference's own run to
    1e-12): fields <= 1e-5 relative L2 at both steps, flux <= 1e-5; graph replay bit-identical to eager."""
    from tests.cases import LOSSY2D, build_lossy2d
    Simulation = simulation_class()
    ref, _, ref_monitor = build_lossy2d(Yee2D, polarization, _oracle_boundaries, phase_dtype=np.complex128)
    sim, _, monitor = build_lossy2d(Simulation, polarization, _product_boundaries)
    graphed, _, graphed_monitor = build_lossy2d(Simulation, polarization, _product_boundaries, graph_steps=16)
    done = 0
    for step in LOSSY2D['steps']:
        print(f'lossy {polarization} to step {step}', flush=True)
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        for name in sim.model.components:
            error = rel_l2(sim.engine.field(name), ref.fields[name])
            print(f'lossy {polarization} step {step} {name}: rel L2 {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
            np.testing.assert_array_equal(graphed.engine.field(name), sim.engine.field(name))
    assert monitor.sample_interval == ref_monitor.sample_interval == 1
    flux_error = np.max(np.abs(monitor.flux() / ref_monitor.flux() - 1.0))
    print(f'lossy {polarization} flux rel error {flux_error:.3e}')
    assert flux_error <= 1e-5
    np.testing.assert_array_equal(graphed_monitor.flux(), monitor.flux())


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_point_currents_match_oracle(gpu, polarization):
    """Point currents (E_z filament currents; H_z magnetic line current injected after the H update and in-plane
    dipole faces, one wrapping at y cell 0, one touching PEC) against the FP64 oracle, which reproduces the
    reference's own run to 1e-12: fields <= 1e-5 relative L2 at both steps; graph replay bit-identical to eager."""
    from tests.cases import POINT2D, build_point2d
    Simulation = simulation_class()
    ref, _ = build_point2d(Yee2D, polarization, _oracle_boundaries)
    sim, _ = build_point2d(Simulation, polarization, _product_boundaries)
    graphed, _ = build_point2d(Simulation, polarization, _product_boundaries, graph_steps=16)
    done = 0
    for step in POINT2D['steps']:
        print(f'point currents {polarization} to step {step}', flush=True)
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        for name in sim.model.components:
            error = rel_l2(sim.engine.field(name), ref.fields[name])
            print(f'{polarization} step {step} {name}: rel L2 {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
            np.testing.assert_array_equal(graphed.engine.field(name), sim.engine.field(name))
    if polarization == 'ez':
        assert not np.any(sim.engine.field('ez')[ref.pec_mask()])


def test_hz_pmc_walls_match_oracle(gpu):
    """H_z with pmc x_lo and y_hi faces (H_z = 0 on the outermost cells after the H update) and a magnetic line current
    beside their corner, against the FP64 oracle: fields within 1e-5 of each field's largest norm over the run (by
    step 400 the pulse has left and the remainder is 5.7 % of the step-150 norm), the wall cells exactly 0, graph
    replay bit-identical."""
    from tests.cases import WALLS2D, build_walls2d
    Simulation = simulation_class()
    ref = build_walls2d(Yee2D, _oracle_boundaries, 'hz')
    sim = build_walls2d(Simulation, _product_boundaries, 'hz')
    graphed = build_walls2d(Simulation, _product_boundaries, 'hz', graph_steps=16)
    done, peak = 0, {}
    for step in WALLS2D['steps']:
        print(f'hz walls to step {step}', flush=True)
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        for name in sim.model.components:
            peak[name] = max(peak.get(name, 0.0), float(np.linalg.norm(ref.fields[name])))
            error = float(np.linalg.norm(sim.engine.field(name) - ref.fields[name])) / peak[name]
            print(f'hz walls step {step} {name}: L2 error / peak norm {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
            np.testing.assert_array_equal(graphed.engine.field(name), sim.engine.field(name))
    hz = sim.engine.field('hz')
    assert not np.any(hz[0]) and not np.any(hz[:, -1]) and np.abs(hz).max() > 0


def test_walls_and_mirrors_match_oracle(gpu):
    """E_z with symmetric, pmc and antisymmetric faces against the FP64 oracle (which reproduces the reference's run to
    1e-12): fields <= 1e-5 relative L2 at both steps; the pinned rows are exact copies (negated on the antisymmetric
    face); graph replay bit-identical to eager stepping."""
    from tests.cases import WALLS2D, build_walls2d
    Simulation = simulation_class()
    ref = build_walls2d(Yee2D, _oracle_boundaries)
    sim = build_walls2d(Simulation, _product_boundaries)
    graphed = build_walls2d(Simulation, _product_boundaries, graph_steps=16)
    done = 0
    for step in WALLS2D['steps']:
        print(f'walls to step {step}', flush=True)
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        for name in sim.model.components:
            error = rel_l2(sim.engine.field(name), ref.fields[name])
            print(f'walls step {step} {name}: rel L2 {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
            np.testing.assert_array_equal(graphed.engine.field(name), sim.engine.field(name))
    ez = sim.engine.field('ez')
    assert np.array_equal(ez[0, 1:], ez[1, 1:]) and np.array_equal(ez[-1], ez[-2])
    assert np.array_equal(ez[1:, 0], -ez[1:, 1]) and np.abs(ez[0]).max() > 0


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_field_probes_match_oracle(gpu, polarization):
    """Field probes of the thin-film case: every sample within 1e-5 of the series' peak of the FP64 oracle's, times
    equal to the oracle's, the last sample equal to the engine's field at the probe's cell bit for bit, rows kept
    while the sample table grows, and graph replay bit-identical to eager stepping."""
    from tests.cases import add_thin_film_probes
    Simulation = simulation_class()
    ref, _, _ = build(Yee2D, THIN_FILM, polarization)
    sim, _, _ = build(Simulation, THIN_FILM, polarization)
    graphed, _, _ = build(Simulation, THIN_FILM, polarization, graph_steps=16)
    ref_probes, probes, graphed_probes = (add_thin_film_probes(s, polarization) for s in (ref, sim, graphed))
    done = 0
    for step in SNAPSHOT_STEPS:   # three runs: the table grows twice
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        for probe in probes:
            times, values = probe.get_data()
            assert len(values) == step and values.dtype == np.float32
            assert values[-1] == sim.engine.field(probe.field_component)[probe.cell]
    for probe, graphed_probe, ref_probe in zip(probes, graphed_probes, ref_probes):
        times, values = probe.get_data()
        ref_times, ref_values = ref_probe.get_data()
        np.testing.assert_array_equal(times, ref_times)
        np.testing.assert_array_equal(graphed_probe.get_data()[1], values)
        peak = np.max(np.abs(ref_values))
        if peak == 0.0:   # H_x (E_z) and E_x (H_z) stay exactly 0 in this y-invariant case
            assert not np.any(values), probe.cell
            continue
        error = np.max(np.abs(values - ref_values)) / peak
        print(f'{polarization} probe {probe.cell} {probe.field_component}: max err / peak {error:.3e}')
        assert error <= 1e-5, (probe.cell, error)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_dft_boxes_match_oracle(gpu, polarization):
    """The PEC case's DFT boxes (full resolution from y cell 0, where H_z's E_x wraps to the last face; downsampled
    by 3) against the FP64 oracle after 500 steps: every spectrum within 1e-5 of its peak (relative L2 <= 1e-5);
    graph replay bit-identical to eager stepping; box cells and H_z re-timing as the reference's."""
    from tests.cases import PEC2D, add_pec2d_dft_boxes, build_pec2d
    Simulation = simulation_class()
    ref, _, _ = build_pec2d(Yee2D, polarization, _oracle_boundaries, phase_dtype=np.complex128)
    sim, _, _ = build_pec2d(Simulation, polarization, _product_boundaries)
    graphed, _, _ = build_pec2d(Simulation, polarization, _product_boundaries, graph_steps=16)
    ref_boxes, boxes, graphed_boxes = (add_pec2d_dft_boxes(s) for s in (ref, sim, graphed))
    steps = PEC2D['steps'][-1]
    ref.run(n_steps=steps)
    sim.run(n_steps=137, print_every=0)   # two runs: the phase tables grow in between
    sim.run(n_steps=steps - 137, print_every=0)
    graphed.run(n_steps=steps, print_every=0)
    assert sim.model.sample_interval == ref.monitors[0].sample_interval
    for index, (box, graphed_box, ref_box) in enumerate(zip(boxes, graphed_boxes, ref_boxes)):
        for name, want in ref_box.spectra.items():
            got = getattr(box, f'{name}_dft')
            np.testing.assert_array_equal(getattr(graphed_box, f'{name}_dft'), got)
            peak_error = np.max(np.abs(got - want)) / np.max(np.abs(want))
            l2 = float(np.linalg.norm(got - want) / np.linalg.norm(want))
            print(f'{polarization} box {index} {name}: max err / peak {peak_error:.3e}, rel L2 {l2:.3e}')
            assert peak_error <= 1e-5 and l2 <= 1e-5, (index, name, peak_error, l2)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_dispersive_poles_match_oracle(gpu, polarization):
    """Drude rectangle (eps_inf background) under a PEC corner and a Lorentz disc on a dielectric, in the fused E_z
    update, against the FP64 oracle (which reproduces the reference's own run to 1e-12): fields <= 1e-5 relative L2
    at 300 and 800 steps, normalised by the field's largest norm so far (by step 800 the pulse has mostly left; the
    oracle itself run in float32 is 1.2e-5 off relative to the remainder, 2e-6 relative to the pulse), flux <= 1e-5;
    PEC cells stay exactly 0; graph replay and a batch bit-identical to eager solo stepping."""
    from tests.cases import POLES2D, build_poles2d
    from opensol_amd.batch import Batch2D
    Simulation = simulation_class()
    ref, ref_monitor = build_poles2d(Yee2D, polarization, phase_dtype=np.complex128)
    sim, monitor = build_poles2d(Simulation, polarization)
    graphed, _ = build_poles2d(Simulation, polarization, graph_steps=16)
    batched, _ = build_poles2d(Simulation, polarization)
    done = 0
    peak = {}
    for step in POLES2D['steps']:
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        for name in sim.model.components:
            peak[name] = max(peak.get(name, 0.0), float(np.linalg.norm(ref.fields[name])))
            error = float(np.linalg.norm(sim.engine.field(name) - ref.fields[name])) / peak[name]
            print(f'{polarization} poles step {step} {name}: rel L2 {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
            np.testing.assert_array_equal(graphed.engine.field(name), sim.engine.field(name))
    if polarization == 'ez':
        assert not np.any(sim.engine.field('ez')[ref.pec_mask()])
    else:
        assert not np.any(sim.engine.field('ex')[~ref.keep_x]) and not np.any(sim.engine.field('ey')[~ref.keep_y])
    flux_error = np.max(np.abs(monitor.flux() / ref_monitor.flux() - 1.0))
    print(f'poles flux rel error {flux_error:.3e}')
    assert flux_error <= 1e-5
    batched.finalize()
    batch = Batch2D([batched], graph_steps=16)
    batch.run([done])
    import torch
    torch.cuda.synchronize()
    batch.close()
    for name in sim.model.components:
        np.testing.assert_array_equal(batched.engine.field(name), sim.engine.field(name))


@pytest.mark.parametrize('case', ['ez', 'hz', 'ez_mirror'])
def test_tfsf_matches_oracle(gpu, case):
    """TFSF plane-wave boxes around a dielectric (E_z, H_z, and E_z opening onto a symmetric y_lo mirror) against the
    FP64 oracle, which reproduces the reference's own runs to 1e-12: each component's L2 error <= 1e-5 of the largest
    norm its field vector (E or H) has had at the checked steps (the pulse has mostly left the grid by step 400),
    flux behind and inside the box <= 1e-5; graph replay bit-identical to eager."""
    from tests.cases import TFSF2D, build_tfsf2d
    Simulation = simulation_class()
    ref, _, ref_monitors = build_tfsf2d(Yee2D, case, _oracle_boundaries, phase_dtype=np.complex128)
    sim, _, monitors = build_tfsf2d(Simulation, case, _product_boundaries)
    graphed, _, graphed_monitors = build_tfsf2d(Simulation, case, _product_boundaries, graph_steps=16)
    done, scales = 0, {}
    for step in TFSF2D['steps']:
        print(f'tfsf {case} to step {step}', flush=True)
        ref.run(n_steps=step - done)
        sim.run(n_steps=step - done, print_every=0)
        graphed.run(n_steps=step - done, print_every=0)
        done = step
        update_vector_scales({name: ref.fields[name] for name in sim.model.components}, scales)
        for name in sim.model.components:
            error = scaled_error(sim.engine.field(name), ref.fields[name], scales, name)
            print(f'tfsf {case} step {step} {name}: rel L2 {error:.3e}')
            assert error <= FIELD_TOL, (step, name, error)
            np.testing.assert_array_equal(graphed.engine.field(name), sim.engine.field(name))
    for monitor, ref_monitor, graphed_monitor in zip(monitors, ref_monitors, graphed_monitors):
        flux_error = np.max(np.abs(monitor.flux() / ref_monitor.flux() - 1.0))
        print(f'tfsf {case} flux rel error {flux_error:.3e}')
        assert flux_error <= 1e-5
        np.testing.assert_array_equal(graphed_monitor.flux(), monitor.flux())


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_empty_tfsf_box_leaks_nothing(gpu, polarization):
    """An empty TFSF box in vacuum on the GPU: the scattered-field region stays below 1e-5 of the incident peak
    (float32 rounding of the 2D update against the host's 1D incident line), the total-field region carries the wave."""
    Simulation = simulation_class()
    sim = Simulation(2e-6, 1.4e-6, 25e-9, npml=10, polarization=polarization)
    box = sim.add_tfsf(0.4e-6, 0.2e-6, 1.6e-6, 1.2e-6, wavelength=600e-9, bandwidth=0.5)
    sim.run(n_steps=150, print_every=0)
    field = sim.engine.field('ez' if polarization == 'ez' else 'ey')
    x0, x1, y0, y1 = box.cells
    inside = np.abs(field[x0 + 2:x1 - 2, y0 + 2:y1 - 2]).max()
    outside = max(np.abs(field[:x0 - 2]).max(), np.abs(field[x1 + 2:]).max(), np.abs(field[:, :y0 - 2]).max(),
                  np.abs(field[:, y1 + 2:]).max())
    print(f'tfsf {polarization}: inside {inside:.3e}, outside {outside:.3e}')
    assert inside > 0.5 and outside < 1e-5 * inside


@pytest.mark.parametrize('name', ['a', 'b'])
def test_hz_walls_and_mirrors_match_oracle(gpu, name):
    """H_z pec and mirror faces (even / odd H_z copies inside the step kernel, the tangential E pinned on the plane of
    the pec and antisymmetric faces, ghost cells on mirrors) against the FP64 oracle, which reproduces the reference's
    own runs to 1e-12: fields <= 1e-5 relative L2 at both steps; the copied rows exact; graph replay bit-identical.
    Case 'b' puts the y_hi wall on column 128 of 129, the first column of a tile of 128."""
    from tests.cases import HZWALLS2D, build_hzwalls2d
    Simulation = simulation_class()
    for lanes in (2, 4):
        sim = build_hzwalls2d(Simulation, name, _product_boundaries, lane_columns=lanes)
        graphed = build_hzwalls2d(Simulation, name, _product_boundaries, lane_columns=lanes, graph_steps=16)
        ref = build_hzwalls2d(Yee2D, name, _oracle_boundaries)
        done = 0
        for step in HZWALLS2D['steps']:
            print(f'H_z walls {name} lanes {lanes} to step {step}', flush=True)
            ref.run(n_steps=step - done)
            sim.run(n_steps=step - done, print_every=0)
            graphed.run(n_steps=step - done, print_every=0)
            done = step
            for field in sim.model.components:
                error = rel_l2(sim.engine.field(field), ref.fields[field])
                print(f'H_z walls {name} lanes {lanes} step {step} {field}: rel L2 {error:.3e}')
                assert error <= FIELD_TOL, (lanes, step, field, error)
                np.testing.assert_array_equal(graphed.engine.field(field), sim.engine.field(field))
        hz = sim.engine.field('hz')
        faces = HZWALLS2D['cases'][name]['faces']
        signs = {'pec': 1.0, 'antisymmetric': 1.0, 'symmetric': -1.0}
        for face, (outer, inner) in zip(('x_lo', 'x_hi', 'y_lo', 'y_hi'),
                                        ((np.s_[0], np.s_[1]), (np.s_[-1], np.s_[-2]), (np.s_[:, 0], np.s_[:, 1]),
                                         (np.s_[:, -1], np.s_[:, -2]))):
            kind = faces[('x_lo', 'x_hi', 'y_lo', 'y_hi').index(face)]
            if kind in signs:
                np.testing.assert_array_equal(hz[outer], signs[kind] * hz[inner])


WALL_SIZES = [(17, 20), (18, 21), (16, 129), (15, 257)]


@pytest.mark.parametrize('lanes', [2, 4])
@pytest.mark.parametrize('shape', WALL_SIZES, ids=[f'{a}x{b}' for a, b in WALL_SIZES])
def test_hz_wall_columns_random_fields(gpu, shape, lanes):
    """Random fields in an H_z grid bounded by pec (x_lo), symmetric (x_hi), antisymmetric (y_lo) and pec (y_hi) faces
    at widths that put column ny - 1 in the same lane as ny - 2, in the next lane, and on a tile edge (129 and 257
    columns): 9 steps against the oracle, <= 5e-6 of each field's peak."""
    Simulation = simulation_class()
    wx, wy = shape
    dx = 25e-9
    faces = ('pec', 'symmetric', 'antisymmetric', 'pec')
    sizes = ((wx - 1) * dx, (wy - 1) * dx)   # one ghost cell on each mirror face
    ref = Yee2D(*sizes, dx, npml=1, boundaries=_oracle_boundaries(faces, 1), polarization='hz')
    sim = Simulation(*sizes, dx, npml=1, boundaries=_product_boundaries(faces, 1), polarization='hz',
                     lane_columns=lanes)
    rng = np.random.default_rng(wy + lanes)
    for x0, y0, x1, y1, eps in random_rectangles(rng, *sizes, 4):
        for target in (ref, sim):
            target.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    ref.finalize()
    sim.finalize()
    assert sim.model.shape == (ref.nx, ref.ny) == shape
    for name in sim.model.components:
        values = rng.standard_normal(shape).astype(np.float32)
        ref.fields[name][...] = values
        sim.engine.set_field(name, values)
    for _ in range(9):
        ref.step()
    sim.run(n_steps=9, print_every=0)
    for name in sim.model.components:
        got, want = sim.engine.field(name), ref.fields[name]
        error = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f'H_z walls {shape} lanes {lanes} {name}: {error:.3e}')
        assert error <= 5e-6, (name, error)
