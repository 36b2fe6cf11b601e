"""3D HIP Yee step (libyee.so through the C ABI) against the FP64 3D oracle: the golden dipole cases end to end (fields
and plane-flux spectra; one with a PEC face, a pmc face and a PEC box), random initial fields over tile shapes that put tile edges, ragged tiles and wraps
everywhere, exact invariance along a periodic axis (halo copies equal their owners bit for bit), graph replay and
launch-geometry invariance (bit-identical)."""
import numpy as np
import pytest

from oracle.yee3d import Yee3D
from tests.cases import scaled_error, update_vector_scales

pytestmark = pytest.mark.gpu

FIELDS = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')
GOLDEN_WINDOW = (0.7e-6, 0.6e-6, 0.55e-6)
GOLDEN_BOXES = (((0.05e-6, 0.05e-6, 0.00e-6), (0.30e-6, 0.35e-6, 0.25e-6), 4.0),
                ((0.40e-6, 0.20e-6, 0.15e-6), (0.70e-6, 0.60e-6, 0.40e-6), 2.25),
                ((0.20e-6, 0.40e-6, 0.30e-6), (0.50e-6, 0.55e-6, 0.55e-6), 6.5))
GOLDEN_CASES = {
    'cpml': dict(faces=None, dipole=(0.36e-6, 0.29e-6, 0.27e-6), orientation='z',
                 planes=(('x', 0.55e-6), ('z', 0.1e-6)), dft_planes=(0.27e-6,),
                 flux_box=((0.2e-6, 0.15e-6, 0.12e-6), (0.5e-6, 0.45e-6, 0.42e-6))),
    'periodic_y': dict(faces={'y_lo': 'periodic', 'y_hi': 'periodic'}, dipole=(0.33e-6, 0.21e-6, 0.26e-6),
                       orientation='x', planes=(('x', 0.15e-6), ('z', 0.45e-6)), dft_planes=(0.26e-6, 0.5e-6)),
    'pec': dict(faces={'z_lo': 'pec', 'z_hi': 'pmc'}, dipole=(0.36e-6, 0.29e-6, 0.27e-6), orientation='z',
                planes=(('x', 0.2e-6), ('z', 0.45e-6)),
                pec_box=((0.45e-6, 0.10e-6, 0.30e-6), (0.60e-6, 0.35e-6, 0.45e-6)), dft_planes=(0.0, 0.27e-6),
                flux_box=((0.25e-6, 0.2e-6, 0.05e-6), (0.45e-6, 0.4e-6, 0.40e-6))),
}


def _build(target, spec, **options):
    sim = target(*GOLDEN_WINDOW, 25e-9, npml=8, boundaries=spec['faces'], **options)
    for lo, hi, eps in GOLDEN_BOXES:
        sim.add_dielectric_box(lo, hi, eps)
    if 'pec_box' in spec:
        sim.add_pec_box(*spec['pec_box'])
    sim.add_dipole(*spec['dipole'], orientation=spec['orientation'], dipole_moment=1e-12, wavelength=0.6e-6,
                   bandwidth=0.5)
    planes = [sim.add_flux_plane(n, p, wavelengths=[0.5e-6, 0.6e-6, 0.7e-6]) for n, p in spec['planes']]
    return sim, planes


@pytest.mark.parametrize('name', sorted(GOLDEN_CASES))
def test_dipole_case_matches_oracle(gpu, name):
    """Fields within 1e-5 (relative L2, normalised by the field norm while the pulse is in the box) after 120 and 300
    steps, flux spectra within 1e-5 of the FP64 oracle."""
    from opensol_amd import Simulation3D
    spec = GOLDEN_CASES[name]
    sim, planes = _build(Simulation3D, spec)
    ref, rplanes = _build(lambda *a, **k: Yee3D(*a, **k), spec)
    dft_planes = [sim.add_dft_plane(z, wavelengths=[0.5e-6, 0.6e-6, 0.7e-6]) for z in spec['dft_planes']]
    ref_dft_planes = [ref.add_dft_plane(z, wavelengths=[0.5e-6, 0.6e-6, 0.7e-6]) for z in spec['dft_planes']]
    boxes = [s.add_flux_box(*spec['flux_box'], wavelengths=[0.5e-6, 0.6e-6, 0.7e-6]) for s in (sim, ref)
             if 'flux_box' in spec]
    ref.finalize()
    sim.run(n_steps=120)
    ref.run(n_steps=120)
    norms = {}
    for f in FIELDS:
        want = ref.fields[f]
        norms[f] = np.linalg.norm(want)
        err = np.linalg.norm(sim.engine.field(f) - want) / norms[f]
        print(f'{name} step 120 {f}: {err:.2e}')
        assert err < 1e-5, f
    sim.run(n_steps=180)
    ref.run(n_steps=180)
    for f in FIELDS:
        err = np.linalg.norm(sim.engine.field(f) - ref.fields[f]) / norms[f]
        print(f'{name} step 300 {f}: {err:.2e}')
        assert err < 1e-5, f
    for plane, rplane in zip(planes, rplanes):
        got, want = plane.flux(), rplane.flux()
        rel = np.max(np.abs(got - want) / np.abs(want))
        print(f'{name} flux {got} oracle {want} rel {rel:.2e}')
        assert rel < 1e-5
    if boxes:   # flux box: six bounded planes, net outward flux within 1e-5 of the oracle's
        got, want = boxes[0].flux(), boxes[1].flux()
        rel = np.max(np.abs(got - want) / np.abs(want))
        print(f'{name} flux box {got} oracle {want} rel {rel:.2e}')
        assert rel < 1e-5
    # DFT planes of E at the cell centres: relative L2 <= 1e-5 per component (exactly 0 where the oracle's is, the
    # tangential E of a PEC face)
    for plane, rplane in zip(dft_planes, ref_dft_planes):
        assert plane.k_grid == rplane.k
        for component in ('ex', 'ey', 'ez'):
            got, want = plane.spectra[component], rplane.spectra[component]
            if not np.any(want):
                assert not np.any(got), (plane.k_grid, component)
                continue
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            print(f'{name} DFT plane k {plane.k_grid} {component}: rel L2 {err:.2e}')
            assert err < 1e-5, (plane.k_grid, component, err)
        np.testing.assert_allclose(plane.intensity(1), rplane.intensity(1), rtol=2e-5,
                                   atol=2e-5 * np.max(rplane.intensity(1)))


# (window cells x, y, z, npml, faces, planes_per_tile)
CASES = [
    ((8, 9, 20), 4, None, 3),                   # nz below one tile, ragged y tile
    ((10, 7, 62), 3, None, 1),                  # exactly one z tile
    ((9, 8, 63), 3, {'z_lo': 'periodic', 'z_hi': 'periodic'}, 4),   # one cell over a tile, z wrap
    ((6, 11, 130), 2, {'x_lo': 'periodic', 'x_hi': 'periodic'}, 2),
    ((7, 6, 70), 3, {'y_lo': 'periodic', 'y_hi': 'periodic'}, 0),
    ((12, 10, 40), 0, {f: 'periodic' for f in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}, 5),
]


def _random_pair(case, seed, **options):
    from opensol_amd import Simulation3D
    cells, npml, faces, planes = case
    dx = 25e-9
    size = tuple(c * dx for c in cells)
    sim = Simulation3D(*size, dx, npml=npml, boundaries=faces, planes_per_tile=planes, **options)
    ref = Yee3D(*size, dx, npml=npml, boundaries=faces)
    rng = np.random.default_rng(seed)
    for _ in range(6):
        lo = rng.uniform(-0.2, 1.0, 3) * np.array(size)
        hi = lo + rng.uniform(0.1, 0.6, 3) * np.array(size)
        eps, sig = rng.uniform(1.0, 9.0), (rng.uniform(0, 2e4) if rng.random() < 0.3 else 0.0)
        sim.add_dielectric_box(tuple(lo), tuple(hi), eps, sig)
        ref.add_dielectric_box(tuple(lo), tuple(hi), eps, sig)
    sim.finalize()
    ref.finalize()
    for f in FIELDS:
        values = rng.standard_normal(ref.shape) * (1.0 if f[0] == 'e' else 1.0 / 376.7)
        ref.fields[f][...] = values
        sim.engine.set_field(f, values)
    return sim, ref


@pytest.mark.parametrize('case', CASES, ids=[str(c[0]) + str(c[2] and sorted(c[2])[0]) for c in CASES])
def test_random_fields_match_oracle(gpu, case):
    sim, ref = _random_pair(case, seed=11)
    sim.run(n_steps=12)
    ref.run(n_steps=12)
    for f in FIELDS:
        want = ref.fields[f]
        err = np.max(np.abs(sim.engine.field(f) - want)) / np.max(np.abs(want))
        assert err < 5e-6, (f, err)


@pytest.mark.parametrize('axis', [0, 1, 2])
def test_invariance_along_periodic_axis_is_exact(gpu, axis):
    """Fields and materials invariant along a periodic axis stay exactly invariant: every halo copy (plane i0-1, row
    j0-1, lanes 0 and 63) rounds as its owner does."""
    from opensol_amd import Simulation3D
    dx = 25e-9
    cells = [14, 13, 70]
    cells[axis] = 5
    faces = {f'{"xyz"[axis]}_lo': 'periodic', f'{"xyz"[axis]}_hi': 'periodic'}
    size = [c * dx for c in cells]
    sim = Simulation3D(*size, dx, npml=3, boundaries=faces, planes_per_tile=2)
    rng = np.random.default_rng(5)
    for _ in range(5):
        lo = rng.uniform(0, 1.0, 3) * np.array(size)
        hi = lo + rng.uniform(0.1, 0.5, 3) * np.array(size)
        lo[axis], hi[axis] = -1.0, 1.0
        sim.add_dielectric_box(tuple(lo), tuple(hi), rng.uniform(1, 8))
    sim.finalize()
    shape = sim.model.shape
    for f in FIELDS:
        line = rng.standard_normal([n if a != axis else 1 for a, n in enumerate(shape)])
        sim.engine.set_field(f, np.broadcast_to(line, shape))
    sim.run(n_steps=10)
    for f in FIELDS:
        values = sim.engine.field(f)
        first = np.take(values, [0], axis=axis)
        assert np.array_equal(values, np.broadcast_to(first, values.shape)), f


def test_graph_replay_and_launch_geometry_are_bit_identical(gpu):
    case = CASES[4]
    runs = []
    for options, planes in (({}, 0), ({'graph_steps': 7}, 0), ({}, 1), ({}, 3)):
        sim, _ = _random_pair(case[:3] + (planes,), seed=3, **options)
        sim.run(n_steps=23)
        runs.append({f: sim.engine.field(f) for f in FIELDS})
    for other in runs[1:]:
        for f in FIELDS:
            assert np.array_equal(runs[0][f], other[f]), f


def test_field_probes_match_oracle(gpu):
    """Field probes of the PEC golden case (every component, at the dipole's cell, inside the PEC box, in the pec
    wall layer and at the window's corners): samples within 1e-5 of the FP64 oracle's, relative to the component's
    peak over the grid, and within 5e-5 in the dipole's own cell, where H is a float32 difference of the large E
    there (the oracle itself run in float32 is off by 1.5e-5 in that cell); times equal, the last sample equal to
    the engine's field bit for bit, graph replay bit-identical."""
    from opensol_amd import Simulation3D
    spec = GOLDEN_CASES['pec']
    spots = [spec['dipole'], (0.5e-6, 0.2e-6, 0.4e-6), (0.3e-6, 0.3e-6, 0.0), (0.0, 0.0, 0.0), GOLDEN_WINDOW,
             (0.1e-6, 0.5e-6, 0.3e-6)]
    runs = []
    for target, options in ((lambda *a, **k: Yee3D(*a, **k), {}), (Simulation3D, {}),
                            (Simulation3D, dict(graph_steps=10))):
        sim, _ = _build(target, spec, **options)
        probes = [sim.add_field_probe(*spot, component=f) for spot in spots for f in FIELDS]
        runs.append((sim, probes))
    component_peak = dict.fromkeys(FIELDS, 0.0)
    for chunk in (37, 83, 80):
        for sim, _ in runs:
            sim.run(n_steps=chunk)
        for f in FIELDS:
            component_peak[f] = max(component_peak[f], float(np.max(np.abs(runs[0][0].fields[f]))))
        sim, probes = runs[1]
        for probe in probes:
            values = probe.get_data()[1]
            assert len(values) == sim.n_step
            assert values[-1] == sim.engine.field(probe.field_component)[probe.cell]
    (ref, ref_probes), (sim, probes), (graphed, graphed_probes) = runs
    for probe, graphed_probe, ref_probe in zip(probes, graphed_probes, ref_probes):
        times, values = probe.get_data()
        ref_times, ref_values = ref_probe.get_data()
        np.testing.assert_array_equal(times, ref_times)
        np.testing.assert_array_equal(graphed_probe.get_data()[1], values)
        peak = np.max(np.abs(ref_values))
        if peak == 0.0:   # pinned PEC faces
            assert not np.any(values), (probe.cell, probe.field_component)
            continue
        error = np.max(np.abs(values - ref_values)) / component_peak[probe.field_component]
        print(f'probe {probe.cell} {probe.field_component}: max err / field peak {error:.3e}')
        tol = 5e-5 if probe.cell == sim.model.sources[0].cell else 1e-5
        assert error <= tol, (probe.cell, probe.field_component, error)


def test_dispersive_poles_match_oracle(gpu):
    """Drude and Lorentz boxes around a dipole (three face records per pole) against the FP64 oracle (which
    reproduces the reference's float32 run to 1.4e-5): fields within 1e-5 relative L2 after 120 and 250 steps,
    normalised by the field's largest norm so far (the pulse leaves the box; the remainder carries the float32
    round-off of the whole run), flux within 1e-5; graph replay bit-identical to eager stepping."""
    from opensol_amd import Simulation3D
    from tests.cases import build_poles3d
    sim, plane = build_poles3d(Simulation3D)
    graphed, _ = build_poles3d(Simulation3D, graph_steps=10)
    ref, ref_plane = build_poles3d(lambda *a, **k: Yee3D(*a, **k))
    norms = {}
    for steps in (120, 130):
        for s in (sim, graphed, ref):
            s.run(n_steps=steps)
        for f in FIELDS:
            norms[f] = max(norms.get(f, 0.0), float(np.linalg.norm(ref.fields[f])))
            err = np.linalg.norm(sim.engine.field(f) - ref.fields[f]) / norms[f]
            print(f'poles step {sim.engine.n_step} {f}: {err:.2e}')
            assert err < 1e-5, f
            np.testing.assert_array_equal(graphed.engine.field(f), sim.engine.field(f))
    rel = np.max(np.abs(plane.flux() / ref_plane.flux() - 1.0))
    print(f'poles flux rel {rel:.2e}')
    assert rel < 1e-5


def test_lossy_case_matches_oracle(gpu):
    """Conductive boxes (sigma_e 2e4 and 3e5 S/m; the float32 face means of eps and sigma in the kernel), a Ricker z
    dipole and a cw x dipole against the FP64 oracle (which matches the reference's float32 run of the same case):
    each component within 1e-5 of the largest norm of its field vector (E or H) after 120 and 250 steps, flux within
    1e-5 (sampled every step, since a cw source has no spectral limit); graph replay bit-identical to eager."""
    from opensol_amd import Simulation3D
    from tests.cases import build_lossy3d
    sim, _, plane = build_lossy3d(Simulation3D)
    graphed, _, _ = build_lossy3d(Simulation3D, graph_steps=10)
    ref, _, ref_plane = build_lossy3d(lambda *a, **k: Yee3D(*a, **k))
    scales = {}
    for steps in (120, 130):
        print(f'lossy 3d: {steps} more steps', flush=True)
        for s in (sim, graphed, ref):
            s.run(n_steps=steps)
        update_vector_scales({f: ref.fields[f] for f in FIELDS}, scales)
        for f in FIELDS:
            err = scaled_error(sim.engine.field(f), ref.fields[f], scales, f)
            print(f'lossy step {sim.engine.n_step} {f}: {err:.2e}')
            assert err < 1e-5, f
            np.testing.assert_array_equal(graphed.engine.field(f), sim.engine.field(f))
    assert sim.model.sample_interval == 1
    rel = np.max(np.abs(plane.flux() / ref_plane.flux() - 1.0))
    print(f'lossy 3d flux rel {rel:.2e}')
    assert rel < 1e-5


@pytest.mark.parametrize('name', ['a', 'b'])
def test_mirror_faces_match_oracle(gpu, name):
    """3D mirror faces (symmetric: tangential H and normal E on the plane pinned to 0 inside the step; antisymmetric:
    the ghost's tangential E = minus its inner neighbour's after the step) against the FP64 oracle, which matches the
    reference's float32 run: each component within 1e-5 of the largest norm of its field vector (E or H) after 100
    and 200 steps, flux within 1e-5; the pinned rows exactly 0 / exact negated copies; graph replay bit-identical to
    eager."""
    from opensol_amd import Simulation3D
    from tests.cases import build_mirrors3d
    sim, _, plane = build_mirrors3d(Simulation3D, name)
    graphed, _, _ = build_mirrors3d(Simulation3D, name, graph_steps=10)
    ref, _, ref_plane = build_mirrors3d(lambda *a, **k: Yee3D(*a, **k), name)
    scales = {}
    for steps in (100, 100):
        print(f'mirrors {name}: {steps} more steps', flush=True)
        for s in (sim, graphed, ref):
            s.run(n_steps=steps)
        update_vector_scales({f: ref.fields[f] for f in FIELDS}, scales)
        for f in FIELDS:
            err = scaled_error(sim.engine.field(f), ref.fields[f], scales, f)
            print(f'mirrors {name} step {sim.engine.n_step} {f}: {err:.2e}')
            assert err < 1e-5, f
            np.testing.assert_array_equal(graphed.engine.field(f), sim.engine.field(f))
    rel = np.max(np.abs(plane.flux() / ref_plane.flux() - 1.0))
    print(f'mirrors {name} flux rel {rel:.2e}')
    assert rel < 1e-5
    for axis, kind, ghost, inner, face in ref.bc.mirror_planes(ref.windows):
        others = [a for a in 'xyz' if a != 'xyz'[axis]]
        if kind == 'symmetric':
            assert not np.any(np.take(sim.engine.field('e' + 'xyz'[axis]), face, axis=axis))
            for o in others:
                assert not np.any(np.take(sim.engine.field('h' + o), face, axis=axis))
        else:
            for o in others:
                values = sim.engine.field('e' + o)
                np.testing.assert_array_equal(np.take(values, ghost, axis=axis), -np.take(values, inner, axis=axis))
