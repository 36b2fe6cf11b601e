"""Host setup of the product (geometry, rectangle painting, CPML tables, curl spacings, coefficient tables, sheet
calibration, waveform and DFT phase tables, clock) against the oracle, on the CPU."""
import numpy as np
import pytest

from opensol_amd.fdtd2d import Model2D, clock_times, steps_for_duration
from opensol_amd.materials import semi_implicit
from oracle.yee2d import EPS0, MU0, Yee2D, steps_for_duration as oracle_steps
from tests.cases import THIN_FILM, build, random_rectangles


@pytest.mark.parametrize('boundaries', [None, 'periodic_y', 'periodic_x', 'periodic'])
def test_geometry_matches_oracle(boundaries):
    for size_x, size_y, dx, npml in ((6e-6, 3e-6, 25e-9, 15), (1.7e-6, 2.3e-6, 20e-9, 7), (163.04e-6, 4e-6, 20e-9, 20)):
        model = Model2D(size_x, size_y, dx, npml=npml, boundaries=boundaries)
        ref = Yee2D(size_x, size_y, dx, npml=npml, boundaries=boundaries)
        assert model.shape == (ref.nx, ref.ny)
        assert model.dt == ref.dt and model.dt.dtype == np.float32
        for axis in (0, 1):
            np.testing.assert_array_equal(model.domain.cell_centers()[axis], ref.geo.centres[axis])
        for x in (0.0, 0.3 * size_x, size_x, 1e-9, size_x * 0.5 + 0.5 * dx):
            assert model.domain.pos_to_cell_axis(0, x) == ref.geo.cell(0, x)


def test_rectangles_paint_like_float32_masks():
    rng = np.random.default_rng(7)
    size_x, size_y, dx = 4e-6, 3e-6, 25e-9
    model = Model2D(size_x, size_y, dx, npml=9)
    ref = Yee2D(size_x, size_y, dx, npml=9)
    for x0, y0, x1, y1, eps in random_rectangles(rng, size_x, size_y, 40, side_range=(0.01, 0.6)):
        model.add_dielectric_rectangle(x0, y0, x1, y1, eps)
        ref.add_dielectric_rectangle(x0, y0, x1, y1, eps)
    # cell-aligned edges (the float32 centre comparison decides them)
    model.add_dielectric_rectangle(1e-6, 1e-6, 2e-6, 2e-6, 3.0)
    ref.add_dielectric_rectangle(1e-6, 1e-6, 2e-6, 2e-6, 3.0)
    np.testing.assert_array_equal(model.materials.property_array(0), ref.materials['eps_r'])


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_cpml_tables_and_spacings_match_oracle(polarization):
    model = Model2D(2e-6, 1.5e-6, 25e-9, npml=11, polarization=polarization)
    ref = Yee2D(2e-6, 1.5e-6, 25e-9, npml=11, polarization=polarization)
    for a, b in ((model.inv_dx_e, ref.inv_dx_e), (model.inv_dx_h, ref.inv_dx_h), (model.inv_dy_e, ref.inv_dy_e),
                 (model.inv_dy_h, ref.inv_dy_h)):
        np.testing.assert_allclose(a, b, rtol=6e-8)
    for side in ref.sides:
        tables = model.cpml[side.axis]
        base = 0 if side.side == 'lo' else ref.bc.npml
        for kind in 'eh':
            rs = side.rows[kind]
            rows = rs['rows']
            psi_rows = (tables.row_e if kind == 'e' else tables.row_h)[rows]
            np.testing.assert_array_equal(psi_rows, np.arange(base, base + ref.bc.npml))
            b = (tables.b_e if kind == 'e' else tables.b_h)[psi_rows]
            g = (tables.g_e if kind == 'e' else tables.g_h)[psi_rows]
            np.testing.assert_allclose(b, rs['b'].ravel(), rtol=6e-8)
            np.testing.assert_allclose(g, (rs['a'] * rs['inv']).ravel(), rtol=6e-8, atol=1e-30)
        outside = np.ones(tables.n, bool)
        outside[side.rows['e']['rows']] = False
    # rows outside every layer carry -1
    assert np.all(model.cpml['x'].row_e[[0, model.nx - 1]] == -1)
    assert np.all(model.cpml['x'].row_h[[model.nx - 1]] == -1)


def test_coefficient_tables_match_oracle_in_float64():
    model, _, _ = build(Model2D, THIN_FILM, 'ez', True)
    ref, _, _ = build(Yee2D, THIN_FILM, 'ez', True)
    model.finalize()
    ref.finalize()
    table = model.materials.table_ez(model.dt)[model.materials.ids]
    for column, want in ((0, ref.ca), (1, ref.cb), (2, ref.da), (3, ref.db)):
        np.testing.assert_allclose(table[..., column], want, rtol=6e-8)
    # H_z: same-id face coefficients and the mean rule across ids
    hz, _, _ = build(Model2D, THIN_FILM, 'hz', True)
    ref_hz, _, _ = build(Yee2D, THIN_FILM, 'hz', True)
    ref_hz.finalize()
    table, table2 = hz.materials.tables_hz(hz.dt)
    ids = hz.materials.ids
    same_x = np.zeros(ids.shape, bool)
    same_x[:, :-1] = ids[:, :-1] == ids[:, 1:]
    np.testing.assert_allclose(table2[ids, 1][same_x], ref_hz.cb_x[same_x], rtol=6e-8)
    values = hz.materials.values()
    left, right = values[ids[:-1, :], 1], values[ids[1:, :], 1]
    _, cb_mean = semi_implicit(0.5 * left + 0.5 * right, 0.0 * left, float(hz.dt), EPS0)
    np.testing.assert_allclose(cb_mean, ref_hz.cb_y[:-1], rtol=1e-12)
    da, db = semi_implicit(np.ones(1), np.zeros(1), float(hz.dt), MU0)
    np.testing.assert_allclose(table[0, :2], [da[0], db[0]], rtol=6e-8)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_sheet_calibration_and_tables_match_oracle(polarization):
    spec = dict(THIN_FILM, monitor=dict(axis='x', position=1.0e-6, span=(0.5e-6, 5e-6), wavelengths=[550e-9]))
    model, src, mon = build(Model2D, spec, polarization, True)
    ref, rsrc, rmon = build(Yee2D, spec, polarization, True, phase_dtype=np.complex128)
    model.finalize()
    ref.finalize()
    np.testing.assert_array_equal(src.profile, rsrc.profile)
    assert (src.position, src.span) == (rsrc.position, rsrc.span)
    assert (mon.position, mon.span, mon.normal) == (rmon.position, rmon.span, rmon.normal)
    assert model.sample_interval == rmon.sample_interval
    # per-step source values and the sample times come from the same float64 clock
    times = clock_times(model.dt, 400)
    half = times[:-1] + 0.5 * float(model.dt)
    np.testing.assert_array_equal(src.wave_table(half), [rsrc.value(t) for t in half])
    m = model.sample_interval
    sample_times = times[m * np.arange(1, 400 // m + 1)]
    table = mon.phase_table(sample_times)
    phases = np.array([rmon.phases(t) for t in sample_times])
    np.testing.assert_allclose(table[..., 0] + 1j * table[..., 1], phases, rtol=0, atol=1e-30)


def test_clock_and_duration_steps_match_runner():
    dt = np.float32(5.837669677904363e-17)
    t = 0.0
    for _ in range(3000):
        t = t + float(dt)
    assert clock_times(dt, 3000)[-1] == t
    for duration in (80e-15, 1e-15, 0.0, 3.3e-14):
        assert steps_for_duration(0.0, dt, duration) == oracle_steps(dt, duration)
    assert steps_for_duration(0.0, dt, 80e-15) == 1370


def test_model_validation_errors():
    with pytest.raises(ValueError):
        Model2D(1e-6, 1e-6, 25e-9, polarization='tm')
    with pytest.raises(ValueError):
        Model2D(1e-6, 1e-6, 25e-9, boundaries='mirror')
    model = Model2D(1e-6, 1e-6, 25e-9, npml=5)
    with pytest.raises(ValueError):
        model.add_line_source(axis='z', position=0.5e-6, wavelength=500e-9)
    with pytest.raises(ValueError):
        model.add_line_source(axis='y', position=2e-6, wavelength=500e-9)   # outside the window
    with pytest.raises(ValueError):
        model.add_flux_monitor(axis='y', position=0.5e-6, span=None, wavelengths=[500e-9])
    with pytest.raises(ValueError):
        model.add_line_source(axis='y', position=0.5e-6)   # neither wavelength nor frequency
    with pytest.raises(ValueError):
        Model2D(1e-6, 1e-6, 25e-9, npml=5, courant=1.5)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_pec_cells_and_source_profile_match_the_oracle(polarization):
    """The PEC cells (region + E_z 'pec' face) and the calibrated source profile (zero on PEC cells / faces) equal the
    oracle's, which reproduces the reference's run."""
    from opensol_amd.boundaries import BoundaryConfig
    from oracle.yee2d import FACES, Boundaries, Yee2D
    from tests.cases import build_pec2d

    def product_bc(faces, npml):
        bc = BoundaryConfig(*faces)
        bc.cpml_npml = npml
        return bc

    model, source, _ = build_pec2d(Model2D, polarization, product_bc)
    ref, ref_source, _ = build_pec2d(Yee2D, polarization, lambda f, n: Boundaries(dict(zip(FACES, f)), npml=n))
    model.finalize()
    ref.finalize()
    np.testing.assert_array_equal(model.materials.ids == model.materials.pec_id(), ref.pec_mask())
    np.testing.assert_allclose(source.profile, ref_source.profile.astype(np.float32), rtol=1e-6, atol=0)
    assert np.all(source.profile[ref_source.profile == 0] == 0)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_point_current_records_match_the_reference(polarization):
    """Point currents: each record's cell or face, field and float32 weight equal the reference's own (the fixture of
    its FP64 run), the H_z magnetic current's wave table samples s(t_n) and the others s(t_n + dt/2); the source
    table is capped and the polarisation checked."""
    import os
    from opensol_amd.boundaries import BoundaryConfig
    from opensol_amd.fdtd2d import clock_times
    from tests.cases import build_point2d

    def product_bc(faces, npml):
        bc = BoundaryConfig(*faces)
        bc.cpml_npml = npml
        return bc

    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'fixtures',
                                  f'point2d_{polarization}.npz'))
    model, sources = build_point2d(Model2D, polarization, product_bc)
    model.finalize()
    for index, source in enumerate(sources):
        records = source if isinstance(source, list) else [source]
        weights = np.concatenate([record.profile for record in records]).astype(np.float64)
        np.testing.assert_array_equal(weights, golden[f'source{index}_weights'])
        if isinstance(source, list):
            faces = [[0 if r.component == 'ex' else 1, r.span[0] if r.axis == 'x' else r.position,
                      r.position if r.axis == 'x' else r.span[0]] for r in records]
            assert faces == golden[f'source{index}_faces'].tolist()
        else:
            assert [source.span[0], source.position] == golden[f'source{index}_cell'].tolist()
    times = clock_times(model.dt, 8)
    for source in model.sources:
        at = times[:-1] if source.field == 'h' else times[:-1] + 0.5 * float(model.dt)
        expected = np.array([source.amplitude * np.float32(source.temporal(t)) for t in at], dtype=np.float32)
        table = source.wave_table(at)
        np.testing.assert_array_equal(table, expected)
    assert [s.field for s in model.sources] == (['e'] * 4 if polarization == 'ez' else ['h'] + ['e'] * 7)
    beam = model.sources[-1]   # the Gaussian beam
    assert list(beam.span) == golden['beam_span'].tolist()
    np.testing.assert_allclose(beam.profile, golden['beam_profile'], rtol=1e-6, atol=0)
    wrong = 'add_magnetic_line_current' if polarization == 'ez' else 'add_point_source'
    with pytest.raises(ValueError, match='needs polarization'):
        getattr(model, wrong)(0.5e-6, 0.5e-6, wavelength=600e-9)
    if polarization == 'hz':
        model.add_electric_dipole(0.3e-6, 0.3e-6, wavelength=600e-9, orientation=45.0)   # 8 + 4 records > 8
        with pytest.raises(ValueError, match='at most 8 source records'):
            model.finalize()


def test_wall_faces_are_padded_like_the_oracle():
    """Wall and mirror faces: mirror faces pad one ghost cell and walls none (the grid equals the oracle's, both
    polarisations); the descriptor's wall factors; an E_z flux line touching a pmc row is refused; an H_z 'pec' face
    is a wall copy, not PEC cells."""
    from opensol_amd.boundaries import BoundaryConfig
    from opensol_amd.fdtd2d import HZ_WALL_SIGNS
    from oracle.yee2d import FACES, Boundaries
    for kind in ('pec', 'symmetric', 'antisymmetric', 'pmc'):
        bc = BoundaryConfig(kind, 'cpml', 'cpml', 'cpml')
        bc.cpml_npml = 5
        hz = Model2D(1e-6, 1e-6, 25e-9, npml=5, boundaries=bc, polarization='hz')
        ref = Yee2D(1e-6, 1e-6, 25e-9, npml=5, polarization='hz',
                    boundaries=Boundaries(dict(zip(FACES, (kind, 'cpml', 'cpml', 'cpml'))), npml=5))
        assert hz.shape == (ref.nx, ref.ny) and not hz.pec_mask().any()
    assert HZ_WALL_SIGNS == {'pmc': 2, 'pec': 1, 'antisymmetric': 1, 'symmetric': -1}
    faces = ('symmetric', 'pmc', 'antisymmetric', 'cpml')
    bc = BoundaryConfig(*faces)
    bc.cpml_npml = 10
    model = Model2D(1.2e-6, 1e-6, 25e-9, npml=10, boundaries=bc)
    ref = Yee2D(1.2e-6, 1e-6, 25e-9, npml=10, boundaries=Boundaries(dict(zip(FACES, faces)), npml=10))
    assert model.shape == (ref.nx, ref.ny) == (49, 51)
    from opensol_amd.fdtd2d import WALL_SIGNS, wall_rows
    assert [WALL_SIGNS.get(k, 0) for k in faces] == [1, 1, -1, 0]
    assert wall_rows(model) == {0: {0, 48}, 1: {0}}
    from opensol_amd.fdtd3d import Model3D   # 3D mirrors pad one ghost cell too
    assert Model3D(0.5e-6, 0.5e-6, 0.5e-6, 25e-9, npml=4, boundaries={'x_lo': 'symmetric'}).shape == (25, 28, 28)
    model.add_flux_monitor('y', 1.2e-6 - 1e-9, (0.2e-6, 0.8e-6), wavelengths=[600e-9])   # x cell 48: the pmc row
    with pytest.raises(ValueError, match='pmc wall row'):
        model.finalize()


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_field_probe_cells_match_the_oracle(polarization):
    """Probe placement (the cell containing the point, window edges included) and the component rule: None is the
    out-of-plane field; only the polarization's own components are accepted (physics/em/placement.py:16, 29-45)."""
    from tests.cases import PROBE_COMPONENTS, THIN_FILM_PROBES, add_thin_film_probes
    model, _, _ = build(Model2D, THIN_FILM, polarization)
    ref, _, _ = build(Yee2D, THIN_FILM, polarization)
    got, want = add_thin_film_probes(model, polarization), add_thin_film_probes(ref, polarization)
    assert [p.cell for p in got] == [p.cell for p in want]
    assert [p.field_component for p in got] == [p.component for p in want]
    assert model.add_field_probe(1e-6, 1e-6).field_component == PROBE_COMPONENTS[polarization][0]
    other = PROBE_COMPONENTS['hz' if polarization == 'ez' else 'ez'][0]
    with pytest.raises(ValueError):
        model.add_field_probe(1e-6, 1e-6, other)
    with pytest.raises(ValueError):
        model.add_field_probe(THIN_FILM['size_x'] + 1e-7, 1e-6)
    assert len(THIN_FILM_PROBES) + 1 == len(model.probes)


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_dft_monitor_cells_match_the_oracle(polarization):
    """DFT box placement (the containing cells of both corners, in either order), the downsampled sample cells
    (physics/em/monitors.py:152-155) and, for H_z, the lower-face partners of E_x and E_y (periodic wrap at index 0)."""
    from tests.cases import PEC2D, add_pec2d_dft_boxes, build_pec2d
    from oracle.yee2d import Boundaries, FACES
    from opensol_amd.boundaries import BoundaryConfig

    def product_boundaries(faces, npml):
        bc = BoundaryConfig(*faces)
        bc.cpml_npml = npml
        return bc

    model, _, _ = build_pec2d(Model2D, polarization, product_boundaries)
    ref, _, _ = build_pec2d(Yee2D, polarization, lambda faces, npml: Boundaries(dict(zip(FACES, faces)), npml=npml))
    boxes, ref_boxes = add_pec2d_dft_boxes(model), add_pec2d_dft_boxes(ref)
    boxes.append(model.add_dft_monitor(1.5e-6, 1e-6, 0.4e-6, 0.0, frequencies=[4e14, 5e14], downsample=7))
    ref_boxes.append(ref.add_dft_monitor(1.5e-6, 1e-6, 0.4e-6, 0.0, frequencies=[4e14, 5e14], downsample=7))
    for box, ref_box in zip(boxes, ref_boxes):
        assert (box.x_samples, box.y_samples) == (ref_box.x_samples, ref_box.y_samples)
        assert box.accumulator_shape()[2:4] == (ref_box.xs.size, ref_box.ys.size)
        assert box.names == tuple(ref_box.spectra)
        if polarization == 'hz':
            # E_x pairs along y, E_y along x; index 0 wraps on the periodic y axis only
            assert box.pairs == ((-1, False), (1, True), (0, False))
    with pytest.raises(ValueError):
        model.add_dft_monitor(0.1e-6, 0.1e-6, 2.5e-6, 0.5e-6, wavelengths=[5e-7])
    with pytest.raises(ValueError):
        model.add_dft_monitor(0.1e-6, 0.1e-6, 0.5e-6, 0.5e-6, wavelengths=[5e-7], downsample=0)


def test_pole_masks_background_and_drive_match_the_oracle():
    """The pole case: Drude rectangle and Lorentz disc masks (float32 centres, as the reference's rect_mask and
    circle_mask), the eps_inf background under the Drude cells, the PEC cells over the Drude corner, and the Courant
    bound with the poles' terms."""
    from tests.cases import build_poles2d
    from opensol_amd import DrudePole
    from opensol_amd.dispersion import pole_step_term
    model, _ = build_poles2d(Model2D)
    ref, _ = build_poles2d(Yee2D)
    model.finalize()
    ref.finalize()
    for (pole, mask), entry in zip(model.poles, ref.poles):
        np.testing.assert_array_equal(mask, entry['mask'])
        coefficients = pole.ade_coefficients(float(model.dt))
        assert tuple(float(c) for c in coefficients) == tuple(float(c) for c in entry['coefficients'])
    # H_z: the drive of each face set is the mask's face average times the face keep, as the oracle's
    model_hz, _ = build_poles2d(Model2D, 'hz')
    ref_hz, _ = build_poles2d(Yee2D, 'hz')
    model_hz.finalize()
    ref_hz.finalize()
    weights = model_hz.pole_weights()
    assert [c for _, c, _ in weights] == [1, 2, 1, 2]
    for (pole, component, weight), entry in zip(weights, [e for e in ref_hz.poles for _ in (0, 1)]):
        drive = pole.ade_coefficients(float(model_hz.dt))[-1]
        want = entry['drive']['ex' if component == 1 else 'ey']
        np.testing.assert_allclose(weight * np.float64(drive), want, rtol=1e-7, atol=0)
    pec = model.materials.pec_flags()[model.materials.ids]
    np.testing.assert_array_equal(pec, ref.pec_mask())
    np.testing.assert_array_equal(model.materials.property_array(0)[~pec], ref.materials['eps_r'][~pec])
    assert pole_step_term(DrudePole(5e15, 1e14), 1e-17) == (float(np.float32(5e15)) * 1e-17) ** 2 / 4
    unstable = Model2D(1e-6, 1e-6, 25e-9, npml=5)
    unstable.add_drude_rectangle(0.2e-6, 0.2e-6, 0.6e-6, 0.6e-6, DrudePole(1.37e16, 1e14, 1.0))
    with pytest.raises(ValueError):
        unstable.check_courant()
    hz = Model2D(1e-6, 1e-6, 25e-9, npml=5, polarization='hz')
    for _ in range(4):
        hz.add_drude_rectangle(0, 0, 1e-7, 1e-7, DrudePole(1e15, 1e13, 2.0))
    with pytest.raises(ValueError):   # four H_z poles fill the eight face records
        hz.add_drude_rectangle(0, 0, 1e-7, 1e-7, DrudePole(1e15, 1e13, 2.0))


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_tfsf_incident_tables_match_the_oracle_line(polarization):
    """The product's host run of the TFSF incident line (float32) equals the oracle's float32 line step by step, bit
    for bit: per step the box columns' incident values and the two face values the box edges read; same cells and
    signed side coefficients."""
    from opensol_amd.fdtd2d import clock_times
    model = Model2D(2e-6, 1.4e-6, 25e-9, npml=10, polarization=polarization)
    box = model.add_tfsf(0.4e-6, 0.2e-6, 1.6e-6, 1.2e-6, wavelength=600e-9, bandwidth=0.5)
    ref = Yee2D(2e-6, 1.4e-6, 25e-9, npml=10, polarization=polarization, dtype=np.float32)
    ref_box = ref.add_tfsf(0.4e-6, 0.2e-6, 1.6e-6, 1.2e-6, wavelength=600e-9, bandwidth=0.5)
    assert box.cells == ref_box.cells
    signs = (-1, 1, 1, -1) if polarization == 'ez' else (1, -1)
    np.testing.assert_array_equal(np.abs(box.coef_h[:len(ref_box.coef_h)]), ref_box.coef_h)
    assert [np.sign(c) for c in box.coef_h[:len(signs)]] == list(signs)
    steps = 260
    times = clock_times(model.dt, steps)
    line, ends = box.tables(times[:-1] + 0.5 * float(model.dt))
    a0, length = ref_box.aux_x0, ref_box.x1 - ref_box.x0
    ref.finalize()
    for n in range(steps):
        if polarization == 'ez':
            np.testing.assert_array_equal(line[n], ref_box.aux_e[a0:a0 + length])
        else:
            np.testing.assert_array_equal(ends[n], ref_box.aux_e[[a0 - 1, a0 + length - 1]])
        ref.step()
        if polarization == 'ez':
            np.testing.assert_array_equal(ends[n], ref_box.aux_h[[a0 - 1, a0 + length - 1]])
        else:
            np.testing.assert_array_equal(line[n], ref_box.aux_h[a0:a0 + length])
    assert np.abs(line).max() > 0.5 if polarization == 'ez' else np.abs(ends).max() > 0.5
