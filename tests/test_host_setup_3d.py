"""Host setup of the 3D product (geometry, box painting, CPML tables, curl spacings, coefficient table, dipole
placement and calibration, flux-plane layout, sample interval) against the 3D oracle, on the CPU."""
import numpy as np
import pytest

from opensol_amd.fdtd3d import Model3D
from opensol_amd.materials import semi_implicit
from opensol_amd.constants import EPS0
from oracle.yee3d import Yee3D

FACES = ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')
BOXES = (((0.05e-6, 0.05e-6, 0.0), (0.30e-6, 0.35e-6, 0.25e-6), 4.0),
         ((0.40e-6, 0.20e-6, 0.15e-6), (0.70e-6, 0.60e-6, 0.40e-6), 2.25),
         ((-1e-6, 0.40e-6, 0.30e-6), (0.50e-6, 0.55e-6, 2e-6), 6.5))


def _pair(faces, npml=6):
    size, dx = (0.7e-6, 0.6e-6, 0.55e-6), 25e-9
    model = Model3D(*size, dx, npml=npml, boundaries=faces)
    ref = Yee3D(*size, dx, npml=npml, boundaries=faces)
    for lo, hi, eps in BOXES:
        model.add_dielectric_box(lo, hi, eps)
        ref.add_dielectric_box(lo, hi, eps)
    return model, ref


CASES = [None, {'y_lo': 'periodic', 'y_hi': 'periodic'}, {'x_lo': 'periodic', 'x_hi': 'periodic',
                                                          'z_lo': 'periodic', 'z_hi': 'periodic'}]


@pytest.mark.parametrize('faces', CASES)
def test_geometry_boxes_and_spacings(faces):
    model, ref = _pair(faces)
    assert model.shape == ref.shape
    assert model.dt == ref.dt and model.dt.dtype == np.float32
    np.testing.assert_array_equal(model.materials.property_array(0), ref.materials['eps_r'])
    for a in range(3):
        np.testing.assert_allclose(model.inv_e[a], ref.inv_e[a], rtol=6e-8)
        np.testing.assert_allclose(model.inv_h[a], ref.inv_h[a], rtol=6e-8)
        if not ref.periodic[a]:
            assert model.inv_h[a][-1] == 0.0


@pytest.mark.parametrize('faces', CASES)
def test_cpml_tables_match_oracle(faces):
    model, ref = _pair(faces)
    for layer in ref.cpml:
        tables = model.cpml['xyz'[layer.axis]]
        row_map = tables.row_e if layer.kind == 'e' else tables.row_h
        np.testing.assert_array_equal(row_map[layer.rows], np.arange(layer.rows.size))
        outside = np.ones(row_map.size, bool)
        outside[layer.rows] = False
        assert np.all(row_map[outside] == -1)
        b = tables.b_e if layer.kind == 'e' else tables.b_h
        g = tables.g_e if layer.kind == 'e' else tables.g_h
        np.testing.assert_allclose(b, layer.b.ravel(), rtol=6e-8)
        np.testing.assert_allclose(g, layer.gain.ravel(), rtol=6e-8, atol=1e-30)
    for a, axis in enumerate('xyz'):
        assert model.cpml[axis].n_rows == sum(l.rows.size for l in ref.cpml if l.axis == a and l.kind == 'e')


def test_coefficient_table_and_faces_match_oracle():
    model, ref = _pair(None)
    model.finalize()
    ref.finalize()
    table = model.table()[model.materials.ids]
    np.testing.assert_allclose(table[..., 0], ref.da, rtol=6e-8)
    np.testing.assert_allclose(table[..., 1], ref.db, rtol=6e-8)
    # same-id faces take the id's (ca, cb); mixed faces the mean the kernel forms from (eps_r, sigma_e)
    for axis in range(3):
        ids = model.materials.ids
        partner = np.take(ids, np.minimum(np.arange(1, ids.shape[axis] + 1), ids.shape[axis] - 1), axis=axis)
        same = ids == partner
        np.testing.assert_allclose(table[..., 2][same], ref.ca[axis][same], rtol=6e-8)
        np.testing.assert_allclose(table[..., 3][same], ref.cb[axis][same], rtol=6e-8)
        values = model.materials.values()
        eps = 0.5 * values[ids, 0] + 0.5 * values[partner, 0]
        _, cb = semi_implicit(eps, 0.0 * eps, float(model.dt), EPS0)
        np.testing.assert_allclose(cb, ref.cb[axis], rtol=1e-12)


@pytest.mark.parametrize('orientation,point', [('z', (0.36e-6, 0.29e-6, 0.27e-6)), ('x', (0.0, 0.6e-6, 0.1e-6)),
                                               ('y', (0.7e-6, 0.0, 0.55e-6))])
def test_dipole_placement_and_weight(orientation, point):
    model, ref = _pair({'y_lo': 'periodic', 'y_hi': 'periodic'})
    src = model.add_dipole(*point, orientation=orientation, wavelength=0.6e-6)
    rsrc = ref.add_dipole(*point, orientation=orientation, wavelength=0.6e-6)
    model.finalize()
    ref.finalize()
    assert src.cell == rsrc.cell and src.component == rsrc.component
    assert src.volume == pytest.approx(rsrc.volume, rel=1e-15)
    assert float(src.weight) == pytest.approx(rsrc.weight, rel=6e-8)
    times = np.arange(50) * float(model.dt) + 0.5 * float(model.dt)
    np.testing.assert_array_equal(src.wave_table(times), np.float32([rsrc.temporal(t) for t in times]))


@pytest.mark.parametrize('normal,position', [('x', 0.55e-6), ('y', 0.05e-6), ('z', 0.3e-6), ('y', 0.55e-6),
                                             ('z', 0.0)])
def test_flux_plane_layout_matches_oracle(normal, position):
    model, ref = _pair({'y_lo': 'periodic', 'y_hi': 'periodic'})
    model.add_dipole(0.3e-6, 0.3e-6, 0.3e-6, wavelength=0.6e-6)
    ref.add_dipole(0.3e-6, 0.3e-6, 0.3e-6, wavelength=0.6e-6)
    plane = model.add_flux_plane(normal, position, wavelengths=[0.5e-6, 0.7e-6])
    rplane = ref.add_flux_plane(normal, position, wavelengths=[0.5e-6, 0.7e-6])
    model.finalize()
    ref.finalize()
    assert plane.edge == rplane.edge and plane.sample_interval == rplane.sample_interval
    for pair, rpair in zip(plane.pairs, rplane.pairs):
        assert 'e' + 'xyz'[pair['e_comp']] == rpair['e'] and 'h' + 'xyz'[pair['h_comp']] == rpair['h']
        assert pair['sign'] == rpair['sign']
        np.testing.assert_allclose(pair['area'], rpair['area'], rtol=1e-15)
        ranges = {rpair['face_axis']: rpair['face_slice'], rpair['cell_axis']: rpair['cell_slice']}
        a, b = plane.tangential
        assert (pair['lo'][0], pair['hi'][0]) == (ranges[a].start, ranges[a].stop)
        assert (pair['lo'][1], pair['hi'][1]) == (ranges[b].start, ranges[b].stop)
    sample_times = np.arange(1, 6) * float(model.dt) * plane.sample_interval
    table = plane.phase_table(sample_times)
    phases = np.array([rplane.phases(t) for t in sample_times])
    np.testing.assert_allclose(table[..., 0] + 1j * table[..., 1], phases, rtol=0, atol=1e-30)


def test_plane_on_the_grid_edge_is_rejected_like_the_oracle():
    model, ref = _pair({'y_lo': 'periodic', 'y_hi': 'periodic'})
    for target in (model, ref):
        with pytest.raises(ValueError):
            target.add_flux_plane('y', 0.0, wavelengths=[0.5e-6])


def test_model3d_validation():
    with pytest.raises(ValueError):
        Model3D(1e-6, 1e-6, 1e-6, 25e-9, boundaries={'x_lo': 'periodic'})
    model = Model3D(0.5e-6, 0.5e-6, 0.5e-6, 25e-9, npml=4)
    with pytest.raises(ValueError):
        model.add_dipole(0.6e-6, 0.1e-6, 0.1e-6, wavelength=1e-6)
    with pytest.raises(ValueError):
        model.add_dipole(0.1e-6, 0.1e-6, 0.1e-6, orientation='w', wavelength=1e-6)
    with pytest.raises(ValueError):
        model.add_flux_plane('q', 0.1e-6, wavelengths=[1e-6])


def test_pec_faces_and_boxes_paint_over_dielectrics():
    """A 'pec' face makes the outermost window layer PEC; a PEC box keeps its cells whatever is painted after it
    (physics/em3d/engine3d.py:339-362, boundaries3d.py:149-158); 'pmc' is the bare grid edge (no CPML, no layer)."""
    from opensol_amd.fdtd3d import Model3D
    dx = 25e-9
    m = Model3D(20 * dx, 16 * dx, 12 * dx, dx, npml=4, boundaries={'z_lo': 'pec', 'z_hi': 'pmc', 'x_hi': 'pec'})
    assert m.shape == (24 + 0, 24, 12)   # x: cpml lo only; z: no CPML
    m.add_pec_box((5 * dx, 5 * dx, 5 * dx), (9 * dx, 9 * dx, 9 * dx))
    m.add_dielectric_box((0, 0, 0), (20 * dx, 16 * dx, 12 * dx), eps_r=3.0)   # painted after the PEC box
    m.finalize()
    ids = m.materials.ids
    pec = m.materials.pec_id()
    assert np.all(ids[:, :, 0] == pec) and np.all(ids[-1] == pec)       # z_lo and x_hi wall layers
    assert not np.any(ids[:-1, :, -1] == pec)                            # pmc: no layer (off the x_hi wall)
    box = ids[4 + 5:4 + 9, 4 + 5:4 + 9, 5:9]
    assert np.all(box == pec)
    assert int((ids == pec).sum()) == 24 * 24 + 24 * 12 - 24 + 4 * 4 * 4
    row = m.table()[pec]
    assert row[2] == 0.0 and row[3] == 0.0 and row[4] == -1.0           # ca = cb = 0, the kernel's PEC mark


def test_pec_face_rule_of_the_dipole_weight():
    """The dipole cb follows the PEC face rule: a face between two PEC cells is pinned, a face touching one PEC cell
    takes the other cell's values (kernels3d.py:61-85)."""
    from opensol_amd.fdtd3d import Model3D
    from opensol_amd.materials import semi_implicit
    from opensol_amd.constants import EPS0
    dx = 25e-9
    m = Model3D(12 * dx, 12 * dx, 12 * dx, dx, npml=3)
    m.add_dielectric_box((0, 0, 0), (12 * dx, 12 * dx, 12 * dx), eps_r=2.0)
    m.add_pec_box((6 * dx, 0, 0), (12 * dx, 12 * dx, 12 * dx))   # window cells x >= 6 (grid x >= 9)
    m.apply_pec()
    _, cb_dielectric = semi_implicit(2.0, 0.0, float(m.dt), EPS0)
    assert m.face_cb(0, (8, 6, 6)) == pytest.approx(float(cb_dielectric), rel=1e-15)   # dielectric | PEC
    assert m.face_cb(0, (10, 6, 6)) == 0.0                                             # PEC | PEC
    assert m.face_cb(1, (10, 6, 6)) == 0.0


def test_dipole_in_a_pec_wall_layer_is_rejected():
    from opensol_amd.fdtd3d import Model3D
    dx = 25e-9
    m = Model3D(12 * dx, 12 * dx, 12 * dx, dx, npml=3, boundaries={'z_lo': 'pec'})
    with pytest.raises(ValueError, match='pec face'):
        m.add_dipole(6 * dx, 6 * dx, 0.2 * dx, orientation='x', wavelength=0.6e-6)
    m.add_dipole(6 * dx, 6 * dx, 0.2 * dx, orientation='z', wavelength=0.6e-6)   # normal to the wall: allowed


def test_courant_check_skips_pec_cells():
    from opensol_amd.fdtd3d import Model3D
    dx = 25e-9
    m = Model3D(8 * dx, 8 * dx, 8 * dx, dx, npml=2, boundaries={'x_lo': 'pec', 'x_hi': 'pec'})
    m.finalize()   # the eps = -1 entry sets no limit


def test_field_probe_cells_match_the_oracle():
    model, ref = _pair({'y_lo': 'periodic', 'y_hi': 'periodic'})
    points = ((0.0, 0.0, 0.0), (0.7e-6, 0.6e-6, 0.55e-6), (0.36e-6, 0.29e-6, 0.27e-6), (0.1e-6, 0.45e-6, 0.5e-6))
    for point in points:
        for component in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
            got, want = model.add_field_probe(*point, component), ref.add_field_probe(*point, component)
            assert got.cell == want.cell and got.field_component == want.component
    with pytest.raises(ValueError):
        model.add_field_probe(0.1e-6, 0.1e-6, 0.1e-6, 'e')
    with pytest.raises(ValueError):
        model.add_field_probe(0.1e-6, 0.1e-6, 0.6e-6)


@pytest.mark.parametrize('z', [0.0, 0.27e-6, 0.55e-6])
def test_dft_plane_layer_matches_the_oracle(z):
    faces = {'z_lo': 'pmc', 'y_lo': 'periodic', 'y_hi': 'periodic'}
    model, ref = _pair(faces)
    plane, ref_plane = model.add_dft_plane(z, wavelengths=[6e-7]), ref.add_dft_plane(z, wavelengths=[6e-7])
    assert plane.k_grid == ref_plane.k
    assert plane.accumulator_shape() == (3, 1, ref_plane.xs.size, ref_plane.ys.size, 1, 4)   # Kahan (sum, comp) pairs
    assert plane.pairs == ((0, False), (1, True), (2, False))
    with pytest.raises(ValueError):
        model.add_dft_plane(0.7e-6, wavelengths=[6e-7])


def test_flux_box_faces_match_the_oracle():
    """Six bounded planes on the window edges nearest the corners, lo faces -1 and hi faces +1, with the oracle's
    pair ranges (physics/em3d/physics3d.py:447-476, monitors3d.py:131-162)."""
    model, ref = _pair({'y_lo': 'periodic', 'y_hi': 'periodic'})
    box = model.add_flux_box((0.5e-6, 0.14e-6, 0.12e-6), (0.2e-6, 0.46e-6, 0.41e-6), wavelengths=[6e-7])
    ref_box = ref.add_flux_box((0.5e-6, 0.14e-6, 0.12e-6), (0.2e-6, 0.46e-6, 0.41e-6), wavelengths=[6e-7])
    assert [s for s, _ in box.faces] == [s for s, _ in ref_box.faces] == [-1.0, 1.0] * 3
    for (_, plane), (_, ref_plane) in zip(box.faces, ref_box.faces):
        assert (plane.normal, plane.edge) == (ref_plane.normal, ref_plane.edge)
        for pair, ref_pair in zip(plane.pairs, ref_plane.pairs):
            a, b = sorted((ref_pair['face_axis'], ref_pair['cell_axis']))
            slices = {ref_pair['face_axis']: ref_pair['face_slice'], ref_pair['cell_axis']: ref_pair['cell_slice']}
            assert pair['lo'] == (slices[a].start, slices[b].start)
            assert pair['hi'] == (slices[a].stop, slices[b].stop)
            np.testing.assert_array_equal(pair['area'], ref_pair['area'])
    with pytest.raises(ValueError):
        model.add_flux_box((0.2e-6, 0.1e-6, 0.1e-6), (0.21e-6, 0.3e-6, 0.3e-6), wavelengths=[6e-7])


@pytest.mark.parametrize('name', ['a', 'b'])
def test_mirror_setup_matches_golden(name):
    """Mirror faces pad one ghost cell: grid shape, painting, dipole cell and weight and the flux plane edge equal the
    reference's golden run (tests/golden/mirrors3d.npz)."""
    import os
    from opensol_amd.fdtd3d import Model3D
    from tests.cases import build_mirrors3d
    golden = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'mirrors3d.npz'))
    model, dipole, plane = build_mirrors3d(Model3D, name)
    model.finalize()
    assert model.shape == tuple(golden[f'{name}_shape'])
    np.testing.assert_array_equal(model.materials.property_array(0), golden[f'{name}_eps_r'])
    assert dipole.cell == tuple(golden[f'{name}_dipole_cell'])
    assert abs(float(dipole.weight) / float(golden[f'{name}_dipole_weight'][0]) - 1.0) < 3e-7
    assert plane.edge == int(golden[f'{name}_plane_edge'])
