"""The 3D oracle against what pins it: golden vectors of the reference's own 3D numpy path (float32 twins, so the
tolerances sit at float32 round-off; one case with a PEC face, a bare pmc face and a PEC box), the exact discrete dispersion of periodic-box eigenmodes of each E component,
and the reduction of a 2D-invariant 3D run to the 2D oracle (which the 2D golden vectors pin bit for bit) along each
axis, CPML included."""
import os

import numpy as np
import pytest

from oracle.yee2d import EPS0, MU0, Yee2D
from oracle.yee3d import Yee3D

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
GOLDEN_CASES = ('cpml', 'periodic_y', 'pec')


def _spec():
    """The case definitions of the golden script (imported at run time: the fixtures stay on the CPU side)."""
    from tests.golden import make_dipole3d_golden
    return make_dipole3d_golden


def _golden_model(name, dtype=np.float64, phase_dtype=np.complex128):
    G = _spec()
    spec = G.CASES[name]
    faces = {face: spec['faces'].get(face, 'cpml') for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    sim = Yee3D(*G.WINDOW, G.DX, npml=G.NPML, boundaries=faces, dtype=dtype, phase_dtype=phase_dtype)
    for lo, hi, eps in G.BOXES:
        sim.add_dielectric_box(lo, hi, eps)
    if 'pec_box' in spec:
        sim.add_pec_box(*spec['pec_box'])
    source = sim.add_dipole(*spec['dipole'], orientation=spec['orientation'], dipole_moment=G.MOMENT,
                            wavelength=G.LAMBDA0, bandwidth=G.BANDWIDTH)
    planes = [sim.add_flux_plane(normal, position, wavelengths=list(G.WAVELENGTHS))
              for normal, position in spec['planes']]
    sim.finalize()
    return sim, source, planes


def _rel(a, b):
    a = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize('name', GOLDEN_CASES)
def test_setup_matches_golden(name):
    golden = np.load(os.path.join(GOLDEN, f'dipole3d_{name}.npz'))
    sim, source, planes = _golden_model(name)
    assert sim.shape == tuple(golden['shape'])
    assert float(sim.dt) == float(golden['dt'])
    np.testing.assert_array_equal(sim.materials['eps_r'], golden['eps_r'])
    if 'pec_cells' in golden:   # PEC box raster and the pec face's wall layer
        np.testing.assert_array_equal(sim.pec_mask(), golden['pec_cells'])
    assert source.cell == tuple(golden['dipole_cell'])
    # the reference rounds cb and moment / volume to float32 before the product
    assert abs(source.weight / float(golden['dipole_weight'][0]) - 1.0) < 3e-7
    assert sim.sample_interval == int(golden['sample_interval'])
    assert [p.edge for p in planes] == list(golden['plane_edges'])


@pytest.mark.parametrize('name', GOLDEN_CASES)
def test_fields_and_flux_match_golden(name):
    """FP64 oracle vs the reference's float32 run: fields within float32 accumulation (1e-4 relative L2 while the pulse
    is in the box, 5e-4 for the absorbed remainder at the end), plane flux spectra within 2e-6."""
    G = _spec()
    golden = np.load(os.path.join(GOLDEN, f'dipole3d_{name}.npz'))
    sim, source, planes = _golden_model(name)
    # field probes on the three slices through the dipole's cell (one per component) and at the window's corners
    x, y, z = G.CASES[name]['dipole']
    sizes = G.WINDOW
    spots = [(x, y, z), (0.1e-6, y, 0.2e-6), (x, 0.5e-6, 0.1e-6), (0.6e-6, 0.05e-6, z), (0.0, 0.0, 0.0), sizes,
             (x, y, 0.4e-6)]
    probes = [sim.add_field_probe(*spot, component=f) for spot, f in zip(spots * 2, G.FIELDS * 2)]
    dft_planes = [sim.add_dft_plane(z, wavelengths=list(G.WAVELENGTHS)) for z in G.CASES[name]['dft_planes']]
    box = (sim.add_flux_box(*G.CASES[name]['flux_box'], wavelengths=list(G.WAVELENGTHS))
           if 'flux_box' in G.CASES[name] else None)
    sim.run(n_steps=G.SLICE_STEP)
    i, j, k = source.cell
    for f in G.FIELDS:
        for axis, values in (('x', sim.fields[f][i]), ('y', sim.fields[f][:, j]), ('z', sim.fields[f][:, :, k])):
            assert _rel(golden[f'{f}_{G.SLICE_STEP}_{axis}'], values) < 1e-4, (f, axis)
    sim.run(n_steps=G.N_STEPS - G.SLICE_STEP)
    assert sim.time == float(golden['time'])
    for f in G.FIELDS:
        assert _rel(golden[f'{f}_{G.N_STEPS}'], sim.fields[f]) < 5e-4, f
    # flux box: the six faces' edges and the net outward flux, within the planes' 2e-6
    if box is not None:
        assert [[face.edge for _, face in box.faces[2 * a:2 * a + 2]] for a in range(3)] == \
            golden['flux_box_edges'].tolist()
        np.testing.assert_allclose(box.flux(), golden['flux_box'], rtol=2e-6)
    # DFT planes of E at the cell centres (the pec case's layer 0 pairs E_z with itself): within 1e-5 relative L2
    # of the reference's complex64 spectra (measured 1.5e-7 to 2.1e-6)
    for index, plane in enumerate(dft_planes):
        assert plane.k == int(golden[f'dft{index}_k'])
        for component in ('ex', 'ey', 'ez'):
            want = golden[f'dft{index}_{component}'].astype(np.complex128)
            if not np.any(want):   # tangential E on the PEC face
                assert not np.any(plane.spectra[component]), (index, component)
                continue
            assert _rel(plane.spectra[component], want) < 1e-5, (index, component)
    # each probe's samples after SLICE_STEP and N_STEPS are the reference's field at its cell (float32 run: within
    # 1e-4 of the field's peak; 1e-3 in the dipole's own cell, where the float32 run accumulates the source's
    # residue; a neighbouring cell is off by 5e-4 to 3e-2 of the peak)
    for probe in probes:
        times, values = probe.get_data()
        assert len(values) == G.N_STEPS and times[-1] == sim.time
        pi, pj, pk = probe.cell
        final = golden[f'{probe.component}_{G.N_STEPS}']
        tol = 1e-3 if probe.cell == source.cell else 1e-4
        assert abs(values[-1] - final[probe.cell]) <= tol * np.abs(final).max(), probe.cell
        for axis, hit, index in (('x', pi == i, (pj, pk)), ('y', pj == j, (pi, pk)), ('z', pk == k, (pi, pj))):
            if hit:
                early = golden[f'{probe.component}_{G.SLICE_STEP}_{axis}']
                assert abs(values[G.SLICE_STEP - 1] - early[index]) <= 1e-4 * np.abs(early).max(), (probe.cell, axis)
    for index, plane in enumerate(planes):
        np.testing.assert_allclose(plane.flux(), golden[f'flux_{index}'], rtol=2e-6)
        for pair_index, (_, e_spectrum, h_spectrum) in enumerate(plane.spectra()):
            for kind, values in (('e', e_spectrum), ('h', h_spectrum)):
                reference = golden[f'{kind}_spectrum_{index}_{pair_index}']
                assert _rel(reference, values) < 5e-6, (index, pair_index, kind)


@pytest.mark.parametrize('component', ['ex', 'ey', 'ez'])
@pytest.mark.parametrize('p,q', [(1, 0), (2, 3)])
def test_periodic_eigenmode_follows_discrete_dispersion(component, p, q):
    """A divergence-free single mode of the periodic Yee box (E_c varying along the two other axes) obeys
    F^{n+1} + F^{n-1} = 2 cos(w dt) F^n with cos(w dt) = 1 - 2 dt^2/(eps0 mu0) sum sin^2(theta/2)/d^2, exactly."""
    dx = 25e-9
    faces = {face: 'periodic' for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    sim = Yee3D(12 * dx, 10 * dx, 8 * dx, dx, boundaries=faces)
    axis = 'xyz'.index(component[1])
    a, b = [other for other in range(3) if other != axis]
    theta = np.zeros(3)
    theta[a], theta[b] = 2 * np.pi * p / sim.shape[a], 2 * np.pi * q / sim.shape[b]
    index = np.indices(sim.shape)
    mode = np.cos(theta[a] * index[a]) * np.cos(theta[b] * index[b])
    sim.fields[component][...] = mode
    d, dt = float(sim.dx32), float(sim.dt)
    cos_wdt = 1.0 - 2.0 * dt ** 2 / (EPS0 * MU0) * np.sum(np.sin(theta / 2) ** 2) / d ** 2
    series = []
    for _ in range(50):
        sim.step()
        series.append(sim.fields[component].copy())
    series = np.array(series)
    assert np.max(np.abs(series[2:] + series[:-2] - 2.0 * cos_wdt * series[1:-1])) < 1e-11 * np.max(np.abs(series))
    last = series[-1]
    coefficient = np.sum(last * mode) / np.sum(mode * mode)
    assert np.max(np.abs(last - coefficient * mode)) < 1e-11
    for other in ('ex', 'ey', 'ez'):
        if other != component:
            assert np.max(np.abs(sim.fields[other])) == 0.0


# (3D invariant axis, the 3D axes that play 2D x and y, 3D fields that play 2D (E_z, H_x, H_y))
REDUCTIONS = [(2, (0, 1), ('ez', 'hx', 'hy')), (0, (1, 2), ('ex', 'hy', 'hz')), (1, (2, 0), ('ey', 'hz', 'hx'))]


@pytest.mark.parametrize('invariant,plane,names', REDUCTIONS)
def test_invariant_run_reduces_to_2d_oracle(invariant, plane, names):
    """A 3D run periodic and invariant along one axis is the 2D E_z-polarized run on the other two (CPML layers,
    curl spacings, coefficients and signs of both in-plane axes), to FP64 round-off."""
    dx, npml = 25e-9, 6
    size = (0.5e-6, 0.4e-6)
    sizes3d = [0.0, 0.0, 0.0]
    sizes3d[plane[0]], sizes3d[plane[1]], sizes3d[invariant] = size[0], size[1], 3 * dx
    faces = {f'{"xyz"[invariant]}_lo': 'periodic', f'{"xyz"[invariant]}_hi': 'periodic'}
    sim = Yee3D(*sizes3d, dx, npml=npml, boundaries=faces)
    # a 2D Courant factor whose float32 step equals the 3D one (0.99 sqrt(2/3) up to a few ulp)
    for delta in range(-40, 41):
        ref = Yee2D(size[0], size[1], dx, npml=npml, courant=0.99 * np.sqrt(2.0 / 3.0) * (1 + delta * 1e-8),
                    phase_dtype=np.complex128)
        if ref.dt == sim.dt:
            break
    assert sim.dt == ref.dt
    rng = np.random.default_rng(3)
    rects = []
    for _ in range(4):   # inside the window: the 3D raster clamps PML centres, the 2D one does not
        x0, y0 = rng.uniform(dx, size[0] - 2 * dx), rng.uniform(dx, size[1] - 2 * dx)
        rects.append((x0, y0, min(x0 + rng.uniform(0.05e-6, 0.3e-6), size[0] - dx),
                      min(y0 + rng.uniform(0.05e-6, 0.3e-6), size[1] - dx), rng.uniform(1.0, 8.0)))
    for x0, y0, x1, y1, eps in rects:
        ref.add_dielectric_rectangle(x0, y0, x1, y1, eps)
        lo, hi = [-1.0] * 3, [1.0] * 3
        lo[plane[0]], lo[plane[1]], hi[plane[0]], hi[plane[1]] = x0, y0, x1, y1
        sim.add_dielectric_box(lo, hi, eps)
    ez0 = rng.standard_normal(ref.fields['ez'].shape)
    ref.fields['ez'][...] = ez0
    order = [plane[0], plane[1], invariant]
    sim.fields[names[0]][...] = np.moveaxis(np.broadcast_to(ez0[:, :, None], ez0.shape + (3,)),
                                            [0, 1, 2], order)
    ref.finalize()
    sim.finalize()
    for _ in range(120):
        ref.step()
        sim.step()
    for name2d, name3d in zip(('ez', 'hx', 'hy'), names):
        got = np.moveaxis(sim.fields[name3d], order, [0, 1, 2])
        want = ref.fields[name2d]
        scale = np.max(np.abs(want))
        for k in range(3):
            assert np.max(np.abs(got[:, :, k] - want)) < 1e-12 * scale, (name2d, name3d, k)


def test_oracle_reproduces_reference_pole_case():
    """Drude and Lorentz boxes around a dipole: masks and eps_r exact, fields within 5e-5 relative L2 of the
    reference's float32 run after 250 steps (measured 1.4e-5, float32 round-off of that run; without the poles the
    fields are off by 1), flux within 1e-6 (measured 3e-7)."""
    from tests.cases import POLES3D, build_poles3d
    golden = np.load(os.path.join(GOLDEN, 'poles3d.npz'))
    sim, plane = build_poles3d(lambda *a, **k: Yee3D(*a, **k))
    assert sim.shape == tuple(golden['shape'])
    np.testing.assert_array_equal(sim.poles[0]['mask'], golden['drude_mask'])
    np.testing.assert_array_equal(sim.poles[1]['mask'], golden['lorentz_mask'])
    np.testing.assert_array_equal(sim.materials['eps_r'], golden['eps_r'])
    sim.run(n_steps=POLES3D['n_steps'])
    for f in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
        error = _rel(sim.fields[f], golden[f].astype(np.float64))
        assert error < 5e-5, (f, error)
    np.testing.assert_allclose(plane.flux(), golden['flux'], rtol=1e-6)


@pytest.mark.parametrize('name', ['a', 'b'])
def test_mirror_faces_match_golden(name):
    """Mirror faces (symmetric: tangential H and normal E on the plane pinned to 0; antisymmetric: the ghost cell's
    tangential E = minus its inner neighbour's) on two corners: the FP64 oracle against the reference's float32 run
    (fields within 5e-4 relative L2 after 200 steps, flux within 2e-5); the pinned rows are exactly 0 / exact images
    in both."""
    from tests.cases import MIRRORS3D, build_mirrors3d
    golden = np.load(os.path.join(GOLDEN, 'mirrors3d.npz'))
    sim, dipole, plane = build_mirrors3d(Yee3D, name)
    sim.finalize()
    assert sim.shape == tuple(golden[f'{name}_shape'])
    np.testing.assert_array_equal(sim.materials['eps_r'], golden[f'{name}_eps_r'])
    assert dipole.cell == tuple(golden[f'{name}_dipole_cell'])
    assert plane.edge == int(golden[f'{name}_plane_edge'])
    sim.run(n_steps=MIRRORS3D['n_steps'])
    for f in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
        assert _rel(golden[f"{name}_{f}"], sim.fields[f]) < 5e-5, f   # measured 1e-6 to 1.3e-5
    np.testing.assert_allclose(plane.flux(), golden[f'{name}_flux'], rtol=2e-5)
    for axis, kind, ghost, inner, face in sim.bc.mirror_planes(sim.windows):
        others = [a for a in 'xyz' if a != 'xyz'[axis]]
        for source in (golden, None):
            get = (lambda c: golden[f'{name}_{c}']) if source is not None else (lambda c: sim.fields[c])
            if kind == 'symmetric':
                assert not np.any(np.take(get('e' + 'xyz'[axis]), face, axis=axis))
                for o in others:
                    assert not np.any(np.take(get('h' + o), face, axis=axis))
            else:
                for o in others:
                    np.testing.assert_array_equal(np.take(get('e' + o), ghost, axis=axis),
                                                  -np.take(get('e' + o), inner, axis=axis))
