"""Write tests/golden/dipole3d_{cpml,periodic_y,pec}.npz: golden vectors of a small 3D FDTD case from the reference's own
numpy path (EMPhysics3D, precision 'single', CPU backend: the float32 numpy twins of the 3D kernels). They pin the 3D
stepping, CPML, dipole calibration and plane-flux composition that no closed form pins end to end.

Case: window 0.7 x 0.6 x 0.55 um, dx 25 nm, 8-cell CPML on every face (or periodic y), eps_r painted directly on the
reference's float32 material array (three boxes, half-open on the float32 cell centres clamped to the window), a
Gaussian point dipole (lambda0 0.6 um, bandwidth 0.5, moment 1e-12 A m) near the centre, two flux planes (normal x
and z) at 3 wavelengths; 300 steps. All six fields are recorded after the last step, and their three planes through the
dipole cell after step 120.

Case 'pec' adds a PEC z_lo face, a bare (pmc) z_hi face and a PEC box over the dielectrics.

The cpml and pec cases also have a flux box around the dipole (net outward flux of its six faces), recorded after
the last step.

Every case also has running-DFT planes of E at the cell centres of one or two z layers (through the dipole; in the
pec case also layer 0, whose E_z pairs with itself), at the flux planes' wavelengths, recorded after the last step.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_dipole3d_golden.py [case ...]
The fixture holds inputs (scalars and the eps_r array of the case) and outputs (field arrays, flux spectra); no source
text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

WINDOW = (0.7e-6, 0.6e-6, 0.55e-6)
DX, NPML = 25e-9, 8
BOXES = (((0.05e-6, 0.05e-6, 0.00e-6), (0.30e-6, 0.35e-6, 0.25e-6), 4.0),
         ((0.40e-6, 0.20e-6, 0.15e-6), (0.70e-6, 0.60e-6, 0.40e-6), 2.25),
         ((0.20e-6, 0.40e-6, 0.30e-6), (0.50e-6, 0.55e-6, 0.55e-6), 6.5))
LAMBDA0, BANDWIDTH, MOMENT = 0.6e-6, 0.5, 1e-12
WAVELENGTHS = (0.5e-6, 0.6e-6, 0.7e-6)
SLICE_STEP = 120
N_STEPS = 300
CASES = {
    'cpml': dict(faces={}, dipole=(0.36e-6, 0.29e-6, 0.27e-6), orientation='z',
                 planes=(('x', 0.55e-6), ('z', 0.1e-6)), dft_planes=(0.27e-6,),
                 flux_box=((0.2e-6, 0.15e-6, 0.12e-6), (0.5e-6, 0.45e-6, 0.42e-6))),
    'periodic_y': dict(faces={'y_lo': 'periodic', 'y_hi': 'periodic'}, dipole=(0.33e-6, 0.21e-6, 0.26e-6),
                       orientation='x', planes=(('x', 0.15e-6), ('z', 0.45e-6)), dft_planes=(0.26e-6,)),
    # a PEC ground face (z_lo), the bare grid edge (z_hi pmc) and a PEC box over the dielectrics
    'pec': dict(faces={'z_lo': 'pec', 'z_hi': 'pmc'}, dipole=(0.36e-6, 0.29e-6, 0.27e-6), orientation='z',
                planes=(('x', 0.2e-6), ('z', 0.45e-6)),
                pec_box=((0.45e-6, 0.10e-6, 0.30e-6), (0.60e-6, 0.35e-6, 0.45e-6)), dft_planes=(0.0, 0.27e-6),
                flux_box=((0.25e-6, 0.2e-6, 0.05e-6), (0.45e-6, 0.4e-6, 0.40e-6))),
}
FIELDS = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')


def box_masks(centres):
    """Cells whose clamped float32 centre lies in [lo, hi) on every axis (the model builder's box raster)."""
    masks = []
    for lo, hi, _ in BOXES:
        axis_masks = [(c >= np.float32(a)) & (c < np.float32(b)) for c, a, b in zip(centres, lo, hi)]
        masks.append(axis_masks[0][:, None, None] & axis_masks[1][None, :, None] & axis_masks[2][None, None, :])
    return masks


def run_case(spec):
    sys.path.insert(0, REFERENCE)
    from open_sol.core.domain3d import Domain3D
    from open_sol.physics.em3d.boundaries3d import BoundaryConfig3D
    from open_sol.physics.em3d.physics3d import EMPhysics3D

    faces = {face: spec['faces'].get(face, 'cpml') for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    bc = BoundaryConfig3D(faces, npml=NPML)
    sides = {face: (NPML if kind == 'cpml' else 0) for face, kind in faces.items()}
    domain = Domain3D.uniform(*WINDOW, DX, npml_sides=sides, boundary_config=bc)
    centres = domain.host_centres(True)
    eps = np.ones(domain.grid.shape, dtype=np.float32)
    for mask, (_, _, value) in zip(box_masks(centres), BOXES):
        eps[mask] = np.float32(value)
    for name, default in (('eps_r', 1.0), ('sigma_e', 0.0), ('mu_r', 1.0)):   # the engine's required arrays
        domain.materials.add_property(name, default)
    domain.materials.eps_r[...] = eps
    em = EMPhysics3D(domain, precision='single')
    pec = np.zeros(domain.grid.shape, dtype=bool)
    if 'pec_box' in spec:
        lo, hi = spec['pec_box']
        axis_masks = [(c >= np.float32(a)) & (c < np.float32(b)) for c, a, b in zip(centres, lo, hi)]
        pec = axis_masks[0][:, None, None] & axis_masks[1][None, :, None] & axis_masks[2][None, None, :]
        em.add_pec(tuple(slice(0, n) for n in domain.grid.shape), pec)
    source = em.add_dipole(*spec['dipole'], orientation=spec['orientation'], dipole_moment=MOMENT,
                           wavelength=LAMBDA0, bandwidth=BANDWIDTH)
    planes = [em.add_flux_plane(normal, position, wavelengths=list(WAVELENGTHS)) for normal, position in spec['planes']]
    dft_planes = [em.add_dft_plane(z, wavelengths=list(WAVELENGTHS)) for z in spec['dft_planes']]
    box = em.add_flux_box(*spec['flux_box'], wavelengths=list(WAVELENGTHS)) if 'flux_box' in spec else None
    em.finalize()
    engine = em.engine
    out = {'shape': np.array(engine.shape), 'dt': np.float64(engine.dt), 'eps_r': eps, 'pec_box_mask': pec,
           'pec_cells': np.asarray(np.asarray(engine.material_id) == engine.pec_id),
           'dipole_cell': np.array(source.cell), 'dipole_weight': np.asarray(source.weight, dtype=np.float64),
           'sample_interval': np.int64(planes[0].sample_interval),
           'plane_edges': np.array([plane.edge for plane in planes])}
    i, j, k = source.cell
    for _ in range(SLICE_STEP):
        em.step()
    for name in FIELDS:
        field = np.array(getattr(engine.fields, name), dtype=np.float32)
        out[f'{name}_{SLICE_STEP}_x'], out[f'{name}_{SLICE_STEP}_y'] = field[i], field[:, j]
        out[f'{name}_{SLICE_STEP}_z'] = field[:, :, k]
    for _ in range(N_STEPS - SLICE_STEP):
        em.step()
    for name in FIELDS:
        out[f'{name}_{N_STEPS}'] = np.array(getattr(engine.fields, name), dtype=np.float32)
    out['time'] = np.float64(engine.time)
    out['n_steps'] = np.int64(engine.n_step)
    if box is not None:
        out['flux_box'] = np.asarray(box.flux(), dtype=np.float64)
        out['flux_box_edges'] = np.array(box.edge_bounds)
    for index, plane in enumerate(dft_planes):
        out[f'dft{index}_k'] = np.int64(plane.k_grid)
        for name in ('ex', 'ey', 'ez'):
            out[f'dft{index}_{name}'] = np.array(plane.spectra[name])
    for index, plane in enumerate(planes):
        out[f'flux_{index}'] = np.asarray(plane.flux(), dtype=np.float64)
        for pair_index, (pair, e_spectrum, h_spectrum) in enumerate(plane.host_pairs()):
            out[f'e_spectrum_{index}_{pair_index}'] = e_spectrum
            out[f'h_spectrum_{index}_{pair_index}'] = h_spectrum
    return out


def main(names=None):
    for name, spec in CASES.items():
        if names and name not in names:
            continue
        out = run_case(spec)
        path = os.path.join(HERE, f'dipole3d_{name}.npz')
        np.savez_compressed(path, **out)
        print(path, out['shape'], 'dt', out['dt'], 'm', out['sample_interval'], 'flux_0', out['flux_0'])


if __name__ == '__main__':
    main(sys.argv[1:])
