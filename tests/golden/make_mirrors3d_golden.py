"""Write tests/golden/mirrors3d.npz: golden vectors of small 3D FDTD cases with mirror faces, from the reference's own
numpy path (EMPhysics3D, precision 'single', CPU backend). They pin the 3D mirror rules: one ghost cell padded beyond
each mirror face; on a 'symmetric' face the tangential H and the normal E on the plane are pinned to 0, on an
'antisymmetric' face the ghost's tangential E is minus its inner neighbour's.

Cases (window 0.55 x 0.5 x 0.45 um, dx 25 nm, 8-cell CPML on the other faces; the spec is tests/cases.py MIRRORS3D):
  'a': x_lo symmetric, y_hi antisymmetric, z_lo antisymmetric; a z dipole near the (x_lo, z_lo) corner;
  'b': y_lo symmetric, z_hi symmetric, x_hi antisymmetric; an x dipole near the (y_lo, z_hi) corner;
each over two dielectric boxes, with a flux plane normal to x; fields after 200 steps, the flux.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_mirrors3d_golden.py
The fixture holds inputs (scalars and eps_r of the cases) and outputs (field arrays, flux spectra); no source text of
the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.cases import MIRRORS3D, poles3d_box_mask   # noqa: E402


def run_case(name):
    sys.path.insert(0, REFERENCE)
    from open_sol.core.domain3d import Domain3D
    from open_sol.physics.em3d.boundaries3d import BoundaryConfig3D
    from open_sol.physics.em3d.physics3d import EMPhysics3D

    spec = MIRRORS3D
    case = spec['cases'][name]
    faces = {face: case['faces'].get(face, 'cpml') for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    bc = BoundaryConfig3D(faces, npml=spec['npml'])
    sides = {face: (spec['npml'] if kind == 'cpml' else 1 if kind in ('symmetric', 'antisymmetric') else 0)
             for face, kind in faces.items()}
    domain = Domain3D.uniform(*spec['window'], spec['dx'], npml_sides=sides, boundary_config=bc)
    centres = domain.host_centres(True)
    eps = np.ones(domain.grid.shape, dtype=np.float32)
    for lo, hi, value in spec['boxes']:
        eps[poles3d_box_mask(centres, lo, hi)] = np.float32(value)
    for prop, default in (('eps_r', 1.0), ('sigma_e', 0.0), ('mu_r', 1.0)):
        domain.materials.add_property(prop, default)
    domain.materials.eps_r[...] = eps
    em = EMPhysics3D(domain, precision='single')
    source = em.add_dipole(*case['dipole'], orientation=case['orientation'], dipole_moment=1e-12,
                           wavelength=spec['wavelength'], bandwidth=0.5)
    plane = em.add_flux_plane('x', spec['plane_x'], wavelengths=list(spec['wavelengths']))
    em.finalize()
    engine = em.engine
    for _ in range(spec['n_steps']):
        em.step()
    out = {'shape': np.array(engine.shape), 'dt': np.float64(engine.dt), 'eps_r': eps,
           'dipole_cell': np.array(source.cell), 'dipole_weight': np.asarray(source.weight, dtype=np.float64),
           'plane_edge': np.int64(plane.edge), 'sample_interval': np.int64(plane.sample_interval),
           'flux': np.asarray(plane.flux(), dtype=np.float64)}
    for field in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
        out[field] = np.array(getattr(engine.fields, field), dtype=np.float32)
    return out


def main():
    out = {}
    for name in MIRRORS3D['cases']:
        for key, value in run_case(name).items():
            out[f'{name}_{key}'] = value
        print(name, out[f'{name}_shape'], 'flux', out[f'{name}_flux'])
    np.savez_compressed(os.path.join(HERE, 'mirrors3d.npz'), **out)


if __name__ == '__main__':
    main()
