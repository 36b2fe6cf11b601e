"""Write tests/fixtures/walls2d_ez.npz: golden vectors of a 2D E_z case bounded by wall and mirror faces, from the
reference's own numpy path (EMPhysics precision='double', CPU backend). It pins the outer-row rules of the E_z
polarisation end to end: the one ghost cell of a mirror face, the even copy of a 'symmetric' face, the odd copy of an
'antisymmetric' face, the 'pmc' wall (the outermost window row copies its neighbour, tangential H = 0 on the plane),
the corner taking both factors, and the re-pin after a source injects into a pinned row; and, for H_z, the 'pmc'
wall (H_z = 0 on the outermost cells, after a magnetic line current beside it).

Case (window 1.2 um x 1 um, dx 25 nm, x_lo 'symmetric', x_hi 'pmc', y_lo 'antisymmetric', y_hi 10-cell CPML): a
dielectric rectangle, a filament current near the (x_lo, y_lo) corner and a Gaussian line source along y whose span
covers the whole grid line (so it injects into the pinned rows). Fields are recorded after 150 and 400 steps.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_walls2d_golden.py
The fixture holds inputs (scalars of the case) and outputs (field arrays); no source text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(os.path.dirname(HERE), 'fixtures')

WINDOW_X, WINDOW_Y, DX, NPML = 1.2e-6, 1e-6, 25e-9, 10
FACES = ('symmetric', 'pmc', 'antisymmetric', 'cpml')
DIELECTRIC = ((0.5e-6, 0.2e-6, 0.8e-6, 0.6e-6), 3.0)
POINT = (0.1e-6, 0.12e-6, 1e-3)
LINE = dict(axis='y', position=0.95e-6)
LAMBDA0, BANDWIDTH = 600e-9, 0.6
STEPS = (150, 400)


HZ_FACES = ('pmc', 'cpml', 'cpml', 'pmc')
HZ_MAGNETIC = (0.05e-6, 0.95e-6, 2.0)


def run_hz():
    """The H_z case: x_lo and y_hi 'pmc' (H_z = 0 on the outermost cells), the dielectric, a magnetic line current
    next to the corner they share and a 45 degree dipole; fields after 150 and 400 steps under hz_* keys."""
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig

    bc = BoundaryConfig(*HZ_FACES)
    bc.cpml_npml = NPML
    domain = Domain.uniform(WINDOW_X, WINDOW_Y, DX, npml=NPML, boundary_config=bc, verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization='hz', boundary_config=bc, precision='double'))
    (x0, y0, x1, y1), eps = DIELECTRIC
    em.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    x, y, km = HZ_MAGNETIC
    em.add_magnetic_line_current(x, y, wavelength=LAMBDA0, bandwidth=BANDWIDTH, magnetic_current=km)
    em.add_electric_dipole(0.6e-6, 0.5e-6, wavelength=LAMBDA0, bandwidth=BANDWIDTH, orientation=45.0)
    out, done = {}, 0
    for target in STEPS:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in ('hz', 'ex', 'ey'):
            out[f'hzcase_{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    return out


def main():
    sys.path.insert(0, REFERENCE)
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig

    bc = BoundaryConfig(*FACES)
    bc.cpml_npml = NPML
    domain = Domain.uniform(WINDOW_X, WINDOW_Y, DX, npml=NPML, boundary_config=bc, verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization='ez', boundary_config=bc, precision='double'))
    (x0, y0, x1, y1), eps = DIELECTRIC
    em.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    x, y, current = POINT
    em.add_point_source(x, y, wavelength=LAMBDA0, bandwidth=BANDWIDTH, current=current)
    em.add_line_source(wavelength=LAMBDA0, bandwidth=BANDWIDTH, **LINE)
    out = {}
    done = 0
    for target in STEPS:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in ('ez', 'hx', 'hy'):
            out[f'{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    engine = em.engine
    out.update({'nx': engine.nx, 'ny': engine.ny, 'dt': float(engine.dt)})
    out.update(run_hz())
    path = os.path.join(FIXTURES, 'walls2d_ez.npz')
    np.savez_compressed(path, **out)
    ez = out['ez_400']
    print(f'{path}: {engine.nx} x {engine.ny}; ez[0] - ez[1] {np.abs(ez[0] - ez[1]).max():.3e}, '
          f'ez[:,0] + ez[:,1] {np.abs(ez[:, 0] + ez[:, 1]).max():.3e}, ez[-1] - ez[-2] {np.abs(ez[-1] - ez[-2]).max():.3e}')


if __name__ == '__main__':
    main()
