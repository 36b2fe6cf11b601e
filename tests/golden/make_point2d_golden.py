"""Write tests/fixtures/point2d_{ez,hz}.npz: golden vectors of 2D FDTD cases driven by point currents, from the
reference's own numpy path (EMPhysics precision='double', CPU backend). They pin the point-current sources end to end:
cell placement, the J = I / (dx dy) density, the cb / db / face-cb calibration (0 where PEC keeps a cell or face),
the injection time (t + dt/2 after the E update, t after the H update for the magnetic line current) and, for the
in-plane dipole, the face shares (with the periodic wrap at y cell 0) and the orientation split.

Cases (window 1.5 um x 1 um, dx 25 nm, 10-cell CPML on x, y periodic; a dielectric of eps_r 2.25 and a PEC rectangle):
  ez: one filament current in the dielectric, one in vacuum, one inside the PEC rectangle (weight 0).
  hz: a magnetic line current in the dielectric; an in-plane dipole at 30 degrees in y cell 0 (its E_x faces wrap to
      the last row); a dipole along y next to the PEC rectangle (its E_x shares fall below the 1e-12 cut; one E_y
      face touches PEC: weight 0).
Both cases also carry a Gaussian beam (a sheet source with the transverse factor exp(-((s - centre) / waist)^2))
over part of a line, whose calibrated profile is recorded. Fields are recorded after 120 and 400 steps, with every
source weight.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_point2d_golden.py
The fixture holds inputs (scalars of the case) and outputs (field arrays, weights); no source text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(os.path.dirname(HERE), 'fixtures')

WINDOW_X, WINDOW_Y, DX, NPML = 1.5e-6, 1e-6, 25e-9, 10
DIELECTRIC = ((0.2e-6, 0.0, 0.7e-6, 1e-6), 2.25)
PEC_RECT = (1.0e-6, 0.4e-6, 1.2e-6, 0.6e-6)
LAMBDA0, BANDWIDTH = 600e-9, 0.6
STEPS = (120, 400)
EZ_POINTS = ((0.4e-6, 0.5e-6, 2e-3), (0.85e-6, 0.3e-6, -1e-3), (1.1e-6, 0.5e-6, 1e-3))
HZ_MAGNETIC = (0.45e-6, 0.55e-6, 2.0)
HZ_DIPOLES = ((0.8e-6, 0.01e-6, 1e-3, 30.0), (0.99e-6, 0.5e-6, 2e-3, 90.0))
BEAM = dict(axis='y', position=1.35e-6, waist=0.2e-6, center=0.6e-6, amplitude=0.5, span=(0.1e-6, 0.9e-6))


def run_case(polarization):
    sys.path.insert(0, REFERENCE)
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig

    bc = BoundaryConfig('cpml', 'cpml', 'periodic', 'periodic')
    bc.cpml_npml = NPML
    domain = Domain.uniform(WINDOW_X, WINDOW_Y, DX, npml=NPML, boundary_config=bc, verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization=polarization, boundary_config=bc,
                                         precision='double'))
    (x0, y0, x1, y1), eps = DIELECTRIC
    em.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    mask = np.asarray(domain.rect_mask(*PEC_RECT), dtype=bool)
    em.add_pec_region(mask)
    if polarization == 'ez':
        sources = [em.add_point_source(x, y, wavelength=LAMBDA0, bandwidth=BANDWIDTH, current=current)
                   for x, y, current in EZ_POINTS]
    else:
        x, y, km = HZ_MAGNETIC
        sources = [em.add_magnetic_line_current(x, y, wavelength=LAMBDA0, bandwidth=BANDWIDTH, magnetic_current=km)]
        sources += [em.add_electric_dipole(x, y, wavelength=LAMBDA0, bandwidth=BANDWIDTH, dipole_moment=moment,
                                           orientation=angle)
                    for x, y, moment, angle in HZ_DIPOLES]
    beam = em.add_gaussian_beam(wavelength=LAMBDA0, bandwidth=BANDWIDTH, **BEAM)
    names = ('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')
    out = {'pec_mask': mask}
    done = 0
    for target in STEPS:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in names:
            out[f'{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    engine = em.engine
    out.update({'nx': engine.nx, 'ny': engine.ny, 'dt': float(engine.dt), 'n_steps': engine.n_step})
    for index, source in enumerate(sources):
        out[f'source{index}_cell'] = np.array([source.ix, source.iy])
        if hasattr(source, 'faces'):
            out[f'source{index}_faces'] = np.array([[0 if c == 'ex' else 1, i, j] for c, i, j, _ in source.faces])
            out[f'source{index}_weights'] = np.asarray(source.face_weights, dtype=np.float64)
        else:
            out[f'source{index}_weights'] = np.asarray(source.weight, dtype=np.float64)
    out['beam_span'] = np.array(beam.span)
    out['beam_profile'] = np.asarray(beam.profile, dtype=np.float64)
    return out


def main():
    for polarization in ('ez', 'hz'):
        out = run_case(polarization)
        path = os.path.join(FIXTURES, f'point2d_{polarization}.npz')
        np.savez_compressed(path, **out)
        weights = {k: out[k] for k in sorted(out) if k.endswith('_weights')}
        print(f'{path}: {out["nx"]} x {out["ny"]}, dt {out["dt"]}, weights {weights}')


if __name__ == '__main__':
    main()
