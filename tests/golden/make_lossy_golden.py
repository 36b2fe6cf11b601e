"""Write tests/golden/lossy2d_{ez,hz}.npz and tests/golden/lossy3d.npz: golden vectors of small FDTD cases with
conductive (sigma_e > 0) media and the 'ricker' and 'cw' waveforms, from the reference's own numpy path (2D:
EMPhysics precision='double'; 3D: EMPhysics3D precision='single', CPU backend). They pin what the lossless golden
cases leave open: the semi-implicit conductive E coefficients (2D cell tables, the H_z face means of eps and sigma,
the 3D face means), the sheet calibration in a conductive medium (|eta| with sigma_e), the Ricker wavelet and the
continuous wave with its switch-on ramp (and the DFT sample interval 1 that a cw source forces).

2D cases (window 2 um x 1 um, dx 25 nm, 10-cell CPML in x, y periodic), E_z and H_z:
  three rectangles painted in order: a mildly conductive one around the first source line (eps 2, sigma_e 5e3 S/m),
  a lossy slab (eps 3, sigma_e 2e4) and a strongly conductive block inside it (eps 1.5, sigma_e 3e5);
  a Ricker line source along y inside the first rectangle (600 nm), a cw line source over part of a line (700 nm,
  amplitude 0.3); a flux line behind the slab at 500 / 600 / 700 nm. Fields after 150 and 500 steps, both source
  profiles, the flux spectra and the two waveforms sampled at t = n dt + dt/2 for n < 600.
3D case (window 0.6 x 0.5 x 0.45 um, dx 25 nm, 8-cell CPML): two boxes with sigma_e (2e4 and 3e5 S/m) painted on
  the reference's float32 material arrays, a Ricker z dipole and a cw x dipole, a flux plane normal to x; fields
  after 250 steps, dipole weights and the plane's flux.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_lossy_golden.py
The fixtures hold inputs (scalars of the cases, the 3D eps / sigma arrays) and outputs (field arrays, flux spectra,
waveform samples); no source text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.cases import LOSSY2D, LOSSY3D, lossy3d_masks   # noqa: E402


def run_2d(polarization):
    sys.path.insert(0, REFERENCE)
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig

    spec = LOSSY2D
    bc = BoundaryConfig(*spec['faces'])
    bc.cpml_npml = spec['npml']
    domain = Domain.uniform(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'], boundary_config=bc,
                            verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization=polarization, boundary_config=bc,
                                         precision='double'))
    for x0, y0, x1, y1, eps, sigma in spec['rectangles']:
        em.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps, sigma_e=sigma)
    sources = [em.add_line_source(**source) for source in spec['sources']]
    monitor = em.add_flux_monitor(**spec['monitor'])
    names = ('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')
    out = {}
    done = 0
    for target in spec['steps']:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in names:
            out[f'{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    engine = em.engine
    dt = float(engine.dt)
    times = (np.arange(spec['waveform_samples'], dtype=np.float64) + 0.5) * dt
    out.update({'nx': engine.nx, 'ny': engine.ny, 'dt': dt, 'n_steps': engine.n_step,
                'sample_interval': monitor.sample_interval, 'flux': np.asarray(monitor.flux(), dtype=np.float64),
                'sigma_e': np.array(domain.materials.sigma_e, dtype=np.float64)})
    for index, source in enumerate(sources):
        out[f'source{index}_profile'] = np.asarray(source.profile, dtype=np.float64)
        out[f'source{index}_waveform'] = np.array([source.temporal(float(t)) for t in times])
    return out


def run_3d():
    sys.path.insert(0, REFERENCE)
    from open_sol.core.domain3d import Domain3D
    from open_sol.physics.em3d.boundaries3d import BoundaryConfig3D
    from open_sol.physics.em3d.physics3d import EMPhysics3D

    spec = LOSSY3D
    faces = {face: 'cpml' for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    bc = BoundaryConfig3D(faces, npml=spec['npml'])
    domain = Domain3D.uniform(*spec['window'], spec['dx'], npml_sides={face: spec['npml'] for face in faces},
                              boundary_config=bc)
    centres = domain.host_centres(True)
    eps = np.ones(domain.grid.shape, dtype=np.float32)
    sigma = np.zeros(domain.grid.shape, dtype=np.float32)
    for mask, (_, _, eps_r, sigma_e) in zip(lossy3d_masks(centres), spec['boxes']):
        eps[mask] = np.float32(eps_r)
        sigma[mask] = np.float32(sigma_e)
    for name, default in (('eps_r', 1.0), ('sigma_e', 0.0), ('mu_r', 1.0)):
        domain.materials.add_property(name, default)
    domain.materials.eps_r[...] = eps
    domain.materials.sigma_e[...] = sigma
    em = EMPhysics3D(domain, precision='single')
    dipoles = [em.add_dipole(*position, orientation=orientation, dipole_moment=moment, wavelength=wavelength,
                             bandwidth=0.5, waveform=waveform)
               for position, orientation, moment, wavelength, waveform in spec['dipoles']]
    plane = em.add_flux_plane('x', spec['plane_x'], wavelengths=list(spec['wavelengths']))
    em.finalize()
    engine = em.engine
    for _ in range(spec['n_steps']):
        em.step()
    out = {'shape': np.array(engine.shape), 'dt': np.float64(engine.dt), 'eps_r': eps, 'sigma_e': sigma,
           'sample_interval': np.int64(plane.sample_interval), 'plane_edge': np.int64(plane.edge),
           'flux': np.asarray(plane.flux(), dtype=np.float64), 'n_steps': np.int64(engine.n_step)}
    for index, dipole in enumerate(dipoles):
        out[f'dipole{index}_cell'] = np.array(dipole.cell)
        out[f'dipole{index}_weight'] = np.asarray(dipole.weight, dtype=np.float64)
    for name in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
        out[name] = np.array(getattr(engine.fields, name), dtype=np.float32)
    return out


def main():
    for polarization in ('ez', 'hz'):
        out = run_2d(polarization)
        path = os.path.join(HERE, f'lossy2d_{polarization}.npz')
        np.savez_compressed(path, **out)
        print(f'{path}: {out["nx"]} x {out["ny"]}, m {out["sample_interval"]}, flux {out["flux"]}')
    out = run_3d()
    path = os.path.join(HERE, 'lossy3d.npz')
    np.savez_compressed(path, **out)
    print(f'{path}: {out["shape"]}, m {out["sample_interval"]}, flux {out["flux"]}')


if __name__ == '__main__':
    main()
