"""Write tests/golden/poles3d.npz: golden vectors of a small 3D FDTD case with dispersive poles, from the reference's
own numpy path (EMPhysics3D, precision 'single', CPU backend). They pin the 3D pole records: the face average of a
pole's cell mask on each E component, J / P from E^n and the corrections after the E update.

Case: window 0.6 x 0.5 x 0.45 um, dx 25 nm, 8-cell CPML on every face, eps_r 2 painted on a box that carries a Drude
pole (omega_p 3e15, gamma 1e14) and a Lorentz pole on a second box (delta_eps 1.5, omega_0 3e15, gamma 2e14), a z
point dipole (lambda0 0.6 um), a flux plane normal to x; 250 steps. All six fields after the last step, the flux,
the masks and the eps_r array.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_poles3d_golden.py
The fixture holds inputs (scalars, masks, eps_r) and outputs (fields, flux); no source text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

WINDOW = (0.6e-6, 0.5e-6, 0.45e-6)
DX, NPML = 25e-9, 8
DRUDE_BOX = ((0.05e-6, 0.05e-6, 0.05e-6), (0.25e-6, 0.30e-6, 0.20e-6))
DRUDE = dict(omega_p=3e15, gamma=1e14)
DRUDE_EPS = 2.0
LORENTZ_BOX = ((0.35e-6, 0.20e-6, 0.25e-6), (0.55e-6, 0.45e-6, 0.40e-6))
LORENTZ = dict(delta_eps=1.5, omega_0=3e15, gamma=2e14)
DIPOLE = (0.30e-6, 0.24e-6, 0.22e-6)
LAMBDA0, BANDWIDTH, MOMENT = 0.6e-6, 0.5, 1e-12
PLANE_X, WAVELENGTHS = 0.45e-6, (0.5e-6, 0.6e-6, 0.7e-6)
N_STEPS = 250
FIELDS = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')


def box_mask(centres, lo, hi):
    axis_masks = [(c >= np.float32(a)) & (c < np.float32(b)) for c, a, b in zip(centres, lo, hi)]
    return axis_masks[0][:, None, None] & axis_masks[1][None, :, None] & axis_masks[2][None, None, :]


def main():
    sys.path.insert(0, REFERENCE)
    from open_sol.core.domain3d import Domain3D
    from open_sol.physics.em.materials import DrudePole, LorentzPole
    from open_sol.physics.em3d.boundaries3d import BoundaryConfig3D
    from open_sol.physics.em3d.physics3d import EMPhysics3D

    faces = {face: 'cpml' for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    bc = BoundaryConfig3D(faces, npml=NPML)
    domain = Domain3D.uniform(*WINDOW, DX, npml_sides={face: NPML for face in faces}, boundary_config=bc)
    centres = domain.host_centres(True)
    drude_mask = box_mask(centres, *DRUDE_BOX)
    lorentz_mask = box_mask(centres, *LORENTZ_BOX)
    eps = np.ones(domain.grid.shape, dtype=np.float32)
    eps[drude_mask] = np.float32(DRUDE_EPS)
    for name, default in (('eps_r', 1.0), ('sigma_e', 0.0), ('mu_r', 1.0)):
        domain.materials.add_property(name, default)
    domain.materials.eps_r[...] = eps
    em = EMPhysics3D(domain, precision='single')
    window = tuple(slice(0, n) for n in domain.grid.shape)
    em.add_dispersive(DrudePole(**DRUDE), window, drude_mask)
    em.add_dispersive(LorentzPole(**LORENTZ), window, lorentz_mask)
    em.add_dipole(*DIPOLE, orientation='z', dipole_moment=MOMENT, wavelength=LAMBDA0, bandwidth=BANDWIDTH)
    plane = em.add_flux_plane('x', PLANE_X, wavelengths=list(WAVELENGTHS))
    em.finalize()
    for _ in range(N_STEPS):
        em.step()
    engine = em.engine
    out = {'shape': np.array(engine.shape), 'dt': np.float64(engine.dt), 'eps_r': eps, 'drude_mask': drude_mask,
           'lorentz_mask': lorentz_mask, 'n_steps': np.int64(engine.n_step),
           'flux': np.asarray(plane.flux(), dtype=np.float64)}
    for name in FIELDS:
        out[name] = np.array(getattr(engine.fields, name), dtype=np.float32)
    path = os.path.join(HERE, 'poles3d.npz')
    np.savez_compressed(path, **out)
    print(path, out['shape'], 'flux', out['flux'], 'max |ez|', float(np.abs(out['ez']).max()))


if __name__ == '__main__':
    main()
