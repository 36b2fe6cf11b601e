"""Write tests/golden/poles2d_{ez,hz}.npz: golden vectors of a small 2D FDTD case with dispersive poles, from the
reference's own numpy path (EMPhysics precision='double', CPU backend). They pin the pole updates (Drude J and
Lorentz P in velocity form from E^n), their E corrections after the Yee update, the eps_inf background of a Drude
rectangle and the PEC keep folded into a pole's drive; in H_z, the poles sit on the E_x / E_y faces with the face
average of the cell mask as their weight.

Case (window 2 um x 1 um, dx 25 nm, 10-cell CPML, y periodic): a Drude rectangle (eps_inf 2) with a PEC rectangle
over one corner of it, a Lorentz disc over a dielectric rectangle (eps_r 2), a Gaussian line source along y and a
flux line behind the media. Fields after 300 and 800 steps, the flux spectrum, the masks and the eps_r array.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_poles2d_golden.py
The fixture holds inputs (scalars of the case, masks, eps_r) and outputs (fields, flux); no source text of the
reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

WINDOW_X, WINDOW_Y, DX, NPML = 2e-6, 1e-6, 25e-9, 10
DRUDE_RECT = (1.0e-6, 0.2e-6, 1.3e-6, 0.7e-6)
DRUDE = dict(omega_p=5e15, gamma=1e14, eps_inf=2.0)
PEC_RECT = (1.2e-6, 0.6e-6, 1.4e-6, 0.8e-6)
DIELECTRIC = ((0.5e-6, 0.1e-6, 0.9e-6, 0.9e-6), 2.0)
LORENTZ_DISC = ((0.7e-6, 0.5e-6), 0.15e-6)
LORENTZ = dict(delta_eps=1.5, omega_0=3e15, gamma=2e14)
SOURCE_X, LAMBDA0, BANDWIDTH = 0.35e-6, 600e-9, 0.8
MONITOR_X, WAVELENGTHS = 1.6e-6, (500e-9, 600e-9, 700e-9)
STEPS = (300, 800)


def main():
    for polarization in ('ez', 'hz'):
        run_case(polarization)


def run_case(polarization):
    sys.path.insert(0, REFERENCE)
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig
    from open_sol.physics.em.materials import DrudePole, LorentzPole

    bc = BoundaryConfig('cpml', 'cpml', 'periodic', 'periodic')
    bc.cpml_npml = NPML
    domain = Domain.uniform(WINDOW_X, WINDOW_Y, DX, npml=NPML, boundary_config=bc, verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization=polarization, boundary_config=bc, precision='double'))
    (x0, y0, x1, y1), eps = DIELECTRIC
    em.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    em.add_drude_rectangle(*DRUDE_RECT, DrudePole(**DRUDE))
    (cx, cy), radius = LORENTZ_DISC
    disc = np.asarray(domain.circle_mask(cx, cy, radius), dtype=bool)
    em.add_dispersive(LorentzPole(**LORENTZ), disc)
    pec = np.asarray(domain.rect_mask(*PEC_RECT), dtype=bool)
    em.add_pec_region(pec)
    em.add_line_source(axis='y', position=SOURCE_X, wavelength=LAMBDA0, bandwidth=BANDWIDTH, waveform='gaussian')
    monitor = em.add_flux_monitor(axis='y', position=MONITOR_X, span=(0.0, WINDOW_Y), wavelengths=list(WAVELENGTHS))
    out = {'drude_mask': np.asarray(domain.rect_mask(*DRUDE_RECT), dtype=bool), 'lorentz_mask': disc,
           'pec_mask': pec}
    done = 0
    for target in STEPS:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in (('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')):
            out[f'{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    engine = em.engine
    out.update({'nx': engine.nx, 'ny': engine.ny, 'dt': float(engine.dt), 'n_steps': engine.n_step,
                'eps_r': np.array(domain.materials.eps_r if polarization == 'ez' else domain.materials.eps_xx,
                                  dtype=np.float64),
                'sample_interval': monitor.sample_interval,
                'flux': np.asarray(monitor.flux(), dtype=np.float64)})
    path = os.path.join(HERE, f'poles2d_{polarization}.npz')
    np.savez_compressed(path, **out)
    print(path, out['nx'], 'x', out['ny'], 'flux', out['flux'])


if __name__ == '__main__':
    main()
