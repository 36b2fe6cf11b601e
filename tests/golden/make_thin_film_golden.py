"""Write tests/golden/thin_film_{ez,hz}.npz: golden vectors of the thin-film FDTD case from the reference's own
numpy path (EMPhysics precision='double', CPU backend), for the end-to-end parity that no closed form pins (CPML
profile, sheet calibration, DFT decimation and the flux formula composed at the 1e-6 level).

The case is examples/em/thin_film.py with: window 6 um x 3 um, dx 25 nm, 15-cell PML in x, periodic y
(270 x 120 cells), slab eps_r 2.25 for x in [2.75, 3.25) um, Gaussian line source along y at x = 1.2 um
(lambda0 600 nm, bandwidth 0.8), flux line along y at x = 4.5 um at 500 and 600 nm, 80 fs; plus the empty-domain
run that normalises the transmission.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_thin_film_golden.py
The fixture holds inputs (scalars of the case) and outputs (field arrays at fixed steps, flux spectra); no source
text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

SNAPSHOT_STEPS = (60, 165, 400)   # field arrays recorded after these steps
WINDOW_X, WINDOW_Y, DX, NPML = 6e-6, 3e-6, 25e-9, 15
SLAB_X0, SLAB_X1, SLAB_EPS = 2.75e-6, 3.25e-6, 2.25
SOURCE_X, LAMBDA0, BANDWIDTH = 1.2e-6, 600e-9, 0.8
MONITOR_X, WAVELENGTHS = 4.5e-6, (500e-9, 600e-9)
DURATION = 80e-15


def run_case(polarization, with_slab):
    sys.path.insert(0, REFERENCE)
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig

    bc = BoundaryConfig.with_periodic_y(NPML)
    domain = Domain.uniform(WINDOW_X, WINDOW_Y, DX, npml=NPML, boundary_config=bc, verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization=polarization, boundary_config=bc,
                                         precision='double'))
    if with_slab:
        em.add_dielectric_rectangle(SLAB_X0, 0.0, SLAB_X1, WINDOW_Y, eps_r=SLAB_EPS)
    em.add_line_source(axis='y', position=SOURCE_X, wavelength=LAMBDA0, bandwidth=BANDWIDTH, waveform='gaussian')
    monitor = em.add_flux_monitor(axis='y', position=MONITOR_X, span=(0.0, WINDOW_Y), wavelengths=list(WAVELENGTHS))

    names = ('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')
    snapshots = {}
    done = 0
    for target in SNAPSHOT_STEPS:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in names:
            snapshots[f'{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    t_end = DURATION
    sim.run(duration=t_end - sim.runner.t, print_every=0)
    engine = em.engine
    out = {
        'nx': engine.nx, 'ny': engine.ny, 'dt': float(engine.dt), 'n_steps': engine.n_step,
        'time': float(engine.time), 'sample_interval': monitor.sample_interval,
        'frequencies': np.asarray(monitor.frequencies, dtype=np.float64),
        'flux': np.asarray(monitor.flux(), dtype=np.float64),
        'e_spectrum': np.asarray(monitor._spectra.host(0)), 'h_spectrum': np.asarray(monitor._spectra.host(1)),
        'source_profile': np.asarray(em.sources[0].profile, dtype=np.float64),
        'source_position': em.sources[0].position,
        'monitor_position': monitor.position, 'monitor_span': np.asarray(monitor.span),
    }
    for name in names:
        out[f'{name}_final'] = np.array(getattr(em.fields, name), dtype=np.float64)
    out.update(snapshots)
    return out


def main():
    for polarization in ('ez', 'hz'):
        film = run_case(polarization, True)
        empty = run_case(polarization, False)
        fixture = {f'film_{key}': value for key, value in film.items()}
        fixture.update({f'empty_{key}': value for key, value in empty.items()})
        fixture['transmission'] = film['flux'] / empty['flux']
        fixture['snapshot_steps'] = np.asarray(SNAPSHOT_STEPS)
        path = os.path.join(HERE, f'thin_film_{polarization}.npz')
        np.savez_compressed(path, **fixture)
        print(f'{path}: n_steps {film["n_steps"]}, m {film["sample_interval"]}, T {fixture["transmission"]}')


if __name__ == '__main__':
    main()
