"""Write tests/golden/pec2d_{ez,hz}.npz: golden vectors of small 2D FDTD cases with perfect electric conductors, from
the reference's own numpy path (EMPhysics precision='double', CPU backend). They pin the PEC rules end to end: E_z
pinned to 0 in PEC cells and on the outermost row of a 'pec' face (E_z), every E face touching a PEC cell pinned
(H_z), and the source weight times the keep mask.

Cases (window 2 um x 1 um, dx 25 nm, 10-cell CPML, y periodic):
  ez: x_lo 'pec' face, x_hi CPML; a PEC rectangle (add_pec_region on the reference's rectangle raster) partly covered
      by a dielectric painted after it; a Gaussian line source along y that crosses the PEC rectangle; a flux line.
  hz: CPML on both x faces; the same PEC rectangle and dielectric; the line source and flux line.
Fields are recorded after 150 and 500 steps, with the flux spectra and the source profile. Two running-DFT box
monitors at the flux line's wavelengths are recorded after 500 steps: one over the whole lower half of the window at
full resolution (it starts at y cell 0, where the periodic wrap pairs the H_z case's E_x with the last face), one
downsampled by 3.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_pec2d_golden.py
The fixture holds inputs (scalars of the case, the PEC mask) and outputs (field arrays, flux spectra); no source text
of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

WINDOW_X, WINDOW_Y, DX, NPML = 2e-6, 1e-6, 25e-9, 10
PEC_RECT = (0.9e-6, 0.3e-6, 1.1e-6, 0.65e-6)
DIELECTRIC = ((1.0e-6, 0.0, 1.6e-6, 1e-6), 3.0)
SOURCE_X, LAMBDA0, BANDWIDTH = 1.0e-6, 600e-9, 0.8
MONITOR_X, WAVELENGTHS = 1.7e-6, (500e-9, 600e-9, 700e-9)
STEPS = (150, 500)
DFT_BOXES = (((0.2e-6, 0.0, 1.8e-6, 0.5e-6), 1), ((0.3e-6, 0.1e-6, 1.9e-6, 0.9e-6), 3))
FACES = {'ez': ('pec', 'cpml', 'periodic', 'periodic'), 'hz': ('cpml', 'cpml', 'periodic', 'periodic')}


def run_case(polarization):
    sys.path.insert(0, REFERENCE)
    from open_sol import Simulation
    from open_sol.core.domain import Domain
    from open_sol.physics.em import EMPhysics, BoundaryConfig

    bc = BoundaryConfig(*FACES[polarization])
    bc.cpml_npml = NPML
    domain = Domain.uniform(WINDOW_X, WINDOW_Y, DX, npml=NPML, boundary_config=bc, verbose=False, backend='cpu')
    sim = Simulation.multiphysics(domain)
    em = sim.add_physics('em', EMPhysics(domain, polarization=polarization, boundary_config=bc,
                                         precision='double'))
    mask = np.asarray(domain.rect_mask(*PEC_RECT), dtype=bool)
    em.add_pec_region(mask)
    (x0, y0, x1, y1), eps = DIELECTRIC
    em.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    em.add_line_source(axis='y', position=SOURCE_X, wavelength=LAMBDA0, bandwidth=BANDWIDTH, waveform='gaussian')
    monitor = em.add_flux_monitor(axis='y', position=MONITOR_X, span=(0.0, WINDOW_Y), wavelengths=list(WAVELENGTHS))
    boxes = [em.add_dft_monitor(*box, wavelengths=list(WAVELENGTHS), downsample=downsample)
             for box, downsample in DFT_BOXES]
    names = ('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')
    out = {'pec_mask': mask}
    done = 0
    for target in STEPS:
        sim.run(n_steps=target - done, print_every=0, flush=False)
        done = target
        for name in names:
            out[f'{name}_{target}'] = np.array(getattr(em.fields, name), dtype=np.float64)
    engine = em.engine
    out.update({'nx': engine.nx, 'ny': engine.ny, 'dt': float(engine.dt), 'n_steps': engine.n_step,
                'sample_interval': monitor.sample_interval,
                'flux': np.asarray(monitor.flux(), dtype=np.float64),
                'source_profile': np.asarray(em.sources[0].profile, dtype=np.float64)})
    for index, box in enumerate(boxes):
        out[f'dft{index}_samples'] = np.array([[box.x_samples.start, box.x_samples.stop, box.x_samples.step],
                                               [box.y_samples.start, box.y_samples.stop, box.y_samples.step]])
        for name in (('ez',) if polarization == 'ez' else ('hz', 'ex', 'ey')):
            out[f'dft{index}_{name}'] = np.array(getattr(box, f'{name}_dft'))
    return out


def main():
    for polarization in ('ez', 'hz'):
        out = run_case(polarization)
        path = os.path.join(HERE, f'pec2d_{polarization}.npz')
        np.savez_compressed(path, **out)
        print(f'{path}: {out["nx"]} x {out["ny"]}, m {out["sample_interval"]}, flux {out["flux"]}, '
              f'PEC cells {int(out["pec_mask"].sum())}')


if __name__ == '__main__':
    main()
