"""Write tests/golden/tfsf2d_{ez,hz,ez_mirror}.npz: golden vectors of 2D TFSF plane-wave boxes from the reference's own
numpy path (EMPhysics precision='double', CPU backend). They pin the total-field / scattered-field box end to end: the
1D incident line (margins, soft source scale, Mur ends), the H-side and E-side box-edge corrections with their
coefficients and timing, and a y mirror face the box opens onto.

Cases (window 2 um x 1.4 um, dx 25 nm, 10-cell CPML; the spec is tests/cases.py TFSF2D): a +x Gaussian plane wave
(600 nm, bandwidth 0.5, amplitude 1 V/m) in the box [0.4, 1.6] x [0.2, 1.2] um around a dielectric rectangle
(eps 4); E_z, H_z, and E_z with a symmetric y_lo mirror and the box [0.4, 1.6] x [0, 1.0] um opening onto it. A flux
line behind the box (scattered field) and one inside it (total field). Fields after 150 and 400 steps, the box cells,
the flux spectra.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_tfsf_golden.py
The fixtures hold inputs (scalars of the cases) and outputs (field arrays, flux spectra); no source text of the
reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.cases import TFSF2D, build_tfsf2d   # noqa: E402


class _ReferenceTarget:
    """The reference's Simulation + EMPhysics(precision='double') behind the constructor / method names that
    tests.cases builders call."""

    def __init__(self, size_x, size_y, dx, npml, boundaries, polarization):
        sys.path.insert(0, REFERENCE)
        from open_sol import Simulation
        from open_sol.core.domain import Domain
        from open_sol.physics.em import EMPhysics
        domain = Domain.uniform(size_x, size_y, dx, npml=npml, boundary_config=boundaries, verbose=False,
                                backend='cpu')
        self.sim = Simulation.multiphysics(domain)
        self.em = self.sim.add_physics('em', EMPhysics(domain, polarization=polarization, boundary_config=boundaries,
                                                      precision='double'))
        self.names = ('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')

    def __getattr__(self, name):
        return getattr(self.em, name)


def _boundaries(faces, npml):
    sys.path.insert(0, REFERENCE)
    from open_sol.physics.em import BoundaryConfig
    bc = BoundaryConfig(*faces)
    bc.cpml_npml = npml
    return bc


def run_case(case):
    target, source, monitors = build_tfsf2d(_ReferenceTarget, case, _boundaries)
    out = {'cells': np.array([source.x0, source.x1, source.y0, source.y1])}
    done = 0
    for step in TFSF2D['steps']:
        target.sim.run(n_steps=step - done, print_every=0, flush=False)
        done = step
        for name in target.names:
            out[f'{name}_{step}'] = np.array(getattr(target.em.fields, name), dtype=np.float64)
    engine = target.em.engine
    out.update({'nx': engine.nx, 'ny': engine.ny, 'dt': float(engine.dt), 'n_steps': engine.n_step,
                'sample_interval': monitors[0].sample_interval})
    for index, monitor in enumerate(monitors):
        out[f'flux{index}'] = np.asarray(monitor.flux(), dtype=np.float64)
    return out


def main():
    for case in ('ez', 'hz', 'ez_mirror'):
        out = run_case(case)
        path = os.path.join(HERE, f'tfsf2d_{case}.npz')
        np.savez_compressed(path, **out)
        print(f'{path}: {out["nx"]} x {out["ny"]}, cells {out["cells"]}, m {out["sample_interval"]}, '
              f'flux {out["flux0"]} / {out["flux1"]}')


if __name__ == '__main__':
    main()
