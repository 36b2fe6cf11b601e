"""Write tests/golden/hzwalls2d.npz: golden vectors of 2D H_z cases bounded by pec and mirror faces, from the
reference's own numpy path (EMPhysics precision='double', CPU backend). They pin the H_z outer-face rules: pec and
antisymmetric faces copy the inner H_z row (H_z even) and pin the tangential E on the plane face to 0, symmetric faces
copy it negated (H_z odd); mirror faces pad one ghost cell, walls none; a corner takes both faces' factors; the rows
are pinned again after the sources inject.

Cases (tests/cases.py HZWALLS2D): 'a' x_lo pec, x_hi symmetric, y_lo antisymmetric, y_hi pec; 'b' CPML on the lo
faces, x_hi symmetric, y_hi pec with 129 columns. Fields after 150 and 400 steps.

Run in this container only (the reference is not present on the GPU box):
    python tests/golden/make_hzwalls2d_golden.py
The fixture holds inputs (scalars of the cases) and outputs (field arrays); no source text of the reference.
"""
import os
import sys

import numpy as np

REFERENCE = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.cases import HZWALLS2D, build_hzwalls2d   # noqa: E402


class _ReferenceTarget:
    """The reference's Simulation + EMPhysics(precision='double') behind the names tests.cases builders call."""

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

    def __getattr__(self, name):
        return getattr(self.em, name)


def _boundaries(faces, npml):
    sys.path.insert(0, REFERENCE)
    from open_sol.physics.em import BoundaryConfig
    bc = BoundaryConfig(*faces)
    bc.cpml_npml = npml
    return bc


def main():
    out = {}
    for name in HZWALLS2D['cases']:
        target = build_hzwalls2d(_ReferenceTarget, name, _boundaries)
        done = 0
        for step in HZWALLS2D['steps']:
            target.sim.run(n_steps=step - done, print_every=0, flush=False)
            done = step
            for field in ('hz', 'ex', 'ey'):
                out[f'{name}_{field}_{step}'] = np.array(getattr(target.em.fields, field), dtype=np.float64)
        out[f'{name}_shape'] = np.array([target.em.engine.nx, target.em.engine.ny])
        print(name, out[f'{name}_shape'], np.abs(out[f'{name}_hz_400']).max())
    np.savez_compressed(os.path.join(HERE, 'hzwalls2d.npz'), **out)


if __name__ == '__main__':
    main()
