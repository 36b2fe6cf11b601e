"""The scripting API on the device: the pole case painted with add_material runs bit for bit as the one painted with
add_drude_rectangle / add_dispersive (itself checked against the oracle), and memory_report counts the engine's
device arrays by owner, each storage once, with nothing left after clear()."""
import io

import numpy as np
import pytest

from tests.cases import POLES2D, build_poles2d

pytestmark = pytest.mark.gpu


def _painted_with_add_material(polarization):
    from opensol_amd import LorentzPole, Simulation
    spec = POLES2D
    sim = Simulation(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'], boundaries=spec['boundaries'],
                     polarization=polarization)
    (x0, y0, x1, y1), eps = spec['dielectric']
    (cx, cy), radius = spec['lorentz_disc']
    sim.add_material(sim.domain.rect_mask(x0, y0, x1, y1), eps_r=eps)
    sim.add_material(sim.domain.rect_mask(*spec['drude_rect']), dispersion=[dict(model='drude', **spec['drude'])])
    sim.add_material(sim.domain.circle_mask(cx, cy, radius), eps_r=eps, dispersion=[LorentzPole(**spec['lorentz'])])
    sim.add_pec_region(sim.domain.rect_mask(*spec['pec_rect']))
    sim.add_line_source(**spec['source'])
    return sim, sim.add_flux_monitor(**spec['monitor'])


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_add_material_pole_case_matches_the_painters(gpu, polarization):
    from opensol_amd import Simulation
    print(f'pole case {polarization}: painters vs add_material, 240 steps', flush=True)
    a, monitor_a = build_poles2d(Simulation, polarization)
    b, monitor_b = _painted_with_add_material(polarization)
    for sim in (a, b):
        sim.run(n_steps=240, print_every=0)
    for name in a.model.components:
        assert np.array_equal(a.engine.field(name), b.engine.field(name)), name
    assert np.array_equal(np.asarray(monitor_a.flux()), np.asarray(monitor_b.flux()))
    a.close()
    b.clear()


def test_memory_report_owners(gpu):
    import torch
    from opensol_amd import Simulation
    print('memory_report of the pole case', flush=True)
    sim, _ = build_poles2d(Simulation, 'ez')
    sim.finalize()
    engine = sim.engine
    out = io.StringIO()
    report = sim.memory_report(file=out)
    owners = report['owners']
    table2 = engine.table2.untyped_storage().nbytes() if engine.table2 is not None else 0
    assert owners['materials'] == engine.material.untyped_storage().nbytes() + \
        engine.table.untyped_storage().nbytes() + table2
    assert owners['em'] >= engine.fields.untyped_storage().nbytes() > 0
    assert owners['couplings'] == 0
    status = report['status']
    assert status['kind'] == 'gpu' and status['budget'] is None
    assert 0 < status['pool_used'] <= status['pool_reserved'] and status['device_free'] < status['device_total']
    assert owners['materials'] + owners['em'] <= status['pool_used']
    assert 'GPU memory' in out.getvalue()
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    sim.clear()
    del engine
    assert torch.cuda.memory_allocated() <= before - owners['em'] - owners['materials']


def test_memory_report_3d(gpu):
    from opensol_amd import Simulation3D
    print('memory_report of a 3D model', flush=True)
    sim = Simulation3D(0.4e-6, 0.35e-6, 0.3e-6, 25e-9, npml=4)
    sim.add_material(sim.domain.box_mask((0.05e-6, 0.1e-6, 0.05e-6), (0.2e-6, 0.3e-6, 0.25e-6)), eps_r=2.5)
    sim.add_dipole(0.2e-6, 0.17e-6, 0.15e-6, orientation='z', dipole_moment=1e-12, wavelength=0.6e-6)
    sim.run(n_steps=10)
    owners = sim.memory_report(file=io.StringIO())['owners']
    assert set(owners) == {'materials', 'em3d', 'couplings'}
    assert owners['materials'] == sim.engine.material.untyped_storage().nbytes() + \
        sim.engine.table.untyped_storage().nbytes()
    assert owners['em3d'] >= sim.engine.fields.untyped_storage().nbytes() > 0
    sim.clear()
