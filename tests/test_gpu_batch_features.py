"""Batched members with the wider features: every member of a Batch2D / Batch3D that carries walls, mirror faces,
TFSF boxes, dispersive poles, field probes or DFT monitors ends with the same fields, flux spectra, probe series and
DFT spectra, bit for bit, as the same model run alone (each alone-run is itself checked against the oracle by the
single-instance tests). The 2D mix holds every E_z and H_z wall kind (H_z 'b' puts a y_hi pec wall on the first
column of a 128-column tile, the case that needs the lane-0 recompute), a TFSF box opening onto a symmetric mirror,
poles in both polarizations and a plain member; members run different step counts and then continue together."""
import numpy as np
import pytest

from tests.cases import (add_pec2d_dft_boxes, build_hzwalls2d, build_mirrors3d, build_pec2d, build_poles2d,
                         build_poles3d, build_tfsf2d, build_walls2d)

pytestmark = pytest.mark.gpu


def _bc(faces, npml):
    from opensol_amd.boundaries import BoundaryConfig
    bc = BoundaryConfig(*faces)
    bc.cpml_npml = npml
    return bc


def _members_2d():
    """[(name, build() -> (sim, readouts), steps)]; readouts are the monitor / probe objects to compare."""
    from opensol_amd import Simulation

    def walls(polarization):
        sim = build_walls2d(Simulation, _bc, polarization)
        probes = [sim.add_field_probe(0.3e-6, 0.4e-6), sim.add_field_probe(0.0, 0.0)]
        box = sim.add_dft_monitor(0.1e-6, 0.1e-6, 0.9e-6, 0.8e-6, wavelengths=[550e-9, 650e-9], downsample=2)
        return sim, probes + [box]

    def pec(polarization):
        sim, _, monitor = build_pec2d(Simulation, polarization, _bc)
        return sim, [monitor] + add_pec2d_dft_boxes(sim)

    def tfsf(case):
        sim, _, monitors = build_tfsf2d(Simulation, case, _bc)
        return sim, list(monitors) + [sim.add_field_probe(1.0e-6, 0.7e-6)]

    def poles(polarization):
        sim, monitor = build_poles2d(Simulation, polarization)
        return sim, [monitor]

    def hzwalls(name):
        sim = build_hzwalls2d(Simulation, name, _bc)
        return sim, [sim.add_field_probe(0.5e-6, 0.5e-6, 'ey')]

    def plain():
        sim = Simulation(1.5e-6, 0.9e-6, 25e-9, npml=8, polarization='ez')
        sim.add_dielectric_rectangle(0.5e-6, 0.2e-6, 0.9e-6, 0.6e-6, eps_r=4.0)
        sim.add_line_source(axis='y', position=0.3e-6, wavelength=600e-9, bandwidth=0.6)
        return sim, [sim.add_flux_monitor(axis='y', position=1.2e-6, span=(0.0, 0.9e-6), wavelengths=[600e-9])]

    return [('walls ez', lambda: walls('ez'), 140), ('walls hz', lambda: walls('hz'), 95),
            ('pec ez', lambda: pec('ez'), 120), ('pec hz', lambda: pec('hz'), 77),
            ('tfsf ez_mirror', lambda: tfsf('ez_mirror'), 150), ('tfsf hz', lambda: tfsf('hz'), 110),
            ('poles ez', lambda: poles('ez'), 160), ('poles hz', lambda: poles('hz'), 60),
            ('hzwalls a', lambda: hzwalls('a'), 130), ('hzwalls b', lambda: hzwalls('b'), 101),
            ('plain', plain, 45)]


def _members_3d():
    from opensol_amd import Simulation3D

    def mirrors(name):
        sim, _, plane = build_mirrors3d(Simulation3D, name)
        return sim, [plane, sim.add_field_probe(0.2e-6, 0.2e-6, 0.2e-6, 'ex'),
                     sim.add_dft_plane(0.2e-6, wavelengths=[0.6e-6])]

    def poles():
        sim, plane = build_poles3d(Simulation3D)
        return sim, [plane, sim.add_field_probe(0.3e-6, 0.24e-6, 0.22e-6, 'ez')]

    def plain():
        sim = Simulation3D(0.4e-6, 0.35e-6, 0.5e-6, 25e-9, npml=6)
        sim.add_dielectric_box((0.1e-6, 0.1e-6, 0.1e-6), (0.3e-6, 0.25e-6, 0.3e-6), 2.5)
        sim.add_dipole(0.2e-6, 0.17e-6, 0.35e-6, orientation='y', dipole_moment=1e-12, wavelength=0.6e-6)
        return sim, [sim.add_flux_plane('z', 0.1e-6, wavelengths=[0.6e-6])]

    return [('mirrors a', lambda: mirrors('a'), 70), ('mirrors b', lambda: mirrors('b'), 45),
            ('poles', poles, 90), ('plain', plain, 30)]


def _readout(sim, readouts):
    sim.engine.synchronize()
    out = {f'field {name}': sim.engine.field(name) for name in sim.model.components} if len(sim.model.shape) == 2 \
        else {f'field {name}': sim.engine.field(name) for name in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')}
    for k, item in enumerate(readouts):
        if hasattr(item, 'get_data'):
            out[f'probe {k}'] = item.get_data()[1]
        elif hasattr(item, 'flux'):
            out[f'flux {k}'] = np.asarray(item.flux())
        else:
            out[f'dft {k}'] = item._spectra()
    return out


def _compare(name, got, want):
    assert got.keys() == want.keys(), name
    for key in want:
        assert np.array_equal(got[key], want[key], equal_nan=True), (name, key)


@pytest.mark.parametrize('three_d', [False, True], ids=['2d', '3d'])
def test_batched_members_with_features_match_solo_runs(gpu, three_d):
    from opensol_amd.batch import Batch2D, Batch3D
    members = _members_3d() if three_d else _members_2d()
    extra = 23
    solo = []
    for name, build, steps in members:
        print(f'solo {name}: {steps} + {extra} steps', flush=True)
        sim, readouts = build()
        sim.run(n_steps=steps, print_every=0)
        first = _readout(sim, readouts)
        sim.run(n_steps=extra, print_every=0)
        solo.append((first, _readout(sim, readouts)))
        sim.close()
    made = [build() for _, build, _ in members]
    print(f'batch of {len(made)} members', flush=True)
    batch = (Batch3D if three_d else Batch2D)([sim for sim, _ in made], graph_steps=8)
    batch.run([steps for _, _, steps in members])
    for (name, _, _), (sim, readouts), (first, _) in zip(members, made, solo):
        _compare(name, _readout(sim, readouts), first)
    batch.run(extra)   # every member continues together
    for (name, _, _), (sim, readouts), (_, second) in zip(members, made, solo):
        _compare(name, _readout(sim, readouts), second)
    batch.close()
    for sim, _ in made:
        sim.close()
