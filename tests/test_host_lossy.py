"""The product's host setup of conductive media and of the Ricker / cw waveforms against the reference's golden runs
(tests/golden/lossy*.npz, written by tests/golden/make_lossy_golden.py), on the CPU: the sigma_e raster, the sheet
calibration in a conductive medium, the float32 waveform tables, the DFT sample interval a cw source forces, the 3D
dipole weights with sigma_e in the face cb, and the semi-implicit coefficient tables against their closed form."""
import os

import numpy as np
import pytest

from opensol_amd.boundaries import BoundaryConfig
from opensol_amd.fdtd2d import Model2D
from opensol_amd.fdtd3d import Model3D
from tests.cases import LOSSY2D, build_lossy2d, build_lossy3d

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
EPS0 = 8.8541878128e-12


def _boundaries(faces, npml):
    bc = BoundaryConfig(*faces)
    bc.cpml_npml = npml
    return bc


@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_lossy_host_setup_matches_golden(polarization):
    golden = np.load(os.path.join(GOLDEN, f'lossy2d_{polarization}.npz'))
    model, sources, _ = build_lossy2d(Model2D, polarization, _boundaries)
    model.finalize()
    assert model.shape == (int(golden['nx']), int(golden['ny']))
    assert float(model.dt) == float(golden['dt'])
    np.testing.assert_array_equal(model.materials.property_array(3).astype(np.float64), golden['sigma_e'])
    assert model.sample_interval == int(golden['sample_interval']) == 1
    times = (np.arange(LOSSY2D['waveform_samples'], dtype=np.float64) + 0.5) * float(model.dt)
    for index, source in enumerate(sources):
        np.testing.assert_array_equal(source.profile.astype(np.float64), golden[f'source{index}_profile'])
        np.testing.assert_array_equal([source.temporal(float(t)) for t in times], golden[f'source{index}_waveform'])


def test_lossy_coefficient_tables_closed_form():
    """ca = (1 - s) / (1 + s), cb = dt / (eps0 eps_r (1 + s)), s = sigma dt / (2 eps0 eps_r), per material id (E_z
    table) and on the H_z faces between two ids (float32 means of eps and sigma)."""
    model, _, _ = build_lossy2d(Model2D, 'ez', _boundaries)
    model.finalize()
    dt = float(model.dt)
    values = model.materials.values().astype(np.float64)
    table = model.materials.table_ez(model.dt).astype(np.float64)
    s = values[:, 3] * dt / (2.0 * EPS0 * values[:, 0])
    np.testing.assert_allclose(table[:, 0], (1.0 - s) / (1.0 + s), rtol=2e-7)
    np.testing.assert_allclose(table[:, 1], dt / (EPS0 * values[:, 0] * (1.0 + s)), rtol=2e-7)
    assert np.max(s) > 0.05


def test_lossy_3d_dipole_weights_match_golden():
    golden = np.load(os.path.join(GOLDEN, 'lossy3d.npz'))
    model, dipoles, plane = build_lossy3d(Model3D)
    model.finalize()
    assert model.shape == tuple(golden['shape'])
    assert float(model.dt) == float(golden['dt'])
    np.testing.assert_array_equal(model.materials.property_array(0), golden['eps_r'])
    np.testing.assert_array_equal(model.materials.property_array(3), golden['sigma_e'])
    for index, dipole in enumerate(dipoles):
        assert dipole.cell == tuple(golden[f'dipole{index}_cell'])
        assert abs(float(dipole.weight) / float(golden[f'dipole{index}_weight'][0]) - 1.0) < 3e-7
    assert model.sample_interval == int(golden['sample_interval']) == 1
    assert plane.edge == int(golden['plane_edge'])
