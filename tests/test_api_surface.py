"""Attribute fall-through of the scripting API (simulation.py:250-270 of the reference): names the Simulation lacks
resolve on its one physics, the EM model; unknown names raise AttributeError naming that physics. No GPU needed."""
import copy

import pytest

from opensol_amd import Simulation, Simulation3D


def test_2d_attributes_fall_through_to_the_model():
    sim = Simulation(3e-6, 1.5e-6, 25e-9, npml=12)
    assert sim.nx == sim.model.nx and sim.ny == sim.model.ny
    assert sim.dt == sim.model.dt
    assert sim.em is sim.model
    # a model method reached through the fall-through acts on the same model
    sim.rect_mask(0.0, 0.0, 1e-6, 1e-6)
    with pytest.raises(AttributeError, match="neither has its only physics 'em'"):
        sim.no_such_attribute
    with pytest.raises(AttributeError):
        sim.__no_such_dunder__
    assert not hasattr(sim, 'no_such_attribute')


def test_3d_attributes_fall_through_to_the_model():
    sim = Simulation3D(1e-6, 1e-6, 1e-6, 50e-9, npml=4)
    assert sim.shape == sim.model.shape
    assert sim.em3d is sim.model
    with pytest.raises(AttributeError, match="'em3d'"):
        sim.no_such_attribute


def test_copy_probe_does_not_recurse():
    # copy.copy looks up dunders on an instance whose __dict__ is still empty: the fall-through must not recurse
    sim = Simulation(3e-6, 1.5e-6, 25e-9, npml=12)
    clone = copy.copy(sim)
    assert clone.model is sim.model
