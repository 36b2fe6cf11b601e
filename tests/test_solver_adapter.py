"""The PhysicsSolver adapter (opensol_amd.solver.FDTDSolver) driven the way the reference's MultiphysicsRunner drives a
solver, and snapshot / restore of resident state.

`runner_advance` below restates the runner's tick loop for one solver (coupling/runner.py:74-118 prepare: dt_sync =
min dt_limit, the hand-off to advance_ticks for a single stride-1 solver without couplings; :174-204 advance: stop
once t >= t_end - dt_sync / 2, step(t, span) per tick, t += dt_sync). CPU tests cover the contract surface; GPU tests
check that every way of driving the engine gives the same bits as Simulation.run, and that a restored snapshot
continues bit for bit."""
import math

import numpy as np
import pytest

from tests.cases import THIN_FILM, build


def runner_advance(solver, n_ticks, t_end=None, t=0.0, hand_off=True):
    dt_sync = min(float(solver.dt_limit()), math.inf)
    assert dt_sync > 0 and math.isfinite(dt_sync)
    stop_time = math.inf if t_end is None else t_end - 0.5 * dt_sync
    if hand_off and callable(getattr(solver, 'advance_ticks', None)):
        return solver.advance_ticks(int(n_ticks), t, dt_sync, stop_time)
    n_done = 0
    while n_done < n_ticks and t < stop_time:
        solver.step(t, dt_sync)
        t = t + dt_sync
        n_done += 1
    return n_done, t


def _box_spectra(box):
    return {name: getattr(box, f'{name}_dft') for name in box.names}


def test_contract_surface_on_the_host():
    from opensol_amd import Simulation
    from opensol_amd.solver import FDTDSolver
    sim, _, _ = build(Simulation, THIN_FILM, 'ez', True)
    solver = FDTDSolver(sim)
    for method in ('prepare', 'step', 'dt_limit', 'state', 'inject_source', 'finalize', 'materials_changed',
                   'scalars', 'advance_ticks', 'solution_state', 'restore_solution'):
        assert callable(getattr(solver, method)), method
    assert solver.name == 'em' and solver.dt_limit() == float(sim.model.dt) and solver.scalars() == {}
    solver.prepare(sim.domain)
    with pytest.raises(NotImplementedError, match="does not accept external source 'ez_source'"):
        solver.inject_source('ez_source', np.zeros(sim.model.shape, np.float32))
    with pytest.raises(ValueError, match='needs an array of shape'):
        solver.inject_source('eps_r', np.ones((3, 3), np.float32))
    eps = np.full(sim.model.shape, 2.0, np.float32)
    eps[:, :10] = 3.5
    version = sim.model.materials.version
    solver.inject_source('eps_r', eps)
    assert sim.model.materials.version > version
    np.testing.assert_array_equal(sim.model.materials.property_array(0), eps)


@pytest.mark.gpu
@pytest.mark.parametrize('hand_off', [True, False])
def test_runner_drives_the_engine_like_simulation_run(gpu, hand_off):
    """A runner-driven solver (hand-off or one step per tick, in chunks, stopping at t_end - dt/2) gives the fields,
    flux and clock of Simulation.run(duration) bit for bit."""
    from opensol_amd import Simulation
    from opensol_amd.solver import FDTDSolver
    duration = 12e-15
    plain, _, plain_monitor = build(Simulation, THIN_FILM, 'ez', True)
    plain.run(duration=duration, print_every=0)
    sim, _, monitor = build(Simulation, THIN_FILM, 'ez', True, graph_steps=16)
    solver = FDTDSolver(sim)
    solver.prepare(sim.domain)
    solver.finalize()
    t, ticks = 0.0, 0
    while True:
        done, t = runner_advance(solver, 37, t_end=duration, t=t, hand_off=hand_off)
        ticks += done
        if done < 37:
            break
    sim.engine.synchronize()
    assert ticks == plain.n_step == sim.n_step and t == plain.time == sim.time
    state = solver.state()
    for name in sim.model.components:
        np.testing.assert_array_equal(state[name].cpu().numpy(), plain.engine.field(name))
    np.testing.assert_array_equal(monitor.flux(), plain_monitor.flux())


@pytest.mark.gpu
def test_injected_permittivity_equals_painting(gpu):
    """inject_source('eps_r', array) before the run equals painting the same values: same bits after 150 steps."""
    from opensol_amd import Simulation
    from opensol_amd.solver import FDTDSolver
    painted, _, painted_monitor = build(Simulation, THIN_FILM, 'ez', True)
    painted.run(n_steps=150, print_every=0)
    sim, _, monitor = build(Simulation, THIN_FILM, 'ez', False)
    solver = FDTDSolver(sim)
    solver.inject_source('eps_r', painted.model.materials.property_array(0))
    runner_advance(solver, 150)
    for name in sim.model.components:
        np.testing.assert_array_equal(sim.engine.field(name), painted.engine.field(name))
    np.testing.assert_array_equal(monitor.flux(), painted_monitor.flux())


@pytest.mark.gpu
@pytest.mark.parametrize('polarization', ['ez', 'hz'])
def test_snapshot_restore_continues_bit_for_bit_2d(gpu, polarization):
    """Snapshot after 200 steps of the pole case (poles, PEC, CPML, flux line) with a DFT box and probes added; 150
    more steps; restore; the same 150 steps again: fields, flux, DFT spectra and probe series equal bit for bit."""
    from opensol_amd import Simulation
    from opensol_amd.solver import FDTDSolver
    from tests.cases import build_poles2d
    sim, monitor = build_poles2d(Simulation, polarization, graph_steps=16)
    box = sim.add_dft_monitor(0.2e-6, 0.1e-6, 1.8e-6, 0.9e-6, wavelengths=[0.55e-6, 0.65e-6])
    probe = sim.add_field_probe(1.1e-6, 0.45e-6)
    solver = FDTDSolver(sim)
    solver.finalize()
    runner_advance(solver, 200)
    snapshot = solver.solution_state()
    runner_advance(solver, 150)
    first = ({n: sim.engine.field(n) for n in sim.model.components}, monitor.flux(), _box_spectra(box),
             probe.get_data())
    solver.restore_solution(snapshot)
    assert sim.n_step == 200
    runner_advance(solver, 150)
    second = ({n: sim.engine.field(n) for n in sim.model.components}, monitor.flux(), _box_spectra(box),
              probe.get_data())
    for name in sim.model.components:
        np.testing.assert_array_equal(first[0][name], second[0][name])
    np.testing.assert_array_equal(first[1], second[1])
    for name in first[2]:
        np.testing.assert_array_equal(first[2][name], second[2][name])
    for a, b in zip(first[3], second[3]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_snapshot_restore_continues_bit_for_bit_3d(gpu):
    from opensol_amd import Simulation3D
    from opensol_amd.solver import FDTDSolver
    from tests.cases import build_poles3d
    sim, plane = build_poles3d(Simulation3D, graph_steps=10)
    solver = FDTDSolver(sim)
    solver.finalize()
    runner_advance(solver, 90)
    snapshot = solver.solution_state()
    runner_advance(solver, 60)
    first = ({n: sim.engine.field(n) for n in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')}, plane.flux())
    solver.restore_solution(snapshot)
    runner_advance(solver, 60)
    for name, values in first[0].items():
        np.testing.assert_array_equal(sim.engine.field(name), values)
    np.testing.assert_array_equal(plane.flux(), first[1])
