"""The reference's solver contract over the HIP engine, and snapshot / restore of an engine's resident state.

FDTDSolver implements the PhysicsSolver contract of the reference (core/solver_base.py:5-42: prepare, step, dt_limit,
state, inject_source, finalize, materials_changed, scalars) for a Simulation or Simulation3D, so the reference's
MultiphysicsRunner (coupling/runner.py:74-218) can drive the HIP stepper as it drives EMPhysics:

  * prepare -> dt_limit gives the runner its dt_sync (the engine's float32 Courant dt, as EMPhysics.dt_limit);
  * each tick the runner calls step(t, span) once (one Yee step; the engine keeps its own clock and dt, as the
    reference's engine does, fdtd_2d_base.py:234-236);
  * advance_ticks(n_ticks, t, dt, stop_time) is the runner's hand-off (runner.py:115-118, 186-190): a single solver at
    stride 1 with no couplings takes whole chunks, here one yee_step call (graph replay) for the whole chunk. The
    runner's float64 clock (t += dt per tick, stop at t >= stop_time) is reproduced tick by tick.

state() returns the current fields as device tensors (the reference returns its run-backend arrays) plus the host
material arrays; inject_source accepts the permittivity and sigma_e slots (written into the material map, which the
next step re-uploads and re-calibrates), and raises the reference's NotImplementedError for any other name (the HIP
path has no per-cell external source term).

solution_state / restore_solution copy every array that evolves with the steps (both field sets, CPML psi, pole
states, flux and DFT accumulators, probe rows, the device clock) to the host and back, as the reference's
solution_state / restore_solution do for its engines (physics/em/solvers/fdtd_2d_base.py:419-456,
physics/em3d/physics3d.py:295-333). A restored engine continues bit for bit as the original would have.
"""
import numpy as np

PERMITTIVITY_COLUMNS = {'ez': {'eps_r': (0, 1, 2)}, 'hz': {'eps_xx': (0,), 'eps_yy': (1,), 'eps_r': (0, 1, 2)},
                        '3d': {'eps_r': (0, 1, 2)}}


def _state_tensors(engine):
    """name -> device tensor of every array that changes as the engine steps."""
    out = {'fields': engine.fields, 'clock': engine.clock}
    for axis, tables in engine.cpml.items():
        if tables is not None:
            out[f'psi_e_{axis}'] = tables['psi_e']
            out[f'psi_h_{axis}'] = tables['psi_h']
    for k, acc in enumerate(engine.accumulators):
        for q, tensor in enumerate(acc if isinstance(acc, (list, tuple)) else [acc]):
            out[f'flux_{k}_{q}'] = tensor
    for k, acc in enumerate(engine.dft_accumulators):
        out[f'dft_{k}'] = acc
    for k, entry in enumerate(engine.poles):
        for q, tensor in enumerate(entry['states']):
            out[f'pole_{k}_{q}'] = tensor
    if engine.probes.samples is not None:
        out['probes'] = engine.probes.samples
    return out


def solution_state(engine):
    """Host copies of an engine's evolving state and its step count (synchronizes the device)."""
    engine.synchronize()
    return {'n_step': int(engine.n_step),
            'arrays': {name: tensor.cpu().numpy().copy() for name, tensor in _state_tensors(engine).items()}}


def restore_solution(engine, state):
    """Write a solution_state back in place (same model; the probe table keeps its own capacity)."""
    torch = engine.torch
    tensors = _state_tensors(engine)
    arrays = state['arrays']
    if set(arrays) != set(tensors):
        raise ValueError(f'snapshot arrays {sorted(arrays)} do not match this engine {sorted(tensors)}')
    n_step = int(state['n_step'])
    engine._ensure_capacity(n_step)
    tensors = _state_tensors(engine)   # the probe table may have grown
    for name, tensor in tensors.items():
        values = arrays[name]
        if name == 'probes':
            rows = min(values.shape[0], tensor.shape[0])
            tensor[:rows].copy_(torch.from_numpy(values[:rows]))
            continue
        if tuple(values.shape) != tuple(tensor.shape):
            raise ValueError(f"snapshot array '{name}' has shape {values.shape}, the engine's {tuple(tensor.shape)}")
        tensor.copy_(torch.from_numpy(values))
    engine.n_step = n_step
    engine.synchronize()


class FDTDSolver:
    """PhysicsSolver contract (core/solver_base.py:5-42) of a Simulation / Simulation3D on the HIP engine."""

    name = 'em'
    coordinate_grids = ('cartesian',)

    def __init__(self, simulation):
        self.simulation = simulation
        self.model = simulation.model
        self.kind = '3d' if len(self.model.shape) == 3 else self.model.polarization
        self.material_schema = {'eps_r': 1.0, 'sigma_e': 0.0, 'mu_r': 1.0}

    # ---- contract ---------------------------------------------------------------------------------------------
    def prepare(self, domain=None) -> None:
        """The engine is built at finalize (the reference's engines are complete at construction,
        fdtd_2d_base.py:227-229)."""

    def finalize(self) -> None:
        self.simulation.finalize()

    def dt_limit(self) -> float:
        return float(self.model.dt)

    def _engine(self):
        sim = self.simulation
        if sim.engine is None or self.model.materials.version != sim._materials_version:
            sim.finalize()
        return sim.engine

    def step(self, t: float = None, dt: float = None) -> None:
        """One Yee step (sources at the engine's own clock; dt is advisory, as in the reference's engines)."""
        self._engine().step(1)

    def advance_ticks(self, n_ticks, t, dt, stop_time):
        """Runner hand-off: run up to n_ticks ticks while t < stop_time, t += dt per tick (float64, as the runner's
        tick loop); one yee_step call for the whole chunk. Returns (ticks run, new t)."""
        n_done = 0
        while n_done < int(n_ticks) and t < stop_time:
            t = t + dt
            n_done += 1
        if n_done:
            self._engine().step(n_done)
        return n_done, t

    def state(self) -> dict:
        """The current fields as device tensors (views of the engine's current state set) and the host material
        arrays."""
        engine = self._engine()
        current = engine.fields[engine.n_step & 1]
        out = {}
        for index, name in enumerate(self.model.components):
            out[name] = current[index] if self.kind == '3d' else current[index, :, :self.model.ny]
        materials = self.model.materials
        if self.kind == 'hz':
            out['eps_xx'], out['eps_yy'] = materials.display_array(0), materials.display_array(1)
        else:
            out['eps_r'] = materials.display_array(0)
        out['sigma_e'] = materials.display_array(3)
        return out

    def inject_source(self, name: str, array) -> None:
        columns = PERMITTIVITY_COLUMNS[self.kind].get(name, (3,) if name == 'sigma_e' else None)
        if columns is None:
            raise NotImplementedError(f"{type(self).__name__} does not accept external source '{name}'")
        values = np.asarray(array.cpu().numpy() if hasattr(array, 'cpu') else array, dtype=np.float32)
        if values.shape != tuple(self.model.shape):
            raise ValueError(f"inject_source('{name}', ...) needs an array of shape {tuple(self.model.shape)}, "
                             f"got {values.shape}")
        self.model.materials.set_property(columns, values)

    def materials_changed(self, names) -> None:
        """The material map changes only through this solver or the model's painting, which bump its version: the
        next step re-uploads ids and tables and re-calibrates the sources."""

    def scalars(self) -> dict:
        return {}

    # ---- snapshots ----------------------------------------------------------------------------------------------
    def solution_state(self):
        return solution_state(self._engine())

    def restore_solution(self, state):
        restore_solution(self._engine(), state)


__all__ = ['FDTDSolver', 'solution_state', 'restore_solution']
