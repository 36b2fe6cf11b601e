"""Batched instances: many Simulations (2D: E_z and H_z mixed) or Simulation3Ds of uneven sizes stepped by one launch
per step.

The reference steps one model at a time (a parameter sweep builds and runs one model per value,
model/study.py:437-447). Here every instance keeps its own device state, clock and step limit, and one kernel launch
per step walks the tiles of all instances (yee2d_batch_create, yee3d_batch_create): instances with fewer steps stop
when their clock reaches their limit while the others go on, and the step loop replays from HIP graphs. An instance's
results are the same alone and in a batch, bit for bit.
"""
import ctypes

import numpy as np

from . import _lib


class _Batch:
    """Steps a list of simulations of one kind together. Sources and monitors must be added before the batch runs;
    each simulation keeps its own fields, monitors and readout."""

    three_d = False

    def __init__(self, simulations, graph_steps=0, autotune=False):
        if not simulations:
            raise ValueError('a batch needs at least one simulation')
        self.simulations = list(simulations)
        self.graph_steps = int(graph_steps)
        self._ctx = _lib.c_ptr()
        self._descs_dev = None
        self._counts = [0] * len(self.simulations)
        import torch
        self.torch = torch
        for sim in self.simulations:
            if (len(sim.model.shape) == 3) != self.three_d:
                raise ValueError(f'{type(self).__name__} steps {"3D" if self.three_d else "2D"} simulations')
            self._set_options(sim)
            if autotune and sim.engine is None:   # the member's size class, tuned once per device (autotune.py)
                self._tune(sim)
            sim.finalize()
        self.lib = _lib.load()

    @property
    def cell_counts(self):
        return [int(np.prod(sim.model.shape)) for sim in self.simulations]

    def run(self, n_steps):
        """Advance instance k by n_steps[k] steps (an int applies to all), asynchronously on the current stream.
        Returns the number of launches (the largest step count)."""
        self.prepare(n_steps)
        return self.launch()

    def prepare(self, n_steps):
        """Copy the descriptors for a run of n_steps[k] steps per instance to the device (synchronizes the current
        stream once); launch() then enqueues the steps. Returns the number of launches the run will take."""
        counts = [int(n_steps)] * len(self.simulations) if np.isscalar(n_steps) else [int(n) for n in n_steps]
        if len(counts) != len(self.simulations) or min(counts) < 0:
            raise ValueError('n_steps must be one non-negative count per simulation')
        self.close()
        self._counts = counts
        longest = max(counts)
        if longest == 0:
            return 0
        descs = (self._desc_type * len(self.simulations))()
        for k, (sim, count) in enumerate(zip(self.simulations, counts)):
            sim.finalize()
            engine = sim.engine
            engine._ensure_capacity(engine.n_step + count)
            descs[k] = self._cached_descriptor(engine)
            descs[k].step_limit = engine.n_step + count
        size = self._bytes(len(self.simulations))
        torch = self.torch
        device = self.simulations[0].engine.device
        self._descs_dev = torch.empty(size, dtype=torch.uint8, device=device)
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(self._create(descs, len(self.simulations), self._descs_dev.data_ptr(), stream,
                                ctypes.byref(self._ctx)), self._create_name)
        if self.graph_steps:
            _lib.check(self.lib.yee_set_graph_steps(self._ctx, self.graph_steps), 'yee_set_graph_steps')
            _lib.check(self.lib.yee_build_graph(self._ctx), 'yee_build_graph')   # capture here, not in launch()
        return longest

    def _cached_descriptor(self, engine):
        """The engine's batch descriptor, rebuilt only when its tables or options change (a chunked caller prepares
        the same instances many times)."""
        key = (type(self).__name__, engine.capacity, engine._materials_version, self._options_key())
        cached = getattr(engine, '_batch_descriptor', None)
        if cached is None or cached[0] != key:
            cached = (key, self._descriptor(engine))
            engine._batch_descriptor = cached
        return cached[1]

    def launch(self):
        """Enqueue the steps prepared by prepare() on the current stream (asynchronous); returns the launches."""
        if not self._ctx:
            return 0
        longest = max(self._counts)
        stream = self.torch.cuda.current_stream(self.simulations[0].engine.device).cuda_stream
        _lib.check(self.lib.yee_step(self._ctx, longest, stream), 'yee_step')
        for sim, count in zip(self.simulations, self._counts):
            sim.engine.n_step += count
        self._counts = [0] * len(self._counts)
        return longest

    def synchronize(self):
        self.torch.cuda.synchronize(self.simulations[0].engine.device)

    def close(self):
        if self._ctx:
            self.lib.yee_destroy(self._ctx)
            self._ctx = _lib.c_ptr()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch2D(_Batch):
    """2D Simulations (E_z and H_z mixed) stepped together; every instance uses `lane_columns` columns per lane.
    autotune=True chooses each not-yet-finalized member's rows per tile by its size class (autotune.py)."""

    _desc_type = _lib.Desc2DC
    _create_name = 'yee2d_batch_create'

    def __init__(self, simulations, graph_steps=0, lane_columns=2, rows_per_tile=0, autotune=False):
        self.lane_columns = int(lane_columns)
        self.rows_per_tile = int(rows_per_tile)
        if autotune and self.rows_per_tile:
            raise ValueError('Batch2D: autotune chooses rows_per_tile; pass one or the other')
        super().__init__(simulations, graph_steps, autotune)

    def _tune(self, sim):
        from .autotune import tune_batch_member
        tune_batch_member(sim, lane_columns=self.lane_columns)

    def _set_options(self, sim):
        if getattr(sim.model, 'polarization', None) not in ('ez', 'hz'):
            raise ValueError('Batch2D steps 2D simulations')
        sim._engine_options['lane_columns'] = self.lane_columns
        if self.rows_per_tile:
            sim._engine_options['rows_per_tile'] = self.rows_per_tile

    def _descriptor(self, engine):
        d = engine.descriptor()
        d.lane_columns = self.lane_columns
        resolved = _lib.Desc2DC()
        _lib.check(self.lib.yee2d_get_desc(engine._ctx, ctypes.byref(resolved)), 'yee2d_get_desc')
        d.rows_per_tile = resolved.rows_per_tile
        return d

    def _options_key(self):
        return (self.lane_columns, self.rows_per_tile)

    def _bytes(self, n):
        return self.lib.yee2d_batch_bytes(n)

    def _create(self, *args):
        return self.lib.yee2d_batch_create(*args)


class Batch3D(_Batch):
    """Simulation3Ds stepped together; planes_per_tile (0: each instance's own) sets the x planes of every tile.
    autotune=True chooses each not-yet-finalized member's planes per tile by its size class (autotune.py)."""

    three_d = True
    _desc_type = _lib.Desc3DC
    _create_name = 'yee3d_batch_create'

    def __init__(self, simulations, graph_steps=0, planes_per_tile=0, autotune=False):
        self.planes_per_tile = int(planes_per_tile)
        if autotune and self.planes_per_tile:
            raise ValueError('Batch3D: autotune chooses planes_per_tile; pass one or the other')
        super().__init__(simulations, graph_steps, autotune)

    def _tune(self, sim):
        from .autotune import tune_batch_member
        tune_batch_member(sim)

    def _set_options(self, sim):
        pass

    def _descriptor(self, engine):
        d = engine.descriptor()
        if self.planes_per_tile:
            d.planes_per_tile = self.planes_per_tile
        return d

    def _options_key(self):
        return self.planes_per_tile

    def _bytes(self, n):
        return self.lib.yee3d_batch_bytes(n)

    def _create(self, *args):
        return self.lib.yee3d_batch_create(*args)


__all__ = ['Batch2D', 'Batch3D']
