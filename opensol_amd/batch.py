"""Batched 2D instances: many Simulations of uneven sizes (E_z and H_z mixed) stepped by one launch per step.

The reference steps one model at a time (a parameter sweep builds and runs one model per value,
model/study.py:437-447). Here every instance keeps its own device state, clock and step limit, and one kernel launch
per step walks the tiles of all instances (yee2d_batch_create): instances with fewer steps stop when their clock
reaches their limit while the others go on, and the step loop replays from HIP graphs. An instance's results are the
same alone and in a batch, bit for bit.
"""
import ctypes

import numpy as np

from . import _lib


class Batch2D:
    """Steps a list of 2D Simulations together. Sources and monitors must be added before the batch runs; each
    Simulation keeps its own fields, monitors and readout."""

    def __init__(self, simulations, graph_steps=0, lane_columns=2, rows_per_tile=0):
        if not simulations:
            raise ValueError('a batch needs at least one simulation')
        self.simulations = list(simulations)
        self.graph_steps = int(graph_steps)
        self.lane_columns = int(lane_columns)
        self.rows_per_tile = int(rows_per_tile)
        self._ctx = _lib.c_ptr()
        self._descs_dev = None
        import torch
        self.torch = torch
        for sim in self.simulations:
            if getattr(sim.model, 'polarization', None) not in ('ez', 'hz'):
                raise ValueError('Batch2D steps 2D simulations')
            sim._engine_options['lane_columns'] = self.lane_columns
            if self.rows_per_tile:
                sim._engine_options['rows_per_tile'] = self.rows_per_tile
            sim.finalize()
        self.lib = _lib.load()

    @property
    def cell_counts(self):
        return [sim.model.nx * sim.model.ny for sim in self.simulations]

    def run(self, n_steps):
        """Advance instance k by n_steps[k] steps (an int applies to all), asynchronously on the current stream.
        Returns the number of launches (the largest step count)."""
        counts = [int(n_steps)] * len(self.simulations) if np.isscalar(n_steps) else [int(n) for n in n_steps]
        if len(counts) != len(self.simulations) or min(counts) < 0:
            raise ValueError('n_steps must be one non-negative count per simulation')
        longest = max(counts)
        if longest == 0:
            return 0
        descs = (_lib.Desc2DC * len(self.simulations))()
        for k, (sim, count) in enumerate(zip(self.simulations, counts)):
            sim.finalize()
            engine = sim.engine
            engine._ensure_capacity(engine.n_step + count)
            d = engine.descriptor()
            d.lane_columns = self.lane_columns
            resolved = _lib.Desc2DC()
            _lib.check(self.lib.yee2d_get_desc(engine._ctx, ctypes.byref(resolved)), 'yee2d_get_desc')
            d.rows_per_tile = resolved.rows_per_tile
            d.step_limit = engine.n_step + count
            descs[k] = d
        size = self.lib.yee2d_batch_bytes(len(self.simulations))
        torch = self.torch
        device = self.simulations[0].engine.device
        self.close()
        self._descs_dev = torch.empty(size, dtype=torch.uint8, device=device)
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(self.lib.yee2d_batch_create(descs, len(self.simulations), self._descs_dev.data_ptr(), stream,
                                               ctypes.byref(self._ctx)), 'yee2d_batch_create')
        if self.graph_steps:
            _lib.check(self.lib.yee_set_graph_steps(self._ctx, self.graph_steps), 'yee_set_graph_steps')
        _lib.check(self.lib.yee_step(self._ctx, longest, stream), 'yee_step')
        for sim, count in zip(self.simulations, counts):
            sim.engine.n_step += count
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
