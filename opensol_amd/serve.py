"""Serving loop for streams of FDTD jobs: continuous batching under an HBM budget, with p50 / p99 latency.

A job is a 2D Simulation or a Simulation3D with a step count. Jobs arrive over time and are admitted while their
device state fits the HBM budget, then stay resident until they finish. Work advances in chunks:

  * all admitted 2D jobs share one launch per step (Batch2D), so a new arrival joins the next chunk instead of
    waiting for the others to finish;
  * each admitted 3D job runs its chunk with its own fused launches on the same stream.

A finished job is read out (its monitors' flux spectra), its buffers are freed and its latency (arrival to result) is
recorded. The plain form, for comparison, runs the same jobs one at a time in arrival order.

Every job's result is the same, bit for bit, whether it runs alone or in the mix: the batched 2D launch and the
single-instance kernels give identical bits, and every instance keeps its own clock.

The reference has no counterpart: it runs one model per call (model/study.py:437-447).
"""
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Job:
    """One FDTD job: `build()` returns a Simulation / Simulation3D with its sources and monitors, `steps` the steps
    to run, `arrival` the arrival time [s] after the stream starts."""
    build: object
    steps: int
    arrival: float
    name: str = ''
    sim: object = None
    done_steps: int = 0
    admitted: float = None
    finished: float = None
    cells: int = 0
    result: list = field(default_factory=list)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def three_d(self):
        return len(self.sim.model.shape) == 3


def state_bytes(sim):
    """Device bytes of an instance: double-buffered fields, ids, CPML psi and monitor accumulators (an estimate made
    before the engine exists)."""
    model = sim.model
    cells = int(np.prod(model.shape))
    if len(model.shape) == 3:
        psi = sum(6 * t.n_rows * cells // n for t, n in zip(model.cpml.values(), model.shape))
        acc = sum(2 * 2 * m.n_freqs * 4 * 8 * int(np.prod(m.accumulator_shape(0)[3:])) for m in model.monitors)
        return cells * (2 * 6 * 4 + 2) + psi * 4 + acc
    psi = sum(4 * t.n_rows * cells // n for t, n in zip(model.cpml.values(), model.shape))
    acc = sum(2 * m.n_freqs * 4 * 8 * m.n_cells for m in model.monitors)
    return cells * (2 * 3 * 4 + 2) + psi * 4 + acc


def _readout(sim):
    return [np.asarray(m.flux()) for m in sim.model.monitors]


class FDTDServer:
    """Continuous-batching server over one GPU. hbm_budget [B] bounds the resident device state; chunk_steps is the
    granularity at which jobs join and leave."""

    def __init__(self, hbm_budget=200e9, chunk_steps=100, graph_steps=0):
        self.hbm_budget = float(hbm_budget)
        self.chunk_steps = int(chunk_steps)
        self.graph_steps = int(graph_steps)
        import torch
        self.torch = torch

    def run(self, jobs, clock=time.perf_counter):
        """Serve `jobs` (sorted by arrival); returns the stats dict. Arrivals are real time from the call."""
        from .batch import Batch2D
        torch = self.torch
        pending = sorted(jobs, key=lambda j: j.arrival)
        active, done = [], []
        used = 0.0
        start = clock()
        cell_updates = 0
        while pending or active:
            now = clock() - start
            # admit every arrived job that fits
            while pending and pending[0].arrival <= now:
                job = pending[0]
                if job.sim is None:
                    job.sim = job.build()
                need = state_bytes(job.sim)
                if active and used + need > self.hbm_budget:
                    break
                pending.pop(0)
                job.sim.finalize()
                job.cells = int(np.prod(job.sim.model.shape))
                job.admitted = now
                job.bytes = need
                used += need
                active.append(job)
            if not active:
                wait = pending[0].arrival - (clock() - start)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            # one chunk: the 2D jobs batched, the 3D jobs each
            jobs2d = [j for j in active if not j.three_d]
            jobs3d = [j for j in active if j.three_d]
            if jobs2d:
                batch = Batch2D([j.sim for j in jobs2d], graph_steps=self.graph_steps)
                counts = [min(self.chunk_steps, j.steps - j.done_steps) for j in jobs2d]
                batch.run(counts)
                for j, c in zip(jobs2d, counts):
                    j.done_steps += c
                    cell_updates += j.cells * c
            for j in jobs3d:
                c = min(self.chunk_steps, j.steps - j.done_steps)
                j.sim.engine.step(c)
                j.done_steps += c
                cell_updates += j.cells * c
            torch.cuda.synchronize()
            if jobs2d:
                batch.close()
            finished_at = clock() - start
            for j in [j for j in active if j.done_steps >= j.steps]:
                j.result = _readout(j.sim)
                j.finished = clock() - start if j.result else finished_at
                j.sim.close()
                j.sim = None
                used -= j.bytes
                active.remove(j)
                done.append(j)
        return self._stats(done, clock() - start, cell_updates)

    def run_plain(self, jobs, clock=time.perf_counter):
        """The plain form: each job alone, in arrival order, to completion."""
        torch = self.torch
        done = []
        start = clock()
        cell_updates = 0
        for job in sorted(jobs, key=lambda j: j.arrival):
            wait = job.arrival - (clock() - start)
            if wait > 0:
                time.sleep(wait)
            if job.sim is None:
                job.sim = job.build()
            job.sim.finalize()
            job.cells = int(np.prod(job.sim.model.shape))
            job.admitted = clock() - start
            job.sim.engine.step(job.steps)
            torch.cuda.synchronize()
            job.done_steps = job.steps
            cell_updates += job.cells * job.steps
            job.result = _readout(job.sim)
            job.finished = clock() - start
            job.sim.close()
            job.sim = None
            done.append(job)
        return self._stats(done, clock() - start, cell_updates)

    @staticmethod
    def _stats(done, makespan, cell_updates):
        latencies = np.array([j.latency for j in done])
        return {'jobs': len(done), 'makespan_s': makespan, 'cell_updates': cell_updates,
                'cell_updates_per_s': cell_updates / makespan if makespan > 0 else 0.0,
                'p50_latency_s': float(np.percentile(latencies, 50)) if latencies.size else None,
                'p99_latency_s': float(np.percentile(latencies, 99)) if latencies.size else None,
                'results': {j.name: j.result for j in done}}


__all__ = ['Job', 'FDTDServer', 'state_bytes']
