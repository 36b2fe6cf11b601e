"""Serving loop for streams of FDTD jobs: continuous batching under an HBM budget, with p50 / p99 latency.

A job is a 2D Simulation or a Simulation3D with a step count. Jobs arrive over time and are admitted while their
device state fits the HBM budget, then stay resident until they finish. Work advances in chunks:

  * all admitted 2D jobs share one launch per step (Batch2D) and all admitted 3D jobs another (Batch3D), on two
    streams, so a new arrival joins the next chunk instead of waiting for the others to finish, and the small
    instances fill the device together;
  * chunks are pipelined: the host reads out finished jobs, builds and admits new arrivals and enqueues the next
    chunk while the current one runs on the device.

A finished job is read out (its monitors' flux spectra), its buffers are freed and its latency (arrival to result) is
recorded. The plain form, for comparison, runs the same jobs one at a time in arrival order.

Every job's result is the same, bit for bit, whether it runs alone or in the mix: the batched launches and the
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
        self._used = 0.0
        import torch
        self.torch = torch

    def _admit(self, pending, active, start, clock):
        """Admit every arrived job whose state fits the budget (at least one job when none is active). The new
        engines are created and zeroed on the current stream."""
        now = clock() - start
        while pending and pending[0].arrival <= now:
            job = pending[0]
            if job.sim is None:
                job.sim = job.build()
            need = state_bytes(job.sim)
            if active and self._used + need > self.hbm_budget:
                break
            pending.pop(0)
            job.sim.finalize()
            # source and phase tables for every step of the job now, so no table is replaced while a chunk runs
            job.sim.engine._ensure_capacity(job.steps)
            job.cells = int(np.prod(job.sim.model.shape))
            job.admitted = now
            job.bytes = need
            self._used += need
            active.append(job)

    def run(self, jobs, clock=time.perf_counter):
        """Serve `jobs` (sorted by arrival); returns the stats dict. Arrivals are real time from the call.

        Each chunk enqueues one batched launch per step for the admitted 2D jobs and one for the admitted 3D jobs, on
        two streams so the two batches overlap on the device. Chunks are pipelined: the host enqueues chunk k, then
        waits for chunk k - 1, reads out the jobs that finished in it, frees their state and admits new arrivals
        while chunk k runs. Which jobs finish in a chunk is known on the host from the step counts, so a finished
        job is never part of a later chunk. Consecutive chunks alternate between two stream pairs, and each chunk's
        streams wait on the previous chunk's events, which keeps every instance's steps in order."""
        from .batch import Batch2D, Batch3D
        torch = self.torch
        if any(j.steps < 1 for j in jobs):
            raise ValueError('every job needs at least one step')
        pending = sorted(jobs, key=lambda j: j.arrival)
        active, done = [], []
        self._used = 0.0
        start = clock()
        cell_updates = 0
        streams = [(torch.cuda.Stream(), torch.cuda.Stream()) for _ in range(2)]   # default priority
        last_event = [None, None]   # end of the previous chunk's 2D and 3D batches
        inflight = None             # (events, batches, jobs finishing in that chunk)
        parity = 0
        phases = dict(admit=0.0, enqueue=0.0, wait=0.0, readout=0.0, chunks=0)
        while pending or active or inflight:
            t0 = clock()
            self._admit(pending, active, start, clock)
            t1 = clock()
            runnable = [j for j in active if j.done_steps < j.steps]
            current = None
            if runnable:
                main = torch.cuda.current_stream()
                events, batches, finishing = [], [], []
                for kind_index, kind in enumerate((Batch2D, Batch3D)):
                    members = [j for j in runnable if j.three_d == kind.three_d]
                    if not members:
                        continue
                    stream = streams[parity][kind_index]
                    with torch.cuda.stream(stream):
                        batch = kind([j.sim for j in members], graph_steps=self.graph_steps)
                        counts = [min(self.chunk_steps, j.steps - j.done_steps) for j in members]
                        batch.prepare(counts)   # synchronizes this stream, which holds only finished chunks
                        stream.wait_stream(main)   # new engines were zeroed on the main stream
                        if last_event[kind_index] is not None:
                            stream.wait_event(last_event[kind_index])
                        batch.launch()
                        event = torch.cuda.Event()
                        event.record(stream)
                    last_event[kind_index] = event
                    events.append(event)
                    batches.append(batch)
                    for j, c in zip(members, counts):
                        j.done_steps += c
                        cell_updates += j.cells * c
                        if j.done_steps >= j.steps:
                            finishing.append(j)
                current = (events, batches, finishing)
                parity ^= 1
                phases['chunks'] += 1
            elif inflight is None:
                wait = pending[0].arrival - (clock() - start)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                phases['admit'] += t1 - t0
                continue
            t2 = clock()
            if inflight is not None:
                events, batches, finishing = inflight
                for event in events:
                    event.synchronize()
                t3 = clock()
                for batch in batches:
                    batch.close()
                for j in finishing:
                    j.result = _readout(j.sim)
                    j.finished = clock() - start
                    j.sim.close()
                    j.sim = None
                    self._used -= j.bytes
                    active.remove(j)
                    done.append(j)
                phases['wait'] += t3 - t2
                phases['readout'] += clock() - t3
            inflight = current
            phases['admit'] += t1 - t0
            phases['enqueue'] += t2 - t1
        torch.cuda.synchronize()
        stats = self._stats(done, clock() - start, cell_updates)
        stats['host_phases_s'] = phases
        return stats

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
