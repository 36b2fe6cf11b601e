"""Serving loop for streams of FDTD jobs: continuous batching under an HBM budget, with p50 / p99 latency.

A job is a 2D Simulation or a Simulation3D with a step count. Jobs arrive over time and are admitted while their
device state fits the HBM budget, then stay resident until they finish. Work advances in chunks:

  * all admitted 2D jobs share one launch per step (Batch2D) and all admitted 3D jobs another (Batch3D), on two
    streams, so a new arrival joins the next chunk instead of waiting for the others to finish, and the small
    instances fill the device together;
  * chunks are pipelined: the host reads out finished jobs, builds and admits new arrivals and enqueues the next
    chunk while the current one runs on the device.

A finished job is read out (its monitors' flux spectra), its buffers are freed and its latency (arrival to result) is
recorded. With autotune (the default) every admitted job's batch-member launch shape is the one tuned for its size
class on this device (autotune.py; tune_classes() does the tuning ahead of a stream). The plain form, for comparison, runs the same jobs one at a time in arrival order.

Failure handling. A job whose device state alone exceeds the HBM budget is refused at arrival (status 'rejected',
with the reason in job.error), as the reference's memory budget refuses an over-sized model instead of spilling
(core/gpu_memory.py:177, 271-290). Job.cancel() may be called from any thread; the server honours it at the next
chunk boundary, like the reference's RunControl.checkpoint between study steps (model/study.py:194-251): a pending
job is dropped, a running one stops after the chunk in flight and its state is freed (status 'cancelled'). After
every chunk each job's current field set is checked on the device for NaN / Inf (yee_flag_nonfinite, read when the
chunk is collected); a flagged job stops there and is reported as 'failed', as is one whose spectra come out
non-finite at the end.

Preemption (preempt_after=seconds). When the job at the head of the queue fits the budget but not the memory left,
and has waited that long, the most recently admitted running jobs are preempted until it fits: each one's evolving
state is copied to the host (solver.solution_state: both field sets, psi, poles, accumulators, probes, clock), its
device state is freed, and it goes back to the queue right behind the new job; re-admitted, it gets a new engine, its
state is written back (restore_solution) and it continues bit for bit. A job is preempted at most once and only a job
that has not run yet triggers preemption, so jobs cannot preempt each other in turn.

Every job's result is the same, bit for bit, whether it runs alone or in the mix: the batched launches and the
single-instance kernels give identical bits, and every instance keeps its own clock.

The reference has no counterpart: it runs one model per call (model/study.py:437-447).
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass(eq=False)
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
    bytes: int = 0        # device state counted against the HBM budget while resident
    result: list = field(default_factory=list)
    status: str = 'pending'   # pending, running, done, rejected, cancelled, failed
    error: str = ''
    preemptions: int = 0
    saved: dict = field(default=None, repr=False)   # host state while preempted
    _cancel: threading.Event = field(default_factory=threading.Event, repr=False)

    def cancel(self):
        """Ask the server to drop this job at its next chunk boundary (thread-safe)."""
        self._cancel.set()

    @property
    def cancelled(self):
        return self._cancel.is_set()

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


def _finish(job, now):
    """Read out a job that ran all its steps; non-finite spectra mark it failed."""
    job.result = _readout(job.sim)
    job.finished = now
    if all(np.all(np.isfinite(r)) for r in job.result):
        job.status = 'done'
    else:
        job.status = 'failed'
        job.error = 'non-finite flux spectra (unstable or diverging fields)'


def choose_victims(head_bytes, free_bytes, running):
    """The running jobs to preempt so that a job of head_bytes fits: the most recently admitted first (the later arrival
    first among jobs admitted together), each at most once (jobs already preempted are skipped), until free_bytes plus
    theirs covers it; [] when even all of them would not."""
    chosen, free = [], float(free_bytes)
    for job in sorted(running, key=lambda j: (j.admitted, j.arrival), reverse=True):
        if free >= head_bytes:
            break
        if job.preemptions == 0:
            chosen.append(job)
            free += job.bytes
    return chosen if free >= head_bytes else []


class FDTDServer:
    """Continuous-batching server over one GPU. hbm_budget [B] bounds the resident device state; chunk_steps is the
    granularity at which jobs join and leave; preempt_after [s], if set, lets a queued job that has waited that long
    preempt running ones; finite_check runs the device NaN / Inf check after every chunk."""

    def __init__(self, hbm_budget=200e9, chunk_steps=100, graph_steps=0, autotune=True, preempt_after=None,
                 finite_check=True):
        self.hbm_budget = float(hbm_budget)
        self.chunk_steps = int(chunk_steps)
        self.graph_steps = int(graph_steps)
        self.autotune = bool(autotune)
        self.preempt_after = None if preempt_after is None else float(preempt_after)
        self.finite_check = bool(finite_check)
        self._used = 0.0
        self.preemptions = 0
        import torch
        self.torch = torch

    def _admit(self, pending, active, start, clock, closed):
        """Admit every arrived job whose state fits the budget. A job larger than the whole budget is refused and a
        cancelled one dropped (both go to `closed`). The new engines are created and zeroed on the current stream."""
        now = clock() - start
        while pending and pending[0].arrival <= now:
            job = pending[0]
            if job.cancelled:
                pending.pop(0)
                job.status, job.finished = 'cancelled', now
                closed.append(job)
                continue
            if job.sim is None:
                job.sim = job.build()
            need = state_bytes(job.sim)
            if need > self.hbm_budget:
                pending.pop(0)
                job.sim = None
                job.status, job.finished = 'rejected', now
                job.error = f'device state {need:.4g} B exceeds the HBM budget {self.hbm_budget:.4g} B'
                closed.append(job)
                continue
            if self._used + need > self.hbm_budget:
                break
            pending.pop(0)
            resumed = job.saved is not None
            if self.autotune and job.sim.engine is None and not resumed:   # the launch shape of the size class
                from .autotune import tune_batch_member
                tune_batch_member(job.sim)
            job.sim.finalize()
            # source and phase tables for every step of the job now, so no table is replaced while a chunk runs
            job.sim.engine._ensure_capacity(job.steps)
            if resumed:   # a preempted job: its state back into the new engine
                from .solver import restore_solution
                restore_solution(job.sim.engine, job.saved)
                job.saved = None
            job.cells = int(np.prod(job.sim.model.shape))
            if not resumed:
                job.admitted = now
            job.bytes = need
            job.status = 'running'
            self._used += need
            active.append(job)

    def tune_classes(self, jobs):
        """Tune the batch-member launch shape of every size class among `jobs` ahead of serving them (each job is
        built once for it and dropped; a class already in the cache costs nothing). Returns the classes seen."""
        from .autotune import class_key, tune_batch_member
        seen = set()
        for job in jobs:
            sim = job.build()
            key = class_key('batch', sim.model)
            if key not in seen:
                seen.add(key)
                tune_batch_member(sim)
            sim.close()
        return sorted(seen)

    def _release(self, job, active, done):
        job.sim.close()
        job.sim = None
        self._used -= job.bytes
        active.remove(job)
        done.append(job)

    def _preemption(self, pending, active, start, clock):
        """The running jobs to preempt for the head of the queue now ([] when none is due)."""
        if self.preempt_after is None or not pending:
            return []
        head = pending[0]
        now = clock() - start
        if head.arrival > now or head.cancelled or head.preemptions or head.saved is not None \
                or now - head.arrival < self.preempt_after:
            return []
        if head.sim is None:
            head.sim = head.build()
        need = state_bytes(head.sim)
        if need > self.hbm_budget or self._used + need <= self.hbm_budget:
            return []   # refused at admission, or it fits anyway
        running = [j for j in active if j.done_steps < j.steps and j.status == 'running']
        return choose_victims(need, self.hbm_budget - self._used, running)

    def _preempt(self, job, pending, active):
        """Move a running job's state to the host, free its device state and queue it behind the head job."""
        from .solver import solution_state
        job.saved = solution_state(job.sim.engine)
        job.sim.close()
        self._used -= job.bytes
        active.remove(job)
        job.preemptions += 1
        job.status = 'pending'
        self.preemptions += 1
        pending.insert(1, job)

    def _collect(self, chunk, active, done, start, clock, phases):
        """Wait for a chunk, then read out its finished jobs, fail the jobs its finite check flagged and free the
        state of every job that stops with it. Returns the flagged jobs that are still in a later chunk."""
        t2 = clock()
        for event in chunk['events']:
            event.synchronize()
        t3 = clock()
        flagged = []
        t_flags = clock()
        for members, steps, flags in chunk['checks']:
            for job, step, bad in zip(members, steps, flags.cpu().numpy()):
                if bad and job.status == 'running':
                    job.status = 'failed'
                    job.error = f'non-finite fields after step {step} (device check at a chunk boundary)'
                    job.finished = clock() - start
                    flagged.append(job)
        t_close = clock()
        for batch in chunk['batches']:
            batch.close()
        t_finish = clock()
        stopping = set(id(j) for j in chunk['stopping'])
        for job in chunk['finishing']:
            if job.status == 'running':
                _finish(job, clock() - start)
        t_release = clock()
        for job in chunk['finishing']:
            if job in active:
                self._release(job, active, done)
        t_end = clock()
        phases['readout_flags'] = phases.get('readout_flags', 0.0) + t_close - t_flags
        phases['readout_close'] = phases.get('readout_close', 0.0) + t_finish - t_close
        phases['readout_spectra'] = phases.get('readout_spectra', 0.0) + t_release - t_finish
        phases['readout_release'] = phases.get('readout_release', 0.0) + t_end - t_release
        for job in chunk['stopping']:
            if job in active:
                job.finished = job.finished if job.finished is not None else clock() - start
                self._release(job, active, done)
        later = []
        for job in flagged:   # not finishing in this chunk: stop stepping it; free it when no chunk holds it
            if job in active and id(job) not in stopping:
                job.done_steps = job.steps
                later.append(job)
        phases['wait'] += t3 - t2
        phases['readout'] += clock() - t3
        return later

    def run(self, jobs, clock=time.perf_counter, on_chunk=None):
        """Serve `jobs` (sorted by arrival); returns the stats dict. Arrivals are real time from the call. on_chunk, if
        given, is called with the list of running jobs at every chunk boundary (a caller may cancel jobs there).

        Each chunk enqueues one batched launch per step for the admitted 2D jobs and one for the admitted 3D jobs, on
        two streams so the two batches overlap on the device. Chunks are pipelined: the host enqueues chunk k, then
        waits for chunk k - 1, reads out the jobs that finished in it, frees their state and admits new arrivals
        while chunk k runs. Which jobs finish in a chunk is known on the host from the step counts, so a finished
        job is never part of a later chunk. Consecutive chunks alternate between two stream pairs, and each chunk's
        streams wait on the previous chunk's events, which keeps every instance's steps in order."""
        from .batch import Batch2D, Batch3D
        from . import _lib
        torch = self.torch
        if any(j.steps < 1 for j in jobs):
            raise ValueError('every job needs at least one step')
        lib = _lib.load() if self.finite_check else None
        pending = sorted(jobs, key=lambda j: j.arrival)
        active, done, closed = [], [], []
        self._used = 0.0
        self.preemptions = 0
        start = clock()
        cell_updates = 0
        streams = None              # two stream pairs (default priority), made when the first chunk runs
        last_event = [None, None]   # end of the previous chunk's 2D and 3D batches
        inflight = None             # the chunk on the device: events, batches, finite checks, finishing / stopping jobs
        parity = 0
        phases = dict(admit=0.0, enqueue=0.0, wait=0.0, readout=0.0, chunks=0)
        while pending or active or inflight:
            t0 = clock()
            victims = self._preemption(pending, active, start, clock)
            if victims:   # settle every victim's state first: the chunk in flight may hold it
                if inflight is not None:
                    held = self._collect(inflight, active, done, start, clock, phases)
                    for job in held:
                        self._release(job, active, done)
                    inflight = None
                for job in self._preemption(pending, active, start, clock):
                    self._preempt(job, pending, active)
            self._admit(pending, active, start, clock, closed)
            t1 = clock()
            if on_chunk is not None:
                on_chunk([j for j in active if j.done_steps < j.steps])
            stopping = [j for j in active if j.cancelled and j.done_steps < j.steps]
            for j in stopping:   # no further steps; freed once the chunk in flight (which may hold it) is done
                j.done_steps = j.steps
                j.status = 'cancelled'
            runnable = [j for j in active if j.done_steps < j.steps]
            current = None
            if runnable and streams is None:
                streams = [(torch.cuda.Stream(), torch.cuda.Stream()) for _ in range(2)]
            if runnable or stopping:
                current = dict(events=[], batches=[], checks=[], finishing=[], stopping=stopping, members=[])
            if runnable:
                main = torch.cuda.current_stream()
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
                        if lib is not None:   # NaN / Inf in each member's state after the chunk
                            flags = torch.zeros(len(members), dtype=torch.int32, device=stream.device)
                            handle = stream.cuda_stream
                            for k, j in enumerate(members):
                                engine = j.sim.engine
                                state = engine.fields[engine.n_step & 1]
                                _lib.check(lib.yee_flag_nonfinite(state.data_ptr(), state.numel(),
                                                                  flags.data_ptr() + 4 * k, handle),
                                           'yee_flag_nonfinite')
                            current['checks'].append((members, [j.sim.engine.n_step for j in members], flags))
                        event = torch.cuda.Event()
                        event.record(stream)
                    last_event[kind_index] = event
                    current['events'].append(event)
                    current['batches'].append(batch)
                    current['members'].extend(members)
                    for j, c in zip(members, counts):
                        j.done_steps += c
                        cell_updates += j.cells * c
                        if j.done_steps >= j.steps:
                            current['finishing'].append(j)
                parity ^= 1
                phases['chunks'] += 1
            elif current is None and inflight is None:
                if pending:
                    wait = pending[0].arrival - (clock() - start)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                phases['admit'] += t1 - t0
                continue
            t2 = clock()
            if inflight is not None:
                held = self._collect(inflight, active, done, start, clock, phases)
                in_current = set(id(j) for j in current['members']) if current is not None else set()
                for job in held:
                    if id(job) in in_current:
                        current['stopping'].append(job)   # freed when the chunk just enqueued is collected
                    else:
                        self._release(job, active, done)
            inflight = current
            phases['admit'] += t1 - t0
            phases['enqueue'] += t2 - t1
        if streams is not None:
            torch.cuda.synchronize()
        stats = self._stats(done + closed, clock() - start, cell_updates)
        stats['host_phases_s'] = phases
        stats['preemptions'] = self.preemptions
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
            _finish(job, clock() - start)
            job.sim.close()
            job.sim = None
            done.append(job)
        return self._stats(done, clock() - start, cell_updates)

    @staticmethod
    def _stats(jobs, makespan, cell_updates):
        """Latency percentiles over the jobs that completed; every job's status, and the errors of the others."""
        done = [j for j in jobs if j.status == 'done']
        latencies = np.array([j.latency for j in done])
        return {'jobs': len(done), 'makespan_s': makespan, 'cell_updates': cell_updates,
                'cell_updates_per_s': cell_updates / makespan if makespan > 0 else 0.0,
                'p50_latency_s': float(np.percentile(latencies, 50)) if latencies.size else None,
                'p99_latency_s': float(np.percentile(latencies, 99)) if latencies.size else None,
                'results': {j.name: j.result for j in done},
                'status': {j.name: j.status for j in jobs},
                'errors': {j.name: j.error for j in jobs if j.error}}


__all__ = ['Job', 'FDTDServer', 'choose_victims', 'state_bytes']
