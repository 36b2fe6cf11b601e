"""Serving on the host side (no GPU needed): a job whose device state alone exceeds the HBM budget is refused at
arrival with the reason, and a job cancelled before it is admitted is dropped; neither touches the device."""
import numpy as np

from opensol_amd.serve import FDTDServer, Job, state_bytes


def _job(k, cancel=False):
    def build():
        from opensol_amd import Simulation
        sim = Simulation(2e-6, 1e-6, 25e-9, npml=8)
        sim.add_line_source(axis='y', position=0.5e-6, wavelength=0.6e-6)
        sim.add_flux_monitor(axis='y', position=1.5e-6, span=(0.0, 1e-6), wavelengths=[0.6e-6])
        return sim
    job = Job(build=build, steps=100, arrival=0.0, name=f'job{k}')
    if cancel:
        job.cancel()
    return job


def test_over_budget_job_is_rejected_and_cancelled_job_dropped():
    need = state_bytes(_job(0).build())
    jobs = [_job(0), _job(1, cancel=True)]
    stats = FDTDServer(hbm_budget=need / 2).run(jobs)
    assert stats['jobs'] == 0 and stats['cell_updates'] == 0
    assert stats['status'] == {'job0': 'rejected', 'job1': 'cancelled'}
    assert 'exceeds the HBM budget' in stats['errors']['job0']
    assert jobs[0].sim is None and np.isfinite(jobs[0].latency)
