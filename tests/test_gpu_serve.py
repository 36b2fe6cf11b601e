"""The serving loop: a stream of 2D and 3D jobs admitted as they arrive (continuous batching of the 2D jobs and of the
3D jobs on two streams, resident state under an HBM budget) gives every job the same flux spectra, bit for bit, as
running it alone; the plain FIFO form gives them too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _job_2d(k):
    def build():
        from opensol_amd import Simulation
        rng = np.random.default_rng(100 + k)
        dx = 25e-9
        sx, sy = rng.integers(30, 80) * dx, rng.integers(20, 60) * dx
        sim = Simulation(sx, sy, dx, npml=8, polarization='ez' if k % 2 == 0 else 'hz')
        for _ in range(4):
            x0, y0 = rng.uniform(0, sx), rng.uniform(0, sy)
            sim.add_dielectric_rectangle(x0, y0, x0 + 0.3 * sx, y0 + 0.3 * sy, eps_r=rng.uniform(1, 8))
        sim.add_line_source(axis='y', position=0.25 * sx, wavelength=0.6e-6, bandwidth=0.5)
        sim.add_flux_monitor(axis='y', position=0.75 * sx, span=(0.0, sy), wavelengths=[0.55e-6, 0.65e-6])
        return sim
    return build


def _job_3d(k):
    def build():
        from opensol_amd import Simulation3D
        dx = 25e-9
        size = (0.5e-6 + 0.1e-6 * k, 0.45e-6, 0.4e-6 + 0.2e-6 * (k % 2))
        sim = Simulation3D(*size, dx, npml=6)
        sim.add_dielectric_box((0.1e-6, 0.1e-6, 0.0), (0.3e-6, 0.3e-6, 0.2e-6), eps_r=3.0 + k)
        sim.add_dipole(0.25e-6, 0.22e-6, 0.2e-6, orientation='z', wavelength=0.6e-6)
        sim.add_flux_plane('x', 0.4e-6, wavelengths=[0.55e-6, 0.65e-6])
        return sim
    return build


def _jobs():
    from opensol_amd.serve import Job
    jobs = []
    for k in range(6):
        jobs.append(Job(build=_job_2d(k), steps=60 + 23 * k, arrival=0.03 * k, name=f'2d{k}'))
    for k in range(3):
        jobs.append(Job(build=_job_3d(k), steps=70 + 40 * k, arrival=0.02 + 0.05 * k, name=f'3d{k}'))
    return jobs


def test_served_results_equal_solo_results(gpu):
    from opensol_amd.serve import FDTDServer
    server = FDTDServer(hbm_budget=1e9, chunk_steps=25)
    served = server.run(_jobs())
    plain = server.run_plain(_jobs())
    assert served['jobs'] == plain['jobs'] == 9
    assert served['cell_updates'] == plain['cell_updates']
    for job in _jobs():
        sim = job.build()
        sim.run(n_steps=job.steps, print_every=0)
        solo = [np.asarray(m.flux()) for m in sim.model.monitors]
        sim.close()
        for a, b, c in zip(solo, served['results'][job.name], plain['results'][job.name]):
            assert np.array_equal(a, b), job.name
            assert np.array_equal(a, c), job.name
    assert served['p50_latency_s'] is not None and served['p99_latency_s'] >= served['p50_latency_s']


def test_budget_holds_jobs_until_memory_frees(gpu):
    from opensol_amd.serve import FDTDServer, state_bytes
    jobs = _jobs()[:4]
    need = max(state_bytes(job.build()) for job in jobs)
    budget = need * 1.5
    server = FDTDServer(hbm_budget=budget, chunk_steps=40)
    stats = server.run(jobs)
    assert stats['jobs'] == 4
    for job in jobs:   # the jobs resident when `job` was admitted fit the budget together
        resident = [o for o in jobs if o.admitted <= job.admitted < o.finished]
        assert sum(o.bytes for o in resident) <= budget


def test_cancelled_job_stops_at_a_chunk_boundary(gpu):
    """A running job cancelled at a chunk boundary stops after the chunk in flight, its state is freed and it reports
    'cancelled'; the other jobs' spectra are unchanged (equal to their solo runs bit for bit); a job cancelled while
    pending never runs."""
    from opensol_amd.serve import FDTDServer
    jobs = _jobs()
    victim, never = jobs[1], jobs[7]
    never.cancel()

    def on_chunk(running):   # cancel the victim once it has run one chunk (of four)
        if any(j is victim for j in running) and victim.done_steps > 0:
            victim.cancel()

    stats = FDTDServer(hbm_budget=1e9, chunk_steps=25).run(jobs, on_chunk=on_chunk)
    assert stats['status'][victim.name] == 'cancelled' and victim.done_steps == victim.steps and victim.sim is None
    assert stats['status'][never.name] == 'cancelled' and never.admitted is None
    assert stats['jobs'] == len(jobs) - 2
    for job in _jobs():
        if job.name in (victim.name, never.name):
            assert job.name not in stats['results']
            continue
        sim = job.build()
        sim.run(n_steps=job.steps, print_every=0)
        solo = [np.asarray(m.flux()) for m in sim.model.monitors]
        sim.close()
        for a, b in zip(solo, stats['results'][job.name]):
            assert np.array_equal(a, b), job.name


def _big_job():
    from opensol_amd import Simulation
    dx = 25e-9
    sim = Simulation(150 * dx, 120 * dx, dx, npml=8, polarization='ez')
    sim.add_dielectric_rectangle(1.0e-6, 1.0e-6, 2.5e-6, 2.0e-6, eps_r=4.0)
    sim.add_line_source(axis='y', position=0.6e-6, wavelength=0.6e-6, bandwidth=0.5)
    sim.add_flux_monitor(axis='y', position=3.2e-6, span=(0.0, 120 * dx), wavelengths=[0.55e-6, 0.65e-6])
    return sim


def test_preempted_job_resumes_bit_for_bit(gpu):
    """Two long jobs fill the budget; a larger job arrives and, with preempt_after=0, preempts the later-admitted one:
    its state goes to the host and its device state is freed, the new job runs, and the preempted job is restored
    into a new engine when memory frees and finishes. All three spectra equal their solo runs bit for bit."""
    from opensol_amd.serve import FDTDServer, Job, state_bytes

    def jobs():
        return [Job(build=_job_2d(0), steps=3000, arrival=0.0, name='first'),
                Job(build=_job_2d(1), steps=3000, arrival=0.005, name='second'),
                Job(build=_big_job, steps=300, arrival=0.05, name='big')]

    sizes = {job.name: state_bytes(job.build()) for job in jobs()}
    assert sizes['big'] > sizes['second']
    budget = sizes['first'] + sizes['big'] + 1.0   # both small jobs fit; the big one only without the second
    stream = jobs()
    stats = FDTDServer(hbm_budget=budget, chunk_steps=25, autotune=False, preempt_after=0.0).run(stream)
    first, second, big = stream
    print('preemptions', stats['preemptions'], 'admitted', [j.admitted for j in stream],
          'finished', [j.finished for j in stream], flush=True)
    assert stats['jobs'] == 3 and stats['preemptions'] == 1
    assert second.preemptions == 1 and first.preemptions == 0 and big.preemptions == 0
    assert big.finished < second.finished
    for job in jobs():
        sim = job.build()
        sim.run(n_steps=job.steps, print_every=0)
        solo = [np.asarray(m.flux()) for m in sim.model.monitors]
        sim.close()
        for a, b in zip(solo, stats['results'][job.name]):
            assert np.array_equal(a, b), job.name


def _nan_job():
    """A job whose state holds a NaN before it is served (planted from the host into one E_z cell)."""
    sim = _job_2d(2)()
    sim.finalize()
    values = sim.engine.field('ez')
    values[values.shape[0] // 2, values.shape[1] // 2] = np.nan
    sim.engine.set_field('ez', values)
    return sim


def test_diverging_job_fails_at_the_first_chunk_boundary(gpu):
    """The device finite check after every chunk flags a job whose fields hold a NaN: it stops after the chunk that
    showed it (status 'failed', the step in the error) instead of running to its end; the other job is unchanged."""
    from opensol_amd.serve import FDTDServer, Job
    good = Job(build=_job_2d(3), steps=200, arrival=0.0, name='good')
    bad = Job(build=_nan_job, steps=200, arrival=0.0, name='bad')
    stats = FDTDServer(hbm_budget=1e9, chunk_steps=25, autotune=False).run([good, bad])
    print(stats['status'], stats['errors'], flush=True)
    assert stats['status'] == {'good': 'done', 'bad': 'failed'}
    assert 'after step 25 (device check' in stats['errors']['bad']
    assert bad.sim is None and stats['jobs'] == 1
    sim = good.build()
    sim.run(n_steps=good.steps, print_every=0)
    for a, b in zip([np.asarray(m.flux()) for m in sim.model.monitors], stats['results']['good']):
        assert np.array_equal(a, b)
    sim.close()
