"""Benchmark of the MI355X FDTD hot path: Yee cell-updates/s on one GPU, with the HBM roofline of the dominant step
kernel and the numpy CPU baseline timed on this host.

Default workload (one step = one full Yee step of the whole grid, CPML included): 2D E_z, 8192 x 8192 cells with a
20-cell CPML on every side (window 8152^2 cells), dx = 20 nm, 400 seeded axis-aligned dielectric rectangles (sides
uniform 0.1-2 um, eps_r uniform in [1, 12], painted in draw order), one Gaussian line source (lambda0 = 1.55 um,
bandwidth 0.5) along y at x = 2 um, 2000 steps by default. Inputs are resident in HBM before timing starts.

--workload yee3d_512: 3D, 512^3 cells with a 12-cell CPML on every face (window 488^3, 12.2 um), dx = 25 nm, 200
seeded dielectric boxes (edges uniform 0.2-3 um, eps_r uniform in [1, 12]), a Gaussian E_z point dipole at the centre
(lambda0 = 1 um, bandwidth 0.5), two flux planes normal to x at the centre -+ 4 um at 0.8 / 0.9 / 1.0 / 1.1 um;
500 steps by default.

--workload batch256: 256 seeded 2D instances (half E_z, half H_z), sides log-uniform in [128, 1024] cells, dx = 25 nm,
12-cell CPML, 20-100 rectangles each (eps_r in [1, 12]), 1000-4000 steps each, one Gaussian line source (lambda0 in
[0.6, 1.6] um) and one flux line at 3 wavelengths each. Reports the cell-updates/s of one batched launch per step
(Batch2D) and of the plain form (each instance run alone, one after another, with the same kernels and graphs).

--workload serve: a seeded stream of jobs with Poisson arrivals at 50, 100, 200 and 400 jobs/s (--jobs per rate): 80 %
2D (as batch256, 500-5000 steps) and 20 % 3D (64^3-192^3 cells, 8-cell CPML, 5-40 boxes, one dipole, one flux plane,
500-5000 steps), served by FDTDServer (continuous batching, HBM budget 200 GB) and by the plain FIFO form (one job at
a time). Reports p50 / p99 latency and aggregate cell-updates/s of both.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 50] [--workload ez8192 | yee3d_512 | batch256 | serve]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CELL_2D = 26          # read E, H_a, H_b + uint16 id, write E, H_a, H_b (fp32)
BYTES_PER_CELL_3D = 50          # read the six fields + uint16 id, write the six fields (fp32)
HBM_PEAK_GBS = 8000.0           # MI355X HBM3E peak
SEED = 20260512
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'traffic.json')


def window_for_cells(n_total, npml, dx):
    """Window size [m] whose grid (with npml CPML cells per side) has exactly n_total cells."""
    from opensol_amd.domain import uniform_cell_count
    size = (n_total - 2 * npml) * dx
    for nudge in (0.0, -1e-7, 1e-7, -1e-6, 1e-6):
        candidate = size + nudge * dx
        if uniform_cell_count(candidate + 2 * npml * dx, dx) == n_total:
            return candidate
    raise ValueError(f'no window size gives {n_total} cells')


def build_ez_workload(n=8192, npml=20, dx=20e-9, rectangles=400, seed=SEED, target=None, **options):
    """The 2D E_z roofline workload as a Simulation (not yet finalized); `target` builds the same model on another
    class with the Simulation constructor and methods (the oracle's Yee2D in the parity tests)."""
    if target is None:
        from opensol_amd import Simulation as target
    size = window_for_cells(n, npml, dx)
    sim = target(size, size, dx, npml=npml, **options)
    rng = np.random.default_rng(seed)
    for _ in range(rectangles):
        cx, cy = rng.uniform(0, size, 2)
        wx, wy = rng.uniform(0.1e-6, 2e-6, 2)
        eps = rng.uniform(1.0, 12.0)
        sim.add_dielectric_rectangle(cx - wx / 2, cy - wy / 2, cx + wx / 2, cy + wy / 2, eps_r=eps)
    sim.add_line_source(axis='y', position=2e-6, wavelength=1.55e-6, bandwidth=0.5)
    shape = sim.model.shape if hasattr(sim, 'model') else (sim.nx, sim.ny)
    assert shape == (n, n), shape
    return sim


def build_batch_instances(count=256, seed=SEED, target=None, **options):
    """The batched-2D workload: (simulations, steps per instance); the first k instances of a count are the same for
    every count. `target` as in build_ez_workload."""
    if target is None:
        from opensol_amd import Simulation as target
    rng = np.random.default_rng(seed)
    dx, npml = 25e-9, 12
    sims, steps = [], []
    for k in range(count):
        nx, ny = (int(round(np.exp(rng.uniform(np.log(128), np.log(1024))))) for _ in range(2))
        sx, sy = window_for_cells(nx, npml, dx), window_for_cells(ny, npml, dx)
        sim = target(sx, sy, dx, npml=npml, polarization='ez' if k % 2 == 0 else 'hz', **options)
        for _ in range(int(rng.integers(20, 101))):
            cx, cy = rng.uniform(0, sx), rng.uniform(0, sy)
            wx, wy = rng.uniform(0.05, 0.4) * sx, rng.uniform(0.05, 0.4) * sy
            sim.add_dielectric_rectangle(cx - wx / 2, cy - wy / 2, cx + wx / 2, cy + wy / 2,
                                         eps_r=rng.uniform(1.0, 12.0))
        wavelength = rng.uniform(0.6e-6, 1.6e-6)
        sim.add_line_source(axis='y', position=0.25 * sx, wavelength=wavelength, bandwidth=0.5)
        sim.add_flux_monitor(axis='y', position=0.75 * sx, span=(0.0, sy),
                             wavelengths=[0.9 * wavelength, wavelength, 1.1 * wavelength])
        sims.append(sim)
        steps.append(int(rng.integers(1000, 4001)))
    return sims, steps


def serve_jobs(rate, count, seed=SEED):
    """Seeded job stream: Poisson arrivals at `rate` jobs/s, 80 % 2D and 20 % 3D jobs."""
    from opensol_amd.serve import Job
    rng = np.random.default_rng(seed + int(rate))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    jobs = []
    for k, arrival in enumerate(arrivals):
        job_seed = int(rng.integers(1 << 30))
        steps = int(rng.integers(500, 5001))
        if rng.random() < 0.8:
            jobs.append(Job(build=_serve_2d(job_seed), steps=steps, arrival=float(arrival), name=f'2d-{k}'))
        else:
            jobs.append(Job(build=_serve_3d(job_seed), steps=steps, arrival=float(arrival), name=f'3d-{k}'))
    return jobs


def _serve_2d(seed):
    def build():
        from opensol_amd import Simulation
        rng = np.random.default_rng(seed)
        dx, npml = 25e-9, 12
        nx, ny = (int(round(np.exp(rng.uniform(np.log(128), np.log(1024))))) for _ in range(2))
        sx, sy = window_for_cells(nx, npml, dx), window_for_cells(ny, npml, dx)
        sim = Simulation(sx, sy, dx, npml=npml, polarization='ez' if rng.random() < 0.5 else 'hz')
        for _ in range(int(rng.integers(20, 101))):
            cx, cy = rng.uniform(0, sx), rng.uniform(0, sy)
            wx, wy = rng.uniform(0.05, 0.4) * sx, rng.uniform(0.05, 0.4) * sy
            sim.add_dielectric_rectangle(cx - wx / 2, cy - wy / 2, cx + wx / 2, cy + wy / 2,
                                         eps_r=rng.uniform(1.0, 12.0))
        wavelength = rng.uniform(0.6e-6, 1.6e-6)
        sim.add_line_source(axis='y', position=0.25 * sx, wavelength=wavelength, bandwidth=0.5)
        sim.add_flux_monitor(axis='y', position=0.75 * sx, span=(0.0, sy),
                             wavelengths=[0.9 * wavelength, wavelength, 1.1 * wavelength])
        return sim
    return build


def _serve_3d(seed):
    def build():
        from opensol_amd import Simulation3D
        rng = np.random.default_rng(seed)
        dx, npml = 25e-9, 8
        n = int(rng.integers(64, 193))
        size = window_for_cells(n, npml, dx)
        sim = Simulation3D(size, size, size, dx, npml=npml)
        for _ in range(int(rng.integers(5, 41))):
            centre = rng.uniform(0, size, 3)
            edges = rng.uniform(0.1, 0.5, 3) * size
            sim.add_dielectric_box(tuple(centre - edges / 2), tuple(centre + edges / 2), eps_r=rng.uniform(1.0, 12.0))
        c = size / 2
        wavelength = rng.uniform(0.6e-6, 1.6e-6)
        sim.add_dipole(c, c, c, orientation='z', wavelength=wavelength, bandwidth=0.5)
        sim.add_flux_plane('x', c + 0.3 * size, wavelengths=[0.9 * wavelength, wavelength, 1.1 * wavelength])
        return sim
    return build


def run_serve(args, torch):
    from opensol_amd.serve import FDTDServer
    server = FDTDServer(hbm_budget=200e9, chunk_steps=100, graph_steps=0, autotune=True)
    tune_start = time.perf_counter()   # the job classes' batch-member shapes, ahead of the timed streams
    classes = server.tune_classes([job for rate in (50, 100, 200, 400) for job in serve_jobs(rate, args.jobs)])
    tune_s = time.perf_counter() - tune_start
    print(f'# autotuned {len(classes)} job classes in {tune_s:.1f} s', file=sys.stderr, flush=True)
    rates = {}
    for rate in (50, 100, 200, 400):
        served = server.run(serve_jobs(rate, args.jobs))
        plain = server.run_plain(serve_jobs(rate, args.jobs))
        for name, stats in (('served', served), ('plain', plain)):
            stats.pop('results')
        rates[str(rate)] = {'served': served, 'plain': plain,
                            'throughput_ratio': served['cell_updates_per_s'] / plain['cell_updates_per_s']}
        print(f'# rate {rate}/s: served p50 {served["p50_latency_s"]:.3f} s p99 {served["p99_latency_s"]:.3f} s '
              f'{served["cell_updates_per_s"]:.3e}/s | plain p50 {plain["p50_latency_s"]:.3f} s p99 '
              f'{plain["p99_latency_s"]:.3f} s {plain["cell_updates_per_s"]:.3e}/s | ratio '
              f'{rates[str(rate)]["throughput_ratio"]:.2f} | served host phases {served["host_phases_s"]}',
              file=sys.stderr, flush=True)
    top = rates['400']['served']
    return {'metric': 'served FDTD stream: aggregate Yee cell-updates/s, p50/p99 latency vs plain FIFO',
            'value': top['cell_updates_per_s'], 'unit': 'cell-updates/s', 'n_gpus': 1, 'steps': None, 'warmup': 0,
            'ms_per_step': None, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
            'data': 'synthetic (seeded Poisson job stream, 80 % 2D / 20 % 3D)',
            'config': {'workload': 'serve_2d3d_poisson', 'jobs_per_rate': args.jobs, 'hbm_budget_B': 200e9,
                       'chunk_steps': 100, 'seed': SEED, 'autotune': {'classes': len(classes), 'seconds': tune_s}},
            'rates': rates, 'roofline': None, 'cpu_baseline': None}


def build_3d_workload(n=512, npml=12, dx=25e-9, boxes=200, seed=SEED, target=None, **options):
    """The 3D roofline workload as a Simulation3D (not yet finalized); `target` as in build_ez_workload (the
    oracle's Yee3D)."""
    if target is None:
        from opensol_amd import Simulation3D as target
    size = window_for_cells(n, npml, dx)
    sim = target(size, size, size, dx, npml=npml, **options)
    rng = np.random.default_rng(seed)
    for _ in range(boxes):
        centre = rng.uniform(0, size, 3)
        edges = rng.uniform(0.2e-6, 3e-6, 3)
        sim.add_dielectric_box(tuple(centre - edges / 2), tuple(centre + edges / 2), eps_r=rng.uniform(1.0, 12.0))
    c = size / 2
    sim.add_dipole(c, c, c, orientation='z', dipole_moment=1e-12, wavelength=1e-6, bandwidth=0.5)
    for x in (c - 4e-6, c + 4e-6):
        sim.add_flux_plane('x', x, wavelengths=[0.8e-6, 0.9e-6, 1.0e-6, 1.1e-6])
    shape = sim.model.shape if hasattr(sim, 'model') else sim.shape
    assert tuple(shape) == (n, n, n), shape
    return sim


def cpu_baseline_3d(sim, steps):
    """The 3D oracle in float32 (the reference's numpy-twin array passes) on the same grid, materials, dipole and
    planes, timed for `steps` steps on one host core."""
    from oracle.yee3d import Yee3D
    model = sim.model
    size = model.domain.sizes
    ref = Yee3D(*size, model.domain.dx, npml=model.bc.cpml_npml, dtype=np.float32)
    assert ref.shape == model.shape
    ref.materials['eps_r'][...] = model.materials.property_array(0)
    c = size[0] / 2
    ref.add_dipole(c, c, c, orientation='z', dipole_moment=1e-12, wavelength=1e-6, bandwidth=0.5)
    for x in (c - 4e-6, c + 4e-6):
        ref.add_flux_plane('x', x, wavelengths=[0.8e-6, 0.9e-6, 1.0e-6, 1.1e-6])
    ref.finalize()
    return _time_on_one_core(ref, steps, model.cell_count,
                             f'{{steps}} steps of the same {model.nx}x{model.ny}x{model.nz} grid (numpy float32, '
                             f'reference numpy-twin array passes)')


def _time_on_one_core(ref, steps, cells, sample):
    affinity = None
    if hasattr(os, 'sched_getaffinity'):
        affinity = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(affinity)})
    try:
        start = time.perf_counter()
        for _ in range(steps):
            ref.step()
        elapsed = time.perf_counter() - start
    finally:
        if affinity is not None:
            os.sched_setaffinity(0, affinity)
    return {'value': cells * steps / elapsed, 'unit': 'cell-updates/s', 'cores': 1, 'kind': 'port',
            'sample': sample.format(steps=steps) + f', {elapsed:.1f} s on 1 core of {_cpu_model()}'}


def cpu_baseline(sim, steps):
    """The numpy restatement of the reference's numpy-twin step (oracle, float32) on the same grid, materials and
    source, timed for `steps` steps on one host core."""
    from oracle.yee2d import Yee2D
    model = sim.model
    size = model.domain.size_x
    ref = Yee2D(size, model.domain.size_y, model.domain.dx, npml=model.bc.cpml_npml, dtype=np.float32)
    assert (ref.nx, ref.ny) == model.shape
    ref.materials['eps_r'][...] = model.materials.property_array(0)
    ref.add_line_source(axis='y', position=2e-6, wavelength=1.55e-6, bandwidth=0.5)
    ref.finalize()
    ref.step()
    return _time_on_one_core(ref, steps, model.nx * model.ny,
                             f'{{steps}} steps of the same {model.nx}x{model.ny} E_z grid (numpy float32, reference '
                             f'numpy-twin array passes)')


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown CPU'


def measured_traffic(workload, field='bytes_per_launch'):
    """HBM bytes per launch (or per step's tile-kernel pair, field 'tiles_bytes_per_step') from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            entry = json.load(f).get(workload)
        return None if entry is None else float(entry[field])
    except (OSError, ValueError, KeyError):
        return None


def _common_fields(args, value, per_step):
    return {'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': per_step * 1e3,
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32'}


def graph_steps_for(steps, longest):
    """Steps per replayed graph for a timed run of `steps` steps: the largest divisor of `steps` not above `longest`,
    so the whole timed region replays captured graphs."""
    longest = max(1, min(int(longest), int(steps)))
    return max(d for d in range(1, longest + 1) if steps % d == 0)


def _time_steps(engine, steps, torch):
    """Seconds for `steps` steps on the library's stream, HIP events on both sides after a device sync."""
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    engine.step(steps)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3


def run_2d(args, torch):
    sim = build_ez_workload(graph_steps=args.graph_steps, rows_per_tile=args.rows_per_tile,
                            lane_columns=args.lane_columns)
    sim.finalize()
    engine = sim.engine
    rows_per_tile, lane_columns = engine.launch_geometry()
    cells = sim.model.nx * sim.model.ny
    print(f'# workload ready: {sim.model.nx}x{sim.model.ny}, {len(sim.model.materials.materials)} materials',
          file=sys.stderr, flush=True)
    engine.prepare(args.warmup + args.steps + args.kernel_steps)   # tables and the step graph, outside the timing
    engine.step(args.warmup)
    elapsed_s = _time_steps(engine, args.steps, torch)
    per_step = elapsed_s / args.steps
    value = cells * args.steps / elapsed_s
    step_achieved = BYTES_PER_CELL_2D * cells / per_step / 1e9
    workload = 'yee2d_ez_8192x8192_cpml20_400rect'
    print(f'# GPU: {per_step * 1e3:.4f} ms/step, {value:.4e} cell-updates/s, {step_achieved:.1f} GB/s algorithmic',
          file=sys.stderr, flush=True)
    # the dominant kernel (bulk tiles): its own launch time from HIP events on the library's stream
    kernels = engine.step_timed(args.kernel_steps) if args.kernel_steps > 0 else None
    roofline = {'bound': 'hbm', 'achieved': None, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': None,
                'traffic': measured_traffic(workload), 'kernel': 'yee2d_step_kernel<EZ, bulk tiles>',
                'bytes_per_cell_update': BYTES_PER_CELL_2D, 'step_achieved': step_achieved,
                'step_frac': step_achieved / HBM_PEAK_GBS}
    if kernels is not None and 'tiles' in kernels:
        # The bulk and edge tile kernels run side by side and together are the step (plus one clock kernel per
        # graph): the measured unit is the pair, timed over the timed region as the step period (fork to fork).
        # yee_step_timed's eager fork-to-join time, with a host synchronization every step, is reported beside it.
        tiles_ms, tiles_cells = kernels['tiles']
        roofline.update(achieved=step_achieved, frac=step_achieved / HBM_PEAK_GBS, launch_ms=per_step * 1e3,
                        cells_per_launch=cells, algorithmic_bytes_per_launch=BYTES_PER_CELL_2D * cells,
                        kernel='yee2d_step_kernel<EZ> bulk + edge tile sets, side by side (step period)',
                        traffic=measured_traffic(workload, 'tiles_bytes_per_step'), eager_tiles_ms=tiles_ms)
        print(f'# kernels: bulk + edge tiles side by side, {per_step * 1e6:.1f} us per step in the timed region '
              f'({step_achieved:.1f} GB/s); eager fork-to-join {tiles_ms * 1e3:.1f} us', file=sys.stderr, flush=True)
    elif kernels is not None:
        bulk_ms, bulk_cells = kernels['bulk']
        achieved = BYTES_PER_CELL_2D * bulk_cells / (bulk_ms * 1e-3) / 1e9
        roofline.update(achieved=achieved, frac=achieved / HBM_PEAK_GBS, launch_ms=bulk_ms, cells_per_launch=bulk_cells,
                        algorithmic_bytes_per_launch=BYTES_PER_CELL_2D * bulk_cells,
                        edge_launch_ms=kernels['edge'][0], edge_cells_per_launch=kernels['edge'][1])
        print(f'# kernels: bulk {bulk_ms * 1e3:.1f} us for {bulk_cells} cells ({achieved:.1f} GB/s), edge '
              f'{kernels["edge"][0] * 1e3:.1f} us for {kernels["edge"][1]} cells', file=sys.stderr, flush=True)
    baseline = None
    if args.cpu_steps > 0:
        baseline = cpu_baseline(sim, args.cpu_steps)
        print(f'# CPU baseline: {baseline["value"]:.4e} cell-updates/s', file=sys.stderr, flush=True)
    result = {'metric': 'Yee cell-updates/s (HBM BW fraction in roofline)', 'value': value, 'unit': 'cell-updates/s',
              **_common_fields(args, value, per_step),
              'data': 'synthetic (seeded rectangles, Gaussian line source)',
              'config': {'workload': workload, 'grid': [sim.model.nx, sim.model.ny], 'cpml': 20, 'dx_m': 20e-9,
                         'rectangles': 400, 'seed': SEED, 'polarization': 'ez', 'graph_steps': args.graph_steps,
                         'rows_per_tile': rows_per_tile, 'lane_columns': lane_columns},
              'roofline': roofline, 'cpu_baseline': baseline}
    return result


def run_features(args, torch):
    """The 8192^2 E_z workload with the monitors and media of the wider path on top: a running-DFT box over the whole
    window (3 wavelengths), a Drude rectangle over the middle third of the rows, 64 field probes. Reports the step time
    with and without them; rocprofv3 gives each feature kernel's own time (profiles/)."""
    from opensol_amd import DrudePole
    plain = build_ez_workload(graph_steps=args.graph_steps)
    plain.finalize()
    plain.engine.prepare(args.warmup + args.steps)
    plain.engine.step(args.warmup)
    plain_s = _time_steps(plain.engine, args.steps, torch) / args.steps
    plain.close()
    del plain
    sim = build_ez_workload(graph_steps=args.graph_steps)
    size = sim.model.domain.sizes[0]
    box = sim.add_dft_monitor(0.0, 0.0, size, size, wavelengths=[1.4e-6, 1.55e-6, 1.7e-6])
    sim.add_drude_rectangle(size / 3, 0.0, 2 * size / 3, size, DrudePole(omega_p=1.5e15, gamma=1e14, eps_inf=4.0))
    rng = np.random.default_rng(SEED)
    for x, y in rng.uniform(0.0, size, (64, 2)):
        sim.add_field_probe(x, y)
    sim.finalize()
    engine = sim.engine
    cells = sim.model.nx * sim.model.ny
    engine.prepare(args.warmup + args.steps)
    engine.step(args.warmup)
    elapsed_s = _time_steps(engine, args.steps, torch)
    per_step = elapsed_s / args.steps
    m = sim.model.sample_interval
    dft_cells = box.mx * box.my
    dft_bytes = dft_cells * (4 + 64 * box.n_freqs)   # E_z read, Kahan (sum, comp) FP64 (re, im) read + write per frequency
    print(f'# features: {per_step * 1e3:.4f} ms/step against {plain_s * 1e3:.4f} plain; DFT box {box.mx}x{box.my} x '
          f'{box.n_freqs} every {m} steps ({dft_bytes / 1e9:.2f} GB per sample)', file=sys.stderr, flush=True)
    value = cells * args.steps / elapsed_s
    return {'metric': 'Yee cell-updates/s (HBM BW fraction in roofline)', 'value': value, 'unit': 'cell-updates/s',
            **_common_fields(args, value, per_step),
            'data': 'synthetic (seeded rectangles, Gaussian line source, Drude rectangle, DFT box, probes)',
            'config': {'workload': 'yee2d_ez_8192x8192_cpml20_400rect_dftbox3_drude_probes64', 'grid': [sim.model.nx,
                       sim.model.ny], 'graph_steps': args.graph_steps, 'sample_interval': m},
            'features': {'plain_ms_per_step': plain_s * 1e3, 'ms_per_step': per_step * 1e3,
                         'overhead': per_step / plain_s - 1.0, 'dft_cells': dft_cells, 'dft_frequencies': box.n_freqs,
                         'dft_bytes_per_sample': dft_bytes, 'drude_rows': int(sim.model.poles[0][1].any(axis=1).sum()),
                         'probes': 64},
            'roofline': None, 'cpu_baseline': None}


def run_3d(args, torch):
    sim = build_3d_workload(graph_steps=args.graph_steps, planes_per_tile=args.planes_per_tile)
    sim.finalize()
    engine = sim.engine
    rows_per_tile, planes_per_tile = engine.launch_geometry()
    cells = sim.model.cell_count
    print(f'# workload ready: {sim.model.shape}, {len(sim.model.materials.materials)} materials, '
          f'm = {sim.model.sample_interval}', file=sys.stderr, flush=True)
    engine.prepare(args.warmup + args.steps + args.kernel_steps)
    engine.step(args.warmup)
    elapsed_s = _time_steps(engine, args.steps, torch)
    per_step = elapsed_s / args.steps
    value = cells * args.steps / elapsed_s
    step_achieved = BYTES_PER_CELL_3D * cells / per_step / 1e9
    workload = 'yee3d_512cubed_cpml12_200box_dipole_2planes'
    print(f'# GPU: {per_step * 1e3:.4f} ms/step, {value:.4e} cell-updates/s, {step_achieved:.1f} GB/s algorithmic',
          file=sys.stderr, flush=True)
    kernels = engine.step_timed(args.kernel_steps) if args.kernel_steps > 0 else None
    roofline = {'bound': 'hbm', 'achieved': None, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': None,
                'traffic': measured_traffic(workload), 'kernel': 'yee3d_step_kernel<4>',
                'bytes_per_cell_update': BYTES_PER_CELL_3D, 'step_achieved': step_achieved,
                'step_frac': step_achieved / HBM_PEAK_GBS}
    if kernels is not None:
        step_ms, step_cells = kernels['step']
        achieved = BYTES_PER_CELL_3D * step_cells / (step_ms * 1e-3) / 1e9
        roofline.update(achieved=achieved, frac=achieved / HBM_PEAK_GBS, launch_ms=step_ms,
                        cells_per_launch=step_cells, algorithmic_bytes_per_launch=BYTES_PER_CELL_3D * step_cells,
                        sample_launch_ms=kernels['sample'][0], sample_cells_per_launch=kernels['sample'][1])
        print(f'# kernels: step {step_ms * 1e3:.1f} us for {step_cells} cells ({achieved:.1f} GB/s), sample '
              f'{kernels["sample"][0] * 1e3:.1f} us', file=sys.stderr, flush=True)
    baseline = None
    if args.cpu_steps > 0:
        baseline = cpu_baseline_3d(sim, args.cpu_steps)
        print(f'# CPU baseline: {baseline["value"]:.4e} cell-updates/s', file=sys.stderr, flush=True)
    result = {'metric': 'Yee cell-updates/s (HBM BW fraction in roofline)', 'value': value, 'unit': 'cell-updates/s',
              **_common_fields(args, value, per_step),
              'data': 'synthetic (seeded boxes, Gaussian point dipole, two flux planes)',
              'config': {'workload': workload, 'grid': list(sim.model.shape), 'cpml': 12, 'dx_m': 25e-9, 'boxes': 200,
                         'seed': SEED, 'graph_steps': args.graph_steps, 'rows_per_tile': rows_per_tile,
                         'planes_per_tile': planes_per_tile, 'sample_interval': sim.model.sample_interval},
              'roofline': roofline, 'cpu_baseline': baseline}
    return result


def run_batch(args, torch):
    from opensol_amd.batch import Batch2D
    plain_sims, steps = build_batch_instances(graph_steps=args.graph_steps)
    batch_sims, _ = build_batch_instances(graph_steps=args.graph_steps)
    cell_updates = sum(s.model.nx * s.model.ny * n for s, n in zip(plain_sims, steps))
    print(f'# workload ready: {len(steps)} instances, {cell_updates:.4e} cell-updates', file=sys.stderr, flush=True)
    for sim, n in zip(plain_sims, steps):   # build the engines, tables and graphs outside the timed region
        sim.finalize()
        sim.engine.prepare(n)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for sim, n in zip(plain_sims, steps):
        sim.engine.step(n)
    stop.record()
    torch.cuda.synchronize()
    plain_s = start.elapsed_time(stop) / 1e3
    print(f'# plain: {plain_s:.3f} s, {cell_updates / plain_s:.4e} cell-updates/s', file=sys.stderr, flush=True)
    tune_start = time.perf_counter()   # each member's rows per tile from its size class (tuned once per device)
    batch = Batch2D(batch_sims, graph_steps=args.graph_steps, autotune=True)
    tune_s = time.perf_counter() - tune_start
    chosen = {}
    for sim in batch_sims:
        rows = int(sim._engine_options.get('rows_per_tile') or 0)
        chosen[str(rows)] = chosen.get(str(rows), 0) + 1
    print(f'# autotuned members in {tune_s:.2f} s: rows per tile {chosen}', file=sys.stderr, flush=True)
    batch.prepare(steps)   # descriptors and the step graph, outside the timed region
    torch.cuda.synchronize()
    start.record()
    launches = batch.launch()
    stop.record()
    torch.cuda.synchronize()
    batch_s = start.elapsed_time(stop) / 1e3
    value = cell_updates / batch_s
    print(f'# batched: {batch_s:.3f} s, {value:.4e} cell-updates/s, {launches} launches, '
          f'{(cell_updates / plain_s and value / (cell_updates / plain_s)):.2f}x plain', file=sys.stderr, flush=True)
    # every instance matches the same instance stepped alone (bit for bit; checked on a sample here and fully in
    # tests/test_gpu_batch.py)
    for k in range(0, len(steps), 37):
        for name in plain_sims[k].model.components:
            assert np.array_equal(plain_sims[k].engine.field(name), batch_sims[k].engine.field(name)), (k, name)
    return {'metric': 'Yee cell-updates/s, batched vs plain one-launch-per-instance', 'value': value,
            'unit': 'cell-updates/s', 'n_gpus': 1, 'steps': max(steps), 'warmup': 0,
            'ms_per_step': batch_s / launches * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'f32', 'data': 'synthetic (seeded instances: rectangles, Gaussian line source, flux line)',
            'config': {'workload': 'batch256_2d_ez_hz_128to1024', 'instances': len(steps), 'seed': SEED,
                       'cell_updates': cell_updates, 'graph_steps': args.graph_steps,
                       'autotune': {'seconds': tune_s, 'rows_per_tile': chosen}},
            'plain': {'value': cell_updates / plain_s, 'seconds': plain_s},
            'batched': {'value': value, 'seconds': batch_s, 'launches': launches},
            'speedup_vs_plain': value / (cell_updates / plain_s),
            'roofline': None, 'cpu_baseline': None}


def main():
    parser = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    parser.add_argument('--gpus', type=int, default=1)
    parser.add_argument('--workload', default='ez8192',
                        choices=('ez8192', 'yee3d_512', 'batch256', 'serve', 'features8192'))
    parser.add_argument('--jobs', type=int, default=200, help='serve: jobs per arrival rate')
    parser.add_argument('--steps', type=int, default=None, help='timed steps (default 2000 2D, 500 3D)')
    parser.add_argument('--warmup', type=int, default=None, help='untimed steps (default 50 2D, 20 3D)')
    parser.add_argument('--graph-steps', type=int, default=None, help='steps per HIP graph (default 50 2D, 10 3D)')
    parser.add_argument('--rows-per-tile', type=int, default=0, help='2D: rows per tile (0: the library default, 8)')
    parser.add_argument('--lane-columns', type=int, default=0, choices=(0, 2, 4), help='2D: 0 = library default')
    parser.add_argument('--planes-per-tile', type=int, default=0, help='3D: x planes per tile (0: the library default, 4)')
    parser.add_argument('--cpu-steps', type=int, default=None,
                        help='steps of the CPU baseline sample (default 8 2D, 1 3D; 0 skips it)')
    parser.add_argument('--kernel-steps', type=int, default=None,
                        help='steps timed per kernel with HIP events for the roofline (default 200 2D, 50 3D; 0 skips)')
    parser.add_argument('--no-sub-records', action='store_true',
                        help='ez8192: skip the 3D 512^3 and batch256 sub-records of the default run')
    args = parser.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: this FDTD path runs on a single MI355X (got --gpus {args.gpus}); '
              f'run it with --gpus 1', file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a GPU', file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    three = args.workload == 'yee3d_512'
    requested = argparse.Namespace(**vars(args))
    resolve_defaults(args, three)
    if args.workload == 'batch256':
        result = run_batch(args, torch)
    elif args.workload == 'serve':
        result = run_serve(args, torch)
    elif args.workload == 'features8192':
        result = run_features(args, torch)
    elif three:
        result = run_3d(args, torch)
    else:
        result = run_2d(args, torch)
        if not args.no_sub_records:
            result['sub_records'] = sub_records(requested, torch)
    print(json.dumps(result))


def resolve_defaults(args, three):
    """Defaults of the 2D or 3D workload for the options left unset; graph_steps divides the timed steps."""
    defaults = dict(steps=(500, 2000), warmup=(20, 50), graph_steps=(10, 50), cpu_steps=(1, 8), kernel_steps=(50, 200))
    for name, (d3, d2) in defaults.items():
        if getattr(args, name) is None:
            setattr(args, name, d3 if three else d2)
    args.graph_steps = graph_steps_for(args.steps, args.graph_steps) if args.graph_steps > 0 else 0
    return args


def sub_records(requested, torch):
    """The 3D 512^3 roofline line (same --steps / --warmup, its own roofline and CPU baseline) and the batch256 ratio
    (its instances' own step counts), measured in the default run after the 2D line."""
    import gc
    out = {}
    gc.collect()
    torch.cuda.empty_cache()
    three = resolve_defaults(argparse.Namespace(**{**vars(requested), 'workload': 'yee3d_512', 'graph_steps': None,
                                                   'cpu_steps': None, 'kernel_steps': None}), True)
    out['yee3d_512'] = run_3d(three, torch)
    gc.collect()
    torch.cuda.empty_cache()
    batch = argparse.Namespace(**{**vars(requested), 'workload': 'batch256', 'graph_steps': 50})
    out['batch256'] = run_batch(batch, torch)
    gc.collect()
    torch.cuda.empty_cache()
    return out


if __name__ == '__main__':
    main()
