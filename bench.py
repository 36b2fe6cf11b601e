"""Benchmark of the MI355X FDTD hot path: Yee cell-updates/s on one GPU, with the HBM roofline of the fused step
kernel and the numpy CPU baseline timed on this host.

Workload (one step = one full Yee step of the whole grid, CPML included): 2D E_z, 8192 x 8192 cells with a 20-cell
CPML on every side (window 8152^2 cells), dx = 20 nm, 400 seeded axis-aligned dielectric rectangles (sides uniform
0.1-2 um, eps_r uniform in [1, 12], painted in draw order), one Gaussian line source (lambda0 = 1.55 um, bandwidth 0.5)
along y at x = 2 um, 2000 steps by default. Inputs are resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 50]
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


def build_ez_workload(n=8192, npml=20, dx=20e-9, rectangles=400, seed=SEED, **options):
    """The 2D E_z roofline workload as a Simulation (not yet finalized) plus its description."""
    from opensol_amd import Simulation
    size = window_for_cells(n, npml, dx)
    sim = Simulation(size, size, dx, npml=npml, **options)
    rng = np.random.default_rng(seed)
    for _ in range(rectangles):
        cx, cy = rng.uniform(0, size, 2)
        wx, wy = rng.uniform(0.1e-6, 2e-6, 2)
        eps = rng.uniform(1.0, 12.0)
        sim.add_dielectric_rectangle(cx - wx / 2, cy - wy / 2, cx + wx / 2, cy + wy / 2, eps_r=eps)
    sim.add_line_source(axis='y', position=2e-6, wavelength=1.55e-6, bandwidth=0.5)
    assert sim.model.shape == (n, n), sim.model.shape
    return sim


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
    cells = model.nx * model.ny
    return {'value': cells * steps / elapsed, 'unit': 'cell-updates/s', 'cores': 1, 'kind': 'port',
            'sample': f'{steps} steps of the same {model.nx}x{model.ny} E_z grid (numpy float32, reference numpy-twin '
                      f'array passes), {elapsed:.1f} s on 1 core of {_cpu_model()}'}


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown CPU'


def measured_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            entry = json.load(f).get(workload)
        return None if entry is None else float(entry['bytes_per_launch'])
    except (OSError, ValueError, KeyError):
        return None


def main():
    parser = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    parser.add_argument('--gpus', type=int, default=1)
    parser.add_argument('--steps', type=int, default=2000)
    parser.add_argument('--warmup', type=int, default=50)
    parser.add_argument('--graph-steps', type=int, default=50)
    parser.add_argument('--rows-per-tile', type=int, default=0, help='0: fill the device wave slots')
    parser.add_argument('--lane-columns', type=int, default=0, choices=(0, 2, 4), help='0: library default')
    parser.add_argument('--cpu-steps', type=int, default=8, help='steps of the CPU baseline sample (0: skip)')
    parser.add_argument('--kernel-steps', type=int, default=200,
                        help='steps timed per kernel with HIP events for the roofline (0: skip)')
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
    sim = build_ez_workload(graph_steps=args.graph_steps, rows_per_tile=args.rows_per_tile,
                            lane_columns=args.lane_columns)
    sim.finalize()
    engine = sim.engine
    rows_per_tile, lane_columns = engine.launch_geometry()
    cells = sim.model.nx * sim.model.ny
    print(f'# workload ready: {sim.model.nx}x{sim.model.ny}, {len(sim.model.materials.materials)} materials',
          file=sys.stderr, flush=True)
    engine.step(args.warmup)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    engine.step(args.steps)
    stop.record()
    torch.cuda.synchronize()
    elapsed_s = start.elapsed_time(stop) / 1e3
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
    if kernels is not None:
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

    result = {
        'metric': 'Yee cell-updates/s (HBM BW fraction in roofline)',
        'value': value,
        'unit': 'cell-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': per_step * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded rectangles, Gaussian line source)',
        'config': {'workload': workload, 'grid': [sim.model.nx, sim.model.ny], 'cpml': 20, 'dx_m': 20e-9,
                   'rectangles': 400, 'seed': SEED, 'polarization': 'ez', 'graph_steps': args.graph_steps,
                   'rows_per_tile': rows_per_tile, 'lane_columns': lane_columns},
        'roofline': roofline,
        'cpu_baseline': baseline,
    }
    print(json.dumps(result))


if __name__ == '__main__':
    main()
