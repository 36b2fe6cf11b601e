"""Profiling helpers for rocprofv3 runs.

    python -m opensol_amd.profiling copy --mib 1024 --repeat 5
        device-to-device copies of a known size: the calibration run that scales rocprofv3 FETCH_SIZE / WRITE_SIZE
        (KiB per dispatch) to HBM bytes for the Yee kernels.
    python -m opensol_amd.profiling stream --mib 4096 --repeat 20
        HBM streaming rates of the runtime's copy, add, fill and sum kernels: the practical ceiling the Yee kernels
        are read against.
    python -m opensol_amd.profiling summarize <counter_collection.csv> [kernel substring]
        per-kernel mean of each counter over the dispatches in a rocprofv3 --pmc output file.
"""
import argparse
import csv
import sys
from collections import defaultdict


def copy_probe(mib, repeat):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('copy probe needs a GPU')
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(repeat):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f'copied {mib} MiB x {repeat}: expected FETCH = WRITE = {mib * 1024} KiB per copy dispatch')


def stream_rates(mib, repeat):
    """HBM streaming rates of the runtime's own kernels on buffers of `mib` MiB, timed with HIP events: device copy
    (read 1 : write 1, the Yee kernels' mix is 14 : 12 in 2D and 26 : 24 in 3D), elementwise scale (1 : 1), add
    (2 : 1), fill (0 : 1) and sum (1 : 0). Rates count bytes read plus bytes written."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('stream rates need a GPU')
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device='cuda')
    b = torch.full_like(a, 2.0)
    c = torch.empty_like(a)
    size = n * 4
    cases = {'copy (1:1)': (lambda: c.copy_(a), 2 * size), 'scale (1:1)': (lambda: torch.mul(a, 2.0, out=c), 2 * size),
             'add (2:1)': (lambda: torch.add(a, b, out=c), 3 * size),
             'fill (0:1)': (lambda: c.fill_(3.0), size), 'sum (1:0)': (lambda: a.sum(), size)}
    for name, (op, moved) in cases.items():
        for _ in range(3):
            op()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeat):
            op()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / repeat
        print(f'{name:>12s}: {ms * 1e3:9.1f} us per pass over {mib} MiB, {moved / (ms * 1e-3) / 1e12:.3f} TB/s '
              f'({moved / (ms * 1e-3) / 8e12 * 100:.1f} % of 8 TB/s)', flush=True)


def summarize(path, kernel_filter=None):
    sums = defaultdict(float)
    counts = defaultdict(int)
    with open(path, newline='') as f:
        for row in csv.DictReader(f):
            name = row.get('Kernel_Name') or row.get('KernelName') or ''
            if kernel_filter and kernel_filter not in name:
                continue
            counter = row.get('Counter_Name') or row.get('CounterName')
            value = row.get('Counter_Value') or row.get('CounterValue')
            if counter is None or value is None:
                continue
            key = (name[:80], counter)
            sums[key] += float(value)
            counts[key] += 1
    for (name, counter), total in sorted(sums.items()):
        print(f'{counter:>12s} mean {total / counts[(name, counter)]:.6g} over {counts[(name, counter)]} '
              f'dispatches  {name}')


def main(argv=None):
    parser = argparse.ArgumentParser()
    sub = parser.add_subparsers(dest='command', required=True)
    copy = sub.add_parser('copy')
    copy.add_argument('--mib', type=int, default=1024)
    copy.add_argument('--repeat', type=int, default=5)
    rates = sub.add_parser('stream')
    rates.add_argument('--mib', type=int, default=4096)
    rates.add_argument('--repeat', type=int, default=20)
    summary = sub.add_parser('summarize')
    summary.add_argument('path')
    summary.add_argument('kernel', nargs='?')
    args = parser.parse_args(argv)
    if args.command == 'copy':
        copy_probe(args.mib, args.repeat)
    elif args.command == 'stream':
        stream_rates(args.mib, args.repeat)
    else:
        summarize(args.path, args.kernel)


if __name__ == '__main__':
    sys.exit(main())
