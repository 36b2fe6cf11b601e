"""Profiling helpers for rocprofv3 runs.

    python -m opensol_amd.profiling copy --mib 1024 --repeat 5
        device-to-device copies of a known size: the calibration run that scales rocprofv3 FETCH_SIZE / WRITE_SIZE
        (KiB per dispatch) to HBM bytes for the Yee kernels.
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
    summary = sub.add_parser('summarize')
    summary.add_argument('path')
    summary.add_argument('kernel', nargs='?')
    args = parser.parse_args(argv)
    if args.command == 'copy':
        copy_probe(args.mib, args.repeat)
    else:
        summarize(args.path, args.kernel)


if __name__ == '__main__':
    sys.exit(main())
