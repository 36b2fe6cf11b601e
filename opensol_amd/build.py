"""Build libyee.so (HIP, gfx950) in-tree: python -m opensol_amd.build"""
import os
import shutil
import subprocess
import sys
import tempfile

PACKAGE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PACKAGE)
CSRC = os.path.join(PACKAGE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
LIBRARY = os.path.join(PACKAGE, 'libyee.so')
SOURCES = ('yee2d.hip', 'yee3d.hip', 'yee_dft.hip', 'yee_api.hip')
ARCH = 'gfx950'
# Packed FP32 instructions are disabled for every HIP source; hipcc -O3 otherwise forms them from float code.
# -ffp-contract=on: FMAs only within one expression, so every copy of a cell update rounds identically.
FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-munsafe-fp-atomics', '-ffp-contract=on',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wall', '-Wno-unused-function']


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found: the HIP toolchain (ROCm) is required to build libyee.so')
    return path


def build(verbose=False):
    """Compile every HIP source (one hipcc process per source, in parallel) and link opensol_amd/libyee.so (always
    recompiles)."""
    compiler = hipcc()
    with tempfile.TemporaryDirectory(prefix='yee_build_') as work:
        jobs = []
        for name in SOURCES:
            obj = os.path.join(work, name.replace('.hip', '.o'))
            cmd = [compiler, *FLAGS, f'-I{INCLUDE}', f'-I{CSRC}', '-c', os.path.join(CSRC, name), '-o', obj]
            if verbose:
                print(' '.join(cmd))
            jobs.append((name, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        failed = []
        for name, obj, proc in jobs:
            out, err = proc.communicate()
            if proc.returncode != 0:
                failed.append(f'{name}:\n{out}\n{err}')
        if failed:
            raise RuntimeError('libyee.so build failed:\n' + '\n'.join(failed))
        tmp = LIBRARY + '.tmp'
        cmd = [compiler, '-shared', '-fPIC', f'--offload-arch={ARCH}', *(obj for _, obj, _ in jobs), '-o', tmp]
        if verbose:
            print(' '.join(cmd))
        result = subprocess.run(cmd, capture_output=True, text=True)
        if result.returncode != 0:
            raise RuntimeError(f'libyee.so link failed:\n{result.stdout}\n{result.stderr}')
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
