"""Index bounds of the 3D wavefront tiles, checked on the CPU with the kernels' own index functions.

opensol_amd/csrc/yee3d_index.h holds the tile decode, the lane and row coordinates, the owner rule and fold() as
constexpr functions, which the HIP kernels use and g++ compiles here. The checker walks every wave slot, lane,
register row and x plane that a launch touches (the prologue plane i0 - 1, the planes of the chunk and the lookahead
plane i + 1), for shapes that stress the wraps: short periodic axes (a 64-lane tile over nz = 5 spans 13 periods),
one-cell axes, extents next to multiples of the 62-cell z tile and the 4-row y tile, and chunk lengths from 1 to a
whole axis. Every folded index must lie inside its axis, and every cell must be owned by exactly one lane of one
tile. (A fold() that wrapped only once read axis_rows[2][57] of a 5-entry table; this check fails on it.)"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'opensol_amd', 'csrc')

CHECKER = r'''
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "yee3d_index.h"
using namespace yee3idx;

static int check(int nx, int ny, int nz, bool px, bool py, bool pz, int P)
{
    const int tiles_j = tiles_along(ny, kRows), tiles_k = tiles_along(nz, kOwned), chunks = tiles_along(nx, P);
    const long tiles = (long)chunks * tiles_j * tiles_k;
    std::vector<int> owned((size_t)nx * ny * nz, 0);
    auto in = [](int v, int n) { return v >= 0 && v < n; };
    for (long slot = 0; slot < tiles; ++slot) {
        const TileCoords c = tile_coords((int)slot, tiles_j, tiles_k);
        if (!in(c.ic, chunks) || !in(c.tj, tiles_j) || !in(c.tk, tiles_k)) return 1;
        const int i0 = c.ic * P, i1 = i0 + P < nx ? i0 + P : nx;
        if (i0 > 0 || px) {
            const int im = i0 > 0 ? i0 - 1 : nx - 1;
            if (!in(im, nx)) return 2;
        }
        for (int i = i0; i < i1; ++i)
            if (!in(fold(i + 1, nx, px), nx)) return 3;
        for (int lane = 0; lane < kLanes; ++lane) {
            const int kz = lane_z(c.tk, lane);
            const int k = fold(kz, nz, pz);
            if (!in(k, nz)) return 4;
            const bool kown = lane_owns(lane, kz, nz);
            if (kown && kz != k) return 5;
            for (int r = 0; r < kRows + 2; ++r) {
                const int jy = row_y(c.tj, r);
                const int j = fold(jy, ny, py);
                if (!in(j, ny)) return 6;
                const bool rown = r >= 1 && r <= kRows && jy < ny;
                if (rown && jy != j) return 7;
                if (kown && rown)
                    for (int i = i0; i < i1; ++i) ++owned[((size_t)i * ny + j) * nz + k];
            }
        }
    }
    for (int v : owned)
        if (v != 1) return 8;
    return 0;
}

int main(int argc, char** argv)
{
    int failures = 0, cases = 0;
    for (int a = 1; a + 6 < argc + 1; a += 7) {
        const int nx = atoi(argv[a]), ny = atoi(argv[a + 1]), nz = atoi(argv[a + 2]);
        const bool px = atoi(argv[a + 3]), py = atoi(argv[a + 4]), pz = atoi(argv[a + 5]);
        const int P = atoi(argv[a + 6]);
        const int code = check(nx, ny, nz, px, py, pz, P);
        ++cases;
        if (code) {
            ++failures;
            printf("FAIL %d %d %d %d %d %d %d: %d\n", nx, ny, nz, px, py, pz, P, code);
        }
    }
    printf("checked %d cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
'''

SHAPES = [  # (nx, ny, nz, periodic x, y, z, planes per tile)
    (7, 6, 70, 0, 1, 0, 4), (12, 10, 40, 1, 0, 0, 4), (9, 22, 12, 0, 1, 0, 3), (16, 16, 70, 0, 0, 0, 4),
    (30, 8, 11, 1, 0, 0, 5), (12, 12, 12, 0, 0, 0, 4), (6, 7, 5, 0, 0, 1, 4), (5, 6, 7, 1, 1, 1, 2),
    (3, 5, 5, 0, 0, 1, 1), (1, 1, 1, 0, 0, 0, 4), (1, 1, 1, 1, 1, 1, 4), (2, 1, 64, 0, 0, 0, 1),
    (4, 3, 62, 0, 0, 0, 4), (4, 5, 63, 0, 0, 1, 4), (4, 4, 124, 0, 0, 0, 3), (4, 4, 125, 0, 0, 1, 4),
    (8, 9, 186, 0, 1, 0, 8), (5, 308, 20, 0, 0, 0, 4), (21, 8, 100, 0, 0, 0, 21), (13, 13, 13, 1, 1, 1, 13),
    (40, 4, 3, 0, 1, 1, 171), (64, 64, 64, 0, 0, 0, 4), (17, 19, 130, 1, 0, 1, 6),
]


@pytest.fixture(scope='module')
def checker(tmp_path_factory):
    compiler = shutil.which('g++') or shutil.which('c++')
    if compiler is None:
        pytest.fail('a host C++ compiler (g++) is needed to compile the index check')
    work = tmp_path_factory.mktemp('index3d')
    source = work / 'check.cpp'
    source.write_text(CHECKER)
    binary = work / 'check'
    subprocess.run([compiler, '-std=c++17', '-O1', f'-I{CSRC}', str(source), '-o', str(binary)], check=True)
    return str(binary)


def test_every_index_in_bounds_and_every_cell_owned_once(checker):
    args = [str(v) for shape in SHAPES for v in shape]
    result = subprocess.run([checker, *args], capture_output=True, text=True)
    assert result.returncode == 0, result.stdout
    assert f'checked {len(SHAPES)} cases, 0 failures' in result.stdout


def test_a_single_wrap_fold_would_be_caught(checker, tmp_path):
    """The check detects the fault it guards against: with fold() wrapping only once (a modified copy of the
    header), the short periodic axis reads outside its table."""
    header = open(os.path.join(CSRC, 'yee3d_index.h')).read()
    full = '''        const int m = v % n;
        return m < 0 ? m + n : m;'''
    assert full in header
    broken = header.replace(full, '        return v < 0 ? v + n : (v >= n ? v - n : v);')
    (tmp_path / 'yee3d_index.h').write_text(broken)
    source = tmp_path / 'check.cpp'
    source.write_text(CHECKER)
    binary = tmp_path / 'check'
    subprocess.run([shutil.which('g++') or 'c++', '-std=c++17', '-O1', f'-I{tmp_path}', str(source), '-o',
                    str(binary)], check=True)
    result = subprocess.run([str(binary), '6', '7', '5', '0', '0', '1', '4'], capture_output=True, text=True)
    assert result.returncode == 1 and 'FAIL 6 7 5 0 0 1 4: 4' in result.stdout
