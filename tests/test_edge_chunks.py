"""Row chunks of the 2D edge tiles, checked on the CPU with the library's own functions.

opensol_amd/csrc/yee_launch.h holds yee_edge_parts (chunks of a long edge tile), yee_edge_split (more chunks when an
instance has few edge tiles) and edge_chunk (the rows of one chunk) as constexpr functions, which the host launch
code and the edge kernel use and g++ compiles here. For every tile length up to the tile's rows (the last tile of a
grid is ragged) the non-empty chunks must cover the tile's rows exactly once; the split must stay between the
long-tile rule and one row per chunk, and reach the wavefront target unless the chunks are single rows."""
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
#include "yee_launch.h"

static int check(int rows, long edge_tiles)
{
    const int parts = yee_edge_split(rows, edge_tiles);
    if (parts < yee_edge_parts(rows) || parts > rows) return 1;
    if (parts < rows && edge_tiles * parts < kEdgeWaveTarget) return 2;
    for (int length = 1; length <= rows; ++length) {   // ragged last tiles included
        const int i0 = 40;
        std::vector<int> seen(length, 0);
        for (int part = 0; part < parts; ++part) {
            const RowSpan s = edge_chunk(i0, i0 + length, part, parts);
            for (int r = s.r0; r < s.r1; ++r) {
                if (r < i0 || r >= i0 + length) return 3;
                ++seen[r - i0];
            }
        }
        for (int v : seen)
            if (v != 1) return 4;
    }
    return 0;
}

int main(int argc, char** argv)
{
    int failures = 0, cases = 0;
    for (int a = 1; a + 1 < argc; a += 2) {
        const int rows = atoi(argv[a]);
        const long edge = atol(argv[a + 1]);
        const int code = check(rows, edge);
        ++cases;
        if (code) {
            ++failures;
            printf("FAIL rows %d edge tiles %ld: %d\n", rows, edge, code);
        }
    }
    if (argc == 1) {   // the split of named cases
        printf("%d %d %d %d %d %d\n", yee_edge_split(8, 2544), yee_edge_split(8, 16), yee_edge_split(64, 1),
               yee_edge_split(64, 100000), yee_edge_split(16, 100000), yee_edge_split(300, 1));
        return 0;
    }
    printf("checked %d cases, %d failures\n", cases, failures);
    return failures ? 1 : 0;
}
'''

ROWS = list(range(1, 71)) + [96, 100, 128, 255, 300, 512]
EDGE_TILES = [1, 2, 3, 16, 100, 255, 256, 257, 1000, 2047, 2048, 2544, 5000, 100000]


@pytest.fixture(scope='module')
def checker(tmp_path_factory):
    compiler = shutil.which('g++') or shutil.which('c++')
    if compiler is None:
        pytest.fail('a host C++ compiler (g++) is needed to compile the chunk check')
    work = tmp_path_factory.mktemp('edge_chunks')
    source = work / 'check.cpp'
    source.write_text(CHECKER)
    binary = work / 'check'
    subprocess.run([compiler, '-std=c++17', '-O1', f'-I{CSRC}', str(source), '-o', str(binary)], check=True)
    return str(binary)


def test_chunks_cover_every_edge_tile_once(checker):
    args = [str(v) for rows in ROWS for edge in EDGE_TILES for v in (rows, edge)]
    result = subprocess.run([checker, *args], capture_output=True, text=True)
    assert result.returncode == 0, result.stdout
    assert f'checked {len(ROWS) * len(EDGE_TILES)} cases, 0 failures' in result.stdout


def test_split_of_named_cases(checker):
    """8192^2 E_z at 8 rows per tile has 2544 edge tiles: one chunk each. A grid of 16 edge tiles gets single-row
    chunks; 64-row tiles keep the long-tile rule's 8 chunks when edge tiles are plentiful."""
    result = subprocess.run([checker], capture_output=True, text=True, check=True)
    assert result.stdout.split() == ['1', '8', '64', '8', '2', '300']


def test_a_floor_chunk_length_would_be_caught(tmp_path):
    """The check detects a chunk length rounded down (a modified copy of the header): rows of ragged tiles would be
    stepped by no chunk."""
    header = open(os.path.join(CSRC, 'yee_launch.h')).read()
    exact = 'const int chunk = (i1 - i0 + parts - 1) / parts;'
    assert exact in header
    (tmp_path / 'yee_launch.h').write_text(header.replace(exact, 'const int chunk = (i1 - i0) / parts;'))
    source = tmp_path / 'check.cpp'
    source.write_text(CHECKER)
    binary = tmp_path / 'check'
    compiler = shutil.which('g++') or shutil.which('c++')
    subprocess.run([compiler, '-std=c++17', '-O1', f'-I{tmp_path}', str(source), '-o', str(binary)], check=True)
    result = subprocess.run([str(binary), '8', '16', '16', '100'], capture_output=True, text=True)
    assert result.returncode == 1 and 'FAIL rows 8' in result.stdout
