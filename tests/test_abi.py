"""The C ABI library builds, loads on a machine without a GPU, and exports every entry point include/yee.h declares
with the struct layouts the ctypes binding assumes (no compute calls here)."""
import ctypes
import os
import re

import pytest

from opensol_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'yee.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:const\s+)?\w+\s*\**\s+\**\s*(yee\w+)\s*\(', text, flags=re.M)))


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(_lib.LIBRARY_PATH):
        from opensol_amd.build import build
        build()
    return _lib.load()


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    for name in ('yee_abi_version', 'yee_sizeof', 'yee_last_error', 'yee2d_create', 'yee_step', 'yee_destroy',
                 'yee2d_batch_create', 'yee2d_batch_bytes', 'yee2d_update', 'yee_set_graph_steps', 'yee3d_create',
                 'yee3d_update', 'yee3d_get_desc', 'yee_step_timed'):
        assert name in names


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), f'libyee.so does not export {name}'
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound, set(declared_functions()) - bound


def test_struct_layouts_match_binding(lib):
    assert lib.yee_abi_version() == _lib.ABI_VERSION
    for which, struct in enumerate(_lib.STRUCTS):
        assert lib.yee_sizeof(which) == ctypes.sizeof(struct)


def test_invalid_descriptors_are_rejected_without_a_gpu(lib):
    ctx = _lib.c_ptr()
    d = _lib.Desc2DC()
    status = lib.yee2d_create(ctypes.byref(d), ctypes.byref(ctx))
    assert status == -1 and not ctx
    assert b'2D grid' in lib.yee_last_error()
    d.nx, d.ny, d.pitch, d.rows_per_tile, d.lane_columns = 16, 16, 18, 4, 4
    assert lib.yee2d_create(ctypes.byref(d), ctypes.byref(ctx)) == -1
    assert b'pitch' in lib.yee_last_error()
    assert lib.yee_step(None, 1, None) == -3
    assert lib.yee_destroy(None) == 0
    d3 = _lib.Desc3DC()
    assert lib.yee3d_create(ctypes.byref(d3), ctypes.byref(ctx)) == -1 and not ctx
    assert b'3D grid' in lib.yee_last_error()
    d3.nx, d3.ny, d3.nz, d3.rows_per_tile, d3.planes_per_tile = 8, 8, 8, 3, 4
    assert lib.yee3d_create(ctypes.byref(d3), ctypes.byref(ctx)) == -1
    assert b'rows_per_tile' in lib.yee_last_error()
