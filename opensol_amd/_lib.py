"""ctypes binding of libyee.so (include/yee.h). Importing the product never falls back to a CPU path: a missing or
mismatched library raises here."""
import ctypes
import os

# OPENSOL_AMD_LIBRARY names another build of the same library (e.g. one compiled with other launch options)
LIBRARY_PATH = os.environ.get('OPENSOL_AMD_LIBRARY') or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                     'libyee.so')
ABI_VERSION = 19
MAX_SOURCES = 8
SOURCE_E, SOURCE_H = 0, 1   # yee_line_source.field
WALL_ZERO = 2   # yee2d_desc.wall_sign of an H_z pmc face
MAX_MONITORS = 4
MAX_DFT = 4
MAX_POLES = 8
MAX_TFSF = 2
MAX_POLES_3D = 12
YEE_DRUDE, YEE_LORENTZ = 1, 2
MAX_SOURCES_3D = 4
MAX_PLANES_3D = 8
STATIC_BATCH_MAX = 256   # instances yee2d_batch_step_static takes (its static device array)
YEE_EZ, YEE_HZ = 0, 1
MIRROR_SYMMETRIC, MIRROR_ANTISYMMETRIC = 1, 2
INT64_MAX = (1 << 63) - 1

c_i32, c_i64, c_f32, c_ptr = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_void_p


class LineSourceC(ctypes.Structure):
    _fields_ = [('axis', c_i32), ('line', c_i32), ('lo', c_i32), ('hi', c_i32), ('field', c_i32), ('pad_', c_i32),
                ('profile', c_ptr), ('wave', c_ptr), ('wave_len', c_i64)]


class FluxLineC(ctypes.Structure):
    _fields_ = [('normal', c_i32), ('position', c_i32), ('lo', c_i32), ('hi', c_i32), ('n_freq', c_i32),
                ('pad_', c_i32), ('phase', c_ptr), ('n_samples', c_i64), ('acc', c_ptr)]


class CpmlAxisC(ctypes.Structure):
    _fields_ = [('n_rows', c_i32), ('pad_', c_i32), ('row_e', c_ptr), ('row_h', c_ptr),
                ('b_e', c_ptr), ('g_e', c_ptr), ('b_h', c_ptr), ('g_h', c_ptr),
                ('psi_e', c_ptr * 2), ('psi_h', c_ptr * 2)]


class ProbesC(ctypes.Structure):
    _fields_ = [('n_probes', c_i32), ('pad_', c_i32), ('cell', c_ptr), ('component', c_ptr), ('samples', c_ptr),
                ('capacity', c_i64)]


class DftRegionC(ctypes.Structure):
    _fields_ = [('n_comp', c_i32), ('n_freq', c_i32), ('start', c_i32 * 3), ('count', c_i32 * 3), ('stride', c_i32 * 3),
                ('comp', c_i32 * 3), ('pair_axis', c_i32 * 3), ('pair_wrap', c_i32 * 3), ('phase', c_ptr),
                ('n_samples', c_i64), ('acc', c_ptr)]


class PoleC(ctypes.Structure):
    _fields_ = [('kind', c_i32), ('component', c_i32), ('i0', c_i32), ('i1', c_i32), ('j0', c_i32), ('j1', c_i32),
                ('damping', c_f32), ('restoring', c_f32), ('inv_dt', c_f32), ('pad_', c_i32), ('drive', c_ptr),
                ('state', c_ptr * 2)]


class TfsfC(ctypes.Structure):
    _fields_ = [('x0', c_i32), ('x1', c_i32), ('y0', c_i32), ('y1', c_i32), ('coef_h', c_f32 * 4), ('coef_e', c_f32 * 4),
                ('line', c_ptr), ('ends', c_ptr), ('n_steps', c_i64)]


class Desc2DC(ctypes.Structure):
    _fields_ = [('nx', c_i32), ('ny', c_i32), ('pitch', c_i32), ('polarization', c_i32), ('periodic_x', c_i32),
                ('periodic_y', c_i32), ('rows_per_tile', c_i32), ('lane_columns', c_i32), ('field', (c_ptr * 3) * 2),
                ('material', c_ptr), ('table', c_ptr), ('table2', c_ptr), ('n_materials', c_i32), ('dt', c_f32), ('dt_over_eps0', c_f32), ('pad_', c_i32),
                ('inv_dx_e', c_ptr), ('inv_dx_h', c_ptr), ('inv_dy_e', c_ptr), ('inv_dy_h', c_ptr),
                ('inv_dx_bulk', c_f32), ('inv_dy_bulk', c_f32), ('layer_x_lo', c_i32), ('layer_x_hi', c_i32),
                ('layer_y_lo', c_i32), ('layer_y_hi', c_i32), ('cpml', CpmlAxisC * 2), ('n_sources', c_i32), ('n_monitors', c_i32), ('sample_interval', c_i32),
                ('pad2_', c_i32), ('source', LineSourceC * MAX_SOURCES), ('monitor', FluxLineC * MAX_MONITORS),
                ('n_dft', c_i32), ('pad3_', c_i32), ('dft', DftRegionC * MAX_DFT),
                ('n_poles', c_i32), ('pad4_', c_i32), ('pole', PoleC * MAX_POLES),
                ('probes', ProbesC), ('wall_sign', c_i32 * 4), ('n_tfsf', c_i32), ('pad5_', c_i32),
                ('tfsf', TfsfC * MAX_TFSF), ('clock', c_ptr), ('step_limit', c_i64)]


class PointSourceC(ctypes.Structure):
    _fields_ = [('component', c_i32), ('i', c_i32), ('j', c_i32), ('k', c_i32), ('weight', c_f32), ('pad_', c_i32),
                ('wave', c_ptr), ('wave_len', c_i64)]


class FluxPlaneC(ctypes.Structure):
    _fields_ = [('normal', c_i32), ('edge', c_i32), ('e_comp', c_i32 * 2), ('h_comp', c_i32 * 2),
                ('lo', (c_i32 * 2) * 2), ('hi', (c_i32 * 2) * 2), ('n_freq', c_i32), ('pad_', c_i32),
                ('phase', c_ptr), ('n_samples', c_i64), ('acc', c_ptr * 2)]


class Cpml3DAxisC(ctypes.Structure):
    _fields_ = [('n_rows', c_i32), ('pad_', c_i32), ('psi_e', c_ptr * 2), ('psi_h', (c_ptr * 2) * 2)]


class AxisRow3DC(ctypes.Structure):
    _fields_ = [('inv_h', c_f32), ('inv_e', c_f32), ('psi_row_h', c_i32), ('b_h', c_f32), ('g_h', c_f32),
                ('psi_row_e', c_i32), ('b_e', c_f32), ('g_e', c_f32)]


class Pole3DC(ctypes.Structure):
    _fields_ = [('kind', c_i32), ('component', c_i32), ('i0', c_i32), ('i1', c_i32), ('j0', c_i32), ('j1', c_i32),
                ('k0', c_i32), ('k1', c_i32), ('damping', c_f32), ('restoring', c_f32), ('inv_dt', c_f32),
                ('pad_', c_i32), ('drive', c_ptr), ('state', c_ptr * 2)]


class Desc3DC(ctypes.Structure):
    _fields_ = [('nx', c_i32), ('ny', c_i32), ('nz', c_i32), ('periodic', c_i32 * 3), ('rows_per_tile', c_i32),
                ('planes_per_tile', c_i32), ('field', (c_ptr * 6) * 2), ('material', c_ptr), ('table', c_ptr),
                ('n_materials', c_i32), ('dt', c_f32), ('axis_rows', c_ptr * 3), ('cpml', Cpml3DAxisC * 3), ('n_sources', c_i32), ('n_planes', c_i32), ('sample_interval', c_i32),
                ('pad_', c_i32), ('source', PointSourceC * MAX_SOURCES_3D), ('plane', FluxPlaneC * MAX_PLANES_3D),
                ('n_dft', c_i32), ('pad2_', c_i32), ('dft', DftRegionC * MAX_DFT),
                ('n_poles', c_i32), ('pad3_', c_i32), ('pole', Pole3DC * MAX_POLES_3D),
                ('probes', ProbesC), ('mirror', c_i32 * 6), ('clock', c_ptr), ('step_limit', c_i64)]


STRUCTS = (Desc2DC, LineSourceC, FluxLineC, CpmlAxisC, Desc3DC, PointSourceC, FluxPlaneC, Cpml3DAxisC, AxisRow3DC,
           ProbesC, DftRegionC, PoleC, Pole3DC, TfsfC)

# Every entry point include/yee.h declares: (name, restype, argtypes)
SIGNATURES = (
    ('yee_abi_version', c_i32, ()),
    ('yee_sizeof', ctypes.c_size_t, (c_i32,)),
    ('yee_last_error', ctypes.c_char_p, ()),
    ('yee2d_create', c_i32, (ctypes.POINTER(Desc2DC), ctypes.POINTER(c_ptr))),
    ('yee_step', c_i32, (c_ptr, c_i64, c_ptr)),
    ('yee_step_timed', c_i32, (c_ptr, c_i64, c_ptr, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_i64))),
    ('yee_set_graph_steps', c_i32, (c_ptr, c_i32)),
    ('yee_build_graph', c_i32, (c_ptr,)),
    ('yee2d_update', c_i32, (c_ptr, ctypes.POINTER(Desc2DC))),
    ('yee2d_get_desc', c_i32, (c_ptr, ctypes.POINTER(Desc2DC))),
    ('yee3d_create', c_i32, (ctypes.POINTER(Desc3DC), ctypes.POINTER(c_ptr))),
    ('yee3d_update', c_i32, (c_ptr, ctypes.POINTER(Desc3DC))),
    ('yee3d_get_desc', c_i32, (c_ptr, ctypes.POINTER(Desc3DC))),
    ('yee_destroy', c_i32, (c_ptr,)),
    ('yee2d_batch_bytes', ctypes.c_size_t, (c_i32,)),
    ('yee2d_batch_create', c_i32, (ctypes.POINTER(Desc2DC), c_i32, c_ptr, c_ptr, ctypes.POINTER(c_ptr))),
    ('yee3d_batch_bytes', ctypes.c_size_t, (c_i32,)),
    ('yee3d_batch_create', c_i32, (ctypes.POINTER(Desc3DC), c_i32, c_ptr, c_ptr, ctypes.POINTER(c_ptr))),
    ('yee2d_batch_step_static', c_i32, (ctypes.POINTER(Desc2DC), c_i32, c_i64, c_ptr)),
    ('yee_flag_nonfinite', c_i32, (c_ptr, c_i64, c_ptr, c_ptr)),
)

_LIB = None


class YeeError(RuntimeError):
    """A libyee.so call returned a non-zero status."""


def load():
    """The loaded library (once per process); RuntimeError when it is missing or its ABI differs."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBRARY_PATH):
        raise RuntimeError(f'{LIBRARY_PATH} is missing: build the HIP library first (python -m opensol_amd.build); '
                           f'there is no CPU fallback')
    lib = ctypes.CDLL(LIBRARY_PATH)
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = list(argtypes)
    if lib.yee_abi_version() != ABI_VERSION:
        raise RuntimeError(f'libyee.so ABI {lib.yee_abi_version()} != binding ABI {ABI_VERSION}')
    for which, struct in enumerate(STRUCTS):
        if lib.yee_sizeof(which) != ctypes.sizeof(struct):
            raise RuntimeError(f'libyee.so struct {struct.__name__} is {lib.yee_sizeof(which)} bytes, the binding '
                               f'{ctypes.sizeof(struct)}: rebuild the library')
    _LIB = lib
    return lib


def check(status, what):
    """Raise YeeError with the library's message for a non-zero status."""
    if status != 0:
        message = load().yee_last_error()
        raise YeeError(f'{what} failed ({status}): {message.decode() if message else ""}')
