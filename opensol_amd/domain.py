"""Uniform tensor grids with CPML padding: cell geometry, time step and the coordinate -> cell rules of the
reference (core/grid.py:9-60, core/grid3d.py:8-16, core/domain.py:14-282, core/domain3d.py:12-114).

A Domain's physical origin is the lower corner of its window (the cells inside the CPML padding). Widths, edges and
centres are float64; the kernel cell width `grid.dx` and the time step `grid.dt` are float32 values, as the reference
keeps them.
"""
import math

import numpy as np

from .constants import C0, DTYPE, FACES, FACES_3D

DEFAULT_COURANT_FACTOR = 0.99


def courant_time_step(spacings, courant_factor=None):
    """float32 dt = courant / (C0 sqrt(sum 1/d^2)) over the finest spacing of each axis (core/grid.py:9-17,
    core/grid3d.py:8-16); courant is a fraction of the CFL limit in (0, 1]."""
    courant_factor = DEFAULT_COURANT_FACTOR if courant_factor is None else courant_factor
    if not 0.0 < float(courant_factor) <= 1.0:
        raise ValueError(f"courant must be a fraction of the CFL limit in (0, 1], got {courant_factor} (it is not "
                         f"C0 dt / dx: 0.99 means dt = 0.99 dx / (C0 sqrt(dims)) on cubic cells)")
    inverse_spacing_squared = sum(1.0 / float(d) ** 2 for d in spacings)
    return DTYPE(float(courant_factor) / (C0 * np.sqrt(inverse_spacing_squared)))


def uniform_cell_count(length: float, dx: float) -> int:
    """Cells of width dx covering length; the -1e-9 removes float round-up (core/grid.py:20-22)."""
    return int(np.ceil(length / dx - 1e-9))


def containing_cell(coord, edges, first, last):
    """Cell i with edges[i] <= coord < edges[i+1] after a 1e-9-cell snap, clamped to [first, last]
    (core/domain.py:275-282)."""
    coord = float(coord)
    if not math.isfinite(coord):
        raise ValueError(f'non-finite coordinate {coord}')
    snap = 1e-9 * (edges[first + 1] - edges[first])
    index = int(np.searchsorted(edges, coord + snap, side='right')) - 1
    return min(max(index, first), last)


class UniformGrid:
    """Cell counts, float32 dx and dt of a uniform grid of `dims` axes (Grid2D / Grid3D)."""

    def __init__(self, totals, dx, courant=None):
        if not dx > 0:
            raise ValueError(f'grid needs dx > 0, got {dx}')
        self.dx = DTYPE(dx)
        self.shape = tuple(uniform_cell_count(total, dx) for total in totals)
        if min(self.shape) < 1:
            raise ValueError(f'grid has no cells: size {totals} m, dx {dx} m')
        # the 2D step takes the float64 width (core/grid.py:34-60); the 3D one the finest float32 cell width
        # (core/grid3d.py:44-56)
        spacing = float(dx) if len(totals) == 2 else float(self.dx)
        self.dt = courant_time_step([spacing] * len(totals), courant)
        self.courant = DTYPE(DEFAULT_COURANT_FACTOR if courant is None else courant)

    @property
    def nx(self):
        return self.shape[0]

    @property
    def ny(self):
        return self.shape[1]

    @property
    def nz(self):
        return self.shape[2]

    @property
    def dy(self):
        return self.dx


class Domain:
    """Uniform window + per-face CPML padding (core/domain.py:180-196). `sides` = padding cells per face."""

    def __init__(self, sizes, dx, sides, courant=None):
        self.dimensions = len(sizes)
        self.axes = 'xyz'[:self.dimensions]
        faces = FACES if self.dimensions == 2 else FACES_3D
        self.npml_sides = {face: int(sides.get(face, 0)) for face in faces}
        self.sizes = tuple(float(s) for s in sizes)
        self.dx = float(dx)
        totals = [size + (self.npml_sides[f'{a}_lo'] + self.npml_sides[f'{a}_hi']) * dx
                  for size, a in zip(self.sizes, self.axes)]
        self.grid = UniformGrid(totals, dx, courant)
        self.shape = self.grid.shape
        self.origins = tuple(self.npml_sides[f'{a}_lo'] for a in self.axes)
        for n, a in zip(self.shape, self.axes):
            if n - self.npml_sides[f'{a}_lo'] - self.npml_sides[f'{a}_hi'] < 1:
                raise ValueError(f'domain has no physical cells along {a}: grid {self.shape}, PML {self.npml_sides}')
        self._widths = tuple(_read_only(np.full(n, self.dx)) for n in self.shape)
        self._edges = tuple(_read_only((np.arange(n + 1) - o) * self.dx) for n, o in zip(self.shape, self.origins))
        self._centres = tuple(_read_only(0.5 * (e[:-1] + e[1:])) for e in self._edges)
        self._centres32 = tuple(c.astype(DTYPE) for c in self._centres)

    # reference vocabulary ------------------------------------------------------------------------------------------
    @property
    def size_x(self):
        return self.sizes[0]

    @property
    def size_y(self):
        return self.sizes[1]

    @property
    def size_z(self):
        return self.sizes[2]

    def physical_window(self):
        """One slice per axis of the cells inside the physical domain."""
        return tuple(slice(self.npml_sides[f'{a}_lo'], n - self.npml_sides[f'{a}_hi'])
                     for a, n in zip(self.axes, self.shape))

    def cell_widths(self):
        return self._widths

    def cell_edges(self):
        return self._edges

    def cell_centers(self):
        return self._centres

    def physical_extent(self):
        extent = []
        for edges, window in zip(self._edges, self.physical_window()):
            extent += [0.0, float(edges[window.stop])]
        return tuple(extent)

    def validate_physical_pos(self, position, label=''):
        """ValueError unless the point lies in the physical window, edges within 1e-9 of the largest side
        (core/domain.py:126-133)."""
        tol = 1e-9 * max(self.sizes)
        if not all(-tol <= float(v) <= s + tol for v, s in zip(position, self.sizes)):
            coords = ', '.join(f'{float(v) * 1e6:.10g}' for v in position)
            extents = ' x '.join(f'{s * 1e6:.10g}' for s in self.sizes)
            raise ValueError(f'{label} position ({coords}) um outside physical domain ({extents}) um')

    def pos_to_cell_axis(self, axis_index, coord):
        window = self.physical_window()[axis_index]
        return containing_cell(coord, self._edges[axis_index], window.start, window.stop - 1)

    def pos_to_cell(self, *coords):
        return tuple(self.pos_to_cell_axis(a, c) for a, c in enumerate(coords))

    def host_centres(self, clamp_to_physical=True):
        """Per-axis float32 centres; clamped: PML cells take the nearest window centre (core/domain.py:90-98)."""
        if not clamp_to_physical:
            return self._centres32
        out = []
        for centres, window in zip(self._centres, self.physical_window()):
            out.append(np.clip(centres, centres[window.start], centres[window.stop - 1]).astype(DTYPE))
        return tuple(out)

    def rect_mask(self, x0, y0, x1, y1):
        """Bool host mask of the cells whose float32 centre lies in [x_min, x_max) x [y_min, y_max), corners in
        any order (core/domain.py:255-260)."""
        x, y = self._centres32[:2]
        inside = [(c >= DTYPE(min(a, b))) & (c < DTYPE(max(a, b))) for c, a, b in ((x, x0, x1), (y, y0, y1))]
        return inside[0][:, None] & inside[1][None, :]

    def circle_mask(self, cx, cy, radius):
        """Bool host mask of the cells whose float32 centre lies within `radius` of (cx, cy), boundary included
        (core/domain.py:262-267)."""
        if not radius >= 0:
            raise ValueError(f'circle radius must be >= 0, got {radius}')
        x, y = self._centres32[:2]
        return (x[:, None] - DTYPE(cx)) ** 2 + (y[None, :] - DTYPE(cy)) ** 2 <= DTYPE(radius) ** 2

    def box_mask(self, lo, hi):
        """Bool host mask of the cells of box_ranges(lo, hi) on a 3D grid (the 3D counterpart of rect_mask)."""
        mask = np.zeros(self.shape, dtype=bool)
        mask[tuple(slice(a, b) for a, b in self.box_ranges(lo, hi))] = True
        return mask

    def box_ranges(self, lows, highs, clamp_to_physical=False):
        """Per-axis half-open index ranges of the cells whose float32 centre lies in [min, max) (the reference's
        rect_mask / box mask compares float32 centres, core/domain.py:90-97, 255-260; the 3D model builder's box raster
        compares the centres clamped to the window, model/geometry.py:1762-1788): a box is a product of ranges, so it
        is painted in O(area)."""
        ranges = []
        for centres, lo, hi in zip(self.host_centres(clamp_to_physical), lows, highs):
            a, b = DTYPE(min(lo, hi)), DTYPE(max(lo, hi))
            ranges.append((int(np.searchsorted(centres, a, side='left')),
                           int(np.searchsorted(centres, b, side='left'))))
        return tuple(ranges)


def _read_only(array):
    array.flags.writeable = False
    return array
