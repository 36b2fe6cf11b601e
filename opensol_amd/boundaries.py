"""Boundary configuration and the CPML tables the kernels read.

Face types, CPML defaults and the grading follow physics/em/boundaries.py:194-356 (BoundaryConfig) and :541-631
(profiles, row layout). The kernel form is the reference's CPMLSided (:689-798): psi = b psi + a (f_hi - f_lo)/d,
field += sign * coefficient * psi, kappa folded into the curl spacings. Here every table is computed in float64 and
rounded once to float32, and the gain a/d is stored per psi row.
"""
import math

import numpy as np

from .constants import DTYPE, EPS0, FACES, FACES_3D

CPML, PERIODIC, PEC, PMC = 'cpml', 'periodic', 'pec', 'pmc'
SYMMETRIC, ANTISYMMETRIC = 'symmetric', 'antisymmetric'
MIRRORS = (SYMMETRIC, ANTISYMMETRIC)
_KINDS = (CPML, PERIODIC, PEC, PMC, SYMMETRIC, ANTISYMMETRIC)


class BoundaryConfig:
    """Per-face types and CPML settings (npml 15, order 3, sigma_factor 1, kappa_max 15, alpha_max 0 S/m).

    Face types: 'cpml', 'periodic', 'pec' (the outermost window cell layer is a perfect electric conductor: 3D and 2D
    E_z) and 'pmc' (3D: the natural grid edge, H tangential 0 half a cell beyond the outermost centres,
    physics/em3d/engine3d.py:12-13; 2D E_z: the outermost row copies its inner neighbour), 'symmetric' / 'antisymmetric'
    (2D E_z mirrors: one ghost cell beyond the window copies the edge cell, negated for 'antisymmetric',
    physics/em/boundaries.py:449-488, 283-291). Bloch phases are not part of this FDTD path."""

    def __init__(self, x_lo='cpml', x_hi='cpml', y_lo='cpml', y_hi='cpml', z_lo=None, z_hi=None):
        faces = {'x_lo': x_lo, 'x_hi': x_hi, 'y_lo': y_lo, 'y_hi': y_hi}
        if z_lo is not None or z_hi is not None:
            faces.update(z_lo=z_lo or 'cpml', z_hi=z_hi or 'cpml')
        self.faces = {}
        for face, kind in faces.items():
            kind = getattr(kind, 'value', kind)
            if kind not in _KINDS:
                raise ValueError(f"face {face}: type {kind!r} is not supported by this FDTD path "
                                 f"(cpml, periodic, pec, pmc, symmetric, antisymmetric)")
            self.faces[face] = kind
        self.cpml_npml = 15
        self.cpml_order = 3
        self.cpml_sigma_factor = 1.0
        self.cpml_kappa_max = 15.0
        self.cpml_alpha_max = 0.0
        self.validate()

    @property
    def dimensions(self):
        return 3 if 'z_lo' in self.faces else 2

    def validate(self):
        """Paired periodic faces and CPML settings in range (physics/em/boundaries.py:232-253)."""
        for axis in 'xyz'[:self.dimensions]:
            lo, hi = self.faces[f'{axis}_lo'], self.faces[f'{axis}_hi']
            if (lo == PERIODIC or hi == PERIODIC) and lo != hi:
                raise ValueError(f'{axis}: periodic BCs must be paired. Got {lo}/{hi}')
        if self.needs_cpml:
            if int(self.cpml_npml) != self.cpml_npml or self.cpml_npml < 1:
                raise ValueError(f'cpml_npml must be an integer >= 1 with a cpml face, got {self.cpml_npml!r}')
            if int(self.cpml_order) != self.cpml_order or self.cpml_order < 1:
                raise ValueError(f'cpml_order must be an integer >= 1, got {self.cpml_order!r}')
            if not self.cpml_sigma_factor > 0.0:
                raise ValueError(f'cpml_sigma_factor must be > 0, got {self.cpml_sigma_factor!r}')
            if not self.cpml_kappa_max >= 1.0:
                raise ValueError(f'cpml_kappa_max must be >= 1, got {self.cpml_kappa_max!r}')
            if not self.cpml_alpha_max >= 0.0:
                raise ValueError(f'cpml_alpha_max must be >= 0 S/m, got {self.cpml_alpha_max!r}')

    def __getattr__(self, name):
        faces = self.__dict__.get('faces', {})
        if name in faces:
            return faces[name]
        raise AttributeError(name)

    @property
    def needs_cpml(self):
        return any(kind == CPML for kind in self.faces.values())

    @property
    def cpml_sides(self):
        return {face: kind == CPML for face, kind in self.faces.items()}

    def periodic(self, axis):
        return self.faces[f'{axis}_lo'] == PERIODIC

    @property
    def periodic_x(self):
        return self.periodic('x')

    @property
    def periodic_y(self):
        return self.periodic('y')

    def pec_faces(self):
        """Faces of type 'pec'."""
        return [face for face, kind in self.faces.items() if kind == PEC]

    @property
    def npml_sides(self):
        """Padding cells per face: cpml_npml on cpml faces, one ghost cell on mirror faces, else 0
        (physics/em/boundaries.py:283-286)."""
        return {face: (self.cpml_npml if kind == CPML else 1 if kind in MIRRORS else 0)
                for face, kind in self.faces.items()}

    @classmethod
    def all_pml(cls, npml=15, dimensions=2):
        bc = cls(*(['cpml'] * 4), *((['cpml'] * 2) if dimensions == 3 else []))
        bc.cpml_npml = npml
        return bc

    @classmethod
    def with_periodic_x(cls, npml_y=15):
        bc = cls('periodic', 'periodic', 'cpml', 'cpml')
        bc.cpml_npml = npml_y
        return bc

    @classmethod
    def with_periodic_y(cls, npml_x=15):
        bc = cls('cpml', 'cpml', 'periodic', 'periodic')
        bc.cpml_npml = npml_x
        return bc

    @classmethod
    def periodic_xy(cls):
        return cls('periodic', 'periodic', 'periodic', 'periodic')

    def __repr__(self):
        return 'BoundaryConfig(' + ', '.join(f"{f}='{k}'" for f, k in self.faces.items()) + ')'


def resolve_boundaries(boundaries, npml):
    """Simulation presets (simulation.py:388-400): None / 'pml', 'periodic_x', 'periodic_y', 'periodic'."""
    if boundaries is None:
        return BoundaryConfig.all_pml(npml)
    if isinstance(boundaries, BoundaryConfig):
        boundaries.validate()
        return boundaries
    presets = {'pml': lambda: BoundaryConfig.all_pml(npml), 'periodic_x': lambda: BoundaryConfig.with_periodic_x(npml),
               'periodic_y': lambda: BoundaryConfig.with_periodic_y(npml), 'periodic': BoundaryConfig.periodic_xy}
    if boundaries not in presets:
        raise ValueError(f"Unknown boundary preset: '{boundaries}'. Options: {list(presets)}")
    return presets[boundaries]()


def side_rows(side, n, npml):
    """(centre rows, hi-face rows) of one CPML side (physics/em/boundaries.py:570-576): lo 1..npml / 0..npml-1,
    hi n-npml-1..n-2 for both kinds."""
    if side == 'lo':
        return slice(1, npml + 1), slice(0, npml)
    return slice(n - npml - 1, n - 1), slice(n - npml - 1, n - 1)


def sigma_max(bc, pml_dx):
    """sigma_factor (order + 1) / (150 pi pml_dx) (physics/em/boundaries.py:579-582)."""
    return bc.cpml_sigma_factor * (bc.cpml_order + 1) / (150.0 * math.pi * float(pml_dx))


class CPMLAxisTables:
    """Host float64 CPML rows of one axis, rounded to float32 for the kernel: psi row maps (-1 outside the
    layers), recursion b and gain g = a/d per psi row, and kappa per grid row for the curl fold."""

    def __init__(self, bc, axis, n, dx, dt):
        npml = int(bc.cpml_npml)
        lo, hi = bc.faces[f'{axis}_lo'] == CPML, bc.faces[f'{axis}_hi'] == CPML
        layers = int(lo) + int(hi)
        self.n = n
        self.n_rows = layers * npml
        self.row_e = np.full(n, -1, dtype=np.int32)
        self.row_h = np.full(n, -1, dtype=np.int32)
        self.kappa_e = np.ones(n)
        self.kappa_h = np.ones(n)
        b_e, g_e, b_h, g_h = (np.zeros(self.n_rows) for _ in range(4))
        if layers == 0:
            self.b_e = self.g_e = self.b_h = self.g_h = np.zeros(0, dtype=DTYPE)
            return
        if n < layers * npml + 2:
            raise ValueError(f'{axis} axis: {n} cells cannot hold {layers} CPML layer(s) of {npml} cells plus the 2 '
                             f'unstretched rows; it needs at least {layers * npml + 2} cells')
        smax = sigma_max(bc, dx)
        inv_bulk = 1.0 / float(dx)

        def graded(depth):
            sigma = smax * depth ** bc.cpml_order
            kappa = 1.0 + (bc.cpml_kappa_max - 1.0) * depth ** bc.cpml_order
            alpha = bc.cpml_alpha_max * (1.0 - depth)
            b = np.exp(-(sigma / kappa + alpha) * float(dt) / EPS0)
            denom = sigma * kappa + kappa ** 2 * alpha
            a = np.divide(sigma * (b - 1.0), denom, out=np.zeros_like(denom), where=denom > 1e-30)
            return kappa, b, a * inv_bulk

        base = 0
        for side, active, depth_e, depth_h in (
                ('lo', lo, np.arange(npml - 1, -1, -1.0) / npml, (np.arange(npml, 0, -1.0) - 0.5) / npml),
                ('hi', hi, np.arange(0, npml, 1.0) / npml, (np.arange(1, npml + 1, 1.0) - 0.5) / npml)):
            if not active:
                continue
            rows_e, rows_h = side_rows(side, n, npml)
            psi_rows = np.arange(base, base + npml, dtype=np.int32)
            kappa, b, g = graded(depth_e)
            self.row_e[rows_e] = psi_rows
            self.kappa_e[rows_e] = kappa
            b_e[psi_rows], g_e[psi_rows] = b, g
            kappa, b, g = graded(depth_h)
            self.row_h[rows_h] = psi_rows
            self.kappa_h[rows_h] = kappa
            b_h[psi_rows], g_h[psi_rows] = b, g
            base += npml
        self.b_e, self.g_e = b_e.astype(DTYPE), g_e.astype(DTYPE)
        self.b_h, self.g_h = b_h.astype(DTYPE), g_h.astype(DTYPE)

    def layer_bounds(self):
        """(lo, hi): every row with a psi row of either kind lies in [0, lo) or [hi, n) (lo side: rows 0..npml,
        hi side: rows n-npml-1..n-2); (0, n) without layers."""
        rows = np.flatnonzero((self.row_e >= 0) | (self.row_h >= 0))
        lo_rows = rows[rows < self.n // 2]
        hi_rows = rows[rows >= self.n // 2]
        lo = int(lo_rows.max()) + 1 if lo_rows.size else 0
        hi = int(hi_rows.min()) if hi_rows.size else self.n
        return lo, hi


def curl_spacings(n, dx, cpml, zero_last_h=False):
    """float32 (inv_e, inv_h) of one axis: 1/w at centres, 1/mean(w_i, w_(i+1)) at hi faces (last = periodic wrap),
    with CPML kappa divided in (physics/em/yee_grid.py:11-45, solvers/fdtd_2d_base.py:76-84,
    boundaries.py:778-798). Kernel widths are the float32 grid dx (yee_grid.py:11-16). zero_last_h: the 3D engine's
    inv_h[n-1] = 0 on a non-periodic axis (physics/em3d/engine3d.py:201-210)."""
    widths = np.full(n, float(DTYPE(dx)))
    inv_e = 1.0 / widths
    inv_h = 1.0 / (0.5 * (widths + np.roll(widths, -1)))
    if zero_last_h:
        inv_h[-1] = 0.0
    if cpml is not None:
        inv_e = inv_e / cpml.kappa_e
        inv_h = inv_h / cpml.kappa_h
    return inv_e.astype(DTYPE), inv_h.astype(DTYPE)
