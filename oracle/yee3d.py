"""ORACLE (test infrastructure only): numpy restatement of the reference's 3D FDTD hot path.

Six-component Yee stepper on (nx, ny, nz) arrays (z contiguous) with the fused CPML psi of every axis, periodic
wraps, width-weighted face coefficients, a calibrated point dipole and plane flux monitors, with the reference's
geometry, clock and quantisation points:

  * dt is the float32 3D CFL step of the float32 cell width (core/grid3d.py:8-16, 59-65); kernel widths are that
    float32 width (physics/em3d/engine3d.py:201, 206-216);
  * material arrays are float32; boxes are painted half-open on the float32 cell centres clamped to the window
    (core/domain.py:90-98, model/geometry.py:1762-1778, 1786-1788; later writes win);
  * E_x / E_y / E_z sit on the hi-x / hi-y / hi-z cell faces and take the mean of the two cells' eps and sigma
    (physics/em3d/kernels3d.py:61-85; physics/em/yee_grid.py:119-126); H sits on edges with the lower cell's mu
    (kernels3d.py:44-56);
  * non-periodic axes: inv_h[n-1] = 0 and the lo neighbour of cell 0 reads 0 (engine3d.py:206-210,
    kernels3d.py:335-400, 538-610);
  * CPML rows and recursion as 2D (physics/em/boundaries.py:551-631), psi gain = a x bulk inverse spacing,
    kappa folded into the curl rows (physics/em3d/boundaries3d.py:208-241);
  * dipole weight = face cb x moment / dual volume on the E face nearest the point (physics/em3d/sources3d.py:31-75,
    engine3d.py:299-311, physics3d.py:403-422);
  * plane flux: E = mean of the two layers across the plane, H on the plane, trapezoid face weights, DFT phases
    e^{-j omega t} m dt (complex64 in the reference, float64 here unless phase_dtype says otherwise), flux =
    cos(pi f dt) sum sign Re(E conj(H e^{j pi f dt})) area (physics/em3d/monitors3d.py:22-128,
    physics/em/monitors.py:22-34, 86-96).

dtype=np.float64 is the reference oracle; dtype=np.float32 runs the same array passes in single precision (the CPU
baseline "port" of the reference's numpy twins). Everything else is float64 host arithmetic.
"""
import math

import numpy as np

from .yee2d import (C0, EPS0, F32, MU0, FluxMonitor, PointProbe, centre_partners, containing_cell, cpml_recursion,
                    cpml_side_rows, dft_sample_interval, make_temporal, resolve_frequency, uniform_cell_count)

AXES = ('x', 'y', 'z')
FACES_3D = ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')
DEFAULT_COURANT = 0.99


def courant_time_step_3d(dx, dy, dz, courant=None):
    """float32 dt = courant / (C0 sqrt(1/dx^2 + 1/dy^2 + 1/dz^2)) (core/grid3d.py:8-16)."""
    courant = DEFAULT_COURANT if courant is None else float(courant)
    if not 0.0 < courant <= 1.0:
        raise ValueError(f'courant must be in (0, 1], got {courant}')
    return F32(courant / (C0 * np.sqrt(1.0 / float(dx) ** 2 + 1.0 / float(dy) ** 2 + 1.0 / float(dz) ** 2)))


class Boundaries3D:
    """Face types ('cpml' / 'periodic' / 'pec' / 'pmc' / 'symmetric' / 'antisymmetric') and CPML settings
    (physics/em3d/boundaries3d.py:17-60). A 'pmc' face is the bare grid edge (engine3d.py:12-13); a 'pec' face makes
    the outermost window cell layer PEC (engine3d.py:13, boundaries3d.py:149-158); a mirror face pads one ghost cell
    outside the window (engine3d.py:14-17)."""

    def __init__(self, faces=None, npml=12, order=3, sigma_factor=1.0, kappa_max=15.0, alpha_max=0.0):
        faces = dict(faces or {})
        self.faces = {face: faces.get(face, 'cpml') for face in FACES_3D}
        for face, kind in self.faces.items():
            if kind not in ('cpml', 'periodic', 'pec', 'pmc', 'symmetric', 'antisymmetric'):
                raise ValueError(f'oracle supports cpml, periodic, pec, pmc and mirror faces, got {face}={kind}')
        for axis in AXES:
            if (self.faces[f'{axis}_lo'] == 'periodic') != (self.faces[f'{axis}_hi'] == 'periodic'):
                raise ValueError(f'{axis}: periodic faces must be paired')
        self.npml, self.order, self.sigma_factor = int(npml), int(order), float(sigma_factor)
        self.kappa_max, self.alpha_max = float(kappa_max), float(alpha_max)

    def cpml(self, face):
        return self.faces[face] == 'cpml'

    def periodic(self, axis):
        return self.faces[f'{axis}_lo'] == 'periodic'

    def npml_sides(self):
        return {face: (self.npml if self.cpml(face) else 1 if self.faces[face] in MIRRORS_3D else 0)
                for face in FACES_3D}

    def mirror_planes(self, windows):
        """[(axis, kind, ghost row, inner row, plane-face row)] of the mirror faces (boundaries3d.py:133-146): the
        plane is the window edge, the hi face of the lo ghost / of the last window cell."""
        planes = []
        for axis, name in enumerate(AXES):
            for side in ('lo', 'hi'):
                kind = self.faces[f'{name}_{side}']
                if kind not in MIRRORS_3D:
                    continue
                cells = windows[axis]
                ghost, inner = (cells.start - 1, cells.start) if side == 'lo' else (cells.stop, cells.stop - 1)
                planes.append((axis, kind, ghost, inner, cells.start - 1 if side == 'lo' else cells.stop - 1))
        return planes


MIRRORS_3D = ('symmetric', 'antisymmetric')


def axis_profiles(bc, axis, n, pml_dx):
    """float64 (sigma, kappa, alpha) at the E rows (centres) and H rows (hi faces) of one axis
    (physics/em/boundaries.py:551-631 through physics/em3d/boundaries3d.py:174-205)."""
    lo, hi = bc.cpml(f'{axis}_lo'), bc.cpml(f'{axis}_hi')
    npml = bc.npml
    e = [np.zeros(n), np.ones(n), np.zeros(n)]
    h = [np.zeros(n), np.ones(n), np.zeros(n)]
    if not (lo or hi):
        return e, h
    if n < (int(lo) + int(hi)) * npml + 2:
        raise ValueError(f'{axis} axis: {n} cells cannot hold the CPML layers')
    sigma_max = bc.sigma_factor * (bc.order + 1) / (150.0 * np.pi * float(pml_dx))

    def graded(depth):
        return (sigma_max * depth ** bc.order, 1.0 + (bc.kappa_max - 1.0) * depth ** bc.order,
                bc.alpha_max * (1.0 - depth))

    for side, active, depth_e, depth_h in (
            ('lo', lo, np.arange(npml - 1, -1, -1.0) / npml, (np.arange(npml, 0, -1.0) - 0.5) / npml),
            ('hi', hi, np.arange(0, npml, 1.0) / npml, (np.arange(1, npml + 1, 1.0) - 0.5) / npml)):
        if not active:
            continue
        rows_e, rows_h = cpml_side_rows(side, n, npml)
        for profile, rows, depth in ((e, rows_e, depth_e), (h, rows_h, depth_h)):
            values = graded(depth)
            for target, value in zip(profile, values):
                target[rows] = value
    return e, h


def along(axis, index):
    """Index tuple selecting `index` along axis (physics/em3d/boundaries3d.py `along`)."""
    key = [slice(None)] * 3
    key[axis] = index
    return tuple(key)


def row_shape(axis):
    shape = [1, 1, 1]
    shape[axis] = -1
    return tuple(shape)


class CpmlAxis:
    """CPML of one axis and kind ('e' / 'h'): grid rows in the layers, b and gain = a x bulk inverse spacing per layer
    row, and psi of the axis' first and second component (physics/em3d/boundaries3d.py:208-241)."""

    def __init__(self, axis, kind, rows, b, gain, shape, dtype):
        self.axis, self.kind = axis, kind
        self.rows = np.asarray(rows, dtype=np.int64)
        self.b = np.asarray(b, dtype=dtype).reshape(row_shape(axis))
        self.gain = np.asarray(gain, dtype=dtype).reshape(row_shape(axis))
        psi_shape = list(shape)
        psi_shape[axis] = self.rows.size
        self.psi = [np.zeros(psi_shape, dtype=dtype), np.zeros(psi_shape, dtype=dtype)]

    def step(self, which, difference):
        """psi = b psi + gain dF on the layer rows of `difference` (a full-grid array); returns psi."""
        psi = self.psi[which]
        psi *= self.b
        psi += self.gain * np.take(difference, self.rows, axis=self.axis)
        return psi

    def add(self, target, coefficient, sign, psi):
        """target[rows] += sign coefficient[rows] psi."""
        key = along(self.axis, self.rows)
        target[key] += sign * coefficient[key] * psi


class DipoleSource:
    """E_component[cell] += weight s(t + dt/2), weight = cb moment / dual volume (physics/em3d/sources3d.py:41-75)."""

    def __init__(self, component, cell, temporal, moment, volume):
        self.component, self.cell = component, tuple(int(v) for v in cell)
        self.temporal, self.moment, self.volume = temporal, float(moment), float(volume)
        self.weight = None

    def inject(self, fields, t, dtype):
        value = self.weight * self.temporal(t)
        fields[self.component][self.cell] += dtype(value)


def face_samples(widths, first_edge, last_edge, wrap=False):
    """(face slice, trapezoid weights) of the samples on cell edges first..last (physics/em3d/monitors3d.py:22-34)."""
    if wrap:
        widths = np.asarray(widths, dtype=np.float64)
        return slice(0, widths.size), 0.5 * (widths + np.roll(widths, -1))
    edges = np.arange(max(first_edge, 1), last_edge + 1)
    weights = np.empty(edges.size)
    for position, edge in enumerate(edges):
        below = widths[edge - 1] if edge > first_edge else 0.0
        above = widths[edge] if edge < last_edge else 0.0
        weights[position] = 0.5 * (below + above)
    return slice(int(edges[0]) - 1, int(edges[-1])), weights


class PlaneFlux:
    """Running DFT of the tangential E (mean of the two layers across the plane) and H on the cell-edge plane `edge`
    normal to `normal` over the tangential edge ranges `bounds` (physics/em3d/monitors3d.py:42-128)."""

    def __init__(self, normal, edge, bounds, frequencies, dt, widths, shape, periodic, phase_dtype):
        if not 1 <= edge <= shape[normal] - 1:
            raise ValueError(f'flux plane at edge {edge} needs a cell on both sides')
        self.normal, self.edge = int(normal), int(edge)
        self.frequencies = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
        self.dt = float(dt)
        self.sample_interval = 1
        self.phase_dtype = phase_dtype
        first, second = (normal + 1) % 3, (normal + 2) % 3
        self.pairs = []
        for e_axis, h_axis, face_axis, cell_axis, sign in ((first, second, first, second, 1.0),
                                                           (second, first, second, first, -1.0)):
            wrap = bool(periodic[face_axis]) and tuple(bounds[face_axis]) == (0, shape[face_axis])
            face_slice, face_weights = face_samples(widths[face_axis], *bounds[face_axis], wrap=wrap)
            lo, hi = bounds[cell_axis]
            cell_slice, cell_weights = slice(lo, hi), np.asarray(widths[cell_axis][lo:hi], dtype=np.float64)
            lower, upper, h_index = [None] * 3, [None] * 3, [None] * 3
            for index in (lower, upper, h_index):
                index[face_axis], index[cell_axis] = face_slice, cell_slice
            lower[normal], upper[normal], h_index[normal] = edge - 1, edge, edge - 1
            by_axis = {face_axis: face_weights, cell_axis: cell_weights}
            low_axis, high_axis = sorted((face_axis, cell_axis))
            area = np.outer(by_axis[low_axis], by_axis[high_axis])
            shape_f = (self.frequencies.size,) + area.shape
            self.pairs.append(dict(e='e' + AXES[e_axis], h='h' + AXES[h_axis], sign=sign, area=area,
                                   lower=tuple(lower), upper=tuple(upper), h_index=tuple(h_index),
                                   face_axis=face_axis, cell_axis=cell_axis, face_slice=face_slice,
                                   cell_slice=cell_slice,
                                   e_dft=np.zeros(shape_f, np.complex128), h_dft=np.zeros(shape_f, np.complex128)))

    def phases(self, t):
        """exp(-j omega t) m dt (physics/em/monitors.py:22-34), rounded to phase_dtype."""
        p = -2.0 * np.pi * self.frequencies * float(t)
        scale = self.dt * self.sample_interval
        if self.phase_dtype == np.complex64:
            return ((np.cos(p) * scale).astype(F32) + 1j * (np.sin(p) * scale).astype(F32)).astype(np.complex128)
        return np.cos(p) * scale + 1j * (np.sin(p) * scale)

    def update(self, fields, t):
        ph = self.phases(t)[:, None, None]
        for pair in self.pairs:
            e = fields[pair['e']]
            sample = 0.5 * (e[pair['lower']].astype(np.float64) + e[pair['upper']].astype(np.float64))
            pair['e_dft'] += ph * sample[None]
            pair['h_dft'] += ph * fields[pair['h']][pair['h_index']].astype(np.float64)[None]

    def spectra(self):
        """Per pair (E spectrum, H spectrum re-timed to the E time) (monitors3d.py:103-110)."""
        retime = np.exp(1j * np.pi * self.frequencies * self.dt)[:, None, None]
        return [(pair, pair['e_dft'], pair['h_dft'] * retime) for pair in self.pairs]

    def flux(self):
        total = np.zeros(self.frequencies.size)
        for pair, e_spectrum, h_spectrum in self.spectra():
            total += pair['sign'] * np.einsum('fij,ij->f', np.real(e_spectrum * np.conj(h_spectrum)), pair['area'])
        return np.cos(np.pi * self.frequencies * self.dt) * total


class FluxBox:
    """Net outward flux through the six faces of a box of edge ranges (monitors3d.py:131-162): the lo face of each
    axis with sign -1, the hi face with +1, each a PlaneFlux over the box's ranges."""

    def __init__(self, faces):
        self.faces = faces

    def flux(self):
        return sum(sign * plane.flux() for sign, plane in self.faces)


class DFTPlane:
    """Running DFT of E_x, E_y, E_z at the cell centres of window layer k (monitors3d.py:165-238): E_x with its lower
    x face, E_y with its lower y face, E_z with the layer below (index 0: the wrapped last one on a periodic axis,
    else itself). complex64 accumulation, as the reference's, when phase_dtype is complex64."""

    def __init__(self, windows, k, frequencies, dt, shape, periodic, phase_dtype):
        self.xs = np.arange(windows[0].start, windows[0].stop)
        self.ys = np.arange(windows[1].start, windows[1].stop)
        self.k = int(k)
        self.x_partners = centre_partners(self.xs, shape[0], periodic[0])
        self.y_partners = centre_partners(self.ys, shape[1], periodic[1])
        self.k_lower = int(centre_partners([self.k], shape[2], periodic[2])[0])
        self.frequencies = np.asarray(frequencies, dtype=np.float64)
        self.omega = 2.0 * np.pi * self.frequencies
        self.dt = float(dt)
        self.phase_dtype = phase_dtype
        self.sample_interval = 1
        acc_dtype = np.complex64 if phase_dtype == np.complex64 else np.complex128
        self.spectra = {name: np.zeros((self.frequencies.size, self.xs.size, self.ys.size), dtype=acc_dtype)
                        for name in ('ex', 'ey', 'ez')}

    phases = FluxMonitor.phases

    def update(self, fields, t):
        ph = self.phases(t)[:, None, None]
        half = F32(0.5)
        box = np.ix_(self.xs, self.ys)
        ex, ey, ez = (fields[name][:, :, self.k] for name in ('ex', 'ey', 'ez'))
        self.spectra['ex'] += ph * (half * (ex[box] + ex[np.ix_(self.x_partners, self.ys)]))
        self.spectra['ey'] += ph * (half * (ey[box] + ey[np.ix_(self.xs, self.y_partners)]))
        self.spectra['ez'] += ph * (half * (ez[box] + fields['ez'][:, :, self.k_lower][box]))

    def intensity(self, index):
        return sum(np.abs(self.spectra[name][index].astype(np.complex128)) ** 2 for name in ('ex', 'ey', 'ez'))


class Yee3D:
    """3D FDTD model (physics/em3d/physics3d.py:144-445, engine3d.py:162-554) stepped on the CPU in `dtype`.
    size_* is the physical window; every cpml face adds npml cells of width dx."""

    def __init__(self, size_x, size_y, size_z, dx, npml=12, courant=None, boundaries=None, dtype=np.float64,
                 phase_dtype=np.complex128):
        self.bc = boundaries if isinstance(boundaries, Boundaries3D) else Boundaries3D(boundaries, npml=npml)
        sides = self.bc.npml_sides()
        self.dx = float(dx)
        self.sizes = (float(size_x), float(size_y), float(size_z))
        self.shape = tuple(uniform_cell_count(size + (sides[f'{a}_lo'] + sides[f'{a}_hi']) * self.dx, self.dx)
                           for size, a in zip(self.sizes, AXES))
        self.nx, self.ny, self.nz = self.shape
        self.dx32 = F32(dx)
        self.dt = courant_time_step_3d(self.dx32, self.dx32, self.dx32, courant)
        self.origins = tuple(sides[f'{a}_lo'] for a in AXES)
        self.windows = tuple(slice(sides[f'{a}_lo'], n - sides[f'{a}_hi']) for a, n in zip(AXES, self.shape))
        self.periodic = tuple(self.bc.periodic(a) for a in AXES)
        # float64 geometry of the domain (core/domain.py:58-81) and float32 kernel widths (engine3d.py:201)
        self.geo_widths = tuple(np.full(n, self.dx) for n in self.shape)
        self.edges = tuple((np.arange(n + 1) - o) * self.dx for n, o in zip(self.shape, self.origins))
        self.centres = tuple(0.5 * (e[:-1] + e[1:]) for e in self.edges)
        self.widths = tuple(np.full(n, float(self.dx32)) for n in self.shape)
        self.dtype = np.dtype(dtype).type
        self.phase_dtype = phase_dtype
        self.fields = {name: np.zeros(self.shape, dtype=self.dtype) for name in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')}
        self.materials = {'eps_r': np.ones(self.shape, F32), 'sigma_e': np.zeros(self.shape, F32),
                          'mu_r': np.ones(self.shape, F32)}
        self.sources, self.monitors, self.probes = [], [], []
        self.pec_boxes = []
        self.poles = []
        self.n_step, self.time = 0, 0.0
        self._finalized = False
        self._build_spacings_and_cpml()

    # ---- setup ------------------------------------------------------------------------------------------------
    def _build_spacings_and_cpml(self):
        """Bulk inverse spacings (inv_h[n-1] = 0 on non-periodic axes), CPML layers, kappa folded into the curl rows
        (engine3d.py:201-216, boundaries3d.py:208-241)."""
        real = self.dtype
        self.inv_e, self.inv_h, self.cpml = [], [], []
        for axis, (n, widths, periodic) in enumerate(zip(self.shape, self.widths, self.periodic)):
            inverse_e = 1.0 / widths
            inverse_h = 1.0 / (0.5 * (widths + np.roll(widths, -1)))
            if not periodic:
                inverse_h[-1] = 0.0
            (sigma_e, kappa_e, alpha_e), (sigma_h, kappa_h, alpha_h) = axis_profiles(
                self.bc, AXES[axis], n, float(max(widths[0], widths[-1])))
            self.inv_e.append((inverse_e / kappa_e).astype(real))
            self.inv_h.append((inverse_h / kappa_h).astype(real))
            name = AXES[axis]
            sides = [side for side in ('lo', 'hi') if self.bc.cpml(f'{name}_{side}')]
            if not sides:
                continue
            for kind, sigma, kappa, alpha, inverse in (('e', sigma_e, kappa_e, alpha_e, inverse_e),
                                                       ('h', sigma_h, kappa_h, alpha_h, inverse_h)):
                rows = []
                for side in sides:
                    rows_e, rows_h = cpml_side_rows(side, n, self.bc.npml)
                    selected = rows_e if kind == 'e' else rows_h
                    rows.extend(range(selected.start, selected.stop))
                rows = np.asarray(rows)
                a, b = cpml_recursion(sigma[rows], kappa[rows], alpha[rows], float(self.dt), np.float64)
                self.cpml.append(CpmlAxis(axis, kind, rows, b, a * inverse[rows], self.shape, real))

    def host_centres(self):
        """float32 cell centres clamped to the window's first / last centre (core/domain.py:90-98)."""
        out = []
        for centres, window in zip(self.centres, self.windows):
            out.append(np.clip(centres, centres[window.start], centres[window.stop - 1]).astype(F32))
        return out

    def box_mask(self, lo, hi):
        """Cells whose clamped float32 centre lies in [min, max) on every axis (model/geometry.py:1762-1788)."""
        masks = []
        for c, a, b in zip(self.host_centres(), lo, hi):
            a, b = sorted((float(a), float(b)))
            masks.append((c >= F32(a)) & (c < F32(b)))
        return masks[0][:, None, None] & masks[1][None, :, None] & masks[2][None, None, :]

    def add_dielectric_box(self, lo, hi, eps_r, sigma_e=0.0):
        mask = self.box_mask(lo, hi)
        self.materials['eps_r'][mask] = F32(eps_r)
        self.materials['sigma_e'][mask] = F32(sigma_e)
        self._finalized = False

    def add_pec_box(self, lo, hi):
        """PEC cells on the box raster, kept over every dielectric (physics3d.py:351-353, engine3d.py:358-360)."""
        self.pec_boxes.append(self.box_mask(lo, hi))
        self._finalized = False

    def pec_wall_layers(self):
        """Outermost window cell layer of every 'pec' face (boundaries3d.py:149-158)."""
        layers = []
        for axis, name in enumerate(AXES):
            for side in ('lo', 'hi'):
                if self.bc.faces[f'{name}_{side}'] == 'pec':
                    window = self.windows[axis]
                    index = window.start if side == 'lo' else window.stop - 1
                    layers.append(along(axis, slice(index, index + 1)))
        return layers

    def pec_mask(self):
        """bool per cell: PEC boxes and pec wall layers."""
        mask = np.zeros(self.shape, dtype=bool)
        for box in self.pec_boxes:
            mask |= box
        for layer in self.pec_wall_layers():
            mask[layer] = True
        return mask

    def validate(self, position, label):
        """core/domain.py:125-133."""
        tol = 1e-9 * max(self.sizes)
        if not all(-tol <= v <= s + tol for v, s in zip(position, self.sizes)):
            raise ValueError(f'{label} position {position} outside the physical domain')

    def cell(self, axis, coord):
        window = self.windows[axis]
        return containing_cell(coord, self.edges[axis], window.start, window.stop - 1)

    def nearest_face(self, axis, coord):
        """E face (hi face of cell f) nearest coord among the window's inner faces (sources3d.py:31-38)."""
        window = self.windows[axis]
        first, last = window.start, window.stop - 2
        if last < first:
            raise ValueError('a dipole along an axis needs at least 2 window cells on that axis')
        return min(range(first, last + 1), key=lambda f: abs(float(self.edges[axis][f + 1]) - float(coord)))

    def nearest_edge(self, axis, coord):
        """Grid edge nearest coord among the window's edges (physics3d.py:492-496)."""
        window = self.windows[axis]
        first, last = window.start, window.stop
        return first + int(np.argmin(np.abs(self.edges[axis][first:last + 1] - float(coord))))

    def face_dual_volume(self, axis, cell):
        """engine3d.py:299-311 (Cartesian)."""
        lengths = []
        for other, widths in enumerate(self.widths):
            index = cell[other]
            partner = (index + 1) % widths.size if self.periodic[other] else min(index + 1, widths.size - 1)
            lengths.append(0.5 * (widths[index] + widths[partner]) if other == axis else widths[index])
        return float(math.prod(lengths))

    def add_dipole(self, x, y, z, orientation='z', dipole_moment=1e-12, wavelength=None, frequency=None,
                   bandwidth=0.5, waveform='gaussian'):
        """physics/em3d/physics3d.py:403-422."""
        if orientation not in AXES:
            raise ValueError(f'dipole orientation must be x, y or z, got {orientation!r}')
        self.validate((x, y, z), 'Dipole')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        axis = AXES.index(orientation)
        cell = [self.cell(a, v) for a, v in enumerate((x, y, z))]
        cell[axis] = self.nearest_face(axis, (x, y, z)[axis])
        for layer in self.pec_wall_layers():   # engine3d.py:379-382, physics3d.py:415-417
            wall_axis = next(a for a, key in enumerate(layer) if key != slice(None))
            if wall_axis != axis and cell[wall_axis] == layer[wall_axis].start:
                raise ValueError('dipole lies in the outermost cell layer of a pec face parallel to it')
        source = DipoleSource('e' + orientation, cell, temporal, dipole_moment, self.face_dual_volume(axis, cell))
        self.sources.append(source)
        self._finalized = False
        return source

    def add_field_probe(self, x, y, z, component='ez'):
        """Samples of one component at the cell containing (x, y, z) after every step (physics3d.py:425-433)."""
        if component not in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'):
            raise ValueError(f'probe component must be one of ex, ey, ez, hx, hy, hz, got {component!r}')
        self.validate((x, y, z), 'Field probe')
        probe = PointProbe(tuple(self.cell(axis, v) for axis, v in enumerate((x, y, z))), component)
        self.probes.append(probe)
        return probe

    def add_flux_box(self, lo, hi, wavelengths=None, frequencies=None):
        """Box between the window edges nearest the corners (physics3d.py:447-476)."""
        if (wavelengths is None) == (frequencies is None):
            raise ValueError('pass exactly one of wavelengths or frequencies')
        freqs = (np.asarray(frequencies, dtype=np.float64) if frequencies is not None
                 else C0 / np.asarray(wavelengths, dtype=np.float64))
        bounds = []
        for axis in range(3):
            first = self.nearest_edge(axis, min(lo[axis], hi[axis]))
            last = self.nearest_edge(axis, max(lo[axis], hi[axis]))
            if last <= first:
                raise ValueError('flux box spans no cell')
            bounds.append((first, last))
        faces = []
        for axis in range(3):
            for side, sign in ((0, -1.0), (1, 1.0)):
                plane = PlaneFlux(axis, bounds[axis][side], bounds, np.atleast_1d(freqs), float(self.dt),
                                  self.widths, self.shape, self.periodic, self.phase_dtype)
                self.monitors.append(plane)
                faces.append((sign, plane))
        self._finalized = False
        return FluxBox(faces)

    def add_dispersive_drude(self, mask, omega_p, gamma):
        """Drude pole on a bool (nx, ny, nz) cell mask over the cells' permittivity (physics3d.py:346-349)."""
        self.poles.append({'pole': ('drude', omega_p, gamma), 'mask': np.asarray(mask, dtype=bool).copy()})
        self._finalized = False

    def add_dispersive_lorentz(self, mask, delta_eps, omega_0, gamma):
        self.poles.append({'pole': ('lorentz', delta_eps, omega_0, gamma),
                           'mask': np.asarray(mask, dtype=bool).copy()})
        self._finalized = False

    def _pole_setup(self):
        """Per pole and E component: the face average of the mask along the component's axis (engine3d.py:461-488,
        uniform cells) times the float32 drive coefficient, and zero states."""
        from .yee2d import drude_coefficients, lorentz_coefficients
        real = self.dtype
        for entry in self.poles:
            kind = entry['pole'][0]
            coefficients = (drude_coefficients(*entry['pole'][1:], self.dt) if kind == 'drude'
                            else lorentz_coefficients(*entry['pole'][1:], self.dt))
            entry['coefficients'] = tuple(real(c) for c in coefficients)
            m = entry['mask'].astype(real)
            entry['drive'] = []
            for axis, n in enumerate(self.shape):
                partner = (np.arange(n) + 1) % n if self.periodic[axis] else np.minimum(np.arange(n) + 1, n - 1)
                weight = real(0.5) * m + real(0.5) * np.take(m, partner, axis=axis)
                entry['drive'].append(weight * real(coefficients[-1]))
            if 'state' not in entry:
                entry['state'] = [[np.zeros(self.shape, dtype=real) for _ in range(1 if kind == 'drude' else 2)]
                                  for _ in range(3)]

    def add_dft_plane(self, z, wavelengths=None, frequencies=None):
        """DFT of E at the cell centres of the window layer containing z (physics3d.py:478-486)."""
        if (wavelengths is None) == (frequencies is None):
            raise ValueError('pass exactly one of wavelengths or frequencies')
        freqs = (np.asarray(frequencies, dtype=np.float64) if frequencies is not None
                 else C0 / np.asarray(wavelengths, dtype=np.float64))
        self.validate((0.0, 0.0, z), 'DFT plane')
        monitor = DFTPlane(self.windows, self.cell(2, z), np.atleast_1d(freqs), float(self.dt), self.shape,
                           self.periodic, self.phase_dtype)
        self.monitors.append(monitor)
        self._finalized = False
        return monitor

    def add_flux_plane(self, normal, position, wavelengths=None, frequencies=None):
        """Flux through the cell-edge plane nearest `position` across the whole window (physics3d.py:434-445)."""
        if (wavelengths is None) == (frequencies is None):
            raise ValueError('pass exactly one of wavelengths or frequencies')
        freqs = (np.asarray(frequencies, dtype=np.float64) if frequencies is not None
                 else C0 / np.asarray(wavelengths, dtype=np.float64))
        axis = AXES.index(normal)
        bounds = [(w.start, w.stop) for w in self.windows]
        edge = self.nearest_edge(axis, position)
        bounds[axis] = (edge, edge)
        monitor = PlaneFlux(axis, edge, bounds, freqs, float(self.dt), self.widths, self.shape, self.periodic,
                            self.phase_dtype)
        self.monitors.append(monitor)
        self._finalized = False
        return monitor

    def face_coefficients(self, axis):
        """(ca, cb) of the E component on the hi face along axis: mean eps / sigma of the cell and its hi partner
        (kernels3d.py:61-85; uniform widths: shares 0.5 / 0.5)."""
        eps, sig = (self.materials[name].astype(np.float64) for name in ('eps_r', 'sigma_e'))
        pec = self.pec_mask()
        n = self.shape[axis]
        partner = np.arange(1, n + 1)
        partner[-1] = 0 if self.periodic[axis] else n - 1
        eps_b, sig_b, pec_b = (np.take(a, partner, axis=axis) for a in (eps, sig, pec))
        # PEC rule (kernels3d.py:61-85): a face touching one PEC cell takes the other cell's values on both sides;
        # a face between two PEC cells is pinned (ca = cb = 0)
        eps_a = np.where(pec, eps_b, eps)
        sig_a = np.where(pec, sig_b, sig)
        eps_b = np.where(pec_b & ~pec, eps, eps_b)
        sig_b = np.where(pec_b & ~pec, sig, sig_b)
        eps_face = (0.5 * eps_a + 0.5 * eps_b) * EPS0
        sig_face = 0.5 * sig_a + 0.5 * sig_b
        pinned = pec & pec_b
        eps_face = np.where(pinned, 1.0, eps_face)   # any finite value: the face is zeroed below
        loss = sig_face * float(self.dt) / (2.0 * eps_face)
        ca = np.where(pinned, 0.0, (1.0 - loss) / (1.0 + loss))
        cb = np.where(pinned, 0.0, (float(self.dt) / eps_face) / (1.0 + loss))
        return ca.astype(self.dtype), cb.astype(self.dtype)

    def finalize(self):
        """Coefficients, dipole calibration and the DFT sample interval (physics3d.py:192-217)."""
        self.ca, self.cb = zip(*(self.face_coefficients(axis) for axis in range(3)))
        mu = self.materials['mu_r'].astype(np.float64) * MU0
        self.da = np.ones(self.shape, dtype=self.dtype)
        self.db = (float(self.dt) / mu).astype(self.dtype)
        for source in self.sources:
            axis = AXES.index(source.component[1])
            source.weight = float(self.cb[axis][source.cell]) * source.moment / source.volume
        self._pole_setup()
        spectral = [m.frequencies for m in self.monitors]
        interval = dft_sample_interval(self.dt, self.sources, spectral)
        for monitor in self.monitors:
            monitor.sample_interval = interval
        self.sample_interval = interval
        self._finalized = True

    # ---- stepping ---------------------------------------------------------------------------------------------
    def _hi_difference(self, f, axis):
        """F[hi neighbour] - F (the last row reads row 0: the wrap, or x 0 through inv_h on non-periodic axes)."""
        return np.roll(f, -1, axis=axis) - f

    def _lo_difference(self, f, axis):
        """F - F[lo neighbour]; row 0 reads the wrap on a periodic axis, 0 otherwise."""
        out = f - np.roll(f, 1, axis=axis)
        if not self.periodic[axis]:
            out[along(axis, 0)] = f[along(axis, 0)]
        return out

    def step(self):
        """One step (engine3d.py:497-554): H with its CPML, E with its CPML, dipoles at t + dt/2, monitors at t + dt."""
        if not self._finalized:
            self.finalize()
        F = self.fields
        ex, ey, ez, hx, hy, hz = (F[n] for n in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'))
        ihx, ihy, ihz = (v.reshape(row_shape(a)) for a, v in enumerate(self.inv_h))
        iex, iey, iez = (v.reshape(row_shape(a)) for a, v in enumerate(self.inv_e))
        da, db = self.da, self.db
        # H (kernels3d.py:335-400)
        dez_y, dey_z = self._hi_difference(ez, 1), self._hi_difference(ey, 2)
        dex_z, dez_x = self._hi_difference(ex, 2), self._hi_difference(ez, 0)
        dey_x, dex_y = self._hi_difference(ey, 0), self._hi_difference(ex, 1)
        hx[...] = da * hx - db * (dez_y * ihy - dey_z * ihz)
        hy[...] = da * hy - db * (dex_z * ihz - dez_x * ihx)
        hz[...] = da * hz - db * (dey_x * ihx - dex_y * ihy)
        h_rows = {0: ((hy, dez_x, 1.0), (hz, dey_x, -1.0)), 1: ((hx, dez_y, -1.0), (hz, dex_y, 1.0)),
                  2: ((hx, dey_z, 1.0), (hy, dex_z, -1.0))}
        for layer in self.cpml:
            if layer.kind != 'h':
                continue
            for which, (target, difference, sign) in enumerate(h_rows[layer.axis]):
                layer.add(target, db, sign, layer.step(which, difference))
        # symmetric mirrors: the tangential H on the plane face is pinned to 0 (engine3d.py:114-126, 513-518)
        mirrors = self.bc.mirror_planes(self.windows)
        for axis, kind, _, _, plane in mirrors:
            if kind == 'symmetric':
                for name in AXES:
                    if name != AXES[axis]:
                        F['h' + name][along(axis, plane)] = 0
        # pole J / P from E^n (engine3d.py:508-522)
        for entry in self.poles:
            for c, field in enumerate((ex, ey, ez)):
                state, drive = entry['state'][c], entry['drive'][c]
                if entry['pole'][0] == 'drude':
                    state[0] += entry['coefficients'][0] * state[0] + drive * field
                else:
                    damping, restoring, _ = entry['coefficients']
                    state[1] += damping * state[1] - restoring * state[0] + drive * field
                    state[0] += state[1]
        # E (kernels3d.py:538-610)
        dhz_y, dhy_z = self._lo_difference(hz, 1), self._lo_difference(hy, 2)
        dhx_z, dhz_x = self._lo_difference(hx, 2), self._lo_difference(hz, 0)
        dhy_x, dhx_y = self._lo_difference(hy, 0), self._lo_difference(hx, 1)
        (cax, cay, caz), (cbx, cby, cbz) = self.ca, self.cb
        ex[...] = cax * ex + cbx * (dhz_y * iey - dhy_z * iez)
        ey[...] = cay * ey + cby * (dhx_z * iez - dhz_x * iex)
        ez[...] = caz * ez + cbz * (dhy_x * iex - dhx_y * iey)
        e_rows = {0: ((ey, cby, dhz_x, -1.0), (ez, cbz, dhy_x, 1.0)), 1: ((ex, cbx, dhz_y, 1.0), (ez, cbz, dhx_y, -1.0)),
                  2: ((ex, cbx, dhy_z, -1.0), (ey, cby, dhx_z, 1.0))}
        for layer in self.cpml:
            if layer.kind != 'e':
                continue
            for which, (target, coefficient, difference, sign) in enumerate(e_rows[layer.axis]):
                layer.add(target, coefficient, sign, layer.step(which, difference))
        # pole corrections (engine3d.py:524-530): Drude E -= cb J, Lorentz E -= cb / dt dP, before the sources
        inv_dt = self.dtype(1.0) / self.dtype(self.dt)
        for entry in self.poles:
            for c, (field, cb) in enumerate(zip((ex, ey, ez), (cbx, cby, cbz))):
                state = entry['state'][c]
                if entry['pole'][0] == 'drude':
                    field -= cb * state[0]
                else:
                    field -= cb * inv_dt * state[1]
        t = self.time
        for source in self.sources:
            source.inject(F, t + 0.5 * float(self.dt), self.dtype)
        # mirror E rows after the sources (engine3d.py:531-540): symmetric, the normal E on the plane face = 0;
        # antisymmetric, the ghost's tangential E = minus its inner neighbour's
        for axis, kind, ghost, inner, plane in mirrors:
            if kind == 'symmetric':
                F['e' + AXES[axis]][along(axis, plane)] = 0
            else:
                for name in AXES:
                    if name != AXES[axis]:
                        component = F['e' + name]
                        component[along(axis, ghost)] = -component[along(axis, inner)]
        self.time = t + float(self.dt)
        self.n_step += 1
        for probe in self.probes:
            probe.record(F, self.time)
        for monitor in self.monitors:
            if self.n_step % monitor.sample_interval == 0:
                monitor.update(F, self.time)

    def run(self, duration=None, n_steps=None):
        """Tick loop stopping at t >= t_end - dt/2 (coupling/runner.py:174-204); returns the steps run."""
        if (duration is None) == (n_steps is None):
            raise ValueError('pass exactly one of duration or n_steps')
        if not self._finalized:
            self.finalize()
        if n_steps is not None:
            for _ in range(int(n_steps)):
                self.step()
            return int(n_steps)
        stop = self.time + float(duration) - 0.5 * float(self.dt)
        done = 0
        while self.time < stop:
            self.step()
            done += 1
        return done
