"""ORACLE (test infrastructure only): numpy restatement of the reference's 2D FDTD hot path.

E_z (E_z, H_x, H_y) and H_z (H_z, E_x, E_y) Yee steppers with CPML psi layers, periodic wraps, calibrated sheet
(line) sources and running-DFT flux monitors, with the reference's geometry, clock and quantisation points:

  * dt and the kernel cell width are float32 values (core/grid.py:9-17, physics/em/yee_grid.py:11-16);
  * material arrays are float32 (core/material_grid.py:38-46), rectangle masks compare float32 cell centres
    (core/domain.py:90-97, 255-260);
  * the sheet weight and calibrated profile are float32 (physics/em/__init__.py:459-473, em/sources.py:383-392);
  * the injected value is float32(amplitude * s(t + dt/2)) times the float32 profile (core/sources.py:146-155,
    221-229);
  * DFT phases are complex64 by default (em/monitors.py:22-34), spectra complex128 (em/monitors.py:86-90).

dtype=np.float64 restates the reference's precision='double' stepper (physics/em/__init__.py:60-86); dtype=np.float32
runs the same array passes as the reference's numpy twins in single precision (the CPU baseline). Everything else is
float64 host arithmetic.
"""
import math

import numpy as np

# core/constants.py:6-9
C0 = 299792458.0
EPS0 = 8.8541878128e-12
MU0 = 1.2566370614e-6
ETA0 = float(np.sqrt(MU0 / EPS0))
F32 = np.float32
FACES = ('x_lo', 'x_hi', 'y_lo', 'y_hi')


# ---------------------------------------------------------------------------------------------------------- grid

def courant_time_step(dx_min, dy_min, courant=None):
    """float32 dt = courant / (C0 sqrt(1/dx^2 + 1/dy^2)) (core/grid.py:9-17)."""
    courant = 0.99 if courant is None else float(courant)
    if not 0.0 < courant <= 1.0:
        raise ValueError(f'courant must be in (0, 1], got {courant}')
    return F32(courant / (C0 * np.sqrt(1.0 / float(dx_min) ** 2 + 1.0 / float(dy_min) ** 2)))


def uniform_cell_count(length, dx):
    """ceil(length / dx - 1e-9) (core/grid.py:20-22)."""
    return int(np.ceil(length / dx - 1e-9))


def containing_cell(coord, edges, first, last):
    """Cell with edges[i] <= coord < edges[i+1] after a 1e-9-cell snap, clamped (core/domain.py:275-282)."""
    coord = float(coord)
    snap = 1e-9 * (edges[first + 1] - edges[first])
    index = int(np.searchsorted(edges, coord + snap, side='right')) - 1
    return min(max(index, first), last)


class Boundaries:
    """Face types and CPML settings (physics/em/boundaries.py:194-356): cpml faces are padded with npml cells."""

    def __init__(self, faces, npml=15, order=3, sigma_factor=1.0, kappa_max=15.0, alpha_max=0.0):
        self.faces = dict(faces)
        for face, kind in self.faces.items():
            if kind not in ('cpml', 'periodic', 'pec', 'pmc', 'symmetric', 'antisymmetric'):
                raise ValueError(f'oracle supports cpml, periodic, pec, pmc and mirror faces, got {face}={kind}')
        for axis in 'xy':
            if (self.faces[f'{axis}_lo'] == 'periodic') != (self.faces[f'{axis}_hi'] == 'periodic'):
                raise ValueError(f'{axis}: periodic faces must be paired')
        self.npml, self.order, self.sigma_factor = int(npml), int(order), float(sigma_factor)
        self.kappa_max, self.alpha_max = float(kappa_max), float(alpha_max)

    @classmethod
    def preset(cls, name, npml):
        """Simulation boundaries presets (simulation.py:388-400)."""
        presets = {None: ('cpml',) * 4, 'pml': ('cpml',) * 4,
                   'periodic_y': ('cpml', 'cpml', 'periodic', 'periodic'),
                   'periodic_x': ('periodic', 'periodic', 'cpml', 'cpml'),
                   'periodic': ('periodic',) * 4}
        if name not in presets:
            raise ValueError(f'unknown boundary preset {name!r}')
        return cls(dict(zip(FACES, presets[name])), npml=npml)

    def cpml(self, face):
        return self.faces[face] == 'cpml'

    def periodic(self, axis):
        return self.faces[f'{axis}_lo'] == 'periodic'

    def npml_sides(self):
        """cpml faces: npml cells; mirror faces: one ghost cell (physics/em/boundaries.py:283-286); walls: none."""
        return {face: (self.npml if self.cpml(face) else 1 if self.faces[face] in ('symmetric', 'antisymmetric')
                       else 0) for face in FACES}

    def wall_sign(self, face):
        """E_z outer-row copy factor of a face: +1 pmc / symmetric (E_z even), -1 antisymmetric, None otherwise
        (physics/em/boundaries.py:478-488)."""
        return {'pmc': 1.0, 'symmetric': 1.0, 'antisymmetric': -1.0}.get(self.faces[face])


class Geometry2D:
    """Uniform window + CPML padding (core/domain.py:180-196): float64 widths/edges/centres, float32 dt and dx."""

    def __init__(self, size_x, size_y, dx, boundaries, courant=None):
        sides = boundaries.npml_sides()
        self.sides = sides
        self.size_x, self.size_y, self.dx = float(size_x), float(size_y), float(dx)
        self.nx = uniform_cell_count(size_x + (sides['x_lo'] + sides['x_hi']) * dx, dx)
        self.ny = uniform_cell_count(size_y + (sides['y_lo'] + sides['y_hi']) * dx, dx)
        self.dx32 = F32(dx)
        self.dt = courant_time_step(dx, dx, courant)
        self.courant = 0.99 if courant is None else float(courant)
        self.origins = (sides['x_lo'], sides['y_lo'])
        self.shape = (self.nx, self.ny)
        self.windows = (slice(sides['x_lo'], self.nx - sides['x_hi']), slice(sides['y_lo'], self.ny - sides['y_hi']))
        # float64 geometry (core/domain.py:54-81, core/grid.py:25-31)
        self.widths = tuple(np.full(n, self.dx) for n in self.shape)
        self.edges = tuple((np.arange(n + 1) - o) * self.dx for n, o in zip(self.shape, self.origins))
        self.centres = tuple(0.5 * (e[:-1] + e[1:]) for e in self.edges)

    def validate(self, x, y, label):
        """core/domain.py:126-133."""
        tol = 1e-9 * max(self.size_x, self.size_y)
        if not (-tol <= x <= self.size_x + tol and -tol <= y <= self.size_y + tol):
            raise ValueError(f'{label} position ({x}, {y}) outside the physical domain')

    def cell(self, axis, coord):
        """core/domain.py:120-124."""
        window = self.windows[axis]
        return containing_cell(coord, self.edges[axis], window.start, window.stop - 1)

    def rect_axes(self, x0, y0, x1, y1):
        """The two axis factors of rect_mask: (cells along x, cells along y) whose float32 centre lies in the range."""
        cx, cy = (c.astype(F32) for c in self.centres)
        x_min, x_max = F32(min(x0, x1)), F32(max(x0, x1))
        y_min, y_max = F32(min(y0, y1)), F32(max(y0, y1))
        return (cx >= x_min) & (cx < x_max), (cy >= y_min) & (cy < y_max)

    def rect_mask(self, x0, y0, x1, y1):
        """Cells whose float32 centre lies in [x_min, x_max) x [y_min, y_max) (core/domain.py:255-260)."""
        mx, my = self.rect_axes(x0, y0, x1, y1)
        return mx[:, None] & my[None, :]


# ----------------------------------------------------------------------------------------------------------- CPML

def cpml_side_rows(side, n, npml):
    """(E rows, H rows) of one side (physics/em/boundaries.py:570-576)."""
    if side == 'lo':
        return slice(1, npml + 1), slice(0, npml)
    return slice(n - npml - 1, n - 1), slice(n - npml - 1, n - 1)


def cpml_axis_profiles(bc, axis, n, pml_dx):
    """float64 (sigma, kappa, alpha) at E rows (centres) and H rows (hi faces) (physics/em/boundaries.py:551-631)."""
    lo, hi = bc.cpml(f'{axis}_lo'), bc.cpml(f'{axis}_hi')
    npml = bc.npml
    sigma_e, kappa_e, alpha_e = np.zeros(n), np.ones(n), np.zeros(n)
    sigma_h, kappa_h, alpha_h = np.zeros(n), np.ones(n), np.zeros(n)
    if not (lo or hi):
        return (sigma_e, kappa_e, alpha_e), (sigma_h, kappa_h, alpha_h)
    if n < (int(lo) + int(hi)) * npml + 2:
        raise ValueError(f'{axis} axis: {n} cells cannot hold the CPML layers')
    sigma_max = bc.sigma_factor * (bc.order + 1) / (150.0 * np.pi * float(pml_dx))

    def graded(depth):
        return (sigma_max * depth ** bc.order, 1.0 + (bc.kappa_max - 1.0) * depth ** bc.order,
                bc.alpha_max * (1.0 - depth))

    if lo:
        e_rows, h_rows = cpml_side_rows('lo', n, npml)
        sigma_e[e_rows], kappa_e[e_rows], alpha_e[e_rows] = graded(np.arange(npml - 1, -1, -1.0) / npml)
        sigma_h[h_rows], kappa_h[h_rows], alpha_h[h_rows] = graded((np.arange(npml, 0, -1.0) - 0.5) / npml)
    if hi:
        e_rows, h_rows = cpml_side_rows('hi', n, npml)
        sigma_e[e_rows], kappa_e[e_rows], alpha_e[e_rows] = graded(np.arange(0, npml, 1.0) / npml)
        sigma_h[h_rows], kappa_h[h_rows], alpha_h[h_rows] = graded((np.arange(1, npml + 1, 1.0) - 0.5) / npml)
    return (sigma_e, kappa_e, alpha_e), (sigma_h, kappa_h, alpha_h)


def cpml_recursion(sigma, kappa, alpha, dt, dtype):
    """(a, b): b = exp(-(sigma/kappa + alpha) dt/eps0), a = sigma (b - 1) / (sigma kappa + kappa^2 alpha)
    (physics/em/boundaries.py:560-567)."""
    b = np.exp(-(sigma / kappa + alpha) * float(dt) / EPS0).astype(dtype)
    denom = sigma * kappa + kappa ** 2 * alpha
    a = np.divide(sigma * (b - 1.0), denom, out=np.zeros_like(denom), where=denom > 1e-30).astype(dtype)
    return a, b


class _Side:
    """One active CPML side (physics/em/boundaries.py:671-730): per row kind its b, a, bulk inverse spacing, rows and
    psi slab."""

    def __init__(self, axis, side, sign, profiles, inv_bulk, dt, n, npml, other_n, dtype):
        self.axis, self.side, self.sign = axis, side, dtype(sign)
        self.rows = {}
        e_rows, h_rows = cpml_side_rows(side, n, npml)
        for kind, rows, profile in (('e', e_rows, profiles[0]), ('h', h_rows, profiles[1])):
            sigma, kappa, alpha = (p[rows] for p in profile)
            a, b = cpml_recursion(sigma, kappa, alpha, dt, dtype)
            inv = np.asarray(inv_bulk[kind][rows], dtype=dtype)
            shape = (npml, other_n) if axis == 'x' else (other_n, npml)
            bcast = (lambda v: v.reshape(-1, 1)) if axis == 'x' else (lambda v: v.reshape(1, -1))
            self.rows[kind] = dict(rows=rows, a=bcast(a), b=bcast(b), inv=bcast(inv), kappa=kappa.astype(dtype),
                                   psi=np.zeros(shape, dtype=dtype))

    def view(self, array, rows):
        return array[rows, :] if self.axis == 'x' else array[:, rows]

    def apply(self, kind, source, coefficient, field):
        """psi = b psi + a (f_hi - f_lo) inv; field += sign coefficient psi on the rows of `kind`
        (physics/em/boundaries.py:27-37, 732-750): 'e' rows difference r - (r-1), 'h' rows (r+1) - r."""
        rs = self.rows[kind]
        rows = rs['rows']
        shift = -1 if kind == 'e' else 1
        other = slice(rows.start + shift, rows.stop + shift)
        upper, lower = (rows, other) if kind == 'e' else (other, rows)
        psi = rs['psi']
        psi *= rs['b']
        psi += rs['a'] * (self.view(source, upper) - self.view(source, lower)) * rs['inv']
        target = self.view(field, rows)
        target += self.sign * self.view(coefficient, rows) * psi


# -------------------------------------------------------------------------------------------------------- sources

class GaussianPulse:
    """exp(-((t-t0)/tau)^2) sin(2 pi f0 (t-t0)), tau = 1/(bw f0 pi), t0 = 4.5 tau (core/sources.py:23-53)."""

    def __init__(self, f0, bandwidth=0.5, t0_factor=4.5):
        self.f0, self.bandwidth = float(f0), float(bandwidth)
        self.tau = 1.0 / (self.bandwidth * self.f0 * np.pi)
        self.t0 = t0_factor * self.tau

    def __call__(self, t):
        delay = float(t) - self.t0
        return math.exp(-(delay / self.tau) ** 2) * math.sin(2.0 * math.pi * self.f0 * delay)

    def spectral_limit(self):
        return self.f0 * (1.0 + 3.72 * self.bandwidth)


class ContinuousWave:
    """(1 - exp(-t/tau_ramp)) sin(2 pi f0 t), tau_ramp = 3/f0 (core/sources.py:56-78)."""

    def __init__(self, f0, ramp_periods=3.0):
        self.f0 = float(f0)
        self.tau_ramp = ramp_periods / self.f0

    def __call__(self, t):
        t = float(t)
        return (1.0 - math.exp(-t / self.tau_ramp)) * math.sin(2.0 * math.pi * self.f0 * t)

    def spectral_limit(self):
        return None


class RickerWavelet:
    """(1 - 2 a^2) exp(-a^2), a = pi f (t - t0), t0 = 4/f (core/sources.py:81-100)."""

    def __init__(self, f_peak, t0_factor=4.0):
        self.f_peak = float(f_peak)
        self.f0 = self.f_peak
        self.t0 = t0_factor / self.f_peak

    def __call__(self, t):
        arg = math.pi * self.f_peak * (float(t) - self.t0)
        return (1.0 - 2.0 * arg * arg) * math.exp(-arg * arg)

    def spectral_limit(self):
        return 4.2 * self.f_peak


def make_temporal(waveform, frequency, bandwidth):
    """physics/em/__init__.py:580-587."""
    if waveform == 'gaussian':
        return GaussianPulse(frequency, bandwidth)
    if waveform == 'cw':
        return ContinuousWave(frequency)
    if waveform == 'ricker':
        return RickerWavelet(frequency)
    raise ValueError(f'unknown waveform {waveform!r}')


def resolve_frequency(wavelength, frequency):
    """physics/em/__init__.py:560-568."""
    if (wavelength is None) == (frequency is None):
        raise ValueError('pass exactly one of wavelength or frequency')
    value = float(frequency if frequency is not None else wavelength)
    return value if frequency is not None else C0 / value


def sheet_admittance(line_eps, mu_r, sigma_m, omega, normal_spacing, dt):
    """float32 |sqrt(eps/mu) sqrt(1 - eps mu (w sin(omega dt/2)/(C0 dt))^2)| (em/sources.py:430-436,
    physics/em/yee_grid.py:58-65)."""
    permeability = mu_r.astype(np.complex128) - 1j * sigma_m.astype(np.float64) / (omega * MU0)
    half_phase = float(normal_spacing) * math.sin(0.5 * omega * float(dt)) / (C0 * float(dt))
    launch = np.sqrt(1.0 - np.asarray(line_eps * permeability, dtype=np.complex128) * half_phase ** 2)
    return np.abs(np.sqrt(line_eps / permeability) * launch).astype(F32)


class LineSource:
    """Calibrated sheet current (physics/em/__init__.py:442-473; core/sources.py:190-229)."""

    def __init__(self, axis, line_cell, span_cells, temporal, amplitude, normal_width, component, shape=None):
        self.axis, self.position, self.span = axis, int(line_cell), span_cells
        self.temporal, self.amplitude = temporal, F32(amplitude)
        self.normal_width = float(normal_width)
        self.component = component
        n_line = span_cells[1] - span_cells[0]
        shape = np.ones(n_line) if shape is None else shape   # Gaussian beam: exp(-((s - centre) / waist)^2)
        self.weight = (2.0 * shape / (ETA0 * self.normal_width)).astype(F32)
        self.profile = np.zeros(n_line, dtype=F32)

    def line(self, array):
        lo, hi = self.span
        return array[lo:hi, self.position] if self.axis == 'x' else array[self.position, lo:hi]

    def value(self, t):
        """float32(amplitude * s(t)) with float32 arithmetic (core/sources.py:221-229)."""
        return F32(self.amplitude * F32(self.temporal(t)))

    def inject(self, fields, t):
        value = self.value(t)
        if value == 0:
            return
        target = self.line(fields[self.component])
        target += value * self.profile


class CellCurrent:
    """Soft current at one cell or face (core/sources.py:162-188 PointSource): weight = coefficient[i, j] times the
    float32 current density in numpy's precision rules (physics/em/sources.py:544-561, 605-616), and each step
    field[i, j] += weight * float32(s(t)). `field` 'e' sources are injected at t + dt/2 after the E update, 'h'
    sources (the H_z magnetic line current, physics/em/sources.py:439-444) at t after the H update."""

    def __init__(self, component, cell, temporal, current_density, coefficient, field):
        self.component, self.cell, self.temporal = component, (int(cell[0]), int(cell[1])), temporal
        self.current_density = F32(current_density)
        self.coefficient, self.field = coefficient, field
        self.weight = np.zeros(1, dtype=F32)

    def calibrate(self, coefficients):
        i, j = self.cell
        np.multiply(coefficients[i:i + 1, j], self.current_density, out=self.weight, casting='same_kind')

    def inject(self, fields, t):
        value = F32(F32(1.0) * F32(self.temporal(t)))
        if value == 0:
            return
        i, j = self.cell
        fields[self.component][i:i + 1, j] += self.weight * value


def aux_line_terms(width, dt, half_omega_dt):
    """(Courant number c dt / width, cos(kappa0 / 2), Mur coefficient) of a uniform vacuum 1D line at the carrier
    (physics/em/sources.py:189-204): sin(kappa0/2) = sin(omega0 dt/2) / S, Mur S' = S cos(kappa0/2) / cos(omega0 dt/2)."""
    courant = C0 * float(dt) / float(width)
    sin_half_kappa = math.sin(half_omega_dt) / courant
    if sin_half_kappa >= 1.0:
        raise ValueError('TFSF: the centre frequency has no propagating wave on the 1D grid')
    cos_half_kappa = math.sqrt(1.0 - sin_half_kappa ** 2)
    mur_courant = courant * cos_half_kappa / math.cos(half_omega_dt)
    return courant, cos_half_kappa, (mur_courant - 1.0) / (mur_courant + 1.0)


class TFSFBox:
    """+x vacuum plane wave on the total-field box [x0, x1) x [y0, y1) of a uniform grid (physics/em/sources.py:207-378,
    _PlaneWaveBox, TFSFSource, TFSFSourceHz). A 1D incident line (20 margin cells, the box columns, 20 margin cells;
    the soft source at aux cell 10, first-order Mur ends matched to the carrier's numerical phase velocity) runs beside
    the 2D grid; after the H update its incident E (time n) corrects the H faces on the box edges and the line's H
    advances; after the E update its incident H (time n + 1/2) corrects the E cells / faces on the box edges and the
    line's E advances with the source value float(2 S cos(kappa0/2) amplitude s(t + dt/2)). E_z: incident (E_z, H_y);
    H_z: incident (E_y, H_z = E_y / eta0). Side coefficients dt / (mu0 or eps0) / dx, 0 on an open (mirror) side."""

    MARGIN, SOURCE_INDEX = 20, 10

    def __init__(self, cells, temporal, dx, dt, amplitude, polarization, open_faces, dtype):
        x0, x1, y0, y1 = (int(v) for v in cells)
        if not (1 <= x0 < x1 and 1 <= y0 < y1):
            raise ValueError(f'TFSF box [{x0}, {x1}) x [{y0}, {y1}) must keep a scattered-field cell on the lo sides')
        self.x0, self.x1, self.y0, self.y1 = x0, x1, y0, y1
        self.cells = (x0, x1, y0, y1)
        self.temporal, self.amplitude = temporal, float(amplitude)
        self.polarization, self.open_faces = polarization, frozenset(open_faces)
        self.real = real = dtype
        self.field = 'tfsf'
        dt = float(dt)
        margin = self.MARGIN
        widths = np.full(2 * margin + (x1 - x0), float(dx))
        face_spacings = 0.5 * (widths[:-1] + widths[1:])
        f0 = temporal.f_peak if isinstance(temporal, RickerWavelet) else temporal.f0
        courant, cos_half_kappa, mur = aux_line_terms(float(dx), dt, 0.5 * 2.0 * np.pi * float(f0) * dt)
        self.source_scale = 2.0 * courant * cos_half_kappa * self.amplitude
        self.mur_lo = self.mur_hi = real(mur)
        self.aux_x0, self.n = margin, widths.size

        def sides(constant, spacings):
            names = ('x_lo', 'x_hi', 'y_lo', 'y_hi')[:len(spacings)]
            return tuple(real(0.0 if name in self.open_faces else dt / (constant * spacing))
                         for name, spacing in zip(names, spacings))
        side_spacings = (float(dx),) * 4
        side_widths = (float(dx),) * 2
        if polarization == 'ez':
            self.aux_e = np.zeros(self.n, dtype=real)   # E_z at the aux cell centres
            self.aux_h = np.zeros(self.n, dtype=real)   # H_y on the aux hi faces
            self.coef_h = sides(MU0, side_spacings)     # H_y rows x0-1 / x1-1, H_x columns y0-1 / y1-1
            self.coef_e = sides(EPS0, side_widths)      # E_z rows x0 / x1-1
            self.aux_coef_h = (dt / (MU0 * face_spacings)).astype(real)
            self.aux_coef_e = (dt / (EPS0 * widths)).astype(real)
        else:
            self.aux_h = np.zeros(self.n, dtype=real)   # H_z at the aux centres (index 0 unused)
            self.aux_e = np.zeros(self.n, dtype=real)   # E_y on the aux hi faces: the soft source and Mur ends
            self.coef_h = sides(MU0, side_widths)       # H_z rows x0 / x1-1
            self.coef_e = sides(EPS0, side_spacings)    # E_y rows x0-1 / x1-1, E_x columns y0-1 / y1-1
            self.aux_coef_h = (-dt / (MU0 * widths[1:])).astype(real)
            self.aux_coef_e = (-dt / (EPS0 * np.append(face_spacings, widths[-1]))).astype(real)

    def _aux_e_update(self, e, h, value):
        """Interior E of the line, the soft source, Mur ends (the tfsf_aux_e_graded kernel, physics/em/sources.py:
        163-187): e[k] += ce[k] (h[k] - h[k-1]) for 1 <= k <= n-2."""
        n = self.n
        old = (e[0], e[1], e[n - 2], e[n - 1])
        e[1:n - 1] += self.aux_coef_e[1:n - 1] * (h[1:n - 1] - h[0:n - 2])
        e[self.SOURCE_INDEX] += value
        e[0] = old[1] + self.mur_lo * (e[1] - old[0])
        e[n - 1] = old[2] + self.mur_hi * (e[n - 2] - old[3])

    def inject_h(self, fields):
        """After the H update: H-side corrections from the incident E at time n, then the line's H update."""
        x0, x1, y0, y1, a0 = self.x0, self.x1, self.y0, self.y1, self.aux_x0
        last = a0 + (x1 - 1 - x0)
        if self.polarization == 'ez':
            hx, hy, inc = fields['hx'], fields['hy'], self.aux_e
            c_xlo, c_xhi, c_ylo, c_yhi = self.coef_h
            hy[x0 - 1, y0:y1] -= c_xlo * inc[a0]
            hy[x1 - 1, y0:y1] += c_xhi * inc[last]
            incident = inc[a0:a0 + (x1 - x0)]
            hx[x0:x1, y0 - 1] += c_ylo * incident
            hx[x0:x1, y1 - 1] -= c_yhi * incident
            self.aux_h[:-1] += self.aux_coef_h * (self.aux_e[1:] - self.aux_e[:-1])
        else:
            hz, inc = fields['hz'], self.aux_e
            c_xlo, c_xhi = self.coef_h
            hz[x0, y0:y1] += c_xlo * inc[a0 - 1]
            hz[x1 - 1, y0:y1] -= c_xhi * inc[last]
            self.aux_h[1:] += self.aux_coef_h * (self.aux_e[1:] - self.aux_e[:-1])

    def inject_e(self, fields, t, keep):
        """After the E update: E-side corrections from the incident H at n + 1/2 (0 in PEC cells / faces, the keep
        re-pin), then the line's E update with the source at t."""
        x0, x1, y0, y1, a0 = self.x0, self.x1, self.y0, self.y1, self.aux_x0
        last = a0 + (x1 - 1 - x0)
        value = self.real(self.source_scale * self.temporal(t))
        if self.polarization == 'ez':
            ez, inc = fields['ez'], self.aux_h
            c_xlo, c_xhi = self.coef_e
            ez[x0, y0:y1] -= c_xlo * inc[a0 - 1]
            ez[x1 - 1, y0:y1] += c_xhi * inc[last]
            if keep is not None:
                ez[x0, y0:y1] *= keep[x0, y0:y1]
                ez[x1 - 1, y0:y1] *= keep[x1 - 1, y0:y1]
            self._aux_e_update(self.aux_e, self.aux_h, value)
        else:
            ex, ey, inc = fields['ex'], fields['ey'], self.aux_h
            c_xlo, c_xhi, c_ylo, c_yhi = self.coef_e
            ey[x0 - 1, y0:y1] += c_xlo * inc[a0]
            ey[x1 - 1, y0:y1] -= c_xhi * inc[last]
            incident = inc[a0:a0 + (x1 - x0)]
            ex[x0:x1, y0 - 1] -= c_ylo * incident
            ex[x0:x1, y1 - 1] += c_yhi * incident
            if keep is not None:
                keep_x, keep_y = keep
                for field, k, rows in ((ey, keep_y, (x0 - 1, x1 - 1)),):
                    for row in rows:
                        field[row, y0:y1] *= k[row, y0:y1]
                for column in (y0 - 1, y1 - 1):
                    ex[x0:x1, column] *= keep_x[x0:x1, column]
            self._aux_e_update(self.aux_e, self.aux_h[1:], value)


def dipole_face_shares(cell, n_cells, periodic):
    """Faces carrying an in-plane dipole along one axis (physics/em/placement.py:268-276): the hi face `cell` and the
    lo face `cell - 1`, half each; on a bounded axis only faces 0..n-2 exist and the survivors share the moment."""
    candidates = [cell, cell - 1]
    if periodic:
        return [(candidates[0], 0.5), (candidates[1] % n_cells, 0.5)]
    kept = [face for face in candidates if 0 <= face <= n_cells - 2]
    if not kept:
        raise ValueError('an electric dipole needs at least 2 cells along a bounded axis')
    return [(face, 1.0 / len(kept)) for face in kept]


class FluxMonitor:
    """Flux line with complex128 E and H spectra phased at the E sample time (em/monitors.py:203-245, 325-379)."""

    def __init__(self, normal, position, span, frequencies, dt, widths, polarization, phase_dtype):
        self.normal, self.position, self.span = normal, int(position), (int(span[0]), int(span[1]))
        self.frequencies = np.asarray(frequencies, dtype=np.float64)
        self.dt = float(dt)
        if polarization == 'ez':
            self.e_component, self.h_component = 'ez', ('hy' if normal == 'x' else 'hx')
            self.sign = -1.0 if normal == 'x' else 1.0
        else:
            self.e_component, self.h_component = ('ey' if normal == 'x' else 'ex'), 'hz'
            self.sign = 1.0 if normal == 'x' else -1.0
        n = self.span[1] - self.span[0]
        self.e_spectrum = np.zeros((self.frequencies.size, n), dtype=np.complex128)
        self.h_spectrum = np.zeros((self.frequencies.size, n), dtype=np.complex128)
        self.widths = np.asarray(widths, dtype=np.float64)[self.span[0]:self.span[1]].copy()
        self.omega = 2.0 * np.pi * self.frequencies
        self.phase_dtype = phase_dtype
        self.sample_interval = 1

    def line(self, array):
        lo, hi = self.span
        return array[self.position, lo:hi] if self.normal == 'x' else array[lo:hi, self.position]

    def phases(self, t):
        """exp(-j omega t) m dt rounded to phase_dtype (em/monitors.py:22-34)."""
        p = -self.omega * float(t)
        scale = self.dt * self.sample_interval
        if self.phase_dtype == np.complex64:
            return ((np.cos(p) * scale).astype(F32) + 1j * (np.sin(p) * scale).astype(F32)).astype(np.complex128)
        return np.cos(p) * scale + 1j * (np.sin(p) * scale)

    def update(self, fields, t):
        ph = self.phases(t)[:, None]
        self.e_spectrum += ph * self.line(fields[self.e_component]).astype(np.float64)
        self.h_spectrum += ph * self.line(fields[self.h_component]).astype(np.float64)

    def flux(self):
        """sign cos(pi f dt) Re(E conj(H e^{j pi f dt})) summed with the cell widths (em/monitors.py:371-379)."""
        h = self.h_spectrum * np.exp(1j * np.pi * self.frequencies * self.dt)[:, None]
        integrand = self.sign * np.cos(np.pi * self.frequencies * self.dt)[:, None] * np.real(
            self.e_spectrum * np.conj(h))
        return integrand @ self.widths


def centre_partners(indices, n, periodic):
    """Lower-face partner of each sampled index: index - 1, and for index 0 the wrapped last face (periodic) or the
    face itself (em/monitors.py:158-169, yee_grid.lower_face)."""
    partners = np.asarray(indices) - 1
    partners[np.asarray(indices) == 0] = n - 1 if periodic else 0
    return partners


class DFTBox:
    """Running DFT over a cell box (em/monitors.py:248-322): E_z, or H_z (phased at t - dt/2) with E_x, E_y averaged
    to the cell centres; every downsample-th cell from the middle of the first block (:152-155). Spectra accumulate
    in complex64, as the reference's, when phase_dtype is complex64, else in complex128."""

    def __init__(self, x_cells, y_cells, frequencies, dt, polarization, periodic, downsample, shape, phase_dtype):
        (x0, x1), (y0, y1) = x_cells, y_cells
        s = int(downsample)
        self.x_samples = slice(x0 + min(s, x1 - x0) // 2, x1, s)
        self.y_samples = slice(y0 + min(s, y1 - y0) // 2, y1, s)
        self.frequencies = np.asarray(frequencies, dtype=np.float64)
        self.omega = 2.0 * np.pi * self.frequencies
        self.dt = float(dt)
        self.polarization = polarization
        self.phase_dtype = phase_dtype
        self.sample_interval = 1
        xs = np.arange(shape[0])[self.x_samples]
        ys = np.arange(shape[1])[self.y_samples]
        self.xs, self.ys = xs, ys
        self.x_partners = centre_partners(xs, shape[0], periodic[0])
        self.y_partners = centre_partners(ys, shape[1], periodic[1])
        names = ('ez',) if polarization == 'ez' else ('hz', 'ex', 'ey')
        acc_dtype = np.complex64 if phase_dtype == np.complex64 else np.complex128
        self.spectra = {name: np.zeros((self.frequencies.size, xs.size, ys.size), dtype=acc_dtype) for name in names}

    phases = FluxMonitor.phases

    def update(self, fields, t):
        ph = self.phases(t)[:, None, None]
        box = np.ix_(self.xs, self.ys)
        half = F32(0.5)
        if self.polarization == 'ez':
            self.spectra['ez'] += ph * fields['ez'][box]
            return
        self.spectra['hz'] += self.phases(t - 0.5 * self.dt)[:, None, None] * fields['hz'][box]
        ex, ey = fields['ex'], fields['ey']
        self.spectra['ex'] += ph * (half * (ex[box] + ex[np.ix_(self.xs, self.y_partners)]))
        self.spectra['ey'] += ph * (half * (ey[box] + ey[np.ix_(self.x_partners, self.ys)]))


def drude_coefficients(omega_p, gamma, dt):
    """(damping, drive) float32 of J += damping J + drive E^n (physics/em/materials.py:132-149)."""
    dt, gamma, wp = float(dt), float(F32(gamma)), float(F32(omega_p))
    denom = 1.0 + gamma * dt / 2.0
    return F32(-gamma * dt / denom), F32((EPS0 * wp ** 2 * dt) / denom)


def lorentz_coefficients(delta_eps, omega_0, gamma, dt):
    """(damping, restoring, drive) float32 of dP += damping dP - restoring P + drive E^n, P += dP
    (physics/em/materials.py:160-180)."""
    dt, w0, g, de = float(dt), float(F32(omega_0)), float(F32(gamma)), float(F32(delta_eps))
    denom = 1.0 + g * dt / 2.0
    return F32(-g * dt / denom), F32(w0 ** 2 * dt ** 2 / denom), F32(EPS0 * de * w0 ** 2 * dt ** 2 / denom)


class PointProbe:
    """Raw Yee samples of one component at one cell after every step, stamped with the time after the step
    (core/monitors.py:194-283; engine3d.py:544-548 records the probe rows after the step and stamps self.time)."""

    def __init__(self, cell, component):
        self.cell = tuple(int(v) for v in cell)
        self.component = component
        self.times, self.values = [], []

    def record(self, fields, t):
        self.times.append(float(t))
        self.values.append(fields[self.component][self.cell])

    def get_data(self):
        return np.array(self.times, dtype=np.float64), np.array(self.values, dtype=np.float64)


def dft_sample_interval(dt, sources, monitor_frequencies):
    """m = max(1, floor(1 / ((max source limit + max monitor frequency) dt))) (em/monitors.py:177-191)."""
    if not sources or not monitor_frequencies:
        return 1
    limits = []
    for source in sources:
        limit = source.temporal.spectral_limit()
        if limit is None:
            return 1
        limits.append(float(limit))
    rate = max(limits) + max(float(np.max(f)) for f in monitor_frequencies)
    return max(1, int(np.floor(1.0 / (rate * float(dt)))))


# --------------------------------------------------------------------------------------------------------- stepper

class Yee2D:
    """2D FDTD model with the reference's Simulation/EMPhysics vocabulary (simulation.py:21-32,
    physics/em/__init__.py:317-518), stepped on the CPU in `dtype`."""

    def __init__(self, size_x, size_y, dx, npml=15, courant=None, boundaries=None, polarization='ez',
                 dtype=np.float64, phase_dtype=np.complex64):
        if polarization not in ('ez', 'hz'):
            raise ValueError(f"polarization must be 'ez' or 'hz', got {polarization!r}")
        self.bc = boundaries if isinstance(boundaries, Boundaries) else Boundaries.preset(boundaries, npml)
        self.geo = geo = Geometry2D(size_x, size_y, dx, self.bc, courant)
        self.polarization = polarization
        self.dtype = np.dtype(dtype).type
        self.phase_dtype = phase_dtype
        self.nx, self.ny, self.dt = geo.nx, geo.ny, geo.dt
        shape = geo.shape
        names = ('ez', 'hx', 'hy') if polarization == 'ez' else ('hz', 'ex', 'ey')
        self.fields = {name: np.zeros(shape, dtype=self.dtype) for name in names}
        if polarization == 'ez':
            self.materials = {'eps_r': np.ones(shape, F32), 'sigma_e': np.zeros(shape, F32)}
        else:
            self.materials = {'eps_xx': np.ones(shape, F32), 'eps_yy': np.ones(shape, F32),
                              'sigma_e': np.zeros(shape, F32)}
        self.materials.update(mu_r=np.ones(shape, F32), sigma_m=np.zeros(shape, F32))
        self.sources, self.monitors, self.probes = [], [], []
        self.pec_regions = []
        self.poles = []
        self.n_step, self.time = 0, 0.0
        self._finalized = False
        self._build_spacings_and_cpml()

    # ---- setup ------------------------------------------------------------------------------------------------
    def _build_spacings_and_cpml(self):
        """Kernel widths = float32 grid.dx (yee_grid.py:11-45); kappa folded into the curl rows
        (fdtd_2d_base.py:76-84, boundaries.py:778-798); CPML sides (boundaries.py:694-730)."""
        geo, real, bc = self.geo, self.dtype, self.bc
        w = float(geo.dx32)
        inv = {}
        for axis, n in (('x', geo.nx), ('y', geo.ny)):
            widths = np.full(n, w)
            inv[axis] = {'e': 1.0 / widths, 'h': 1.0 / (0.5 * (widths + np.roll(widths, -1)))}
        self._inv_bulk = inv
        plus, minus = 1.0, -1.0
        x_sign, y_sign = (plus, minus) if self.polarization == 'ez' else (minus, plus)
        self.sides = []
        curl = {f'{axis}_{kind}': inv[axis][kind].astype(real) for axis in 'xy' for kind in 'eh'}
        for face in FACES:
            if not bc.cpml(face):
                continue
            axis, side = face[0], face[2:]
            n, other = (geo.nx, geo.ny) if axis == 'x' else (geo.ny, geo.nx)
            profiles = cpml_axis_profiles(bc, axis, n, w)
            s = _Side(axis, side, x_sign if axis == 'x' else y_sign, profiles, inv[axis], float(geo.dt), n, bc.npml,
                      other, real)
            self.sides.append(s)
            for kind in 'eh':
                rs = s.rows[kind]
                curl[f'{axis}_{kind}'][rs['rows']] /= rs['kappa']
        self.inv_dx_e, self.inv_dx_h = curl['x_e'], curl['x_h']
        self.inv_dy_e, self.inv_dy_h = curl['y_e'], curl['y_h']

    def add_dielectric_rectangle(self, x0, y0, x1, y1, eps_r, sigma_e=0.0):
        """physics/em/__init__.py:317-327, core/material_grid.py:56-83 (later writes win)."""
        # the mask mx x my is a block of cells (centres increase along each axis): write it through index ranges,
        # which sets the same cells as the boolean mask in a fraction of the time on large grids
        mx, my = self.geo.rect_axes(x0, y0, x1, y1)
        rows, columns = np.nonzero(mx)[0], np.nonzero(my)[0]
        if rows.size and columns.size:
            block = (slice(rows[0], rows[-1] + 1), slice(columns[0], columns[-1] + 1))
        else:
            block = (slice(0, 0), slice(0, 0))
        names = ('eps_r',) if self.polarization == 'ez' else ('eps_xx', 'eps_yy')
        for name in names:
            self.materials[name][block] = F32(eps_r)
        self.materials['sigma_e'][block] = F32(sigma_e)
        self._finalized = False

    def add_drude_rectangle(self, x0, y0, x1, y1, omega_p, gamma, eps_inf=1.0):
        """Drude pole on the rectangle with eps_r = eps_inf there (physics/em/__init__.py:329-340)."""
        self._add_pole(('drude', omega_p, gamma), self.geo.rect_mask(x0, y0, x1, y1), eps_inf)

    def add_dispersive_drude(self, mask, omega_p, gamma):
        self._add_pole(('drude', omega_p, gamma), mask, None)

    def add_dispersive_lorentz(self, mask, delta_eps, omega_0, gamma):
        """Lorentz pole over the cells' own eps_r (physics/em/__init__.py:342-355)."""
        self._add_pole(('lorentz', delta_eps, omega_0, gamma), mask, None)

    def _add_pole(self, pole, mask, eps_inf):
        mask = np.asarray(mask, dtype=bool)
        if eps_inf is not None:
            for name in (('eps_r',) if self.polarization == 'ez' else ('eps_xx', 'eps_yy')):
                self.materials[name][mask] = F32(eps_inf)
        self.poles.append({'pole': pole, 'mask': mask.copy()})
        self._finalized = False

    def add_pec_region(self, mask):
        """PEC cells (physics/em/__init__.py:372-376): E_z pinned to 0 in them (ca = cb = 0,
        solvers/fdtd_ez_2d.py:180-233), or every E face touching them pinned (H_z, solvers/fdtd_hz_2d.py:475-495,
        yee_grid.py:145-150); sources inject weight x keep."""
        self.pec_regions.append(np.asarray(mask, dtype=bool).copy())
        self._finalized = False

    def pec_mask(self):
        """PEC regions plus the outermost row of every 'pec' face: E_z wall rows pinned to 0 x the inner row
        (boundaries.py:57-61, 105-128)."""
        mask = np.zeros(self.geo.shape, dtype=bool)
        for region in self.pec_regions:
            mask |= region
        for face, kind in self.bc.faces.items():
            if kind == 'pec' and self.polarization == 'ez':   # H_z 'pec' faces are wall copies (_hz_walls)
                index = 0 if face.endswith('lo') else -1
                if face[0] == 'x':
                    mask[index, :] = True
                else:
                    mask[:, index] = True
        return mask

    def add_gaussian_beam(self, axis, position, waist, center, wavelength=None, frequency=None, bandwidth=0.5,
                          amplitude=1.0, waveform='gaussian', span=None):
        """Sheet current with the transverse factor exp(-((s - center) / waist)^2) over the float64 centres s of the
        line cells (physics/em/__init__.py:450-457, placement.py:246-249)."""
        if not waist > 0:
            raise ValueError('Gaussian beam waist must be > 0')
        return self.add_line_source(axis, position, wavelength, frequency, bandwidth, amplitude, waveform, span,
                                    beam=(float(waist), float(center)))

    def add_line_source(self, axis, position, wavelength=None, frequency=None, bandwidth=0.5, amplitude=1.0,
                        waveform='gaussian', span=None, beam=None):
        """Sheet current of amplitude [V/m] per side (physics/em/__init__.py:442-473, placement.py:223-251)."""
        geo = self.geo
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        along_axis = 0 if axis == 'x' else 1
        if axis == 'x':
            geo.validate(0.0, position, 'Line source (y)')
            line_cell = geo.cell(1, position)
        elif axis == 'y':
            geo.validate(position, 0.0, 'Line source (x)')
            line_cell = geo.cell(0, position)
        else:
            raise ValueError(f"axis must be 'x' or 'y', got {axis!r}")
        n_along = geo.shape[along_axis]
        if span is None:
            span_cells = (0, n_along)
        else:
            start, end = sorted((float(span[0]), float(span[1])))
            lo, hi = geo.cell(along_axis, start), geo.cell(along_axis, end) + 1
            window = geo.windows[along_axis]
            if lo == window.start and self.bc.cpml(f'{axis}_lo'):
                lo = 0
            if hi == window.stop and self.bc.cpml(f'{axis}_hi'):
                hi = n_along
            span_cells = (lo, hi)
        if self.polarization == 'ez':
            normal_width = float(geo.widths[1 - along_axis][line_cell])
            component = 'ez'
        else:
            inv_face = self._inv_bulk['y' if axis == 'x' else 'x']['h']
            n_normal = geo.shape[1 - along_axis]
            if line_cell == n_normal - 1 and not self.bc.periodic('y' if axis == 'x' else 'x'):
                raise ValueError('H_z sheet on the outer grid wall')
            normal_width = 1.0 / float(inv_face[line_cell])
            component = 'ex' if axis == 'x' else 'ey'
        shape = None
        if beam is not None:
            waist, centre = beam
            along = geo.centres[along_axis][span_cells[0]:span_cells[1]]
            shape = np.exp(-((along - centre) / waist) ** 2)
        source = LineSource(axis, line_cell, span_cells, temporal, amplitude, normal_width, component, shape)
        self.sources.append(source)
        self._finalized = False
        return source

    def add_tfsf(self, x0, y0, x1, y1, wavelength=None, frequency=None, bandwidth=0.5, amplitude=1.0,
                 waveform='gaussian'):
        """+x plane wave of incident peak `amplitude` [V/m] (E_z on ez; E_y on hz) inside the total-field box of cells
        [cell(x_lo), cell(x_hi) + 1) x [cell(y_lo), cell(y_hi) + 1) (physics/em/__init__.py:475-490, placement.py:
        97-150): one scattered-field cell must stay inside the window on every side, except a y mirror face of the
        wave's parity (E_z symmetric, H_z antisymmetric), which the box opens onto (no correction on that side)."""
        geo = self.geo
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        for x, y, label in ((x0, y0, 'corner 0'), (x1, y1, 'corner 1')):
            geo.validate(x, y, f'TFSF box {label}')
        x_lo, x_hi = sorted((x0, x1))
        y_lo, y_hi = sorted((y0, y1))
        cells = (geo.cell(0, x_lo), geo.cell(0, x_hi) + 1, geo.cell(1, y_lo), geo.cell(1, y_hi) + 1)
        wx, wy = geo.windows
        reached = {face for face, hit in (('x_lo', cells[0] <= wx.start), ('x_hi', cells[1] >= wx.stop),
                                          ('y_lo', cells[2] <= wy.start), ('y_hi', cells[3] >= wy.stop)) if hit}
        parity = 'symmetric' if self.polarization == 'ez' else 'antisymmetric'
        open_faces = {face for face in reached if face[0] == 'y' and self.bc.faces[face] == parity}
        if {'x_lo', 'x_hi'} & reached or (reached - open_faces):
            raise ValueError(f'TFSF box cells {cells} need a scattered-field cell inside the window on every side')
        source = TFSFBox(cells, temporal, float(geo.dx32), float(self.dt), amplitude, self.polarization, open_faces,
                         self.dtype)
        self.sources.append(source)
        self._finalized = False
        return source

    def _point(self, x, y, label):
        self.geo.validate(x, y, label)
        return self.geo.cell(0, x), self.geo.cell(1, y)

    def add_point_source(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, current=1e-3,
                         waveform='gaussian'):
        """E_z filament current I [A]: J = I / (dx dy) with the float32 kernel width, weight = cb J
        (physics/em/__init__.py:411-418, placement.py:48-52)."""
        if self.polarization != 'ez':
            raise ValueError('add_point_source needs polarization ez')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        cell = self._point(x, y, 'Point source')
        w = float(self.geo.dx32)
        source = CellCurrent('ez', cell, temporal, float(current) / (w * w), 'cb', 'e')
        self.sources.append(source)
        self._finalized = False
        return source

    def add_magnetic_line_current(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, magnetic_current=1.0,
                                  waveform='gaussian'):
        """H_z magnetic line current K_m [V]: M = K_m / (dx dy), weight = db M, added to H_z after the H update at t
        (physics/em/__init__.py:420-429)."""
        if self.polarization != 'hz':
            raise ValueError('add_magnetic_line_current needs polarization hz')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        cell = self._point(x, y, 'Magnetic line current')
        w = float(self.geo.dx32)
        source = CellCurrent('hz', cell, temporal, float(magnetic_current) / (w * w), 'db', 'h')
        self.sources.append(source)
        self._finalized = False
        return source

    def add_electric_dipole(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, dipole_moment=1e-3,
                            orientation=0.0, waveform='gaussian'):
        """H_z in-plane dipole M [A] along `orientation` [deg]: E_x faces (ix, j) carry f M cos / (dx_i dual dy_j),
        E_y faces (i, iy) f M sin / (dual dx_i dy_j); weights cb_x / cb_y of the face (physics/em/__init__.py:431-440,
        placement.py:279-299, sources.py:605-616)."""
        if self.polarization != 'hz':
            raise ValueError('add_electric_dipole needs polarization hz')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        ix, iy = self._point(x, y, 'Electric dipole')
        w = float(self.geo.dx32)
        angle = math.radians(float(orientation))
        mx, my = float(dipole_moment) * math.cos(angle), float(dipole_moment) * math.sin(angle)
        faces = [('ex', (ix, j), share * mx * self._inv_bulk['y']['h'][j] / w)
                 for j, share in dipole_face_shares(iy, self.ny, self.bc.periodic('y'))]
        faces += [('ey', (i, iy), share * my * self._inv_bulk['x']['h'][i] / w)
                  for i, share in dipole_face_shares(ix, self.nx, self.bc.periodic('x'))]
        biggest = max(abs(density) for *_, density in faces)
        faces = [face for face in faces if abs(face[2]) > 1e-12 * biggest]
        if not faces:
            raise ValueError('Electric dipole needs a finite non-zero moment and orientation')
        added = [CellCurrent(name, cell, temporal, density, 'cb_x' if name == 'ex' else 'cb_y', 'e')
                 for name, cell, density in faces]
        self.sources.extend(added)
        self._finalized = False
        return added

    def add_field_probe(self, x, y, field_component=None):
        """Samples of one component at the cell containing (x, y); None = the out-of-plane field
        (physics/em/__init__.py:493-498, placement.py:16, 29-45)."""
        names = tuple(self.fields)
        component = names[0] if field_component is None else field_component
        if component not in names:
            raise ValueError(f'field probe component of an {self.polarization!r} physics must be one of {names}, '
                             f'got {component!r}')
        self.geo.validate(x, y, 'Field probe')
        probe = PointProbe((self.geo.cell(0, x), self.geo.cell(1, y)), component)
        self.probes.append(probe)
        return probe

    def add_dft_monitor(self, x0, y0, x1, y1, wavelengths=None, frequencies=None, downsample=1):
        """Cells [cell(x_lo), cell(x_hi)] x [cell(y_lo), cell(y_hi)] (placement.py:97-105, physics/em/__init__.py:
        500-509)."""
        geo = self.geo
        if (wavelengths is None) == (frequencies is None):
            raise ValueError('pass exactly one of wavelengths or frequencies')
        freqs = (np.asarray(frequencies, dtype=np.float64) if frequencies is not None
                 else C0 / np.asarray(wavelengths, dtype=np.float64))
        geo.validate(x0, y0, 'DFT monitor corner 0')
        geo.validate(x1, y1, 'DFT monitor corner 1')
        x_lo, x_hi = sorted((x0, x1))
        y_lo, y_hi = sorted((y0, y1))
        monitor = DFTBox((geo.cell(0, x_lo), geo.cell(0, x_hi) + 1), (geo.cell(1, y_lo), geo.cell(1, y_hi) + 1),
                         np.atleast_1d(freqs), float(geo.dt), self.polarization,
                         (self.bc.periodic('x'), self.bc.periodic('y')), downsample, geo.shape, self.phase_dtype)
        self.monitors.append(monitor)
        self._finalized = False
        return monitor

    def add_flux_monitor(self, axis, position, span, wavelengths=None, frequencies=None):
        """placement.py:55-70, physics/em/__init__.py:511-518."""
        geo = self.geo
        if (wavelengths is None) == (frequencies is None):
            raise ValueError('pass exactly one of wavelengths or frequencies')
        freqs = (np.asarray(frequencies, dtype=np.float64) if frequencies is not None
                 else C0 / np.asarray(wavelengths, dtype=np.float64))
        lo, hi = sorted((float(span[0]), float(span[1])))
        if axis == 'x':
            normal, line_cell, cells = 'y', geo.cell(1, position), (geo.cell(0, lo), geo.cell(0, hi) + 1)
            widths = geo.widths[0]
        else:
            normal, line_cell, cells = 'x', geo.cell(0, position), (geo.cell(1, lo), geo.cell(1, hi) + 1)
            widths = geo.widths[1]
        monitor = FluxMonitor(normal, line_cell, cells, np.atleast_1d(freqs), float(geo.dt), widths,
                              self.polarization, self.phase_dtype)
        self.monitors.append(monitor)
        self._finalized = False
        return monitor

    def finalize(self):
        """Coefficients (fdtd_ez_2d_kernels.py:310-353, fdtd_hz_2d_kernels.py:190-278), source calibration
        (em/sources.py:562-602) and the DFT sample interval (physics/em/__init__.py:181-203)."""
        real, m = self.dtype, self.materials
        dt, eps0, mu0 = real(self.dt), real(EPS0), real(MU0)

        def coefficients(eps_rel, sigma, permittivity):
            eps = eps_rel.astype(real) * permittivity
            k = sigma.astype(real) * dt / (real(2.0) * eps)
            denom = real(1.0) + k
            return (real(1.0) - k) / denom, (dt / eps) / denom

        self.da, self.db = coefficients(m['mu_r'], m['sigma_m'], mu0)
        pec = self.pec_mask()
        if self.polarization == 'ez':
            self.ca, self.cb = coefficients(m['eps_r'], m['sigma_e'], eps0)
            self.keep = ~pec
            self.ca = self.ca * self.keep
            self.cb = self.cb * self.keep
            # poles (fdtd_ez_2d.py:123-163): drive = mask x PEC keep x the float32 drive coefficient
            for entry in self.poles:
                kind = entry['pole'][0]
                coefficients = (drude_coefficients(*entry['pole'][1:], self.dt) if kind == 'drude'
                                else lorentz_coefficients(*entry['pole'][1:], self.dt))
                entry['coefficients'] = tuple(real(c) for c in coefficients)
                entry['drive'] = (entry['mask'] & self.keep).astype(real) * real(coefficients[-1])
                if 'state' not in entry:
                    entry['state'] = [np.zeros(self.geo.shape, dtype=real) for _ in range(1 if kind == 'drude' else 2)]
        else:
            half = real(0.5)
            pj = np.roll(np.arange(self.ny), -1) if self.bc.periodic('y') else np.minimum(np.arange(self.ny) + 1,
                                                                                         self.ny - 1)
            pi = np.roll(np.arange(self.nx), -1) if self.bc.periodic('x') else np.minimum(np.arange(self.nx) + 1,
                                                                                         self.nx - 1)
            exx, eyy, sig = (m[name].astype(real) for name in ('eps_xx', 'eps_yy', 'sigma_e'))
            self.ca_x, self.cb_x = coefficients(exx * half + exx[:, pj] * half, sig * half + sig[:, pj] * half, eps0)
            self.ca_y, self.cb_y = coefficients(eyy * half + eyy[pi] * half, sig * half + sig[pi] * half, eps0)
            self.keep_x, self.keep_y = ~(pec | pec[:, pj]), ~(pec | pec[pi])   # faces touching a PEC cell
            self.ca_x, self.cb_x = self.ca_x * self.keep_x, self.cb_x * self.keep_x
            self.ca_y, self.cb_y = self.ca_y * self.keep_y, self.cb_y * self.keep_y
            # poles on the E faces (fdtd_hz_2d.py:441-473, 539-561): drive = the mask's face average (yee_grid.py:
            # 128-142, uniform cells) x the face keep x the drive coefficient
            for entry in self.poles:
                kind = entry['pole'][0]
                coefficients = (drude_coefficients(*entry['pole'][1:], self.dt) if kind == 'drude'
                                else lorentz_coefficients(*entry['pole'][1:], self.dt))
                entry['coefficients'] = tuple(real(c) for c in coefficients)
                m = entry['mask'].astype(real)
                entry['drive'] = {'ex': (half * m + half * m[:, pj]) * self.keep_x * real(coefficients[-1]),
                                  'ey': (half * m + half * m[pi]) * self.keep_y * real(coefficients[-1])}
                if 'state' not in entry:
                    entry['state'] = {name: [np.zeros(self.geo.shape, dtype=real)
                                             for _ in range(1 if kind == 'drude' else 2)] for name in ('ex', 'ey')}
        self._calibrate()
        if self.n_step == 0:
            interval = dft_sample_interval(self.dt, self.sources, [mon.frequencies for mon in self.monitors])
            for monitor in self.monitors:
                monitor.sample_interval = interval
        self._finalized = True

    def _calibrate(self):
        """profile = float32(cb_line admittance weight) (em/sources.py:383-392, 582-602)."""
        m = self.materials
        for source in self.sources:
            if isinstance(source, TFSFBox):   # vacuum plane wave: no calibration
                continue
            if isinstance(source, CellCurrent):
                source.calibrate(getattr(self, source.coefficient))
                continue
            omega = 2.0 * np.pi * source.temporal.f0
            if self.polarization == 'ez':
                cb_line = source.line(self.cb)
                eps = source.line(m['eps_r']).astype(np.complex128) - 1j * source.line(m['sigma_e']).astype(
                    np.float64) / (omega * EPS0)
            else:
                cb_face, cell_eps = (self.cb_x, m['eps_xx']) if source.component == 'ex' else (self.cb_y, m['eps_yy'])
                cb_line = source.line(cb_face)
                n_normal = self.geo.shape[1 if source.axis == 'x' else 0]
                lower_row, upper_row = source.position, (source.position + 1) % n_normal
                lo, hi = source.span

                def row_eps(row):
                    line = (slice(lo, hi), row) if source.axis == 'x' else (row, slice(lo, hi))
                    return cell_eps[line].astype(np.complex128) - 1j * m['sigma_e'][line].astype(np.float64) / (
                        omega * EPS0)
                w = float(self.geo.dx32)
                eps = (row_eps(lower_row) * w + row_eps(upper_row) * w) / (w + w)
            admittance = sheet_admittance(eps, source.line(m['mu_r']), source.line(m['sigma_m']), omega,
                                          source.normal_width, float(self.dt))
            source.profile[...] = cb_line * admittance * source.weight   # cb_line is 0 on PEC cells / faces

    # ---- stepping ---------------------------------------------------------------------------------------------
    def step(self):
        """One Yee step: H, H CPML, E, E CPML, electric sources at t + dt/2, clock, monitors at t + dt
        (fdtd_ez_2d.py:71-119, fdtd_hz_2d.py:144-189, fdtd_2d_base.py:470-488)."""
        if not self._finalized:
            self.finalize()
        t_start = float(self.time)
        t_half = t_start + 0.5 * float(self.dt)
        t_full = t_start + float(self.dt)
        if self.polarization == 'ez':
            self._step_ez()
        else:
            self._step_hz(t_start)
        pec = self.pec_mask()
        keep = None if not pec.any() else (self.keep if self.polarization == 'ez' else (self.keep_x, self.keep_y))
        for source in self.sources:
            field = getattr(source, 'field', 'e')
            if field == 'e':
                source.inject(self.fields, t_half)
            elif field == 'tfsf':
                source.inject_e(self.fields, t_half, keep)
        if self.polarization == 'ez':
            self._ez_walls()
        else:
            self._exy_walls()
        self.n_step += 1
        self.time = t_full
        for probe in self.probes:
            probe.record(self.fields, t_full)
        for monitor in self.monitors:
            if self.n_step % monitor.sample_interval == 0:
                monitor.update(self.fields, t_full)

    HZ_CENTRE_ROWS = {'x_lo': (np.s_[0, :], np.s_[1, :]), 'x_hi': (np.s_[-1, :], np.s_[-2, :]),
                      'y_lo': (np.s_[:, 0], np.s_[:, 1]), 'y_hi': (np.s_[:, -1], np.s_[:, -2])}
    PLANE_FACE_ROWS = {'x_lo': np.s_[0, :], 'x_hi': np.s_[-2, :], 'y_lo': np.s_[:, 0], 'y_hi': np.s_[:, -2]}

    def _hz_walls(self):
        """Outer H_z rows (boundaries.py:82-93, 517-527): pec / antisymmetric faces H_z even, then pmc faces 0, then
        symmetric faces H_z odd, each kind over the faces in order."""
        hz = self.fields['hz']
        for kinds, sign in ((('pec', 'antisymmetric'), 1.0), (('pmc',), 0.0), (('symmetric',), -1.0)):
            for face in FACES:
                if self.bc.faces[face] in kinds:
                    outer, inner = self.HZ_CENTRE_ROWS[face]
                    hz[outer] = 0 if sign == 0 else self.dtype(sign) * hz[inner]

    def _exy_walls(self):
        """Tangential E = 0 on the plane face of the pec / antisymmetric faces (E_y rows 0 / nx-2 on x faces, E_x
        columns 0 / ny-2 on y faces; apply_exy_walls, boundaries.py:529-538), after the sources."""
        for face in FACES:
            if self.bc.faces[face] in ('pec', 'antisymmetric'):
                self.fields['ey' if face[0] == 'x' else 'ex'][self.PLANE_FACE_ROWS[face]] = 0

    def _ez_walls(self):
        """Outer E_z rows after the sources (the re-pin of fdtd_ez_2d.py:105-117): x faces, then y faces, each row
        set to its factor times its inner neighbour (physics/em/boundaries.py:130-141 host twin, :478-488)."""
        ez = self.fields['ez']
        for face, outer, inner in (('x_lo', np.s_[0, :], np.s_[1, :]), ('x_hi', np.s_[-1, :], np.s_[-2, :]),
                                   ('y_lo', np.s_[:, 0], np.s_[:, 1]), ('y_hi', np.s_[:, -1], np.s_[:, -2])):
            sign = self.bc.wall_sign(face)
            if sign is not None:
                ez[outer] = self.dtype(sign) * ez[inner]

    def _step_ez(self):
        f = self.fields
        ez, hx, hy = f['ez'], f['hx'], f['hy']
        da, db, ca, cb = self.da, self.db, self.ca, self.cb
        px, py = self.bc.periodic('x'), self.bc.periodic('y')
        # update_h (fdtd_ez_2d_kernels.py:23-60)
        hx[:, :-1] = da[:, :-1] * hx[:, :-1] - db[:, :-1] * (ez[:, 1:] - ez[:, :-1]) * self.inv_dy_h[:-1]
        if py:
            hx[:, -1] = da[:, -1] * hx[:, -1] - db[:, -1] * (ez[:, 0] - ez[:, -1]) * self.inv_dy_h[-1]
        hy[:-1] = da[:-1] * hy[:-1] + db[:-1] * (ez[1:] - ez[:-1]) * self.inv_dx_h[:-1, None]
        if px:
            hy[-1] = da[-1] * hy[-1] + db[-1] * (ez[0] - ez[-1]) * self.inv_dx_h[-1]
        for side in self.sides:   # apply_h: x sides (ez, db, hy), y sides (ez, db, hx) (boundaries.py:752-756)
            side.apply('h', ez, db, hy if side.axis == 'x' else hx)
        # pmc planes: tangential H = 0 on face 0 / n-2 (apply_h_walls, boundaries.py:463-473)
        for face, field, row in (('x_lo', hy, np.s_[0, :]), ('x_hi', hy, np.s_[-2, :]),
                                 ('y_lo', hx, np.s_[:, 0]), ('y_hi', hx, np.s_[:, -2])):
            if self.bc.faces[face] == 'pmc':
                field[row] = 0
        # TFSF H-side corrections after the H update, then the pmc planes again (fdtd_ez_2d.py:89-98)
        tfsf = [source for source in self.sources if getattr(source, 'field', 'e') == 'tfsf']
        for source in tfsf:
            source.inject_h(f)
        for face, field, row in (('x_lo', hy, np.s_[0, :]), ('x_hi', hy, np.s_[-2, :]),
                                 ('y_lo', hx, np.s_[:, 0]), ('y_hi', hx, np.s_[:, -2])):
            if tfsf and self.bc.faces[face] == 'pmc':
                field[row] = 0
        # pole updates from E^n (fdtd_ez_2d.py:318-345)
        for entry in self.poles:
            if entry['pole'][0] == 'drude':
                damping, _ = entry['coefficients']
                j = entry['state'][0]
                j += damping * j + entry['drive'] * ez
            else:
                damping, restoring, _ = entry['coefficients']
                p, dp = entry['state']
                dp += damping * dp - restoring * p + entry['drive'] * ez
                p += dp
        # update_e (fdtd_ez_2d_kernels.py:125-140)
        curl = np.empty_like(ez)
        curl[1:] = (hy[1:] - hy[:-1]) * self.inv_dx_e[1:, None]
        curl[0] = ((hy[0] - hy[-1]) if px else hy[0]) * self.inv_dx_e[0]
        curl[:, 1:] -= (hx[:, 1:] - hx[:, :-1]) * self.inv_dy_e[1:]
        curl[:, 0] -= ((hx[:, 0] - hx[:, -1]) if py else hx[:, 0]) * self.inv_dy_e[0]
        ez *= ca
        curl *= cb
        ez += curl
        for side in self.sides:   # apply_e: x sides (hy, cb, ez), y sides (hx, cb, ez) (boundaries.py:758-762)
            side.apply('e', hy if side.axis == 'x' else hx, cb, ez)
        # pole corrections (fdtd_ez_2d.py:347-358): Drude E -= cb J_total, Lorentz E -= cb / dt dP
        drude = [entry['state'][0] for entry in self.poles if entry['pole'][0] == 'drude']
        if drude:
            ez -= cb * sum(drude[1:], drude[0].copy())
        inv_dt = self.dtype(1.0) / self.dtype(self.dt)
        for entry in self.poles:
            if entry['pole'][0] == 'lorentz':
                ez -= cb * inv_dt * entry['state'][1]

    def _step_hz(self, t_start):
        f = self.fields
        hz, ex, ey = f['hz'], f['ex'], f['ey']
        px, py = self.bc.periodic('x'), self.bc.periodic('y')
        # update_hz (fdtd_hz_2d_kernels.py:65-80)
        curl = np.empty_like(hz)
        curl[:, 1:] = (ex[:, 1:] - ex[:, :-1]) * self.inv_dy_e[1:]
        curl[:, 0] = ((ex[:, 0] - ex[:, -1]) if py else ex[:, 0]) * self.inv_dy_e[0]
        curl[1:] -= (ey[1:] - ey[:-1]) * self.inv_dx_e[1:, None]
        curl[0] -= ((ey[0] - ey[-1]) if px else ey[0]) * self.inv_dx_e[0]
        hz *= self.da
        curl *= self.db
        hz += curl
        for side in self.sides:   # apply_hz: x sides (ey, db, hz), y sides (ex, db, hz) (boundaries.py:766-768)
            side.apply('e', ey if side.axis == 'x' else ex, self.db, hz)
        # magnetic line currents at t_start after the H update (fdtd_hz_2d.py:144-189, inject_magnetic)
        for source in self.sources:
            field = getattr(source, 'field', 'e')
            if field == 'h':
                source.inject(f, t_start)
            elif field == 'tfsf':
                source.inject_h(f)
        # H_z walls after the magnetic injections (apply_hz_walls, boundaries.py:517-527; re-pinned after injection,
        # fdtd_hz_2d.py:165-170): pec / antisymmetric faces copy the inner row (H_z even), pmc faces zero it,
        # symmetric faces copy it negated (H_z odd); each kind over the faces in order (_set_outer_rows :82-93)
        self._hz_walls()
        # pole updates on the E faces from E^n (fdtd_hz_2d.py:286-310)
        for entry in self.poles:
            for name, field in (('ex', ex), ('ey', ey)):
                state, drive = entry['state'][name], entry['drive'][name]
                if entry['pole'][0] == 'drude':
                    state[0] += entry['coefficients'][0] * state[0] + drive * field
                else:
                    damping, restoring, _ = entry['coefficients']
                    state[1] += damping * state[1] - restoring * state[0] + drive * field
                    state[0] += state[1]
        # update_exy (fdtd_hz_2d_kernels.py:146-155)
        ca_x, cb_x, ca_y, cb_y = self.ca_x, self.cb_x, self.ca_y, self.cb_y
        ex[:, :-1] = ca_x[:, :-1] * ex[:, :-1] + cb_x[:, :-1] * (hz[:, 1:] - hz[:, :-1]) * self.inv_dy_h[:-1]
        if py:
            ex[:, -1] = ca_x[:, -1] * ex[:, -1] + cb_x[:, -1] * (hz[:, 0] - hz[:, -1]) * self.inv_dy_h[-1]
        ey[:-1] = ca_y[:-1] * ey[:-1] - cb_y[:-1] * (hz[1:] - hz[:-1]) * self.inv_dx_h[:-1, None]
        if px:
            ey[-1] = ca_y[-1] * ey[-1] - cb_y[-1] * (hz[0] - hz[-1]) * self.inv_dx_h[-1]
        for side in self.sides:   # apply_exy: x sides (hz, cb_y, ey), y sides (hz, cb_x, ex) (boundaries.py:770-772)
            if side.axis == 'x':
                side.apply('h', hz, cb_y, ey)
            else:
                side.apply('h', hz, cb_x, ex)
        # pole corrections (fdtd_hz_2d.py:312-323): Drude E -= cb J, Lorentz E -= cb / dt dP, per face set
        inv_dt = self.dtype(1.0) / self.dtype(self.dt)
        for entry in self.poles:
            for field, cb, name in ((ex, cb_x, 'ex'), (ey, cb_y, 'ey')):
                state = entry['state'][name]
                if entry['pole'][0] == 'drude':
                    field -= cb * state[0]
                else:
                    field -= cb * inv_dt * state[1]

    def run(self, duration=None, n_steps=None):
        """Tick loop stopping at t >= t_end - dt/2 (coupling/runner.py:174-204); returns the steps run."""
        if (duration is None) == (n_steps is None):
            raise ValueError('pass exactly one of duration or n_steps')
        if not self._finalized:
            self.finalize()
        dt = float(self.dt)
        if n_steps is not None:
            for _ in range(int(n_steps)):
                self.step()
            return int(n_steps)
        stop = self.time + float(duration) - 0.5 * dt
        done = 0
        while self.time < stop:
            self.step()
            done += 1
        return done

    def field_window(self, name):
        """Physical-window slice of a field (viz/visualization.py:617-620)."""
        return self.fields[name][self.geo.windows]


def steps_for_duration(dt, duration, t_start=0.0):
    """Number of ticks the runner takes for `duration` from t_start (coupling/runner.py:182-203)."""
    dt = float(dt)
    stop = t_start + float(duration) - 0.5 * dt
    t, n = float(t_start), 0
    while t < stop:
        t = t + dt
        n += 1
    return n
