"""Calibrated sheet (line) sources and their per-step waveform tables.

A line source adds float32(amplitude * s(t + dt/2)) * profile to its field line after the E update of every step
(core/sources.py:146-155, 190-229; physics/em/sources.py:469-480 for the H_z in-plane sheet). The profile is the
reference's sheet calibration (physics/em/sources.py:383-436, 562-602): cb of the line cells times the local admittance
|sqrt(eps/mu) cos(kappa0/2)| times 2 shape / (eta0 w_perp), rounded to float32. The kernel reads the per-step value
from a device table built here from the same float64 clock the reference accumulates.
"""
import math

import numpy as np

from .constants import C0, DTYPE, EPS0, ETA0, MU0


def resolve_frequency(wavelength, frequency) -> float:
    """Source frequency [Hz] from exactly one of wavelength [m] / frequency [Hz] (physics/em/__init__.py:560-568)."""
    if (wavelength is None) == (frequency is None):
        raise ValueError(f'pass exactly one of wavelength or frequency, got wavelength={wavelength}, '
                         f'frequency={frequency}')
    value = float(frequency if frequency is not None else wavelength)
    if not (np.isfinite(value) and value > 0):
        raise ValueError(f'source wavelength / frequency must be finite and > 0, got {value}')
    return value if frequency is not None else C0 / value


def resolve_frequencies(wavelengths, frequencies):
    """Monitor frequencies [Hz] (1D float64) from exactly one of wavelengths / frequencies
    (physics/em/__init__.py:571-577)."""
    if (wavelengths is None) == (frequencies is None):
        raise ValueError('pass exactly one of wavelengths or frequencies')
    values = np.atleast_1d(np.array(frequencies if frequencies is not None else wavelengths, dtype=np.float64))
    if values.ndim != 1 or values.size == 0 or not np.all(np.isfinite(values) & (values > 0)):
        raise ValueError(f'monitor {"frequencies" if frequencies is not None else "wavelengths"} must be a non-empty '
                         f'1D list of finite positive values, got {values}')
    return values if frequencies is not None else C0 / values


def sheet_admittance(line_eps, mu_r, sigma_m, omega, normal_spacing, dt):
    """float32 |sqrt(eps/mu) sqrt(1 - eps mu (w sin(omega dt/2) / (C0 dt))^2)| of the line cells
    (physics/em/sources.py:430-436, physics/em/yee_grid.py:58-65)."""
    permeability = np.asarray(mu_r, dtype=np.complex128) - 1j * np.asarray(sigma_m, dtype=np.float64) / (omega * MU0)
    half_phase = float(normal_spacing) * math.sin(0.5 * float(omega) * float(dt)) / (C0 * float(dt))
    launch = np.sqrt(1.0 - np.asarray(line_eps * permeability, dtype=np.complex128) * half_phase ** 2)
    return np.abs(np.sqrt(line_eps / permeability) * launch).astype(DTYPE)


class LineSource:
    """Sheet current on one grid line: axis 'x' = line along x at y-cell `position` (axis 'y': along y at x-cell).
    component: 'ez' (E_z sheet through the line-cell centres) or 'ex' / 'ey' (H_z: hi faces of the line cells)."""

    field = 'e'   # injected after the E update

    def __init__(self, axis, position, temporal, amplitude, span, normal_width, component, transverse_shape=None):
        if axis not in ('x', 'y'):
            raise ValueError(f"LineSource axis must be 'x' or 'y', got {axis!r}")
        self.axis = axis
        self.position = int(position)
        self.temporal = temporal
        self.amplitude = DTYPE(amplitude)
        self.span = (int(span[0]), int(span[1]))
        if not 0 <= self.span[0] < self.span[1]:
            raise ValueError(f'LineSource span must be a half-open cell range lo < hi, got {span}')
        self.normal_width = float(normal_width)
        self.component = component
        n_line = self.span[1] - self.span[0]
        shape = np.ones(n_line) if transverse_shape is None else np.asarray(transverse_shape, dtype=np.float64)
        self.weight = (2.0 * shape / (ETA0 * self.normal_width)).astype(DTYPE)
        self.profile = np.zeros(n_line, dtype=DTYPE)

    @property
    def n_cells(self):
        return self.span[1] - self.span[0]

    def value(self, t):
        """float32(amplitude * s(t)) in float32 arithmetic (core/sources.py:221-229)."""
        return DTYPE(self.amplitude * DTYPE(self.temporal(float(t))))

    def wave_table(self, half_step_times):
        """float32 value per step from the times t_n + dt/2."""
        return np.array([self.value(t) for t in half_step_times], dtype=DTYPE)

    def line(self, array):
        lo, hi = self.span
        return array[lo:hi, self.position] if self.axis == 'x' else array[self.position, lo:hi]

    def calibrate(self, cb_line, line_eps, mu_line, sigma_m_line, dt):
        """profile = float32(cb * admittance(omega0) * weight) (physics/em/sources.py:582-602)."""
        omega = 2.0 * np.pi * float(self.temporal.f0)
        admittance = sheet_admittance(line_eps, mu_line, sigma_m_line, omega, self.normal_width, dt)
        self.profile[...] = np.asarray(cb_line, dtype=np.float64) * admittance * self.weight


class PointCurrent(LineSource):
    """Current at one cell or face of a 2D model, kept as a one-cell line (axis, line cell, span (cell, cell + 1)):
    profile = float32(coefficient * float32(current density)), the coefficient being cb of the E_z cell (filament
    current I / cell area, physics/em/sources.py:544-561 PointCurrentCalibration), cb of the E_x / E_y face (one face
    of an in-plane electric dipole, :605-616 DipoleCalibration) or db of the H_z cell (magnetic line current
    K_m / cell area). field 'e' adds profile * float32(s(t_n + dt/2)) after the E update, field 'h' adds
    profile * float32(s(t_n)) to H_z after the H update (physics/em/sources.py:439-444, core/sources.py:162-188)."""

    def __init__(self, axis, line, cell, temporal, current_density, component, field='e'):
        super().__init__(axis, line, temporal, 1.0, (cell, cell + 1), 1.0, component)
        if field not in ('e', 'h'):
            raise ValueError(f"point current field must be 'e' or 'h', got {field!r}")
        self.current_density = DTYPE(current_density)
        self.field = field

    def calibrate(self, coefficient, *unused):
        self.profile[...] = np.asarray(coefficient, dtype=np.float64) * np.float64(self.current_density)


def dipole_faces(cell, n_cells, periodic):
    """(face, moment fraction) of the hi face `cell` and the lo face `cell - 1` along one axis, half each; on a
    non-periodic axis a face outside [0, n_cells - 2] (the outer face is never updated) passes its half to the other
    (physics/em/placement.py:268-276)."""
    if periodic:
        return [(cell, 0.5), ((cell - 1) % n_cells, 0.5)]
    faces = [face for face in (cell, cell - 1) if 0 <= face <= n_cells - 2]
    if not faces:
        raise ValueError(f'an electric dipole needs at least 2 cells along a non-periodic axis, got {n_cells}')
    return [(face, 1.0 / len(faces)) for face in faces]


def electric_dipole_faces(ix, iy, n_cells, periodic, width, dipole_moment, orientation):
    """[('ex' | 'ey', i, j, J [A/m^2]), ...] of an in-plane dipole of moment M [A] along `orientation` [deg, ccw from
    +x] in cell (ix, iy) of a uniform grid of cell width `width`: a face carrying the fraction f of M_axis has
    J = f M_axis / (width * dual spacing); faces below 1e-12 of the largest are dropped (physics/em/placement.py:
    279-299)."""
    theta = math.radians(float(orientation))
    moment = (float(dipole_moment) * math.cos(theta), float(dipole_moment) * math.sin(theta))
    inv_dual = 1.0 / (0.5 * (width + width))
    faces = [('ex', ix, j, f * moment[0] * inv_dual / width) for j, f in dipole_faces(iy, n_cells[1], periodic[1])]
    faces += [('ey', i, iy, f * moment[1] * inv_dual / width) for i, f in dipole_faces(ix, n_cells[0], periodic[0])]
    largest = max(abs(face[3]) for face in faces)
    faces = [face for face in faces if abs(face[3]) > 1e-12 * largest]
    if not faces:
        raise ValueError(f'Electric dipole needs a finite non-zero dipole_moment and a finite orientation, got '
                         f'dipole_moment={dipole_moment}, orientation={orientation}')
    return faces


def complex_permittivity(eps_rel, sigma_e, omega):
    """eps_rel - j sigma_e / (omega eps0) (physics/em/dispersion.py:10-21, no poles)."""
    return np.asarray(eps_rel, dtype=np.complex128) - 1j * np.asarray(sigma_e, dtype=np.float64) / (omega * EPS0)


class PointDipole:
    """Electric point dipole of moment I l [A m] on the E face `cell` along `component` ('ex' / 'ey' / 'ez'):
    E[cell] += float32(weight * float32(s(t + dt/2))), weight = float32(cb of the face * moment / dual volume)
    (physics/em3d/sources3d.py:41-75, physics/em3d/engine3d.py:299-311)."""

    def __init__(self, component, cell, temporal, moment, volume):
        if component not in ('ex', 'ey', 'ez'):
            raise ValueError(f"dipole component must be 'ex', 'ey' or 'ez', got {component!r}")
        self.component = component
        self.axis = 'xyz'.index(component[1])
        self.cell = tuple(int(v) for v in cell)
        self.temporal = temporal
        self.moment = float(moment)
        self.volume = float(volume)
        self.weight = DTYPE(0.0)

    def calibrate(self, cb):
        self.weight = DTYPE(float(cb) * self.moment / self.volume)

    def wave_table(self, half_step_times):
        """float32 s(t_n + dt/2) per step."""
        return np.array([self.temporal(t) for t in half_step_times], dtype=np.float64).astype(DTYPE)


class PlaneWaveBox:
    """+x vacuum plane wave in the total-field cells [x0, x1) x [y0, y1) of a uniform 2D grid (physics/em/sources.py:
    207-378: _PlaneWaveBox, TFSFSource, TFSFSourceHz). A 1D incident line (20 margin cells, the box columns, 20 margin
    cells; the soft source at line cell 10 with value float32(2 S cos(kappa0/2) amplitude s(t + dt/2)); first-order
    Mur ends matched to the carrier's numerical phase velocity) runs on the host in float32 with the reference's
    update order, since it depends on nothing in the grid; tables() records, per step n, the incident values the box
    edges need: E_z: line = E_z^n of the box columns, ends = H_y^(n+1/2) at the faces x0-1, x1-1; H_z: line =
    H_z^(n+1/2) of the box columns, ends = E_y^n at the faces x0-1, x1-1. The kernel adds them with the signed side
    coefficients dt / (mu0 or eps0) / dx (0 on a side that opens onto a mirror face)."""

    MARGIN, SOURCE_INDEX = 20, 10
    field = 'tfsf'

    def __init__(self, cells, temporal, dx, dt, amplitude, polarization, open_faces=()):
        x0, x1, y0, y1 = (int(v) for v in cells)
        if not (1 <= x0 < x1 and 1 <= y0 < y1):
            raise ValueError(f'TFSF box [{x0}, {x1}) x [{y0}, {y1}) must be non-empty and keep one scattered-field '
                             f'cell on the lo sides (x0 >= 1, y0 >= 1)')
        self.x0, self.x1, self.y0, self.y1 = x0, x1, y0, y1
        self.temporal, self.amplitude = temporal, float(amplitude)
        self.polarization, self.open_faces = polarization, frozenset(open_faces)
        dt, dx = float(dt), float(dx)
        margin = self.MARGIN
        widths = np.full(2 * margin + (x1 - x0), dx)
        face_spacings = 0.5 * (widths[:-1] + widths[1:])
        f0 = getattr(temporal, 'f_peak', None) or temporal.f0
        half_omega_dt = math.pi * float(f0) * dt
        courant = C0 * dt / dx
        sin_half_kappa = math.sin(half_omega_dt) / courant
        if sin_half_kappa >= 1.0:
            raise ValueError(f'TFSF: the centre frequency has no propagating wave on the 1D grid (sin(k dx/2) = '
                             f'{sin_half_kappa:.3g} >= 1); refine the mesh')
        cos_half_kappa = math.sqrt(1.0 - sin_half_kappa ** 2)
        mur_courant = courant * cos_half_kappa / math.cos(half_omega_dt)
        self.mur = DTYPE((mur_courant - 1.0) / (mur_courant + 1.0))
        self.source_scale = 2.0 * courant * cos_half_kappa * self.amplitude
        self.aux_x0, self.n = margin, widths.size

        def sides(constant, count):
            names = ('x_lo', 'x_hi', 'y_lo', 'y_hi')[:count]
            return [DTYPE(0.0 if name in self.open_faces else dt / (constant * dx)) for name in names]
        if polarization == 'ez':
            h, e = sides(MU0, 4), sides(EPS0, 2)
            # H_y rows x0-1 (-) / x1-1 (+), H_x columns y0-1 (+) / y1-1 (-); E_z rows x0 (-) / x1-1 (+)
            self.coef_h = (-h[0], h[1], h[2], -h[3])
            self.coef_e = (-e[0], e[1], DTYPE(0.0), DTYPE(0.0))
            self.aux_coef_h = (dt / (MU0 * face_spacings)).astype(DTYPE)
            self.aux_coef_e = (dt / (EPS0 * widths)).astype(DTYPE)
        else:
            h, e = sides(MU0, 2), sides(EPS0, 4)
            # H_z rows x0 (+) / x1-1 (-); E_y rows x0-1 (+) / x1-1 (-), E_x columns y0-1 (-) / y1-1 (+)
            self.coef_h = (h[0], -h[1], DTYPE(0.0), DTYPE(0.0))
            self.coef_e = (e[0], -e[1], -e[2], e[3])
            self.aux_coef_h = (-dt / (MU0 * widths[1:])).astype(DTYPE)
            self.aux_coef_e = (-dt / (EPS0 * np.append(face_spacings, widths[-1]))).astype(DTYPE)

    @property
    def cells(self):
        return self.x0, self.x1, self.y0, self.y1

    def _aux_e_update(self, e, h, value):
        n = self.n
        old = (e[0], e[1], e[n - 2], e[n - 1])
        e[1:n - 1] += self.aux_coef_e[1:n - 1] * (h[1:n - 1] - h[0:n - 2])
        e[self.SOURCE_INDEX] += value
        e[0] = old[1] + self.mur * (e[1] - old[0])
        e[n - 1] = old[2] + self.mur * (e[n - 2] - old[3])

    def tables(self, half_step_times):
        """(line (steps, x1 - x0), ends (steps, 2)) float32 incident values for the steps whose half-step times are
        given (the source is evaluated at t_n + dt/2 on the engine's float64 clock)."""
        steps = len(half_step_times)
        length = self.x1 - self.x0
        a0, last = self.aux_x0, self.aux_x0 + length - 1
        line = np.zeros((steps, length), dtype=DTYPE)
        ends = np.zeros((steps, 2), dtype=DTYPE)
        aux_e = np.zeros(self.n, dtype=DTYPE)
        aux_h = np.zeros(self.n, dtype=DTYPE)
        scale = self.source_scale
        for step, t in enumerate(half_step_times):
            value = DTYPE(scale * self.temporal(float(t)))
            if self.polarization == 'ez':
                line[step] = aux_e[a0:a0 + length]
                aux_h[:-1] += self.aux_coef_h * (aux_e[1:] - aux_e[:-1])
                ends[step] = (aux_h[a0 - 1], aux_h[last])
                self._aux_e_update(aux_e, aux_h, value)
            else:
                ends[step] = (aux_e[a0 - 1], aux_e[last])
                aux_h[1:] += self.aux_coef_h * (aux_e[1:] - aux_e[:-1])
                line[step] = aux_h[a0:a0 + length]
                self._aux_e_update(aux_e, aux_h[1:], value)
        return line, ends
