"""2D FDTD models (host setup) and their device engine over libyee.so.

Model2D gathers what the reference spreads over Domain / MaterialGrid / EMPhysics / BoundaryManager / CPMLSided and
the source and monitor objects (physics/em/__init__.py:63-557): geometry, a material-index grid, CPML tables, curl
spacings, calibrated sheet sources and flux monitors, all host float64 rounded once to float32. Engine2D owns the
device buffers (torch tensors on the ROCm device: memory only) and steps them with one fused HIP launch per step.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .boundaries import CPMLAxisTables, curl_spacings, resolve_boundaries
from .constants import C0, DTYPE, EPS0, MU0
from .domain import Domain
from .materials import MaterialMap, semi_implicit
from .dispersion import DrudePole, LorentzPole, bounding_box, pole_step_term
from .monitors import DFTMonitor, FieldProbe, FluxMonitor, ProbeTable, dft_sample_interval
from .sources import (LineSource, PlaneWaveBox, PointCurrent, complex_permittivity, electric_dipole_faces,
                      resolve_frequencies, resolve_frequency)
from .waveforms import make_temporal

POLARIZATIONS = ('ez', 'hz')
COMPONENTS = {'ez': ('ez', 'hx', 'hy'), 'hz': ('hz', 'ex', 'ey')}


class Model2D:
    """Host description of one 2D FDTD model (E_z or H_z polarisation) on a uniform grid."""

    def __init__(self, size_x, size_y, dx, npml=15, courant=None, boundaries=None, polarization='ez'):
        if polarization not in POLARIZATIONS:
            raise ValueError(f"polarization must be 'ez' (E_z, H_x, H_y) or 'hz' (H_z, E_x, E_y), got "
                             f"{polarization!r}")
        if size_x is None or size_y is None or dx is None:
            raise ValueError('a 2D model needs size_x, size_y and dx')
        self.polarization = polarization
        self.bc = resolve_boundaries(boundaries, npml)
        if self.bc.dimensions != 2:
            raise ValueError('a 2D model needs a 2D BoundaryConfig')
        self.domain = Domain((size_x, size_y), dx, self.bc.npml_sides, courant)
        self.grid = self.domain.grid
        self.nx, self.ny = self.grid.shape
        self.dt = self.grid.dt
        self.materials = MaterialMap((self.nx, self.ny))
        self.sources = []
        self.tfsf = []      # PlaneWaveBox records (TFSF boxes)
        self.monitors = []
        dx32 = float(self.grid.dx)
        self.cpml = {axis: CPMLAxisTables(self.bc, axis, n, dx32, float(self.dt))
                     for axis, n in (('x', self.nx), ('y', self.ny))}
        self.inv_dx_e, self.inv_dx_h = curl_spacings(self.nx, dx32, self.cpml['x'])
        self.inv_dy_e, self.inv_dy_h = curl_spacings(self.ny, dx32, self.cpml['y'])
        # the curl spacing outside the CPML layers (uniform grid: every bulk entry of all four arrays)
        self.inv_bulk = tuple(float(a[b]) for a, b in ((self.inv_dx_e, self.nx // 2), (self.inv_dy_e, self.ny // 2)))
        self.sample_interval = 1
        self.pec_regions = []   # bool (nx, ny) masks, applied over the dielectrics at finalize
        self.probes = []
        self.dft_monitors = []
        self.poles = []   # (DrudePole or LorentzPole, bool (nx, ny) mask)

    @property
    def shape(self):
        return (self.nx, self.ny)

    @property
    def components(self):
        return COMPONENTS[self.polarization]

    @property
    def cell_count(self):
        return self.nx * self.ny

    # ---- materials ------------------------------------------------------------------------------------------------
    def add_dielectric_rectangle(self, x0, y0, x1, y1, eps_r, sigma_e=0.0):
        """Cells whose float32 centre lies in [x_min, x_max) x [y_min, y_max) get eps_r (every permittivity array
        of the polarisation) and sigma_e (physics/em/__init__.py:317-336; core/domain.py:255-260)."""
        ranges = self.domain.box_ranges((x0, y0), (x1, y1))
        self.materials.paint_box(ranges, self._material(eps_r, sigma_e))

    def add_dielectric_circle(self, cx, cy, radius, eps_r, sigma_e=0.0):
        """Cells whose float32 centre lies within `radius` of (cx, cy) (core/domain.py:262-268)."""
        if not radius >= 0:
            raise ValueError(f'circle radius must be >= 0, got {radius}')
        x, y = (c.astype(DTYPE) for c in self.domain.cell_centers())
        mask = (x[:, None] - DTYPE(cx)) ** 2 + (y[None, :] - DTYPE(cy)) ** 2 <= DTYPE(radius) ** 2
        self.materials.paint_mask(mask, self._material(eps_r, sigma_e))

    def _material(self, eps_r, sigma_e):
        return self.materials.material_id(eps_r, eps_r, eps_r, sigma_e, 1.0, 0.0)

    def add_field_probe(self, x, y, field_component=None):
        """Raw Yee samples of one component at the cell containing (x, y); None = the out-of-plane field
        (physics/em/__init__.py:493-498, placement.py:29-45)."""
        components = COMPONENTS[self.polarization]
        field_component = components[0] if field_component is None else field_component
        if field_component not in components:
            raise ValueError(f"field probe component of an {self.polarization!r} physics must be one of "
                             f"{components}, got {field_component!r}")
        self.domain.validate_physical_pos((x, y), 'Field probe')
        probe = FieldProbe(self.domain.pos_to_cell(x, y), field_component)
        self.probes.append(probe)
        return probe

    # ---- dispersive poles (E_z) ----------------------------------------------------------------------------------
    def rect_mask(self, x0, y0, x1, y1):
        """bool (nx, ny): cells whose float32 centre lies in [x_min, x_max) x [y_min, y_max) (core/domain.py:255-260)."""
        return self.domain.rect_mask(x0, y0, x1, y1)

    def circle_mask(self, cx, cy, radius):
        """bool (nx, ny): cells whose float32 centre lies within `radius` of (cx, cy) (core/domain.py:262-267)."""
        return self.domain.circle_mask(cx, cy, radius)

    def add_drude_rectangle(self, x0, y0, x1, y1, drude):
        """Drude metal on the rectangle's cells, its eps_inf written as their permittivity
        (physics/em/__init__.py:329-340)."""
        self._add_pole(drude, self.rect_mask(x0, y0, x1, y1), background=True)

    def add_drude_circle(self, cx, cy, radius, drude):
        self._add_pole(drude, self.circle_mask(cx, cy, radius), background=True)

    def add_dispersive(self, pole, mask):
        """Drude or Lorentz pole on the bool (nx, ny) mask over the cells' own permittivity
        (physics/em/__init__.py:342-355)."""
        self._add_pole(pole, mask, background=False)

    def _add_pole(self, pole, mask, background):
        if not isinstance(pole, (DrudePole, LorentzPole)):
            raise TypeError(f'add_dispersive needs a DrudePole or LorentzPole, got {type(pole).__name__}')
        mask = np.asarray(mask)
        if mask.dtype != np.bool_ or mask.shape != self.shape:
            raise ValueError(f'pole mask must be a bool array of shape {self.shape}, got {mask.dtype} {mask.shape}')
        limit = _lib.MAX_POLES if self.polarization == 'ez' else _lib.MAX_POLES // 2   # H_z: one record per face set
        if len(self.poles) >= limit:
            raise ValueError(f'at most {limit} dispersive poles per 2D {self.polarization} model')
        if background:   # eps_r = eps_inf under the mask, the other material values kept
            materials = self.materials
            for old in np.unique(materials.ids[mask]):
                values = materials.materials[int(old)]
                new = materials.material_id(pole.eps_inf, pole.eps_inf, pole.eps_inf, *values[3:])
                materials.ids[mask & (materials.ids == old)] = new
        self.poles.append((pole, mask.copy()))
        self.materials.version += 1

    def pole_weights(self):
        """[(pole, component, float32 weight array)]: E_z, the pole's cell mask on its cells (component 0); H_z, the
        mask's face average on the E_x (1, hi-y) and E_y (2, hi-x) faces (physics/em/yee_grid.py:128-142, uniform
        cells: half of each adjacent cell; the last face of a non-periodic axis pairs its cell with itself), times the
        PEC keep (a cell or face touching PEC carries no current)."""
        pec = self.materials.pec_flags()[self.materials.ids]
        out = []
        for pole, mask in self.poles:
            if self.polarization == 'ez':
                out.append((pole, 0, (mask & ~pec).astype(DTYPE)))
                continue
            m = mask.astype(DTYPE)
            for component, axis in ((1, 1), (2, 0)):
                n = self.shape[axis]
                partner = (np.arange(n) + 1) % n if self.bc.periodic('xy'[axis]) else np.minimum(np.arange(n) + 1,
                                                                                                n - 1)
                take = (slice(None), partner) if axis == 1 else (partner, slice(None))
                weight = DTYPE(0.5) * m + DTYPE(0.5) * m[take]
                weight[pec | pec[take]] = 0.0
                out.append((pole, component, weight))
        return out

    def pole_term(self):
        """float64 (nx, ny): the summed stability terms of the poles on each cell (0 without)."""
        term = np.zeros(self.shape)
        for pole, mask in self.poles:
            term[mask] += pole_step_term(pole, float(self.dt))
        return term

    def add_pec_region(self, mask):
        """Perfect electric conductor on the bool (nx, ny) cell mask (physics/em/__init__.py:372-376): E_z pinned to 0
        in its cells (E_z), every E face touching it pinned to 0 (H_z); kept over every dielectric."""
        mask = np.asarray(mask)
        if mask.dtype != np.bool_ or mask.shape != self.shape:
            raise ValueError(f'PEC mask must be a bool array of shape {self.shape}, got {mask.dtype} {mask.shape}')
        self.pec_regions.append(mask.copy())
        self.materials.version += 1

    def add_pec_rectangle(self, x0, y0, x1, y1):
        """PEC on the cells whose float32 centre lies in [x_min, x_max) x [y_min, y_max) (the raster of
        add_dielectric_rectangle)."""
        mask = np.zeros(self.shape, dtype=bool)
        mask[tuple(slice(lo, hi) for lo, hi in self.domain.box_ranges((x0, y0), (x1, y1)))] = True
        self.add_pec_region(mask)

    def pec_mask(self):
        """bool (nx, ny): PEC regions and, for E_z, the outermost window row of every 'pec' face (an H_z 'pec' face is
        a wall copy in the step kernel instead, physics/em/boundaries.py:500-538)."""
        mask = np.zeros(self.shape, dtype=bool)
        for region in self.pec_regions:
            mask |= region
        window = self.domain.physical_window()
        for face in (self.bc.pec_faces() if self.polarization == 'ez' else ()):
            axis = 'xy'.index(face[0])
            index = window[axis].start if face.endswith('lo') else window[axis].stop - 1
            if axis == 0:
                mask[index, :] = True
            else:
                mask[:, index] = True
        return mask

    def apply_pec(self):
        """Paint the PEC cells over the dielectrics; the material version changes only when a cell changes."""
        mask = self.pec_mask()
        if mask.any():
            self.materials.paint_pec([mask])

    def check_courant(self):
        """ValueError when eps * mu_r < (C0 dt)^2 (1/dx^2 + 1/dy^2) for a material in use (physics/em/__init__.py:
        242-281; eps = eps_r for E_z, min(eps_xx, eps_yy) for H_z)."""
        used = self.materials.used()
        values = self.materials.values()[used[~self.materials.pec_flags()[used]]]   # PEC cells set no limit
        if values.size == 0:
            return
        eps = values[:, 0] if self.polarization == 'ez' else np.minimum(values[:, 0], values[:, 1])
        dx2 = float(self.domain.dx) ** 2
        required = np.float64(DTYPE((C0 * float(self.dt)) ** 2 / dx2)) * 2.0
        if self.poles:   # per cell: eps mu_r >= the light term + mu_r x the poles' terms (physics/em/__init__.py:242-270)
            cells = self.materials.values()[self.materials.ids]
            keep = ~self.materials.pec_flags()[self.materials.ids]
            eps_cells = cells[..., 0] if self.polarization == 'ez' else np.minimum(cells[..., 0], cells[..., 1])
            ratio = eps_cells * cells[..., 4] / (required + cells[..., 4] * self.pole_term())
            worst = float(ratio[keep].min()) if keep.any() else 1.0
            if worst < 1.0 - 1e-6:
                factor = math.floor(float(self.grid.courant) * math.sqrt(max(worst, 0.0)) * 1e3) / 1e3
                raise ValueError(f'a cell with dispersive poles is unstable at this time step (eps mu_r / requirement '
                                 f'= {worst:.4g}); lower courant_factor to <= {factor:.3f}')
        ratio = eps * values[:, 4] / required
        worst = float(ratio.min())
        if worst < 1.0 - 1e-6:
            factor = math.floor(float(self.grid.courant) * math.sqrt(max(worst, 0.0)) * 1e3) / 1e3
            raise ValueError(f'eps * mu_r = {float((eps * values[:, 4]).min()):.4g} is below (C0 dt)^2 (1/dx^2 + '
                             f'1/dy^2) = {required:.4g}, so the time step is unstable; lower courant_factor to <= '
                             f'{factor:.3f}')

    # ---- sources and monitors -------------------------------------------------------------------------------------
    def add_line_source(self, axis, position, wavelength=None, frequency=None, bandwidth=0.5, amplitude=1.0,
                        waveform='gaussian', span=None, waist=None, center=None):
        """Sheet current radiating `amplitude` [V/m] to each side (physics/em/__init__.py:442-473,
        placement.py:223-251): axis 'x' = line along x at y = position; span None = the whole grid line; a span end
        on a cpml window edge continues through the PML."""
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        domain, bc = self.domain, self.bc
        if axis not in ('x', 'y'):
            raise ValueError(f"Line source axis must be 'x' or 'y', got {axis!r}")
        along_axis, normal_axis = (0, 1) if axis == 'x' else (1, 0)
        point = [0.0, 0.0]
        point[normal_axis] = position
        domain.validate_physical_pos(point, f"Line source ({'xy'[normal_axis]})")
        line_cell = domain.pos_to_cell_axis(normal_axis, position)
        n_along = self.grid.shape[along_axis]
        if span is None:
            span_cells = (0, n_along)
        else:
            start, end = sorted((float(span[0]), float(span[1])))
            for coord in (start, end):
                p = list(point)
                p[along_axis] = coord
                domain.validate_physical_pos(p, 'Line source span')
            lo, hi = domain.pos_to_cell_axis(along_axis, start), domain.pos_to_cell_axis(along_axis, end) + 1
            window = domain.physical_window()[along_axis]
            if lo == window.start and bc.faces[f'{axis}_lo'] == 'cpml':
                lo = 0
            if hi == window.stop and bc.faces[f'{axis}_hi'] == 'cpml':
                hi = n_along
            span_cells = (lo, hi)
        if self.polarization == 'ez':
            normal_width = float(domain.cell_widths()[normal_axis][line_cell])
            component = 'ez'
        else:
            n_normal = self.grid.shape[normal_axis]
            periodic = bc.periodic('xy'[normal_axis])
            if line_cell == n_normal - 1 and not periodic:
                raise ValueError(f'Line source: the H_z sheet sits on the hi face of cell {line_cell}, the outer grid '
                                 f'wall; move the line at least one cell inside')
            # dual spacing of the hi face = 1 / the float64 inverse face spacing (placement.py:253-265)
            w = float(self.grid.dx)
            normal_width = 1.0 / (1.0 / (0.5 * (w + w)))
            component = 'ex' if axis == 'x' else 'ey'
        shape = None
        if waist is not None:
            centres = domain.cell_centers()[along_axis][span_cells[0]:span_cells[1]]
            shape = np.exp(-((centres - float(center)) / float(waist)) ** 2)
        source = LineSource(axis, line_cell, temporal, amplitude, span_cells, normal_width, component, shape)
        self.sources.append(source)
        return source

    def add_tfsf(self, x0, y0, x1, y1, wavelength=None, frequency=None, bandwidth=0.5, amplitude=1.0,
                 waveform='gaussian'):
        """+x plane wave of incident peak `amplitude` [V/m] (E_z on ez; E_y on hz, H_z = E_y / eta0) inside the
        total-field box of cells [cell(x_lo), cell(x_hi) + 1) x [cell(y_lo), cell(y_hi) + 1) (physics/em/__init__.py:
        475-490, placement.py:97-150): one scattered-field cell must stay inside the window on every side, except a y
        mirror face of the wave's parity (E_z symmetric), which the box opens onto."""
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        domain = self.domain
        for x, y, label in ((x0, y0, 'corner 0'), (x1, y1, 'corner 1')):
            domain.validate_physical_pos((x, y), f'TFSF box {label}')
        x_lo, x_hi = sorted((x0, x1))
        y_lo, y_hi = sorted((y0, y1))
        cells = (domain.pos_to_cell_axis(0, x_lo), domain.pos_to_cell_axis(0, x_hi) + 1,
                 domain.pos_to_cell_axis(1, y_lo), domain.pos_to_cell_axis(1, y_hi) + 1)
        wx, wy = domain.physical_window()
        reached = {face for face, hit in (('x_lo', cells[0] <= wx.start), ('x_hi', cells[1] >= wx.stop),
                                          ('y_lo', cells[2] <= wy.start), ('y_hi', cells[3] >= wy.stop)) if hit}
        parity = 'symmetric' if self.polarization == 'ez' else 'antisymmetric'
        open_faces = {face for face in reached if face[0] == 'y' and self.bc.faces[face] == parity}
        if {'x_lo', 'x_hi'} & reached or (reached - open_faces):
            raise ValueError(f'TFSF box cells [{cells[0]}, {cells[1]}) x [{cells[2]}, {cells[3]}) need at least one '
                             f'scattered-field cell inside the physical window [{wx.start}, {wx.stop}) x '
                             f'[{wy.start}, {wy.stop}) on every side, except a y mirror face (ez symmetric, hz '
                             f'antisymmetric) the box opens onto')
        if len(self.tfsf) >= _lib.MAX_TFSF:
            raise ValueError(f'at most {_lib.MAX_TFSF} TFSF boxes per 2D model')
        box = PlaneWaveBox(cells, temporal, float(self.grid.dx), float(self.dt), amplitude, self.polarization,
                           open_faces)
        self.tfsf.append(box)
        return box

    def add_gaussian_beam(self, axis, position, waist, center, wavelength=None, frequency=None, bandwidth=0.5,
                          amplitude=1.0, waveform='gaussian', span=None):
        """Line source with the transverse factor exp(-((s - center) / waist)^2) over the float64 centres s of its
        cells; `amplitude` is the peak field at the waist (physics/em/__init__.py:450-457, placement.py:246-249)."""
        if not waist > 0:
            raise ValueError(f'Gaussian beam waist must be > 0, got {waist}')
        return self.add_line_source(axis, position, wavelength, frequency, bandwidth, amplitude, waveform, span,
                                    waist=waist, center=center)

    def _point_cell(self, x, y, label):
        """(ix, iy, cell area from the float32 grid width) of the cell containing (x, y) (placement.py:41-52)."""
        self.domain.validate_physical_pos((x, y), label)
        ix, iy = self.domain.pos_to_cell(x, y)
        width = float(self.grid.dx)
        return ix, iy, width * width

    def _require(self, polarization, name):
        if self.polarization != polarization:
            raise ValueError(f'{name} needs polarization {polarization!r}, this model is {self.polarization!r}')

    def add_point_source(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, current=1e-3,
                         waveform='gaussian'):
        """z filament current `current` [A] at the cell containing (x, y) (E_z only): E_z += cb * I / (dx dy) *
        s(t + dt/2) after the E update (physics/em/__init__.py:411-418)."""
        self._require('ez', 'add_point_source')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        ix, iy, area = self._point_cell(x, y, 'Point source')
        source = PointCurrent('x', iy, ix, temporal, float(current) / area, 'ez')
        self.sources.append(source)
        return source

    def add_magnetic_line_current(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, magnetic_current=1.0,
                                  waveform='gaussian'):
        """z magnetic line current K_m = `magnetic_current` [V] at the cell containing (x, y) (H_z only):
        H_z += db * K_m / (dx dy) * s(t) after the H update (physics/em/__init__.py:420-429)."""
        self._require('hz', 'add_magnetic_line_current')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        ix, iy, area = self._point_cell(x, y, 'Magnetic line current')
        source = PointCurrent('x', iy, ix, temporal, float(magnetic_current) / area, 'hz', field='h')
        self.sources.append(source)
        return source

    def add_electric_dipole(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, dipole_moment=1e-3,
                            orientation=0.0, waveform='gaussian'):
        """In-plane electric dipole M = `dipole_moment` [A] along `orientation` [deg, ccw from +x] in the cell
        containing (x, y) (H_z only), shared over the E_x / E_y faces around the cell; one source record per face
        (physics/em/__init__.py:431-440). Returns the list of face sources."""
        self._require('hz', 'add_electric_dipole')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        ix, iy, _ = self._point_cell(x, y, 'Electric dipole')
        faces = electric_dipole_faces(ix, iy, self.grid.shape, (self.bc.periodic('x'), self.bc.periodic('y')),
                                      float(self.grid.dx), dipole_moment, orientation)
        added = []
        for component, i, j, density in faces:
            if component == 'ex':   # E_x face (i, j): hi y face of cell (i, j), a line along x at y-cell j
                added.append(PointCurrent('x', j, i, temporal, density, 'ex'))
            else:                   # E_y face (i, j): hi x face of cell (i, j), a line along y at x-cell i
                added.append(PointCurrent('y', i, j, temporal, density, 'ey'))
        self.sources.extend(added)
        return added

    def add_flux_monitor(self, axis, position, span, wavelengths=None, frequencies=None):
        """Poynting flux through a line (axis 'x' = along x at y = position, flux > 0 toward +y); span = (start, end)
        along it [m] (placement.py:55-70, physics/em/__init__.py:511-518)."""
        if span is None or len(span) != 2:
            raise ValueError(f'Flux monitor needs span = (start, end) along the line [m], got {span!r}')
        freqs = resolve_frequencies(wavelengths, frequencies)
        lo, hi = sorted((float(span[0]), float(span[1])))
        domain = self.domain
        widths = domain.cell_widths()
        if axis == 'x':
            domain.validate_physical_pos((lo, position), 'Flux monitor start')
            domain.validate_physical_pos((hi, position), 'Flux monitor end')
            normal, line_cell = 'y', domain.pos_to_cell_axis(1, position)
            cells = (domain.pos_to_cell_axis(0, lo), domain.pos_to_cell_axis(0, hi) + 1)
            along_widths = widths[0]
        elif axis == 'y':
            domain.validate_physical_pos((position, lo), 'Flux monitor start')
            domain.validate_physical_pos((position, hi), 'Flux monitor end')
            normal, line_cell = 'x', domain.pos_to_cell_axis(0, position)
            cells = (domain.pos_to_cell_axis(1, lo), domain.pos_to_cell_axis(1, hi) + 1)
            along_widths = widths[1]
        else:
            raise ValueError(f"Flux monitor axis (along the line) must be 'x' or 'y', got {axis!r}")
        monitor = FluxMonitor(normal, line_cell, cells, freqs, float(self.dt), along_widths, self.polarization)
        self.monitors.append(monitor)
        return monitor

    def add_dft_monitor(self, x0, y0, x1, y1, wavelengths=None, frequencies=None, downsample=1):
        """Running DFT over the cells [cell(x_lo), cell(x_hi)] x [cell(y_lo), cell(y_hi)] of a physical box; downsample
        s > 1 samples every s-th cell along x and y (physics/em/__init__.py:500-509, placement.py:97-105)."""
        domain = self.domain
        domain.validate_physical_pos((x0, y0), 'DFT monitor corner 0')
        domain.validate_physical_pos((x1, y1), 'DFT monitor corner 1')
        x_lo, x_hi = sorted((float(x0), float(x1)))
        y_lo, y_hi = sorted((float(y0), float(y1)))
        monitor = DFTMonitor(slice(domain.pos_to_cell_axis(0, x_lo), domain.pos_to_cell_axis(0, x_hi) + 1),
                             slice(domain.pos_to_cell_axis(1, y_lo), domain.pos_to_cell_axis(1, y_hi) + 1),
                             resolve_frequencies(wavelengths, frequencies), float(self.dt), self.polarization,
                             self.bc.periodic('x'), self.bc.periodic('y'), downsample)
        self.dft_monitors.append(monitor)
        return monitor

    def finalize(self):
        """Courant check, sheet calibration and the DFT sample interval shared by every flux and DFT monitor
        (physics/em/__init__.py:181-203)."""
        if len(self.sources) > _lib.MAX_SOURCES:
            raise ValueError(f'at most {_lib.MAX_SOURCES} source records per 2D model (line sources, point currents, '
                             f'dipole faces)')
        pinned = wall_rows(self)
        for monitor in self.monitors:   # the fused flux sample reads E before the wall rows are copied
            normal = 'xy'.index(monitor.normal)
            if monitor.position in pinned[normal] or any(c in pinned[1 - normal] for c in range(*monitor.span)):
                raise ValueError('a flux monitor may not touch a pmc wall row (its E there is pinned after the step)')
        if len(self.monitors) > _lib.MAX_MONITORS:
            raise ValueError(f'at most {_lib.MAX_MONITORS} flux monitors per 2D model')
        if len(self.dft_monitors) > _lib.MAX_DFT:
            raise ValueError(f'at most {_lib.MAX_DFT} DFT monitors per 2D model')
        self.apply_pec()
        self.check_courant()
        self.calibrate()
        spectral = self.monitors + self.dft_monitors
        self.sample_interval = dft_sample_interval(float(self.dt), self.sources + self.tfsf,
                                                   [m.frequencies for m in spectral])
        for monitor in spectral:
            monitor.sample_interval = self.sample_interval

    def calibrate(self):
        """Sheet profiles from the float64 coefficients of the line cells (physics/em/sources.py:582-602); PEC cells
        (E_z) and faces touching PEC (H_z) inject nothing (fdtd_ez_2d.py:175-177: the weight times the keep mask)."""
        values = self.materials.values().copy()
        pec = self.materials.pec_flags()
        values[pec] = (1.0, 1.0, 1.0, 0.0, 1.0, 0.0)   # calibrated as vacuum, then zeroed by the keep mask
        dt = float(self.dt)
        for source in self.sources:
            if source.field == 'h':   # H_z magnetic line current: db of the cell, from the cell's own material
                ids = source.line(self.materials.ids).astype(np.int64)
                raw = self.materials.values()
                _, db = semi_implicit(raw[ids, 4], raw[ids, 5], dt, MU0)
                source.calibrate(db)
                continue
            omega = 2.0 * np.pi * float(source.temporal.f0)
            ids = source.line(self.materials.ids).astype(np.int64)
            mu, sigma_m = values[ids, 4], values[ids, 5]
            if self.polarization == 'ez':
                _, cb = semi_implicit(values[ids, 0], values[ids, 3], dt, EPS0)
                line_eps = complex_permittivity(values[ids, 0], values[ids, 3], omega)
                keep = ~pec[ids]
            else:
                column = 0 if source.component == 'ex' else 1
                normal_axis = 1 if source.axis == 'x' else 0
                n_normal = self.grid.shape[normal_axis]
                upper = (source.position + 1) % n_normal
                if upper == 0 and not self.bc.periodic('xy'[normal_axis]):
                    upper = source.position
                lo, hi = source.span
                upper_ids = (self.materials.ids[lo:hi, upper] if source.axis == 'x'
                             else self.materials.ids[upper, lo:hi]).astype(np.int64)
                eps_face = values[ids, column] * 0.5 + values[upper_ids, column] * 0.5
                sig_face = values[ids, 3] * 0.5 + values[upper_ids, 3] * 0.5
                _, cb = semi_implicit(eps_face, sig_face, dt, EPS0)
                w = float(self.grid.dx)
                line_eps = (complex_permittivity(values[ids, column], values[ids, 3], omega) * w
                            + complex_permittivity(values[upper_ids, column], values[upper_ids, 3], omega) * w
                            ) / (w + w)
                keep = ~(pec[ids] | pec[upper_ids])
            source.calibrate(np.where(keep, cb, 0.0), line_eps, mu, sigma_m, dt)


# E_z outer rows copied from their inner neighbour after each step: pmc and symmetric faces keep E_z even, an
# antisymmetric face odd (physics/em/boundaries.py:478-488); 'pec' rows are PEC cells instead.
WALL_SIGNS = {'pmc': 1, 'symmetric': 1, 'antisymmetric': -1}
# H_z outer rows inside the step kernel (physics/em/boundaries.py:82-93, 500-538): pmc zeroes them, pec and
# antisymmetric faces copy the inner row (H_z even, tangential E on the plane face pinned), symmetric faces copy it
# negated.
HZ_WALL_SIGNS = {'pmc': _lib.WALL_ZERO, 'pec': 1, 'antisymmetric': 1, 'symmetric': -1}


def wall_rows(model):
    """{axis: set of outer rows} of the model's E_z wall rows copied after the step kernel (the fused flux sample reads
    E before that copy); H_z walls act inside the step kernel, before any sample."""
    rows = {0: set(), 1: set()}
    for face, kind in model.bc.faces.items():
        if kind in WALL_SIGNS and model.polarization == 'ez':
            axis = 'xy'.index(face[0])
            rows[axis].add(0 if face.endswith('lo') else model.grid.shape[axis] - 1)
    return rows


def clock_times(dt, n_steps):
    """float64 times after 0..n_steps steps from t = 0, accumulated t += dt as the reference's clocks do
    (solvers/fdtd_2d_base.py:470-488, coupling/runner.py:191-203); cumsum adds sequentially."""
    return np.concatenate(([0.0], np.cumsum(np.full(int(n_steps), float(dt)))))


def steps_for_duration(time, dt, duration):
    """Ticks the runner takes from `time` for `duration`: it stops once t >= t_end - dt/2 (coupling/runner.py:182-203)."""
    dt = float(dt)
    stop = float(time) + float(duration) - 0.5 * dt
    t, n = float(time), 0
    while t < stop:
        t = t + dt
        n += 1
    return n


class Engine2D:
    """Device state and libyee context of one Model2D. torch provides the HBM buffers and the stream."""

    def __init__(self, model, device=None, rows_per_tile=0, graph_steps=0, capacity=1024, lane_columns=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('Engine2D needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU path')
        self.torch = torch
        self.lib = _lib.load()
        self.model = model
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else device)
        self.rows_per_tile = int(rows_per_tile)
        self.lane_columns = int(lane_columns)
        self.graph_steps = int(graph_steps)
        self.n_step = 0
        self.capacity = 0
        self._times = None
        nx, ny = model.shape
        self.pitch = pitch = -(-ny // 4) * 4   # rows padded to whole 16 B groups; padding columns stay 0
        dev, f32 = self.device, torch.float32
        self.fields = torch.zeros((2, 3, nx, pitch), dtype=f32, device=dev)
        self.clock = torch.zeros(1, dtype=torch.int64, device=dev)
        self.inv = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in
                    (model.inv_dx_e, model.inv_dx_h, model.inv_dy_e, model.inv_dy_h)]
        self.cpml = {}
        for axis, other in (('x', ny), ('y', nx)):
            tables = model.cpml[axis]
            if tables.n_rows == 0:
                self.cpml[axis] = None
                continue
            shape = (2, tables.n_rows, pitch) if axis == 'x' else (2, nx, tables.n_rows)
            self.cpml[axis] = dict(
                row_e=torch.from_numpy(tables.row_e).to(dev), row_h=torch.from_numpy(tables.row_h).to(dev),
                b_e=torch.from_numpy(tables.b_e).to(dev), g_e=torch.from_numpy(tables.g_e).to(dev),
                b_h=torch.from_numpy(tables.b_h).to(dev), g_h=torch.from_numpy(tables.g_h).to(dev),
                psi_e=torch.zeros(shape, dtype=f32, device=dev), psi_h=torch.zeros(shape, dtype=f32, device=dev))
        self.profiles = [torch.from_numpy(s.profile.copy()).to(dev) for s in model.sources]
        self.accumulators = [torch.zeros(m.accumulator_shape(), dtype=torch.float64, device=dev)
                             for m in model.monitors]
        for index, monitor in enumerate(model.monitors):
            monitor.engine, monitor.index = self, index
        self.dft_accumulators = [torch.zeros(m.accumulator_shape(), dtype=torch.float64, device=dev)
                                 for m in model.dft_monitors]
        for index, monitor in enumerate(model.dft_monitors):
            monitor.engine, monitor.index = self, index
        self.waves = [None] * len(model.sources)
        self.tfsf_tables = [None] * len(model.tfsf)
        self.phases = [None] * len(model.monitors)
        self.dft_phases = [None] * len(model.dft_monitors)
        self.poles = []
        for pole, component, weight in model.pole_weights():
            box = bounding_box(weight > 0)
            if box is None:
                continue
            (i0, i1), (j0, j1) = box
            coefficients = pole.ade_coefficients(float(model.dt))
            drive = torch.from_numpy(np.ascontiguousarray(weight[i0:i1, j0:j1] * coefficients[-1])).to(dev)
            states = [torch.zeros((i1 - i0, j1 - j0), dtype=f32, device=dev)
                      for _ in range(1 if isinstance(pole, DrudePole) else 2)]
            self.poles.append(dict(pole=pole, component=component, box=box, coefficients=coefficients, drive=drive,
                                   states=states))
        components = COMPONENTS[model.polarization]
        self.probes = ProbeTable(model.probes, [p.cell[0] * pitch + p.cell[1] for p in model.probes],
                                 [components.index(p.field_component) for p in model.probes], dev)
        for index, probe in enumerate(model.probes):
            probe.engine, probe.index = self, index
        self._upload_materials()
        self._ctx = _lib.c_ptr()
        self._ensure_capacity(capacity, create=True)
        if self.graph_steps:
            _lib.check(self.lib.yee_set_graph_steps(self._ctx, self.graph_steps), 'yee_set_graph_steps')

    # ---- device tables --------------------------------------------------------------------------------------------
    def _upload_materials(self):
        torch, model = self.torch, self.model
        ids = np.zeros((model.nx, self.pitch), dtype=np.uint16)
        ids[:, :model.ny] = model.materials.ids
        self.material = torch.from_numpy(ids.view(np.int16)).to(self.device)
        if model.polarization == 'ez':
            self.table = torch.from_numpy(model.materials.table_ez(model.dt)).to(self.device)
            self.table2 = None
        else:
            table, table2 = model.materials.tables_hz(model.dt)
            self.table = torch.from_numpy(table).to(self.device)
            self.table2 = torch.from_numpy(table2).to(self.device)
        self._materials_version = model.materials.version

    def _ensure_capacity(self, steps, create=False):
        """Waveform and DFT phase tables covering absolute steps [0, steps)."""
        if steps <= self.capacity and not create:
            return
        torch, model = self.torch, self.model
        capacity = max(int(steps), 2 * self.capacity, 1024)
        times = clock_times(model.dt, capacity)
        half = times[:-1] + 0.5 * float(model.dt)
        self.waves = [torch.from_numpy(s.wave_table(times[:-1] if s.field == 'h' else half)).to(self.device)
                      for s in model.sources]
        self.tfsf_tables = [tuple(torch.from_numpy(a).to(self.device) for a in box.tables(half)) for box in model.tfsf]
        m = model.sample_interval
        n_samples = capacity // m
        sample_times = times[m * np.arange(1, n_samples + 1)]
        self.phases = [torch.from_numpy(np.ascontiguousarray(mon.phase_table(sample_times))).to(self.device)
                       for mon in model.monitors]
        self.dft_phases = [torch.from_numpy(np.ascontiguousarray(mon.phase_table(sample_times))).to(self.device)
                           for mon in model.dft_monitors]
        self.probes.ensure(capacity)
        self._times = times
        self.capacity = capacity
        self._rebind(create)

    def probe_series(self, index):
        """(times after each step [s], float32 values) of probe `index` over the steps run so far."""
        return self._times[1:self.n_step + 1].copy(), self.probes.series(index, self.n_step)

    def descriptor(self):
        """The yee2d_desc of the current buffers (for the single-instance context or a batch)."""
        model = self.model
        d = _lib.Desc2DC()
        d.nx, d.ny, d.pitch = model.nx, model.ny, self.pitch
        d.polarization = _lib.YEE_EZ if model.polarization == 'ez' else _lib.YEE_HZ
        d.periodic_x, d.periodic_y = int(model.bc.periodic('x')), int(model.bc.periodic('y'))
        d.rows_per_tile = self.rows_per_tile
        d.lane_columns = self.lane_columns
        for s in range(2):
            for c in range(3):
                d.field[s][c] = self.fields[s, c].data_ptr()
        d.material = self.material.data_ptr()
        d.table = self.table.data_ptr()
        d.table2 = self.table2.data_ptr() if self.table2 is not None else None
        d.n_materials = len(model.materials.materials)
        d.dt = float(model.dt)
        d.dt_over_eps0 = float(np.float32(float(model.dt) / EPS0))
        d.inv_dx_e, d.inv_dx_h, d.inv_dy_e, d.inv_dy_h = (t.data_ptr() for t in self.inv)
        d.inv_dx_bulk, d.inv_dy_bulk = model.inv_bulk
        d.layer_x_lo, d.layer_x_hi = model.cpml['x'].layer_bounds()
        d.layer_y_lo, d.layer_y_hi = model.cpml['y'].layer_bounds()
        for a, axis in enumerate('xy'):
            c = self.cpml[axis]
            if c is None:
                d.cpml[a].n_rows = 0
                continue
            d.cpml[a].n_rows = model.cpml[axis].n_rows
            for name in ('row_e', 'row_h', 'b_e', 'g_e', 'b_h', 'g_h'):
                setattr(d.cpml[a], name, c[name].data_ptr())
            for s in range(2):
                d.cpml[a].psi_e[s] = c['psi_e'][s].data_ptr()
                d.cpml[a].psi_h[s] = c['psi_h'][s].data_ptr()
        d.n_sources = len(model.sources)
        for k, source in enumerate(model.sources):
            src = d.source[k]
            src.axis = 0 if source.axis == 'x' else 1
            src.line = source.position
            src.lo, src.hi = source.span
            src.field = _lib.SOURCE_H if source.field == 'h' else _lib.SOURCE_E
            src.profile = self.profiles[k].data_ptr()
            src.wave = self.waves[k].data_ptr()
            src.wave_len = self.waves[k].numel()
        for f, face in enumerate(('x_lo', 'x_hi', 'y_lo', 'y_hi')):
            if model.polarization == 'ez':
                d.wall_sign[f] = WALL_SIGNS.get(model.bc.faces[face], 0)
            else:
                d.wall_sign[f] = HZ_WALL_SIGNS.get(model.bc.faces[face], 0)
        d.n_monitors = len(model.monitors)
        d.sample_interval = model.sample_interval
        for k, monitor in enumerate(model.monitors):
            mon = d.monitor[k]
            mon.normal = 0 if monitor.normal == 'x' else 1
            mon.position = monitor.position
            mon.lo, mon.hi = monitor.span
            mon.n_freq = monitor.n_freqs
            mon.phase = self.phases[k].data_ptr()
            mon.n_samples = self.phases[k].shape[0]
            mon.acc = self.accumulators[k].data_ptr()
        d.n_dft = len(model.dft_monitors)
        for k, monitor in enumerate(model.dft_monitors):
            monitor.fill(d.dft[k], COMPONENTS[model.polarization], self.dft_phases[k], self.dft_accumulators[k])
        d.n_poles = len(self.poles)
        for k, entry in enumerate(self.poles):
            pl = d.pole[k]
            (pl.i0, pl.i1), (pl.j0, pl.j1) = entry['box']
            drude = isinstance(entry['pole'], DrudePole)
            pl.kind = _lib.YEE_DRUDE if drude else _lib.YEE_LORENTZ
            pl.component = entry['component']
            pl.damping = float(entry['coefficients'][0])
            pl.restoring = 0.0 if drude else float(entry['coefficients'][1])
            pl.inv_dt = float(np.float32(1.0) / np.float32(model.dt))
            pl.drive = entry['drive'].data_ptr()
            for q, state in enumerate(entry['states']):
                pl.state[q] = state.data_ptr()
        self.probes.fill(d.probes)
        d.n_tfsf = len(model.tfsf)
        for k, (box, (line, ends)) in enumerate(zip(model.tfsf, self.tfsf_tables)):
            t = d.tfsf[k]
            t.x0, t.x1, t.y0, t.y1 = box.cells
            for q in range(4):
                t.coef_h[q], t.coef_e[q] = float(box.coef_h[q]), float(box.coef_e[q])
            t.line, t.ends = line.data_ptr(), ends.data_ptr()
            t.n_steps = line.shape[0]
        d.clock = self.clock.data_ptr()
        d.step_limit = _lib.INT64_MAX
        return d

    def _rebind(self, create=False):
        d = self.descriptor()
        if create:
            _lib.check(self.lib.yee2d_create(ctypes_byref(d), ctypes_byref(self._ctx)), 'yee2d_create')
        else:
            _lib.check(self.lib.yee2d_update(self._ctx, ctypes_byref(d)), 'yee2d_update')

    def launch_geometry(self):
        """(rows_per_tile, lane_columns) the library resolved for this instance."""
        d = _lib.Desc2DC()
        _lib.check(self.lib.yee2d_get_desc(self._ctx, ctypes_byref(d)), 'yee2d_get_desc')
        return d.rows_per_tile, d.lane_columns

    def materials_changed(self):
        """Re-upload the id grid and coefficient tables after the model's materials changed."""
        self._upload_materials()
        self.profiles = [self.torch.from_numpy(s.profile.copy()).to(self.device) for s in self.model.sources]
        self._rebind()

    # ---- stepping -------------------------------------------------------------------------------------------------
    def stream_handle(self):
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def prepare(self, n_steps):
        """Grow the device tables to cover the next n_steps steps and instantiate the step graph now, so that stepping
        them afterwards only launches (no table upload, no capture)."""
        if self.model.materials.version != self._materials_version:
            self.materials_changed()
        self._ensure_capacity(self.n_step + int(n_steps))
        _lib.check(self.lib.yee_build_graph(self._ctx), 'yee_build_graph')

    def step(self, n_steps):
        """Enqueue n_steps steps on the current stream (asynchronous)."""
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        if self.model.materials.version != self._materials_version:
            self.materials_changed()
        self._ensure_capacity(self.n_step + n_steps)
        _lib.check(self.lib.yee_step(self._ctx, n_steps, self.stream_handle()), 'yee_step')
        self.n_step += n_steps

    def step_timed(self, n_steps):
        """Run n_steps steps with HIP event pairs around the step's tile kernels (synchronous): {'tiles': (average ms,
        cells)} for the bulk and edge tile sets launched side by side from the fork to the join (as yee_step launches
        them), else {'bulk' / 'edge': (average ms per launch, cells per launch)}."""
        n_steps = int(n_steps)
        if self.model.materials.version != self._materials_version:
            self.materials_changed()
        self._ensure_capacity(self.n_step + n_steps)
        kernel_ms = (ctypes.c_double * 2)()
        cells = (ctypes.c_int64 * 2)()
        _lib.check(self.lib.yee_step_timed(self._ctx, n_steps, self.stream_handle(), kernel_ms, cells),
                   'yee_step_timed')
        self.n_step += n_steps
        if cells[1] == 0 and cells[0] == self.model.nx * self.model.ny and kernel_ms[1] == 0.0:
            # both tile sets side by side (the launch yee_step makes): one class, every cell
            return {'tiles': (kernel_ms[0] / max(n_steps, 1), int(cells[0]))}
        return {name: (kernel_ms[k] / max(n_steps, 1), int(cells[k])) for k, name in enumerate(('bulk', 'edge'))}

    @property
    def time(self):
        return float(self._times[self.n_step]) if self.n_step < self._times.size else float(
            clock_times(self.model.dt, self.n_step)[-1])

    def synchronize(self):
        self.torch.cuda.synchronize(self.device)

    def field(self, name):
        """Host float32 copy of a field component of the current state set."""
        index = self.model.components.index(name)
        return self.fields[self.n_step & 1, index, :, :self.model.ny].cpu().numpy()

    def set_field(self, name, values):
        index = self.model.components.index(name)
        self.fields[self.n_step & 1, index, :, :self.model.ny].copy_(
            self.torch.as_tensor(np.asarray(values, dtype=np.float32)))

    def read_accumulator(self, index):
        return self.accumulators[index].cpu().numpy()

    def read_dft_accumulator(self, index):
        return self.dft_accumulators[index].cpu().numpy()

    def close(self):
        if self._ctx:
            self.lib.yee_destroy(self._ctx)
            self._ctx = _lib.c_ptr()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ctypes_byref(obj):
    import ctypes
    return ctypes.byref(obj)
