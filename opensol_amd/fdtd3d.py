"""3D FDTD models (host setup) and their device engine over libyee.so.

Model3D gathers what the reference spreads over Domain3D / MaterialGrid / EMPhysics3D / FDTDEngine3D / cpml_slabs
and the dipole and plane-flux objects (physics/em3d/physics3d.py:144-445, engine3d.py:162-311): geometry, a
material-index grid, per-axis CPML tables and curl spacings, calibrated point dipoles and flux planes, all host
float64 rounded once to float32. Engine3D owns the device buffers (torch tensors on the ROCm device: memory only) and
steps them with one fused HIP launch per step (plus the plane-sample launch).
"""
import ctypes
import math

import numpy as np

from . import _lib
from .boundaries import BoundaryConfig, CPMLAxisTables, curl_spacings
from .constants import C0, DTYPE, EPS0, MU0
from .domain import Domain
from .fdtd2d import clock_times, steps_for_duration
from .materials import MaterialMap, semi_implicit
from .dispersion import DrudePole, LorentzPole, pole_step_term
from .monitors import DFTPlane3D, FieldProbe, FluxBox, FluxPlane, ProbeTable, dft_sample_interval
from .sources import PointDipole, resolve_frequencies, resolve_frequency
from .waveforms import make_temporal

AXES = ('x', 'y', 'z')
COMPONENTS = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')


def resolve_boundaries_3d(boundaries, npml):
    """None: CPML on every face; a dict of face -> 'cpml' / 'periodic' / 'pec' / 'pmc' / 'symmetric' /
    'antisymmetric' (missing faces: cpml); or a BoundaryConfig with z faces (physics/em3d/boundaries3d.py:17-60). A
    mirror face pads one ghost cell beyond the window (physics/em3d/engine3d.py:14-17)."""
    if boundaries is None:
        return BoundaryConfig.all_pml(npml, dimensions=3)
    if isinstance(boundaries, BoundaryConfig):
        if boundaries.dimensions != 3:
            raise ValueError('a 3D model needs a BoundaryConfig with z faces')
        boundaries.validate()
        return boundaries
    if isinstance(boundaries, dict):
        faces = {face: boundaries.get(face, 'cpml') for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
        bc = BoundaryConfig(**faces)
        bc.cpml_npml = npml
        bc.validate()
        return bc
    raise ValueError(f'3D boundaries must be None, a face dict or a BoundaryConfig, got {boundaries!r}')


class Model3D:
    """Host description of one 3D FDTD model on a uniform grid; size_* is the physical window, every cpml face adds
    npml cells of width dx (core/domain3d.py:55-63)."""

    def __init__(self, size_x, size_y, size_z, dx, npml=12, courant=None, boundaries=None):
        self.bc = resolve_boundaries_3d(boundaries, npml)
        self.domain = Domain((size_x, size_y, size_z), dx, self.bc.npml_sides, courant)
        self.grid = self.domain.grid
        self.shape = self.grid.shape
        self.nx, self.ny, self.nz = self.shape
        self.dt = self.grid.dt
        self.periodic = tuple(self.bc.periodic(a) for a in AXES)
        self.materials = MaterialMap(self.shape)
        self.sources = []
        self.monitors = []
        dx32 = float(self.grid.dx)
        self.kernel_widths = tuple(np.full(n, dx32) for n in self.shape)
        self.cpml = {}
        self.inv_e, self.inv_h = [], []
        for axis, n, periodic in zip(AXES, self.shape, self.periodic):
            tables = CPMLAxisTables(self.bc, axis, n, dx32, float(self.dt))
            self.cpml[axis] = tables
            inv_e, inv_h = curl_spacings(n, dx32, tables, zero_last_h=not periodic)
            self.inv_e.append(inv_e)
            self.inv_h.append(inv_h)
        self.sample_interval = 1
        self.pec_boxes = []   # index ranges of PEC boxes, painted over the dielectrics at finalize
        self.poles = []       # (DrudePole or LorentzPole, bool (nx, ny, nz) mask)
        self.probes = []
        self.dft_monitors = []

    def axis_rows(self, axis):
        """(n, 8) float32 records of one axis (the yee3d_axis_row layout): inv_h, inv_e, psi row of the hi-face
        kind (int32 bits, -1 outside the layers), its b and gain, psi row of the centre kind, its b and gain."""
        name = AXES[axis]
        tables = self.cpml[name]
        n = self.shape[axis]
        out = np.zeros((n, 8), dtype=DTYPE)
        ints = out.view(np.int32)
        out[:, 0], out[:, 1] = self.inv_h[axis], self.inv_e[axis]
        ints[:, 2], ints[:, 5] = tables.row_h, tables.row_e
        for column, rows, b, g in ((3, tables.row_h, tables.b_h, tables.g_h), (6, tables.row_e, tables.b_e, tables.g_e)):
            inside = rows >= 0
            out[inside, column] = b[rows[inside]]
            out[inside, column + 1] = g[rows[inside]]
        return out

    @property
    def cell_count(self):
        return self.nx * self.ny * self.nz

    # ---- materials ------------------------------------------------------------------------------------------------
    def add_dielectric_box(self, lo, hi, eps_r, sigma_e=0.0):
        """Cells whose float32 centre, clamped to the window, lies in [min, max) on every axis get eps_r and sigma_e
        (the 3D model builder's box raster, model/geometry.py:1762-1788; later writes win)."""
        ranges = self.domain.box_ranges(lo, hi, clamp_to_physical=True)
        self.materials.paint_box(ranges, self.materials.material_id(eps_r, eps_r, eps_r, sigma_e, 1.0, 0.0))

    def add_pec_box(self, lo, hi):
        """PEC cells on the box raster of add_dielectric_box (physics/em3d/physics3d.py:351-353). PEC regions are
        applied over every dielectric, whatever the order of the calls (engine3d.py:339-362)."""
        self.pec_boxes.append(self.domain.box_ranges(lo, hi, clamp_to_physical=True))
        self.materials.version += 1   # a built engine re-applies the regions on its next step

    def pec_wall_layers(self):
        """Index tuples of the outermost window cell layer of every 'pec' face (physics/em3d/boundaries3d.py:149-158)."""
        window = self.domain.physical_window()
        layers = []
        for face in self.bc.pec_faces():
            axis = AXES.index(face[0])
            index = window[axis].start if face.endswith('lo') else window[axis].stop - 1
            key = [slice(None)] * 3
            key[axis] = slice(index, index + 1)
            layers.append(tuple(key))
        return layers

    def pec_wall_cell(self, cell, axes):
        """Whether a cell lies in the pec wall layer of a face normal to one of `axes` (engine3d.py:379-382)."""
        for face in self.bc.pec_faces():
            axis = AXES.index(face[0])
            window = self.domain.physical_window()[axis]
            index = window.start if face.endswith('lo') else window.stop - 1
            if axis in axes and cell[axis] == index:
                return True
        return False

    def apply_pec(self):
        """Paint the PEC boxes, then the pec wall layers, over the dielectrics (engine3d.py:358-362); the material
        version changes only when a cell changes."""
        if not self.pec_boxes and not self.bc.pec_faces():
            return
        keys = [tuple(slice(lo, hi) for lo, hi in ranges) for ranges in self.pec_boxes] + self.pec_wall_layers()
        self.materials.paint_pec([key for key in keys if self.materials.ids[key].size])

    def add_dispersive(self, pole, window, mask):
        """Drude or Lorentz pole on the bool mask of the cells of `window` (a slice per axis) over their own
        permittivity (physics/em3d/physics3d.py:346-349)."""
        if not isinstance(pole, (DrudePole, LorentzPole)):
            raise TypeError(f'add_dispersive needs a DrudePole or LorentzPole, got {type(pole).__name__}')
        full = np.zeros(self.shape, dtype=bool)
        window = tuple(window)
        mask = np.asarray(mask)
        if len(window) != 3 or mask.dtype != np.bool_ or mask.shape != full[window].shape:
            raise ValueError(f'pole mask must be a bool array of the window shape {full[window].shape}')
        if 3 * (len(self.poles) + 1) > _lib.MAX_POLES_3D:
            raise ValueError(f'at most {_lib.MAX_POLES_3D // 3} dispersive poles per 3D model')
        full[window] = mask
        self.poles.append((pole, full))
        self.materials.version += 1

    def pole_weights(self):
        """[(pole, component, float32 weight)]: per E component the face average of the pole's cell mask along that
        axis (physics/em3d/engine3d.py:461-488; uniform cells: half of each adjacent cell, the last face of a
        non-periodic axis pairs its cell with itself)."""
        out = []
        for pole, mask in self.poles:
            m = mask.astype(DTYPE)
            for axis in range(3):
                n = self.shape[axis]
                partner = (np.arange(n) + 1) % n if self.periodic[axis] else np.minimum(np.arange(n) + 1, n - 1)
                out.append((pole, axis, DTYPE(0.5) * m + DTYPE(0.5) * np.take(m, partner, axis=axis)))
        return out

    def check_courant(self):
        """ValueError when eps_r mu_r < (C0 dt)^2 sum 1/d^2 (+ mu_r x the poles' terms in their cells) for a material
        in use (physics/em3d/engine3d.py:824-852); PEC cells do not step E and set no limit."""
        used = self.materials.used()
        values = self.materials.values()[used[~self.materials.pec_flags()[used]]]
        if values.size == 0:
            return
        d = float(self.grid.dx)
        required = (C0 * float(self.dt)) ** 2 * 3.0 / d ** 2
        worst = float((values[:, 0] * values[:, 4]).min() / required)
        if self.poles:
            term = np.zeros(self.shape)
            for pole, mask in self.poles:
                term[mask] += pole_step_term(pole, float(self.dt))
            cells = self.materials.values()[self.materials.ids]
            keep = ~self.materials.pec_flags()[self.materials.ids]
            ratio = cells[..., 0] * cells[..., 4] / (required + cells[..., 4] * term)
            worst = min(worst, float(ratio[keep].min()) if keep.any() else 1.0)
        if worst < 1.0 - 1e-6:
            factor = math.floor(float(self.grid.courant) * math.sqrt(max(worst, 0.0)) * 1e3) / 1e3
            raise ValueError(f'eps_r * mu_r is below (C0 dt)^2 sum 1/d^2, so the time step is unstable; lower '
                             f'courant_factor to <= {factor:.3f}')

    # ---- sources and monitors -------------------------------------------------------------------------------------
    def nearest_face(self, axis, coordinate):
        """E face (hi face of cell f) nearest the coordinate among the window's inner faces
        (physics/em3d/sources3d.py:31-38)."""
        window = self.domain.physical_window()[axis]
        first, last = window.start, window.stop - 2
        if last < first:
            raise ValueError('a dipole along an axis needs at least 2 window cells on that axis')
        edges = self.domain.cell_edges()[axis]
        return min(range(first, last + 1), key=lambda f: abs(float(edges[f + 1]) - float(coordinate)))

    def face_dual_volume(self, axis, cell):
        """Edge between the face's two centres times the cell face [m^3] (physics/em3d/engine3d.py:299-311)."""
        lengths = []
        for other, widths in enumerate(self.kernel_widths):
            index = cell[other]
            partner = (index + 1) % widths.size if self.periodic[other] else min(index + 1, widths.size - 1)
            lengths.append(0.5 * (widths[index] + widths[partner]) if other == axis else widths[index])
        return float(math.prod(lengths))

    def add_dipole(self, x, y, z, orientation='z', dipole_moment=1e-12, wavelength=None, frequency=None,
                   bandwidth=0.5, waveform='gaussian'):
        """Point dipole of moment I l [A m] along x / y / z on the E face nearest (x, y, z): the face along the dipole
        nearest the coordinate, the containing cell on the other two axes (physics/em3d/physics3d.py:403-422)."""
        if orientation not in AXES:
            raise ValueError(f"dipole orientation must be one of {AXES}, got {orientation!r}")
        self.domain.validate_physical_pos((x, y, z), 'Dipole')
        temporal = make_temporal(waveform, resolve_frequency(wavelength, frequency), bandwidth)
        axis = AXES.index(orientation)
        cell = list(self.domain.pos_to_cell(x, y, z))
        cell[axis] = self.nearest_face(axis, (x, y, z)[axis])
        if self.pec_wall_cell(cell, tuple(other for other in range(3) if other != axis)):
            raise ValueError(f'dipole at ({x:.4g}, {y:.4g}, {z:.4g}) m lies in the outermost cell layer of a pec face '
                             f'parallel to it, where its tangential E is pinned to 0; move it inside')
        source = PointDipole('e' + orientation, cell, temporal, dipole_moment, self.face_dual_volume(axis, cell))
        self.sources.append(source)
        return source

    def add_field_probe(self, x, y, z, component='ez'):
        """Raw Yee samples of one component at the cell containing (x, y, z) (physics/em3d/physics3d.py:425-433)."""
        if component not in COMPONENTS:
            raise ValueError(f'probe component must be one of {list(COMPONENTS)}, got {component!r}')
        self.domain.validate_physical_pos((x, y, z), 'Field probe')
        probe = FieldProbe(self.domain.pos_to_cell(x, y, z), component)
        self.probes.append(probe)
        return probe

    def nearest_edge(self, axis, coordinate):
        """Grid edge nearest the coordinate among the window's edges (physics/em3d/physics3d.py:492-496)."""
        window = self.domain.physical_window()[axis]
        edges = self.domain.cell_edges()[axis]
        return window.start + int(np.argmin(np.abs(edges[window.start:window.stop + 1] - float(coordinate))))

    def add_flux_plane(self, normal, position, wavelengths=None, frequencies=None):
        """Flux through the cell-edge plane nearest `position` [m] along `normal` across the whole physical window;
        > 0 toward +normal (physics/em3d/physics3d.py:434-445)."""
        if normal not in AXES:
            raise ValueError(f"flux plane normal must be one of {AXES}, got {normal!r}")
        axis = AXES.index(normal)
        bounds = [(w.start, w.stop) for w in self.domain.physical_window()]
        edge = self.nearest_edge(axis, position)
        bounds[axis] = (edge, edge)
        monitor = FluxPlane(axis, edge, bounds, resolve_frequencies(wavelengths, frequencies), float(self.dt),
                            self.kernel_widths, self.shape, self.periodic)
        self.monitors.append(monitor)
        return monitor

    def add_flux_box(self, lo, hi, wavelengths=None, frequencies=None):
        """Net outward flux through the box between the window cell edges nearest the corners lo, hi [m]: six bounded
        flux planes (physics/em3d/physics3d.py:447-476)."""
        bounds = []
        for axis in range(3):
            first = self.nearest_edge(axis, min(lo[axis], hi[axis]))
            last = self.nearest_edge(axis, max(lo[axis], hi[axis]))
            if last <= first:
                raise ValueError(f'flux box spans no cell along {AXES[axis]}: corners {lo[axis]:.4g} and '
                                 f'{hi[axis]:.4g} m snap to the same cell edge')
            bounds.append((first, last))
        if len(self.monitors) + 6 > _lib.MAX_PLANES_3D:
            raise ValueError(f'a flux box needs six of the {_lib.MAX_PLANES_3D} flux planes of a 3D model')
        freqs = resolve_frequencies(wavelengths, frequencies)
        faces = []
        for axis in range(3):
            for side, sign in ((0, -1.0), (1, 1.0)):
                plane = FluxPlane(axis, bounds[axis][side], bounds, freqs, float(self.dt), self.kernel_widths,
                                  self.shape, self.periodic)
                self.monitors.append(plane)
                faces.append((sign, plane))
        return FluxBox(faces)

    def add_dft_plane(self, z, wavelengths=None, frequencies=None):
        """Running DFT of E at the cell centres of the window layer containing height z [m]
        (physics/em3d/physics3d.py:478-486)."""
        self.domain.validate_physical_pos((0.0, 0.0, z), 'DFT plane')
        monitor = DFTPlane3D(self.domain.physical_window(), self.domain.pos_to_cell_axis(2, z),
                             resolve_frequencies(wavelengths, frequencies), float(self.dt), self.periodic)
        self.dft_monitors.append(monitor)
        return monitor

    # ---- coefficients ---------------------------------------------------------------------------------------------
    def table(self):
        """(n_materials, 8) float32 per id: (da, db, ca, cb, eps_r, sigma_e, 0, 0) -- the H edge coefficients
        (kernels3d.py:44-56), the E face coefficients between two cells of the id, and the float32 values from which
        a face between two ids forms its mean (kernels3d.py:61-85)."""
        v = self.materials.values()
        dt = float(self.dt)
        da, db = semi_implicit(v[:, 4], v[:, 5], dt, MU0)
        ca, cb = semi_implicit(v[:, 0], v[:, 3], dt, EPS0)
        pec = self.materials.pec_flags()
        ca[pec] = cb[pec] = 0.0   # a face between two PEC cells is pinned; eps_r = -1 marks the id for the kernel
        zeros = np.zeros_like(da)
        return np.stack([da, db, ca, cb, v[:, 0], v[:, 3], zeros, zeros], axis=1).astype(DTYPE)

    def face_cb(self, axis, cell):
        """float64 cb of the E face along axis at cell: mean eps_r / sigma_e of the cell and its hi partner."""
        values = self.materials.values()
        cell = tuple(int(v) for v in cell)
        n = self.shape[axis]
        partner = list(cell)
        partner[axis] = (cell[axis] + 1) % n if self.periodic[axis] else min(cell[axis] + 1, n - 1)
        a, b = values[int(self.materials.ids[cell])], values[int(self.materials.ids[tuple(partner)])]
        if a[0] < 0 and b[0] < 0:   # PEC rule (kernels3d.py:61-85): two PEC cells pin the face
            return 0.0
        if a[0] < 0:   # one PEC cell: the face takes the other cell's values
            a = b
        elif b[0] < 0:
            b = a
        _, cb = semi_implicit(0.5 * a[0] + 0.5 * b[0], 0.5 * a[3] + 0.5 * b[3], float(self.dt), EPS0)
        return float(cb)

    def finalize(self):
        """Courant check, dipole calibration and the DFT sample interval (physics/em3d/physics3d.py:192-217)."""
        if len(self.sources) > _lib.MAX_SOURCES_3D:
            raise ValueError(f'at most {_lib.MAX_SOURCES_3D} dipoles per 3D model')
        if len(self.monitors) > _lib.MAX_PLANES_3D:
            raise ValueError(f'at most {_lib.MAX_PLANES_3D} flux planes per 3D model')
        if len(self.dft_monitors) > _lib.MAX_DFT:
            raise ValueError(f'at most {_lib.MAX_DFT} DFT planes per 3D model')
        self.apply_pec()
        self.check_courant()
        for source in self.sources:
            source.calibrate(self.face_cb(source.axis, source.cell))
        spectral = self.monitors + self.dft_monitors
        self.sample_interval = dft_sample_interval(float(self.dt), self.sources, [m.frequencies for m in spectral])
        for monitor in spectral:
            monitor.sample_interval = self.sample_interval


class Engine3D:
    """Device state and libyee context of one Model3D. torch provides the HBM buffers and the stream."""

    def __init__(self, model, device=None, planes_per_tile=0, graph_steps=0, capacity=1024):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('Engine3D needs a ROCm GPU (torch.cuda.is_available() is False); there is no CPU path')
        self.torch = torch
        self.lib = _lib.load()
        self.model = model
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else device)
        self.planes_per_tile = int(planes_per_tile)
        self.graph_steps = int(graph_steps)
        self.n_step = 0
        self.capacity = 0
        self._times = None
        dev, f32 = self.device, torch.float32
        shape = model.shape
        self.fields = torch.zeros((2, 6) + shape, dtype=f32, device=dev)
        self.clock = torch.zeros(1, dtype=torch.int64, device=dev)
        self.axis_rows = [torch.from_numpy(model.axis_rows(a)).to(dev) for a in range(3)]
        self.cpml = {}
        for a, axis in enumerate(AXES):
            tables = model.cpml[axis]
            if tables.n_rows == 0:
                self.cpml[axis] = None
                continue
            psi_shape = list(shape)
            psi_shape[a] = tables.n_rows
            self.cpml[axis] = dict(
                psi_e=torch.zeros([2] + psi_shape, dtype=f32, device=dev),
                psi_h=torch.zeros([2, 2] + psi_shape, dtype=f32, device=dev))
        self.accumulators = [[torch.zeros(m.accumulator_shape(q), dtype=torch.float64, device=dev) for q in range(2)]
                             for m in model.monitors]
        for index, monitor in enumerate(model.monitors):
            monitor.engine, monitor.index = self, index
        self.dft_accumulators = [torch.zeros(m.accumulator_shape(), dtype=torch.float64, device=dev)
                                 for m in model.dft_monitors]
        for index, monitor in enumerate(model.dft_monitors):
            monitor.engine, monitor.index = self, index
        self.waves = [None] * len(model.sources)
        self.phases = [None] * len(model.monitors)
        self.dft_phases = [None] * len(model.dft_monitors)
        self.poles = []
        for pole, component, weight in model.pole_weights():
            nonzero = [np.nonzero(np.any(weight > 0, axis=tuple(b for b in range(3) if b != a)))[0] for a in range(3)]
            if nonzero[0].size == 0:
                continue
            box = tuple((int(v[0]), int(v[-1]) + 1) for v in nonzero)
            coefficients = pole.ade_coefficients(float(model.dt))
            crop = tuple(slice(lo, hi) for lo, hi in box)
            drive = torch.from_numpy(np.ascontiguousarray(weight[crop] * coefficients[-1])).to(dev)
            states = [torch.zeros(drive.shape, dtype=torch.float32, device=dev)
                      for _ in range(1 if isinstance(pole, DrudePole) else 2)]
            self.poles.append(dict(pole=pole, component=component, box=box, coefficients=coefficients, drive=drive,
                                   states=states))
        ny, nz = shape[1], shape[2]
        self.probes = ProbeTable(model.probes, [(p.cell[0] * ny + p.cell[1]) * nz + p.cell[2] for p in model.probes],
                                 [COMPONENTS.index(p.field_component) for p in model.probes], dev)
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
        self.material = torch.from_numpy(model.materials.ids.view(np.int16)).to(self.device)
        self.table = torch.from_numpy(model.table()).to(self.device)
        self._materials_version = model.materials.version

    def _ensure_capacity(self, steps, create=False):
        """Waveform and DFT phase tables covering absolute steps [0, steps)."""
        if steps <= self.capacity and not create:
            return
        torch, model = self.torch, self.model
        capacity = max(int(steps), 2 * self.capacity, 1024)
        times = clock_times(model.dt, capacity)
        half = times[:-1] + 0.5 * float(model.dt)
        self.waves = [torch.from_numpy(s.wave_table(half)).to(self.device) for s in model.sources]
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
        model = self.model
        d = _lib.Desc3DC()
        d.nx, d.ny, d.nz = model.shape
        for a in range(3):
            d.periodic[a] = int(model.periodic[a])
        d.rows_per_tile = 0
        d.planes_per_tile = self.planes_per_tile
        for s in range(2):
            for c in range(6):
                d.field[s][c] = self.fields[s, c].data_ptr()
        d.material = self.material.data_ptr()
        d.table = self.table.data_ptr()
        d.n_materials = len(model.materials.materials)
        d.dt = float(model.dt)
        for a, axis in enumerate(AXES):
            d.axis_rows[a] = self.axis_rows[a].data_ptr()
            c = self.cpml[axis]
            if c is None:
                d.cpml[a].n_rows = 0
                continue
            d.cpml[a].n_rows = model.cpml[axis].n_rows
            for q in range(2):
                d.cpml[a].psi_e[q] = c['psi_e'][q].data_ptr()
                for s in range(2):
                    d.cpml[a].psi_h[s][q] = c['psi_h'][s, q].data_ptr()
        d.n_sources = len(model.sources)
        for k, source in enumerate(model.sources):
            src = d.source[k]
            src.component = source.axis
            src.i, src.j, src.k = source.cell
            src.weight = float(source.weight)
            src.wave = self.waves[k].data_ptr()
            src.wave_len = self.waves[k].numel()
        d.n_planes = len(model.monitors)
        d.sample_interval = model.sample_interval
        for k, monitor in enumerate(model.monitors):
            pl = d.plane[k]
            pl.normal, pl.edge = monitor.normal, monitor.edge
            for q, pair in enumerate(monitor.pairs):
                pl.e_comp[q], pl.h_comp[q] = pair['e_comp'], pair['h_comp']
                for t in range(2):
                    pl.lo[q][t], pl.hi[q][t] = pair['lo'][t], pair['hi'][t]
                pl.acc[q] = self.accumulators[k][q].data_ptr()
            pl.n_freq = monitor.n_freqs
            pl.phase = self.phases[k].data_ptr()
            pl.n_samples = self.phases[k].shape[0]
        d.n_dft = len(model.dft_monitors)
        for k, monitor in enumerate(model.dft_monitors):
            monitor.fill(d.dft[k], COMPONENTS, self.dft_phases[k], self.dft_accumulators[k])
        d.n_poles = len(self.poles)
        for k, entry in enumerate(self.poles):
            pl = d.pole[k]
            (pl.i0, pl.i1), (pl.j0, pl.j1), (pl.k0, pl.k1) = entry['box']
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
        kinds = {'symmetric': _lib.MIRROR_SYMMETRIC, 'antisymmetric': _lib.MIRROR_ANTISYMMETRIC}
        for f, face in enumerate(('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')):
            d.mirror[f] = kinds.get(model.bc.faces[face], 0)
        d.clock = self.clock.data_ptr()
        d.step_limit = _lib.INT64_MAX
        return d

    def _rebind(self, create=False):
        d = self.descriptor()
        if create:
            _lib.check(self.lib.yee3d_create(ctypes.byref(d), ctypes.byref(self._ctx)), 'yee3d_create')
        else:
            _lib.check(self.lib.yee3d_update(self._ctx, ctypes.byref(d)), 'yee3d_update')

    def launch_geometry(self):
        """(rows_per_tile, planes_per_tile) the library resolved for this instance."""
        d = _lib.Desc3DC()
        _lib.check(self.lib.yee3d_get_desc(self._ctx, ctypes.byref(d)), 'yee3d_get_desc')
        return d.rows_per_tile, d.planes_per_tile

    def materials_changed(self):
        self._upload_materials()
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
        """Run n_steps steps with a HIP event pair around every kernel (synchronous): {'step' / 'sample': (average
        ms per launch, cells per launch)}."""
        n_steps = int(n_steps)
        if self.model.materials.version != self._materials_version:
            self.materials_changed()
        self._ensure_capacity(self.n_step + n_steps)
        kernel_ms = (ctypes.c_double * 2)()
        cells = (ctypes.c_int64 * 2)()
        _lib.check(self.lib.yee_step_timed(self._ctx, n_steps, self.stream_handle(), kernel_ms, cells),
                   'yee_step_timed')
        self.n_step += n_steps
        return {name: (kernel_ms[k] / max(n_steps, 1), int(cells[k])) for k, name in enumerate(('step', 'sample'))}

    @property
    def time(self):
        return float(self._times[self.n_step]) if self.n_step < self._times.size else float(
            clock_times(self.model.dt, self.n_step)[-1])

    def synchronize(self):
        self.torch.cuda.synchronize(self.device)

    def field(self, name):
        """Host float32 copy of a field component of the current state set."""
        return self.fields[self.n_step & 1, COMPONENTS.index(name)].cpu().numpy()

    def set_field(self, name, values):
        self.fields[self.n_step & 1, COMPONENTS.index(name)].copy_(
            self.torch.as_tensor(np.asarray(values, dtype=np.float32)))

    def read_plane_accumulator(self, index, pair):
        return self.accumulators[index][pair].cpu().numpy()

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


__all__ = ['Model3D', 'Engine3D', 'resolve_boundaries_3d', 'steps_for_duration']
