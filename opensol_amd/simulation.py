"""Simulation: the reference's 2D EM-FDTD scripting API (simulation.py:18-399, physics/em/__init__.py:63-557) on the
MI355X Yee stepper.

    sim = Simulation(size_x=6e-6, size_y=3e-6, dx=25e-9, npml=15, boundaries='periodic_y')
    sim.add_dielectric_rectangle(2.75e-6, 0, 3.25e-6, 3e-6, eps_r=2.25)
    source = sim.add_line_source(axis='y', position=1.2e-6, wavelength=600e-9, bandwidth=0.8)
    monitor = sim.add_flux_monitor(axis='y', position=4.5e-6, span=(0, 3e-6), wavelengths=[500e-9, 600e-9])
    sim.run(duration=80e-15)
    spectrum = monitor.flux()

Stepping runs on the GPU only (libyee.so through ctypes); the host keeps the float64 clock, the calibration and the
flux readout, as the reference's host code does.
"""
import time as _time
from types import MappingProxyType

import numpy as np

from .constants import DTYPE
from .fdtd2d import Engine2D, Model2D, steps_for_duration

MATERIAL_PROPERTIES = ('eps_r', 'sigma_e')   # the material properties of the FDTD path


def material_values(material, dispersion, properties):
    """(eps_r, sigma_e, poles) of add_material's arguments (simulation.py:104-177) on the FDTD path: numeric eps_r /
    sigma_e and a dispersion list of DrudePole / LorentzPole objects or their model-JSON dicts; without eps_r the
    cells take the background permittivity of the poles. Library materials are refused."""
    from .dispersion import DrudePole, LorentzPole, background_permittivity, pole_from_dict
    if material is not None:
        raise ValueError('add_material: library materials are not part of the FDTD path; give eps_r, sigma_e and '
                         'dispersion')
    unknown = sorted(set(properties) - set(MATERIAL_PROPERTIES))
    if unknown:
        raise ValueError(f'unknown material properties {unknown}; valid: {", ".join(MATERIAL_PROPERTIES)}')
    for name, v in properties.items():
        if v is not None and (isinstance(v, (str, bytes, bool)) or np.ndim(v) != 0):
            raise ValueError(f'material property {name}={v!r}: add_material takes numbers')
    poles = [pole if isinstance(pole, (DrudePole, LorentzPole)) else pole_from_dict(pole) for pole in dispersion or []]
    eps = properties.get('eps_r')
    sigma = properties.get('sigma_e')
    return (background_permittivity(poles) if eps is None else float(eps),
            0.0 if sigma is None else float(sigma), poles)


def _device_bytes(obj, seen, visited=None):
    """Bytes of the torch device tensors reachable from obj through containers and this package's objects, each
    storage once (seen holds counted storage addresses); modules, libraries and foreign objects are not walked."""
    import torch
    visited = set() if visited is None else visited
    if isinstance(obj, torch.Tensor):
        if not obj.is_cuda:
            return 0
        storage = obj.untyped_storage()
        if storage.data_ptr() in seen:
            return 0
        seen.add(storage.data_ptr())
        return int(storage.nbytes())
    if id(obj) in visited:
        return 0
    if isinstance(obj, dict):
        items = list(obj.values())
    elif isinstance(obj, (list, tuple)):
        items = obj
    elif type(obj).__module__.startswith(__package__ + '.') and hasattr(obj, '__dict__'):
        items = list(vars(obj).values())
    else:
        return 0
    visited.add(id(obj))
    return sum(_device_bytes(item, seen, visited) for item in items)


def _format_bytes(n):
    return f'{n / 2**30:.2f} GiB' if n >= 2**30 else f'{n / 2**20:.1f} MiB'


def memory_report(sim, file=None, physics='em'):
    """Print (to file, default stdout) and return the memory status and the device bytes this simulation owns
    (simulation.py:205-245): 'materials' the per-cell material ids and coefficient tables, `physics` ('em', 'em3d')
    every other array of the engine (both field sets, CPML psi, pole, source and monitor tables, accumulators, probes), 'couplings' 0 (none
    on this path); a storage shared by two owners counts for the first. The status dict has the reference's keys
    (pool_used, pool_reserved, device_free, device_total, budget, other, host_available, kind); no budget is set on
    this path, and host_owners is empty (the FDTD path keeps no CPU factorisations)."""
    import psutil
    import torch
    if sim.model is None:
        raise RuntimeError('simulation cleared')
    engine = sim.engine
    seen = set()
    owners = {'materials': 0, physics: 0, 'couplings': 0}
    if engine is not None:
        for name in ('material', 'table', 'table2'):
            owners['materials'] += _device_bytes(getattr(engine, name, None), seen)
        owners[physics] = _device_bytes(engine, seen)
    host = psutil.virtual_memory()
    if torch.cuda.is_available():
        free, total = torch.cuda.mem_get_info()
        reserved = int(torch.cuda.memory_reserved())
        status = {'pool_used': int(torch.cuda.memory_allocated()), 'pool_reserved': reserved, 'device_free': int(free),
                  'device_total': int(total), 'budget': None, 'other': max(0, int(total) - int(free) - reserved),
                  'host_available': int(host.available), 'kind': 'gpu'}
        lines = ['GPU memory (pool counters cover the whole process)',
                 *(f'  {key.replace("_", " "):<14} {_format_bytes(status[key])}'
                   for key in ('pool_used', 'pool_reserved', 'other', 'device_free', 'device_total')),
                 '  budget         not set', 'Owned by this simulation (device arrays)']
    else:
        status = {'pool_used': 0, 'pool_reserved': 0, 'device_free': int(host.available), 'device_total':
                  int(host.total), 'budget': None, 'other': int(host.total - host.available),
                  'host_available': int(host.available), 'kind': 'cpu'}
        lines = ['Host memory (no GPU visible)', f'  available      {_format_bytes(status["device_free"])}',
                 f'  total          {_format_bytes(status["device_total"])}', 'Owned by this simulation (device arrays)']
    lines += [f'  {name:<14} {_format_bytes(size)}' for name, size in owners.items()]
    print('\n'.join(lines), file=file)
    return {'status': status, 'owners': owners, 'host_owners': {}}


def mesh_info(sim):
    """Print the window and grid cell counts, the CPML padding per face, the cell sizes and the grid's time step
    (simulation.py:359-378; the grid here is always uniform)."""
    if sim.model is None:
        raise RuntimeError('simulation cleared')
    domain = sim.domain
    counts = [w.stop - w.start for w in domain.physical_window()]
    extent = domain.physical_extent()[1::2]
    axes = domain.axes
    print('\n--- Mesh info (uniform) ---')
    print(f"  Physical domain: {' x '.join(f'{e * 1e6:.3f}' for e in extent)} µm, "
          f"{' x '.join(str(n) for n in counts)} = {int(np.prod(counts)):,} cells")
    print(f"  Total grid: {' x '.join(str(n) for n in domain.shape)} = {int(np.prod(domain.shape)):,} cells, "
          f'PML cells per side {domain.npml_sides}')
    print('  ' + ', '.join(f'd{a}: {w.min() * 1e9:.2f}-{w.max() * 1e9:.2f} nm' for a, w in zip(axes,
                                                                                         domain.cell_widths())))
    print(f'  Grid dt = {float(domain.grid.dt):.4e} s')


def _clear(sim):
    """Close the engine, unbind the model's monitors and probes from it (so that handles the caller keeps hold no
    device arrays) and drop the model."""
    sim.close()
    model = sim.model
    if model is not None:
        for group in ('monitors', 'dft_monitors', 'probes'):
            for item in getattr(model, group, ()):
                item.engine = None
    sim.model = None


def _add_material(sim, mask, material, dispersion, properties, add_pole):
    if sim.model is None:
        raise RuntimeError('simulation cleared')
    model = sim.model
    mask = np.asarray(mask)
    if mask.dtype != np.bool_ or mask.shape != tuple(model.shape):
        raise ValueError(f'add_material needs a bool mask of the full grid {tuple(model.shape)} (domain.rect_mask / '
                         f'circle_mask), got {mask.dtype} {mask.shape}')
    eps, sigma, poles = material_values(material, dispersion, properties)
    if poles and sim.engine is not None:
        raise ValueError('add_material with dispersion: add dispersive materials before the first run()')
    if any(np.any(cells & mask) for _, cells in model.poles):
        raise ValueError('add_material: the mask overlaps dispersive cells (poles are never removed): paint '
                         'dispersive materials after the materials they replace')
    materials = model.materials
    material = materials.material_id(eps, eps, eps, sigma)
    count = len(model.poles)
    try:   # the poles first, so that a refused one (the per-model pole limit) leaves the model as it was
        for pole in poles:
            add_pole(pole, mask)
    except Exception:
        del model.poles[count:]
        raise
    materials.paint_mask(mask, material)


class Simulation:
    """2D E_z / H_z FDTD model with the reference's method names; attributes of the model fall through."""

    def __init__(self, size_x: float = None, size_y: float = None, dx: float = None, npml: int = 15,
                 courant: float = None, boundaries=None, *, polarization: str = 'ez', device=None,
                 rows_per_tile: int = 0, graph_steps: int = 0, lane_columns: int = 0):
        if size_x is None or size_y is None or dx is None:
            raise ValueError('Simulation() needs (size_x, size_y, dx, ...) for the 2D EM-FDTD path')
        self.model = Model2D(size_x, size_y, dx, npml=npml, courant=courant, boundaries=boundaries,
                             polarization=polarization)
        self.domain = self.model.domain
        self.boundary_config = self.model.bc
        self.polarization = polarization
        self.engine = None
        self._engine_options = dict(device=device, rows_per_tile=rows_per_tile, graph_steps=graph_steps,
                                    lane_columns=lane_columns)
        self._materials_version = None
        self.callbacks = []

    # ---- accessors ------------------------------------------------------------------------------------------------
    @property
    def grid(self):
        return self.domain.grid

    @property
    def sources(self):
        return self.model.sources

    @property
    def monitors(self):
        return self.model.monitors

    @property
    def n_step(self):
        return 0 if self.engine is None else self.engine.n_step

    @property
    def time(self):
        return 0.0 if self.engine is None else self.engine.time

    @property
    def t(self):
        return self.time

    @property
    def tick(self):
        return self.n_step

    # ---- physics access (simulation.py:250-270) -------------------------------------------------------------------
    @property
    def em(self):
        """The one attached physics: the 2D EM model (the reference's sim.em)."""
        return self.model

    def __getattr__(self, name):
        # Only reached when normal lookup fails: names of the EM model (the sole physics) fall through to it. Dunder
        # lookups (copy / pickle probes) never do.
        if name.startswith('__'):
            raise AttributeError(name)
        model = self.__dict__.get('model')
        if model is None:
            raise AttributeError(f"'Simulation' has no attribute '{name}' and no physics are attached.")
        try:
            return getattr(model, name)
        except AttributeError:
            raise AttributeError(f"'Simulation' has no attribute '{name}' and neither has its only physics "
                                 f"'em'.") from None

    # ---- model building -------------------------------------------------------------------------------------------
    def add_dielectric_rectangle(self, x0, y0, x1, y1, eps_r, sigma_e=0.0):
        self.model.add_dielectric_rectangle(x0, y0, x1, y1, eps_r, sigma_e)

    def add_dielectric_circle(self, cx, cy, radius, eps_r, sigma_e=0.0):
        self.model.add_dielectric_circle(cx, cy, radius, eps_r, sigma_e)

    def add_pec_region(self, mask):
        """PEC on a bool (nx, ny) cell mask (physics/em/__init__.py:372-376), kept over every dielectric."""
        self.model.add_pec_region(mask)

    def add_drude_rectangle(self, x0, y0, x1, y1, drude):
        """Drude metal on the rectangle (cell centres in [x0, x1) x [y0, y1)), eps_r = its eps_inf there
        (physics/em/__init__.py:329-340); E_z only."""
        self._require_unbuilt('add_drude_rectangle')
        self.model.add_drude_rectangle(x0, y0, x1, y1, drude)

    def add_drude_circle(self, cx, cy, radius, drude):
        self._require_unbuilt('add_drude_circle')
        self.model.add_drude_circle(cx, cy, radius, drude)

    def add_dispersive(self, pole, mask):
        """Drude or Lorentz pole on a bool (nx, ny) cell mask over the cells' own eps_r (physics/em/__init__.py:
        342-355); E_z only."""
        self._require_unbuilt('add_dispersive')
        self.model.add_dispersive(pole, mask)

    def add_pec_rectangle(self, x0, y0, x1, y1):
        """PEC on the rectangle raster of add_dielectric_rectangle (this build's convenience over add_pec_region)."""
        self.model.add_pec_rectangle(x0, y0, x1, y1)

    def add_material(self, mask, material=None, dispersion=None, **properties):
        """Paint eps_r / sigma_e (numbers) and optional Drude / Lorentz dispersion on a bool (nx, ny) full-grid cell
        mask (domain.rect_mask / circle_mask); later calls win cells (simulation.py:104-177). Poles need an unbuilt
        simulation and may not overlap earlier dispersive cells, as the reference refuses (its poles are never
        removed)."""
        _add_material(self, mask, material, dispersion, properties, lambda pole, cells: self.model.add_dispersive(
            pole, cells))

    def add_line_source(self, axis, position, wavelength=None, frequency=None, bandwidth=0.5, amplitude=1.0,
                        waveform='gaussian', span=None):
        self._require_unbuilt('add_line_source')
        return self.model.add_line_source(axis, position, wavelength, frequency, bandwidth, amplitude, waveform, span)

    def add_point_source(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, current=1e-3,
                         waveform='gaussian'):
        """z filament current [A] at the cell containing (x, y), E_z only (physics/em/__init__.py:411-418)."""
        self._require_unbuilt('add_point_source')
        return self.model.add_point_source(x, y, wavelength, frequency, bandwidth, current, waveform)

    def add_magnetic_line_current(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, magnetic_current=1.0,
                                  waveform='gaussian'):
        """z magnetic line current [V] at the cell containing (x, y), H_z only (physics/em/__init__.py:420-429)."""
        self._require_unbuilt('add_magnetic_line_current')
        return self.model.add_magnetic_line_current(x, y, wavelength, frequency, bandwidth, magnetic_current,
                                                    waveform)

    def add_electric_dipole(self, x, y, wavelength=None, frequency=None, bandwidth=0.5, dipole_moment=1e-3,
                            orientation=0.0, waveform='gaussian'):
        """In-plane electric dipole [A] along `orientation` [deg] in the cell containing (x, y), H_z only
        (physics/em/__init__.py:431-440)."""
        self._require_unbuilt('add_electric_dipole')
        return self.model.add_electric_dipole(x, y, wavelength, frequency, bandwidth, dipole_moment, orientation,
                                              waveform)

    def add_gaussian_beam(self, axis, position, waist, center, wavelength=None, frequency=None, bandwidth=0.5,
                          amplitude=1.0, waveform='gaussian', span=None):
        """Line source with transverse factor exp(-((s - center)/waist)^2) (physics/em/__init__.py:450-457)."""
        self._require_unbuilt('add_gaussian_beam')
        return self.model.add_gaussian_beam(axis, position, waist, center, wavelength, frequency, bandwidth, amplitude,
                                            waveform, span)

    def add_field_probe(self, x, y, field_component=None):
        """Raw Yee samples of one component at the cell containing (x, y), one per step; get_data() -> (times,
        values) (physics/em/__init__.py:493-498)."""
        self._require_unbuilt('add_field_probe')
        return self.model.add_field_probe(x, y, field_component)

    def add_tfsf(self, x0, y0, x1, y1, wavelength=None, frequency=None, bandwidth=0.5, amplitude=1.0,
                 waveform='gaussian'):
        """+x plane wave of incident peak `amplitude` [V/m] inside the total-field box (x0, y0)-(x1, y1)
        (physics/em/__init__.py:475-490)."""
        self._require_unbuilt('add_tfsf')
        return self.model.add_tfsf(x0, y0, x1, y1, wavelength, frequency, bandwidth, amplitude, waveform)

    def add_dft_monitor(self, x0, y0, x1, y1, wavelengths=None, frequencies=None, downsample=1):
        """Running DFT over the cell box of (x0, y0)-(x1, y1); downsample s > 1 samples every s-th cell along x and y
        (physics/em/__init__.py:500-509)."""
        self._require_unbuilt('add_dft_monitor')
        return self.model.add_dft_monitor(x0, y0, x1, y1, wavelengths, frequencies, downsample)

    def add_flux_monitor(self, axis, position, span, wavelengths=None, frequencies=None):
        self._require_unbuilt('add_flux_monitor')
        return self.model.add_flux_monitor(axis, position, span, wavelengths, frequencies)

    def _require_unbuilt(self, what):
        if self.engine is not None:
            raise ValueError(f'{what}: add sources and monitors before the first run()')

    # ---- running --------------------------------------------------------------------------------------------------
    def finalize(self):
        """Courant check, calibration, DFT interval and (first call) the device engine."""
        model = self.model
        if model is None:
            raise RuntimeError('simulation cleared')
        if self.engine is None:
            model.finalize()
            self.engine = Engine2D(model, **self._engine_options)
        elif model.materials.version != self._materials_version:
            model.apply_pec()
            model.check_courant()
            model.calibrate()
            self.engine.materials_changed()
        self._materials_version = model.materials.version
        return {'em': []}

    def run(self, duration: float = None, n_steps: int = None, print_every: int = 500, flush: bool = True):
        """Advance by `duration` [s] (stopping once t >= t_end - dt/2) or exactly `n_steps` steps
        (simulation.py:272-327, coupling/runner.py:174-204)."""
        if (duration is None) == (n_steps is None):
            raise ValueError('Simulation.run needs exactly one of duration [s] or n_steps')
        if n_steps is not None and int(n_steps) < 0:
            raise ValueError(f'n_steps must be >= 0, got {n_steps}')
        if duration is not None and not float(duration) >= 0:
            raise ValueError(f'duration must be >= 0, got {duration}')
        self.finalize()
        engine = self.engine
        count = int(n_steps) if n_steps is not None else steps_for_duration(engine.time, self.model.dt, duration)
        if print_every > 0:
            span = f'n_steps={count}' if n_steps is not None else f'duration={float(duration):.4e} s'
            print(f'\nOpen_sol (MI355X): physics=[em], dt_sync={float(self.model.dt):.4e} s, {span}')
        wall_start = _time.perf_counter()
        if self.callbacks:
            for _ in range(count):
                engine.step(1)
                for callback in self.callbacks:
                    callback(self)
        else:
            engine.step(count)
        if flush:
            engine.synchronize()
        if print_every > 0:
            engine.synchronize()
            print(f'Done: {count} ticks in {_time.perf_counter() - wall_start:.2f} s (t = {engine.time:.4e} s)')
        return count

    # ---- readout --------------------------------------------------------------------------------------------------
    def get_field_snapshot(self, component=None):
        """float32 field over the physical window (viz/visualization.py:617-620, renderer.py:78-99)."""
        if self.engine is None:
            self.finalize()
        component = component or self.model.components[0]
        if component in ('eps_r', 'eps_xx', 'eps_yy', 'sigma_e', 'mu_r', 'sigma_m'):
            column = {'eps_r': 0, 'eps_xx': 0, 'eps_yy': 1, 'sigma_e': 3, 'mu_r': 4, 'sigma_m': 5}[component]
            return self.model.materials.display_array(column, self.domain.physical_window())
        if component not in self.model.components:
            raise KeyError(f"no field '{component}'; available: {self.model.components}")
        return np.asarray(self.engine.field(component)[self.domain.physical_window()], dtype=DTYPE)

    def get_eps_r_snapshot(self):
        return self.get_field_snapshot('eps_r' if self.polarization == 'ez' else 'eps_xx')

    @property
    def physics(self):
        """Read-only name -> physics mapping (simulation.py:72-75): the one EM model."""
        return MappingProxyType({} if self.model is None else {'em': self.model})

    def memory_report(self, file=None) -> dict:
        return memory_report(self, file)

    def mesh_info(self):
        mesh_info(self)

    def clear(self):
        """Free the device state and drop the model; a cleared Simulation is dead (simulation.py:191-203)."""
        _clear(self)
        self.callbacks = []

    def state(self):
        """Host copies of the current fields."""
        if self.engine is None:
            self.finalize()
        return {name: self.engine.field(name) for name in self.model.components}

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None


class Simulation3D:
    """3D FDTD model with the reference's EMPhysics3D method names (physics/em3d/physics3d.py:144-445) and the
    Simulation run loop; size_* is the physical window, every cpml face adds npml cells.

        sim = Simulation3D(12.2e-6, 12.2e-6, 12.2e-6, dx=25e-9, npml=12)
        sim.add_dielectric_box((1e-6, 1e-6, 1e-6), (3e-6, 2e-6, 4e-6), eps_r=4.0)
        sim.add_dipole(6.1e-6, 6.1e-6, 6.1e-6, orientation='z', wavelength=1e-6, bandwidth=0.5)
        plane = sim.add_flux_plane('x', 10.1e-6, wavelengths=[0.8e-6, 1e-6])
        sim.run(n_steps=500)
        spectrum = plane.flux()
    """

    def __init__(self, size_x, size_y, size_z, dx, npml: int = 12, courant: float = None, boundaries=None, *,
                 device=None, planes_per_tile: int = 0, graph_steps: int = 0):
        from .fdtd3d import Model3D
        self.model = Model3D(size_x, size_y, size_z, dx, npml=npml, courant=courant, boundaries=boundaries)
        self.domain = self.model.domain
        self.boundary_config = self.model.bc
        self.engine = None
        self._engine_options = dict(device=device, planes_per_tile=planes_per_tile, graph_steps=graph_steps)
        self._materials_version = None

    @property
    def grid(self):
        return self.domain.grid

    @property
    def sources(self):
        return self.model.sources

    @property
    def monitors(self):
        return self.model.monitors

    @property
    def n_step(self):
        return 0 if self.engine is None else self.engine.n_step

    @property
    def time(self):
        return 0.0 if self.engine is None else self.engine.time

    @property
    def em3d(self):
        """The one attached physics: the 3D EM model."""
        return self.model

    def __getattr__(self, name):
        # names of the 3D EM model (the sole physics) fall through to it, as Simulation's do (simulation.py:250-270)
        if name.startswith('__'):
            raise AttributeError(name)
        model = self.__dict__.get('model')
        if model is None:
            raise AttributeError(f"'Simulation3D' has no attribute '{name}' and no physics are attached.")
        try:
            return getattr(model, name)
        except AttributeError:
            raise AttributeError(f"'Simulation3D' has no attribute '{name}' and neither has its only physics "
                                 f"'em3d'.") from None

    def add_dielectric_box(self, lo, hi, eps_r, sigma_e=0.0):
        self.model.add_dielectric_box(lo, hi, eps_r, sigma_e)

    def add_pec_box(self, lo, hi):
        """PEC cells on the box raster of add_dielectric_box, applied over every dielectric
        (physics/em3d/physics3d.py:351-353)."""
        self.model.add_pec_box(lo, hi)

    def add_material(self, mask, material=None, dispersion=None, **properties):
        """Paint eps_r / sigma_e (numbers) and optional Drude / Lorentz dispersion on a bool (nx, ny, nz) full-grid
        cell mask; later calls win cells (simulation.py:104-177). Poles as in Simulation.add_material."""
        window = tuple(slice(0, n) for n in self.model.shape)
        _add_material(self, mask, material, dispersion, properties,
                      lambda pole, cells: self.model.add_dispersive(pole, window, cells))

    @property
    def physics(self):
        """Read-only name -> physics mapping: the one 3D EM model, 'em3d' as its attribute name."""
        return MappingProxyType({} if self.model is None else {'em3d': self.model})

    def memory_report(self, file=None) -> dict:
        return memory_report(self, file, 'em3d')

    def mesh_info(self):
        mesh_info(self)

    def clear(self):
        _clear(self)

    def add_dipole(self, x, y, z, orientation='z', dipole_moment=1e-12, wavelength=None, frequency=None,
                   bandwidth=0.5, waveform='gaussian'):
        self._require_unbuilt('add_dipole')
        return self.model.add_dipole(x, y, z, orientation, dipole_moment, wavelength, frequency, bandwidth, waveform)

    def add_field_probe(self, x, y, z, component='ez'):
        """Raw Yee samples of one component at the cell containing (x, y, z), one per step
        (physics/em3d/physics3d.py:425-433)."""
        self._require_unbuilt('add_field_probe')
        return self.model.add_field_probe(x, y, z, component)

    def add_dispersive(self, pole, window, mask):
        """Drude or Lorentz pole on a bool mask of the cells of `window` (one slice per axis), over their own
        permittivity (physics/em3d/physics3d.py:346-349)."""
        self._require_unbuilt('add_dispersive')
        self.model.add_dispersive(pole, window, mask)

    def add_flux_box(self, lo, hi, wavelengths=None, frequencies=None):
        """Net outward flux through the box between the cell edges nearest the corners lo, hi (physics3d.py:447-451)."""
        self._require_unbuilt('add_flux_box')
        return self.model.add_flux_box(lo, hi, wavelengths, frequencies)

    def add_dft_plane(self, z, wavelengths=None, frequencies=None):
        """Running DFT of E at the cell centres of the window layer containing height z (physics3d.py:478-486)."""
        self._require_unbuilt('add_dft_plane')
        return self.model.add_dft_plane(z, wavelengths, frequencies)

    def add_flux_plane(self, normal, position, wavelengths=None, frequencies=None):
        self._require_unbuilt('add_flux_plane')
        return self.model.add_flux_plane(normal, position, wavelengths, frequencies)

    def _require_unbuilt(self, what):
        if self.engine is not None:
            raise ValueError(f'{what}: add sources and monitors before the first run()')

    def finalize(self):
        from .fdtd3d import Engine3D
        model = self.model
        if model is None:
            raise RuntimeError('simulation cleared')
        if self.engine is None:
            model.finalize()
            self.engine = Engine3D(model, **self._engine_options)
        elif model.materials.version != self._materials_version:
            model.finalize()
            self.engine.materials_changed()
        self._materials_version = model.materials.version
        return {'em3d': []}

    def run(self, duration: float = None, n_steps: int = None, print_every: int = 0, flush: bool = True):
        """Advance by `duration` [s] (stopping once t >= t_end - dt/2) or exactly `n_steps` steps."""
        if (duration is None) == (n_steps is None):
            raise ValueError('Simulation3D.run needs exactly one of duration [s] or n_steps')
        if n_steps is not None and int(n_steps) < 0:
            raise ValueError(f'n_steps must be >= 0, got {n_steps}')
        self.finalize()
        engine = self.engine
        count = int(n_steps) if n_steps is not None else steps_for_duration(engine.time, self.model.dt, duration)
        wall_start = _time.perf_counter()
        engine.step(count)
        if flush or print_every > 0:
            engine.synchronize()
        if print_every > 0:
            print(f'Done: {count} ticks in {_time.perf_counter() - wall_start:.2f} s (t = {engine.time:.4e} s)')
        return count

    def get_field_snapshot(self, component='ez'):
        """float32 field over the physical window."""
        if self.engine is None:
            self.finalize()
        return np.asarray(self.engine.field(component)[self.domain.physical_window()], dtype=DTYPE)

    def state(self):
        if self.engine is None:
            self.finalize()
        return {name: self.engine.field(name) for name in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')}

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
