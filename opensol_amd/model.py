"""Model-JSON front end of the FDTD path: a model file of the reference's headless API (the em_fdtd_ez, em_fdtd_hz and
em_fdtd_3d physics types) is built onto Simulation / Simulation3D, its time-dependent study (with a parameter sweep)
runs on the HIP path, and the results come back and are saved in the reference's results-directory layout.

Covered, from the reference's model vocabulary (open_sol/model/spec.py, physics/em/definition_fdtd.py,
physics/em3d/definition.py, physics/em/definition_common.py:1167-1287):

  * parameters: expressions over the earlier parameters and the builtin constants and units
    (model/expressions.py:26-32), evaluated in dependency order (:616-680);
  * domain size_x, size_y (size_z), a uniform mesh of cell dx; physics settings bc_<face>, pml_cells, pml_order,
    pml_sigma_factor, pml_kappa_max, pml_alpha_max, courant_factor (definition_common.py:139-190, 660-668);
  * geometry kinds rectangle, circle, line, point and box (model/geometry.py:111-262); materials with eps_r and
    sigma_e painted on rectangles, circles (2D) or boxes (3D), in the order the materials are listed;
  * 2D features line_source, gaussian_beam, point_source (E_z filament current; H_z magnetic line current or
    in-plane dipole), flux_monitor, field_probe, dft_monitor and pec_domain; 3D features dipole_source, field_probe, flux_plane,
    flux_box, dft_plane and pec_domain;
  * the time_dependent study (duration; the runner stops at t >= t_end - dt/2) and its sweep over one parameter;
  * results: flux_spectrum (with power_per_cw where every source shares one waveform, physics/em/results.py:
    93-122), timeseries and dft_fields per feature, the last run's fields on the window, results.npz in the
    'open_sol.results' version-1 layout (io/results_io.py:399-433, 517-532) and summary.json with the per-result
    digests (:156-225).

Not covered (a ValueError names them): other physics types, features, geometry kinds and study types, couplings,
graded meshes, field snapshots during the run and derived values.
"""
import ast
import json
import math
import operator
import os
import time

import numpy as np

from .constants import C0, EPS0, MU0

# model/expressions.py:26-32
CONSTANTS = {
    'pi': math.pi, 'c_const': float(C0), 'eps0_const': float(EPS0), 'mu0_const': float(MU0),
    'h_const': 6.62607015e-34, 'hbar_const': 6.62607015e-34 / (2.0 * math.pi), 'e_const': 1.602176634e-19,
    'kB_const': 1.380649e-23, 'NA_const': 6.02214076e23,
    'nm': 1e-9, 'um': 1e-6, 'mm': 1e-3, 'cm': 1e-2,
    'fs': 1e-15, 'ps': 1e-12, 'ns': 1e-9, 'us': 1e-6, 'ms': 1e-3,
    'THz': 1e12, 'GHz': 1e9, 'MHz': 1e6, 'kHz': 1e3,
}
_FUNCTIONS = {
    'sin': math.sin, 'cos': math.cos, 'tan': math.tan, 'asin': math.asin, 'acos': math.acos, 'atan': math.atan,
    'atan2': math.atan2, 'sinh': math.sinh, 'cosh': math.cosh, 'tanh': math.tanh, 'exp': math.exp, 'log': math.log,
    'log10': math.log10, 'sqrt': math.sqrt, 'abs': abs, 'min': min, 'max': max, 'floor': math.floor,
    'ceil': math.ceil, 'hypot': math.hypot, 'pow': math.pow, 'sign': lambda v: float(np.sign(v)),
}
_BINARY = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.Pow: operator.pow, ast.Mod: operator.mod, ast.FloorDiv: operator.floordiv}
_UNARY = {ast.UAdd: operator.pos, ast.USub: operator.neg}
_ALIASES = {'μm': 'um', 'μs': 'us'}
RESULTS_FORMAT, RESULTS_FORMAT_VERSION = 'open_sol.results', 1
FACES_2D = ('x_lo', 'x_hi', 'y_lo', 'y_hi')
FACES_3D = FACES_2D + ('z_lo', 'z_hi')
_VALID_SPECTRUM_FRACTION = 1e-6   # physics/em/results.py:23


# ---- expressions ----------------------------------------------------------------------------------------------------
def _names(text):
    tree = ast.parse(str(text).strip(), mode='eval')
    return {node.id for node in ast.walk(tree) if isinstance(node, ast.Name)} - set(_FUNCTIONS)


def evaluate(value, names):
    """A number, or the float value of an expression over `names` (parameters) and the builtin constants; ValueError
    for anything but arithmetic, the listed functions and known names."""
    if isinstance(value, bool):
        raise ValueError(f'expected a number or expression, got {value!r}')
    if isinstance(value, (int, float)):
        return float(value)
    if not isinstance(value, str):
        raise ValueError(f'expected a number or expression, got {type(value).__name__}')
    text = value.strip()
    for alias, name in _ALIASES.items():
        text = text.replace(alias, name)
    tree = ast.parse(text, mode='eval')

    def walk(node):
        if isinstance(node, ast.Expression):
            return walk(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) and not isinstance(node.value,
                                                                                                       bool):
            return float(node.value)
        if isinstance(node, ast.Name):
            if node.id in names:
                return float(names[node.id])
            if node.id in CONSTANTS:
                return CONSTANTS[node.id]
            raise ValueError(f'unknown name {node.id!r} in {value!r}')
        if isinstance(node, ast.BinOp) and type(node.op) in _BINARY:
            return float(_BINARY[type(node.op)](walk(node.left), walk(node.right)))
        if isinstance(node, ast.UnaryOp) and type(node.op) in _UNARY:
            return float(_UNARY[type(node.op)](walk(node.operand)))
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCTIONS \
                and not node.keywords:
            return float(_FUNCTIONS[node.func.id](*(walk(arg) for arg in node.args)))
        raise ValueError(f'{value!r}: only numbers, parameter names, + - * / ** % // and the functions '
                         f'{sorted(_FUNCTIONS)} are allowed')

    result = walk(tree)
    if not math.isfinite(result):
        raise ValueError(f'{value!r} evaluates to {result}')
    return result


def evaluate_parameters(parameters, overrides=None):
    """{name: float} of the model's parameters in dependency order (model/expressions.py:616-680); overrides
    {name: expression} replace expressions first (a sweep value)."""
    raw = {}
    for item in parameters:
        name = item['name']
        if name in raw:
            raise ValueError(f'parameter {name!r} is defined more than once')
        if name in CONSTANTS or name in _FUNCTIONS:
            raise ValueError(f'parameter name {name!r} is reserved')
        raw[name] = item['expression']
    for name, expression in (overrides or {}).items():
        if name not in raw:
            raise KeyError(f'unknown parameter {name!r}')
        raw[name] = expression
    values, pending = {}, dict(raw)
    while pending:
        ready = [name for name, expr in pending.items()
                 if isinstance(expr, (int, float)) or not (_names(expr) - set(values) - set(CONSTANTS))]
        if not ready:
            raise ValueError(f'parameters {sorted(pending)} depend on unknown names or on each other')
        for name in ready:
            values[name] = evaluate(pending.pop(name), values)
    return {name: values[name] for name in raw}


# ---- spec access ----------------------------------------------------------------------------------------------------
def load(path):
    """The model dict of a model JSON file."""
    with open(path, encoding='utf-8-sig') as handle:
        return json.load(handle)


def _physics(spec):
    physics = spec.get('physics') or []
    if len(physics) != 1:
        raise ValueError(f'this front end runs one physics per model, got {len(physics)}')
    entry = physics[0]
    if entry.get('type') not in ('em_fdtd_ez', 'em_fdtd_hz', 'em_fdtd_3d'):
        raise ValueError(f"physics type {entry.get('type')!r} is not an FDTD type of this path "
                         f"(em_fdtd_ez, em_fdtd_hz, em_fdtd_3d)")
    if spec.get('couplings'):
        raise ValueError('couplings are not part of the FDTD path')
    return entry


def _wavelengths(params, values):
    """core/wavelengths.py:5-10: count wavelengths linearly from start to stop; one gives their centre."""
    start = values('wavelength_start', 400e-9)
    stop = values('wavelength_stop', 800e-9)
    count = int(values('n_wavelengths', 100))
    if count < 1:
        raise ValueError('n_wavelengths must be >= 1')
    if count == 1:
        return np.array([0.5 * (start + stop)])
    if not stop > start:
        raise ValueError(f'wavelength_stop ({stop}) must be > wavelength_start ({start})')
    return np.linspace(start, stop, count)


class _Built:
    """One build of the model at one set of parameter values: the simulation and its feature objects."""

    def __init__(self, spec, names):
        from .boundaries import BoundaryConfig
        from .simulation import Simulation, Simulation3D
        physics = _physics(spec)
        self.physics_name = physics.get('name', 'em')
        self.kind = physics['type']
        self.three_d = self.kind == 'em_fdtd_3d'
        value = lambda v: evaluate(v, names)   # noqa: E731
        settings = physics.get('settings') or {}
        track = settings.get('track_absorption', False)
        if track if isinstance(track, bool) else value(track):
            raise ValueError('absorption tracking (absorbed_power) is not part of the FDTD path')

        def setting(key, default):
            raw = settings.get(key, default)
            return raw if isinstance(raw, str) and key.startswith('bc_') else value(raw)

        mesh = spec.get('mesh') or {}
        if mesh.get('mode', 'uniform') != 'uniform' or mesh.get('refinements'):
            raise ValueError('only a uniform mesh is part of the FDTD path')
        dx = value(mesh.get('dx', '25*nm'))
        domain = spec['domain']
        faces = FACES_3D if self.three_d else FACES_2D
        bc = BoundaryConfig(**{face: setting(f'bc_{face}', 'cpml') for face in faces})
        bc.cpml_npml = int(setting('pml_cells', 15))
        bc.cpml_order = int(setting('pml_order', 3))
        bc.cpml_sigma_factor = setting('pml_sigma_factor', 1.0)
        bc.cpml_kappa_max = setting('pml_kappa_max', 15.0)
        bc.cpml_alpha_max = setting('pml_alpha_max', 0.0)
        courant = setting('courant_factor', 0.99)
        sizes = [value(domain['size_x']), value(domain['size_y'])]
        if self.three_d:
            sizes.append(value(domain['size_z']))
            self.sim = Simulation3D(*sizes, dx, npml=bc.cpml_npml, courant=courant, boundaries=bc)
        else:
            self.sim = Simulation(*sizes, dx, npml=bc.cpml_npml, courant=courant, boundaries=bc,
                                  polarization=self.kind[-2:])
        self.geometry = {}
        for item in spec.get('geometry') or []:
            self.geometry[item['name']] = (item['kind'], {k: value(v) for k, v in (item.get('params') or {}).items()})
        self._owner = np.full(self.sim.model.shape, -1, dtype=np.int32)   # the material each cell was painted with
        dispersive = []
        for index, material in enumerate(spec.get('materials') or []):
            poles = self._paint(material, value, index)
            if poles:
                dispersive.append((index, poles))
        for index, poles in dispersive:   # definition_common.py:1009-1041: each pole on its material's cells
            for pole in poles:
                if self.three_d:
                    window = tuple(slice(0, n) for n in self.sim.model.shape)
                    self.sim.add_dispersive(pole, window, self._owner == index)
                else:
                    self.sim.add_dispersive(pole, self._owner == index)
        self.features = []
        for feature in physics.get('features') or []:
            if feature.get('enabled', True):
                self.features.append((feature['name'], feature['type'], self._apply(feature, value)))

    def _shape(self, name, kinds):
        if name not in self.geometry:
            raise ValueError(f'selection names unknown geometry {name!r}')
        kind, params = self.geometry[name]
        if kind not in kinds:
            raise ValueError(f'geometry {name!r} is a {kind}; expected {" or ".join(kinds)}')
        return kind, params

    def _single(self, feature, kinds):
        selection = feature.get('selection') or []
        if len(selection) != 1:
            raise ValueError(f"feature {feature['name']!r} needs exactly one selected object")
        return self._shape(selection[0], kinds)

    def _paint(self, material, value, index):
        """Paint one material in list order; returns its dispersive poles (DrudePole / LorentzPole from the
        'dispersion' list, definition_common.py:1079-1099), whose background permittivity is the eps_r property,
        else the first Drude eps_inf, else 1 (:1102-1111)."""
        from .dispersion import background_permittivity, pole_from_dict
        properties = material.get('properties') or {}
        unknown = set(properties) - {'eps_r', 'sigma_e'}
        if unknown:
            raise ValueError(f"material {material.get('name')!r}: only eps_r, sigma_e and dispersion are part of "
                             f"the FDTD path")
        poles = [pole_from_dict(pole, value) for pole in material.get('dispersion') or []]
        eps = value(properties['eps_r']) if 'eps_r' in properties else background_permittivity(poles)
        sigma = value(properties.get('sigma_e', 0.0))
        for name in material.get('selection') or []:
            kind, p = self._shape(name, ('box',) if self.three_d else ('rectangle', 'circle'))
            if kind == 'rectangle':
                self.sim.add_dielectric_rectangle(p['x0'], p['y0'], p['x1'], p['y1'], eps_r=eps, sigma_e=sigma)
                ranges = self.sim.domain.box_ranges((p['x0'], p['y0']), (p['x1'], p['y1']))
                self._owner[tuple(slice(lo, hi) for lo, hi in ranges)] = index
            elif kind == 'circle':
                self.sim.model.add_dielectric_circle(p['cx'], p['cy'], p['r'], eps_r=eps, sigma_e=sigma)
                x, y = (c.astype(np.float32) for c in self.sim.domain.cell_centers())
                self._owner[(x[:, None] - np.float32(p['cx'])) ** 2 + (y[None, :] - np.float32(p['cy'])) ** 2
                            <= np.float32(p['r']) ** 2] = index
            else:
                lo, hi = (p['x0'], p['y0'], p['z0']), (p['x1'], p['y1'], p['z1'])
                self.sim.add_dielectric_box(lo, hi, eps, sigma)
                ranges = self.sim.domain.box_ranges(lo, hi, clamp_to_physical=True)
                self._owner[tuple(slice(a, b) for a, b in ranges)] = index
        return poles

    def _apply(self, feature, value):
        """definition_common.py:1167-1287 (2D) and physics/em3d/definition.py:472-520 (3D)."""
        params = feature.get('params') or {}
        get = lambda key, default=None: value(params[key]) if key in params else default   # noqa: E731
        text = lambda key, default: params.get(key, default)   # noqa: E731
        kind, sim = feature['type'], self.sim
        source_time = dict(wavelength=get('wavelength', 550e-9), bandwidth=get('bandwidth', 0.5),
                           waveform=text('waveform', 'gaussian'))
        if kind == 'pec_domain':
            for name in feature.get('selection') or []:
                shape, p = self._shape(name, ('box',) if self.three_d else ('rectangle',))
                if self.three_d:
                    sim.add_pec_box((p['x0'], p['y0'], p['z0']), (p['x1'], p['y1'], p['z1']))
                else:
                    sim.add_pec_rectangle(p['x0'], p['y0'], p['x1'], p['y1'])
            return None
        if kind == 'field_probe':
            _, p = self._single(feature, ('point',))
            if self.three_d:
                return sim.add_field_probe(p['x'], p['y'], p.get('z', 0.0), text('component', 'ez'))
            return sim.add_field_probe(p['x'], p['y'], text('component', self.kind[-2:]))
        if not self.three_d:
            if kind in ('line_source', 'gaussian_beam', 'flux_monitor'):
                _, p = self._single(feature, ('line',))
                if p['x0'] == p['x1']:
                    axis, position, span = 'y', 0.5 * (p['x0'] + p['x1']), (p['y0'], p['y1'])
                elif p['y0'] == p['y1']:
                    axis, position, span = 'x', 0.5 * (p['y0'] + p['y1']), (p['x0'], p['x1'])
                else:
                    raise ValueError(f"feature {feature['name']!r}: its line must be horizontal or vertical")
                if kind == 'line_source':
                    return sim.add_line_source(axis, position, amplitude=get('amplitude', 1.0), span=span,
                                               **source_time)
                if kind == 'gaussian_beam':   # definition_common.py:1227-1232
                    center = get('center')
                    center = 0.5 * (span[0] + span[1]) if center is None else center
                    return sim.add_gaussian_beam(axis, position, get('waist', 1e-6), center,
                                                 amplitude=get('amplitude', 1.0), span=span, **source_time)
                return sim.add_flux_monitor(axis, position, tuple(sorted(span)), wavelengths=_wavelengths(params, get))
            if kind == 'point_source':   # definition_common.py:1199-1217
                _, p = self._single(feature, ('point',))
                if self.kind.endswith('ez'):
                    return sim.add_point_source(p['x'], p['y'], current=get('current', 1e-3), **source_time)
                if text('excitation', 'magnetic_current') == 'magnetic_current':
                    return sim.add_magnetic_line_current(p['x'], p['y'], magnetic_current=get('magnetic_current', 1.0),
                                                         **source_time)
                return sim.add_electric_dipole(p['x'], p['y'], dipole_moment=get('dipole_moment', 1e-3),
                                               orientation=get('orientation', 0.0), **source_time)
            if kind == 'dft_monitor':
                _, p = self._single(feature, ('rectangle',))
                return sim.add_dft_monitor(p['x0'], p['y0'], p['x1'], p['y1'], wavelengths=_wavelengths(params, get),
                                           downsample=int(get('downsample', 1)))
        else:
            if kind == 'dipole_source':
                _, p = self._single(feature, ('point',))
                return sim.add_dipole(p['x'], p['y'], p.get('z', 0.0), orientation=text('dipole_axis', 'z'),
                                      dipole_moment=get('current_moment', 1e-12), **source_time)
            if kind == 'flux_plane':
                return sim.add_flux_plane(text('normal', 'z'), get('position', 0.0),
                                          wavelengths=_wavelengths(params, get))
            if kind == 'dft_plane':
                return sim.add_dft_plane(get('position', 0.0), wavelengths=_wavelengths(params, get))
            if kind == 'flux_box':
                _, p = self._single(feature, ('box',))
                return sim.add_flux_box((p['x0'], p['y0'], p['z0']), (p['x1'], p['y1'], p['z1']),
                                        wavelengths=_wavelengths(params, get))
        raise ValueError(f"feature type {kind!r} is not part of the {self.kind} FDTD path here")

    # ---- results ----
    def results(self):
        """{'<physics>.<feature>': result dict} of every result-bearing feature."""
        out = {}
        for name, kind, obj in self.features:
            key = f'{self.physics_name}.{name}'
            if kind in ('flux_monitor', 'flux_plane', 'flux_box'):
                out[key] = self._flux_result(obj)
            elif kind == 'field_probe':
                times, values = obj.get_data()
                out[key] = {'kind': 'timeseries', 't': times, 'values': values, 'component': obj.field_component,
                            'unit': 'A/m' if obj.field_component.startswith('h') else 'V/m'}
            elif kind in ('dft_monitor', 'dft_plane'):
                out[key] = self._dft_result(obj)
        return out

    def _source_spectrum(self, frequency, sample_interval):
        """(S(omega), valid) of the one waveform every source shares, else None (physics/em/results.py:55-100;
        core/sources.py:103-116: S = sum over the sampled steps n of s(n dt) exp(-j omega n dt) m dt)."""
        sources = list(self.sim.model.sources)
        if not sources:
            return None
        first = sources[0].temporal
        if any(type(s.temporal) is not type(first) or vars(s.temporal) != vars(first) for s in sources):
            return None
        dt = float(self.sim.model.dt)
        n_step, m = self.sim.engine.n_step, int(sample_interval)

        def spectrum(freqs):
            times = np.arange(m, n_step + 1, m, dtype=np.float64) * dt
            weighted = np.asarray(first(times), dtype=np.float64) * (dt * m)
            return np.exp(-2j * np.pi * np.outer(np.atleast_1d(freqs), times)) @ weighted

        values = spectrum(frequency)
        carrier = float(np.abs(spectrum([first.f0]))[0] ** 2)
        valid = (np.abs(values) ** 2 >= _VALID_SPECTRUM_FRACTION * carrier) & (carrier > 0.0)
        if type(first).__name__ == 'ContinuousWave':
            valid &= np.abs(frequency - first.f0) * n_step * dt < 1.0
        return values, valid

    def _flux_result(self, monitor):
        frequency = np.array(monitor.frequencies, dtype=np.float64)
        result = {'kind': 'flux_spectrum', 'frequency': frequency, 'wavelength': C0 / frequency,
                  'power': np.asarray(monitor.flux(), dtype=np.float64), 'unit': 'a.u.', 'mirror_factor': 1.0}
        shared = self._source_spectrum(frequency, monitor.sample_interval)
        if shared is not None:
            values, valid = shared
            scale = np.where(valid, 1.0 / (2.0 * np.where(valid, np.abs(values) ** 2, 1.0)), np.nan)
            result['power_per_cw'] = result['power'] * scale
        return result

    def _dft_result(self, monitor):
        """'dft_fields' of a DFT box (physics/em/results.py:125-150) or 3D DFT plane (physics/em3d/monitors3d.py:
        273-295), with 'intensity_per_cw_factor' = 1 / |S(omega)|^2 when every source shares one waveform."""
        frequency = np.array(monitor.frequencies, dtype=np.float64)
        centres = self.sim.domain.cell_centers()
        edges = self.sim.domain.cell_edges()
        if self.three_d:
            xs, ys = monitor.window
            intensity = np.stack([monitor.intensity(i) for i in range(monitor.n_freqs)])
            extent = (float(edges[0][xs.start]), float(edges[0][xs.stop]), float(edges[1][ys.start]),
                      float(edges[1][ys.stop]))
            result = {'kind': 'dft_fields', 'frequency': frequency, 'wavelength': C0 / frequency,
                      'x_centers': np.array(centres[0][xs]), 'y_centers': np.array(centres[1][ys]), 'extent': extent,
                      'shape': tuple(monitor.shape), 'quantity': 'abs(E_x)^2 + abs(E_y)^2 + abs(E_z)^2',
                      'z': float(centres[2][monitor.k_grid]), 'intensity_all': intensity}
        else:
            if monitor.polarization == 'ez':
                intensity = np.abs(monitor.ez_dft) ** 2
            else:
                intensity = np.abs(monitor.ex_dft) ** 2 + np.abs(monitor.ey_dft) ** 2
            xs, ys = monitor.x_slice, monitor.y_slice
            extent = (float(edges[0][xs.start]), float(edges[0][xs.stop]), float(edges[1][ys.start]),
                      float(edges[1][ys.stop]))
            result = {'kind': 'dft_fields', 'frequency': frequency, 'wavelength': C0 / frequency,
                      'x_centers': np.array(centres[0][monitor.x_samples]),
                      'y_centers': np.array(centres[1][monitor.y_samples]), 'extent': extent,
                      'shape': (int(monitor.mx), int(monitor.my)), 'downsample': int(monitor.downsample),
                      'quantity': 'abs(E_z)^2' if monitor.polarization == 'ez' else 'abs(E_x)^2 + abs(E_y)^2',
                      'intensity_all': intensity}
        shared = self._source_spectrum(frequency, monitor.sample_interval)
        if shared is not None:
            values, valid = shared
            result['intensity_per_cw_factor'] = np.where(valid, 1.0 / np.where(valid, np.abs(values) ** 2, 1.0),
                                                         np.nan)
        return result


# ---- stored fields and snapshots (model/study.py:916-967, 2041-2130, 2623-2645; physics/em/renderer.py:12-312) -----
_SNAPSHOT_PIXELS = 250_000        # window cells per stored field (model/study.py:44)
_SNAPSHOT_BYTE_LIMIT = 1 << 30    # intermediate snapshots stop past this many bytes (:45)


def _field_info(name, label, unit, cmap, signed=False, primary=False, parity='', staggered=''):
    """A FieldInfo record (core/renderer.py) as results.npz stores it."""
    return {'name': name, 'label': label, 'unit': unit, 'signed': signed, 'cmap': cmap, 'primary': primary,
            'requires_study': '', 'parity': parity, 'staggered': staggered, 'plane_signs': ''}


_MODULI = [_field_info('e_abs', '|E|', 'V/m', 'inferno'), _field_info('e_intensity', '|E|^2', 'V^2/m^2', 'inferno'),
           _field_info('h_abs', '|H|', 'A/m', 'viridis')]
_MEDIA = [_field_info('absorbed_power', 'Absorbed power density', 'W/m^3', 'magma'),
          _field_info('carrier_density', 'Free-carrier density', '1/m^3', 'viridis'),
          _field_info('population_inversion', 'Population inversion N2 - N1', '1/m^3', 'RdBu_r', signed=True,
                      parity='even'),
          _field_info('crystallinity', 'Crystallinity', '1', 'viridis')]
_SIGMA = _field_info('sigma_e', 'Electric conductivity', 'S/m', 'cividis')
# The fields a time-dependent study stores of a 2D FDTD physics (every output field but the frequency-domain ones),
# primary first, in the renderer's order (physics/em/renderer.py:159-215).
STORED_FIELDS = {
    'ez': [_field_info('ez', 'E_z', 'V/m', 'RdBu_r', signed=True, primary=True, parity='polar_z'),
           _field_info('hx', 'H_x (+y face)', 'A/m', 'RdBu_r', signed=True, parity='axial_x', staggered='y'),
           _field_info('hy', 'H_y (+x face)', 'A/m', 'RdBu_r', signed=True, parity='axial_y', staggered='x'),
           *_MODULI, _field_info('eps_r', 'Relative permittivity', '1', 'YlOrBr'), _SIGMA, *_MEDIA],
    'hz': [_field_info('hz', 'H_z', 'A/m', 'RdBu_r', signed=True, primary=True, parity='axial_z'),
           _field_info('ex', 'E_x', 'V/m', 'RdBu_r', signed=True, parity='polar_x'),
           _field_info('ey', 'E_y', 'V/m', 'RdBu_r', signed=True, parity='polar_y'),
           *_MODULI, _field_info('eps_xx', 'Relative permittivity xx', '1', 'YlOrBr'),
           _field_info('eps_yy', 'Relative permittivity yy', '1', 'YlOrBr'), _SIGMA, *_MEDIA],
}


# em_fdtd_3d (physics/em3d/renderer3d.py:14-16, 79-92): the raw Yee components, the moduli at the cell centres, the
# E_z permittivity array and the absorbed power, stored primary first.
_RAW_3D = (('ez', 'E_z (+z face)', ''), ('ex', 'E_x (+x face)', 'x'), ('ey', 'E_y (+y face)', 'y'),
           ('hx', 'H_x (+y +z edge)', 'y'), ('hy', 'H_y (+x +z edge)', 'x'), ('hz', 'H_z (+x +y edge)', 'xy'))
STORED_FIELDS['3d'] = [
    *(_field_info(name, label, 'V/m' if name[0] == 'e' else 'A/m', 'RdBu_r', signed=True, primary=name == 'ez',
                  parity=f"{'polar' if name[0] == 'e' else 'axial'}_{name[1]}", staggered=staggered)
      for name, label, staggered in _RAW_3D),
    *_MODULI, _field_info('eps_r', 'Relative permittivity (eps_zz)', '1', 'YlOrBr'), _MEDIA[0]]
# grid axes a raw 3D component is averaged along to reach the cell centre (E: its own axis; H: its edge's two axes)
CENTRE_AXES_3D = {'ex': (0,), 'ey': (1,), 'ez': (2,), 'hx': (1, 2), 'hy': (0, 2), 'hz': (0, 1)}


def _next_output_time(t_now, t_start, interval, dt):
    """The first output time after t_now on the interval's grid, at least one tick ahead (model/study.py:2688-2691)."""
    k = math.floor((t_now + 0.5 * dt - t_start) / interval) + 1
    return max(t_start + k * interval, t_now + dt)


def snapshot_schedule(dt, duration, output_interval=0.0, n_snapshots=20):
    """[(tick, t)] of the snapshots a time-dependent study records, the last one the final state: the study's tick
    loop (t += dt in float64, stopping at t >= t_end - dt/2) with a record whenever t >= next output - dt/2
    (model/study.py:916-967). output_interval 0 spaces n_snapshots evenly over the duration."""
    dt = float(dt)
    half = 0.5 * dt
    interval = float(output_interval)
    if interval <= 0.0 and int(n_snapshots) > 1:
        interval = float(duration) / (int(n_snapshots) - 1)
    t, tick, t_final = 0.0, 0, float(duration)
    next_output = 0.0 if interval > 0.0 else math.inf
    out = []
    while t < t_final - half:
        if t >= next_output - half:
            out.append((tick, t))
            next_output = _next_output_time(t, 0.0, interval, dt)
        stop = min(t_final, next_output) - half
        while t < stop:
            t += dt
            tick += 1
    out.append((tick, t))
    return out


def _to_centres(face, axis, periodic):
    """Hi-face samples averaged with their lower neighbour onto the cell centres, in float32; the first cell pairs
    with the last on a periodic axis, else with itself (physics/em/fields.py:53-66)."""
    face = np.asarray(face, dtype=np.float32)
    lower = np.roll(face, 1, axis=axis)
    if not periodic:
        first = [slice(None)] * face.ndim
        first[axis] = 0
        lower[tuple(first)] = face[tuple(first)]
    return (face + lower) * np.float32(0.5)


def _nearest_cells(edges, n_pixels):
    """int64 index of the cell containing each of n_pixels uniform pixel centres over [edges[0], edges[-1]]
    (model/field_access_tensor.py:835-841)."""
    span = float(edges[-1] - edges[0])
    centres = edges[0] + (np.arange(n_pixels) + 0.5) * (span / n_pixels)
    return np.clip(np.searchsorted(edges, centres, side='right') - 1, 0, len(edges) - 2).astype(np.int64)


def _display_view(u_edges, v_edges, max_pixels=_SNAPSHOT_PIXELS):
    """None when a window plane of these cell edges is stored natively (at most max_pixels cells), else (u index,
    v index, native shape) of its nearest-cell resampling on a uniform grid of at most max_pixels pixels
    (model/field_access_tensor.py:74-89, 245-267)."""
    n_u, n_v = len(u_edges) - 1, len(v_edges) - 1
    if n_u * n_v <= max_pixels:
        return None
    width, height = float(u_edges[-1] - u_edges[0]), float(v_edges[-1] - v_edges[0])
    finest = float(min(np.diff(u_edges).min(), np.diff(v_edges).min()))
    pixel = max(finest, math.sqrt(width * height / max_pixels))
    pixels_u, pixels_v = max(1, int(width / pixel)), max(1, int(height / pixel))
    if pixels_u * pixels_v > max_pixels:   # float rounding guard
        pixels_u = max(1, max_pixels // pixels_v)
    return _nearest_cells(u_edges, pixels_u), _nearest_cells(v_edges, pixels_v), [n_u, n_v]


def _shown(values, view):
    """A window plane as a snapshot stores it: itself, or its nearest-cell resampling."""
    if view is None:
        return values
    return np.ascontiguousarray(values[view[0][:, None], view[1][None, :]])


class _Recorder:
    """Stored-field snapshots of a 2D run (model/study.py:2041-2130): each stored field over the native window cells,
    float32, with the cell centres, or, for a window above _SNAPSHOT_PIXELS cells, resampled to the nearest cell of a
    uniform grid of at most that many pixels (no centres); intermediate snapshots stop once they hold more than
    _SNAPSHOT_BYTE_LIMIT bytes (the final one is always taken, and its native values make the field group).
    `requested`: the study's stored_fields keys, else every stored field."""

    def __init__(self, built, requested=()):
        sim = built.sim
        self.sim = sim
        self.physics = built.physics_name
        self.polarization = sim.model.polarization
        infos = {f'{self.physics}.{info["name"]}': info for info in STORED_FIELDS[self.polarization]}
        for key in requested:
            if key not in infos:
                raise ValueError(f"stored field '{key}' is not an output field of this model; available: "
                                 f"{sorted(infos)}")
        self.keys = list(requested) or list(infos)
        self.infos = {key: infos[key] for key in self.keys}
        self.window = tuple(sim.domain.physical_window())
        edges, centres = sim.domain.cell_edges(), sim.domain.cell_centers()
        self.edges = [np.asarray(e[w.start:w.stop + 1], dtype=np.float64) for e, w in zip(edges, self.window)]
        self.centres = [np.asarray(c[w], dtype=np.float64) for c, w in zip(centres, self.window)]
        self.extent = [float(self.edges[0][0]), float(self.edges[0][-1]), float(self.edges[1][0]),
                       float(self.edges[1][-1])]
        self.view = _display_view(*self.edges)
        self.snapshots = []
        self.final_fields = None
        self.stored_bytes = 0
        self.full = False

    def _values(self):
        """{component: float32 window array} of the current state (physics/em/renderer.py:159-240)."""
        sim, w = self.sim, self.window
        bc = sim.model.bc
        px, py = bc.periodic('x'), bc.periodic('y')
        materials = sim.model.materials
        out = {}
        if self.polarization == 'ez':
            ez, hx, hy = (sim.engine.field(n) for n in ('ez', 'hx', 'hy'))
            out['ez'], out['hx'], out['hy'] = ez[w], hx[w], hy[w]
            out['e_abs'] = np.abs(ez[w])
            out['e_intensity'] = ez[w] * ez[w]
            hx_c, hy_c = _to_centres(hx, 1, py)[w], _to_centres(hy, 0, px)[w]
            out['h_abs'] = np.sqrt(hx_c * hx_c + hy_c * hy_c)
            out['eps_r'] = materials.display_array(0, w)
        else:
            hz, ex, ey = (sim.engine.field(n) for n in ('hz', 'ex', 'ey'))
            ex_c, ey_c = _to_centres(ex, 1, py)[w], _to_centres(ey, 0, px)[w]
            out['hz'], out['ex'], out['ey'] = hz[w], ex_c, ey_c
            intensity = ex_c * ex_c + ey_c * ey_c
            out['e_abs'] = np.sqrt(intensity)
            out['e_intensity'] = intensity
            out['h_abs'] = np.abs(hz[w])
            out['eps_xx'] = materials.display_array(0, w)
            out['eps_yy'] = materials.display_array(1, w)
        out['sigma_e'] = materials.display_array(3, w)
        shape = tuple(s.stop - s.start for s in w)
        # no absorption tracking, carriers, gain or phase-change media on this path: these read zero, as the
        # reference's do without them (physics/em/renderer.py:62-69, 122-150)
        for name in ('absorbed_power', 'carrier_density', 'population_inversion', 'crystallinity'):
            out[name] = np.zeros(shape, dtype=np.float32)
        return {name: np.ascontiguousarray(values, dtype=np.float32) for name, values in out.items()}

    def record(self, final=False):
        """Store the current state (skipped once full unless final)."""
        if self.full and not final:
            return
        values = self._values()
        native = {key: values[key.split('.', 1)[1]] for key in self.keys}
        fields = {key: _shown(v, self.view) for key, v in native.items()}
        cuts = [{key: _shown(cut[key.split('.', 1)[1]], view) for key in self.keys}
                for cut, view in zip(self._cut_values(), self._cut_views())]
        self.snapshots.append({'time': float(self.sim.time), 'label': None, 'time_label': self.time_label(),
                               'fields': fields, 'cuts': cuts})
        if final:
            self.final_fields = native   # the field group: native window cells (api.py:3560-3575)
        self.stored_bytes += _snapshot_bytes(self.snapshots[-1:])
        if not (final or self.full) and self.stored_bytes > _SNAPSHOT_BYTE_LIMIT:
            self.full = True

    def time_label(self):
        return f't = {float(self.sim.time) * 1e15:.1f} fs'

    plane = None

    def _cut_values(self):
        return []   # 2D: no cut planes

    def _cut_views(self):
        return []

    def cut_records(self):
        return []

    def layout(self):
        """What every snapshot of this run shares: infos, extent, centres, edges, the cut plane (3D), the native shape
        of a resampled plane (None when native) and the stored cuts' own layouts."""
        return {'infos': self.infos, 'extent': self.extent, 'centres': self.centres, 'edges': self.edges,
                'plane': self.plane, 'axes': ['x', 'y'], 'native_shape': None if self.view is None else self.view[2],
                'cuts': []}


class _Recorder3D(_Recorder):
    """Stored-field snapshots of a 3D run: the mid-height z cut of each stored field (model/field_access_tensor.py:
    186-204: the window cell containing size_z / 2), float32 over the window's x and y cells, plus one cut per stored
    cut plane ([axis, position]: the window cell containing the position, model/study.py:2250-2258), over the other
    two axes in grid order. Only the two layers a cut needs are copied from the device."""

    def __init__(self, built, requested=(), cuts=()):
        sim = built.sim
        self.sim = sim
        self.physics = built.physics_name
        domain = sim.model.domain
        infos = {f'{self.physics}.{info["name"]}': info for info in STORED_FIELDS['3d']}
        for key in requested:
            if key not in infos:
                raise ValueError(f"stored field '{key}' is not an output field of this model; available: "
                                 f"{sorted(infos)}")
        self.keys = list(requested) or list(infos)
        self.infos = {key: infos[key] for key in self.keys}
        self.window = tuple(domain.physical_window())
        edges, centres = domain.cell_edges(), domain.cell_centers()
        self.planes = []
        for axis_name, position in [('z', 0.5 * domain.size_z)] + list(cuts):
            axis = 'xyz'.index(axis_name)
            k = domain.pos_to_cell_axis(axis, position)
            others = [b for b in range(3) if b != axis]
            plane_edges = [np.asarray(edges[b][self.window[b].start:self.window[b].stop + 1], dtype=np.float64)
                           for b in others]
            self.planes.append({'axis': axis, 'k': k, 'plane': [axis_name, float(centres[axis][k])],
                                'axes': ['xyz'[b] for b in others], 'edges': plane_edges,
                                'centres': [np.asarray(centres[b][self.window[b]], dtype=np.float64) for b in others],
                                'extent': [float(plane_edges[0][0]), float(plane_edges[0][-1]),
                                           float(plane_edges[1][0]), float(plane_edges[1][-1])],
                                'view': _display_view(*plane_edges)})
        default = self.planes[0]
        self.plane, self.edges, self.centres, self.extent, self.view = (
            default['plane'], default['edges'], default['centres'], default['extent'], default['view'])
        self.final_fields = None
        self.snapshots = []
        self.stored_bytes = 0
        self.full = False

    def _centre(self, layers, name, axis=2):
        """float32 window cut normal to `axis` of a raw component averaged to the cell centres (physics/em3d/
        renderer3d.py:115-147): the window plus the lower row along each averaged axis (cell 0: the wrapped last row on
        a periodic axis, else its own), one pair sum per averaged axis in axis order, then the product with 1 / 2^axes.
        layers[name] holds the grid with two layers along `axis`: (lower, cut)."""
        model = self.sim.model
        averaged = CENTRE_AXES_3D[name]
        index = []
        for b in range(3):
            if b == axis:
                index.append(np.asarray([0, 1] if b in averaged else [1]))
                continue
            cells = self.window[b]
            rows = list(range(cells.start, cells.stop))
            if b in averaged:
                lower = cells.start - 1 if cells.start > 0 else (model.shape[b] - 1 if model.periodic[b] else 0)
                rows = [lower] + rows
            index.append(np.asarray(rows))
        values = layers[name][np.ix_(*index)].astype(np.float32)
        for b in averaged:
            upper, lower = [slice(None)] * 3, [slice(None)] * 3
            upper[b], lower[b] = slice(1, None), slice(None, -1)
            values = values[tuple(upper)] + values[tuple(lower)]
        values = values * np.float32(1.0 / 2 ** len(averaged))
        return np.take(values, 0, axis=axis)

    def _plane_values(self, plane):
        """{component: float32 cut} of the current state on one plane (physics/em3d/renderer3d.py:96-113)."""
        from .fdtd3d import COMPONENTS
        model = self.sim.model
        engine = self.sim.engine
        axis, k = plane['axis'], plane['k']
        k_lower = k - 1 if k > 0 else (model.shape[axis] - 1 if model.periodic[axis] else k)
        state = engine.fields[engine.n_step & 1]
        pick = [slice(None)] * 4
        pick[1 + axis] = [k_lower, k]
        cut = state[tuple(pick)].cpu().numpy()   # (6, ...) with two layers along the cut axis
        layers = {name: cut[c] for c, name in enumerate(COMPONENTS)}
        window = list(self.window)
        window[axis] = 1   # the cut layer
        out = {name: layers[name][tuple(window)] for name in COMPONENTS}
        e2 = sum(self._centre(layers, name, axis) ** 2 for name in ('ex', 'ey', 'ez'))
        h2 = sum(self._centre(layers, name, axis) ** 2 for name in ('hx', 'hy', 'hz'))
        out['e_intensity'] = e2
        out['e_abs'] = np.sqrt(e2)
        out['h_abs'] = np.sqrt(h2)
        cells = list(self.window)   # the E_z permittivity of the cut's cells only (no full-grid array)
        cells[axis] = k
        out['eps_r'] = model.materials.display_array(2, tuple(cells))
        # no absorption tracking on this path: zeros, as the reference's without it (renderer3d.py:104-113)
        out['absorbed_power'] = np.zeros(out['ez'].shape, dtype=np.float32)
        return {name: np.ascontiguousarray(values, dtype=np.float32) for name, values in out.items()}

    def _values(self):
        return self._plane_values(self.planes[0])

    def _cut_values(self):
        return [self._plane_values(plane) for plane in self.planes[1:]]

    def _cut_views(self):
        return [plane['view'] for plane in self.planes[1:]]

    def layout(self):
        out = super().layout()
        out['cuts'] = [{'infos': self.infos, 'extent': p['extent'], 'centres': p['centres'], 'plane': p['plane'],
                        'axes': p['axes'], 'native_shape': None if p['view'] is None else p['view'][2]}
                       for p in self.planes[1:]]
        return out

    def cut_records(self):
        return [list(plane['plane']) for plane in self.planes[1:]]


def _snapshot_bytes(snapshots):
    """Host bytes of the stored values of snapshots, cuts included (model/study.py:662-671)."""
    return sum(v.nbytes for snapshot in snapshots for fields in [snapshot['fields'], *snapshot.get('cuts', [])]
               for v in fields.values())


# ---- digests (io/results_io.py:156-225) ----------------------------------------------------------------------------
def _finite(values):
    values = np.asarray(values, dtype=np.float64).ravel()
    return values[np.isfinite(values)]


def result_digest(result):
    kind = result.get('kind')
    if kind == 'timeseries':
        times, values = np.asarray(result['t'], np.float64), np.asarray(result['values'], np.float64)
        finite = _finite(values)
        return {'kind': kind, 'component': result.get('component'), 'unit': result.get('unit'),
                'n_samples': int(values.size), 't_start': float(times[0]) if times.size else None,
                't_stop': float(times[-1]) if times.size else None,
                'min': float(finite.min()) if finite.size else None, 'max': float(finite.max()) if finite.size else None,
                'final': float(values[-1]) if values.size else None}
    if kind == 'flux_spectrum':
        wavelength, power = np.asarray(result['wavelength'], np.float64), np.asarray(result['power'], np.float64)
        finite = _finite(power)
        peak = int(np.nanargmax(np.where(np.isfinite(power), power, -np.inf))) if finite.size else None
        return {'kind': kind, 'unit': result.get('unit'), 'n_wavelengths': int(wavelength.size),
                'wavelength_min': float(wavelength.min()), 'wavelength_max': float(wavelength.max()),
                'power_min': float(finite.min()) if finite.size else None,
                'power_max': float(finite.max()) if finite.size else None,
                'wavelength_at_max': None if peak is None else float(wavelength[peak])}
    if kind == 'dft_fields':
        wavelength = np.asarray(result['wavelength'], np.float64)
        intensity_max = [float(_finite(frame).max()) if _finite(frame).size else None
                         for frame in result.get('intensity_all', [])]
        return {'kind': kind, 'quantity': result.get('quantity'), 'shape': [int(n) for n in result['shape']],
                'n_wavelengths': int(wavelength.size), 'wavelengths': wavelength.tolist(),
                'intensity_max': intensity_max or None}
    return {'kind': kind, 'keys': sorted(str(key) for key in result)}


def _json_safe(value):
    if value is None or isinstance(value, (bool, int, str)):
        return value
    if isinstance(value, (float, np.floating)):
        value = float(value)
        return value if math.isfinite(value) else ('nan' if math.isnan(value) else ('inf' if value > 0 else '-inf'))
    if isinstance(value, np.integer):
        return int(value)
    if isinstance(value, dict):
        return {str(k): _json_safe(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_json_safe(v) for v in value]
    if isinstance(value, np.ndarray):
        return _json_safe(value.tolist())
    raise TypeError(f'{type(value).__name__} is not JSON serialisable')


def _is_array(value):
    return isinstance(value, np.ndarray) and value.ndim >= 1 and value.dtype.kind in 'biufc'


def _split(result, prefix, arrays):
    """Non-array entries of a result; its arrays go into `arrays` under prefix + key (io/results_io.py:517-532)."""
    entries = {}
    for key, value in result.items():
        if '/' in str(key):
            raise ValueError(f"result key {key!r} contains '/'")
        if _is_array(value):
            arrays[prefix + str(key)] = value
        else:
            entries[key] = value
    return entries


def _field_entries(fields, layout, time_label, prefix, arrays):
    """Metadata entries of one snapshot cut's {key: values}; values and native cell centres go into arrays under
    prefix + key (io/results_io.py:467-481)."""
    entries = {}
    for key, values in fields.items():
        arrays[f'{prefix}{key}/values'] = values
        if layout['native_shape'] is None:   # native cells: their centres (a resampled plane has none)
            arrays[f'{prefix}{key}/x_centers'] = layout['centres'][0]
            arrays[f'{prefix}{key}/y_centers'] = layout['centres'][1]
        entries[key] = {'extent': layout['extent'], 'info': layout['infos'][key], 'time_label': time_label,
                        'native_shape': layout['native_shape'], 'plane': layout['plane'],
                        'axes': list(layout['axes']), 'unfolded': False, 'mesh': 'mesh', 'triangulation': None}
    return entries


def _snapshot_entry(snapshot, layout, prefix, arrays):
    """Metadata entry {time, label, fields, cuts, segments} of one snapshot (io/results_io.py:441-464): the default
    cut's fields under prefix, stored cut c under prefix + 'cut<c>/'."""
    time_label = snapshot['time_label']
    return {'time': snapshot['time'], 'label': snapshot['label'],
            'fields': _field_entries(snapshot['fields'], layout, time_label, prefix, arrays),
            'cuts': [_field_entries(fields, cut_layout, time_label, f'{prefix}cut{c}/', arrays)
                     for c, (fields, cut_layout) in enumerate(zip(snapshot.get('cuts', []), layout['cuts']))],
            'segments': {}}


class Results:
    """Outcome of run(): status, the last run's results, stored fields and snapshots, every swept run's results (and,
    with keep_fields, snapshots), summary() and save(). result(physics, feature) / sweep_result(physics, feature) as
    the reference's Results. fields: {key: (values, x edges, y edges)} of the final state; snapshots: [{time, label,
    time_label, fields {key: values}}]; sweep_snapshots: [(layout, snapshots)] per swept value."""

    def __init__(self, status, spec, results, sweep_results, sweep, fields, summary, snapshots=(), layout=None,
                 sweep_snapshots=()):
        self.status = status
        self.spec = spec
        self.results = results
        self.sweep_results = sweep_results
        self.sweep = sweep
        self.fields = fields
        self.snapshots = list(snapshots)
        self.layout = layout
        self.sweep_snapshots = list(sweep_snapshots)
        self._summary = summary

    def summary(self):
        return self._summary

    def result(self, physics, feature=None):
        return self.results[f'{physics}.{feature}' if feature is not None else physics]

    def sweep_result(self, physics, feature=None):
        key = f'{physics}.{feature}' if feature is not None else physics
        return [run[key] for run in self.sweep_results]

    def save(self, directory):
        """results.npz (format 'open_sol.results' version 1; io/results_io.py:398-438), model.json and summary.json;
        returns the files dict."""
        os.makedirs(directory, exist_ok=True)
        arrays = {}
        metadata = {'format': RESULTS_FORMAT, 'version': RESULTS_FORMAT_VERSION, 'fields': {}, 'results': {},
                    'sweep_results': [], 'sector_results': [], 'sector_combined_results': {}, 'snapshots': [],
                    'sweep_snapshots': []}
        t_end = self._summary['t_end']
        for key, (values, x_edges, y_edges) in self.fields.items():
            arrays[f'field/{key}/values'] = np.asarray(values)
            arrays[f'field/{key}/x_edges'] = np.asarray(x_edges, dtype=np.float64)
            arrays[f'field/{key}/y_edges'] = np.asarray(y_edges, dtype=np.float64)
            physics, component = key.split('.', 1)
            metadata['fields'][key] = {'physics': physics, 'component': component,
                                       'info': self.layout['infos'][key] if self.layout else None, 'time': t_end,
                                       'time_label': f't = {t_end * 1e15:.1f} fs',
                                       'plane': self.layout['plane'] if self.layout else None, 'mesh': 'mesh',
                                       'triangulation': None}
        for key, result in self.results.items():
            metadata['results'][key] = _split(result, f'result/{key}/', arrays)
        for index, run in enumerate(self.sweep_results if self.sweep else []):
            metadata['sweep_results'].append({key: _split(result, f'sweep/{index}/{key}/', arrays)
                                              for key, result in run.items()})
        metadata['snapshots'] = [_snapshot_entry(snapshot, self.layout, f'snapshot/{index}/', arrays)
                                 for index, snapshot in enumerate(self.snapshots)]
        metadata['sweep_snapshots'] = [[_snapshot_entry(snapshot, layout, f'sweep_snapshot/{value}/{index}/', arrays)
                                        for index, snapshot in enumerate(snapshots)]
                                       for value, (layout, snapshots) in enumerate(self.sweep_snapshots)]
        files = {'summary': 'summary.json', 'arrays': 'results.npz', 'model': 'model.json'}
        with open(os.path.join(directory, 'results.npz'), 'wb') as handle:
            np.savez(handle, **arrays, metadata=np.array(json.dumps(_json_safe(metadata))))
        with open(os.path.join(directory, 'model.json'), 'w', encoding='utf-8') as handle:
            json.dump(self.spec, handle, indent=2)
        summary = dict(self._summary, files=files)
        with open(os.path.join(directory, 'summary.json'), 'w', encoding='utf-8') as handle:
            json.dump(_json_safe(summary), handle, indent=2, ensure_ascii=True, allow_nan=False)
            handle.write('\n')
        return files


def _stored_cuts(study, names):
    """[(axis, position [m])] of the study's stored_cuts (model/study.py:2250-2258)."""
    planes = []
    for cut in study.get('stored_cuts') or []:
        if not (isinstance(cut, (list, tuple)) and len(cut) == 2 and cut[0] in ('x', 'y', 'z')):
            raise ValueError(f"stored cut must be ['x' | 'y' | 'z', position in m], got {cut!r}")
        planes.append((cut[0], evaluate(cut[1], names)))
    return planes


def _advance(built, settings, names, requested, cuts=()):
    """One time-dependent study of a build (model/study.py:916-967): the snapshot ticks of snapshot_schedule, with a
    record at each and the final state last (3D: the mid-height z cut). Returns the recorder."""
    sim = built.sim
    duration = evaluate(settings.get('duration', 200e-15), names)
    interval = evaluate(settings.get('output_interval', 0.0), names)
    count = int(evaluate(settings.get('n_snapshots', 20), names))
    if not float(interval) >= 0.0 or count < 1:
        raise ValueError(f'output_interval must be >= 0 and n_snapshots >= 1, got {interval!r}, {count!r}')
    sim.finalize()
    if cuts and not built.three_d:
        raise ValueError('stored cuts need a 3D model (em_fdtd_3d); this model is 2D')
    recorder = _Recorder3D(built, requested, cuts) if built.three_d else _Recorder(built, requested)
    schedule = snapshot_schedule(sim.model.dt, duration, interval, count)
    for index, (tick, _) in enumerate(schedule):
        if tick > sim.n_step:
            sim.run(n_steps=tick - sim.n_step, print_every=0)
        last = index == len(schedule) - 1
        if last or not recorder.full:
            recorder.record(final=last)
    return recorder


class _KeptSweepSnapshots:
    """keep_fields: every swept value's snapshots while they total at most _SNAPSHOT_BYTE_LIMIT bytes, from the value
    that passes it on only each value's final snapshot (model/study.py:674-693)."""

    def __init__(self):
        self.stored_bytes = 0
        self.finals_only = False
        self.kept = []

    def take(self, recorder):
        snapshots = recorder.snapshots
        if not self.finals_only and self.stored_bytes + _snapshot_bytes(snapshots) > _SNAPSHOT_BYTE_LIMIT:
            self.finals_only = True
        if self.finals_only:
            snapshots = snapshots[-1:]
        self.stored_bytes += _snapshot_bytes(snapshots)
        self.kept.append((recorder.layout(), snapshots))


def run(spec, out_dir=None, parameters=None, progress=None):
    """Build and run a model (a dict or a path to its JSON) on the HIP path; returns Results and, with out_dir, saves
    them there. parameters {name: expression} override the model's. A study sweep runs one build per value; 2D runs
    store the stored-field snapshots of the study's output schedule (model/study.py:916-967, 2041-2130)."""
    if isinstance(spec, (str, os.PathLike)):
        spec = load(spec)
    study = spec.get('study') or {'type': 'time_dependent', 'settings': {'duration': 200e-15}}
    if study.get('type') != 'time_dependent':
        raise ValueError(f"study type {study.get('type')!r}: only 'time_dependent' is part of the FDTD path")
    if spec.get('derived_values'):
        raise ValueError('derived values are not part of the FDTD path')
    settings = study.get('settings') or {}
    requested = list(study.get('stored_fields') or [])
    overrides = dict(parameters or {})
    base = evaluate_parameters(spec.get('parameters') or [], overrides)
    sweep = study.get('sweep')
    runs = [overrides]
    kept = None
    if sweep:
        if sweep.get('parameter') not in base:
            raise ValueError(f"sweep parameter {sweep.get('parameter')!r} is not a model parameter")
        values = [evaluate(v, base) for v in sweep.get('values') or []]
        if not values:
            raise ValueError(f"sweep of {sweep['parameter']!r} has no values")
        runs = [dict(overrides, **{sweep['parameter']: v}) for v in values]
        kept = _KeptSweepSnapshots() if sweep.get('keep_fields', True) else None
        sweep = {'parameter': sweep['parameter'], 'values': values}
    wall = time.perf_counter()
    sweep_results, built, names, recorder = [], None, base, None
    for index, run_overrides in enumerate(runs):
        names = evaluate_parameters(spec.get('parameters') or [], run_overrides)
        if built is not None:
            built.sim.close()
        built = _Built(spec, names)
        recorder = _advance(built, settings, names, requested, _stored_cuts(study, names))
        sweep_results.append(built.results())
        if kept is not None:
            kept.take(recorder)
        if progress is not None:
            progress({'run': index, 'of': len(runs), 'steps': built.sim.engine.n_step})
    results = sweep_results[-1]
    snapshots = recorder.snapshots
    summary = {'status': 'completed', 'message': '', 'parameters': names,
               'results': {key: result_digest(result) for key, result in results.items()},
               't_end': float(built.sim.time), 'ticks': int(built.sim.engine.n_step), 'tick_unit': 'step',
               'snapshot_times': [s['time'] for s in snapshots], 'snapshot_labels': [s['label'] for s in snapshots],
               'stored_fields': list(recorder.keys), 'stored_cuts': recorder.cut_records(),
               'wall_time': time.perf_counter() - wall}
    if sweep:
        summary['sweep'] = dict(sweep, derived={}, reference={}, error={}, curve_error={},
                                results=[{key: result_digest(r) for key, r in run.items()} for run in sweep_results],
                                failed=[], converged=[None] * len(runs), derived_series=[None] * len(runs))
        if kept is not None and kept.kept:
            summary['sweep']['snapshot_times'] = [[s['time'] for s in snaps] for _, snaps in kept.kept]
            summary['sweep']['snapshot_labels'] = [[s['label'] for s in snaps] for _, snaps in kept.kept]
    # the field group: the stored fields of the final state on the native window cells
    fields = {key: (values, recorder.edges[0], recorder.edges[1]) for key, values in recorder.final_fields.items()}
    out = Results('completed', spec, results, sweep_results, sweep, fields, summary, snapshots, recorder.layout(),
                  kept.kept if kept is not None else ())
    built.sim.close()
    if out_dir is not None:
        out.save(out_dir)
    return out


__all__ = ['CONSTANTS', 'Results', 'STORED_FIELDS', 'evaluate', 'evaluate_parameters', 'load', 'result_digest', 'run',
           'snapshot_schedule']
