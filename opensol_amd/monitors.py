"""Running-DFT flux monitors (host side): phase tables, device accumulator layout and the flux readout.

The kernel samples E and H on the monitor line after every m-th step (physics/em/monitors.py:177-191 for m, :203-224
for the sampling) and adds exp(-j omega t) m dt * field into FP64 Kahan (sum, compensation) pairs; the reference adds
complex64 phases into complex128 sums (:22-34, :86-90). The readout is the reference's FluxMonitor.flux
(:371-379): H (sampled dt/2 before E) is re-timed by exp(j pi f dt), and the spectral power is
sign cos(pi f dt) Re(E conj(H)) summed with the cell widths along the line.
"""
import numpy as np


def dft_sample_interval(dt, sources, monitor_frequencies) -> int:
    """m = max(1, floor(1 / ((max source spectral limit + max monitor frequency) dt))); 1 without a source, without
    a monitor, or with a source whose spectrum has no limit (physics/em/monitors.py:177-191)."""
    if not sources or not monitor_frequencies:
        return 1
    limits = []
    for source in sources:
        spectral_limit = getattr(getattr(source, 'temporal', None), 'spectral_limit', None)
        limit = spectral_limit() if spectral_limit is not None else None
        if limit is None:
            return 1
        limits.append(float(limit))
    rate = max(limits) + max(float(np.max(f)) for f in monitor_frequencies)
    return max(1, int(np.floor(1.0 / (rate * float(dt)))))


class FluxMonitor:
    """Spectral Poynting flux through a line, > 0 toward +normal. normal 'x': line along y at x-cell `position`;
    normal 'y': line along x at y-cell `position`. E_z: (E_z, H_y) / (E_z, H_x), sign -1 / +1; H_z: (E_y, H_z) /
    (E_x, H_z), sign +1 / -1 (physics/em/monitors.py:325-369)."""

    def __init__(self, normal, position, span, frequencies, dt, widths, polarization):
        if normal not in ('x', 'y'):
            raise ValueError(f"FluxMonitor normal must be 'x' or 'y', got {normal!r}")
        self.normal = normal
        self.position = int(position)
        self.span = (int(span[0]), int(span[1]))
        if not 0 <= self.span[0] < self.span[1]:
            raise ValueError(f'FluxMonitor span must be a half-open cell range lo < hi, got {span}')
        self.frequencies = np.asarray(frequencies, dtype=np.float64)
        self.omega = 2.0 * np.pi * self.frequencies
        self.dt = float(dt)
        self.n_freqs = self.frequencies.size
        self.polarization = polarization
        if polarization == 'ez':
            self.sign = -1.0 if normal == 'x' else 1.0
        else:
            self.sign = 1.0 if normal == 'x' else -1.0
        widths = np.asarray(widths, dtype=np.float64)
        if widths.size < self.span[1]:
            raise ValueError(f'FluxMonitor widths (shape {widths.shape}) must cover span {self.span}')
        self.cell_widths = widths[self.span[0]:self.span[1]].copy()
        self.sample_interval = 1
        self.engine = None    # set when the device engine binds the monitor
        self.index = None

    @property
    def n_cells(self):
        return self.span[1] - self.span[0]

    def accumulator_shape(self):
        """(2 (E, H), n_freq, 4 (sum re, sum im, comp re, comp im), n_cells) float64 on the device."""
        return (2, self.n_freqs, 4, self.n_cells)

    def phase_table(self, sample_times):
        """(n_samples, n_freq, 2) float64: exp(-j omega t) m dt at each sample time."""
        phase = -np.outer(np.asarray(sample_times, dtype=np.float64), self.omega)
        scale = self.dt * self.sample_interval
        return np.stack([np.cos(phase) * scale, np.sin(phase) * scale], axis=-1)

    def spectra(self):
        """Host complex128 (E, H) spectra (n_freq, n_cells) at the E sample time (sum minus compensation)."""
        acc = self.engine.read_accumulator(self.index)
        values = acc[:, :, 0, :] - acc[:, :, 2, :] + 1j * (acc[:, :, 1, :] - acc[:, :, 3, :])
        return values[0], values[1]

    def flux(self):
        """Spectral power per unit length (n_freqs,) float64 (physics/em/monitors.py:371-379)."""
        e_spectra, h_spectra = self.spectra()
        h_spectra = h_spectra * np.exp(1j * np.pi * self.frequencies * self.dt)[:, None]
        e_time_average = np.cos(np.pi * self.frequencies * self.dt)
        integrand = self.sign * e_time_average[:, None] * np.real(e_spectra * np.conj(h_spectra))
        return integrand @ self.cell_widths


def face_samples(widths, first_edge, last_edge, wrap=False):
    """(first face, last face + 1, trapezoid weights) of the samples on cell edges first..last: face f is the hi face
    of cell f (edge f + 1), edge 0 has no face; wrap (a whole periodic axis): every face, dual widths
    (physics/em3d/monitors3d.py:22-34)."""
    widths = np.asarray(widths, dtype=np.float64)
    if wrap:
        return 0, widths.size, 0.5 * (widths + np.roll(widths, -1))
    edges = np.arange(max(first_edge, 1), last_edge + 1)
    weights = np.empty(edges.size)
    for position, edge in enumerate(edges):
        below = widths[edge - 1] if edge > first_edge else 0.0
        above = widths[edge] if edge < last_edge else 0.0
        weights[position] = 0.5 * (below + above)
    return int(edges[0]) - 1, int(edges[-1]), weights


class FluxPlane:
    """Spectral Poynting flux through the cell-edge plane `edge` normal to axis `normal` (0 / 1 / 2) over the edge
    ranges `bounds` of the tangential axes, > 0 toward +normal (physics/em3d/monitors3d.py:42-128). Two pairs: (E of
    the first tangential axis, H of the second) with sign +1 and (E second, H first) with sign -1, first = (normal +
    1) % 3. E is the mean of the two layers across the plane, H the layer below it; both are phased at the E time."""

    def __init__(self, normal, edge, bounds, frequencies, dt, widths, shape, periodic):
        if not 1 <= edge <= shape[normal] - 1:
            raise ValueError(f"flux plane at {'xyz'[normal]} edge {edge} needs a cell on both sides (edges 1 to "
                             f"{shape[normal] - 1})")
        self.normal, self.edge = int(normal), int(edge)
        self.frequencies = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
        self.omega = 2.0 * np.pi * self.frequencies
        self.n_freqs = self.frequencies.size
        self.dt = float(dt)
        self.sample_interval = 1
        self.tangential = tuple(a for a in range(3) if a != normal)   # grid order
        first, second = (normal + 1) % 3, (normal + 2) % 3
        self.pairs = []
        for e_axis, h_axis, face_axis, cell_axis, sign in ((first, second, first, second, 1.0),
                                                           (second, first, second, first, -1.0)):
            wrap = bool(periodic[face_axis]) and tuple(bounds[face_axis]) == (0, shape[face_axis])
            f_lo, f_hi, face_weights = face_samples(widths[face_axis], *bounds[face_axis], wrap=wrap)
            c_lo, c_hi = bounds[cell_axis]
            cell_weights = np.asarray(widths[cell_axis][c_lo:c_hi], dtype=np.float64)
            ranges = {face_axis: (f_lo, f_hi), cell_axis: (c_lo, c_hi)}
            weights = {face_axis: face_weights, cell_axis: cell_weights}
            a, b = self.tangential
            self.pairs.append(dict(e_comp=e_axis, h_comp=h_axis, sign=sign, lo=(ranges[a][0], ranges[b][0]),
                                   hi=(ranges[a][1], ranges[b][1]), area=np.outer(weights[a], weights[b])))
        self.engine = None
        self.index = None

    def accumulator_shape(self, pair):
        """(2 (E, H), n_freq, 4 (sum re, sum im, comp re, comp im), cells a, cells b) float64 on the device."""
        p = self.pairs[pair]
        return (2, self.n_freqs, 4, p['hi'][0] - p['lo'][0], p['hi'][1] - p['lo'][1])

    def phase_table(self, sample_times):
        """(n_samples, n_freq, 2) float64: exp(-j omega t) m dt at each sample time."""
        phase = -np.outer(np.asarray(sample_times, dtype=np.float64), self.omega)
        scale = self.dt * self.sample_interval
        return np.stack([np.cos(phase) * scale, np.sin(phase) * scale], axis=-1)

    def spectra(self):
        """Per pair (pair, E spectrum, H spectrum re-timed to the E time), complex128 (n_freq, cells a, cells b)
        (monitors3d.py:103-110)."""
        retime = np.exp(1j * np.pi * self.frequencies * self.dt)[:, None, None]
        out = []
        for q, pair in enumerate(self.pairs):
            acc = self.engine.read_plane_accumulator(self.index, q)
            values = acc[:, :, 0] - acc[:, :, 2] + 1j * (acc[:, :, 1] - acc[:, :, 3])
            out.append((pair, values[0], values[1] * retime))
        return out

    def flux(self):
        """Spectral power (n_freqs,) float64: cos(pi f dt) sum sign Re(E conj(H)) area (monitors3d.py:112-118)."""
        total = np.zeros(self.n_freqs)
        for pair, e_spectrum, h_spectrum in self.spectra():
            total += pair['sign'] * np.einsum('fij,ij->f', np.real(e_spectrum * np.conj(h_spectrum)), pair['area'])
        return np.cos(np.pi * self.frequencies * self.dt) * total


class FieldProbe:
    """Raw Yee samples of one field component at one cell, one after every step (core/monitors.py:194-283; the
    reference's FieldProbeRows sample every probe in one launch per step, :42-130). get_data() gives the times after
    each recorded step [s] (float64) and the values (float32), as FieldProbe.get_data."""

    def __init__(self, cell, field_component):
        self.cell = tuple(int(v) for v in cell)
        self.field_component = field_component
        self.engine = None
        self.index = None

    @property
    def ix(self):
        return self.cell[0]

    @property
    def iy(self):
        return self.cell[1]

    @property
    def iz(self):
        return self.cell[2] if len(self.cell) > 2 else None

    def get_data(self):
        if self.engine is None:
            return np.empty(0, dtype=np.float64), np.empty(0, dtype=np.float32)
        return self.engine.probe_series(self.index)


class ProbeTable:
    """Device table of an engine's field probes: flat cell indices (int64), component indices (int32) and the
    (capacity, n_probes) float32 samples the probe launch writes at row n after step n (the yee_probes record)."""

    def __init__(self, probes, cells, components, device):
        import torch
        self.torch = torch
        self.device = device
        self.n = len(probes)
        self.cell = torch.tensor(cells, dtype=torch.int64, device=device) if self.n else None
        self.component = torch.tensor(components, dtype=torch.int32, device=device) if self.n else None
        self.samples = None
        self.capacity = 0

    def ensure(self, capacity):
        """Room for rows [0, capacity); the rows recorded so far are kept."""
        if not self.n or capacity <= self.capacity:
            return
        grown = self.torch.zeros((int(capacity), self.n), dtype=self.torch.float32, device=self.device)
        if self.samples is not None:
            grown[:self.capacity] = self.samples
        self.samples = grown
        self.capacity = int(capacity)

    def fill(self, probes_c):
        """Fill a yee_probes record (ProbesC)."""
        probes_c.n_probes = self.n
        if not self.n:
            return
        probes_c.cell = self.cell.data_ptr()
        probes_c.component = self.component.data_ptr()
        probes_c.samples = self.samples.data_ptr()
        probes_c.capacity = self.capacity

    def series(self, index, n_step):
        return self.samples[:n_step, index].cpu().numpy()


class FluxBox:
    """Net outward flux through the six cell-edge faces of a box (physics/em3d/monitors3d.py:131-162): one bounded
    FluxPlane per face over the box's edge ranges, summed with sign -1 on the lo face and +1 on the hi face of each
    axis."""

    def __init__(self, faces):
        self.faces = list(faces)   # [(sign, FluxPlane)], x lo, x hi, y lo, y hi, z lo, z hi
        self.frequencies = self.faces[0][1].frequencies
        self.n_freqs = self.faces[0][1].n_freqs

    @property
    def sample_interval(self):
        return self.faces[0][1].sample_interval

    def flux(self):
        return sum(sign * plane.flux() for sign, plane in self.faces)


def dft_sample_slice(start, stop, stride):
    """Cells a DFT box samples along one axis of [start, stop): every stride-th cell from the middle of the first
    stride-cell block (physics/em/monitors.py:152-155)."""
    return slice(start + min(stride, stop - start) // 2, stop, stride)


class _DFTRegion:
    """Device record of a running DFT over a box of sampled cells (the yee_dft_region record): FP64 Kahan (sum re,
    sum im, compensation re, compensation im) accumulators [component][frequency][cells], phases exp(-j omega t) m dt at the sample times (the flux monitors'
    tables). Subclasses name the components and their pairing."""

    names = ()

    def _setup(self, frequencies, dt):
        self.frequencies = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
        if self.frequencies.ndim != 1 or not self.frequencies.size or not np.all(np.isfinite(self.frequencies)) \
                or np.any(self.frequencies <= 0):
            raise ValueError(f'DFT frequencies must be positive and finite, got {frequencies!r}')
        self.omega = 2.0 * np.pi * self.frequencies
        self.n_freqs = self.frequencies.size
        self.dt = float(dt)
        self.sample_interval = 1
        self.engine = None
        self.index = None

    @property
    def sample_counts(self):
        return tuple(len(range(s.start, s.stop, s.step)) for s in self.samples)

    def accumulator_shape(self):
        return (len(self.names), self.n_freqs) + self.sample_counts + (4,)

    def phase_table(self, sample_times):
        phase = -np.outer(np.asarray(sample_times, dtype=np.float64), self.omega)
        scale = self.dt * self.sample_interval
        return np.stack([np.cos(phase) * scale, np.sin(phase) * scale], axis=-1)

    def fill(self, region, components, phase, acc):
        """Fill a yee_dft_region (DftRegionC) for an engine whose field components are `components`."""
        samples = list(self.samples) + [slice(0, 1, 1)] * (3 - len(self.samples))
        region.n_comp = len(self.names)
        region.n_freq = self.n_freqs
        for a, s in enumerate(samples):
            region.start[a] = s.start
            region.count[a] = len(range(s.start, s.stop, s.step))
            region.stride[a] = s.step
        for c, name in enumerate(self.names):
            axis, wrap = self.pairs[c]
            region.comp[c] = components.index(name)
            region.pair_axis[c] = axis
            region.pair_wrap[c] = int(wrap)
        region.phase = phase.data_ptr()
        region.n_samples = phase.shape[0]
        region.acc = acc.data_ptr()

    def _spectra(self):
        """complex128 (component, n_freq, *sample counts) spectra from the device accumulators."""
        if self.engine is None:
            return np.zeros(self.accumulator_shape()[:-1], dtype=np.complex128)
        acc = self.engine.read_dft_accumulator(self.index)
        return (acc[..., 0] - acc[..., 2]) + 1j * (acc[..., 1] - acc[..., 3])


class DFTMonitor(_DFTRegion):
    """Running DFT over the cell box (x_slice, y_slice) at each frequency (physics/em/monitors.py:248-322): E_z
    ('ez'), or H_z plus E_x / E_y averaged to the cell centres ('hz': E_x with its lower y face, E_y with its lower x
    face; at index 0 the wrapped last face on a periodic axis, else the face itself); downsample s > 1 samples every
    s-th cell along x and y. ez_dft / hz_dft / ex_dft / ey_dft are complex128 (n_freq, mx, my) spectra (the
    reference's are complex64); H_z is re-timed from the E sample time to its own, t - dt/2."""

    def __init__(self, x_slice, y_slice, frequencies, dt, polarization='ez', periodic_x=False, periodic_y=False,
                 downsample=1):
        for label, cells in (('x_slice', x_slice), ('y_slice', y_slice)):
            if cells.step not in (None, 1) or cells.stop is None or not (cells.start or 0) < cells.stop:
                raise ValueError(f'DFTMonitor {label} must be a non-empty unit-step slice, got {cells}')
        if isinstance(downsample, bool) or int(downsample) != downsample or downsample < 1:
            raise ValueError(f'DFTMonitor downsample must be an int >= 1, got {downsample!r}')
        if polarization not in ('ez', 'hz'):
            raise ValueError(f"DFTMonitor polarization must be 'ez' or 'hz', got {polarization!r}")
        self._setup(frequencies, dt)
        self.polarization = polarization
        self.x_slice = slice(x_slice.start or 0, x_slice.stop)
        self.y_slice = slice(y_slice.start or 0, y_slice.stop)
        self.downsample = int(downsample)
        self.x_samples = dft_sample_slice(self.x_slice.start, self.x_slice.stop, self.downsample)
        self.y_samples = dft_sample_slice(self.y_slice.start, self.y_slice.stop, self.downsample)
        self.samples = (self.x_samples, self.y_samples)
        self.mx, self.my = self.sample_counts
        if polarization == 'ez':
            self.names, self.pairs = ('ez',), ((-1, False),)
        else:
            self.names = ('hz', 'ex', 'ey')
            self.pairs = ((-1, False), (1, bool(periodic_y)), (0, bool(periodic_x)))

    def _component(self, name):
        values = self._spectra()[self.names.index(name)]
        if name == 'hz':   # sampled with the E-time phase; the reference's H_z phase is at t - dt/2
            values = values * np.exp(1j * np.pi * self.frequencies * self.dt)[:, None, None]
        return values

    def __getattr__(self, name):
        if name in ('ez_dft', 'hz_dft', 'ex_dft', 'ey_dft') and name[:2] in self.__dict__.get('names', ()):
            return self._component(name[:2])
        raise AttributeError(name)


class DFTPlane3D(_DFTRegion):
    """Running DFT of E_x, E_y, E_z averaged to the cell centres of z layer k_grid over the window
    (physics/em3d/monitors3d.py:165-238): E_x with its lower x face, E_y with its lower y face, E_z with the layer
    below (at index 0: the wrapped last face or layer on a periodic axis, else the face itself)."""

    names = ('ex', 'ey', 'ez')

    def __init__(self, window, k_grid, frequencies, dt, periodic):
        self._setup(frequencies, dt)
        self.window = (slice(window[0].start, window[0].stop), slice(window[1].start, window[1].stop))
        self.k_grid = int(k_grid)
        self.shape = tuple(w.stop - w.start for w in self.window)
        self.samples = (slice(self.window[0].start, self.window[0].stop, 1),
                        slice(self.window[1].start, self.window[1].stop, 1), slice(self.k_grid, self.k_grid + 1, 1))
        self.pairs = ((0, bool(periodic[0])), (1, bool(periodic[1])), (2, bool(periodic[2])))

    @property
    def spectra(self):
        """{'ex', 'ey', 'ez'}: complex128 (n_freq, nx window, ny window) centre phasors."""
        values = self._spectra()[..., 0]
        return {name: values[c] for c, name in enumerate(self.names)}

    def spectrum(self, name, index):
        return self.spectra[name][index]

    def intensity(self, index):
        """float64 |E_x|^2 + |E_y|^2 + |E_z|^2 at frequency `index`."""
        return sum(np.abs(self.spectrum(name, index)) ** 2 for name in self.names)
