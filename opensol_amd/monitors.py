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
