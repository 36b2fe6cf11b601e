"""Shared 2D test cases: the same model built on the oracle (oracle.yee2d) and on the product (opensol_amd)."""
import numpy as np

THIN_FILM = dict(size_x=6e-6, size_y=3e-6, dx=25e-9, npml=15, boundaries='periodic_y',
                 slab=(2.75e-6, 0.0, 3.25e-6, 3e-6, 2.25), source=dict(axis='y', position=1.2e-6,
                                                                       wavelength=600e-9, bandwidth=0.8),
                 monitor=dict(axis='y', position=4.5e-6, span=(0.0, 3e-6), wavelengths=[500e-9, 600e-9]),
                 duration=80e-15)
SNAPSHOT_STEPS = (60, 165, 400)


def build(target, spec, polarization, with_slab=True, **options):
    """Build spec on `target` (the oracle Yee2D class or opensol_amd.Simulation); returns (sim, source, monitor)."""
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'], boundaries=spec['boundaries'],
                 polarization=polarization, **options)
    if with_slab and spec.get('slab'):
        x0, y0, x1, y1, eps = spec['slab']
        sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    for x0, y0, x1, y1, eps in spec.get('rectangles', ()):
        sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    source = sim.add_line_source(**spec['source']) if spec.get('source') else None
    monitor = sim.add_flux_monitor(**spec['monitor']) if spec.get('monitor') else None
    return sim, source, monitor


def random_rectangles(rng, size_x, size_y, count, eps_range=(1.0, 12.0), side_range=(0.1, 0.5)):
    """(x0, y0, x1, y1, eps) rectangles with centres uniform over the window (painted in order, later wins)."""
    out = []
    for _ in range(count):
        cx, cy = rng.uniform(0, size_x), rng.uniform(0, size_y)
        wx, wy = rng.uniform(*side_range) * size_x, rng.uniform(*side_range) * size_y
        out.append((cx - wx / 2, cy - wy / 2, cx + wx / 2, cy + wy / 2, float(rng.uniform(*eps_range))))
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    norm = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / norm) if norm > 0 else float(np.linalg.norm(a))
