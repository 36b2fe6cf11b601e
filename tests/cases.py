"""Shared 2D test cases: the same model built on the oracle (oracle.yee2d) and on the product (opensol_amd)."""
import numpy as np

THIN_FILM = dict(size_x=6e-6, size_y=3e-6, dx=25e-9, npml=15, boundaries='periodic_y',
                 slab=(2.75e-6, 0.0, 3.25e-6, 3e-6, 2.25), source=dict(axis='y', position=1.2e-6,
                                                                       wavelength=600e-9, bandwidth=0.8),
                 monitor=dict(axis='y', position=4.5e-6, span=(0.0, 3e-6), wavelengths=[500e-9, 600e-9]),
                 duration=80e-15)
SNAPSHOT_STEPS = (60, 165, 400)
# field probes of the thin-film case: (x, y, index into the polarization's components); window corners, the source
# line, the film's first cell and the flux line
THIN_FILM_PROBES = ((1.2e-6, 1.5e-6, 0), (3.0e-6, 0.7e-6, 1), (4.5e-6, 2.9e-6, 2), (0.0, 0.0, 0), (6e-6, 3e-6, 0),
                    (2.76e-6, 1.0e-6, 0), (4.5e-6, 1.2e-6, 1), (5.0e-6, 0.3e-6, 2))
PROBE_COMPONENTS = {'ez': ('ez', 'hx', 'hy'), 'hz': ('hz', 'ex', 'ey')}

# 2D cases with perfect electric conductors (the golden script tests/golden/make_pec2d_golden.py writes the same
# cases from the reference): a PEC rectangle under a dielectric painted after it, a line source crossing the PEC, and
# for E_z a 'pec' x_lo face.
PEC2D = dict(size_x=2e-6, size_y=1e-6, dx=25e-9, npml=10, pec_rect=(0.9e-6, 0.3e-6, 1.1e-6, 0.65e-6),
             rectangles=((1.0e-6, 0.0, 1.6e-6, 1e-6, 3.0),),
             source=dict(axis='y', position=1.0e-6, wavelength=600e-9, bandwidth=0.8),
             monitor=dict(axis='y', position=1.7e-6, span=(0.0, 1e-6), wavelengths=[500e-9, 600e-9, 700e-9]),
             faces={'ez': ('pec', 'cpml', 'periodic', 'periodic'), 'hz': ('cpml', 'cpml', 'periodic', 'periodic')},
             steps=(150, 500),
             dft_boxes=(((0.2e-6, 0.0, 1.8e-6, 0.5e-6), 1), ((0.3e-6, 0.1e-6, 1.9e-6, 0.9e-6), 3)))


def add_pec2d_dft_boxes(sim):
    """The PEC case's two DFT box monitors (full resolution from y cell 0; downsampled by 3)."""
    wavelengths = PEC2D['monitor']['wavelengths']
    return [sim.add_dft_monitor(*box, wavelengths=wavelengths, downsample=s) for box, s in PEC2D['dft_boxes']]


def build_pec2d(target, polarization, boundary_type, **options):
    """The PEC case on `target` (oracle Yee2D with oracle Boundaries, or Simulation with BoundaryConfig);
    boundary_type builds the face configuration from (faces, npml)."""
    spec = PEC2D
    bc = boundary_type(spec['faces'][polarization], spec['npml'])
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'], boundaries=bc,
                 polarization=polarization, **options)
    x0, y0, x1, y1 = spec['pec_rect']
    sim.add_pec_region(pec_rect_mask(sim, x0, y0, x1, y1))
    for x0, y0, x1, y1, eps in spec['rectangles']:
        sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    source = sim.add_line_source(**spec['source'])
    monitor = sim.add_flux_monitor(**spec['monitor'])
    return sim, source, monitor


# 2D cases driven by point currents (tests/golden/make_point2d_golden.py writes them from the reference): a dielectric
# and a PEC rectangle; E_z filament currents (one in the dielectric, one in vacuum, one inside the PEC); H_z a magnetic
# line current, an in-plane dipole at 30 degrees in y cell 0 (faces wrap) and one along y touching the PEC.
POINT2D = dict(size_x=1.5e-6, size_y=1e-6, dx=25e-9, npml=10, faces=('cpml', 'cpml', 'periodic', 'periodic'),
               dielectric=((0.2e-6, 0.0, 0.7e-6, 1e-6), 2.25), pec_rect=(1.0e-6, 0.4e-6, 1.2e-6, 0.6e-6),
               wavelength=600e-9, bandwidth=0.6, steps=(120, 400),
               ez_points=((0.4e-6, 0.5e-6, 2e-3), (0.85e-6, 0.3e-6, -1e-3), (1.1e-6, 0.5e-6, 1e-3)),
               hz_magnetic=(0.45e-6, 0.55e-6, 2.0),
               hz_dipoles=((0.8e-6, 0.01e-6, 1e-3, 30.0), (0.99e-6, 0.5e-6, 2e-3, 90.0)),
               beam=dict(axis='y', position=1.35e-6, waist=0.2e-6, center=0.6e-6, amplitude=0.5,
                         span=(0.1e-6, 0.9e-6)))


def build_point2d(target, polarization, boundary_type, **options):
    """The point-current case (with its Gaussian beam last) on `target`; returns (sim, sources) with one entry per
    reference point source (a dipole's entry is its list of face sources)."""
    spec = POINT2D
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'],
                 boundaries=boundary_type(spec['faces'], spec['npml']), polarization=polarization, **options)
    (x0, y0, x1, y1), eps = spec['dielectric']
    sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    sim.add_pec_region(pec_rect_mask(sim, *spec['pec_rect']))
    pulse = dict(wavelength=spec['wavelength'], bandwidth=spec['bandwidth'])
    if polarization == 'ez':
        sources = [sim.add_point_source(x, y, current=current, **pulse) for x, y, current in spec['ez_points']]
    else:
        x, y, km = spec['hz_magnetic']
        sources = [sim.add_magnetic_line_current(x, y, magnetic_current=km, **pulse)]
        sources += [sim.add_electric_dipole(x, y, dipole_moment=moment, orientation=angle, **pulse)
                    for x, y, moment, angle in spec['hz_dipoles']]
    sim.add_gaussian_beam(**spec['beam'], **pulse)
    return sim, sources


# 2D E_z case bounded by wall and mirror faces (tests/golden/make_walls2d_golden.py writes it from the reference):
# x_lo symmetric, x_hi pmc, y_lo antisymmetric, y_hi CPML; a dielectric, a filament current near the (x_lo, y_lo)
# corner and a line source over the whole grid line (into the pinned rows).
WALLS2D = dict(size_x=1.2e-6, size_y=1e-6, dx=25e-9, npml=10, faces=('symmetric', 'pmc', 'antisymmetric', 'cpml'),
               dielectric=((0.5e-6, 0.2e-6, 0.8e-6, 0.6e-6), 3.0), point=(0.1e-6, 0.12e-6, 1e-3),
               line=dict(axis='y', position=0.95e-6), wavelength=600e-9, bandwidth=0.6, steps=(150, 400))


def build_walls2d(target, boundary_type, polarization='ez', **options):
    """The wall and mirror case on `target` (oracle Yee2D or opensol_amd Simulation); H_z: x_lo and y_hi pmc, a
    magnetic line current beside their corner and a 45 degree dipole."""
    spec = WALLS2D
    faces = spec['faces'] if polarization == 'ez' else ('pmc', 'cpml', 'cpml', 'pmc')
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'],
                 boundaries=boundary_type(faces, spec['npml']), polarization=polarization, **options)
    (x0, y0, x1, y1), eps = spec['dielectric']
    sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    pulse = dict(wavelength=spec['wavelength'], bandwidth=spec['bandwidth'])
    if polarization == 'ez':
        x, y, current = spec['point']
        sim.add_point_source(x, y, current=current, **pulse)
        sim.add_line_source(**spec['line'], **pulse)
    else:
        sim.add_magnetic_line_current(0.05e-6, 0.95e-6, magnetic_current=2.0, **pulse)
        sim.add_electric_dipole(0.6e-6, 0.5e-6, orientation=45.0, **pulse)
    return sim


def pec_rect_mask(sim, x0, y0, x1, y1):
    """Cells whose float64 centre lies in [x0, x1) x [y0, y1) (the reference's Domain.rect_mask, core/domain.py:
    255-260), on either target."""
    geo = getattr(sim, 'geo', None)
    if geo is not None:
        xc, yc = geo.centres
    else:
        xc, yc = sim.domain.cell_centers()
    return ((xc[:, None] >= x0) & (xc[:, None] < x1)) & ((yc[None, :] >= y0) & (yc[None, :] < y1))


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


def add_thin_film_probes(sim, polarization):
    return [sim.add_field_probe(x, y, PROBE_COMPONENTS[polarization][c]) for x, y, c in THIN_FILM_PROBES]


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


def update_vector_scales(fields, scales):
    """Running maximum, over the checked steps, of the L2 norm of the whole E and of the whole H field (all their
    components in `fields`). Float32 rounding is relative to the whole field, so a component that symmetry keeps
    near zero (H_x of a z dipole between mirrors, H_z of a centred z dipole) or a field that has decayed since its
    peak is measured against this scale, not against its own current norm."""
    for kind in ('e', 'h'):
        names = [name for name in fields if name[0] == kind]
        if names:
            norm = float(np.sqrt(sum(np.linalg.norm(np.asarray(fields[n], np.float64)) ** 2 for n in names)))
            scales[kind] = max(scales.get(kind, 0.0), norm)
    return scales


def scaled_error(a, b, scales, name):
    """||a - b|| over the running scale of the field vector that component `name` belongs to."""
    diff = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(np.linalg.norm(diff) / scales[name[0]])

# 2D E_z case with dispersive poles (tests/golden/make_poles2d_golden.py writes it from the reference): a Drude
# rectangle (eps_inf 2) under a PEC corner, a Lorentz disc over a dielectric rectangle, a line source and a flux line.
POLES2D = dict(size_x=2e-6, size_y=1e-6, dx=25e-9, npml=10, boundaries='periodic_y',
               drude_rect=(1.0e-6, 0.2e-6, 1.3e-6, 0.7e-6), drude=dict(omega_p=5e15, gamma=1e14, eps_inf=2.0),
               pec_rect=(1.2e-6, 0.6e-6, 1.4e-6, 0.8e-6), dielectric=((0.5e-6, 0.1e-6, 0.9e-6, 0.9e-6), 2.0),
               lorentz_disc=((0.7e-6, 0.5e-6), 0.15e-6), lorentz=dict(delta_eps=1.5, omega_0=3e15, gamma=2e14),
               source=dict(axis='y', position=0.35e-6, wavelength=600e-9, bandwidth=0.8),
               monitor=dict(axis='y', position=1.6e-6, span=(0.0, 1e-6), wavelengths=[500e-9, 600e-9, 700e-9]),
               steps=(300, 800))


def build_poles2d(target, polarization='ez', **options):
    """The pole case on the oracle Yee2D, opensol_amd.Simulation or Model2D; returns (sim, monitor)."""
    spec = POLES2D
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'], boundaries=spec['boundaries'],
                 polarization=polarization, **options)
    (x0, y0, x1, y1), eps = spec['dielectric']
    sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    (cx, cy), radius = spec['lorentz_disc']
    if hasattr(sim, 'geo'):   # the oracle
        sim.add_drude_rectangle(*spec['drude_rect'], **spec['drude'])
        xc, yc = (c.astype(np.float32) for c in sim.geo.centres)
        disc = (xc[:, None] - np.float32(cx)) ** 2 + (yc[None, :] - np.float32(cy)) ** 2 <= np.float32(radius) ** 2
        sim.add_dispersive_lorentz(disc, **spec['lorentz'])
        sim.add_pec_region(sim.geo.rect_mask(*spec['pec_rect']))
    else:
        from opensol_amd import DrudePole, LorentzPole
        sim.add_drude_rectangle(*spec['drude_rect'], DrudePole(**spec['drude']))
        model = getattr(sim, 'model', sim)   # a Simulation or a bare Model2D
        sim.add_dispersive(LorentzPole(**spec['lorentz']), model.circle_mask(cx, cy, radius))
        sim.add_pec_region(model.rect_mask(*spec['pec_rect']))
    sim.add_line_source(**spec['source'])
    monitor = sim.add_flux_monitor(**spec['monitor'])
    return sim, monitor


# 3D case with dispersive poles (tests/golden/make_poles3d_golden.py writes it from the reference): eps_r 2 on a box
# with a Drude pole, a Lorentz pole on a second box, a z dipole and a flux plane normal to x; 8-cell CPML.
POLES3D = dict(window=(0.6e-6, 0.5e-6, 0.45e-6), dx=25e-9, npml=8,
               drude_box=((0.05e-6, 0.05e-6, 0.05e-6), (0.25e-6, 0.30e-6, 0.20e-6)), drude=(3e15, 1e14), drude_eps=2.0,
               lorentz_box=((0.35e-6, 0.20e-6, 0.25e-6), (0.55e-6, 0.45e-6, 0.40e-6)), lorentz=(1.5, 3e15, 2e14),
               dipole=(0.30e-6, 0.24e-6, 0.22e-6), plane_x=0.45e-6, wavelengths=[0.5e-6, 0.6e-6, 0.7e-6],
               n_steps=250)


def poles3d_box_mask(centres, lo, hi):
    """Cells whose float32 centre (clamped to the window) lies in the half-open box."""
    axis_masks = [(np.asarray(c, np.float32) >= np.float32(a)) & (np.asarray(c, np.float32) < np.float32(b))
                  for c, a, b in zip(centres, lo, hi)]
    return axis_masks[0][:, None, None] & axis_masks[1][None, :, None] & axis_masks[2][None, None, :]


def build_poles3d(target, **options):
    """The 3D pole case on the oracle Yee3D or on opensol_amd.Simulation3D; returns (sim, plane)."""
    spec = POLES3D
    sim = target(*spec['window'], spec['dx'], npml=spec['npml'], **options)
    if hasattr(sim, 'host_centres'):   # the oracle
        centres = sim.host_centres()
        drude = poles3d_box_mask(centres, *spec['drude_box'])
        lorentz = poles3d_box_mask(centres, *spec['lorentz_box'])
        sim.materials['eps_r'][drude] = np.float32(spec['drude_eps'])
        sim.add_dispersive_drude(drude, *spec['drude'])
        sim.add_dispersive_lorentz(lorentz, *spec['lorentz'])
    else:
        from opensol_amd import DrudePole, LorentzPole
        sim.add_dielectric_box(*spec['drude_box'], spec['drude_eps'])
        centres = sim.model.domain.host_centres(True)
        shape = sim.model.shape
        window = tuple(slice(0, n) for n in shape)
        sim.add_dispersive(DrudePole(*spec['drude']), window, poles3d_box_mask(centres, *spec['drude_box']))
        sim.add_dispersive(LorentzPole(*spec['lorentz']), window, poles3d_box_mask(centres, *spec['lorentz_box']))
    sim.add_dipole(*spec['dipole'], orientation='z', dipole_moment=1e-12, wavelength=0.6e-6, bandwidth=0.5)
    plane = sim.add_flux_plane('x', spec['plane_x'], wavelengths=spec['wavelengths'])
    return sim, plane


# 2D cases with conductive media and the ricker / cw waveforms (tests/golden/make_lossy_golden.py writes them from the
# reference): three rectangles with sigma_e > 0 painted in order (the first one around the Ricker line source, so its
# sheet calibration sees sigma_e), a cw line source over part of a line, a flux line behind the slab.
LOSSY2D = dict(size_x=2e-6, size_y=1e-6, dx=25e-9, npml=10, faces=('cpml', 'cpml', 'periodic', 'periodic'),
               rectangles=((0.3e-6, 0.0, 0.7e-6, 1e-6, 2.0, 5e3), (0.9e-6, 0.0, 1.5e-6, 1e-6, 3.0, 2e4),
                           (1.1e-6, 0.2e-6, 1.3e-6, 0.6e-6, 1.5, 3e5)),
               sources=(dict(axis='y', position=0.5e-6, wavelength=600e-9, waveform='ricker'),
                        dict(axis='y', position=0.35e-6, wavelength=700e-9, waveform='cw', amplitude=0.3,
                             span=(0.2e-6, 0.7e-6))),
               monitor=dict(axis='y', position=1.7e-6, span=(0.0, 1e-6), wavelengths=[500e-9, 600e-9, 700e-9]),
               steps=(150, 500), waveform_samples=600)


def build_lossy2d(target, polarization, boundary_type, **options):
    """The conductive-media case on `target` (oracle Yee2D with oracle Boundaries, or Simulation with BoundaryConfig);
    returns (sim, sources, monitor)."""
    spec = LOSSY2D
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'],
                 boundaries=boundary_type(spec['faces'], spec['npml']), polarization=polarization, **options)
    for x0, y0, x1, y1, eps, sigma in spec['rectangles']:
        sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps, sigma_e=sigma)
    sources = [sim.add_line_source(**source) for source in spec['sources']]
    monitor = sim.add_flux_monitor(**spec['monitor'])
    return sim, sources, monitor


# 3D case with conductive boxes, a Ricker z dipole and a cw x dipole (tests/golden/make_lossy_golden.py).
LOSSY3D = dict(window=(0.6e-6, 0.5e-6, 0.45e-6), dx=25e-9, npml=8,
               boxes=(((0.05e-6, 0.05e-6, 0.05e-6), (0.35e-6, 0.30e-6, 0.25e-6), 3.0, 2e4),
                      ((0.30e-6, 0.20e-6, 0.20e-6), (0.50e-6, 0.45e-6, 0.40e-6), 2.0, 3e5)),
               dipoles=(((0.27e-6, 0.24e-6, 0.22e-6), 'z', 1e-12, 0.6e-6, 'ricker'),
                        ((0.20e-6, 0.30e-6, 0.30e-6), 'x', 5e-13, 0.7e-6, 'cw')),
               plane_x=0.45e-6, wavelengths=(0.5e-6, 0.6e-6, 0.7e-6), n_steps=250)


def lossy3d_masks(centres):
    """One mask per LOSSY3D box: cells whose float32 centre (clamped to the window) lies in the half-open box."""
    return [poles3d_box_mask(centres, lo, hi) for lo, hi, _, _ in LOSSY3D['boxes']]


def build_lossy3d(target, **options):
    """The 3D conductive case on the oracle Yee3D or on opensol_amd.Simulation3D; returns (sim, dipoles, plane)."""
    spec = LOSSY3D
    sim = target(*spec['window'], spec['dx'], npml=spec['npml'], **options)
    if hasattr(sim, 'host_centres'):   # the oracle: paint the arrays as the golden script does
        for mask, (_, _, eps, sigma) in zip(lossy3d_masks(sim.host_centres()), spec['boxes']):
            sim.materials['eps_r'][mask] = np.float32(eps)
            sim.materials['sigma_e'][mask] = np.float32(sigma)
    else:
        for lo, hi, eps, sigma in spec['boxes']:
            sim.add_dielectric_box(lo, hi, eps, sigma_e=sigma)
    dipoles = [sim.add_dipole(*position, orientation=orientation, dipole_moment=moment, wavelength=wavelength,
                              bandwidth=0.5, waveform=waveform)
               for position, orientation, moment, wavelength, waveform in spec['dipoles']]
    plane = sim.add_flux_plane('x', spec['plane_x'], wavelengths=list(spec['wavelengths']))
    return sim, dipoles, plane


# 2D TFSF plane-wave boxes (tests/golden/make_tfsf_golden.py writes them from the reference): a +x Gaussian plane wave
# in the total-field box around a dielectric scatterer, CPML around; a flux line in the scattered-field region behind
# the box and one inside it. Case 'ez_mirror' puts a symmetric mirror on y_lo, which the box opens onto.
TFSF2D = dict(size_x=2e-6, size_y=1.4e-6, dx=25e-9, npml=10, box=(0.4e-6, 0.2e-6, 1.6e-6, 1.2e-6),
              scatterer=((0.8e-6, 0.5e-6, 1.1e-6, 0.9e-6), 4.0), wavelength=600e-9, bandwidth=0.5, amplitude=1.0,
              monitors=(dict(axis='y', position=1.8e-6, span=(0.0, 1.4e-6), wavelengths=[500e-9, 600e-9, 700e-9]),
                        dict(axis='y', position=1.3e-6, span=(0.3e-6, 1.1e-6), wavelengths=[500e-9, 600e-9, 700e-9])),
              faces={'ez': ('cpml',) * 4, 'hz': ('cpml',) * 4, 'ez_mirror': ('cpml', 'cpml', 'symmetric', 'cpml')},
              mirror_box=(0.4e-6, 0.0, 1.6e-6, 1.0e-6), steps=(150, 400))


def build_tfsf2d(target, case, boundary_type, **options):
    """The TFSF case `case` ('ez', 'hz', 'ez_mirror') on `target`; returns (sim, source, monitors)."""
    spec = TFSF2D
    polarization = 'hz' if case == 'hz' else 'ez'
    sim = target(spec['size_x'], spec['size_y'], spec['dx'], npml=spec['npml'],
                 boundaries=boundary_type(spec['faces'][case], spec['npml']), polarization=polarization, **options)
    (x0, y0, x1, y1), eps = spec['scatterer']
    sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    box = spec['mirror_box'] if case == 'ez_mirror' else spec['box']
    source = sim.add_tfsf(*box, wavelength=spec['wavelength'], bandwidth=spec['bandwidth'],
                          amplitude=spec['amplitude'])
    monitors = [sim.add_flux_monitor(**monitor) for monitor in spec['monitors']]
    return sim, source, monitors


# 3D cases with mirror faces (tests/golden/make_mirrors3d_golden.py writes them from the reference): a ghost cell
# padded beyond each mirror face, CPML on the other faces, two dielectric boxes, a dipole near the mirrors' corner and
# a flux plane normal to x.
MIRRORS3D = dict(window=(0.55e-6, 0.5e-6, 0.45e-6), dx=25e-9, npml=8,
                 boxes=(((0.05e-6, 0.05e-6, 0.05e-6), (0.25e-6, 0.25e-6, 0.20e-6), 3.0),
                        ((0.30e-6, 0.25e-6, 0.20e-6), (0.50e-6, 0.45e-6, 0.40e-6), 2.0)),
                 wavelength=0.6e-6, plane_x=0.4e-6, wavelengths=(0.5e-6, 0.6e-6, 0.7e-6), n_steps=200,
                 cases={'a': dict(faces={'x_lo': 'symmetric', 'y_hi': 'antisymmetric', 'z_lo': 'antisymmetric'},
                                  dipole=(0.08e-6, 0.22e-6, 0.07e-6), orientation='z'),
                        'b': dict(faces={'y_lo': 'symmetric', 'z_hi': 'symmetric', 'x_hi': 'antisymmetric'},
                                  dipole=(0.2e-6, 0.06e-6, 0.38e-6), orientation='x')})


def build_mirrors3d(target, name, **options):
    """Mirror case `name` on the oracle Yee3D or on opensol_amd.Simulation3D; returns (sim, dipole, plane)."""
    spec = MIRRORS3D
    case = spec['cases'][name]
    faces = {face: case['faces'].get(face, 'cpml') for face in ('x_lo', 'x_hi', 'y_lo', 'y_hi', 'z_lo', 'z_hi')}
    sim = target(*spec['window'], spec['dx'], npml=spec['npml'], boundaries=faces, **options)
    if hasattr(sim, 'host_centres'):   # the oracle: paint the array as the golden script does
        for lo, hi, value in spec['boxes']:
            sim.materials['eps_r'][poles3d_box_mask(sim.host_centres(), lo, hi)] = np.float32(value)
    else:
        for lo, hi, value in spec['boxes']:
            sim.add_dielectric_box(lo, hi, value)
    dipole = sim.add_dipole(*case['dipole'], orientation=case['orientation'], dipole_moment=1e-12,
                            wavelength=spec['wavelength'], bandwidth=0.5)
    plane = sim.add_flux_plane('x', spec['plane_x'], wavelengths=list(spec['wavelengths']))
    return sim, dipole, plane


# 2D H_z cases bounded by pec and mirror faces (tests/golden/make_hzwalls2d_golden.py writes them from the reference):
# 'a' x_lo pec, x_hi symmetric, y_lo antisymmetric, y_hi pec around a dielectric, with a magnetic line current by the
# (x_lo, y_lo) corner, an in-plane dipole by the y_hi wall and a line source along y; 'b' CPML on the lo faces, x_hi
# symmetric and a y_hi pec wall at column 128 of 129 (the first column of a wavefront tile of 128 columns), a
# magnetic line current beside that wall and a line source along x.
HZWALLS2D = dict(dx=25e-9, npml=10, wavelength=600e-9, bandwidth=0.6, steps=(150, 400),
                 cases={'a': dict(size=(1.0e-6, 1.0e-6), faces=('pec', 'symmetric', 'antisymmetric', 'pec'),
                                  dielectric=((0.3e-6, 0.3e-6, 0.6e-6, 0.7e-6), 3.0),
                                  magnetic=(0.04e-6, 0.06e-6, 2.0), dipole=(0.5e-6, 0.96e-6, 1e-3, 60.0),
                                  line=dict(axis='y', position=0.8e-6, span=(0.1e-6, 0.9e-6))),
                        'b': dict(size=(1.0e-6, 2.975e-6), faces=('cpml', 'symmetric', 'cpml', 'pec'),
                                  dielectric=((0.2e-6, 1.0e-6, 0.8e-6, 2.5e-6), 2.5),
                                  magnetic=(0.5e-6, 2.94e-6, 2.0), dipole=(0.7e-6, 2.96e-6, 1e-3, 0.0),
                                  line=dict(axis='x', position=2.0e-6, span=(0.0, 1.0e-6)))})


def build_hzwalls2d(target, name, boundary_type, **options):
    """H_z wall case `name` on `target` (oracle Yee2D or opensol_amd Simulation)."""
    spec = HZWALLS2D
    case = spec['cases'][name]
    sim = target(*case['size'], spec['dx'], npml=spec['npml'], boundaries=boundary_type(case['faces'], spec['npml']),
                 polarization='hz', **options)
    (x0, y0, x1, y1), eps = case['dielectric']
    sim.add_dielectric_rectangle(x0, y0, x1, y1, eps_r=eps)
    pulse = dict(wavelength=spec['wavelength'], bandwidth=spec['bandwidth'])
    x, y, km = case['magnetic']
    sim.add_magnetic_line_current(x, y, magnetic_current=km, **pulse)
    x, y, moment, angle = case['dipole']
    sim.add_electric_dipole(x, y, dipole_moment=moment, orientation=angle, **pulse)
    sim.add_line_source(**case['line'], **pulse)
    return sim
