/*
 * yee.h — C ABI of libyee.so, the MI355X (gfx950) FDTD Yee stepper.
 *
 * The library steps caller-owned device buffers (raw HBM pointers, e.g. from torch tensors): it never allocates or
 * frees device memory, never synchronises inside a step, and reads every per-step scalar (injected source value, DFT
 * phase, step index) from device tables, so a step sequence can be replayed from a HIP graph.
 *
 * State is double-buffered: after n steps the fields and CPML psi live in set n & 1. The absolute step index n of an
 * instance is its device clock (*clock), advanced by the library after each batch of launches.
 *
 * Layouts: 2D arrays are (nx, pitch) row-major (y contiguous, index i*pitch + j, columns ny..pitch-1 are padding that
 * stays 0); 3D arrays (nx, ny, nz) with z contiguous.
 * E_z polarisation: E_z at cell centres, H_x on hi-y faces, H_y on hi-x faces. H_z polarisation: H_z at centres,
 * E_x on hi-y faces, E_y on hi-x faces. 3D: E_x/E_y/E_z on the hi-x/y/z faces of cell (i, j, k), H on the edges.
 *
 * Every function returns 0 on success or a negative YEE_E* code; yee_last_error() gives a thread-local message.
 * A context is used from one host thread; distinct contexts are independent.
 */
#ifndef OPENSOL_AMD_YEE_H
#define OPENSOL_AMD_YEE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YEE_ABI_VERSION 19
#define YEE_MAX_SOURCES 8
#define YEE_MAX_MONITORS 4
#define YEE_MAX_DFT 4
#define YEE_MAX_POLES 8
#define YEE_DRUDE 1
#define YEE_LORENTZ 2

#define YEE_OK 0
#define YEE_EINVAL (-1)   /* invalid argument or descriptor */
#define YEE_EHIP (-2)     /* a HIP runtime call failed */
#define YEE_ESTATE (-3)   /* call out of order (e.g. stepping a destroyed context) */

/* Polarisations of a 2D instance. */
#define YEE_EZ 0   /* (E_z, H_x, H_y) */
#define YEE_HZ 1   /* (H_z, E_x, E_y) */

/*
 * Soft line (sheet) source: field[line cells] += wave[n] * profile after the E update of step n.
 * Replaces core/sources.py:221-229 (LineSource.inject -> _add_scaled_profile :146-155) and, for H_z,
 * physics/em/sources.py:469-480 (TangentialSheetCurrentSource.inject_electric on E_x / E_y).
 * axis 0: line along x at y-cell `line` (E_z, or E_x for H_z); axis 1: line along y at x-cell `line` (E_z, or E_y).
 * A point current is a one-cell line: the E_z filament current (core/sources.py:162-188 PointSource.inject,
 * physics/em/__init__.py:411-418), one face of an H_z in-plane electric dipole (physics/em/sources.py:447-466,
 * :431-440) or, with field YEE_SOURCE_H, the H_z magnetic line current (physics/em/sources.py:439-444, :420-429),
 * which adds wave[n] * profile to H_z after the H update of step n (wave sampled at t_n, not t_n + dt/2).
 */
#define YEE_WALL_ZERO 2   /* wall_sign of an H_z pmc face */
#define YEE_SOURCE_E 0   /* the E component named by the polarisation and axis, after the E update */
#define YEE_SOURCE_H 1   /* H_z (H_z instances only), after the H update */
typedef struct yee_line_source {
    int32_t axis;
    int32_t line;
    int32_t lo, hi;          /* half-open cell span along the line */
    int32_t field;           /* YEE_SOURCE_E or YEE_SOURCE_H */
    int32_t pad_;
    const float* profile;    /* (hi - lo) calibrated float32 profile */
    const float* wave;       /* float32 amplitude * s(t_n + dt/2) per absolute step n */
    int64_t wave_len;        /* entries in wave; steps n >= wave_len inject nothing */
} yee_line_source;

/*
 * Running-DFT flux line with FP64 Kahan-compensated accumulators, sampled after step n when (n + 1) % m == 0.
 * Replaces physics/em/monitors.py:22-34 (_dft_phase), :86-90 (_dft_accum_line), :203-224 (_LineSpectra.update)
 * for FluxMonitor (:325-369). normal 0: line along y at x-cell `position`; normal 1: line along x at y-cell.
 * E_z samples (E_z, H_y) or (E_z, H_x); H_z samples (E_y, H_z) or (E_x, H_z).
 */
typedef struct yee_flux_line {
    int32_t normal;
    int32_t position;
    int32_t lo, hi;          /* half-open cells along the line */
    int32_t n_freq;
    int32_t pad_;
    const double* phase;     /* [sample s][n_freq][2]: exp(-j omega t) m dt at t = t_{(s+1) m} */
    int64_t n_samples;
    double* acc;             /* [2 (E, H)][n_freq][4 (sum re, sum im, comp re, comp im)][hi - lo] */
} yee_flux_line;

/*
 * CPML of one axis: psi rows of the cell-centre ('e') and hi-face ('h') kinds. Replaces
 * physics/em/boundaries.py:689-772 (CPMLSided: psi = b psi + a (f_hi - f_lo)/d; field += sign coefficient psi),
 * with gain g = a / d of the bulk spacing per psi row.
 */
typedef struct yee_cpml_axis {
    int32_t n_rows;          /* psi rows per kind (0: no CPML on this axis) */
    int32_t pad_;
    const int32_t* row_e;    /* (n) psi row of each centre row, -1 outside the layers */
    const int32_t* row_h;    /* (n) psi row of each hi-face row, -1 outside */
    const float* b_e;
    const float* g_e;
    const float* b_h;
    const float* g_h;
    float* psi_e[2];         /* ping-pong psi of the centre rows: x axis (n_rows, pitch), y axis (nx, n_rows) */
    float* psi_h[2];         /* ping-pong psi of the hi-face rows, same layouts */
} yee_cpml_axis;

/*
 * One 2D instance. Replaces the 2D engine step of physics/em/solvers/fdtd_ez_2d.py:71-119 and
 * fdtd_hz_2d.py:144-189 (Yee kernels fdtd_ez_2d_kernels.py:23-167, fdtd_hz_2d_kernels.py:28-180, CPML, sources,
 * monitors) with one fused launch per step.
 * table per material id: E_z (ca, cb, da, db); H_z (da, db, eps_xx, eps_yy) and table2 (ca_x, cb_x, ca_y, cb_y of a
 * face between two cells of that id, sigma_e, 0, 0, 0); a face between two ids takes the mean eps / sigma (shares 1/2
 * on uniform grids) and forms ca, cb in the kernel (fdtd_hz_2d_kernels.py:190-278).
 */
/*
 * Field probes of an instance: after step n (on the device clock), samples[n * n_probes + k] = the field component
 * component[k] at the flat element index cell[k] of the new state set, for n < capacity (NaN for a record outside
 * the grid or the components). One small launch per step (also inside the step graphs; a batch samples all its
 * members' probes in one launch). Replaces the FieldProbe rows of the reference (core/monitors.py:42-130: one
 * probe_rows launch per step; physics/em/__init__.py:493-498, physics/em3d/physics3d.py:425-433).
 */
typedef struct yee_probes {
    int32_t n_probes;
    int32_t pad_;
    const int64_t* cell;        /* 2D: i * pitch + j; 3D: (i * ny + j) * nz + k */
    const int32_t* component;   /* index into the instance's field components */
    float* samples;             /* (capacity, n_probes) float32 */
    int64_t capacity;           /* rows of samples: steps n >= capacity record nothing */
} yee_probes;

/*
 * Running DFT of up to three field components over a box of sampled cells, with FP64 accumulators, sampled after step
 * n when (n + 1) % m == 0 (m = the instance's sample_interval). Sampled cell t of the box has grid index
 * start[a] + stride[a] * t_a on each axis a (t_a < count[a]; a 2D instance uses axes 0 and 1 with count[2] = 1).
 * Spectrum c samples component comp[c] of the new state set: the value at the cell, or (pair_axis[c] = a >= 0) the
 * float32 mean 0.5 (f[cell] + f[lower]) with the lower neighbour along a, where the lower neighbour of index 0 is the
 * last index (pair_wrap[c] = 1: periodic axis) or the cell itself (0). acc[c][f][t] += v exp(-j omega t_s) m dt.
 * Replaces physics/em/monitors.py:248-322 (DFTMonitor: E_z, or H_z and E_x, E_y at the cell centres, with a
 * downsample stride) and physics/em3d/monitors3d.py:165-238 (DFTPlane3D: E at the centres of one z layer). The
 * reference accumulates complex64; H_z, sampled at t - dt/2 there, is re-timed by the caller at readout.
 */
typedef struct yee_dft_region {
    int32_t n_comp;          /* 1..3 spectra */
    int32_t n_freq;
    int32_t start[3];
    int32_t count[3];
    int32_t stride[3];
    int32_t comp[3];         /* index into the instance's field components */
    int32_t pair_axis[3];    /* -1: the value at the cell; 0..2: mean with the lower neighbour along that axis */
    int32_t pair_wrap[3];
    const double* phase;     /* [sample s][n_freq][2]: exp(-j omega t) m dt at t = t_{(s+1) m} */
    int64_t n_samples;
    double* acc;             /* [n_comp][n_freq][count0 * count1 * count2][4 (sum re, sum im, comp re, comp im)]:
                                FP64 Kahan pairs; the spectrum is sum - comp */
} yee_dft_region;

/*
 * Dispersive pole of a 2D instance on one E component over the index box [i0, i1) x [j0, j1), in the fused E update
 * of the step kernel (auxiliary differential equations in velocity form, from E^n before the update; corrections after
 * the Yee E update with its CPML, before the sources):
 *   Drude (YEE_DRUDE):     J += damping J + drive E^n;                     E -= cb J
 *   Lorentz (YEE_LORENTZ): dP += damping dP - restoring P + drive E^n; P += dP;  E -= cb inv_dt dP
 * component 0: E_z at the cell centres (E_z instances); 1: E_x on the hi-y faces, 2: E_y on the hi-x faces (H_z
 * instances: one record per face set of a pole). cb is the cell's or face's E coefficient (0 on PEC); drive holds the
 * pole's drive coefficient times its cell mask, or the mask's face average (0 outside, so the state stays 0) and the
 * PEC keep. The state is read and written in place by the owner of its cell or face.
 * Replaces physics/em/solvers/fdtd_ez_2d.py:123-163, 318-358, fdtd_hz_2d.py:286-323, 441-473, 539-561 and the kernels
 * of fdtd_ez_2d_kernels.py:176-264; coefficients physics/em/materials.py:132-190.
 */
typedef struct yee_pole {
    int32_t kind;            /* YEE_DRUDE or YEE_LORENTZ */
    int32_t component;       /* 0 E_z, 1 E_x faces, 2 E_y faces */
    int32_t i0, i1, j0, j1;
    float damping, restoring, inv_dt;
    int32_t pad_;
    const float* drive;      /* (i1 - i0) * (j1 - j0), row-major */
    float* state[2];         /* Drude: J (state[1] unused); Lorentz: P, dP; (i1 - i0) * (j1 - j0) each */
} yee_pole;

/*
 * TFSF plane-wave box of a 2D instance: a +x vacuum plane wave in the total-field cells [x0, x1) x [y0, y1). The
 * incident values come from the reference's 1D incident line run on the host for every step (the line depends on
 * nothing in the grid): `line` holds one value per box column, `ends` the two values at the faces x0 - 1 and x1 - 1.
 * The corrections are added to every copy of a corrected value inside the step kernel:
 *   E_z (incident E_z, H_y): after the H update, H_y of rows x0-1 / x1-1 (columns [y0, y1)) += coef_h[0] / coef_h[1] *
 *     line[n][0] / line[n][L-1], H_x of columns y0-1 / y1-1 (rows i in [x0, x1)) += coef_h[2] / coef_h[3] *
 *     line[n][i - x0] (line = incident E_z at time n); after the E update, E_z of rows x0 / x1-1 += coef_e[0] /
 *     coef_e[1] * ends[n][0] / ends[n][1] (incident H_y at n + 1/2);
 *   H_z (incident E_y, H_z): after the H update, H_z of rows x0 / x1-1 += coef_h[0] / coef_h[1] * ends[n][0] /
 *     ends[n][1] (incident E_y at n); after the E update, E_y of rows x0-1 / x1-1 += coef_e[0] / coef_e[1] *
 *     line[n][0] / line[n][L-1], E_x of columns y0-1 / y1-1 += coef_e[2] / coef_e[3] * line[n][i - x0] (incident H_z
 *     at n + 1/2).
 * Coefficients carry their signs (0 on a side that opens onto a mirror face); E-side corrections skip PEC cells and
 * pinned faces (cb = 0). Replaces physics/em/sources.py:207-378 (TFSFSource, TFSFSourceHz and their correction
 * kernels :27-130).
 */
typedef struct yee_tfsf {
    int32_t x0, x1, y0, y1;
    float coef_h[4];
    float coef_e[4];
    const float* line;       /* [step][x1 - x0] */
    const float* ends;       /* [step][2] */
    int64_t n_steps;         /* rows of both tables; steps past them add nothing */
} yee_tfsf;

#define YEE_MAX_TFSF 2

typedef struct yee2d_desc {
    int32_t nx, ny;
    int32_t pitch;           /* row stride in elements of every (nx, *) array: a multiple of 4, >= ny */
    int32_t polarization;    /* YEE_EZ or YEE_HZ */
    int32_t periodic_x, periodic_y;
    int32_t rows_per_tile;   /* x rows marched by one wavefront tile (0: 4, balanced over nx) */
    int32_t lane_columns;    /* y columns per lane: 2 or 4 (0: 2); a tile spans 64 * lane_columns columns */
    float* field[2][3];      /* [set][component]: E_z (ez, hx, hy) or H_z (hz, ex, ey) */
    const uint16_t* material;
    const float* table;      /* (n_materials, 4) */
    const float* table2;     /* H_z: (n_materials, 8) */
    int32_t n_materials;
    float dt;                /* float32 time step (H_z face coefficients) */
    float dt_over_eps0;      /* float32 dt / eps0 (H_z face coefficients) */
    int32_t pad_;
    const float* inv_dx_e;   /* (nx) 1/width at centres, CPML kappa folded in */
    const float* inv_dx_h;   /* (nx) 1/spacing at hi faces (last: periodic wrap) */
    const float* inv_dy_e;   /* (ny) */
    const float* inv_dy_h;   /* (ny) */
    float inv_dx_bulk;       /* inv_d* outside the CPML layers (uniform grid) */
    float inv_dy_bulk;
    int32_t layer_x_lo;      /* rows [0, layer_x_lo) and [layer_x_hi, nx) hold every x CPML row (both kinds) */
    int32_t layer_x_hi;
    int32_t layer_y_lo;      /* columns, likewise for y */
    int32_t layer_y_hi;
    yee_cpml_axis cpml[2];   /* x, y */
    int32_t n_sources;
    int32_t n_monitors;
    int32_t sample_interval; /* DFT decimation m (physics/em/monitors.py:177-191) */
    int32_t pad2_;
    yee_line_source source[YEE_MAX_SOURCES];
    yee_flux_line monitor[YEE_MAX_MONITORS];
    int32_t n_dft;
    int32_t pad3_;
    yee_dft_region dft[YEE_MAX_DFT];
    int32_t n_poles;
    int32_t pad4_;
    yee_pole pole[YEE_MAX_POLES];
    yee_probes probes;
    /* Outer-face rows per face (x_lo, x_hi, y_lo, y_hi). E_z: 0 none, +1 the row copies its inner neighbour after
     * the step (pmc, symmetric mirror), -1 the negated copy (antisymmetric mirror); x faces first, a corner taking
     * both factors (physics/em/boundaries.py:478-488, wall_centre_rows :105-128). H_z (inside the step kernel, after
     * the H update, the magnetic currents and TFSF): 0 none, YEE_WALL_ZERO H_z = 0 on the outermost cells (pmc),
     * +1 the outermost H_z cells copy their inner neighbour and the tangential E on the plane face (face row 0 / n - 2)
     * is 0 after the sources (pec, antisymmetric mirror), -1 the negated copy (symmetric mirror); a corner takes both
     * faces' factors (:82-93, 500-538). */
    int32_t wall_sign[4];
    int32_t n_tfsf;
    int32_t pad5_;
    yee_tfsf tfsf[YEE_MAX_TFSF];
    int64_t* clock;          /* device int64: absolute step index n of the next step */
    int64_t step_limit;      /* steps n >= step_limit are skipped (batched instances); INT64_MAX for none */
} yee2d_desc;

/*
 * Point dipole of a 3D instance: E_component[i, j, k] += float32(weight * wave[n]) after the E update of step n.
 * Replaces physics/em3d/sources3d.py:41-75 (DipoleSource3D.inject; weight = face cb x moment / dual volume, the
 * calibration of :56-61 and physics/em3d/engine3d.py:299-311 done on the host).
 */
typedef struct yee_point_source {
    int32_t component;       /* 0 E_x, 1 E_y, 2 E_z */
    int32_t i, j, k;
    float weight;
    int32_t pad_;
    const float* wave;       /* float32 s(t_n + dt/2) per absolute step n */
    int64_t wave_len;
} yee_point_source;

/*
 * Running-DFT flux plane of a 3D instance (FP64 Kahan accumulators), sampled after step n when (n + 1) % m == 0.
 * Replaces physics/em3d/monitors3d.py:90-101 (PlaneFlux3D.update: E = mean of the two layers across the plane,
 * H on the plane). Pair p samples E component e_comp[p] on layers edge - 1 and edge along `normal` and H component
 * h_comp[p] on layer edge - 1, over index ranges [lo[p][a], hi[p][a]) of the two tangential axes a (in grid axis
 * order: the lower-numbered tangential axis first).
 */
typedef struct yee_flux_plane {
    int32_t normal;
    int32_t edge;
    int32_t e_comp[2];       /* 0..2 = E_x..E_z */
    int32_t h_comp[2];       /* 0..2 = H_x..H_z */
    int32_t lo[2][2];        /* [pair][tangential axis in grid order] */
    int32_t hi[2][2];
    int32_t n_freq;
    int32_t pad_;
    const double* phase;     /* [sample s][n_freq][2]: exp(-j omega t) m dt at t = t_{(s+1) m} */
    int64_t n_samples;
    double* acc[2];          /* per pair: [2 (E, H)][n_freq][4 (sum re, sum im, comp re, comp im)][cells a][cells b] */
} yee_flux_plane;

/*
 * Per-index constants of one axis of a 3D instance, one 32-byte record per grid index (cell centre rows 'e', hi-face
 * rows 'h'): the kappa-folded inverse spacings (inv_h[n-1] = 0 on a non-periodic axis, physics/em3d/engine3d.py:
 * 201-216) and the CPML psi row of each kind (-1 outside the layers) with its recursion b and gain = a x bulk inverse
 * spacing (physics/em3d/boundaries3d.py:208-241).
 */
typedef struct yee3d_axis_row {
    float inv_h;
    float inv_e;
    int32_t psi_row_h;
    float b_h;
    float g_h;
    int32_t psi_row_e;
    float b_e;
    float g_e;
} yee3d_axis_row;

/*
 * CPML psi of one axis of a 3D instance: psi of the axis' first and second component. Layouts: x axis (n_rows, ny,
 * nz), y axis (nx, n_rows, nz), z axis (nx, ny, n_rows). E psi is updated in place; H psi is ping-pong ([set] = step
 * parity, like the fields).
 */
typedef struct yee3d_cpml_axis {
    int32_t n_rows;          /* psi rows per kind (0: no CPML on this axis) */
    int32_t pad_;
    float* psi_e[2];         /* [first, second component] */
    float* psi_h[2][2];      /* [set][first, second component] */
} yee3d_cpml_axis;

#define YEE_MIRROR_SYMMETRIC 1
#define YEE_MIRROR_ANTISYMMETRIC 2
#define YEE3D_MAX_SOURCES 4
#define YEE3D_MAX_PLANES 8     /* a flux box is six bounded planes */

/*
 * One 3D instance. Replaces the 3D engine step of physics/em3d/engine3d.py:497-554 (update_h_3d / update_e_3d with
 * their fused CPML, kernels3d.py:335-400, 538-610; dipole sources; plane flux monitors) with one fused E/H launch
 * per step plus a plane-sample launch. table per material id: (da, db, ca, cb, eps_r, sigma_e, 0, 0): H edge
 * coefficients of the id, the E face coefficients of a face between two cells of the id, and the float32 eps_r and
 * sigma_e from which a face between two ids forms its mean coefficients (kernels3d.py:61-85, shares 1/2).
 */
/*
 * Dispersive pole record of a 3D instance on one E component (0 E_x, 1 E_y, 2 E_z: the hi faces along that axis) over
 * the index box [i0, i1) x [j0, j1) x [k0, k1): the updates and corrections of yee_pole, in the fused E update of the
 * step kernel. drive holds the pole's drive coefficient times the face average of its cell mask along the component's
 * axis. One pole is three records. Replaces physics/em3d/engine3d.py:431-530 (add_pole, _pole_region) and the kernels
 * of kernels3d.py:967-1004.
 */
typedef struct yee3d_pole {
    int32_t kind;            /* YEE_DRUDE or YEE_LORENTZ */
    int32_t component;       /* 0 E_x, 1 E_y, 2 E_z */
    int32_t i0, i1, j0, j1, k0, k1;
    float damping, restoring, inv_dt;
    int32_t pad_;
    const float* drive;      /* (i1 - i0) * (j1 - j0) * (k1 - k0), k fastest */
    float* state[2];         /* Drude: J; Lorentz: P, dP */
} yee3d_pole;

#define YEE3D_MAX_POLES 12

typedef struct yee3d_desc {
    int32_t nx, ny, nz;
    int32_t periodic[3];
    int32_t rows_per_tile;   /* y rows per wavefront tile: 4 (0: 4) */
    int32_t planes_per_tile; /* x planes marched by one wavefront (0: 4, balanced over nx) */
    float* field[2][6];      /* [set][ex, ey, ez, hx, hy, hz], (nx, ny, nz) z contiguous */
    const uint16_t* material;
    const float* table;      /* (n_materials, 8) */
    int32_t n_materials;
    float dt;
    const yee3d_axis_row* axis_rows[3];   /* per axis (n) records */
    yee3d_cpml_axis cpml[3];
    int32_t n_sources;
    int32_t n_planes;
    int32_t sample_interval;
    int32_t pad_;
    yee_point_source source[YEE3D_MAX_SOURCES];
    yee_flux_plane plane[YEE3D_MAX_PLANES];
    int32_t n_dft;
    int32_t pad2_;
    yee_dft_region dft[YEE_MAX_DFT];
    int32_t n_poles;
    int32_t pad3_;
    yee3d_pole pole[YEE3D_MAX_POLES];
    yee_probes probes;
    /* Mirror faces (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi): 0 none, YEE_MIRROR_SYMMETRIC, YEE_MIRROR_ANTISYMMETRIC. The
     * face pads one ghost cell (index 0 on a lo face, n - 1 on a hi face); the plane is the hi face of cell 0 / n - 2.
     * Symmetric: the tangential H on the plane is 0 after the H update (in the step kernel) and the normal E on the
     * plane 0 after the sources; antisymmetric: the ghost's tangential E = minus its inner neighbour's after the
     * sources (yee3d_mirror_kernel, zeros first, then images in face order). Replaces physics/em3d/engine3d.py:
     * 114-126, 513-540 (_mirror_rows and the mirror-row writes). */
    int32_t mirror[6];
    int64_t* clock;
    int64_t step_limit;
} yee3d_desc;

typedef struct yee_ctx yee_ctx;

/* Library ABI version (YEE_ABI_VERSION) and sizeof checks for bindings. */
int32_t yee_abi_version(void);
/* 0: yee2d_desc, 1: yee_line_source, 2: yee_flux_line, 3: yee_cpml_axis, 4: yee3d_desc, 5: yee_point_source,
 * 6: yee_flux_plane, 7: yee3d_cpml_axis, 8: yee3d_axis_row, 9: yee_probes, 10: yee_dft_region, 11: yee_pole,
 * 12: yee3d_pole, 13: yee_tfsf */
size_t yee_sizeof(int32_t which);
const char* yee_last_error(void);

/*
 * Create a stepping context for one 2D instance (host memory only; the descriptor is copied).
 * Replaces the engine construction of physics/em/solvers/fdtd_2d_base.py:120-223 for the stepping part.
 */
int32_t yee2d_create(const yee2d_desc* desc, yee_ctx** out);

/*
 * Advance by n_steps steps on `stream` (hipStream_t; 0 = the null stream). Launches only: no allocation, no sync.
 * Whole chunks of graph_steps steps replay a cached HIP graph when graph_steps > 0.
 * Replaces coupling/runner.py:174-204 (advance) -> EMPhysics.step (physics/em/__init__.py:163-168).
 */
int32_t yee_step(yee_ctx* ctx, int64_t n_steps, void* stream);

/*
 * Profiling form of yee_step: the same launches, one step at a time and never from a graph, with a HIP event pair
 * around every kernel. kernel_ms[c] receives the summed kernel time of class c over the n_steps steps and cells[c]
 * the cells one launch of that class updates (single 2D instance: class 0 = bulk tiles, 1 = edge tiles, or, when both
 * sets exist and run side by side as yee_step launches them, class 0 = both, every cell, from the fork to the join,
 * and class 1 empty; 3D: class 0 = the fused step kernel, every cell, 1 = the plane-sample kernel). Synchronizes `stream`. No counterpart in the
 * reference (its timing is wall clock around Simulation.run).
 */
int32_t yee_step_timed(yee_ctx* ctx, int64_t n_steps, void* stream, double* kernel_ms, int64_t* cells);

/* Steps per replayed HIP graph (0 disables graphs; default 0). Drops cached graphs. */
int32_t yee_set_graph_steps(yee_ctx* ctx, int32_t graph_steps);

/*
 * Capture and instantiate the graph of graph_steps steps now, without launching it, so that later yee_step calls only
 * replay it (no-op when graph_steps is 0 or a graph is cached). The reference captures its step graphs lazily on the
 * first step (physics/em/solvers/fdtd_2d_base.py:558-599); this moves that capture ahead of a timed region.
 */
int32_t yee_build_graph(yee_ctx* ctx);

/* Replace the descriptor (new buffers or tables); drops cached graphs. */
int32_t yee2d_update(yee_ctx* ctx, const yee2d_desc* desc);

/* The descriptor a single 2D context steps, with the defaults it resolved (rows_per_tile, lane_columns). */
int32_t yee2d_get_desc(const yee_ctx* ctx, yee2d_desc* out);

/*
 * Create a stepping context for one 3D instance (host memory only; the descriptor is copied). yee_step advances it.
 * Replaces the stepping part of physics/em3d/engine3d.py:162-263 (FDTDEngine3D construction).
 */
int32_t yee3d_create(const yee3d_desc* desc, yee_ctx** out);
int32_t yee3d_update(yee_ctx* ctx, const yee3d_desc* desc);
int32_t yee3d_get_desc(const yee_ctx* ctx, yee3d_desc* out);

int32_t yee_destroy(yee_ctx* ctx);

/*
 * Batched 2D instances (mixed E_z / H_z, uneven sizes) stepped by one launch per step over their tiles.
 * descs_dev: caller device buffer of yee2d_batch_bytes(n) bytes that receives the descriptors, each member's tile sets
 * and the bulk and edge tile prefixes (copied on `stream`, which is synchronized once before the call returns). A step
 * is two launches: the members' bulk tiles in the lean loops and their edge tiles (CPML, source, monitor and TFSF
 * lines, E_z and H_z pole rows, H_z wall rows) in the general loops; then, when any member has them, one launch each
 * over all members for E_z wall rows, DFT regions and field probes, in the single-instance order. Members share
 * lane_columns; every other feature of a single instance is allowed. Each instance keeps its own clock and
 * step_limit, so instances of different lengths share launches and finished ones skip their tiles. No counterpart in
 * the reference, which steps one model at a time (model/study.py:437-447 builds one model per sweep value).
 */
size_t yee2d_batch_bytes(int32_t n);
int32_t yee2d_batch_create(const yee2d_desc* descs, int32_t n, void* descs_dev, void* stream, yee_ctx** out);

/*
 * Batched 3D instances of uneven sizes stepped by one launch per step over their tiles (members with dispersive poles
 * take the E update that carries them), then one launch per mirror kind and face present in any member (symmetric
 * faces first, as a single instance orders them), one sample launch over all their flux-plane cells, and one launch
 * each over all members for DFT planes and field probes. descs_dev: caller device buffer of yee3d_batch_bytes(n) bytes
 * that receives the descriptors and the block and plane-cell prefixes (copied on `stream`, which is synchronized once
 * before the call returns). Each instance keeps its own clock, step_limit and planes_per_tile (4 rows per tile). No
 * counterpart in the reference, which steps one model at a time (model/study.py:437-447).
 */
size_t yee3d_batch_bytes(int32_t n);
int32_t yee3d_batch_create(const yee3d_desc* descs, int32_t n, void* descs_dev, void* stream, yee_ctx** out);

/*
 * yee2d_batch_step_static: the no-allocation form of yee2d_batch_create + yee_step. No counterpart in the reference;
 * it is for callers that cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc): it
 * calls no hipMalloc / hipFree and takes no workspace. Steps n batched 2D instances (n <= YEE_STATIC_BATCH_MAX) by
 * n_steps launches on `stream`, as a batch context would: each instance stops at its step_limit and keeps its own
 * clock. The descriptors (lane_columns and rows_per_tile set, as yee2d_get_desc resolves them; the batch's other
 * rules apply), their tile sets and the tile prefixes are copied on `stream` into a static device array of the
 * library, one per GPU (yee2d_batch_bytes(YEE_STATIC_BATCH_MAX) bytes, held until the process ends),
 * and `stream` is synchronized once after the copies, before the launches are enqueued. Every call on a GPU shares
 * that GPU's array: calls on one stream are ordered, but calls that overlap on one GPU (other streams or threads)
 * race. The current device must be the stream's device.
 */
#define YEE_STATIC_BATCH_MAX 256
int32_t yee2d_batch_step_static(const yee2d_desc* descs, int32_t n, int64_t n_steps, void* stream);

/*
 * yee_flag_nonfinite: enqueue on `stream` a check of data[0 .. count) (device floats, e.g. an instance's current field
 * set) that sets *flag (a device int32) to 1 when any value is NaN or infinite; the flag is never cleared, so the
 * caller zeroes it and reads it after the stream has passed the check. No counterpart in the reference, which checks
 * Courant stability before a run and nothing during it (physics/em/__init__.py:242-281); it lets a server stop a
 * diverging job at the next chunk boundary instead of at its end.
 */
int32_t yee_flag_nonfinite(const float* data, int64_t count, int32_t* flag, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPENSOL_AMD_YEE_H */
