// C ABI of libyee.so: stepping contexts over caller-owned device buffers (see include/yee.h).
// The library allocates no device memory; a context holds the host copy of its descriptor(s), the launch geometry
// and (optionally) an instantiated HIP graph that replays graph_steps steps plus the clock advance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "yee.h"
#include "yee3d_index.h"
#include "yee_launch.h"

template <int POL, int VEC, bool EDGE, bool POLES, bool LEFT>
__global__ void yee2d_step_kernel(yee2d_desc d, yee2d_tile_sets sets, int32_t offset);
template <int VEC, bool EDGE, bool LEFT>
__global__ void yee2d_batch_kernel(const yee2d_desc* __restrict__ descs, const yee2d_tile_sets* __restrict__ sets,
                                   const int32_t* __restrict__ prefix, int32_t n_inst, int32_t offset);
template <int J, bool POLES>
__global__ void yee3d_step_kernel(yee3d_desc d, int32_t tiles_j, int32_t tiles_k, int32_t tiles, int32_t offset);
template <int J>
__global__ void yee3d_batch_kernel(const yee3d_desc* __restrict__ descs, const int32_t* __restrict__ prefix,
                                   int32_t n_inst, int32_t offset);
__global__ void yee3d_sample_kernel(yee3d_desc d, int64_t cells_total, int32_t offset);
__global__ void yee3d_batch_sample_kernel(const yee3d_desc* __restrict__ descs, const int64_t* __restrict__ prefix,
                                          int32_t n_inst, int32_t offset);
__global__ void yee_clock_advance_batch3d(const yee3d_desc* __restrict__ descs, int32_t n_inst, int64_t k);
__global__ void yee_clock_advance(int64_t* clock, int64_t k, int64_t limit);
__global__ void yee2d_probe_kernel(yee2d_desc d, int32_t offset);
__global__ void yee2d_wall_kernel(yee2d_desc d, int32_t offset);
__global__ void yee3d_probe_kernel(yee3d_desc d, int32_t offset);
__global__ void yee3d_mirror_kernel(yee3d_desc d, int32_t face, int32_t offset);
__global__ void yee2d_dft_kernel(yee2d_desc d, int32_t offset);
__global__ void yee3d_dft_kernel(yee3d_desc d, int32_t offset);
__global__ void yee_clock_advance_batch(const yee2d_desc* __restrict__ descs, int32_t n_inst, int64_t k);
__global__ void yee2d_batch_probe_kernel(const yee2d_desc* __restrict__ descs, int32_t offset);
__global__ void yee2d_batch_wall_kernel(const yee2d_desc* __restrict__ descs, int32_t offset);
__global__ void yee2d_batch_dft_kernel(const yee2d_desc* __restrict__ descs, int32_t offset);
__global__ void yee3d_batch_dft_kernel(const yee3d_desc* __restrict__ descs, int32_t offset);
__global__ void yee3d_batch_mirror_kernel(const yee3d_desc* __restrict__ descs, int32_t face, int32_t kind,
                                          int32_t offset);
__global__ void yee3d_batch_probe_kernel(const yee3d_desc* __restrict__ descs, int32_t offset);
__global__ void yee_nonfinite_kernel(const float* __restrict__ data, int64_t count, int32_t* flag);

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t code, const std::string& message)
{
    g_last_error = message;
    return code;
}

#define YEE_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t err_ = (call);                                                                      \
        if (err_ != hipSuccess)                                                                        \
            return fail(YEE_EHIP, std::string(#call) + ": " + hipGetErrorString(err_));                \
    } while (0)

// Rows per 2D tile picked automatically. Each tile re-reads one edge row, yet short tiles measured fastest: on the
// 8192^2 E_z step (2 columns per lane) 8 rows took 353.3 us against 359.3 at 16, 370.7 at 32, 370.8 at 6 and
// 410.8 at 4 (and 371-377 us at 10, 12, 20 and 24 rows).
constexpr int kAutoRows = 8;

enum CtxKind { kSingle2D = 1, kBatch2D = 2, kSingle3D = 3, kBatch3D = 4 };

constexpr int kRows3D = yee3idx::kRows;    // y rows per 3D wavefront tile (yee3d_index.h)
constexpr int kOwned3D = yee3idx::kOwned;  // z cells a 3D wavefront tile owns (yee3d_index.h)
constexpr int kAutoPlanes = 4;     // default x planes per 3D tile

}  // namespace

struct yee_ctx {
    int kind = 0;
    yee2d_desc d2{};
    // batch
    std::vector<yee2d_desc> descs;
    yee2d_desc* descs_dev = nullptr;
    int32_t* prefix_dev = nullptr;
    int32_t n_inst = 0;
    // 2D batch: member tile sets and the bulk / general tile prefixes in the device buffer, and the tile counts
    yee2d_tile_sets* sets_dev = nullptr;
    int32_t* general_prefix_dev = nullptr;
    int64_t general_tiles = 0;
    // launch geometry
    int64_t tiles = 0;
    int blocks = 0;        // blocks of the bulk kernel (4 tiles each)
    int edge_blocks = 0;   // blocks of the edge kernel (sets.edge_parts wavefronts per tile, 4 per block)
    yee2d_tile_sets sets{};
    // 3D
    yee3d_desc d3{};
    int32_t tiles_j3 = 0, tiles_k3 = 0;
    int64_t sample_cells = 0;   // plane cells x pairs of the sample launch
    // 3D batch: descriptors, block prefix and plane-cell prefix in the caller's device buffer
    std::vector<yee3d_desc> descs3;
    yee3d_desc* descs3_dev = nullptr;
    int32_t* block_prefix_dev = nullptr;
    int64_t* sample_prefix_dev = nullptr;
    // graph replay
    int32_t graph_steps = 0;
    hipGraphExec_t graph = nullptr;
    hipStream_t capture_stream = nullptr;
    // 2D: the edge tile kernel runs beside the bulk one on a second stream (fork / join events)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
};

namespace {

// Field probe table: the counts and pointers are checked here. The device cells and components are not read on the
// host; the probe kernels sample NaN for a record outside the grid or the components.
int32_t check_probes(const yee_probes& p)
{
    if (p.n_probes < 0) return fail(YEE_EINVAL, "n_probes must be >= 0");
    if (p.n_probes > 0 && (!p.cell || !p.component || !p.samples || p.capacity < 0))
        return fail(YEE_EINVAL, "field probes need cell, component and samples arrays and a capacity >= 0");
    return YEE_OK;
}

// DFT regions of an instance with grid extents n_axis (2D: n_axis[2] = 1) and n_fields components per state set.
int32_t check_dft(const yee_dft_region* regions, int32_t n_dft, const int32_t* n_axis, int32_t dims, int32_t n_fields,
                  int32_t sample_interval)
{
    if (n_dft < 0 || n_dft > YEE_MAX_DFT) return fail(YEE_EINVAL, "n_dft out of range");
    if (n_dft > 0 && sample_interval < 1) return fail(YEE_EINVAL, "sample_interval must be >= 1");
    for (int32_t k = 0; k < n_dft; ++k) {
        const yee_dft_region& r = regions[k];
        if (r.n_comp < 1 || r.n_comp > 3) return fail(YEE_EINVAL, "DFT region n_comp must be 1..3");
        if (r.n_freq < 1 || !r.phase || !r.acc || r.n_samples < 0) return fail(YEE_EINVAL, "DFT region arrays missing");
        int64_t cells = 1;
        for (int a = 0; a < 3; ++a) {
            if (r.count[a] < 1 || r.stride[a] < 1 || r.start[a] < 0)
                return fail(YEE_EINVAL, "DFT region needs count, stride >= 1 and start >= 0");
            if ((int64_t)r.start[a] + (int64_t)r.stride[a] * (r.count[a] - 1) >= n_axis[a])
                return fail(YEE_EINVAL, "DFT region outside the grid");
            cells *= r.count[a];
        }
        if (cells > ((int64_t)1 << 38)) return fail(YEE_EINVAL, "DFT region too large");
        for (int c = 0; c < r.n_comp; ++c) {
            if (r.comp[c] < 0 || r.comp[c] >= n_fields) return fail(YEE_EINVAL, "DFT component out of range");
            if (r.pair_axis[c] < -1 || r.pair_axis[c] >= dims) return fail(YEE_EINVAL, "DFT pair axis out of range");
            if (r.pair_wrap[c] != 0 && r.pair_wrap[c] != 1) return fail(YEE_EINVAL, "DFT pair wrap must be 0 or 1");
        }
    }
    return YEE_OK;
}

// Blocks of 256 lanes along x of a DFT launch: enough for the largest region (blockIdx.y selects the region), at most
// kDftBlocks (the kernel strides over larger regions).
constexpr int64_t kDftBlocks = 2048;
int dft_blocks(const yee_dft_region* regions, int32_t n_dft)
{
    int64_t cells = 0;
    for (int32_t k = 0; k < n_dft; ++k)
        cells = std::max(cells, (int64_t)regions[k].count[0] * regions[k].count[1] * regions[k].count[2]);
    return (int)std::min(kDftBlocks, (cells + 255) / 256);
}

int32_t check_2d(const yee2d_desc* d)
{
    if (!d) return fail(YEE_EINVAL, "null descriptor");
    if (check_probes(d->probes) != YEE_OK) return YEE_EINVAL;
    if (d->nx < 2 || d->ny < 2) return fail(YEE_EINVAL, "2D grid needs nx, ny >= 2");
    const int32_t extents[3] = {d->nx, d->ny, 1};
    if (check_dft(d->dft, d->n_dft, extents, 2, 3, d->sample_interval) != YEE_OK) return YEE_EINVAL;
    if (d->n_poles < 0 || d->n_poles > YEE_MAX_POLES) return fail(YEE_EINVAL, "n_poles out of range");
    for (int p = 0; p < d->n_poles; ++p) {
        const yee_pole& pl = d->pole[p];
        if (pl.kind != YEE_DRUDE && pl.kind != YEE_LORENTZ) return fail(YEE_EINVAL, "pole kind must be 1 or 2");
        if (d->polarization == YEE_EZ ? pl.component != 0 : (pl.component != 1 && pl.component != 2))
            return fail(YEE_EINVAL, "pole component must be 0 (E_z) or 1, 2 (E_x, E_y faces of H_z)");
        if (pl.i0 < 0 || pl.i1 > d->nx || pl.i0 >= pl.i1 || pl.j0 < 0 || pl.j1 > d->ny || pl.j0 >= pl.j1)
            return fail(YEE_EINVAL, "pole box outside the grid");
        if (!pl.drive || !pl.state[0] || (pl.kind == YEE_LORENTZ && !pl.state[1]))
            return fail(YEE_EINVAL, "pole arrays missing");
    }
    if (d->pitch < d->ny || d->pitch % 4 != 0) return fail(YEE_EINVAL, "pitch must be a multiple of 4 and >= ny");
    if (d->polarization != YEE_EZ && d->polarization != YEE_HZ) return fail(YEE_EINVAL, "polarization must be 0 or 1");
    if (d->rows_per_tile < 1) return fail(YEE_EINVAL, "rows_per_tile must be >= 1");
    if (d->lane_columns != 2 && d->lane_columns != 4) return fail(YEE_EINVAL, "lane_columns must be 2 or 4");
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 3; ++c)
            if (!d->field[s][c]) return fail(YEE_EINVAL, "null field buffer");
    if (!d->material || !d->table || d->n_materials < 1) return fail(YEE_EINVAL, "material ids / table missing");
    if (d->polarization == YEE_HZ && !d->table2) return fail(YEE_EINVAL, "H_z needs table2");
    if (!d->inv_dx_e || !d->inv_dx_h || !d->inv_dy_e || !d->inv_dy_h) return fail(YEE_EINVAL, "spacings missing");
    const int n_axis[2] = {d->nx, d->ny};
    for (int a = 0; a < 2; ++a) {
        const yee_cpml_axis& c = d->cpml[a];
        if (c.n_rows < 0) return fail(YEE_EINVAL, "negative CPML rows");
        if (c.n_rows > 0) {
            if (!c.row_e || !c.row_h || !c.b_e || !c.g_e || !c.b_h || !c.g_h || !c.psi_e[0] || !c.psi_e[1] ||
                !c.psi_h[0] || !c.psi_h[1])
                return fail(YEE_EINVAL, "CPML arrays missing");
            if (c.n_rows > n_axis[a]) return fail(YEE_EINVAL, "more CPML rows than cells");
        }
    }
    if (d->n_sources < 0 || d->n_sources > YEE_MAX_SOURCES) return fail(YEE_EINVAL, "n_sources out of range");
    for (int s = 0; s < d->n_sources; ++s) {
        const yee_line_source& src = d->source[s];
        if (src.axis != 0 && src.axis != 1) return fail(YEE_EINVAL, "source axis must be 0 or 1");
        const int n_line = src.axis == 1 ? d->nx : d->ny;   // the fixed coordinate's extent
        const int n_along = src.axis == 1 ? d->ny : d->nx;
        if (src.line < 0 || src.line >= n_line) return fail(YEE_EINVAL, "source line outside the grid");
        if (src.lo < 0 || src.hi > n_along || src.lo >= src.hi) return fail(YEE_EINVAL, "source span out of range");
        if (!src.profile || !src.wave) return fail(YEE_EINVAL, "source profile / wave missing");
        if (src.field != YEE_SOURCE_E && !(src.field == YEE_SOURCE_H && d->polarization == YEE_HZ))
            return fail(YEE_EINVAL, "source field must be YEE_SOURCE_E (or YEE_SOURCE_H on an H_z instance)");
    }
    for (int f = 0; f < 4; ++f) {
        const int32_t sign = d->wall_sign[f];
        if (sign == 0) continue;
        const int axis_periodic = f < 2 ? d->periodic_x : d->periodic_y;
        const int n_axis = f < 2 ? d->nx : d->ny;
        if (d->polarization == YEE_EZ && sign != 1 && sign != -1)
            return fail(YEE_EINVAL, "E_z wall_sign must be -1, 0 or 1");
        if (d->polarization == YEE_HZ && sign != YEE_WALL_ZERO && sign != 1 && sign != -1)
            return fail(YEE_EINVAL, "H_z wall_sign must be -1, 0, 1 or YEE_WALL_ZERO");
        if (axis_periodic || n_axis < (d->polarization == YEE_HZ ? 3 : 2))
            return fail(YEE_EINVAL, "wall rows need a non-periodic axis (3+ cells for H_z, 2+ for E_z)");
    }
    if (d->n_tfsf < 0 || d->n_tfsf > YEE_MAX_TFSF) return fail(YEE_EINVAL, "n_tfsf out of range");
    for (int b = 0; b < d->n_tfsf; ++b) {
        const yee_tfsf& t = d->tfsf[b];
        // one scattered-field cell on every side: the corrected rows x0-1 .. x1-1 and columns y0-1 .. y1-1 exist
        if (t.x0 < 1 || t.x1 > d->nx - 1 || t.x0 >= t.x1 || t.y0 < 1 || t.y1 > d->ny - 1 || t.y0 >= t.y1)
            return fail(YEE_EINVAL, "TFSF box needs 1 <= x0 < x1 <= nx - 1 and 1 <= y0 < y1 <= ny - 1");
        if (!t.line || !t.ends || t.n_steps < 0) return fail(YEE_EINVAL, "TFSF incident tables missing");
    }
    if (d->n_monitors < 0 || d->n_monitors > YEE_MAX_MONITORS) return fail(YEE_EINVAL, "n_monitors out of range");
    if (d->n_monitors > 0 && d->sample_interval < 1) return fail(YEE_EINVAL, "sample_interval must be >= 1");
    for (int m = 0; m < d->n_monitors; ++m) {
        const yee_flux_line& mon = d->monitor[m];
        if (mon.normal != 0 && mon.normal != 1) return fail(YEE_EINVAL, "monitor normal must be 0 or 1");
        const int n_pos = mon.normal == 0 ? d->nx : d->ny;
        const int n_along = mon.normal == 0 ? d->ny : d->nx;
        if (mon.position < 0 || mon.position >= n_pos) return fail(YEE_EINVAL, "monitor line outside the grid");
        if (mon.lo < 0 || mon.hi > n_along || mon.lo >= mon.hi) return fail(YEE_EINVAL, "monitor span out of range");
        if (mon.n_freq < 1 || !mon.phase || !mon.acc) return fail(YEE_EINVAL, "monitor arrays missing");
    }
    if (!d->clock) return fail(YEE_EINVAL, "clock missing");
    return YEE_OK;
}

int32_t check_3d(const yee3d_desc* d)
{
    if (!d) return fail(YEE_EINVAL, "null descriptor");
    if (check_probes(d->probes) != YEE_OK) return YEE_EINVAL;
    const int n[3] = {d->nx, d->ny, d->nz};
    for (int a = 0; a < 3; ++a)
        if (n[a] < 2) return fail(YEE_EINVAL, "3D grid needs nx, ny, nz >= 2");
    if (check_dft(d->dft, d->n_dft, n, 3, 6, d->sample_interval) != YEE_OK) return YEE_EINVAL;
    if (d->n_poles < 0 || d->n_poles > YEE3D_MAX_POLES) return fail(YEE_EINVAL, "n_poles out of range");
    for (int p = 0; p < d->n_poles; ++p) {
        const yee3d_pole& pl = d->pole[p];
        if (pl.kind != YEE_DRUDE && pl.kind != YEE_LORENTZ) return fail(YEE_EINVAL, "pole kind must be 1 or 2");
        if (pl.component < 0 || pl.component > 2) return fail(YEE_EINVAL, "pole component must be 0, 1 or 2");
        if (pl.i0 < 0 || pl.i1 > d->nx || pl.i0 >= pl.i1 || pl.j0 < 0 || pl.j1 > d->ny || pl.j0 >= pl.j1 ||
            pl.k0 < 0 || pl.k1 > d->nz || pl.k0 >= pl.k1)
            return fail(YEE_EINVAL, "pole box outside the grid");
        if (!pl.drive || !pl.state[0] || (pl.kind == YEE_LORENTZ && !pl.state[1]))
            return fail(YEE_EINVAL, "pole arrays missing");
    }
    if ((int64_t)d->nx * d->ny * d->nz > ((int64_t)1 << 40)) return fail(YEE_EINVAL, "3D grid too large");
    if ((int64_t)d->ny * d->nz >= ((int64_t)1 << 29)) return fail(YEE_EINVAL, "3D plane (ny * nz) too large");
    if (d->rows_per_tile != kRows3D) return fail(YEE_EINVAL, "rows_per_tile must be 4 (or 0)");
    if (d->planes_per_tile < 1) return fail(YEE_EINVAL, "planes_per_tile must be >= 1");
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 6; ++c)
            if (!d->field[s][c]) return fail(YEE_EINVAL, "null field buffer");
    if (!d->material || !d->table || d->n_materials < 1) return fail(YEE_EINVAL, "material ids / table missing");
    for (int a = 0; a < 3; ++a) {
        if (!d->axis_rows[a]) return fail(YEE_EINVAL, "axis rows missing");
        if (reinterpret_cast<uintptr_t>(d->axis_rows[a]) % 32 != 0) return fail(YEE_EINVAL, "axis rows not 32 B aligned");
        const yee3d_cpml_axis& c = d->cpml[a];
        if (c.n_rows < 0) return fail(YEE_EINVAL, "negative CPML rows");
        if (c.n_rows > 0) {
            if (d->periodic[a]) return fail(YEE_EINVAL, "a periodic axis has no CPML");
            if (!c.psi_e[0] || !c.psi_e[1] || !c.psi_h[0][0] || !c.psi_h[0][1] || !c.psi_h[1][0] || !c.psi_h[1][1])
                return fail(YEE_EINVAL, "CPML arrays missing");
            if (c.n_rows > n[a]) return fail(YEE_EINVAL, "more CPML rows than cells");
        }
    }
    for (int f = 0; f < 6; ++f) {
        const int32_t kind = d->mirror[f];
        if (kind == 0) continue;
        if (kind != YEE_MIRROR_SYMMETRIC && kind != YEE_MIRROR_ANTISYMMETRIC)
            return fail(YEE_EINVAL, "mirror must be 0, YEE_MIRROR_SYMMETRIC or YEE_MIRROR_ANTISYMMETRIC");
        if (d->periodic[f >> 1] || n[f >> 1] < 3) return fail(YEE_EINVAL, "a mirror face needs a non-periodic axis of 3+ cells");
    }
    if (d->n_sources < 0 || d->n_sources > YEE3D_MAX_SOURCES) return fail(YEE_EINVAL, "n_sources out of range");
    for (int s = 0; s < d->n_sources; ++s) {
        const yee_point_source& src = d->source[s];
        if (src.component < 0 || src.component > 2) return fail(YEE_EINVAL, "source component must be 0..2");
        if (src.i < 0 || src.i >= d->nx || src.j < 0 || src.j >= d->ny || src.k < 0 || src.k >= d->nz)
            return fail(YEE_EINVAL, "source cell outside the grid");
        if (!src.wave) return fail(YEE_EINVAL, "source wave missing");
    }
    if (d->n_planes < 0 || d->n_planes > YEE3D_MAX_PLANES) return fail(YEE_EINVAL, "n_planes out of range");
    if (d->n_planes > 0 && d->sample_interval < 1) return fail(YEE_EINVAL, "sample_interval must be >= 1");
    for (int p = 0; p < d->n_planes; ++p) {
        const yee_flux_plane& pl = d->plane[p];
        if (pl.normal < 0 || pl.normal > 2) return fail(YEE_EINVAL, "plane normal must be 0..2");
        if (pl.edge < 1 || pl.edge > n[pl.normal] - 1) return fail(YEE_EINVAL, "plane edge needs a cell on both sides");
        const int axes[2] = {pl.normal == 0 ? 1 : 0, pl.normal == 2 ? 1 : 2};
        for (int q = 0; q < 2; ++q) {
            if (pl.e_comp[q] < 0 || pl.e_comp[q] > 2 || pl.h_comp[q] < 0 || pl.h_comp[q] > 2)
                return fail(YEE_EINVAL, "plane components must be 0..2");
            for (int t = 0; t < 2; ++t)
                if (pl.lo[q][t] < 0 || pl.hi[q][t] > n[axes[t]] || pl.lo[q][t] >= pl.hi[q][t])
                    return fail(YEE_EINVAL, "plane range out of the grid");
            if (!pl.acc[q]) return fail(YEE_EINVAL, "plane accumulators missing");
        }
        if (pl.n_freq < 1 || !pl.phase) return fail(YEE_EINVAL, "plane phases missing");
    }
    if (!d->clock) return fail(YEE_EINVAL, "clock missing");
    return YEE_OK;
}

int64_t tiles_j(const yee2d_desc& d)
{
    const int cols = 64 * d.lane_columns;
    return (d.ny + cols - 1) / cols;
}

int64_t tiles_2d(const yee2d_desc& d)
{
    return tiles_j(d) * ((d.nx + d.rows_per_tile - 1) / d.rows_per_tile);
}

using StepKernel = void (*)(yee2d_desc, yee2d_tile_sets, int32_t);

// Bulk-tile (edge = false) or edge-tile kernel of a descriptor.
// The 3D step kernel of an instance: the one with the pole update only when it has poles (its registers and its
// descriptor reads differ).
typedef void (*StepKernel3D)(yee3d_desc, int32_t, int32_t, int32_t, int32_t);
StepKernel3D step_kernel_3d(const yee3d_desc& d)
{
    return d.n_poles > 0 ? yee3d_step_kernel<kRows3D, true> : yee3d_step_kernel<kRows3D, false>;
}

// An H_z instance whose y_hi copy wall (pec or mirror) column ny - 1 is the first column of a tile: its edge tiles
// run the general loop with lane 0's recompute of column ny - 2.
bool needs_left_hz(const yee2d_desc& d)
{
    const int32_t w = d.wall_sign[3];
    return d.polarization == YEE_HZ && (w == 1 || w == -1) && (d.ny - 1) % (64 * d.lane_columns) == 0;
}

StepKernel step_kernel(const yee2d_desc& d, bool edge)
{
    if (d.polarization == YEE_EZ) {
        const bool poles = d.n_poles > 0;
        if (d.lane_columns == 2)
            return edge ? yee2d_step_kernel<YEE_EZ, 2, true, false, false>
                        : (poles ? yee2d_step_kernel<YEE_EZ, 2, false, true, false>
                                 : yee2d_step_kernel<YEE_EZ, 2, false, false, false>);
        return edge ? yee2d_step_kernel<YEE_EZ, 4, true, false, false>
                    : (poles ? yee2d_step_kernel<YEE_EZ, 4, false, true, false>
                             : yee2d_step_kernel<YEE_EZ, 4, false, false, false>);
    }
    const bool left = needs_left_hz(d);
    if (d.lane_columns == 2)
        return edge ? (left ? yee2d_step_kernel<YEE_HZ, 2, true, false, true>
                            : yee2d_step_kernel<YEE_HZ, 2, true, false, false>)
                    : yee2d_step_kernel<YEE_HZ, 2, false, false, false>;
    return edge ? (left ? yee2d_step_kernel<YEE_HZ, 4, true, false, true> : yee2d_step_kernel<YEE_HZ, 4, true, false, false>)
                : yee2d_step_kernel<YEE_HZ, 4, false, false, false>;
}

// Defaults of a 2D descriptor: 2 columns per lane; kAutoRows rows per tile.
void apply_defaults(yee2d_desc& d)
{
    if (d.lane_columns == 0) d.lane_columns = 2;
    if (d.rows_per_tile == 0 && d.nx > 0 && d.ny > 0 && (d.lane_columns == 2 || d.lane_columns == 4)) {
        const int64_t rows = std::min<int64_t>(kAutoRows, d.nx);
        // even out the tiles: the same count with the shortest rows
        const int64_t tiles_x = (d.nx + rows - 1) / rows;
        d.rows_per_tile = (int32_t)((d.nx + tiles_x - 1) / tiles_x);
    }
}

// Edge tile rows and columns of a descriptor: tile rows whose rows (with the row before and the row after) reach an x
// CPML layer or hold a source / monitor line along y, tile columns likewise for y. Too many lines: every tile is
// an edge tile. E_z poles run in both loops; H_z pole rows are edge rows.
// batch: the member of a batched launch, whose lean loop has no pole update, so E_z pole rows are edge rows too.
yee2d_tile_sets tile_sets(const yee2d_desc& d, bool batch = false)
{
    yee2d_tile_sets s{};
    const int cols = 64 * d.lane_columns;
    s.tiles_j = (d.ny + cols - 1) / cols;
    s.tiles_i = (d.nx + d.rows_per_tile - 1) / d.rows_per_tile;
    std::vector<int32_t> rows, columns;
    for (int ti = 0; ti < s.tiles_i; ++ti) {
        const int i0 = ti * d.rows_per_tile, i1 = std::min(d.nx, i0 + d.rows_per_tile);
        const int a = i0 > 0 ? i0 - 1 : 0, b = i1 < d.nx ? i1 : d.nx - 1;
        bool edge = a < d.layer_x_lo || b >= d.layer_x_hi;
        for (int k = 0; k < d.n_sources; ++k)
            edge = edge || (d.source[k].axis == 1 && d.source[k].line >= a && d.source[k].line <= b);
        for (int k = 0; k < d.n_monitors; ++k)
            edge = edge || (d.monitor[k].normal == 0 && d.monitor[k].position >= a && d.monitor[k].position <= b);
        if (d.polarization == YEE_HZ)   // H_z wall rows are pinned in the general loop
            edge = edge || (d.wall_sign[0] != 0 && a == 0) || (d.wall_sign[1] != 0 && b >= d.nx - 1);
        for (int k = 0; (d.polarization == YEE_HZ || batch) && k < d.n_poles; ++k)   // lean loops without poles
            edge = edge || (d.pole[k].i0 < i1 && d.pole[k].i1 > i0);
        for (int k = 0; k < d.n_tfsf; ++k) {   // TFSF corrected rows x0-1, x0, x1-1
            const yee_tfsf& t = d.tfsf[k];
            for (const int r : {t.x0 - 1, t.x0, t.x1 - 1}) edge = edge || (r >= a && r <= b);
        }
        if (edge) rows.push_back(ti);
    }
    for (int tj = 0; tj < s.tiles_j; ++tj) {
        const int j0 = tj * cols, j1 = std::min(d.ny, j0 + cols);
        const int a = j0 > 0 ? j0 - 1 : 0, b = j1 < d.ny ? j1 : d.ny - 1;
        bool edge = a < d.layer_y_lo || b >= d.layer_y_hi;
        if (d.polarization == YEE_HZ)
            edge = edge || (d.wall_sign[2] != 0 && a == 0) || (d.wall_sign[3] != 0 && b >= d.ny - 1);
        for (int k = 0; k < d.n_sources; ++k)
            edge = edge || (d.source[k].axis == 0 && d.source[k].line >= a && d.source[k].line <= b);
        for (int k = 0; k < d.n_monitors; ++k)
            edge = edge || (d.monitor[k].normal == 1 && d.monitor[k].position >= a && d.monitor[k].position <= b);
        for (int k = 0; k < d.n_tfsf; ++k) {   // TFSF corrected columns y0-1, y1-1
            const yee_tfsf& t = d.tfsf[k];
            for (const int c : {t.y0 - 1, t.y1 - 1}) edge = edge || (c >= a && c <= b);
        }
        if (edge) columns.push_back(tj);
    }
    if ((int)rows.size() > YEE_MAX_EDGE_LINES || (int)columns.size() > YEE_MAX_EDGE_LINES) {
        s.all_edge = 1;
        return s;
    }
    s.n_rows = (int32_t)rows.size();
    s.n_cols = (int32_t)columns.size();
    for (int k = 0; k < s.n_rows; ++k) s.rows[k] = rows[k];
    for (int k = 0; k < s.n_cols; ++k) s.cols[k] = columns[k];
    return s;
}

void set_geometry(yee_ctx* ctx)
{
    const yee2d_desc& d = ctx->d2;
    ctx->sets = tile_sets(d);
    const yee2d_tile_sets& s = ctx->sets;
    ctx->tiles = tiles_2d(d);
    const int64_t bulk = s.all_edge ? 0 : (int64_t)(s.tiles_i - s.n_rows) * std::max(0, s.tiles_j - s.n_cols);
    const int64_t edge = ctx->tiles - bulk;
    ctx->blocks = (int)((bulk + 3) / 4);
    ctx->sets.edge_parts = yee_edge_split(d.rows_per_tile, edge);
    ctx->edge_blocks = (int)((edge * ctx->sets.edge_parts + 3) / 4);
}

// Defaults of a 3D descriptor: 4 rows per tile and kAutoPlanes x planes per tile (balanced over the x extent).
// Short tiles re-read one extra E plane per kAutoPlanes planes but keep more tiles in flight; 4 measured fastest on
// 512^3 (2.05 ms per step against 2.15 ms at 8 and slower beyond).
void apply_defaults_3d(yee3d_desc& d)
{
    if (d.rows_per_tile == 0) d.rows_per_tile = kRows3D;
    if (d.planes_per_tile == 0 && d.nx > 0) {
        const int64_t planes = std::min<int64_t>(kAutoPlanes, d.nx);
        const int64_t chunks = (d.nx + planes - 1) / planes;
        d.planes_per_tile = (int32_t)((d.nx + chunks - 1) / chunks);
    }
}

// Wavefront tiles of a 3D instance: x chunks x y tiles x z tiles.
int64_t tiles_3d(const yee3d_desc& d)
{
    const int64_t chunks = (d.nx + d.planes_per_tile - 1) / d.planes_per_tile;
    return chunks * ((d.ny + kRows3D - 1) / kRows3D) * ((d.nz + kOwned3D - 1) / kOwned3D);
}

// Plane cells x pairs sampled by the sample launch of a 3D instance.
int64_t sample_cells_3d(const yee3d_desc& d)
{
    int64_t cells = 0;
    for (int p = 0; p < d.n_planes; ++p)
        for (int q = 0; q < 2; ++q)
            cells += (int64_t)(d.plane[p].hi[q][0] - d.plane[p].lo[q][0]) * (d.plane[p].hi[q][1] - d.plane[p].lo[q][1]);
    return cells;
}

void set_geometry_3d(yee_ctx* ctx)
{
    const yee3d_desc& d = ctx->d3;
    ctx->tiles_j3 = (d.ny + kRows3D - 1) / kRows3D;
    ctx->tiles_k3 = (d.nz + kOwned3D - 1) / kOwned3D;
    ctx->tiles = tiles_3d(d);
    ctx->blocks = (int)((ctx->tiles + 3) / 4);
    ctx->sample_cells = sample_cells_3d(d);
}

void drop_graph(yee_ctx* ctx)
{
    if (ctx->graph) {
        (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
    }
}

// The mirror-row launches of a 3D instance after its step kernel: symmetric faces first, then antisymmetric ones.
void launch_mirrors_3d(const yee3d_desc& d, int32_t offset, hipStream_t stream)
{
    for (const int kind : {YEE_MIRROR_SYMMETRIC, YEE_MIRROR_ANTISYMMETRIC}) {
        for (int face = 0; face < 6; ++face) {
            if (d.mirror[face] != kind) continue;
            const int a = face >> 1;
            const int64_t cells = (a == 0 ? (int64_t)d.ny * d.nz : a == 1 ? (int64_t)d.nx * d.nz : (int64_t)d.nx * d.ny);
            hipLaunchKernelGGL(yee3d_mirror_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, stream, d,
                               face, offset);
        }
    }
}

// The launches a 3D batch's members need after each step: per mirror kind (symmetric, antisymmetric) and face, the
// blocks of the largest such face of any member (0: none has it); DFT regions and probes as in BatchAux2D.
struct BatchAux3D {
    int mirror_blocks[2][6] = {};
    int dft_blocks = 0, dft_regions = 0;
    bool probes = false;
};

BatchAux3D batch_aux_3d(const yee3d_desc* descs, int32_t n)
{
    BatchAux3D a;
    for (int32_t k = 0; k < n; ++k) {
        const yee3d_desc& d = descs[k];
        for (int face = 0; face < 6; ++face) {
            const int kind = d.mirror[face] == YEE_MIRROR_SYMMETRIC ? 0 : d.mirror[face] == YEE_MIRROR_ANTISYMMETRIC ? 1 : -1;
            if (kind < 0) continue;
            const int axis = face >> 1;
            const int64_t cells = axis == 0 ? (int64_t)d.ny * d.nz : axis == 1 ? (int64_t)d.nx * d.nz
                                                                               : (int64_t)d.nx * d.ny;
            a.mirror_blocks[kind][face] = std::max(a.mirror_blocks[kind][face], (int)((cells + 255) / 256));
        }
        if (d.n_dft > 0) {
            a.dft_blocks = std::max(a.dft_blocks, dft_blocks(d.dft, d.n_dft));
            a.dft_regions = std::max(a.dft_regions, (int)d.n_dft);
        }
        a.probes = a.probes || d.probes.n_probes > 0;
    }
    return a;
}

bool has_walls(const yee2d_desc& d)
{
    return d.wall_sign[0] != 0 || d.wall_sign[1] != 0 || d.wall_sign[2] != 0 || d.wall_sign[3] != 0;
}

// The E_z copy launch (H_z walls act inside the H update).
bool has_ez_walls(const yee2d_desc& d)
{
    return d.polarization == YEE_EZ && has_walls(d);
}

// One lane per column (x faces) and per row (y faces).
int wall_blocks(const yee2d_desc& d) { return (int)(((int64_t)d.nx + d.ny + 255) / 256); }

// Device buffer of a 2D batch: the member descriptors, their tile sets, then the bulk and the edge tile prefixes
// (tile counts of members 0..k-1, n + 1 entries each).
struct Batch2DLayout {
    size_t sets, bulk_prefix, edge_prefix, bytes;
};

constexpr Batch2DLayout batch2d_layout(int32_t n)
{
    const size_t sets = sizeof(yee2d_desc) * (size_t)n;
    const size_t bulk = sets + sizeof(yee2d_tile_sets) * (size_t)n;
    const size_t edge = bulk + sizeof(int32_t) * (size_t)(n + 1);
    return Batch2DLayout{sets, bulk, edge, edge + sizeof(int32_t) * (size_t)(n + 1)};
}

// The launches a 2D batch's members need after each step, over all members at once: E_z walls (blocks along x of the
// largest member's wall lanes), DFT regions (blocks of the largest region, most regions of a member) and probes. Zero
// blocks: no member needs that launch.
struct BatchAux2D {
    int wall_blocks = 0, dft_blocks = 0, dft_regions = 0;
    bool probes = false;
    bool left = false;   // an H_z member with a y_hi copy wall on a tile edge: the edge kernel's LEFT instantiation
};

BatchAux2D batch_aux_2d(const yee2d_desc* descs, int32_t n)
{
    BatchAux2D a;
    for (int32_t k = 0; k < n; ++k) {
        const yee2d_desc& d = descs[k];
        if (has_ez_walls(d)) a.wall_blocks = std::max(a.wall_blocks, wall_blocks(d));
        if (d.n_dft > 0) {
            a.dft_blocks = std::max(a.dft_blocks, dft_blocks(d.dft, d.n_dft));
            a.dft_regions = std::max(a.dft_regions, (int)d.n_dft);
        }
        a.probes = a.probes || d.probes.n_probes > 0;
        a.left = a.left || needs_left_hz(d);
    }
    return a;
}

// A 2D batch on the device: where its arrays are, how many tiles of each class it has and its after-step launches.
struct Batch2DArgs {
    const yee2d_desc* descs;
    const yee2d_tile_sets* sets;
    const int32_t* bulk_prefix;
    const int32_t* edge_prefix;
    int32_t n;
    int64_t bulk_tiles, edge_tiles;
    int32_t lane_columns;
    BatchAux2D aux;
};

Batch2DArgs batch2d_args(void* base, int32_t n, int64_t bulk_tiles, int64_t edge_tiles, int32_t lane_columns,
                         const BatchAux2D& aux)
{
    const Batch2DLayout l = batch2d_layout(n);
    char* b = static_cast<char*>(base);
    return Batch2DArgs{reinterpret_cast<const yee2d_desc*>(b), reinterpret_cast<const yee2d_tile_sets*>(b + l.sets),
                       reinterpret_cast<const int32_t*>(b + l.bulk_prefix),
                       reinterpret_cast<const int32_t*>(b + l.edge_prefix), n, bulk_tiles, edge_tiles, lane_columns,
                       aux};
}

// count launches of the batched bulk-tile kernel (the lean loops) and of the batched edge-tile kernel (the general
// loops, one wavefront per tile: the many members keep the device busy), each step followed by the members' E_z
// walls, DFT regions and probes in the single-instance order, then the clocks.
int32_t launch_batch_2d(const Batch2DArgs& b, int32_t count, hipStream_t stream)
{
    auto bulk = b.lane_columns == 2 ? yee2d_batch_kernel<2, false, false> : yee2d_batch_kernel<4, false, false>;
    auto edge = b.lane_columns == 2
                    ? (b.aux.left ? yee2d_batch_kernel<2, true, true> : yee2d_batch_kernel<2, true, false>)
                    : (b.aux.left ? yee2d_batch_kernel<4, true, true> : yee2d_batch_kernel<4, true, false>);
    const int bulk_blocks = (int)((b.bulk_tiles + 3) / 4);
    const int edge_blocks = (int)((b.edge_tiles + 3) / 4);
    const BatchAux2D& a = b.aux;
    for (int32_t k = 0; k < count; ++k) {
        if (bulk_blocks > 0)
            hipLaunchKernelGGL(bulk, dim3(bulk_blocks), dim3(256), 0, stream, b.descs, b.sets, b.bulk_prefix, b.n, k);
        if (edge_blocks > 0)
            hipLaunchKernelGGL(edge, dim3(edge_blocks), dim3(256), 0, stream, b.descs, b.sets, b.edge_prefix, b.n, k);
        if (a.wall_blocks > 0)
            hipLaunchKernelGGL(yee2d_batch_wall_kernel, dim3(a.wall_blocks, b.n), dim3(256), 0, stream, b.descs, k);
        if (a.dft_blocks > 0)
            hipLaunchKernelGGL(yee2d_batch_dft_kernel, dim3(a.dft_blocks, a.dft_regions, b.n), dim3(256), 0, stream,
                               b.descs, k);
        if (a.probes) hipLaunchKernelGGL(yee2d_batch_probe_kernel, dim3(b.n), dim3(256), 0, stream, b.descs, k);
    }
    YEE_HIP(hipGetLastError());
    const int threads = 256;
    hipLaunchKernelGGL(yee_clock_advance_batch, dim3((b.n + threads - 1) / threads), dim3(threads), 0, stream,
                       b.descs, b.n, (int64_t)count);
    YEE_HIP(hipGetLastError());
    return YEE_OK;
}

// The second stream and fork / join events of a 2D context whose step has both tile sets (made once, before any
// capture). Default priority, non-blocking.
bool concurrent_sets(const yee_ctx* ctx) { return ctx->kind == kSingle2D && ctx->blocks > 0 && ctx->edge_blocks > 0; }

int32_t ensure_side(yee_ctx* ctx)
{
    if (!concurrent_sets(ctx) || ctx->side) return YEE_OK;
    YEE_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    YEE_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    YEE_HIP(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
    return YEE_OK;
}

// Launch `count` steps (offsets 0..count-1 from the device clock) followed by the clock advance.
int32_t launch_steps(yee_ctx* ctx, int32_t count, hipStream_t stream)
{
    if (count <= 0) return YEE_OK;
    if (ctx->kind == kSingle2D) {
        const StepKernel bulk = step_kernel(ctx->d2, false);
        const StepKernel edge = step_kernel(ctx->d2, true);
        const bool fork = concurrent_sets(ctx);
        if (fork) {
            int32_t status = ensure_side(ctx);
            if (status != YEE_OK) return status;
        }
        for (int32_t k = 0; k < count; ++k) {
            if (fork) {
                // the two tile sets are disjoint and both read state set n, write set n + 1: the edge tiles (few,
                // latency-bound) run beside the bulk ones; everything after the step waits for both
                YEE_HIP(hipEventRecord(ctx->fork_ev, stream));
                YEE_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
                hipLaunchKernelGGL(edge, dim3(ctx->edge_blocks), dim3(256), 0, ctx->side, ctx->d2, ctx->sets, k);
                hipLaunchKernelGGL(bulk, dim3(ctx->blocks), dim3(256), 0, stream, ctx->d2, ctx->sets, k);
                YEE_HIP(hipEventRecord(ctx->join_ev, ctx->side));
                YEE_HIP(hipStreamWaitEvent(stream, ctx->join_ev, 0));
            } else {
                if (ctx->blocks > 0)
                    hipLaunchKernelGGL(bulk, dim3(ctx->blocks), dim3(256), 0, stream, ctx->d2, ctx->sets, k);
                if (ctx->edge_blocks > 0)
                    hipLaunchKernelGGL(edge, dim3(ctx->edge_blocks), dim3(256), 0, stream, ctx->d2, ctx->sets, k);
            }
            if (has_ez_walls(ctx->d2))
                hipLaunchKernelGGL(yee2d_wall_kernel, dim3(wall_blocks(ctx->d2)), dim3(256), 0, stream, ctx->d2, k);
            if (ctx->d2.n_dft > 0)
                hipLaunchKernelGGL(yee2d_dft_kernel, dim3(dft_blocks(ctx->d2.dft, ctx->d2.n_dft), ctx->d2.n_dft),
                                   dim3(256), 0, stream, ctx->d2, k);
            if (ctx->d2.probes.n_probes > 0)
                hipLaunchKernelGGL(yee2d_probe_kernel, dim3(1), dim3(256), 0, stream, ctx->d2, k);
        }
        YEE_HIP(hipGetLastError());
        hipLaunchKernelGGL(yee_clock_advance, dim3(1), dim3(64), 0, stream, ctx->d2.clock, (int64_t)count,
                           ctx->d2.step_limit);
        YEE_HIP(hipGetLastError());
        return YEE_OK;
    }
    if (ctx->kind == kSingle3D) {
        const int sample_blocks = (int)((ctx->sample_cells + 255) / 256);
        for (int32_t k = 0; k < count; ++k) {
            hipLaunchKernelGGL(step_kernel_3d(ctx->d3), dim3(ctx->blocks), dim3(256), 0, stream, ctx->d3,
                               ctx->tiles_j3, ctx->tiles_k3, (int32_t)ctx->tiles, k);
            launch_mirrors_3d(ctx->d3, k, stream);
            if (sample_blocks > 0)
                hipLaunchKernelGGL(yee3d_sample_kernel, dim3(sample_blocks), dim3(256), 0, stream, ctx->d3,
                                   ctx->sample_cells, k);
            if (ctx->d3.n_dft > 0)
                hipLaunchKernelGGL(yee3d_dft_kernel, dim3(dft_blocks(ctx->d3.dft, ctx->d3.n_dft), ctx->d3.n_dft),
                                   dim3(256), 0, stream, ctx->d3, k);
            if (ctx->d3.probes.n_probes > 0)
                hipLaunchKernelGGL(yee3d_probe_kernel, dim3(1), dim3(256), 0, stream, ctx->d3, k);
        }
        YEE_HIP(hipGetLastError());
        hipLaunchKernelGGL(yee_clock_advance, dim3(1), dim3(64), 0, stream, ctx->d3.clock, (int64_t)count,
                           ctx->d3.step_limit);
        YEE_HIP(hipGetLastError());
        return YEE_OK;
    }
    if (ctx->kind == kBatch3D) {
        const int sample_blocks = (int)((ctx->sample_cells + 255) / 256);
        const BatchAux3D a = batch_aux_3d(ctx->descs3.data(), ctx->n_inst);
        for (int32_t k = 0; k < count; ++k) {
            hipLaunchKernelGGL(yee3d_batch_kernel<kRows3D>, dim3(ctx->blocks), dim3(256), 0, stream, ctx->descs3_dev,
                               ctx->block_prefix_dev, ctx->n_inst, k);
            // mirror faces as a single instance orders them: every symmetric face, then the antisymmetric ones
            for (int kind = 0; kind < 2; ++kind)
                for (int face = 0; face < 6; ++face)
                    if (a.mirror_blocks[kind][face] > 0)
                        hipLaunchKernelGGL(yee3d_batch_mirror_kernel, dim3(a.mirror_blocks[kind][face], ctx->n_inst),
                                           dim3(256), 0, stream, ctx->descs3_dev, face,
                                           kind == 0 ? YEE_MIRROR_SYMMETRIC : YEE_MIRROR_ANTISYMMETRIC, k);
            if (sample_blocks > 0)
                hipLaunchKernelGGL(yee3d_batch_sample_kernel, dim3(sample_blocks), dim3(256), 0, stream,
                                   ctx->descs3_dev, ctx->sample_prefix_dev, ctx->n_inst, k);
            if (a.dft_blocks > 0)
                hipLaunchKernelGGL(yee3d_batch_dft_kernel, dim3(a.dft_blocks, a.dft_regions, ctx->n_inst), dim3(256),
                                   0, stream, ctx->descs3_dev, k);
            if (a.probes)
                hipLaunchKernelGGL(yee3d_batch_probe_kernel, dim3(ctx->n_inst), dim3(256), 0, stream,
                                   ctx->descs3_dev, k);
        }
        YEE_HIP(hipGetLastError());
        const int threads = 256;
        hipLaunchKernelGGL(yee_clock_advance_batch3d, dim3((ctx->n_inst + threads - 1) / threads), dim3(threads), 0,
                           stream, ctx->descs3_dev, ctx->n_inst, (int64_t)count);
        YEE_HIP(hipGetLastError());
        return YEE_OK;
    }
    if (ctx->kind == kBatch2D)
        return launch_batch_2d(batch2d_args(ctx->descs_dev, ctx->n_inst, ctx->tiles, ctx->general_tiles,
                                            ctx->descs[0].lane_columns, batch_aux_2d(ctx->descs.data(), ctx->n_inst)),
                               count, stream);
    return fail(YEE_ESTATE, "context has no instances");
}

int32_t build_graph(yee_ctx* ctx)
{
    if (!ctx->capture_stream) YEE_HIP(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
    int32_t side_status = ensure_side(ctx);
    if (side_status != YEE_OK) return side_status;
    YEE_HIP(hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
    int32_t status = launch_steps(ctx, ctx->graph_steps, ctx->capture_stream);
    hipGraph_t graph = nullptr;
    hipError_t end = hipStreamEndCapture(ctx->capture_stream, &graph);
    if (status != YEE_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return status;
    }
    if (end != hipSuccess) return fail(YEE_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(end));
    hipError_t inst = hipGraphInstantiate(&ctx->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (inst != hipSuccess) {
        ctx->graph = nullptr;
        return fail(YEE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(inst));
    }
    return YEE_OK;
}

// One member k of a 2D batch (defaults already resolved): the single-instance checks and the batch's one restriction,
// a shared lane_columns (every member's probes, DFT regions, walls, TFSF boxes and poles are stepped as alone).
int32_t check_batch_member_2d(const yee2d_desc& d, int32_t lane_columns, int32_t k)
{
    int32_t status = check_2d(&d);
    if (status == YEE_OK && d.lane_columns != lane_columns)
        status = fail(YEE_EINVAL, "batched instances must share lane_columns");
    if (status != YEE_OK) return fail(status, "instance " + std::to_string(k) + ": " + g_last_error);
    return YEE_OK;
}

// The device image of a 2D batch (batch2d_layout(n).bytes bytes at `image`) from member descriptors whose defaults
// are resolved: checks every member, then writes the descriptors, their tile sets and the two tile prefixes.
int32_t stage_batch_2d(const yee2d_desc* descs, int32_t n, char* image, int64_t* bulk_tiles, int64_t* edge_tiles)
{
    const Batch2DLayout l = batch2d_layout(n);
    auto* sets = reinterpret_cast<yee2d_tile_sets*>(image + l.sets);
    auto* bulk = reinterpret_cast<int32_t*>(image + l.bulk_prefix);
    auto* edge = reinterpret_cast<int32_t*>(image + l.edge_prefix);
    bulk[0] = edge[0] = 0;
    for (int32_t k = 0; k < n; ++k) {
        int32_t status = check_batch_member_2d(descs[k], descs[0].lane_columns, k);
        if (status != YEE_OK) return status;
        std::memcpy(image + sizeof(yee2d_desc) * (size_t)k, &descs[k], sizeof(yee2d_desc));
        sets[k] = tile_sets(descs[k], true);
        const yee2d_tile_sets& t = sets[k];
        const int64_t all = (int64_t)t.tiles_i * t.tiles_j;
        const int64_t b = t.all_edge ? 0 : (int64_t)(t.tiles_i - t.n_rows) * std::max(0, t.tiles_j - t.n_cols);
        const int64_t next_bulk = (int64_t)bulk[k] + b, next_edge = (int64_t)edge[k] + (all - b);
        if (next_bulk > INT32_MAX / 4 || next_edge > INT32_MAX / 4)
            return fail(YEE_EINVAL, "batch has too many tiles");
        bulk[k + 1] = (int32_t)next_bulk;
        edge[k + 1] = (int32_t)next_edge;
    }
    *bulk_tiles = bulk[n];
    *edge_tiles = edge[n];
    return YEE_OK;
}

// ---- the no-allocation batch entry point -------------------------------------------------------------------------
// Descriptors, tile sets and tile prefixes of yee2d_batch_step_static live in this file-scope device array, laid out as
// yee2d_batch_bytes(n) lays out a caller's buffer. Every GPU holds its own copy until the process ends. In host code
// `scratch` names a host-side stand-in, not a device address: the address comes from hipGetSymbolAddress, looked up
// once per device (with that device current) on the first call there, and only that address is used.
constexpr size_t kStaticBatchBytes = batch2d_layout(YEE_STATIC_BATCH_MAX).bytes;
static __device__ __align__(256) unsigned char scratch[kStaticBatchBytes];
// the host image the static entry point copies from (no heap allocation; used under g_scratch_mutex)
alignas(256) char g_static_image[kStaticBatchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

// The device address of `scratch` on the current device.
int32_t scratch_address(void** out)
{
    int device = -1;
    YEE_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) return fail(YEE_EINVAL, "device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void* ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || !ptr)
            return fail(YEE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        g_scratch_address[device] = ptr;
    }
    *out = g_scratch_address[device];
    return YEE_OK;
}

}  // namespace

extern "C" {

int32_t yee_abi_version(void) { return YEE_ABI_VERSION; }

size_t yee_sizeof(int32_t which)
{
    switch (which) {
        case 0: return sizeof(yee2d_desc);
        case 1: return sizeof(yee_line_source);
        case 2: return sizeof(yee_flux_line);
        case 3: return sizeof(yee_cpml_axis);
        case 4: return sizeof(yee3d_desc);
        case 5: return sizeof(yee_point_source);
        case 6: return sizeof(yee_flux_plane);
        case 7: return sizeof(yee3d_cpml_axis);
        case 8: return sizeof(yee3d_axis_row);
        case 9: return sizeof(yee_probes);
        case 10: return sizeof(yee_dft_region);
        case 11: return sizeof(yee_pole);
        case 12: return sizeof(yee3d_pole);
        case 13: return sizeof(yee_tfsf);
        default: return 0;
    }
}

const char* yee_last_error(void) { return g_last_error.c_str(); }

int32_t yee2d_create(const yee2d_desc* desc, yee_ctx** out)
{
    if (!out) return fail(YEE_EINVAL, "null output pointer");
    *out = nullptr;
    yee2d_desc d = desc ? *desc : yee2d_desc{};
    if (desc) apply_defaults(d);
    int32_t status = check_2d(desc ? &d : nullptr);
    if (status != YEE_OK) return status;
    yee_ctx* ctx = new yee_ctx();
    ctx->kind = kSingle2D;
    ctx->d2 = d;
    set_geometry(ctx);
    *out = ctx;
    return YEE_OK;
}

int32_t yee2d_update(yee_ctx* ctx, const yee2d_desc* desc)
{
    if (!ctx || ctx->kind != kSingle2D) return fail(YEE_ESTATE, "not a single 2D context");
    yee2d_desc d = desc ? *desc : yee2d_desc{};
    if (desc) apply_defaults(d);
    int32_t status = check_2d(desc ? &d : nullptr);
    if (status != YEE_OK) return status;
    drop_graph(ctx);
    ctx->d2 = d;
    set_geometry(ctx);
    return YEE_OK;
}

int32_t yee2d_batch_create(const yee2d_desc* descs, int32_t n, void* descs_dev, void* stream, yee_ctx** out)
{
    if (!out) return fail(YEE_EINVAL, "null output pointer");
    *out = nullptr;
    if (!descs || n < 1 || !descs_dev) return fail(YEE_EINVAL, "batch needs descriptors and a device buffer");
    std::vector<yee2d_desc> host(descs, descs + n);
    for (int32_t k = 0; k < n; ++k) {
        if (host[k].lane_columns == 0) host[k].lane_columns = host[0].lane_columns ? host[0].lane_columns : 2;
        if (host[k].rows_per_tile == 0) host[k].rows_per_tile = 32;
    }
    const Batch2DLayout l = batch2d_layout(n);
    std::vector<char> image(l.bytes);
    int64_t bulk_tiles = 0, edge_tiles = 0;
    int32_t status = stage_batch_2d(host.data(), n, image.data(), &bulk_tiles, &edge_tiles);
    if (status != YEE_OK) return status;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    YEE_HIP(hipMemcpyAsync(descs_dev, image.data(), l.bytes, hipMemcpyHostToDevice, s));
    YEE_HIP(hipStreamSynchronize(s));   // the host image is the copy's source
    yee_ctx* ctx = new yee_ctx();
    ctx->kind = kBatch2D;
    ctx->descs = std::move(host);
    ctx->descs_dev = reinterpret_cast<yee2d_desc*>(descs_dev);
    ctx->n_inst = n;
    ctx->tiles = bulk_tiles;
    ctx->general_tiles = edge_tiles;
    ctx->blocks = (int)((bulk_tiles + 3) / 4);
    *out = ctx;
    return YEE_OK;
}

size_t yee2d_batch_bytes(int32_t n) { return n < 1 ? 0 : batch2d_layout(n).bytes; }

int32_t yee2d_batch_step_static(const yee2d_desc* descs, int32_t n, int64_t n_steps, void* stream)
{
    if (!descs || n < 1) return fail(YEE_EINVAL, "batch needs descriptors");
    if (n > YEE_STATIC_BATCH_MAX)
        return fail(YEE_EINVAL, "yee2d_batch_step_static takes at most " + std::to_string(YEE_STATIC_BATCH_MAX) +
                                    " instances (its static device array), got " + std::to_string(n));
    if (n_steps < 0) return fail(YEE_EINVAL, "n_steps must be >= 0");
    for (int32_t k = 0; k < n; ++k)
        if (descs[k].lane_columns == 0 || descs[k].rows_per_tile == 0)
            return fail(YEE_EINVAL, "instance " + std::to_string(k) +
                                        ": yee2d_batch_step_static copies descriptors as given and needs lane_columns "
                                        "and rows_per_tile set");
    if (n_steps == 0) return YEE_OK;
    void* base = nullptr;
    int32_t status = scratch_address(&base);
    if (status != YEE_OK) return status;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int64_t bulk_tiles = 0, edge_tiles = 0;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        status = stage_batch_2d(descs, n, g_static_image, &bulk_tiles, &edge_tiles);
        if (status != YEE_OK) return status;
        // stream-ordered after this stream's earlier work (an earlier call's launches read the same array); the
        // host image is reused by the next call, so the copy completes before the lock is released
        YEE_HIP(hipMemcpyAsync(base, g_static_image, batch2d_layout(n).bytes, hipMemcpyHostToDevice, s));
        YEE_HIP(hipStreamSynchronize(s));
    }
    const Batch2DArgs args = batch2d_args(base, n, bulk_tiles, edge_tiles, descs[0].lane_columns,
                                          batch_aux_2d(descs, n));
    for (int64_t remaining = n_steps; remaining > 0;) {
        const int32_t chunk = remaining > 4096 ? 4096 : (int32_t)remaining;
        status = launch_batch_2d(args, chunk, s);
        if (status != YEE_OK) return status;
        remaining -= chunk;
    }
    return YEE_OK;
}

int32_t yee_flag_nonfinite(const float* data, int64_t count, int32_t* flag, void* stream)
{
    if (count < 0 || (count > 0 && (!data || !flag)))
        return fail(YEE_EINVAL, "yee_flag_nonfinite needs data, flag and count >= 0");
    if (count == 0) return YEE_OK;
    const int64_t blocks = std::min<int64_t>(2048, (count + 255) / 256);
    hipLaunchKernelGGL(yee_nonfinite_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), data, count, flag);
    YEE_HIP(hipGetLastError());
    return YEE_OK;
}

int32_t yee2d_get_desc(const yee_ctx* ctx, yee2d_desc* out)
{
    if (!ctx || ctx->kind != kSingle2D || !out) return fail(YEE_ESTATE, "not a single 2D context");
    *out = ctx->d2;
    return YEE_OK;
}

int32_t yee_set_graph_steps(yee_ctx* ctx, int32_t graph_steps)
{
    if (!ctx) return fail(YEE_ESTATE, "null context");
    if (graph_steps < 0) return fail(YEE_EINVAL, "graph_steps must be >= 0");
    drop_graph(ctx);
    ctx->graph_steps = graph_steps;
    return YEE_OK;
}

int32_t yee_build_graph(yee_ctx* ctx)
{
    if (!ctx || !ctx->kind) return fail(YEE_ESTATE, "null or destroyed context");
    if (ctx->graph_steps <= 0 || ctx->graph) return YEE_OK;
    return build_graph(ctx);
}

// Cells in the bulk tile set: the non-edge tile rows times the non-edge tile columns (the set is their product).
int64_t bulk_cells(const yee2d_desc& d, const yee2d_tile_sets& s)
{
    if (s.all_edge) return 0;
    const int cols = 64 * d.lane_columns;
    int64_t rows = 0, columns = 0;
    for (int ti = 0, q = 0; ti < s.tiles_i; ++ti) {
        if (q < s.n_rows && s.rows[q] == ti) { ++q; continue; }
        rows += std::min(d.nx, (ti + 1) * d.rows_per_tile) - ti * d.rows_per_tile;
    }
    for (int tj = 0, q = 0; tj < s.tiles_j; ++tj) {
        if (q < s.n_cols && s.cols[q] == tj) { ++q; continue; }
        columns += std::min(d.ny, (tj + 1) * cols) - tj * cols;
    }
    return rows * columns;
}

// yee_step_timed of a 3D context: class 0 = the fused step kernel (every cell), 1 = the plane-sample kernel.
int32_t step_timed_3d(yee_ctx* ctx, int64_t n_steps, hipStream_t s, double* kernel_ms, int64_t* cells)
{
    const yee3d_desc& d = ctx->d3;
    cells[0] = (int64_t)d.nx * d.ny * d.nz;
    cells[1] = ctx->sample_cells;
    kernel_ms[0] = kernel_ms[1] = 0.0;
    const int sample_blocks = (int)((ctx->sample_cells + 255) / 256);
    hipEvent_t ev[3];
    for (auto& e : ev) YEE_HIP(hipEventCreate(&e));
    int32_t status = YEE_OK;
    for (int64_t k = 0; k < n_steps && status == YEE_OK; ++k) {
        (void)hipEventRecord(ev[0], s);
        hipLaunchKernelGGL(step_kernel_3d(ctx->d3), dim3(ctx->blocks), dim3(256), 0, s, d, ctx->tiles_j3,
                           ctx->tiles_k3, (int32_t)ctx->tiles, 0);
        (void)hipEventRecord(ev[1], s);
        launch_mirrors_3d(d, 0, s);
        if (sample_blocks > 0)
            hipLaunchKernelGGL(yee3d_sample_kernel, dim3(sample_blocks), dim3(256), 0, s, d, ctx->sample_cells, 0);
        (void)hipEventRecord(ev[2], s);
        if (d.n_dft > 0)
            hipLaunchKernelGGL(yee3d_dft_kernel, dim3(dft_blocks(d.dft, d.n_dft), d.n_dft), dim3(256), 0, s, d, 0);
        if (d.probes.n_probes > 0) hipLaunchKernelGGL(yee3d_probe_kernel, dim3(1), dim3(256), 0, s, d, 0);
        hipLaunchKernelGGL(yee_clock_advance, dim3(1), dim3(64), 0, s, d.clock, (int64_t)1, d.step_limit);
        hipError_t e = hipEventSynchronize(ev[2]);
        float a = 0.f, b = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&a, ev[0], ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[1], ev[2]);
        if (e != hipSuccess) {
            status = fail(YEE_EHIP, hipGetErrorString(e));
            break;
        }
        kernel_ms[0] += a;
        kernel_ms[1] += b;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (status == YEE_OK) YEE_HIP(hipStreamSynchronize(s));
    return status;
}

int32_t yee_step_timed(yee_ctx* ctx, int64_t n_steps, void* stream, double* kernel_ms, int64_t* cells)
{
    if (!ctx || !ctx->kind) return fail(YEE_ESTATE, "null or destroyed context");
    if (ctx->kind != kSingle2D && ctx->kind != kSingle3D)
        return fail(YEE_EINVAL, "yee_step_timed times a single-instance context");
    if (n_steps < 0 || !kernel_ms || !cells) return fail(YEE_EINVAL, "n_steps must be >= 0 and outputs non-null");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (ctx->kind == kSingle3D) return step_timed_3d(ctx, n_steps, s, kernel_ms, cells);
    const StepKernel bulk = step_kernel(ctx->d2, false);
    const StepKernel edge = step_kernel(ctx->d2, true);
    // tile sets launched side by side (as yee_step launches them): one class, every cell, from the fork to the join
    const bool fork = concurrent_sets(ctx);
    if (fork) {
        int32_t status = ensure_side(ctx);
        if (status != YEE_OK) return status;
    }
    cells[0] = fork ? (int64_t)ctx->d2.nx * ctx->d2.ny : bulk_cells(ctx->d2, ctx->sets);
    cells[1] = fork ? 0 : (int64_t)ctx->d2.nx * ctx->d2.ny - cells[0];
    kernel_ms[0] = kernel_ms[1] = 0.0;
    hipEvent_t ev[3];
    for (auto& e : ev) YEE_HIP(hipEventCreate(&e));
    int32_t status = YEE_OK;
    for (int64_t k = 0; k < n_steps && status == YEE_OK; ++k) {
        (void)hipEventRecord(ev[0], s);
        if (fork) {
            (void)hipEventRecord(ctx->fork_ev, s);
            (void)hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0);
            hipLaunchKernelGGL(edge, dim3(ctx->edge_blocks), dim3(256), 0, ctx->side, ctx->d2, ctx->sets, 0);
            hipLaunchKernelGGL(bulk, dim3(ctx->blocks), dim3(256), 0, s, ctx->d2, ctx->sets, 0);
            (void)hipEventRecord(ctx->join_ev, ctx->side);
            (void)hipStreamWaitEvent(s, ctx->join_ev, 0);
            (void)hipEventRecord(ev[1], s);
        } else {
            if (ctx->blocks > 0)
                hipLaunchKernelGGL(bulk, dim3(ctx->blocks), dim3(256), 0, s, ctx->d2, ctx->sets, 0);
            (void)hipEventRecord(ev[1], s);
            if (ctx->edge_blocks > 0)
                hipLaunchKernelGGL(edge, dim3(ctx->edge_blocks), dim3(256), 0, s, ctx->d2, ctx->sets, 0);
        }
        (void)hipEventRecord(ev[2], s);
        if (has_ez_walls(ctx->d2))
            hipLaunchKernelGGL(yee2d_wall_kernel, dim3(wall_blocks(ctx->d2)), dim3(256), 0, s, ctx->d2, 0);
        if (ctx->d2.n_dft > 0)
            hipLaunchKernelGGL(yee2d_dft_kernel, dim3(dft_blocks(ctx->d2.dft, ctx->d2.n_dft), ctx->d2.n_dft),
                               dim3(256), 0, s, ctx->d2, 0);
        if (ctx->d2.probes.n_probes > 0)
            hipLaunchKernelGGL(yee2d_probe_kernel, dim3(1), dim3(256), 0, s, ctx->d2, 0);
        hipLaunchKernelGGL(yee_clock_advance, dim3(1), dim3(64), 0, s, ctx->d2.clock, (int64_t)1,
                           ctx->d2.step_limit);
        hipError_t e = hipEventSynchronize(ev[2]);
        float a = 0.f, b = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&a, ev[0], ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[1], ev[2]);
        if (e != hipSuccess) {
            status = fail(YEE_EHIP, hipGetErrorString(e));
            break;
        }
        kernel_ms[0] += a;
        if (!fork) kernel_ms[1] += b;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (status == YEE_OK) YEE_HIP(hipStreamSynchronize(s));
    return status;
}

int32_t yee_step(yee_ctx* ctx, int64_t n_steps, void* stream)
{
    if (!ctx || !ctx->kind) return fail(YEE_ESTATE, "null or destroyed context");
    if (n_steps < 0) return fail(YEE_EINVAL, "n_steps must be >= 0");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int64_t remaining = n_steps;
    if (ctx->graph_steps > 0 && remaining >= ctx->graph_steps) {
        if (!ctx->graph) {
            int32_t status = build_graph(ctx);
            if (status != YEE_OK) return status;
        }
        while (remaining >= ctx->graph_steps) {
            YEE_HIP(hipGraphLaunch(ctx->graph, s));
            remaining -= ctx->graph_steps;
        }
    }
    while (remaining > 0) {
        const int32_t chunk = remaining > 4096 ? 4096 : (int32_t)remaining;
        int32_t status = launch_steps(ctx, chunk, s);
        if (status != YEE_OK) return status;
        remaining -= chunk;
    }
    return YEE_OK;
}

int32_t yee3d_create(const yee3d_desc* desc, yee_ctx** out)
{
    if (!out) return fail(YEE_EINVAL, "null output pointer");
    *out = nullptr;
    yee3d_desc d = desc ? *desc : yee3d_desc{};
    if (desc) apply_defaults_3d(d);
    int32_t status = check_3d(desc ? &d : nullptr);
    if (status != YEE_OK) return status;
    yee_ctx* ctx = new yee_ctx();
    ctx->kind = kSingle3D;
    ctx->d3 = d;
    set_geometry_3d(ctx);
    *out = ctx;
    return YEE_OK;
}

int32_t yee3d_update(yee_ctx* ctx, const yee3d_desc* desc)
{
    if (!ctx || ctx->kind != kSingle3D) return fail(YEE_ESTATE, "not a 3D context");
    yee3d_desc d = desc ? *desc : yee3d_desc{};
    if (desc) apply_defaults_3d(d);
    int32_t status = check_3d(desc ? &d : nullptr);
    if (status != YEE_OK) return status;
    drop_graph(ctx);
    ctx->d3 = d;
    set_geometry_3d(ctx);
    return YEE_OK;
}

int32_t yee3d_batch_create(const yee3d_desc* descs, int32_t n, void* descs_dev, void* stream, yee_ctx** out)
{
    if (!out) return fail(YEE_EINVAL, "null output pointer");
    *out = nullptr;
    if (!descs || n < 1 || !descs_dev) return fail(YEE_EINVAL, "batch needs descriptors and a device buffer");
    std::vector<yee3d_desc> host(descs, descs + n);
    std::vector<int64_t> sample_prefix(n + 1, 0);
    std::vector<int32_t> block_prefix(n + 1, 0);
    for (int32_t k = 0; k < n; ++k) {
        apply_defaults_3d(host[k]);
        const int32_t status = check_3d(&host[k]);
        if (status != YEE_OK) return fail(status, "instance " + std::to_string(k) + ": " + g_last_error);
        if (host[k].rows_per_tile != kRows3D) return fail(YEE_EINVAL, "batched 3D instances use 4 rows per tile");
        const int64_t next = block_prefix[k] + (tiles_3d(host[k]) + 3) / 4;
        if (next > INT32_MAX / 2) return fail(YEE_EINVAL, "batch has too many tiles");
        block_prefix[k + 1] = (int32_t)next;
        sample_prefix[k + 1] = sample_prefix[k] + sample_cells_3d(host[k]);
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char* base = reinterpret_cast<char*>(descs_dev);
    auto* dev_descs = reinterpret_cast<yee3d_desc*>(base);
    auto* dev_samples = reinterpret_cast<int64_t*>(base + sizeof(yee3d_desc) * n);
    auto* dev_blocks = reinterpret_cast<int32_t*>(base + sizeof(yee3d_desc) * n + sizeof(int64_t) * (n + 1));
    YEE_HIP(hipMemcpyAsync(dev_descs, host.data(), sizeof(yee3d_desc) * n, hipMemcpyHostToDevice, s));
    YEE_HIP(hipMemcpyAsync(dev_samples, sample_prefix.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s));
    YEE_HIP(hipMemcpyAsync(dev_blocks, block_prefix.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, s));
    YEE_HIP(hipStreamSynchronize(s));   // the host vectors are the copies' sources
    yee_ctx* ctx = new yee_ctx();
    ctx->kind = kBatch3D;
    ctx->descs3 = std::move(host);
    ctx->descs3_dev = dev_descs;
    ctx->sample_prefix_dev = dev_samples;
    ctx->block_prefix_dev = dev_blocks;
    ctx->n_inst = n;
    ctx->blocks = block_prefix[n];
    ctx->tiles = (int64_t)block_prefix[n] * 4;
    ctx->sample_cells = sample_prefix[n];
    *out = ctx;
    return YEE_OK;
}

size_t yee3d_batch_bytes(int32_t n)
{
    return n < 1 ? 0 : sizeof(yee3d_desc) * (size_t)n + (sizeof(int64_t) + sizeof(int32_t)) * (size_t)(n + 1);
}

int32_t yee3d_get_desc(const yee_ctx* ctx, yee3d_desc* out)
{
    if (!ctx || ctx->kind != kSingle3D || !out) return fail(YEE_ESTATE, "not a 3D context");
    *out = ctx->d3;
    return YEE_OK;
}

int32_t yee_destroy(yee_ctx* ctx)
{
    if (!ctx) return YEE_OK;
    drop_graph(ctx);
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    ctx->kind = 0;
    delete ctx;
    return YEE_OK;
}

}  // extern "C"
