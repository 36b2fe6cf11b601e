// Fused 2D Yee step for gfx950: one launch advances E_z (E_z, H_x, H_y) or H_z (H_z, E_x, E_y) instances by one
// full step (H update, H CPML, E update, E CPML, sheet-source injection, FP64 Kahan DFT flux samples).
//
// Work decomposition. A wavefront tile owns OWN = 63 consecutive y columns (one lane per column, y is contiguous in
// HBM so every row access is one coalesced 256 B request) and marches rows_per_tile rows along x, carrying the
// previous row in registers (register blocking along the slow axis). The 64th lane is a halo column: E_z tiles place
// it on the left (H_x of column j0-1 feeds the E update of column j0), H_z tiles on the right (H_z of column j0+63
// feeds E_x of column j0+62). Halo lanes run the same code as owned lanes but store nothing, so cross-lane neighbours
// come from wave shuffles and no LDS or barrier is needed. Neighbouring tiles read each other's halo rows/columns,
// which is race-free because every state array (fields and CPML psi) is double-buffered: step n reads set n & 1 and
// writes set (n & 1) ^ 1.
//
// HBM traffic per cell-update: read E, H_a, H_b (12 B) + uint16 material id (2 B), write 12 B = 26 B; the per-id
// coefficient table stays in L1/L2. Tile-boundary rows/columns are re-read (1/rows_per_tile and 1/63 extra) from L2.
//
// Reference semantics (file:line of the reference package open_sol/):
//   E_z H update       physics/em/solvers/fdtd_ez_2d_kernels.py:23-60   (update_h_kernel / _update_h_host)
//   E_z E update       physics/em/solvers/fdtd_ez_2d_kernels.py:88-140  (update_e_kernel / _update_e_host)
//   H_z H update       physics/em/solvers/fdtd_hz_2d_kernels.py:28-80   (update_hz_kernel)
//   H_z E update       physics/em/solvers/fdtd_hz_2d_kernels.py:115-155 (update_exy_kernel)
//   face coefficients  physics/em/solvers/fdtd_hz_2d_kernels.py:190-278 (e_face_coefficients, shares 1/2)
//   CPML               physics/em/boundaries.py:27-37, 689-772          (psi recursion, signs ez x+/y-, hz x-/y+)
//   sources            core/sources.py:146-155, 221-229; physics/em/sources.py:469-480
//   DFT                physics/em/monitors.py:22-34, 86-90, 203-224     (complex128 spectra; here Kahan FP64)
//   step order         physics/em/solvers/fdtd_ez_2d.py:71-119, fdtd_hz_2d.py:144-189, fdtd_2d_base.py:470-488
#include "yee_device.h"

namespace yee {

constexpr int OWN = YEE_WAVE - 1;   // owned columns per wavefront tile

__host__ __device__ inline int tiles_j_2d(int ny) { return (ny + OWN - 1) / OWN; }
__host__ __device__ inline int tiles_i_2d(int nx, int rows) { return (nx + rows - 1) / rows; }

// ------------------------------------------------------------------------------------------------------------ E_z
__device__ __forceinline__ void ez_tile(const yee2d_desc& d, int tj, int ti, int64_t n)
{
    const int lane = threadIdx.x & (YEE_WAVE - 1);
    const int nx = d.nx, ny = d.ny;
    const int i0 = ti * d.rows_per_tile;
    const int i1 = min(nx, i0 + d.rows_per_tile);
    const bool px = d.periodic_x != 0, py = d.periodic_y != 0;
    const int par = (int)(n & 1);

    // lane -> column; lane 0 is the halo column j0 - 1 (wrapped on a periodic axis, absent otherwise)
    const int col = tj * OWN - 1 + lane;
    const bool owner = lane > 0 && col < ny;
    int j = col;
    if (col < 0) j = py ? ny - 1 : -1;
    const bool valid = j >= 0 && j < ny;
    const int jc = valid ? j : 0;

    const float* __restrict__ ez_o = d.field[par][0];
    const float* __restrict__ hx_o = d.field[par][1];
    const float* __restrict__ hy_o = d.field[par][2];
    float* __restrict__ ez_n = d.field[par ^ 1][0];
    float* __restrict__ hx_n = d.field[par ^ 1][1];
    float* __restrict__ hy_n = d.field[par ^ 1][2];
    const float4* __restrict__ tab = reinterpret_cast<const float4*>(d.table);
    const uint16_t* __restrict__ mat = d.material;

    // column constants: H_x is updated on columns < ny-1 (every column when periodic), its hi neighbour wraps to 0
    const bool hx_updated = valid && (j < ny - 1 || py);
    const int jr = (j + 1 < ny) ? j + 1 : 0;
    const bool direct_right = hx_updated && (lane == YEE_WAVE - 1 || j == ny - 1);
    const float idy_h = d.inv_dy_h[jc];
    const float idy_e = d.inv_dy_e[jc];

    const yee_cpml_axis& cx = d.cpml[0];
    const yee_cpml_axis& cy = d.cpml[1];
    const int nry = cy.n_rows;
    int ry_h = -1, ry_e = -1;
    float by_h = 0.f, gy_h = 0.f, by_e = 0.f, gy_e = 0.f;
    if (nry > 0 && valid) {
        ry_h = cy.row_h[jc];
        ry_e = cy.row_e[jc];
        if (ry_h >= 0) { by_h = cy.b_h[ry_h]; gy_h = cy.g_h[ry_h]; }
        if (ry_e >= 0) { by_e = cy.b_e[ry_e]; gy_e = cy.g_e[ry_e]; }
    }
    const float* __restrict__ psi_hy_o = cy.psi_h[par];
    float* __restrict__ psi_hy_n = cy.psi_h[par ^ 1];
    const float* __restrict__ psi_ey_o = cy.psi_e[par];
    float* __restrict__ psi_ey_n = cy.psi_e[par ^ 1];

    const bool sample = d.n_monitors > 0 && ((n + 1) % d.sample_interval) == 0;
    const int64_t sample_index = (n + 1) / (d.sample_interval > 0 ? d.sample_interval : 1) - 1;

    // prologue: E of row i0, H_y of row i0-1 after this step (0 below a non-periodic grid edge)
    float e_cur = valid ? ez_o[(int64_t)i0 * ny + jc] : 0.f;
    float hy_prev = 0.f;
    if (i0 > 0 || px) {
        const int ip = i0 > 0 ? i0 - 1 : nx - 1;
        const int64_t p0 = (int64_t)ip * ny + jc;
        const float e_p = valid ? ez_o[p0] : 0.f;
        const float4 c = tab[mat[p0]];
        const float de = e_cur - e_p;
        float v = c.z * hy_o[p0] + c.w * de * d.inv_dx_h[ip];
        const int rx = cx.n_rows > 0 ? cx.row_h[ip] : -1;
        if (rx >= 0) {
            const float psi = cx.b_h[rx] * cx.psi_h[par][(int64_t)rx * ny + jc] + cx.g_h[rx] * de;
            v += c.w * psi;
        }
        hy_prev = valid ? v : 0.f;
    }

    for (int i = i0; i < i1; ++i) {
        const int64_t r0 = (int64_t)i * ny;
        const bool hy_updated = (i < nx - 1) || px;
        const int inext = (i + 1 < nx) ? i + 1 : 0;
        const float e_next = (valid && hy_updated) ? ez_o[(int64_t)inext * ny + jc] : 0.f;
        const float hx_old = hx_o[r0 + jc];
        const float hy_old = hy_o[r0 + jc];
        const float4 c = tab[mat[r0 + jc]];   // (ca, cb, da, db)

        // E of the hi-y neighbour: the next lane, or a direct read at the tile edge / periodic wrap
        float e_right = shfl_from_next(e_cur);
        if (direct_right) e_right = ez_o[r0 + jr];

        // H_y on the hi-x face (+ x CPML, sign +)
        float hy_v = hy_old;
        if (hy_updated) {
            const float de = e_next - e_cur;
            hy_v = c.z * hy_old + c.w * de * d.inv_dx_h[i];
            const int rx = cx.n_rows > 0 ? cx.row_h[i] : -1;
            if (rx >= 0) {
                const int64_t p = (int64_t)rx * ny + jc;
                const float psi = cx.b_h[rx] * cx.psi_h[par][p] + cx.g_h[rx] * de;
                if (owner) cx.psi_h[par ^ 1][p] = psi;
                hy_v += c.w * psi;
            }
        }
        // H_x on the hi-y face (+ y CPML, sign -)
        float hx_v = valid ? hx_old : 0.f;
        if (hx_updated) {
            const float de = e_right - e_cur;
            hx_v = c.z * hx_old - c.w * de * idy_h;
            if (ry_h >= 0) {
                const int64_t p = (int64_t)i * nry + ry_h;
                const float psi = by_h * psi_hy_o[p] + gy_h * de;
                if (owner) psi_hy_n[p] = psi;
                hx_v -= c.w * psi;
            }
        }
        // H_x of the lo-y neighbour after this step (lane 0 of a periodic/interior tile holds column j0-1)
        const float hx_left = shfl_from_prev(hx_v);

        // E_z at the centre (+ x CPML sign +, y CPML sign -)
        const float dhy = hy_v - hy_prev;
        const float dhx = hx_v - hx_left;
        float ez_v = c.x * e_cur + c.y * (dhy * d.inv_dx_e[i] - dhx * idy_e);
        const int rxe = cx.n_rows > 0 ? cx.row_e[i] : -1;
        if (rxe >= 0) {
            const int64_t p = (int64_t)rxe * ny + jc;
            const float psi = cx.b_e[rxe] * cx.psi_e[par][p] + cx.g_e[rxe] * dhy;
            if (owner) cx.psi_e[par ^ 1][p] = psi;
            ez_v += c.y * psi;
        }
        if (ry_e >= 0) {
            const int64_t p = (int64_t)i * nry + ry_e;
            const float psi = by_e * psi_ey_o[p] + gy_e * dhx;
            if (owner) psi_ey_n[p] = psi;
            ez_v -= c.y * psi;
        }
        // sheet sources at t + dt/2: E_z += float32(value * profile)
        for (int s = 0; s < d.n_sources; ++s) {
            const yee_line_source& src = d.source[s];
            const bool on_line = src.axis == 1 ? (i == src.line && j >= src.lo && j < src.hi)
                                               : (j == src.line && i >= src.lo && i < src.hi);
            if (owner && on_line) {
                const int k = src.axis == 1 ? j - src.lo : i - src.lo;
                ez_v += __fmul_rn(wave_value(src, n), src.profile[k]);
            }
        }
        if (owner) {
            ez_n[r0 + j] = ez_v;
            hx_n[r0 + j] = hx_v;
            hy_n[r0 + j] = hy_v;
        }
        if (sample) {
            for (int m = 0; m < d.n_monitors; ++m) {
                const yee_flux_line& mon = d.monitor[m];
                if (mon.normal == 0) {
                    if (owner && i == mon.position && j >= mon.lo && j < mon.hi)
                        dft_accumulate(mon, sample_index, j - mon.lo, ez_v, hy_v);
                } else if (owner && j == mon.position && i >= mon.lo && i < mon.hi) {
                    dft_accumulate(mon, sample_index, i - mon.lo, ez_v, hx_v);
                }
            }
        }
        hy_prev = hy_v;
        e_cur = e_next;
    }
}

// ------------------------------------------------------------------------------------------------------------ H_z
// E face coefficients of a face between cells of ids a and b: the per-id table holds the coefficients of a face
// between two cells of that id (computed in FP64 on the host); a mixed face takes the mean eps / sigma (shares 1/2,
// fdtd_hz_2d_kernels.py:190-278) and forms ca, cb here in float32.
__device__ __forceinline__ void face_coefficients(const yee2d_desc& d, int ida, int idb, int comp, float* ca,
                                                  float* cb)
{
    const float* __restrict__ t2 = d.table2;
    if (ida == idb) {
        *ca = t2[8 * ida + 2 * comp];
        *cb = t2[8 * ida + 2 * comp + 1];
        return;
    }
    const float4* __restrict__ tab = reinterpret_cast<const float4*>(d.table);
    const float ea = comp == 0 ? tab[ida].z : tab[ida].w;
    const float eb = comp == 0 ? tab[idb].z : tab[idb].w;
    const float sa = t2[8 * ida + 4], sb = t2[8 * idb + 4];
    const float eps_face = 0.5f * ea + 0.5f * eb;
    const float sig_face = 0.5f * sa + 0.5f * sb;
    const float k = sig_face * (0.5f * d.dt_over_eps0) / eps_face;
    const float den = 1.0f + k;
    *ca = (1.0f - k) / den;
    *cb = (d.dt_over_eps0 / eps_face) / den;
}

__device__ __forceinline__ void hz_tile(const yee2d_desc& d, int tj, int ti, int64_t n)
{
    const int lane = threadIdx.x & (YEE_WAVE - 1);
    const int nx = d.nx, ny = d.ny;
    const int i0 = ti * d.rows_per_tile;
    const int i1 = min(nx, i0 + d.rows_per_tile);
    const bool px = d.periodic_x != 0, py = d.periodic_y != 0;
    const int par = (int)(n & 1);

    // lane -> column; lane 63 is the halo column j0 + 63 (column 0 after the last one on a periodic axis)
    const int col = tj * OWN + lane;
    const bool owner = lane < YEE_WAVE - 1 && col < ny;
    int j = col;
    if (col >= ny) j = (py && col == ny) ? 0 : -1;
    const bool valid = j >= 0;
    const int jc = valid ? j : 0;

    const float* __restrict__ hz_o = d.field[par][0];
    const float* __restrict__ ex_o = d.field[par][1];
    const float* __restrict__ ey_o = d.field[par][2];
    float* __restrict__ hz_n = d.field[par ^ 1][0];
    float* __restrict__ ex_n = d.field[par ^ 1][1];
    float* __restrict__ ey_n = d.field[par ^ 1][2];
    const float4* __restrict__ tab = reinterpret_cast<const float4*>(d.table);
    const uint16_t* __restrict__ mat = d.material;

    const bool ex_updated = valid && (j < ny - 1 || py);
    const int jl = j > 0 ? j - 1 : (py ? ny - 1 : -1);
    const bool direct_left = valid && lane == 0 && jl >= 0;
    const float idy_h = d.inv_dy_h[jc];
    const float idy_e = d.inv_dy_e[jc];

    const yee_cpml_axis& cx = d.cpml[0];
    const yee_cpml_axis& cy = d.cpml[1];
    const int nry = cy.n_rows;
    int ry_h = -1, ry_e = -1;
    float by_h = 0.f, gy_h = 0.f, by_e = 0.f, gy_e = 0.f;
    if (nry > 0 && valid) {
        ry_h = cy.row_h[jc];
        ry_e = cy.row_e[jc];
        if (ry_h >= 0) { by_h = cy.b_h[ry_h]; gy_h = cy.g_h[ry_h]; }
        if (ry_e >= 0) { by_e = cy.b_e[ry_e]; gy_e = cy.g_e[ry_e]; }
    }

    const bool sample = d.n_monitors > 0 && ((n + 1) % d.sample_interval) == 0;
    const int64_t sample_index = (n + 1) / (d.sample_interval > 0 ? d.sample_interval : 1) - 1;

    // H_z of row k after this step from the old E (ey_below = E_y of row k-1, 0 below a non-periodic edge).
    // store: write H_z and its psi (the row is owned by this tile).
    auto hz_row = [&](int k, float hz_old, float ex_k, float ey_k, float ey_below, int id_k, bool store) -> float {
        const int64_t rk = (int64_t)k * ny;
        float ex_left = shfl_from_prev(ex_k);
        if (direct_left) ex_left = ex_o[rk + jl];
        if (jl < 0) ex_left = 0.f;
        const float dex = ex_k - ex_left;
        const float dey = ey_k - ey_below;
        const float4 c = tab[id_k];   // (da, db, eps_xx, eps_yy)
        float v = c.x * hz_old + c.y * (dex * idy_e - dey * d.inv_dx_e[k]);
        const int rx = cx.n_rows > 0 ? cx.row_e[k] : -1;
        if (rx >= 0) {   // x sides: H_z += -db psi
            const int64_t p = (int64_t)rx * ny + jc;
            const float psi = cx.b_e[rx] * cx.psi_e[par][p] + cx.g_e[rx] * dey;
            if (store && owner) cx.psi_e[par ^ 1][p] = psi;
            v -= c.y * psi;
        }
        if (ry_e >= 0) {   // y sides: H_z += +db psi
            const int64_t p = (int64_t)k * nry + ry_e;
            const float psi = by_e * cy.psi_e[par][p] + gy_e * dex;
            if (store && owner) cy.psi_e[par ^ 1][p] = psi;
            v += c.y * psi;
        }
        if (store && owner) hz_n[rk + j] = v;
        return valid ? v : 0.f;
    };

    // prologue: row i0 and E_y of row i0-1
    int64_t r0 = (int64_t)i0 * ny;
    float hz_old_cur = hz_o[r0 + jc];
    float ex_cur = valid ? ex_o[r0 + jc] : 0.f;
    float ey_cur = valid ? ey_o[r0 + jc] : 0.f;
    int id_cur = mat[r0 + jc];
    float ey_below = 0.f;
    if ((i0 > 0 || px) && valid) ey_below = ey_o[(int64_t)(i0 > 0 ? i0 - 1 : nx - 1) * ny + jc];
    float hz_cur = hz_row(i0, hz_old_cur, ex_cur, ey_cur, ey_below, id_cur, true);

    for (int i = i0; i < i1; ++i) {
        r0 = (int64_t)i * ny;
        const bool ey_updated = (i < nx - 1) || px;
        // H_z of the next row (row 0 after the last one on a periodic axis); owned when inside the tile
        float hz_next = 0.f;
        int id_next = id_cur;
        float ex_next = 0.f, ey_next = 0.f;
        if (ey_updated) {
            const int k = (i + 1 < nx) ? i + 1 : 0;
            const int64_t rk = (int64_t)k * ny;
            const float hz_old_next = hz_o[rk + jc];
            ex_next = valid ? ex_o[rk + jc] : 0.f;
            ey_next = valid ? ey_o[rk + jc] : 0.f;
            id_next = mat[rk + jc];
            hz_next = hz_row(k, hz_old_next, ex_next, ey_next, ey_cur, id_next, i + 1 < i1);
        }
        // E_x on the hi-y face: H_z of the next lane (+ y CPML 'h' rows, sign +)
        const float hz_right = shfl_from_next(hz_cur);
        const int id_right = __shfl_down(id_cur, 1, YEE_WAVE);
        float ex_v = ex_cur;
        if (ex_updated) {
            float ca, cb;
            face_coefficients(d, id_cur, id_right, 0, &ca, &cb);
            const float dh = hz_right - hz_cur;
            ex_v = ca * ex_cur + cb * dh * idy_h;
            if (ry_h >= 0) {
                const int64_t p = (int64_t)i * nry + ry_h;
                const float psi = by_h * cy.psi_h[par][p] + gy_h * dh;
                if (owner) cy.psi_h[par ^ 1][p] = psi;
                ex_v += cb * psi;
            }
        }
        // E_y on the hi-x face: H_z of the next row (+ x CPML 'h' rows, sign -)
        float ey_v = ey_cur;
        if (ey_updated) {
            float ca, cb;
            face_coefficients(d, id_cur, id_next, 1, &ca, &cb);
            const float dh = hz_next - hz_cur;
            ey_v = ca * ey_cur - cb * dh * d.inv_dx_h[i];
            const int rx = cx.n_rows > 0 ? cx.row_h[i] : -1;
            if (rx >= 0) {
                const int64_t p = (int64_t)rx * ny + jc;
                const float psi = cx.b_h[rx] * cx.psi_h[par][p] + cx.g_h[rx] * dh;
                if (owner) cx.psi_h[par ^ 1][p] = psi;
                ey_v -= cb * psi;
            }
        }
        // sheet current on the hi faces at t + dt/2 (E_y for a line along y, E_x for a line along x)
        for (int s = 0; s < d.n_sources; ++s) {
            const yee_line_source& src = d.source[s];
            if (!owner) continue;
            const float w = wave_value(src, n);
            if (src.axis == 1) {
                if (i == src.line && j >= src.lo && j < src.hi) ey_v += __fmul_rn(w, src.profile[j - src.lo]);
            } else if (j == src.line && i >= src.lo && i < src.hi) {
                ex_v += __fmul_rn(w, src.profile[i - src.lo]);
            }
        }
        if (owner) {
            ex_n[r0 + j] = ex_v;
            ey_n[r0 + j] = ey_v;
        }
        if (sample) {
            for (int m = 0; m < d.n_monitors; ++m) {
                const yee_flux_line& mon = d.monitor[m];
                if (mon.normal == 0) {
                    if (owner && i == mon.position && j >= mon.lo && j < mon.hi)
                        dft_accumulate(mon, sample_index, j - mon.lo, ey_v, hz_cur);
                } else if (owner && j == mon.position && i >= mon.lo && i < mon.hi) {
                    dft_accumulate(mon, sample_index, i - mon.lo, ex_v, hz_cur);
                }
            }
        }
        hz_cur = hz_next;
        ex_cur = ex_next;
        ey_cur = ey_next;
        id_cur = id_next;
    }
}

__device__ __forceinline__ void tile_2d(const yee2d_desc& d, int local_tile, int64_t n)
{
    const int tj_count = tiles_j_2d(d.ny);
    const int tj = local_tile % tj_count;
    const int ti = local_tile / tj_count;
    if (d.polarization == YEE_EZ)
        ez_tile(d, tj, ti, n);
    else
        hz_tile(d, tj, ti, n);
}

}  // namespace yee

// One instance: grid = ceil(tiles / 4) blocks of 4 independent wavefront tiles.
__global__ void __launch_bounds__(256) yee2d_step_kernel(yee2d_desc d, int32_t offset)
{
    const int64_t n = *d.clock + offset;
    if (n >= d.step_limit) return;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int total = yee::tiles_j_2d(d.ny) * yee::tiles_i_2d(d.nx, d.rows_per_tile);
    if (tile >= total) return;
    yee::tile_2d(d, tile, n);
}

// Many instances in one launch: tile t belongs to the instance whose tile range [prefix[k], prefix[k+1]) holds it.
__global__ void __launch_bounds__(256) yee2d_batch_kernel(const yee2d_desc* __restrict__ descs,
                                                         const int32_t* __restrict__ prefix, int32_t n_inst,
                                                         int32_t offset)
{
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= prefix[n_inst]) return;
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= tile) lo = mid; else hi = mid;
    }
    const yee2d_desc& d = descs[lo];
    const int64_t n = *d.clock + offset;
    if (n >= d.step_limit) return;
    yee::tile_2d(d, tile - prefix[lo], n);
}

// Advance one clock by k steps (bounded by the instance's step limit); a single lane writes with a vector store.
__global__ void yee_clock_advance(int64_t* clock, int64_t k, int64_t limit)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t v = *clock + k;
        *clock = v < limit ? v : limit;
    }
}

__global__ void yee_clock_advance_batch(const yee2d_desc* __restrict__ descs, int32_t n_inst, int64_t k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        int64_t* clock = descs[i].clock;
        const int64_t v = *clock + k;
        *clock = v < descs[i].step_limit ? v : descs[i].step_limit;
    }
}
