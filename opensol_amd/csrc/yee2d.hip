// Fused 2D Yee step for gfx950: one launch advances E_z (E_z, H_x, H_y) or H_z (H_z, E_x, E_y) instances by one
// full step (H update, H CPML, E update, E CPML, sheet-source injection, FP64 Kahan DFT flux samples).
//
// Work decomposition. A wavefront tile owns 64 * VEC consecutive y columns, VEC per lane (y is contiguous in HBM:
// rows are padded to a multiple of 4 floats, so every lane access is an aligned 8 or 16 B vector), and marches
// rows_per_tile rows along x with register blocking along that slow axis. Cross-lane neighbours come from wave
// shuffles; the one neighbour across the tile edge (H_x of column j0-1 for E_z, H_z of column j0+64*VEC for H_z) is
// recomputed by the edge lane from direct reads. Neighbouring tiles read each other's edge rows/columns, which is
// race-free because every state array (fields and CPML psi) is double-buffered: step n reads set n & 1 and writes set
// (n & 1) ^ 1. Every copy of a cell goes through the same update helper (and the library is built with
// -ffp-contract=on), so a recomputed halo equals its owner's value bit for bit.
//
// Memory pipeline. The row index is wave-uniform (scalar registers), so per-row constants come from scalar loads,
// and the per-id coefficient table sits in LDS: the only vector-memory operations are the field / id rows, the CPML
// psi rows and the stores. Each iteration first issues the loads its own row depends on (psi), then the next row's
// field loads, then computes; the loop is unrolled over the rotating row buffers so no register copy forces a wait
// on the prefetch (gfx9 counts vector loads and stores in one in-order vmcnt).
//
// HBM traffic per cell-update: read E, H_a, H_b (12 B) + uint16 material id (2 B), write 12 B = 26 B.
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

__host__ __device__ inline int tiles_j_2d(int ny, int vec) { return (ny + vec * YEE_WAVE - 1) / (vec * YEE_WAVE); }
__host__ __device__ inline int tiles_i_2d(int nx, int rows) { return (nx + rows - 1) / rows; }

// ------------------------------------------------------------------------------------------------ LDS tables
constexpr int LDS_IDS_EZ = 1024;   // E_z: (ca, cb, da, db) per id, 16 KiB
constexpr int LDS_IDS_HZ = 512;    // H_z: (da, db, eps_xx, eps_yy), (ca_x, cb_x, ca_y, cb_y), sigma_e per id, 18 KiB

__shared__ float4 lds_tab_a[LDS_IDS_EZ];
__shared__ float4 lds_tab_b[LDS_IDS_HZ];
__shared__ float lds_tab_sig[LDS_IDS_HZ];

// Stage the coefficient table of one instance in LDS (whole block; every thread reaches the barrier).
template <int POL>
__device__ __forceinline__ bool stage_tables(const yee2d_desc& d)
{
    const int cap = POL == YEE_EZ ? LDS_IDS_EZ : LDS_IDS_HZ;
    const bool use = d.n_materials <= cap;
    if (use) {
        const float4* a = reinterpret_cast<const float4*>(d.table);
        for (int k = threadIdx.x; k < d.n_materials; k += blockDim.x) {
            lds_tab_a[k] = a[k];
            if (POL == YEE_HZ) {
                const float* t2 = d.table2 + 8 * k;
                lds_tab_b[k] = make_float4(t2[0], t2[1], t2[2], t2[3]);
                lds_tab_sig[k] = t2[4];
            }
        }
    }
    __syncthreads();
    return use;
}

template <bool LDS>
__device__ __forceinline__ float4 tab_a(const yee2d_desc& d, int id)
{
    if constexpr (LDS) return lds_tab_a[id];
    else return reinterpret_cast<const float4*>(d.table)[id];
}

template <bool LDS>
__device__ __forceinline__ float4 tab_b(const yee2d_desc& d, int id)
{
    if constexpr (LDS) return lds_tab_b[id];
    else return *reinterpret_cast<const float4*>(d.table2 + 8 * id);
}

template <bool LDS>
__device__ __forceinline__ float tab_sig(const yee2d_desc& d, int id)
{
    if constexpr (LDS) return lds_tab_sig[id];
    else return d.table2[8 * id + 4];
}

// ------------------------------------------------------------------------------------------------ vector access
template <int VEC>
__device__ __forceinline__ void ldv(const float* __restrict__ p, float out[VEC])
{
    if constexpr (VEC == 4) {
        const float4 v = *reinterpret_cast<const float4*>(p);
        out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
    } else {
        const float2 v = *reinterpret_cast<const float2*>(p);
        out[0] = v.x; out[1] = v.y;
    }
}

// Stores of the new state are non-temporal (the nt cache policy): nothing reads them again in this step, and the
// next step's reads come after the whole 8192^2 state has streamed past the caches. Streaming them keeps L2 for the
// halo rows and columns that neighbouring tiles re-read: 0.3451 against 0.3520 ms per 8192^2 E_z step.
template <int VEC>
__device__ __forceinline__ void stv(float* __restrict__ p, const float v[VEC])
{
    if constexpr (VEC == 4) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 x = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(x, reinterpret_cast<f4*>(p));
    } else {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const f2 x = {v[0], v[1]};
        __builtin_nontemporal_store(x, reinterpret_cast<f2*>(p));
    }
}

// VEC uint16 ids kept packed (two per register) until use, so a prefetch is not waited on early
template <int VEC>
__device__ __forceinline__ void ldids(const uint16_t* __restrict__ p, uint32_t out[VEC / 2])
{
    if constexpr (VEC == 4) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        out[0] = v.x; out[1] = v.y;
    } else {
        out[0] = *reinterpret_cast<const uint32_t*>(p);
    }
}

__device__ __forceinline__ int id_at(const uint32_t* packed, int k)
{
    return (packed[k >> 1] >> (16 * (k & 1))) & 0xffff;
}

// element k of a register array without dynamic indexing (which would move the array to scratch memory)
template <int VEC>
__device__ __forceinline__ float pickv(const float v[VEC], int k)
{
    float out = v[0];
#pragma unroll
    for (int q = 1; q < VEC; ++q) out = k == q ? v[q] : out;
    return out;
}

// ------------------------------------------------------------------------------------------ single-cell updates
// E_z polarisation, c = (ca, cb, da, db): H_y on a hi-x face = da H_y + db dE_z/dx (x CPML: psi = b psi + g dE,
// H_y += db psi); H_x on a hi-y face = da H_x - db dE_z/dy (y CPML: H_x -= db psi).
__device__ __forceinline__ float ez_hy(float4 c, float hy, float de, float idx, bool layer, float b, float g,
                                       float psi_old, float* psi_new)
{
    float v = c.z * hy + c.w * de * idx;
    if (layer) {
        const float psi = b * psi_old + g * de;
        *psi_new = psi;
        v += c.w * psi;
    }
    return v;
}

__device__ __forceinline__ float ez_hx(float4 c, float hx, float de, float idy, bool layer, float b, float g,
                                       float psi_old, float* psi_new)
{
    float v = c.z * hx - c.w * de * idy;
    if (layer) {
        const float psi = b * psi_old + g * de;
        *psi_new = psi;
        v -= c.w * psi;
    }
    return v;
}

// E_z at a centre = ca E_z + cb (dH_y/dx - dH_x/dy), before its CPML corrections.
__device__ __forceinline__ float ez_centre(float4 c, float e, float dhy, float dhx, float idx, float idy)
{
    return c.x * e + c.y * (dhy * idx - dhx * idy);
}

// H_z polarisation, c = (da, db, eps_xx, eps_yy): H_z at a centre = da H_z + db (dE_x/dy - dE_y/dx)
// (x CPML: H_z -= db psi_x, y CPML: H_z += db psi_y).
__device__ __forceinline__ float hz_centre(float4 c, float hz, float dex, float dey, float idy, float idx,
                                           bool layer_x, float bx, float gx, float psi_x_old, float* psi_x_new,
                                           bool layer_y, float by, float gy, float psi_y_old, float* psi_y_new)
{
    float v = c.x * hz + c.y * (dex * idy - dey * idx);
    if (layer_x) {
        const float psi = bx * psi_x_old + gx * dey;
        *psi_x_new = psi;
        v -= c.y * psi;
    }
    if (layer_y) {
        const float psi = by * psi_y_old + gy * dex;
        *psi_y_new = psi;
        v += c.y * psi;
    }
    return v;
}

// E_x on a hi-y face = ca E_x + cb dH_z/dy; E_y on a hi-x face = ca E_y - cb dH_z/dx (before CPML).
__device__ __forceinline__ float hz_ex(float ca, float cb, float ex, float dh, float idy) { return ca * ex + cb * dh * idy; }
__device__ __forceinline__ float hz_ey(float ca, float cb, float ey, float dh, float idx) { return ca * ey - cb * dh * idx; }

// E face coefficients between cells of ids a and b (comp 0: E_x face, 1: E_y face). Same id: the table's FP64
// coefficients (zero for the PEC id); mixed: mean eps / sigma (shares 1/2, fdtd_hz_2d_kernels.py:190-278) formed here
// in float32, or zero when either cell is PEC.
template <bool LDS>
__device__ __forceinline__ void face_coefficients(const yee2d_desc& d, int ida, int idb, int comp, float* ca,
                                                  float* cb)
{
    if (ida == idb) {
        const float4 s = tab_b<LDS>(d, ida);
        *ca = comp == 0 ? s.x : s.z;
        *cb = comp == 0 ? s.y : s.w;
        return;
    }
    const float4 a = tab_a<LDS>(d, ida), b = tab_a<LDS>(d, idb);
    const float ea = comp == 0 ? a.z : a.w;
    const float eb = comp == 0 ? b.z : b.w;
    if (ea < 0.f || eb < 0.f) {   // a PEC id (eps = -1): every E face touching a PEC cell is pinned
        *ca = 0.f;                // (physics/em/solvers/fdtd_hz_2d.py:475-495, yee_grid.py:145-150)
        *cb = 0.f;
        return;
    }
    const float eps_face = 0.5f * ea + 0.5f * eb;
    const float sig_face = 0.5f * tab_sig<LDS>(d, ida) + 0.5f * tab_sig<LDS>(d, idb);
    const float k = sig_face * (0.5f * d.dt_over_eps0) / eps_face;
    const float den = 1.0f + k;
    *ca = (1.0f - k) / den;
    *cb = (d.dt_over_eps0 / eps_face) / den;
}

// Per-column constants of one lane's columns: curl spacings and the y CPML psi row of each column (-1 outside).
template <int VEC>
struct Columns {
    int jb;                 // first column of the lane
    bool any;               // the lane holds at least one real column
    bool y_layer;           // some column of the lane lies in a y CPML layer
    float idy_e[VEC], idy_h[VEC];
    int ry_e[VEC], ry_h[VEC];

    __device__ __forceinline__ void load(const yee2d_desc& d, int first)
    {
        jb = first;
        any = jb < d.ny;
        y_layer = false;
        const yee_cpml_axis& cy = d.cpml[1];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int j = jb + k;
            const int jc = j < d.ny ? j : d.ny - 1;
            idy_e[k] = d.inv_dy_e[jc];
            idy_h[k] = d.inv_dy_h[jc];
            ry_e[k] = (cy.n_rows > 0 && j < d.ny) ? cy.row_e[j] : -1;
            ry_h[k] = (cy.n_rows > 0 && j < d.ny) ? cy.row_h[j] : -1;
            y_layer = y_layer || ry_e[k] >= 0 || ry_h[k] >= 0;
        }
    }
};

// Per-row x CPML constants (wave-uniform: scalar loads).
struct RowCpml {
    int r;          // psi row, -1 outside the x layers
    float b, g;
};

__device__ __forceinline__ RowCpml row_cpml(const yee_cpml_axis& c, int i, bool centre)
{
    RowCpml out{-1, 0.f, 0.f};
    if (c.n_rows > 0) {
        out.r = centre ? c.row_e[i] : c.row_h[i];
        if (out.r >= 0) {
            out.b = centre ? c.b_e[out.r] : c.b_h[out.r];
            out.g = centre ? c.g_e[out.r] : c.g_h[out.r];
        }
    }
    return out;
}

// Pole records of one E component (0 E_z cells, 1 E_x faces, 2 E_y faces) on row i: each record whose box holds the
// row updates its state from E^n (e_n) and corrects the updated E (e) with the cell's or face's coefficient cb.
template <int VEC>
__device__ __forceinline__ void apply_pole_records(const yee2d_desc& d, int component, int i, int jb,
                                                   const float e_n[VEC], const float cb[VEC], float e[VEC])
{
    for (int p = 0; p < d.n_poles; ++p) {
        const yee_pole& pl = d.pole[p];
        if (pl.component != component || i < pl.i0 || i >= pl.i1) continue;
        const int64_t base = (int64_t)(i - pl.i0) * (pl.j1 - pl.j0) - pl.j0;
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int j = jb + k;
            if (j < pl.j0 || j >= pl.j1) continue;
            const int64_t q = base + j;
            const float drive = pl.drive[q];
            if (pl.kind == YEE_DRUDE) {
                float jc = pl.state[0][q];
                jc += pl.damping * jc + drive * e_n[k];
                pl.state[0][q] = jc;
                e[k] -= cb[k] * jc;
            } else {
                const float pc = pl.state[0][q];
                float dpc = pl.state[1][q];
                dpc += pl.damping * dpc - pl.restoring * pc + drive * e_n[k];
                pl.state[0][q] = pc + dpc;
                pl.state[1][q] = dpc;
                e[k] -= cb[k] * pl.inv_dt * dpc;
            }
        }
    }
}

// Dispersive poles of row i (E_z cells), before the sources (physics/em/solvers/fdtd_ez_2d.py:262-358). The loop is
// uniform; rows outside every pole box skip it.
template <int VEC>
__device__ __forceinline__ void apply_poles(const yee2d_desc& d, int i, int jb, const float e_n[VEC],
                                            const float4 c[VEC], float ez[VEC])
{
    float cb[VEC];
#pragma unroll
    for (int k = 0; k < VEC; ++k) cb[k] = c[k].y;
    apply_pole_records<VEC>(d, 0, i, jb, e_n, cb, ez);
}

// ------------------------------------------------------------------------------------------------ TFSF boxes
// Box-edge corrections of the TFSF records (physics/em/sources.py:27-130); the incident tables come from the host's
// run of the 1D incident line. Every copy of a corrected H value (owner, prologue row, halo column) goes through these
// helpers, so the copies stay bit-identical. E-side corrections skip PEC cells and pinned faces (cb = 0), as the
// reference's keep re-pin after injection does.
__device__ __forceinline__ float tfsf_line(const yee_tfsf& t, int64_t n, int k)
{
    return n < t.n_steps ? t.line[n * (t.x1 - t.x0) + k] : 0.f;
}
__device__ __forceinline__ float tfsf_end(const yee_tfsf& t, int64_t n, int e)
{
    return n < t.n_steps ? t.ends[2 * n + e] : 0.f;
}

// E_z: H_y on the face row i (rows x0-1, x1-1; columns [y0, y1)) from the incident E_z at time n
__device__ __forceinline__ float tfsf_ez_hy(const yee2d_desc& d, int64_t n, int i, int j, float v)
{
    for (int b = 0; b < d.n_tfsf; ++b) {
        const yee_tfsf& t = d.tfsf[b];
        if (j < t.y0 || j >= t.y1) continue;
        if (i == t.x0 - 1) v += t.coef_h[0] * tfsf_line(t, n, 0);
        if (i == t.x1 - 1) v += t.coef_h[1] * tfsf_line(t, n, t.x1 - 1 - t.x0);
    }
    return v;
}
// E_z: H_x on the face column j (columns y0-1, y1-1; rows [x0, x1)) from the incident E_z at time n
__device__ __forceinline__ float tfsf_ez_hx(const yee2d_desc& d, int64_t n, int i, int j, float v)
{
    for (int b = 0; b < d.n_tfsf; ++b) {
        const yee_tfsf& t = d.tfsf[b];
        if (i < t.x0 || i >= t.x1) continue;
        if (j == t.y0 - 1) v += t.coef_h[2] * tfsf_line(t, n, i - t.x0);
        if (j == t.y1 - 1) v += t.coef_h[3] * tfsf_line(t, n, i - t.x0);
    }
    return v;
}
// E_z: E_z of rows x0, x1-1 (columns [y0, y1)) from the incident H_y at time n + 1/2
__device__ __forceinline__ float tfsf_ez_e(const yee2d_desc& d, int64_t n, int i, int j, float v, float cb)
{
    if (cb == 0.f) return v;
    for (int b = 0; b < d.n_tfsf; ++b) {
        const yee_tfsf& t = d.tfsf[b];
        if (j < t.y0 || j >= t.y1) continue;
        if (i == t.x0) v += t.coef_e[0] * tfsf_end(t, n, 0);
        if (i == t.x1 - 1) v += t.coef_e[1] * tfsf_end(t, n, 1);
    }
    return v;
}
// H_z: H_z of rows x0, x1-1 (columns [y0, y1)) from the incident E_y at time n
__device__ __forceinline__ float tfsf_hz_h(const yee2d_desc& d, int64_t n, int i, int j, float v)
{
    for (int b = 0; b < d.n_tfsf; ++b) {
        const yee_tfsf& t = d.tfsf[b];
        if (j < t.y0 || j >= t.y1) continue;
        if (i == t.x0) v += t.coef_h[0] * tfsf_end(t, n, 0);
        if (i == t.x1 - 1) v += t.coef_h[1] * tfsf_end(t, n, 1);
    }
    return v;
}
// H_z: E_y on the face rows x0-1, x1-1 (columns [y0, y1)) from the incident H_z at time n + 1/2
__device__ __forceinline__ float tfsf_hz_ey(const yee2d_desc& d, int64_t n, int i, int j, float v, float cb)
{
    if (cb == 0.f) return v;
    for (int b = 0; b < d.n_tfsf; ++b) {
        const yee_tfsf& t = d.tfsf[b];
        if (j < t.y0 || j >= t.y1) continue;
        if (i == t.x0 - 1) v += t.coef_e[0] * tfsf_line(t, n, 0);
        if (i == t.x1 - 1) v += t.coef_e[1] * tfsf_line(t, n, t.x1 - 1 - t.x0);
    }
    return v;
}
// H_z: E_x on the face columns y0-1, y1-1 (rows [x0, x1)) from the incident H_z at time n + 1/2
__device__ __forceinline__ float tfsf_hz_ex(const yee2d_desc& d, int64_t n, int i, int j, float v, float cb)
{
    if (cb == 0.f) return v;
    for (int b = 0; b < d.n_tfsf; ++b) {
        const yee_tfsf& t = d.tfsf[b];
        if (i < t.x0 || i >= t.x1) continue;
        if (j == t.y0 - 1) v += t.coef_e[2] * tfsf_line(t, n, i - t.x0);
        if (j == t.y1 - 1) v += t.coef_e[3] * tfsf_line(t, n, i - t.x0);
    }
    return v;
}

__device__ __forceinline__ int wave_index()
{
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// ------------------------------------------------------------------------------------------------------------ E_z
// Loads of row i issued one row ahead: E of row i+1 (the H_y difference), H_x, H_y, ids of row i, and the edge
// values: lane 0 reads column jl (E, H_x, id), the right-edge lane reads E of column jr.
template <int VEC>
struct EzIn {
    float e_next[VEC], hx[VEC], hy[VEC];
    uint32_t ids[VEC / 2];
    float e_l, hx_l, e_r;
    uint32_t id_l;
};

template <int VEC, bool LDS>
struct EzTile {
    const yee2d_desc& d;
    int64_t n;
    int par, lane, nx, ny, P, jb_c, jl, jr, nry;
    bool px, py, left_halo, right_halo, sample;
    int64_t sample_index;
    Columns<VEC> col;
    float idy_h_l;
    int ry_l;
    float e_cur[VEC], hy_prev[VEC];

    __device__ __forceinline__ EzTile(const yee2d_desc& desc, int64_t step) : d(desc), n(step) {}

    __device__ __forceinline__ void load(int i, EzIn<VEC>& r) const
    {
        const bool hy_updated = (i < nx - 1) || px;
        const int inext = (i + 1 < nx) ? i + 1 : 0;
        const int64_t row = (int64_t)i * P;
        if (hy_updated) ldv<VEC>(d.field[par][0] + (int64_t)inext * P + jb_c, r.e_next);
        else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) r.e_next[k] = 0.f;
        }
        ldv<VEC>(d.field[par][1] + row + jb_c, r.hx);
        ldv<VEC>(d.field[par][2] + row + jb_c, r.hy);
        ldids<VEC>(d.material + row + jb_c, r.ids);
        r.e_l = r.hx_l = r.e_r = 0.f;
        r.id_l = 0;
        if (left_halo) {
            r.e_l = d.field[par][0][row + jl];
            r.hx_l = d.field[par][1][row + jl];
            r.id_l = d.material[row + jl];
        }
        if (right_halo) r.e_r = d.field[par][0][row + jr];
    }

    // one row: the row's own psi reads first, then the next row's prefetch, then the update
    __device__ __forceinline__ void step(int i, const EzIn<VEC>& cur, EzIn<VEC>& nxt, bool prefetch)
    {
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
        const int64_t row = (int64_t)i * P;
        const bool hy_updated = (i < nx - 1) || px;
        const RowCpml xh = row_cpml(cx, i, false);
        const RowCpml xe = row_cpml(cx, i, true);

        // (1) psi of this row
        float psi_xh[VEC] = {}, psi_xe[VEC] = {}, psi_yh[VEC] = {}, psi_ye[VEC] = {};
        float psi_l = 0.f;
        if (xh.r >= 0 && hy_updated) ldv<VEC>(cx.psi_h[par] + (int64_t)xh.r * P + jb_c, psi_xh);
        if (xe.r >= 0) ldv<VEC>(cx.psi_e[par] + (int64_t)xe.r * P + jb_c, psi_xe);
        if (col.y_layer) {
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                if (col.ry_h[k] >= 0) psi_yh[k] = cy.psi_h[par][(int64_t)i * nry + col.ry_h[k]];
                if (col.ry_e[k] >= 0) psi_ye[k] = cy.psi_e[par][(int64_t)i * nry + col.ry_e[k]];
            }
        }
        if (ry_l >= 0) psi_l = cy.psi_h[par][(int64_t)i * nry + ry_l];
        // (2) next row
        if (prefetch) load(i + 1, nxt);

        // (3) update
        const float idx_h = d.inv_dx_h[i];
        const float idx_e = d.inv_dx_e[i];
        float4 c[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) c[k] = tab_a<LDS>(d, id_at(cur.ids, k));   // (ca, cb, da, db)

        const float e_sh = shfl_from_next(e_cur[0]);
        float e_right[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int j = col.jb + k;
            if (k < VEC - 1 && j + 1 < ny) e_right[k] = e_cur[k + 1];
            else if (j + 1 < ny) e_right[k] = lane == YEE_WAVE - 1 ? cur.e_r : e_sh;
            else e_right[k] = cur.e_r;   // periodic wrap to column 0 (unused otherwise)
        }

        float hy_v[VEC], hx_v[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            hy_v[k] = cur.hy[k];
            const int j = col.jb + k;
            if (hy_updated) {
                hy_v[k] = ez_hy(c[k], cur.hy[k], cur.e_next[k] - e_cur[k], idx_h, xh.r >= 0, xh.b, xh.g, psi_xh[k],
                                &psi_xh[k]);
                if (d.n_tfsf > 0) hy_v[k] = tfsf_ez_hy(d, n, i, j, hy_v[k]);
            }
            hx_v[k] = j < ny ? cur.hx[k] : 0.f;
            if (j < ny && (j < ny - 1 || py)) {
                const bool layer = col.ry_h[k] >= 0;
                hx_v[k] = ez_hx(c[k], cur.hx[k], e_right[k] - e_cur[k], col.idy_h[k], layer,
                                layer ? cy.b_h[col.ry_h[k]] : 0.f, layer ? cy.g_h[col.ry_h[k]] : 0.f, psi_yh[k],
                                &psi_yh[k]);
                if (d.n_tfsf > 0) hx_v[k] = tfsf_ez_hx(d, n, i, j, hx_v[k]);
            }
        }
        // H_x left of the first column: the previous lane, or lane 0's recompute of column jl
        float hx_left = shfl_from_prev(hx_v[VEC - 1]);
        if (left_halo) {
            float unused;
            hx_left = ez_hx(tab_a<LDS>(d, cur.id_l), cur.hx_l, e_cur[0] - cur.e_l, idy_h_l, ry_l >= 0,
                            ry_l >= 0 ? cy.b_h[ry_l] : 0.f, ry_l >= 0 ? cy.g_h[ry_l] : 0.f, psi_l, &unused);
            if (d.n_tfsf > 0) hx_left = tfsf_ez_hx(d, n, i, jl, hx_left);
        } else if (lane == 0) {
            hx_left = 0.f;   // column 0 of a non-periodic axis reads H_x = 0 below it
        }

        float ez_v[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int j = col.jb + k;
            const float dhy = hy_v[k] - hy_prev[k];
            const float dhx = hx_v[k] - (k == 0 ? hx_left : hx_v[k - 1]);
            float v = ez_centre(c[k], e_cur[k], dhy, dhx, idx_e, col.idy_e[k]);
            if (xe.r >= 0) {
                psi_xe[k] = xe.b * psi_xe[k] + xe.g * dhy;
                v += c[k].y * psi_xe[k];
            }
            if (col.ry_e[k] >= 0) {
                psi_ye[k] = cy.b_e[col.ry_e[k]] * psi_ye[k] + cy.g_e[col.ry_e[k]] * dhx;
                v -= c[k].y * psi_ye[k];
            }
            ez_v[k] = j < ny ? v : 0.f;
        }
        // dispersive poles from E^n, then their corrections of the updated E (before the sources)
        apply_poles<VEC>(d, i, col.jb, e_cur, c, ez_v);
        // TFSF E-side corrections from the incident H at t + dt/2
        if (d.n_tfsf > 0) {
#pragma unroll
            for (int k = 0; k < VEC; ++k)
                if (col.jb + k < ny) ez_v[k] = tfsf_ez_e(d, n, i, col.jb + k, ez_v[k], c[k].y);
        }
        // sheet sources at t + dt/2: E_z += float32(value * profile)
        for (int s = 0; s < d.n_sources; ++s) {
            const yee_line_source& src = d.source[s];
            const float w = wave_value(src, n);
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                const int j = col.jb + k;
                const bool on_line = src.axis == 1 ? (i == src.line && j >= src.lo && j < src.hi)
                                                   : (j == src.line && i >= src.lo && i < src.hi);
                if (on_line) ez_v[k] += __fmul_rn(w, src.profile[src.axis == 1 ? j - src.lo : i - src.lo]);
            }
        }
        // (4) stores
        if (col.any) {
            const int jb = col.jb;
            stv<VEC>(d.field[par ^ 1][0] + row + jb, ez_v);
            stv<VEC>(d.field[par ^ 1][1] + row + jb, hx_v);
            stv<VEC>(d.field[par ^ 1][2] + row + jb, hy_v);
            if (xh.r >= 0 && hy_updated) stv<VEC>(cx.psi_h[par ^ 1] + (int64_t)xh.r * P + jb, psi_xh);
            if (xe.r >= 0) stv<VEC>(cx.psi_e[par ^ 1] + (int64_t)xe.r * P + jb, psi_xe);
            if (col.y_layer) {
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    const int j = jb + k;
                    if (col.ry_h[k] >= 0 && (j < ny - 1 || py))
                        cy.psi_h[par ^ 1][(int64_t)i * nry + col.ry_h[k]] = psi_yh[k];
                    if (col.ry_e[k] >= 0) cy.psi_e[par ^ 1][(int64_t)i * nry + col.ry_e[k]] = psi_ye[k];
                }
            }
        }
        if (sample) {
            for (int m = 0; m < d.n_monitors; ++m) {
                const yee_flux_line& mon = d.monitor[m];
#pragma unroll
                for (int k = 0; k < VEC; ++k) {
                    const int j = col.jb + k;
                    if (mon.normal == 0) {
                        if (i == mon.position && j >= mon.lo && j < mon.hi)
                            dft_accumulate(mon, sample_index, j - mon.lo, ez_v[k], hy_v[k]);
                    } else if (j == mon.position && i >= mon.lo && i < mon.hi) {
                        dft_accumulate(mon, sample_index, i - mon.lo, ez_v[k], hx_v[k]);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            hy_prev[k] = hy_v[k];
            e_cur[k] = cur.e_next[k];
        }
    }

    __device__ __forceinline__ void run(int tj, int i0, int i1)
    {
        lane = threadIdx.x & (YEE_WAVE - 1);
        nx = d.nx; ny = d.ny; P = d.pitch;
        px = d.periodic_x != 0; py = d.periodic_y != 0;
        par = (int)(n & 1);
        col.load(d, tj * VEC * YEE_WAVE + VEC * lane);
        const int jb = col.jb;
        jb_c = col.any ? jb : 0;
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
        nry = cy.n_rows;

        // left edge: H_x of column jl feeds the E update of column jb (lane 0 only)
        jl = jb > 0 ? jb - 1 : (py ? ny - 1 : -1);
        left_halo = lane == 0 && col.any && jl >= 0;
        idy_h_l = 0.f;
        ry_l = -1;
        if (left_halo) {
            idy_h_l = d.inv_dy_h[jl];
            if (nry > 0) ry_l = cy.row_h[jl];
        }
        // right edge: E of column jr for the H_x of the lane's last real column (tile edge or periodic wrap)
        const int last = min(jb + VEC - 1, ny - 1);
        const bool wraps = py && last == ny - 1 && col.any;
        jr = wraps ? 0 : jb + VEC;
        right_halo = col.any && (wraps || (lane == YEE_WAVE - 1 && jb + VEC < ny));

        sample = d.n_monitors > 0 && ((n + 1) % d.sample_interval) == 0;
        sample_index = (n + 1) / (d.sample_interval > 0 ? d.sample_interval : 1) - 1;

        // prologue: E of row i0, H_y of row i0-1 after this step (0 below a non-periodic grid edge)
        ldv<VEC>(d.field[par][0] + (int64_t)i0 * P + jb_c, e_cur);
#pragma unroll
        for (int k = 0; k < VEC; ++k) hy_prev[k] = 0.f;
        if ((i0 > 0 || px) && col.any) {
            const int ip = i0 > 0 ? i0 - 1 : nx - 1;
            const int64_t p0 = (int64_t)ip * P + jb_c;
            float e_p[VEC], hy_o[VEC], psi_o[VEC] = {};
            uint32_t idp[VEC / 2];
            ldv<VEC>(d.field[par][0] + p0, e_p);
            ldv<VEC>(d.field[par][2] + p0, hy_o);
            ldids<VEC>(d.material + p0, idp);
            const RowCpml xh = row_cpml(cx, ip, false);
            if (xh.r >= 0) ldv<VEC>(cx.psi_h[par] + (int64_t)xh.r * P + jb_c, psi_o);
            const float idx = d.inv_dx_h[ip];
#pragma unroll
            for (int k = 0; k < VEC; ++k) {
                float unused;
                hy_prev[k] = ez_hy(tab_a<LDS>(d, id_at(idp, k)), hy_o[k], e_cur[k] - e_p[k], idx, xh.r >= 0, xh.b,
                                   xh.g, psi_o[k], &unused);
                if (d.n_tfsf > 0) hy_prev[k] = tfsf_ez_hy(d, n, ip, jb + k, hy_prev[k]);
            }
        }

        EzIn<VEC> a, b;
        load(i0, a);
        int i = i0;
        for (; i + 1 < i1; i += 2) {
            step(i, a, b, true);
            step(i + 1, b, a, i + 2 < i1);
        }
        if (i < i1) step(i, a, b, false);
    }
};

// ------------------------------------------------------------------------------------------------------------ H_z
// Loads of row k issued ahead: H_z, E_x, E_y, ids, lane 0's E_x of column jl, and the right-edge lane's values of
// column jr (H_z, E_x, E_y, id) for recomputing H_z there.
template <int VEC>
struct HzIn {
    float hz[VEC], ex[VEC], ey[VEC];
    uint32_t ids[VEC / 2];
    float ex_l, hz_r, ex_r, ey_r;
    uint32_t id_r;
};

// psi reads of one H_z row (x row of centres, y columns, right-edge column)
template <int VEC>
struct HzPsi {
    float x[VEC], y[VEC], x_r, y_r;
};

// H_z outer-face rules of a face (wall_sign): YEE_WALL_ZERO pins the outermost H_z cells to 0 (pmc); +1 copies the
// inner neighbour into them (pec, antisymmetric mirror: H_z even) and pins the tangential E on the plane face (face
// row 0 / n - 2, the inner face of the outermost cell); -1 copies it negated (symmetric mirror: H_z odd)
// (physics/em/boundaries.py:82-93, 500-538).
__device__ __forceinline__ float hz_copy_sign(int32_t wall) { return wall == 1 ? 1.f : (wall == -1 ? -1.f : 0.f); }

// LEFT: the tile whose first column is the y_hi copy-wall column ny - 1 (a tile edge) recomputes H_z of column
// ny - 2 in lane 0 from memory, for the copy (a separate kernel instantiation; the host selects it).
template <int VEC, bool LDS, bool LEFT = false>
struct HzTile {
    const yee2d_desc& d;
    int64_t n;
    int par, lane, nx, ny, P, jb_c, jl, jr, kr, nry, i0, i1;
    bool px, py, left_halo, right_halo, sample;
    float s_xlo, s_xhi, s_ylo, s_yhi;   // copy-wall signs (0: no copy wall)
    bool left_hz;                       // LEFT: this lane recomputes H_z of column jl = ny - 2
    int64_t sample_index;
    Columns<VEC> col;
    float idy_e_r;
    int ry_e_r;
    float hz_cur[VEC], hz_cur_r;   // H_z after this step of the current row (and of the right-edge column)

    __device__ __forceinline__ HzTile(const yee2d_desc& desc, int64_t step) : d(desc), n(step) {}

    // the row after r, wrapped on a periodic axis; -1 past a non-periodic edge
    __device__ __forceinline__ int row_after(int r) const { return r < nx ? r : (px ? r - nx : -1); }

    __device__ __forceinline__ void load(int k, HzIn<VEC>& r) const
    {
        const int64_t row = (int64_t)k * P;
        ldv<VEC>(d.field[par][0] + row + jb_c, r.hz);
        ldv<VEC>(d.field[par][1] + row + jb_c, r.ex);
        ldv<VEC>(d.field[par][2] + row + jb_c, r.ey);
        ldids<VEC>(d.material + row + jb_c, r.ids);
        r.ex_l = r.hz_r = r.ex_r = r.ey_r = 0.f;
        r.id_r = 0;
        if (left_halo) r.ex_l = d.field[par][1][row + jl];
        if (right_halo) {
            r.hz_r = d.field[par][0][row + jr];
            r.ex_r = d.field[par][1][row + jr];
            r.ey_r = d.field[par][2][row + jr];
            r.id_r = d.material[row + jr];
        }
    }

    __device__ __forceinline__ void load_hz_psi(int k, const RowCpml& xe, HzPsi<VEC>& p) const
    {
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
#pragma unroll
        for (int q = 0; q < VEC; ++q) { p.x[q] = 0.f; p.y[q] = 0.f; }
        p.x_r = p.y_r = 0.f;
        if (xe.r >= 0) {
            ldv<VEC>(cx.psi_e[par] + (int64_t)xe.r * P + jb_c, p.x);
            if (right_halo) p.x_r = cx.psi_e[par][(int64_t)xe.r * P + jr];
        }
        if (col.y_layer) {
#pragma unroll
            for (int q = 0; q < VEC; ++q)
                if (col.ry_e[q] >= 0) p.y[q] = cy.psi_e[par][(int64_t)k * nry + col.ry_e[q]];
        }
        if (ry_e_r >= 0) p.y_r = cy.psi_e[par][(int64_t)k * nry + ry_e_r];
    }

    // H_z of row k after this step from its old values r (ey_below: E_y of row k-1, 0 below a non-periodic edge);
    // store: write H_z and its psi (the row belongs to this tile)
    __device__ __forceinline__ void hz_row(int k, const HzIn<VEC>& r, const float ey_below[VEC], float ey_below_r,
                                           const RowCpml& xe, HzPsi<VEC>& p, bool store, float out[VEC],
                                           float* out_r)
    {
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
        const float idx_e = d.inv_dx_e[k];
        float ex_left = shfl_from_prev(r.ex[VEC - 1]);
        if (lane == 0) ex_left = left_halo ? r.ex_l : 0.f;
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
            const int j = col.jb + q;
            const bool layer_y = col.ry_e[q] >= 0;
            float v = hz_centre(tab_a<LDS>(d, id_at(r.ids, q)), r.hz[q],
                                r.ex[q] - (q == 0 ? ex_left : r.ex[q - 1]), r.ey[q] - ey_below[q],
                                col.idy_e[q], idx_e, xe.r >= 0, xe.b, xe.g, p.x[q], &p.x[q], layer_y,
                                layer_y ? cy.b_e[col.ry_e[q]] : 0.f, layer_y ? cy.g_e[col.ry_e[q]] : 0.f,
                                p.y[q], &p.y[q]);
            if (d.n_sources > 0) v = magnetic_currents(k, j, v);
            if (d.n_tfsf > 0) v = tfsf_hz_h(d, n, k, j, v);
            if (wall_cell(k, j)) v = 0.f;
            out[q] = j < ny ? v : 0.f;
        }
        // copy walls on the y faces: column 0 from column 1 (the same lane), column ny - 1 from column ny - 2 (the same
        // lane, the previous lane, or lane 0's recompute when ny - 1 starts a tile)
        if (s_ylo != 0.f && col.jb == 0) out[0] = s_ylo * out[1];
        if (s_yhi != 0.f) {
            const float prev_lane = shfl_from_prev(out[VEC - 1]);
            float left_value = 0.f;
            if constexpr (LEFT) {
                if (left_hz) left_value = hz_recompute(k, jl, xe);
            }
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
                if (col.jb + q != ny - 1) continue;
                const float inner = q > 0 ? out[q > 0 ? q - 1 : 0] : (lane > 0 ? prev_lane : left_value);
                out[q] = s_yhi * inner;
            }
        }
        if (store && col.any) {
            const int64_t row = (int64_t)k * P + col.jb;
            stv<VEC>(d.field[par ^ 1][0] + row, out);
            if (xe.r >= 0) stv<VEC>(cx.psi_e[par ^ 1] + (int64_t)xe.r * P + col.jb, p.x);
            if (col.y_layer) {
#pragma unroll
                for (int q = 0; q < VEC; ++q)
                    if (col.ry_e[q] >= 0) cy.psi_e[par ^ 1][(int64_t)k * nry + col.ry_e[q]] = p.y[q];
            }
        }
        *out_r = 0.f;
        if (right_halo) {
            float unused_x, unused_y;
            *out_r = hz_centre(tab_a<LDS>(d, r.id_r), r.hz_r, r.ex_r - pickv<VEC>(r.ex, kr), r.ey_r - ey_below_r,
                               idy_e_r, idx_e, xe.r >= 0, xe.b, xe.g, p.x_r, &unused_x, ry_e_r >= 0,
                               ry_e_r >= 0 ? cy.b_e[ry_e_r] : 0.f, ry_e_r >= 0 ? cy.g_e[ry_e_r] : 0.f, p.y_r,
                               &unused_y);
            if (d.n_sources > 0) *out_r = magnetic_currents(k, jr, *out_r);
            if (d.n_tfsf > 0) *out_r = tfsf_hz_h(d, n, k, jr, *out_r);
            if (wall_cell(k, jr)) *out_r = 0.f;
            if (s_yhi != 0.f && jr == ny - 1) *out_r = s_yhi * pickv<VEC>(out, VEC - 1);   // the copy of column ny - 2
        }
    }

    // H_z of cell (k, j) after this step from memory (old values), its CPML, magnetic currents, TFSF and pmc rows:
    // the LEFT recompute of column ny - 2 (the same operands as hz_row, so the same bits).
    __device__ __forceinline__ float hz_recompute(int k, int j, const RowCpml& xe) const
    {
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
        const int64_t row = (int64_t)k * P;
        const float hz_o = d.field[par][0][row + j];
        const float ex_j = d.field[par][1][row + j];
        const int jm = j > 0 ? j - 1 : (py ? ny - 1 : -1);
        const float ex_m = jm >= 0 ? d.field[par][1][row + jm] : 0.f;
        const float ey_j = d.field[par][2][row + j];
        const int kb = k > 0 ? k - 1 : (px ? nx - 1 : -1);
        const float ey_b = kb >= 0 ? d.field[par][2][(int64_t)kb * P + j] : 0.f;
        const int id = d.material[row + j];
        const int ry = nry > 0 ? cy.row_e[j] : -1;
        const float psx = xe.r >= 0 ? cx.psi_e[par][(int64_t)xe.r * P + j] : 0.f;
        const float psy = ry >= 0 ? cy.psi_e[par][(int64_t)k * nry + ry] : 0.f;
        float unused_x, unused_y;
        float v = hz_centre(tab_a<LDS>(d, id), hz_o, ex_j - ex_m, ey_j - ey_b, d.inv_dy_e[j], d.inv_dx_e[k],
                            xe.r >= 0, xe.b, xe.g, psx, &unused_x, ry >= 0, ry >= 0 ? cy.b_e[ry] : 0.f,
                            ry >= 0 ? cy.g_e[ry] : 0.f, psy, &unused_y);
        if (d.n_sources > 0) v = magnetic_currents(k, j, v);
        if (d.n_tfsf > 0) v = tfsf_hz_h(d, n, k, j, v);
        if (wall_cell(k, j)) v = 0.f;
        return v;
    }

    // an outermost H_z cell of a pmc face (pinned to 0 after the H update and the magnetic currents)
    __device__ __forceinline__ bool wall_cell(int k, int j) const
    {
        return (d.wall_sign[0] != 0 && k == 0) || (d.wall_sign[1] != 0 && k == nx - 1) ||
               (d.wall_sign[2] != 0 && j == 0) || (d.wall_sign[3] != 0 && j == ny - 1);
    }

    // magnetic line currents on H_z cell (k, j) after the H update, at t_n: H_z += float32(value * profile). Every
    // copy of the cell (owner row, lookahead row, right halo column) adds the same term.
    __device__ __forceinline__ float magnetic_currents(int k, int j, float v) const
    {
        for (int s = 0; s < d.n_sources; ++s) {
            const yee_line_source& src = d.source[s];
            if (src.field != YEE_SOURCE_H) continue;
            const bool on_line = src.axis == 1 ? (k == src.line && j >= src.lo && j < src.hi)
                                               : (j == src.line && k >= src.lo && k < src.hi);
            if (on_line) v += __fmul_rn(wave_value(src, n), src.profile[src.axis == 1 ? j - src.lo : k - src.lo]);
        }
        return v;
    }

    // one row i: psi reads of H_z row i+1 and of the faces of row i, the prefetch of row i+2, then H_z of row i+1
    // (from nxt) and the faces of row i (from cur)
    __device__ __forceinline__ void step(int i, const HzIn<VEC>& cur, const HzIn<VEC>& nxt, HzIn<VEC>& pre)
    {
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
        const int64_t row = (int64_t)i * P;
        const int k1 = row_after(i + 1);
        const bool ey_updated = k1 >= 0;
        const RowCpml xe = ey_updated ? row_cpml(cx, k1, true) : RowCpml{-1, 0.f, 0.f};
        const RowCpml xh = row_cpml(cx, i, false);

        // (1) psi of H_z row k1 and of the faces of row i
        HzPsi<VEC> p;
        if (ey_updated) load_hz_psi(k1, xe, p);
        float psi_xh[VEC] = {}, psi_yh[VEC] = {};
        if (xh.r >= 0 && ey_updated) ldv<VEC>(cx.psi_h[par] + (int64_t)xh.r * P + jb_c, psi_xh);
        if (col.y_layer) {
#pragma unroll
            for (int q = 0; q < VEC; ++q)
                if (col.ry_h[q] >= 0) psi_yh[q] = cy.psi_h[par][(int64_t)i * nry + col.ry_h[q]];
        }
        // (2) the row after the lookahead row
        const int k2 = (i + 1 < i1) ? row_after(i + 2) : -1;
        if (k2 >= 0) load(k2, pre);

        // (3) H_z of row k1 (stored when inside the tile), then the faces of row i
        float hz_next[VEC] = {}, hz_next_r = 0.f;
        if (ey_updated) hz_row(k1, nxt, cur.ey, cur.ey_r, xe, p, i + 1 < i1, hz_next, &hz_next_r);
        // copy walls on the x faces: row 0 from row 1, row nx - 1 from row nx - 2 (after their y-face copies, so a
        // corner takes both factors), stored again over the rows written above
        if (s_xlo != 0.f && i == 0 && ey_updated) {
#pragma unroll
            for (int q = 0; q < VEC; ++q) hz_cur[q] = s_xlo * hz_next[q];
            hz_cur_r = s_xlo * hz_next_r;
            if (col.any) stv<VEC>(d.field[par ^ 1][0] + row + col.jb, hz_cur);
        }
        if (s_xhi != 0.f && k1 == nx - 1) {
#pragma unroll
            for (int q = 0; q < VEC; ++q) hz_next[q] = s_xhi * hz_cur[q];
            hz_next_r = s_xhi * hz_cur_r;
            if (i + 1 < i1 && col.any) stv<VEC>(d.field[par ^ 1][0] + (int64_t)k1 * P + col.jb, hz_next);
        }

        const float hz_sh = shfl_from_next(hz_cur[0]);
        const int id_sh = __shfl_down(id_at(cur.ids, 0), 1, YEE_WAVE);
        float ex_v[VEC], ey_v[VEC], cb_x[VEC], cb_y[VEC];
        const float idx_h = d.inv_dx_h[i];
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
            const int j = col.jb + q;
            const int id_q = id_at(cur.ids, q);
            // E_x on the hi-y face: H_z of the next column (+ y CPML 'h' rows, sign +)
            ex_v[q] = j < ny ? cur.ex[q] : 0.f;
            cb_x[q] = cb_y[q] = 0.f;
            if (j < ny && (j < ny - 1 || py)) {
                float hz_right;
                int id_right;
                if (q < VEC - 1 && j + 1 < ny) { hz_right = hz_cur[q + 1]; id_right = id_at(cur.ids, q + 1); }
                else if (j + 1 < ny && lane != YEE_WAVE - 1) { hz_right = hz_sh; id_right = id_sh; }
                else { hz_right = hz_cur_r; id_right = (int)cur.id_r; }   // tile edge or periodic wrap
                float ca, cb;
                face_coefficients<LDS>(d, id_q, id_right, 0, &ca, &cb);
                cb_x[q] = cb;
                const float dh = hz_right - hz_cur[q];
                ex_v[q] = hz_ex(ca, cb, cur.ex[q], dh, col.idy_h[q]);
                if (col.ry_h[q] >= 0) {
                    psi_yh[q] = cy.b_h[col.ry_h[q]] * psi_yh[q] + cy.g_h[col.ry_h[q]] * dh;
                    ex_v[q] += cb * psi_yh[q];
                }
            }
            // E_y on the hi-x face: H_z of the next row (+ x CPML 'h' rows, sign -)
            ey_v[q] = j < ny ? cur.ey[q] : 0.f;
            if (ey_updated && j < ny) {
                float ca, cb;
                face_coefficients<LDS>(d, id_q, id_at(nxt.ids, q), 1, &ca, &cb);
                cb_y[q] = cb;
                const float dh = hz_next[q] - hz_cur[q];
                ey_v[q] = hz_ey(ca, cb, cur.ey[q], dh, idx_h);
                if (xh.r >= 0) {
                    psi_xh[q] = xh.b * psi_xh[q] + xh.g * dh;
                    ey_v[q] -= cb * psi_xh[q];
                }
            }
        }
        // dispersive poles of the E faces from E^n, then their corrections (before the sources)
        if (d.n_poles > 0) {
            apply_pole_records<VEC>(d, 1, i, col.jb, cur.ex, cb_x, ex_v);
            apply_pole_records<VEC>(d, 2, i, col.jb, cur.ey, cb_y, ey_v);
        }
        // TFSF E-side corrections from the incident H_z at t + dt/2
        if (d.n_tfsf > 0) {
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
                const int j = col.jb + q;
                if (j >= ny) continue;
                ex_v[q] = tfsf_hz_ex(d, n, i, j, ex_v[q], cb_x[q]);
                if (ey_updated) ey_v[q] = tfsf_hz_ey(d, n, i, j, ey_v[q], cb_y[q]);
            }
        }
        // sheet current on the hi faces at t + dt/2 (E_y for a line along y, E_x for a line along x)
        for (int s = 0; s < d.n_sources; ++s) {
            const yee_line_source& src = d.source[s];
            if (src.field != YEE_SOURCE_E) continue;
            const float w = wave_value(src, n);
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
                const int j = col.jb + q;
                if (src.axis == 1) {
                    if (i == src.line && j >= src.lo && j < src.hi) ey_v[q] += __fmul_rn(w, src.profile[j - src.lo]);
                } else if (j == src.line && i >= src.lo && i < src.hi) {
                    ex_v[q] += __fmul_rn(w, src.profile[i - src.lo]);
                }
            }
        }
        // tangential E pinned on the plane face of the pec / antisymmetric faces, after the TFSF corrections and
        // the sheet currents (a dipole face on the plane is zeroed again)
        if ((d.wall_sign[0] == 1 && i == 0) || (d.wall_sign[1] == 1 && i == nx - 2)) {
#pragma unroll
            for (int q = 0; q < VEC; ++q) ey_v[q] = 0.f;
        }
        if (d.wall_sign[2] == 1 || d.wall_sign[3] == 1) {
#pragma unroll
            for (int q = 0; q < VEC; ++q) {
                const int j = col.jb + q;
                if ((d.wall_sign[2] == 1 && j == 0) || (d.wall_sign[3] == 1 && j == ny - 2)) ex_v[q] = 0.f;
            }
        }
        // (4) stores
        if (col.any) {
            const int jb = col.jb;
            stv<VEC>(d.field[par ^ 1][1] + row + jb, ex_v);
            stv<VEC>(d.field[par ^ 1][2] + row + jb, ey_v);
            if (xh.r >= 0 && ey_updated) stv<VEC>(cx.psi_h[par ^ 1] + (int64_t)xh.r * P + jb, psi_xh);
            if (col.y_layer) {
#pragma unroll
                for (int q = 0; q < VEC; ++q) {
                    const int j = jb + q;
                    if (col.ry_h[q] >= 0 && (j < ny - 1 || py))
                        cy.psi_h[par ^ 1][(int64_t)i * nry + col.ry_h[q]] = psi_yh[q];
                }
            }
        }
        if (sample) {
            for (int m = 0; m < d.n_monitors; ++m) {
                const yee_flux_line& mon = d.monitor[m];
#pragma unroll
                for (int q = 0; q < VEC; ++q) {
                    const int j = col.jb + q;
                    if (mon.normal == 0) {
                        if (i == mon.position && j >= mon.lo && j < mon.hi)
                            dft_accumulate(mon, sample_index, j - mon.lo, ey_v[q], hz_cur[q]);
                    } else if (j == mon.position && i >= mon.lo && i < mon.hi) {
                        dft_accumulate(mon, sample_index, i - mon.lo, ex_v[q], hz_cur[q]);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < VEC; ++q) hz_cur[q] = hz_next[q];
        hz_cur_r = hz_next_r;
    }

    __device__ __forceinline__ void run(int tj, int row0, int row1)
    {
        lane = threadIdx.x & (YEE_WAVE - 1);
        nx = d.nx; ny = d.ny; P = d.pitch;
        px = d.periodic_x != 0; py = d.periodic_y != 0;
        par = (int)(n & 1);
        i0 = row0;
        i1 = row1;
        col.load(d, tj * VEC * YEE_WAVE + VEC * lane);
        const int jb = col.jb;
        jb_c = col.any ? jb : 0;
        const yee_cpml_axis& cx = d.cpml[0];
        const yee_cpml_axis& cy = d.cpml[1];
        nry = cy.n_rows;

        jl = jb > 0 ? jb - 1 : (py ? ny - 1 : -1);
        left_halo = lane == 0 && col.any && jl >= 0;
        const int last = min(jb + VEC - 1, ny - 1);
        const bool wraps = py && last == ny - 1 && col.any;
        jr = wraps ? 0 : jb + VEC;
        kr = wraps ? last - jb : VEC - 1;
        right_halo = col.any && (wraps || (lane == YEE_WAVE - 1 && jb + VEC < ny));
        idy_e_r = 0.f;
        ry_e_r = -1;
        if (right_halo) {
            idy_e_r = d.inv_dy_e[jr];
            if (nry > 0) ry_e_r = cy.row_e[jr];
        }
        s_xlo = hz_copy_sign(d.wall_sign[0]);
        s_xhi = hz_copy_sign(d.wall_sign[1]);
        s_ylo = hz_copy_sign(d.wall_sign[2]);
        s_yhi = hz_copy_sign(d.wall_sign[3]);
        left_hz = LEFT && lane == 0 && col.jb == ny - 1 && s_yhi != 0.f;
        sample = d.n_monitors > 0 && ((n + 1) % d.sample_interval) == 0;
        sample_index = (n + 1) / (d.sample_interval > 0 ? d.sample_interval : 1) - 1;

        // prologue: row i0 (and E_y of row i0-1), H_z of row i0, the loads of row i0+1
        HzIn<VEC> a, b, c;
        load(i0, a);
        float ey_below[VEC] = {};
        float ey_below_r = 0.f;
        if (i0 > 0 || px) {
            const int ip = i0 > 0 ? i0 - 1 : nx - 1;
            ldv<VEC>(d.field[par][2] + (int64_t)ip * P + jb_c, ey_below);
            if (right_halo) ey_below_r = d.field[par][2][(int64_t)ip * P + jr];
        }
        {
            const RowCpml xe = row_cpml(cx, i0, true);
            HzPsi<VEC> p;
            load_hz_psi(i0, xe, p);
            hz_row(i0, a, ey_below, ey_below_r, xe, p, true, hz_cur, &hz_cur_r);
        }
        const int r1 = row_after(i0 + 1);
        if (r1 >= 0) load(r1, b);

        int i = i0;
        for (; i + 2 < i1; i += 3) {
            step(i, a, b, c);
            step(i + 1, b, c, a);
            step(i + 2, c, a, b);
        }
        if (i < i1) step(i, a, b, c);
        if (i + 1 < i1) step(i + 1, b, c, a);
    }
};

// ------------------------------------------------------------------------------------------------- bulk tiles
// A tile whose rows and columns (with the halo row/column on each side) lie outside every CPML layer, and that no
// source or monitor line crosses (the host lists such tiles), runs these lean loops: bulk spacings as scalars, no
// CPML, source or monitor code, seven live pointers. They evaluate the same helpers with the same operands as the
// general path, so both classes of tile produce identical bits.
template <int VEC>
struct EzBulkIn {
    float e_next[VEC], hx[VEC], hy[VEC];
    uint32_t ids[VEC / 2];
    float e_l, hx_l, e_r;
    uint32_t id_l;
};

template <int VEC, bool LDS, bool POLES>
__device__ __forceinline__ void ez_bulk_tile(const yee2d_desc& d, int tj, int ti, int64_t n)
{
    const int lane = threadIdx.x & (YEE_WAVE - 1);
    const int nx = d.nx, ny = d.ny, P = d.pitch;
    const bool px = d.periodic_x != 0, py = d.periodic_y != 0;
    const int par = (int)(n & 1);
    const int i0 = ti * d.rows_per_tile;
    const int i1 = min(nx, i0 + d.rows_per_tile);
    const int jb = tj * VEC * YEE_WAVE + VEC * lane;
    const bool any = jb < ny;
    const int jb_c = any ? jb : 0;
    const float* __restrict__ e_o = d.field[par][0];
    const float* __restrict__ hx_o = d.field[par][1];
    const float* __restrict__ hy_o = d.field[par][2];
    float* __restrict__ e_n = d.field[par ^ 1][0];
    float* __restrict__ hx_n = d.field[par ^ 1][1];
    float* __restrict__ hy_n = d.field[par ^ 1][2];
    const uint16_t* __restrict__ mat = d.material;
    const float idx = d.inv_dx_bulk, idy = d.inv_dy_bulk;
    const int jl = jb > 0 ? jb - 1 : (py ? ny - 1 : -1);
    const bool left_halo = lane == 0 && any && jl >= 0;
    const int last = min(jb + VEC - 1, ny - 1);
    const bool wraps = py && last == ny - 1 && any;
    const int jr = wraps ? 0 : jb + VEC;
    const bool right_halo = any && (wraps || (lane == YEE_WAVE - 1 && jb + VEC < ny));

    auto load = [&](int i, EzBulkIn<VEC>& r) {
        const bool hy_updated = (i < nx - 1) || px;
        const int inext = (i + 1 < nx) ? i + 1 : 0;
        const int64_t row = (int64_t)i * P;
        if (hy_updated) ldv<VEC>(e_o + (int64_t)inext * P + jb_c, r.e_next);
        else {
#pragma unroll
            for (int k = 0; k < VEC; ++k) r.e_next[k] = 0.f;
        }
        ldv<VEC>(hx_o + row + jb_c, r.hx);
        ldv<VEC>(hy_o + row + jb_c, r.hy);
        ldids<VEC>(mat + row + jb_c, r.ids);
        r.e_l = r.hx_l = r.e_r = 0.f;
        r.id_l = 0;
        if (left_halo) { r.e_l = e_o[row + jl]; r.hx_l = hx_o[row + jl]; r.id_l = mat[row + jl]; }
        if (right_halo) r.e_r = e_o[row + jr];
    };

    float e_cur[VEC], hy_prev[VEC];
    ldv<VEC>(e_o + (int64_t)i0 * P + jb_c, e_cur);
#pragma unroll
    for (int k = 0; k < VEC; ++k) hy_prev[k] = 0.f;
    if ((i0 > 0 || px) && any) {
        const int ip = i0 > 0 ? i0 - 1 : nx - 1;
        const int64_t p0 = (int64_t)ip * P + jb_c;
        float e_p[VEC], hy_o_p[VEC];
        uint32_t idp[VEC / 2];
        ldv<VEC>(e_o + p0, e_p);
        ldv<VEC>(hy_o + p0, hy_o_p);
        ldids<VEC>(mat + p0, idp);
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            float unused;
            hy_prev[k] = ez_hy(tab_a<LDS>(d, id_at(idp, k)), hy_o_p[k], e_cur[k] - e_p[k], idx, false, 0.f, 0.f,
                               0.f, &unused);
        }
    }

    auto step = [&](int i, const EzBulkIn<VEC>& cur, EzBulkIn<VEC>& nxt, bool prefetch) {
        if (prefetch) load(i + 1, nxt);
        const int64_t row = (int64_t)i * P;
        const bool hy_updated = (i < nx - 1) || px;
        float4 c[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) c[k] = tab_a<LDS>(d, id_at(cur.ids, k));
        const float e_sh = shfl_from_next(e_cur[0]);
        float hy_v[VEC], hx_v[VEC], ez_v[VEC];
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const int j = jb + k;
            float e_right;
            if (k < VEC - 1 && j + 1 < ny) e_right = e_cur[k + 1];
            else if (j + 1 < ny) e_right = lane == YEE_WAVE - 1 ? cur.e_r : e_sh;
            else e_right = cur.e_r;
            float unused;
            hy_v[k] = hy_updated ? ez_hy(c[k], cur.hy[k], cur.e_next[k] - e_cur[k], idx, false, 0.f, 0.f, 0.f,
                                         &unused)
                                 : cur.hy[k];
            hx_v[k] = j < ny ? cur.hx[k] : 0.f;
            if (j < ny && (j < ny - 1 || py))
                hx_v[k] = ez_hx(c[k], cur.hx[k], e_right - e_cur[k], idy, false, 0.f, 0.f, 0.f, &unused);
        }
        float hx_left = shfl_from_prev(hx_v[VEC - 1]);
        if (left_halo) {
            float unused;
            hx_left = ez_hx(tab_a<LDS>(d, cur.id_l), cur.hx_l, e_cur[0] - cur.e_l, idy, false, 0.f, 0.f, 0.f,
                            &unused);
        } else if (lane == 0) {
            hx_left = 0.f;
        }
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            const float dhx = hx_v[k] - (k == 0 ? hx_left : hx_v[k - 1]);
            ez_v[k] = jb + k < ny ? ez_centre(c[k], e_cur[k], hy_v[k] - hy_prev[k], dhx, idx, idy) : 0.f;
        }
        if constexpr (POLES) apply_poles<VEC>(d, i, jb, e_cur, c, ez_v);
        if (any) {
            stv<VEC>(e_n + row + jb, ez_v);
            stv<VEC>(hx_n + row + jb, hx_v);
            stv<VEC>(hy_n + row + jb, hy_v);
        }
#pragma unroll
        for (int k = 0; k < VEC; ++k) {
            hy_prev[k] = hy_v[k];
            e_cur[k] = cur.e_next[k];
        }
    };

    EzBulkIn<VEC> a, b;
    load(i0, a);
    int i = i0;
    for (; i + 1 < i1; i += 2) {
        step(i, a, b, true);
        step(i + 1, b, a, i + 2 < i1);
    }
    if (i < i1) step(i, a, b, false);
}

template <int VEC>
struct HzBulkIn {
    float hz[VEC], ex[VEC], ey[VEC];
    uint32_t ids[VEC / 2];
    float ex_l, hz_r, ex_r, ey_r;
    uint32_t id_r;
};

template <int VEC, bool LDS>
__device__ __forceinline__ void hz_bulk_tile(const yee2d_desc& d, int tj, int ti, int64_t n)
{
    const int lane = threadIdx.x & (YEE_WAVE - 1);
    const int nx = d.nx, ny = d.ny, P = d.pitch;
    const bool px = d.periodic_x != 0, py = d.periodic_y != 0;
    const int par = (int)(n & 1);
    const int i0 = ti * d.rows_per_tile;
    const int i1 = min(nx, i0 + d.rows_per_tile);
    const int jb = tj * VEC * YEE_WAVE + VEC * lane;
    const bool any = jb < ny;
    const int jb_c = any ? jb : 0;
    const float* __restrict__ hz_o = d.field[par][0];
    const float* __restrict__ ex_o = d.field[par][1];
    const float* __restrict__ ey_o = d.field[par][2];
    float* __restrict__ hz_n = d.field[par ^ 1][0];
    float* __restrict__ ex_n = d.field[par ^ 1][1];
    float* __restrict__ ey_n = d.field[par ^ 1][2];
    const uint16_t* __restrict__ mat = d.material;
    const float idx = d.inv_dx_bulk, idy = d.inv_dy_bulk;
    const int jl = jb > 0 ? jb - 1 : (py ? ny - 1 : -1);
    const bool left_halo = lane == 0 && any && jl >= 0;
    const int last = min(jb + VEC - 1, ny - 1);
    const bool wraps = py && last == ny - 1 && any;
    const int jr = wraps ? 0 : jb + VEC;
    const int kr = wraps ? last - jb : VEC - 1;
    const bool right_halo = any && (wraps || (lane == YEE_WAVE - 1 && jb + VEC < ny));
    auto row_after = [&](int r) -> int { return r < nx ? r : (px ? r - nx : -1); };

    auto load = [&](int k, HzBulkIn<VEC>& r) {
        const int64_t row = (int64_t)k * P;
        ldv<VEC>(hz_o + row + jb_c, r.hz);
        ldv<VEC>(ex_o + row + jb_c, r.ex);
        ldv<VEC>(ey_o + row + jb_c, r.ey);
        ldids<VEC>(mat + row + jb_c, r.ids);
        r.ex_l = r.hz_r = r.ex_r = r.ey_r = 0.f;
        r.id_r = 0;
        if (left_halo) r.ex_l = ex_o[row + jl];
        if (right_halo) { r.hz_r = hz_o[row + jr]; r.ex_r = ex_o[row + jr]; r.ey_r = ey_o[row + jr]; r.id_r = mat[row + jr]; }
    };
    auto hz_row = [&](int k, const HzBulkIn<VEC>& r, const float ey_below[VEC], float ey_below_r, bool store,
                      float out[VEC], float* out_r) {
        float ex_left = shfl_from_prev(r.ex[VEC - 1]);
        if (lane == 0) ex_left = left_halo ? r.ex_l : 0.f;
        float unused_x, unused_y;
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
            const float v = hz_centre(tab_a<LDS>(d, id_at(r.ids, q)), r.hz[q],
                                      r.ex[q] - (q == 0 ? ex_left : r.ex[q - 1]), r.ey[q] - ey_below[q], idy, idx,
                                      false, 0.f, 0.f, 0.f, &unused_x, false, 0.f, 0.f, 0.f, &unused_y);
            out[q] = jb + q < ny ? v : 0.f;
        }
        if (store && any) stv<VEC>(hz_n + (int64_t)k * P + jb, out);
        *out_r = 0.f;
        if (right_halo)
            *out_r = hz_centre(tab_a<LDS>(d, r.id_r), r.hz_r, r.ex_r - pickv<VEC>(r.ex, kr), r.ey_r - ey_below_r, idy,
                               idx, false, 0.f, 0.f, 0.f, &unused_x, false, 0.f, 0.f, 0.f, &unused_y);
    };

    float hz_cur[VEC], hz_cur_r;
    HzBulkIn<VEC> a, b, c;
    load(i0, a);
    {
        float ey_below[VEC] = {};
        float ey_below_r = 0.f;
        if (i0 > 0 || px) {
            const int ip = i0 > 0 ? i0 - 1 : nx - 1;
            ldv<VEC>(ey_o + (int64_t)ip * P + jb_c, ey_below);
            if (right_halo) ey_below_r = ey_o[(int64_t)ip * P + jr];
        }
        hz_row(i0, a, ey_below, ey_below_r, true, hz_cur, &hz_cur_r);
    }
    const int r1 = row_after(i0 + 1);
    if (r1 >= 0) load(r1, b);

    auto step = [&](int i, const HzBulkIn<VEC>& cur, const HzBulkIn<VEC>& nxt, HzBulkIn<VEC>& pre) {
        const int64_t row = (int64_t)i * P;
        const int k1 = row_after(i + 1);
        const bool ey_updated = k1 >= 0;
        const int k2 = (i + 1 < i1) ? row_after(i + 2) : -1;
        if (k2 >= 0) load(k2, pre);
        float hz_next[VEC] = {}, hz_next_r = 0.f;
        if (ey_updated) hz_row(k1, nxt, cur.ey, cur.ey_r, i + 1 < i1, hz_next, &hz_next_r);
        const float hz_sh = shfl_from_next(hz_cur[0]);
        const int id_sh = __shfl_down(id_at(cur.ids, 0), 1, YEE_WAVE);
        float ex_v[VEC], ey_v[VEC];
#pragma unroll
        for (int q = 0; q < VEC; ++q) {
            const int j = jb + q;
            const int id_q = id_at(cur.ids, q);
            ex_v[q] = j < ny ? cur.ex[q] : 0.f;
            if (j < ny && (j < ny - 1 || py)) {
                float hz_right;
                int id_right;
                if (q < VEC - 1 && j + 1 < ny) { hz_right = hz_cur[q + 1]; id_right = id_at(cur.ids, q + 1); }
                else if (j + 1 < ny && lane != YEE_WAVE - 1) { hz_right = hz_sh; id_right = id_sh; }
                else { hz_right = hz_cur_r; id_right = (int)cur.id_r; }
                float ca, cb;
                face_coefficients<LDS>(d, id_q, id_right, 0, &ca, &cb);
                ex_v[q] = hz_ex(ca, cb, cur.ex[q], hz_right - hz_cur[q], idy);
            }
            ey_v[q] = j < ny ? cur.ey[q] : 0.f;
            if (ey_updated && j < ny) {
                float ca, cb;
                face_coefficients<LDS>(d, id_q, id_at(nxt.ids, q), 1, &ca, &cb);
                ey_v[q] = hz_ey(ca, cb, cur.ey[q], hz_next[q] - hz_cur[q], idx);
            }
        }
        if (any) {
            stv<VEC>(ex_n + row + jb, ex_v);
            stv<VEC>(ey_n + row + jb, ey_v);
        }
#pragma unroll
        for (int q = 0; q < VEC; ++q) hz_cur[q] = hz_next[q];
        hz_cur_r = hz_next_r;
    };

    int i = i0;
    for (; i + 2 < i1; i += 3) {
        step(i, a, b, c);
        step(i + 1, b, c, a);
        step(i + 2, c, a, b);
    }
    if (i < i1) step(i, a, b, c);
    if (i + 1 < i1) step(i + 1, b, c, a);
}

constexpr int POL_ANY = 2;   // polarisation read from the descriptor (batched launches mixing E_z and H_z)

// Tile geometry of a local tile index (tiles run along y first, then x).
struct TileBox {
    int tj, ti, i0, i1, j0, j1;
};

template <int VEC>
__device__ __forceinline__ TileBox tile_box(const yee2d_desc& d, int local_tile)
{
    TileBox t;
    const int tj_count = tiles_j_2d(d.ny, VEC);
    t.tj = local_tile % tj_count;
    t.ti = local_tile / tj_count;
    t.i0 = t.ti * d.rows_per_tile;
    t.i1 = min(d.nx, t.i0 + d.rows_per_tile);
    t.j0 = t.tj * VEC * YEE_WAVE;
    t.j1 = min(d.ny, t.j0 + VEC * YEE_WAVE);
    return t;
}

// The k-th index of [0, n) not in the sorted list `skip` (length m).
__device__ __forceinline__ int nth_outside(const int32_t* skip, int m, int k)
{
    int v = k;
    for (int q = 0; q < m; ++q)
        if (skip[q] <= v) ++v;
    return v;
}

// Edge tile `item`: the full edge rows first (rows x all columns), then the edge columns of the other rows.
__device__ __forceinline__ bool edge_tile(const yee2d_tile_sets& s, int item, int* ti, int* tj)
{
    if (s.all_edge) {
        *ti = item / s.tiles_j;
        *tj = item % s.tiles_j;
        return item < s.tiles_i * s.tiles_j;
    }
    const int full = s.n_rows * s.tiles_j;
    if (item < full) {
        *ti = s.rows[item / s.tiles_j];
        *tj = item % s.tiles_j;
        return true;
    }
    const int k = item - full;
    if (s.n_cols == 0 || k >= (s.tiles_i - s.n_rows) * s.n_cols) return false;
    *ti = nth_outside(s.rows, s.n_rows, k / s.n_cols);
    *tj = s.cols[k % s.n_cols];
    return true;
}

// Bulk tile `item`: row-major over the rows and columns outside the edge lists.
__device__ __forceinline__ bool bulk_tile(const yee2d_tile_sets& s, int item, int* ti, int* tj)
{
    const int cols = s.tiles_j - s.n_cols;
    if (s.all_edge || cols <= 0 || item >= (s.tiles_i - s.n_rows) * cols) return false;
    *ti = nth_outside(s.rows, s.n_rows, item / cols);
    *tj = nth_outside(s.cols, s.n_cols, item % cols);
    return true;
}

template <int POL, int VEC, bool EDGE, bool LDS, bool POLES = false, bool LEFT = false>
__device__ __forceinline__ void tile_2d(const yee2d_desc& d, const TileBox& t, int64_t n, int part = 0,
                                        int parts = 1)
{
    const bool ez = POL == YEE_EZ || (POL == POL_ANY && d.polarization == YEE_EZ);
    if constexpr (EDGE) {
        const RowSpan span = edge_chunk(t.i0, t.i1, part, parts);
        const int r0 = span.r0;
        int r1 = span.r1;
        if (!ez && (d.wall_sign[1] == 1 || d.wall_sign[1] == -1)) {
            // an H_z x_hi copy wall: row nx - 1 copies row nx - 2, so the chunk holding row nx - 2 takes row nx - 1
            // too and no chunk starts at it
            if (r0 == d.nx - 1 && r0 > 0) return;
            if (r1 == d.nx - 1) r1 = d.nx;
        }
        if (r0 >= r1) return;
        if (ez) {
            EzTile<VEC, LDS> tile(d, n);
            tile.run(t.tj, r0, r1);
        } else {
            HzTile<VEC, LDS, LEFT> tile(d, n);
            tile.run(t.tj, r0, r1);
        }
    } else {
        if (ez)
            ez_bulk_tile<VEC, LDS, POLES>(d, t.tj, t.ti, n);
        else
            hz_bulk_tile<VEC, LDS>(d, t.tj, t.ti, n);
    }
}

}  // namespace yee

namespace yee {

// One block of the bulk (EDGE = false) or edge (EDGE = true) tile set; `block` counts blocks within that set.
template <int POL, int VEC, bool EDGE, bool POLES, bool LEFT = false>
__device__ __forceinline__ void step_block(const yee2d_desc& d, const yee2d_tile_sets& sets, int64_t n, int block)
{
    const int parts = EDGE ? sets.edge_parts : 1;
    const int slot = block * 4 + wave_index();
    const int item = slot / parts;
    int ti = 0, tj = 0;
    const bool mine = EDGE ? edge_tile(sets, item, &ti, &tj) : bulk_tile(sets, item, &ti, &tj);
    if (!__syncthreads_or(mine)) return;
    const bool lds = stage_tables<POL>(d);
    if (!mine) return;
    const TileBox t = tile_box<VEC>(d, ti * sets.tiles_j + tj);
    if (lds)
        tile_2d<POL, VEC, EDGE, true, POLES, LEFT>(d, t, n, slot % parts, parts);
    else
        tile_2d<POL, VEC, EDGE, false, POLES, LEFT>(d, t, n, slot % parts, parts);
}

}  // namespace yee

// One instance, two kernels over disjoint tile sets. The host lists the edge tile rows and columns (those touching a
// CPML layer, with the halo row / column, or a source or monitor line); every other tile is a bulk tile. EDGE = false
// steps the bulk tiles with the lean loop, EDGE = true the edge tiles with the general one, each edge tile split into
// sets.edge_parts row chunks (yee_edge_split). Tiles are enumerated compactly: no block is launched without work. (One
// launch holding both paths runs slower: it carries the general path's registers into the bulk tiles.)
//
// The descriptor is read in place in the kernel argument segment (the first argument sits at offset 0) through the
// constant address space. Bound to the argument itself, the general H_z path took its address in a way that made the
// compiler copy all 736 bytes to scratch and turn every field access into a flat load (168 VGPRs + 744 B/lane of
// scratch); read in place, the fields are scalar loads (127 VGPRs, no scratch).
template <int POL, int VEC, bool EDGE, bool POLES, bool LEFT>
__global__ void __launch_bounds__(256) yee2d_step_kernel(yee2d_desc desc, yee2d_tile_sets sets, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    const yee2d_desc& d = *(const yee2d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    if (d.clock != desc.clock) return;   // the argument layout is not the one assumed: do nothing
    const int64_t n = *d.clock + offset;
    if (n >= d.step_limit) return;
    yee::step_block<POL, VEC, EDGE, POLES, LEFT>(d, sets, n, blockIdx.x);
}

// POLES: the E_z bulk loop with the dispersive-pole update (a separate instantiation, so the loop without poles keeps
// its own register allocation); the general loop always carries it. LEFT: the H_z general loop with the lane-0
// recompute of column ny - 2 for a y_hi copy wall on a tile edge.
#define YEE_STEP_KERNEL(POL, VEC, EDGE, POLES, LEFT) \
    template __global__ void yee2d_step_kernel<POL, VEC, EDGE, POLES, LEFT>(yee2d_desc, yee2d_tile_sets, int32_t);
#define YEE_STEP_KERNELS(POL, VEC) \
    YEE_STEP_KERNEL(POL, VEC, false, false, false) YEE_STEP_KERNEL(POL, VEC, true, false, false)
YEE_STEP_KERNELS(YEE_EZ, 2)
YEE_STEP_KERNELS(YEE_EZ, 4)
YEE_STEP_KERNELS(YEE_HZ, 2)
YEE_STEP_KERNELS(YEE_HZ, 4)
YEE_STEP_KERNEL(YEE_HZ, 2, true, false, true)
YEE_STEP_KERNEL(YEE_HZ, 4, true, false, true)
YEE_STEP_KERNEL(YEE_EZ, 2, false, true, false)
YEE_STEP_KERNEL(YEE_EZ, 4, false, true, false)

// Many instances (E_z and H_z mixed) in one launch, one kernel per tile class: with EDGE = false every member's bulk
// tiles run the lean loops, with EDGE = true its edge tiles run the general loops (sources, monitors, CPML, poles,
// TFSF and H_z walls), one wavefront per tile. Tile t of
// the class belongs to the member whose range [prefix[k], prefix[k+1]) holds it; the member's tile sets (the host's
// lists for that member, E_z pole rows among the edge rows) give the tile. Tables are read from global memory (they
// stay in L1/L2).
template <int VEC, bool EDGE, bool LEFT>
__global__ void __launch_bounds__(256) yee2d_batch_kernel(const yee2d_desc* __restrict__ descs,
                                                         const yee2d_tile_sets* __restrict__ sets,
                                                         const int32_t* __restrict__ prefix, int32_t n_inst,
                                                         int32_t offset)
{
    const int tile = blockIdx.x * 4 + yee::wave_index();
    if (tile >= prefix[n_inst]) return;
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= tile) lo = mid; else hi = mid;
    }
    // Read the descriptor and the tile sets through the constant address space: their fields become scalar loads, as
    // a kernel argument's.
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    typedef const __attribute__((address_space(4))) yee2d_tile_sets ConstSets;
    const yee2d_desc& d = *(const yee2d_desc*)((ConstDesc*)descs + lo);
    const yee2d_tile_sets& s = *(const yee2d_tile_sets*)((ConstSets*)sets + lo);
    const int64_t n = *d.clock + offset;
    if (n >= d.step_limit) return;
    int ti = 0, tj = 0;
    const int item = tile - prefix[lo];
    if (!(EDGE ? yee::edge_tile(s, item, &ti, &tj) : yee::bulk_tile(s, item, &ti, &tj))) return;
    const yee::TileBox t = yee::tile_box<VEC>(d, ti * s.tiles_j + tj);
    if constexpr (EDGE && LEFT) {
        // an H_z y_hi copy wall on a tile edge: the general loop with the lane-0 recompute of column ny - 2
        const int32_t w = d.wall_sign[3];
        if (d.polarization == YEE_HZ && (w == 1 || w == -1) && (d.ny - 1) % (64 * VEC) == 0) {
            yee::tile_2d<yee::POL_ANY, VEC, true, false, false, true>(d, t, n);
            return;
        }
    }
    yee::tile_2d<yee::POL_ANY, VEC, EDGE, false>(d, t, n);
}

// LEFT: the edge kernel of a batch with an H_z member that has a y_hi copy wall on a tile edge; its members take the
// general loop with the lane-0 recompute (a separate instantiation: that loop's registers stay out of the plain one).
#define YEE_BATCH_KERNEL(VEC, EDGE, LEFT)                                                                        \
    template __global__ void yee2d_batch_kernel<VEC, EDGE, LEFT>(const yee2d_desc* __restrict__,                \
                                                                 const yee2d_tile_sets* __restrict__,           \
                                                                 const int32_t* __restrict__, int32_t, int32_t);
YEE_BATCH_KERNEL(2, false, false)
YEE_BATCH_KERNEL(2, true, false)
YEE_BATCH_KERNEL(2, true, true)
YEE_BATCH_KERNEL(4, false, false)
YEE_BATCH_KERNEL(4, true, false)
YEE_BATCH_KERNEL(4, true, true)

namespace yee {

// Field probes of one instance after step n: samples[n * n_probes + k] = field[new set][component[k]][cell[k]]
// (core/monitors.py:42-130), cell = i * pitch + j; the block's lanes stride over the probes.
__device__ __forceinline__ void probes_2d(const yee2d_desc& d, int64_t n)
{
    const yee_probes& p = d.probes;
    if (n >= d.step_limit || n >= p.capacity) return;
    const int set = (int)((n & 1) ^ 1);
    const int64_t cells = (int64_t)d.nx * d.pitch;
    for (int k = threadIdx.x; k < p.n_probes; k += blockDim.x) {
        const int32_t c = p.component[k];
        const int64_t cell = p.cell[k];
        // a record outside the grid or the components samples NaN rather than reading out of bounds
        p.samples[n * p.n_probes + k] = (c >= 0 && c < 3 && cell >= 0 && cell < cells)
                                            ? d.field[set][c][cell] : __builtin_nanf("");
    }
}

// E_z outer-face rows of one instance after step n (physics/em/boundaries.py:105-128 wall_centre_rows, :478-488):
// lane k < ny pins column k of the pinned x faces from row 1 (nx - 2), taking the y factor and the inner column at a
// pinned y corner; lane ny + i pins the y faces of row i (corner rows belong to the x lanes); lanes beyond do
// nothing. Every lane reads only rows and columns no lane writes, so the lanes need no ordering.
__device__ __forceinline__ void ez_walls(const yee2d_desc& d, int64_t n, int64_t k)
{
    if (n >= d.step_limit) return;
    float* ez = d.field[(n & 1) ^ 1][0];
    const int64_t P = d.pitch;
    const float s_xlo = (float)d.wall_sign[0], s_xhi = (float)d.wall_sign[1];
    const float s_ylo = (float)d.wall_sign[2], s_yhi = (float)d.wall_sign[3];
    if (k < d.ny) {
        int64_t j = k, inner = k;
        float yf = 1.0f;
        if (j == 0 && d.wall_sign[2] != 0) { inner = 1; yf = s_ylo; }
        else if (j == d.ny - 1 && d.wall_sign[3] != 0) { inner = d.ny - 2; yf = s_yhi; }
        if (d.wall_sign[0] != 0) ez[j] = (s_xlo * yf) * ez[P + inner];
        if (d.wall_sign[1] != 0) ez[(int64_t)(d.nx - 1) * P + j] = (s_xhi * yf) * ez[(int64_t)(d.nx - 2) * P + inner];
    } else if (k < (int64_t)d.ny + d.nx) {
        const int64_t i = k - d.ny;
        if ((i == 0 && d.wall_sign[0] != 0) || (i == d.nx - 1 && d.wall_sign[1] != 0)) return;
        const int64_t row = i * P;
        if (d.wall_sign[2] != 0) ez[row] = s_ylo * ez[row + 1];
        if (d.wall_sign[3] != 0) ez[row + d.ny - 1] = s_yhi * ez[row + d.ny - 2];
    }
}

__device__ __forceinline__ bool has_ez_walls(const yee2d_desc& d)
{
    return d.polarization == YEE_EZ &&
           (d.wall_sign[0] != 0 || d.wall_sign[1] != 0 || d.wall_sign[2] != 0 || d.wall_sign[3] != 0);
}

}  // namespace yee

// Field probes after step n = *clock + offset. One block; the descriptor is read in place as in yee2d_step_kernel.
__global__ void __launch_bounds__(256) yee2d_probe_kernel(yee2d_desc desc, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    const yee2d_desc& d = *(const yee2d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    if (d.clock != desc.clock) return;
    yee::probes_2d(d, *d.clock + offset);
}

// E_z outer-face rows after step n = *clock + offset, one lane per column and row.
__global__ void __launch_bounds__(256) yee2d_wall_kernel(yee2d_desc desc, int32_t offset)
{
    yee::ez_walls(desc, *desc.clock + offset, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The probes of a batch's members after their step *clock + offset: block k serves member k (members without probes
// return at once). Descriptors are read through the constant address space.
__global__ void __launch_bounds__(256) yee2d_batch_probe_kernel(const yee2d_desc* __restrict__ descs, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    const yee2d_desc& d = *(const yee2d_desc*)((ConstDesc*)descs + blockIdx.x);
    if (d.probes.n_probes > 0) yee::probes_2d(d, *d.clock + offset);
}

// The E_z walls of a batch's members: blockIdx.y is the member, blockIdx.x * 256 + lane its column / row lane (the
// grid spans the largest member; members without E_z walls return at once).
__global__ void __launch_bounds__(256) yee2d_batch_wall_kernel(const yee2d_desc* __restrict__ descs, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    const yee2d_desc& d = *(const yee2d_desc*)((ConstDesc*)descs + blockIdx.y);
    if (yee::has_ez_walls(d)) yee::ez_walls(d, *d.clock + offset, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// flag[0] = 1 when any of data[0 .. count) is NaN or infinite (grid-stride; lanes that find one store the same value,
// so the stores need no ordering). The flag is never cleared here.
__global__ void __launch_bounds__(256) yee_nonfinite_kernel(const float* __restrict__ data, int64_t count,
                                                           int32_t* flag)
{
    bool bad = false;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        bad = bad || !__builtin_isfinite(data[t]);
    if (bad) flag[0] = 1;
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
