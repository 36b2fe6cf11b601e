// Fused 3D Yee step for gfx950: one launch advances a 3D instance by one full step (H update with its CPML on every
// axis, E update with its CPML, point dipole injection); a second, small launch takes the FP64 Kahan DFT samples of
// the flux planes on sample steps.
//
// Work decomposition. Arrays are (nx, ny, nz) with z contiguous. A wavefront tile spans 64 consecutive z cells, one
// per lane, of which lanes 1..62 own their cell: lane 0 holds the cell below (its H feeds the E of lane 1) and lane 63
// the cell above (its E feeds the H of lane 62), so every z neighbour is a wave shuffle and no lane runs extra code.
// A lane carries J rows along y in registers (rows j0..j0+J-1 owned, plus the halo row j0-1 whose H_x, H_z feed the
// E of row j0, plus row j0+J whose E feeds the H of row j0+J-1), and the tile marches planes i0..i1-1 along x with the
// new H of the previous plane kept in registers (the plane i0-1 is recomputed once at the start). Halo values are
// recomputed from the previous state, which is race-free because the fields and the H CPML psi are double-buffered
// (step n reads set n & 1 and writes set (n & 1) ^ 1); E psi is read and written only by the owner of its cell.
// Every copy of an H value goes through the same update helper (the library is built with -ffp-contract=on), so a
// recomputed halo equals its owner's value bit for bit.
//
// Boundaries: a periodic axis wraps its indices; on a non-periodic axis indices are clamped to the grid, which makes
// the hi neighbour of the last row the row itself (face partner = the cell itself, and inv_h[n-1] = 0 cancels the
// curl term), and the lo neighbour of row 0 is forced to 0.
//
// HBM traffic per cell-update: read six fields (24 B) + uint16 material id (2 B), write six fields (24 B) = 50 B.
//
// Reference semantics (file:line of the reference package open_sol/):
//   H update + CPML    physics/em3d/kernels3d.py:335-400 (update_h_3d), :44-56 (edge_coefficients)
//   E update + CPML    physics/em3d/kernels3d.py:538-610 (update_e_3d), :61-85 (face_coefficients)
//   CPML rows          physics/em3d/boundaries3d.py:208-241 (cpml_slabs)
//   dipole             physics/em3d/sources3d.py:41-75
//   plane flux         physics/em3d/monitors3d.py:90-101; physics/em/monitors.py:22-34, 92-96
//   step order         physics/em3d/engine3d.py:497-554
#include "yee_device.h"
#include "yee3d_index.h"

namespace yee3 {

constexpr int LANES = yee3idx::kLanes;
constexpr int OWNED = yee3idx::kOwned;   // z cells a wavefront tile owns (lanes 1..62)
constexpr int LDS_IDS = 512;       // ids whose table is staged in LDS: (da, db, ca, cb) + (eps_r, sigma_e), 12 KiB
constexpr float EPS0F = 8.8541878128e-12f;

__shared__ float4 lds3_coef[LDS_IDS];
__shared__ float2 lds3_mat[LDS_IDS];

__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Stage the table of one instance in LDS (whole block; every thread reaches the barrier).
__device__ __forceinline__ bool stage_tables(const yee3d_desc& d)
{
    const bool use = d.n_materials <= LDS_IDS;
    if (use) {
        for (int k = threadIdx.x; k < d.n_materials; k += blockDim.x) {
            const float* t = d.table + 8 * k;
            lds3_coef[k] = make_float4(t[0], t[1], t[2], t[3]);
            lds3_mat[k] = make_float2(t[4], t[5]);
        }
    }
    __syncthreads();
    return use;
}

template <bool LDS>
__device__ __forceinline__ float4 coef(const yee3d_desc& d, int id)
{
    if constexpr (LDS) return lds3_coef[id];
    else return *reinterpret_cast<const float4*>(d.table + 8 * id);
}

template <bool LDS>
__device__ __forceinline__ float2 mat(const yee3d_desc& d, int id)
{
    if constexpr (LDS) return lds3_mat[id];
    else return *reinterpret_cast<const float2*>(d.table + 8 * id + 4);
}

// (ca, cb) of an E face between cells of ids a and b: the id's own entry when they agree, else the coefficients of the
// float32 mean eps_r and sigma_e (shares 1/2; kernels3d.py:61-85). With k = sigma dt / (2 eps):
// ca = (1 - k) / (1 + k) = (eps - sigma dt/2) / (eps + sigma dt/2) and cb = (dt / eps) / (1 + k) = dt / (eps + sigma
// dt/2), formed with one reciprocal (within 2 ulp of the reference's float32 divisions).
// PEC ids carry eps_r = -1 and (ca, cb) = (0, 0): a face between two PEC cells is pinned by the id's own entry, and a
// face between a PEC cell and another cell takes the other cell's values (kernels3d.py:61-85), i.e. its own entry.
template <bool LDS>
__device__ __forceinline__ float2 face(const yee3d_desc& d, int a, int b)
{
    if (a == b) {
        const float4 c = coef<LDS>(d, a);
        return make_float2(c.z, c.w);
    }
    const float2 ma = mat<LDS>(d, a), mb = mat<LDS>(d, b);
    if (ma.x < 0.f || mb.x < 0.f) {
        const float4 c = coef<LDS>(d, ma.x < 0.f ? b : a);
        return make_float2(c.z, c.w);
    }
    const float eps = (0.5f * ma.x + 0.5f * mb.x) * EPS0F;
    const float loss = (0.5f * ma.y + 0.5f * mb.y) * (0.5f * d.dt);
    const float inv = __builtin_amdgcn_rcpf(eps + loss);
    return make_float2((eps - loss) * inv, d.dt * inv);
}

__device__ __forceinline__ float from_next(float v) { return __shfl_down(v, 1, LANES); }
__device__ __forceinline__ float from_prev(float v) { return __shfl_up(v, 1, LANES); }
__device__ __forceinline__ int from_next_i(int v) { return __shfl_down(v, 1, LANES); }

using yee3idx::fold;

// psi recursion of one layer cell: b psi + g dF.
__device__ __forceinline__ float psi_next(float b, float psi, float g, float difference) { return b * psi + g * difference; }

// ---- H components (shared by owners and halo copies) ----
// H_x: -db (dEz/dy - dEy/dz); CPML y first (-), z first (+).
__device__ __forceinline__ float hx_yee(float da, float db, float hx, float dez_y, float ihy, float dey_z, float ihz)
{
    return da * hx - db * (dez_y * ihy - dey_z * ihz);
}
// H_y: -db (dEx/dz - dEz/dx); CPML x first (+), z second (-).
__device__ __forceinline__ float hy_yee(float da, float db, float hy, float dex_z, float ihz, float dez_x, float ihx)
{
    return da * hy - db * (dex_z * ihz - dez_x * ihx);
}
// H_z: -db (dEy/dx - dEx/dy); CPML x second (-), y second (+).
__device__ __forceinline__ float hz_yee(float da, float db, float hz, float dey_x, float ihx, float dex_y, float ihy)
{
    return da * hz - db * (dey_x * ihx - dex_y * ihy);
}

// CPML row of one axis at one index: psi row (-1 outside the layers), b, gain.
struct Row {
    int r;
    float b, g;
};

// The per-index record of an axis read at a wave-uniform index: through the constant address space, so it is a scalar
// load (counted apart from the vector loads of the field rows, which it would otherwise wait behind).
__device__ __forceinline__ yee3d_axis_row uniform_row(const yee3d_axis_row* rows, int index)
{
    typedef const __attribute__((address_space(4))) float ConstF;
    typedef const __attribute__((address_space(4))) int32_t ConstI;
    ConstF* f = (ConstF*)(rows) + 8 * index;
    ConstI* n = (ConstI*)(rows) + 8 * index;
    yee3d_axis_row out;
    out.inv_h = f[0];
    out.inv_e = f[1];
    out.psi_row_h = n[2];
    out.b_h = f[3];
    out.g_h = f[4];
    out.psi_row_e = n[5];
    out.b_e = f[6];
    out.g_e = f[7];
    return out;
}

__device__ __forceinline__ Row h_row(const yee3d_axis_row& a) { return Row{a.psi_row_h, a.b_h, a.g_h}; }
__device__ __forceinline__ Row e_row(const yee3d_axis_row& a) { return Row{a.psi_row_e, a.b_e, a.g_e}; }

__device__ __forceinline__ int64_t psi_x(const yee3d_desc& d, int r, int j, int k)
{
    return ((int64_t)r * d.ny + j) * d.nz + k;
}
__device__ __forceinline__ int64_t psi_y(const yee3d_desc& d, int i, int r, int k)
{
    return ((int64_t)i * d.cpml[1].n_rows + r) * d.nz + k;
}
__device__ __forceinline__ int64_t psi_z(const yee3d_desc& d, int i, int j, int r)
{
    return ((int64_t)i * d.ny + j) * d.cpml[2].n_rows + r;
}

// Values of one cell column the H update needs from the current plane.
struct Cell {
    float ex, ey, ez;
    int id;
};

// New H of one cell.
struct HNew {
    float hx, hy, hz;
};

// H update of one cell at (i, j, k) from E at the cell (ex, ey, ez), at j+1 (ez_yp, ex_yp), at k+1 (ey_zp, ex_zp) and
// at i+1 (ey_xp, ez_xp), with the psi of every layer the cell lies in (read from set `par`; the owner, store = true,
// writes the new psi to set par ^ 1).
__device__ __forceinline__ HNew h_cell(const yee3d_desc& d, int par, bool store, float da, float db, float hx,
                                       float hy, float hz, float ex, float ey, float ez, float ez_yp, float ex_yp,
                                       float ey_zp, float ex_zp, float ey_xp, float ez_xp, float ihx, float ihy,
                                       float ihz, const Row& rx, const Row& ry, const Row& rz, int i, int j, int k)
{
    const float dez_y = ez_yp - ez, dey_z = ey_zp - ey;
    const float dex_z = ex_zp - ex, dez_x = ez_xp - ez;
    const float dey_x = ey_xp - ey, dex_y = ex_yp - ex;
    float vx = hx_yee(da, db, hx, dez_y, ihy, dey_z, ihz);
    float vy = hy_yee(da, db, hy, dex_z, ihz, dez_x, ihx);
    float vz = hz_yee(da, db, hz, dey_x, ihx, dex_y, ihy);
    const int out = par ^ 1;
    if (rx.r >= 0) {   // x layers: H_y from E_z (+), H_z from E_y (-)
        const int64_t p = psi_x(d, rx.r, j, k);
        const float a = psi_next(rx.b, d.cpml[0].psi_h[par][0][p], rx.g, dez_x);
        const float b = psi_next(rx.b, d.cpml[0].psi_h[par][1][p], rx.g, dey_x);
        if (store) {
            d.cpml[0].psi_h[out][0][p] = a;
            d.cpml[0].psi_h[out][1][p] = b;
        }
        vy += db * a;
        vz -= db * b;
    }
    if (ry.r >= 0) {   // y layers: H_x from E_z (-), H_z from E_x (+)
        const int64_t p = psi_y(d, i, ry.r, k);
        const float a = psi_next(ry.b, d.cpml[1].psi_h[par][0][p], ry.g, dez_y);
        const float b = psi_next(ry.b, d.cpml[1].psi_h[par][1][p], ry.g, dex_y);
        if (store) {
            d.cpml[1].psi_h[out][0][p] = a;
            d.cpml[1].psi_h[out][1][p] = b;
        }
        vx -= db * a;
        vz += db * b;
    }
    if (rz.r >= 0) {   // z layers: H_x from E_y (+), H_y from E_x (-)
        const int64_t p = psi_z(d, i, j, rz.r);
        const float a = psi_next(rz.b, d.cpml[2].psi_h[par][0][p], rz.g, dey_z);
        const float b = psi_next(rz.b, d.cpml[2].psi_h[par][1][p], rz.g, dex_z);
        if (store) {
            d.cpml[2].psi_h[out][0][p] = a;
            d.cpml[2].psi_h[out][1][p] = b;
        }
        vx += db * a;
        vy -= db * b;
    }
    if (d.mirror[0] | d.mirror[1] | d.mirror[2] | d.mirror[3] | d.mirror[4] | d.mirror[5]) {
        // symmetric mirror planes (the hi face of cell 0 / n - 2): tangential H = 0 (engine3d.py:513-518)
        if ((d.mirror[0] == YEE_MIRROR_SYMMETRIC && i == 0) || (d.mirror[1] == YEE_MIRROR_SYMMETRIC && i == d.nx - 2))
            vy = vz = 0.f;
        if ((d.mirror[2] == YEE_MIRROR_SYMMETRIC && j == 0) || (d.mirror[3] == YEE_MIRROR_SYMMETRIC && j == d.ny - 2))
            vx = vz = 0.f;
        if ((d.mirror[4] == YEE_MIRROR_SYMMETRIC && k == 0) || (d.mirror[5] == YEE_MIRROR_SYMMETRIC && k == d.nz - 2))
            vx = vy = 0.f;
    }
    return HNew{vx, vy, vz};
}

// Pole records on cell (i, j, k): each record whose box holds the cell updates its state from E^n of its component
// and corrects the updated value with that face's cb (physics/em3d/engine3d.py:508-530).
__device__ __forceinline__ void apply_poles_3d(const yee3d_desc& d, int i, int j, int k, const float e_old[3],
                                               const float cb[3], float e_new[3])
{
    for (int p = 0; p < d.n_poles; ++p) {
        const yee3d_pole& pl = d.pole[p];
        if (i < pl.i0 || i >= pl.i1 || j < pl.j0 || j >= pl.j1 || k < pl.k0 || k >= pl.k1) continue;
        const int64_t q = ((int64_t)(i - pl.i0) * (pl.j1 - pl.j0) + (j - pl.j0)) * (pl.k1 - pl.k0) + (k - pl.k0);
        const int c = pl.component;
        const float drive = pl.drive[q];
        if (pl.kind == YEE_DRUDE) {
            float jc = pl.state[0][q];
            jc += pl.damping * jc + drive * e_old[c];
            pl.state[0][q] = jc;
            e_new[c] -= cb[c] * jc;
        } else {
            const float pc = pl.state[0][q];
            float dpc = pl.state[1][q];
            dpc += pl.damping * dpc - pl.restoring * pc + drive * e_old[c];
            pl.state[0][q] = pc + dpc;
            pl.state[1][q] = dpc;
            e_new[c] -= cb[c] * pl.inv_dt * dpc;
        }
    }
}

// E update of one owned cell with its CPML (psi in place), dispersive poles (POLES) and point sources.
template <bool LDS, bool POLES>
__device__ __forceinline__ void e_cell(const yee3d_desc& d, int64_t n, float& ex, float& ey, float& ez, float2 fx,
                                       float2 fy, float2 fz, float dhz_y, float dhy_z, float dhx_z, float dhz_x,
                                       float dhy_x, float dhx_y, float iex, float iey, float iez, const Row& rx,
                                       const Row& ry, const Row& rz, int i, int j, int k)
{
    float vx = fx.x * ex + fx.y * (dhz_y * iey - dhy_z * iez);
    float vy = fy.x * ey + fy.y * (dhx_z * iez - dhz_x * iex);
    float vz = fz.x * ez + fz.y * (dhy_x * iex - dhx_y * iey);
    if (rx.r >= 0) {   // x layers: E_y from H_z (-), E_z from H_y (+)
        const int64_t p = psi_x(d, rx.r, j, k);
        float* pf = d.cpml[0].psi_e[0];
        float* ps = d.cpml[0].psi_e[1];
        const float a = psi_next(rx.b, pf[p], rx.g, dhz_x);
        const float b = psi_next(rx.b, ps[p], rx.g, dhy_x);
        pf[p] = a;
        ps[p] = b;
        vy -= fy.y * a;
        vz += fz.y * b;
    }
    if (ry.r >= 0) {   // y layers: E_x from H_z (+), E_z from H_x (-)
        const int64_t p = psi_y(d, i, ry.r, k);
        float* pf = d.cpml[1].psi_e[0];
        float* ps = d.cpml[1].psi_e[1];
        const float a = psi_next(ry.b, pf[p], ry.g, dhz_y);
        const float b = psi_next(ry.b, ps[p], ry.g, dhx_y);
        pf[p] = a;
        ps[p] = b;
        vx += fx.y * a;
        vz -= fz.y * b;
    }
    if (rz.r >= 0) {   // z layers: E_x from H_y (-), E_y from H_x (+)
        const int64_t p = psi_z(d, i, j, rz.r);
        float* pf = d.cpml[2].psi_e[0];
        float* ps = d.cpml[2].psi_e[1];
        const float a = psi_next(rz.b, pf[p], rz.g, dhy_z);
        const float b = psi_next(rz.b, ps[p], rz.g, dhx_z);
        pf[p] = a;
        ps[p] = b;
        vx -= fx.y * a;
        vy += fy.y * b;
    }
    if constexpr (POLES) {
        const float e_old[3] = {ex, ey, ez};
        const float cb[3] = {fx.y, fy.y, fz.y};
        float e_new[3] = {vx, vy, vz};
        apply_poles_3d(d, i, j, k, e_old, cb, e_new);
        vx = e_new[0];
        vy = e_new[1];
        vz = e_new[2];
    }
    for (int s = 0; s < d.n_sources; ++s) {
        const yee_point_source& src = d.source[s];
        if (src.i == i && src.j == j && src.k == k && n < src.wave_len) {
            const float add = __fmul_rn(src.weight, src.wave[n]);
            if (src.component == 0) vx += add;
            else if (src.component == 1) vy += add;
            else vz += add;
        }
    }
    ex = vx;
    ey = vy;
    ez = vz;
}

// Raw buffer access to one x plane of an array: the plane's base and size sit in a buffer resource (scalar
// registers), the row offset in a scalar register and the lane's z offset in a vector register, so a load or store
// needs no vector address arithmetic; an access outside the plane reads 0 and writes nothing.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int RSRC_DWORD3 = 0x00020000;   // 32-bit data format, no swizzle (gfx9 buffer resource word 3)

__device__ __forceinline__ Rsrc plane_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, RSRC_DWORD3);
}
__device__ __forceinline__ float ldf(Rsrc r, uint32_t lane_bytes, uint32_t row_bytes)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_bytes, (int)row_bytes, 0));
}
__device__ __forceinline__ int ldid(Rsrc r, uint32_t lane_bytes, uint32_t row_bytes)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b16(r, (int)lane_bytes, (int)row_bytes, 0);
}
// Default cache policy: non-temporal stores (nt) measured 2.240 against 2.085 ms per 512^3 step, while the 2D stores
// gain from nt. (Unlike the 2D rows, a 62-cell z row writes partial 64 B segments at both ends; that difference is
// the likely reason, not a measured one.)
__device__ __forceinline__ void stf(Rsrc r, uint32_t lane_bytes, uint32_t row_bytes, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)lane_bytes, (int)row_bytes, 0);
}

// One wavefront tile: planes [i0, i1), rows j0..j0+J-1, z cells tk*OWNED .. +OWNED-1.
template <int J, bool LDS, bool POLES>
__device__ void tile3d(const yee3d_desc& d, int64_t n, int i0, int i1, int tj, int tk)
{
    static_assert(J == yee3idx::kRows, "the row index math assumes kRows rows per lane");
    const int lane = threadIdx.x & (LANES - 1);
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const bool px = d.periodic[0] != 0, py = d.periodic[1] != 0, pz = d.periodic[2] != 0;
    const int64_t sx = (int64_t)ny * nz;
    const uint32_t plane_bytes = (uint32_t)(sx * 4), plane_id_bytes = (uint32_t)(sx * 2);
    const int par = (int)(n & 1);

    // z of this lane: byte offsets within a row
    const int kraw = yee3idx::lane_z(tk, lane);
    const int k = fold(kraw, nz, pz);
    const uint32_t vk4 = (uint32_t)k * 4u, vk2 = (uint32_t)k * 2u;
    const bool k_zero = !pz && kraw < 0;                      // lo neighbour of cell 0: H reads 0
    const bool k_own = yee3idx::lane_owns(lane, kraw, nz);
    const yee3d_axis_row az = d.axis_rows[2][k];
    const float ihz = az.inv_h, iez = az.inv_e;
    const Row rz_h = h_row(az), rz_e = e_row(az);

    // rows: 0 = halo j0-1, 1..J owned, J+1 = the row above (E only)
    int jr[J + 2];
    uint32_t roff[J + 2];   // row byte offsets j * nz * 4 within a field plane (wave-uniform)
    float ihy[J + 2], iey[J + 2];
    bool own[J + 2];
#pragma unroll
    for (int r = 0; r < J + 2; ++r) {
        const int jraw = yee3idx::row_y(tj, r);
        jr[r] = fold(jraw, ny, py);
        roff[r] = (uint32_t)jr[r] * (uint32_t)nz * 4u;
        own[r] = r >= 1 && r <= J && jraw < ny;
        const yee3d_axis_row ay = uniform_row(d.axis_rows[1], jr[r]);
        ihy[r] = ay.inv_h;
        iey[r] = ay.inv_e;
    }
    const bool row0_zero = !py && tj == 0;   // the halo row below j = 0 of a non-periodic y axis: its H reads 0

    // E (and ids) of plane i, rows 0..J+1 (the row above needs only E_x, E_z)
    auto load_plane = [&](int i, Cell (&c)[J + 2]) {
        const int64_t po = (int64_t)i * sx;
        const Rsrc rx = plane_rsrc(d.field[par][0] + po, plane_bytes);
        const Rsrc ry = plane_rsrc(d.field[par][1] + po, plane_bytes);
        const Rsrc rz = plane_rsrc(d.field[par][2] + po, plane_bytes);
        const Rsrc rid = plane_rsrc(d.material + po, plane_id_bytes);
#pragma unroll
        for (int r = 0; r < J + 2; ++r) {
            c[r].ex = ldf(rx, vk4, roff[r]);
            c[r].ey = r <= J ? ldf(ry, vk4, roff[r]) : 0.f;
            c[r].ez = ldf(rz, vk4, roff[r]);
            c[r].id = ldid(rid, vk2, roff[r] >> 1);
        }
    };

    // old H of plane i, rows r_lo..J (the halo row feeds E through H_x and H_z only)
    auto load_hold = [&](int i, float (&hold)[J + 1][3], int r_lo) {
        const int64_t po = (int64_t)i * sx;
        const Rsrc rx = plane_rsrc(d.field[par][3] + po, plane_bytes);
        const Rsrc ry = plane_rsrc(d.field[par][4] + po, plane_bytes);
        const Rsrc rz = plane_rsrc(d.field[par][5] + po, plane_bytes);
#pragma unroll
        for (int r = 0; r < J + 1; ++r) {
            if (r < r_lo) continue;
            hold[r][0] = ldf(rx, vk4, roff[r]);
            hold[r][1] = r > 0 ? ldf(ry, vk4, roff[r]) : 0.f;
            hold[r][2] = ldf(rz, vk4, roff[r]);
        }
    };

    // new H of plane i (cur) from E of plane i+1 (nxt), rows r_lo..J
    auto h_plane = [&](int i, const Cell (&cur)[J + 2], const Cell (&nxt)[J + 2], HNew (&out)[J + 1],
                       const float (&hold)[J + 1][3], int r_lo, bool store) {
        const yee3d_axis_row ax = uniform_row(d.axis_rows[0], i);
        const Row rx = h_row(ax);
#pragma unroll
        for (int r = 0; r < J + 1; ++r) {
            if (r < r_lo) continue;
            const Row ry = h_row(uniform_row(d.axis_rows[1], jr[r]));
            const float ey_zp = from_next(cur[r].ey);
            const float ex_zp = from_next(cur[r].ex);
            const float4 c = coef<LDS>(d, cur[r].id);
            out[r] = h_cell(d, par, store && own[r] && k_own, c.x, c.y, hold[r][0], hold[r][1], hold[r][2], cur[r].ex,
                            cur[r].ey, cur[r].ez, cur[r + 1].ez, cur[r + 1].ex, ey_zp, ex_zp, nxt[r].ey, nxt[r].ez,
                            ax.inv_h, ihy[r], ihz, rx, ry, rz_h, i, jr[r], k);
        }
    };

    Cell cur[J + 2], nxt[J + 2];
    float hold[J + 1][3];
    HNew hn[J + 1];
    float hy_prev[J + 1], hz_prev[J + 1];   // new H_y, H_z of plane i-1, rows 1..J

    // prologue: new H of plane i0 - 1 (rows 1..J), or 0 below cell 0 of a non-periodic x axis
    load_plane(i0, cur);
    if (i0 > 0 || px) {
        const int im = i0 > 0 ? i0 - 1 : nx - 1;
        Cell prev[J + 2];
        load_plane(im, prev);
        load_hold(im, hold, 1);
        h_plane(im, prev, cur, hn, hold, 1, false);
#pragma unroll
        for (int r = 1; r < J + 1; ++r) {
            hy_prev[r] = hn[r].hy;
            hz_prev[r] = hn[r].hz;
        }
    } else {
#pragma unroll
        for (int r = 0; r < J + 1; ++r) hy_prev[r] = hz_prev[r] = 0.f;
    }

    for (int i = i0; i < i1; ++i) {
        const int ip = fold(i + 1, nx, px);
        load_plane(ip, nxt);
        load_hold(i, hold, 0);
        h_plane(i, cur, nxt, hn, hold, 0, true);
        // zero the H that stands for "below cell 0" on non-periodic axes
        if (row0_zero) { hn[0].hx = 0.f; hn[0].hz = 0.f; }
        if (k_zero) {
#pragma unroll
            for (int r = 0; r < J + 1; ++r) { hn[r].hx = 0.f; hn[r].hy = 0.f; }
        }
        const yee3d_axis_row ax = uniform_row(d.axis_rows[0], i);
        const Row rx_e = e_row(ax);
        const int64_t po = (int64_t)i * sx;
        const Rsrc wex = plane_rsrc(d.field[par ^ 1][0] + po, plane_bytes);
        const Rsrc wey = plane_rsrc(d.field[par ^ 1][1] + po, plane_bytes);
        const Rsrc wez = plane_rsrc(d.field[par ^ 1][2] + po, plane_bytes);
        const Rsrc whx = plane_rsrc(d.field[par ^ 1][3] + po, plane_bytes);
        const Rsrc why = plane_rsrc(d.field[par ^ 1][4] + po, plane_bytes);
        const Rsrc whz = plane_rsrc(d.field[par ^ 1][5] + po, plane_bytes);
#pragma unroll
        for (int r = 1; r < J + 1; ++r) {
            const float hx_km = from_prev(hn[r].hx);
            const float hy_km = from_prev(hn[r].hy);
            const int id_zp = from_next_i(cur[r].id);
            const float dhz_y = hn[r].hz - hn[r - 1].hz;
            const float dhy_z = hn[r].hy - hy_km;
            const float dhx_z = hn[r].hx - hx_km;
            const float dhz_x = hn[r].hz - hz_prev[r];
            const float dhy_x = hn[r].hy - hy_prev[r];
            const float dhx_y = hn[r].hx - hn[r - 1].hx;
            if (own[r] && k_own) {
                const int id = cur[r].id;
                const float2 fx = face<LDS>(d, id, nxt[r].id);
                const float2 fy = face<LDS>(d, id, cur[r + 1].id);
                const float2 fz = face<LDS>(d, id, id_zp);
                const Row ry_e = e_row(uniform_row(d.axis_rows[1], jr[r]));
                float ex = cur[r].ex, ey = cur[r].ey, ez = cur[r].ez;
                e_cell<LDS, POLES>(d, n, ex, ey, ez, fx, fy, fz, dhz_y, dhy_z, dhx_z, dhz_x, dhy_x, dhx_y, ax.inv_e,
                                   iey[r],
                            iez, rx_e, ry_e, rz_e, i, jr[r], k);
                stf(wex, vk4, roff[r], ex);
                stf(wey, vk4, roff[r], ey);
                stf(wez, vk4, roff[r], ez);
                stf(whx, vk4, roff[r], hn[r].hx);
                stf(why, vk4, roff[r], hn[r].hy);
                stf(whz, vk4, roff[r], hn[r].hz);
            }
            hy_prev[r] = hn[r].hy;
            hz_prev[r] = hn[r].hz;
        }
#pragma unroll
        for (int r = 0; r < J + 2; ++r) cur[r] = nxt[r];
    }
}

}  // namespace yee3

namespace yee3 {

// The tiles of one block of one instance: tile t = (i chunk, j tile, k tile), k fastest, one per wavefront. Every
// thread of the block reaches the table barrier before a wave without a tile leaves.
template <int J, bool POLES>
__device__ __forceinline__ void step_block(const yee3d_desc& d, int64_t n, int tile, int tiles, int tiles_j,
                                           int tiles_k)
{
    const bool mine = tile < tiles;
    if (!__syncthreads_or(mine)) return;
    const bool lds = stage_tables(d);
    if (!mine) return;
    const yee3idx::TileCoords c = yee3idx::tile_coords(tile, tiles_j, tiles_k);
    const int tk = c.tk, tj = c.tj, ic = c.ic;
    const int i0 = ic * d.planes_per_tile;
    const int i1 = min(d.nx, i0 + d.planes_per_tile);
    if (lds)
        tile3d<J, true, POLES>(d, n, i0, i1, tj, tk);
    else
        tile3d<J, false, POLES>(d, n, i0, i1, tj, tk);
}

}  // namespace yee3

// One 3D instance, four tiles per block. POLES: the E update carries the dispersive-pole records; that instantiation
// reads the descriptor in place at the kernel-argument pointer (constant address space), since the records are
// indexed at run time.
template <int J, bool POLES>
__global__ void __launch_bounds__(256) yee3d_step_kernel(yee3d_desc desc, int32_t tiles_j, int32_t tiles_k,
                                                        int32_t tiles, int32_t offset)
{
    if constexpr (POLES) {
        typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
        const yee3d_desc& d = *(const yee3d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
        if (d.clock != desc.clock) return;
        const int64_t n = *d.clock + offset;
        if (n >= d.step_limit) return;
        yee3::step_block<J, true>(d, n, blockIdx.x * 4 + yee3::wave_index(), tiles, tiles_j, tiles_k);
    } else {
        const int64_t n = *desc.clock + offset;
        if (n >= desc.step_limit) return;
        yee3::step_block<J, false>(desc, n, blockIdx.x * 4 + yee3::wave_index(), tiles, tiles_j, tiles_k);
    }
}

// Many 3D instances of uneven sizes in one launch: block b belongs to the instance whose block range
// [prefix[k], prefix[k+1]) holds it, so a block stages one instance's table. Each instance runs on its own clock
// and stops at its own step limit; members with dispersive poles take the E update that carries them.
template <int J>
__global__ void __launch_bounds__(256) yee3d_batch_kernel(const yee3d_desc* __restrict__ descs,
                                                         const int32_t* __restrict__ prefix, int32_t n_inst,
                                                         int32_t offset)
{
    const int block = blockIdx.x;
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= block) lo = mid; else hi = mid;
    }
    // Read the descriptor through the constant address space: its fields become scalar loads, as a kernel argument's.
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)((ConstDesc*)descs + lo);
    const int64_t n = *d.clock + offset;
    if (n >= d.step_limit) return;
    const int tiles_j = (d.ny + J - 1) / J;
    const int tiles_k = (d.nz + yee3::OWNED - 1) / yee3::OWNED;
    const int tiles = ((d.nx + d.planes_per_tile - 1) / d.planes_per_tile) * tiles_j * tiles_k;
    const int tile = (block - prefix[lo]) * 4 + yee3::wave_index();
    if (d.n_poles > 0)   // the E update with the member's dispersive-pole records (block-uniform: one member per block)
        yee3::step_block<J, true>(d, n, tile, tiles, tiles_j, tiles_k);
    else
        yee3::step_block<J, false>(d, n, tile, tiles, tiles_j, tiles_k);
}

template __global__ void yee3d_step_kernel<4, false>(yee3d_desc, int32_t, int32_t, int32_t, int32_t);
template __global__ void yee3d_step_kernel<4, true>(yee3d_desc, int32_t, int32_t, int32_t, int32_t);
template __global__ void yee3d_batch_kernel<4>(const yee3d_desc* __restrict__, const int32_t* __restrict__, int32_t,
                                               int32_t);

// FP64 Kahan DFT samples of the flux planes after step n (when (n + 1) % m == 0): one thread per plane cell and pair;
// E = mean of the layers edge - 1 and edge along the normal, H on layer edge - 1 (monitors3d.py:90-101).
namespace yee3 {

// Sample of plane cell t (cells of plane 0 pair 0, then pair 1, then plane 1 ...) after step n.
__device__ __forceinline__ void sample_cell(const yee3d_desc& d, int64_t n, int64_t t)
{
    if (n >= d.step_limit || d.n_planes == 0) return;
    const int m = d.sample_interval > 0 ? d.sample_interval : 1;
    if ((n + 1) % m != 0) return;
    const int64_t sample = (n + 1) / m - 1;
    const int set = (int)((n & 1) ^ 1);
    for (int p = 0; p < d.n_planes; ++p) {
        const yee_flux_plane& pl = d.plane[p];
        for (int q = 0; q < 2; ++q) {
            const int na = pl.hi[q][0] - pl.lo[q][0], nb = pl.hi[q][1] - pl.lo[q][1];
            const int64_t count = (int64_t)na * nb;
            if (t >= count) {
                t -= count;
                continue;
            }
            if (sample >= pl.n_samples) return;
            const int a = pl.lo[q][0] + (int)(t / nb), b = pl.lo[q][1] + (int)(t % nb);
            const int ta = pl.normal == 0 ? 1 : 0, tb = pl.normal == 2 ? 1 : 2;   // tangential axes in grid order
            int idx[3];
            idx[ta] = a;
            idx[tb] = b;
            idx[pl.normal] = pl.edge - 1;
            const int64_t lower = ((int64_t)idx[0] * d.ny + idx[1]) * d.nz + idx[2];
            idx[pl.normal] = pl.edge;
            const int64_t upper = ((int64_t)idx[0] * d.ny + idx[1]) * d.nz + idx[2];
            const float* e = d.field[set][pl.e_comp[q]];
            const float* h = d.field[set][3 + pl.h_comp[q]];
            const double e_value = 0.5 * ((double)e[lower] + (double)e[upper]);
            const double h_value = (double)h[lower];
            const int nf = pl.n_freq;
            const double* ph = pl.phase + sample * (int64_t)nf * 2;
            for (int f = 0; f < nf; ++f) {
                const double pr = ph[2 * f], pi = ph[2 * f + 1];
                double* acc_e = pl.acc[q] + ((int64_t)(0 * nf + f) * 4) * count + t;
                double* acc_h = pl.acc[q] + ((int64_t)(1 * nf + f) * 4) * count + t;
                yee::kahan_add(acc_e, acc_e + 2 * count, pr * e_value);
                yee::kahan_add(acc_e + count, acc_e + 3 * count, pi * e_value);
                yee::kahan_add(acc_h, acc_h + 2 * count, pr * h_value);
                yee::kahan_add(acc_h + count, acc_h + 3 * count, pi * h_value);
            }
            return;
        }
    }
}

}  // namespace yee3

__global__ void __launch_bounds__(256) yee3d_sample_kernel(yee3d_desc d, int64_t cells_total, int32_t offset)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells_total) return;
    yee3::sample_cell(d, *d.clock + offset, t);
}

// Samples of many instances in one launch: thread t belongs to the instance whose plane-cell range
// [prefix[k], prefix[k+1]) holds it.
__global__ void __launch_bounds__(256) yee3d_batch_sample_kernel(const yee3d_desc* __restrict__ descs,
                                                                const int64_t* __restrict__ prefix, int32_t n_inst,
                                                                int32_t offset)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= prefix[n_inst]) return;
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= t) lo = mid; else hi = mid;
    }
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)((ConstDesc*)descs + lo);
    yee3::sample_cell(d, *d.clock + offset, t - prefix[lo]);
}

namespace yee3 {

// Mirror face `face` (0 x_lo .. 5 z_hi) of one instance after step n, cell t of the face: symmetric, the normal E on
// the plane (the hi face of cell 0 / n - 2) = 0; antisymmetric, the ghost cell's (0 / n - 1) tangential E = minus its
// inner neighbour's (1 / n - 2) (physics/em3d/engine3d.py:531-540). The host launches the symmetric faces first, then
// the antisymmetric ones in face order, as the reference writes its zero rows before its images.
__device__ __forceinline__ void mirror_cell(const yee3d_desc& d, int64_t n, int face, int64_t t)
{
    if (n >= d.step_limit) return;
    const int a = face >> 1, hi = face & 1;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;   // the two tangential axes, in grid order
    const int32_t ext[3] = {d.nx, d.ny, d.nz};
    if (t >= (int64_t)ext[b] * ext[c]) return;
    int idx[3];
    idx[b] = (int)(t / ext[c]);
    idx[c] = (int)(t % ext[c]);
    const int64_t stride[3] = {(int64_t)d.ny * d.nz, d.nz, 1};
    const int set = (int)((n & 1) ^ 1);
    if (d.mirror[face] == YEE_MIRROR_SYMMETRIC) {
        idx[a] = hi ? ext[a] - 2 : 0;
        d.field[set][a][idx[0] * stride[0] + idx[1] * stride[1] + idx[2]] = 0.f;
    } else if (d.mirror[face] == YEE_MIRROR_ANTISYMMETRIC) {
        idx[a] = hi ? ext[a] - 1 : 0;
        const int64_t ghost = idx[0] * stride[0] + idx[1] * stride[1] + idx[2];
        const int64_t inner = ghost + (hi ? -stride[a] : stride[a]);
        d.field[set][b][ghost] = -d.field[set][b][inner];
        d.field[set][c][ghost] = -d.field[set][c][inner];
    }
}

// Field probes of one instance after step n: samples[n * n_probes + k] = field[new set][component[k]][cell[k]]
// (core/monitors.py:42-130, physics/em3d/physics3d.py:425-433); the block's lanes stride over the probes.
__device__ __forceinline__ void probes_3d(const yee3d_desc& d, int64_t n)
{
    const yee_probes& p = d.probes;
    if (n >= d.step_limit || n >= p.capacity) return;
    const int set = (int)((n & 1) ^ 1);
    const int64_t cells = (int64_t)d.nx * d.ny * d.nz;
    for (int k = threadIdx.x; k < p.n_probes; k += blockDim.x) {
        const int32_t c = p.component[k];
        const int64_t cell = p.cell[k];
        // a record outside the grid or the components samples NaN rather than reading out of bounds
        p.samples[n * p.n_probes + k] = (c >= 0 && c < 6 && cell >= 0 && cell < cells)
                                            ? d.field[set][c][cell] : __builtin_nanf("");
    }
}

}  // namespace yee3

// Mirror face `face` after step n = *clock + offset, one lane per cell of the face. The descriptor is read in place at
// the kernel-argument pointer: the field table is indexed by runtime axes.
__global__ void __launch_bounds__(256) yee3d_mirror_kernel(yee3d_desc desc, int32_t face, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    if (d.clock != desc.clock) return;
    yee3::mirror_cell(d, *d.clock + offset, face, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Field probes after step n = *clock + offset. One block; the descriptor is read in place.
__global__ void __launch_bounds__(256) yee3d_probe_kernel(yee3d_desc desc, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    if (d.clock != desc.clock) return;
    yee3::probes_3d(d, *d.clock + offset);
}

// Face `face` of the batch's members whose mirror there is of `kind`: blockIdx.y is the member, blockIdx.x * 256 +
// lane the face cell (the grid spans the largest face; the rest return at once). The host launches the kinds and faces
// in the single-instance order.
__global__ void __launch_bounds__(256) yee3d_batch_mirror_kernel(const yee3d_desc* __restrict__ descs, int32_t face,
                                                                int32_t kind, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)((ConstDesc*)descs + blockIdx.y);
    if (d.mirror[face] == kind)
        yee3::mirror_cell(d, *d.clock + offset, face, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The probes of a batch's members: block k serves member k (members without probes return at once).
__global__ void __launch_bounds__(256) yee3d_batch_probe_kernel(const yee3d_desc* __restrict__ descs, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)((ConstDesc*)descs + blockIdx.x);
    if (d.probes.n_probes > 0) yee3::probes_3d(d, *d.clock + offset);
}

// Advance the clocks of a 3D batch by k steps, each bounded by its instance's step limit.
__global__ void yee_clock_advance_batch3d(const yee3d_desc* __restrict__ descs, int32_t n_inst, int64_t k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_inst) {
        int64_t* clock = descs[i].clock;
        const int64_t v = *clock + k;
        *clock = v < descs[i].step_limit ? v : descs[i].step_limit;
    }
}
