// Running-DFT box monitors for gfx950: one launch per step (skipped at once except on sample steps) accumulates the
// FP64 spectra of every DFT region of an instance, blockIdx.y per region and a grid-stride loop over its sampled
// cells. The grid is capped at one full occupancy of the device (2048 blocks of 256 lanes), so a step that takes no
// sample costs a few microseconds however large the box is.
//
// On sample step n ((n + 1) % m == 0, sample s = (n + 1) / m - 1) each lane reads its components from the new state
// set (one float each, or two for a face pair averaged to the cell centre) and adds v * phase[s][f] to its FP64 Kahan
// (sum, compensation) pairs of the real and imaginary parts at every frequency f. Accumulators are laid out
// [component][frequency][cell][4 (sum re, sum im, comp re, comp im)], so a wavefront reads and writes 64 consecutive
// 32-byte records: the kernel streams 64 B per cell, component and frequency plus the field reads, and is HBM bound
// on the sample steps.
//
// The descriptor is read in place at the kernel-argument pointer (constant address space): the field table is
// indexed by the runtime component, which would otherwise copy the whole descriptor to private memory.
//
// Reference semantics (file:line of the reference package open_sol/):
//   2D box monitor     physics/em/monitors.py:248-322 (DFTMonitor), :152-169 (dft_sample_slice, _centre_pairs)
//   3D plane monitor   physics/em3d/monitors3d.py:165-238 (DFTPlane3D)
//   phase              physics/em/monitors.py:22-34 (exp(-j omega t) m dt)
#include "yee.h"
#include <hip/hip_runtime.h>

namespace {

// FP64 Kahan-compensated accumulation of x into (sum, comp) held in registers: the true running sum is sum - comp
// (Kahan 1965; hipcc compiles without fast-math, so the compensation survives).
__device__ __forceinline__ void kahan(double& sum, double& comp, double x)
{
    const double y = x - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// Accumulate one sampled cell `t` of region `r`. field: the new state set's component pointers; n_axis: grid extent
// per axis; axis_stride: element stride per axis.
__device__ __forceinline__ void dft_cell(const yee_dft_region& r, float* const* field, const int32_t* n_axis,
                                         const int64_t* axis_stride, int64_t sample, int64_t t)
{
    const int64_t cells = (int64_t)r.count[0] * r.count[1] * r.count[2];
    if (t >= cells) return;
    const int64_t t2 = t % r.count[2];
    const int64_t rest = t / r.count[2];
    const int64_t t1 = rest % r.count[1];
    const int64_t t0 = rest / r.count[1];
    const int64_t index[3] = {r.start[0] + t0 * r.stride[0], r.start[1] + t1 * r.stride[1],
                              r.start[2] + t2 * r.stride[2]};
    const int64_t cell = index[0] * axis_stride[0] + index[1] * axis_stride[1] + index[2] * axis_stride[2];
    const double* phase = r.phase + sample * r.n_freq * 2;
    for (int c = 0; c < r.n_comp; ++c) {
        const float* f = field[r.comp[c]];
        float v = f[cell];
        const int a = r.pair_axis[c];
        if (a >= 0) {
            const int64_t lower = index[a] > 0 ? cell - axis_stride[a]
                                               : (r.pair_wrap[c] ? cell + (int64_t)(n_axis[a] - 1) * axis_stride[a]
                                                                 : cell);
            v = 0.5f * (v + f[lower]);
        }
        double4* acc = reinterpret_cast<double4*>(r.acc) + (int64_t)c * r.n_freq * cells + t;
        const double x = (double)v;
        for (int q = 0; q < r.n_freq; ++q) {
            double4 a = acc[(int64_t)q * cells];   // (sum re, sum im, comp re, comp im)
            kahan(a.x, a.z, x * phase[2 * q]);
            kahan(a.y, a.w, x * phase[2 * q + 1]);
            acc[(int64_t)q * cells] = a;
        }
    }
}

// Sample index of step n, or -1 when step n takes no DFT sample.
__device__ __forceinline__ int64_t sample_of(int64_t n, int32_t m)
{
    return ((n + 1) % m) == 0 ? (n + 1) / m - 1 : -1;
}

// Region `region` of a 2D instance after step n (cell index i * pitch + j); the grid's x blocks stride over its cells.
__device__ __forceinline__ void dft_region_2d(const yee2d_desc& d, int64_t n, int region)
{
    if (n >= d.step_limit) return;
    const yee_dft_region& r = d.dft[region];
    const int64_t s = sample_of(n, d.sample_interval);
    if (s < 0 || s >= r.n_samples) return;
    const int set = (int)((n & 1) ^ 1);
    const int32_t n_axis[3] = {d.nx, d.ny, 1};
    const int64_t axis_stride[3] = {d.pitch, 1, 0};
    const int64_t cells = (int64_t)r.count[0] * r.count[1] * r.count[2];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += (int64_t)gridDim.x * blockDim.x)
        dft_cell(r, d.field[set], n_axis, axis_stride, s, t);
}

// Region `region` of a 3D instance after step n (cell index (i * ny + j) * nz + k).
__device__ __forceinline__ void dft_region_3d(const yee3d_desc& d, int64_t n, int region)
{
    if (n >= d.step_limit) return;
    const yee_dft_region& r = d.dft[region];
    const int64_t s = sample_of(n, d.sample_interval);
    if (s < 0 || s >= r.n_samples) return;
    const int set = (int)((n & 1) ^ 1);
    const int32_t n_axis[3] = {d.nx, d.ny, d.nz};
    const int64_t axis_stride[3] = {(int64_t)d.ny * d.nz, d.nz, 1};
    const int64_t cells = (int64_t)r.count[0] * r.count[1] * r.count[2];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += (int64_t)gridDim.x * blockDim.x)
        dft_cell(r, d.field[set], n_axis, axis_stride, s, t);
}

}  // namespace

// DFT regions of a 2D instance after step n = *clock + offset, blockIdx.y per region.
__global__ void __launch_bounds__(256) yee2d_dft_kernel(yee2d_desc desc, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    const yee2d_desc& d = *(const yee2d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    if (d.clock != desc.clock) return;
    dft_region_2d(d, *d.clock + offset, blockIdx.y);
}

// DFT regions of a 3D instance after step n = *clock + offset, blockIdx.y per region.
__global__ void __launch_bounds__(256) yee3d_dft_kernel(yee3d_desc desc, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)(ConstDesc*)__builtin_amdgcn_kernarg_segment_ptr();
    if (d.clock != desc.clock) return;
    dft_region_3d(d, *d.clock + offset, blockIdx.y);
}

// The DFT regions of a batch's members: blockIdx.z is the member, blockIdx.y its region (the grid spans the most
// regions and the largest region of any member; the rest return at once). Descriptors are read through the constant
// address space.
__global__ void __launch_bounds__(256) yee2d_batch_dft_kernel(const yee2d_desc* __restrict__ descs, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee2d_desc ConstDesc;
    const yee2d_desc& d = *(const yee2d_desc*)((ConstDesc*)descs + blockIdx.z);
    if ((int)blockIdx.y < d.n_dft) dft_region_2d(d, *d.clock + offset, blockIdx.y);
}

__global__ void __launch_bounds__(256) yee3d_batch_dft_kernel(const yee3d_desc* __restrict__ descs, int32_t offset)
{
    typedef const __attribute__((address_space(4))) yee3d_desc ConstDesc;
    const yee3d_desc& d = *(const yee3d_desc*)((ConstDesc*)descs + blockIdx.z);
    if ((int)blockIdx.y < d.n_dft) dft_region_3d(d, *d.clock + offset, blockIdx.y);
}
