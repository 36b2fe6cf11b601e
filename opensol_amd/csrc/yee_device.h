// Device helpers shared by the 2D and 3D Yee kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "yee.h"
#include "yee_launch.h"

#define YEE_WAVE 64

namespace yee {

// FP64 Kahan-compensated accumulation of x into (sum, comp): the true running sum is sum - comp.
// The order of operations is the published Kahan (1965) summation; no reassociation is allowed here, and hipcc
// compiles without fast-math, so the compensation survives.
__device__ __forceinline__ void kahan_add(double* __restrict__ sum, double* __restrict__ comp, double x)
{
    const double s = *sum;
    const double y = x - *comp;
    const double t = s + y;
    *comp = (t - s) - y;
    *sum = t;
}

// acc layout of one flux line: [2 (E, H)][n_freq][4 (sum re, sum im, comp re, comp im)][n_cells].
__device__ __forceinline__ void dft_accumulate(const yee_flux_line& mon, int64_t sample, int cell, float e_value,
                                               float h_value)
{
    const int n_cells = mon.hi - mon.lo;
    const int nf = mon.n_freq;
    const double* ph = mon.phase + sample * (int64_t)nf * 2;
    const double e = (double)e_value, h = (double)h_value;
    for (int f = 0; f < nf; ++f) {
        const double pr = ph[2 * f], pi = ph[2 * f + 1];
        double* a = mon.acc + ((int64_t)(0 * nf + f) * 4) * n_cells + cell;
        kahan_add(a, a + 2 * (int64_t)n_cells, pr * e);
        kahan_add(a + n_cells, a + 3 * (int64_t)n_cells, pi * e);
        double* b = mon.acc + ((int64_t)(1 * nf + f) * 4) * n_cells + cell;
        kahan_add(b, b + 2 * (int64_t)n_cells, pr * h);
        kahan_add(b + n_cells, b + 3 * (int64_t)n_cells, pi * h);
    }
}

// Source value injected after the E update of step n (0 past the table).
__device__ __forceinline__ float wave_value(const yee_line_source& s, int64_t n)
{
    return (n >= 0 && n < s.wave_len) ? s.wave[n] : 0.0f;
}

__device__ __forceinline__ float shfl_from_next(float v) { return __shfl_down(v, 1, YEE_WAVE); }
__device__ __forceinline__ float shfl_from_prev(float v) { return __shfl_up(v, 1, YEE_WAVE); }

}  // namespace yee
