// Streaming ceiling probe (not part of libyee.so): three float32 input arrays to three output arrays, the read/write
// mix of a Yee step (every state array read once and written once), with 4, 8 or 16 B per lane and several grid
// sizes. It measures the HBM bandwidth a plain streaming kernel reaches on this device, the practical roofline the
// Yee kernels are compared against.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(call)                                                                           \
    do {                                                                                      \
        hipError_t err_ = (call);                                                             \
        if (err_ != hipSuccess) {                                                             \
            std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(err_));          \
            std::exit(1);                                                                     \
        }                                                                                     \
    } while (0)

template <int V> struct Lanes;
template <> struct Lanes<1> { typedef float T; };
template <> struct Lanes<2> { typedef float2 T; };
template <> struct Lanes<4> { typedef float4 T; };

__device__ inline float mix(float a, float b) { return a * 1.0001f + b; }
__device__ inline float2 mix(float2 a, float2 b) { return make_float2(mix(a.x, b.x), mix(a.y, b.y)); }
__device__ inline float4 mix(float4 a, float4 b)
{
    return make_float4(mix(a.x, b.x), mix(a.y, b.y), mix(a.z, b.z), mix(a.w, b.w));
}

// out_x = f(a, b), out_y = f(b, c), out_z = f(c, a) over n_vec elements of V floats (grid-stride loop).
template <int V>
__global__ void __launch_bounds__(256) stream3(const float* a, const float* b, const float* c, float* x, float* y,
                                               float* z, long n_vec)
{
    typedef typename Lanes<V>::T T;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n_vec; i += stride) {
        const T va = reinterpret_cast<const T*>(a)[i];
        const T vb = reinterpret_cast<const T*>(b)[i];
        const T vc = reinterpret_cast<const T*>(c)[i];
        reinterpret_cast<T*>(x)[i] = mix(va, vb);
        reinterpret_cast<T*>(y)[i] = mix(vb, vc);
        reinterpret_cast<T*>(z)[i] = mix(vc, va);
    }
}

int main(int argc, char** argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : (1L << 28);   // floats per array (default 1 GiB each)
    const int repeat = 5;
    float* buf[6];
    for (auto& p : buf) {
        CHECK(hipMalloc(&p, n * 4));
        CHECK(hipMemset(p, 0, n * 4));
    }
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    double best = 0.0;
    for (int blocks : {1024, 4096, 16384}) {
        for (int v : {1, 2, 4}) {
            const long n_vec = n / v;
            auto launch = [&]() {
                if (v == 1)
                    hipLaunchKernelGGL(stream3<1>, dim3(blocks), dim3(256), 0, 0, buf[0], buf[1], buf[2], buf[3],
                                       buf[4], buf[5], n_vec);
                else if (v == 2)
                    hipLaunchKernelGGL(stream3<2>, dim3(blocks), dim3(256), 0, 0, buf[0], buf[1], buf[2], buf[3],
                                       buf[4], buf[5], n_vec);
                else
                    hipLaunchKernelGGL(stream3<4>, dim3(blocks), dim3(256), 0, 0, buf[0], buf[1], buf[2], buf[3],
                                       buf[4], buf[5], n_vec);
            };
            launch();
            CHECK(hipDeviceSynchronize());
            CHECK(hipEventRecord(e0));
            for (int r = 0; r < repeat; ++r) launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= repeat;
            const double gbs = 6.0 * n * 4 / (ms * 1e-3) / 1e9;
            if (gbs > best) best = gbs;
            std::printf("blocks %6d, %2d B per lane: %.3f ms per launch, %.1f GB/s\n", blocks, 4 * v, ms, gbs);
        }
    }
    std::printf("best: %.1f GB/s (%.1f %% of 8000 GB/s) for 3 x %ld B read + 3 x %ld B written per launch\n", best,
                best / 80.0, n * 4, n * 4);
    for (auto& p : buf) CHECK(hipFree(p));
    return 0;
}
