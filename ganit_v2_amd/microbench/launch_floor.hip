// Launch floor and streaming-read rates of one GPU (the two numbers that bound a PDHG iteration made of two
// dependent kernels): back-to-back empty kernels at several grid sizes, and grid-stride / slab-per-block float4
// reads of 21, 32, 56 MB and 1 GiB, each timed with HIP events over repeated launches.  Every access is inside
// the one 1 GiB buffer by construction (grid-stride loops bounded by n4).
// Built by ganit_v2_amd/build.py as bin/ganit_launch_floor; run it with no arguments.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

__global__ void k_empty(float* out) { if (threadIdx.x == 1023) out[0] = 1.f; }

__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ p, size_t n4, float* out) {
    float s = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x) {
        float4 v = p[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == -1.2345f) out[0] = s;
}

// one pass per workgroup over a contiguous slab (like a row block), 8 float4 per thread unrolled
__global__ __launch_bounds__(256) void k_slab(const float4* __restrict__ p, size_t n4, float* out) {
    const size_t per = (n4 + gridDim.x - 1) / gridDim.x;
    const size_t b0 = blockIdx.x * per, b1 = b0 + per < n4 ? b0 + per : n4;
    float s = 0.f;
    for (size_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        float4 v = p[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == -1.2345f) out[0] = s;
}

int main() {
    const size_t maxb = (size_t)1 << 30;
    float4* p; float* out;
    CK(hipMalloc(&p, maxb));
    CK(hipMalloc(&out, 4096));
    CK(hipMemset(p, 0, maxb));
    hipStream_t st; CK(hipStreamCreate(&st));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    auto timeit = [&](auto launch, int reps) -> float {
        for (int i = 0; i < 5; ++i) launch();
        float ms = -1.f;  // a failed call leaves a negative time
        if (hipEventRecord(e0, st) != hipSuccess) return ms;
        for (int i = 0; i < reps; ++i) launch();
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return -1.f;
        return ms * 1000.f / reps;
    };
    for (int g : {1, 256, 1450, 4096}) {
        float us = timeit([&] { hipLaunchKernelGGL(k_empty, dim3(g), dim3(256), 0, st, out); }, 500);
        printf("empty grid %5d: %.2f us/launch\n", g, us);
    }
    for (size_t mb : {21, 32, 56, 1024}) {
        const size_t n4 = mb * 1000000 / 16;
        for (int g : {1024, 1450, 2048, 4096, 8192}) {
            float us = timeit([&] { hipLaunchKernelGGL(k_read, dim3(g), dim3(256), 0, st, p, n4, out); }, mb > 100 ? 20 : 300);
            float us2 = timeit([&] { hipLaunchKernelGGL(k_slab, dim3(g), dim3(256), 0, st, p, n4, out); }, mb > 100 ? 20 : 300);
            printf("read %4zu MB grid %5d: stride %.2f us (%.0f GB/s)  slab %.2f us (%.0f GB/s)\n", mb, g, us,
                   n4 * 16 / (us * 1e3), us2, n4 * 16 / (us2 * 1e3));
        }
    }
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    return 0;
}
