// Cost of a grid-wide barrier inside one cooperative launch (agent-scope release/acquire on a counter),
// against the launch-to-launch floor: the measurement behind keeping one kernel launch per half-iteration
// rather than a persistent kernel with grid barriers.  Built by ganit_v2_amd/build.py as bin/ganit_grid_barrier.  Every spin is bounded: a barrier that does not complete within the
// limit records a failure and moves on, so the kernel always exits.  Data written before each barrier is
// read back after it by a workgroup on another XCD (double-buffered slots) and checked.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); return 1; } } while (0)

__global__ __launch_bounds__(256) void k_bar(unsigned* ctr, int nbar, int limit, int* fail, float* buf, int per_wg,
                                             int* bad) {
    const int G = gridDim.x, b = blockIdx.x;
    const int nb = (b + 1) % G;  // next workgroup: another XCD when G is a multiple of 8
    __shared__ int stop;
    for (int k = 0; k < nbar; ++k) {
        if (threadIdx.x == 0) stop = __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (stop) return;  // a barrier timed out somewhere: every block leaves
        float* mine = buf + ((size_t)(k & 1) * G + b) * per_wg;
        for (int i = threadIdx.x; i < per_wg; i += 256) mine[i] = (float)(k * 8 + (b & 7));
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = (unsigned)(k + 1) * (unsigned)G;
            int s = 0;
            while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++s > limit) {
                    atomicAdd(fail, 1);
                    break;
                }
            }
        }
        __syncthreads();
        const float* theirs = buf + ((size_t)(k & 1) * G + nb) * per_wg;
        int wrong = 0;
        for (int i = threadIdx.x; i < per_wg; i += 256) wrong += theirs[i] != (float)(k * 8 + (nb & 7));
        if (wrong) atomicAdd(bad, wrong);
    }
}

__global__ void k_empty(int* p) { if (threadIdx.x == 1023) p[0] = 1; }

int main() {
    int dev = 0, cus = 0, coop = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    int per_cu = 0;
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bar, 256, 0));
    printf("CUs %d cooperative %d max resident k_bar blocks per CU %d\n", cus, coop, per_cu);
    if (!coop) return 0;
    unsigned* ctr; int *fail, *bad; float* buf;
    const int maxg = cus * per_cu, max_per = 2048;
    CK(hipMalloc(&ctr, 256));
    CK(hipMalloc(&fail, 4));
    CK(hipMalloc(&bad, 4));
    CK(hipMalloc(&buf, sizeof(float) * 2 * (size_t)maxg * max_per));
    hipStream_t st; CK(hipStreamCreate(&st));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    // launch floor for reference
    for (int i = 0; i < 10; ++i) hipLaunchKernelGGL(k_empty, dim3(1024), dim3(256), 0, st, fail);
    CK(hipEventRecord(e0, st));
    for (int i = 0; i < 500; ++i) hipLaunchKernelGGL(k_empty, dim3(1024), dim3(256), 0, st, fail);
    CK(hipEventRecord(e1, st));
    CK(hipEventSynchronize(e1));
    float ms = 0; CK(hipEventElapsedTime(&ms, e0, e1));
    printf("empty kernel, 1024 blocks: %.2f us per launch\n", ms * 1e3 / 500);
    const int limit = 1 << 16, nbar = 2000;
    for (int G : {256, 512, 1024}) {
        if (G > maxg) continue;
        for (int per : {0, 512, 2048}) {
            float best = 1e30f;
            int f_h = 0, b_h = 0;
            for (int rep = 0; rep < 3; ++rep) {
                CK(hipMemsetAsync(ctr, 0, 4, st));
                CK(hipMemsetAsync(fail, 0, 4, st));
                CK(hipMemsetAsync(bad, 0, 4, st));
                int nbar_v = nbar, lim = limit, per_v = per;
                void* args[] = {&ctr, &nbar_v, &lim, &fail, &buf, &per_v, &bad};
                CK(hipEventRecord(e0, st));
                CK(hipLaunchCooperativeKernel((const void*)k_bar, dim3(G), dim3(256), args, 0, st));
                CK(hipEventRecord(e1, st));
                CK(hipEventSynchronize(e1));
                CK(hipEventElapsedTime(&ms, e0, e1));
                best = ms < best ? ms : best;
                int f1, b1;
                CK(hipMemcpy(&f1, fail, 4, hipMemcpyDeviceToHost));
                CK(hipMemcpy(&b1, bad, 4, hipMemcpyDeviceToHost));
                f_h += f1; b_h += b1;
            }
            printf("grid %5d, %5d floats per block: %.2f us per barrier step (spin failures %d, wrong reads %d)\n", G,
                   per, best * 1e3 / nbar, f_h, b_h);
            fflush(stdout);
        }
    }
    CK(hipDeviceSynchronize());
    return 0;
}
