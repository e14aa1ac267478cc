// Small kernels shared by the captured batches: fills instead of memset nodes
// (a graph replays them reliably) and a device-side counter increment.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace th {

__global__ void zero_u64_kernel(unsigned long long *p, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        p[k] = 0ull;
}
__global__ void zero_u32_kernel(uint32_t *p, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        p[k] = 0u;
}
// Clears one batch's tallies and counters in one launch.
__global__ void clear_batch_kernel(uint32_t *a, int64_t na, uint32_t *b, int64_t nb, unsigned long long *c, int64_t nc,
                                   uint32_t *d, int64_t nd) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < na + nb + nc + nd; k += stride) {
        if (k < na) a[k] = 0u;
        else if (k < na + nb) b[k - na] = 0u;
        else if (k < na + nb + nc) c[k - na - nb] = 0ull;
        else d[k - na - nb - nc] = 0u;
    }
}
__global__ void bump_kernel(unsigned long long *p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p += 1ull;
}

}  // namespace th
