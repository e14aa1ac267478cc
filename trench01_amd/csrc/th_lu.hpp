// Triangular solves with the sparse LU factors of the Poisson operator, for a
// block of right-hand sides: the Green's columns (one Poisson solve per
// surface-charge column) computed in FP64 on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace th {

constexpr int kLuLanes = 64;   // right-hand sides per workgroup: one wave, one lane each

__device__ inline double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int kLuGroup = 32;   // column loads issued before the first of them is used

// One chunk (up to 64 entries) of a row: acc minus val * x[idx] over the
// chunk, subtracted in order.  The chunk's (index, value) pairs sit one per
// lane (my_i, my_v) and are broadcast with readlane; the x loads of a group
// go out together, so a chunk costs one or two memory latencies instead of
// one per entry.  Lanes past the chunk hold index 0 (a valid row), and their
// loads are not used.
__device__ inline double chunk_update(double acc, int cnt, int my_i, double my_v, const double *xc, int ld) {
#pragma unroll
    for (int g = 0; g < kLuLanes; g += kLuGroup) {
        if (g >= cnt) break;
        double xv[kLuGroup];
#pragma unroll
        for (int t = 0; t < kLuGroup; ++t) xv[t] = xc[(size_t)__builtin_amdgcn_readlane(my_i, g + t) * ld];
#pragma unroll
        for (int t = 0; t < kLuGroup; ++t)
            if (g + t < cnt) acc -= readlane_f64(my_v, g + t) * xv[t];
    }
    return acc;
}

struct LuChunk {
    int cnt, my_i;
    double my_v;
};

__device__ inline LuChunk load_chunk(int base, int end, const int *__restrict__ idx, const double *__restrict__ val) {
    const int lane = threadIdx.x;
    const int cnt = min(kLuLanes, end - base);
    return LuChunk{cnt, lane < cnt ? idx[base + lane] : 0, lane < cnt ? val[base + lane] : 0.0};
}

// One level of a triangle: rows[blockIdx.x] for the 64 columns of slab
// blockIdx.y.  A level's rows depend only on rows of earlier levels, which
// earlier launches on the same stream have finished; x[i] becomes
// x[i] - sum over the row of val * x[idx], divided by diag[i] for U (diag
// nullptr: L's unit diagonal).  The host builds the levels from the
// validated factors, so every row and index is in [0, n).
__global__ __launch_bounds__(kLuLanes) void lu_level_kernel(const int *__restrict__ rows, const int *__restrict__ ptr,
                                                            const int *__restrict__ idx,
                                                            const double *__restrict__ val,
                                                            const double *__restrict__ diag, int ld, double *x) {
    const int i = rows[blockIdx.x];
    double *xc = x + (size_t)blockIdx.y * kLuLanes + threadIdx.x;
    const int q1 = ptr[i + 1];
    double acc = xc[(size_t)i * ld];
    for (int base = ptr[i]; base < q1; base += kLuLanes) {
        const LuChunk c = load_chunk(base, q1, idx, val);
        acc = chunk_update(acc, c.cnt, c.my_i, c.my_v, xc, ld);
    }
    xc[(size_t)i * ld] = diag ? acc / diag[i] : acc;
}

}  // namespace th
