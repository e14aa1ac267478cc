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

// Solves one triangle in place, rows in dependency order (first, first +
// step, ... for n rows).  The first chunk of the next row is loaded while the
// current row is being reduced; rows longer than a chunk load the rest as
// they go.
template <bool UPPER>
__device__ inline void triangle(int n, const int *__restrict__ ptr, const int *__restrict__ idx,
                                const double *__restrict__ val, const double *__restrict__ diag, double *xc,
                                int ld) {
    int i = UPPER ? n - 1 : 0;
    LuChunk c = load_chunk(ptr[i], ptr[i + 1], idx, val);
    for (int r = 0; r < n; ++r, i += UPPER ? -1 : 1) {
        const int q0 = ptr[i], q1 = ptr[i + 1];
        const int j = UPPER ? i - 1 : i + 1;
        const LuChunk next = r + 1 < n ? load_chunk(ptr[j], ptr[j + 1], idx, val) : LuChunk{0, 0, 0.0};
        double acc = chunk_update(xc[(size_t)i * ld], c.cnt, c.my_i, c.my_v, xc, ld);
        for (int base = q0 + kLuLanes; base < q1; base += kLuLanes) {
            const LuChunk more = load_chunk(base, q1, idx, val);
            acc = chunk_update(acc, more.cnt, more.my_i, more.my_v, xc, ld);
        }
        xc[(size_t)i * ld] = UPPER ? acc / diag[i] : acc;
        c = next;
    }
}

// x holds n rows of ld right-hand-side values (row-major, ld a multiple of
// kLuLanes) and is overwritten by U^-1 L^-1 x.  L is unit lower triangular
// (its strictly lower part in CSR: lp, li, lv), U upper triangular (strictly
// upper part in CSR: up, ui, uv; diagonal ud).  Each lane owns one column of
// x and walks the rows in dependency order, so every x[j] it reads is final
// and was written by the lane itself.  The host checks that li[q] < i and
// ui[q] > i for the entries of row i, which also keeps every access in
// bounds.
__global__ __launch_bounds__(kLuLanes) void lu_solve_kernel(int n, const int *__restrict__ lp,
                                                            const int *__restrict__ li,
                                                            const double *__restrict__ lv,
                                                            const int *__restrict__ up,
                                                            const int *__restrict__ ui,
                                                            const double *__restrict__ uv,
                                                            const double *__restrict__ ud, int ld, double *x) {
    double *xc = x + (size_t)blockIdx.x * kLuLanes + threadIdx.x;
    triangle<false>(n, lp, li, lv, nullptr, xc, ld);
    triangle<true>(n, up, ui, uv, ud, xc, ld);
}

}  // namespace th
