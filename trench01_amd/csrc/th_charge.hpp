// Per-batch charge and field update on the GPU (after the push):
//
//   deposit   rho_k += (n_i,k - n_e,k) q / dx^2 on surface cells, conductor
//             cells excluded (trench_charging_2d.py:1239-1243)
//   leak      Q = rho dx^2 (+ dt G phi_ext);  Q' = M Q;  rho = Q' / dx^2   (:1246-1256)
//   driven    rho[conductor] = 0                                             (:1257-1258)
//   floating  drift = sum (rho - rho_prev) g_cond dx^2; dV = conductor_step;
//             rho[c_cell] += C (dV - drift) / dx^2                          (:1260-1266)
//   potential phi = G Q + phi_ext on the kept rows, G the Green's matrix
//             (one column per distinct surface response; the reference's
//             solve(-rho/eps0) restricted to the rows the push reads)       (:1268, :795-797);
//             FP64, or G split into two 16-bit halves (BF16 or scaled FP16)
//             with FP32 products and accumulation (precision variant)
//   field     E = -grad(phi) from vacuum cells only, packed with vac and phi
//             into the push's cell records                                  (:429-454, :1269)
//
// Charge state lives on the m surface cells: deposits, leak and the
// conductor's charge all land there, so rho is an m-vector.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace th {

constexpr int kUpdBlock = 1024;   // the m-vector update runs in one workgroup

struct ChargeDev {
    int m, n_cols, n_rows, nx, ny;
    const int32_t *col_of;      // [m] surface index -> Green's column, -1: none
    const int32_t *col_ptr;     // [n_cols + 1] column c's surface cells are col_cells[col_ptr[c] .. col_ptr[c+1])
    const int32_t *col_cells;   // [<= m] surface indices grouped by column, ascending within a column
    const int32_t *rows;        // [n_rows] flat cell of each kept row
    const int32_t *surf;        // [m] flat cell of each surface index
    const double *green;        // [n_rows][n_cols]
    const double *external;     // [N]
    const double *leak_M;       // [m][m] or null
    const double *leak_c;       // [m] or null
    const uint8_t *region;      // [m] leak-monitor region (bottom half of the trench)
    const uint8_t *cond_s;      // [m] conductor surface cells
    const double *g_cond_s;     // [m] conductor response at surface cells (floating)
    const uint8_t *vac;         // [N]
    const double *wface;        // [N]
    double *rho;                // [m] charge density [C/m^3]
    double *rho_prev;           // [m]
    double *qcol;               // [n_cols] charge per Green's column [C/m]
    double *phi;                // [N] potential (kept rows; others stay 0)
    double *diag;               // per-batch scalars, see kDiag*
    double4 *cells;             // push records {Ex*vac, Ey*vac, vac, phi}
    float4 *cells32;            // the same records in float for the FP32 push
    const uint32_t *counts;     // [2][m] deposit tallies of the batch
    // split Green's matrix: G ~ (hi + lo * lo_scale) * inv_scale, rows padded to n_pad (multiple of 8)
    const uint16_t *g_hi, *g_lo;
    int n_pad, green_mode;      // 0 FP64, 1 BF16 hi/lo, 2 FP16 hi/lo
    float lo_scale;
    double inv_scale;
    double q_macro, dx, C_cond, Te, dt_batch, top_voltage, substrate_voltage;
    int c_idx, floating, driven, has_leak, top_dirichlet, leak_has_c;
};

enum : int { kDiagLeakDQ = 0, kDiagDrift, kDiagDV, kDiagNi, kDiagNe, kDiagCount };

// numpy.logaddexp for finite arguments.
__device__ __forceinline__ double logaddexp_d(double a, double b) {
    if (a == b) return a + 0.693147180559945309417232121458176568;
    double t = a - b;
    return t > 0.0 ? a + log1p(exp(-t)) : b + log1p(exp(t));
}

// Exact one-batch solution of the floating-conductor ODE
// dv/dt = (q n_i/C + drift - q n_e/C exp(v/Te)) / dt (trench_charging_2d.py:815-830).
__device__ __forceinline__ double conductor_step_d(double n_i, double n_e, double q, double C, double Te,
                                                   double drift) {
    double x_i = q * n_i / C + drift, x_e = q * n_e / C;
    if (n_e == 0.0) return x_i;
    double k = x_i / Te;
    double frac = k != 0.0 ? -expm1(-k) / k : 1.0;
    return -Te * logaddexp_d(-k, log(x_e / Te * frac));
}

__device__ __forceinline__ double block_sum(double v, double *scratch) {
    // wave reduction then across the 16 waves of the workgroup
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += scratch[w];
        scratch[32] = s;
    }
    __syncthreads();
    return scratch[32];
}

// Deposit of one batch's tallies; with a leak operator also the leak's input
// q_new = rho dx^2 (+ dt G phi_ext) and q_old = rho dx^2 (the region outflow
// terms start from it).  q_old / q_new: LDS or global, m each.
__device__ __forceinline__ void charge_deposit(const ChargeDev &d, double *q_old, double *q_new) {
    const double dx2 = d.dx * d.dx;
    const int m = d.m;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        double r = d.rho[k];
        d.rho_prev[k] = r;
        double drho = (double)((int64_t)d.counts[k] - (int64_t)d.counts[m + k]) * d.q_macro / dx2;
        if (d.cond_s[k]) drho = 0.0;
        r += drho;
        d.rho[k] = r;
        if (d.has_leak) {
            const double q = r * dx2;
            q_old[k] = q;
            q_new[k] = d.leak_has_c ? q + d.leak_c[k] : q;
        }
    }
}

// Leak (backward Euler over the surface conductance network) for rows k =
// first, first + step, ...: one wave per row of M, lanes read the row
// coalesced, partial sums combined by a fixed shuffle tree.  Row k's region
// outflow term replaces q_old[k].  q_new: all m entries, read only.
__device__ __forceinline__ void leak_rows(const ChargeDev &d, const double *q_new, double *q_old, int first,
                                          int step) {
    const double dx2 = d.dx * d.dx;
    const int m = d.m, lane = threadIdx.x & 63;
    for (int k = first; k < m; k += step) {
        const double *row = d.leak_M + (int64_t)k * m;
        double acc = 0.0;
        for (int j = lane; j < m; j += 64) acc += row[j] * q_new[j];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
        if (lane == 0) {
            d.rho[k] = acc / dx2;
            q_old[k] = d.region[k] ? q_old[k] - acc : 0.0;   // region outflow terms
        }
    }
}

// The rest of the update after the leak (one workgroup): the leak monitor
// (sum of the region outflow terms in q_old, global or LDS), a driven
// conductor, the floating conductor's step, and the charge per Green's
// column.  sh: LDS for 2m doubles; scratch: 40 doubles of LDS.
__device__ __forceinline__ void charge_finish(const ChargeDev &d, const double *outflow, double *sh, double *scratch) {
    const double dx2 = d.dx * d.dx;
    const int m = d.m;
    double leak_dq = 0.0;
    if (d.has_leak) {
        double part = 0.0;
        for (int k = threadIdx.x; k < m; k += blockDim.x) part += outflow[k];
        leak_dq = block_sum(part, scratch);
    }
    if (d.driven)
        for (int k = threadIdx.x; k < m; k += blockDim.x)
            if (d.cond_s[k]) d.rho[k] = 0.0;
    __syncthreads();
    if (d.floating) {
        double part = 0.0, ni = 0.0, ne = 0.0;
        for (int k = threadIdx.x; k < m; k += blockDim.x) {
            part += (d.rho[k] - d.rho_prev[k]) * d.g_cond_s[k];
            if (d.cond_s[k]) {
                ni += (double)d.counts[k];
                ne += (double)d.counts[m + k];
            }
        }
        double drift = block_sum(part, scratch) * dx2;
        double n_i = block_sum(ni, scratch), n_e = block_sum(ne, scratch);
        if (threadIdx.x == 0) {
            double dv = conductor_step_d(n_i, n_e, d.q_macro, d.C_cond, d.Te, drift);
            d.rho[d.c_idx] += d.C_cond * (dv - drift) / dx2;
            d.diag[kDiagDrift] = drift;
            d.diag[kDiagDV] = dv;
            d.diag[kDiagNi] = n_i;
            d.diag[kDiagNe] = n_e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) d.diag[kDiagLeakDQ] = leak_dq;
    // charge per Green's column, one wave per column: lanes stride the
    // column's surface cells (one for a dielectric cell, all of the floating
    // conductor's for its column), partial sums combined by a fixed shuffle tree
    double *rho_s = sh;
    for (int k = threadIdx.x; k < m; k += blockDim.x) rho_s[k] = d.rho[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    for (int c = threadIdx.x >> 6; c < d.n_cols; c += waves) {
        double acc = 0.0;
        for (int q = d.col_ptr[c] + lane; q < d.col_ptr[c + 1]; q += 64) acc += rho_s[d.col_cells[q]] * dx2;
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
        if (lane == 0) d.qcol[c] = acc;
    }
}

// Deposit, leak, conductor and column charges of one batch over the m-vector
// (one workgroup).
__device__ __forceinline__ void charge_update(const ChargeDev &d, double *sh, double *scratch) {
    double *q_old = sh, *q_new = sh + d.m;
    charge_deposit(d, q_old, q_new);
    __syncthreads();
    if (d.has_leak) {
        leak_rows(d, q_new, q_old, threadIdx.x >> 6, blockDim.x >> 6);
        __syncthreads();
    }
    charge_finish(d, q_old, sh, scratch);
}

// The same update in three launches when there is a leak operator: the m x m
// GEMV streams M (2.2 MB at 520 surface cells) with one wave per row over many
// workgroups instead of one workgroup's latency-bound reads.  q: global
// scratch of 2m doubles (q_old, q_new).
__global__ void __launch_bounds__(kUpdBlock) charge_deposit_kernel(ChargeDev d, double *q) {
    charge_deposit(d, q, q + d.m);
}
__global__ void __launch_bounds__(256) leak_rows_kernel(ChargeDev d, double *q) {
    extern __shared__ double qn[];
    for (int k = threadIdx.x; k < d.m; k += blockDim.x) qn[k] = q[d.m + k];
    __syncthreads();
    leak_rows(d, qn, q, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), gridDim.x * (blockDim.x >> 6));
}
__global__ void __launch_bounds__(kUpdBlock) charge_finish_kernel(ChargeDev d, const double *q) {
    extern __shared__ double sh[];
    __shared__ double scratch[40];
    charge_finish(d, q, sh, scratch);
}

__global__ void __launch_bounds__(kUpdBlock) charge_update_kernel(ChargeDev d) {
    extern __shared__ double sh[];          // q_old[m], q_new[m]
    __shared__ double scratch[40];
    charge_update(d, sh, scratch);
}

// Replicas of one batch from one charge state: workgroup r applies the update
// of replica r's tallies to its own copy of the charge, so each replica's
// result equals that of charge_update_kernel run on its tallies alone.
struct ReplicaDev {
    const uint32_t *counts;     // [R][2][m]
    double *rho, *rho_prev;     // [R][m]
    double *diag;               // [R][kDiagCount]
    double *qcol;               // [R][n_cols]
};

__global__ void __launch_bounds__(kUpdBlock) charge_update_replicas_kernel(ChargeDev d, ReplicaDev r) {
    extern __shared__ double sh[];
    __shared__ double scratch[40];
    const int b = blockIdx.x;
    ChargeDev e = d;
    e.rho = r.rho + (int64_t)b * d.m;
    e.rho_prev = r.rho_prev + (int64_t)b * d.m;
    e.counts = r.counts + (int64_t)b * 2 * d.m;
    e.diag = r.diag + (int64_t)b * kDiagCount;
    e.qcol = r.qcol + (int64_t)b * d.n_cols;
    for (int k = threadIdx.x; k < d.m; k += blockDim.x) e.rho[k] = d.rho[k];
    __syncthreads();
    charge_update(e, sh, scratch);
}

// Per-replica tallies from the push's per-particle end records: particles
// [offset + r n, offset + (r+1) n) belong to replica r (blockIdx.y).  Counts
// [R][2][m] per surface cell; scal [R][4] = {escaped ions, escaped electrons,
// lost, absorbed off the surface cells}.  Both are cleared by the caller.
__global__ void __launch_bounds__(256) replica_tally_kernel(const int8_t *__restrict__ status,
                                                            const int32_t *__restrict__ hit,
                                                            const int8_t *__restrict__ spc,
                                                            const int32_t *__restrict__ smap, int m, int64_t offset,
                                                            int64_t n, uint32_t *counts, unsigned long long *scal) {
    extern __shared__ uint32_t lds[];   // [2][m]
    __shared__ unsigned long long lscal[4];
    const int r = blockIdx.y;
    for (int k = threadIdx.x; k < 2 * m; k += blockDim.x) lds[k] = 0u;
    if (threadIdx.x < 4) lscal[threadIdx.x] = 0ull;
    __syncthreads();
    const int64_t base = offset + (int64_t)r * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = base + i;
        const int st = status[q];
        const int sp = spc[q] != 0;
        if (st == 1) {
            const int k = smap[hit[q]];
            if (k >= 0) atomicAdd(&lds[sp * m + k], 1u);
            else atomicAdd(&lscal[3], 1ull);
        } else if (st == 2) {
            atomicAdd(&lscal[sp], 1ull);
        } else {
            atomicAdd(&lscal[2], 1ull);
        }
    }
    __syncthreads();
    uint32_t *out = counts + (int64_t)r * 2 * m;
    for (int k = threadIdx.x; k < 2 * m; k += blockDim.x)
        if (lds[k]) atomicAdd(&out[k], lds[k]);
    if (threadIdx.x < 4 && lscal[threadIdx.x]) atomicAdd(&scal[4 * r + threadIdx.x], lscal[threadIdx.x]);
}

// phi[rows[r]] = sum_c G[r][c] qcol[c] + external[rows[r]]: one wave per row,
// lanes stride the row (coalesced), column charges staged in LDS.
__global__ void __launch_bounds__(256) green_gemv_kernel(ChargeDev d) {
    extern __shared__ double q[];
    for (int c = threadIdx.x; c < d.n_cols; c += blockDim.x) q[c] = d.qcol[c];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int waves = blockDim.x >> 6;
    for (int r = blockIdx.x * waves + (threadIdx.x >> 6); r < d.n_rows; r += gridDim.x * waves) {
        const double *g = d.green + (int64_t)r * d.n_cols;
        double acc = 0.0;
        for (int c = lane; c < d.n_cols; c += 64) acc += g[c] * q[c];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
        if (lane == 0) {
            int cell = d.rows[r];
            d.phi[cell] = acc + d.external[cell];
        }
    }
}

// Adjoint of the potential update phi = G Q + phi_ext: the gradient of
// sum_cells w * phi with respect to the column charges, G^T w over the kept
// rows.  Workgroup b sums rows [b R, (b+1) R) for every column into
// partial[b][c] (coalesced: consecutive threads read consecutive columns of a
// row); green_adjoint_sum_kernel adds the slices in order, so the result does
// not depend on the launch.
__global__ void __launch_bounds__(256) green_adjoint_partial_kernel(ChargeDev d, const double *__restrict__ w,
                                                                     int rows_per_block, double *partial) {
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(r0 + rows_per_block, d.n_rows);
    for (int c = threadIdx.x; c < d.n_cols; c += blockDim.x) {
        double acc = 0.0;
        for (int r = r0; r < r1; ++r) acc += d.green[(int64_t)r * d.n_cols + c] * w[d.rows[r]];
        partial[(int64_t)blockIdx.x * d.n_cols + c] = acc;
    }
}
__global__ void green_adjoint_sum_kernel(const double *__restrict__ partial, int blocks, int n_cols, double *out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cols) return;
    double acc = 0.0;
    for (int b = 0; b < blocks; ++b) acc += partial[(int64_t)b * n_cols + c];
    out[c] = acc;
}

enum : int { kGreenF64 = 0, kGreenBF16 = 1, kGreenF16 = 2, kGreenBF16Mfma = 3, kGreenF16Mfma = 4 };

__device__ __forceinline__ float bf16_to_float(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ uint32_t float_to_bf16(float f) {   // round to nearest even (finite f)
    uint32_t u = __float_as_uint(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
template <int MODE>
__device__ __forceinline__ float half_to_float(uint32_t h) {
    if constexpr (MODE == kGreenBF16) {
        return bf16_to_float(h);
    } else {
        _Float16 v;
        uint16_t b = (uint16_t)h;
        __builtin_memcpy(&v, &b, 2);
        return (float)v;
    }
}

// Splits G (scaled by `scale`) into hi + lo * lo_scale in 16-bit halves.
template <int MODE>
__global__ void green_split_kernel(const double *__restrict__ green, int n_rows, int n_cols, int n_pad, double scale,
                                   float lo_scale, uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n_rows * n_pad) return;
    int r = (int)(idx / n_pad), c = (int)(idx - (int64_t)r * n_pad);
    double g = c < n_cols ? green[(int64_t)r * n_cols + c] * scale : 0.0;
    uint16_t h, l;
    if constexpr (MODE == kGreenBF16) {
        h = (uint16_t)float_to_bf16((float)g);
        l = (uint16_t)float_to_bf16((float)((g - (double)bf16_to_float(h)) / (double)lo_scale));
    } else {
        _Float16 hv = (_Float16)(float)g;
        _Float16 lv = (_Float16)(float)((g - (double)(float)hv) / (double)lo_scale);
        __builtin_memcpy(&h, &hv, 2);
        __builtin_memcpy(&l, &lv, 2);
    }
    hi[idx] = h;
    lo[idx] = l;
}

// phi[rows[r]] = inv_scale * sum_c (hi + lo * lo_scale)[r][c] qcol[c] + external:
// 16 lanes per row, 16-byte loads of eight halves, FP32 products and
// accumulation, the row sum returned to FP64 for the external potential.
template <int MODE>
__global__ void __launch_bounds__(256) green_gemv_split_kernel(ChargeDev d) {
    extern __shared__ float qf[];
    for (int c = threadIdx.x; c < d.n_pad; c += blockDim.x) qf[c] = c < d.n_cols ? (float)d.qcol[c] : 0.0f;
    __syncthreads();
    const int sub = threadIdx.x & 15;
    const int rows_per_block = blockDim.x >> 4;
    const int vecs = d.n_pad >> 3;
    for (int r = blockIdx.x * rows_per_block + (threadIdx.x >> 4); r < d.n_rows; r += gridDim.x * rows_per_block) {
        const uint4 *h = reinterpret_cast<const uint4 *>(d.g_hi + (int64_t)r * d.n_pad);
        const uint4 *l = reinterpret_cast<const uint4 *>(d.g_lo + (int64_t)r * d.n_pad);
        float acc = 0.0f;
        for (int v = sub; v < vecs; v += 16) {
            uint4 hv = h[v], lv = l[v];
            const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w}, lw[4] = {lv.x, lv.y, lv.z, lv.w};
            const float *q = qf + 8 * v;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float g0 = half_to_float<MODE>(hw[k] & 0xffffu) + half_to_float<MODE>(lw[k] & 0xffffu) * d.lo_scale;
                float g1 = half_to_float<MODE>(hw[k] >> 16) + half_to_float<MODE>(lw[k] >> 16) * d.lo_scale;
                acc = fmaf(g0, q[2 * k], acc);
                acc = fmaf(g1, q[2 * k + 1], acc);
            }
        }
        for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 16);
        if (sub == 0) {
            int cell = d.rows[r];
            d.phi[cell] = (double)acc * d.inv_scale + d.external[cell];
        }
    }
}

// The split Green's matrix on matrix cores: Phi = G Q for 1 to 16 charge
// vectors Q (columns of qcols [R][n_cols], C/m), with G = (hi + lo lo_scale)
// inv_scale as in green_gemv_split_kernel and each charge split into two
// halves of the same 16-bit type (BF16, or FP16 after scaling the charges by a
// power of two into its range): three MFMA products per K block
// (hi*hi + hi*lo + lo*hi, FP32 accumulation; lo*lo is below FP32's last bit).
// v_mfma_f32_16x16x32_{bf16,f16}: lane l holds A[l%16][8(l/16)+e],
// B[8(l/16)+e][l%16], C[4(l/16)+v][l%16].  A wave computes 16 rows x 16
// right-hand sides; G streams from HBM in 16-byte loads per lane (n_pad is a
// multiple of 32), the charge halves sit transposed in LDS.  batch: the batch
// update (R = 1), phi[rows[r]] = G q + external; else out[j][r] = (G Q)[r][j].
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int MODE>
__device__ __forceinline__ void split_half(float v, uint16_t &hi, uint16_t &lo) {
    if constexpr (MODE == kGreenBF16) {
        hi = (uint16_t)float_to_bf16(v);
        lo = (uint16_t)float_to_bf16(v - bf16_to_float(hi));
    } else {
        _Float16 h = (_Float16)v;
        _Float16 l = (_Float16)(v - (float)h);
        __builtin_memcpy(&hi, &h, 2);
        __builtin_memcpy(&lo, &l, 2);
    }
}

template <int MODE>
__device__ __forceinline__ f32x4 mfma16(s16x8 a, s16x8 b, f32x4 c) {
    if constexpr (MODE == kGreenBF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
}

template <int MODE>
__global__ void __launch_bounds__(256) green_mfma_kernel(ChargeDev d, const double *__restrict__ qcols, int R,
                                                         double *__restrict__ out, int batch) {
    extern __shared__ __align__(16) uint16_t bsh[];   // [2][16][K]: charge halves hi, lo, transposed (row j = right-hand side j)
    __shared__ float red[4];
    const int K = d.n_pad, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nq = R * d.n_cols;
    // FP16: charges scaled by 2^(14 - ceil(log2 max|q|)) (the same in every workgroup)
    float qs = 1.0f;
    if constexpr (MODE == kGreenF16) {
        float mx = 0.0f;
        for (int i = tid; i < nq; i += blockDim.x) mx = fmaxf(mx, fabsf((float)qcols[i]));
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        if (lane == 0) red[wave] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        qs = mx > 0.0f ? ldexpf(1.0f, 14 - (int)ceilf(log2f(mx))) : 1.0f;
    }
    for (int idx = tid; idx < 16 * K; idx += blockDim.x) {
        const int j = idx / K, k = idx - j * K;
        const float v = (j < R && k < d.n_cols) ? (float)qcols[(int64_t)j * d.n_cols + k] * qs : 0.0f;
        uint16_t hi, lo;
        split_half<MODE>(v, hi, lo);
        bsh[idx] = hi;
        bsh[16 * K + idx] = lo;
    }
    __syncthreads();
    const int n_tiles = (d.n_rows + 15) >> 4;
    const double scale = d.inv_scale / (double)qs;
    for (int tile = blockIdx.x * 4 + wave; tile < n_tiles; tile += gridDim.x * 4) {
        const int r = min(tile * 16 + (lane & 15), d.n_rows - 1);   // rows past the end repeat the last (not stored)
        const s16x8 *gh = reinterpret_cast<const s16x8 *>(d.g_hi + (int64_t)r * K + 8 * (lane >> 4));
        const s16x8 *gl = reinterpret_cast<const s16x8 *>(d.g_lo + (int64_t)r * K + 8 * (lane >> 4));
        const s16x8 *bh = reinterpret_cast<const s16x8 *>(bsh + (lane & 15) * K + 8 * (lane >> 4));
        const s16x8 *bl = reinterpret_cast<const s16x8 *>(bsh + 16 * K + (lane & 15) * K + 8 * (lane >> 4));
        f32x4 acc1 = {0.0f, 0.0f, 0.0f, 0.0f}, acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kb = 0; kb < (K >> 5); ++kb) {   // 32 columns per block: 4 vectors of 8 halves
            const s16x8 ah = gh[4 * kb], al = gl[4 * kb], b_hi = bh[4 * kb], b_lo = bl[4 * kb];
            acc1 = mfma16<MODE>(ah, b_hi, acc1);
            acc1 = mfma16<MODE>(ah, b_lo, acc1);
            acc2 = mfma16<MODE>(al, b_hi, acc2);
        }
        const int j = lane & 15;
        if (j < R) {
            for (int v = 0; v < 4; ++v) {
                const int row = tile * 16 + 4 * (lane >> 4) + v;
                if (row >= d.n_rows) break;
                const double val = (double)(acc1[v] + d.lo_scale * acc2[v]) * scale;
                if (batch) {
                    const int cell = d.rows[row];
                    d.phi[cell] = val + d.external[cell];
                } else {
                    out[(int64_t)j * d.n_rows + row] = val;
                }
            }
        }
    }
}

// Vacuum-only gradient and packing of the push records (trench_charging_2d.py:429-454).
__global__ void field_pack_kernel(ChargeDev d) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int N = d.nx * d.ny;
    if (c >= N) return;
    const int ny = d.ny, nx = d.nx;
    const int ix = c / ny, iy = c - ix * ny;
    const bool v = d.vac[c] != 0;
    const double dx = d.dx;
    const double pc = d.phi[c];
    double ex = 0.0, ey = 0.0;
    if (v) {
        int cw = ((ix + nx - 1) % nx) * ny + iy, ce = ((ix + 1) % nx) * ny + iy;
        bool vw = d.vac[cw] != 0, ve = d.vac[ce] != 0;
        double g = 0.0;
        if (vw && ve) g = (d.phi[ce] - d.phi[cw]) / (2 * dx);
        else if (ve) g = (d.phi[ce] - pc) / dx;
        else if (vw) g = (pc - d.phi[cw]) / dx;
        ex = -g;
        bool vs = iy > 0 && d.vac[c - 1] != 0;
        bool vn = iy < ny - 1 ? d.vac[c + 1] != 0 : true;
        double ps = iy > 0 ? d.phi[c - 1] : 2 * d.substrate_voltage - pc;
        double pn = iy < ny - 1 ? d.phi[c + 1] : (d.top_dirichlet ? 2 * d.top_voltage - pc : pc);
        g = 0.0;
        if (vs && vn) g = (pn - ps) / (2 * dx);
        else if (vn) g = (pn - pc) / dx;
        else if (vs) g = (pc - ps) / dx;
        ey = -g;
    }
    d.cells[c] = make_double4(v ? ex : 0.0, v ? ey : 0.0, v ? 1.0 : 0.0, pc);
    d.cells32[c] = make_float4(v ? (float)ex : 0.0f, v ? (float)ey : 0.0f, v ? 1.0f : 0.0f, (float)pc);
}

}  // namespace th

namespace th {

// Per-batch monitors: probe potentials, integrated charge per region, the
// absolute charge of one region, and the ion / electron counts on the trench
// floor (trench_charging_2d.py:1272-1285, :871-874).
struct MonitorDev {
    int n_probes, n_regions, abs_region, n_bottom;
    const int32_t *probe_cells;   // [n_probes] flat cells
    const int32_t *region_ptr;    // [n_regions + 1] CSR offsets into region_idx
    const int32_t *region_idx;    // surface indices
    const int32_t *bottom_s;      // [n_bottom] surface indices of the floor cells
    double *out;                  // [n_probes + n_regions + 3]
};

// End of a captured batch: its monitors, scalar tallies and update
// diagnostics into row (batch - first) of the run's history, then batch + 1.
// run = {batch, first}.  Row: monitors [n_mon], {escaped ions, escaped
// electrons, lost, steps, absorbed off the surface}, {leak_dq, drift, dv,
// n_i_cond, n_e_cond}.
__global__ void batch_record_kernel(const double *__restrict__ mon, int n_mon, const unsigned long long *__restrict__ scal,
                                    const double *__restrict__ diag, unsigned long long *run, double *hist, int row,
                                    int cap) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const unsigned long long b = run[0], r = b - run[1];
    if (r < (unsigned long long)cap) {
        double *o = hist + r * (unsigned long long)row;
        for (int k = 0; k < n_mon; ++k) o[k] = mon[k];
        o[n_mon + 0] = (double)scal[1];   // kScalEsc0
        o[n_mon + 1] = (double)scal[2];   // kScalEsc1
        o[n_mon + 2] = (double)scal[3];   // kScalLost
        o[n_mon + 3] = (double)scal[4];   // kScalSteps
        o[n_mon + 4] = (double)scal[5];   // kScalOff
        for (int k = 0; k < kDiagCount; ++k) o[n_mon + 5 + k] = diag[k];
    }
    run[0] = b + 1;
}

__global__ void __launch_bounds__(256) monitor_kernel(ChargeDev d, MonitorDev mo) {
    __shared__ double scratch[40];
    const double dx2 = d.dx * d.dx;
    for (int i = threadIdx.x; i < mo.n_probes; i += blockDim.x) mo.out[i] = d.phi[mo.probe_cells[i]];
    for (int r = 0; r < mo.n_regions; ++r) {
        double part = 0.0, apart = 0.0;
        for (int j = mo.region_ptr[r] + threadIdx.x; j < mo.region_ptr[r + 1]; j += blockDim.x) {
            double v = d.rho[mo.region_idx[j]];
            part += v;
            apart += fabs(v);
        }
        double sum = block_sum(part, scratch);
        double asum = r == mo.abs_region ? block_sum(apart, scratch) : 0.0;
        if (threadIdx.x == 0) {
            mo.out[mo.n_probes + r] = sum * dx2;
            if (r == mo.abs_region) mo.out[mo.n_probes + mo.n_regions] = asum * dx2;
        }
    }
    double ni = 0.0, ne = 0.0;
    for (int j = threadIdx.x; j < mo.n_bottom; j += blockDim.x) {
        ni += (double)d.counts[mo.bottom_s[j]];
        ne += (double)d.counts[d.m + mo.bottom_s[j]];
    }
    double si = block_sum(ni, scratch), se = block_sum(ne, scratch);
    if (threadIdx.x == 0) {
        mo.out[mo.n_probes + mo.n_regions + 1] = si;
        mo.out[mo.n_probes + mo.n_regions + 2] = se;
    }
}

}  // namespace th
