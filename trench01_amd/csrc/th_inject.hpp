// Device-side injection of one batch at the top boundary.
//
// Same distributions as the host injection (trench_charging_2d.py:458-489):
// n ions then n electrons at y = Ly - 1e-3 dx, x uniform over the period;
// electrons Maxwellian, flux-weighted in the normal direction (Rayleigh
// speed); ions from the chosen energy distribution.  The random stream is
// Philox4x32-10 (J. K. Salmon et al., "Parallel random numbers: as easy as
// 1, 2, 3", SC'11) keyed by the run seed and counted by (particle, batch), so
// a batch is reproducible and independent of launch shape -- but it is a
// different stream from the host's PCG64, so results agree with the host
// path statistically, not draw for draw.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace th {

struct Philox {
    static __device__ __forceinline__ void round(uint32_t c[4], uint32_t k[2]) {
        const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
        uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k[0], n2 = hi0 ^ c[3] ^ k[1];
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
    static __device__ __forceinline__ void block(uint32_t c[4], uint64_t seed) {
        uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        for (int r = 0; r < 10; ++r) {
            round(c, k);
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
    }
};

// Uniform in (0, 1) from two 32-bit words (53 random bits, centred in its bin).
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    uint64_t v = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    v &= (1ull << 53) - 1ull;
    return ((double)v + 0.5) * 0x1p-53;
}

struct InjectDev {
    int64_t n;              // particles per species
    double lx, y0;          // period, injection height
    double qm_i, qm_e;      // charge/mass of ions and electrons
    int ion_kind;           // 0 energy + lateral thermal normal, 1 mono + polar angle, 2 arcsine + polar angle
    double e_ion, e_lo, e_hi;        // [eV]
    double sigma_theta;     // polar-angle sigma [rad]
    double vth_i;           // lateral thermal speed sqrt(Ti e / m_i)
    double v_per_sqrt_ev;   // sqrt(2 e / m_i)
    double s_e;             // sqrt(Te e / m_e)
    uint64_t seed, batch;
    double *x, *y, *vx, *vy, *qm;
    int8_t *spc;
};

// Particle k of a batch (k < n: ions, then n electrons), drawn from the
// Philox block (k, batch, key).
__device__ __forceinline__ void draw_particle(const InjectDev &d, int64_t k, uint64_t key, double &x, double &vx,
                                              double &vy, double &qm, int8_t &sp) {
    uint32_t c0[4] = {(uint32_t)k, (uint32_t)(k >> 32), (uint32_t)d.batch, 0u};
    uint32_t c1[4] = {(uint32_t)k, (uint32_t)(k >> 32), (uint32_t)d.batch, 1u};
    Philox::block(c0, key);
    Philox::block(c1, key);
    double u0 = u01(c0[0], c0[1]), u1 = u01(c0[2], c0[3]);
    double u2 = u01(c1[0], c1[1]), u3 = u01(c1[2], c1[3]);
    // standard normal from (u2, u3), Box-Muller
    double gauss = sqrt(-2.0 * log(u2)) * cos(6.283185307179586476925286766559 * u3);
    x = u0 * d.lx;
    if (k < d.n) {
        sp = 0;
        qm = d.qm_i;
        if (d.ion_kind == 0) {
            vy = -d.v_per_sqrt_ev * sqrt(d.e_ion);
            vx = d.vth_i * gauss;
        } else {
            double e = d.ion_kind == 1 ? d.e_ion
                                       : d.e_lo + (d.e_hi - d.e_lo) * pow(sin(1.5707963267948966 * u1), 2.0);
            double v = d.v_per_sqrt_ev * sqrt(e), th = d.sigma_theta * gauss;
            vx = v * sin(th);
            vy = -v * cos(th);
        }
    } else {
        sp = 1;
        qm = d.qm_e;
        vx = d.s_e * gauss;
        vy = -d.s_e * sqrt(-2.0 * log(u1));
    }
}

__device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t batch) {
    return seed ^ (batch >> 32) * 0x9E3779B97F4A7C15ull;
}

// inject_kernel with the batch index read from device memory (captured
// batches advance it themselves).
__global__ void inject_at_kernel(InjectDev d, const unsigned long long *__restrict__ batch) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * d.n) return;
    d.batch = *batch;
    double x, vx, vy, qm;
    int8_t sp;
    draw_particle(d, k, stream_key(d.seed, d.batch), x, vx, vy, qm, sp);
    d.x[k] = x; d.y[k] = d.y0; d.vx[k] = vx; d.vy[k] = vy; d.qm[k] = qm; d.spc[k] = sp;
}

__global__ void inject_kernel(InjectDev d) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * d.n) return;
    double x, vx, vy, qm;
    int8_t sp;
    draw_particle(d, k, stream_key(d.seed, d.batch), x, vx, vy, qm, sp);
    d.x[k] = x; d.y[k] = d.y0; d.vx[k] = vx; d.vy[k] = vy; d.qm[k] = qm; d.spc[k] = sp;
}

}  // namespace th
