// FP64 helpers shared by the HIP push kernels.  Written as __host__ __device__
// so the identical source can also be compiled for the CPU and checked there
// against the C library.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define TH_HD __host__ __device__ __forceinline__
#else
#define TH_HD static inline
#endif

// Hypot core from C. F. Borges, "An Improved Algorithm for hypot(a,b)",
// arXiv:1904.09481 (2019): h = sqrt(a^2+b^2) followed by one correction step
// computed without FMA (Borges' non-fused variant).  This is the algorithm of
// glibc >= 2.35's binary64 hypot on x86-64, i.e. the function NumPy's np.hypot
// calls, so the push reproduces the reference's CPU path bit for bit.
// Requires ax >= ay >= 0 with squares that neither overflow nor underflow.
TH_HD double th_hypot_core(double ax, double ay) {
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        double d = h - ay;
        t1 = ax * (2.0 * d - ax);
        t2 = (d - 2.0 * (ax - ay)) * d;
    } else {
        double d = h - ax;
        t1 = 2.0 * d * (ax - 2.0 * ay);
        t2 = (4.0 * d - ay) * ay + d * d;
    }
    return h - (t1 + t2) / (2.0 * h);
}

// Full hypot: ordering, widely-separated operands, and power-of-two rescaling
// for very large (> 2^511) or very small (< 2^-459) operands, where the
// correction terms would overflow or lose bits to underflow (exact scalings).
TH_HD double th_hypot_full(double x, double y) {
    if (!isfinite(x) || !isfinite(y)) {
        if (isinf(x) || isinf(y)) return INFINITY;
        return x + y;
    }
    x = fabs(x);
    y = fabs(y);
    double ax = x < y ? y : x, ay = x < y ? x : y;
    // One core call for all three ranges: scaling by 2^-600 / 2^600 is exact,
    // and multiplying the result by 2^600 equals dividing it by 2^-600.
    double down = 1.0, up = 1.0;
    if (ax > 0x1p+511) {
        if (ay <= ax * 0x1p-54) return ax + ay;
        down = 0x1p-600;
        up = 0x1p+600;
    } else if (ay < 0x1p-459) {
        if (ax >= ay * 0x1p+54) return ax + ay;
        down = 0x1p+600;
        up = 0x1p-600;
    } else if (ay <= ax * 0x1p-54) {
        return ax + ay;
    }
    return th_hypot_core(ax * down, ay * down) * up;
}

// The same function with the common case first: operands in the core's range
// (no rescaling, not widely separated, finite, non-zero) go straight to the
// core, which is what th_hypot_full computes for them (its scalings are then
// by 1.0).  NaN fails every comparison and takes the full path.
TH_HD double th_hypot(double x, double y) {
    double fx = fabs(x), fy = fabs(y);
    double ax = fx < fy ? fy : fx, ay = fx < fy ? fx : fy;
    if (__builtin_expect(ax <= 0x1p+511 && ay >= 0x1p-459 && ay > ax * 0x1p-54, 1))
        return th_hypot_core(ax, ay);
    return th_hypot_full(x, y);
}

