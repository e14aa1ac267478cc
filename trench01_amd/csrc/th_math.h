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

// Square root and quotient without the range steps of the IEEE sequences.
//
// On gfx950 the compiler evaluates a correctly rounded FP64 sqrt(x) as: scale
// x below 2^-767 by 2^256; r = rsq(x); g = x r, h = r / 2; one refinement
// e = 1/2 - h g, g += g e, h += h e; two corrections d = x - g g, g += d h;
// scale g back by 2^-128 if x was scaled; return x itself for +-0 and +inf.
// For x in [2^-767, DBL_MAX] both scalings are by 2^0 and the final select
// keeps g, so th_sqrt_n -- the same instructions without those steps --
// returns the same double, i.e. the correctly rounded square root.
//
// A correctly rounded a / b is evaluated as: d = div_scale(b), n =
// div_scale(a), each the operand itself or the operand scaled by 2^+-128;
// r = rcp(d) refined twice by r += r (1 - d r); q = n r; q = fma(n - d q, r,
// q), scaled back when n was scaled; then div_fixup, which gives the result
// its sign and handles zero, infinite and NaN operands.  div_scale scales
// only for a = 0 or b = 0 (then div_fixup alone decides the result), a
// denormal b, |b| >= 2^1022 (1/b denormal), |a| < 2^-969, a quotient in the
// denormal range, or an exponent of a exceeding that of b by 768 or more.
// Outside those cases th_div_n -- the same instructions without the two
// div_scale steps -- returns the same double; it is used only where the
// operands are bounded away from them (stated at each use).  The host
// versions are the library operations, which are correctly rounded too.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double th_sqrt_n(double x) {
    double r = __builtin_amdgcn_rsq(x);
    double g = x * r;
    double h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
__device__ __forceinline__ double th_div_n(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    double q = a * r;
    const double rem = fma(-b, q, a);
    q = fma(rem, r, q);
    return __builtin_amdgcn_div_fixup(q, b, a);
}
#else
TH_HD double th_sqrt_n(double x) { return sqrt(x); }
TH_HD double th_div_n(double a, double b) { return a / b; }
#endif

// The push's CFL step cfl_dx / (hv + sqrt(two_dx ha) + 1) with the square
// root and the division in their range-free forms where the operands allow
// (velocities and accelerations of any physical particle), else the library
// operations: the same double either way.
TH_HD double th_cfl_dt(double cfl_dx, double two_dx, double hv, double ha) {
    const double s = two_dx * ha;
    // With s in [2^-767, 2^1000] the root lies in [2^-384, 2^500]; with 0 <= hv
    // <= 2^799 the divisor lies in [1, 2^800), and with cfl_dx in [2^-100,
    // 2^100] the quotient in [2^-900, 2^100]: no operand is near a scaling case.
    if (s >= 0x1p-767 && s <= 0x1p+1000 && hv <= 0x1p+799 && cfl_dx >= 0x1p-100 && cfl_dx <= 0x1p+100)
        return th_div_n(cfl_dx, hv + th_sqrt_n(s) + 1.0);
    return cfl_dx / (hv + sqrt(s) + 1.0);
}

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

// th_hypot_core with the range-free square root and quotient, for 2^-383 <=
// ay <= ax <= 2^511 and ay > ax 2^-54: the square's argument lies in
// [2^-766, 2^1023]; the correction's divisor 2h in [2^-382, 2^513]; its
// dividend t1 + t2 is at most h^2 2^-50 (the correction is below one ulp of
// h) and is 0 or at least h^2 2^-160 (a sum of rounded products of operands
// no smaller than ay >= h 2^-54 is a multiple of one ulp of ay^2), so the
// quotient lies in [h 2^-161, h 2^-51] and no operand is near a scaling case.
// Both correction forms are evaluated and Borges' test selects one (no branch:
// the push's waves hold operands of both kinds).
TH_HD double th_hypot_core_n(double ax, double ay) {
    const double h = th_sqrt_n(ax * ax + ay * ay);
    const double d1 = h - ay;
    const double t11 = ax * (2.0 * d1 - ax);
    const double t21 = (d1 - 2.0 * (ax - ay)) * d1;
    const double d2 = h - ax;
    const double t12 = 2.0 * d2 * (ax - 2.0 * ay);
    const double t22 = (4.0 * d2 - ay) * ay + d2 * d2;
    const int first = h <= 2.0 * ay;
    const double t1 = first ? t11 : t12, t2 = first ? t21 : t22;
    return h - th_div_n(t1 + t2, 2.0 * h);
}

// The same function with the common case first: operands in the core's range
// (no rescaling, finite, non-zero) go straight to the core.  For those that
// are not widely separated this is what th_hypot_full computes (its scalings
// are then by 1.0).  For widely separated ones (ay <= ax 2^-54, where
// th_hypot_full returns ax + ay, i.e. ax, since ay < ulp(ax) / 2) the core
// returns ax too: RN(ax^2) + RN(ay^2) rounds to RN(ax^2), the correctly
// rounded square root of RN(ax^2) is ax (binary floating point with
// round-to-nearest), d = 0, t1 + t2 = -RN(ay^2), and ax + RN(ay^2) / (2 ax)
// rounds to ax.
TH_HD double th_hypot(double x, double y) {
    const double fx = fabs(x), fy = fabs(y);
    const double ax = fmax(fx, fy), ay = fmin(fx, fy);   // NaN operands take the full path
    if (__builtin_expect(!__builtin_isunordered(x, y) && ax <= 0x1p+511 && ay >= 0x1p-383, 1))
        return th_hypot_core_n(ax, ay);
    return th_hypot_full(x, y);
}

// th_hypot of two independent pairs: both cores are evaluated unconditionally
// in one straight-line block (their dependent chains interleave), and a pair
// outside the core's range -- rare -- is then recomputed by th_hypot_full.
// The same doubles as two th_hypot calls.
TH_HD void th_hypot2(double x0, double y0, double x1, double y1, double *h0, double *h1) {
    const double fx0 = fabs(x0), fy0 = fabs(y0), fx1 = fabs(x1), fy1 = fabs(y1);
    const double a0 = fmax(fx0, fy0), b0 = fmin(fx0, fy0);
    const double a1 = fmax(fx1, fy1), b1 = fmin(fx1, fy1);
    const int ok0 = !__builtin_isunordered(x0, y0) && a0 <= 0x1p+511 && b0 >= 0x1p-383;
    const int ok1 = !__builtin_isunordered(x1, y1) && a1 <= 0x1p+511 && b1 >= 0x1p-383;
    double r0 = th_hypot_core_n(a0, b0), r1 = th_hypot_core_n(a1, b1);
    if (__builtin_expect(!(ok0 && ok1), 0)) {
        if (!ok0) r0 = th_hypot_full(x0, y0);
        if (!ok1) r1 = th_hypot_full(x1, y1);
    }
    *h0 = r0;
    *h1 = r1;
}

