// Particle push kernels for gfx950 (CDNA4, wave64).
//
// One lane owns one particle at a time.  Lanes whose particle terminated
// (absorbed, escaped, or out of steps) are refilled from a global work
// counter with one atomic per wave (ballot + prefix popcount), so a wave stays
// populated although path lengths run from 1 to max_steps.  Deposit tallies
// are integer atomics into a per-workgroup LDS histogram over the surface
// cells (exact and independent of order), flushed once per workgroup.
//
// The FP64 step reproduces the reference push bit for bit
// (trench_trace.cu:17-59, :92-185; trench_charging_2d.py:492-702): same
// operation order, contraction off, glibc-equivalent hypot (th_math.h).
// Cell indices are clamped to the grid on both sides (the reference clamps x
// only from above; positions are never negative there, so results agree, and
// a caller's out-of-range initial x cannot index outside the grid here).
// Bit-preserving restructurings used here:
//   * x/dx and y/dx of the new position are computed once per step and reused
//     for the next step's interpolation (x/dx - 0.5) and for the previous-cell
//     indices of wall entry; a reflected particle restores them with x0, y0;
//   * the periodic wrap takes a - L or a + L where that equals the reference's
//     fmod form exactly (Sterbenz), else falls back to fmod (the RF clock
//     wraps the same way);
//   * cfl*dx, 2*dx, nx*dx and ny*dx are formed once on the host (same products);
//   * x/dx and y/dx use Markstein's correction (P. Markstein, "Computation of
//     elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
//     Dev. 34(1), 1990): with y = RN(1/dx) and a faithful quotient q, one fused
//     residual step RN(q + (x - q dx) y) is the correctly rounded x/dx, i.e. the
//     same double as the IEEE division (positions never over- or underflow).
//
// RF full-cycle push (trench_charging_2d.py:556-602, :629-632, :654-682;
// trench_rf.py:5-11): each particle carries a phase clock and its flight
// time; the field is the frozen field plus the two unit-voltage basis fields
// weighted by the voltage wave at mid-step, and the step is capped by dt_max
// and by an acceleration bound over the whole period.  The bound's two-term
// product is formed as fma(b0, |e0|, b1 |e1|), the rounding NumPy's matrix
// product gives all but the last n % 4 live particles of a step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "th_math.h"

namespace th {

constexpr int kBlock = 256;          // 4 waves per workgroup
constexpr int kStepsPerCheck = 4;    // default steps between refill checks
enum : int { kScalQueue = 0, kScalEsc0, kScalEsc1, kScalLost, kScalSteps, kScalOff, kScalCount };

struct Grid {
    const double4 *__restrict__ cells;   // {Ex*vac, Ey*vac, vac, phi} per cell
    // Interpolation corners per lattice key (i+1, j+1), i in [-1, nx-1],
    // j in [-1, ny-1]: {Ex*vac, Ey*vac} of cells (i0,j0), (i1,j0), (i0,j1),
    // (i1,j1) after the periodic wrap / clip, and their vacuum flags (bits 0-3).
    const double2 *__restrict__ corners;  // [(nx+1)(ny+1)][4]
    const uint8_t *__restrict__ corner_vac;   // [(nx+1)(ny+1)]
    const float4 *__restrict__ cells32;  // FP32 push: {Ex*vac, Ey*vac, vac, phi} in float
    const double *__restrict__ wface;    // barrier face weight per cell
    const uint8_t *__restrict__ vac;     // 1 vacuum, 0 solid
    const int32_t *__restrict__ smap;    // cell -> surface index, -1 elsewhere
    int nx, ny, m;
    double dx, inv_dx, cfl_dx, two_dx, lx, ly;   // 1/dx, cfl*dx, 2*dx, nx*dx, ny*dx formed on the host
    // RF full-cycle push only.
    const double4 *__restrict__ rfcells;  // {E0x*vac, E0y*vac, E1x*vac, E1y*vac}: substrate / mask basis
    const double2 *__restrict__ rfphi;    // {phi0, phi1} basis potentials (barrier term)
    const double2 *__restrict__ wave;     // [n_wave] {substrate, mask} deviation from the period mean
    int n_wave;
    double period, dt_max, bound0, bound1;   // bound = max |wave| per column
    // Columns [tile_c0, tile_c0 + tile_w) that the tile push stages in LDS
    // (tile_w = 0: none).
    int tile_c0, tile_w;

    // Record access of the FP64 step (shared with the ensemble's lane view).
    static constexpr bool taper = true;   // the TAPERED template flag decides
    static constexpr bool tiled = false;  // records from global memory only
    __device__ __forceinline__ double4 cell(int c) const { return cells[c]; }
    __device__ __forceinline__ const double2 *corner(int key) const { return corners + 4 * (int64_t)key; }
    __device__ __forceinline__ unsigned corner_flags(int key) const { return corner_vac[key]; }
    __device__ __forceinline__ bool is_vac(int c) const { return vac[c] != 0; }
    __device__ __forceinline__ bool is_vac_at(int ix, int iy) const { return vac[ix * ny + iy] != 0; }
    __device__ __forceinline__ double face(int c) const { return wface[c]; }
    __device__ __forceinline__ double len_x() const { return lx; }
    __device__ __forceinline__ double len_y() const { return ly; }
};

// The single-instance grid with the columns [tile_c0, tile_c0 + tile_w)
// staged in LDS: per cell {Ex*vac, Ey*vac} (the doubles of the corner
// records) and one vacuum bit.  A step whose four interpolation cells lie in
// those columns reads them from LDS instead of its global corner record; the
// vacuum test of a cell in them reads the bit.  The values are the same, so
// the step's results are.
struct TileView : Grid {
    static constexpr bool tiled = true;
    const double2 *te;     // [tile_w * ny], column-major like the grid (cell (c0 + u, j) at u * ny + j)
    const uint32_t *tv;    // vacuum bits of the same cells, 32 per word
    __device__ __forceinline__ bool holds(int i) const { return i >= tile_c0 && i + 1 < tile_c0 + tile_w; }
    __device__ __forceinline__ bool tile_bit(int k) const { return (tv[k >> 5] >> (k & 31)) & 1u; }
    __device__ __forceinline__ bool is_vac_at(int ix, int iy) const {
        const unsigned u = (unsigned)(ix - tile_c0);
        if (u < (unsigned)tile_w) return tile_bit((int)u * ny + iy);
        return vac[ix * ny + iy] != 0;
    }
};

// Record-mode outputs (trench_trace.cu:149-183, :189-192): the end point of
// every particle (clipped to the face it crossed), its final velocity, and
// the path of the probe particles (slots[pid] >= 0).
struct Record {
    double *end;             // [n][2]
    double *vel;             // [n][2]
    const int32_t *slots;    // [n] probe index or -1
    double *paths;           // [n_probe][stride][2]
    int32_t *path_len;       // [n_probe]
    int stride;              // max_steps + 1
    double *flight;          // RF: [n] flight time to the recorded point
    double *clock;           // RF: [n] phase clock at the recorded point, in [0, period)
};

// A particle carried from one pass of the push to the next (multi-pass
// schedule): its full state and step count, so the next pass continues it
// bit for bit (40 bytes); the RF push also carries its phase clock and
// flight time (56 bytes).
struct Survivor {
    double x, y, vx, vy;
    int32_t pid, k;
};
struct SurvivorRF : Survivor {
    double clock, elapsed;   // RF phase clock and flight time
};

// Tangent-mode inputs and outputs (push64_tangent_kernel): the direction's
// initial-state tangents per particle and field tangent per cell (each null =
// zero), and per particle the tangent of the recorded end point and of the
// final velocity, the face crossed at the end, and the step count.
// One deposit of the relaxed-barrier push (push64_soft_kernel): weight a
// (tangent da) absorbed in cell `cell` across an x-face (face 1) or a y-face
// (face 0) at the crossing point (x, y) with tangent (dx, dy).  kind 0: the
// absorbed share of a relaxed barrier decision; 1: the particle's final
// absorption (its remaining weight).
struct SoftEvent {
    double x, y, dx, dy, a, da;
    int32_t pid, cell;
    int8_t face, kind, sp, pad0;
    int32_t pad1;
};

// Relaxed-barrier mode: at a barrier test the particle reflects with
// probability p = 1 / (1 + exp(t)), t = (vn^2 / 2 - q/m dphi) / (|q/m| width)
// (the sharp test reflects for t < 0), depositing the share 1 - p of its
// weight where it would have been absorbed and going on reflected with the
// share p.  For |t| >= cut the decision is the sharp one.  Deposits are
// face-smoothed and accumulated per species and surface cell with their
// tangents; each is also listed in `events` when that is non-null.
struct SoftIO {
    const double *dphi;       // [nx*ny] potential tangent per cell, or null
    double inv_width, cut;    // 1 / width [1/V], |t| beyond which the decision is sharp
    double *dep, *ddep;       // [2][m] accumulated (null: not the soft push)
    SoftEvent *events;        // [ev_cap] or null
    unsigned *n_events;       // events produced (may exceed ev_cap: the excess is not stored)
    unsigned ev_cap;
};

struct TangentIO {
    const double *dx0, *dy0, *dvx0, *dvy0;   // [n] or null
    const double2 *dfield;                   // [nx*ny] d{Ex*vac, Ey*vac} or null
    double *dend, *dvel;                     // [n][2] (soft push: may be null)
    int8_t *face;                            // [n] 1 x-face, 0 y-face, -1 none
    int32_t *steps;                          // [n]
    SoftIO soft;                             // relaxed-barrier push only
};

struct Work {
    const double *__restrict__ x, *__restrict__ y, *__restrict__ vx, *__restrict__ vy, *__restrict__ qm;
    const int8_t *__restrict__ spc;
    const double *__restrict__ t0;       // RF: launch time per particle
    int64_t begin, end;                  // particles [begin, end)
    int8_t *status;                      // nullable
    int32_t *hit;                        // nullable
    uint32_t *counts;                    // [2][m] per-surface-cell tallies
    uint32_t *overflow;                  // [2][nx*ny] tallies for non-surface cells
    unsigned long long *scal;            // kScalCount counters
    unsigned long long *queue;           // this launch's work counter (zero at launch)
    int max_steps, refill;
    int order, prio_steps, check_steps;
    Record rec;                          // record mode only
    // Multi-pass schedule: particles still live after step_cap steps are
    // appended to `out` instead of being counted; a later pass reads them from
    // `in` (its length in *in_count).  step_cap >= max_steps: single pass.
    int step_cap;
    const void *in;                      // Survivor, or SurvivorRF for the RF push
    const unsigned *in_count;
    void *out;
    unsigned *out_count;
    TangentIO tan;                       // tangent mode only (single pass)
};

__device__ __forceinline__ int wrap_cell(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (i < 0 && i >= -n) return i + n;
    if (i >= n && i < 2 * n) return i - n;
    int r = i % n;
    return r < 0 ? r + n : r;
}
__device__ __forceinline__ int clip_cell(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

// wrap_cell(i, n) and wrap_cell(i + 1, n): branch-free for the lattice
// indices of positions inside the period (i in [-1, n-1]).
__device__ __forceinline__ void wrap_pair(int i, int n, int &i0, int &i1) {
    if (__builtin_expect((unsigned)(i + 1) <= (unsigned)n, 1)) {
        i0 = i < 0 ? n - 1 : i;
        i1 = i + 1 == n ? 0 : i + 1;
    } else {
        i0 = wrap_cell(i, n);
        i1 = wrap_cell(i + 1, n);
    }
}

// NumPy's float %, i.e. r = fmod(a, L); r < 0 ? r + L : r, for L > 0.
__device__ __forceinline__ double wrap_pos(double a, double L) {
    if (a >= 0.0 && a < L) return a;
    if (a >= L && a < 2.0 * L) return a - L;   // exact: L <= a < 2L
    if (a < 0.0 && a > -L) return a + L;       // fmod(a, L) == a here
    double r = fmod(a, L);
    return r < 0.0 ? r + L : r;
}

// NumPy's float remainder for a divisor L > 0 (npy_divmod): fmod, then a
// negative remainder is shifted by L and a zero one is +0.
__device__ __forceinline__ double py_mod(double a, double L) {
    if (a >= 0.0 && a < L) return a;
    if (a >= L && a < 2.0 * L) return a - L;   // exact: L <= a < 2L
    double r = fmod(a, L);
    if (r != 0.0) return r < 0.0 ? r + L : r;
    return 0.0;
}

// Voltage deviations at time `at`: periodic linear interpolation of the wave
// (trench_rf.py:5-11).
__device__ __forceinline__ void rf_coefficients(const Grid &g, double at, double &vs, double &vm) {
    double f = py_mod(at, g.period) / g.period * (double)g.n_wave;
    double fl = floor(f);
    int i = (int)fl % g.n_wave;
    int j = i + 1 == g.n_wave ? 0 : i + 1;
    double w = f - fl;
    double2 a = g.wave[i], b = g.wave[j];
    vs = a.x * (1.0 - w) + b.x * w;
    vm = a.y * (1.0 - w) + b.y * w;
}

// Per-particle RF state.
struct RfClock {
    double clock, elapsed;   // phase clock in [0, period), flight time
    double ft, ck;           // record mode: flight time and clock at the recorded point
};

// x / dx, correctly rounded, from the host-rounded reciprocal (see above).
__device__ __forceinline__ double div_dx(double a, double dx, double inv_dx) {
    double q = a * inv_dx;
    double r = fma(-q, dx, a);
    q = fma(r, inv_dx, q);
    r = fma(-q, dx, a);
    return fma(r, inv_dx, q);
}

template <typename Real>
struct StateT {
    Real x, y, vx, vy, qm;
    Real q, r;   // x/dx, y/dx of the current position
};
using State64 = StateT<double>;
using State32 = StateT<float>;

typedef double th_dvec4 __attribute__((ext_vector_type(4)));

// The interpolation record a lane holds for its next step (single-instance
// FP64 push): the lattice key it belongs to (-1: none) and the record itself.
// Each step issues the next record's loads as soon as the new position is
// known, before its classification, so their latency overlaps that work
// instead of opening the next step.  The record depends only on the key within
// one launch (one field): a lane that takes a new particle, or whose particle
// a reflection sends back, uses it when the key matches and loads otherwise.
struct CornerCache {
    int key = -1;
    uint8_t flags = 0;
    th_dvec4 a, b;
};

// Step hooks of the FP64 push that do nothing: the push kernels' policy.
struct NoTangent {
    template <class V>
    __device__ __forceinline__ void field(const V &, int, int, int, int, double, double, double, double, double,
                                          double, double2, double2, double2, double2, double, double, double, double,
                                          bool, double, double, double) {}
    __device__ __forceinline__ void push(double, double, double, double, double, double, double) {}
    __device__ __forceinline__ void reflect(bool) {}
    __device__ __forceinline__ void crossing(bool, double, double, double, double) {}
    __device__ __forceinline__ void point() {}
};

// Forward-mode derivative of one particle's path along one direction theta
// (pathwise: the discrete events of the path -- cells, wall hits, barrier
// decisions, step count -- are held fixed).  The direction enters through the
// initial state (dx, dy, dvx, dvy) and through a field tangent: per cell
// d(Ex, Ey)/d(theta), vacuum-masked, or null for a fixed field.  The hooks
// follow the FP64 step's arithmetic term by term:
//   E = sum w_k e_k / sum w_k v_k over the four corners, w bilinear in
//       (x/dx - 0.5, y/dx - 0.5), the denominator's clamp at 1e-12 held;
//   dt = cfl dx / (|v| + sqrt(2 dx |q/m E|) + 1);
//   v' = v + a dt,  s = (v + v') dt / 2,  x' = x + s (the periodic wrap is a
//   constant shift);
//   a barrier reflection restores the start point and flips one velocity
//   component; the recorded crossing point x0 + f s with f = (face - x0)/s.
struct Tangent64 {
    double dx = 0.0, dy = 0.0, dvx = 0.0, dvy = 0.0;   // d(state)/d(theta)
    double dx0 = 0.0, dy0 = 0.0;                       // at the step's start
    double dex = 0.0, dey = 0.0;                       // d(E) at the step
    double dsx = 0.0, dsy = 0.0;                       // d(step displacement)
    double dpx = 0.0, dpy = 0.0;                       // d(recorded point)
    int face = -1;                                     // crossing: 1 x-face, 0 y-face, -1 none
    const double2 *__restrict__ dfield = nullptr;

    template <class V>
    __device__ __forceinline__ void field(const V &g, int i0, int i1, int j0, int j1, double tx, double ty,
                                          double w00, double w10, double w01, double w11, double2 e00, double2 e10,
                                          double2 e01, double2 e11, double v00, double v10, double v01, double v11,
                                          bool clamped, double sw, double ex, double ey) {
        dx0 = dx;
        dy0 = dy;
        const double dtx = dx * g.inv_dx, dty = dy * g.inv_dx;
        const double d00 = -dtx * (1.0 - ty) - (1.0 - tx) * dty, d10 = dtx * (1.0 - ty) - tx * dty;
        const double d01 = -dtx * ty + (1.0 - tx) * dty, d11 = dtx * ty + tx * dty;
        double nx_ = d00 * e00.x + d10 * e10.x + d01 * e01.x + d11 * e11.x;
        double ny_ = d00 * e00.y + d10 * e10.y + d01 * e01.y + d11 * e11.y;
        if (dfield) {
            const int ny = g.ny;
            const double2 f00 = dfield[i0 * ny + j0], f10 = dfield[i1 * ny + j0];
            const double2 f01 = dfield[i0 * ny + j1], f11 = dfield[i1 * ny + j1];
            nx_ += w00 * f00.x + w10 * f10.x + w01 * f01.x + w11 * f11.x;
            ny_ += w00 * f00.y + w10 * f10.y + w01 * f01.y + w11 * f11.y;
        }
        const double dsw = clamped ? 0.0 : d00 * v00 + d10 * v10 + d01 * v01 + d11 * v11;
        dex = (nx_ - ex * dsw) / sw;
        dey = (ny_ - ey * dsw) / sw;
        face = -1;
    }
    // vx, vy: the step's start velocity; (ax, ay) = q/m E; dt as computed.
    __device__ __forceinline__ void push(double vx, double vy, double qm, double ax, double ay, double dt,
                                         double two_dx) {
        const double dax = qm * dex, day = qm * dey;
        const double hv = th_hypot(vx, vy), ha = th_hypot(ax, ay);
        const double root = sqrt(two_dx * ha);
        const double dhv = hv > 0.0 ? (vx * dvx + vy * dvy) / hv : 0.0;
        const double dha = ha > 0.0 ? (ax * dax + ay * day) / ha : 0.0;
        const double droot = root > 0.0 ? 0.5 * two_dx * dha / root : 0.0;
        const double ddt = -dt * (dhv + droot) / (hv + root + 1.0);
        const double dvxn = dvx + dax * dt + ax * ddt, dvyn = dvy + day * dt + ay * ddt;
        const double vxn = vx + ax * dt, vyn = vy + ay * dt;
        dsx = 0.5 * ((dvx + dvxn) * dt + (vx + vxn) * ddt);
        dsy = 0.5 * ((dvy + dvyn) * dt + (vy + vyn) * ddt);
        dx = dx0 + dsx;
        dy = dy0 + dsy;
        dvx = dvxn;
        dvy = dvyn;
    }
    __device__ __forceinline__ void reflect(bool face_x) {
        dx = dx0;
        dy = dy0;
        if (face_x) dvx = -dvx; else dvy = -dvy;
    }
    // The step's end was recorded at fraction f of the step, where it crosses
    // an x-face (x_face) or a y-face; f_raw before the clip to [0, 1].
    __device__ __forceinline__ void crossing(bool x_face, double f_raw, double f, double sx, double sy) {
        double df = 0.0;
        if (f_raw > 0.0 && f_raw < 1.0) df = x_face ? -(dx0 + f_raw * dsx) / sx : -(dy0 + f_raw * dsy) / sy;
        dpx = dx0 + df * sx + f * dsx;
        dpy = dy0 + df * sy + f * dsy;
        face = x_face ? 1 : 0;
    }
    __device__ __forceinline__ void point() {
        dpx = dx;
        dpy = dy;
    }
};

// Face smoothing of a deposit in cell c (surface index k) crossed through an
// x-face (xf) or a y-face at the point (xq, yq): the neighbour along the face
// on the side of the point's coordinate u along it (u in [0, 1] in cell
// units) takes the share w_nb = |u - 1/2| when it is a surface cell (index
// knb) with vacuum in front of it on the particle's side, the hit cell the
// rest; slope = d w_nb / d(coordinate along the face) [1/m] (zero when u was
// clipped to the face's end or no neighbour shares).  Returns false for a
// cell off the surface.
template <class V>
__device__ __forceinline__ bool face_share(const V &g, int c, bool xf, double xq, double yq, int &k, int &knb,
                                           double &w_nb, double &slope) {
    k = g.smap[c];
    if (k < 0) return false;
    const int nx = g.nx, ny = g.ny, cx = c / ny, cy = c - (c / ny) * ny;
    const double ex = xq / g.dx, ey = yq / g.dx;
    double along = xf ? ey - (double)cy : py_mod(ex - (double)cx, (double)nx);
    const double across = xf ? py_mod(ex - (double)cx, (double)nx) : ey - (double)cy;
    const bool moving = along > 0.0 && along < 1.0;   // a clipped crossing point stays put
    along = fmin(fmax(along, 0.0), 1.0);
    const int side = py_mod(rint(across), (double)nx) == 1.0 ? 1 : -1;   // from the cell after (+1) or before
    const int step = along >= 0.5 ? 1 : -1;
    const int nbx = xf ? cx : ((cx + step) % nx + nx) % nx, nby = xf ? cy + step : cy;
    const int fx = xf ? ((nbx + side) % nx + nx) % nx : nbx, fy = xf ? nby : nby + side;
    bool valid = nby >= 0 && nby < ny && fy >= 0 && fy < ny;
    knb = k;
    if (valid) {
        const int nb = nbx * ny + nby;
        valid = !g.is_vac(nb) && g.is_vac(fx * ny + fy) && g.smap[nb] >= 0;
        if (valid) knb = g.smap[nb];
    }
    w_nb = valid ? fabs(along - 0.5) : 0.0;
    const double sgn = along > 0.5 ? 1.0 : (along < 0.5 ? -1.0 : 0.0);
    slope = valid && moving ? sgn / g.dx : 0.0;
    return true;
}

// Face-smoothed deposit of weight a (tangent da) absorbed in cell c at the
// crossing point (xq, yq), tangent (dxq, dyq) (face_share): the shares and
// their tangents (the share's tangent is slope times the tangent of the
// coordinate along the face) accumulated into dep / ddep ([2][m], species sp).
template <class V>
__device__ __forceinline__ void soft_face_deposit(const V &g, double *dep, double *ddep, int sp, int c, bool xf,
                                                  double xq, double yq, double dxq, double dyq, double a,
                                                  double da) {
    int k, knb;
    double w_nb, slope;
    if (!face_share(g, c, xf, xq, yq, k, knb, w_nb, slope)) return;
    const double dw = slope * (xf ? dyq : dxq);
    const int m = g.m;
    atomicAdd(&dep[sp * m + k], a * (1.0 - w_nb));
    atomicAdd(&ddep[sp * m + k], da * (1.0 - w_nb) - a * dw);
    if (knb != k || w_nb != 0.0) {
        atomicAdd(&dep[sp * m + knb], a * w_nb);
        atomicAdd(&ddep[sp * m + knb], da * w_nb + a * dw);
    }
}

// Transpose of one face-smoothed deposit for the cotangent gbar ([2][m]) of
// the deposits: the cotangent of the weight a (abar) and of the crossing
// point (pbx, pby).
template <class V>
__device__ __forceinline__ void soft_face_vjp(const V &g, const double *gbar, int sp, int c, bool xf, double xq,
                                              double yq, double a, double &abar, double &pbx, double &pby) {
    int k, knb;
    double w_nb, slope;
    abar = pbx = pby = 0.0;
    if (!face_share(g, c, xf, xq, yq, k, knb, w_nb, slope)) return;
    const double gk = gbar[sp * g.m + k], gn = gbar[sp * g.m + knb];
    abar = (1.0 - w_nb) * gk + w_nb * gn;
    const double pa = a * slope * (gn - gk);
    if (xf) pby = pa; else pbx = pa;
}

// Tangent of the relaxed-barrier push (SoftIO): Tangent64's path derivative
// plus the particle's weight W and its tangent.  At a barrier test with
// |t| < cut the particle deposits W (1 - p) where it hit and goes on
// reflected with W p; the tangent of p = 1 / (1 + exp(t)) is -p (1 - p) dt,
// dt = (vn dvn - q/m wface (dphi_c - dphi_v)) / (|q/m| width), vn the normal
// velocity at the step's end and dphi the potential tangent (held fixed
// between cut and infinity, where the sharp decision is taken).
struct SoftTangent : Tangent64 {
    const double *__restrict__ dphi = nullptr;
    double inv_width = 0.0, cut = 0.0;
    double W = 1.0, dW = 0.0;
    bool pend = false;           // this step's barrier decision was relaxed
    double pp = 0.0, dpp = 0.0;  // its reflection probability and tangent
    int sp = 0, pid = 0;
    double *dep = nullptr, *ddep = nullptr;   // LDS accumulators [2][m]
    SoftEvent *events = nullptr;
    unsigned *n_events = nullptr;
    unsigned ev_cap = 0;

    // have = vn^2 / 2, need = q/m wface (phi_c - phi_v): the sharp test
    // reflects when have < need.
    __device__ __forceinline__ bool reflects(double have, double need, double vn, double qm, double wf, int c, int v,
                                             bool face_x) {
        pend = false;
        const double aq = fabs(qm);
        const double t = (have - need) / aq * inv_width;
        if (!(t > -cut && t < cut)) return have < need;
        const double p = 1.0 / (1.0 + exp(t));
        const double dvn = face_x ? dvx : dvy;
        const double dneed = dphi ? qm * wf * (dphi[c] - dphi[v]) : 0.0;
        const double dt = (vn * dvn - dneed) / aq * inv_width;
        pp = p;
        dpp = -p * (1.0 - p) * dt;
        pend = true;
        return true;   // the share p goes on reflected
    }
    // Deposit of weight a at the recorded crossing point (dpx, dpy set by crossing()).
    template <class V>
    __device__ __forceinline__ void deposit(const V &g, int c, bool xf, double xq, double yq, double a, double da,
                                            int kind) {
        if (events) {
            const unsigned slot = atomicAdd(n_events, 1u);
            if (slot < ev_cap) {
                SoftEvent e;
                e.x = xq; e.y = yq; e.dx = dpx; e.dy = dpy; e.a = a; e.da = da;
                e.pid = pid; e.cell = c; e.face = xf ? 1 : 0; e.kind = (int8_t)kind; e.sp = (int8_t)sp;
                e.pad0 = 0; e.pad1 = 0;
                events[slot] = e;
            }
        }
        soft_face_deposit(g, dep, ddep, sp, c, xf, xq, yq, dpx, dpy, a, da);
    }
    // The relaxed decision's absorbed share, deposited at the crossing point
    // of the hit face; the weight goes on with the reflected path.
    template <class V>
    __device__ __forceinline__ void relaxed(const V &g, int c, bool xf, double xq, double yq) {
        deposit(g, c, xf, xq, yq, W * (1.0 - pp), dW * (1.0 - pp) - W * dpp, 0);
        dW = dW * pp + W * dpp;
        W = W * pp;
        pend = false;
        dpx = dx;
        dpy = dy;
        face = -1;
    }
};

// Policies that take the relaxed barrier decision (reflects / pend / relaxed hooks).
template <class T> struct is_soft_policy : std::false_type {};
template <> struct is_soft_policy<SoftTangent> : std::true_type {};

// One FP64 push step.  Returns 0 (continue), 1 (absorbed at *cell), 2
// (escaped).  With RECORD, (xp, yp) is the step's recorded point.  With RF the
// voltage wave acts during the step and rc advances.  V reads one instance's
// records: Grid itself, or an ensemble lane's view (th_ensemble.hpp); cell
// indices are the instance's own.  T receives the step's intermediate values
// (NoTangent, or Tangent64 for the derivative of the path).
template <bool BARRIER, bool TAPERED, bool RECORD, bool RF, class V, class T, bool CACHE = false>
__device__ __forceinline__ int step64t(const V &g, State64 &s, int &cell, double &xp, double &yp, RfClock &rc,
                                       T &tan, CornerCache *cc = nullptr) {
    static_assert(std::is_same<T, NoTangent>::value || (RECORD && !RF), "the tangent follows the FP64 record step");
    const int nx = g.nx, ny = g.ny;
    // Vacuum-weighted bilinear field at the cell-centred lattice.
    double fx = s.q - 0.5, fy = s.r - 0.5;
    double fi = floor(fx), fj = floor(fy);
    int i = (int)fi, j = (int)fj;
    double tx = fx - fi, ty = fy - fj;   // (double)i == fi for every position the grid can hold
    double w00 = (1.0 - tx) * (1.0 - ty), w10 = tx * (1.0 - ty), w01 = (1.0 - tx) * ty, w11 = tx * ty;
    double2 e00, e10, e01, e11;
    double v00, v10, v01, v11;              // vacuum weights 1.0 / 0.0
    int i0, i1, j0, j1;
    unsigned fl;                            // vacuum flags of (i0,j0), (i1,j0), (i0,j1), (i1,j1): bits 0-3
    if (__builtin_expect((unsigned)(i + 1) <= (unsigned)nx && (unsigned)(j + 1) <= (unsigned)ny, 1)) {
        i0 = i < 0 ? nx - 1 : i;
        i1 = i + 1 == nx ? 0 : i + 1;
        j0 = j < 0 ? 0 : j;
        j1 = j + 1 >= ny ? ny - 1 : j + 1;
        unsigned f = 0u;
        bool staged = false;
        if constexpr (V::tiled) {
            if (g.holds(i)) {   // four cells in the LDS tile (i0 = i, i1 = i + 1, no wrap)
                const int b0 = (i - g.tile_c0) * ny;
                e00 = g.te[b0 + j0]; e10 = g.te[b0 + ny + j0];
                e01 = g.te[b0 + j1]; e11 = g.te[b0 + ny + j1];
                f = (g.tile_bit(b0 + j0) ? 1u : 0u) | (g.tile_bit(b0 + ny + j0) ? 2u : 0u) |
                    (g.tile_bit(b0 + j1) ? 4u : 0u) | (g.tile_bit(b0 + ny + j1) ? 8u : 0u);
                staged = true;
            }
        }
        if (!staged) {
            // one 64-byte corner record and one flag byte per step
            // 24-bit multiply (grids are under 2^24 cells per side)
            const int key = (int)__umul24((unsigned)(i + 1), (unsigned)(ny + 1)) + (j + 1);
            th_dvec4 a, b;
            // two 32-byte vector loads (kept whole: element-wise double2 reads get split into 8-byte loads)
            const th_dvec4 *c = (const th_dvec4 *)__builtin_assume_aligned(g.corner(key), 64);
            if constexpr (CACHE) {
                if (cc->key != key) {   // no record prefetched for this key (new particle, reflection)
                    cc->a = c[0];
                    cc->b = c[1];
                    cc->flags = g.corner_flags(key);
                }
                a = cc->a;
                b = cc->b;
                f = cc->flags;
            } else {
                a = c[0];
                b = c[1];
                f = g.corner_flags(key);
            }
            e00 = make_double2(a.x, a.y); e10 = make_double2(a.z, a.w);
            e01 = make_double2(b.x, b.y); e11 = make_double2(b.z, b.w);
        }
        fl = f;
        v00 = (f & 1u) ? 1.0 : 0.0; v10 = (f & 2u) ? 1.0 : 0.0;
        v01 = (f & 4u) ? 1.0 : 0.0; v11 = (f & 8u) ? 1.0 : 0.0;
    } else {                                // positions outside the period or grid
        wrap_pair(i, nx, i0, i1);
        j0 = clip_cell(j, ny);
        j1 = clip_cell(j + 1, ny);
        double4 c00 = g.cell(i0 * ny + j0), c10 = g.cell(i1 * ny + j0);
        double4 c01 = g.cell(i0 * ny + j1), c11 = g.cell(i1 * ny + j1);
        e00 = make_double2(c00.x, c00.y); e10 = make_double2(c10.x, c10.y);
        e01 = make_double2(c01.x, c01.y); e11 = make_double2(c11.x, c11.y);
        v00 = c00.z; v10 = c10.z; v01 = c01.z; v11 = c11.z;
        fl = (v00 != 0.0 ? 1u : 0u) | (v10 != 0.0 ? 2u : 0u) | (v01 != 0.0 ? 4u : 0u) | (v11 != 0.0 ? 8u : 0u);
    }
    // sw starts at its first term: 0 + w v == w v for w v >= +0; with v in
    // {0, 1} (vacuum flag bit k of fl) and w >= 0, w v is a select of w or +0
    double ex = 0.0, ey = 0.0, sw = (fl & 1u) ? w00 : 0.0;
    ex += w00 * e00.x; ey += w00 * e00.y;
    ex += w10 * e10.x; ey += w10 * e10.y; sw += (fl & 2u) ? w10 : 0.0;
    ex += w01 * e01.x; ey += w01 * e01.y; sw += (fl & 4u) ? w01 : 0.0;
    ex += w11 * e11.x; ey += w11 * e11.y; sw += (fl & 8u) ? w11 : 0.0;
    const bool clamped = sw < 1e-12;
    sw = clamped ? 1e-12 : sw;
    // Dividing by a weight sum of exactly 1 (four vacuum cells, as for most
    // steps) returns the numerator: a wave whose lanes all have one skips the
    // divisions.
    if (__ballot(sw != 1.0) != 0ull) {
        ex = ex / sw;
        ey = ey / sw;
    }
    tan.field(g, i0, i1, j0, j1, tx, ty, w00, w10, w01, w11, e00, e10, e01, e11, v00, v10, v01, v11, clamped, sw,
              ex, ey);
    // CFL-limited trapezoidal push.
    double ax, ay, dt;
    if constexpr (RF) {
        double4 b00 = g.rfcells[i0 * ny + j0], b10 = g.rfcells[i1 * ny + j0];
        double4 b01 = g.rfcells[i0 * ny + j1], b11 = g.rfcells[i1 * ny + j1];
        double e0x = 0.0, e0y = 0.0, e1x = 0.0, e1y = 0.0;
        e0x += w00 * b00.x; e0y += w00 * b00.y; e1x += w00 * b00.z; e1y += w00 * b00.w;
        e0x += w10 * b10.x; e0y += w10 * b10.y; e1x += w10 * b10.z; e1y += w10 * b10.w;
        e0x += w01 * b01.x; e0y += w01 * b01.y; e1x += w01 * b01.z; e1y += w01 * b01.w;
        e0x += w11 * b11.x; e0y += w11 * b11.y; e1x += w11 * b11.z; e1y += w11 * b11.w;
        e0x = e0x / sw; e0y = e0y / sw; e1x = e1x / sw; e1y = e1y / sw;
        // Acceleration bound over the period, then the field at mid-step.
        double bx = fabs(ex) + fma(g.bound0, fabs(e0x), g.bound1 * fabs(e1x));
        double by = fabs(ey) + fma(g.bound0, fabs(e0y), g.bound1 * fabs(e1y));
        double hb, hv;
        th_hypot2(bx, by, s.vx, s.vy, &hb, &hv);
        dt = th_cfl_dt(g.cfl_dx, g.two_dx, hv, fabs(s.qm) * hb);
        if (!(dt < g.dt_max) && !isnan(dt)) dt = g.dt_max;
        double vs, vm;
        rf_coefficients(g, rc.clock + dt / 2.0, vs, vm);
        ex = ex + (vs * e0x + vm * e1x);
        ey = ey + (vs * e0y + vm * e1y);
        ax = s.qm * ex;
        ay = s.qm * ey;
        rc.clock = py_mod(rc.clock + dt, g.period);
        rc.elapsed = rc.elapsed + dt;
    } else {
        ax = s.qm * ex;
        ay = s.qm * ey;
        double hv, ha;
        th_hypot2(s.vx, s.vy, ax, ay, &hv, &ha);
        dt = th_cfl_dt(g.cfl_dx, g.two_dx, hv, ha);
        tan.push(s.vx, s.vy, s.qm, ax, ay, dt, g.two_dx);
    }
    double vxn = s.vx + ax * dt, vyn = s.vy + ay * dt;
    const double x0 = s.x, y0 = s.y, q0 = s.q, r0 = s.r;
    double sx = 0.5 * (s.vx + vxn) * dt, sy = 0.5 * (s.vy + vyn) * dt;
    double xu = x0 + sx;
    s.x = wrap_pos(xu, g.len_x());
    s.y = y0 + sy;
    s.vx = vxn;
    s.vy = vyn;
    s.q = div_dx(s.x, g.dx, g.inv_dx);
    s.r = div_dx(s.y, g.dx, g.inv_dx);
    if constexpr (CACHE) {
        // The next step's record (a position outside the period or grid loads
        // key 0's record and marks it unusable).
        const int ni = (int)floor(s.q - 0.5), nj = (int)floor(s.r - 0.5);
        const bool inside = (unsigned)(ni + 1) <= (unsigned)nx && (unsigned)(nj + 1) <= (unsigned)ny;
        const int nkey = inside ? (int)__umul24((unsigned)(ni + 1), (unsigned)(ny + 1)) + (nj + 1) : 0;
        const th_dvec4 *c = (const th_dvec4 *)__builtin_assume_aligned(g.corner(nkey), 64);
        cc->a = c[0];
        cc->b = c[1];
        cc->flags = g.corner_flags(nkey);
        cc->key = inside ? nkey : -1;
    }
    // Classification of the end point.  The end cell is nearly always one of
    // the four interpolation cells (a step moves less than half a cell at
    // cfl <= 0.5): then its vacuum flag is one of fl's bits, the same byte of
    // the vacuum map, and the map is read only otherwise.
    bool escaped = s.y >= g.len_y();
    int ix = clip_cell((int)s.q, nx), iy = clip_cell((int)s.r, ny);
    const int xe = ix == i0 ? 0 : (ix == i1 ? 1 : -1);
    const int ye = iy == j0 ? 0 : (iy == j1 ? 2 : -1);
    bool hit = false;
    if (!escaped) hit = (xe >= 0 && ye >= 0) ? !((fl >> (xe + ye)) & 1u) : !g.is_vac_at(ix, iy);
    bool corner = false;
    if (TAPERED && g.taper && !hit && !escaped) {
        const int ix0c = clip_cell((int)q0, nx), iy0c = clip_cell((int)r0, ny);
        corner = ix != ix0c && iy != iy0c;
        // A diagonal move whose two side cells (ix, iy0) and (ix0, iy) are both
        // vacuum crosses no solid cell: the wall-entry search below would
        // return the end cell, not hit.  Without a periodic wrap of the step,
        // and with x/dx inside the grid, its dix / diy are ix - ix0 and
        // iy - iy0, so it is skipped when the interpolation's four cells hold
        // both side cells and flag them vacuum.
        if (corner && xu >= 0.0 && xu < g.len_x() && s.q < (double)nx) {
            const int xs = ix0c == i0 ? 0 : (ix0c == i1 ? 1 : -1);
            const int ys = iy0c == j0 ? 0 : (iy0c == j1 ? 2 : -1);
            if (xe >= 0 && xs >= 0 && ye >= 0 && ys >= 0 && ((fl >> (xe + ys)) & 1u) && ((fl >> (xs + ye)) & 1u))
                corner = false;
        }
    }
    int cx = ix, cy = iy;
    bool face_x = false;
    if ((hit && (BARRIER || TAPERED || RECORD)) || corner) {
        // First solid cell crossed during the step, and the vacuum cell before it.
        int ix0 = clip_cell((int)q0, nx), iy0 = clip_cell((int)r0, ny);
        int dix = (int)floor(xu / g.dx) - ix0;
        int diy = clip_cell((int)floor(s.r), ny) - iy0;   // (y0 + sy)/dx is s.r
        double tfx = dix != 0 ? ((double)(ix0 + (dix > 0)) * g.dx - x0) / sx : INFINITY;
        double tfy = diy != 0 ? ((double)(iy0 + (diy > 0)) * g.dx - y0) / sy : INFINITY;
        bool x_first = tfx < tfy;
        int ax_ = x_first ? wrap_cell(ix0 + dix, nx) : ix0;
        int ay_ = x_first ? iy0 : iy0 + diy;
        bool last = dix != 0 && diy != 0 && g.is_vac_at(ax_, ay_);
        cx = last ? wrap_cell(ix0 + dix, nx) : ax_;
        cy = last ? iy0 + diy : ay_;
        int vcx = last ? ax_ : ix0, vcy = last ? ay_ : iy0;
        face_x = last ? !x_first : x_first;
        hit = !g.is_vac_at(cx, cy);
        if (BARRIER && hit) {
            double vn = face_x ? s.vx : s.vy;
            int c = cx * ny + cy, v = vcx * ny + vcy;
            double dphi = g.face(c) * (g.cell(c).w - g.cell(v).w);
            if constexpr (RF) {   // the wall potential at the step's end time
                double cs, cm;
                rf_coefficients(g, rc.clock, cs, cm);
                double2 pc = g.rfphi[c], pv = g.rfphi[v];
                dphi = dphi + g.wface[c] * (cs * (pc.x - pv.x) + cm * (pc.y - pv.y));
            }
            bool reflect;
            if constexpr (is_soft_policy<T>::value)
                reflect = tan.reflects(0.5 * vn * vn, s.qm * dphi, vn, s.qm, g.face(c), c, v, face_x);
            else
                reflect = 0.5 * vn * vn < s.qm * dphi;
            if (reflect) {   // cannot climb the wall barrier: reflect
                if (face_x) s.vx = -s.vx; else s.vy = -s.vy;
                s.x = x0; s.y = y0; s.q = q0; s.r = r0;
                hit = false;
                tan.reflect(face_x);
            }
        }
        ix = cx;
        iy = cy;
    }
    if (RECORD) {
        xp = s.x;
        yp = s.y;
        double tf = 1.0;   // fraction of the step to the recorded point
        tan.point();
        // the point where the step crosses the hit cell's face
        auto face_point = [&](double &raw, double &frac, double &xq, double &yq) {
            if (face_x) {
                double xb = (double)(cx + (sx < 0.0)) * g.dx;
                xb += nearbyint((x0 - xb) / g.len_x()) * g.len_x();
                raw = (xb - x0) / sx;
            } else {
                raw = ((double)(cy + (sy < 0.0)) * g.dx - y0) / sy;
            }
            frac = fmin(fmax(raw, 0.0), 1.0);
            xq = wrap_pos(x0 + frac * sx, g.len_x());
            yq = y0 + frac * sy;
        };
        if (hit) {
            double raw, frac;
            face_point(raw, frac, xp, yp);
            tf = frac;
            tan.crossing(face_x, raw, frac, sx, sy);
        }
        if constexpr (is_soft_policy<T>::value) {
            if (tan.pend) {   // the relaxed decision's absorbed share, at the face it would have crossed
                double raw, frac, xq, yq;
                face_point(raw, frac, xq, yq);
                tan.crossing(face_x, raw, frac, sx, sy);
                tan.relaxed(g, cx * ny + cy, face_x, xq, yq);
            }
        }
        if (escaped) {
            double raw = (g.len_y() - y0) / sy;
            double frac = fmin(fmax(raw, 0.0), 1.0);
            tf = frac;
            xp = wrap_pos(x0 + frac * sx, g.len_x());
            yp = g.len_y();
            tan.crossing(false, raw, frac, sx, sy);
        }
        if constexpr (RF) {
            rc.ft = rc.elapsed - (1.0 - tf) * dt;
            rc.ck = py_mod(rc.clock - (1.0 - tf) * dt, g.period);
        }
    }
    if (hit) { cell = ix * ny + iy; return 1; }
    return escaped ? 2 : 0;
}

template <bool BARRIER, bool TAPERED, bool RECORD, bool RF, class V>
__device__ __forceinline__ int step64(const V &g, State64 &s, int &cell, double &xp, double &yp, RfClock &rc) {
    NoTangent none;
    return step64t<BARRIER, TAPERED, RECORD, RF>(g, s, cell, xp, yp, rc, none);
}

__device__ __forceinline__ float wrap_posf(float a, float L) {
    if (a >= 0.0f && a < L) return a;
    if (a >= L && a < 2.0f * L) return a - L;
    if (a < 0.0f && a > -L) return a + L;
    float r = fmodf(a, L);
    return r < 0.0f ? r + L : r;
}

// One FP32 push step: the same algorithm in single precision (SI units,
// float field records, x/dx as a multiply by 1/dx, library hypotf).  Not
// bit-compatible with the reference; its tallies agree statistically.
template <bool BARRIER, bool TAPERED>
__device__ __forceinline__ int step32(const Grid &g, State32 &s, int &cell) {
    const int nx = g.nx, ny = g.ny;
    const float dx = (float)g.dx, inv_dx = (float)g.inv_dx, lx = (float)g.lx, ly = (float)g.ly;
    float fx = s.q - 0.5f, fy = s.r - 0.5f;
    float fi = floorf(fx), fj = floorf(fy);
    int i = (int)fi, j = (int)fj;
    float tx = fx - fi, ty = fy - fj;
    int i0, i1;
    wrap_pair(i, nx, i0, i1);
    int j0 = clip_cell(j, ny), j1 = clip_cell(j + 1, ny);
    float4 c00 = g.cells32[i0 * ny + j0], c10 = g.cells32[i1 * ny + j0];
    float4 c01 = g.cells32[i0 * ny + j1], c11 = g.cells32[i1 * ny + j1];
    float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty), w01 = (1.0f - tx) * ty, w11 = tx * ty;
    float ex = w00 * c00.x + w10 * c10.x + w01 * c01.x + w11 * c11.x;
    float ey = w00 * c00.y + w10 * c10.y + w01 * c01.y + w11 * c11.y;
    float sw = w00 * c00.z + w10 * c10.z + w01 * c01.z + w11 * c11.z;
    // The hardware's one-ulp reciprocal and square root (this push is not
    // bit-compatible anyway): |v| and |E| as square roots of sums of squares,
    // which stay finite for any speed below 1e19 m/s and field below 1e19 V/m
    // (hypotf otherwise); |a| = |q/m| |E| (|q/m| E can exceed the float range
    // when squared).
    float inv_sw = __builtin_amdgcn_rcpf(fmaxf(sw, 1e-12f));
    ex *= inv_sw;
    ey *= inv_sw;
    float ax = s.qm * ex, ay = s.qm * ey;
    const float v2 = s.vx * s.vx + s.vy * s.vy, e2 = ex * ex + ey * ey;
    const float hv = v2 < 1e38f ? __builtin_amdgcn_sqrtf(v2) : hypotf(s.vx, s.vy);
    const float ha = fabsf(s.qm) * (e2 < 1e38f ? __builtin_amdgcn_sqrtf(e2) : hypotf(ex, ey));
    float dt = (float)g.cfl_dx * __builtin_amdgcn_rcpf(hv + __builtin_amdgcn_sqrtf((float)g.two_dx * ha) + 1.0f);
    float vxn = s.vx + ax * dt, vyn = s.vy + ay * dt;
    const float x0 = s.x, y0 = s.y, q0 = s.q, r0 = s.r;
    float sx = 0.5f * (s.vx + vxn) * dt, sy = 0.5f * (s.vy + vyn) * dt;
    float xu = x0 + sx;
    s.x = wrap_posf(xu, lx);
    s.y = y0 + sy;
    s.vx = vxn;
    s.vy = vyn;
    s.q = s.x * inv_dx;
    s.r = s.y * inv_dx;
    // End cell and diagonal moves as in step64t: vacuum flags from the four
    // interpolation cells' records where they hold the cell, else the map.
    bool escaped = s.y >= ly;
    int ix = clip_cell((int)s.q, nx), iy = clip_cell((int)s.r, ny);
    const unsigned fl = (c00.z != 0.0f ? 1u : 0u) | (c10.z != 0.0f ? 2u : 0u) | (c01.z != 0.0f ? 4u : 0u) |
                        (c11.z != 0.0f ? 8u : 0u);
    const int xe = ix == i0 ? 0 : (ix == i1 ? 1 : -1);
    const int ye = iy == j0 ? 0 : (iy == j1 ? 2 : -1);
    bool hit = false;
    if (!escaped) hit = (xe >= 0 && ye >= 0) ? !((fl >> (xe + ye)) & 1u) : !g.vac[ix * ny + iy];
    bool corner = false;
    if (TAPERED && !hit && !escaped) {
        const int ix0c = clip_cell((int)q0, nx), iy0c = clip_cell((int)r0, ny);
        corner = ix != ix0c && iy != iy0c;
        if (corner && xu >= 0.0f && xu < lx && s.q < (float)nx) {
            const int xs = ix0c == i0 ? 0 : (ix0c == i1 ? 1 : -1);
            const int ys = iy0c == j0 ? 0 : (iy0c == j1 ? 2 : -1);
            if (xe >= 0 && xs >= 0 && ye >= 0 && ys >= 0 && ((fl >> (xe + ys)) & 1u) && ((fl >> (xs + ye)) & 1u))
                corner = false;
        }
    }
    if ((hit && (BARRIER || TAPERED)) || corner) {
        int ix0 = clip_cell((int)q0, nx), iy0 = clip_cell((int)r0, ny);
        int dix = (int)floorf(xu * inv_dx) - ix0;
        int diy = clip_cell((int)floorf(s.r), ny) - iy0;
        float tfx = dix != 0 ? ((float)(ix0 + (dix > 0)) * dx - x0) / sx : INFINITY;
        float tfy = diy != 0 ? ((float)(iy0 + (diy > 0)) * dx - y0) / sy : INFINITY;
        bool x_first = tfx < tfy;
        int ax_ = x_first ? wrap_cell(ix0 + dix, nx) : ix0;
        int ay_ = x_first ? iy0 : iy0 + diy;
        bool last = dix != 0 && diy != 0 && g.vac[ax_ * ny + ay_];
        int cx = last ? wrap_cell(ix0 + dix, nx) : ax_;
        int cy = last ? iy0 + diy : ay_;
        int vcx = last ? ax_ : ix0, vcy = last ? ay_ : iy0;
        bool face_x = last ? !x_first : x_first;
        hit = !g.vac[cx * ny + cy];
        if (BARRIER && hit) {
            float vn = face_x ? s.vx : s.vy;
            int c = cx * ny + cy, v = vcx * ny + vcy;
            float dphi = (float)g.wface[c] * (g.cells32[c].w - g.cells32[v].w);
            if (0.5f * vn * vn < s.qm * dphi) {
                if (face_x) s.vx = -s.vx; else s.vy = -s.vy;
                s.x = x0; s.y = y0; s.q = q0; s.r = r0;
                hit = false;
            }
        }
        ix = cx;
        iy = cy;
    }
    if (hit) { cell = ix * ny + iy; return 1; }
    return escaped ? 2 : 0;
}

// Records the end of one particle: per-particle outputs and LDS tallies.
__device__ __forceinline__ void finish(const Grid &g, const Work &w, uint32_t *lds_counts,
                                       unsigned long long *lds_scal, int64_t pid, int sp, int st,
                                       int cell) {
    if (w.status) w.status[pid] = (int8_t)st;
    if (w.hit) w.hit[pid] = st == 1 ? cell : -1;
    if (st == 1) {
        int k = g.smap[cell];
        if (k >= 0) {
            atomicAdd(&lds_counts[sp * g.m + k], 1u);
        } else {
            atomicAdd(&w.overflow[(int64_t)sp * g.nx * g.ny + cell], 1u);
            atomicAdd(&lds_scal[kScalOff], 1ull);
        }
    } else if (st == 2) {
        atomicAdd(&lds_scal[kScalEsc0 + sp], 1ull);
    } else {
        atomicAdd(&lds_scal[kScalLost], 1ull);
    }
}

enum : int {
    kModeF64 = 0, kModeF64Record = 1, kModeF32 = 2, kModeF64Rf = 3, kModeF64RfRecord = 4, kModeF64Tangent = 5,
    kModeF64Soft = 6
};

// LDS bytes of the tile push beyond the tallies: the tallies [2][m] (u32)
// rounded up to 16 bytes, then tile_w * ny double2 and their vacuum bits.
__host__ __device__ __forceinline__ size_t tile_offset(int m) { return ((size_t)2 * m * 4 + 15) & ~(size_t)15; }
__host__ __device__ __forceinline__ size_t tile_bytes(int w, int ny) {
    return (size_t)w * ny * 16 + (((size_t)w * ny + 31) / 32) * 4;
}

// Persistent refill loop shared by every push variant.  TILE: the FP64 push
// with the grid's tile columns staged in LDS (TileView).
template <int MODE, bool BARRIER, bool TAPERED, bool TILE = false>
__device__ __forceinline__ void push_body(const Grid &g, const Work &w) {
    using Real = typename std::conditional<MODE == kModeF32, float, double>::type;
    constexpr bool SOFT = MODE == kModeF64Soft;
    constexpr bool TANGENT = MODE == kModeF64Tangent || SOFT;
    constexpr bool RECORD = MODE == kModeF64Record || MODE == kModeF64RfRecord || TANGENT;
    constexpr bool RF = MODE == kModeF64Rf || MODE == kModeF64RfRecord;
    static_assert(!TILE || MODE == kModeF64, "the tile push is the plain FP64 push");
    using Tan = typename std::conditional<SOFT, SoftTangent,
                                          typename std::conditional<TANGENT, Tangent64, NoTangent>::type>::type;
    using Surv = typename std::conditional<RF, SurvivorRF, Survivor>::type;
    extern __shared__ uint32_t lds_counts[];   // [2][m] (then the tile, TILE only; the soft deposits, SOFT only)
    __shared__ unsigned long long lds_scal[kScalCount];
    for (int k = threadIdx.x; k < 2 * g.m; k += kBlock) lds_counts[k] = 0u;
    double *lds_dep = nullptr;   // SOFT: deposits [2][m], then their tangents [2][m]
    if constexpr (SOFT) {
        lds_dep = reinterpret_cast<double *>(reinterpret_cast<char *>(lds_counts) + tile_offset(g.m));
        for (int k = threadIdx.x; k < 4 * g.m; k += kBlock) lds_dep[k] = 0.0;
    }
    if (threadIdx.x < kScalCount) lds_scal[threadIdx.x] = 0ull;
    TileView tview;
    if constexpr (TILE) {
        static_cast<Grid &>(tview) = g;
        double2 *te = reinterpret_cast<double2 *>(reinterpret_cast<char *>(lds_counts) + tile_offset(g.m));
        uint32_t *tv = reinterpret_cast<uint32_t *>(te + (size_t)g.tile_w * g.ny);
        const int cells = g.tile_w * g.ny, base = g.tile_c0 * g.ny;   // the tile's columns are contiguous
        for (int k = threadIdx.x; k < cells; k += kBlock) {
            const double4 r = g.cells[base + k];
            te[k] = make_double2(r.x, r.y);
        }
        for (int wd = threadIdx.x; wd < (cells + 31) / 32; wd += kBlock) {
            uint32_t bits = 0u;
            for (int b = 0; b < 32; ++b) {
                const int k = wd * 32 + b;
                if (k < cells && g.vac[base + k]) bits |= 1u << b;
            }
            tv[wd] = bits;
        }
        tview.te = te;
        tview.tv = tv;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int64_t q_begin = w.in ? 0 : w.begin;
    const int64_t q_end = w.in ? (int64_t)*w.in_count : w.end;
    int pid = -1;       // >= 0 particle (pools hold < 2^31), -1 idle, -2 no work left
    StateT<Real> s;
    int sp = 0, k = 0;            // species, step count
    int slot = -1, plen = 0;      // record mode: probe slot and path length
    double xp = 0.0, yp = 0.0;    // record mode: last recorded point
    RfClock rc{0.0, 0.0, 0.0, 0.0};
    Tan tn;
    if constexpr (TANGENT) tn.dfield = w.tan.dfield;
    if constexpr (SOFT) {
        tn.dphi = w.tan.soft.dphi;
        tn.inv_width = w.tan.soft.inv_width;
        tn.cut = w.tan.soft.cut;
        tn.dep = lds_dep;
        tn.ddep = lds_dep + 2 * g.m;
        tn.events = w.tan.soft.events;
        tn.n_events = w.tan.soft.n_events;
        tn.ev_cap = w.tan.soft.ev_cap;
    }
    CornerCache cache;
    // Steps this lane traced: each particle adds its step count when it ends
    // here and subtracts the count it arrived with (64-bit: exact).
    unsigned long long my_steps = 0ull;
    for (;;) {
        unsigned long long idle = __ballot(pid == -1);
        unsigned long long busy = __ballot(pid >= 0);
        if (idle != 0ull && (__popcll(idle) >= w.refill || busy == 0ull)) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(w.queue, (unsigned long long)__popcll(idle));
            base = __shfl(base, leader);
            if (pid == -1) {
                int64_t q = (int64_t)(base + (unsigned long long)__popcll(idle & below));
                q = w.order ? q_end - 1 - q : q_begin + q;
                if (q >= q_begin && q < q_end) {
                    if (w.in) {
                        const Surv &v = static_cast<const Surv *>(w.in)[q];
                        pid = v.pid;
                        s.x = (Real)v.x; s.y = (Real)v.y; s.vx = (Real)v.vx; s.vy = (Real)v.vy;
                        k = v.k;
                        my_steps -= (unsigned long long)k;
                        if constexpr (RF) {
                            rc.clock = v.clock;
                            rc.elapsed = v.elapsed;
                        }
                    } else {
                        pid = (int)q;
                        s.x = (Real)w.x[q]; s.y = (Real)w.y[q]; s.vx = (Real)w.vx[q]; s.vy = (Real)w.vy[q];
                        k = 0;
                        if constexpr (RF) {
                            rc.clock = py_mod(w.t0[q], g.period);
                            rc.elapsed = 0.0;
                        }
                    }
                    s.qm = (Real)w.qm[pid];
                    if constexpr (MODE == kModeF32) {
                        s.q = s.x * (float)g.inv_dx;
                        s.r = s.y * (float)g.inv_dx;
                    } else {
                        s.q = div_dx(s.x, g.dx, g.inv_dx);
                        s.r = div_dx(s.y, g.dx, g.inv_dx);
                    }
                    sp = w.spc[pid] != 0;   // a species byte other than 0 / 1 cannot index past the tallies
                    if constexpr (TANGENT) {   // single pass: every particle starts here
                        tn.dx = w.tan.dx0 ? w.tan.dx0[pid] : 0.0;
                        tn.dy = w.tan.dy0 ? w.tan.dy0[pid] : 0.0;
                        tn.dvx = w.tan.dvx0 ? w.tan.dvx0[pid] : 0.0;
                        tn.dvy = w.tan.dvy0 ? w.tan.dvy0[pid] : 0.0;
                        tn.dpx = tn.dx;
                        tn.dpy = tn.dy;
                        tn.face = -1;
                        if constexpr (SOFT) {
                            tn.W = 1.0;
                            tn.dW = 0.0;
                            tn.pend = false;
                            tn.sp = sp;
                            tn.pid = pid;
                        }
                    }
                    if constexpr (RECORD) {
                        xp = s.x;
                        yp = s.y;
                        slot = TANGENT ? -1 : w.rec.slots[pid];
                        plen = 0;
                        if (slot >= 0) {
                            double *path = w.rec.paths + 2 * (int64_t)slot * w.rec.stride;
                            path[0] = s.x;
                            path[1] = s.y;
                            plen = 1;
                        }
                    }
                } else {
                    pid = -2;
                }
            }
            busy = __ballot(pid >= 0);
        }
        if (busy == 0ull) break;
        if (w.prio_steps > 0) {
            // Waves carrying long-lived particles hold the launch's critical path:
            // let them issue first.
            if (__ballot(pid >= 0 && k >= w.prio_steps) != 0ull) __builtin_amdgcn_s_setprio(2);
            else __builtin_amdgcn_s_setprio(0);
        }
        if (pid >= 0) {
            for (int t = 0; t < w.check_steps; ++t) {
                int cell = -1;
                int st;
                if constexpr (MODE == kModeF32) {
                    st = step32<BARRIER, TAPERED>(g, s, cell);
                } else if constexpr (TILE) {
                    st = step64t<BARRIER, TAPERED, false, false>(tview, s, cell, xp, yp, rc, tn);
                } else if constexpr (MODE == kModeF64) {
                    st = step64t<BARRIER, TAPERED, false, false, Grid, Tan, true>(g, s, cell, xp, yp, rc, tn, &cache);
                } else {
                    st = step64t<BARRIER, TAPERED, RECORD, RF>(g, s, cell, xp, yp, rc, tn);
                    if constexpr (RECORD) {
                        if (slot >= 0) {
                            double *path = w.rec.paths + 2 * ((int64_t)slot * w.rec.stride + plen);
                            path[0] = xp;
                            path[1] = yp;
                            ++plen;
                        }
                    }
                }
                ++k;
                if (st != 0 || k >= w.max_steps) {
                    finish(g, w, lds_counts, lds_scal, pid, sp, st != 0 ? st : 3, cell);
                    if (RECORD && (!SOFT || w.rec.end)) {
                        w.rec.end[2 * pid] = xp;
                        w.rec.end[2 * pid + 1] = yp;
                        w.rec.vel[2 * pid] = s.vx;
                        w.rec.vel[2 * pid + 1] = s.vy;
                        if (slot >= 0) w.rec.path_len[slot] = plen;
                        if constexpr (RF) {
                            w.rec.flight[pid] = rc.ft;
                            w.rec.clock[pid] = rc.ck;
                        }
                    }
                    if constexpr (SOFT) {   // the remaining weight where the path ends on a wall
                        if (st == 1) tn.deposit(g, cell, tn.face == 1, xp, yp, tn.W, tn.dW, 1);
                    }
                    if constexpr (TANGENT) {
                        if (!SOFT || w.tan.dend) {
                            w.tan.dend[2 * pid] = tn.dpx;
                            w.tan.dend[2 * pid + 1] = tn.dpy;
                            w.tan.dvel[2 * pid] = tn.dvx;
                            w.tan.dvel[2 * pid + 1] = tn.dvy;
                            w.tan.face[pid] = (int8_t)tn.face;
                        }
                        if (!SOFT || w.tan.steps) w.tan.steps[pid] = k;
                    }
                    my_steps += (unsigned long long)k;
                    pid = -1;
                    break;
                }
                if (k >= w.step_cap) {   // carry to the next pass
                    unsigned slot_out = atomicAdd(w.out_count, 1u);
                    Surv v;
                    v.x = (double)s.x; v.y = (double)s.y; v.vx = (double)s.vx; v.vy = (double)s.vy;
                    if constexpr (RF) {
                        v.clock = rc.clock;
                        v.elapsed = rc.elapsed;
                    }
                    v.pid = pid;
                    v.k = k;
                    static_cast<Surv *>(w.out)[slot_out] = v;
                    my_steps += (unsigned long long)k;
                    pid = -1;
                    break;
                }
            }
        }
    }
    atomicAdd(&lds_scal[kScalSteps], my_steps);
    __syncthreads();
    for (int c = threadIdx.x; c < 2 * g.m; c += kBlock)
        if (lds_counts[c]) atomicAdd(&w.counts[c], lds_counts[c]);
    if constexpr (SOFT) {
        for (int c = threadIdx.x; c < 2 * g.m; c += kBlock) {
            if (lds_dep[c] != 0.0) unsafeAtomicAdd(&w.tan.soft.dep[c], lds_dep[c]);
            if (lds_dep[2 * g.m + c] != 0.0) unsafeAtomicAdd(&w.tan.soft.ddep[c], lds_dep[2 * g.m + c]);
        }
    }
    if (threadIdx.x >= kScalEsc0 && threadIdx.x < kScalCount && lds_scal[threadIdx.x])
        atomicAdd(&w.scal[threadIdx.x], lds_scal[threadIdx.x]);
}

// FP64 push.  At 4 waves per SIMD it keeps its 107 VGPRs unspilled; 5 waves
// (95 VGPRs, two spilled) measured 10.85 against 9.7-9.85 ms per AR 8 batch of
// 1e6 particles, 6 waves (80 VGPRs, 16 spilled) 11.9 ms.
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) push64_kernel(Grid g, Work w) {
    push_body<kModeF64, BARRIER, TAPERED>(g, w);
}

// FP64 push with the tile columns of the field staged in LDS (for passes
// whose few, long-lived particles wait on the latency of each step's loads).
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) push64_tile_kernel(Grid g, Work w) {
    push_body<kModeF64, BARRIER, TAPERED, true>(g, w);
}

// FP64 push with record mode (diagnostics; its extra registers stay unspilled).
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) push64_record_kernel(Grid g, Work w) {
    push_body<kModeF64Record, BARRIER, TAPERED>(g, w);
}

// FP64 RF full-cycle push, without and with record mode.
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) push64_rf_kernel(Grid g, Work w) {
    push_body<kModeF64Rf, BARRIER, TAPERED>(g, w);
}
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) push64_rf_record_kernel(Grid g, Work w) {
    push_body<kModeF64RfRecord, BARRIER, TAPERED>(g, w);
}

// FP64 record push with the forward-mode derivative of every path (single pass).
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) push64_tangent_kernel(Grid g, Work w) {
    push_body<kModeF64Tangent, BARRIER, TAPERED>(g, w);
}

// FP64 record push with a relaxed barrier decision and the forward-mode
// derivative of the face-smoothed weighted deposit (single pass; SoftIO).
template <bool TAPERED>
__global__ void __launch_bounds__(kBlock) push64_soft_kernel(Grid g, Work w) {
    push_body<kModeF64Soft, true, TAPERED>(g, w);
}

// FP32 push.
template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(kBlock) push32_kernel(Grid g, Work w) {
    push_body<kModeF32, BARRIER, TAPERED>(g, w);
}

// Corner records of every lattice key (see Grid::corners) from the cell
// records of one nx x ny grid whose cells start at cell0 and whose keys start
// at key0 in their arrays (both 0 for a single instance).
__global__ void pack_corners_kernel(const double4 *__restrict__ cells, const uint8_t *__restrict__ vac, int nx,
                                    int ny, int cell0, int key0, double2 *__restrict__ corners,
                                    uint8_t *__restrict__ corner_vac) {
    int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= (nx + 1) * (ny + 1)) return;
    int i = key / (ny + 1) - 1, j = key % (ny + 1) - 1;
    int i0 = i < 0 ? nx - 1 : i, i1 = i + 1 == nx ? 0 : i + 1;
    int j0 = j < 0 ? 0 : j, j1 = j + 1 >= ny ? ny - 1 : j + 1;
    int c[4] = {i0 * ny + j0, i1 * ny + j0, i0 * ny + j1, i1 * ny + j1};
    unsigned f = 0;
    for (int k = 0; k < 4; ++k) {
        double4 r = cells[cell0 + c[k]];
        corners[4 * (int64_t)(key0 + key) + k] = make_double2(r.x, r.y);
        f |= (vac[cell0 + c[k]] ? 1u : 0u) << k;
    }
    corner_vac[key0 + key] = (uint8_t)f;
}

// Packs the RF basis (two unit-voltage responses) into per-cell records.
__global__ void pack_rf_kernel(const double *__restrict__ ex, const double *__restrict__ ey,
                               const double *__restrict__ phi, const uint8_t *__restrict__ vac, int n,
                               double4 *__restrict__ rfcells, double2 *__restrict__ rfphi) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    bool v = vac[c] != 0;
    rfcells[c] = make_double4(v ? ex[c] : 0.0, v ? ey[c] : 0.0, v ? ex[n + c] : 0.0, v ? ey[n + c] : 0.0);
    rfphi[c] = make_double2(phi ? phi[c] : 0.0, phi ? phi[n + c] : 0.0);
}

// Packs the frozen field of one batch into per-cell records.
__global__ void pack_cells_kernel(const double *__restrict__ ex, const double *__restrict__ ey,
                                  const double *__restrict__ phi, const uint8_t *__restrict__ vac,
                                  int n, double4 *__restrict__ cells,
                                  float4 *__restrict__ cells32) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    bool v = vac[c] != 0;
    double e_x = v ? ex[c] : 0.0, e_y = v ? ey[c] : 0.0;
    cells[c] = make_double4(e_x, e_y, v ? 1.0 : 0.0, phi ? phi[c] : 0.0);
    cells32[c] = make_float4((float)e_x, (float)e_y, v ? 1.0f : 0.0f, phi ? (float)phi[c] : 0.0f);
}

__global__ void selftest_math_kernel(const double *a, const double *b, int64_t n, double *out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double x = a[k], y = b[k];
    out[9 * k + 0] = th_hypot(x, y);
    out[9 * k + 1] = sqrt(x);
    out[9 * k + 2] = x / y;
    out[9 * k + 3] = fmod(x, y);
    out[9 * k + 4] = th_sqrt_n(fabs(x));
    out[9 * k + 5] = th_div_n(x, y);
    out[9 * k + 6] = th_cfl_dt(1.75e-9, 1e-8, fabs(x), fabs(y));
    double h0, h1;
    th_hypot2(x, y, y * 0x1p-500, x, &h0, &h1);   // the second pair often outside the core's range
    out[9 * k + 7] = h0;
    out[9 * k + 8] = h1;
}

__global__ void copy_kernel(const double4 *__restrict__ src, double4 *__restrict__ dst, int64_t n) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

}  // namespace th
