// Reverse-mode (adjoint) pathwise derivative of the FP64 push.
//
// For each particle, given the cotangent of its recorded end point and of its
// final velocity, the kernel returns the cotangent of its initial state and
// adds the cotangent of the vacuum-masked field {Ex*vac, Ey*vac} at every cell
// the path's interpolations read: the transpose of the forward-mode tangent
// push (Tangent64 in th_push.hpp), with the same convention -- the discrete
// events of the path (cells, wall hits, barrier decisions, step count) are
// held fixed.
//
// One lane owns one particle.  A forward sweep re-traces the path with the
// push's own step (step64t, record mode) and stores the state at the start of
// every step on a tape in HBM (32 bytes per step; a wave's 64 particles share
// a region laid out step-major, so one step's stores and loads are contiguous);
// the backward sweep re-runs each step from its tape entry with a policy that
// captures the step's intermediate values, then applies the transposed
// hooks:
//   field   E = sum w_k e_k / sum w_k v_k   ->  e_k-bar += w_k Ebar / sw,
//           x-bar += d(w)/dx terms (the clamp of sum w v at 1e-12 held);
//   push    dt = cfl dx / (|v| + sqrt(2 dx |q/m E|) + 1), v' = v + a dt,
//           s = (v + v') dt / 2, x' = x + s;
//   reflect x' = x0, one velocity component flipped;
//   crossing the recorded point x0 + f s, f = (face - x0) / s.
// The field cotangents are summed in registers while a lane's consecutive
// steps share their four interpolation cells, and then with FP64 atomics
// (their order, and so the last bits of the sums, may vary between runs).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "th_push.hpp"

namespace th {

// Intermediate values of one FP64 record step, as step64t hands them to its
// tangent hooks.
struct StepCapture {
    int i0, i1, j0, j1;
    double tx, ty, w00, w10, w01, w11;
    double2 e00, e10, e01, e11;
    double v00, v10, v01, v11;
    bool clamped;
    double sw, ex, ey;
    double vx, vy, qm, ax, ay, dt, two_dx;
    bool reflected = false, face_x = false;
    bool crossed = false, x_face = false;
    double f_raw = 0.0, f = 0.0, sx = 0.0, sy = 0.0;

    template <class V>
    __device__ __forceinline__ void field(const V &, int i0_, int i1_, int j0_, int j1_, double tx_, double ty_,
                                          double w00_, double w10_, double w01_, double w11_, double2 e00_,
                                          double2 e10_, double2 e01_, double2 e11_, double v00_, double v10_,
                                          double v01_, double v11_, bool clamped_, double sw_, double ex_,
                                          double ey_) {
        i0 = i0_; i1 = i1_; j0 = j0_; j1 = j1_;
        tx = tx_; ty = ty_;
        w00 = w00_; w10 = w10_; w01 = w01_; w11 = w11_;
        e00 = e00_; e10 = e10_; e01 = e01_; e11 = e11_;
        v00 = v00_; v10 = v10_; v01 = v01_; v11 = v11_;
        clamped = clamped_; sw = sw_; ex = ex_; ey = ey_;
        reflected = false;
        crossed = false;
    }
    __device__ __forceinline__ void push(double vx_, double vy_, double qm_, double ax_, double ay_, double dt_,
                                         double two_dx_) {
        vx = vx_; vy = vy_; qm = qm_; ax = ax_; ay = ay_; dt = dt_; two_dx = two_dx_;
    }
    __device__ __forceinline__ void reflect(bool fx) {
        reflected = true;
        face_x = fx;
    }
    __device__ __forceinline__ void crossing(bool xf, double fr, double ff, double sx_, double sy_) {
        crossed = true;
        x_face = xf;
        f_raw = fr;
        f = ff;
        sx = sx_;
        sy = sy_;
    }
    __device__ __forceinline__ void point() { crossed = false; }
};

// Cotangent of a particle's state {x, y, vx, vy}.
struct Bar4 {
    double x, y, vx, vy;
};

// Field cotangent of the four interpolation cells a lane's consecutive steps
// share, summed in registers and added to the grid's with FP64 atomics only
// when the cells change (about every third step) and at the path's end.
struct FieldBar {
    int k00 = -1, k10 = -1, k01 = -1, k11 = -1;
    double a00x = 0.0, a00y = 0.0, a10x = 0.0, a10y = 0.0, a01x = 0.0, a01y = 0.0, a11x = 0.0, a11y = 0.0;
    __device__ __forceinline__ void flush(double *fbar) {
        if (k00 < 0 || !fbar) return;
        unsafeAtomicAdd(&fbar[2 * k00], a00x);
        unsafeAtomicAdd(&fbar[2 * k00 + 1], a00y);
        unsafeAtomicAdd(&fbar[2 * k10], a10x);
        unsafeAtomicAdd(&fbar[2 * k10 + 1], a10y);
        unsafeAtomicAdd(&fbar[2 * k01], a01x);
        unsafeAtomicAdd(&fbar[2 * k01 + 1], a01y);
        unsafeAtomicAdd(&fbar[2 * k11], a11x);
        unsafeAtomicAdd(&fbar[2 * k11 + 1], a11y);
        k00 = -1;
    }
    __device__ __forceinline__ void add(int c00, int c10, int c01, int c11, double w00, double w10, double w01,
                                        double w11, double bx, double by, double *fbar) {
        if (c00 != k00 || c10 != k10 || c01 != k01 || c11 != k11) {
            flush(fbar);
            k00 = c00; k10 = c10; k01 = c01; k11 = c11;
            a00x = a00y = a10x = a10y = a01x = a01y = a11x = a11y = 0.0;
        }
        a00x += w00 * bx; a00y += w00 * by;
        a10x += w10 * bx; a10y += w10 * by;
        a01x += w01 * bx; a01y += w01 * by;
        a11x += w11 * bx; a11y += w11 * by;
    }
};

// Transpose of one step.  `out` is the cotangent of the state after the step;
// `pend` is the cotangent of the point the step recorded: on the last step
// its end point (clipped to the face it crossed), on a step with a relaxed
// barrier decision the crossing point of its deposit (else zero).  (bvn_x,
// bvn_y) is an extra cotangent of the step's end velocity before a
// reflection flips it (the relaxed decision's dependence on the normal
// velocity).  Returns the cotangent of the state before the step and adds the
// field cotangent of the step's four corners to fb (then fbar, [N] {Ex, Ey},
// which may be null).
__device__ __forceinline__ Bar4 step_transpose(const StepCapture &c, Bar4 out, double pend_x, double pend_y,
                                               bool last, int ny, double inv_dx, double *fbar, FieldBar &fb,
                                               double bvn_x = 0.0, double bvn_y = 0.0) {
    double bsx = 0.0, bsy = 0.0, bx0 = 0.0, by0 = 0.0;
    double bvxn = out.vx, bvyn = out.vy;
    double bxo = out.x, byo = out.y;
    if (last || c.crossed) {
        if (c.crossed) {
            // p = x0 + f s (+ df s), df = -(x0 + f_raw s - face)/s on the crossed axis
            bx0 += pend_x;
            by0 += pend_y;
            const double bf = c.sx * pend_x + c.sy * pend_y;
            bsx += c.f * pend_x;
            bsy += c.f * pend_y;
            if (c.f_raw > 0.0 && c.f_raw < 1.0) {
                if (c.x_face) {
                    bx0 += -bf / c.sx;
                    bsx += -c.f_raw * bf / c.sx;
                } else {
                    by0 += -bf / c.sy;
                    bsy += -c.f_raw * bf / c.sy;
                }
            }
        } else {
            bxo += pend_x;
            byo += pend_y;
        }
    }
    if (c.reflected) {   // x' = x0, the crossed face's velocity component flipped
        bx0 += bxo;
        by0 += byo;
        if (c.face_x) bvxn = -bvxn; else bvyn = -bvyn;
    } else {             // x' = x0 + s
        bsx += bxo;
        bsy += byo;
        bx0 += bxo;
        by0 += byo;
    }
    bvxn += bvn_x;
    bvyn += bvn_y;
    // push: vxn = vx + ax dt, s = (v + vn) dt / 2
    const double vxn = c.vx + c.ax * c.dt, vyn = c.vy + c.ay * c.dt;
    double bvx = 0.5 * c.dt * bsx, bvy = 0.5 * c.dt * bsy;
    bvxn += 0.5 * c.dt * bsx;
    bvyn += 0.5 * c.dt * bsy;
    double bdt = 0.5 * (c.vx + vxn) * bsx + 0.5 * (c.vy + vyn) * bsy;
    bvx += bvxn;
    bvy += bvyn;
    double bax = c.dt * bvxn, bay = c.dt * bvyn;
    bdt += c.ax * bvxn + c.ay * bvyn;
    // dt = cfl dx / (hv + root + 1), hv = |v|, root = sqrt(2 dx ha), ha = |a|
    const double hv = th_hypot(c.vx, c.vy), ha = th_hypot(c.ax, c.ay);
    const double root = sqrt(c.two_dx * ha);
    const double k = -c.dt / (hv + root + 1.0);
    const double bhv = k * bdt, broot = k * bdt;
    const double bha = root > 0.0 ? 0.5 * c.two_dx * broot / root : 0.0;
    if (ha > 0.0) {
        bax += c.ax / ha * bha;
        bay += c.ay / ha * bha;
    }
    if (hv > 0.0) {
        bvx += c.vx / hv * bhv;
        bvy += c.vy / hv * bhv;
    }
    // a = q/m E
    const double bex = c.qm * bax, bey = c.qm * bay;
    // field: E = n / sw with n = sum w e, sw = sum w v (clamped: constant)
    const double bnx = bex / c.sw, bny = bey / c.sw;
    const double bsw = c.clamped ? 0.0 : -(c.ex * bex + c.ey * bey) / c.sw;
    if (fbar)
        fb.add(c.i0 * ny + c.j0, c.i1 * ny + c.j0, c.i0 * ny + c.j1, c.i1 * ny + c.j1, c.w00, c.w10, c.w01, c.w11,
               bnx, bny, fbar);
    const double b00 = c.e00.x * bnx + c.e00.y * bny + c.v00 * bsw;
    const double b10 = c.e10.x * bnx + c.e10.y * bny + c.v10 * bsw;
    const double b01 = c.e01.x * bnx + c.e01.y * bny + c.v01 * bsw;
    const double b11 = c.e11.x * bnx + c.e11.y * bny + c.v11 * bsw;
    const double btx = -(1.0 - c.ty) * b00 + (1.0 - c.ty) * b10 - c.ty * b01 + c.ty * b11;
    const double bty = -(1.0 - c.tx) * b00 - c.tx * b10 + (1.0 - c.tx) * b01 + c.tx * b11;
    Bar4 in;
    in.x = bx0 + btx * inv_dx;
    in.y = by0 + bty * inv_dx;
    in.vx = bvx;
    in.vy = bvy;
    return in;
}

struct AdjointIO {
    const double *x, *y, *vx, *vy, *qm;   // initial states [n]
    const int32_t *order;                  // lanes take particles in this order [n]
    const int32_t *steps;                  // path length of each particle [n]
    const int64_t *group_off;              // tape offset of each group of 64 positions of `order`
    double4 *tape;                         // per group, step-major: [step][64 lanes] start state {x, y, vx, vy}
    const double *bar_end, *bar_vel;       // [n][2] each, nullable
    double *bar_state;                     // [n][4] out
    double *bar_field;                     // [N][2] accumulated, nullable
    unsigned long long *mismatch;          // particles whose re-traced length differs from steps
    int64_t begin, end;                    // positions in `order` handled by this launch
    int max_steps;
};

template <bool BARRIER, bool TAPERED>
__global__ void __launch_bounds__(256) push64_adjoint_kernel(Grid g, AdjointIO a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // begin and stride are multiples of 64, so a wave's lanes hold the 64
    // positions of one group
    for (int64_t o = a.begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < a.end; o += stride) {
        const int pid = a.order[o];
        const int n_steps = a.steps[pid];
        double4 *tape = a.tape + a.group_off[o >> 6] + (o & 63);
        State64 s;
        s.x = a.x[pid]; s.y = a.y[pid]; s.vx = a.vx[pid]; s.vy = a.vy[pid]; s.qm = a.qm[pid];
        s.q = div_dx(s.x, g.dx, g.inv_dx);
        s.r = div_dx(s.y, g.dx, g.inv_dx);
        RfClock rc{0.0, 0.0, 0.0, 0.0};
        double xp = 0.0, yp = 0.0;
        int cell = -1, k = 0, st = 0;
        NoTangent none;
        // forward sweep: the tape holds the start state of every step (never
        // more than n_steps entries, at most the group's longest path)
        while (k < n_steps) {
            tape[64 * (int64_t)k] = make_double4(s.x, s.y, s.vx, s.vy);
            st = step64t<BARRIER, TAPERED, true, false>(g, s, cell, xp, yp, rc, none);
            ++k;
            if (st != 0 || k >= a.max_steps) break;
        }
        const bool ended = st != 0 || k >= a.max_steps;
        if (k != n_steps || !ended) {
            atomicAdd(a.mismatch, 1ull);
            if (a.bar_state)
                for (int c = 0; c < 4; ++c) a.bar_state[4 * (int64_t)pid + c] = 0.0;
            continue;
        }
        // backward sweep
        Bar4 b{0.0, 0.0, a.bar_vel ? a.bar_vel[2 * (int64_t)pid] : 0.0,
               a.bar_vel ? a.bar_vel[2 * (int64_t)pid + 1] : 0.0};
        const double pe_x = a.bar_end ? a.bar_end[2 * (int64_t)pid] : 0.0;
        const double pe_y = a.bar_end ? a.bar_end[2 * (int64_t)pid + 1] : 0.0;
        FieldBar fb;
        for (int j = n_steps - 1; j >= 0; --j) {
            const double4 t0 = tape[64 * (int64_t)j];
            State64 r;
            r.x = t0.x; r.y = t0.y; r.vx = t0.z; r.vy = t0.w; r.qm = s.qm;
            r.q = div_dx(r.x, g.dx, g.inv_dx);
            r.r = div_dx(r.y, g.dx, g.inv_dx);
            StepCapture cap;
            int cj = -1;
            (void)step64t<BARRIER, TAPERED, true, false>(g, r, cj, xp, yp, rc, cap);
            const bool last = j == n_steps - 1;
            b = step_transpose(cap, b, last ? pe_x : 0.0, last ? pe_y : 0.0, last, g.ny, g.inv_dx, a.bar_field, fb);
        }
        fb.flush(a.bar_field);
        if (a.bar_state) {
            double *o4 = a.bar_state + 4 * (int64_t)pid;
            o4[0] = b.x; o4[1] = b.y; o4[2] = b.vx; o4[3] = b.vy;
        }
    }
}

// StepCapture with the relaxed barrier decision of SoftTangent (the same
// margin t, probability p and cut, so a re-traced path takes the same
// branches): a step whose decision was relaxed records p, the normal velocity
// and the two cells of the potential step, and its deposit's cell, face and
// crossing point (the crossing itself through StepCapture::crossing).  W is
// the weight carried along the path (a forward sweep multiplies it by p at
// every relaxed decision).
struct SoftCapture : StepCapture {
    double inv_width = 0.0, cut = 0.0;
    double W = 1.0;
    bool pend = false, event = false;
    double p = 0.0, vn = 0.0, aq = 0.0, qmwf = 0.0;
    int ec = 0, ev = 0, e_cell = 0;
    bool e_xf = false;
    double exq = 0.0, eyq = 0.0;

    template <class V>
    __device__ __forceinline__ void field(const V &g, int i0_, int i1_, int j0_, int j1_, double tx_, double ty_,
                                          double w00_, double w10_, double w01_, double w11_, double2 e00_,
                                          double2 e10_, double2 e01_, double2 e11_, double v00_, double v10_,
                                          double v01_, double v11_, bool clamped_, double sw_, double ex_,
                                          double ey_) {
        StepCapture::field(g, i0_, i1_, j0_, j1_, tx_, ty_, w00_, w10_, w01_, w11_, e00_, e10_, e01_, e11_, v00_,
                           v10_, v01_, v11_, clamped_, sw_, ex_, ey_);
        pend = false;
        event = false;
    }
    __device__ __forceinline__ bool reflects(double have, double need, double vn_, double qm, double wf, int c, int v,
                                             bool face_x_) {
        (void)face_x_;
        pend = false;
        const double a = fabs(qm);
        const double t = (have - need) / a * inv_width;
        if (!(t > -cut && t < cut)) return have < need;
        p = 1.0 / (1.0 + exp(t));
        vn = vn_;
        aq = a;
        qmwf = qm * wf;
        ec = c;
        ev = v;
        pend = true;
        return true;
    }
    template <class V>
    __device__ __forceinline__ void relaxed(const V &, int c, bool xf, double xq, double yq) {
        event = true;
        e_cell = c;
        e_xf = xf;
        exq = xq;
        eyq = yq;
        pend = false;
        W = W * p;
    }
};
template <> struct is_soft_policy<SoftCapture> : std::true_type {};

// Inputs and outputs of the relaxed push's adjoint (th_trace_soft_adjoint).
struct SoftAdjointIO {
    const double *x, *y, *vx, *vy, *qm;    // initial states [n] (the range's own indices)
    const int8_t *spc;                     // [n]
    const int32_t *order;                  // lanes take particles in this order [n]
    const int32_t *steps;                  // path length of each particle [n]
    const int64_t *group_off;              // tape offset of each group of 64 positions of `order`
    double4 *tape;                         // per group, step-major start states
    const double *gbar;                    // [2][m] cotangent of the deposits
    double inv_width, cut;
    double *bar_state;                     // [n][4] out
    double *bar_field;                     // [N][2] accumulated, nullable
    double *bar_phi;                       // [N] accumulated, nullable
    unsigned long long *mismatch;
    int64_t begin, end;
    int max_steps;
};

// Adjoint of the relaxed-barrier push: for the cotangent gbar of its
// face-smoothed deposits, the cotangent of every particle's initial state,
// of the field and of the potential (through the relaxed decisions).  Along
// a path the weight obeys W' = W p at a relaxed decision, which deposits
// a = W (1 - p); backwards, with abar the deposit's cotangent and Wbar that
// of the weight after the decision: Wbar <- abar (1 - p) + Wbar p, pbar =
// W (Wbar - abar), tbar = -p (1 - p) pbar, which reaches the normal velocity
// (dt/dvn = vn / (|q/m| w)) and the potentials of the two cells of the
// barrier (dt/dphi_c = -q/m wface / (|q/m| w) = -dt/dphi_v); the deposit's
// crossing point is transposed like the last step's recorded point.  The
// weight before a decision is recovered as W / p.
template <bool TAPERED>
__global__ void __launch_bounds__(256) push64_soft_adjoint_kernel(Grid g, SoftAdjointIO a) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t o = a.begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < a.end; o += stride) {
        const int pid = a.order[o];
        const int n_steps = a.steps[pid];
        const int sp = a.spc[pid] != 0;
        double4 *tape = a.tape + a.group_off[o >> 6] + (o & 63);
        State64 s;
        s.x = a.x[pid]; s.y = a.y[pid]; s.vx = a.vx[pid]; s.vy = a.vy[pid]; s.qm = a.qm[pid];
        s.q = div_dx(s.x, g.dx, g.inv_dx);
        s.r = div_dx(s.y, g.dx, g.inv_dx);
        RfClock rc{0.0, 0.0, 0.0, 0.0};
        double xp = 0.0, yp = 0.0;
        int cell = -1, k = 0, st = 0;
        SoftCapture fw;
        fw.inv_width = a.inv_width;
        fw.cut = a.cut;
        while (k < n_steps) {
            tape[64 * (int64_t)k] = make_double4(s.x, s.y, s.vx, s.vy);
            st = step64t<true, TAPERED, true, false>(g, s, cell, xp, yp, rc, fw);
            ++k;
            if (st != 0 || k >= a.max_steps) break;
        }
        const bool ended = st != 0 || k >= a.max_steps;
        if (k != n_steps || !ended) {
            atomicAdd(a.mismatch, 1ull);
            if (a.bar_state)
                for (int c = 0; c < 4; ++c) a.bar_state[4 * (int64_t)pid + c] = 0.0;
            continue;
        }
        // the final absorption deposits the remaining weight at the recorded point
        double W = fw.W, Wbar = 0.0, pfx = 0.0, pfy = 0.0;
        if (st == 1) soft_face_vjp(g, a.gbar, sp, cell, fw.x_face, xp, yp, W, Wbar, pfx, pfy);
        Bar4 b{0.0, 0.0, 0.0, 0.0};
        FieldBar fb;
        for (int j = n_steps - 1; j >= 0; --j) {
            const double4 t0 = tape[64 * (int64_t)j];
            State64 r;
            r.x = t0.x; r.y = t0.y; r.vx = t0.z; r.vy = t0.w; r.qm = s.qm;
            r.q = div_dx(r.x, g.dx, g.inv_dx);
            r.r = div_dx(r.y, g.dx, g.inv_dx);
            SoftCapture cap;
            cap.inv_width = a.inv_width;
            cap.cut = a.cut;
            int cj = -1;
            (void)step64t<true, TAPERED, true, false>(g, r, cj, xp, yp, rc, cap);
            const bool last = j == n_steps - 1;
            double px = last ? pfx : 0.0, py = last ? pfy : 0.0, bvx = 0.0, bvy = 0.0;
            if (cap.event) {
                const double p = cap.p, Wb = W / p;   // the weight before this decision
                double ab, ex, ey;
                soft_face_vjp(g, a.gbar, sp, cap.e_cell, cap.e_xf, cap.exq, cap.eyq, Wb * (1.0 - p), ab, ex, ey);
                const double tbar = -p * (1.0 - p) * (Wb * (Wbar - ab));
                const double u = tbar * a.inv_width / cap.aq;
                if (cap.face_x) bvx = u * cap.vn; else bvy = u * cap.vn;
                if (a.bar_phi) {
                    unsafeAtomicAdd(&a.bar_phi[cap.ec], -u * cap.qmwf);
                    unsafeAtomicAdd(&a.bar_phi[cap.ev], u * cap.qmwf);
                }
                Wbar = ab * (1.0 - p) + Wbar * p;
                W = Wb;
                px += ex;
                py += ey;
            }
            b = step_transpose(cap, b, px, py, last, g.ny, g.inv_dx, a.bar_field, fb, bvx, bvy);
        }
        fb.flush(a.bar_field);
        if (a.bar_state) {
            double *o4 = a.bar_state + 4 * (int64_t)pid;
            o4[0] = b.x; o4[1] = b.y; o4[2] = b.vx; o4[3] = b.vy;
        }
    }
}

}  // namespace th
