/*
 * ORACLE -- test infrastructure only.  CPU restatement of trench01's FP64
 * particle push, used by tests, smoke() and the bench's cpu_baseline leg as
 * the checker.  The product (trench01_amd) never links or calls this file.
 *
 * Restates, in one scalar loop per particle, the algorithm of
 *   trench_trace.cu:17-39     interpolate  (vacuum-weighted bilinear field)
 *   trench_trace.cu:41-59     wall_entry   (first solid cell crossed in a step)
 *   trench_trace.cu:92-185    push / classify / barrier / absorb loop
 *   trench_charging_2d.py:492-509, :512-539, :542-702  (trace_cpu, same algorithm
 *                              vectorised; its NumPy operation order is what
 *                              the FP64 arithmetic below reproduces)
 *   trench_cuda.py:118-122    tally conventions (bincount per species)
 *
 * Arithmetic contract: every expression is evaluated left to right exactly as
 * the reference writes it, with contraction off (-ffp-contract=off), and hypot
 * is glibc's hypot -- the function NumPy's np.hypot calls -- so results are
 * bit-identical with trace_cpu on identical inputs.
 *
 * Record mode (trench_trace.cu:149-183, :189-192; trench_charging_2d.py:565-577,
 * :652-686) is restated: end points clipped to the crossed face, final
 * velocities, and probe paths.
 *
 * The RF full-cycle push (trench_charging_2d.py:556-602, :629-632, :654-682;
 * trench_rf.py:5-11; trench_trace.cu:61-68, :97-110, :131-136) is restated in
 * oracle_trace_rf, step-synchronously over the live particles as trace_cpu
 * runs it.  trace_cpu's acceleration bound is a NumPy matrix product
 * (voltage_bound @ |basis field|); the OpenBLAS kernel behind it forms
 * fma(b0, a0, b1*a1) for whole blocks of four elements and fma(b1, a1, b0*a0)
 * for the last n % 4 (single-threaded; checked on random data).  blas_tail=1
 * reproduces that per-position rounding, which depends on the particle's place
 * among the live ones; blas_tail=0 uses the block form for every particle,
 * the order a per-particle kernel can keep.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int wrap_i(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }
static inline int clip_i(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }
static inline double wrap_x(double a, double len) {
    double r = fmod(a, len);          /* NumPy float %: fmod, then shift negatives */
    return r < 0.0 ? r + len : r;
}

typedef struct {
    const double *ex, *ey, *phi, *wface;
    const uint8_t *vac;
    int nx, ny, tapered, barrier, record;
    double dx, cfl, lx, ly;
} grid_t;

/* Record-mode outputs of one particle (all nullable except when record is on). */
typedef struct {
    double *end;      /* [2] end point */
    double *vel;      /* [2] final velocity */
    double *path;     /* [2 * (max_steps + 1)] probe path, or NULL */
    int32_t *plen;    /* path length */
} rec_t;

/* trench_charging_2d.py:492-509 / trench_trace.cu:17-39 */
static void field_at(const grid_t *g, double x, double y, double *ox, double *oy) {
    double fx = x / g->dx - 0.5, fy = y / g->dx - 0.5;
    int i = (int)floor(fx), j = (int)floor(fy);
    double tx = fx - i, ty = fy - j;
    int i0 = wrap_i(i, g->nx), i1 = wrap_i(i + 1, g->nx);
    int j0 = clip_i(j, g->ny), j1 = clip_i(j + 1, g->ny);
    int c[4] = {i0 * g->ny + j0, i1 * g->ny + j0, i0 * g->ny + j1, i1 * g->ny + j1};
    double w[4] = {(1.0 - tx) * (1.0 - ty), tx * (1.0 - ty), (1.0 - tx) * ty, tx * ty};
    double ax = 0.0, ay = 0.0, aw = 0.0;
    for (int k = 0; k < 4; ++k) {
        int v = g->vac[c[k]] != 0;
        ax += w[k] * (v ? g->ex[c[k]] : 0.0);
        ay += w[k] * (v ? g->ey[c[k]] : 0.0);
        aw += w[k] * (double)v;
    }
    if (aw < 1e-12) aw = 1e-12;
    *ox = ax / aw;
    *oy = ay / aw;
}

/* trench_charging_2d.py:512-539 / trench_trace.cu:41-59 */
static void first_solid(const grid_t *g, double x0, double y0, double sx, double sy,
                        int *cx, int *cy, int *vx_, int *vy_, int *face_x) {
    int nx = g->nx, ny = g->ny;
    int ix0 = (int)(x0 / g->dx); if (ix0 > nx - 1) ix0 = nx - 1;
    int iy0 = clip_i((int)(y0 / g->dx), ny);
    int dix = (int)floor((x0 + sx) / g->dx) - ix0;
    int diy = clip_i((int)floor((y0 + sy) / g->dx), ny) - iy0;
    double tfx = dix != 0 ? ((ix0 + (dix > 0)) * g->dx - x0) / sx : INFINITY;
    double tfy = diy != 0 ? ((iy0 + (diy > 0)) * g->dx - y0) / sy : INFINITY;
    int x_first = tfx < tfy;
    int ax = x_first ? wrap_i(ix0 + dix, nx) : ix0;
    int ay = x_first ? iy0 : iy0 + diy;
    int last = dix != 0 && diy != 0 && g->vac[ax * ny + ay];
    *cx = last ? wrap_i(ix0 + dix, nx) : ax;
    *cy = last ? iy0 + diy : ay;
    *vx_ = last ? ax : ix0;
    *vy_ = last ? ay : iy0;
    *face_x = last ? !x_first : x_first;
}

/* One particle to termination.  Returns the status (1 absorbed, 2 escaped,
 * 3 still live after max_steps) and writes the hit cell and step count. */
static int push_one(const grid_t *g, int max_steps, double x, double y, double vx, double vy,
                    double qm, int32_t *hit, int64_t *nsteps, const rec_t *rec) {
    const double dx = g->dx;
    const int nx = g->nx, ny = g->ny;
    int status = 3;
    int plen = 0;
    double xp = x, yp = y;
    if (rec && rec->path) { rec->path[0] = x; rec->path[1] = y; plen = 1; }
    *nsteps = max_steps;
    for (int k = 0; k < max_steps; ++k) {
        double ex, ey;
        field_at(g, x, y, &ex, &ey);
        double ax = qm * ex, ay = qm * ey;
        double dt = g->cfl * dx / (hypot(vx, vy) + sqrt(2.0 * dx * hypot(ax, ay)) + 1.0);
        double vxn = vx + ax * dt, vyn = vy + ay * dt;
        double x0 = x, y0 = y;
        double sx = 0.5 * (vx + vxn) * dt, sy = 0.5 * (vy + vyn) * dt;
        x = wrap_x(x + sx, g->lx);
        y = y + sy;
        vx = vxn; vy = vyn;
        int escaped = y >= g->ly;
        int ix = (int)(x / dx); if (ix > nx - 1) ix = nx - 1;
        int iy = clip_i((int)(y / dx), ny);
        int hitting = !escaped && !g->vac[ix * ny + iy];
        int corner = 0;
        if (g->tapered && !hitting && !escaped) {
            int ix0 = (int)(x0 / dx); if (ix0 > nx - 1) ix0 = nx - 1;
            corner = ix != ix0 && iy != clip_i((int)(y0 / dx), ny);
        }
        int cx = ix, cy = iy, face_x = 0;
        if ((hitting && (g->barrier || g->record || g->tapered)) || corner) {
            int vcx, vcy;
            first_solid(g, x0, y0, sx, sy, &cx, &cy, &vcx, &vcy, &face_x);
            hitting = !g->vac[cx * ny + cy];
            if (hitting && g->barrier) {
                double vn = face_x ? vx : vy;
                int c = cx * ny + cy, v = vcx * ny + vcy;
                double dphi = g->wface[c] * (g->phi[c] - g->phi[v]);
                if (0.5 * vn * vn < qm * dphi) {          /* cannot climb the wall barrier */
                    if (face_x) vx = -vx; else vy = -vy;
                    x = x0; y = y0;
                    hitting = 0;
                }
            }
            ix = cx; iy = cy;
        }
        xp = x; yp = y;
        if (rec && hitting) {           /* end point on the crossed face */
            double frac;
            if (face_x) {
                double xb = (cx + (sx < 0.0)) * dx;
                xb += nearbyint((x0 - xb) / g->lx) * g->lx;
                frac = (xb - x0) / sx;
            } else {
                frac = ((cy + (sy < 0.0)) * dx - y0) / sy;
            }
            frac = fmin(fmax(frac, 0.0), 1.0);
            xp = wrap_x(x0 + frac * sx, g->lx);
            yp = y0 + frac * sy;
        }
        if (rec && escaped) {
            double frac = fmin(fmax((g->ly - y0) / sy, 0.0), 1.0);
            xp = wrap_x(x0 + frac * sx, g->lx);
            yp = g->ly;
        }
        if (rec && rec->path) { rec->path[2 * plen] = xp; rec->path[2 * plen + 1] = yp; ++plen; }
        if (hitting) { *hit = ix * ny + iy; *nsteps = k + 1; status = 1; break; }
        if (escaped) { *nsteps = k + 1; status = 2; break; }
    }
    if (rec) {
        rec->end[0] = xp; rec->end[1] = yp;
        rec->vel[0] = vx; rec->vel[1] = vy;
        if (rec->path) *rec->plen = plen;
    }
    return status;
}

/*
 * Trace n particles.  Outputs: status[n] (int8), hit[n] (int32, -1 unless
 * absorbed), optional steps[n]; tallies cnt_i/cnt_e[nx*ny] (int64, zeroed
 * here), escaped[2], *lost and *total_steps.  Record mode when endpoints and
 * final_velocity are given ([n][2] each); slots[n] (-1 or a probe index),
 * paths[n_probe][max_steps+1][2] and path_lengths[n_probe] then hold the
 * probe paths.  nthreads <= 0 uses OpenMP's default.  Returns 0, or -1 on a
 * species value other than 0/1.
 */
int oracle_trace(const double *x, const double *y, const double *vx, const double *vy,
                 const double *qm, const int8_t *spc, int64_t n,
                 const double *ex, const double *ey, const uint8_t *vac, int nx, int ny,
                 const double *phi, const double *wface,
                 double dx, double cfl, int max_steps, int tapered, int nthreads,
                 int8_t *status, int32_t *hit, int64_t *steps,
                 int64_t *cnt_i, int64_t *cnt_e, int64_t *escaped, int64_t *lost,
                 int64_t *total_steps,
                 double *endpoints, double *final_velocity, const int32_t *slots,
                 double *paths, int32_t *path_lengths) {
    for (int64_t k = 0; k < n; ++k)
        if (spc[k] != 0 && spc[k] != 1) return -1;
    int record = endpoints != NULL && final_velocity != NULL;
    grid_t g = {ex, ey, phi, wface, vac, nx, ny, tapered, phi != NULL && wface != NULL, record,
                dx, cfl, nx * dx, ny * dx};
    int64_t tot = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tot)
#endif
    for (int64_t k = 0; k < n; ++k) {
        int32_t h = -1;
        int64_t s = 0;
        rec_t rec, *rp = NULL;
        if (record) {
            int slot = slots ? slots[k] : -1;
            rec.end = endpoints + 2 * k;
            rec.vel = final_velocity + 2 * k;
            rec.path = slot >= 0 ? paths + (int64_t)slot * 2 * (max_steps + 1) : NULL;
            rec.plen = slot >= 0 ? path_lengths + slot : NULL;
            rp = &rec;
        }
        status[k] = (int8_t)push_one(&g, max_steps, x[k], y[k], vx[k], vy[k], qm[k], &h, &s, rp);
        hit[k] = h;
        if (steps) steps[k] = s;
        tot += s;
    }
    memset(cnt_i, 0, sizeof(int64_t) * (size_t)nx * ny);
    memset(cnt_e, 0, sizeof(int64_t) * (size_t)nx * ny);
    escaped[0] = escaped[1] = 0;
    *lost = 0;
    for (int64_t k = 0; k < n; ++k) {
        if (status[k] == 1) (spc[k] ? cnt_e : cnt_i)[hit[k]] += 1;
        else if (status[k] == 2) escaped[spc[k]] += 1;
        else *lost += 1;
    }
    *total_steps = tot;
    return 0;
}

/* ---- RF full-cycle push ------------------------------------------------- */

typedef struct {
    const double *ex[2], *ey[2];   /* basis fields of unit substrate / mask voltage */
    const double *phi[2];          /* basis potentials (barrier term) */
    const double *wave;            /* [n_wave][2] voltage deviations from the period mean */
    int64_t n_wave;
    double period, dt_max, bound[2];
} rf_t;

/* NumPy's float remainder (npy_divmod): the result takes the divisor's sign. */
static double py_mod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

/* trench_rf.py:5-11 -- periodic linear interpolation of the voltage wave. */
static void rf_coefficients(const rf_t *rf, double at, double *vs, double *vm) {
    double f = py_mod(at, rf->period) / rf->period * (double)rf->n_wave;
    double fl = floor(f);
    int64_t i = ((int64_t)fl) % rf->n_wave;
    if (i < 0) i += rf->n_wave;
    int64_t j = (i + 1) % rf->n_wave;
    double w = f - fl;
    *vs = rf->wave[2 * i] * (1.0 - w) + rf->wave[2 * j] * w;
    *vm = rf->wave[2 * i + 1] * (1.0 - w) + rf->wave[2 * j + 1] * w;
}

/* interp_field on an arbitrary field pair (trench_charging_2d.py:492-509). */
static void field_pair_at(const grid_t *g, const double *fx_, const double *fy_, double x, double y,
                          double *ox, double *oy) {
    grid_t h = *g;
    h.ex = fx_;
    h.ey = fy_;
    field_at(&h, x, y, ox, oy);
}

typedef struct {
    double x, y, vx, vy, qm, clock, elapsed;
    int64_t id;
    int plen;
} rf_state_t;

/* One RF step of one particle (trench_charging_2d.py:584-686).  tail selects
 * the rounding of the matrix-product bound (see the file header).  Returns 0
 * (live), 1 (absorbed, *cell) or 2 (escaped); in record mode also the step's
 * recorded point, flight time and phase clock. */
static int rf_step(const grid_t *g, const rf_t *rf, rf_state_t *s, int tail, int *cell,
                   double *xp_out, double *yp_out, double *flight, double *clock_out) {
    const double dx = g->dx;
    const int nx = g->nx, ny = g->ny;
    double ex, ey, e0x, e0y, e1x, e1y;
    field_at(g, s->x, s->y, &ex, &ey);
    double speed = hypot(s->vx, s->vy);
    field_pair_at(g, rf->ex[0], rf->ey[0], s->x, s->y, &e0x, &e0y);
    field_pair_at(g, rf->ex[1], rf->ey[1], s->x, s->y, &e1x, &e1y);
    double b0 = rf->bound[0], b1 = rf->bound[1];
    double bx, by;
    if (!tail) {
        bx = fabs(ex) + fma(b0, fabs(e0x), b1 * fabs(e1x));
        by = fabs(ey) + fma(b0, fabs(e0y), b1 * fabs(e1y));
    } else {
        bx = fabs(ex) + fma(b1, fabs(e1x), b0 * fabs(e0x));
        by = fabs(ey) + fma(b1, fabs(e1y), b0 * fabs(e0y));
    }
    double accel_bound = fabs(s->qm) * hypot(bx, by);
    double dt = g->cfl * dx / (speed + sqrt(2.0 * dx * accel_bound) + 1.0);
    if (!(dt < rf->dt_max) && !isnan(dt)) dt = rf->dt_max;   /* np.minimum(dt_max, dt) */
    double vs, vm;
    rf_coefficients(rf, s->clock + dt / 2.0, &vs, &vm);
    ex = ex + (vs * e0x + vm * e1x);
    ey = ey + (vs * e0y + vm * e1y);
    double ax = s->qm * ex, ay = s->qm * ey;
    s->clock = py_mod(s->clock + dt, rf->period);
    s->elapsed = s->elapsed + dt;
    double vx = s->vx, vy = s->vy;
    double vxn = vx + ax * dt, vyn = vy + ay * dt;
    double x0 = s->x, y0 = s->y;
    double sx = 0.5 * (vx + vxn) * dt, sy = 0.5 * (vy + vyn) * dt;
    double x = wrap_x(x0 + sx, g->lx);
    double y = y0 + sy;
    vx = vxn; vy = vyn;
    int escaped = y >= g->ly;
    int ix = (int)(x / dx); if (ix > nx - 1) ix = nx - 1;
    int iy = clip_i((int)(y / dx), ny);
    int hitting = !escaped && !g->vac[ix * ny + iy];
    int corner = 0;
    if (g->tapered && !hitting && !escaped) {
        int ix0 = (int)(x0 / dx); if (ix0 > nx - 1) ix0 = nx - 1;
        corner = ix != ix0 && iy != clip_i((int)(y0 / dx), ny);
    }
    int cx = ix, cy = iy, face_x = 0;
    if ((hitting && (g->barrier || g->record || g->tapered)) || corner) {
        int vcx, vcy;
        first_solid(g, x0, y0, sx, sy, &cx, &cy, &vcx, &vcy, &face_x);
        hitting = !g->vac[cx * ny + cy];
        if (hitting && g->barrier) {
            double vn = face_x ? vx : vy;
            int c = cx * ny + cy, v = vcx * ny + vcy;
            double dphi = g->wface[c] * (g->phi[c] - g->phi[v]);
            double cs, cm;
            rf_coefficients(rf, s->clock, &cs, &cm);
            dphi = dphi + g->wface[c] * (cs * (rf->phi[0][c] - rf->phi[0][v]) +
                                         cm * (rf->phi[1][c] - rf->phi[1][v]));
            if (0.5 * vn * vn < s->qm * dphi) {
                if (face_x) vx = -vx; else vy = -vy;
                x = x0; y = y0;
                hitting = 0;
            }
        }
        ix = cx; iy = cy;
    }
    s->x = x; s->y = y; s->vx = vx; s->vy = vy;
    if (g->record) {
        double xp = x, yp = y, tf = 1.0;
        if (hitting) {
            double frac;
            if (face_x) {
                double xb = (cx + (sx < 0.0)) * dx;
                xb += nearbyint((x0 - xb) / g->lx) * g->lx;
                frac = (xb - x0) / sx;
            } else {
                frac = ((cy + (sy < 0.0)) * dx - y0) / sy;
            }
            frac = fmin(fmax(frac, 0.0), 1.0);
            tf = frac;
            xp = wrap_x(x0 + frac * sx, g->lx);
            yp = y0 + frac * sy;
        }
        if (escaped) {
            double frac = fmin(fmax((g->ly - y0) / sy, 0.0), 1.0);
            tf = frac;
            xp = wrap_x(x0 + frac * sx, g->lx);
            yp = g->ly;
        }
        *xp_out = xp;
        *yp_out = yp;
        *flight = s->elapsed - (1.0 - tf) * dt;
        *clock_out = py_mod(s->clock - (1.0 - tf) * dt, rf->period);
    }
    if (hitting) { *cell = ix * ny + iy; return 1; }
    return escaped ? 2 : 0;
}

/*
 * RF push of n particles (trench_charging_2d.py:542-702 with rf).  Arguments as
 * oracle_trace, plus the basis fields rf_ex/rf_ey/rf_phi [2][nx*ny] (rf_phi
 * only with a barrier), the wave [n_wave][2], period, dt_max, the voltage
 * bound [2] (max |wave| per column) and each particle's launch time.  Record
 * mode adds flight_time[n] and phase_clock[n] (the clock at the recorded
 * point, in [0, period)).  Single-threaded.  Returns 0, or -1 on a species
 * value other than 0/1.
 */
int oracle_trace_rf(const double *x, const double *y, const double *vx, const double *vy,
                    const double *qm, const int8_t *spc, int64_t n,
                    const double *ex, const double *ey, const uint8_t *vac, int nx, int ny,
                    const double *phi, const double *wface,
                    double dx, double cfl, int max_steps, int tapered,
                    const double *rf_ex, const double *rf_ey, const double *rf_phi,
                    const double *wave, int64_t n_wave, double period, double dt_max,
                    const double *bound, const double *initial_time, int blas_tail,
                    int8_t *status, int32_t *hit, int64_t *steps,
                    int64_t *cnt_i, int64_t *cnt_e, int64_t *escaped, int64_t *lost,
                    int64_t *total_steps,
                    double *endpoints, double *final_velocity, const int32_t *slots,
                    double *paths, int32_t *path_lengths,
                    double *flight_time, double *phase_clock) {
    for (int64_t k = 0; k < n; ++k)
        if (spc[k] != 0 && spc[k] != 1) return -1;
    int record = endpoints != NULL && final_velocity != NULL;
    int64_t cells = (int64_t)nx * ny;
    grid_t g = {ex, ey, phi, wface, vac, nx, ny, tapered, phi != NULL && wface != NULL, record,
                dx, cfl, nx * dx, ny * dx};
    rf_t rf = {{rf_ex, rf_ex + cells}, {rf_ey, rf_ey + cells},
               {rf_phi, rf_phi ? rf_phi + cells : NULL}, wave, n_wave, period, dt_max,
               {bound[0], bound[1]}};
    rf_state_t *live = (rf_state_t *)malloc(sizeof(rf_state_t) * (size_t)(n > 0 ? n : 1));
    int64_t n_live = n;
    for (int64_t k = 0; k < n; ++k) {
        rf_state_t s = {x[k], y[k], vx[k], vy[k], qm[k], py_mod(initial_time[k], period), 0.0, k, 0};
        status[k] = 3;
        hit[k] = -1;
        if (steps) steps[k] = max_steps;
        if (record) {
            endpoints[2 * k] = x[k]; endpoints[2 * k + 1] = y[k];
            final_velocity[2 * k] = vx[k]; final_velocity[2 * k + 1] = vy[k];
            flight_time[k] = 0.0;
            phase_clock[k] = s.clock;
            int slot = slots ? slots[k] : -1;
            if (slot >= 0) {
                double *path = paths + (int64_t)slot * 2 * (max_steps + 1);
                path[0] = x[k]; path[1] = y[k];
                s.plen = 1;
            }
        }
        live[k] = s;
    }
    int64_t tot = 0;
    for (int step = 0; step < max_steps && n_live > 0; ++step) {
        int64_t tail_from = blas_tail ? n_live - n_live % 4 : n_live;
        int64_t kept = 0;
        for (int64_t r = 0; r < n_live; ++r) {
            rf_state_t s = live[r];
            int cell = -1;
            double xp = 0.0, yp = 0.0, ft = 0.0, ck = 0.0;
            int st = rf_step(&g, &rf, &s, r >= tail_from, &cell, &xp, &yp, &ft, &ck);
            int64_t k = s.id;
            ++tot;
            if (record) {
                endpoints[2 * k] = xp; endpoints[2 * k + 1] = yp;
                final_velocity[2 * k] = s.vx; final_velocity[2 * k + 1] = s.vy;
                flight_time[k] = ft;
                phase_clock[k] = ck;
                int slot = slots ? slots[k] : -1;
                if (slot >= 0) {
                    double *path = paths + ((int64_t)slot * (max_steps + 1) + s.plen) * 2;
                    path[0] = xp; path[1] = yp;
                    ++s.plen;
                    path_lengths[slot] = s.plen;
                }
            }
            if (st != 0) {
                status[k] = (int8_t)st;
                if (st == 1) hit[k] = cell;
                if (steps) steps[k] = step + 1;
            } else {
                live[kept++] = s;
            }
        }
        n_live = kept;
    }
    if (record && slots) {
        for (int64_t r = 0; r < n_live; ++r) {
            int slot = slots[live[r].id];
            if (slot >= 0) path_lengths[slot] = live[r].plen;
        }
    }
    free(live);
    memset(cnt_i, 0, sizeof(int64_t) * (size_t)cells);
    memset(cnt_e, 0, sizeof(int64_t) * (size_t)cells);
    escaped[0] = escaped[1] = 0;
    *lost = 0;
    for (int64_t k = 0; k < n; ++k) {
        if (status[k] == 1) (spc[k] ? cnt_e : cnt_i)[hit[k]] += 1;
        else if (status[k] == 2) escaped[spc[k]] += 1;
        else *lost += 1;
    }
    *total_steps = tot;
    return 0;
}

/* glibc hypot on arrays, for tests that pin the device hypot against it. */
void oracle_hypot(const double *a, const double *b, double *out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) out[k] = hypot(a[k], b[k]);
}
