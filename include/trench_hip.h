/*
 * trench_hip.h -- C-ABI of libtrench_hip.so, the MI355X (gfx950) backend of
 * trench01's per-batch charging hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes
 * (TH_OK == 0, negative on error, message in th_last_error()).  No exceptions
 * cross the boundary.  Every entry point selects the handle's device itself,
 * so calls may come from any host thread (one call at a time per handle).
 *
 * Grid arrays are (nx, ny) C-order float64 / uint8 (cell index ix*ny + iy),
 * exactly as the reference passes them (trench_charging_2d.py:185-226,
 * trench_cuda.py:61).  Particle arrays are length-n float64 (x, y, vx, vy, qm)
 * plus int8 species (0 ion, 1 electron) (trench_charging_2d.py:458-489).
 */
#ifndef TRENCH_HIP_H
#define TRENCH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TH_OK 0
#define TH_EINVAL (-1)   /* bad argument (shape, range, species, null pointer) */
#define TH_EHIP (-2)     /* HIP runtime failure */
#define TH_ENOMEM (-3)   /* device allocation failed */
#define TH_ESTATE (-4)   /* handle used in the wrong order (e.g. trace before fields) */

typedef struct th_tracer th_tracer;

/* Push options; mirror of the Params fields the reference kernel receives
 * (trench_cuda.py:107-116). */
typedef struct th_trace_params {
    double dx;            /* cell size [m] (Params.dx) */
    double cfl;           /* 0 < cfl <= 1 (Params.cfl) */
    int32_t max_steps;    /* >= 1 (Params.max_steps) */
    int32_t tapered;      /* bool(trench_taper_deg or (mask_t > 0 and mask_taper_deg)) */
    int32_t barrier;      /* 1: wall_model "barrier" (phi and wface must be set) */
    int32_t precision;    /* 0: FP64, bit-identical with the reference; 1: FP32 push */
    int32_t refill;       /* wave refill threshold in lanes, 0 = library default */
    int32_t chunk;        /* particles per launch when a stop flag is given, 0 = default */
    /* Schedule (no effect on results): */
    int32_t order;        /* 0 = particles in index order, 1 = reversed (electrons first) */
    int32_t prio_steps;   /* raise a wave's issue priority once a particle is this old, 0 = off */
    int32_t check_steps;  /* push steps between refill checks, 0 = library default */
    int32_t pass_steps;   /* multi-pass schedule: first pass stops particles at this many steps and
                             later passes continue the survivors with doubling caps; 0 = one pass */
    int32_t rf;           /* 1: RF full-cycle push (th_set_rf, launch times per particle), FP64 only */
    int32_t blocks_per_cu; /* schedule: resident workgroups per CU at most, 0 = as many as fit */
    int32_t tile_pass;    /* schedule: passes from this one on (1-based) stage the field of the columns
                             around the trench in LDS (FP64, non-RF, multi-pass only); 0 = never */
} th_trace_params;

/* Scalar tallies of one trace call (trench_cuda.py:119-122). */
typedef struct th_tally {
    int64_t escaped[2];   /* escaped through the top, per species */
    int64_t lost;         /* still live after max_steps, or not traced after a stop */
    int64_t steps;        /* sum over particles of push steps taken */
    int64_t off_surface;  /* absorbed in a solid cell that is not a surface cell */
    int32_t cancelled;    /* 1 if the stop flag ended the call early */
    int32_t reserved;
} th_tally;

/* Thread-local message of the last failing call on this thread. */
const char *th_last_error(void);

/* Number of visible HIP devices. */
int th_device_count(int32_t *count);

/* Replaces CudaTracer.__init__ (trench_cuda.py:24-39): validates the device,
 * uploads the vacuum map and builds the surface-cell index used by the
 * LDS-privatised deposit tallies (surface cells as in
 * trench_charging_2d.py:734-744). */
int th_tracer_create(const uint8_t *vac, int32_t nx, int32_t ny, int32_t device, th_tracer **out);
void th_tracer_destroy(th_tracer *t);

/* Device name (trench_cuda.py:37-39). */
int th_tracer_device_name(const th_tracer *t, char *buf, int32_t len);

/* Number of surface cells m and their flat indices (ascending), for callers
 * that keep charge as an m-vector. */
int th_tracer_surface(const th_tracer *t, int32_t *m, int32_t *flat /* nullable, length m */);

/* Uploads the frozen field of one batch: the field uploads of
 * trench_cuda.py:61 and :80.  phi and wface may both be NULL (absorb wall
 * model); otherwise both are required. */
int th_set_fields(th_tracer *t, const double *Ex, const double *Ey, const double *phi,
                  const double *wface);

/* One tracer call on host arrays: replaces CudaTracer._trace
 * (trench_cuda.py:50-138) for the non-record, non-RF case.  Outputs the
 * per-cell counts cnt_i / cnt_e (int64, nx*ny, overwritten), the scalar tally,
 * and optionally the per-particle status (1 absorbed, 2 escaped, 3 live) and
 * absorbed cell (-1 if none).  stop, if non-NULL, is polled between launches;
 * once it reads non-zero the untraced particles are reported lost. */
int th_trace(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
             const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
             const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally,
             int8_t *status /* nullable */, int32_t *hit /* nullable */);

/* Record mode: th_trace plus, per particle, the end point (clipped to the
 * face it crossed) and final velocity ([n][2] each), and the paths of the
 * probe particles: slots[n] holds each particle's probe index or -1,
 * paths is [n_probe][max_steps+1][2], path_lengths[n_probe].  Replaces the
 * record branch of CudaTracer._trace (trench_cuda.py:81-97, :123-137;
 * kernel trench_trace.cu:149-183, :189-192).  status and hit are required.
 * Untraced particles (after a stop) keep their initial point and velocity. */
int th_trace_record(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                    const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
                    const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally,
                    int8_t *status, int32_t *hit, double *endpoints, double *final_velocity,
                    const int32_t *slots, int32_t n_probe, double *paths, int32_t *path_lengths);

/* RF full-cycle push (field_mode "circuit_rf"): the basis of unit substrate
 * and mask voltage responses -- Ex, Ey and (for the barrier wall model) phi,
 * each [2][nx*ny] -- the voltage wave [n_wave][2] (deviation from the period
 * mean, substrate then mask), the period and the step cap dt_max.  Replaces
 * the rf uploads of CudaTracer._trace (trench_cuda.py:65-71) and validates as
 * trench_rf.validate (trench_rf.py:21-32).  n_wave = 0 clears the RF state. */
int th_set_rf(th_tracer *t, const double *Ex, const double *Ey, const double *phi /* nullable */,
              const double *wave, int64_t n_wave, double period, double dt_max);

/* th_trace with the RF full-cycle push (params.rf = 1): each particle starts
 * its phase clock at initial_time[k] mod period (trench_cuda.py:69;
 * trench_charging_2d.py:556-602, :629-632; kernel trench_trace.cu:97-110,
 * :131-136). */
int th_trace_rf(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                const double *qm, const int8_t *spc, const double *initial_time, int64_t n,
                const th_trace_params *p, const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e,
                th_tally *tally, int8_t *status /* nullable */, int32_t *hit /* nullable */);

/* th_trace_record with the RF push: also the flight time to the recorded
 * point and the phase clock there, in [0, period), per particle
 * (trench_charging_2d.py:575-577, :680-682; trench_cuda.py:134-137).
 * Untraced particles keep flight time 0 and their launch clock. */
int th_trace_rf_record(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                       const double *qm, const int8_t *spc, const double *initial_time, int64_t n,
                       const th_trace_params *p, const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e,
                       th_tally *tally, int8_t *status, int32_t *hit, double *endpoints, double *final_velocity,
                       const int32_t *slots, int32_t n_probe, double *paths, int32_t *path_lengths,
                       double *flight_time, double *phase_clock);

/* Tangent (forward-mode, pathwise) push: the record-mode FP64 trace of n
 * particles -- status, hit, endpoints and final_velocity equal those of
 * th_trace_record -- together with the derivative of every path along one
 * direction theta: initial-state tangents dx0, dy0, dvx0, dvy0 per particle
 * (each nullable = 0) and a field tangent dEx, dEy per cell (both nullable =
 * fixed field).  Outputs per particle: steps, the face crossed by the recorded
 * end point (1 x-face, 0 y-face, -1 none: still live), d_endpoints and
 * d_final_velocity ([n][2] each).  The discrete events of each path (cells,
 * wall hits, barrier decisions, step count) are held fixed.  No reference
 * counterpart: the derivative that the differentiation of the charging loop
 * asks for (gradient of the deposit with respect to the ion distribution and
 * the surface charge), built on the push of trench_trace.cu:17-185. */
int th_trace_tangent(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                     const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
                     const double *dx0, const double *dy0, const double *dvx0, const double *dvy0,
                     const double *dEx, const double *dEy, th_tally *tally, int8_t *status, int32_t *hit,
                     int32_t *steps, int8_t *face, double *endpoints, double *d_endpoints, double *final_velocity,
                     double *d_final_velocity);

/* One deposit of the relaxed-barrier push: weight a (tangent da) of particle
 * pid (pool index) absorbed in cell `cell` across an x-face (face 1) or a
 * y-face (face 0) at the crossing point (x, y), tangent (dx, dy); kind 0: the
 * absorbed share of a relaxed barrier decision, 1: the final absorption of
 * the particle's remaining weight; species 0 ion, 1 electron. */
typedef struct th_soft_event {
    double x, y, dx, dy, a, da;
    int32_t pid, cell;
    int8_t face, kind, species, pad0;
    int32_t pad1;
} th_soft_event;

/* Relaxed-barrier push with its forward-mode derivative, over the resident
 * pool[offset, offset+n) and the fields last set (barrier wall model, FP64).
 * The reference's barrier test (trench_charging_2d.py:623-640,
 * trench_trace.cu:128-143) reflects a particle when vn^2/2 < q/m dphi; here it
 * reflects with probability p = 1 / (1 + exp(t)), t = (vn^2/2 - q/m dphi) /
 * (|q/m| width) (width in volts), by splitting: the share 1 - p of the
 * particle's weight is deposited where it hit and the share p goes on along
 * the reflected path.  For |t| >= cut the sharp decision is taken (the
 * relaxed model's weights there differ from 0 or 1 by less than exp(-cut)).
 * Every deposit is face-smoothed (the hit cell and its neighbour along the
 * crossed face share it linearly in the crossing point).  deposit[2][m] and
 * d_deposit[2][m] (species-major, surface-cell order of th_tracer_surface)
 * receive the weighted deposits and their derivative along the direction
 * given by the initial-state tangents dx0..dvy0 (n each, nullable), the field
 * tangent dEx, dEy and the potential tangent dphi (nx*ny each, nullable).
 * events (nullable, event_capacity entries) lists each deposit; *n_events
 * (nullable) is the number produced, which may exceed the capacity (the
 * excess is not stored).  status / hit / steps (n each) are nullable.  The
 * path events other than the relaxed decisions (cells, walls, step counts)
 * are held fixed, as in th_trace_tangent.  No reference counterpart: the
 * relaxation makes the barrier decision differentiable, for Jacobian-vector
 * products of the charging map in a Newton-Krylov steady solve. */
int th_trace_soft(th_tracer *t, int64_t offset, int64_t n, const th_trace_params *p, double width, double cut,
                  const double *dx0, const double *dy0, const double *dvx0, const double *dvy0, const double *dEx,
                  const double *dEy, const double *dphi, th_tally *tally, double *deposit, double *d_deposit,
                  int8_t *status, int32_t *hit, int32_t *steps, th_soft_event *events, int64_t event_capacity,
                  int64_t *n_events);

/* Adjoint of th_trace_soft: for the cotangent gbar[2][m] of the relaxed push's
 * deposits (species-major, surface-cell order), the cotangent of every
 * particle's initial state bar_state[n][4] (x, y, vx, vy), of the vacuum-masked
 * field bar_field[nx*ny][2] (nullable) and of the potential bar_phi[nx*ny]
 * (nullable; through the relaxed barrier decisions), summed over particles and
 * steps, so that for any direction sum(gbar * d_deposit) = <bar_state, d
 * state0> + <bar_field, dE> + <bar_phi, dphi>.  Same particles (pool[offset,
 * offset+n)), fields, width and cut as the th_trace_soft call whose per-
 * particle step counts are `steps` (a different re-traced length is an
 * error).  Each path is re-traced once onto a tape in device memory (32
 * bytes per step, in chunks of at most 4 GiB) and swept back; the weight
 * before a relaxed decision is recovered from the weight after it.  *ms
 * (nullable) receives the kernels' time.  No reference counterpart, as
 * th_trace_soft. */
int th_trace_soft_adjoint(th_tracer *t, int64_t offset, int64_t n, const th_trace_params *p, double width,
                          double cut, const int32_t *steps, const double *gbar, double *bar_state, double *bar_field,
                          double *bar_phi, double *ms);

/* Adjoint (reverse-mode, pathwise) push: the transpose of th_trace_tangent's
 * derivative.  For n particles traced through the fields last set, given
 * the cotangent of each recorded end point bar_end[n][2] and of each final
 * velocity bar_vel[n][2] (each nullable = 0), returns the cotangent of each
 * initial state bar_state[n][4] (x, y, vx, vy) and, if bar_field is non-NULL,
 * the cotangent of the vacuum-masked field {Ex*vac, Ey*vac} per cell
 * bar_field[nx*ny][2], summed over all particles and steps: for any
 * direction, sum over particles of <bar_end, d end> + <bar_vel, d vel> equals
 * <bar_state, d state0> + <bar_field, d field>.  steps[n] are the particles'
 * path lengths from th_trace_tangent (or th_trace_record) of the same
 * particles and field; a particle whose re-traced length differs is an
 * error.  The discrete events of each path are held fixed, as in
 * th_trace_tangent.  Each path is re-traced once forward onto a tape in
 * device memory (32 bytes per step; in chunks of at most 4 GiB) and swept
 * back.  *ms (nullable) receives the kernels' time.  No reference
 * counterpart, as th_trace_tangent (built on the push of
 * trench_trace.cu:17-185). */
int th_trace_adjoint(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                     const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p, const int32_t *steps,
                     const double *bar_end, const double *bar_vel, double *bar_state, double *bar_field, double *ms);

/* Device-resident particle pool, so a batch can be traced with its states
 * already in HBM.  th_particles_reserve sizes the pool (contents are lost);
 * th_particles_upload copies n host particles to pool[offset, offset+n).
 * No reference counterpart: the reference uploads states on every call
 * (trench_cuda.py:60). */
int th_particles_reserve(th_tracer *t, int64_t capacity);
int th_particles_upload(th_tracer *t, int64_t offset, const double *x, const double *y, const double *vx,
                        const double *vy, const double *qm, const int8_t *spc, int64_t n);

/* Launch times of pool[offset, offset+n) for RF traces of the resident pool
 * (th_trace_resident and th_charging_step with params.rf = 1); the reference
 * draws them per batch (trench_charging_2d.py:1231-1234). */
int th_particles_upload_times(th_tracer *t, int64_t offset, const double *initial_time, int64_t n);

/* Traces pool[offset, offset+n) with the fields last set.  Same outputs as
 * th_trace; cnt_i / cnt_e may be NULL to leave the tallies on the device
 * only (read them with th_read_counts). */
int th_trace_resident(th_tracer *t, int64_t offset, int64_t n, const th_trace_params *p,
                      int64_t *cnt_i /* nullable */, int64_t *cnt_e /* nullable */, th_tally *tally,
                      int8_t *status /* nullable */, int32_t *hit /* nullable */);

/* Copies pool[offset, offset+n) back to the host (any pointer may be NULL). */
int th_particles_read(th_tracer *t, int64_t offset, int64_t n, double *x, double *y, double *vx, double *vy,
                      double *qm, int8_t *spc);

/* The field the push currently uses, as Ex, Ey on the grid (nx*ny each; zero
 * on solid cells, where the push never reads it). */
int th_read_field(th_tracer *t, double *Ex, double *Ey);

/* Per-surface-cell counts of the last trace: counts[0..m) ions, [m..2m) electrons. */
int th_read_counts(th_tracer *t, int32_t *counts);

/* GPU time of the push launches of the last trace call, from HIP events
 * recorded on the library's stream around them [ms]. */
int th_last_push_ms(const th_tracer *t, double *ms);

/* ---------------------------------------------------------------------------
 * Device-side injection.  Writes n ions then n electrons into
 * pool[offset, offset + 2n) with the distributions of the reference's host
 * injection (trench_charging_2d.py:458-489), drawn from a Philox4x32-10
 * stream keyed by (seed, batch, particle).  No reference counterpart (the
 * reference draws on the host with PCG64): statistically equivalent, not the
 * same draws.
 */
typedef struct th_inject_params {
    int64_t n;             /* particles per species */
    double lx, y0;         /* nx*dx, ny*dx - 1e-3*dx */
    double qm_i, qm_e;     /* e/m_i, -e/m_e */
    int32_t ion_kind;      /* 0: E = e_ion, lateral normal vth_i; 1: mono e_ion + polar angle;
                              2: arcsine IED on [e_lo, e_hi] + polar angle */
    int32_t reserved;
    double e_ion, e_lo, e_hi, sigma_theta, vth_i, v_per_sqrt_ev, s_e;
} th_inject_params;

int th_inject(th_tracer *t, int64_t offset, const th_inject_params *ip, uint64_t seed, uint64_t batch);

/* ---------------------------------------------------------------------------
 * Device-resident charging state: the per-batch path after the push
 * (trench_charging_2d.py:1239-1269) on the GPU.  Charge lives on the m surface
 * cells of the tracer; the potential is phi = G Q + phi_ext on the kept rows,
 * with G the Green's matrix (columns of the reference's LU solve, B at
 * :795-797, merged where surface cells share one response).
 */
typedef struct th_charging th_charging;

typedef struct th_charging_setup {
    int32_t n_cols;              /* distinct Green's columns */
    int32_t n_rows;              /* kept rows (cells whose potential the push reads) */
    const int32_t *col_of;       /* [m] surface index -> column, -1 for none */
    const int32_t *rows;         /* [n_rows] flat cell per row */
    const double *green;         /* [n_rows][n_cols], V per C/m */
    const double *external;      /* [nx*ny] potential of the applied voltages */
    const double *wface;         /* [nx*ny] barrier face weight (:1085) */
    const double *leak_M;        /* [m][m] backward-Euler leak operator (:768-803), NULL = no leak */
    const double *leak_c;        /* [m] dt G phi_ext on surface cells (:1250-1251), NULL = zero */
    const uint8_t *region;       /* [m] leak monitor region (:800-802), NULL = none */
    const uint8_t *cond_s;       /* [m] 1 on conductor surface cells */
    const double *g_cond_s;      /* [m] conductor response at surface cells (:807-812), floating only */
    double q_macro, dx, C_cond, Te, dt_batch, top_voltage, substrate_voltage;
    int32_t c_idx;               /* surface index of the conductor's charge cell (:1090) */
    int32_t floating, driven, top_dirichlet;
} th_charging_setup;

typedef struct th_batch_diag {
    th_tally tally;
    double leak_dq;              /* charge leaving the leak monitor region [C/m] */
    double drift, dv;            /* conductor drift and potential change [V] */
    double n_i_cond, n_e_cond;   /* macro-particles absorbed on the conductor */
    double push_ms, update_ms;   /* GPU time of the push and of the charge/field update */
} th_batch_diag;

int th_charging_create(th_tracer *t, const th_charging_setup *s, th_charging **out);
void th_charging_destroy(th_charging *c);

/* Precision of the potential update (no reference counterpart; the
 * reference solves with its FP64 LU, trench_charging_2d.py:1268): 0 FP64
 * Green's matrix; 1 the matrix split into BF16 hi + lo halves; 2 split into
 * FP16 hi + lo halves of the matrix scaled into FP16 range.  The split modes
 * form products and sums in FP32 and add the external potential in FP64.
 * 3 and 4: the BF16 / FP16 split on matrix cores (MFMA, FP32 accumulation),
 * with each charge split into two halves of the same type (FP16: after
 * scaling by a power of two into its range).  Recomputes the potential and
 * field of the current charge. */
int th_charging_set_green(th_charging *c, int32_t mode);

/* G Q for R (1..16) charge vectors at once on matrix cores, in the split
 * precision of mode 3 or 4 (th_charging_set_green): qcols[R][n_cols] column
 * charges [C/m], out[R][n_rows] potentials on the kept rows without the
 * external potential [V].  Reads G once for all R.  *ms (nullable): kernel
 * time.  No reference counterpart (the reference solves one right-hand side
 * per LU solve, trench_charging_2d.py:1268). */
int th_charging_potentials(th_charging *c, int32_t R, const double *qcols, double *out, double *ms);

/* Charge per macro-particle of the following batches [C/m] (the set-up's
 * q_macro, trench_charging_2d.py:1093-1094): a batch of f * n_per_batch
 * particles per species at q_macro / f has the expected deposit of the full
 * batch.  No reference counterpart (the reference's batch size is fixed). */
int th_charging_set_q_macro(th_charging *c, double q_macro);

/* Sets the surface charge density rho (m values, C/m^3) and recomputes the
 * potential and the push's field records from it. */
int th_charging_set_rho(th_charging *c, const double *rho_s);

/* One batch: traces pool[offset, offset+n) through the current field, then
 * deposit, leak, conductor, potential and field update, on the device. */
int th_charging_step(th_charging *c, int64_t offset, int64_t n, const th_trace_params *p, th_batch_diag *out);

/* Replicas of one batch from the current charge: pool[offset + r n,
 * offset + (r+1) n) holds replica r's particles (r < replicas).  One push
 * traces all of them through the current field; each replica's tallies then
 * update its own copy of the charge (deposit, leak, conductor), exactly as
 * th_charging_step would from the same charge: rho_out[replicas][m] receives
 * the replicas' charge densities, bit-identical with `replicas` separate
 * th_charging_step calls each preceded by th_charging_set_rho of the current
 * charge.  scal_out (nullable) receives per replica {escaped ions, escaped
 * electrons, lost, absorbed off the surface cells}; out the push's totals.
 * The charging state (charge, potential, field) is left unchanged.  No
 * reference counterpart: the batch body of trench_charging_2d.py:1225-1266
 * run for several particle sets at once, so the push's tail of long paths
 * is paid once for all of them. */
int th_charging_step_replicas(th_charging *c, int64_t offset, int64_t n, int32_t replicas,
                              const th_trace_params *p, double *rho_out, int64_t *scal_out, th_batch_diag *out);

/* `batches` (at most 1024) consecutive batches of the device-resident loop
 * without the host: each batch injects 2 * ip->n particles into
 * pool[offset, offset + 2 ip->n) on the device (th_inject's distributions and
 * stream, batch indices first_batch, first_batch + 1, ...), traces them and
 * updates charge, potential and field exactly as th_inject + th_charging_step,
 * and records a row: the monitors of th_charging_monitors (none if unset),
 * then {escaped ions, escaped electrons, lost, steps, absorbed off the
 * surface}, then {leak_dq, drift, dv, n_i_cond, n_e_cond} (th_batch_diag).
 * rows[batches][th_charging_row_size].  The batch is captured once as a
 * hipGraph (fill kernels, no memset nodes; re-captured when the offset,
 * schedule, injection, precision or monitors change) and replayed; *ms
 * (nullable) receives the GPU time of the replays.  No RF push.  No
 * reference counterpart: the batch loop of trench_charging_2d.py:1225-1285
 * with its host steps between batches removed. */
int th_charging_row_size(const th_charging *c, int32_t *size);
int th_charging_run(th_charging *c, int64_t offset, const th_inject_params *ip, uint64_t seed, uint64_t first_batch,
                    int32_t batches, const th_trace_params *p, double *rows, double *ms);

/* Per-batch monitors on the device (trench_charging_2d.py:1272-1285): after
 * every th_charging_step the library gathers the potential at n_probes cells,
 * the integrated charge sum(rho) dx^2 of n_regions regions (CSR lists of
 * surface indices), sum(|rho|) dx^2 of region abs_region, and the ion and
 * electron counts on the n_bottom floor cells (surface indices), and keeps
 * them on the host: th_charging_monitors copies the last batch's
 * n_probes + n_regions + 3 values (no device access). */
int th_charging_set_monitors(th_charging *c, int32_t n_probes, const int32_t *probe_cells, int32_t n_regions,
                             const int32_t *region_ptr, const int32_t *region_idx, int32_t abs_region,
                             int32_t n_bottom, const int32_t *bottom_s);
int th_charging_monitors(const th_charging *c, double *out);

/* Adjoint of the potential update: out[n_cols] = G^T w, the gradient of
 * sum over cells of w[cell] * phi[cell] (w on the nx*ny grid; only the kept
 * rows enter) with respect to the charge of each Green's column [V per C/m].
 * No reference counterpart (the reference has no derivatives); the transpose
 * of the G Q product behind phi = solve(-rho/EPS0) + external
 * (trench_charging_2d.py:1268, :795-797). */
int th_charging_green_adjoint(th_charging *c, const double *w, double *out);

/* th_charging_green_adjoint without any device allocation: same inputs,
 * outputs and kernels (bit-identical results), with the weights, the partial
 * sums and the result held in a fixed per-device array of 8 MiB (2^20
 * doubles) that the library keeps from load until the process ends.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Rejects with TH_EINVAL a handle whose nx*ny +
 * (ceil(n_rows/64) + 1) * n_cols exceeds 2^20.  Calls on one device share that
 * array, so calls that overlap on one device (from several threads or
 * handles) race; sequential calls do not. */
int th_charging_green_adjoint_noalloc(th_charging *c, const double *w, double *out);

/* Reads the state: rho (m), potential on the grid (nx*ny; cells outside the
 * kept rows read 0), and the last batch's per-surface-cell counts (2m). */
int th_charging_read(th_charging *c, double *rho_s, double *phi /* nullable */, int32_t *counts /* nullable */);

/* Device self-check of the FP64 arithmetic the push relies on: for each k,
 * out[9k..9k+8] = {hypot(a,b), sqrt(a), a/b, fmod(a,b), sqrt(|a|) and a/b in
 * their range-free forms, the CFL step 1.75e-9 / (|a| + sqrt(1e-8 |b|) + 1),
 * and hypot(a,b), hypot(b 2^-500, a) as one paired evaluation} computed on
 * the GPU with the kernels' own routines.  No reference
 * counterpart; used by tests to pin the device arithmetic against the host C
 * library. */
int th_selftest_math(th_tracer *t, const double *a, const double *b, int64_t n, double *out);

/* Streams `bytes` (a multiple of 32) from one device buffer to another with a
 * plain grid-stride copy kernel, `repeats` times, and reports the average
 * kernel time [ms].  No reference counterpart: it calibrates the HBM traffic
 * counters (FETCH_SIZE / WRITE_SIZE) against a known byte count. */
int th_probe_copy(th_tracer *t, int64_t bytes, int32_t repeats, double *ms);

/* Solves L U X = B in FP64 on `device` for k right-hand sides: B and X are
 * n-by-k, row-major, on the host.  L is unit lower triangular, given by its
 * strictly lower part in CSR (l_ptr[n+1], l_idx, l_val); U is upper
 * triangular, given by its strictly upper part in CSR (u_ptr, u_idx, u_val)
 * and its diagonal u_diag[n].  The caller applies the LU's row and column
 * permutations.  Replaces the host LU solve of the Green's columns, one
 * Poisson system per surface-charge column (make_solver's lu.solve on the
 * unit-charge block, trench_charging_2d.py:297-309, :795-797); the
 * factorisation itself stays on the host.  Each triangle's rows are grouped
 * by dependency level and solved one level per launch.  *ms (nullable)
 * receives the time of the launches. */
int th_lu_solve(int32_t device, int32_t n, const int32_t *l_ptr, const int32_t *l_idx, const double *l_val,
                const int32_t *u_ptr, const int32_t *u_idx, const double *u_val, const double *u_diag, int32_t k,
                const double *b, double *x, double *ms);

/* ---------------------------------------------------------------------------
 * Ensemble push: many trench instances (own grid, fields, taper, particle
 * count, ion distributions) traced together by one persistent launch per
 * pass.  No reference counterpart -- the reference traces one geometry per
 * call (CudaTracer.__call__, trench_cuda.py:41-138, one launch per 256-step
 * chunk); every instance's counts, escapes, losses and steps equal a
 * single-instance th_trace of the same particles bit for bit.  FP64, no RF,
 * one dx / cfl / max_steps / wall model for the whole ensemble.
 */
typedef struct th_ensemble th_ensemble;

int th_ensemble_create(int32_t device, th_ensemble **out);
void th_ensemble_destroy(th_ensemble *e);

/* Appends an instance: vacuum map and fields as th_tracer_create /
 * th_set_fields (phi, wface nullable together), `tapered` as
 * th_trace_params.tapered, and its batch size in particles (ions first, then
 * electrons).  Instances are added before the first upload, injection or run;
 * *index receives the instance's index. */
int th_ensemble_add(th_ensemble *e, const uint8_t *vac, int32_t nx, int32_t ny, const double *Ex, const double *Ey,
                    const double *phi, const double *wface, int32_t tapered, int64_t n_particles, int32_t *index);
int th_ensemble_set_fields(th_ensemble *e, int32_t index, const double *Ex, const double *Ey, const double *phi,
                           const double *wface);
/* Surface cells of an instance (as th_tracer_surface). */
int th_ensemble_surface(const th_ensemble *e, int32_t index, int32_t *m, int32_t *flat);
/* Device injection of an instance's batch (th_inject's distributions; ip->n
 * is half the instance's particle count). */
int th_ensemble_set_inject(th_ensemble *e, int32_t index, const th_inject_params *ip);
/* Host particle states of an instance (n = its particle count). */
int th_ensemble_upload(th_ensemble *e, int32_t index, const double *x, const double *y, const double *vx,
                       const double *vy, const double *qm, const int8_t *spc, int64_t n);

/* Runs `batches` batches and accumulates the tallies of every instance
 * (cleared first).  flags bit 0: inject every batch on the device (batch
 * indices first_batch, first_batch + 1, ...), else trace the uploaded states
 * each batch; bit 1: keep the absorbing cell per particle (modes 0 and 2).
 * mode 0: one launch per pass over all instances; 1: the same batch captured
 * once as a hipGraph and replayed; 2: one launch per instance and pass (the
 * reference's one-geometry-per-call schedule).  Results do not depend on the
 * mode; the elapsed time is th_ensemble_last_ms. */
int th_ensemble_run(th_ensemble *e, const th_trace_params *p, uint64_t seed, uint64_t first_batch, int32_t batches,
                    int32_t mode, int32_t flags);
/* Tallies of one instance after th_ensemble_run: counts on its grid (nx*ny,
 * nullable together), scalars, per-particle status of the last batch (1
 * absorbed, 2 escaped, 3 lost; nullable) and instance-local absorbing cell or
 * -1 (nullable; needs flags bit 1). */
int th_ensemble_read(th_ensemble *e, int32_t index, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally, int8_t *status,
                     int32_t *hit);
/* Tallies of all instances at once: counts [2][total surface slots] (ions,
 * then electrons; instance i's slots follow those of instances < i, in
 * th_ensemble_surface order) and per instance {escaped ions, escaped
 * electrons, lost, steps, absorbed off the surface cells} (nullable each). */
int th_ensemble_totals(th_ensemble *e, int64_t *counts, int64_t *scal);
/* The particle states of an instance in the pool (after injection: the batch's initial states). */
int th_ensemble_read_particles(th_ensemble *e, int32_t index, double *x, double *y, double *vx, double *vy,
                               double *qm, int8_t *spc);
int th_ensemble_last_ms(const th_ensemble *e, double *ms);

#ifdef __cplusplus
}
#endif
#endif
