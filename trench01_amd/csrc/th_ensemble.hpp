// Ensemble push: many trench instances (each its own grid, fields, taper and
// particle count) traced by one persistent launch.
//
// The records of all instances sit in shared arenas (cells, interpolation
// corners, vacuum map, wall weights, surface map); an instance is a table row
// with its grid size and its offsets into the arenas and the particle pool.
// The work queue runs over the whole pool, so a lane that finishes a particle
// of one instance takes the next particle of whatever instance comes next:
// small instances do not leave the chip idle and long-lived particles of one
// instance overlap with the bulk of the others.  Per step a lane reads its
// instance through a LaneView (uniform arena bases in scalar registers, the
// instance's grid size and offsets in vector registers) with the same FP64
// step as the single-instance push (step64 in th_push.hpp), so every instance
// gets the counts the single-instance push would give it, bit for bit.
//
// Tallies: the surface slots of all instances are numbered globally (an
// instance's slots are contiguous); absorptions are global integer atomics
// into [2][slots] counters, exact and independent of order, replicated so
// that workgroups spread over several copies (summed on readback).  LDS
// privatisation as in the single-instance push does not apply: the slots of
// hundreds of instances do not fit in LDS, and a workgroup's lanes hold
// particles of many instances at once.  Escapes, losses and step counts are
// summed per instance after the push from per-particle end records.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "th_inject.hpp"
#include "th_push.hpp"
#include "th_util.hpp"

namespace th {

enum : int { kInstEsc0 = 0, kInstEsc1, kInstLost, kInstSteps, kInstOff, kInstCount };

// One instance: grid size, offsets into the arenas and the pool.
struct Instance {
    int32_t nx, ny;
    int32_t cell0;     // first cell of its records in the cell arenas
    int32_t key0;      // first lattice key of its corner records
    int32_t tapered;   // corner crossings count (tapered walls)
    int32_t pad;
    int64_t p0, p1;    // pool range [p0, p1): ions, then electrons
};

struct EnsembleGrid {
    const double4 *__restrict__ cells;       // {Ex*vac, Ey*vac, vac, phi} per arena cell
    const double2 *__restrict__ corners;     // [keys][4]
    const uint8_t *__restrict__ corner_vac;  // [keys]
    const double *__restrict__ wface;        // per arena cell
    const uint8_t *__restrict__ vac;         // per arena cell
    const int32_t *__restrict__ slot;        // arena cell -> global surface slot, -1 elsewhere
    const Instance *__restrict__ inst;       // [n_inst]
    const int64_t *__restrict__ starts;      // [n_inst + 1]: p0 of every instance, then the pool size
    int n_inst, slots;
    double dx, inv_dx, cfl_dx, two_dx;
};

struct EnsembleWork {
    const double *__restrict__ x, *__restrict__ y, *__restrict__ vx, *__restrict__ vy, *__restrict__ qm;
    const int8_t *__restrict__ spc;
    int64_t begin, end;              // pool range of this launch (first pass)
    int8_t *status;                  // per particle: end status (1 absorbed, 2 escaped, 3 lost, 4 absorbed off-surface)
    int32_t *steps;                  // per particle: steps taken in total
    int32_t *hit;                    // nullable: instance-local absorbing cell
    uint32_t *counts;                // [replicas][2][slots]
    int replicas;                    // workgroup b adds into replica b % replicas
    unsigned long long *inst_scal;   // [n_inst][kInstCount]
    unsigned long long *queue;       // this launch's work counter (zero at launch)
    int max_steps, refill, check_steps, step_cap;
    const Survivor *in;              // later passes: survivors of the previous one
    const unsigned *in_count;
    Survivor *out;
    unsigned *out_count;
};

// The FP64 step's view of one lane's instance (see step64).
struct LaneView {
    const double4 *__restrict__ cells;
    const double2 *__restrict__ corners;
    const uint8_t *__restrict__ corner_vac;
    const double *__restrict__ wface;
    const uint8_t *__restrict__ vac;
    double dx, inv_dx, cfl_dx, two_dx;
    int nx, ny, cell0, key0;
    bool taper;
    __device__ __forceinline__ double4 cell(int c) const { return cells[cell0 + c]; }
    __device__ __forceinline__ const double2 *corner(int key) const { return corners + 4 * (int64_t)(key0 + key); }
    __device__ __forceinline__ unsigned corner_flags(int key) const { return corner_vac[key0 + key]; }
    static constexpr bool tiled = false;
    __device__ __forceinline__ bool is_vac(int c) const { return vac[cell0 + c] != 0; }
    __device__ __forceinline__ bool is_vac_at(int ix, int iy) const { return vac[cell0 + ix * ny + iy] != 0; }
    __device__ __forceinline__ double face(int c) const { return wface[cell0 + c]; }
    // nx*dx and ny*dx: the products the single-instance push forms on the host
    // (recomputed per use rather than held in four more vector registers)
    __device__ __forceinline__ double len_x() const { return (double)nx * dx; }
    __device__ __forceinline__ double len_y() const { return (double)ny * dx; }
};

// Index of the instance holding pool index q: the last i with starts[i] <= q.
__device__ __forceinline__ int find_instance(const int64_t *starts, int n_inst, int64_t q) {
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= q) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// A particle's end is recorded at the lane's next refill, not in the step
// loop: its status byte, its total step count and -- if absorbed on a surface
// cell -- one global atomic into the deposit counters.  Global stores and
// atomics inside the step loop, even on rarely taken branches, halve the push
// rate (they share the wait counter of the loads the step waits on); at the
// refill they overlap the refill's own queue atomic.  A lane ends at most one
// particle between refills, so one pending record per lane suffices.  Escapes,
// losses and steps per instance are summed afterwards by
// ensemble_reduce_kernel from the per-particle records.
enum : int8_t { kEndAbsorbed = 1, kEndEscaped = 2, kEndLost = 3, kEndOffSurface = 4 };

constexpr int kLdsStarts = 1024;   // instances whose pool starts are searched in LDS

template <bool BARRIER>
__device__ __forceinline__ void ensemble_body(const EnsembleGrid &g, const EnsembleWork &w) {
    // pool starts of the instances in LDS for the refill's binary search
    __shared__ int64_t lds_starts[kLdsStarts + 1];
    const bool starts_in_lds = g.n_inst <= kLdsStarts;
    if (starts_in_lds)
        for (int i = threadIdx.x; i <= g.n_inst; i += blockDim.x) lds_starts[i] = g.starts[i];
    __syncthreads();
    const int64_t *starts = starts_in_lds ? lds_starts : g.starts;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    // deposit counters of this workgroup's replica: fewer atomics meet on one address
    uint32_t *counts = w.counts + (int64_t)(blockIdx.x % w.replicas) * 2 * g.slots;
    const int64_t q_begin = w.in ? 0 : w.begin;
    const int64_t q_end = w.in ? (int64_t)*w.in_count : w.end;
    LaneView v;
    v.cells = g.cells; v.corners = g.corners; v.corner_vac = g.corner_vac; v.wface = g.wface; v.vac = g.vac;
    v.dx = g.dx; v.inv_dx = g.inv_dx; v.cfl_dx = g.cfl_dx; v.two_dx = g.two_dx;
    v.nx = 1; v.ny = 1; v.cell0 = 0; v.key0 = 0; v.taper = false;
    // pid >= 0: particle in flight; -1 idle; -2 no work left; <= -3: particle
    // -3 - pid reached the pass's step cap and is written to the survivor list
    // at the next refill (state intact meanwhile).
    int64_t pid = -1;
    int32_t end_pid = -1, end_k = 0, end_dep = -1;   // pending end record (end_pid < 0: none)
    int end_st = 0;
    State64 s;
    int sp = 0, k = 0;
    double xp = 0.0, yp = 0.0;
    RfClock rc{0.0, 0.0, 0.0, 0.0};
    for (;;) {
        unsigned long long idle = __ballot(pid == -1 || pid <= -3);
        unsigned long long busy = __ballot(pid >= 0);
        if (idle != 0ull && (__popcll(idle) >= w.refill || busy == 0ull)) {
            unsigned long long carry = __ballot(pid <= -3);
            if (carry != 0ull) {   // one slot reservation for the wave's survivors
                int cl = __ffsll((long long)carry) - 1;
                unsigned o = 0;
                if (lane == cl) o = atomicAdd(w.out_count, (unsigned)__popcll(carry));
                o = __shfl(o, cl);
                if (pid <= -3) {
                    Survivor c;
                    c.x = s.x; c.y = s.y; c.vx = s.vx; c.vy = s.vy;
                    c.pid = (int32_t)(-3 - pid);
                    c.k = k;
                    w.out[o + (unsigned)__popcll(carry & below)] = c;
                    pid = -1;
                }
            }
            if (end_pid >= 0) {
                if (end_dep >= 0) atomicAdd(&counts[end_dep], 1u);
                w.status[end_pid] = (int8_t)end_st;
                w.steps[end_pid] = end_k;
                end_pid = -1;
            }
            int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(w.queue, (unsigned long long)__popcll(idle));
            base = __shfl(base, leader);
            if (pid == -1) {
                int64_t q = q_begin + (int64_t)(base + (unsigned long long)__popcll(idle & below));
                if (q < q_end) {
                    if (w.in) {
                        const Survivor &c = w.in[q];
                        pid = c.pid;
                        s.x = c.x; s.y = c.y; s.vx = c.vx; s.vy = c.vy;
                        k = c.k;
                    } else {
                        pid = q;
                        s.x = w.x[q]; s.y = w.y[q]; s.vx = w.vx[q]; s.vy = w.vy[q];
                        k = 0;
                    }
                    const Instance &d = g.inst[find_instance(starts, g.n_inst, pid)];
                    v.nx = d.nx; v.ny = d.ny; v.cell0 = d.cell0; v.key0 = d.key0; v.taper = d.tapered != 0;
                    s.qm = w.qm[pid];
                    s.q = div_dx(s.x, g.dx, g.inv_dx);
                    s.r = div_dx(s.y, g.dx, g.inv_dx);
                    sp = w.spc[pid] != 0;   // a species byte other than 0 / 1 cannot index past the tallies
                } else {
                    pid = -2;
                }
            }
            busy = __ballot(pid >= 0);
        }
        if (busy == 0ull) break;
        if (pid >= 0) {
            for (int t = 0; t < w.check_steps; ++t) {
                int cell = -1;
                int st = step64<BARRIER, true, false, false>(v, s, cell, xp, yp, rc);
                ++k;
                if (st != 0 || k >= w.max_steps) {
                    end_pid = (int32_t)pid;
                    end_k = k;
                    end_st = st != 0 ? st : kEndLost;
                    end_dep = -1;
                    if (st == 1) {
                        int slot = g.slot[v.cell0 + cell];
                        if (slot >= 0) end_dep = sp * g.slots + slot;
                        else end_st = kEndOffSurface;
                        if (w.hit) w.hit[pid] = cell;
                    }
                    pid = -1;
                    break;
                }
                if (k >= w.step_cap) {   // carry to the next pass (written at the next refill)
                    pid = -3 - pid;
                    break;
                }
            }
        }
    }
}

// A contiguous range of one instance's particles, the unit of the
// reduction below.
struct Chunk {
    int64_t begin, end;
    int32_t inst, pad;
};

// Per-instance sums of the end records of one batch, added to the
// instance's scalar tallies: escapes per species, losses, absorptions off
// the surface cells, steps.  One workgroup per chunk (grid-stride), so large
// instances are summed by many workgroups.
__global__ void __launch_bounds__(256) ensemble_reduce_kernel(const Chunk *__restrict__ chunks, int n_chunks,
                                                               const int8_t *__restrict__ status,
                                                               const int32_t *__restrict__ steps,
                                                               const int8_t *__restrict__ spc,
                                                               unsigned long long *__restrict__ inst_scal) {
    __shared__ unsigned long long part[4][kInstCount];
    for (int c = blockIdx.x; c < n_chunks; c += gridDim.x) {
        const Chunk ch = chunks[c];
        unsigned long long acc[kInstCount] = {0ull, 0ull, 0ull, 0ull, 0ull};
        for (int64_t q = ch.begin + threadIdx.x; q < ch.end; q += blockDim.x) {
            int st = status[q];
            acc[kInstEsc0] += (st == kEndEscaped && spc[q] == 0) ? 1ull : 0ull;
            acc[kInstEsc1] += (st == kEndEscaped && spc[q] != 0) ? 1ull : 0ull;
            acc[kInstLost] += st == kEndLost ? 1ull : 0ull;
            acc[kInstOff] += st == kEndOffSurface ? 1ull : 0ull;
            acc[kInstSteps] += (unsigned long long)steps[q];
        }
        for (int f = 0; f < kInstCount; ++f)
            for (int off = 32; off > 0; off >>= 1) acc[f] += __shfl_xor(acc[f], off, 64);
        if ((threadIdx.x & 63) == 0)
            for (int f = 0; f < kInstCount; ++f) part[threadIdx.x >> 6][f] = acc[f];
        __syncthreads();
        if (threadIdx.x < kInstCount) {
            unsigned long long v = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                   part[3][threadIdx.x];
            if (v) atomicAdd(&inst_scal[(int64_t)ch.inst * kInstCount + threadIdx.x], v);
        }
        __syncthreads();
    }
}

// Two register budgets: 5 waves per SIMD (at most 102 VGPRs; a few values
// spill around the rare wall-entry path) or 4 (up to 128 VGPRs, no spills).
// Which is faster depends on the mix of instances; the launch picks by the
// schedule's workgroups-per-CU cap.
template <bool BARRIER>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
ensemble_push_kernel(EnsembleGrid g, EnsembleWork w) {
    ensemble_body<BARRIER>(g, w);
}
template <bool BARRIER>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
ensemble_push4_kernel(EnsembleGrid g, EnsembleWork w) {
    ensemble_body<BARRIER>(g, w);
}

// Device injection of every instance's batch: the particle at pool index k
// belongs to instance i = find_instance(k); it is particle k - p0 of that
// instance's batch (ions first), drawn with the instance's distributions from
// a Philox stream keyed by the seed and the instance index.  The batch index
// is read from *batch (so a captured graph advances it by itself).
__global__ void ensemble_inject_kernel(const InjectDev *__restrict__ params, const int64_t *__restrict__ starts,
                                       int n_inst, uint64_t seed, const unsigned long long *__restrict__ batch,
                                       double *x, double *y, double *vx, double *vy, double *qm, int8_t *spc) {
    const int64_t total = starts[n_inst];
    const uint64_t b = *batch;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        int i = find_instance(starts, n_inst, k);
        InjectDev d = params[i];
        d.batch = b;
        double px, pvx, pvy, pqm;
        int8_t sp;
        draw_particle(d, k - starts[i], stream_key(seed ^ (0x632BE59BD9B4E019ull * (uint64_t)(i + 1)), b), px, pvx,
                      pvy, pqm, sp);
        x[k] = px; y[k] = d.y0; vx[k] = pvx; vy[k] = pvy; qm[k] = pqm; spc[k] = sp;
    }
}

}  // namespace th
