// C-ABI of libtrench_hip.so (declarations and contracts in include/trench_hip.h).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/trench_hip.h"
#include "th_adjoint.hpp"
#include "th_charge.hpp"
#include "th_inject.hpp"
#include "th_lu.hpp"
#include "th_push.hpp"
#include "th_util.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define TH_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? TH_ENOMEM : TH_EHIP, "%s: %s (%s:%d)",    \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);                    \
    } while (0)

template <class T>
int dev_alloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    TH_HIP(hipMalloc((void **)p, count * sizeof(T)));
    return TH_OK;
}

// Device memory of th_charging_green_adjoint_noalloc: the weights staged from
// the host, the per-workgroup partial sums and the result, 8 MiB per device,
// held from module load until the process ends.  Host code never passes
// `scratch` itself to a kernel or a copy (on the host it names a stand-in,
// not device memory): it uses the address hipGetSymbolAddress returns for each
// device, resolved once per device with that device current.
constexpr size_t kScratchDoubles = (size_t)1 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.
int scratch_address(int device, double **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return fail(TH_EINVAL, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        TH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(TH_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return TH_OK;
}

}  // namespace

struct th_tracer {
    int device = 0, nx = 0, ny = 0, N = 0, m = 0, cus = 0;
    hipStream_t stream = nullptr;
    std::vector<uint8_t> vac;
    std::vector<int32_t> surf;   // flat indices of surface cells, ascending
    uint8_t *d_vac = nullptr;
    int32_t *d_smap = nullptr;
    double4 *d_cells = nullptr;
    float4 *d_cells32 = nullptr;
    double2 *d_corners = nullptr;      // interpolation corners per lattice key
    uint8_t *d_corner_vac = nullptr;
    double *d_wface = nullptr, *d_tmp = nullptr;   // d_tmp: 3*N upload staging
    bool fields = false, barrier_fields = false;
    // RF full-cycle push: basis records, voltage wave, clock parameters
    double4 *d_rfcells = nullptr;
    double2 *d_rfphi = nullptr, *d_wave = nullptr;
    int n_wave = 0;
    double period = 0.0, dt_max = 0.0, bound[2] = {0.0, 0.0};
    bool rf = false, rf_phi = false;
    // particle buffers
    int64_t cap = 0;
    double *d_x = nullptr, *d_y = nullptr, *d_vx = nullptr, *d_vy = nullptr, *d_qm = nullptr;
    int8_t *d_spc = nullptr, *d_status = nullptr;
    int32_t *d_hit = nullptr;
    double *d_t0 = nullptr;                          // RF launch times
    th::SurvivorRF *d_surv[2] = {nullptr, nullptr}; // multi-pass carry lists (ping-pong; the non-RF
                                                    // push stores 40-byte Survivor records in them)
    unsigned *d_surv_count = nullptr;                // per-pass list lengths
    // tallies
    uint32_t *d_counts = nullptr, *d_overflow = nullptr;
    unsigned long long *d_scal = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // around the push launches
    double push_ms = 0.0;
    int blocks_per_cu = 0;                     // launch cap of the current trace (0: occupancy limit)
    int tile_c0 = 0, tile_w = 0;               // columns the tile push stages in LDS (0: no tile)
    char name[256] = "";
};

namespace {

// Surface cells: solid cells with a vacuum cell among their 8 neighbours,
// x periodic, outside the y range counts as solid (trench_charging_2d.py:734-744).
void surface_cells(const std::vector<uint8_t> &vac, int nx, int ny, std::vector<int32_t> &out) {
    out.clear();
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            if (vac[ix * ny + iy]) continue;
            bool near = false;
            for (int sx = -1; sx <= 1 && !near; ++sx)
                for (int sy = -1; sy <= 1 && !near; ++sy) {
                    if (!sx && !sy) continue;
                    int jy = iy + sy;
                    if (jy < 0 || jy >= ny) continue;
                    int jx = ((ix + sx) % nx + nx) % nx;
                    near = vac[jx * ny + jy] != 0;
                }
            if (near) out.push_back(ix * ny + iy);
        }
}

constexpr int kMaxPasses = 32;
// d_scal: the push's kScalCount counters, then one work counter per pass.
constexpr int kScalAll = th::kScalCount + kMaxPasses + 1;

void free_particles(th_tracer *t) {
    for (void *p : {(void *)t->d_x, (void *)t->d_y, (void *)t->d_vx, (void *)t->d_vy, (void *)t->d_qm,
                    (void *)t->d_spc, (void *)t->d_status, (void *)t->d_hit, (void *)t->d_surv[0],
                    (void *)t->d_surv[1], (void *)t->d_surv_count, (void *)t->d_t0})
        if (p) (void)hipFree(p);
    t->d_x = t->d_y = t->d_vx = t->d_vy = t->d_qm = t->d_t0 = nullptr;
    t->d_spc = t->d_status = nullptr;
    t->d_hit = nullptr;
    t->d_surv[0] = t->d_surv[1] = nullptr;
    t->d_surv_count = nullptr;
    t->cap = 0;
}

int reserve_particles(th_tracer *t, int64_t n) {
    if (n <= t->cap && n > 0) return TH_OK;
    if (n > INT32_MAX) return fail(TH_EINVAL, "at most 2^31-1 particles per pool");
    free_particles(t);
    int64_t cap = std::max<int64_t>(n, 1024);
    int rc;
    if ((rc = dev_alloc(&t->d_x, cap)) || (rc = dev_alloc(&t->d_y, cap)) || (rc = dev_alloc(&t->d_vx, cap)) ||
        (rc = dev_alloc(&t->d_vy, cap)) || (rc = dev_alloc(&t->d_qm, cap)) || (rc = dev_alloc(&t->d_spc, cap)) ||
        (rc = dev_alloc(&t->d_status, cap)) || (rc = dev_alloc(&t->d_hit, cap)) || (rc = dev_alloc(&t->d_t0, cap)) ||
        (rc = dev_alloc(&t->d_surv[0], cap)) || (rc = dev_alloc(&t->d_surv[1], cap)) ||
        (rc = dev_alloc(&t->d_surv_count, kMaxPasses + 1))) {
        free_particles(t);
        return rc;
    }
    t->cap = cap;
    return TH_OK;
}

// Rebuilds the interpolation corners after the cell records changed.
int pack_corners(th_tracer *t) {
    int keys = (t->nx + 1) * (t->ny + 1);
    hipLaunchKernelGGL(th::pack_corners_kernel, dim3((keys + 255) / 256), dim3(256), 0, t->stream, t->d_cells,
                       t->d_vac, t->nx, t->ny, 0, 0, t->d_corners, t->d_corner_vac);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

th::Grid make_grid(const th_tracer *t, const th_trace_params *p) {
    th::Grid g;
    g.cells = t->d_cells;
    g.cells32 = t->d_cells32;
    g.corners = t->d_corners;
    g.corner_vac = t->d_corner_vac;
    g.wface = t->d_wface;
    g.vac = t->d_vac;
    g.smap = t->d_smap;
    g.nx = t->nx;
    g.ny = t->ny;
    g.m = t->m;
    g.dx = p->dx;
    g.inv_dx = 1.0 / p->dx;
    g.cfl_dx = p->cfl * p->dx;
    g.two_dx = 2.0 * p->dx;
    g.lx = t->nx * p->dx;
    g.ly = t->ny * p->dx;
    g.rfcells = t->d_rfcells;
    g.rfphi = t->d_rfphi;
    g.wave = t->d_wave;
    g.n_wave = t->n_wave;
    g.period = t->period;
    g.dt_max = t->dt_max;
    g.bound0 = t->bound[0];
    g.bound1 = t->bound[1];
    g.tile_c0 = t->tile_c0;
    g.tile_w = t->tile_w;
    return g;
}

// Persistent launch: as many workgroups as can be resident at once (the
// refill loop keeps them busy), never more than the particles need.
template <typename K>
int launch_kernel(th_tracer *t, K kernel, const th::Grid &g, const th::Work &w, int64_t want, size_t lds) {
    int per_cu = 0;
    TH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, th::kBlock, lds));
    if (t->blocks_per_cu > 0) per_cu = std::min(per_cu, t->blocks_per_cu);
    if (per_cu < 1) per_cu = 1;
    int blocks = (int)std::min<int64_t>(want, (int64_t)t->cus * per_cu);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(th::kBlock), lds, t->stream, g, w);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

template <bool B, bool T>
int launch_variant(th_tracer *t, const th::Grid &g, const th::Work &w, int64_t want, size_t lds, int precision,
                   bool record, bool rf) {
    if (w.tan.soft.dep) {   // relaxed-barrier push: barrier walls only, deposits in LDS after the counts
        if (!B) return fail(TH_EINVAL, "the relaxed-barrier push needs the barrier wall model");
        const size_t soft_lds = th::tile_offset(t->m) + sizeof(double) * 4 * (size_t)t->m;
        auto kernel = th::push64_soft_kernel<T>;
        TH_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)soft_lds));
        return launch_kernel(t, kernel, g, w, want, soft_lds);
    }
    if (w.tan.dend) return launch_kernel(t, th::push64_tangent_kernel<B, T>, g, w, want, lds);
    if (rf)
        return record ? launch_kernel(t, th::push64_rf_record_kernel<B, T>, g, w, want, lds)
                      : launch_kernel(t, th::push64_rf_kernel<B, T>, g, w, want, lds);
    if (precision == 1) return launch_kernel(t, th::push32_kernel<B, T>, g, w, want, lds);
    if (record) return launch_kernel(t, th::push64_record_kernel<B, T>, g, w, want, lds);
    return launch_kernel(t, th::push64_kernel<B, T>, g, w, want, lds);
}

// The FP64 push with the tile columns in LDS: one workgroup per CU (the tile
// fills most of the CU's LDS).
template <bool B, bool T>
int launch_tile_variant(th_tracer *t, const th::Grid &g, const th::Work &w, int64_t want) {
    const size_t lds = th::tile_offset(t->m) + th::tile_bytes(t->tile_w, t->ny);
    auto kernel = th::push64_tile_kernel<B, T>;
    TH_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return launch_kernel(t, kernel, g, w, want, lds);
}

int launch_tile(th_tracer *t, const th::Grid &g, const th::Work &w, bool barrier, bool tapered) {
    int64_t n = w.end - w.begin;
    if (n <= 0 && !w.in) return TH_OK;
    int64_t want = w.in ? (int64_t)t->cus : (n + th::kBlock - 1) / th::kBlock;
    if (barrier)
        return tapered ? launch_tile_variant<true, true>(t, g, w, want) : launch_tile_variant<true, false>(t, g, w, want);
    return tapered ? launch_tile_variant<false, true>(t, g, w, want) : launch_tile_variant<false, false>(t, g, w, want);
}

int launch_push(th_tracer *t, const th::Grid &g, const th::Work &w, bool barrier, bool tapered, int precision,
                bool record, bool rf) {
    int64_t n = w.end - w.begin;
    if (n <= 0) return TH_OK;
    size_t lds = sizeof(uint32_t) * 2 * (size_t)t->m;
    int64_t want = (n + th::kBlock - 1) / th::kBlock;
    if (barrier)
        return tapered ? launch_variant<true, true>(t, g, w, want, lds, precision, record, rf)
                       : launch_variant<true, false>(t, g, w, want, lds, precision, record, rf);
    return tapered ? launch_variant<false, true>(t, g, w, want, lds, precision, record, rf)
                   : launch_variant<false, false>(t, g, w, want, lds, precision, record, rf);
}

}  // namespace

extern "C" {

const char *th_last_error(void) { return g_err; }

int th_device_count(int32_t *count) {
    if (!count) return fail(TH_EINVAL, "count is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return TH_OK;
}

int th_tracer_create(const uint8_t *vac, int32_t nx, int32_t ny, int32_t device, th_tracer **out) {
    if (!out) return fail(TH_EINVAL, "out is NULL");
    *out = nullptr;
    if (!vac || nx < 1 || ny < 1) return fail(TH_EINVAL, "invalid vacuum map %dx%d", nx, ny);
    if ((int64_t)nx * ny > (1 << 30) || nx >= (1 << 24) || ny >= (1 << 24)) return fail(TH_EINVAL, "grid too large");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    if (device < 0 || device >= count)
        return fail(TH_EINVAL, "HIP device %d is not available (%d visible)", device, count);
    TH_HIP(hipSetDevice(device));
    th_tracer *t = new th_tracer();
    t->device = device;
    t->nx = nx;
    t->ny = ny;
    t->N = nx * ny;
    t->vac.assign(vac, vac + (size_t)t->N);
    for (auto &v : t->vac) v = v ? 1 : 0;
    surface_cells(t->vac, nx, ny, t->surf);
    t->m = (int)t->surf.size();
    std::vector<int32_t> smap((size_t)t->N, -1);
    for (int k = 0; k < t->m; ++k) smap[t->surf[k]] = k;
    hipDeviceProp_t prop;
    int rc = TH_OK;
    auto bad = [&](int code) { th_tracer_destroy(t); return code; };
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bad(fail(TH_EHIP, "hipGetDeviceProperties failed"));
    // the marketing name can be empty on some drivers; the ISA name is always set
    snprintf(t->name, sizeof t->name, "%s", prop.name[0] ? prop.name : prop.gcnArchName);
    t->cus = prop.multiProcessorCount;
    if (2 * (size_t)t->m * sizeof(uint32_t) > 96 * 1024)
        return bad(fail(TH_EINVAL, "%d surface cells exceed the LDS tally capacity", t->m));
    {
        // LDS tile of the late passes: the widest band of whole columns whose
        // field and vacuum bits fit the CU's LDS beside the tallies, centred on
        // the vacuum cells of the grid's lower half (the trench, where the
        // particles that live longest spend most of their steps).
        const long lds = (long)prop.maxSharedMemoryPerMultiProcessor - (long)th::tile_offset(t->m) - 512;
        int w = 0;
        while (w < nx && lds > 0 && th::tile_bytes(w + 1, ny) <= (size_t)lds) ++w;
        double sum = 0.0;
        long count = 0;
        for (int ix = 0; ix < nx; ++ix)
            for (int iy = 0; iy < ny / 2; ++iy)
                if (t->vac[(size_t)ix * ny + iy]) {
                    sum += ix + 0.5;
                    ++count;
                }
        const int centre = count ? (int)(sum / count) : nx / 2;
        t->tile_w = w >= 2 ? w : 0;
        t->tile_c0 = t->tile_w ? std::max(0, std::min(nx - t->tile_w, centre - t->tile_w / 2)) : 0;
    }
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(TH_EHIP, "hipStreamCreate failed"));
    if (hipEventCreate(&t->ev0) != hipSuccess || hipEventCreate(&t->ev1) != hipSuccess)
        return bad(fail(TH_EHIP, "hipEventCreate failed"));
    if ((rc = dev_alloc(&t->d_vac, t->N)) || (rc = dev_alloc(&t->d_smap, t->N)) ||
        (rc = dev_alloc(&t->d_cells, t->N)) || (rc = dev_alloc(&t->d_cells32, t->N)) ||
        (rc = dev_alloc(&t->d_corners, 4 * (size_t)(nx + 1) * (ny + 1))) ||
        (rc = dev_alloc(&t->d_corner_vac, (size_t)(nx + 1) * (ny + 1))) ||
        (rc = dev_alloc(&t->d_wface, t->N)) || (rc = dev_alloc(&t->d_tmp, 3 * (size_t)t->N)) ||
        (rc = dev_alloc(&t->d_counts, 2 * (size_t)std::max(t->m, 1))) ||
        (rc = dev_alloc(&t->d_overflow, 2 * (size_t)t->N)) || (rc = dev_alloc(&t->d_scal, kScalAll)))
        return bad(rc);
    if (hipMemcpy(t->d_vac, t->vac.data(), t->N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->d_smap, smap.data(), sizeof(int32_t) * t->N, hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail(TH_EHIP, "upload of the vacuum map failed"));
    double *sp = nullptr;   // per-device set-up of the allocation-free entry point's scratch
    if ((rc = scratch_address(device, &sp))) return bad(rc);
    *out = t;
    return TH_OK;
}

void th_tracer_destroy(th_tracer *t) {
    if (!t) return;
    (void)hipSetDevice(t->device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    free_particles(t);
    for (void *p : {(void *)t->d_vac, (void *)t->d_smap, (void *)t->d_cells, (void *)t->d_cells32,
                    (void *)t->d_corners, (void *)t->d_corner_vac,
                    (void *)t->d_wface, (void *)t->d_tmp, (void *)t->d_counts, (void *)t->d_overflow,
                    (void *)t->d_scal, (void *)t->d_rfcells, (void *)t->d_rfphi, (void *)t->d_wave})
        if (p) (void)hipFree(p);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

int th_tracer_device_name(const th_tracer *t, char *buf, int32_t len) {
    if (!t || !buf || len < 1) return fail(TH_EINVAL, "invalid argument");
    snprintf(buf, (size_t)len, "%s", t->name);
    return TH_OK;
}

int th_tracer_surface(const th_tracer *t, int32_t *m, int32_t *flat) {
    if (!t || !m) return fail(TH_EINVAL, "invalid argument");
    *m = t->m;
    if (flat) memcpy(flat, t->surf.data(), sizeof(int32_t) * (size_t)t->m);
    return TH_OK;
}

int th_set_fields(th_tracer *t, const double *Ex, const double *Ey, const double *phi, const double *wface) {
    if (!t || !Ex || !Ey) return fail(TH_EINVAL, "Ex and Ey are required");
    if ((phi == nullptr) != (wface == nullptr)) return fail(TH_EINVAL, "phi and wface go together");
    TH_HIP(hipSetDevice(t->device));
    size_t bytes = sizeof(double) * (size_t)t->N;
    TH_HIP(hipMemcpyAsync(t->d_tmp, Ex, bytes, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipMemcpyAsync(t->d_tmp + t->N, Ey, bytes, hipMemcpyHostToDevice, t->stream));
    if (phi) {
        TH_HIP(hipMemcpyAsync(t->d_tmp + 2 * (size_t)t->N, phi, bytes, hipMemcpyHostToDevice, t->stream));
        TH_HIP(hipMemcpyAsync(t->d_wface, wface, bytes, hipMemcpyHostToDevice, t->stream));
    }
    int threads = 256, blocks = (t->N + threads - 1) / threads;
    // FP32 records hold E/dx so the FP32 push works in cell units; dx is not
    // known here, so they are scaled when the push is launched (1.0 for now).
    hipLaunchKernelGGL(th::pack_cells_kernel, dim3(blocks), dim3(threads), 0, t->stream, t->d_tmp,
                       t->d_tmp + t->N, phi ? t->d_tmp + 2 * (size_t)t->N : nullptr, t->d_vac, t->N,
                       t->d_cells, t->d_cells32);
    TH_HIP(hipGetLastError());
    int rc = pack_corners(t);
    if (rc) return rc;
    TH_HIP(hipStreamSynchronize(t->stream));
    t->fields = true;
    t->barrier_fields = phi != nullptr;
    return TH_OK;
}

}  // extern "C"

namespace {

int check_params(const th_tracer *t, const th_trace_params *p) {
    if (!(p->cfl > 0.0 && p->cfl <= 1.0) || p->max_steps < 1)
        return fail(TH_EINVAL, "tracing needs 0 < cfl <= 1 and max_steps >= 1");
    if (!(p->dx > 0.0)) return fail(TH_EINVAL, "dx must be positive");
    if (p->precision != 0 && p->precision != 1) return fail(TH_EINVAL, "precision %d is not available", p->precision);
    if (p->blocks_per_cu < 0) return fail(TH_EINVAL, "blocks_per_cu must be >= 0");
    if (!t->fields) return fail(TH_ESTATE, "th_set_fields must precede tracing");
    if (p->barrier && !t->barrier_fields) return fail(TH_ESTATE, "barrier wall model needs phi and wface");
    if (p->rf) {
        if (!t->rf) return fail(TH_ESTATE, "th_set_rf must precede an RF trace");
        if (p->precision != 0) return fail(TH_EINVAL, "the RF push is FP64 only");
        if (p->barrier && !t->rf_phi) return fail(TH_ESTATE, "an RF barrier trace needs the basis potentials");
    }
    return TH_OK;
}

int upload(th_tracer *t, int64_t off, const double *x, const double *y, const double *vx, const double *vy,
           const double *qm, const int8_t *spc, int64_t n) {
    for (int64_t k = 0; k < n; ++k)
        if (spc[k] != 0 && spc[k] != 1) return fail(TH_EINVAL, "species must be 0 (ion) or 1 (electron)");
    size_t b8 = sizeof(double) * (size_t)n;
    TH_HIP(hipMemcpyAsync(t->d_x + off, x, b8, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipMemcpyAsync(t->d_y + off, y, b8, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipMemcpyAsync(t->d_vx + off, vx, b8, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipMemcpyAsync(t->d_vy + off, vy, b8, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipMemcpyAsync(t->d_qm + off, qm, b8, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipMemcpyAsync(t->d_spc + off, spc, (size_t)n, hipMemcpyHostToDevice, t->stream));
    return TH_OK;
}

// Clears the tallies and launches the push over pool[off, off+n) on the
// tracer's stream.  With a stop flag the range is cut into chunks and the flag
// is read between them; *traced is the end of the traced prefix.
int launch_trace(th_tracer *t, int64_t off, int64_t n, const th_trace_params *p, const volatile int32_t *stop,
                 bool want_status, bool want_hit, int64_t *traced, th_tally *tally,
                 const th::Record *rec = nullptr, const th::TangentIO *tan = nullptr) {
    t->blocks_per_cu = p->blocks_per_cu;
    TH_HIP(hipMemsetAsync(t->d_counts, 0, sizeof(uint32_t) * 2 * (size_t)std::max(t->m, 1), t->stream));
    TH_HIP(hipMemsetAsync(t->d_overflow, 0, sizeof(uint32_t) * 2 * (size_t)t->N, t->stream));
    TH_HIP(hipMemsetAsync(t->d_scal, 0, sizeof(unsigned long long) * kScalAll, t->stream));   // incl. pass queues
    *traced = off;
    TH_HIP(hipEventRecord(t->ev0, t->stream));
    if (n > 0) {
        if (want_status) TH_HIP(hipMemsetAsync(t->d_status + off, 3, (size_t)n, t->stream));
        if (want_hit) TH_HIP(hipMemsetAsync(t->d_hit + off, 0xff, sizeof(int32_t) * (size_t)n, t->stream));
        th::Grid g = make_grid(t, p);
        th::Work w;
        w.x = t->d_x; w.y = t->d_y; w.vx = t->d_vx; w.vy = t->d_vy; w.qm = t->d_qm; w.spc = t->d_spc;
        w.t0 = t->d_t0;
        w.status = want_status ? t->d_status : nullptr;
        w.hit = want_hit ? t->d_hit : nullptr;
        w.counts = t->d_counts;
        w.overflow = t->d_overflow;
        w.scal = t->d_scal;
        w.max_steps = p->max_steps;
        w.refill = p->refill > 0 ? std::min(p->refill, 64) : 16;
        w.order = p->order != 0;
        w.prio_steps = p->prio_steps > 0 ? p->prio_steps : 0;
        w.check_steps = p->check_steps > 0 ? std::min(p->check_steps, 256) : th::kStepsPerCheck;
        w.rec = rec ? *rec : th::Record{};
        w.tan = tan ? *tan : th::TangentIO{};
        w.step_cap = p->max_steps;
        w.in = nullptr;
        w.in_count = nullptr;
        w.out = nullptr;
        w.out_count = nullptr;
        w.queue = t->d_scal + th::kScalQueue;
        int rc;
        const bool passes = !stop && !rec && !tan && p->pass_steps > 0 && p->pass_steps < p->max_steps;
        if (passes) {
            // Multi-pass: caps pass_steps, 2*pass_steps, ... up to max_steps; pass k
            // continues the survivors of pass k-1.  The passes are stream-ordered;
            // each reads its list length from device memory.
            TH_HIP(hipMemsetAsync(t->d_surv_count, 0, sizeof(unsigned) * (kMaxPasses + 1), t->stream));
            w.begin = off;
            w.end = off + n;
            int64_t cap = p->pass_steps;
            for (int pass = 0;; ++pass) {
                bool last = cap >= p->max_steps || pass == kMaxPasses - 1;
                w.step_cap = last ? p->max_steps : (int)cap;
                w.in = pass ? t->d_surv[(pass - 1) & 1] : nullptr;
                w.in_count = pass ? t->d_surv_count + pass - 1 : nullptr;
                w.out = last ? nullptr : t->d_surv[pass & 1];
                w.out_count = last ? nullptr : t->d_surv_count + pass;
                w.queue = t->d_scal + th::kScalCount + pass;   // cleared above
                const bool tile = p->tile_pass > 0 && pass + 1 >= p->tile_pass && t->tile_w > 0 && p->precision == 0 &&
                                  !p->rf;
                if ((rc = tile ? launch_tile(t, g, w, p->barrier != 0, p->tapered != 0)
                               : launch_push(t, g, w, p->barrier != 0, p->tapered != 0, p->precision, false,
                                             p->rf != 0)))
                    return rc;
                if (last) break;
                cap = std::min<int64_t>(2 * cap, p->max_steps);
            }
            *traced = off + n;
        } else {
            int64_t chunk = stop ? (p->chunk > 0 ? p->chunk : 65536) : n;
            for (int64_t b = off; b < off + n; b += chunk) {
                if (stop && *stop) {
                    tally->cancelled = 1;
                    break;
                }
                w.begin = b;
                w.end = std::min(off + n, b + chunk);
                TH_HIP(hipMemsetAsync(t->d_scal, 0, sizeof(unsigned long long), t->stream));   // work counter
                if ((rc = launch_push(t, g, w, p->barrier != 0, p->tapered != 0, p->precision, rec != nullptr,
                                      p->rf != 0)))
                    return rc;
                if (stop) TH_HIP(hipStreamSynchronize(t->stream));
                *traced = w.end;
            }
        }
    }
    TH_HIP(hipEventRecord(t->ev1, t->stream));
    return TH_OK;
}

void fill_tally(const unsigned long long *scal, int64_t untraced, th_tally *tally) {
    tally->escaped[0] = (int64_t)scal[th::kScalEsc0];
    tally->escaped[1] = (int64_t)scal[th::kScalEsc1];
    tally->lost = (int64_t)scal[th::kScalLost] + untraced;
    tally->steps = (int64_t)scal[th::kScalSteps];
    tally->off_surface = (int64_t)scal[th::kScalOff];
}

// Traces pool[off, off+n) and collects the tallies (see th_trace).
int trace_range(th_tracer *t, int64_t off, int64_t n, const th_trace_params *p, const volatile int32_t *stop,
                int64_t *cnt_i, int64_t *cnt_e, th_tally *tally, int8_t *status, int32_t *hit) {
    memset(tally, 0, sizeof *tally);
    if (cnt_i) memset(cnt_i, 0, sizeof(int64_t) * (size_t)t->N);
    if (cnt_e) memset(cnt_e, 0, sizeof(int64_t) * (size_t)t->N);
    int64_t traced = off;
    int rc = launch_trace(t, off, n, p, stop, status != nullptr, hit != nullptr, &traced, tally);
    if (rc) return rc;
    std::vector<uint32_t> counts(2 * (size_t)std::max(t->m, 1));
    unsigned long long scal[th::kScalCount];
    TH_HIP(hipMemcpyAsync(scal, t->d_scal, sizeof scal, hipMemcpyDeviceToHost, t->stream));
    if (cnt_i || cnt_e)
        TH_HIP(hipMemcpyAsync(counts.data(), t->d_counts, sizeof(uint32_t) * counts.size(),
                              hipMemcpyDeviceToHost, t->stream));
    if (status && n) TH_HIP(hipMemcpyAsync(status, t->d_status + off, (size_t)n, hipMemcpyDeviceToHost, t->stream));
    if (hit && n)
        TH_HIP(hipMemcpyAsync(hit, t->d_hit + off, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    float ms = 0.0f;
    TH_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->push_ms = ms;
    if (cnt_i && cnt_e) {
        for (int k = 0; k < t->m; ++k) {
            cnt_i[t->surf[k]] += counts[k];
            cnt_e[t->surf[k]] += counts[t->m + k];
        }
        if (scal[th::kScalOff]) {
            std::vector<uint32_t> over(2 * (size_t)t->N);
            TH_HIP(hipMemcpy(over.data(), t->d_overflow, sizeof(uint32_t) * over.size(), hipMemcpyDeviceToHost));
            for (int c = 0; c < t->N; ++c) {
                cnt_i[c] += over[c];
                cnt_e[c] += over[t->N + c];
            }
        }
    }
    fill_tally(scal, off + n - traced, tally);
    return TH_OK;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// NumPy's float remainder for a divisor L > 0 (npy_divmod).
double host_py_mod(double a, double L) {
    double r = fmod(a, L);
    if (r != 0.0) return r < 0.0 ? r + L : r;
    return 0.0;
}

int upload_times(th_tracer *t, int64_t off, const double *t0, int64_t n) {
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(t0[k])) return fail(TH_EINVAL, "launch times must be finite");
    if (n) TH_HIP(hipMemcpyAsync(t->d_t0 + off, t0, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, t->stream));
    return TH_OK;
}

// th_trace / th_trace_rf on host arrays (t0: launch times, RF only).
int trace_host(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
               const double *qm, const int8_t *spc, const double *t0, int64_t n, const th_trace_params *p,
               const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally, int8_t *status,
               int32_t *hit) {
    if (!t || !p || !cnt_i || !cnt_e || !tally) return fail(TH_EINVAL, "invalid argument");
    if (n < 0) return fail(TH_EINVAL, "negative particle count");
    if (n > 0 && (!x || !y || !vx || !vy || !qm || !spc)) return fail(TH_EINVAL, "particle array is NULL");
    if (p->rf && n > 0 && !t0) return fail(TH_EINVAL, "an RF trace needs the launch times");
    int rc = check_params(t, p);
    if (rc) return rc;
    for (int64_t k = 0; k < n; ++k)
        if (spc[k] != 0 && spc[k] != 1) return fail(TH_EINVAL, "species must be 0 (ion) or 1 (electron)");
    if (status)
        for (int64_t k = 0; k < n; ++k) status[k] = 3;
    if (hit)
        for (int64_t k = 0; k < n; ++k) hit[k] = -1;
    TH_HIP(hipSetDevice(t->device));
    if (n > 0) {
        if ((rc = reserve_particles(t, n))) return rc;
        if ((rc = upload(t, 0, x, y, vx, vy, qm, spc, n))) return rc;
        if (p->rf && (rc = upload_times(t, 0, t0, n))) return rc;
    }
    return trace_range(t, 0, n, p, stop, cnt_i, cnt_e, tally, status, hit);
}

}  // namespace

extern "C" {

int th_trace(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
             const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
             const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally, int8_t *status,
             int32_t *hit) {
    if (p && p->rf) return fail(TH_EINVAL, "an RF trace takes launch times: use th_trace_rf");
    return trace_host(t, x, y, vx, vy, qm, spc, nullptr, n, p, stop, cnt_i, cnt_e, tally, status, hit);
}

int th_trace_rf(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                const double *qm, const int8_t *spc, const double *initial_time, int64_t n,
                const th_trace_params *p, const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e,
                th_tally *tally, int8_t *status, int32_t *hit) {
    if (p && !p->rf) return fail(TH_EINVAL, "th_trace_rf needs params.rf = 1");
    return trace_host(t, x, y, vx, vy, qm, spc, initial_time, n, p, stop, cnt_i, cnt_e, tally, status, hit);
}

int th_set_rf(th_tracer *t, const double *Ex, const double *Ey, const double *phi, const double *wave,
              int64_t n_wave, double period, double dt_max) {
    if (!t) return fail(TH_EINVAL, "invalid argument");
    if (n_wave == 0) {
        t->rf = false;
        return TH_OK;
    }
    if (!Ex || !Ey || !wave || n_wave < 2 || n_wave > (1 << 24))
        return fail(TH_EINVAL, "RF needs the basis fields and a wave of 2 to 2^24 phases");
    if (!(std::isfinite(period) && period > 0.0 && std::isfinite(dt_max) && dt_max > 0.0))
        return fail(TH_EINVAL, "RF period and dt_max must be positive and finite");
    const size_t N = (size_t)t->N;
    double bound[2] = {0.0, 0.0};
    for (int64_t k = 0; k < 2 * n_wave; ++k) {
        if (!std::isfinite(wave[k])) return fail(TH_EINVAL, "RF wave must be finite");
        bound[k & 1] = std::max(bound[k & 1], std::fabs(wave[k]));
    }
    for (size_t c = 0; c < 2 * N; ++c)
        if (!std::isfinite(Ex[c]) || !std::isfinite(Ey[c]) || (phi && !std::isfinite(phi[c])))
            return fail(TH_EINVAL, "RF basis must be finite");
    TH_HIP(hipSetDevice(t->device));
    TH_HIP(hipStreamSynchronize(t->stream));
    int rc;
    if (!t->d_rfcells && ((rc = dev_alloc(&t->d_rfcells, N)) || (rc = dev_alloc(&t->d_rfphi, N)))) return rc;
    if (n_wave > t->n_wave || !t->d_wave) {
        if (t->d_wave) (void)hipFree(t->d_wave);
        t->d_wave = nullptr;
        t->n_wave = 0;
        if ((rc = dev_alloc(&t->d_wave, (size_t)n_wave))) return rc;
    }
    double *stage = nullptr;
    if ((rc = dev_alloc(&stage, 6 * N))) return rc;
    hipError_t e = hipMemcpyAsync(stage, Ex, sizeof(double) * 2 * N, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(stage + 2 * N, Ey, sizeof(double) * 2 * N, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && phi)
        e = hipMemcpyAsync(stage + 4 * N, phi, sizeof(double) * 2 * N, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->d_wave, wave, sizeof(double) * 2 * (size_t)n_wave, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) {
        int threads = 256, blocks = (t->N + threads - 1) / threads;
        hipLaunchKernelGGL(th::pack_rf_kernel, dim3(blocks), dim3(threads), 0, t->stream, stage, stage + 2 * N,
                           phi ? stage + 4 * N : nullptr, t->d_vac, t->N, t->d_rfcells, t->d_rfphi);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return fail(TH_EHIP, "RF upload: %s", hipGetErrorString(e));
    t->n_wave = (int)n_wave;
    t->period = period;
    t->dt_max = dt_max;
    t->bound[0] = bound[0];
    t->bound[1] = bound[1];
    t->rf = true;
    t->rf_phi = phi != nullptr;
    return TH_OK;
}

int th_particles_upload_times(th_tracer *t, int64_t offset, const double *initial_time, int64_t n) {
    if (!t || offset < 0 || n < 0 || (n > 0 && !initial_time)) return fail(TH_EINVAL, "invalid argument");
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    TH_HIP(hipSetDevice(t->device));
    int rc = upload_times(t, offset, initial_time, n);
    if (rc) return rc;
    TH_HIP(hipStreamSynchronize(t->stream));
    return TH_OK;
}

int th_particles_reserve(th_tracer *t, int64_t capacity) {
    if (!t || capacity < 0) return fail(TH_EINVAL, "invalid argument");
    TH_HIP(hipSetDevice(t->device));
    TH_HIP(hipStreamSynchronize(t->stream));
    if (capacity <= t->cap) return TH_OK;
    return reserve_particles(t, capacity);
}

int th_particles_upload(th_tracer *t, int64_t offset, const double *x, const double *y, const double *vx,
                        const double *vy, const double *qm, const int8_t *spc, int64_t n) {
    if (!t || offset < 0 || n < 0) return fail(TH_EINVAL, "invalid argument");
    if (n > 0 && (!x || !y || !vx || !vy || !qm || !spc)) return fail(TH_EINVAL, "particle array is NULL");
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    TH_HIP(hipSetDevice(t->device));
    int rc = upload(t, offset, x, y, vx, vy, qm, spc, n);
    if (rc) return rc;
    TH_HIP(hipStreamSynchronize(t->stream));
    return TH_OK;
}

int th_trace_resident(th_tracer *t, int64_t offset, int64_t n, const th_trace_params *p, int64_t *cnt_i,
                      int64_t *cnt_e, th_tally *tally, int8_t *status, int32_t *hit) {
    if (!t || !p || !tally || offset < 0 || n < 0) return fail(TH_EINVAL, "invalid argument");
    if ((cnt_i == nullptr) != (cnt_e == nullptr)) return fail(TH_EINVAL, "cnt_i and cnt_e go together");
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    int rc = check_params(t, p);
    if (rc) return rc;
    TH_HIP(hipSetDevice(t->device));
    return trace_range(t, offset, n, p, nullptr, cnt_i, cnt_e, tally, status, hit);
}

int th_read_counts(th_tracer *t, int32_t *counts) {
    if (!t || !counts) return fail(TH_EINVAL, "invalid argument");
    TH_HIP(hipSetDevice(t->device));
    TH_HIP(hipMemcpy(counts, t->d_counts, sizeof(uint32_t) * 2 * (size_t)t->m, hipMemcpyDeviceToHost));
    return TH_OK;
}

int th_last_push_ms(const th_tracer *t, double *ms) {
    if (!t || !ms) return fail(TH_EINVAL, "invalid argument");
    *ms = t->push_ms;
    return TH_OK;
}

int th_selftest_math(th_tracer *t, const double *a, const double *b, int64_t n, double *out) {
    if (!t || n < 0 || (n > 0 && (!a || !b || !out))) return fail(TH_EINVAL, "invalid argument");
    if (n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    double *d = nullptr;
    TH_HIP(hipMalloc((void **)&d, sizeof(double) * 11 * (size_t)n));
    int rc = TH_OK;
    hipError_t e = hipMemcpyAsync(d, a, sizeof(double) * n, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d + n, b, sizeof(double) * n, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) {
        int threads = 256;
        int64_t blocks = (n + threads - 1) / threads;
        hipLaunchKernelGGL(th::selftest_math_kernel, dim3((unsigned)blocks), dim3(threads), 0, t->stream, d,
                           d + n, n, d + 2 * n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d + 2 * n, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) rc = fail(TH_EHIP, "selftest: %s", hipGetErrorString(e));
    (void)hipFree(d);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Device-resident charging state.

struct th_charging {
    th_tracer *t = nullptr;
    th::ChargeDev d{};
    int32_t *col_of = nullptr, *rows = nullptr, *surf = nullptr, *col_ptr = nullptr, *col_cells = nullptr;
    double *green = nullptr, *external = nullptr, *leak_M = nullptr, *leak_c = nullptr, *g_cond_s = nullptr;
    double *leak_q = nullptr;   // [2m] the leak's input and outflow terms between the update's launches
    uint8_t *region = nullptr, *cond_s = nullptr;
    double *rho = nullptr, *rho_prev = nullptr, *qcol = nullptr, *phi = nullptr, *diag = nullptr;
    hipEvent_t ev2 = nullptr;
    int update_lds = 0;
    // split Green's matrix (precision variant), built on first use
    uint16_t *g_hi = nullptr, *g_lo = nullptr;
    int green_built = 0;
    double green_absmax = 0.0;
    // replicas of one batch (th_charging_step_replicas), grown on demand
    int rep_cap = 0;
    uint32_t *rep_counts = nullptr;
    unsigned long long *rep_scal = nullptr;
    double *rep_rho = nullptr, *rep_prev = nullptr, *rep_diag = nullptr, *rep_qcol = nullptr;
    // one device-injected batch captured as a hipGraph (th_charging_run)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    th_trace_params graph_params{};
    th::InjectDev graph_inject{};
    th::ChargeDev graph_d{};
    th::MonitorDev graph_mon{};
    int64_t graph_offset = -1;
    unsigned long long *d_run = nullptr;   // {batch index, first batch of the run}
    double *d_hist = nullptr;              // [kRunRows][row] per-batch records
    // monitors
    th::MonitorDev mon{};
    int32_t *mon_buf = nullptr;     // probe cells, region ptr, region idx, bottom (one allocation)
    double *mon_out = nullptr;
    std::vector<double> mon_host;
};

namespace {

template <class T>
int upload_new(T **dst, const T *src, size_t count, hipStream_t s) {
    int rc = dev_alloc(dst, count);
    if (rc) return rc;
    if (src) TH_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    else TH_HIP(hipMemsetAsync(*dst, 0, sizeof(T) * std::max<size_t>(count, 1), s));
    return TH_OK;
}

// G Q on matrix cores for R <= 16 charge vectors qcols [R][n_cols] (device):
// into phi (batch) or out [R][n_rows].  One wave per 16 rows, four per
// workgroup, the charge halves in LDS.
int launch_green_mfma(th_charging *c, const double *qcols, int R, double *out, bool batch) {
    th_tracer *t = c->t;
    const int tiles = (c->d.n_rows + 15) / 16;
    const int blocks = std::max(1, std::min((tiles + 3) / 4, t->cus * 8));
    const size_t lds = sizeof(uint16_t) * 2 * 16 * (size_t)c->d.n_pad;
    const bool bf16 = c->d.green_mode == th::kGreenBF16Mfma || c->d.green_mode == th::kGreenBF16;
    if (bf16)
        hipLaunchKernelGGL(th::green_mfma_kernel<th::kGreenBF16>, dim3(blocks), dim3(256), lds, t->stream, c->d, qcols,
                           R, out, batch ? 1 : 0);
    else
        hipLaunchKernelGGL(th::green_mfma_kernel<th::kGreenF16>, dim3(blocks), dim3(256), lds, t->stream, c->d, qcols,
                           R, out, batch ? 1 : 0);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

// Potential from the column charges, then the push's field records.
int launch_field(th_charging *c) {
    th_tracer *t = c->t;
    if (c->d.green_mode == th::kGreenF64) {
        int waves = 4, rows_per_block = waves;
        int blocks = (int)std::min<int64_t>((c->d.n_rows + rows_per_block - 1) / rows_per_block, (int64_t)t->cus * 8);
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(th::green_gemv_kernel, dim3(blocks), dim3(64 * waves), sizeof(double) * c->d.n_cols,
                           t->stream, c->d);
    } else if (c->d.green_mode == th::kGreenBF16Mfma || c->d.green_mode == th::kGreenF16Mfma) {
        int rc = launch_green_mfma(c, c->d.qcol, 1, nullptr, true);
        if (rc) return rc;
    } else {
        int rows_per_block = 256 / 16;
        int blocks = (int)std::min<int64_t>((c->d.n_rows + rows_per_block - 1) / rows_per_block, (int64_t)t->cus * 8);
        if (blocks < 1) blocks = 1;
        size_t lds = sizeof(float) * (size_t)c->d.n_pad;
        if (c->d.green_mode == th::kGreenBF16)
            hipLaunchKernelGGL(th::green_gemv_split_kernel<th::kGreenBF16>, dim3(blocks), dim3(256), lds, t->stream,
                               c->d);
        else
            hipLaunchKernelGGL(th::green_gemv_split_kernel<th::kGreenF16>, dim3(blocks), dim3(256), lds, t->stream,
                               c->d);
    }
    TH_HIP(hipGetLastError());
    hipLaunchKernelGGL(th::field_pack_kernel, dim3((t->N + 255) / 256), dim3(256), 0, t->stream, c->d);
    TH_HIP(hipGetLastError());
    return pack_corners(t);
}

}  // namespace

extern "C" {

void th_charging_destroy(th_charging *c) {
    if (!c) return;
    (void)hipSetDevice(c->t->device);
    (void)hipStreamSynchronize(c->t->stream);
    for (void *p : {(void *)c->col_of, (void *)c->col_ptr, (void *)c->col_cells, (void *)c->rows, (void *)c->surf,
                    (void *)c->green, (void *)c->external,
                    (void *)c->leak_M, (void *)c->leak_c, (void *)c->g_cond_s, (void *)c->region,
                    (void *)c->cond_s, (void *)c->rho, (void *)c->rho_prev, (void *)c->qcol, (void *)c->phi,
                    (void *)c->diag, (void *)c->g_hi, (void *)c->g_lo, (void *)c->rep_counts, (void *)c->rep_scal,
                    (void *)c->leak_q,
                    (void *)c->rep_rho, (void *)c->rep_prev, (void *)c->rep_diag, (void *)c->rep_qcol})
        if (p) (void)hipFree(p);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->d_run) (void)hipFree(c->d_run);
    if (c->d_hist) (void)hipFree(c->d_hist);
    if (c->mon_buf) (void)hipFree(c->mon_buf);
    if (c->mon_out) (void)hipFree(c->mon_out);
    delete c;
}

int th_charging_create(th_tracer *t, const th_charging_setup *s, th_charging **out) {
    if (!t || !s || !out) return fail(TH_EINVAL, "invalid argument");
    *out = nullptr;
    const int m = t->m;
    if (s->n_cols < 1 || s->n_rows < 1 || !s->col_of || !s->rows || !s->green || !s->external || !s->wface ||
        !s->cond_s)
        return fail(TH_EINVAL, "incomplete charging set-up");
    if (s->floating && (!s->g_cond_s || s->c_idx < 0 || s->c_idx >= m || !(s->C_cond > 0.0)))
        return fail(TH_EINVAL, "floating conductor needs g_cond_s, C_cond > 0 and a charge cell");
    if (!(s->dx > 0.0) || !(s->Te > 0.0)) return fail(TH_EINVAL, "dx and Te must be positive");
    for (int k = 0; k < m; ++k)
        if (s->col_of[k] < -1 || s->col_of[k] >= s->n_cols) return fail(TH_EINVAL, "col_of out of range");
    for (int r = 0; r < s->n_rows; ++r)
        if (s->rows[r] < 0 || s->rows[r] >= t->N) return fail(TH_EINVAL, "row cell out of range");
    if ((size_t)2 * m * sizeof(double) > 64 * 1024) return fail(TH_EINVAL, "too many surface cells for the update");
    TH_HIP(hipSetDevice(t->device));
    th_charging *c = new th_charging();
    c->t = t;
    auto bad = [&](int code) { th_charging_destroy(c); return code; };
    hipStream_t st = t->stream;
    int rc;
    // surface cells grouped by Green's column (ascending within a column)
    std::vector<int32_t> col_ptr((size_t)s->n_cols + 1, 0), col_cells;
    for (int k = 0; k < m; ++k)
        if (s->col_of[k] >= 0) ++col_ptr[(size_t)s->col_of[k] + 1];
    for (int q = 0; q < s->n_cols; ++q) col_ptr[(size_t)q + 1] += col_ptr[q];
    col_cells.resize((size_t)std::max(col_ptr[s->n_cols], 1));
    {
        std::vector<int32_t> fill(col_ptr.begin(), col_ptr.end() - 1);
        for (int k = 0; k < m; ++k)
            if (s->col_of[k] >= 0) col_cells[(size_t)fill[s->col_of[k]]++] = k;
    }
    if ((rc = upload_new(&c->col_of, s->col_of, m, st)) || (rc = upload_new(&c->rows, s->rows, s->n_rows, st)) ||
        (rc = upload_new(&c->col_ptr, col_ptr.data(), col_ptr.size(), st)) ||
        (rc = upload_new(&c->col_cells, col_cells.data(), col_cells.size(), st)) ||
        (rc = upload_new(&c->surf, t->surf.data(), m, st)) ||
        (rc = upload_new(&c->green, s->green, (size_t)s->n_rows * s->n_cols, st)) ||
        (rc = upload_new(&c->external, s->external, t->N, st)) ||
        (rc = upload_new(&c->cond_s, s->cond_s, m, st)) ||
        (rc = upload_new(&c->region, s->region, m, st)) ||
        (rc = upload_new(&c->g_cond_s, s->g_cond_s, m, st)) ||
        (rc = upload_new(&c->rho, (const double *)nullptr, m, st)) ||
        (rc = upload_new(&c->rho_prev, (const double *)nullptr, m, st)) ||
        (rc = upload_new(&c->qcol, (const double *)nullptr, s->n_cols, st)) ||
        (rc = upload_new(&c->phi, (const double *)nullptr, t->N, st)) ||
        (rc = upload_new(&c->diag, (const double *)nullptr, th::kDiagCount, st)))
        return bad(rc);
    if (s->leak_M && (rc = upload_new(&c->leak_M, s->leak_M, (size_t)m * m, st))) return bad(rc);
    if (s->leak_M && (rc = upload_new(&c->leak_q, (const double *)nullptr, 2 * (size_t)m, st))) return bad(rc);
    if (s->leak_c && (rc = upload_new(&c->leak_c, s->leak_c, m, st))) return bad(rc);
    if (hipMemcpyAsync(t->d_wface, s->wface, sizeof(double) * t->N, hipMemcpyHostToDevice, st) != hipSuccess)
        return bad(fail(TH_EHIP, "wface upload failed"));
    if (hipEventCreate(&c->ev2) != hipSuccess) return bad(fail(TH_EHIP, "hipEventCreate failed"));
    th::ChargeDev &d = c->d;
    d.m = m; d.n_cols = s->n_cols; d.n_rows = s->n_rows; d.nx = t->nx; d.ny = t->ny;
    d.col_of = c->col_of; d.col_ptr = c->col_ptr; d.col_cells = c->col_cells; d.rows = c->rows; d.surf = c->surf; d.green = c->green; d.external = c->external;
    d.leak_M = c->leak_M; d.leak_c = c->leak_c; d.region = c->region; d.cond_s = c->cond_s;
    d.g_cond_s = c->g_cond_s; d.vac = t->d_vac; d.wface = t->d_wface;
    d.rho = c->rho; d.rho_prev = c->rho_prev; d.qcol = c->qcol; d.phi = c->phi; d.diag = c->diag;
    d.cells = t->d_cells; d.cells32 = t->d_cells32; d.counts = t->d_counts;
    d.q_macro = s->q_macro; d.dx = s->dx; d.C_cond = s->C_cond; d.Te = s->Te; d.dt_batch = s->dt_batch;
    d.green_mode = th::kGreenF64;
    d.n_pad = (s->n_cols + 31) / 32 * 32;   // whole K blocks of the MFMA update
    d.lo_scale = 1.0f;
    d.inv_scale = 1.0;
    for (size_t k = 0; k < (size_t)s->n_rows * s->n_cols; ++k)
        c->green_absmax = std::max(c->green_absmax, std::fabs(s->green[k]));
    d.top_voltage = s->top_voltage; d.substrate_voltage = s->substrate_voltage;
    d.c_idx = s->c_idx; d.floating = s->floating; d.driven = s->driven; d.has_leak = s->leak_M != nullptr;
    d.top_dirichlet = s->top_dirichlet; d.leak_has_c = s->leak_c != nullptr;
    c->update_lds = (int)(2 * (size_t)std::max(m, 1) * sizeof(double));
    if (hipStreamSynchronize(st) != hipSuccess) return bad(fail(TH_EHIP, "set-up upload failed"));
    *out = c;
    return TH_OK;
}

int th_charging_set_green(th_charging *c, int32_t mode) {
    if (!c) return fail(TH_EINVAL, "invalid argument");
    if (mode < th::kGreenF64 || mode > th::kGreenF16Mfma)
        return fail(TH_EINVAL, "Green's precision %d is not available", mode);
    th_tracer *t = c->t;
    TH_HIP(hipSetDevice(t->device));
    th::ChargeDev &d = c->d;
    double scale = 1.0;
    float lo_scale = 1.0f;
    // the MFMA modes use the same split matrix as the scalar ones
    const int half = mode == th::kGreenBF16Mfma ? th::kGreenBF16 : mode == th::kGreenF16Mfma ? th::kGreenF16 : mode;
    if (half == th::kGreenF16) {
        // G * scale spans at most 2^14 (FP16 overflows at 65504); lo carries 11 more bits
        int e = c->green_absmax > 0.0 ? (int)std::ceil(std::log2(c->green_absmax)) : 0;
        scale = std::ldexp(1.0, 14 - e);
        lo_scale = std::ldexp(1.0f, -11);
    }
    if (half != th::kGreenF64 && c->green_built != half) {
        size_t count = (size_t)d.n_rows * d.n_pad;
        int rc;
        if (!c->g_hi || !c->g_lo) {   // both halves or neither: a failed allocation leaves no half behind
            if ((rc = dev_alloc(&c->g_hi, count)) || (rc = dev_alloc(&c->g_lo, count))) {
                if (c->g_hi) (void)hipFree(c->g_hi);
                c->g_hi = nullptr;
                c->g_lo = nullptr;
                return rc;
            }
        }
        int blocks = (int)((count + 255) / 256);
        if (half == th::kGreenBF16)
            hipLaunchKernelGGL(th::green_split_kernel<th::kGreenBF16>, dim3(blocks), dim3(256), 0, t->stream, c->green,
                               d.n_rows, d.n_cols, d.n_pad, scale, lo_scale, c->g_hi, c->g_lo);
        else
            hipLaunchKernelGGL(th::green_split_kernel<th::kGreenF16>, dim3(blocks), dim3(256), 0, t->stream, c->green,
                               d.n_rows, d.n_cols, d.n_pad, scale, lo_scale, c->g_hi, c->g_lo);
        TH_HIP(hipGetLastError());
        c->green_built = half;
    }
    d.green_mode = mode;
    d.g_hi = c->g_hi;
    d.g_lo = c->g_lo;
    d.lo_scale = lo_scale;
    d.inv_scale = 1.0 / scale;      // a power of two
    int rc = launch_field(c);       // potential and field of the current charge in the new precision
    if (rc) return rc;
    TH_HIP(hipStreamSynchronize(t->stream));
    return TH_OK;
}

int th_charging_potentials(th_charging *c, int32_t R, const double *qcols, double *out, double *ms) {
    if (!c || R < 1 || R > 16 || !qcols || !out) return fail(TH_EINVAL, "invalid argument (1 <= R <= 16)");
    const int mode = c->d.green_mode;
    if (mode != th::kGreenBF16Mfma && mode != th::kGreenF16Mfma)
        return fail(TH_ESTATE, "th_charging_potentials needs a split Green's matrix on matrix cores (mode 3 or 4)");
    th_tracer *t = c->t;
    TH_HIP(hipSetDevice(t->device));
    double *dq = nullptr, *dout = nullptr;
    int rc = dev_alloc(&dq, (size_t)R * c->d.n_cols);
    if (rc) return rc;
    if ((rc = dev_alloc(&dout, (size_t)R * c->d.n_rows))) {
        (void)hipFree(dq);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(dq, qcols, sizeof(double) * R * c->d.n_cols, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipEventRecord(t->ev0, t->stream);
    if (e == hipSuccess) rc = launch_green_mfma(c, dq, R, dout, false);
    if (e == hipSuccess && !rc) e = hipEventRecord(t->ev1, t->stream);
    if (e == hipSuccess && !rc)
        e = hipMemcpyAsync(out, dout, sizeof(double) * R * c->d.n_rows, hipMemcpyDeviceToHost, t->stream);
    hipError_t sync = hipStreamSynchronize(t->stream);
    float el = 0.0f;
    if (e == hipSuccess && !rc && sync == hipSuccess && ms && hipEventElapsedTime(&el, t->ev0, t->ev1) == hipSuccess)
        *ms = el;
    (void)hipFree(dq);
    (void)hipFree(dout);
    if (rc) return rc;
    if (e != hipSuccess || sync != hipSuccess)
        return fail(TH_EHIP, "potentials: %s", hipGetErrorString(e != hipSuccess ? e : sync));
    return TH_OK;
}

int th_charging_set_q_macro(th_charging *c, double q_macro) {
    if (!c || !(q_macro > 0.0) || !std::isfinite(q_macro)) return fail(TH_EINVAL, "q_macro must be positive");
    c->d.q_macro = q_macro;
    return TH_OK;
}

int th_charging_set_rho(th_charging *c, const double *rho_s) {
    if (!c || !rho_s) return fail(TH_EINVAL, "invalid argument");
    th_tracer *t = c->t;
    TH_HIP(hipSetDevice(t->device));
    TH_HIP(hipMemcpyAsync(c->rho, rho_s, sizeof(double) * c->d.m, hipMemcpyHostToDevice, t->stream));
    // column charges on the host (one-off), then potential and records on the device
    std::vector<int32_t> col(c->d.m);
    TH_HIP(hipMemcpyAsync(col.data(), c->col_of, sizeof(int32_t) * c->d.m, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    std::vector<double> q(c->d.n_cols, 0.0);
    const double dx2 = c->d.dx * c->d.dx;
    for (int k = 0; k < c->d.m; ++k)
        if (col[k] >= 0) q[col[k]] += rho_s[k] * dx2;
    TH_HIP(hipMemcpyAsync(c->qcol, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice, t->stream));
    int rc = launch_field(c);
    if (rc) return rc;
    TH_HIP(hipStreamSynchronize(t->stream));
    t->fields = true;
    t->barrier_fields = true;
    return TH_OK;
}

// The batch's charge update on the tracer's stream: one workgroup, or with a
// leak operator three launches so that the leak's GEMV streams M over many
// workgroups (same arithmetic, same order).
static int launch_update(th_charging *c) {
    th_tracer *t = c->t;
    if (c->leak_q) {
        const int m = c->d.m;
        hipLaunchKernelGGL(th::charge_deposit_kernel, dim3(1), dim3(th::kUpdBlock), 0, t->stream, c->d, c->leak_q);
        hipLaunchKernelGGL(th::leak_rows_kernel, dim3((m + 3) / 4), dim3(256), sizeof(double) * (size_t)m, t->stream,
                           c->d, c->leak_q);
        hipLaunchKernelGGL(th::charge_finish_kernel, dim3(1), dim3(th::kUpdBlock), c->update_lds, t->stream, c->d,
                           (const double *)c->leak_q);
    } else {
        hipLaunchKernelGGL(th::charge_update_kernel, dim3(1), dim3(th::kUpdBlock), c->update_lds, t->stream, c->d);
    }
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int th_charging_step(th_charging *c, int64_t offset, int64_t n, const th_trace_params *p, th_batch_diag *out) {
    if (!c || !p || !out || offset < 0 || n < 0) return fail(TH_EINVAL, "invalid argument");
    th_tracer *t = c->t;
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    int rc = check_params(t, p);
    if (rc) return rc;
    TH_HIP(hipSetDevice(t->device));
    memset(out, 0, sizeof *out);
    int64_t traced = offset;
    if ((rc = launch_trace(t, offset, n, p, nullptr, false, false, &traced, &out->tally))) return rc;
    if ((rc = launch_update(c))) return rc;
    if ((rc = launch_field(c))) return rc;
    if (c->mon_out) {
        hipLaunchKernelGGL(th::monitor_kernel, dim3(1), dim3(256), 0, t->stream, c->d, c->mon);
        TH_HIP(hipGetLastError());
    }
    TH_HIP(hipEventRecord(c->ev2, t->stream));
    unsigned long long scal[th::kScalCount];
    double diag[th::kDiagCount];
    TH_HIP(hipMemcpyAsync(scal, t->d_scal, sizeof scal, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipMemcpyAsync(diag, c->diag, sizeof diag, hipMemcpyDeviceToHost, t->stream));
    if (c->mon_out)
        TH_HIP(hipMemcpyAsync(c->mon_host.data(), c->mon_out, sizeof(double) * c->mon_host.size(),
                              hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    fill_tally(scal, offset + n - traced, &out->tally);
    out->leak_dq = diag[th::kDiagLeakDQ];
    out->drift = diag[th::kDiagDrift];
    out->dv = diag[th::kDiagDV];
    out->n_i_cond = diag[th::kDiagNi];
    out->n_e_cond = diag[th::kDiagNe];
    float a = 0.0f, b = 0.0f;
    TH_HIP(hipEventElapsedTime(&a, t->ev0, t->ev1));
    TH_HIP(hipEventElapsedTime(&b, t->ev1, c->ev2));
    out->push_ms = a;
    out->update_ms = b;
    t->push_ms = a;
    return TH_OK;
}

}  // extern "C"

namespace {

int reserve_replicas(th_charging *c, int R) {
    if (R <= c->rep_cap) return TH_OK;
    for (void *p : {(void *)c->rep_counts, (void *)c->rep_scal, (void *)c->rep_rho, (void *)c->rep_prev,
                    (void *)c->rep_diag, (void *)c->rep_qcol})
        if (p) (void)hipFree(p);
    c->rep_counts = nullptr;
    c->rep_scal = nullptr;
    c->rep_rho = c->rep_prev = c->rep_diag = c->rep_qcol = nullptr;
    c->rep_cap = 0;
    const size_t m = (size_t)std::max(c->d.m, 1);
    int rc;
    if ((rc = dev_alloc(&c->rep_counts, (size_t)R * 2 * m)) || (rc = dev_alloc(&c->rep_scal, (size_t)R * 4)) ||
        (rc = dev_alloc(&c->rep_rho, (size_t)R * m)) || (rc = dev_alloc(&c->rep_prev, (size_t)R * m)) ||
        (rc = dev_alloc(&c->rep_diag, (size_t)R * th::kDiagCount)) ||
        (rc = dev_alloc(&c->rep_qcol, (size_t)R * c->d.n_cols)))
        return rc;
    c->rep_cap = R;
    return TH_OK;
}

}  // namespace

extern "C" {

int th_charging_step_replicas(th_charging *c, int64_t offset, int64_t n, int32_t replicas, const th_trace_params *p,
                              double *rho_out, int64_t *scal_out, th_batch_diag *out) {
    if (!c || !p || !rho_out || !out || offset < 0 || n < 0 || replicas < 1 || replicas > 4096)
        return fail(TH_EINVAL, "invalid argument");
    th_tracer *t = c->t;
    const int64_t total = n * replicas;
    if (offset + total > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    int rc = check_params(t, p);
    if (rc) return rc;
    TH_HIP(hipSetDevice(t->device));
    if ((rc = reserve_replicas(c, replicas))) return rc;
    memset(out, 0, sizeof *out);
    const int m = t->m;
    int64_t traced = offset;
    // one push over every replica's particles, keeping each particle's end
    if ((rc = launch_trace(t, offset, total, p, nullptr, true, true, &traced, &out->tally))) return rc;
    TH_HIP(hipMemsetAsync(c->rep_counts, 0, sizeof(uint32_t) * 2 * (size_t)std::max(m, 1) * replicas, t->stream));
    TH_HIP(hipMemsetAsync(c->rep_scal, 0, sizeof(unsigned long long) * 4 * (size_t)replicas, t->stream));
    if (n > 0 && m > 0) {
        const int per = (int)std::min<int64_t>((n + 255) / 256, std::max(1, 4 * t->cus / replicas));
        hipLaunchKernelGGL(th::replica_tally_kernel, dim3(per, replicas), dim3(256), sizeof(uint32_t) * 2 * m,
                           t->stream, t->d_status, t->d_hit, t->d_spc, t->d_smap, m, offset, n, c->rep_counts,
                           c->rep_scal);
        TH_HIP(hipGetLastError());
    }
    th::ReplicaDev r{c->rep_counts, c->rep_rho, c->rep_prev, c->rep_diag, c->rep_qcol};
    hipLaunchKernelGGL(th::charge_update_replicas_kernel, dim3(replicas), dim3(th::kUpdBlock), c->update_lds,
                       t->stream, c->d, r);
    TH_HIP(hipGetLastError());
    TH_HIP(hipEventRecord(c->ev2, t->stream));
    unsigned long long scal[th::kScalCount];
    std::vector<unsigned long long> rscal(4 * (size_t)replicas);
    TH_HIP(hipMemcpyAsync(scal, t->d_scal, sizeof scal, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipMemcpyAsync(rscal.data(), c->rep_scal, sizeof(unsigned long long) * rscal.size(),
                          hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipMemcpyAsync(rho_out, c->rep_rho, sizeof(double) * (size_t)m * replicas, hipMemcpyDeviceToHost,
                          t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    fill_tally(scal, offset + total - traced, &out->tally);
    if (scal_out)
        for (size_t k = 0; k < rscal.size(); ++k) scal_out[k] = (int64_t)rscal[k];
    float a = 0.0f, b = 0.0f;
    TH_HIP(hipEventElapsedTime(&a, t->ev0, t->ev1));
    TH_HIP(hipEventElapsedTime(&b, t->ev1, c->ev2));
    out->push_ms = a;
    out->update_ms = b;
    t->push_ms = a;
    return TH_OK;
}

}  // extern "C"

namespace {

static_assert(th::kScalEsc0 == 1 && th::kScalEsc1 == 2 && th::kScalLost == 3 && th::kScalSteps == 4 &&
                  th::kScalOff == 5,
              "batch_record_kernel reads the push's scalar tallies by these indices");

constexpr int kRunRows = 1024;   // batches per th_charging_run call at most

int run_row(const th_charging *c) { return c->mon.n_probes + c->mon.n_regions + 3 + 5 + th::kDiagCount; }

// One batch as kernels only (capturable): injection at the device batch
// index, tally clears by fills, the push passes, deposit / leak / conductor,
// potential and field, monitors, and the batch's record.  The same kernels
// in the same order as th_inject + th_charging_step, so the results are the
// same.
int charging_batch_kernels(th_charging *c, int64_t offset, int64_t n, const th_trace_params *p,
                           const th::InjectDev &inj) {
    th_tracer *t = c->t;
    int rc;
    hipLaunchKernelGGL(th::inject_at_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, t->stream, inj,
                       c->d_run);
    TH_HIP(hipGetLastError());
    // tallies, counters, the passes' work counters and survivor counts: one fill
    hipLaunchKernelGGL(th::clear_batch_kernel, dim3(64), dim3(256), 0, t->stream, t->d_counts,
                       2 * (int64_t)std::max(t->m, 1), t->d_overflow, 2 * (int64_t)t->N, t->d_scal, (int64_t)kScalAll,
                       t->d_surv_count, (int64_t)(kMaxPasses + 1));
    TH_HIP(hipGetLastError());
    t->blocks_per_cu = p->blocks_per_cu;
    th::Grid g = make_grid(t, p);
    th::Work w;
    w.x = t->d_x; w.y = t->d_y; w.vx = t->d_vx; w.vy = t->d_vy; w.qm = t->d_qm; w.spc = t->d_spc;
    w.t0 = t->d_t0;
    w.status = nullptr;
    w.hit = nullptr;
    w.counts = t->d_counts;
    w.overflow = t->d_overflow;
    w.scal = t->d_scal;
    w.max_steps = p->max_steps;
    w.refill = p->refill > 0 ? std::min(p->refill, 64) : 16;
    w.order = p->order != 0;
    w.prio_steps = p->prio_steps > 0 ? p->prio_steps : 0;
    w.check_steps = p->check_steps > 0 ? std::min(p->check_steps, 256) : th::kStepsPerCheck;
    w.rec = th::Record{};
    w.tan = th::TangentIO{};
    w.begin = offset;
    w.end = offset + n;
    const bool passes = p->pass_steps > 0 && p->pass_steps < p->max_steps;
    int64_t cap = passes ? p->pass_steps : p->max_steps;
    for (int pass = 0;; ++pass) {
        const bool last = cap >= p->max_steps || pass == kMaxPasses - 1;
        w.step_cap = last ? p->max_steps : (int)cap;
        w.in = pass ? t->d_surv[(pass - 1) & 1] : nullptr;
        w.in_count = pass ? t->d_surv_count + pass - 1 : nullptr;
        w.out = last ? nullptr : t->d_surv[pass & 1];
        w.out_count = last ? nullptr : t->d_surv_count + pass;
        w.queue = t->d_scal + th::kScalCount + pass;   // cleared above
        const bool tile = p->tile_pass > 0 && pass + 1 >= p->tile_pass && t->tile_w > 0 && p->precision == 0;
        if ((rc = tile ? launch_tile(t, g, w, p->barrier != 0, p->tapered != 0)
                       : launch_push(t, g, w, p->barrier != 0, p->tapered != 0, p->precision, false, false)))
            return rc;
        if (last) break;
        cap = std::min<int64_t>(2 * cap, p->max_steps);
    }
    if ((rc = launch_update(c))) return rc;
    if ((rc = launch_field(c))) return rc;
    const int n_mon = c->mon_out ? c->mon.n_probes + c->mon.n_regions + 3 : 0;
    if (c->mon_out) {
        hipLaunchKernelGGL(th::monitor_kernel, dim3(1), dim3(256), 0, t->stream, c->d, c->mon);
        TH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(th::batch_record_kernel, dim3(1), dim3(64), 0, t->stream, c->mon_out, n_mon, t->d_scal, c->diag,
                       c->d_run, c->d_hist, run_row(c), kRunRows);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

void free_charging_graph(th_charging *c) {
    if (c->exec) (void)hipGraphExecDestroy(c->exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->exec = nullptr;
    c->graph = nullptr;
    c->graph_offset = -1;
}

// Captures the batch once per (offset, schedule, injection, charging state
// layout, monitors): everything the captured kernels take by value.
int charging_capture(th_charging *c, int64_t offset, int64_t n, const th_trace_params *p, const th::InjectDev &inj) {
    if (c->exec && c->graph_offset == offset && memcmp(&c->graph_params, p, sizeof *p) == 0 &&
        memcmp(&c->graph_inject, &inj, sizeof inj) == 0 && memcmp(&c->graph_d, &c->d, sizeof c->d) == 0 &&
        memcmp(&c->graph_mon, &c->mon, sizeof c->mon) == 0)
        return TH_OK;
    free_charging_graph(c);
    th_tracer *t = c->t;
    TH_HIP(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
    int rc = charging_batch_kernels(c, offset, n, p, inj);
    hipGraph_t graph = nullptr;
    hipError_t end = hipStreamEndCapture(t->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (end != hipSuccess) return fail(TH_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(end));
    c->graph = graph;
    hipError_t inst = hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0);
    if (inst != hipSuccess) {
        free_charging_graph(c);
        return fail(TH_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(inst));
    }
    c->graph_offset = offset;
    c->graph_params = *p;
    c->graph_inject = inj;
    c->graph_d = c->d;
    c->graph_mon = c->mon;
    return TH_OK;
}

th::InjectDev inject_dev(th_tracer *t, int64_t offset, const th_inject_params *ip, uint64_t seed, uint64_t batch) {
    th::InjectDev d{};
    d.n = ip->n; d.lx = ip->lx; d.y0 = ip->y0; d.qm_i = ip->qm_i; d.qm_e = ip->qm_e; d.ion_kind = ip->ion_kind;
    d.e_ion = ip->e_ion; d.e_lo = ip->e_lo; d.e_hi = ip->e_hi; d.sigma_theta = ip->sigma_theta;
    d.vth_i = ip->vth_i; d.v_per_sqrt_ev = ip->v_per_sqrt_ev; d.s_e = ip->s_e;
    d.seed = seed;
    d.batch = batch;
    d.x = t->d_x + offset; d.y = t->d_y + offset; d.vx = t->d_vx + offset; d.vy = t->d_vy + offset;
    d.qm = t->d_qm + offset; d.spc = t->d_spc + offset;
    return d;
}

}  // namespace

extern "C" {

int th_charging_row_size(const th_charging *c, int32_t *size) {
    if (!c || !size) return fail(TH_EINVAL, "invalid argument");
    *size = run_row(c);
    return TH_OK;
}

int th_charging_run(th_charging *c, int64_t offset, const th_inject_params *ip, uint64_t seed, uint64_t first_batch,
                    int32_t batches, const th_trace_params *p, double *rows, double *ms) {
    if (!c || !ip || !p || offset < 0 || batches < 0 || (batches > 0 && !rows)) return fail(TH_EINVAL, "invalid argument");
    if (batches > kRunRows) return fail(TH_EINVAL, "at most %d batches per call", kRunRows);
    if (ip->ion_kind < 0 || ip->ion_kind > 2) return fail(TH_EINVAL, "unknown ion distribution %d", ip->ion_kind);
    if (ip->n < 1) return fail(TH_EINVAL, "a run needs particles");
    if (p->rf) return fail(TH_EINVAL, "the captured batch has no RF push (launch times come from the host)");
    th_tracer *t = c->t;
    const int64_t n = 2 * ip->n;
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    int rc = check_params(t, p);
    if (rc) return rc;
    if (ms) *ms = 0.0;
    if (batches == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    const int row = run_row(c);
    if (row > 64) return fail(TH_EINVAL, "%d monitors exceed the run record", row);
    if (!c->d_run && (rc = dev_alloc(&c->d_run, 2))) return rc;
    if (!c->d_hist && (rc = dev_alloc(&c->d_hist, (size_t)kRunRows * 64))) return rc;
    const th::InjectDev inj = inject_dev(t, offset, ip, seed, 0);
    if ((rc = charging_capture(c, offset, n, p, inj))) return rc;
    const unsigned long long run[2] = {first_batch, first_batch};
    TH_HIP(hipMemcpyAsync(c->d_run, run, sizeof run, hipMemcpyHostToDevice, t->stream));
    TH_HIP(hipEventRecord(t->ev0, t->stream));
    for (int b = 0; b < batches; ++b) TH_HIP(hipGraphLaunch(c->exec, t->stream));
    TH_HIP(hipEventRecord(t->ev1, t->stream));
    TH_HIP(hipMemcpyAsync(rows, c->d_hist, sizeof(double) * (size_t)batches * row, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    if (c->mon_out) {   // the last batch's monitors, as after th_charging_step
        const size_t k = c->mon_host.size();
        memcpy(c->mon_host.data(), rows + (size_t)(batches - 1) * row, sizeof(double) * k);
    }
    float el = 0.0f;
    TH_HIP(hipEventElapsedTime(&el, t->ev0, t->ev1));
    if (ms) *ms = el;
    t->fields = true;
    return TH_OK;
}

int th_charging_read(th_charging *c, double *rho_s, double *phi, int32_t *counts) {
    if (!c) return fail(TH_EINVAL, "invalid argument");
    th_tracer *t = c->t;
    TH_HIP(hipSetDevice(t->device));
    if (rho_s) TH_HIP(hipMemcpyAsync(rho_s, c->rho, sizeof(double) * c->d.m, hipMemcpyDeviceToHost, t->stream));
    if (phi) TH_HIP(hipMemcpyAsync(phi, c->phi, sizeof(double) * t->N, hipMemcpyDeviceToHost, t->stream));
    if (counts)
        TH_HIP(hipMemcpyAsync(counts, t->d_counts, sizeof(uint32_t) * 2 * t->m, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    return TH_OK;
}

}  // extern "C"

namespace {

// G^T w with buf = [N] weights, [blocks][n_cols] partial sums, [n_cols] result
// (green_adjoint_buffer doubles): upload, two kernels, readback, wait.  Every
// entry of buf that is read is written first in the same call.
int green_adjoint_rows_per_block() { return 64; }

size_t green_adjoint_buffer(const th_charging *c) {
    const int rpb = green_adjoint_rows_per_block();
    const size_t blocks = (size_t)std::max(1, (c->d.n_rows + rpb - 1) / rpb);
    return (size_t)c->t->N + blocks * c->d.n_cols + c->d.n_cols;
}

int green_adjoint_into(th_charging *c, const double *w, double *out, double *buf) {
    th_tracer *t = c->t;
    const int rows_per_block = green_adjoint_rows_per_block();
    const int blocks = std::max(1, (c->d.n_rows + rows_per_block - 1) / rows_per_block);
    double *dw = buf, *partial = buf + t->N, *dout = partial + (size_t)blocks * c->d.n_cols;
    hipError_t e = hipMemcpyAsync(dw, w, sizeof(double) * t->N, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(th::green_adjoint_partial_kernel, dim3(blocks), dim3(256), 0, t->stream, c->d, dw,
                           rows_per_block, partial);
        hipLaunchKernelGGL(th::green_adjoint_sum_kernel, dim3((c->d.n_cols + 255) / 256), dim3(256), 0, t->stream,
                           partial, blocks, c->d.n_cols, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(double) * c->d.n_cols, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) return fail(TH_EHIP, "green adjoint: %s", hipGetErrorString(e));
    return TH_OK;
}

}  // namespace

extern "C" {

int th_charging_green_adjoint(th_charging *c, const double *w, double *out) {
    if (!c || !w || !out) return fail(TH_EINVAL, "invalid argument");
    TH_HIP(hipSetDevice(c->t->device));
    double *buf = nullptr;
    int rc = dev_alloc(&buf, green_adjoint_buffer(c));
    if (rc) return rc;
    rc = green_adjoint_into(c, w, out, buf);
    if (rc) (void)hipStreamSynchronize(c->t->stream);   // nothing may still use buf when it is freed
    (void)hipFree(buf);
    return rc;
}

int th_charging_green_adjoint_noalloc(th_charging *c, const double *w, double *out) {
    if (!c || !w || !out) return fail(TH_EINVAL, "invalid argument");
    const size_t need = green_adjoint_buffer(c);
    if (need > kScratchDoubles)
        return fail(TH_EINVAL,
                    "G^T w on %d cells x %d columns needs %zu doubles of scratch; the allocation-free variant holds "
                    "%zu (use th_charging_green_adjoint)",
                    c->t->N, c->d.n_cols, need, kScratchDoubles);
    TH_HIP(hipSetDevice(c->t->device));
    double *buf = nullptr;
    int rc = scratch_address(c->t->device, &buf);
    if (rc) return rc;
    return green_adjoint_into(c, w, out, buf);
}

int th_inject(th_tracer *t, int64_t offset, const th_inject_params *ip, uint64_t seed, uint64_t batch) {
    if (!t || !ip || offset < 0 || ip->n < 0) return fail(TH_EINVAL, "invalid argument");
    if (ip->ion_kind < 0 || ip->ion_kind > 2) return fail(TH_EINVAL, "unknown ion distribution %d", ip->ion_kind);
    if (offset + 2 * ip->n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    if (ip->n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    th::InjectDev d;
    d.n = ip->n; d.lx = ip->lx; d.y0 = ip->y0; d.qm_i = ip->qm_i; d.qm_e = ip->qm_e; d.ion_kind = ip->ion_kind;
    d.e_ion = ip->e_ion; d.e_lo = ip->e_lo; d.e_hi = ip->e_hi; d.sigma_theta = ip->sigma_theta;
    d.vth_i = ip->vth_i; d.v_per_sqrt_ev = ip->v_per_sqrt_ev; d.s_e = ip->s_e; d.seed = seed; d.batch = batch;
    d.x = t->d_x + offset; d.y = t->d_y + offset; d.vx = t->d_vx + offset; d.vy = t->d_vy + offset;
    d.qm = t->d_qm + offset; d.spc = t->d_spc + offset;
    int64_t blocks = (2 * ip->n + 255) / 256;
    hipLaunchKernelGGL(th::inject_kernel, dim3((unsigned)blocks), dim3(256), 0, t->stream, d);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

}  // extern "C"

extern "C" {

int th_particles_read(th_tracer *t, int64_t offset, int64_t n, double *x, double *y, double *vx, double *vy,
                      double *qm, int8_t *spc) {
    if (!t || offset < 0 || n < 0 || offset + n > t->cap) return fail(TH_EINVAL, "invalid range");
    if (n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    size_t b8 = sizeof(double) * (size_t)n;
    if (x) TH_HIP(hipMemcpyAsync(x, t->d_x + offset, b8, hipMemcpyDeviceToHost, t->stream));
    if (y) TH_HIP(hipMemcpyAsync(y, t->d_y + offset, b8, hipMemcpyDeviceToHost, t->stream));
    if (vx) TH_HIP(hipMemcpyAsync(vx, t->d_vx + offset, b8, hipMemcpyDeviceToHost, t->stream));
    if (vy) TH_HIP(hipMemcpyAsync(vy, t->d_vy + offset, b8, hipMemcpyDeviceToHost, t->stream));
    if (qm) TH_HIP(hipMemcpyAsync(qm, t->d_qm + offset, b8, hipMemcpyDeviceToHost, t->stream));
    if (spc) TH_HIP(hipMemcpyAsync(spc, t->d_spc + offset, (size_t)n, hipMemcpyDeviceToHost, t->stream));
    TH_HIP(hipStreamSynchronize(t->stream));
    return TH_OK;
}

int th_read_field(th_tracer *t, double *Ex, double *Ey) {
    if (!t || !Ex || !Ey) return fail(TH_EINVAL, "invalid argument");
    if (!t->fields) return fail(TH_ESTATE, "no field has been set");
    TH_HIP(hipSetDevice(t->device));
    std::vector<double> rec(4 * (size_t)t->N);
    TH_HIP(hipMemcpy(rec.data(), t->d_cells, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
    for (int c = 0; c < t->N; ++c) {
        Ex[c] = rec[4 * (size_t)c];
        Ey[c] = rec[4 * (size_t)c + 1];
    }
    return TH_OK;
}

}  // extern "C"

extern "C" int th_probe_copy(th_tracer *t, int64_t bytes, int32_t repeats, double *ms) {
    if (!t || !ms || bytes <= 0 || bytes % 32 || repeats < 1) return fail(TH_EINVAL, "invalid argument");
    TH_HIP(hipSetDevice(t->device));
    double4 *a = nullptr, *b = nullptr;
    TH_HIP(hipMalloc((void **)&a, (size_t)bytes));
    if (hipMalloc((void **)&b, (size_t)bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(TH_ENOMEM, "copy probe allocation failed");
    }
    int64_t n = bytes / 32;
    int blocks = t->cus * 8;
    hipError_t e = hipMemsetAsync(a, 0, (size_t)bytes, t->stream);
    if (e == hipSuccess) e = hipEventRecord(t->ev0, t->stream);
    for (int r = 0; r < repeats && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(th::copy_kernel, dim3(blocks), dim3(256), 0, t->stream, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(t->ev1, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    float f = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, t->ev0, t->ev1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(TH_EHIP, "copy probe: %s", hipGetErrorString(e));
    *ms = f / repeats;
    return TH_OK;
}

namespace {

// A triangular factor's off-diagonal part in CSR: ptr from 0 to nnz, not
// decreasing; row i's column indices below i (lower) or above i (upper).
bool triangle_ok(int32_t n, const int32_t *ptr, const int32_t *idx, const double *val, bool lower) {
    if (ptr[0] != 0) return false;
    for (int32_t i = 0; i < n; ++i) {
        if (ptr[i + 1] < ptr[i]) return false;
        for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) {
            if (lower ? (idx[q] < 0 || idx[q] >= i) : (idx[q] <= i || idx[q] >= n)) return false;
            if (!std::isfinite(val[q])) return false;
        }
    }
    return true;
}

// Rows of a validated triangle grouped by dependency level: level(i) = 1 +
// the largest level of the rows its entries read (0 for none), rows visited
// in dependency order.  order holds the rows level by level; level_ptr has
// one more entry than there are levels.
void triangle_levels(int32_t n, const int32_t *ptr, const int32_t *idx, bool upper, std::vector<int32_t> &order,
                     std::vector<int32_t> &level_ptr) {
    std::vector<int32_t> level((size_t)n, 0);
    int32_t top = 0;
    for (int32_t r = 0; r < n; ++r) {
        const int32_t i = upper ? n - 1 - r : r;
        int32_t l = 0;
        for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) l = std::max(l, level[idx[q]] + 1);
        level[i] = l;
        top = std::max(top, l);
    }
    level_ptr.assign((size_t)top + 2, 0);
    for (int32_t i = 0; i < n; ++i) ++level_ptr[level[i] + 1];
    for (size_t l = 1; l < level_ptr.size(); ++l) level_ptr[l] += level_ptr[l - 1];
    order.resize((size_t)n);
    std::vector<int32_t> fill(level_ptr.begin(), level_ptr.end() - 1);
    for (int32_t i = 0; i < n; ++i) order[fill[level[i]]++] = i;
    // every level holds a row: a row of level l > 0 reads a row of level l - 1
}

}  // namespace

extern "C" int th_lu_solve(int32_t device, int32_t n, const int32_t *l_ptr, const int32_t *l_idx,
                           const double *l_val, const int32_t *u_ptr, const int32_t *u_idx, const double *u_val,
                           const double *u_diag, int32_t k, const double *b, double *x, double *ms) {
    if (n < 1 || k < 0 || !l_ptr || !u_ptr || !u_diag || (k > 0 && (!b || !x))) return fail(TH_EINVAL, "invalid argument");
    if ((l_ptr[n] > 0 && (!l_idx || !l_val)) || (u_ptr[n] > 0 && (!u_idx || !u_val)))
        return fail(TH_EINVAL, "factor array is NULL");
    if (!triangle_ok(n, l_ptr, l_idx, l_val, true)) return fail(TH_EINVAL, "L is not strictly lower triangular CSR");
    if (!triangle_ok(n, u_ptr, u_idx, u_val, false)) return fail(TH_EINVAL, "U is not strictly upper triangular CSR");
    for (int32_t i = 0; i < n; ++i)
        if (!(std::fabs(u_diag[i]) > 0.0) || !std::isfinite(u_diag[i])) return fail(TH_EINVAL, "U has a zero pivot");
    if (ms) *ms = 0.0;
    if (k == 0) return TH_OK;
    TH_HIP(hipSetDevice(device));
    const int slabs = (k + th::kLuLanes - 1) / th::kLuLanes;
    const size_t ld = (size_t)slabs * th::kLuLanes;
    const size_t nl = (size_t)l_ptr[n], nu = (size_t)u_ptr[n];
    std::vector<int32_t> l_order, l_levels, u_order, u_levels;
    triangle_levels(n, l_ptr, l_idx, false, l_order, l_levels);
    triangle_levels(n, u_ptr, u_idx, true, u_order, u_levels);
    int32_t *d_lp = nullptr, *d_li = nullptr, *d_up = nullptr, *d_ui = nullptr, *d_lo = nullptr, *d_uo = nullptr;
    double *d_lv = nullptr, *d_uv = nullptr, *d_ud = nullptr, *d_x = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;   // its own stream: the legacy null stream would wait for every blocking stream
    auto release = [&]() {   // after an error too: nothing enqueued may still use the buffers
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        for (void *q : {(void *)d_lp, (void *)d_li, (void *)d_up, (void *)d_ui, (void *)d_lv, (void *)d_uv,
                        (void *)d_ud, (void *)d_x, (void *)d_lo, (void *)d_uo})
            if (q) (void)hipFree(q);
        for (hipEvent_t v : ev)
            if (v) (void)hipEventDestroy(v);
    };
    int rc;
    if ((rc = dev_alloc(&d_lp, (size_t)n + 1)) || (rc = dev_alloc(&d_li, nl)) || (rc = dev_alloc(&d_lv, nl)) ||
        (rc = dev_alloc(&d_up, (size_t)n + 1)) || (rc = dev_alloc(&d_ui, nu)) || (rc = dev_alloc(&d_uv, nu)) ||
        (rc = dev_alloc(&d_ud, (size_t)n)) || (rc = dev_alloc(&d_x, (size_t)n * ld)) ||
        (rc = dev_alloc(&d_lo, (size_t)n)) || (rc = dev_alloc(&d_uo, (size_t)n))) {
        release();
        return rc;
    }
    hipError_t e = hipStreamCreate(&st);
    if (e == hipSuccess) e = hipMemcpy(d_lp, l_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nl) e = hipMemcpy(d_li, l_idx, sizeof(int32_t) * nl, hipMemcpyHostToDevice);
    if (e == hipSuccess && nl) e = hipMemcpy(d_lv, l_val, sizeof(double) * nl, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_up, u_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nu) e = hipMemcpy(d_ui, u_idx, sizeof(int32_t) * nu, hipMemcpyHostToDevice);
    if (e == hipSuccess && nu) e = hipMemcpy(d_uv, u_val, sizeof(double) * nu, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ud, u_diag, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_lo, l_order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_uo, u_order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
    // rows padded to ld on the host (one linear copy each way); the padding columns start at zero and stay zero
    std::vector<double> pad;
    if (ld != (size_t)k) {
        pad.assign((size_t)n * ld, 0.0);
        for (size_t i = 0; i < (size_t)n; ++i) memcpy(pad.data() + i * ld, b + i * k, sizeof(double) * k);
    }
    if (e == hipSuccess)
        e = hipMemcpy(d_x, pad.empty() ? b : pad.data(), sizeof(double) * n * ld, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], st);
    // one stream-ordered launch per level: L's levels, then U's; a level's rows run in parallel
    for (size_t l = 0; l + 1 < l_levels.size() && e == hipSuccess; ++l) {
        hipLaunchKernelGGL(th::lu_level_kernel, dim3(l_levels[l + 1] - l_levels[l], slabs), dim3(th::kLuLanes), 0,
                           st, d_lo + l_levels[l], d_lp, d_li, d_lv, (const double *)nullptr, (int)ld, d_x);
        e = hipGetLastError();
    }
    for (size_t l = 0; l + 1 < u_levels.size() && e == hipSuccess; ++l) {
        hipLaunchKernelGGL(th::lu_level_kernel, dim3(u_levels[l + 1] - u_levels[l], slabs), dim3(th::kLuLanes), 0,
                           st, d_uo + u_levels[l], d_up, d_ui, d_uv, (const double *)d_ud, (int)ld, d_x);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev[1], st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pad.empty() ? x : pad.data(), d_x, sizeof(double) * n * ld, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && !pad.empty())
        for (size_t i = 0; i < (size_t)n; ++i) memcpy(x + i * k, pad.data() + i * ld, sizeof(double) * k);
    float f = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, ev[0], ev[1]);
    release();
    if (e != hipSuccess) return fail(TH_EHIP, "LU solve: %s", hipGetErrorString(e));
    if (ms) *ms = f;
    return TH_OK;
}

namespace {

// th_trace_record / th_trace_rf_record (t0, flight, clock: RF only).
int trace_record_host(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                      const double *qm, const int8_t *spc, const double *t0, int64_t n, const th_trace_params *p,
                      const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally,
                      int8_t *status, int32_t *hit, double *endpoints, double *final_velocity,
                      const int32_t *slots, int32_t n_probe, double *paths, int32_t *path_lengths,
                      double *flight, double *clock) {
    if (!t || !p || !cnt_i || !cnt_e || !tally || !status || !hit || !endpoints || !final_velocity)
        return fail(TH_EINVAL, "invalid argument");
    if (p->rf && (!flight || !clock || (n > 0 && !t0)))
        return fail(TH_EINVAL, "an RF record trace needs launch times, flight_time and phase_clock");
    if (n < 0 || n_probe < 0 || (n_probe > 0 && (!slots || !paths || !path_lengths)))
        return fail(TH_EINVAL, "invalid probe arguments");
    if (n > 0 && (!x || !y || !vx || !vy || !qm || !spc)) return fail(TH_EINVAL, "particle array is NULL");
    if (p->precision != 0) return fail(TH_EINVAL, "record mode is FP64 only");
    int rc = check_params(t, p);
    if (rc) return rc;
    for (int64_t k = 0; k < n; ++k) {
        if (spc[k] != 0 && spc[k] != 1) return fail(TH_EINVAL, "species must be 0 (ion) or 1 (electron)");
        if (slots && (slots[k] < -1 || slots[k] >= n_probe)) return fail(TH_EINVAL, "probe slot out of range");
    }
    const int stride = p->max_steps + 1;
    // Untraced particles (after a stop) keep their initial point and velocity,
    // and probe paths start with the initial point (trench_cuda.py:89-97).
    for (int64_t k = 0; k < n; ++k) {
        endpoints[2 * k] = x[k];
        endpoints[2 * k + 1] = y[k];
        final_velocity[2 * k] = vx[k];
        final_velocity[2 * k + 1] = vy[k];
        status[k] = 3;
        hit[k] = -1;
        if (p->rf) {
            flight[k] = 0.0;
            clock[k] = host_py_mod(t0[k], t->period);
        }
        if (slots && slots[k] >= 0) {
            paths[2 * (int64_t)slots[k] * stride] = x[k];
            paths[2 * (int64_t)slots[k] * stride + 1] = y[k];
            path_lengths[slots[k]] = 1;
        }
    }
    memset(tally, 0, sizeof *tally);
    memset(cnt_i, 0, sizeof(int64_t) * (size_t)t->N);
    memset(cnt_e, 0, sizeof(int64_t) * (size_t)t->N);
    if (n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    if ((rc = reserve_particles(t, n))) return rc;
    if ((rc = upload(t, 0, x, y, vx, vy, qm, spc, n))) return rc;
    if (p->rf && (rc = upload_times(t, 0, t0, n))) return rc;
    double *d_end = nullptr, *d_vel = nullptr, *d_paths = nullptr, *d_flight = nullptr, *d_clock = nullptr;
    int32_t *d_slots = nullptr, *d_len = nullptr;
    size_t path_doubles = 2 * (size_t)std::max(n_probe, 1) * stride;
    auto release = [&]() {
        for (void *q : {(void *)d_end, (void *)d_vel, (void *)d_paths, (void *)d_slots, (void *)d_len,
                        (void *)d_flight, (void *)d_clock})
            if (q) (void)hipFree(q);
    };
    if ((rc = dev_alloc(&d_end, 2 * (size_t)n)) || (rc = dev_alloc(&d_vel, 2 * (size_t)n)) ||
        (rc = dev_alloc(&d_paths, path_doubles)) || (rc = dev_alloc(&d_slots, (size_t)n)) ||
        (rc = dev_alloc(&d_len, (size_t)std::max(n_probe, 1))) ||
        (p->rf && ((rc = dev_alloc(&d_flight, (size_t)n)) || (rc = dev_alloc(&d_clock, (size_t)n))))) {
        release();
        return rc;
    }
    std::vector<int32_t> sl((size_t)n, -1);
    if (slots) memcpy(sl.data(), slots, sizeof(int32_t) * (size_t)n);
    hipError_t e = hipMemcpyAsync(d_end, endpoints, sizeof(double) * 2 * n, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_vel, final_velocity, sizeof(double) * 2 * n, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_slots, sl.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && n_probe > 0)
        e = hipMemcpyAsync(d_paths, paths, sizeof(double) * 2 * (size_t)n_probe * stride, hipMemcpyHostToDevice,
                           t->stream);
    if (e == hipSuccess && n_probe > 0)
        e = hipMemcpyAsync(d_len, path_lengths, sizeof(int32_t) * n_probe, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && p->rf)
        e = hipMemcpyAsync(d_flight, flight, sizeof(double) * n, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && p->rf)
        e = hipMemcpyAsync(d_clock, clock, sizeof(double) * n, hipMemcpyHostToDevice, t->stream);
    if (e != hipSuccess) {
        release();
        return fail(TH_EHIP, "record upload: %s", hipGetErrorString(e));
    }
    th::Record rec{d_end, d_vel, d_slots, d_paths, d_len, stride, d_flight, d_clock};
    int64_t traced = 0;
    rc = launch_trace(t, 0, n, p, stop, true, true, &traced, tally, &rec);
    if (rc) {
        (void)hipStreamSynchronize(t->stream);
        release();
        return rc;
    }
    std::vector<uint32_t> counts(2 * (size_t)std::max(t->m, 1));
    unsigned long long scal[th::kScalCount];
    e = hipMemcpyAsync(scal, t->d_scal, sizeof scal, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts.data(), t->d_counts, sizeof(uint32_t) * counts.size(), hipMemcpyDeviceToHost,
                           t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(status, t->d_status, (size_t)n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hit, t->d_hit, sizeof(int32_t) * n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(endpoints, d_end, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(final_velocity, d_vel, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && n_probe > 0)
        e = hipMemcpyAsync(paths, d_paths, sizeof(double) * 2 * (size_t)n_probe * stride, hipMemcpyDeviceToHost,
                           t->stream);
    if (e == hipSuccess && n_probe > 0)
        e = hipMemcpyAsync(path_lengths, d_len, sizeof(int32_t) * n_probe, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && p->rf)
        e = hipMemcpyAsync(flight, d_flight, sizeof(double) * n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && p->rf)
        e = hipMemcpyAsync(clock, d_clock, sizeof(double) * n, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    release();
    if (e != hipSuccess) return fail(TH_EHIP, "record readback: %s", hipGetErrorString(e));
    for (int k = 0; k < t->m; ++k) {
        cnt_i[t->surf[k]] += counts[k];
        cnt_e[t->surf[k]] += counts[t->m + k];
    }
    if (scal[th::kScalOff]) {
        std::vector<uint32_t> over(2 * (size_t)t->N);
        TH_HIP(hipMemcpy(over.data(), t->d_overflow, sizeof(uint32_t) * over.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < t->N; ++c) {
            cnt_i[c] += over[c];
            cnt_e[c] += over[t->N + c];
        }
    }
    fill_tally(scal, n - traced, tally);
    return TH_OK;
}

}  // namespace

extern "C" int th_trace_record(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                               const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
                               const volatile int32_t *stop, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally,
                               int8_t *status, int32_t *hit, double *endpoints, double *final_velocity,
                               const int32_t *slots, int32_t n_probe, double *paths, int32_t *path_lengths) {
    if (p && p->rf) return fail(TH_EINVAL, "an RF trace takes launch times: use th_trace_rf_record");
    return trace_record_host(t, x, y, vx, vy, qm, spc, nullptr, n, p, stop, cnt_i, cnt_e, tally, status, hit,
                             endpoints, final_velocity, slots, n_probe, paths, path_lengths, nullptr, nullptr);
}

extern "C" int th_trace_rf_record(th_tracer *t, const double *x, const double *y, const double *vx,
                                  const double *vy, const double *qm, const int8_t *spc, const double *initial_time,
                                  int64_t n, const th_trace_params *p, const volatile int32_t *stop, int64_t *cnt_i,
                                  int64_t *cnt_e, th_tally *tally, int8_t *status, int32_t *hit, double *endpoints,
                                  double *final_velocity, const int32_t *slots, int32_t n_probe, double *paths,
                                  int32_t *path_lengths, double *flight_time, double *phase_clock) {
    if (p && !p->rf) return fail(TH_EINVAL, "th_trace_rf_record needs params.rf = 1");
    return trace_record_host(t, x, y, vx, vy, qm, spc, initial_time, n, p, stop, cnt_i, cnt_e, tally, status, hit,
                             endpoints, final_velocity, slots, n_probe, paths, path_lengths, flight_time,
                             phase_clock);
}

extern "C" int th_trace_tangent(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                                const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
                                const double *dx0, const double *dy0, const double *dvx0, const double *dvy0,
                                const double *dEx, const double *dEy, th_tally *tally, int8_t *status, int32_t *hit,
                                int32_t *steps, int8_t *face, double *endpoints, double *d_endpoints,
                                double *final_velocity, double *d_final_velocity) {
    if (!t || !p || !tally || !status || !hit || !steps || !face || !endpoints || !d_endpoints || !final_velocity ||
        !d_final_velocity)
        return fail(TH_EINVAL, "invalid argument");
    if (n < 0) return fail(TH_EINVAL, "invalid particle count");
    if (n > 0 && (!x || !y || !vx || !vy || !qm || !spc)) return fail(TH_EINVAL, "particle array is NULL");
    if ((dEx == nullptr) != (dEy == nullptr)) return fail(TH_EINVAL, "dEx and dEy go together");
    if (p->precision != 0 || p->rf) return fail(TH_EINVAL, "the tangent push follows the FP64 push without RF");
    int rc = check_params(t, p);
    if (rc) return rc;
    for (int64_t k = 0; k < n; ++k)
        if (spc[k] != 0 && spc[k] != 1) return fail(TH_EINVAL, "species must be 0 (ion) or 1 (electron)");
    memset(tally, 0, sizeof *tally);
    if (n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    if ((rc = reserve_particles(t, n))) return rc;
    if ((rc = upload(t, 0, x, y, vx, vy, qm, spc, n))) return rc;
    // per-particle buffers: 4 initial tangents, end point, its tangent, velocity, its tangent
    double *d_buf = nullptr;
    double2 *d_field = nullptr;
    int8_t *d_face = nullptr;
    int32_t *d_steps = nullptr;
    auto release = [&]() {
        for (void *q : {(void *)d_buf, (void *)d_field, (void *)d_face, (void *)d_steps})
            if (q) (void)hipFree(q);
    };
    const size_t nn = (size_t)n;
    if ((rc = dev_alloc(&d_buf, 12 * nn)) || (dEx && (rc = dev_alloc(&d_field, (size_t)t->N))) ||
        (rc = dev_alloc(&d_face, nn)) || (rc = dev_alloc(&d_steps, nn))) {
        release();
        return rc;
    }
    double *d_init[4] = {d_buf, d_buf + nn, d_buf + 2 * nn, d_buf + 3 * nn};
    const double *h_init[4] = {dx0, dy0, dvx0, dvy0};
    double *d_end = d_buf + 4 * nn, *d_dend = d_buf + 6 * nn, *d_vel = d_buf + 8 * nn, *d_dvel = d_buf + 10 * nn;
    hipError_t e = hipSuccess;
    for (int c = 0; c < 4 && e == hipSuccess; ++c)
        if (h_init[c]) e = hipMemcpyAsync(d_init[c], h_init[c], sizeof(double) * nn, hipMemcpyHostToDevice, t->stream);
    std::vector<double2> field;
    if (dEx) {   // vacuum-masked, as the push's own records
        field.resize((size_t)t->N);
        for (int c = 0; c < t->N; ++c)
            field[c] = t->vac[c] ? make_double2(dEx[c], dEy[c]) : make_double2(0.0, 0.0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_field, field.data(), sizeof(double2) * field.size(), hipMemcpyHostToDevice,
                               t->stream);
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(t->stream);
        release();
        return fail(TH_EHIP, "tangent upload: %s", hipGetErrorString(e));
    }
    th::Record rec{d_end, d_vel, nullptr, nullptr, nullptr, p->max_steps + 1, nullptr, nullptr};
    th::TangentIO tan{h_init[0] ? d_init[0] : nullptr, h_init[1] ? d_init[1] : nullptr,
                      h_init[2] ? d_init[2] : nullptr, h_init[3] ? d_init[3] : nullptr, d_field, d_dend, d_dvel,
                      d_face, d_steps};
    int64_t traced = 0;
    rc = launch_trace(t, 0, n, p, nullptr, true, true, &traced, tally, &rec, &tan);
    if (rc) {
        (void)hipStreamSynchronize(t->stream);
        release();
        return rc;
    }
    unsigned long long scal[th::kScalCount];
    e = hipMemcpyAsync(scal, t->d_scal, sizeof scal, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(status, t->d_status, nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hit, t->d_hit, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(steps, d_steps, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(face, d_face, nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(endpoints, d_end, sizeof(double) * 2 * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_endpoints, d_dend, sizeof(double) * 2 * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(final_velocity, d_vel, sizeof(double) * 2 * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_final_velocity, d_dvel, sizeof(double) * 2 * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    release();
    if (e != hipSuccess) return fail(TH_EHIP, "tangent readback: %s", hipGetErrorString(e));
    float ms = 0.0f;
    TH_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->push_ms = ms;
    fill_tally(scal, n - traced, tally);
    return TH_OK;
}

static_assert(sizeof(th_soft_event) == sizeof(th::SoftEvent) && offsetof(th_soft_event, da) == offsetof(th::SoftEvent, da) &&
                  offsetof(th_soft_event, pid) == offsetof(th::SoftEvent, pid) &&
                  offsetof(th_soft_event, face) == offsetof(th::SoftEvent, face) &&
                  offsetof(th_soft_event, species) == offsetof(th::SoftEvent, sp),
              "th_soft_event mirrors th::SoftEvent");

extern "C" int th_trace_soft(th_tracer *t, int64_t offset, int64_t n, const th_trace_params *p, double width,
                             double cut, const double *dx0, const double *dy0, const double *dvx0, const double *dvy0,
                             const double *dEx, const double *dEy, const double *dphi, th_tally *tally,
                             double *deposit, double *d_deposit, int8_t *status, int32_t *hit, int32_t *steps,
                             th_soft_event *events, int64_t event_capacity, int64_t *n_events) {
    if (!t || !p || !tally || !deposit || !d_deposit || offset < 0 || n < 0) return fail(TH_EINVAL, "invalid argument");
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    if ((dEx == nullptr) != (dEy == nullptr)) return fail(TH_EINVAL, "dEx and dEy go together");
    if (!(width > 0.0) || !(cut > 0.0) || !std::isfinite(width) || !std::isfinite(cut))
        return fail(TH_EINVAL, "width and cut must be positive and finite");
    if (event_capacity < 0 || event_capacity > (int64_t)0xffffffffu || (event_capacity > 0 && !events))
        return fail(TH_EINVAL, "invalid event buffer");
    if (!p->barrier) return fail(TH_EINVAL, "the relaxed-barrier push needs the barrier wall model");
    if (p->precision != 0 || p->rf) return fail(TH_EINVAL, "the relaxed-barrier push follows the FP64 push without RF");
    int rc = check_params(t, p);
    if (rc) return rc;
    memset(tally, 0, sizeof *tally);
    const size_t m2 = 2 * (size_t)t->m;
    memset(deposit, 0, sizeof(double) * m2);
    memset(d_deposit, 0, sizeof(double) * m2);
    if (n_events) *n_events = 0;
    if (n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    // device buffers: initial tangents (pool-indexed), field and potential tangents, deposits, events
    const size_t span = (size_t)(offset + n), nn = (size_t)n;
    double *d_init = nullptr, *d_dphi = nullptr, *d_dep = nullptr;
    double2 *d_field = nullptr;
    int32_t *d_steps = nullptr;
    th::SoftEvent *d_ev = nullptr;
    unsigned *d_nev = nullptr;
    auto release = [&]() {
        for (void *q : {(void *)d_init, (void *)d_dphi, (void *)d_dep, (void *)d_field, (void *)d_steps, (void *)d_ev,
                        (void *)d_nev})
            if (q) (void)hipFree(q);
    };
    const double *h_init[4] = {dx0, dy0, dvx0, dvy0};
    const bool any_init = dx0 || dy0 || dvx0 || dvy0;
    if ((any_init && (rc = dev_alloc(&d_init, 4 * span))) || (dEx && (rc = dev_alloc(&d_field, (size_t)t->N))) ||
        (dphi && (rc = dev_alloc(&d_dphi, (size_t)t->N))) || (rc = dev_alloc(&d_dep, 2 * m2)) ||
        (steps && (rc = dev_alloc(&d_steps, span))) ||
        (event_capacity > 0 && (rc = dev_alloc(&d_ev, (size_t)event_capacity))) || (rc = dev_alloc(&d_nev, 1))) {
        release();
        return rc;
    }
    hipError_t e = hipMemsetAsync(d_dep, 0, sizeof(double) * 2 * m2, t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_nev, 0, sizeof(unsigned), t->stream);
    for (int c = 0; c < 4 && e == hipSuccess; ++c)
        if (h_init[c])
            e = hipMemcpyAsync(d_init + c * span + offset, h_init[c], sizeof(double) * nn, hipMemcpyHostToDevice,
                               t->stream);
    std::vector<double2> field;
    if (dEx && e == hipSuccess) {   // vacuum-masked, as the push's own records
        field.resize((size_t)t->N);
        for (int c = 0; c < t->N; ++c)
            field[c] = t->vac[c] ? make_double2(dEx[c], dEy[c]) : make_double2(0.0, 0.0);
        e = hipMemcpyAsync(d_field, field.data(), sizeof(double2) * field.size(), hipMemcpyHostToDevice, t->stream);
    }
    if (dphi && e == hipSuccess)
        e = hipMemcpyAsync(d_dphi, dphi, sizeof(double) * (size_t)t->N, hipMemcpyHostToDevice, t->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(t->stream);
        release();
        return fail(TH_EHIP, "relaxed push upload: %s", hipGetErrorString(e));
    }
    th::Record rec{};
    th::TangentIO tan{};
    for (int c = 0; c < 4; ++c) {
        const double *q = h_init[c] ? d_init + c * span : nullptr;
        if (c == 0) tan.dx0 = q; else if (c == 1) tan.dy0 = q; else if (c == 2) tan.dvx0 = q; else tan.dvy0 = q;
    }
    tan.dfield = d_field;
    tan.steps = d_steps;
    tan.soft.dphi = d_dphi;
    tan.soft.inv_width = 1.0 / width;
    tan.soft.cut = cut;
    tan.soft.dep = d_dep;
    tan.soft.ddep = d_dep + m2;
    tan.soft.events = d_ev;
    tan.soft.n_events = d_nev;
    tan.soft.ev_cap = (unsigned)event_capacity;
    int64_t traced = 0;
    rc = launch_trace(t, offset, n, p, nullptr, status != nullptr, hit != nullptr, &traced, tally, &rec, &tan);
    if (rc) {
        (void)hipStreamSynchronize(t->stream);
        release();
        return rc;
    }
    unsigned long long scal[th::kScalCount];
    unsigned nev = 0;
    e = hipMemcpyAsync(scal, t->d_scal, sizeof scal, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&nev, d_nev, sizeof nev, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(deposit, d_dep, sizeof(double) * m2, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_deposit, d_dep + m2, sizeof(double) * m2, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && status)
        e = hipMemcpyAsync(status, t->d_status + offset, nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && hit)
        e = hipMemcpyAsync(hit, t->d_hit + offset, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && steps)
        e = hipMemcpyAsync(steps, d_steps + offset, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e == hipSuccess && events && nev > 0)
        e = hipMemcpy(events, d_ev, sizeof(th::SoftEvent) * std::min<size_t>(nev, (size_t)event_capacity),
                      hipMemcpyDeviceToHost);
    release();
    if (e != hipSuccess) return fail(TH_EHIP, "relaxed push readback: %s", hipGetErrorString(e));
    if (n_events) *n_events = (int64_t)nev;
    float ms = 0.0f;
    TH_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->push_ms = ms;
    fill_tally(scal, offset + n - traced, tally);
    return TH_OK;
}

extern "C" int th_trace_adjoint(th_tracer *t, const double *x, const double *y, const double *vx, const double *vy,
                                const double *qm, const int8_t *spc, int64_t n, const th_trace_params *p,
                                const int32_t *steps, const double *bar_end, const double *bar_vel, double *bar_state,
                                double *bar_field, double *ms) {
    if (!t || !p || !steps || !bar_state) return fail(TH_EINVAL, "invalid argument");
    if (n < 0) return fail(TH_EINVAL, "invalid particle count");
    if (n > 0 && (!x || !y || !vx || !vy || !qm || !spc)) return fail(TH_EINVAL, "particle array is NULL");
    if (p->precision != 0 || p->rf) return fail(TH_EINVAL, "the adjoint push follows the FP64 push without RF");
    int rc = check_params(t, p);
    if (rc) return rc;
    for (int64_t k = 0; k < n; ++k) {
        if (spc[k] != 0 && spc[k] != 1) return fail(TH_EINVAL, "species must be 0 (ion) or 1 (electron)");
        if (steps[k] < 1 || steps[k] > p->max_steps)
            return fail(TH_EINVAL, "steps[%lld] = %d is outside [1, max_steps]", (long long)k, steps[k]);
    }
    if (bar_field) memset(bar_field, 0, sizeof(double) * 2 * (size_t)t->N);
    if (ms) *ms = 0.0;
    if (n == 0) {
        return TH_OK;
    }
    TH_HIP(hipSetDevice(t->device));
    if ((rc = reserve_particles(t, n))) return rc;
    if ((rc = upload(t, 0, x, y, vx, vy, qm, spc, n))) return rc;
    // Lanes take particles longest first (waves of similar path lengths).  The
    // 64 consecutive positions of `order` that one wave traces share a tape
    // region of 64 x (their longest path) entries, step-major, so that the
    // wave's tape stores and loads of one step are contiguous; the tape is
    // filled chunk by chunk (whole groups), at most kTapeSteps entries (32
    // bytes each) at once.
    const size_t nn = (size_t)n;
    std::vector<int32_t> order(nn);
    for (size_t k = 0; k < nn; ++k) order[k] = (int32_t)k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return steps[a] > steps[b]; });
    constexpr int64_t kTapeSteps = (int64_t)1 << 27;   // 4 GiB of tape per chunk
    const size_t ng = (nn + 63) / 64;
    std::vector<int64_t> goff(ng);
    std::vector<std::pair<int64_t, int64_t>> chunks;
    int64_t run = 0, begin = 0, tape_steps = 0;
    for (size_t gi = 0; gi < ng; ++gi) {
        const int64_t len = 64 * (int64_t)steps[order[64 * gi]];   // the group's longest path (sorted)
        if (run + len > kTapeSteps && run > 0) {
            chunks.emplace_back(begin, (int64_t)(64 * gi));
            tape_steps = std::max(tape_steps, run);
            begin = (int64_t)(64 * gi);
            run = 0;
        }
        goff[gi] = run;
        run += len;
    }
    chunks.emplace_back(begin, (int64_t)nn);
    tape_steps = std::max(tape_steps, run);
    int32_t *d_order = nullptr, *d_steps = nullptr;
    int64_t *d_off = nullptr;
    double *d_bar = nullptr, *d_field = nullptr;
    double4 *d_tape = nullptr;
    unsigned long long *d_mis = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(t->stream);
        for (void *q : {(void *)d_order, (void *)d_steps, (void *)d_off, (void *)d_bar, (void *)d_field,
                        (void *)d_tape, (void *)d_mis})
            if (q) (void)hipFree(q);
    };
    if ((rc = dev_alloc(&d_order, nn)) || (rc = dev_alloc(&d_steps, nn)) || (rc = dev_alloc(&d_off, ng)) ||
        (rc = dev_alloc(&d_bar, 8 * nn)) || (bar_field && (rc = dev_alloc(&d_field, 2 * (size_t)t->N))) ||
        (rc = dev_alloc(&d_tape, (size_t)tape_steps)) || (rc = dev_alloc(&d_mis, 1))) {
        release();
        return rc;
    }
    double *d_end = d_bar, *d_vel = d_bar + 2 * nn, *d_state = d_bar + 4 * nn;
    hipError_t e = hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_steps, steps, sizeof(int32_t) * nn, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, goff.data(), sizeof(int64_t) * ng, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && bar_end)
        e = hipMemcpyAsync(d_end, bar_end, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && bar_vel)
        e = hipMemcpyAsync(d_vel, bar_vel, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && d_field) e = hipMemsetAsync(d_field, 0, sizeof(double) * 2 * (size_t)t->N, t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_mis, 0, sizeof(unsigned long long), t->stream);
    if (e != hipSuccess) {
        release();
        return fail(TH_EHIP, "adjoint upload: %s", hipGetErrorString(e));
    }
    th::Grid g = make_grid(t, p);
    th::AdjointIO a{t->d_x, t->d_y, t->d_vx, t->d_vy, t->d_qm, d_order, d_steps, d_off, d_tape,
                    bar_end ? d_end : nullptr, bar_vel ? d_vel : nullptr, d_state, d_field, d_mis, 0, 0,
                    p->max_steps};
    e = hipEventRecord(t->ev0, t->stream);
    for (const auto &ch : chunks) {
        if (e != hipSuccess) break;
        a.begin = ch.first;
        a.end = ch.second;
        const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((ch.second - ch.first + 255) / 256,
                                                                          (int64_t)t->cus * 8));
        auto kernel = p->barrier ? (p->tapered ? th::push64_adjoint_kernel<true, true>
                                               : th::push64_adjoint_kernel<true, false>)
                                 : (p->tapered ? th::push64_adjoint_kernel<false, true>
                                               : th::push64_adjoint_kernel<false, false>);
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, t->stream, g, a);
        e = hipGetLastError();
        if (e != hipSuccess) break;
    }
    if (e == hipSuccess) e = hipEventRecord(t->ev1, t->stream);
    unsigned long long mismatch = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(bar_state, d_state, sizeof(double) * 4 * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && bar_field)
        e = hipMemcpyAsync(bar_field, d_field, sizeof(double) * 2 * (size_t)t->N, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&mismatch, d_mis, sizeof mismatch, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    float el = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&el, t->ev0, t->ev1);
    release();
    if (e != hipSuccess) return fail(TH_EHIP, "adjoint push: %s", hipGetErrorString(e));
    if (ms) *ms = el;
    if (mismatch)
        return fail(TH_EINVAL, "%llu particle(s) took a different number of steps than `steps` gives: the steps "
                               "must come from a trace of the same particles and field",
                    mismatch);
    return TH_OK;
}

extern "C" int th_trace_soft_adjoint(th_tracer *t, int64_t offset, int64_t n, const th_trace_params *p,
                                     double width, double cut, const int32_t *steps, const double *gbar,
                                     double *bar_state, double *bar_field, double *bar_phi, double *ms) {
    if (!t || !p || !steps || !gbar || !bar_state || offset < 0 || n < 0) return fail(TH_EINVAL, "invalid argument");
    if (offset + n > t->cap) return fail(TH_EINVAL, "pool holds %lld particles", (long long)t->cap);
    if (!(width > 0.0) || !(cut > 0.0) || !std::isfinite(width) || !std::isfinite(cut))
        return fail(TH_EINVAL, "width and cut must be positive and finite");
    if (!p->barrier) return fail(TH_EINVAL, "the relaxed-barrier push needs the barrier wall model");
    if (p->precision != 0 || p->rf) return fail(TH_EINVAL, "the relaxed-barrier push follows the FP64 push without RF");
    int rc = check_params(t, p);
    if (rc) return rc;
    for (int64_t k = 0; k < n; ++k)
        if (steps[k] < 1 || steps[k] > p->max_steps)
            return fail(TH_EINVAL, "steps[%lld] = %d is outside [1, max_steps]", (long long)k, steps[k]);
    if (bar_field) memset(bar_field, 0, sizeof(double) * 2 * (size_t)t->N);
    if (bar_phi) memset(bar_phi, 0, sizeof(double) * (size_t)t->N);
    if (ms) *ms = 0.0;
    if (n == 0) return TH_OK;
    TH_HIP(hipSetDevice(t->device));
    // tape layout and chunks as th_trace_adjoint: longest paths first, 64
    // positions of `order` per wave sharing a step-major tape region
    const size_t nn = (size_t)n, m2 = 2 * (size_t)t->m;
    std::vector<int32_t> order(nn);
    for (size_t k = 0; k < nn; ++k) order[k] = (int32_t)k;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return steps[a] > steps[b]; });
    constexpr int64_t kTapeSteps = (int64_t)1 << 27;
    const size_t ng = (nn + 63) / 64;
    std::vector<int64_t> goff(ng);
    std::vector<std::pair<int64_t, int64_t>> chunks;
    int64_t run = 0, begin = 0, tape_steps = 0;
    for (size_t gi = 0; gi < ng; ++gi) {
        const int64_t len = 64 * (int64_t)steps[order[64 * gi]];
        if (run + len > kTapeSteps && run > 0) {
            chunks.emplace_back(begin, (int64_t)(64 * gi));
            tape_steps = std::max(tape_steps, run);
            begin = (int64_t)(64 * gi);
            run = 0;
        }
        goff[gi] = run;
        run += len;
    }
    chunks.emplace_back(begin, (int64_t)nn);
    tape_steps = std::max(tape_steps, run);
    int32_t *d_order = nullptr, *d_steps = nullptr;
    int64_t *d_off = nullptr;
    double *d_state = nullptr, *d_field = nullptr, *d_phi = nullptr, *d_g = nullptr;
    double4 *d_tape = nullptr;
    unsigned long long *d_mis = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(t->stream);
        for (void *q : {(void *)d_order, (void *)d_steps, (void *)d_off, (void *)d_state, (void *)d_field,
                        (void *)d_phi, (void *)d_g, (void *)d_tape, (void *)d_mis})
            if (q) (void)hipFree(q);
    };
    if ((rc = dev_alloc(&d_order, nn)) || (rc = dev_alloc(&d_steps, nn)) || (rc = dev_alloc(&d_off, ng)) ||
        (rc = dev_alloc(&d_state, 4 * nn)) || (bar_field && (rc = dev_alloc(&d_field, 2 * (size_t)t->N))) ||
        (bar_phi && (rc = dev_alloc(&d_phi, (size_t)t->N))) || (rc = dev_alloc(&d_g, m2)) ||
        (rc = dev_alloc(&d_tape, (size_t)tape_steps)) || (rc = dev_alloc(&d_mis, 1))) {
        release();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_steps, steps, sizeof(int32_t) * nn, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, goff.data(), sizeof(int64_t) * ng, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_g, gbar, sizeof(double) * m2, hipMemcpyHostToDevice, t->stream);
    if (e == hipSuccess && d_field) e = hipMemsetAsync(d_field, 0, sizeof(double) * 2 * (size_t)t->N, t->stream);
    if (e == hipSuccess && d_phi) e = hipMemsetAsync(d_phi, 0, sizeof(double) * (size_t)t->N, t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_mis, 0, sizeof(unsigned long long), t->stream);
    if (e != hipSuccess) {
        release();
        return fail(TH_EHIP, "relaxed adjoint upload: %s", hipGetErrorString(e));
    }
    th::Grid g = make_grid(t, p);
    th::SoftAdjointIO a{t->d_x + offset, t->d_y + offset, t->d_vx + offset, t->d_vy + offset, t->d_qm + offset,
                        t->d_spc + offset, d_order, d_steps, d_off, d_tape, d_g, 1.0 / width, cut, d_state, d_field,
                        d_phi, d_mis, 0, 0, p->max_steps};
    e = hipEventRecord(t->ev0, t->stream);
    for (const auto &ch : chunks) {
        if (e != hipSuccess) break;
        a.begin = ch.first;
        a.end = ch.second;
        const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((ch.second - ch.first + 255) / 256,
                                                                          (int64_t)t->cus * 8));
        auto kernel = p->tapered ? th::push64_soft_adjoint_kernel<true> : th::push64_soft_adjoint_kernel<false>;
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, t->stream, g, a);
        e = hipGetLastError();
        if (e != hipSuccess) break;
    }
    if (e == hipSuccess) e = hipEventRecord(t->ev1, t->stream);
    unsigned long long mismatch = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(bar_state, d_state, sizeof(double) * 4 * nn, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && bar_field)
        e = hipMemcpyAsync(bar_field, d_field, sizeof(double) * 2 * (size_t)t->N, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess && bar_phi)
        e = hipMemcpyAsync(bar_phi, d_phi, sizeof(double) * (size_t)t->N, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&mismatch, d_mis, sizeof mismatch, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    float el = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&el, t->ev0, t->ev1);
    release();
    if (e != hipSuccess) return fail(TH_EHIP, "relaxed adjoint push: %s", hipGetErrorString(e));
    if (ms) *ms = el;
    if (mismatch)
        return fail(TH_EINVAL, "%llu particle(s) took a different number of steps than `steps` gives: the steps "
                               "must come from th_trace_soft of the same particles, field, width and cut",
                    mismatch);
    return TH_OK;
}

extern "C" int th_charging_set_monitors(th_charging *c, int32_t n_probes, const int32_t *probe_cells,
                                        int32_t n_regions, const int32_t *region_ptr, const int32_t *region_idx,
                                        int32_t abs_region, int32_t n_bottom, const int32_t *bottom_s) {
    if (!c || n_probes < 0 || n_regions < 0 || n_bottom < 0) return fail(TH_EINVAL, "invalid argument");
    if ((n_probes && !probe_cells) || !region_ptr || (n_bottom && !bottom_s))
        return fail(TH_EINVAL, "monitor array is NULL");
    if (abs_region < -1 || abs_region >= n_regions) return fail(TH_EINVAL, "abs_region out of range");
    th_tracer *t = c->t;
    for (int i = 0; i < n_probes; ++i)
        if (probe_cells[i] < 0 || probe_cells[i] >= t->N) return fail(TH_EINVAL, "probe cell out of range");
    if (region_ptr[0] != 0) return fail(TH_EINVAL, "region_ptr must start at 0");
    for (int r = 0; r < n_regions; ++r)
        if (region_ptr[r + 1] < region_ptr[r]) return fail(TH_EINVAL, "region_ptr must not decrease");
    int32_t n_idx = region_ptr[n_regions];
    if (n_idx && !region_idx) return fail(TH_EINVAL, "region_idx is NULL");
    for (int j = 0; j < n_idx; ++j)
        if (region_idx[j] < 0 || region_idx[j] >= t->m) return fail(TH_EINVAL, "region index out of range");
    for (int j = 0; j < n_bottom; ++j)
        if (bottom_s[j] < 0 || bottom_s[j] >= t->m) return fail(TH_EINVAL, "floor index out of range");
    TH_HIP(hipSetDevice(t->device));
    TH_HIP(hipStreamSynchronize(t->stream));
    size_t total = (size_t)n_probes + n_regions + 1 + n_idx + n_bottom;
    std::vector<int32_t> buf(std::max<size_t>(total, 1));
    size_t o = 0;
    if (n_probes) memcpy(buf.data() + o, probe_cells, sizeof(int32_t) * n_probes);
    size_t o_ptr = o + n_probes;
    memcpy(buf.data() + o_ptr, region_ptr, sizeof(int32_t) * (n_regions + 1));
    size_t o_idx = o_ptr + n_regions + 1;
    if (n_idx) memcpy(buf.data() + o_idx, region_idx, sizeof(int32_t) * n_idx);
    size_t o_bot = o_idx + n_idx;
    if (n_bottom) memcpy(buf.data() + o_bot, bottom_s, sizeof(int32_t) * n_bottom);
    // the new buffers replace the old ones only once they are complete: on an error the handle keeps its
    // previous monitors
    int32_t *mon_buf = nullptr;
    double *mon_out = nullptr;
    const size_t n_out = (size_t)n_probes + n_regions + 3;
    int rc = dev_alloc(&mon_buf, buf.size());
    if (!rc) rc = dev_alloc(&mon_out, n_out);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpy(mon_buf, buf.data(), sizeof(int32_t) * buf.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemset(mon_out, 0, sizeof(double) * n_out);
    if (rc || e != hipSuccess) {
        if (mon_buf) (void)hipFree(mon_buf);
        if (mon_out) (void)hipFree(mon_out);
        return rc ? rc : fail(TH_EHIP, "monitor upload: %s", hipGetErrorString(e));
    }
    if (c->mon_buf) (void)hipFree(c->mon_buf);
    if (c->mon_out) (void)hipFree(c->mon_out);
    c->mon_buf = mon_buf;
    c->mon_out = mon_out;
    c->mon.n_probes = n_probes;
    c->mon.n_regions = n_regions;
    c->mon.abs_region = abs_region;
    c->mon.n_bottom = n_bottom;
    c->mon.probe_cells = c->mon_buf + o;
    c->mon.region_ptr = c->mon_buf + o_ptr;
    c->mon.region_idx = c->mon_buf + o_idx;
    c->mon.bottom_s = c->mon_buf + o_bot;
    c->mon.out = c->mon_out;
    c->mon_host.assign(n_out, 0.0);
    return TH_OK;
}

extern "C" int th_charging_monitors(const th_charging *c, double *out) {
    if (!c || !out) return fail(TH_EINVAL, "invalid argument");
    if (!c->mon_out) return fail(TH_ESTATE, "no monitors set");
    memcpy(out, c->mon_host.data(), sizeof(double) * c->mon_host.size());
    return TH_OK;
}

// Ensemble push (many trench instances per launch): same translation unit.
#include "th_ensemble_capi.hpp"
