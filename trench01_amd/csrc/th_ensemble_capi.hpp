// C-ABI of the ensemble push (declarations in include/trench_hip.h).  Part of
// the th_capi.hip translation unit: it uses that file's error helpers,
// surface-cell rule and pass limit.
#include "th_ensemble.hpp"

namespace {
constexpr int kCountReplicas = 16;        // copies of the deposit counters (see th_ensemble.hpp)
constexpr int64_t kReduceChunk = 1 << 15;   // particles per workgroup of the end-record reduction
}

struct th_ensemble {
    int device = 0, cus = 0;
    int per_cu[2][2] = {{1, 1}, {1, 1}};   // resident workgroups per CU: [register budget 5 / 4][barrier]
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // instance table and host staging (until the arenas are built)
    std::vector<th::Instance> inst;
    std::vector<std::vector<int32_t>> surf;   // surface cells per instance, ascending flat index
    std::vector<int32_t> slot0;               // first global surface slot per instance
    std::vector<th::InjectDev> inject;
    std::vector<uint8_t> has_inject, has_phi;
    std::vector<double4> h_cells;
    std::vector<uint8_t> h_vac;
    std::vector<double> h_wface;
    std::vector<int32_t> h_slot;
    int64_t n_cells = 0, n_keys = 0, n_pool = 0;
    int slots = 0;
    bool built = false;
    // arenas
    double4 *d_cells = nullptr;
    double2 *d_corners = nullptr;
    uint8_t *d_corner_vac = nullptr, *d_vac = nullptr;
    double *d_wface = nullptr;
    int32_t *d_slot = nullptr;
    th::Instance *d_inst = nullptr;
    int64_t *d_starts = nullptr;
    th::InjectDev *d_inject = nullptr;
    th::Chunk *d_chunks = nullptr;   // reduction ranges (kReduceChunk particles of one instance)
    int n_chunks = 0;
    // pool and tallies
    double *d_x = nullptr, *d_y = nullptr, *d_vx = nullptr, *d_vy = nullptr, *d_qm = nullptr;
    int8_t *d_spc = nullptr, *d_status = nullptr;   // status: end record per particle
    int32_t *d_hit = nullptr, *d_steps = nullptr;
    th::Survivor *d_surv[2] = {nullptr, nullptr};
    unsigned *d_surv_count = nullptr;        // [kMaxPasses + 1]
    unsigned long long *d_queue = nullptr;   // [kMaxPasses] work counters
    unsigned long long *d_batch = nullptr;   // batch index read by the injection
    uint32_t *d_counts = nullptr;            // [kCountReplicas][2][slots]
    unsigned long long *d_scal = nullptr;    // [n_inst][kInstCount]
    // captured batch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    th_trace_params graph_params{};
    uint64_t graph_seed = 0;
    int graph_flags = -1;
};

namespace {

void ensemble_free_graph(th_ensemble *e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    e->graph_flags = -1;
}

// Cell records {Ex*vac, Ey*vac, vac, phi} and wall weights of one instance.
void pack_host_cells(const th_ensemble *e, int index, const double *Ex, const double *Ey, const double *phi,
                     const double *wface, double4 *cells, double *wf) {
    const th::Instance &d = e->inst[index];
    const int64_t N = (int64_t)d.nx * d.ny;
    const uint8_t *vac = e->h_vac.data() + d.cell0;
    for (int64_t c = 0; c < N; ++c) {
        bool v = vac[c] != 0;
        cells[c] = make_double4(v ? Ex[c] : 0.0, v ? Ey[c] : 0.0, v ? 1.0 : 0.0, phi ? phi[c] : 0.0);
        wf[c] = wface ? wface[c] : 0.0;
    }
}

int ensemble_corners(th_ensemble *e, int index) {
    const th::Instance &d = e->inst[index];
    int keys = (d.nx + 1) * (d.ny + 1);
    hipLaunchKernelGGL(th::pack_corners_kernel, dim3((keys + 255) / 256), dim3(256), 0, e->stream, e->d_cells,
                       e->d_vac, d.nx, d.ny, d.cell0, d.key0, e->d_corners, e->d_corner_vac);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

void ensemble_free_device(th_ensemble *e) {
    for (void **q : {(void **)&e->d_cells, (void **)&e->d_corners, (void **)&e->d_corner_vac, (void **)&e->d_vac,
                     (void **)&e->d_steps, (void **)&e->d_chunks, (void **)&e->d_wface, (void **)&e->d_slot,
                     (void **)&e->d_inst, (void **)&e->d_starts, (void **)&e->d_inject, (void **)&e->d_x,
                     (void **)&e->d_y, (void **)&e->d_vx, (void **)&e->d_vy, (void **)&e->d_qm, (void **)&e->d_spc,
                     (void **)&e->d_status, (void **)&e->d_hit, (void **)&e->d_surv[0], (void **)&e->d_surv[1],
                     (void **)&e->d_surv_count, (void **)&e->d_queue, (void **)&e->d_batch, (void **)&e->d_counts,
                     (void **)&e->d_scal}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
}

int ensemble_build_device(th_ensemble *e);

// Arenas, pool and tallies on the device; on failure nothing stays allocated.
int ensemble_build(th_ensemble *e) {
    if (e->built) return TH_OK;
    int rc = ensemble_build_device(e);
    if (rc) ensemble_free_device(e);
    return rc;
}

int ensemble_build_device(th_ensemble *e) {
    if (e->inst.empty()) return fail(TH_ESTATE, "the ensemble has no instances");
    const int n = (int)e->inst.size();
    const int64_t pool = std::max<int64_t>(e->n_pool, 1);
    int rc;
    if ((rc = dev_alloc(&e->d_cells, e->n_cells)) || (rc = dev_alloc(&e->d_corners, 4 * e->n_keys)) ||
        (rc = dev_alloc(&e->d_corner_vac, e->n_keys)) || (rc = dev_alloc(&e->d_vac, e->n_cells)) ||
        (rc = dev_alloc(&e->d_wface, e->n_cells)) || (rc = dev_alloc(&e->d_slot, e->n_cells)) ||
        (rc = dev_alloc(&e->d_inst, n)) || (rc = dev_alloc(&e->d_starts, n + 1)) ||
        (rc = dev_alloc(&e->d_inject, n)) || (rc = dev_alloc(&e->d_x, pool)) || (rc = dev_alloc(&e->d_y, pool)) ||
        (rc = dev_alloc(&e->d_vx, pool)) || (rc = dev_alloc(&e->d_vy, pool)) || (rc = dev_alloc(&e->d_qm, pool)) ||
        (rc = dev_alloc(&e->d_spc, pool)) || (rc = dev_alloc(&e->d_status, pool)) ||
        (rc = dev_alloc(&e->d_hit, pool)) || (rc = dev_alloc(&e->d_steps, pool)) ||
        (rc = dev_alloc(&e->d_surv[0], pool)) ||
        (rc = dev_alloc(&e->d_surv[1], pool)) || (rc = dev_alloc(&e->d_surv_count, kMaxPasses + 1)) ||
        (rc = dev_alloc(&e->d_queue, kMaxPasses)) || (rc = dev_alloc(&e->d_batch, 1)) ||
        (rc = dev_alloc(&e->d_counts, (size_t)kCountReplicas * 2 * std::max(e->slots, 1))) ||
        (rc = dev_alloc(&e->d_scal, (size_t)n * th::kInstCount)))
        return rc;
    std::vector<int64_t> starts(n + 1);
    for (int i = 0; i < n; ++i) starts[i] = e->inst[i].p0;
    starts[n] = e->n_pool;
    std::vector<th::Chunk> chunks;
    for (int i = 0; i < n; ++i)
        for (int64_t b = e->inst[i].p0; b < e->inst[i].p1; b += kReduceChunk)
            chunks.push_back(th::Chunk{b, std::min(e->inst[i].p1, b + kReduceChunk), i, 0});
    e->n_chunks = (int)chunks.size();
    if ((rc = dev_alloc(&e->d_chunks, std::max<size_t>(chunks.size(), 1)))) return rc;
    if (!chunks.empty())
        TH_HIP(hipMemcpy(e->d_chunks, chunks.data(), sizeof(th::Chunk) * chunks.size(), hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_cells, e->h_cells.data(), sizeof(double4) * e->n_cells, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_vac, e->h_vac.data(), e->n_cells, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_wface, e->h_wface.data(), sizeof(double) * e->n_cells, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_slot, e->h_slot.data(), sizeof(int32_t) * e->n_cells, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_inst, e->inst.data(), sizeof(th::Instance) * n, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_starts, starts.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(e->d_inject, e->inject.data(), sizeof(th::InjectDev) * n, hipMemcpyHostToDevice));
    // a defined pool before the first upload or injection (all particles at rest at the origin)
    for (double *a : {e->d_x, e->d_y, e->d_vx, e->d_vy, e->d_qm})
        TH_HIP(hipMemset(a, 0, sizeof(double) * (size_t)pool));
    TH_HIP(hipMemset(e->d_spc, 0, (size_t)pool));
    TH_HIP(hipMemset(e->d_status, 3, (size_t)pool));
    TH_HIP(hipMemset(e->d_steps, 0, sizeof(int32_t) * (size_t)pool));
    TH_HIP(hipMemset(e->d_hit, 0xff, sizeof(int32_t) * (size_t)pool));
    for (int i = 0; i < n; ++i)
        if ((rc = ensemble_corners(e, i))) return rc;
    TH_HIP(hipStreamSynchronize(e->stream));
    std::vector<double4>().swap(e->h_cells);   // the device holds them now
    std::vector<double>().swap(e->h_wface);
    e->built = true;
    return TH_OK;
}

th::EnsembleGrid ensemble_grid(const th_ensemble *e, const th_trace_params *p) {
    th::EnsembleGrid g;
    g.cells = e->d_cells;
    g.corners = e->d_corners;
    g.corner_vac = e->d_corner_vac;
    g.wface = e->d_wface;
    g.vac = e->d_vac;
    g.slot = e->d_slot;
    g.inst = e->d_inst;
    g.starts = e->d_starts;
    g.n_inst = (int)e->inst.size();
    g.slots = e->slots;
    g.dx = p->dx;
    g.inv_dx = 1.0 / p->dx;   // the host-formed constants of the single-instance push
    g.cfl_dx = p->cfl * p->dx;
    g.two_dx = 2.0 * p->dx;
    return g;
}

int ensemble_launch(th_ensemble *e, const th_trace_params *p, const th::EnsembleGrid &g,
                    const th::EnsembleWork &w, int64_t want) {
    // a cap of at most 4 workgroups (16 waves) per CU takes the 4-wave register budget
    const int budget = p->blocks_per_cu > 0 && p->blocks_per_cu <= 4 ? 1 : 0;
    int per_cu = e->per_cu[budget][p->barrier ? 1 : 0];
    if (p->blocks_per_cu > 0) per_cu = std::min(per_cu, p->blocks_per_cu);
    int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)e->cus * per_cu));
    auto kernel = budget ? (p->barrier ? th::ensemble_push4_kernel<true> : th::ensemble_push4_kernel<false>)
                         : (p->barrier ? th::ensemble_push_kernel<true> : th::ensemble_push_kernel<false>);
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(th::kBlock), 0, e->stream, g, w);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int zero_u64(th_ensemble *e, unsigned long long *ptr, int64_t n) {
    hipLaunchKernelGGL(th::zero_u64_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0,
                       e->stream, ptr, n);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

int zero_u32(th_ensemble *e, uint32_t *ptr, int64_t n) {
    hipLaunchKernelGGL(th::zero_u32_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0,
                       e->stream, ptr, n);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

// The pass schedule over pool[begin, end): caps pass_steps, 2*pass_steps, ...
// up to max_steps, each pass continuing the previous pass's survivors; the
// pass counters must be zero.
int ensemble_passes(th_ensemble *e, const th_trace_params *p, const th::EnsembleGrid &g, th::EnsembleWork w,
                    int64_t begin, int64_t end) {
    if (end <= begin) return TH_OK;
    w.begin = begin;
    w.end = end;
    const int64_t want = (end - begin + th::kBlock - 1) / th::kBlock;
    const bool passes = p->pass_steps > 0 && p->pass_steps < p->max_steps;
    int64_t cap = passes ? p->pass_steps : p->max_steps;
    int rc;
    for (int pass = 0;; ++pass) {
        bool last = !passes || cap >= p->max_steps || pass == kMaxPasses - 1;
        w.step_cap = last ? p->max_steps : (int)cap;
        w.in = pass ? e->d_surv[(pass - 1) & 1] : nullptr;
        w.in_count = pass ? e->d_surv_count + pass - 1 : nullptr;
        w.out = last ? nullptr : e->d_surv[pass & 1];
        w.out_count = last ? nullptr : e->d_surv_count + pass;
        w.queue = e->d_queue + pass;
        if ((rc = ensemble_launch(e, p, g, w, want))) return rc;
        if (last) break;
        cap = std::min<int64_t>(2 * cap, p->max_steps);
    }
    return TH_OK;
}

enum : int { kEnsInject = 1, kEnsStatus = 2 };
enum : int { kEnsOneLaunch = 0, kEnsGraph = 1, kEnsPerInstance = 2 };

// One batch on the stream: device injection (optional), the push passes (one
// launch per pass over all instances, or per instance and pass), batch
// counter + 1.  Only kernels, so the sequence can be captured.
int ensemble_batch(th_ensemble *e, const th_trace_params *p, uint64_t seed, int flags, bool per_instance) {
    const int n = (int)e->inst.size();
    int rc;
    if (flags & kEnsInject) {
        int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((e->n_pool + 255) / 256, (int64_t)e->cus * 16));
        hipLaunchKernelGGL(th::ensemble_inject_kernel, dim3(blocks), dim3(256), 0, e->stream, e->d_inject,
                           e->d_starts, n, seed, e->d_batch, e->d_x, e->d_y, e->d_vx, e->d_vy, e->d_qm, e->d_spc);
        TH_HIP(hipGetLastError());
    }
    th::EnsembleGrid g = ensemble_grid(e, p);
    th::EnsembleWork w;
    w.x = e->d_x; w.y = e->d_y; w.vx = e->d_vx; w.vy = e->d_vy; w.qm = e->d_qm; w.spc = e->d_spc;
    w.status = e->d_status;
    w.steps = e->d_steps;
    w.hit = (flags & kEnsStatus) ? e->d_hit : nullptr;
    w.counts = e->d_counts;
    w.replicas = kCountReplicas;
    w.inst_scal = e->d_scal;
    w.queue = nullptr;
    w.max_steps = p->max_steps;
    w.refill = p->refill > 0 ? std::min(p->refill, 64) : 16;
    w.check_steps = p->check_steps > 0 ? std::min(p->check_steps, 256) : th::kStepsPerCheck;
    w.step_cap = p->max_steps;
    w.in = nullptr;
    w.in_count = nullptr;
    w.out = nullptr;
    w.out_count = nullptr;
    if (per_instance) {
        for (int i = 0; i < n; ++i) {
            if ((rc = zero_u64(e, e->d_queue, kMaxPasses)) || (rc = zero_u32(e, e->d_surv_count, kMaxPasses + 1)) ||
                (rc = ensemble_passes(e, p, g, w, e->inst[i].p0, e->inst[i].p1)))
                return rc;
        }
    } else {
        if ((rc = zero_u64(e, e->d_queue, kMaxPasses)) || (rc = zero_u32(e, e->d_surv_count, kMaxPasses + 1)) ||
            (rc = ensemble_passes(e, p, g, w, 0, e->n_pool)))
            return rc;
    }
    if (e->n_chunks > 0) {
        hipLaunchKernelGGL(th::ensemble_reduce_kernel, dim3((unsigned)std::min(e->n_chunks, e->cus * 8)), dim3(256), 0,
                           e->stream, e->d_chunks, e->n_chunks, e->d_status, e->d_steps, e->d_spc, e->d_scal);
        TH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(th::bump_kernel, dim3(1), dim3(64), 0, e->stream, e->d_batch);
    TH_HIP(hipGetLastError());
    return TH_OK;
}

bool same_params(const th_trace_params &a, const th_trace_params &b) { return memcmp(&a, &b, sizeof a) == 0; }

int ensemble_capture(th_ensemble *e, const th_trace_params *p, uint64_t seed, int flags) {
    if (e->exec && e->graph_flags == flags && e->graph_seed == seed && same_params(e->graph_params, *p))
        return TH_OK;
    ensemble_free_graph(e);
    TH_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = ensemble_batch(e, p, seed, flags, false);
    hipGraph_t graph = nullptr;
    hipError_t end = hipStreamEndCapture(e->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (end != hipSuccess) return fail(TH_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(end));
    e->graph = graph;
    hipError_t inst = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
    if (inst != hipSuccess) {
        ensemble_free_graph(e);
        return fail(TH_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(inst));
    }
    e->graph_params = *p;
    e->graph_seed = seed;
    e->graph_flags = flags;
    return TH_OK;
}

int check_index(const th_ensemble *e, int32_t index) {
    if (!e) return fail(TH_EINVAL, "ensemble is NULL");
    if (index < 0 || index >= (int)e->inst.size()) return fail(TH_EINVAL, "no instance %d", index);
    return TH_OK;
}

}  // namespace

extern "C" {

int th_ensemble_create(int32_t device, th_ensemble **out) {
    if (!out) return fail(TH_EINVAL, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    if (device < 0 || device >= count)
        return fail(TH_EINVAL, "HIP device %d is not available (%d visible)", device, count);
    TH_HIP(hipSetDevice(device));
    th_ensemble *e = new th_ensemble();
    e->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        th_ensemble_destroy(e);
        return fail(TH_EHIP, "hipGetDeviceProperties failed");
    }
    e->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
        th_ensemble_destroy(e);
        return fail(TH_EHIP, "stream / event creation failed");
    }
    int occ[4] = {0, 0, 0, 0};
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[0], th::ensemble_push_kernel<false>, th::kBlock, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[1], th::ensemble_push_kernel<true>, th::kBlock, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[2], th::ensemble_push4_kernel<false>, th::kBlock, 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[3], th::ensemble_push4_kernel<true>, th::kBlock, 0) !=
            hipSuccess) {
        th_ensemble_destroy(e);
        return fail(TH_EHIP, "occupancy query failed");
    }
    for (int k = 0; k < 4; ++k) e->per_cu[k / 2][k % 2] = std::max(occ[k], 1);
    *out = e;
    return TH_OK;
}

void th_ensemble_destroy(th_ensemble *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ensemble_free_graph(e);
    ensemble_free_device(e);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int th_ensemble_add(th_ensemble *e, const uint8_t *vac, int32_t nx, int32_t ny, const double *Ex, const double *Ey,
                    const double *phi, const double *wface, int32_t tapered, int64_t n_particles, int32_t *index) {
    if (!e || !vac || !Ex || !Ey || !index) return fail(TH_EINVAL, "invalid argument");
    if (e->built) return fail(TH_ESTATE, "instances are added before the first upload, injection or run");
    if ((phi == nullptr) != (wface == nullptr)) return fail(TH_EINVAL, "phi and wface go together");
    if (nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 26)) return fail(TH_EINVAL, "invalid grid %dx%d", nx, ny);
    if (n_particles < 0) return fail(TH_EINVAL, "negative particle count");
    const int64_t N = (int64_t)nx * ny, keys = (int64_t)(nx + 1) * (ny + 1);
    if (e->n_cells + N > INT32_MAX || e->n_keys + keys > INT32_MAX / 4)
        return fail(TH_EINVAL, "the ensemble's grids exceed 2^31 cells");
    if (e->n_pool + n_particles > INT32_MAX) return fail(TH_EINVAL, "at most 2^31-1 particles per ensemble batch");
    if (e->slots + N > (1 << 30)) return fail(TH_EINVAL, "the ensemble's surface cells exceed 2^30");
    th::Instance d;
    d.nx = nx;
    d.ny = ny;
    d.cell0 = (int32_t)e->n_cells;
    d.key0 = (int32_t)e->n_keys;
    d.tapered = tapered ? 1 : 0;
    d.pad = 0;
    d.p0 = e->n_pool;
    d.p1 = e->n_pool + n_particles;
    std::vector<uint8_t> v(vac, vac + N);
    for (auto &c : v) c = c ? 1 : 0;
    std::vector<int32_t> surf;
    surface_cells(v, nx, ny, surf);
    const int i = (int)e->inst.size();
    e->inst.push_back(d);
    e->slot0.push_back(e->slots);
    e->h_vac.insert(e->h_vac.end(), v.begin(), v.end());
    e->h_slot.resize((size_t)(e->n_cells + N), -1);
    for (size_t k = 0; k < surf.size(); ++k) e->h_slot[(size_t)e->n_cells + surf[k]] = e->slots + (int32_t)k;
    e->slots += (int)surf.size();
    e->surf.push_back(std::move(surf));
    e->h_cells.resize((size_t)(e->n_cells + N));
    e->h_wface.resize((size_t)(e->n_cells + N));
    pack_host_cells(e, i, Ex, Ey, phi, wface, e->h_cells.data() + e->n_cells, e->h_wface.data() + e->n_cells);
    e->has_phi.push_back(phi != nullptr);
    th::InjectDev none;
    memset(&none, 0, sizeof none);
    e->inject.push_back(none);
    e->has_inject.push_back(0);
    e->n_cells += N;
    e->n_keys += keys;
    e->n_pool += n_particles;
    *index = i;
    return TH_OK;
}

int th_ensemble_set_fields(th_ensemble *e, int32_t index, const double *Ex, const double *Ey, const double *phi,
                           const double *wface) {
    int rc = check_index(e, index);
    if (rc) return rc;
    if (!Ex || !Ey) return fail(TH_EINVAL, "Ex and Ey are required");
    if ((phi == nullptr) != (wface == nullptr)) return fail(TH_EINVAL, "phi and wface go together");
    const th::Instance &d = e->inst[index];
    const int64_t N = (int64_t)d.nx * d.ny;
    e->has_phi[index] = phi != nullptr;
    if (!e->built) {
        pack_host_cells(e, index, Ex, Ey, phi, wface, e->h_cells.data() + d.cell0, e->h_wface.data() + d.cell0);
        return TH_OK;
    }
    TH_HIP(hipSetDevice(e->device));
    std::vector<double4> cells((size_t)N);
    std::vector<double> wf((size_t)N);
    pack_host_cells(e, index, Ex, Ey, phi, wface, cells.data(), wf.data());
    TH_HIP(hipMemcpyAsync(e->d_cells + d.cell0, cells.data(), sizeof(double4) * N, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_wface + d.cell0, wf.data(), sizeof(double) * N, hipMemcpyHostToDevice, e->stream));
    if ((rc = ensemble_corners(e, index))) return rc;
    TH_HIP(hipStreamSynchronize(e->stream));
    return TH_OK;
}

int th_ensemble_surface(const th_ensemble *e, int32_t index, int32_t *m, int32_t *flat) {
    int rc = check_index(e, index);
    if (rc) return rc;
    if (!m) return fail(TH_EINVAL, "m is NULL");
    *m = (int32_t)e->surf[index].size();
    if (flat) memcpy(flat, e->surf[index].data(), sizeof(int32_t) * e->surf[index].size());
    return TH_OK;
}

int th_ensemble_set_inject(th_ensemble *e, int32_t index, const th_inject_params *ip) {
    int rc = check_index(e, index);
    if (rc) return rc;
    if (!ip) return fail(TH_EINVAL, "ip is NULL");
    if (ip->ion_kind < 0 || ip->ion_kind > 2) return fail(TH_EINVAL, "unknown ion distribution %d", ip->ion_kind);
    const th::Instance &d = e->inst[index];
    if (2 * ip->n != d.p1 - d.p0)
        return fail(TH_EINVAL, "instance %d holds %lld particles, not 2 x %lld", index, (long long)(d.p1 - d.p0),
                    (long long)ip->n);
    th::InjectDev &q = e->inject[index];
    q.n = ip->n; q.lx = ip->lx; q.y0 = ip->y0; q.qm_i = ip->qm_i; q.qm_e = ip->qm_e; q.ion_kind = ip->ion_kind;
    q.e_ion = ip->e_ion; q.e_lo = ip->e_lo; q.e_hi = ip->e_hi; q.sigma_theta = ip->sigma_theta;
    q.vth_i = ip->vth_i; q.v_per_sqrt_ev = ip->v_per_sqrt_ev; q.s_e = ip->s_e;
    q.seed = 0; q.batch = 0;
    q.x = q.y = q.vx = q.vy = q.qm = nullptr;
    q.spc = nullptr;
    e->has_inject[index] = 1;
    if (e->built) {
        TH_HIP(hipSetDevice(e->device));
        TH_HIP(hipMemcpy(e->d_inject + index, &q, sizeof q, hipMemcpyHostToDevice));
    }
    return TH_OK;
}

int th_ensemble_upload(th_ensemble *e, int32_t index, const double *x, const double *y, const double *vx,
                       const double *vy, const double *qm, const int8_t *spc, int64_t n) {
    int rc = check_index(e, index);
    if (rc) return rc;
    const th::Instance &d = e->inst[index];
    if (n != d.p1 - d.p0) return fail(TH_EINVAL, "instance %d holds %lld particles", index, (long long)(d.p1 - d.p0));
    if (n > 0 && (!x || !y || !vx || !vy || !qm || !spc)) return fail(TH_EINVAL, "particle array is NULL");
    for (int64_t k = 0; k < n; ++k)
        if (spc[k] != 0 && spc[k] != 1) return fail(TH_EINVAL, "species must be 0 (ion) or 1 (electron)");
    TH_HIP(hipSetDevice(e->device));
    if ((rc = ensemble_build(e))) return rc;
    if (n == 0) return TH_OK;
    size_t b8 = sizeof(double) * (size_t)n;
    TH_HIP(hipMemcpyAsync(e->d_x + d.p0, x, b8, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_y + d.p0, y, b8, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_vx + d.p0, vx, b8, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_vy + d.p0, vy, b8, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_qm + d.p0, qm, b8, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_spc + d.p0, spc, (size_t)n, hipMemcpyHostToDevice, e->stream));
    TH_HIP(hipStreamSynchronize(e->stream));
    return TH_OK;
}

int th_ensemble_run(th_ensemble *e, const th_trace_params *p, uint64_t seed, uint64_t first_batch, int32_t batches,
                    int32_t mode, int32_t flags) {
    if (!e || !p) return fail(TH_EINVAL, "invalid argument");
    if (batches < 1) return fail(TH_EINVAL, "batches must be >= 1");
    if (mode < kEnsOneLaunch || mode > kEnsPerInstance) return fail(TH_EINVAL, "unknown mode %d", mode);
    if (flags & ~(kEnsInject | kEnsStatus)) return fail(TH_EINVAL, "unknown flags 0x%x", flags);
    if (mode == kEnsGraph && (flags & kEnsStatus)) return fail(TH_EINVAL, "per-particle status is not captured");
    if (!(p->cfl > 0.0 && p->cfl <= 1.0) || p->max_steps < 1 || !(p->dx > 0.0))
        return fail(TH_EINVAL, "tracing needs dx > 0, 0 < cfl <= 1 and max_steps >= 1");
    if (p->precision != 0 || p->rf) return fail(TH_EINVAL, "the ensemble push is FP64 without RF");
    if (p->blocks_per_cu < 0) return fail(TH_EINVAL, "blocks_per_cu must be >= 0");
    if (e->inst.empty()) return fail(TH_ESTATE, "the ensemble has no instances");
    for (size_t i = 0; i < e->inst.size(); ++i) {
        if (p->barrier && !e->has_phi[i]) return fail(TH_ESTATE, "barrier wall model: instance %zu lacks phi", i);
        if ((flags & kEnsInject) && !e->has_inject[i] && e->inst[i].p1 > e->inst[i].p0)
            return fail(TH_ESTATE, "device injection: instance %zu has no distributions", i);
    }
    TH_HIP(hipSetDevice(e->device));
    int rc = ensemble_build(e);
    if (rc) return rc;
    if (mode == kEnsGraph && (rc = ensemble_capture(e, p, seed, flags))) return rc;
    const int n = (int)e->inst.size();
    TH_HIP(hipMemsetAsync(e->d_counts, 0, sizeof(uint32_t) * kCountReplicas * 2 * (size_t)std::max(e->slots, 1),
                          e->stream));
    TH_HIP(hipMemsetAsync(e->d_scal, 0, sizeof(unsigned long long) * n * th::kInstCount, e->stream));
    TH_HIP(hipMemcpyAsync(e->d_batch, &first_batch, sizeof first_batch, hipMemcpyHostToDevice, e->stream));
    if (flags & kEnsStatus)
        TH_HIP(hipMemsetAsync(e->d_hit, 0xff, sizeof(int32_t) * (size_t)std::max<int64_t>(e->n_pool, 1), e->stream));
    TH_HIP(hipStreamSynchronize(e->stream));   // first_batch lives on this stack frame
    TH_HIP(hipEventRecord(e->ev0, e->stream));
    for (int b = 0; b < batches; ++b) {
        if (mode == kEnsGraph) {
            TH_HIP(hipGraphLaunch(e->exec, e->stream));
        } else if ((rc = ensemble_batch(e, p, seed, flags, mode == kEnsPerInstance))) {
            return rc;
        }
    }
    TH_HIP(hipEventRecord(e->ev1, e->stream));
    TH_HIP(hipStreamSynchronize(e->stream));
    float ms = 0.0f;
    TH_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_ms = ms;
    return TH_OK;
}

int th_ensemble_read(th_ensemble *e, int32_t index, int64_t *cnt_i, int64_t *cnt_e, th_tally *tally, int8_t *status,
                     int32_t *hit) {
    int rc = check_index(e, index);
    if (rc) return rc;
    if ((cnt_i == nullptr) != (cnt_e == nullptr)) return fail(TH_EINVAL, "cnt_i and cnt_e go together");
    if (!e->built) return fail(TH_ESTATE, "nothing traced yet");
    TH_HIP(hipSetDevice(e->device));
    const th::Instance &d = e->inst[index];
    const std::vector<int32_t> &surf = e->surf[index];
    const int m = (int)surf.size();
    const int64_t N = (int64_t)d.nx * d.ny, np = d.p1 - d.p0;
    if (cnt_i) {
        memset(cnt_i, 0, sizeof(int64_t) * (size_t)N);
        memset(cnt_e, 0, sizeof(int64_t) * (size_t)N);
        std::vector<uint32_t> c((size_t)std::max(m, 1));
        for (int r = 0; r < kCountReplicas && m; ++r)
            for (int sp = 0; sp < 2; ++sp) {
                const uint32_t *src = e->d_counts + ((size_t)r * 2 + sp) * e->slots + e->slot0[index];
                TH_HIP(hipMemcpy(c.data(), src, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
                int64_t *dst = sp ? cnt_e : cnt_i;
                for (int k = 0; k < m; ++k) dst[surf[k]] += c[k];
            }
    }
    if (tally) {
        unsigned long long s[th::kInstCount];
        TH_HIP(hipMemcpy(s, e->d_scal + (size_t)index * th::kInstCount, sizeof s, hipMemcpyDeviceToHost));
        memset(tally, 0, sizeof *tally);
        tally->escaped[0] = (int64_t)s[th::kInstEsc0];
        tally->escaped[1] = (int64_t)s[th::kInstEsc1];
        tally->lost = (int64_t)s[th::kInstLost];
        tally->steps = (int64_t)s[th::kInstSteps];
        tally->off_surface = (int64_t)s[th::kInstOff];
    }
    if (status && np) {
        TH_HIP(hipMemcpy(status, e->d_status + d.p0, (size_t)np, hipMemcpyDeviceToHost));
        for (int64_t q = 0; q < np; ++q)
            if (status[q] == th::kEndOffSurface) status[q] = th::kEndAbsorbed;
    }
    if (hit && np) TH_HIP(hipMemcpy(hit, e->d_hit + d.p0, sizeof(int32_t) * (size_t)np, hipMemcpyDeviceToHost));
    return TH_OK;
}

int th_ensemble_totals(th_ensemble *e, int64_t *counts, int64_t *scal) {
    if (!e) return fail(TH_EINVAL, "ensemble is NULL");
    if (!e->built) return fail(TH_ESTATE, "nothing traced yet");
    TH_HIP(hipSetDevice(e->device));
    const size_t per = 2 * (size_t)e->slots;
    if (counts && per) {
        std::vector<uint32_t> c(kCountReplicas * per);
        TH_HIP(hipMemcpy(c.data(), e->d_counts, sizeof(uint32_t) * c.size(), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < per; ++j) {
            int64_t v = 0;
            for (int r = 0; r < kCountReplicas; ++r) v += c[r * per + j];
            counts[j] = v;
        }
    }
    if (scal) {
        std::vector<unsigned long long> t(e->inst.size() * th::kInstCount);
        TH_HIP(hipMemcpy(t.data(), e->d_scal, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < t.size(); ++j) scal[j] = (int64_t)t[j];
    }
    return TH_OK;
}

int th_ensemble_read_particles(th_ensemble *e, int32_t index, double *x, double *y, double *vx, double *vy,
                               double *qm, int8_t *spc) {
    int rc = check_index(e, index);
    if (rc) return rc;
    if (!e->built) return fail(TH_ESTATE, "nothing uploaded or injected yet");
    const th::Instance &d = e->inst[index];
    const int64_t n = d.p1 - d.p0;
    if (n == 0) return TH_OK;
    if (!x || !y || !vx || !vy || !qm || !spc) return fail(TH_EINVAL, "particle array is NULL");
    TH_HIP(hipSetDevice(e->device));
    size_t b8 = sizeof(double) * (size_t)n;
    TH_HIP(hipMemcpy(x, e->d_x + d.p0, b8, hipMemcpyDeviceToHost));
    TH_HIP(hipMemcpy(y, e->d_y + d.p0, b8, hipMemcpyDeviceToHost));
    TH_HIP(hipMemcpy(vx, e->d_vx + d.p0, b8, hipMemcpyDeviceToHost));
    TH_HIP(hipMemcpy(vy, e->d_vy + d.p0, b8, hipMemcpyDeviceToHost));
    TH_HIP(hipMemcpy(qm, e->d_qm + d.p0, b8, hipMemcpyDeviceToHost));
    TH_HIP(hipMemcpy(spc, e->d_spc + d.p0, (size_t)n, hipMemcpyDeviceToHost));
    return TH_OK;
}

int th_ensemble_last_ms(const th_ensemble *e, double *ms) {
    if (!e || !ms) return fail(TH_EINVAL, "invalid argument");
    *ms = e->last_ms;
    return TH_OK;
}

}  // extern "C"
