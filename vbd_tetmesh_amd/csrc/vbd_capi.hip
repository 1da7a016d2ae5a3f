// C ABI of the VBD solver (declared in include/vbd_capi.h): handle lifecycle, upload,
// VBD_GPU_Solver-style solves, the device time step captured in a hipGraph, launch
// configuration / autotuning / timing, and the host-side mesh preparation helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "vbd_capi.h"
#include "vbd_internal.h"

using namespace vbd;

struct vbd_solver {
    int device = 0;
    hipStream_t stream = nullptr;     // the stream all work is enqueued on
    hipStream_t ownStream = nullptr;  // the handle's own non-blocking stream
    std::string err;
    bool inited = false;
    DevMesh m;
    vbd_step_params params{};
    bool paramsSet = false;
    int lpv = 8;
    int unroll = 1;
    int block = 256;
    int materialModel = VBD_MATERIAL_STABLE_NEO_HOOKEAN;
    bool graphEnabled = true;
    // graphs of vbd_solve calls and of the converged-step / adjoint-solve iteration blocks, keyed
    // by their content; at most VBD_MAX_CACHED_GRAPHS, least recently used evicted
    struct BlockGraph {
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::map<std::string, BlockGraph> blockGraphs;
    std::map<std::string, uint64_t> blockSeen;  // keys run once without a graph (captured on reuse)
    uint64_t blockClock = 0;
    double* h_res = nullptr;  // pinned host copy of a block's residual
    hipGraphExec_t graphExec = nullptr;
    bool graphValid = false;
    double* d_energy = nullptr;   // 2 doubles: energy results
    double* d_epart = nullptr;    // per-block energy partial sums
    int energyEvery = 0;          // vbd_step's energy log: every n-th iteration (0: off)
    double* d_elog = nullptr;     // the log of the last step: substeps * ceil(iterations / n)
    int elogCap = 0;
    std::vector<int32_t> meshVertOffsets;  // batched uploads: per-mesh vertex offsets
    // differentiation scratch, allocated on first use (slot order, nV entries)
    f4* d_w = nullptr;
    f4* d_rhs = nullptr;
    f4* d_out = nullptr;
    f4* d_wsave[2] = {nullptr, nullptr};
    double* d_r2 = nullptr;
    double* d_sum = nullptr;   // 2 doubles
    float* d_tgrad = nullptr;  // 2 * nT
    float* d_x3 = nullptr;     // 3 * nV staging
    float* h_x3 = nullptr;     // 3 * nV pinned host staging of large transfers
    hipEvent_t xev[16] = {};   // one per chunk of a staged device-to-host copy
    // FP64 differentiation state (vbd_set_differentiation_precision(h, 64)); fp64Fresh: it holds
    // the current state (the FP32 arrays are its rounding); cleared by every entry point that
    // changes the FP32 state, after which the next FP64 call promotes the FP32 state
    int persistent = 0;          // 1 or a grid cap (8k): one persistent launch per sweep (tiled stable Neo-Hookean)
    unsigned* d_bar = nullptr;   // its grid barrier (PERSIST_BAR_WORDS words: counts, generation, timeout flag)
    int diffBits = 32;
    int convergedMethod = VBD_CONVERGED_VBD;  // vbd_set_converged_method
    int lastNewton = 0, lastInner = 0, lastSweeps = 0;  // the last converged step's work (vbd_last_step_stats)
    Dev64 d64;
    bool fp64Fresh = false;
    double* d_x3d = nullptr;   // 3 * nV staging of FP64 state transfers
    double* d_tgrad64 = nullptr;  // 2 * nT FP64 parameter gradients
};

// Scratch of vbd_total_energy_noalloc: the per-block partial sums of the energy reduction (and
// the second stage of its fixed-order sum), one copy per GPU, never allocated or freed.  Host
// code uses the address looked up per device (scratch_ptr), never the symbol itself.
static __device__ __align__(256) double scratch[VBD_NOALLOC_SCRATCH_DOUBLES];

namespace {

thread_local std::string g_err;

// Device address of `scratch` on each GPU, looked up once per device by vbd_create.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// With `device` current: looks up its copy of `scratch` unless that was done already.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    double* p = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

double* scratch_ptr(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

int fail(vbd_solver* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    g_err = msg;
    return code;
}

int hip_fail(vbd_solver* h, const char* what, hipError_t e) {
    return fail(h, VBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_CHECK(h, expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail((h), #expr, _e); \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

void drop_graph(vbd_solver* h) {
    // a replay enqueued by an earlier asynchronous call may still be running
    if ((h->graphExec || !h->blockGraphs.empty()) && h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->graphExec) (void)hipGraphExecDestroy(h->graphExec);
    h->graphExec = nullptr;
    h->graphValid = false;
    for (auto& kv : h->blockGraphs) (void)hipGraphExecDestroy(kv.second.exec);
    h->blockGraphs.clear();
    h->blockSeen.clear();
}

// Keys remembered as run once: a key seen again within this many distinct keys is captured.
constexpr size_t BLOCK_SEEN_CAP = 4 * VBD_MAX_CACHED_GRAPHS;

template <class M>
void evict_oldest(M& map) {
    auto old = map.begin();
    for (auto it = map.begin(); it != map.end(); ++it)
        if (it->second < old->second) old = it;
    map.erase(old);
}

// Runs the work `enqueue` puts on the handle's stream, then waits (if `wait`).  With graphs
// enabled, a key's first use launches the work directly and a later use captures it once and
// replays the graph; the same kernels run either way, so results are identical.  A caller that
// changes the baked-in scalars on every call therefore never pays a capture, and the cache keeps
// at most VBD_MAX_CACHED_GRAPHS graphs (least recently used evicted).  The key must name
// everything the work bakes into arguments that can change between calls (the buffers are fixed
// per handle; set_params, launch configuration, material changes and re-initialisation drop the
// cache).
template <class Fn>
hipError_t run_block(vbd_solver* h, const std::string& key, Fn&& enqueue, bool wait = true) {
    hipError_t e = hipSuccess;
    const uint64_t now = ++h->blockClock;
    auto it = h->graphEnabled ? h->blockGraphs.find(key) : h->blockGraphs.end();
    if (it != h->blockGraphs.end()) {
        it->second.used = now;
        e = hipGraphLaunch(it->second.exec, h->stream);
    } else if (!h->graphEnabled || h->blockSeen.find(key) == h->blockSeen.end()) {
        if (h->graphEnabled) {
            if (h->blockSeen.size() >= BLOCK_SEEN_CAP) evict_oldest(h->blockSeen);
            h->blockSeen.emplace(key, now);
        }
        e = enqueue();
    } else {
        h->blockSeen.erase(key);
        if (h->blockGraphs.size() >= (size_t)VBD_MAX_CACHED_GRAPHS) {
            auto old = h->blockGraphs.begin();
            for (auto jt = h->blockGraphs.begin(); jt != h->blockGraphs.end(); ++jt)
                if (jt->second.used < old->second.used) old = jt;
            // a replay of the evicted graph may still be in flight
            e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) return e;
            (void)hipGraphExecDestroy(old->second.exec);
            h->blockGraphs.erase(old);
        }
        hipGraph_t graph = nullptr;
        e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        e = enqueue();
        hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
        if (e == hipSuccess) e = e2;
        hipGraphExec_t exec = nullptr;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return e;
        h->blockGraphs.emplace(key, vbd_solver::BlockGraph{exec, now});
        e = hipGraphLaunch(exec, h->stream);
    }
    if (e == hipSuccess && wait) e = hipStreamSynchronize(h->stream);
    return e;
}

// Key part for a block of n iterations starting at iteration it0 with the given omegas.
std::string block_key(char kind, int it0, const std::vector<float>& omegas, float extra) {
    std::string k(1, kind);
    char buf[32];
    snprintf(buf, sizeof(buf), "%d:%d:%08x:", (int)omegas.size(), it0 & 1, (unsigned)__builtin_bit_cast(uint32_t, extra));
    k += buf;
    for (float w : omegas) {
        snprintf(buf, sizeof(buf), "%08x", (unsigned)__builtin_bit_cast(uint32_t, w));
        k += buf;
    }
    return k;
}

void free_scratch(vbd_solver* h) {
    drop_graph(h);  // block graphs reference the scratch buffers
    if (h->h_res) (void)hipHostFree(h->h_res);
    h->h_res = nullptr;
    if (h->h_x3) (void)hipHostFree(h->h_x3);
    h->h_x3 = nullptr;
    for (hipEvent_t& ev : h->xev) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    void* ps[] = {h->d_w,  h->d_rhs, h->d_out,   h->d_wsave[0], h->d_wsave[1],
                  h->d_r2, h->d_sum, h->d_tgrad, h->d_x3,       h->d_epart};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    h->d_w = h->d_rhs = h->d_out = h->d_wsave[0] = h->d_wsave[1] = nullptr;
    h->d_r2 = h->d_sum = h->d_epart = nullptr;
    h->d_tgrad = h->d_x3 = nullptr;
    free_dev64(h->d64);
    if (h->d_x3d) (void)hipFree(h->d_x3d);
    if (h->d_tgrad64) (void)hipFree(h->d_tgrad64);
    h->d_x3d = nullptr;
    h->d_tgrad64 = nullptr;
    h->fp64Fresh = false;
}

// VBDsolver::getAcceleratorOmega (VBD.cpp:302-311): float arguments, double literals, so the
// quotients are formed in double and rounded to float on return.
float omega_ref(int order, float rho, float prev) {
    switch (order) {
        case 1: return 1.0f;
        case 2: return (float)(2.0 / (2.0 - (double)(rho * rho)));
        default: return (float)(4.0 / (4.0 - (double)((rho * rho) * prev)));
    }
}

// 1 / h^2 as VBDsolver::solve forms it (VBD.cpp:116-117)
float inv_h2_ref(float h) { return (float)(1.0 / (double)(h * h)); }

ContactArgs contact_of(const vbd_step_params& p) {
    ContactArgs c;
    c.on = p.contact_on;
    c.axis = p.contact_axis;
    c.k = p.contact_k;
    c.plane = p.contact_plane;
    return c;
}

bool finite_array(const float* a, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

// One color launch with the handle's configuration: lpv > 0 selects the per-vertex lane-group
// kernel, lpv == 0 the tiled kernel (unroll = records per thread; compact records for the
// stable Neo-Hookean material, full records for the corotated one).
hipError_t launch_color_cfg(vbd_solver* h, int c, float sdt, float mu, float lambda, const ContactArgs& ct, f4* save) {
    const bool corot = h->materialModel == VBD_MATERIAL_COROTATED_GPU;
    if (h->lpv == 0)
        return launch_color_tiled(h->m, c, h->unroll, h->block, sdt, mu, lambda, ct, save, h->stream, corot);
    if (corot) {
        // per-vertex lane groups on the full records; the configured lanes when 8/16/32
        const bool own = h->lpv == 8 || h->lpv == 16 || h->lpv == 32;
        return launch_color_corotated(h->m, c, own ? h->lpv : 16, own ? h->block : 256, sdt, mu, lambda, ct, save,
                                      h->stream);
    }
    return launch_color(h->m, c, h->lpv, h->unroll, h->block, sdt, mu, lambda, ct, save, h->stream);
}

// One Gauss-Seidel sweep over all colors.
hipError_t enqueue_sweep(vbd_solver* h, float sdt, float mu, float lambda, const ContactArgs& ct, f4* save) {
    for (int c = 0; c < h->m.nColors; ++c) {
        hipError_t e = launch_color_cfg(h, c, sdt, mu, lambda, ct, save);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Polls of a persistent sweep's grid barrier before it gives up (each poll sleeps about 50 ns
// and reads across the device): half a second at the least, far above any legitimate wait (a
// color's slowest block), and a bound on a launch whose blocks are not all resident.
constexpr unsigned PERSIST_SPIN = 1u << 23;

bool persistent_on(const vbd_solver* h) {
    return h->persistent && h->lpv == 0 && h->materialModel == VBD_MATERIAL_STABLE_NEO_HOOKEAN && h->d_bar;
}

// Grid of the persistent sweep, a multiple of 8 and at most the largest color's tile count:
// mode 1 takes 2 blocks per CU (the fastest grid measured on the sphere, the batch and the cube;
// a grid of the occupancy API's full count, 4 blocks of 512 threads per CU for this kernel, was
// seen not to become co-resident on MI355X, its barrier timing out, where 3 per CU was); a cap
// 8k takes at most 8k blocks and at most the occupancy API's count.  0 when not even 8 fit.
int persistent_grid(const vbd_solver* h, const ContactArgs& ct, bool save_on) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess) return 0;
    const int occ = persistent_occupancy(h->m, h->unroll, h->block, ct, save_on);
    int tiles = 0;
    for (int c : h->m.colorTileCount) tiles = std::max(tiles, c);
    int g = std::min(occ * cus, (tiles + 7) & ~7);
    g = std::min(g, h->persistent >= 8 ? h->persistent : 2 * cus);
    g &= ~7;
    return g >= 8 ? g : 0;
}

// Iteration loop of VBDsolver::step (VBD.cpp:75-96) / VBD_GPU_Solver::solveVBD (.cu:338-361):
// save[it & 1] holds each vertex's start-of-sweep position; after the sweep, omega > 1 blends
// against save[(it + 1) & 1], the start of the previous sweep, which is the prevPrevPos of the
// reference (the two buffers alternate instead of being copied).  The blend also writes its
// result into save[(it + 1) & 1], the next sweep's start, so only the sweeps that no blend
// precedes (the first two) save from their color launches.
// With `elog`, the diagnostic energy (VBD.cpp:92-94) after iterations 0, every, 2 every, ...
// goes to elog[it / every] (mu, lambda: the energy's material, the step parameters).
hipError_t enqueue_iterations(vbd_solver* h, int iters, float sdt, float mu, float lambda, bool accel, float rho,
                              const ContactArgs& ct, double* elog = nullptr, int every = 0) {
    float omega = 1.0f;
    const int pgrid = persistent_on(h) ? persistent_grid(h, ct, accel) : 0;
    bool kept = false;  // the previous blend wrote this sweep's start into save[it & 1]
    for (int it = 0; it < iters; ++it) {
        f4* save = accel && !kept ? h->m.save[it & 1] : nullptr;
        if (pgrid > 0) {  // one launch: every color, then the blend
            if (accel) omega = omega_ref(it + 1, rho, omega);
            hipError_t e = launch_sweep_persistent(h->m, h->unroll, h->block, pgrid, sdt, mu, lambda, ct, save,
                                                   h->m.save[(it + 1) & 1], accel ? omega : 1.0f, h->d_bar,
                                                   PERSIST_SPIN, h->stream);
            if (e == hipSuccess && elog && every > 0 && it % every == 0)
                e = launch_energy(h->m, 0, mu, lambda, h->d_epart, elog + it / every, h->stream);
            if (e != hipSuccess) return e;
            continue;
        }
        hipError_t e = enqueue_sweep(h, sdt, mu, lambda, ct, save);
        if (e != hipSuccess) return e;
        if (accel) {
            omega = omega_ref(it + 1, rho, omega);
            kept = omega > 1.0f;
            if (kept) {
                e = launch_blend(h->m, omega, h->m.save[(it + 1) & 1], h->stream, true);
                if (e != hipSuccess) return e;
            }
        }
        if (elog && every > 0 && it % every == 0) {
            e = launch_energy(h->m, 0, mu, lambda, h->d_epart, elog + it / every, h->stream);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

int energy_log_slots(const vbd_solver* h) {
    if (h->energyEvery <= 0 || !h->paramsSet) return 0;
    return h->params.substeps * ((h->params.num_iterations + h->energyEvery - 1) / h->energyEvery);
}

// One VBDsolver::step (VBD.cpp:60-101): substeps of forwardStep, iterations, updateVelocity.
hipError_t enqueue_step(vbd_solver* h) {
    const vbd_step_params& p = h->params;
    const float hstep = p.dt / (float)p.substeps;
    const float sdt = inv_h2_ref(hstep);
    const ContactArgs ct = contact_of(p);
    for (int sub = 0; sub < p.substeps; ++sub) {
        hipError_t e = launch_predict(h->m, hstep, p.gravity, h->stream);
        if (e != hipSuccess) return e;
        if (p.start_from_inertia) {
            e = launch_copy_inertia(h->m, h->stream);
            if (e != hipSuccess) return e;
        }
        const int per = h->energyEvery > 0 ? (p.num_iterations + h->energyEvery - 1) / h->energyEvery : 0;
        e = enqueue_iterations(h, p.num_iterations, sdt, p.mu, p.lambda, p.use_acceleration != 0, p.acceleration_rho,
                               ct, h->d_elog ? h->d_elog + (size_t)sub * per : nullptr, h->energyEvery);
        if (e != hipSuccess) return e;
        e = launch_velocity(h->m, hstep, p.velocity_damping, h->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int check_ready(vbd_solver* h) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (!h->inited) return fail(h, VBD_ERR_STATE, "vbd_init has not been called on this handle");
    return VBD_OK;
}

// The 3 * nV float staging buffer of the caller-order transfers, kept for the handle's life
// (a per-call hipMalloc / hipFree pair would also synchronise the device on every frame).
hipError_t ensure_stage3(vbd_solver* h) {
    if (h->d_x3) return hipSuccess;
    return hipMalloc(&h->d_x3, sizeof(float) * 3 * (size_t)(h->m.nV > 0 ? h->m.nV : 1));
}

// Large transfers of caller (pageable) xyz arrays go through pinned host staging in chunks:
// the host copies one chunk while the DMA engine moves another, instead of the runtime's
// pageable path (device to host: 1.42 ms for the 128^3 cube's 25.8 MB, against 0.46 ms of DMA
// from pinned memory and 0.5 ms of host copy).  Small ones take one runtime copy.
constexpr size_t XFER_STAGED_MIN = (size_t)1 << 20;

hipError_t ensure_host_stage(vbd_solver* h) {
    hipError_t e = hipSuccess;
    if (!h->h_x3) e = hipHostMalloc(&h->h_x3, sizeof(float) * 3 * (size_t)(h->m.nV > 0 ? h->m.nV : 1), hipHostMallocDefault);
    for (hipEvent_t& ev : h->xev)
        if (e == hipSuccess && !ev) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    return e;
}

size_t xfer_chunk(size_t bytes) {
    const size_t n = sizeof(((vbd_solver*)nullptr)->xev) / sizeof(hipEvent_t);
    return ((bytes + n - 1) / n + 4095) & ~(size_t)4095;
}

// host xyz (caller order) -> d_x3, enqueued on the stream
hipError_t h2d_xyz(vbd_solver* h, const float* src) {
    const size_t bytes = sizeof(float) * 3 * (size_t)h->m.nV;
    if (bytes < XFER_STAGED_MIN) return hipMemcpyAsync(h->d_x3, src, bytes, hipMemcpyHostToDevice, h->stream);
    hipError_t e = ensure_host_stage(h);
    // the staging buffer may still feed an earlier copy of this handle
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    const size_t chunk = xfer_chunk(bytes);
    for (size_t off = 0; e == hipSuccess && off < bytes; off += chunk) {
        const size_t len = std::min(chunk, bytes - off);
        std::memcpy((char*)h->h_x3 + off, (const char*)src + off, len);
        e = hipMemcpyAsync((char*)h->d_x3 + off, (const char*)h->h_x3 + off, len, hipMemcpyHostToDevice, h->stream);
    }
    return e;
}

// d_x3 -> host xyz (caller order); returns when the host array holds it
hipError_t d2h_xyz(vbd_solver* h, float* dst) {
    const size_t bytes = sizeof(float) * 3 * (size_t)h->m.nV;
    hipError_t e = hipSuccess;
    if (bytes < XFER_STAGED_MIN) {
        e = hipMemcpyAsync(dst, h->d_x3, bytes, hipMemcpyDeviceToHost, h->stream);
        return e == hipSuccess ? hipStreamSynchronize(h->stream) : e;
    }
    e = ensure_host_stage(h);
    const size_t chunk = xfer_chunk(bytes);
    int k = 0;
    for (size_t off = 0; e == hipSuccess && off < bytes; off += chunk, ++k) {
        e = hipMemcpyAsync((char*)h->h_x3 + off, (const char*)h->d_x3 + off, std::min(chunk, bytes - off),
                           hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipEventRecord(h->xev[k], h->stream);
    }
    k = 0;
    for (size_t off = 0; e == hipSuccess && off < bytes; off += chunk, ++k) {
        e = hipEventSynchronize(h->xev[k]);
        if (e == hipSuccess) std::memcpy((char*)dst + off, (const char*)h->h_x3 + off, std::min(chunk, bytes - off));
    }
    return e;
}

int upload_vertex_array(vbd_solver* h, const float* src, f4* dst) {
    const size_t n = (size_t)h->m.nV;
    HIP_CHECK(h, ensure_stage3(h));
    hipError_t e = h2d_xyz(h, src);
    if (e == hipSuccess) e = launch_scatter_pos((int)n, h->d_x3, h->m.inv, dst, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_fail(h, "upload", e);
    return VBD_OK;
}

// A persistent sweep whose grid barrier timed out leaves its flag raised: reported (and the
// barrier reset) at the next synchronising call, since the positions it left are not a sweep's.
int check_persistent(vbd_solver* h) {
    if (!h->d_bar) return VBD_OK;
    unsigned flag = 0;
    HIP_CHECK(h, hipMemcpyAsync(&flag, h->d_bar + PERSIST_FLAG, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    if (!flag) return VBD_OK;
    HIP_CHECK(h, hipMemsetAsync(h->d_bar, 0, PERSIST_BAR_WORDS * sizeof(unsigned), h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return fail(h, VBD_ERR_HIP, "persistent sweep: a grid barrier timed out (blocks not co-resident); the positions "
                                "are not valid");
}

int download_vertex_array(vbd_solver* h, const f4* src, float* dst) {
    const size_t n = (size_t)h->m.nV;
    HIP_CHECK(h, ensure_stage3(h));
    hipError_t e = launch_gather_pos((int)n, src, h->m.inv, h->d_x3, h->stream);
    if (e == hipSuccess) e = d2h_xyz(h, dst);
    if (e != hipSuccess) return hip_fail(h, "download", e);
    return check_persistent(h);
}

// Valence-histogram key of the autotuner cache: free-vertex count bucket (log2), mean
// valence, and the share of vertices above 16 tets, quantised.
std::string tune_key(const DevMesh& m, bool contact, int material) {
    double sum = 0, n = 0, big = 0;
    for (size_t d = 0; d < m.valenceHist.size(); ++d) {
        sum += (double)d * m.valenceHist[d];
        n += m.valenceHist[d];
        if (d > 16) big += m.valenceHist[d];
    }
    const double mean = n > 0 ? sum / n : 0;
    char buf[128];
    snprintf(buf, sizeof(buf), "n%d_c%d_mean%d_big%d_mat%d_ct%d_m%d", (int)std::log2(std::max(1.0, n)),
             m.nColors, (int)std::lround(mean * 2), (int)std::lround(n > 0 ? 10 * big / n : 0), (int)m.perTetMat,
             (int)contact, material);
    return buf;
}

struct TuneEntry {
    int l, u, b;
};
std::mutex g_tune_mu;
std::map<std::string, TuneEntry> g_tune_cache;

// Launch configuration picked from the valence histogram alone: the smallest lane count
// that covers the mean valence in at most three strides.
int heuristic_lpv(const DevMesh& m) {
    double sum = 0, n = 0;
    for (size_t d = 0; d < m.valenceHist.size(); ++d) {
        sum += (double)d * m.valenceHist[d];
        n += m.valenceHist[d];
    }
    const double mean = n > 0 ? sum / n : 8;
    int l = 4;
    while (l < 32 && l * 3 < mean) l *= 2;
    return l;
}

// The tile cut (and its padded records) for a tiled configuration; the corotated material also
// needs the full records padded.
hipError_t tiles_for(vbd_solver* h, int records, int max_verts, std::string& err) {
    return build_tiles(h->m, records, max_verts, h->stream, err, h->materialModel == VBD_MATERIAL_COROTATED_GPU);
}

// The 52-byte full pair records are read by the per-vertex kernels (lanes > 0), the corotated
// material and the differentiation passes; the default tiled stable Neo-Hookean kernel reads the
// 32-byte compact records only.  They are built on first use and released when the handle's
// configuration stops reading them.
bool needs_full_records(const vbd_solver* h) {
    return h->lpv != 0 || h->materialModel == VBD_MATERIAL_COROTATED_GPU || h->d_w != nullptr;
}

void release_full_records_if_unused(vbd_solver* h) {
    if (needs_full_records(h) || (!h->m.pa && !h->m.tpa)) return;
    drop_graph(h);  // captured launches may reference them
    release_full_records(h->m, h->stream);
}

hipError_t ensure_energy_scratch(vbd_solver* h) {
    if (h->d_epart) return hipSuccess;
    return hipMalloc(&h->d_epart, sizeof(double) * ((size_t)energy_partials(h->m) + SUM_STAGE));
}

// Default launch configuration: the tiled kernel with the first tile in the preference order
// below that holds the largest valence; else the per-vertex kernel with the histogram
// heuristic.
hipError_t choose_default_config(vbd_solver* h, std::string& err) {
    // block 512 with 1 record per thread when the mean valence is 20 or more (a 256-record tile
    // would hold fewer than 13 vertices): the autotuner's choice on the 128^3 cube and the 96^3
    // sphere (k = 23.4, 22.9), 3.4% faster than 256 x 2 on the cube; block 256 with 1 on the
    // batched beams (k = 16.9)
    double sum = 0, n = 0;
    for (size_t d = 0; d < h->m.valenceHist.size(); ++d) {
        sum += (double)d * h->m.valenceHist[d];
        n += h->m.valenceHist[d];
    }
    const bool wide = n > 0 && sum / n >= 20.0;
    const int narrow_first[][2] = {{256, 1}, {256, 2}, {512, 1}, {512, 2}};
    const int wide_first[][2] = {{512, 1}, {256, 2}, {256, 1}, {512, 2}};
    for (const auto& c : wide ? wide_first : narrow_first) {
        std::string err;
        if (h->m.tiledOk && h->m.maxValence <= c[0] * c[1] &&
            tiles_for(h, c[0] * c[1], c[0] / 8 - 1, err) == hipSuccess) {
            h->lpv = 0;
            h->unroll = c[1];
            h->block = c[0];
            return hipSuccess;
        }
    }
    h->lpv = heuristic_lpv(h->m);
    h->unroll = 1;
    h->block = 256;
    return ensure_full_records(h->m, h->stream, err);
}

}  // namespace

extern "C" {

int32_t vbd_capi_version(void) { return VBD_CAPI_VERSION; }

const char* vbd_global_last_error(void) { return g_err.c_str(); }

const char* vbd_last_error(const vbd_solver* h) { return h ? h->err.c_str() : g_err.c_str(); }

int vbd_create(vbd_solver** out, int32_t device) {
    if (!out) return fail(nullptr, VBD_ERR_ARG, "out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(nullptr, "hipGetDeviceCount", e);
    if (device < 0 || device >= n) return fail(nullptr, VBD_ERR_ARG, "device index out of range");
    vbd_solver* h = new (std::nothrow) vbd_solver();
    if (!h) return fail(nullptr, VBD_ERR_ALLOC, "host allocation failed");
    h->device = device;
    DeviceGuard g(device);
    e = hipStreamCreateWithFlags(&h->ownStream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(nullptr, "hipStreamCreateWithFlags", e);
    }
    h->stream = h->ownStream;
    e = hipMalloc(&h->d_energy, sizeof(double) * 2);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(h->ownStream);
        delete h;
        return hip_fail(nullptr, "hipMalloc", e);
    }
    e = scratch_setup(device);  // this device's copy of the no-allocation scratch
    if (e != hipSuccess) {
        (void)hipFree(h->d_energy);
        (void)hipStreamDestroy(h->ownStream);
        delete h;
        return hip_fail(nullptr, "hipGetSymbolAddress(scratch)", e);
    }
    *out = h;
    return VBD_OK;
}

void vbd_destroy(vbd_solver* h) {
    if (!h) return;
    DeviceGuard g(h->device);
    (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_scratch(h);
    free_device_mesh(h->m);
    if (h->d_energy) (void)hipFree(h->d_energy);
    if (h->d_elog) (void)hipFree(h->d_elog);
    if (h->d_bar) (void)hipFree(h->d_bar);
    (void)hipStreamSynchronize(h->ownStream);
    (void)hipStreamDestroy(h->ownStream);
    delete h;
}

int vbd_set_stream(vbd_solver* h, void* stream) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    DeviceGuard g(h->device);
    hipStream_t next = stream ? (hipStream_t)stream : h->ownStream;
    if (stream) {
        int dev = -1;
        hipError_t e = hipStreamGetDevice(next, &dev);
        if (e != hipSuccess) return hip_fail(h, "hipStreamGetDevice", e);
        if (dev != h->device) return fail(h, VBD_ERR_ARG, "the stream belongs to another device");
    }
    // work already enqueued on the previous stream finishes before anything on the new one
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    h->stream = next;
    return VBD_OK;
}

int vbd_init(vbd_solver* h, const vbd_mesh_desc* d) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (!d) return fail(h, VBD_ERR_ARG, "mesh is null");
    const int nV = d->num_verts, nT = d->num_tets, nC = d->num_colors;
    if (nV < 0 || nT < 0 || nC < 0) return fail(h, VBD_ERR_ARG, "negative size");
    if ((int64_t)nT * 4 >= (int64_t)1 << 31) return fail(h, VBD_ERR_ARG, "4*num_tets must fit int32");
    if (nV >= (1 << 27)) return fail(h, VBD_ERR_ARG, "num_verts must be below 2^27 (32-bit byte offsets of 16-byte positions)");
    if ((nV > 0 && (!d->positions || !d->mass || !d->is_fixed)) ||
        (nT > 0 && (!d->tets || !d->dminv || !d->rest_volume)) || !d->color_offsets || (nV > 0 && !d->color_verts))
        return fail(h, VBD_ERR_ARG, "required mesh array is null");
    if ((d->tet_mu == nullptr) != (d->tet_lambda == nullptr))
        return fail(h, VBD_ERR_ARG, "tet_mu and tet_lambda must both be set or both be null");
    if (d->slot_order < 0 || d->slot_order > 2) return fail(h, VBD_ERR_ARG, "slot_order must be 0, 1 or 2");
    // host-side validation of the index arrays
    for (int64_t i = 0; i < (int64_t)4 * nT; ++i)
        if (d->tets[i] < 0 || d->tets[i] >= nV) return fail(h, VBD_ERR_ARG, "tet vertex index out of range");
    for (int t = 0; t < nT; ++t) {
        const int32_t* v = d->tets + 4 * (size_t)t;
        if (v[0] == v[1] || v[0] == v[2] || v[0] == v[3] || v[1] == v[2] || v[1] == v[3] || v[2] == v[3])
            return fail(h, VBD_ERR_ARG, "tet with repeated vertex");
    }
    if (d->color_offsets[0] != 0 || d->color_offsets[nC] != nV)
        return fail(h, VBD_ERR_ARG, "color offsets must start at 0 and end at num_verts");
    // offsets first (monotonic and ending at nV, so no color reads past the caller's list)
    for (int c = 0; c < nC; ++c)
        if (d->color_offsets[c + 1] < d->color_offsets[c]) return fail(h, VBD_ERR_ARG, "color offsets decrease");
    std::vector<int32_t> colorOf(nV > 0 ? nV : 1, -1);
    for (int c = 0; c < nC; ++c) {
        for (int i = d->color_offsets[c]; i < d->color_offsets[c + 1]; ++i) {
            const int v = d->color_verts[i];
            if (v < 0 || v >= nV || colorOf[v] != -1)
                return fail(h, VBD_ERR_ARG, "color lists must hold every vertex exactly once");
            colorOf[v] = c;
        }
    }
    // a valid coloring: no two corners of a tet share a color (same-color vertices must be
    // independent for the within-color parallel solve to equal the sequential sweep)
    for (int t = 0; t < nT; ++t) {
        const int32_t* v = d->tets + 4 * (size_t)t;
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b)
                if (colorOf[v[a]] == colorOf[v[b]])
                    return fail(h, VBD_ERR_ARG, "invalid coloring: two corners of a tet share a color");
    }
    if (!finite_array(d->positions, 3 * (size_t)nV) || (d->inertia && !finite_array(d->inertia, 3 * (size_t)nV)) ||
        (d->velocity && !finite_array(d->velocity, 3 * (size_t)nV)) || !finite_array(d->mass, nV) ||
        !finite_array(d->dminv, 9 * (size_t)nT) || !finite_array(d->rest_volume, nT))
        return fail(h, VBD_ERR_NUMERIC, "non-finite value in mesh input");

    DeviceGuard g(h->device);
    (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_scratch(h);
    free_device_mesh(h->m);
    h->inited = false;

    // stage host arrays on the device, build, free staging
    struct Stage {
        void* p = nullptr;
        ~Stage() { if (p) (void)hipFree(p); }
    };
    Stage s_tets, s_col, s_fix, s_dm, s_vol, s_mu, s_lam, s_x, s_y, s_v, s_m;
    auto up = [&](Stage& st, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(&st.p, bytes ? bytes : 4);
        if (e != hipSuccess) return e;
        if (bytes) return hipMemcpyAsync(st.p, src, bytes, hipMemcpyHostToDevice, h->stream);
        return hipSuccess;
    };
    HIP_CHECK(h, up(s_tets, d->tets, sizeof(int32_t) * 4 * (size_t)nT));
    HIP_CHECK(h, up(s_col, colorOf.data(), sizeof(int32_t) * (size_t)nV));
    HIP_CHECK(h, up(s_fix, d->is_fixed, (size_t)nV));
    HIP_CHECK(h, up(s_dm, d->dminv, sizeof(float) * 9 * (size_t)nT));
    HIP_CHECK(h, up(s_vol, d->rest_volume, sizeof(float) * (size_t)nT));
    if (d->tet_mu) {
        HIP_CHECK(h, up(s_mu, d->tet_mu, sizeof(float) * (size_t)nT));
        HIP_CHECK(h, up(s_lam, d->tet_lambda, sizeof(float) * (size_t)nT));
    }
    HIP_CHECK(h, up(s_x, d->positions, sizeof(float) * 3 * (size_t)nV));
    if (d->inertia) HIP_CHECK(h, up(s_y, d->inertia, sizeof(float) * 3 * (size_t)nV));
    if (d->velocity) HIP_CHECK(h, up(s_v, d->velocity, sizeof(float) * 3 * (size_t)nV));
    HIP_CHECK(h, up(s_m, d->mass, sizeof(float) * (size_t)nV));
    // Slot order: free vertices color-major, pinned last; inside a group either ascending caller
    // id or a Morton (Z-order) key of the positions, so that consecutive slots -- and the
    // neighbour gathers of consecutive vertices -- are close in space.  Results do not depend on
    // the order inside a color (same-color vertices share no tet).
    std::vector<int32_t> perm, slotOff;
    if (d->slot_order != 1 && nV > 0) {
        std::vector<uint64_t> key(nV);
        float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
        for (int v = 0; v < nV; ++v)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], d->positions[3 * (size_t)v + c]);
                hi[c] = std::max(hi[c], d->positions[3 * (size_t)v + c]);
            }
        auto spread = [](uint64_t x) {  // 21 bits -> every third bit
            x &= 0x1fffff;
            x = (x | x << 32) & 0x1f00000000ffffULL;
            x = (x | x << 16) & 0x1f0000ff0000ffULL;
            x = (x | x << 8) & 0x100f00f00f00f00fULL;
            x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
            x = (x | x << 2) & 0x1249249249249249ULL;
            return x;
        };
        for (int v = 0; v < nV; ++v) {
            uint64_t q[3];
            for (int c = 0; c < 3; ++c) {
                const float ext = hi[c] - lo[c];
                const float t = ext > 0 ? (d->positions[3 * (size_t)v + c] - lo[c]) / ext : 0.f;
                q[c] = (uint64_t)std::min(2097151.0f, std::max(0.0f, t * 2097151.0f));
            }
            key[v] = spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2;
        }
        std::vector<int32_t> byKey(nV);
        for (int v = 0; v < nV; ++v) byKey[v] = v;
        std::sort(byKey.begin(), byKey.end(),
                  [&](int a, int b) { return key[a] < key[b] || (key[a] == key[b] && a < b); });
        bool morton = d->slot_order == 2;
        if (d->slot_order == 0) {
            // auto: keep the caller's order unless the Morton order brings the four corners of
            // a tet closer together on average (shuffled or scattered ids)
            std::vector<int32_t> rank(nV);
            for (int i = 0; i < nV; ++i) rank[byKey[i]] = i;
            double spanId = 0, spanKey = 0;
            for (int t = 0; t < nT; ++t) {
                const int32_t* c = d->tets + 4 * (size_t)t;
                int lo1 = c[0], hi1 = c[0], lo2 = rank[c[0]], hi2 = rank[c[0]];
                for (int k = 1; k < 4; ++k) {
                    lo1 = std::min(lo1, c[k]);
                    hi1 = std::max(hi1, c[k]);
                    lo2 = std::min(lo2, rank[c[k]]);
                    hi2 = std::max(hi2, rank[c[k]]);
                }
                spanId += hi1 - lo1;
                spanKey += hi2 - lo2;
            }
            morton = spanKey < spanId;
        }
        if (morton) {
            // stable counting sort of the Morton order by group (color, pinned last)
            slotOff.assign(nC + 2, 0);
            std::vector<int64_t> start(nC + 2, 0);
            for (int v = 0; v < nV; ++v) start[(d->is_fixed[v] ? nC : colorOf[v]) + 1]++;
            for (int g = 0; g <= nC; ++g) start[g + 1] += start[g];
            for (int c = 0; c <= nC; ++c) slotOff[c] = (int32_t)start[c];
            perm.resize(nV);
            for (int i = 0; i < nV; ++i) {
                const int v = byKey[i];
                perm[start[d->is_fixed[v] ? nC : colorOf[v]]++] = v;
            }
        }
    }
    std::string err;
    hipError_t e = build_device_mesh(h->m, nV, nT, nC, (const i4*)s_tets.p, (const int*)s_col.p,
                                     (const uint8_t*)s_fix.p, (const float*)s_dm.p, (const float*)s_vol.p,
                                     (const float*)s_mu.p, (const float*)s_lam.p, (const float*)s_x.p,
                                     (const float*)s_y.p, (const float*)s_v.p, (const float*)s_m.p,
                                     perm.empty() ? nullptr : perm.data(), slotOff.empty() ? nullptr : slotOff.data(),
                                     h->stream, err);
    if (e != hipSuccess) return fail(h, e == hipErrorOutOfMemory ? VBD_ERR_ALLOC : VBD_ERR_HIP, "mesh build: " + err);
    e = choose_default_config(h, err);
    if (e != hipSuccess) {
        free_device_mesh(h->m);
        return fail(h, e == hipErrorOutOfMemory ? VBD_ERR_ALLOC : VBD_ERR_HIP, "launch configuration: " + err);
    }
    h->meshVertOffsets.assign({0, nV});
    h->inited = true;
    return VBD_OK;
}

int vbd_init_batch(vbd_solver* h, int32_t num_meshes, const vbd_mesh_desc* meshes) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (num_meshes <= 0 || !meshes) return fail(h, VBD_ERR_ARG, "no meshes");
    const bool mat = meshes[0].tet_mu != nullptr;
    int64_t nV = 0, nT = 0;
    int nC = 0;
    for (int i = 0; i < num_meshes; ++i) {
        const vbd_mesh_desc& d = meshes[i];
        if (d.num_verts < 0 || d.num_tets < 0 || d.num_colors < 0) return fail(h, VBD_ERR_ARG, "negative size");
        if ((d.tet_mu != nullptr) != mat || (d.tet_lambda != nullptr) != mat)
            return fail(h, VBD_ERR_ARG, "per-tet materials must be given for all meshes or none");
        if (!d.color_offsets) return fail(h, VBD_ERR_ARG, "color offsets null");
        if (d.num_verts > 0 && !d.color_verts) return fail(h, VBD_ERR_ARG, "color vertex list null");
        // each mesh's color offsets are read before the union is validated: check them first
        if (d.color_offsets[0] != 0 || d.color_offsets[d.num_colors] != d.num_verts)
            return fail(h, VBD_ERR_ARG, "color offsets must start at 0 and end at num_verts");
        for (int c = 0; c < d.num_colors; ++c)
            if (d.color_offsets[c + 1] < d.color_offsets[c]) return fail(h, VBD_ERR_ARG, "color offsets decrease");
        nV += d.num_verts;
        nT += d.num_tets;
        nC = std::max(nC, (int)d.num_colors);
    }
    if (nV >= ((int64_t)1 << 31) || 4 * nT >= ((int64_t)1 << 31)) return fail(h, VBD_ERR_ARG, "batch too large");
    std::vector<float> x(3 * nV), y, v, mass(nV), dm(9 * nT), vol(nT), mu, lam;
    std::vector<uint8_t> fixed(nV);
    std::vector<int32_t> tets(4 * nT), coff(nC + 1, 0), cverts;
    bool anyY = false, anyV = false;
    for (int i = 0; i < num_meshes; ++i) {
        anyY |= meshes[i].inertia != nullptr;
        anyV |= meshes[i].velocity != nullptr;
    }
    if (anyY) y.resize(3 * nV);
    if (anyV) v.assign(3 * nV, 0.f);
    if (mat) {
        mu.resize(nT);
        lam.resize(nT);
    }
    std::vector<int32_t> voff(num_meshes + 1, 0);
    int64_t vo = 0, to = 0;
    for (int i = 0; i < num_meshes; ++i) {
        const vbd_mesh_desc& d = meshes[i];
        const size_t nv = d.num_verts, nt = d.num_tets;
        if ((nv && (!d.positions || !d.mass || !d.is_fixed)) || (nt && (!d.tets || !d.dminv || !d.rest_volume)))
            return fail(h, VBD_ERR_ARG, "required mesh array is null");
        std::memcpy(&x[3 * vo], d.positions, sizeof(float) * 3 * nv);
        if (anyY) std::memcpy(&y[3 * vo], d.inertia ? d.inertia : d.positions, sizeof(float) * 3 * nv);
        if (anyV && d.velocity) std::memcpy(&v[3 * vo], d.velocity, sizeof(float) * 3 * nv);
        std::memcpy(&mass[vo], d.mass, sizeof(float) * nv);
        std::memcpy(&fixed[vo], d.is_fixed, nv);
        for (size_t k = 0; k < 4 * nt; ++k) {
            if (d.tets[k] < 0 || d.tets[k] >= (int)nv) return fail(h, VBD_ERR_ARG, "tet vertex index out of range");
            tets[4 * to + k] = d.tets[k] + (int32_t)vo;
        }
        std::memcpy(&dm[9 * to], d.dminv, sizeof(float) * 9 * nt);
        std::memcpy(&vol[to], d.rest_volume, sizeof(float) * nt);
        if (mat) {
            std::memcpy(&mu[to], d.tet_mu, sizeof(float) * nt);
            std::memcpy(&lam[to], d.tet_lambda, sizeof(float) * nt);
        }
        if (d.color_offsets[d.num_colors] != (int)nv) return fail(h, VBD_ERR_ARG, "color offsets do not end at num_verts");
        vo += nv;
        to += nt;
        voff[i + 1] = (int32_t)vo;
    }
    // colors merged by index: merged color c = union over meshes of each mesh's color c
    cverts.reserve(nV);
    for (int c = 0; c < nC; ++c) {
        for (int i = 0; i < num_meshes; ++i) {
            const vbd_mesh_desc& d = meshes[i];
            if (c >= d.num_colors) continue;
            for (int k = d.color_offsets[c]; k < d.color_offsets[c + 1]; ++k) cverts.push_back(d.color_verts[k] + voff[i]);
        }
        coff[c + 1] = (int32_t)cverts.size();
    }
    vbd_mesh_desc u{};
    u.num_verts = (int32_t)nV;
    u.num_tets = (int32_t)nT;
    u.positions = x.data();
    u.inertia = anyY ? y.data() : nullptr;
    u.velocity = anyV ? v.data() : nullptr;
    u.mass = mass.data();
    u.is_fixed = fixed.data();
    u.tets = tets.data();
    u.dminv = dm.data();
    u.rest_volume = vol.data();
    u.tet_mu = mat ? mu.data() : nullptr;
    u.tet_lambda = mat ? lam.data() : nullptr;
    u.num_colors = nC;
    u.color_offsets = coff.data();
    u.color_verts = cverts.data();
    u.slot_order = meshes[0].slot_order;
    int rc = vbd_init(h, &u);
    if (rc == VBD_OK) h->meshVertOffsets = voff;
    return rc;
}

int vbd_get_info(const vbd_solver* h, vbd_info* info) {
    if (!h || !info) return fail(nullptr, VBD_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    if (!h->inited) return VBD_OK;
    info->num_verts = h->m.nV;
    info->num_tets = h->m.nT;
    info->num_free = h->m.nFree;
    info->num_colors = h->m.nColors;
    info->num_pairs = h->m.nPairs;
    info->max_valence = h->m.maxValence;
    info->per_tet_material = h->m.perTetMat;
    info->lanes_per_vertex = h->lpv;
    info->pairs_per_lane = h->unroll;
    info->block_size = h->block;
    info->device_bytes = h->m.bytes;
    info->cached_graphs = (int32_t)h->blockGraphs.size() + (h->graphExec ? 1 : 0);
    info->full_records = h->m.pa != nullptr;
    if (persistent_on(h) && h->paramsSet) {
        DeviceGuard g(h->device);
        info->persistent_grid = persistent_grid(h, contact_of(h->params), h->params.use_acceleration != 0);
    }
    return VBD_OK;
}

int vbd_upload_positions(vbd_solver* h, const float* x) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!x && h->m.nV) return fail(h, VBD_ERR_ARG, "null positions");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    return upload_vertex_array(h, x, h->m.pos);
}

int vbd_upload_inertia(vbd_solver* h, const float* y) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!y && h->m.nV) return fail(h, VBD_ERR_ARG, "null inertia");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    return upload_vertex_array(h, y, h->m.inert);
}

int vbd_upload_velocities(vbd_solver* h, const float* v) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!v && h->m.nV) return fail(h, VBD_ERR_ARG, "null velocities");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    return upload_vertex_array(h, v, h->m.vel);
}

int vbd_download_positions(vbd_solver* h, float* x) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!x && h->m.nV) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    return download_vertex_array(h, h->m.pos, x);
}

int vbd_download_velocities(vbd_solver* h, float* v) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!v && h->m.nV) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    return download_vertex_array(h, h->m.vel, v);
}

int vbd_download_inertia(vbd_solver* h, float* y) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!y && h->m.nV) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    return download_vertex_array(h, h->m.inert, y);
}

int vbd_download_previous_positions(vbd_solver* h, float* x) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!x && h->m.nV) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    return download_vertex_array(h, h->m.xprev, x);
}

int vbd_get_adjacency(vbd_solver* h, int32_t* offsets, int32_t* entries) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!offsets || (!entries && h->m.nT)) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    HIP_CHECK(h, hipMemcpyAsync(offsets, h->m.adjOff, sizeof(int32_t) * (h->m.nV + 1), hipMemcpyDeviceToHost, h->stream));
    if (h->m.nT)
        HIP_CHECK(h, hipMemcpyAsync(entries, h->m.adjEntry, sizeof(int32_t) * 4 * (size_t)h->m.nT,
                                    hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_get_color_lists(vbd_solver* h, int32_t* offsets, int32_t* verts) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!offsets || (!verts && h->m.nV)) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    for (int c = 0; c <= h->m.nColors; ++c) offsets[c] = h->m.colorFullOff[c];
    if (h->m.nV)
        HIP_CHECK(h, hipMemcpyAsync(verts, h->m.colorList, sizeof(int32_t) * h->m.nV, hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_solve_iteration(vbd_solver* h, float dt_sqr_reciprocal, float mu, float lambda) {
    int rc = check_ready(h);
    if (rc) return rc;
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    ContactArgs ct{};
    if (h->paramsSet) ct = contact_of(h->params);
    HIP_CHECK(h, enqueue_sweep(h, dt_sqr_reciprocal, mu, lambda, ct, nullptr));
    return VBD_OK;
}

int vbd_solve(vbd_solver* h, int32_t num_iterations, float dt_sqr_reciprocal, float mu, float lambda,
              int32_t use_acceleration, float rho, int32_t start_from_inertia) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (num_iterations < 0) return fail(h, VBD_ERR_ARG, "negative iteration count");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    ContactArgs ct{};
    if (h->paramsSet) ct = contact_of(h->params);
    if (num_iterations == 0 && !start_from_inertia) return VBD_OK;
    // the whole call is one cached graph (asynchronous, like the reference solveVBD)
    const std::vector<float> args = {(float)num_iterations, dt_sqr_reciprocal, mu, lambda, (float)use_acceleration,
                                     rho, (float)start_from_inertia};
    HIP_CHECK(h, run_block(h, block_key('S', 0, args, 0.f), [&]() -> hipError_t {
        if (start_from_inertia) {
            hipError_t e = launch_copy_inertia(h->m, h->stream);
            if (e != hipSuccess) return e;
        }
        return enqueue_iterations(h, num_iterations, dt_sqr_reciprocal, mu, lambda, use_acceleration != 0, rho, ct);
    }, false));
    return VBD_OK;
}

int vbd_set_params(vbd_solver* h, const vbd_step_params* p) {
    if (!h || !p) return fail(h, VBD_ERR_ARG, "null argument");
    if (!(p->dt > 0.f) || p->substeps <= 0 || p->num_iterations < 0)
        return fail(h, VBD_ERR_ARG, "dt must be > 0, substeps > 0, num_iterations >= 0");
    if (p->contact_on && (p->contact_axis < 0 || p->contact_axis > 2))
        return fail(h, VBD_ERR_ARG, "contact axis must be 0, 1 or 2");
    h->params = *p;
    h->paramsSet = true;
    drop_graph(h);
    return VBD_OK;
}

int vbd_forward_step(vbd_solver* h) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    HIP_CHECK(h, launch_predict(h->m, h->params.dt / (float)h->params.substeps, h->params.gravity, h->stream));
    return VBD_OK;
}

int vbd_update_velocity(vbd_solver* h) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    HIP_CHECK(h, launch_velocity(h->m, h->params.dt / (float)h->params.substeps, h->params.velocity_damping, h->stream));
    return VBD_OK;
}

int vbd_set_graph_enabled(vbd_solver* h, int32_t enabled) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    h->graphEnabled = enabled != 0;
    DeviceGuard g(h->device);
    drop_graph(h);
    return VBD_OK;
}

int vbd_step(vbd_solver* h, int32_t num_frames) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    if (num_frames < 0) return fail(h, VBD_ERR_ARG, "negative frame count");
    // an empty mesh has nothing to update; with every vertex pinned the step still runs
    // forwardStep (x_prev = x, y = x for pinned vertices, VBD.cpp:103-112) and the energy record
    if (h->m.nV == 0) return VBD_OK;
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    if (h->energyEvery > 0) {  // buffers of the energy log exist before a capture
        HIP_CHECK(h, ensure_energy_scratch(h));
        const int n = energy_log_slots(h);
        if (h->elogCap < n) {
            drop_graph(h);
            if (h->d_elog) (void)hipFree(h->d_elog);
            h->d_elog = nullptr;
            h->elogCap = 0;
            HIP_CHECK(h, hipMalloc(&h->d_elog, sizeof(double) * (size_t)n));
            h->elogCap = n;
        }
    }
    if (!h->graphEnabled) {
        for (int f = 0; f < num_frames; ++f) HIP_CHECK(h, enqueue_step(h));
        return VBD_OK;
    }
    if (!h->graphValid) {
        hipGraph_t graph = nullptr;
        HIP_CHECK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_step(h);
        hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
        if (e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_fail(h, "capture of the time step", e);
        }
        if (e2 != hipSuccess) return hip_fail(h, "hipStreamEndCapture", e2);
        e = hipGraphInstantiate(&h->graphExec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            h->graphExec = nullptr;
            return hip_fail(h, "hipGraphInstantiate", e);
        }
        h->graphValid = true;
    }
    for (int f = 0; f < num_frames; ++f) HIP_CHECK(h, hipGraphLaunch(h->graphExec, h->stream));
    return VBD_OK;
}

int vbd_set_energy_log(vbd_solver* h, int32_t every) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    DeviceGuard g(h->device);
    drop_graph(h);
    h->energyEvery = every > 0 ? every : 0;
    return VBD_OK;
}

int vbd_energy_log(vbd_solver* h, double* out, int32_t capacity, int32_t* count) {
    int rc = check_ready(h);
    if (rc) return rc;
    const int n = h->d_elog ? std::min(energy_log_slots(h), h->elogCap) : 0;
    if (count) *count = n;
    if (capacity < 0 || (!out && capacity > 0)) return fail(h, VBD_ERR_ARG, "bad output buffer");
    const int k = std::min(n, (int)capacity);
    DeviceGuard g(h->device);
    if (k > 0) HIP_CHECK(h, hipMemcpyAsync(out, h->d_elog, sizeof(double) * (size_t)k, hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_synchronize(vbd_solver* h) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    DeviceGuard g(h->device);
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return check_persistent(h);
}

int vbd_set_persistent(vbd_solver* h, int32_t mode) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (mode < 0 || (mode > 1 && (mode & 7))) return fail(h, VBD_ERR_ARG, "persistent mode must be 0, 1 or a multiple of 8");
    DeviceGuard g(h->device);
    drop_graph(h);
    if (mode && !h->d_bar) {
        HIP_CHECK(h, hipMalloc(&h->d_bar, PERSIST_BAR_WORDS * sizeof(unsigned)));
        HIP_CHECK(h, hipMemsetAsync(h->d_bar, 0, PERSIST_BAR_WORDS * sizeof(unsigned), h->stream));
        HIP_CHECK(h, hipStreamSynchronize(h->stream));
    }
    h->persistent = mode;
    return VBD_OK;
}

int vbd_set_material_model(vbd_solver* h, int32_t model) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (model != VBD_MATERIAL_STABLE_NEO_HOOKEAN && model != VBD_MATERIAL_COROTATED_GPU)
        return fail(h, VBD_ERR_ARG, "unknown material model");
    DeviceGuard g(h->device);
    drop_graph(h);
    h->materialModel = model;
    if (h->inited) {
        std::string err;
        if (h->lpv == 0 && tiles_for(h, h->unroll * h->block, h->block / 8 - 1, err) != hipSuccess)
            return fail(h, VBD_ERR_STATE, err);
        if (needs_full_records(h) && ensure_full_records(h->m, h->stream, err) != hipSuccess)
            return fail(h, VBD_ERR_ALLOC, err);
        release_full_records_if_unused(h);
    }
    return VBD_OK;
}

int vbd_set_launch_config(vbd_solver* h, int32_t lanes_per_vertex, int32_t block_size, int32_t pairs_per_lane) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (lanes_per_vertex == 0) {
        if (!tiled_config_supported(pairs_per_lane, block_size))
            return fail(h, VBD_ERR_ARG,
                        "unsupported tiled configuration: block_size 64-512 (multiple of 64), pairs_per_lane 1 or 2");
        if (!h->inited) return fail(h, VBD_ERR_STATE, "the tiled kernel needs an initialised mesh");
        DeviceGuard g(h->device);
        std::string err;
        if (tiles_for(h, block_size * pairs_per_lane, block_size / 8 - 1, err) != hipSuccess)
            return fail(h, VBD_ERR_ARG, err);
    } else if (!launch_config_supported(lanes_per_vertex, pairs_per_lane, block_size) ||
               block_size % lanes_per_vertex) {
        return fail(h, VBD_ERR_ARG,
                    "unsupported launch configuration: lanes_per_vertex 4/8/16 with pairs_per_lane 1-3, or 32/64 with "
                    "pairs_per_lane 1; block_size 64/128/192/256; or lanes_per_vertex 0 (tiled)");
    }
    DeviceGuard g(h->device);
    if (lanes_per_vertex > 0 && h->inited) {  // the per-vertex kernels read the full records
        std::string err;
        if (ensure_full_records(h->m, h->stream, err) != hipSuccess) return fail(h, VBD_ERR_ALLOC, err);
    }
    h->lpv = lanes_per_vertex;
    h->unroll = pairs_per_lane;
    h->block = block_size;
    drop_graph(h);
    if (h->inited) release_full_records_if_unused(h);
    return VBD_OK;
}

int vbd_time_iterations(vbd_solver* h, int32_t iterations, double* color_ms, int32_t* color_launches,
                        double* blend_ms, int32_t* blend_launches) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (iterations <= 0) return fail(h, VBD_ERR_ARG, "iterations must be positive");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    const vbd_step_params p = h->params;
    const bool accel = h->paramsSet && p.use_acceleration;
    const float hstep = h->paramsSet ? p.dt / (float)p.substeps : 1.0f / 60.0f;
    const float sdt = inv_h2_ref(hstep);
    const float mu = h->paramsSet ? p.mu : 1e5f, lam = h->paramsSet ? p.lambda : 4e5f;
    const ContactArgs ct = h->paramsSet ? contact_of(p) : ContactArgs{};
    struct Ev {
        hipEvent_t a, b;
        bool blend;
    };
    std::vector<Ev> evs;
    evs.reserve((size_t)iterations * (h->m.nColors + 1));
    auto mk = [&](bool blend) -> hipError_t {
        Ev e{};
        e.blend = blend;
        hipError_t r = hipEventCreate(&e.a);
        if (r != hipSuccess) return r;
        r = hipEventCreate(&e.b);
        if (r != hipSuccess) { (void)hipEventDestroy(e.a); return r; }
        evs.push_back(e);
        return hipSuccess;
    };
    hipError_t e = hipSuccess;
    float omega = 1.0f;
    bool kept = false;  // as in the step: a blend writes the next sweep's start
    for (int it = 0; it < iterations && e == hipSuccess; ++it) {
        f4* save = accel && !kept ? h->m.save[it & 1] : nullptr;
        for (int c = 0; c < h->m.nColors && e == hipSuccess; ++c) {
            if (h->m.colorSlotOff[c + 1] == h->m.colorSlotOff[c]) continue;
            if ((e = mk(false)) != hipSuccess) break;
            (void)hipEventRecord(evs.back().a, h->stream);
            e = launch_color_cfg(h, c, sdt, mu, lam, ct, save);
            (void)hipEventRecord(evs.back().b, h->stream);
        }
        if (accel && e == hipSuccess) {
            omega = omega_ref(it + 1, p.acceleration_rho, omega);
            kept = omega > 1.0f;
            if (kept) {
                if ((e = mk(true)) != hipSuccess) break;
                (void)hipEventRecord(evs.back().a, h->stream);
                e = launch_blend(h->m, omega, h->m.save[(it + 1) & 1], h->stream, true);
                (void)hipEventRecord(evs.back().b, h->stream);
            }
        }
    }
    hipError_t es = hipStreamSynchronize(h->stream);
    double cms = 0, bms = 0;
    int cn = 0, bn = 0;
    for (auto& ev : evs) {
        float ms = 0.f;
        if (e == hipSuccess && es == hipSuccess) (void)hipEventElapsedTime(&ms, ev.a, ev.b);
        if (ev.blend) { bms += ms; ++bn; } else { cms += ms; ++cn; }
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    if (e != hipSuccess) return hip_fail(h, "timed launch", e);
    if (es != hipSuccess) return hip_fail(h, "hipStreamSynchronize", es);
    if (color_ms) *color_ms = cms;
    if (color_launches) *color_launches = cn;
    if (blend_ms) *blend_ms = bms;
    if (blend_launches) *blend_launches = bn;
    return VBD_OK;
}

namespace {
// Milliseconds of `iterations` sweeps (with the Chebyshev blends when the parameters enable
// them) launched back to back between two events, as a replayed step runs them.
hipError_t time_sweeps(vbd_solver* h, int iterations, double* ms) {
    const vbd_step_params p = h->params;
    const bool accel = h->paramsSet && p.use_acceleration;
    const float hstep = h->paramsSet ? p.dt / (float)p.substeps : 1.0f / 60.0f;
    const float sdt = inv_h2_ref(hstep);
    const float mu = h->paramsSet ? p.mu : 1e5f, lam = h->paramsSet ? p.lambda : 4e5f;
    const ContactArgs ct = h->paramsSet ? contact_of(p) : ContactArgs{};
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, h->stream);
    if (e == hipSuccess) e = enqueue_iterations(h, iterations, sdt, mu, lam, accel, p.acceleration_rho, ct);
    if (e == hipSuccess) e = hipEventRecord(b, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    *ms = t;
    return e;
}
}  // namespace

int vbd_autotune(vbd_solver* h, int32_t iterations, int32_t* best_lanes, int32_t* best_unroll, int32_t* best_block,
                 float* best_ms) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (iterations <= 0) return fail(h, VBD_ERR_ARG, "iterations must be positive");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    const std::string key = tune_key(h->m, h->paramsSet && h->params.contact_on, h->materialModel);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        std::string terr;
        if (it != g_tune_cache.end() &&
            (it->second.l != 0 || tiles_for(h, it->second.u * it->second.b, it->second.b / 8 - 1, terr) == hipSuccess) &&
            (it->second.l == 0 || ensure_full_records(h->m, h->stream, terr) == hipSuccess)) {
            h->lpv = it->second.l;
            h->unroll = it->second.u;
            h->block = it->second.b;
            drop_graph(h);
            release_full_records_if_unused(h);
            if (best_lanes) *best_lanes = h->lpv;
            if (best_unroll) *best_unroll = h->unroll;
            if (best_block) *best_block = h->block;
            if (best_ms) *best_ms = -1.f;
            return VBD_OK;
        }
    }
    // the per-vertex candidates read the full records (released below if the winner does not)
    {
        std::string ferr;
        if (ensure_full_records(h->m, h->stream, ferr) != hipSuccess) return fail(h, VBD_ERR_ALLOC, ferr);
    }
    // keep the state: timing sweeps move the positions and the Chebyshev buffers
    f4* keep = nullptr;
    HIP_CHECK(h, hipMalloc(&keep, sizeof(f4) * (h->m.nV ? h->m.nV : 1)));
    hipError_t e = launch_copy_f4(h->m.nV, h->m.pos, keep, h->stream);
    struct Cand {
        int l, u;
    };
    // per-vertex kernel: (lanes, pairs per lane) x blocks 64/128/256; tiled kernel (lanes 0):
    // (records per thread) x blocks 128/256/512
    const Cand cands[] = {{4, 1}, {4, 2}, {4, 3}, {8, 1}, {8, 2}, {8, 3}, {16, 1}, {16, 2}, {32, 1}, {0, 1}, {0, 2}};
    const int cand_b[] = {64, 128, 256, 512};
    int bl = h->lpv, bu = h->unroll, bb = h->block;
    double best = 1e30, current = 1e30;  // current: the configuration set when autotune was called
    const int ol = h->lpv, ou = h->unroll, ob = h->block;
    // the current tile cut stays allocated (out of the mesh) while candidates build their own
    TileCut kept;
    detach_tiles(h->m, kept);
    for (const Cand& cd : cands) {
        for (int b : cand_b) {
            if (e != hipSuccess) break;
            if (cd.l > 0 && (b > 256 || b % cd.l)) continue;
            // the corotated model runs the tiled kernel or the per-vertex kernel with lanes 8/16/32
            if (h->materialModel == VBD_MATERIAL_COROTATED_GPU && cd.l != 0 &&
                (cd.u != 1 || (cd.l != 8 && cd.l != 16 && cd.l != 32)))
                continue;
            if (cd.l == 0) {
                if (b < 128 || !tiled_config_supported(cd.u, b) || !h->m.tiledOk || h->m.maxValence > cd.u * b)
                    continue;
                std::string terr;
                if (tiles_for(h, cd.u * b, b / 8 - 1, terr) != hipSuccess) continue;
            }
            h->lpv = cd.l;
            h->unroll = cd.u;
            h->block = b;
            // one untimed warm sweep, then the sweeps back to back (as a replayed step runs them)
            double cms = 0;
            hipError_t te = time_sweeps(h, 1, &cms);
            if (te == hipSuccess) te = time_sweeps(h, iterations, &cms);
            rc = te == hipSuccess ? VBD_OK : hip_fail(h, "autotune timing", te);
            if (rc != VBD_OK) {
                h->lpv = ol;
                h->unroll = ou;
                h->block = ob;
                attach_tiles(h->m, kept, h->stream);
                (void)launch_copy_f4(h->m.nV, keep, h->m.pos, h->stream);
                (void)hipStreamSynchronize(h->stream);
                (void)hipFree(keep);
                return rc;
            }
            if (cd.l == ol && cd.u == ou && b == ob) current = cms;
            if (cms < best) {
                best = cms;
                bl = cd.l;
                bu = cd.u;
                bb = b;
            }
            e = launch_copy_f4(h->m.nV, keep, h->m.pos, h->stream);
        }
    }
    // keep the current configuration unless a candidate is at least 1% faster: near-ties are
    // timing noise, and a stable choice keeps repeated runs on the same configuration
    if (current < 1e30 && best > 0.99 * current) {
        bl = ol;
        bu = ou;
        bb = ob;
        best = current;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(keep);
    if (e != hipSuccess) {
        h->lpv = ol;
        h->unroll = ou;
        h->block = ob;
        attach_tiles(h->m, kept, h->stream);
        drop_graph(h);
        return hip_fail(h, "autotune", e);
    }
    if (bl == ol && bu == ou && bb == ob && kept.tileStart) {
        attach_tiles(h->m, kept, h->stream);  // the current configuration won: its own cut back
    } else {
        free_tile_cut(kept);
        if (bl == 0) {
            std::string terr;
            if (tiles_for(h, bu * bb, bb / 8 - 1, terr) != hipSuccess) return fail(h, VBD_ERR_STATE, terr);
        }
    }
    h->lpv = bl;
    h->unroll = bu;
    h->block = bb;
    drop_graph(h);
    release_full_records_if_unused(h);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {bl, bu, bb};
    }
    if (best_lanes) *best_lanes = bl;
    if (best_unroll) *best_unroll = bu;
    if (best_block) *best_block = bb;
    if (best_ms) *best_ms = (float)(best / iterations);
    return VBD_OK;
}

int vbd_time_color_launches(vbd_solver* h, int32_t sweeps, double* ms_per_launch, int32_t* launches) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (sweeps <= 0) return fail(h, VBD_ERR_ARG, "sweeps must be positive");
    DeviceGuard g(h->device);
    h->fp64Fresh = false;  // the FP32 state changes
    const vbd_step_params p = h->params;
    const float hstep = h->paramsSet ? p.dt / (float)p.substeps : 1.0f / 60.0f;
    const float sdt = inv_h2_ref(hstep);
    const float mu = h->paramsSet ? p.mu : 1e5f, lam = h->paramsSet ? p.lambda : 4e5f;
    const ContactArgs ct = h->paramsSet ? contact_of(p) : ContactArgs{};
    int per_sweep = 0;
    for (int c = 0; c < h->m.nColors; ++c)
        if (h->m.colorSlotOff[c + 1] > h->m.colorSlotOff[c]) ++per_sweep;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t ee = hipSuccess;
        // the step's color launches after its first two sweeps: the blends write each sweep's
        // start, so the color launches save nothing
        for (int it = 0; it < sweeps && ee == hipSuccess; ++it) ee = enqueue_sweep(h, sdt, mu, lam, ct, nullptr);
        e = hipStreamEndCapture(h->stream, &graph);
        if (ee != hipSuccess) e = ee;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipGraphLaunch(exec, h->stream);  // untimed: warm caches and clocks
    if (e == hipSuccess) e = hipEventRecord(a, h->stream);
    if (e == hipSuccess) e = hipGraphLaunch(exec, h->stream);
    if (e == hipSuccess) e = hipEventRecord(b, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(h, "vbd_time_color_launches", e);
    const int n = sweeps * per_sweep;
    if (ms_per_launch) *ms_per_launch = n > 0 ? (double)ms / n : 0.0;
    if (launches) *launches = n;
    return VBD_OK;
}

int vbd_layout_bytes(const vbd_solver* h, double* color_bytes, double* blend_bytes) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (!h->inited) return fail(const_cast<vbd_solver*>(h), VBD_ERR_STATE, "not initialised");
    const bool compact = h->lpv == 0 && h->materialModel == VBD_MATERIAL_STABLE_NEO_HOOKEAN;
    const double per_rec = (compact ? 32.0 : 52.0) + (h->m.perTetMat ? 8.0 : 0.0);
    if (color_bytes) *color_bytes = 56.0 * h->m.nFree + per_rec * (double)h->m.nPairs;
    if (blend_bytes) *blend_bytes = 64.0 * h->m.nFree;
    return VBD_OK;
}

// Algorithmic bytes (fixed yardstick; neighbour-position gathers excluded):
//   color launches, per solved vertex: x read 12 + x write 12 + y 12 + mass 4 + CSR row 4 +
//   fixed flag 1 = 45; per (vertex, incident tet): CSR entry 4 + tet 16 + DmInv 36 + V0 4 = 60,
//   + 8 (mu, lambda) with per-tet materials.  Chebyshev adds 36 per vertex (read x_{n-2},
//   write x_{n-1}, write blended x), attributed to the blend launches.
int vbd_algorithmic_bytes(const vbd_solver* h, double* color_bytes, double* blend_bytes) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (!h->inited) return fail(const_cast<vbd_solver*>(h), VBD_ERR_STATE, "not initialised");
    const double per_pair = 60.0 + (h->m.perTetMat ? 8.0 : 0.0);
    if (color_bytes) *color_bytes = 45.0 * h->m.nFree + per_pair * (double)h->m.nPairs;
    if (blend_bytes) *blend_bytes = 36.0 * h->m.nFree;
    return VBD_OK;
}

int vbd_total_energy(vbd_solver* h, float mu, float lambda, double* energy) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!energy) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    HIP_CHECK(h, ensure_energy_scratch(h));
    HIP_CHECK(h, launch_energy(h->m, 0, mu, lambda, h->d_epart, h->d_energy, h->stream));
    HIP_CHECK(h, hipMemcpyAsync(energy, h->d_energy, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_total_energy_noalloc(vbd_solver* h, float mu, float lambda, double* energy_out) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!energy_out) return fail(h, VBD_ERR_ARG, "null output");
    if (energy_partials(h->m) + SUM_STAGE > VBD_NOALLOC_SCRATCH_DOUBLES)
        return fail(h, VBD_ERR_ARG, "mesh too large for the no-allocation scratch (at most 33488896 tets and free vertices)");
    double* ptr = scratch_ptr(h->device);
    if (!ptr) return fail(h, VBD_ERR_STATE, "no-allocation scratch not set up on this device");
    DeviceGuard g(h->device);
    HIP_CHECK(h, launch_energy(h->m, 0, mu, lambda, ptr, energy_out, h->stream));
    return VBD_OK;
}

int vbd_incremental_potential(vbd_solver* h, double* potential, double* elastic) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    DeviceGuard g(h->device);
    const vbd_step_params& p = h->params;
    const float hstep = p.dt / (float)p.substeps;
    double out[2] = {0.0, 0.0};
    HIP_CHECK(h, ensure_energy_scratch(h));
    HIP_CHECK(h, launch_energy(h->m, h->materialModel == VBD_MATERIAL_COROTATED_GPU ? 2 : 1, p.mu, p.lambda,
                               h->d_epart, h->d_energy, h->stream));
    HIP_CHECK(h, launch_inertia_energy(h->m, 1.0 / ((double)hstep * hstep), contact_of(p), h->d_epart,
                                       h->d_energy + 1, h->stream));
    HIP_CHECK(h, hipMemcpyAsync(out, h->d_energy, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    if (elastic) *elastic = out[0];
    if (potential) *potential = out[0] + out[1];
    return VBD_OK;
}

int vbd_time_energy(vbd_solver* h, int32_t mode, int32_t reps, double* ms_per_eval, double* bytes_per_eval) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (mode != 0 && mode != 1) return fail(h, VBD_ERR_ARG, "mode must be 0 (diagnostic) or 1 (incremental potential)");
    if (reps <= 0) return fail(h, VBD_ERR_ARG, "reps must be positive");
    if (mode == 1 && !h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    DeviceGuard g(h->device);
    HIP_CHECK(h, ensure_energy_scratch(h));
    const vbd_step_params& p = h->params;
    const float mu = h->paramsSet ? p.mu : 1e5f, lam = h->paramsSet ? p.lambda : 4e5f;
    const int tmode = mode == 0 ? 0 : (h->materialModel == VBD_MATERIAL_COROTATED_GPU ? 2 : 1);
    const double sdt = h->paramsSet ? 1.0 / ((double)(p.dt / (float)p.substeps) * (p.dt / (float)p.substeps)) : 0.0;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, h->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        e = launch_energy(h->m, tmode, mu, lam, h->d_epart, h->d_energy, h->stream);
        if (e == hipSuccess && mode == 1)
            e = launch_inertia_energy(h->m, sdt, contact_of(p), h->d_epart, h->d_energy + 1, h->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(b, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(h, "vbd_time_energy", e);
    if (ms_per_eval) *ms_per_eval = (double)ms / reps;
    if (bytes_per_eval)
        *bytes_per_eval = (56.0 + (h->m.perTetMat ? 8.0 : 0.0)) * h->m.nT + 12.0 * h->m.nV +
                          (mode == 1 ? 28.0 * h->m.nFree : 0.0);
    return VBD_OK;
}

// ---------------------------------------------------------------------------------------
// differentiation entry points
// ---------------------------------------------------------------------------------------
namespace {

// d_r2 holds max(nV, 1) per-vertex values, then SUM_STAGE doubles for the two-stage sum
double* r2_stage(vbd_solver* h) { return h->d_r2 + (h->m.nV > 0 ? h->m.nV : 1); }

hipError_t ensure_scratch(vbd_solver* h) {
    const size_t n = (size_t)(h->m.nV > 0 ? h->m.nV : 1);
    const size_t t = (size_t)(h->m.nT > 0 ? h->m.nT : 1);
    hipError_t e = hipSuccess;
    std::string err;  // the differentiation passes read the full records
    if ((e = ensure_full_records(h->m, h->stream, err)) != hipSuccess) return e;
    if (!h->d_w && (e = hipMalloc(&h->d_w, sizeof(f4) * n)) != hipSuccess) return e;
    if (!h->d_rhs && (e = hipMalloc(&h->d_rhs, sizeof(f4) * n)) != hipSuccess) return e;
    if (!h->d_out && (e = hipMalloc(&h->d_out, sizeof(f4) * n)) != hipSuccess) return e;
    if (!h->d_wsave[0] && (e = hipMalloc(&h->d_wsave[0], sizeof(f4) * n)) != hipSuccess) return e;
    if (!h->d_wsave[1] && (e = hipMalloc(&h->d_wsave[1], sizeof(f4) * n)) != hipSuccess) return e;
    if (!h->d_r2 && (e = hipMalloc(&h->d_r2, sizeof(double) * (n + SUM_STAGE))) != hipSuccess) return e;
    if (!h->d_sum && (e = hipMalloc(&h->d_sum, sizeof(double) * 2)) != hipSuccess) return e;
    if (!h->d_tgrad && (e = hipMalloc(&h->d_tgrad, sizeof(float) * 2 * t)) != hipSuccess) return e;
    if (!h->d_x3 && (e = hipMalloc(&h->d_x3, sizeof(float) * 3 * n)) != hipSuccess) return e;
    if (!h->h_res && (e = hipHostMalloc(&h->h_res, sizeof(double) * 2, hipHostMallocDefault)) != hipSuccess) return e;
    return e;
}

// Lanes per vertex of the FP32 differentiation passes: the launch configuration's when it is 8
// or 32, else 32 (one lane per record of a typical vertex; the 16^3 adjoint ran 10% faster than
// with 16).
int pass_lpv(const vbd_solver* h) { return (h->lpv == 8 || h->lpv == 32) ? h->lpv : 32; }

// host 3N (caller order) -> slot f4 buffer; pinned slots zeroed when zero_pinned
hipError_t upload_vec(vbd_solver* h, const float* src, f4* dst, bool zero_pinned) {
    hipError_t e = h2d_xyz(h, src);
    if (e == hipSuccess) e = launch_scatter_pos(h->m.nV, h->d_x3, h->m.inv, dst, h->stream);
    if (e == hipSuccess && zero_pinned && h->m.nV > h->m.nFree)
        e = hipMemsetAsync(dst + h->m.nFree, 0, sizeof(f4) * (size_t)(h->m.nV - h->m.nFree), h->stream);
    return e;
}

hipError_t download_vec(vbd_solver* h, const f4* src, float* dst) {
    hipError_t e = launch_gather_pos(h->m.nV, src, h->m.inv, h->d_x3, h->stream);
    if (e == hipSuccess) e = d2h_xyz(h, dst);
    return e;
}

// sum of the first n entries of d_r2 -> host (deterministic)
hipError_t sum_r2(vbd_solver* h, int n, double* out) {
    hipError_t e = launch_sum_fixed(n, h->d_r2, h->d_sum, h->stream, r2_stage(h));
    if (e == hipSuccess) e = hipMemcpyAsync(out, h->d_sum, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e;
}

float step_mu(const vbd_solver* h) { return h->paramsSet ? h->params.mu : 0.f; }
float step_lambda(const vbd_solver* h) { return h->paramsSet ? h->params.lambda : 0.f; }
float step_sdt(const vbd_solver* h) { return inv_h2_ref(h->params.dt / (float)h->params.substeps); }

hipError_t forward_residual(vbd_solver* h, double* norm2) {
    const ContactArgs ct = contact_of(h->params);
    hipError_t e = launch_pass(3, h->m, 0, h->m.nFree, pass_lpv(h), 256, nullptr, nullptr, h->d_out, h->d_r2, nullptr,
                               step_mu(h), step_lambda(h), step_sdt(h), ct, h->stream);
    if (e == hipSuccess) e = sum_r2(h, h->m.nFree, norm2);
    return e;
}

// Enqueues |r|^2 of the per-vertex residual already in d_r2 into the pinned h_res[0] (no wait).
hipError_t enqueue_sum_to_host(vbd_solver* h) {
    hipError_t e = launch_sum_fixed(h->m.nFree, h->d_r2, h->d_sum, h->stream, r2_stage(h));
    if (e == hipSuccess) e = hipMemcpyAsync(h->h_res, h->d_sum, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    return e;
}

// ---- FP64 differentiation state ----
bool fp64(const vbd_solver* h) { return h->diffBits == 64; }

double step_sdt64(const vbd_solver* h) {
    const double hs = (double)(h->params.dt / (float)h->params.substeps);
    return 1.0 / (hs * hs);
}

// The FP64 arrays exist and hold the current state (promoted from the FP32 state when it changed).
hipError_t ensure_fp64(vbd_solver* h) {
    hipError_t e = ensure_scratch(h);
    if (e == hipSuccess) e = alloc_dev64(h->d64, h->m.nV);
    if (e == hipSuccess && !h->d_x3d) e = hipMalloc(&h->d_x3d, sizeof(double) * 3 * (size_t)(h->m.nV > 0 ? h->m.nV : 1));
    if (e != hipSuccess || h->fp64Fresh) return e;
    const int n = h->m.nV;
    if ((e = launch_f4_to_d4(n, h->m.pos, h->d64.pos, h->stream)) != hipSuccess) return e;
    if ((e = launch_f4_to_d4(n, h->m.inert, h->d64.inert, h->stream)) != hipSuccess) return e;
    if ((e = launch_f4_to_d4(n, h->m.vel, h->d64.vel, h->stream)) != hipSuccess) return e;
    if ((e = launch_f4_to_d4(n, h->m.xprev, h->d64.xprev, h->stream)) != hipSuccess) return e;
    h->fp64Fresh = true;
    return e;
}

// The FP32 state as the rounding of the FP64 state.
hipError_t round_fp64(vbd_solver* h) {
    const int n = h->m.nV;
    hipError_t e = launch_d4_to_f4(n, h->d64.pos, h->m.pos, h->stream);
    if (e == hipSuccess) e = launch_d4_to_f4(n, h->d64.inert, h->m.inert, h->stream);
    if (e == hipSuccess) e = launch_d4_to_f4(n, h->d64.vel, h->m.vel, h->stream);
    if (e == hipSuccess) e = launch_d4_to_f4(n, h->d64.xprev, h->m.xprev, h->stream);
    return e;
}

// host 3N floats (caller order) -> slot double4 buffer, pinned slots zeroed
hipError_t upload_vec64(vbd_solver* h, const float* src, d4* dst) {
    hipError_t e = h2d_xyz(h, src);
    if (e == hipSuccess) e = launch_scatter_fd(h->m.nV, h->d_x3, h->m.inv, dst, h->stream);
    if (e == hipSuccess && h->m.nV > h->m.nFree)
        e = hipMemsetAsync(dst + h->m.nFree, 0, sizeof(d4) * (size_t)(h->m.nV - h->m.nFree), h->stream);
    return e;
}

hipError_t download_vec64(vbd_solver* h, const d4* src, float* dst) {
    hipError_t e = launch_gather_df(h->m.nV, src, h->m.inv, h->d_x3, h->stream);
    if (e == hipSuccess) e = d2h_xyz(h, dst);
    return e;
}

// host 3N doubles (caller order) <-> slot double4 buffer (pinned slots zeroed on upload)
hipError_t upload_vec64d(vbd_solver* h, const double* src, d4* dst) {
    const size_t bytes = sizeof(double) * 3 * (size_t)h->m.nV;
    hipError_t e = bytes ? hipMemcpyAsync(h->d_x3d, src, bytes, hipMemcpyHostToDevice, h->stream) : hipSuccess;
    if (e == hipSuccess) e = launch_scatter_d(h->m.nV, h->d_x3d, h->m.inv, dst, h->stream);
    if (e == hipSuccess && h->m.nV > h->m.nFree)
        e = hipMemsetAsync(dst + h->m.nFree, 0, sizeof(d4) * (size_t)(h->m.nV - h->m.nFree), h->stream);
    return e;
}

hipError_t download_vec64d(vbd_solver* h, const d4* src, double* dst) {
    const size_t bytes = sizeof(double) * 3 * (size_t)h->m.nV;
    hipError_t e = launch_gather_d(h->m.nV, src, h->m.inv, h->d_x3d, h->stream);
    if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, h->d_x3d, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e;
}

hipError_t forward_residual64(vbd_solver* h, double* norm2) {
    hipError_t e = launch_pass64(1, h->m, h->d64, 0, h->m.nFree, nullptr, nullptr, h->d64.out, h->d_r2, nullptr,
                                 step_mu(h), step_lambda(h), step_sdt64(h), contact_of(h->params), h->stream);
    if (e == hipSuccess) e = sum_r2(h, h->m.nFree, norm2);
    return e;
}


}  // namespace

extern "C" {

int vbd_set_materials(vbd_solver* h, const float* tet_mu, const float* tet_lambda) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!h->m.perTetMat) return fail(h, VBD_ERR_STATE, "mesh was uploaded without per-tet materials");
    if ((!tet_mu || !tet_lambda) && h->m.nT) return fail(h, VBD_ERR_ARG, "null material array");
    if (!finite_array(tet_mu, h->m.nT) || !finite_array(tet_lambda, h->m.nT))
        return fail(h, VBD_ERR_NUMERIC, "non-finite material");
    DeviceGuard g(h->device);
    HIP_CHECK(h, hipMemcpyAsync(h->m.tmu, tet_mu, sizeof(float) * h->m.nT, hipMemcpyHostToDevice, h->stream));
    HIP_CHECK(h, hipMemcpyAsync(h->m.tlam, tet_lambda, sizeof(float) * h->m.nT, hipMemcpyHostToDevice, h->stream));
    HIP_CHECK(h, launch_update_pmat(h->m, h->stream));
    // the tiled kernel reads the padded copy of the records of its tile cut
    HIP_CHECK(h, launch_pad_pmat(h->m, h->stream));
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_residual_norm(vbd_solver* h, double* norm, double* gravity_load) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    DeviceGuard g(h->device);
    HIP_CHECK(h, fp64(h) ? ensure_fp64(h) : ensure_scratch(h));
    double r2 = 0, l2 = 0;
    HIP_CHECK(h, fp64(h) ? forward_residual64(h, &r2) : forward_residual(h, &r2));
    HIP_CHECK(h, launch_gravity_load2(h->m, h->params.gravity, h->d_r2, h->stream));
    HIP_CHECK(h, sum_r2(h, h->m.nFree, &l2));
    if (norm) *norm = std::sqrt(r2);
    if (gravity_load) *gravity_load = std::sqrt(l2);
    return VBD_OK;
}

static hipError_t cg64_run(vbd_solver* h, double bn, int max_it, int check_every, double tol, int* iters,
                           bool* broke);

// Newton method of the converged step: CG tolerance relative to |grad G| (a CPU prototype on the
// 16^3 workload: 1e-1 -> 7.7 Newton steps and 190 CG iterations per step, 1e-2 -> 4.1 and 196,
// 1e-3 -> 3.0 and 208), and the steps before it hands the state to VBD sweeps.
constexpr double NEWTON_ETA = 1e-2;
constexpr int NEWTON_MAX_STEPS = 20;

int vbd_step_converged(vbd_solver* h, int32_t max_iterations, int32_t check_every, float abs_tol, float rel_tol,
                       int32_t stall_checks, int32_t* iterations, double* final_norm) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    if (max_iterations < 1 || check_every < 1) return fail(h, VBD_ERR_ARG, "max_iterations and check_every must be >= 1");
    if (h->params.substeps != 1) return fail(h, VBD_ERR_ARG, "converged steps support substeps == 1");
    if (h->convergedMethod == VBD_CONVERGED_NEWTON && !fp64(h))
        return fail(h, VBD_ERR_STATE, "the Newton method of the converged step needs precision 64");
    DeviceGuard g(h->device);
    const bool f64 = fp64(h);
    HIP_CHECK(h, f64 ? ensure_fp64(h) : ensure_scratch(h));
    if (!f64) h->fp64Fresh = false;  // the FP32 state changes
    const vbd_step_params& p = h->params;
    const float hstep = p.dt;
    const float sdt = inv_h2_ref(hstep);
    const double sdt64 = step_sdt64(h);
    const ContactArgs ct = contact_of(p);
    double l2 = 0;
    HIP_CHECK(h, launch_gravity_load2(h->m, p.gravity, h->d_r2, h->stream));
    HIP_CHECK(h, sum_r2(h, h->m.nFree, &l2));
    const double tol = std::max((double)abs_tol, (double)rel_tol * std::sqrt(l2));
    if (f64) {
        HIP_CHECK(h, launch_predict64(h->m, h->d64, (double)hstep, p.gravity, h->stream));
        if (p.start_from_inertia) HIP_CHECK(h, launch_copy_d4(h->m.nFree, h->d64.inert, h->d64.pos, h->stream));
    } else {
        HIP_CHECK(h, launch_predict(h->m, hstep, p.gravity, h->stream));
        if (p.start_from_inertia) HIP_CHECK(h, launch_copy_inertia(h->m, h->stream));
    }
    double r2 = 0;
    int newton = 0, inner = 0;
    bool sweeps = true;
    if (h->convergedMethod == VBD_CONVERGED_NEWTON) {
        // Newton on G: (M/h^2 + K(x)) dx = -grad G by block-Jacobi CG to NEWTON_ETA |grad G|, then
        // x += alpha dx with alpha halved (at most 4 times) while |grad G| does not fall.  A CG
        // breakdown, a failed line search or NEWTON_MAX_STEPS steps hand the state to the VBD
        // sweeps below.
        const int n = h->m.nFree;
        Dev64& d = h->d64;
        auto residual = [&](double* r) -> hipError_t {
            hipError_t e = launch_pass64(1, h->m, d, 0, n, nullptr, nullptr, d.out, h->d_r2, nullptr, p.mu, p.lambda,
                                         sdt64, ct, h->stream);
            double q = 0.0;
            if (e == hipSuccess) e = sum_r2(h, n, &q);
            *r = std::sqrt(q);
            return e;
        };
        double r = 0.0;
        HIP_CHECK(h, residual(&r));
        sweeps = false;
        while (r > tol && std::isfinite(r)) {
            if (newton >= NEWTON_MAX_STEPS) {
                sweeps = true;
                break;
            }
            HIP_CHECK(h, launch_scale_d4(n, d.out, d.rhs, -1.0, h->stream));
            HIP_CHECK(h, hipMemsetAsync(d.w, 0, sizeof(d4) * (size_t)h->m.nV, h->stream));
            int its = 0;
            bool broke = false;
            HIP_CHECK(h, cg64_run(h, r, max_iterations, check_every, NEWTON_ETA * r, &its, &broke));
            inner += its;
            ++newton;
            if (broke) {
                sweeps = true;
                break;
            }
            HIP_CHECK(h, launch_copy_d4(n, d.pos, d.save[0], h->stream));
            double alpha = 1.0, rn = 0.0;
            bool ok = false;
            for (int ls = 0; ls < 5 && !ok; ++ls, alpha *= 0.5) {
                HIP_CHECK(h, launch_newton_move64(n, d.pos, d.save[0], d.w, alpha, h->stream));
                HIP_CHECK(h, residual(&rn));
                ok = std::isfinite(rn) && rn < r;
            }
            if (!ok) {  // back to the start of this Newton step
                HIP_CHECK(h, launch_copy_d4(n, d.save[0], d.pos, h->stream));
                sweeps = true;
                break;
            }
            r = rn;
        }
        if (!std::isfinite(r)) {
            h->lastNewton = newton;
            h->lastInner = inner;
            h->lastSweeps = 0;
            if (iterations) *iterations = newton;
            if (final_norm) *final_norm = r;
            return fail(h, VBD_ERR_NUMERIC, "converged step: the residual became non-finite");
        }
        r2 = r * r;
    }
    float omega = 1.0f;
    int it = 0, stalled = 0;
    double best = 1e300;
    const bool accel = p.use_acceleration != 0;
    std::vector<float> om;
    while (sweeps && it < max_iterations) {
        // one block: the iterations up to the next check, then the residual
        const int n = std::min(check_every - it % check_every, max_iterations - it);
        om.assign(n, 1.0f);
        for (int j = 0; j < n && accel; ++j) om[j] = omega = omega_ref(it + j + 1, p.acceleration_rho, omega);
        const int it0 = it;
        HIP_CHECK(h, run_block(h, block_key(f64 ? 'D' : 'F', it0, om, 0.f), [&]() -> hipError_t {
            for (int j = 0; j < n; ++j) {
                const int k = it0 + j;
                hipError_t e = hipSuccess;
                if (f64) {
                    for (int c = 0; c < h->m.nColors && e == hipSuccess; ++c) {
                        const int s0 = h->m.colorSlotOff[c], cnt = h->m.colorSlotOff[c + 1] - s0;
                        e = launch_pass64(0, h->m, h->d64, s0, cnt, nullptr, nullptr, nullptr, nullptr,
                                          accel ? h->d64.save[k & 1] : nullptr, p.mu, p.lambda, sdt64, ct, h->stream);
                    }
                    if (e == hipSuccess && accel && om[j] > 1.0f)
                        e = launch_blend64(h->m.nFree, h->d64.pos, h->d64.save[(k + 1) & 1], om[j], h->stream);
                } else {
                    e = enqueue_sweep(h, sdt, p.mu, p.lambda, ct, accel ? h->m.save[k & 1] : nullptr);
                    if (e == hipSuccess && accel && om[j] > 1.0f)
                        e = launch_blend(h->m, om[j], h->m.save[(k + 1) & 1], h->stream);
                }
                if (e != hipSuccess) return e;
            }
            hipError_t e = f64 ? launch_pass64(1, h->m, h->d64, 0, h->m.nFree, nullptr, nullptr, h->d64.out, h->d_r2,
                                               nullptr, p.mu, p.lambda, sdt64, ct, h->stream)
                               : launch_pass(3, h->m, 0, h->m.nFree, pass_lpv(h), 256, nullptr, nullptr, h->d_out,
                                             h->d_r2, nullptr, p.mu, p.lambda, sdt, ct, h->stream);
            return e == hipSuccess ? enqueue_sum_to_host(h) : e;
        }));
        it += n;
        r2 = h->h_res[0];
        const double r = std::sqrt(r2);
        if (!std::isfinite(r)) {
            if (iterations) *iterations = it;
            if (final_norm) *final_norm = r;
            return fail(h, VBD_ERR_NUMERIC, "converged step: the residual became non-finite");
        }
        if (r <= tol) break;
        // stagnation at the float32 floor: no 1% gain on the best residual for stall_checks checks
        if (r < 0.99 * best) {
            best = r;
            stalled = 0;
        } else if (stall_checks > 0 && ++stalled >= stall_checks) {
            break;
        }
    }
    if (f64) {
        HIP_CHECK(h, launch_velocity64(h->m, h->d64, (double)hstep, (double)p.velocity_damping, h->stream));
        HIP_CHECK(h, round_fp64(h));
    } else {
        HIP_CHECK(h, launch_velocity(h->m, hstep, p.velocity_damping, h->stream));
    }
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    h->lastNewton = newton;
    h->lastInner = inner;
    h->lastSweeps = it;
    if (iterations) *iterations = newton + it;
    if (final_norm) *final_norm = std::sqrt(r2);
    return VBD_OK;
}

int vbd_set_converged_method(vbd_solver* h, int32_t method) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (method != VBD_CONVERGED_VBD && method != VBD_CONVERGED_NEWTON)
        return fail(h, VBD_ERR_ARG, "method must be VBD_CONVERGED_VBD or VBD_CONVERGED_NEWTON");
    h->convergedMethod = method;
    return VBD_OK;
}

int vbd_last_step_stats(const vbd_solver* h, int32_t* newton_steps, int32_t* inner_iterations, int32_t* sweeps) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (newton_steps) *newton_steps = h->lastNewton;
    if (inner_iterations) *inner_iterations = h->lastInner;
    if (sweeps) *sweeps = h->lastSweeps;
    return VBD_OK;
}

static int hessian_vector_core(vbd_solver* h, const float* w, float* out, const double* w_d, double* out_d) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    if (((!w && !w_d) || (!out && !out_d)) && h->m.nV) return fail(h, VBD_ERR_ARG, "null vector");
    DeviceGuard g(h->device);
    if (fp64(h)) {
        HIP_CHECK(h, ensure_fp64(h));
        HIP_CHECK(h, w_d ? upload_vec64d(h, w_d, h->d64.w) : upload_vec64(h, w, h->d64.w));
        HIP_CHECK(h, hipMemsetAsync(h->d64.out, 0, sizeof(d4) * (size_t)(h->m.nV > 0 ? h->m.nV : 1), h->stream));
        HIP_CHECK(h, launch_pass64(4, h->m, h->d64, 0, h->m.nFree, h->d64.w, nullptr, h->d64.out, nullptr, nullptr,
                                   step_mu(h), step_lambda(h), step_sdt64(h), contact_of(h->params), h->stream));
        HIP_CHECK(h, out_d ? download_vec64d(h, h->d64.out, out_d) : download_vec64(h, h->d64.out, out));
        return VBD_OK;
    }
    HIP_CHECK(h, ensure_scratch(h));
    HIP_CHECK(h, upload_vec(h, w, h->d_w, true));
    HIP_CHECK(h, hipMemsetAsync(h->d_out, 0, sizeof(f4) * (size_t)(h->m.nV > 0 ? h->m.nV : 1), h->stream));
    HIP_CHECK(h, launch_pass(2, h->m, 0, h->m.nFree, pass_lpv(h), 256, h->d_w, nullptr, h->d_out, nullptr, nullptr,
                             step_mu(h), step_lambda(h), step_sdt(h), contact_of(h->params), h->stream));
    HIP_CHECK(h, download_vec(h, h->d_out, out));
    return VBD_OK;
}

int vbd_hessian_vector(vbd_solver* h, const float* w, float* out) { return hessian_vector_core(h, w, out, nullptr, nullptr); }

int vbd_hessian_vector_f64(vbd_solver* h, const double* w, double* out) {
    if (h && !fp64(h)) return fail(h, VBD_ERR_STATE, "the FP64 vectors need vbd_set_differentiation_precision(h, 64)");
    return hessian_vector_core(h, nullptr, nullptr, w, out);
}

static int adjoint_solve_core(vbd_solver* h, const float* rhs, float* w, const double* rhs_d, double* w_d,
                              int32_t max_iterations, int32_t check_every, float rel_tol, float rho,
                              int32_t stall_checks, int32_t* iterations, double* rel_residual) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    if (((!w && !w_d) || (!rhs && !rhs_d)) && h->m.nV) return fail(h, VBD_ERR_ARG, "null vector");
    if (max_iterations < 1 || check_every < 1) return fail(h, VBD_ERR_ARG, "max_iterations and check_every must be >= 1");
    if (!(rho >= 0.f && rho < 1.f)) return fail(h, VBD_ERR_ARG, "rho must be in [0, 1)");
    DeviceGuard g(h->device);
    const bool f64 = fp64(h);
    HIP_CHECK(h, f64 ? ensure_fp64(h) : ensure_scratch(h));
    const ContactArgs ct = contact_of(h->params);
    const float mu = step_mu(h), lam = step_lambda(h), sdt = step_sdt(h);
    const double sdt64 = step_sdt64(h);
    const int lpv = pass_lpv(h);
    double b2 = 0, r2 = 0;
    if (f64) {
        HIP_CHECK(h, rhs_d ? upload_vec64d(h, rhs_d, h->d64.rhs) : upload_vec64(h, rhs, h->d64.rhs));
        HIP_CHECK(h, w_d ? upload_vec64d(h, w_d, h->d64.w) : upload_vec64(h, w, h->d64.w));
        HIP_CHECK(h, launch_norm2_d4(h->m.nFree, h->d64.rhs, h->d_r2, h->stream));
    } else {
        HIP_CHECK(h, upload_vec(h, rhs, h->d_rhs, true));
        HIP_CHECK(h, upload_vec(h, w, h->d_w, true));
        HIP_CHECK(h, launch_norm2_f4(h->m.nFree, h->d_rhs, h->d_r2, h->stream));
    }
    HIP_CHECK(h, sum_r2(h, h->m.nFree, &b2));
    const double bn = std::sqrt(b2);
    float omega = 1.0f;
    int it = 0, stalled = 0;
    double best = 1e300;
    const bool accel = rho > 0.f;
    std::vector<float> om;
    double first = 0.0;  // residual at the first check
    while (it < max_iterations && bn > 0) {
        const int n = std::min(check_every - it % check_every, max_iterations - it);
        om.assign(n, 1.0f);
        for (int j = 0; j < n && accel; ++j) om[j] = omega = omega_ref(it + j + 1, rho, omega);
        const int it0 = it;
        HIP_CHECK(h, run_block(h, block_key(f64 ? 'E' : 'A', it0, om, rho), [&]() -> hipError_t {
            for (int j = 0; j < n; ++j) {
                const int k = it0 + j;
                for (int c = 0; c < h->m.nColors; ++c) {
                    const int s0 = h->m.colorSlotOff[c], cnt = h->m.colorSlotOff[c + 1] - s0;
                    hipError_t e = f64 ? launch_pass64(2, h->m, h->d64, s0, cnt, h->d64.w, h->d64.rhs, nullptr, nullptr,
                                                       accel ? h->d64.wsave[k & 1] : nullptr, mu, lam, sdt64, ct,
                                                       h->stream)
                                       : launch_pass(0, h->m, s0, cnt, lpv, 256, h->d_w, h->d_rhs, nullptr, nullptr,
                                                     accel ? h->d_wsave[k & 1] : nullptr, mu, lam, sdt, ct, h->stream);
                    if (e != hipSuccess) return e;
                }
                if (accel && om[j] > 1.0f) {
                    hipError_t e = f64 ? launch_blend64(h->m.nFree, h->d64.w, h->d64.wsave[(k + 1) & 1], om[j], h->stream)
                                       : launch_blend_vec(h->m.nFree, h->d_w, h->d_wsave[(k + 1) & 1], om[j], h->stream);
                    if (e != hipSuccess) return e;
                }
            }
            hipError_t e = f64 ? launch_pass64(3, h->m, h->d64, 0, h->m.nFree, h->d64.w, h->d64.rhs, h->d64.out, h->d_r2,
                                               nullptr, mu, lam, sdt64, ct, h->stream)
                               : launch_pass(1, h->m, 0, h->m.nFree, lpv, 256, h->d_w, h->d_rhs, h->d_out, h->d_r2,
                                             nullptr, mu, lam, sdt, ct, h->stream);
            return e == hipSuccess ? enqueue_sum_to_host(h) : e;
        }));
        it += n;
        r2 = h->h_res[0];
        const double r = std::sqrt(r2);
        if (r <= rel_tol * bn) break;
        // Chebyshev with rho above the iteration's spectral radius diverges, fast (the residual
        // explodes) or slowly (it creeps up until the stall test would end the solve): report both
        const bool blowup = !std::isfinite(r) || r > 1e3 * std::max(bn, first) || (accel && r > 1e2 * best);
        const bool stall = !(r < 0.99 * best) && stall_checks > 0 && stalled + 1 >= stall_checks;
        if (blowup || (stall && accel && r > 10.0 * best)) {
            if (iterations) *iterations = it;
            if (rel_residual) *rel_residual = r / bn;
            return fail(h, VBD_ERR_NUMERIC,
                        accel ? "adjoint solve diverged (Chebyshev rho too large); w is unchanged"
                              : "adjoint solve diverged without acceleration (M/h^2 + K is not positive definite at "
                                "this state: inverted or degenerate elements, or a step that did not converge); w is "
                                "unchanged");
        }
        if (first == 0.0) first = r;
        if (r < 0.99 * best) {
            best = r;
            stalled = 0;
        } else if (stall_checks > 0 && ++stalled >= stall_checks) {
            break;
        }
    }
    HIP_CHECK(h, w_d ? download_vec64d(h, h->d64.w, w_d)
                     : (f64 ? download_vec64(h, h->d64.w, w) : download_vec(h, h->d_w, w)));
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = bn > 0 ? std::sqrt(r2) / bn : 0.0;
    return VBD_OK;
}

int vbd_adjoint_solve(vbd_solver* h, const float* rhs, float* w, int32_t max_iterations, int32_t check_every,
                      float rel_tol, float rho, int32_t stall_checks, int32_t* iterations, double* rel_residual) {
    return adjoint_solve_core(h, rhs, w, nullptr, nullptr, max_iterations, check_every, rel_tol, rho, stall_checks,
                              iterations, rel_residual);
}

int vbd_adjoint_solve_f64(vbd_solver* h, const double* rhs, double* w, int32_t max_iterations, int32_t check_every,
                          float rel_tol, float rho, int32_t stall_checks, int32_t* iterations, double* rel_residual) {
    if (h && !fp64(h)) return fail(h, VBD_ERR_STATE, "the FP64 vectors need vbd_set_differentiation_precision(h, 64)");
    return adjoint_solve_core(h, nullptr, nullptr, rhs, w, max_iterations, check_every, rel_tol, rho, stall_checks,
                              iterations, rel_residual);
}

// Conjugate gradients with the block-Jacobi preconditioner on A = M/h^2 + K (+ contact) at the
// FP64 state, in the single-reduction form (the kernels' comment in vbd_fp64.hip): rhs in
// d64.rhs (norm bn), x in d64.w (its start on entry).  Per iteration three launches: the update
// of the directions, x, r and z with r.z and r.r per slot; w = A z with w.z per slot; one launch
// summing the three (fixed order).  Blocks of check_every iterations replay one cached graph,
// keyed by the block length, the parity of its first iteration (which selects the scalar slots)
// and whether it starts the solve; the host reads r.r and the breakdown flag after each block.
// Stops when |r| <= tol or after max_it iterations; *broke when gamma or the alpha denominator
// was not positive (A not positive definite) or |r| grew 1e3-fold.
static hipError_t cg64_run(vbd_solver* h, double bn, int max_it, int check_every, double tol, int* iters,
                           bool* broke) {
    const ContactArgs ct = contact_of(h->params);
    const float mu = step_mu(h), lam = step_lambda(h);
    const double sdt64 = step_sdt64(h);
    const int n = h->m.nFree;
    Dev64& d = h->d64;
    *iters = 0;
    *broke = false;
    if (!(bn > 0) || n == 0) return hipSuccess;
    hipError_t e = alloc_cg64(d, h->m.nV);
    if (e != hipSuccess) return e;
    // per-slot products: r.z, w.z, r.r (each with SUM_STAGE doubles of two-stage scratch after it)
    const size_t stride = (size_t)h->m.nV + SUM_STAGE;
    double* pRz = h->d_r2;
    double* pWz = d.cg_part;
    double* pRr = d.cg_part + stride;
    auto sums = [&](int gammaSlot) -> hipError_t {
        if (n <= SUM_TWO_STAGE)
            return launch_sum3_fixed(n, pRz, d.cg_scal + gammaSlot, pWz, d.cg_scal + CG_D, pRr, d.cg_scal + CG_RR,
                                     h->stream);
        hipError_t f = launch_sum_fixed(n, pRz, d.cg_scal + gammaSlot, h->stream, pRz + n);
        if (f == hipSuccess) f = launch_sum_fixed(n, pWz, d.cg_scal + CG_D, h->stream, pWz + n);
        if (f == hipSuccess) f = launch_sum_fixed(n, pRr, d.cg_scal + CG_RR, h->stream, pRr + n);
        return f;
    };
    // the products A z read z at pinned neighbours: zero there (the iterations write free slots only)
    if (h->m.nV > n) e = hipMemsetAsync(d.cg_z + n, 0, sizeof(d4) * (size_t)(h->m.nV - n), h->stream);
    // D^-1, r = rhs - A x, z = D^-1 r, p = s = 0, w = A z, gamma_0, delta_0
    if (e == hipSuccess)
        e = launch_pass64(5, h->m, d, 0, n, nullptr, nullptr, d.cg_da, nullptr, d.cg_db, mu, lam, sdt64, ct, h->stream);
    if (e == hipSuccess)
        e = launch_pass64(3, h->m, d, 0, n, d.w, d.rhs, d.cg_r, pRz, nullptr, mu, lam, sdt64, ct, h->stream);
    double r02 = 0.0;
    if (e == hipSuccess) e = sum_r2(h, n, &r02);
    if (e != hipSuccess) return e;
    const double first = std::sqrt(r02);  // a start that already meets the tolerance is returned as is
    if (first <= tol) return hipSuccess;
    e = launch_cg_init64(n, d, pRz, pRr, h->stream);
    if (e == hipSuccess)
        e = launch_pass64(4, h->m, d, 0, n, d.cg_z, nullptr, d.out, pWz, nullptr, mu, lam, sdt64, ct, h->stream);
    if (e == hipSuccess) e = sums(CG_G);
    if (e != hipSuccess) return e;
    int it = 0;
    while (it < max_it) {
        const int nb = std::min(check_every - it % check_every, max_it - it);
        const int it0 = it;
        e = run_block(h, block_key(it0 == 0 ? 'H' : 'G', it0, std::vector<float>(nb, 1.0f), 0.f), [&]() -> hipError_t {
            hipError_t f = hipSuccess;
            for (int j = 0; j < nb && f == hipSuccess; ++j) {
                const int k = it0 + j;
                f = launch_cg_update64(n, k, d, d.w, d.out, pRz, pRr, h->stream);
                if (f == hipSuccess)
                    f = launch_pass64(4, h->m, d, 0, n, d.cg_z, nullptr, d.out, pWz, nullptr, mu, lam, sdt64, ct,
                                      h->stream);
                if (f == hipSuccess) f = sums(CG_G + ((k + 1) & 1));
            }
            if (f == hipSuccess)
                f = hipMemcpyAsync(h->h_res, d.cg_scal + CG_RR, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream);
            return f;
        });
        if (e != hipSuccess) return e;
        it += nb;
        *iters = it;
        const double r = std::sqrt(h->h_res[0]), bad = h->h_res[1];  // CG_BAD follows CG_RR
        if (bad != 0.0 || !std::isfinite(r) || r > 1e3 * std::max(bn, first)) {
            *broke = true;
            return hipSuccess;
        }
        if (r <= tol) break;
    }
    return hipSuccess;
}

// The adjoint system solved by cg64_run from the caller's rhs and start; the reported residual
// is recomputed from rhs - A w (the recurrence's r drifts from it in rounding).
int vbd_adjoint_solve_cg_f64(vbd_solver* h, const double* rhs, double* w, int32_t max_iterations,
                             int32_t check_every, float rel_tol, int32_t* iterations, double* rel_residual) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!fp64(h)) return fail(h, VBD_ERR_STATE, "the FP64 vectors need vbd_set_differentiation_precision(h, 64)");
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if (!h->paramsSet) return fail(h, VBD_ERR_STATE, "vbd_set_params has not been called");
    if ((!w || !rhs) && h->m.nV) return fail(h, VBD_ERR_ARG, "null vector");
    if (max_iterations < 1 || check_every < 1) return fail(h, VBD_ERR_ARG, "max_iterations and check_every must be >= 1");
    if (!(rel_tol >= 0.f)) return fail(h, VBD_ERR_ARG, "rel_tol must be >= 0");
    DeviceGuard g(h->device);
    HIP_CHECK(h, ensure_fp64(h));
    const int n = h->m.nFree;
    Dev64& d = h->d64;
    HIP_CHECK(h, upload_vec64d(h, rhs, d.rhs));
    HIP_CHECK(h, upload_vec64d(h, w, d.w));
    double b2 = 0.0;
    HIP_CHECK(h, launch_norm2_d4(n, d.rhs, h->d_r2, h->stream));
    HIP_CHECK(h, sum_r2(h, n, &b2));
    const double bn = std::sqrt(b2);
    int it = 0;
    bool broke = false;
    HIP_CHECK(h, cg64_run(h, bn, max_iterations, check_every, (double)rel_tol * bn, &it, &broke));
    if (broke) {
        if (iterations) *iterations = it;
        if (rel_residual) *rel_residual = std::sqrt(h->h_res[0]) / bn;
        return fail(h, VBD_ERR_NUMERIC,
                    "conjugate gradients broke down (A is not positive definite at this state); w is unchanged");
    }
    double r2 = 0.0;
    if (bn > 0 && n > 0) {
        HIP_CHECK(h, launch_pass64(3, h->m, d, 0, n, d.w, d.rhs, d.out, h->d_r2, nullptr, step_mu(h), step_lambda(h),
                                   step_sdt64(h), contact_of(h->params), h->stream));
        HIP_CHECK(h, sum_r2(h, n, &r2));
    }
    HIP_CHECK(h, download_vec64d(h, d.w, w));
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = bn > 0 ? std::sqrt(r2) / bn : 0.0;
    return VBD_OK;
}

int vbd_parameter_gradients(vbd_solver* h, const float* w, float* dmu, float* dlambda) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if ((!w || !dmu || !dlambda) && h->m.nT) return fail(h, VBD_ERR_ARG, "null argument");
    DeviceGuard g(h->device);
    if (fp64(h)) {
        HIP_CHECK(h, ensure_fp64(h));
        HIP_CHECK(h, upload_vec64(h, w, h->d64.w));
        HIP_CHECK(h, launch_param_grad64(h->m, h->d64, h->d64.w, h->d_tgrad, h->d_tgrad + h->m.nT, h->stream));
    } else {
        HIP_CHECK(h, ensure_scratch(h));
        HIP_CHECK(h, upload_vec(h, w, h->d_w, true));
        HIP_CHECK(h, launch_param_grad(h->m, h->d_w, h->d_tgrad, h->d_tgrad + h->m.nT, h->stream));
    }
    if (h->m.nT) {
        HIP_CHECK(h, hipMemcpyAsync(dmu, h->d_tgrad, sizeof(float) * h->m.nT, hipMemcpyDeviceToHost, h->stream));
        HIP_CHECK(h, hipMemcpyAsync(dlambda, h->d_tgrad + h->m.nT, sizeof(float) * h->m.nT, hipMemcpyDeviceToHost, h->stream));
    }
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_parameter_gradients_f64(vbd_solver* h, const double* w, double* dmu, double* dlambda) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!fp64(h)) return fail(h, VBD_ERR_STATE, "the FP64 vectors need vbd_set_differentiation_precision(h, 64)");
    if (h->materialModel != VBD_MATERIAL_STABLE_NEO_HOOKEAN)
        return fail(h, VBD_ERR_STATE, "differentiation supports the stable Neo-Hookean material only");
    if ((!w || !dmu || !dlambda) && h->m.nT) return fail(h, VBD_ERR_ARG, "null argument");
    DeviceGuard g(h->device);
    HIP_CHECK(h, ensure_fp64(h));
    if (!h->d_tgrad64)
        HIP_CHECK(h, hipMalloc(&h->d_tgrad64, sizeof(double) * 2 * (size_t)(h->m.nT > 0 ? h->m.nT : 1)));
    HIP_CHECK(h, upload_vec64d(h, w, h->d64.w));
    HIP_CHECK(h, launch_param_grad64(h->m, h->d64, h->d64.w, h->d_tgrad64, h->d_tgrad64 + h->m.nT, h->stream));
    if (h->m.nT) {
        HIP_CHECK(h, hipMemcpyAsync(dmu, h->d_tgrad64, sizeof(double) * h->m.nT, hipMemcpyDeviceToHost, h->stream));
        HIP_CHECK(h, hipMemcpyAsync(dlambda, h->d_tgrad64 + h->m.nT, sizeof(double) * h->m.nT, hipMemcpyDeviceToHost,
                                    h->stream));
    }
    HIP_CHECK(h, hipStreamSynchronize(h->stream));
    return VBD_OK;
}

int vbd_set_differentiation_precision(vbd_solver* h, int32_t bits) {
    if (!h) return fail(nullptr, VBD_ERR_ARG, "null handle");
    if (bits != 32 && bits != 64) return fail(h, VBD_ERR_ARG, "precision must be 32 or 64 bits");
    if (bits == h->diffBits) return VBD_OK;
    DeviceGuard g(h->device);
    drop_graph(h);
    h->diffBits = bits;
    h->fp64Fresh = false;  // entering FP64 promotes the FP32 state; leaving keeps its rounding
    return VBD_OK;
}

int vbd_download_state_f64(vbd_solver* h, double* positions, double* velocities) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!fp64(h)) return fail(h, VBD_ERR_STATE, "the FP64 state needs vbd_set_differentiation_precision(h, 64)");
    if ((!positions || !velocities) && h->m.nV) return fail(h, VBD_ERR_ARG, "null output");
    DeviceGuard g(h->device);
    HIP_CHECK(h, ensure_fp64(h));
    const size_t bytes = sizeof(double) * 3 * (size_t)h->m.nV;
    double* outs[2] = {positions, velocities};
    const d4* srcs[2] = {h->d64.pos, h->d64.vel};
    for (int k = 0; k < 2 && h->m.nV; ++k) {
        HIP_CHECK(h, launch_gather_d(h->m.nV, srcs[k], h->m.inv, h->d_x3d, h->stream));
        HIP_CHECK(h, hipMemcpyAsync(outs[k], h->d_x3d, bytes, hipMemcpyDeviceToHost, h->stream));
        HIP_CHECK(h, hipStreamSynchronize(h->stream));
    }
    return VBD_OK;
}

int vbd_upload_state_f64(vbd_solver* h, const double* positions, const double* velocities) {
    int rc = check_ready(h);
    if (rc) return rc;
    if (!fp64(h)) return fail(h, VBD_ERR_STATE, "the FP64 state needs vbd_set_differentiation_precision(h, 64)");
    const size_t n3 = 3 * (size_t)h->m.nV;
    for (const double* a : {positions, velocities})
        if (a)
            for (size_t i = 0; i < n3; ++i)
                if (!std::isfinite(a[i])) return fail(h, VBD_ERR_NUMERIC, "non-finite value in the FP64 state");
    DeviceGuard g(h->device);
    HIP_CHECK(h, ensure_fp64(h));
    const size_t bytes = sizeof(double) * n3;
    const double* ins[2] = {positions, velocities};
    d4* dsts[2] = {h->d64.pos, h->d64.vel};
    f4* dst32[2] = {h->m.pos, h->m.vel};
    for (int k = 0; k < 2 && h->m.nV; ++k) {
        if (!ins[k]) continue;
        HIP_CHECK(h, hipMemcpyAsync(h->d_x3d, ins[k], bytes, hipMemcpyHostToDevice, h->stream));
        HIP_CHECK(h, launch_scatter_d(h->m.nV, h->d_x3d, h->m.inv, dsts[k], h->stream));
        HIP_CHECK(h, launch_d4_to_f4(h->m.nV, dsts[k], dst32[k], h->stream));
        HIP_CHECK(h, hipStreamSynchronize(h->stream));
    }
    return VBD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// host helpers (evaluated without fused multiply-adds, like the reference's scalar code)
// ---------------------------------------------------------------------------------------
#pragma clang fp contract(off)

// TetMesh::initialize (TetMesh.h:213-275) and computeRestVolume (TetMesh.h:291-298):
// V = |e0 . (e1 x e2)| / 6, mass += rho V / 4 per corner, adjacency appended in tet order.
int vbd_mesh_initialize(int32_t nV, const float* rest, int32_t nT, const int32_t* tets, float density, float* mass,
                        float* tet_rest_volume, int32_t* adj_offsets, int32_t* adj_tets) {
    if (nV < 0 || nT < 0 || (nV && (!rest || !mass || !adj_offsets)) || (nT && (!tets || !tet_rest_volume || !adj_tets)))
        return fail(nullptr, VBD_ERR_ARG, "vbd_mesh_initialize: bad argument");
    if (!adj_offsets) return fail(nullptr, VBD_ERR_ARG, "vbd_mesh_initialize: adj_offsets null");
    std::vector<int32_t> deg(nV, 0);
    for (int i = 0; i < nV; ++i) mass[i] = 0.0f;
    for (int t = 0; t < nT; ++t) {
        const int32_t* v = tets + 4 * (size_t)t;
        for (int j = 0; j < 4; ++j)
            if (v[j] < 0 || v[j] >= nV) return fail(nullptr, VBD_ERR_ARG, "vbd_mesh_initialize: index out of range");
        float e0[3], e1[3], e2[3];
        for (int c = 0; c < 3; ++c) {
            e0[c] = rest[3 * (size_t)v[1] + c] - rest[3 * (size_t)v[0] + c];
            e1[c] = rest[3 * (size_t)v[2] + c] - rest[3 * (size_t)v[0] + c];
            e2[c] = rest[3 * (size_t)v[3] + c] - rest[3 * (size_t)v[0] + c];
        }
        const float cx = e1[1] * e2[2] - e1[2] * e2[1];
        const float cy = e1[2] * e2[0] - e1[0] * e2[2];
        const float cz = e1[0] * e2[1] - e1[1] * e2[0];
        const float vol = std::fabs(e0[0] * cx + (e0[1] * cy + e0[2] * cz)) / 6.0f;
        tet_rest_volume[t] = vol;
        const float tm = vol * density;
        for (int j = 0; j < 4; ++j) {
            deg[v[j]]++;
            mass[v[j]] += tm / 4.0f;
        }
    }
    adj_offsets[0] = 0;
    for (int i = 0; i < nV; ++i) adj_offsets[i + 1] = adj_offsets[i] + deg[i];
    std::vector<int32_t> cur(adj_offsets, adj_offsets + nV);
    for (int t = 0; t < nT; ++t)
        for (int j = 0; j < 4; ++j) adj_tets[cur[tets[4 * (size_t)t + j]]++] = t;
    return VBD_OK;
}

// GraphColoring::colorVertices (GraphColoring.cpp:28-66): greedy first-fit in ascending
// vertex order over the graph of shared tets; groups list vertices in ascending order.  A
// per-color stamp array replaces the reference's per-vertex unordered_sets (same result).
int vbd_color_vertices(int32_t nV, const int32_t* tets, const int32_t* adj_offsets, const int32_t* adj_tets,
                       int32_t* color_of, int32_t* color_offsets, int32_t* color_verts, int32_t* num_colors) {
    if (nV < 0 || (nV && (!tets || !adj_offsets || !adj_tets || !color_of || !color_verts)) || !color_offsets ||
        !num_colors)
        return fail(nullptr, VBD_ERR_ARG, "vbd_color_vertices: bad argument");
    std::vector<int32_t> stamp;
    std::vector<int32_t> count;
    for (int v = 0; v < nV; ++v) color_of[v] = -1;
    for (int v = 0; v < nV; ++v) {
        for (int e = adj_offsets[v]; e < adj_offsets[v + 1]; ++e) {
            const int32_t* tet = tets + 4 * (size_t)adj_tets[e];
            for (int i = 0; i < 4; ++i) {
                const int u = tet[i];
                if (u == v) continue;
                const int c = color_of[u];
                if (c >= 0) stamp[c] = v;
            }
        }
        int c = 0;
        while (c < (int)stamp.size() && stamp[c] == v) ++c;
        if (c == (int)stamp.size()) {
            stamp.push_back(-1);
            count.push_back(0);
        }
        color_of[v] = c;
        count[c]++;
    }
    const int nc = (int)count.size();
    color_offsets[0] = 0;
    for (int c = 0; c < nc; ++c) color_offsets[c + 1] = color_offsets[c] + count[c];
    std::vector<int32_t> cur(color_offsets, color_offsets + nc);
    for (int v = 0; v < nV; ++v) color_verts[cur[color_of[v]]++] = v;
    *num_colors = nc;
    return VBD_OK;
}

// VBDsolver::precomputeTetData (VBD.cpp:36-58): Dm = [x1-x0, x2-x0, x3-x0] at rest,
// DmInv = adj(Dm)/det(Dm) (cofactor column 0 dotted with column 0 for det), V0 =
// |d3 . (d1 x d2)| / 6 (Energy.cpp:24-32).  DmInv is written row-major.
int vbd_precompute_tet_data(const float* rest, int32_t nT, const int32_t* tets, float* dminv, float* rest_volume) {
    if (nT < 0 || (nT && (!rest || !tets || !dminv || !rest_volume)))
        return fail(nullptr, VBD_ERR_ARG, "vbd_precompute_tet_data: bad argument");
    for (int t = 0; t < nT; ++t) {
        const int32_t* v = tets + 4 * (size_t)t;
        float A[3][3];
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c) A[c][j] = rest[3 * (size_t)v[j + 1] + c] - rest[3 * (size_t)v[0] + c];
        auto cof = [&](int i, int j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
        };
        const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
        const float det = c00 * A[0][0] + (c10 * A[1][0] + c20 * A[2][0]);
        const float inv = 1.0f / det;
        float* D = dminv + 9 * (size_t)t;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[3 * j + i] = cof(i, j) * inv;
        const float d1[3] = {A[0][0], A[1][0], A[2][0]}, d2[3] = {A[0][1], A[1][1], A[2][1]},
                    d3[3] = {A[0][2], A[1][2], A[2][2]};
        const float cx = d1[1] * d2[2] - d1[2] * d2[1];
        const float cy = d1[2] * d2[0] - d1[0] * d2[2];
        const float cz = d1[0] * d2[1] - d1[1] * d2[0];
        rest_volume[t] = std::fabs(d3[0] * cx + (d3[1] * cy + d3[2] * cz)) / 6.0f;
    }
    return VBD_OK;
}

float vbd_accelerator_omega(int32_t order, float rho, float prev_omega) { return omega_ref(order, rho, prev_omega); }

}  // extern "C"
