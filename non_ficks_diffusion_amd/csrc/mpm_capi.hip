// Host side of libmpm_gpu: scene context, device memory, launch sequence
// of the substep and the C ABI of include/mpm_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mpm_gpu.h"
#include "mpm_internal.h"
#include "mpm_kernels.h"

namespace {
thread_local std::string g_err;
}

int mpm_set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHECK(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

enum KClass {
    KC_BIN = 0, KC_SCAN, KC_SETUP, KC_REORDER, KC_P2G, KC_NORMALIZE, KC_INIT_VOLUME, KC_PFORCE, KC_GRID,
    KC_CG_FORCE, KC_CG_NODE, KC_G2P, KC_MAXSPEED, KC_COUNT
};
const char* kClassNames[KC_COUNT] = {"bin",          "scan",         "block_setup", "reorder", "p2g",
                                     "normalize",    "init_volume",  "constitutive_force", "grid_update",
                                     "cg_force",     "cg_node",      "g2p",         "max_speed"};

}  // namespace

struct mpm_context {
    mpm_scene_desc desc{};
    int D = 3;
    int64_t pcap = 0;
    int bcap = 0, nb = 0, nslots = 0, nparts = 0;
    int64_t gcap = 0;
    hipStream_t stream = nullptr;
    float* pbuf[2] = {nullptr, nullptr};
    int* idbuf[2] = {nullptr, nullptr};
    int* keybuf[2] = {nullptr, nullptr};
    int cur = 0;
    int* prank = nullptr;
    float* rcache = nullptr;
    float* pop = nullptr;
    int *slot_count = nullptr, *slot_touched = nullptr, *slot_active = nullptr;
    int *block_slot = nullptr, *block_pstart = nullptr, *block_nbr = nullptr, *scan_part = nullptr;
    int *cell_count = nullptr, *cell_start = nullptr;
    float *g_m = nullptr, *g_v = nullptr, *g_vs = nullptr, *g_q = nullptr, *g_s = nullptr, *g_r = nullptr;
    unsigned* g_flags = nullptr;
    mpm::Counters* ctr = nullptr;
    mpm::CGState* cg = nullptr;
    double* red = nullptr;
    unsigned* wg_part = nullptr;
    int cg_predict = 0;
    float last_dt = 0.0f;  // dt of the last substep (operator timing)  // iterations launched before the done flag was seen, last solve
    float* tiles = nullptr;
    int* h_done = nullptr;  // pinned
    hipEvent_t done_ev[2] = {nullptr, nullptr};
    int grid_particle = 0, grid_block = 0, grid_node = 0;
    bool have_particles = false, waiting = false, initialized = false;
    bool binned = false;  // particles sit in the block ranges of the last binning
    std::vector<void*> allocations;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_class;
    size_t ev_used = 0;
    double ms[KC_COUNT] = {};
    int64_t launches[KC_COUNT] = {};
};

namespace {

template <class T>
int dalloc(mpm_context* c, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    e = hipMemset(q, 0, bytes);
    if (e != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    c->allocations.push_back(q);
    *p = (T*)q;
    return MPM_OK;
}

int nfields(int D) { return 2 * D + D * D + 3; }

mpm::KP make_kp(const mpm_context* c, float dt) {
    mpm::KP p{};
    const mpm_scene_desc& d = c->desc;
    p.counts = d.counts;
    p.nb = c->nb;
    p.nslots = c->nslots;
    p.dx = 1.0f / (float)d.counts;
    p.inv_dx = (float)d.counts;
    p.mat.mu0 = d.material.mu0;
    p.mat.lambda0 = d.material.lambda0;
    p.mat.sy = d.material.stretching_yield;
    p.mat.cy = d.material.compression_yield;
    p.mat.xi = d.material.hardening_factor;
    p.mat.hmax = d.material.hardening_max;
    p.mat.plastic = d.material.plastic;
    for (int a = 0; a < 3; ++a) p.gravity[a] = d.gravity[a];
    p.flip = d.flip;
    p.nplanes = d.nplanes;
    for (int i = 0; i < d.nplanes; ++i) {
        for (int a = 0; a < 3; ++a) {
            p.pn[i][a] = d.planes[i].normal[a];
            p.ps[i][a] = d.planes[i].surface[a];
        }
        p.pmu[i] = d.planes[i].mu;
    }
    p.crop_ghost_x = d.crop_ghost_x;
    p.project_mode = d.project_mode;
    p.cg_tol = (double)d.cg_tolerance;
    p.cg_max = d.cg_max_iterations;
    p.cg_min = d.cg_min_iterations;
    p.pcur = c->pbuf[c->cur];
    p.palt = c->pbuf[c->cur ^ 1];
    p.idcur = c->idbuf[c->cur];
    p.idalt = c->idbuf[c->cur ^ 1];
    p.keycur = c->keybuf[c->cur];
    p.keyalt = c->keybuf[c->cur ^ 1];
    p.prank = c->prank;
    p.rcache = c->rcache;
    p.pop = c->pop;
    p.pcap = c->pcap;
    p.slot_count = c->slot_count;
    p.slot_touched = c->slot_touched;
    p.slot_active = c->slot_active;
    p.block_slot = c->block_slot;
    p.block_pstart = c->block_pstart;
    p.block_nbr = c->block_nbr;
    p.cell_count = c->cell_count;
    p.cell_start = c->cell_start;
    p.bcap = c->bcap;
    p.scan_part = c->scan_part;
    p.nparts = c->nparts;
    p.g_m = c->g_m;
    p.g_v = c->g_v;
    p.g_vs = c->g_vs;
    p.g_q = c->g_q;
    p.g_s = c->g_s;
    p.g_r = c->g_r;
    p.g_flags = c->g_flags;
    p.gcap = c->gcap;
    p.ctr = c->ctr;
    p.cg = c->cg;
    p.red = c->red;
    p.red_stride = c->grid_node;
    p.red_count = c->grid_node;
    p.wg_part = c->wg_part;
    p.wg_part_count = c->grid_block;
    p.tiles = c->tiles;
    p.dt = dt;
    return p;
}

template <class F>
void timed(mpm_context* c, int cls, F&& f) {
    if (!c->timing) {
        f();
        return;
    }
    if (c->ev_used + 2 > c->ev.size()) {
        for (int k = 0; k < 256; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) break;
            c->ev.push_back(e);
        }
        c->ev_class.resize(c->ev.size() / 2 + 1);
    }
    (void)hipEventRecord(c->ev[c->ev_used], c->stream);
    f();
    (void)hipEventRecord(c->ev[c->ev_used + 1], c->stream);
    c->ev_class[c->ev_used / 2] = cls;
    c->ev_used += 2;
}

int collect_timing(mpm_context* c) {
    if (!c->timing || c->ev_used == 0) return MPM_OK;
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
        const int cls = c->ev_class[k / 2];
        c->ms[cls] += ms;
        c->launches[cls] += 1;
    }
    c->ev_used = 0;
    return MPM_OK;
}

// ---- stage sequences --------------------------------------------------------
template <int D>
void stage_bin(mpm_context* c) {
    mpm::KP p = make_kp(c, 0.0f);
    timed(c, KC_BIN, [&] { mpm::launch_bin<D>(p, c->stream, c->grid_particle); });
    timed(c, KC_SCAN, [&] { mpm::launch_scan<D>(p, c->stream); });
    timed(c, KC_SETUP, [&] { mpm::launch_block_setup<D>(p, c->stream, c->grid_block); });
    timed(c, KC_REORDER, [&] { mpm::launch_reorder<D>(p, c->stream, c->grid_particle); });
    c->cur ^= 1;
    c->binned = true;
}

template <int D>
void stage_p2g(mpm_context* c) {
    mpm::KP p = make_kp(c, 0.0f);
    timed(c, KC_P2G, [&] { mpm::launch_p2g<D>(p, c->stream, c->grid_block); });
    timed(c, KC_NORMALIZE, [&] { mpm::launch_normalize<D>(p, c->stream, c->grid_node); });
}

template <int D>
void stage_waiting(mpm_context* c) {
    if (!c->waiting) return;
    mpm::KP p = make_kp(c, 0.0f);
    timed(c, KC_INIT_VOLUME, [&] { mpm::launch_init_volume<D>(p, c->stream, c->grid_block); });
    c->waiting = false;
}

// Apply_Force's implicit solve (MPM_Example.cpp:700-711): CG set-up, restart
// (r = b - A x, Project) and up to cg_max_iterations iterations.
template <int D>
int stage_cg(mpm_context* c, float dt, bool do_init) {
    mpm::KP p = make_kp(c, dt);
    if (do_init)
        timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, false, false, true, false); });
    timed(c, KC_CG_FORCE, [&] { mpm::launch_cg_force<D>(p, c->stream, c->grid_block, true); });
    timed(c, KC_CG_NODE, [&] { mpm::launch_cg_restart_node<D>(p, c->stream, c->grid_node); });
    const int max_it = c->desc.cg_max_iterations;
    auto iteration = [&](int k) {
        timed(c, KC_CG_NODE, [&] { mpm::launch_cg_update_s<D>(p, c->stream, c->grid_node, k); });
        timed(c, KC_CG_FORCE, [&] { mpm::launch_cg_force<D>(p, c->stream, c->grid_block, false); });
        timed(c, KC_CG_NODE, [&] { mpm::launch_cg_node1<D>(p, c->stream, c->grid_node); });
        timed(c, KC_CG_NODE, [&] { mpm::launch_cg_node2<D>(p, c->stream, c->grid_node, k); });
    };
    // Iterations after convergence return at once, so launching a few too
    // many is harmless.  The first batch is the previous solve's iteration
    // count; after it the host polls the done flag one chunk behind the GPU
    // (double-buffered events), so the queue never runs dry while it waits.
    const int chunk = 4;
    int first = std::min(max_it, std::max(c->cg_predict, chunk));
    int slot = 0, launched = 0;
    int boundary[2] = {0, 0};
    bool pending[2] = {false, false};
    while (launched < max_it) {
        const int n = launched == 0 ? first : std::min(chunk, max_it - launched);
        for (int j = 0; j < n; ++j) iteration(launched + j);
        launched += n;
        HIPCHECK(hipMemcpyAsync(&c->h_done[slot], &c->cg->done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipEventRecord(c->done_ev[slot], c->stream));
        pending[slot] = true;
        boundary[slot] = launched;
        slot ^= 1;
        if (pending[slot]) {
            HIPCHECK(hipEventSynchronize(c->done_ev[slot]));
            pending[slot] = false;
            if (c->h_done[slot]) {
                // done by the first boundary: try one fewer next time
                c->cg_predict = boundary[slot] == first ? first - 1 : boundary[slot];
                break;
            }
        }
    }
    if (launched >= max_it) c->cg_predict = max_it;
    timed(c, KC_CG_NODE, [&] { mpm::launch_cg_finalize(p, c->stream, max_it); });
    return MPM_OK;
}

template <int D>
void stage_g2p(mpm_context* c, float dt) {
    mpm::KP p = make_kp(c, dt);
    timed(c, KC_G2P, [&] { mpm::launch_g2p<D>(p, c->stream, c->grid_block); });
    c->binned = false;
}

template <int D>
int run_stage(mpm_context* c, int stage, float dt) {
    mpm::KP p = make_kp(c, dt);
    switch (stage) {
        case MPM_STAGE_BIN: stage_bin<D>(c); break;
        case MPM_STAGE_P2G: stage_p2g<D>(c); break;
        case MPM_STAGE_WAITING: stage_waiting<D>(c); break;
        case MPM_STAGE_CONSTITUTIVE:
            timed(c, KC_PFORCE, [&] { mpm::launch_particle_force<D>(p, c->stream, c->grid_block, true, false); });
            break;
        case MPM_STAGE_EXPLICIT_FORCE:
            timed(c, KC_PFORCE, [&] { mpm::launch_particle_force<D>(p, c->stream, c->grid_block, false, true); });
            timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, true, false, false, true); });
            break;
        case MPM_STAGE_COLLISION:
            timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, false, true, false, false); });
            break;
        case MPM_STAGE_CG: return stage_cg<D>(c, dt, true);
        case MPM_STAGE_G2P: stage_g2p<D>(c, dt); break;
        default: return mpm_set_error(MPM_E_INVALID, "unknown stage");
    }
    return MPM_OK;
}

// MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116) with the constitutive update,
// explicit force, collision and CG set-up fused.
template <int D>
int substep(mpm_context* c, float dt) {
    stage_bin<D>(c);
    stage_p2g<D>(c);
    mpm::KP p = make_kp(c, dt);
    timed(c, KC_PFORCE, [&] { mpm::launch_particle_force<D>(p, c->stream, c->grid_block, true, true); });
    timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, true, true, true, false); });
    int rc = stage_cg<D>(c, dt, false);
    if (rc != MPM_OK) return rc;
    stage_g2p<D>(c, dt);
    return MPM_OK;
}

template <int D>
int initialize(mpm_context* c) {
    stage_bin<D>(c);
    stage_p2g<D>(c);
    stage_waiting<D>(c);
    return MPM_OK;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return MPM_OK;
}

int read_counters(mpm_context* c, mpm::Counters& h, mpm::CGState* cg = nullptr) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(&h, c->ctr, sizeof(h), hipMemcpyDeviceToHost));
    if (cg) HIPCHECK(hipMemcpy(cg, c->cg, sizeof(*cg), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int fill_stats(mpm_context* c, mpm_step_stats* s) {
    mpm::Counters h{};
    mpm::CGState g{};
    int rc = read_counters(c, h, &g);
    if (rc) return rc;
    s->cg_iterations = g.iter;
    s->cg_multiplies = g.multiplies;
    s->cg_converged = g.converged;
    s->cg_residual = g.conv;
    s->active_blocks = h.nblocks;
    s->live_particles = (int64_t)h.n_live - h.n_removed;
    float v2;
    std::memcpy(&v2, &h.max_v2, sizeof(v2));
    s->max_speed = std::sqrt((double)v2);
    if (h.overflow) return mpm_set_error(MPM_E_CAPACITY, "active blocks exceed the block capacity");
    if (h.nonfinite) return mpm_set_error(MPM_E_NONFINITE, "non-finite particle position");
    return MPM_OK;
}

}  // namespace

extern "C" {

const char* mpm_last_error(void) { return g_err.c_str(); }

int mpm_device_info(char* buf, int32_t len) {
    if (!buf || len <= 0) return mpm_set_error(MPM_E_INVALID, "mpm_device_info: bad buffer");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, (size_t)len, "%s;%d;%s", prop.name, prop.multiProcessorCount, prop.gcnArchName);
    return MPM_OK;
}

int mpm_create(const mpm_scene_desc* desc, int64_t capacity, mpm_handle* out) {
    if (!desc || !out) return mpm_set_error(MPM_E_INVALID, "mpm_create: null argument");
    if (desc->dim != 2 && desc->dim != 3) return mpm_set_error(MPM_E_INVALID, "mpm_create: dim must be 2 or 3");
    if (desc->counts < 4 || desc->counts > 4096) return mpm_set_error(MPM_E_INVALID, "mpm_create: counts out of range");
    if (desc->nplanes < 0 || desc->nplanes > MPM_MAX_PLANES) return mpm_set_error(MPM_E_INVALID, "mpm_create: too many planes");
    if (capacity < 1 || capacity > (int64_t)1 << 30) return mpm_set_error(MPM_E_INVALID, "mpm_create: bad capacity");
    mpm_context* c = new mpm_context;
    c->desc = *desc;
    c->D = desc->dim;
    const int B = c->D == 3 ? 4 : 8;
    c->nb = (desc->counts + 2 * B) / B + 1;
    c->nslots = c->D == 3 ? c->nb * c->nb * c->nb : c->nb * c->nb;
    if ((int64_t)c->nslots * 64 > ((int64_t)1 << 31) - 1) {
        delete c;
        return mpm_set_error(MPM_E_INVALID, "mpm_create: grid too large");
    }
    c->pcap = capacity;
    // Block capacity: every slot of small grids, else capacity/8 blocks -- about
    // 16x the active blocks (home blocks plus halo) of a scene packed at 8
    // particles per cell.  Overflow makes the substep a no-op and is reported.
    c->bcap = (int)std::min<int64_t>(c->nslots, std::max<int64_t>(4096, capacity / 8));
    c->gcap = (int64_t)c->bcap * 64;
    c->nparts = (c->nslots + mpm::SCAN_TILE - 1) / mpm::SCAN_TILE;
    const int D = c->D;
    int rc = MPM_OK;
    auto A = [&](int r) { if (rc == MPM_OK) rc = r; };
    if (hipStreamCreate(&c->stream) != hipSuccess) rc = mpm_set_error(MPM_E_HIP, "hipStreamCreate failed");
    for (int k = 0; k < 2; ++k) {
        A(dalloc(c, &c->pbuf[k], (size_t)nfields(D) * c->pcap));
        A(dalloc(c, &c->idbuf[k], (size_t)c->pcap));
        A(dalloc(c, &c->keybuf[k], (size_t)c->pcap));
    }
    A(dalloc(c, &c->prank, (size_t)c->pcap));
    A(dalloc(c, &c->rcache, (size_t)D * D * c->pcap));
    A(dalloc(c, &c->pop, (size_t)mpm::op_record_floats(D) * c->pcap));
    A(dalloc(c, &c->slot_count, (size_t)c->nslots));
    A(dalloc(c, &c->slot_touched, (size_t)c->nslots));
    A(dalloc(c, &c->slot_active, (size_t)c->nslots));
    A(dalloc(c, &c->block_slot, (size_t)c->bcap));
    A(dalloc(c, &c->block_pstart, (size_t)c->bcap + 1));
    A(dalloc(c, &c->block_nbr, (size_t)c->bcap * (D == 3 ? 27 : 9)));
    A(dalloc(c, &c->cell_count, (size_t)c->nslots * 64));
    A(dalloc(c, &c->cell_start, (size_t)c->bcap * 65 + 1));
    A(dalloc(c, &c->tiles, (size_t)c->bcap * (D + 1) * (D == 3 ? 216 : 100)));
    A(dalloc(c, &c->scan_part, (size_t)2 * c->nparts));
    A(dalloc(c, &c->g_m, (size_t)c->gcap));
    A(dalloc(c, &c->g_v, (size_t)D * c->gcap));
    A(dalloc(c, &c->g_vs, (size_t)D * c->gcap));
    A(dalloc(c, &c->g_q, (size_t)D * c->gcap));
    A(dalloc(c, &c->g_s, (size_t)D * c->gcap));
    A(dalloc(c, &c->g_r, (size_t)D * c->gcap));
    A(dalloc(c, &c->g_flags, (size_t)c->gcap));
    A(dalloc(c, &c->ctr, 1));
    A(dalloc(c, &c->cg, 1));
    c->grid_particle = (int)std::min<int64_t>(std::max<int64_t>((c->pcap + mpm::WG - 1) / mpm::WG, 8), 4096);
    c->grid_block = (int)std::min<int64_t>(((int64_t)c->bcap + 7) / 8 * 8, 8192);
    // node kernels: one thread per node up to 1024 work-groups (a multiple of 8 for the XCD ranges)
    c->grid_node = (int)std::min<int64_t>(((c->gcap + mpm::WG - 1) / mpm::WG + 7) / 8 * 8, 1024);
    A(dalloc(c, &c->red, (size_t)3 * c->grid_node));
    A(dalloc(c, &c->wg_part, (size_t)2 * c->grid_block));
    if (rc == MPM_OK) {
        void* hp = nullptr;
        if (hipHostMalloc(&hp, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess) rc = mpm_set_error(MPM_E_HIP, "hipHostMalloc failed");
        c->h_done = (int*)hp;
        for (int k = 0; k < 2 && rc == MPM_OK; ++k)
            if (hipEventCreateWithFlags(&c->done_ev[k], hipEventDisableTiming) != hipSuccess) rc = mpm_set_error(MPM_E_HIP, "hipEventCreate failed");
    }
    if (rc != MPM_OK) {
        std::string msg = g_err;
        mpm_destroy(c);
        return mpm_set_error(rc, msg);
    }
    *out = c;
    return MPM_OK;
}

int mpm_destroy(mpm_handle c) {
    if (!c) return MPM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocations) (void)hipFree(p);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; ++k) if (c->done_ev[k]) (void)hipEventDestroy(c->done_ev[k]);
    if (c->h_done) (void)hipHostFree(c->h_done);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MPM_OK;
}

int mpm_upload_particles(mpm_handle c, int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                         const float* mass, const float* volume) {
    if (!c || !x || !v || !mass) return mpm_set_error(MPM_E_INVALID, "mpm_upload_particles: null argument");
    if (n < 0 || n > c->pcap) return mpm_set_error(MPM_E_CAPACITY, "mpm_upload_particles: n exceeds capacity");
    const int D = c->D, NF = nfields(D);
    std::vector<float> h((size_t)NF * c->pcap, 0.0f);
    auto f = [&](int field, int64_t i) -> float& { return h[(size_t)field * c->pcap + i]; };
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) {
            f(a, i) = x[i * D + a];
            f(D + a, i) = v[i * D + a];
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) f(2 * D + a * D + b, i) = F ? F[(i * D + a) * D + b] : (a == b ? 1.0f : 0.0f);
        f(2 * D + D * D, i) = Jp ? Jp[i] : 1.0f;
        f(2 * D + D * D + 1, i) = mass[i];
        f(2 * D + D * D + 2, i) = volume ? volume[i] : 0.0f;
    }
    std::vector<int> ids((size_t)std::max<int64_t>(n, 1));
    std::iota(ids.begin(), ids.end(), 0);
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->cur = 0;
    HIPCHECK(hipMemcpy(c->pbuf[0], h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    if (n > 0) HIPCHECK(hipMemcpy(c->idbuf[0], ids.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(c->keybuf[0], 0, (size_t)c->pcap * sizeof(int)));
    mpm::Counters hc{};
    hc.n_live = (int)n;
    hc.n_prev = (int)n;
    HIPCHECK(hipMemcpy(c->ctr, &hc, sizeof(hc), hipMemcpyHostToDevice));
    mpm::KP p = make_kp(c, 0.0f);
    if (D == 3) mpm::launch_max_speed<3>(p, c->stream, c->grid_particle);
    else mpm::launch_max_speed<2>(p, c->stream, c->grid_particle);
    c->have_particles = true;
    c->waiting = volume == nullptr;
    c->initialized = false;
    c->binned = false;
    return check_launch("mpm_upload_particles");
}

int mpm_download_particles(mpm_handle c, int64_t cap, float* x, float* v, float* F, float* Jp, float* mass,
                           float* volume, int32_t* ids, int64_t* n_out) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_download_particles: null handle");
    mpm::Counters hc{};
    int rc = read_counters(c, hc);
    if (rc) return rc;
    const int D = c->D, NF = nfields(D);
    const int64_t n = hc.n_live;
    std::vector<float> h((size_t)NF * std::max<int64_t>(n, 1));
    std::vector<int> id((size_t)std::max<int64_t>(n, 1)), key((size_t)std::max<int64_t>(n, 1));
    for (int fld = 0; fld < NF && n > 0; ++fld)
        HIPCHECK(hipMemcpy(&h[(size_t)fld * n], c->pbuf[c->cur] + (size_t)fld * c->pcap, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    if (n > 0) {
        HIPCHECK(hipMemcpy(id.data(), c->idbuf[c->cur], (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(key.data(), c->keybuf[c->cur], (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> order;
    order.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i) if (key[(size_t)i] >= 0) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return id[(size_t)a] < id[(size_t)b]; });
    const int64_t live = (int64_t)order.size();
    if (n_out) *n_out = live;
    if (live > cap) return mpm_set_error(MPM_E_CAPACITY, "mpm_download_particles: cap too small");
    auto g = [&](int fld, int64_t i) { return h[(size_t)fld * n + i]; };
    for (int64_t k = 0; k < live; ++k) {
        const int64_t i = order[(size_t)k];
        for (int a = 0; a < D; ++a) {
            if (x) x[k * D + a] = g(a, i);
            if (v) v[k * D + a] = g(D + a, i);
        }
        if (F) for (int a = 0; a < D * D; ++a) F[k * D * D + a] = g(2 * D + a, i);
        if (Jp) Jp[k] = g(2 * D + D * D, i);
        if (mass) mass[k] = g(2 * D + D * D + 1, i);
        if (volume) volume[k] = g(2 * D + D * D + 2, i);
        if (ids) ids[k] = id[(size_t)i];
    }
    return MPM_OK;
}

int mpm_initialize(mpm_handle c) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_initialize: null handle");
    if (!c->have_particles) return mpm_set_error(MPM_E_NOT_READY, "mpm_initialize: no particles uploaded");
    int rc = c->D == 3 ? initialize<3>(c) : initialize<2>(c);
    if (rc) return rc;
    c->initialized = true;
    rc = check_launch("mpm_initialize");
    if (rc) return rc;
    return collect_timing(c);
}

int mpm_substep(mpm_handle c, float dt, mpm_step_stats* stats) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_substep: null handle");
    if (!c->initialized) return mpm_set_error(MPM_E_NOT_READY, "mpm_substep: call mpm_initialize first");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return mpm_set_error(MPM_E_INVALID, "mpm_substep: dt must be positive");
    int rc = c->D == 3 ? substep<3>(c, dt) : substep<2>(c, dt);
    if (rc) return rc;
    c->last_dt = dt;
    rc = check_launch("mpm_substep");
    if (rc) return rc;
    rc = collect_timing(c);
    if (rc) return rc;
    if (stats) return fill_stats(c, stats);
    return MPM_OK;
}

int mpm_max_particle_velocity(mpm_handle c, float* out) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_max_particle_velocity: null argument");
    mpm::Counters hc{};
    int rc = read_counters(c, hc);
    if (rc) return rc;
    float v2;
    std::memcpy(&v2, &hc.max_v2, sizeof(v2));
    *out = std::sqrt(v2);
    return MPM_OK;
}

int mpm_advance_to_time(mpm_handle c, double time_in, double target_in, int32_t* substeps_out, double* time_out,
                        mpm_step_stats* stats) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_advance_to_time: null handle");
    const float min_dt = c->desc.min_dt, max_dt = c->desc.max_dt, cfl = c->desc.cfl;
    const float dx_min = 1.0f / (float)c->desc.counts;
    float time = (float)time_in;
    const float target_time = (float)target_in;
    bool done = false;
    int n = 0;
    while (!done) {
        float max_v = 0.0f;
        int rc = mpm_max_particle_velocity(c, &max_v);
        if (rc) return rc;
        float dt = std::max(min_dt, std::min(max_dt, cfl * dx_min / std::max(max_v, 1e-2f)));
        if (target_time - time < dt * 1.001f) {
            dt = target_time - time;
            done = true;
        } else if (target_time - time < 2.0f * dt) {
            dt = (target_time - time) * 0.5f;
        }
        if (!(dt > 0.0f)) break;
        rc = mpm_substep(c, dt, nullptr);
        if (rc) return rc;
        time += dt;
        ++n;
    }
    if (substeps_out) *substeps_out = n;
    if (time_out) *time_out = (double)time;
    if (stats) return fill_stats(c, stats);
    return MPM_OK;
}

int mpm_synchronize(mpm_handle c) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_synchronize: null handle");
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_run_stage(mpm_handle c, int32_t stage, float dt) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_run_stage: null handle");
    if (!c->have_particles) return mpm_set_error(MPM_E_NOT_READY, "mpm_run_stage: no particles uploaded");
    if (stage != MPM_STAGE_BIN && stage != MPM_STAGE_CONSTITUTIVE && !c->binned)
        return mpm_set_error(MPM_E_NOT_READY, "mpm_run_stage: particles moved since the last MPM_STAGE_BIN");
    int rc = c->D == 3 ? run_stage<3>(c, stage, dt) : run_stage<2>(c, stage, dt);
    if (rc) return rc;
    if (stage == MPM_STAGE_WAITING || stage == MPM_STAGE_P2G) c->initialized = true;
    rc = check_launch("mpm_run_stage");
    if (rc) return rc;
    return collect_timing(c);
}

int mpm_read_grid(mpm_handle c, int32_t channel, float* out, int64_t out_len) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_read_grid: null argument");
    const int D = c->D, B = D == 3 ? 4 : 8, n = c->desc.counts;
    const float* src = nullptr;
    int width = D;
    switch (channel) {
        case MPM_CH_MASS: src = c->g_m; width = 1; break;
        case MPM_CH_VELOCITY: src = c->g_v; break;
        case MPM_CH_VELOCITY_STAR: src = c->g_vs; break;
        case MPM_CH_FORCE: src = c->g_q; break;
        case MPM_CH_CG_R: src = c->g_r; break;
        case MPM_CH_CG_S: src = c->g_s; break;
        case MPM_CH_CG_Q: src = c->g_q; break;
        default: return mpm_set_error(MPM_E_INVALID, "mpm_read_grid: unknown channel");
    }
    int64_t total = 1;
    for (int a = 0; a < D; ++a) total *= n + 2;
    if (out_len < total * width) return mpm_set_error(MPM_E_INVALID, "mpm_read_grid: output too small");
    mpm::Counters hc{};
    int rc = read_counters(c, hc);
    if (rc) return rc;
    const int nbk = std::min(hc.nblocks, c->bcap);
    std::vector<int> slots((size_t)std::max(nbk, 1));
    std::vector<float> vals((size_t)std::max(nbk, 1) * 64 * width);
    if (nbk > 0) {
        HIPCHECK(hipMemcpy(slots.data(), c->block_slot, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
        for (int w = 0; w < width; ++w)
            HIPCHECK(hipMemcpy(&vals[(size_t)w * nbk * 64], src + (size_t)w * c->gcap, (size_t)nbk * 64 * sizeof(float), hipMemcpyDeviceToHost));
    }
    std::fill(out, out + total * width, 0.0f);
    for (int b = 0; b < nbk; ++b) {
        int bb[3];
        int s = slots[(size_t)b];
        for (int a = D - 1; a >= 0; --a) { bb[a] = s % c->nb; s /= c->nb; }
        for (int l = 0; l < 64; ++l) {
            int lc[3];
            if (D == 3) { lc[0] = l >> 4; lc[1] = (l >> 2) & 3; lc[2] = l & 3; }
            else { lc[0] = l >> 3; lc[1] = l & 7; }
            int64_t o = 0;
            bool ok = true;
            for (int a = 0; a < D; ++a) {
                const int i = B * bb[a] + lc[a] - B;
                ok = ok && i >= -1 && i <= n;
                o = o * (n + 2) + (i + 1);
            }
            if (!ok) continue;
            for (int w = 0; w < width; ++w) out[o * width + w] = vals[(size_t)w * nbk * 64 + (size_t)b * 64 + l];
        }
    }
    return MPM_OK;
}

int mpm_read_flags(mpm_handle c, uint32_t* out, int64_t out_len) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_read_flags: null argument");
    const int D = c->D, B = D == 3 ? 4 : 8, n = c->desc.counts;
    int64_t total = 1;
    for (int a = 0; a < D; ++a) total *= n + 2;
    if (out_len < total) return mpm_set_error(MPM_E_INVALID, "mpm_read_flags: output too small");
    mpm::Counters hc{};
    int rc = read_counters(c, hc);
    if (rc) return rc;
    const int nbk = std::min(hc.nblocks, c->bcap);
    std::vector<int> slots((size_t)std::max(nbk, 1));
    std::vector<unsigned> fl((size_t)std::max(nbk, 1) * 64);
    if (nbk > 0) {
        HIPCHECK(hipMemcpy(slots.data(), c->block_slot, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(fl.data(), c->g_flags, (size_t)nbk * 64 * sizeof(unsigned), hipMemcpyDeviceToHost));
    }
    std::fill(out, out + total, 0u);
    for (int b = 0; b < nbk; ++b) {
        int bb[3];
        int s = slots[(size_t)b];
        for (int a = D - 1; a >= 0; --a) { bb[a] = s % c->nb; s /= c->nb; }
        for (int l = 0; l < 64; ++l) {
            int lc[3];
            if (D == 3) { lc[0] = l >> 4; lc[1] = (l >> 2) & 3; lc[2] = l & 3; }
            else { lc[0] = l >> 3; lc[1] = l & 7; }
            int64_t o = 0;
            bool ok = true;
            for (int a = 0; a < D; ++a) {
                const int i = B * bb[a] + lc[a] - B;
                ok = ok && i >= -1 && i <= n;
                o = o * (n + 2) + (i + 1);
            }
            if (ok) out[o] = fl[(size_t)b * 64 + l];
        }
    }
    return MPM_OK;
}

int mpm_read_bins(mpm_handle c, int64_t cap_blocks, int32_t* block_coords, int32_t* block_particle_counts,
                  int64_t cap_particles, int32_t* particle_block, int32_t* particle_ids, int32_t* nblocks,
                  int64_t* nparticles) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_read_bins: null handle");
    mpm::Counters hc{};
    int rc = read_counters(c, hc);
    if (rc) return rc;
    const int D = c->D;
    const int nbk = std::min(hc.nblocks, c->bcap);
    const int64_t np = hc.n_live;
    if (nblocks) *nblocks = nbk;
    if (nparticles) *nparticles = np;
    std::vector<int> slots((size_t)std::max(nbk, 1)), pstart((size_t)nbk + 1);
    if (nbk > 0) {
        HIPCHECK(hipMemcpy(slots.data(), c->block_slot, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(pstart.data(), c->block_pstart, ((size_t)nbk + 1) * sizeof(int), hipMemcpyDeviceToHost));
    } else {
        pstart[0] = 0;
    }
    if (block_coords || block_particle_counts) {
        if (cap_blocks < nbk) return mpm_set_error(MPM_E_CAPACITY, "mpm_read_bins: cap_blocks too small");
        for (int b = 0; b < nbk; ++b) {
            int s = slots[(size_t)b];
            int bb[3];
            for (int a = D - 1; a >= 0; --a) { bb[a] = s % c->nb; s /= c->nb; }
            if (block_coords) for (int a = 0; a < D; ++a) block_coords[(int64_t)b * D + a] = bb[a];
            if (block_particle_counts) block_particle_counts[b] = pstart[(size_t)b + 1] - pstart[(size_t)b];
        }
    }
    if (particle_block || particle_ids) {
        if (cap_particles < np) return mpm_set_error(MPM_E_CAPACITY, "mpm_read_bins: cap_particles too small");
        if (particle_ids && np > 0)
            HIPCHECK(hipMemcpy(particle_ids, c->idbuf[c->cur], (size_t)np * sizeof(int), hipMemcpyDeviceToHost));
        if (particle_block)
            for (int b = 0; b < nbk; ++b)
                for (int i = pstart[(size_t)b]; i < pstart[(size_t)b + 1]; ++i) particle_block[i] = b;
    }
    return MPM_OK;
}

int mpm_cg_stats(mpm_handle c, mpm_step_stats* stats) {
    if (!c || !stats) return mpm_set_error(MPM_E_INVALID, "mpm_cg_stats: null argument");
    return fill_stats(c, stats);
}

int mpm_enable_timing(mpm_handle c, int32_t on) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_enable_timing: null handle");
    int rc = collect_timing(c);
    if (rc) return rc;
    c->timing = on != 0;
    for (int k = 0; k < KC_COUNT; ++k) { c->ms[k] = 0.0; c->launches[k] = 0; }
    return MPM_OK;
}

int mpm_kernel_times(mpm_handle c, int32_t nclasses, double* ms, int64_t* launches) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_kernel_times: null handle");
    int rc = collect_timing(c);
    if (rc) return rc;
    for (int k = 0; k < nclasses && k < KC_COUNT; ++k) {
        if (ms) ms[k] = c->ms[k];
        if (launches) launches[k] = c->launches[k];
    }
    return MPM_OK;
}

int mpm_time_cg_operator(mpm_handle c, int32_t reps, float* avg_ms) {
    if (!c || !avg_ms || reps <= 0) return mpm_set_error(MPM_E_INVALID, "mpm_time_cg_operator: bad argument");
    if (c->last_dt <= 0.0f) return mpm_set_error(MPM_E_NOT_READY, "mpm_time_cg_operator: run a substep first");
    // the operator kernel returns at once after a finished solve: clear the
    // done flag for the timed launches and restore it afterwards
    int done = 0;
    HIPCHECK(hipMemcpyAsync(&done, &c->cg->done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemsetAsync(&c->cg->done, 0, sizeof(int), c->stream));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    mpm::KP p = make_kp(c, c->last_dt);
    HIPCHECK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) {
        if (c->D == 3) mpm::launch_cg_force<3>(p, c->stream, c->grid_block, false);
        else mpm::launch_cg_force<2>(p, c->stream, c->grid_block, false);
    }
    HIPCHECK(hipEventRecord(e1, c->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    HIPCHECK(hipMemcpyAsync(&c->cg->done, &done, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    int rc = check_launch("mpm_time_cg_operator");
    if (rc) return rc;
    *avg_ms = ms / (float)reps;
    return MPM_OK;
}

const char* mpm_kernel_class_name(int32_t cls) { return (cls >= 0 && cls < KC_COUNT) ? kClassNames[cls] : ""; }
int32_t mpm_kernel_class_count(void) { return KC_COUNT; }

}  // extern "C"
