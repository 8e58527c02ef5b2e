// Host side of libmpm_gpu: batch context (a single scene is a batch of one),
// device memory, launch sequence of the substep and the C ABI of
// include/mpm_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstddef>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mpm_gpu.h"
#include "mpm_context.h"
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

const char* kClassNames[KC_COUNT] = {"bin",          "scan",         "block_setup", "reorder", "p2g",
                                     "normalize",    "init_volume",  "constitutive_force", "grid_update",
                                     "cg_force",     "cg_node",      "g2p",         "max_speed"};

}  // namespace

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

int nfields(int D) { return 2 * D + D * D + 4; }  // X V F Jp m V0 material
constexpr int MAX_NODE_GROUPS = 2048;
constexpr int MAX_GRAPH_CG_ITERATIONS = 256;
int block_edge(int D) { return D == 3 ? 4 : 8; }
int slots_per_axis(int D, int counts) { const int B = block_edge(D); return (counts + 2 * B) / B + 1; }

mpm::SceneParams scene_params(const mpm_scene_desc& d, int slot_base, int nslots) {
    mpm::SceneParams p{};
    p.counts = d.counts;
    p.nb = slots_per_axis(d.dim, d.counts);
    p.slot_base = slot_base;
    p.nslots = nslots;
    p.dx = 1.0f / (float)d.counts;
    p.inv_dx = (float)d.counts;
    p.nmat = std::max(1, std::min((int)d.nmaterials, mpm::MAX_MATERIALS));
    for (int k = 0; k < p.nmat; ++k) {
        const mpm_material& m = k == 0 ? d.material : d.extra_materials[k - 1];
        mpm::Material& o = p.mats[k];
        o.mu0 = m.mu0;
        o.lambda0 = m.lambda0;
        o.sy = m.stretching_yield;
        o.cy = m.compression_yield;
        o.xi = m.hardening_factor;
        o.hmax = m.hardening_max;
        o.plastic = m.plastic;
        o.eos = m.eos;
        o.K = m.bulk_modulus;
        o.gamma = m.gamma;
        o.rho0 = m.density;
    }
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
    p.dt = 0.0f;
    p.active = 0;
    p.waiting = 0;
    return p;
}

mpm::KP make_kp(const mpm_context* c) {
    mpm::KP p{};
    p.scn = c->scn;
    p.sst = c->sst;
    p.cg = c->cg;
    p.slot_scene = c->slot_scene;
    p.nscenes = c->nscenes;
    p.solo = c->solo >= 0 ? c->solo : 0;
    p.cg_target = (c->nscenes == 1 || c->solo >= 0) ? 1 : c->nscenes;
    p.pcur = c->pbuf[c->cur];
    p.palt = c->pbuf[c->cur ^ 1];
    p.idcur = c->idbuf[c->cur];
    p.idalt = c->idbuf[c->cur ^ 1];
    p.keycur = c->keybuf[c->cur];
    p.keyalt = c->keybuf[c->cur ^ 1];
    p.scur = c->scbuf[c->cur];
    p.salt = c->scbuf[c->cur ^ 1];
    p.prank = c->prank;
    p.cell_order = c->cell_order;
    p.rcache = c->rcache;
    p.pop = c->pop;
    p.pcap = c->pcap;
    p.nslots = c->nslots;
    p.slot_count = c->slot_count;
    p.slot_touched = c->slot_touched;
    p.slot_active = c->slot_active;
    p.block_slot = c->block_slot;
    p.block_scene = c->block_scene;
    p.block_pstart = c->block_pstart;
    p.block_nbr = c->block_nbr;
    p.block_tnbr = c->block_tnbr;
    p.block_org = c->block_org;
    p.block_order = c->block_order;
    p.block_full = c->block_full;
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
    p.tiles = c->tiles;
    p.ctr = c->ctr;
    p.bsum = c->bsum;
    p.bmax = c->bmax;
    p.bsq = c->bsq;
    p.bu = c->bu;
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

// dt and activity of every scene for the next kernels.  dts == nullptr: every
// scene gets dt_all; force_active makes every scene active whatever its dt
// (grid build, stage-by-stage calls).
void set_step(mpm_context* c, const float* dts, float dt_all, bool force_active, bool clear_waiting) {
    const mpm::KP p = make_kp(c);
    for (int first = 0; first < c->nscenes; first += mpm::STEP_ARGS_MAX) {
        mpm::StepArgs a{};
        a.first = first;
        a.n = std::min(mpm::STEP_ARGS_MAX, c->nscenes - first);
        a.force_active = force_active ? 1 : 0;
        a.clear_waiting = clear_waiting ? 1 : 0;
        for (int i = 0; i < a.n; ++i) a.dt[i] = dts ? dts[first + i] : dt_all;
        mpm::launch_set_step(p, c->stream, a);
    }
}

// The CG operator in the handle's mode: four waves per block (a cell's
// particles split over them) or one wave per block (a lane runs its cell).
template <int D>
void launch_operator_as(const mpm_context* c, const mpm::KP& p, int mode, int grid, bool restart, int k, bool one) {
    if (mode == 1) mpm::launch_cg_force_cell<D>(p, c->stream, grid, restart, k, one);
    else mpm::launch_cg_force<D>(p, c->stream, grid, restart, k, one);
}
template <int D>
void launch_operator(const mpm_context* c, const mpm::KP& p, bool restart, int k, bool one) {
    launch_operator_as<D>(c, p, c->op_mode, c->grid_op, restart, k, one);
}

// ---- stage sequences --------------------------------------------------------
template <int D>
void stage_bin(mpm_context* c) {
    mpm::KP p = make_kp(c);
    timed(c, KC_BIN, [&] { mpm::launch_bin<D>(p, c->stream, c->grid_particle); });
    timed(c, KC_SCAN, [&] { mpm::launch_scan(p, c->stream); });
    timed(c, KC_SETUP, [&] {
        mpm::launch_block_setup<D>(p, c->stream, c->grid_block);
        mpm::launch_setup_order(p, c->stream, c->grid_particle);
    });
    timed(c, KC_REORDER, [&] { mpm::launch_reorder<D>(p, c->stream, c->grid_particle); });
    c->cur ^= 1;
    c->records_fresh = false;
    c->binned = true;
    c->ever_binned = true;
}

template <int D>
void stage_p2g(mpm_context* c) {
    mpm::KP p = make_kp(c);
    timed(c, KC_P2G, [&] { mpm::launch_p2g<D>(p, c->stream, c->grid_block); });
    timed(c, KC_NORMALIZE, [&] { mpm::launch_normalize<D>(p, c->stream, c->grid_node); });
}

template <int D>
void stage_waiting(mpm_context* c) {
    mpm::KP p = make_kp(c);
    timed(c, KC_INIT_VOLUME, [&] { mpm::launch_init_volume<D>(p, c->stream, c->grid_block); });
}

// particles from which a batch substep with one active scene is launched
// (single-scene CG kernels, polled) rather than replayed from the batch graph
constexpr int64_t SOLO_LAUNCHED_MIN = 65536;
// iterations per launch batch of the polled CG after the first; the batch
// right after the first (the predicted count) covers the host's reading of
// the first batch's poll and is usually all no-ops
constexpr int CG_CHUNK = 4;
constexpr int CG_COVER = 1;
// A done poll copies Counters::overflow .. cg_iters (five ints) into a slot
// of POLL_INTS pinned ints: [0] overflow, [1] finished solves, [4] the
// largest iteration count among them.
constexpr int POLL_INTS = 8;
static_assert(offsetof(mpm::Counters, cg_ndone) == offsetof(mpm::Counters, overflow) + sizeof(int) &&
                  offsetof(mpm::Counters, cg_iters) == offsetof(mpm::Counters, overflow) + 4 * sizeof(int),
              "Counters layout of the done poll");

int max_cg_iterations(const mpm_context* c) {
    if (c->pool) return c->pool_cg_max;
    int m = 0;
    for (const SceneHost& s : c->sc) m = std::max(m, s.desc.cg_max_iterations);
    return m;
}

// Apply_Force's implicit solve (MPM_Example.cpp:700-711) for every active
// scene: CG set-up, restart (r = b - A x, Project) and iterations until every
// scene's solve is done (converged or at its cap).  Batches reduce each
// scene's block partials in per-scene kernels (five launches per iteration);
// one scene needs no reducer launches (four per iteration).
template <int D>
int stage_cg(mpm_context* c, bool do_init, bool fixed = false) {
    mpm::KP p = make_kp(c);
    const bool one = c->nscenes == 1 || c->solo >= 0;
    const int gn = c->grid_node;
    if (do_init)
        timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, gn, false, false, true, false); });
    timed(c, KC_CG_FORCE, [&] { launch_operator<D>(c, p, true, 0, one); });
    timed(c, KC_CG_NODE, [&] {
        mpm::launch_cg_restart_node<D>(p, c->stream, gn, one);
        if (!one) mpm::launch_cg_reduce_rr(p, c->stream, 0);
    });
    const int max_it = max_cg_iterations(c);
    auto iteration = [&](int k) {
        if (one) timed(c, KC_CG_NODE, [&] { mpm::launch_cg_update_s<D>(p, c->stream, gn, k); });
        timed(c, KC_CG_FORCE, [&] { launch_operator<D>(c, p, false, k, one); });
        timed(c, KC_CG_NODE, [&] {
            mpm::launch_cg_node1<D>(p, c->stream, gn, k, one);
            if (!one) mpm::launch_cg_reduce_sq(p, c->stream, k);
            mpm::launch_cg_node2<D>(p, c->stream, gn, k, one);
            if (!one) mpm::launch_cg_reduce_rr(p, c->stream, k + 1);
        });
    };
    if (fixed) {
        // captured substeps: every iteration up to the largest cap, no polling
        for (int k = 0; k < max_it; ++k) iteration(k);
        if (one) mpm::launch_cg_update_s<D>(p, c->stream, gn, max_it);
        return MPM_OK;
    }
    // Iterations after every solve is done return at once, so launching a few
    // too many is harmless.  The first batch is the previous solve's iteration
    // count; after it the host polls the done count one chunk behind the GPU
    // (double-buffered events), so the queue never runs dry while it waits.
    const int first = std::min(max_it, std::max(c->cg_predict, CG_CHUNK));
    const int target = one ? 1 : c->nscenes;
    int slot = 0, launched = 0;
    int boundary[2] = {0, 0};
    bool pending[2] = {false, false};
    bool finished = false;
    while (launched < max_it) {
        const int n = launched == 0 ? first : std::min(launched == first ? CG_COVER : CG_CHUNK, max_it - launched);
        for (int j = 0; j < n; ++j) iteration(launched + j);
        launched += n;
        // one wave stores the overflow flag, the done count and the
        // iterations into the pinned slot (a kernel in the stream's order: no
        // runtime copy)
        mpm::launch_poll(p, c->stream, c->d_done + POLL_INTS * slot);
        HIPCHECK(hipEventRecord(c->done_ev[slot], c->stream));
        pending[slot] = true;
        boundary[slot] = launched;
        slot ^= 1;
        if (pending[slot]) {
            HIPCHECK(hipEventSynchronize(c->done_ev[slot]));
            pending[slot] = false;
            const int* h = &c->h_done[POLL_INTS * slot];
            c->overflow_seen = c->overflow_seen || h[0] != 0;
            if (h[1] >= target) {
                // the next solve's first batch: this one's iterations (iteration
                // k's convergence test ends a solve of k updates)
                c->cg_predict = h[4] + 1;
                finished = true;
                break;
            }
        }
    }
    if (!finished) {
        // one scene: the convergence test after the last iteration (batches ran it in reduce_rr)
        if (one) timed(c, KC_CG_NODE, [&] { mpm::launch_cg_update_s<D>(p, c->stream, gn, max_it); });
        // the polls still in flight are read once the rest of the substep is queued
        c->poll_open = true;
        c->poll_older = slot;  // the slot written first of the two
        c->poll_target = target;
        c->poll_max_it = max_it;
        for (int k = 0; k < 2; ++k) {
            c->poll_pending[k] = pending[k];
            c->poll_boundary[k] = boundary[k];
        }
    }
    return MPM_OK;
}

// Reads the done polls of a solve that ran to its last chunk: the overflow
// flag they carry makes the substep report MPM_E_CAPACITY, and the
// iteration count they carry sets the next solve's first batch, so the
// prediction comes back down after a solve that needed many.
int finish_cg_poll(mpm_context* c) {
    if (!c->poll_open) return MPM_OK;
    c->poll_open = false;
    int iters = -1;
    for (int j = 0; j < 2; ++j) {
        const int slot = (c->poll_older + j) & 1;
        if (!c->poll_pending[slot]) continue;
        c->poll_pending[slot] = false;
        HIPCHECK(hipEventSynchronize(c->done_ev[slot]));
        const int* h = &c->h_done[POLL_INTS * slot];
        c->overflow_seen = c->overflow_seen || h[0] != 0;
        if (iters < 0 && h[1] >= c->poll_target) iters = h[4];
    }
    // seen done: its iterations; never seen done (the cap's test runs after
    // the last poll): the cap
    c->cg_predict = iters >= 0 ? iters + 1 : c->poll_max_it;
    return MPM_OK;
}

template <int D>
void stage_g2p(mpm_context* c) {
    mpm::KP p = make_kp(c);
    timed(c, KC_G2P, [&] { mpm::launch_g2p<D>(p, c->stream, c->grid_block); });
    c->binned = false;
}

template <int D>
int run_stage(mpm_context* c, int stage, float dt) {
    c->solo = -1;
    set_step(c, nullptr, dt, true, false);
    mpm::KP p = make_kp(c);
    switch (stage) {
        case MPM_STAGE_BIN: stage_bin<D>(c); break;
        case MPM_STAGE_P2G: stage_p2g<D>(c); break;
        case MPM_STAGE_WAITING:
            stage_waiting<D>(c);
            set_step(c, nullptr, dt, true, true);
            break;
        case MPM_STAGE_CONSTITUTIVE:
            timed(c, KC_PFORCE, [&] { mpm::launch_particle_force<D>(p, c->stream, c->grid_block, true, false); });
            c->records_fresh = c->binned;
            break;
        case MPM_STAGE_EXPLICIT_FORCE:
            timed(c, KC_PFORCE, [&] { mpm::launch_particle_force<D>(p, c->stream, c->grid_block, false, true); });
            timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, true, false, false, true); });
            break;
        case MPM_STAGE_COLLISION:
            timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, false, true, false, false); });
            break;
        case MPM_STAGE_CG: {
            const int rc = stage_cg<D>(c, true);
            return rc ? rc : finish_cg_poll(c);
        }
        case MPM_STAGE_G2P: stage_g2p<D>(c); break;
        default: return mpm_set_error(MPM_E_INVALID, "unknown stage");
    }
    return MPM_OK;
}

// MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116) for every scene with
// dt > 0, with the constitutive update, explicit force, collision and CG
// set-up fused.
// capture: the launch sequence of a graph (dts from d_dt, fixed CG count).
template <int D>
int substep(mpm_context* c, const float* dts, bool capture = false) {
    // a batch or arena substep with one active scene runs the single-scene
    // CG kernels on it (captured substeps keep the batch kernels: their
    // active set changes between replays)
    c->solo = -1;
    struct SoloReset {  // only this substep's launches see the solo scene
        mpm_context* c;
        ~SoloReset() { c->solo = -1; }
    } solo_reset{c};
    if (!capture && c->nscenes > 1) {
        int active = 0, last = -1;
        for (int s = 0; s < c->nscenes; ++s)
            if (dts[s] > 0.0f) { ++active; last = s; }
        if (active == 1) c->solo = last;
    }
    if (capture) mpm::launch_set_step_dev(make_kp(c), c->stream, c->d_dt);
    else set_step(c, dts, 0.0f, false, false);
    c->overflow_seen = false;
    stage_bin<D>(c);
    stage_p2g<D>(c);
    mpm::KP p = make_kp(c);
    const bool was_emitted = c->emitted;
    std::vector<int> was_waiting;
    if (c->pool || c->emitted) {
        // scenes attached or particles added since the last substep:
        // Process_Waiting_Particles (every other scene's blocks return at once)
        timed(c, KC_INIT_VOLUME, [&] { mpm::launch_init_volume<D>(p, c->stream, c->grid_block); });
        mpm::launch_clear_waiting(p, c->stream);
        if (!capture) {
            c->emitted = false;
            for (mpm::SceneParams& t : c->table) {
                was_waiting.push_back(t.waiting);
                t.waiting = 0;
            }
        }
    }
    timed(c, KC_PFORCE, [&] { mpm::launch_particle_force<D>(p, c->stream, c->grid_block, true, true); });
    c->records_fresh = true;
    timed(c, KC_GRID, [&] { mpm::launch_grid_update<D>(p, c->stream, c->grid_node, true, true, true, false); });
    int rc = stage_cg<D>(c, false, capture);
    if (rc != MPM_OK) return rc;
    stage_g2p<D>(c);
    if (!capture) {
        rc = finish_cg_poll(c);
        if (rc != MPM_OK) return rc;
    }
    // a binning over the block capacity made the substep a no-op (seen when
    // the CG's done poll read the counters; stats always read them); waiting
    // particles keep waiting (k_clear_waiting left the device flags set)
    if (c->overflow_seen) {
        c->emitted = was_emitted;
        for (size_t k = 0; k < was_waiting.size(); ++k) c->table[k].waiting = was_waiting[k];
        return mpm_set_error(MPM_E_CAPACITY, "active blocks exceed the block capacity");
    }
    return MPM_OK;
}

template <int D>
int initialize(mpm_context* c) {
    set_step(c, nullptr, 0.0f, true, false);
    stage_bin<D>(c);
    stage_p2g<D>(c);
    stage_waiting<D>(c);
    set_step(c, nullptr, 0.0f, true, true);
    return MPM_OK;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return MPM_OK;
}

int read_counters(mpm_context* c, mpm::Counters& h) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(&h, c->ctr, sizeof(h), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int read_scene(mpm_context* c, int s, mpm::SceneState& st, mpm::CGState* cg) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(&st, c->sst + s, sizeof(st), hipMemcpyDeviceToHost));
    if (cg) HIPCHECK(hipMemcpy(cg, c->cg + s, sizeof(*cg), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int fill_stats(mpm_context* c, int s, mpm_step_stats* out) {
    mpm::Counters h{};
    mpm::SceneState st{};
    mpm::CGState g{};
    int rc = read_counters(c, h);
    if (rc) return rc;
    rc = read_scene(c, s, st, &g);
    if (rc) return rc;
    out->cg_iterations = g.iter;
    out->cg_multiplies = g.multiplies;
    out->cg_converged = g.converged;
    out->cg_residual = g.conv;
    out->active_blocks = st.block_end - st.block_begin;
    out->live_particles = c->ever_binned ? (int64_t)(st.pend - st.pbegin) - st.n_removed : c->sc[(size_t)s].n_uploaded;
    float v2;
    std::memcpy(&v2, &st.max_v2, sizeof(v2));
    out->max_speed = std::sqrt((double)v2);
    if (h.overflow) return mpm_set_error(MPM_E_CAPACITY, "active blocks exceed the block capacity");
    if (st.nonfinite) return mpm_set_error(MPM_E_NONFINITE, "non-finite particle state (the particles were removed)");
    return MPM_OK;
}

// Sizes that follow from a context's D, pcap, nslots and bcap: grid nodes,
// scan parts, reduction partials and the default launch shapes.
void size_context(mpm_context* c) {
    c->gcap = (int64_t)c->bcap * 64;
    c->nparts = (c->nslots + mpm::SCAN_TILE - 1) / mpm::SCAN_TILE;
    // reduction partials: per block (batches) or per node work-group (one scene)
    c->nred = (size_t)std::max(c->bcap, MAX_NODE_GROUPS);
    c->grid_particle = (int)std::min<int64_t>(((c->pcap + mpm::WG - 1) / mpm::WG + 7) / 8 * 8, 4096);
    c->grid_block = (int)std::min<int64_t>(((int64_t)c->bcap + 7) / 8 * 8, 8192);
    c->grid_op = c->grid_block;
    c->op_mode = 0;
    // node kernels: one wave per block, up to 1024 work-groups (a multiple of 8 for the XCD ranges)
    c->grid_node = (int)std::min<int64_t>(((int64_t)c->bcap / (mpm::WG / 64) + 7) / 8 * 8 + 8, 1024);
}

// Every device buffer of a sized context, in one fixed order: a handle
// allocates each of them (allocate_context), the allocation-free substep
// carves them from its static scratch (mpm_substep_noalloc).
template <class Take>
void layout_buffers(mpm_context* c, Take&& take) {
    const int D = c->D;
    const size_t nscenes = (size_t)c->nscenes, pcap = (size_t)c->pcap, nslots = (size_t)c->nslots,
                 bcap = (size_t)c->bcap, gcap = (size_t)c->gcap;
    for (int k = 0; k < 2; ++k) {
        take(&c->pbuf[k], (size_t)nfields(D) * pcap);
        take(&c->idbuf[k], pcap);
        take(&c->keybuf[k], pcap);
        take(&c->scbuf[k], pcap);
    }
    take(&c->prank, pcap);
    take(&c->cell_order, pcap);
    take(&c->rcache, (size_t)D * D * pcap);
    take(&c->pop, (size_t)mpm::op_record_floats(D) * pcap);
    take(&c->slot_count, nslots);
    take(&c->slot_touched, nslots);
    take(&c->slot_active, nslots);
    take(&c->slot_scene, nslots);
    take(&c->block_slot, bcap);
    take(&c->block_scene, bcap);
    take(&c->block_pstart, bcap + 1);
    take(&c->block_nbr, bcap * (D == 3 ? 27 : 9));
    take(&c->block_tnbr, bcap * (D == 3 ? 27 : 9));
    take(&c->block_org, bcap * 3);
    take(&c->block_order, 2 * bcap);
    take(&c->block_full, bcap);
    take(&c->cell_count, nslots * 64);
    take(&c->cell_start, bcap * 65 + 1);
    take(&c->tiles, bcap * (D + 1) * (D == 3 ? 216 : 100));
    take(&c->scan_part, (size_t)2 * c->nparts);
    take(&c->g_m, gcap);
    take(&c->g_v, (size_t)D * gcap);
    take(&c->g_vs, (size_t)D * gcap);
    take(&c->g_q, (size_t)D * gcap);
    take(&c->g_s, (size_t)D * gcap);
    take(&c->g_r, (size_t)D * gcap);
    take(&c->g_flags, gcap);
    take(&c->ctr, 1);
    take(&c->scn, nscenes);
    take(&c->sst, nscenes);
    take(&c->cg, nscenes);
    take(&c->bsum, c->nred);
    take(&c->bmax, c->nred);
    take(&c->bsq, c->nred);
    take(&c->bu, (size_t)2 * bcap);
    take(&c->d_dt, nscenes);
    take(&c->frame, nscenes);
}

// Device buffers, stream, events and launch shapes of a context whose D,
// nscenes, pcap, nslots and bcap are set.
int allocate_context(mpm_context* c) {
    const int nscenes = c->nscenes;
    size_context(c);
    int rc = MPM_OK;
    auto A = [&](int r) { if (rc == MPM_OK) rc = r; };
    if (hipStreamCreate(&c->stream) != hipSuccess) rc = mpm_set_error(MPM_E_HIP, "hipStreamCreate failed");
    layout_buffers(c, [&](auto** ptr, size_t count) { A(dalloc(c, ptr, count)); });
    if (rc == MPM_OK) {
        void* hp = nullptr;
        // coherent: k_poll stores the counters into it from the device
        if (hipHostMalloc(&hp, 2 * POLL_INTS * sizeof(int), hipHostMallocCoherent) != hipSuccess)
            rc = mpm_set_error(MPM_E_HIP, "hipHostMalloc failed");
        c->h_done = (int*)hp;
        if (rc == MPM_OK && hipHostGetDevicePointer((void**)&c->d_done, hp, 0) != hipSuccess)
            rc = mpm_set_error(MPM_E_HIP, "hipHostGetDevicePointer failed");
        for (int k = 0; k < 2 && rc == MPM_OK; ++k)
            if (hipEventCreateWithFlags(&c->done_ev[k], hipEventDisableTiming) != hipSuccess) rc = mpm_set_error(MPM_E_HIP, "hipEventCreate failed");
        hp = nullptr;
        if (rc == MPM_OK && hipHostMalloc(&hp, sizeof(float) * (size_t)nscenes, hipHostMallocDefault) != hipSuccess)
            rc = mpm_set_error(MPM_E_HIP, "hipHostMalloc failed");
        c->h_dt = (float*)hp;
        if (rc == MPM_OK && hipEventCreateWithFlags(&c->dt_ev, hipEventDisableTiming) != hipSuccess)
            rc = mpm_set_error(MPM_E_HIP, "hipEventCreate failed");
        hp = nullptr;
        // the overflow flag through the frame-done count (Counters::overflow .. frame_ndone)
        if (rc == MPM_OK && hipHostMalloc(&hp, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess)
            rc = mpm_set_error(MPM_E_HIP, "hipHostMalloc failed");
        c->h_fdone = (int*)hp;
        if (rc == MPM_OK && hipEventCreateWithFlags(&c->fdone_ev, hipEventDisableTiming) != hipSuccess)
            rc = mpm_set_error(MPM_E_HIP, "hipEventCreate failed");
    }
    if (rc == MPM_OK && hipMemset(c->keybuf[0], 0xff, (size_t)c->pcap * sizeof(int)) != hipSuccess)
        rc = mpm_set_error(MPM_E_HIP, "key buffer initialisation failed");
    return rc;
}

// Scene table entry of a free arena entry: no slots, no particles, inactive;
// its CG counts as done at once (cap 0).
mpm::SceneParams free_entry_params() {
    mpm::SceneParams p{};
    p.counts = 4;
    p.nb = 1;
    p.dx = 0.25f;
    p.inv_dx = 4.0f;
    return p;
}

int upload_tables(mpm_context* c, const std::vector<int>& slot_scene, int n_live, int slot_hw) {
    HIPCHECK(hipMemcpy(c->scn, c->table.data(), sizeof(mpm::SceneParams) * (size_t)c->nscenes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->slot_scene, slot_scene.data(), sizeof(int) * slot_scene.size(), hipMemcpyHostToDevice));
    mpm::Counters hc{};
    hc.n_live = n_live;
    hc.n_prev = n_live;
    hc.slot_hw = slot_hw;
    HIPCHECK(hipMemcpy(c->ctr, &hc, sizeof(hc), hipMemcpyHostToDevice));
    return MPM_OK;
}

int check_desc(const mpm_scene_desc& d, const char* who) {
    if (d.dim != 2 && d.dim != 3) return mpm_set_error(MPM_E_INVALID, std::string(who) + ": dim must be 2 or 3");
    if (d.counts < 4 || d.counts > 4096) return mpm_set_error(MPM_E_INVALID, std::string(who) + ": counts out of range");
    if (d.nplanes < 0 || d.nplanes > MPM_MAX_PLANES) return mpm_set_error(MPM_E_INVALID, std::string(who) + ": too many planes");
    if (d.nmaterials < 0 || d.nmaterials > MPM_MAX_MATERIALS)
        return mpm_set_error(MPM_E_INVALID, std::string(who) + ": nmaterials out of range");
    for (int k = 0; k < std::max(1, (int)d.nmaterials); ++k) {
        const mpm_material& m = k == 0 ? d.material : d.extra_materials[k - 1];
        // an EOS particle's pressure needs a positive density, bulk modulus and exponent
        if (m.eos && !(m.density > 0.0f && m.bulk_modulus >= 0.0f && m.gamma > 0.0f && std::isfinite(m.bulk_modulus) &&
                       std::isfinite(m.gamma) && std::isfinite(m.density)))
            return mpm_set_error(MPM_E_INVALID, std::string(who) + ": EOS material needs density > 0, K >= 0, gamma > 0");
    }
    return MPM_OK;
}
static_assert(mpm::MAX_MATERIALS == MPM_MAX_MATERIALS, "material table size");

// Particle inputs: finite positions, velocities and (when given) F and
// volumes, a positive Jp (det Fp, or an EOS particle's density) and a finite
// positive mass.
int check_particles(const char* who, int D, int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                    const float* mass, const float* volume) {
    for (int64_t i = 0; i < n; ++i) {
        bool ok = std::isfinite(mass[i]) && mass[i] > 0.0f;
        for (int a = 0; a < D; ++a) ok = ok && std::isfinite(x[i * D + a]) && std::isfinite(v[i * D + a]);
        if (F)
            for (int k = 0; k < D * D; ++k) ok = ok && std::isfinite(F[i * D * D + k]);
        if (Jp) ok = ok && std::isfinite(Jp[i]) && Jp[i] > 0.0f;
        if (volume) ok = ok && std::isfinite(volume[i]) && volume[i] >= 0.0f;
        if (!ok)
            return mpm_set_error(MPM_E_INVALID, std::string(who) + ": particle " + std::to_string(i) +
                                                    " has a non-finite value, or a mass or Jp that is not positive");
    }
    return MPM_OK;
}

// Material entry of particle i (material == nullptr: 0), checked against the
// scene's table; the Jp slot of an EOS particle starts at its density.
int particle_material(const mpm_scene_desc& d, const int32_t* material, int64_t i, int& id, bool& eos, float& rho0) {
    id = material ? material[i] : 0;
    const int nmat = std::max(1, (int)d.nmaterials);
    if (id < 0 || id >= nmat) return mpm_set_error(MPM_E_INVALID, "particle material outside the scene's table");
    const mpm_material& m = id == 0 ? d.material : d.extra_materials[id - 1];
    eos = m.eos != 0;
    rho0 = m.density;
    return MPM_OK;
}

int slots_of_desc(const mpm_scene_desc& d) {
    const int nb = slots_per_axis(d.dim, d.counts);
    return d.dim == 3 ? nb * nb * nb : nb * nb;
}

int create_batch(const mpm_scene_desc* descs, const int64_t* capacities, int nscenes, mpm_handle* out) {
    if (!descs || !capacities || !out || nscenes < 1) return mpm_set_error(MPM_E_INVALID, "mpm_batch_create: bad argument");
    const int D = descs[0].dim;
    int64_t pcap = 0, slots = 0;
    for (int s = 0; s < nscenes; ++s) {
        const mpm_scene_desc& d = descs[s];
        int rc = check_desc(d, "mpm_create");
        if (rc) return rc;
        if (d.dim != D) return mpm_set_error(MPM_E_INVALID, "mpm_batch_create: scenes of one batch share dim");
        if (capacities[s] < 1 || capacities[s] > (int64_t)1 << 30) return mpm_set_error(MPM_E_INVALID, "mpm_create: bad capacity");
        slots += slots_of_desc(d);
        pcap += capacities[s];
    }
    // keys are slot*64 + cell in 32 bits
    if (slots * 64 > ((int64_t)1 << 31) - 1) return mpm_set_error(MPM_E_INVALID, "mpm_create: grid too large");
    if (pcap > ((int64_t)1 << 31) - 1) return mpm_set_error(MPM_E_INVALID, "mpm_create: too many particles");
    mpm_context* c = new mpm_context;
    c->D = D;
    c->nscenes = nscenes;
    c->sc.resize((size_t)nscenes);
    c->table.resize((size_t)nscenes);
    std::vector<int> slot_scene((size_t)slots);
    int64_t sb = 0, pb = 0;
    for (int s = 0; s < nscenes; ++s) {
        const mpm_scene_desc& d = descs[s];
        const int ns = slots_of_desc(d);
        c->sc[(size_t)s].desc = d;
        c->sc[(size_t)s].capacity = capacities[s];
        c->sc[(size_t)s].pbase = pb;
        c->table[(size_t)s] = scene_params(d, (int)sb, ns);
        std::fill(slot_scene.begin() + sb, slot_scene.begin() + sb + ns, s);
        sb += ns;
        pb += capacities[s];
    }
    c->nslots = (int)sb;
    c->pcap = pcap;
    // Block capacity: every slot of small grids, else pcap/16 blocks -- about
    // 8x the active blocks (home blocks plus halo) of scenes packed at 8
    // particles per cell.  Overflow makes the substep a no-op and is reported.
    c->bcap = (int)std::min<int64_t>(c->nslots, std::max<int64_t>(4096, pcap / 16));
    int rc = allocate_context(c);
    // the particle buffer spans every scene's range; empty entries carry key -1
    if (rc == MPM_OK) rc = upload_tables(c, slot_scene, (int)c->pcap, c->nslots);
    if (rc != MPM_OK) {
        std::string msg = g_err;
        mpm_destroy(c);
        return mpm_set_error(rc, msg);
    }
    *out = c;
    return MPM_OK;
}

// ---- serving arena ------------------------------------------------------------
int alloc_slots(mpm_context* c, int count) {
    for (size_t k = 0; k < c->free_slots.size(); ++k) {
        auto& r = c->free_slots[k];
        if (r.second >= count) {
            const int first = r.first;
            r.first += count;
            r.second -= count;
            if (r.second == 0) c->free_slots.erase(c->free_slots.begin() + (long)k);
            return first;
        }
    }
    return -1;
}

void free_slot_range(mpm_context* c, int first, int count) {
    auto& f = c->free_slots;
    auto it = std::lower_bound(f.begin(), f.end(), std::make_pair(first, 0));
    it = f.insert(it, {first, count});
    size_t k = (size_t)(it - f.begin());
    if (k + 1 < f.size() && f[k].first + f[k].second == f[k + 1].first) {  // merge with the next
        f[k].second += f[k + 1].second;
        f.erase(f.begin() + (long)k + 1);
    }
    if (k > 0 && f[k - 1].first + f[k - 1].second == f[k].first) {  // merge with the previous
        f[k - 1].second += f[k].second;
        f.erase(f.begin() + (long)k);
    }
}

// End of the highest allocated slot range of an arena, for the binning's scan.
int update_slot_hw(mpm_context* c) {
    int hw = 0;
    for (int e = 0; e < c->nscenes; ++e)
        if (c->entry_used[(size_t)e]) hw = std::max(hw, c->table[(size_t)e].slot_base + c->table[(size_t)e].nslots);
    HIPCHECK(hipMemcpy(&c->ctr->slot_hw, &hw, sizeof(int), hipMemcpyHostToDevice));
    return MPM_OK;
}

int fill_slot_scene(mpm_context* c, int first, int count, int value) {
    std::vector<int> h((size_t)count, value);
    HIPCHECK(hipMemcpy(c->slot_scene + first, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return MPM_OK;
}

// Writes scene s's particles into its range of the (unbinned) buffer.
int upload_scene(mpm_context* c, int s, int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                 const float* mass, const float* volume, const int32_t* material) {
    SceneHost& sh = c->sc[(size_t)s];
    if (n < 0 || n > sh.capacity) return mpm_set_error(MPM_E_CAPACITY, "mpm_upload_particles: n exceeds capacity");
    {
        const int rc = check_particles("mpm_upload_particles", c->D, n, x, v, F, Jp, mass, volume);
        if (rc) return rc;
    }
    const int D = c->D, NF = nfields(D);
    const int64_t cap = sh.capacity;
    std::vector<float> h((size_t)NF * (size_t)cap, 0.0f);
    auto f = [&](int field, int64_t i) -> float& { return h[(size_t)field * (size_t)cap + (size_t)i]; };
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) {
            f(a, i) = x[i * D + a];
            f(D + a, i) = v[i * D + a];
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) f(2 * D + a * D + b, i) = F ? F[(i * D + a) * D + b] : (a == b ? 1.0f : 0.0f);
        int id;
        bool eos;
        float rho0;
        const int rc = particle_material(sh.desc, material, i, id, eos, rho0);
        if (rc) return rc;
        f(2 * D + D * D, i) = Jp ? Jp[i] : (eos ? rho0 : 1.0f);
        f(2 * D + D * D + 1, i) = mass[i];
        f(2 * D + D * D + 2, i) = volume ? volume[i] : 0.0f;
        std::memcpy(&f(2 * D + D * D + 3, i), &id, sizeof(int));
    }
    std::vector<int> ids((size_t)cap), keys((size_t)cap), scene((size_t)cap, s);
    for (int64_t i = 0; i < cap; ++i) {
        ids[(size_t)i] = (int)i;
        keys[(size_t)i] = i < n ? 0 : -1;
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    const int64_t b = sh.pbase;
    for (int fld = 0; fld < NF; ++fld)
        HIPCHECK(hipMemcpy(c->pbuf[0] + (size_t)fld * c->pcap + b, &h[(size_t)fld * cap], (size_t)cap * sizeof(float),
                           hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->idbuf[0] + b, ids.data(), (size_t)cap * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->keybuf[0] + b, keys.data(), (size_t)cap * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->scbuf[0] + b, scene.data(), (size_t)cap * sizeof(int), hipMemcpyHostToDevice));
    const int waiting = volume == nullptr ? 1 : 0;
    HIPCHECK(hipMemcpy(&c->scn[s].waiting, &waiting, sizeof(int), hipMemcpyHostToDevice));
    c->table[(size_t)s].waiting = waiting;
    mpm::SceneState st{};
    HIPCHECK(hipMemcpy(c->sst + s, &st, sizeof(st), hipMemcpyHostToDevice));
    sh.uploaded = true;
    sh.n_uploaded = n;
    return MPM_OK;
}

// Per-scene max |V|^2 of the freshly uploaded buffer.
int upload_max_speed(mpm_context* c) {
    mpm::KP p = make_kp(c);
    if (c->D == 3) mpm::launch_max_speed<3>(p, c->stream, c->grid_particle);
    else mpm::launch_max_speed<2>(p, c->stream, c->grid_particle);
    return check_launch("max speed");
}

int download_scene(mpm_context* c, int s, int64_t cap, float* x, float* v, float* F, float* Jp, float* mass,
                   float* volume, int32_t* ids, int64_t* n_out, int32_t* material = nullptr) {
    mpm::Counters hc{};
    int rc = read_counters(c, hc);
    if (rc) return rc;
    int64_t b0 = 0, b1 = 0;
    if (c->ever_binned && !hc.overflow) {
        mpm::SceneState st{};
        rc = read_scene(c, s, st, nullptr);
        if (rc) return rc;
        b0 = st.pbegin;
        b1 = st.pend;
    } else if (c->ever_binned) {  // an overflowing binning kept the buffer as it was
        b0 = 0;
        b1 = hc.n_live;
    } else {
        b0 = c->sc[(size_t)s].pbase;
        b1 = b0 + c->sc[(size_t)s].capacity;
    }
    const int D = c->D, NF = nfields(D);
    const int64_t n = std::max<int64_t>(b1 - b0, 0);
    std::vector<float> h((size_t)NF * (size_t)std::max<int64_t>(n, 1));
    std::vector<int> id((size_t)std::max<int64_t>(n, 1)), key((size_t)std::max<int64_t>(n, 1)),
        scn((size_t)std::max<int64_t>(n, 1));
    if (n > 0) {
        for (int fld = 0; fld < NF; ++fld)
            HIPCHECK(hipMemcpy(&h[(size_t)fld * n], c->pbuf[c->cur] + (size_t)fld * c->pcap + b0, (size_t)n * sizeof(float),
                               hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(id.data(), c->idbuf[c->cur] + b0, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(key.data(), c->keybuf[c->cur] + b0, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(scn.data(), c->scbuf[c->cur] + b0, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    }
    std::vector<int64_t> order;
    order.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        if (key[(size_t)i] >= 0 && scn[(size_t)i] == s) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return id[(size_t)a] < id[(size_t)b]; });
    const int64_t live = (int64_t)order.size();
    if (n_out) *n_out = live;
    if (live > cap) return mpm_set_error(MPM_E_CAPACITY, "mpm_download_particles: cap too small");
    auto g = [&](int fld, int64_t i) { return h[(size_t)fld * n + (size_t)i]; };
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
        if (material) {
            const float mf = g(2 * D + D * D + 3, i);
            std::memcpy(&material[k], &mf, sizeof(int32_t));
        }
        if (ids) ids[k] = id[(size_t)i];
    }
    return MPM_OK;
}

void invalidate_graphs(mpm_context* c) {
    for (int k = 0; k < 2; ++k) {
        if (c->gexec[k]) {
            (void)hipGraphExecDestroy(c->gexec[k]);
            c->gexec[k] = nullptr;
        }
        if (c->fexec[k]) {
            (void)hipGraphExecDestroy(c->fexec[k]);
            c->fexec[k] = nullptr;
        }
    }
}

// Device frame loop (MPM_Driver::Advance_To_Target_Time for every scene, the
// dt rule on the device): a graph of FRAME_K substeps, each one k_frame_dt
// (the next dt of every scene from its last G2P) and the captured substep,
// replayed until every scene reached its frame end.  The host reads one
// count per replay instead of the max speeds after every substep.  FRAME_K is
// even, so a replay leaves the particle buffers at the parity it found.
constexpr int FRAME_K = 4;

int capture_frame_graph(mpm_context* c, int parity) {
    const bool timing = c->timing;
    c->timing = false;
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = MPM_OK;
    for (int k = 0; k < FRAME_K && rc == MPM_OK; ++k) {
        mpm::launch_frame_dt(make_kp(c), c->stream, c->frame, c->d_dt);
        rc = c->D == 3 ? substep<3>(c, nullptr, true) : substep<2>(c, nullptr, true);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = timing;
    c->cur = parity;  // capture records launches only
    if (rc != MPM_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc != MPM_OK ? rc : mpm_set_error(MPM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&c->fexec[parity], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->fexec[parity] = nullptr;
        return mpm_set_error(MPM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    return MPM_OK;
}

// Advances every scene s from time[s] to target[s] (float, as the host loop)
// with the device dt rule; writes the times reached and the substeps taken.
int device_frames(mpm_context* c, const float* time, const float* target, float* time_out, int* steps_out) {
    const int n = c->nscenes;
    std::vector<mpm::FrameState> f((size_t)n);
    int ndone = 0;
    for (int s = 0; s < n; ++s) {
        const mpm_scene_desc& d = c->sc[(size_t)s].desc;
        mpm::FrameState& st = f[(size_t)s];
        st.time = time[s];
        st.target = target[s];
        st.cfl = d.cfl;
        st.dx = 1.0f / (float)d.counts;
        st.min_dt = d.min_dt;
        st.max_dt = d.max_dt;
        st.last_dt = c->sc[(size_t)s].last_dt;
        st.done = !(target[s] > time[s]) ? 1 : 0;
        st.steps = 0;
        ndone += st.done;
    }
    HIPCHECK(hipMemcpyAsync(c->frame, f.data(), sizeof(mpm::FrameState) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    while (ndone < n) {
        const int parity = c->cur;
        if (!c->fexec[parity]) {
            const int rc = capture_frame_graph(c, parity);
            if (rc) return rc;
        }
        HIPCHECK(hipGraphLaunch(c->fexec[parity], c->stream));
        c->binned = false;
        c->ever_binned = true;
        c->stepped = true;
        // one copy reads the overflow flag and the frame-done count
        static_assert(offsetof(mpm::Counters, frame_ndone) == offsetof(mpm::Counters, overflow) + 3 * sizeof(int),
                      "layout");
        HIPCHECK(hipMemcpyAsync(c->h_fdone, &c->ctr->overflow, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipEventRecord(c->fdone_ev, c->stream));
        HIPCHECK(hipEventSynchronize(c->fdone_ev));
        // a binning over the block capacity made the replayed substeps no-ops
        if (c->h_fdone[0] != 0) return mpm_set_error(MPM_E_CAPACITY, "active blocks exceed the block capacity");
        ndone = c->h_fdone[3];
    }
    HIPCHECK(hipMemcpy(f.data(), c->frame, sizeof(mpm::FrameState) * (size_t)n, hipMemcpyDeviceToHost));
    for (int s = 0; s < n; ++s) {
        const mpm::FrameState& st = f[(size_t)s];
        if (time_out) time_out[s] = st.time;
        if (steps_out) steps_out[s] = st.steps;
        if (st.steps > 0) c->sc[(size_t)s].last_dt = st.last_dt;
    }
    return check_launch("device frame loop");
}

// One substep as a graph replay: the dts go to d_dt through pinned staging,
// then the graph of the current buffer parity (captured on first use) runs.
int graph_substep(mpm_context* c, const float* dts) {
    if (c->dt_pending) {  // the staging buffer is free once the last copy ran
        HIPCHECK(hipEventSynchronize(c->dt_ev));
        c->dt_pending = false;
    }
    std::memcpy(c->h_dt, dts, sizeof(float) * (size_t)c->nscenes);
    HIPCHECK(hipMemcpyAsync(c->d_dt, c->h_dt, sizeof(float) * (size_t)c->nscenes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipEventRecord(c->dt_ev, c->stream));
    c->dt_pending = true;
    const int parity = c->cur;
    if (!c->gexec[parity]) {
        const bool timing = c->timing;
        c->timing = false;
        HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = c->D == 3 ? substep<3>(c, nullptr, true) : substep<2>(c, nullptr, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->timing = timing;
        c->cur = parity;  // capture records launches only: the buffers have not moved
        if (rc != MPM_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc != MPM_OK ? rc : mpm_set_error(MPM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        }
        const hipError_t ei = hipGraphInstantiate(&c->gexec[parity], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->gexec[parity] = nullptr;
            return mpm_set_error(MPM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
    }
    HIPCHECK(hipGraphLaunch(c->gexec[parity], c->stream));
    c->cur = parity ^ 1;  // the replay binned once
    c->binned = false;
    c->ever_binned = true;
    return MPM_OK;
}

int do_substep(mpm_context* c, const float* dts) {
    // a substep after mpm_add_particles runs launched: it adds the waiting
    // pass; so does a batch substep with one large active scene, which then
    // takes the single-scene CG kernels with the CG's done poll
    int active = 0, last = -1;
    for (int s = 0; s < c->nscenes; ++s)
        if (dts[s] > 0.0f) { ++active; last = s; }
    const bool solo = c->nscenes > 1 && active == 1 && c->sc[(size_t)last].n_uploaded >= SOLO_LAUNCHED_MIN;
    int rc = c->graph && !c->emitted && !solo ? graph_substep(c, dts)
             : c->D == 3                      ? substep<3>(c, dts)
                                              : substep<2>(c, dts);
    if (rc) return rc;
    for (int s = 0; s < c->nscenes; ++s)
        if (dts[s] > 0.0f) c->sc[(size_t)s].last_dt = dts[s];
    c->stepped = true;
    rc = check_launch("mpm_substep");
    if (rc) return rc;
    return collect_timing(c);
}

// Max |V| per scene after the last G2P.  Also reports a block-capacity
// overflow of the last binning (the substeps it made no-ops may have been
// queued without a poll: replayed graphs).
int max_velocities(mpm_context* c, float* out) {
    HIPCHECK(hipStreamSynchronize(c->stream));
    std::vector<mpm::SceneState> st((size_t)c->nscenes);
    HIPCHECK(hipMemcpy(st.data(), c->sst, sizeof(mpm::SceneState) * st.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < c->nscenes; ++s) {
        float v2;
        std::memcpy(&v2, &st[(size_t)s].max_v2, sizeof(v2));
        out[s] = std::sqrt(v2);
    }
    int overflow = 0;
    HIPCHECK(hipMemcpy(&overflow, &c->ctr->overflow, sizeof(int), hipMemcpyDeviceToHost));
    if (overflow) return mpm_set_error(MPM_E_CAPACITY, "active blocks exceed the block capacity");
    return MPM_OK;
}

// ---- launch-shape autotuner ----------------------------------------------
// Size bucket: dimension, batch of one or more, log2 of particle capacity
// and of block slots.
using TuneKey = std::tuple<int, int, int, int>;
std::mutex g_tune_mutex;
std::map<TuneKey, mpm_launch_config> g_tune_cache;

int ilog2(int64_t v) {
    int r = 0;
    while (((int64_t)1 << (r + 1)) <= v) ++r;
    return r;
}
TuneKey tune_key(const mpm_context* c) { return TuneKey(c->D, c->nscenes == 1 ? 1 : 0, ilog2(c->pcap), ilog2(c->nslots)); }
int round8(int64_t v) { return (int)std::max<int64_t>(8, (v + 7) / 8 * 8); }

int check_config(const mpm_context* c, const mpm_launch_config& g) {
    auto ok = [](int v, int hi) { return v >= 8 && v <= hi && v % 8 == 0; };
    if (!ok(g.block_groups, 1 << 20) || !ok(g.node_groups, (int)std::min<size_t>(c->nred, MAX_NODE_GROUPS)) ||
        !ok(g.particle_groups, 1 << 20) || !ok(g.operator_groups, 1 << 20))
        return mpm_set_error(MPM_E_INVALID, "launch config: work-group counts must be multiples of 8 in range "
                                            "(node_groups at most 2048)");
    if (g.operator_mode != 0 && g.operator_mode != 1)
        return mpm_set_error(MPM_E_INVALID, "launch config: operator_mode must be 0 or 1");
    return MPM_OK;
}

mpm_launch_config current_config(const mpm_context* c) {
    mpm_launch_config g{};
    g.block_groups = c->grid_block;
    g.node_groups = c->grid_node;
    g.particle_groups = c->grid_particle;
    g.operator_mode = c->op_mode;
    g.operator_groups = c->grid_op;
    return g;
}

void apply_config(mpm_context* c, const mpm_launch_config& g) {
    if (g.block_groups != c->grid_block || g.node_groups != c->grid_node || g.particle_groups != c->grid_particle ||
        g.operator_mode != c->op_mode || g.operator_groups != c->grid_op)
        invalidate_graphs(c);
    c->grid_block = g.block_groups;
    c->grid_node = g.node_groups;
    c->grid_particle = g.particle_groups;
    c->op_mode = g.operator_mode;
    c->grid_op = g.operator_groups;
}

// Re-opens every scene's CG solve so that operator launches do work.
int reopen_cg(mpm_context* c) {
    std::vector<mpm::CGState> g((size_t)c->nscenes);
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(g.data(), c->cg, sizeof(mpm::CGState) * g.size(), hipMemcpyDeviceToHost));
    for (auto& x : g) x.done = 0;
    HIPCHECK(hipMemcpy(c->cg, g.data(), sizeof(mpm::CGState) * g.size(), hipMemcpyHostToDevice));
    const int zero = 0;
    HIPCHECK(hipMemcpy(&c->ctr->cg_ndone, &zero, sizeof(int), hipMemcpyHostToDevice));
    return MPM_OK;
}

// Median time of five runs of seq (after one untimed run); prep runs
// before each, outside the timed span.
template <class Prep, class Seq>
int time_median(mpm_context* c, Prep&& prep, Seq&& seq, float& ms_out) {
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    float t[5];
    int rc = MPM_OK;
    for (int r = -1; r < 5 && rc == MPM_OK; ++r) {
        rc = prep();
        if (rc) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = mpm_set_error(MPM_E_HIP, "hipEventRecord"); break; }
        seq();
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = mpm_set_error(MPM_E_HIP, "autotune timing failed");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r >= 0) t[r] = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    std::sort(t, t + 5);
    ms_out = t[2];
    return check_launch("autotune");
}

// Times each candidate work-group count on a sequence of the kernels it
// shapes, weighted as in a substep of a few CG iterations:
//   per-block:    P2G + particle force + 8 CG operator applications;
//   per-node:     normalize + grid update + 8 CG restart node passes;
//   per-particle: binning (bin, scan, block set-up, reorder) twice.
// Every kernel of a sequence is idempotent on the binned particles, so the
// particle state does not change.
template <int D>
int autotune(mpm_context* c, const float* dts, mpm_launch_config& best) {
    struct TimingOff {  // no per-class events while tuning
        mpm_context* c;
        bool saved;
        ~TimingOff() { c->timing = saved; }
    } timing_off{c, c->timing};
    c->timing = false;
    set_step(c, dts, 0.0f, false, false);
    stage_bin<D>(c);
    mpm::Counters h{};
    int rc = read_counters(c, h);
    if (rc) return rc;
    if (h.overflow) return mpm_set_error(MPM_E_CAPACITY, "mpm_autotune: active blocks exceed the block capacity");
    const bool one = c->nscenes == 1;
    mpm_launch_config cur = current_config(c);
    best = cur;
    // a grid and CG set-up on the binned particles
    {
        mpm::KP p = make_kp(c);
        mpm::launch_p2g<D>(p, c->stream, c->grid_block);
        mpm::launch_normalize<D>(p, c->stream, c->grid_node);
        mpm::launch_particle_force<D>(p, c->stream, c->grid_block, false, true);
        mpm::launch_grid_update<D>(p, c->stream, c->grid_node, true, true, true, false);
        launch_operator<D>(c, p, true, 0, one);
        mpm::launch_cg_restart_node<D>(p, c->stream, c->grid_node, one);
    }
    auto no_prep = [] { return MPM_OK; };
    auto reopen = [&] { return reopen_cg(c); };
    // the CG operator: four waves per block, or one wave per block with a
    // lane's cell accumulated in registers
    std::vector<int> cand;
    for (int g : {1024, 2048, 4096, 8192, 16384, 32768})
        if (g <= round8(h.nblocks) || g == 1024) cand.push_back(g);
    if (std::find(cand.begin(), cand.end(), cur.block_groups) == cand.end()) cand.push_back(cur.block_groups);
    std::vector<std::pair<int, int>> ops;
    for (int g : cand) ops.push_back({0, g});
    for (int g : {256, 512, 1024, 2048, 4096})
        if (g <= round8(h.nblocks / 4 + 8) || g == 256) ops.push_back({1, g});
    if (std::find(ops.begin(), ops.end(), std::make_pair(cur.operator_mode, cur.operator_groups)) == ops.end())
        ops.push_back({cur.operator_mode, cur.operator_groups});
    float best_ms = FLT_MAX;
    for (const auto& og : ops) {
        const mpm::KP p = make_kp(c);
        float ms = 0.0f;
        rc = time_median(c, reopen, [&] {
            for (int k = 0; k < 8; ++k) launch_operator_as<D>(c, p, og.first, og.second, false, 1, one);
        }, ms);
        if (rc) return rc;
        if (ms < best_ms) {
            best_ms = ms;
            best.operator_mode = og.first;
            best.operator_groups = og.second;
            best.operator_ms = ms;
        }
    }
    // per-block kernels
    best_ms = FLT_MAX;
    for (int g : cand) {
        const mpm::KP p = make_kp(c);
        float ms = 0.0f;
        rc = time_median(c, no_prep, [&] {
            mpm::launch_p2g<D>(p, c->stream, g);
            mpm::launch_particle_force<D>(p, c->stream, g, false, true);
        }, ms);
        if (rc) return rc;
        if (ms < best_ms) { best_ms = ms; best.block_groups = g; best.block_ms = ms; }
    }
    // per-node kernels
    cand.clear();
    const int node_hi = (int)std::min<size_t>(c->nred, MAX_NODE_GROUPS);
    for (int g : {128, 256, 512, 1024, 2048})
        if (g <= node_hi && (g <= round8(h.nblocks / 4 + 8) || g == 128)) cand.push_back(g);
    if (std::find(cand.begin(), cand.end(), cur.node_groups) == cand.end()) cand.push_back(cur.node_groups);
    best_ms = FLT_MAX;
    for (int g : cand) {
        const mpm::KP p = make_kp(c);
        float ms = 0.0f;
        rc = time_median(c, no_prep, [&] {
            mpm::launch_normalize<D>(p, c->stream, g);
            mpm::launch_grid_update<D>(p, c->stream, g, true, true, true, false);
            for (int k = 0; k < 8; ++k) mpm::launch_cg_restart_node<D>(p, c->stream, g, one);
        }, ms);
        if (rc) return rc;
        if (ms < best_ms) { best_ms = ms; best.node_groups = g; best.node_ms = ms; }
    }
    // per-particle kernels: 1 to 16 particles per thread
    cand.clear();
    for (int ppt : {1, 2, 4, 8, 16}) {
        const int g = round8(((int64_t)h.n_live + (int64_t)mpm::WG * ppt - 1) / ((int64_t)mpm::WG * ppt));
        if (g <= (1 << 20) && std::find(cand.begin(), cand.end(), g) == cand.end()) cand.push_back(g);
    }
    if (std::find(cand.begin(), cand.end(), cur.particle_groups) == cand.end()) cand.push_back(cur.particle_groups);
    best_ms = FLT_MAX;
    const int keep_particle = c->grid_particle;
    for (int g : cand) {
        float ms = 0.0f;
        c->grid_particle = g;
        rc = time_median(c, no_prep, [&] {
            stage_bin<D>(c);
            stage_bin<D>(c);
        }, ms);
        if (rc) { c->grid_particle = keep_particle; return rc; }
        if (ms < best_ms) { best_ms = ms; best.particle_groups = g; best.particle_ms = ms; }
    }
    c->grid_particle = keep_particle;
    // the grid and CG buffers hold tuning scratch; the particles stay binned
    c->binned = true;
    return MPM_OK;
}

// The reference's dt rule (MPM_Driver.cpp:55-80) for one scene, in float.
float next_dt(const mpm_scene_desc& d, float max_v, float time, float target, bool& last) {
    const float dx_min = 1.0f / (float)d.counts;
    float dt = std::max(d.min_dt, std::min(d.max_dt, d.cfl * dx_min / std::max(max_v, 1e-2f)));
    last = false;
    if (target - time < dt * 1.001f) {
        dt = target - time;
        last = true;
    } else if (target - time < 2.0f * dt) {
        dt = (target - time) * 0.5f;
    }
    return dt;
}

// Dense views of scene 0 over node indices [-1, counts]^dim.
int dense_index(const mpm_context* c, int slot, int l, int64_t& o) {
    const int D = c->D, B = block_edge(D), n = c->sc[0].desc.counts;
    const int nb = c->table[0].nb;
    int bb[3];
    int s = slot - c->table[0].slot_base;
    for (int a = D - 1; a >= 0; --a) { bb[a] = s % nb; s /= nb; }
    int lc[3];
    if (D == 3) { lc[0] = l >> 4; lc[1] = (l >> 2) & 3; lc[2] = l & 3; }
    else { lc[0] = l >> 3; lc[1] = l & 7; }
    o = 0;
    for (int a = 0; a < D; ++a) {
        const int i = B * bb[a] + lc[a] - B;
        if (i < -1 || i > n) return 0;
        o = o * (n + 2) + (i + 1);
    }
    return 1;
}

}  // namespace

// ---- internal entry points of the request server -----------------------------
namespace mpm_host {

UnitBytes unit_bytes(int D) {
    const int64_t nf = nfields(D), DD = (int64_t)D * D;
    UnitBytes u;
    // two particle buffers (fields, id, key, scene) + rank, in-cell order, rotation cache, operator record
    u.particle = 2 * (nf + 3) * 4 + 4 + 4 + DD * 4 + (int64_t)mpm::op_record_floats(D) * 4;
    // count, touched, active, owner + 64 cell counters
    u.slot = 4 * 4 + 64 * 4;
    // block tables, neighbours, cell ranges, tiles, 64 nodes of grid channels, partials
    const int64_t nbr = D == 3 ? 27 : 9, tile = D == 3 ? 216 : 100;
    u.block = 3 * 4 + 2 * nbr * 4 + 6 * 4 + 65 * 4 + (D + 1) * tile * 4 + 64 * (1 + 5 * D + 1) * 4 + 3 * 8 + 2 * 4;
    return u;
}

int pool_create(int D, int max_scenes, int64_t pcap, int64_t nslots, int64_t bcap, int cg_max, mpm_context** out) {
    if (D != 2 && D != 3) return mpm_set_error(MPM_E_INVALID, "arena: dim must be 2 or 3");
    if (max_scenes < 1 || max_scenes > (1 << 20)) return mpm_set_error(MPM_E_INVALID, "arena: bad scene count");
    if (pcap < 1 || pcap > ((int64_t)1 << 31) - 1) return mpm_set_error(MPM_E_INVALID, "arena: bad particle capacity");
    if (nslots < 1 || nslots * 64 > ((int64_t)1 << 31) - 1) return mpm_set_error(MPM_E_INVALID, "arena: bad slot capacity");
    if (bcap < 1 || bcap > nslots) return mpm_set_error(MPM_E_INVALID, "arena: block capacity must be in [1, slots]");
    if (cg_max < 1 || cg_max > MAX_GRAPH_CG_ITERATIONS) return mpm_set_error(MPM_E_INVALID, "arena: CG cap must be in [1, 256]");
    mpm_context* c = new mpm_context;
    c->D = D;
    c->nscenes = max_scenes;
    c->pool = true;
    c->pool_cg_max = cg_max;
    c->sc.resize((size_t)max_scenes);
    c->table.assign((size_t)max_scenes, free_entry_params());
    c->entry_used.assign((size_t)max_scenes, 0);
    c->pcap = pcap;
    c->nslots = (int)nslots;
    c->bcap = (int)bcap;
    c->free_slots.push_back({0, (int)nslots});
    int rc = allocate_context(c);
    if (rc == MPM_OK) rc = upload_tables(c, std::vector<int>((size_t)nslots, -1), 0, 0);
    if (rc != MPM_OK) {
        std::string msg = g_err;
        mpm_destroy(c);
        return mpm_set_error(rc, msg);
    }
    c->initialized = true;
    *out = c;
    return MPM_OK;
}

int slots_of(const mpm_scene_desc& d) { return slots_of_desc(d); }

int pool_largest_free_slots(const mpm_context* c) {
    int m = 0;
    for (const auto& r : c->free_slots) m = std::max(m, r.second);
    return m;
}

int live_particles(mpm_context* c, int64_t* n) {
    mpm::Counters h{};
    int rc = read_counters(c, h);
    if (rc) return rc;
    *n = h.n_live;
    if (h.overflow) return mpm_set_error(MPM_E_CAPACITY, "active blocks exceed the block capacity");
    return MPM_OK;
}

int pool_attach(mpm_context* c, const mpm_scene_desc& d, int64_t n, const float* x, const float* v,
                const float* mass, const int32_t* material, int* entry) {
    if (!c->pool) return mpm_set_error(MPM_E_INVALID, "arena: not an arena");
    int rc = check_desc(d, "arena attach");
    if (rc) return rc;
    if (d.dim != c->D) return mpm_set_error(MPM_E_INVALID, "arena attach: dimension differs from the arena's");
    if (d.cg_max_iterations < 1 || d.cg_max_iterations > c->pool_cg_max)
        return mpm_set_error(MPM_E_INVALID, "arena attach: CG cap above the arena's");
    if (n < 1 || !x || !v || !mass) return mpm_set_error(MPM_E_INVALID, "arena attach: no particles");
    rc = check_particles("arena attach", c->D, n, x, v, nullptr, nullptr, mass, nullptr);
    if (rc) return rc;
    int e = -1;
    for (int k = 0; k < c->nscenes; ++k)
        if (!c->entry_used[(size_t)k]) { e = k; break; }
    if (e < 0) return mpm_set_error(MPM_E_CAPACITY, "arena attach: no free scene entry");
    int64_t n_live = 0;
    rc = live_particles(c, &n_live);
    if (rc) return rc;
    if (n_live + n > c->pcap) return mpm_set_error(MPM_E_CAPACITY, "arena attach: particle space exhausted");
    const int ns = slots_of_desc(d);
    const int first = alloc_slots(c, ns);
    if (first < 0) return mpm_set_error(MPM_E_CAPACITY, "arena attach: no free slot range");
    mpm::SceneParams sp = scene_params(d, first, ns);
    sp.waiting = 1;
    c->table[(size_t)e] = sp;
    SceneHost& sh = c->sc[(size_t)e];
    sh = SceneHost{};
    sh.desc = d;
    sh.capacity = n;
    sh.pbase = n_live;
    sh.n_uploaded = n;
    sh.uploaded = true;
    // particles appended after the live ones in the current buffer
    const int D = c->D, NF = nfields(D);
    std::vector<float> h((size_t)NF * (size_t)n);
    auto f = [&](int field, int64_t i) -> float& { return h[(size_t)field * (size_t)n + (size_t)i]; };
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) {
            f(a, i) = x[i * D + a];
            f(D + a, i) = v[i * D + a];
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) f(2 * D + a * D + b, i) = a == b ? 1.0f : 0.0f;
        int id;
        bool eos;
        float rho0;
        rc = particle_material(d, material, i, id, eos, rho0);
        if (rc) return rc;
        f(2 * D + D * D, i) = eos ? rho0 : 1.0f;
        f(2 * D + D * D + 1, i) = mass[i];
        f(2 * D + D * D + 2, i) = 0.0f;
        std::memcpy(&f(2 * D + D * D + 3, i), &id, sizeof(int));
    }
    std::vector<int> ids((size_t)n), keys((size_t)n, 0), scene((size_t)n, e);
    std::iota(ids.begin(), ids.end(), 0);
    const int cur = c->cur;
    for (int fld = 0; fld < NF; ++fld)
        HIPCHECK(hipMemcpy(c->pbuf[cur] + (size_t)fld * c->pcap + n_live, &h[(size_t)fld * n], (size_t)n * sizeof(float),
                           hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->idbuf[cur] + n_live, ids.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->keybuf[cur] + n_live, keys.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->scbuf[cur] + n_live, scene.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->scn + e, &sp, sizeof(sp), hipMemcpyHostToDevice));
    mpm::SceneState st{};
    HIPCHECK(hipMemcpy(c->sst + e, &st, sizeof(st), hipMemcpyHostToDevice));
    rc = fill_slot_scene(c, first, ns, e);
    if (rc) return rc;
    const int nl = (int)(n_live + n);
    HIPCHECK(hipMemcpy(&c->ctr->n_live, &nl, sizeof(int), hipMemcpyHostToDevice));
    c->entry_used[(size_t)e] = 1;
    c->binned = false;
    rc = update_slot_hw(c);
    if (rc) return rc;
    // its max |V| for the first dt (the other scenes' maxima do not change)
    rc = upload_max_speed(c);
    if (rc) return rc;
    *entry = e;
    return MPM_OK;
}

int pool_detach(mpm_context* c, int e) {
    if (!c->pool || e < 0 || e >= c->nscenes || !c->entry_used[(size_t)e])
        return mpm_set_error(MPM_E_INVALID, "arena detach: no such scene");
    mpm::launch_remove_scene(make_kp(c), c->stream, c->grid_particle, e);
    int rc = check_launch("arena detach");
    if (rc) return rc;
    const mpm::SceneParams old = c->table[(size_t)e];
    rc = fill_slot_scene(c, old.slot_base, old.nslots, -1);
    if (rc) return rc;
    free_slot_range(c, old.slot_base, old.nslots);
    c->table[(size_t)e] = free_entry_params();
    HIPCHECK(hipMemcpy(c->scn + e, &c->table[(size_t)e], sizeof(mpm::SceneParams), hipMemcpyHostToDevice));
    mpm::SceneState st{};
    HIPCHECK(hipMemcpy(c->sst + e, &st, sizeof(st), hipMemcpyHostToDevice));
    c->sc[(size_t)e] = SceneHost{};
    c->entry_used[(size_t)e] = 0;
    return update_slot_hw(c);
}

int check_particles(const char* who, int D, int64_t n, const float* x, const float* v, const float* F,
                    const float* Jp, const float* mass, const float* volume) {
    return ::check_particles(who, D, n, x, v, F, Jp, mass, volume);
}
int substep(mpm_context* c, const float* dts) { return do_substep(c, dts); }
int max_velocities(mpm_context* c, float* out) { return ::max_velocities(c, out); }
int download(mpm_context* c, int scene, int64_t cap, float* x, float* v, float* F, float* Jp, float* mass,
             float* volume, int32_t* ids, int64_t* n_out) {
    return download_scene(c, scene, cap, x, v, F, Jp, mass, volume, ids, n_out);
}
int stats(mpm_context* c, int scene, mpm_step_stats* out) { return fill_stats(c, scene, out); }
float next_dt(const mpm_scene_desc& d, float max_v, float time, float target, bool& last) {
    return ::next_dt(d, max_v, time, target, last);
}

}  // namespace mpm_host

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
    return create_batch(desc, &capacity, 1, out);
}

int mpm_destroy(mpm_handle c) {
    if (!c) return MPM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : c->allocations) (void)hipFree(p);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; ++k) if (c->done_ev[k]) (void)hipEventDestroy(c->done_ev[k]);
    if (c->h_done) (void)hipHostFree(c->h_done);
    invalidate_graphs(c);
    if (c->dt_ev) (void)hipEventDestroy(c->dt_ev);
    if (c->h_dt) (void)hipHostFree(c->h_dt);
    if (c->fdone_ev) (void)hipEventDestroy(c->fdone_ev);
    if (c->h_fdone) (void)hipHostFree(c->h_fdone);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MPM_OK;
}

int mpm_upload_particles(mpm_handle c, int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                         const float* mass, const float* volume) {
    return mpm_upload_particles_ex(c, n, x, v, F, Jp, mass, volume, nullptr);
}

int mpm_upload_particles_ex(mpm_handle c, int64_t n, const float* x, const float* v, const float* F, const float* Jp,
                            const float* mass, const float* volume, const int32_t* material) {
    // a scene may start empty (the fluid-source test): n = 0 needs no arrays
    if (!c || (n > 0 && (!x || !v || !mass))) return mpm_set_error(MPM_E_INVALID, "mpm_upload_particles: null argument");
    if (c->nscenes != 1) return mpm_set_error(MPM_E_INVALID, "mpm_upload_particles: batch handle, use mpm_batch_upload");
    // a new particle set: the buffer starts over in its unbinned layout
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->cur = 0;
    HIPCHECK(hipMemset(c->keybuf[0], 0xff, (size_t)c->pcap * sizeof(int)));
    mpm::Counters hc{};
    hc.slot_hw = c->nslots;
    hc.n_live = (int)c->pcap;
    hc.n_prev = (int)c->pcap;
    HIPCHECK(hipMemcpy(c->ctr, &hc, sizeof(hc), hipMemcpyHostToDevice));
    c->initialized = c->binned = c->ever_binned = c->stepped = false;
    int rc = upload_scene(c, 0, n, x, v, F, Jp, mass, volume, material);
    if (rc) return rc;
    return upload_max_speed(c);
}

// MPM_Example::Add_Fluid_Source / Allocate_Particle (MPM_Example.cpp:925-1027):
// n particles join scene `scene` of a running handle as waiting particles
// (F = I, Jp = 1 or an EOS material's density, volume from the grid at the
// next substep), after every live particle, with the next ids of the scene.
int mpm_add_particles(mpm_handle c, int32_t scene, int64_t n, const float* x, const float* v, const float* mass,
                      const int32_t* material) {
    if (!c || !x || !v || !mass) return mpm_set_error(MPM_E_INVALID, "mpm_add_particles: null argument");
    if (scene < 0 || scene >= c->nscenes) return mpm_set_error(MPM_E_INVALID, "mpm_add_particles: no such scene");
    if (c->pool) return mpm_set_error(MPM_E_INVALID, "mpm_add_particles: arena scenes are the server's");
    if (!c->initialized) return mpm_set_error(MPM_E_NOT_READY, "mpm_add_particles: call mpm_initialize first");
    if (n < 0) return mpm_set_error(MPM_E_INVALID, "mpm_add_particles: negative count");
    if (n == 0) return MPM_OK;
    SceneHost& sh = c->sc[(size_t)scene];
    const mpm_scene_desc& d = sh.desc;
    int rc = check_particles("mpm_add_particles", c->D, n, x, v, nullptr, nullptr, mass, nullptr);
    if (rc) return rc;
    int64_t n_live = 0;
    rc = mpm_host::live_particles(c, &n_live);
    if (rc) return rc;
    if (n_live + n > c->pcap) return mpm_set_error(MPM_E_CAPACITY, "mpm_add_particles: particle capacity exhausted");
    const int D = c->D, NF = nfields(D);
    std::vector<float> h((size_t)NF * (size_t)n);
    auto f = [&](int field, int64_t i) -> float& { return h[(size_t)field * (size_t)n + (size_t)i]; };
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) {
            f(a, i) = x[i * D + a];
            f(D + a, i) = v[i * D + a];
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) f(2 * D + a * D + b, i) = a == b ? 1.0f : 0.0f;
        int id;
        bool eos;
        float rho0;
        rc = particle_material(d, material, i, id, eos, rho0);
        if (rc) return rc;
        f(2 * D + D * D, i) = eos ? rho0 : 1.0f;
        f(2 * D + D * D + 1, i) = mass[i];
        f(2 * D + D * D + 2, i) = 0.0f;
        std::memcpy(&f(2 * D + D * D + 3, i), &id, sizeof(int));
    }
    std::vector<int> ids((size_t)n), keys((size_t)n, 0), sc_of((size_t)n, scene);
    std::iota(ids.begin(), ids.end(), (int)sh.n_uploaded);
    const int cur = c->cur;
    for (int fld = 0; fld < NF; ++fld)
        HIPCHECK(hipMemcpy(c->pbuf[cur] + (size_t)fld * c->pcap + n_live, &h[(size_t)fld * n], (size_t)n * sizeof(float),
                           hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->idbuf[cur] + n_live, ids.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->keybuf[cur] + n_live, keys.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->scbuf[cur] + n_live, sc_of.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    const int nl = (int)(n_live + n);
    HIPCHECK(hipMemcpy(&c->ctr->n_live, &nl, sizeof(int), hipMemcpyHostToDevice));
    c->table[(size_t)scene].waiting = 1;
    HIPCHECK(hipMemcpy(c->scn + scene, &c->table[(size_t)scene], sizeof(mpm::SceneParams), hipMemcpyHostToDevice));
    sh.n_uploaded += n;
    c->emitted = true;
    c->binned = false;
    return MPM_OK;
}

int mpm_download_particles(mpm_handle c, int64_t cap, float* x, float* v, float* F, float* Jp, float* mass,
                           float* volume, int32_t* ids, int64_t* n_out) {
    return mpm_download_particles_ex(c, cap, x, v, F, Jp, mass, volume, ids, nullptr, n_out);
}

int mpm_download_particles_ex(mpm_handle c, int64_t cap, float* x, float* v, float* F, float* Jp, float* mass,
                              float* volume, int32_t* ids, int32_t* material, int64_t* n_out) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_download_particles: null handle");
    return download_scene(c, 0, cap, x, v, F, Jp, mass, volume, ids, n_out, material);
}

int mpm_initialize(mpm_handle c) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_initialize: null handle");
    for (const SceneHost& s : c->sc)
        if (!s.uploaded) return mpm_set_error(MPM_E_NOT_READY, "mpm_initialize: no particles uploaded");
    int rc = c->D == 3 ? initialize<3>(c) : initialize<2>(c);
    if (rc) return rc;
    mpm::Counters h{};
    rc = read_counters(c, h);
    if (rc) return rc;
    // the particles need more active blocks than the handle stores: nothing
    // was built, and the handle waits for an upload that fits
    if (h.overflow) return mpm_set_error(MPM_E_CAPACITY, "mpm_initialize: active blocks exceed the block capacity");
    c->initialized = true;
    rc = check_launch("mpm_initialize");
    if (rc) return rc;
    return collect_timing(c);
}

int mpm_substep(mpm_handle c, float dt, mpm_step_stats* stats) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_substep: null handle");
    if (!c->initialized) return mpm_set_error(MPM_E_NOT_READY, "mpm_substep: call mpm_initialize first");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return mpm_set_error(MPM_E_INVALID, "mpm_substep: dt must be positive");
    std::vector<float> dts((size_t)c->nscenes, dt);
    int rc = do_substep(c, dts.data());
    if (rc) return rc;
    if (stats) return fill_stats(c, 0, stats);
    return MPM_OK;
}

int mpm_max_particle_velocity(mpm_handle c, float* out) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_max_particle_velocity: null argument");
    std::vector<float> v((size_t)c->nscenes);
    int rc = max_velocities(c, v.data());
    if (rc) return rc;
    *out = v[0];
    return MPM_OK;
}

int mpm_advance_to_time(mpm_handle c, double time_in, double target_in, int32_t* substeps_out, double* time_out,
                        mpm_step_stats* stats) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_advance_to_time: null handle");
    if (c->nscenes != 1) return mpm_set_error(MPM_E_INVALID, "mpm_advance_to_time: batch handle, use mpm_batch_advance_to_time");
    float time = (float)time_in;
    const float target_time = (float)target_in;
    bool done = false;
    int n = 0;
    if (c->graph && !c->emitted) {  // the device frame loop
        int rc = device_frames(c, &time, &target_time, &time, &n);
        if (rc) return rc;
        done = true;
    }
    while (!done) {
        float max_v = 0.0f;
        int rc = mpm_max_particle_velocity(c, &max_v);
        if (rc) return rc;
        const float dt = next_dt(c->sc[0].desc, max_v, time, target_time, done);
        if (!(dt > 0.0f)) break;
        rc = mpm_substep(c, dt, nullptr);
        if (rc) return rc;
        time += dt;
        ++n;
    }
    if (substeps_out) *substeps_out = n;
    if (time_out) *time_out = (double)time;
    if (stats) return fill_stats(c, 0, stats);
    return MPM_OK;
}

int mpm_synchronize(mpm_handle c) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_synchronize: null handle");
    HIPCHECK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_run_stage(mpm_handle c, int32_t stage, float dt) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_run_stage: null handle");
    for (const SceneHost& s : c->sc)
        if (!s.uploaded) return mpm_set_error(MPM_E_NOT_READY, "mpm_run_stage: no particles uploaded");
    if (stage != MPM_STAGE_BIN && stage != MPM_STAGE_CONSTITUTIVE && !c->binned)
        return mpm_set_error(MPM_E_NOT_READY, "mpm_run_stage: particles moved since the last MPM_STAGE_BIN");
    // the CG operator reads the records the constitutive update wrote, in the
    // order of the binning it ran on
    if (stage == MPM_STAGE_CG && !c->records_fresh)
        return mpm_set_error(MPM_E_NOT_READY, "mpm_run_stage: MPM_STAGE_CG needs MPM_STAGE_CONSTITUTIVE after the last MPM_STAGE_BIN");
    int rc = c->D == 3 ? run_stage<3>(c, stage, dt) : run_stage<2>(c, stage, dt);
    if (rc) return rc;
    if (stage == MPM_STAGE_WAITING || stage == MPM_STAGE_P2G) c->initialized = true;
    if (stage == MPM_STAGE_G2P && dt > 0.0f) {
        for (SceneHost& s : c->sc) s.last_dt = dt;
        c->stepped = true;
    }
    rc = check_launch("mpm_run_stage");
    if (rc) return rc;
    return collect_timing(c);
}

int mpm_read_grid(mpm_handle c, int32_t channel, float* out, int64_t out_len) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_read_grid: null argument");
    const int D = c->D, n = c->sc[0].desc.counts;
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
    mpm::SceneState st{};
    int rc = read_scene(c, 0, st, nullptr);
    if (rc) return rc;
    const int b0 = st.block_begin, nbk = std::max(st.block_end - st.block_begin, 0);
    std::vector<int> slots((size_t)std::max(nbk, 1));
    std::vector<float> vals((size_t)std::max(nbk, 1) * 64 * width);
    if (nbk > 0) {
        HIPCHECK(hipMemcpy(slots.data(), c->block_slot + b0, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
        for (int w = 0; w < width; ++w)
            HIPCHECK(hipMemcpy(&vals[(size_t)w * nbk * 64], src + (size_t)w * c->gcap + (size_t)b0 * 64,
                               (size_t)nbk * 64 * sizeof(float), hipMemcpyDeviceToHost));
    }
    std::fill(out, out + total * width, 0.0f);
    for (int b = 0; b < nbk; ++b)
        for (int l = 0; l < 64; ++l) {
            int64_t o;
            if (!dense_index(c, slots[(size_t)b], l, o)) continue;
            for (int w = 0; w < width; ++w) out[o * width + w] = vals[(size_t)w * nbk * 64 + (size_t)b * 64 + l];
        }
    return MPM_OK;
}

int mpm_read_flags(mpm_handle c, uint32_t* out, int64_t out_len) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_read_flags: null argument");
    const int D = c->D, n = c->sc[0].desc.counts;
    int64_t total = 1;
    for (int a = 0; a < D; ++a) total *= n + 2;
    if (out_len < total) return mpm_set_error(MPM_E_INVALID, "mpm_read_flags: output too small");
    mpm::SceneState st{};
    int rc = read_scene(c, 0, st, nullptr);
    if (rc) return rc;
    const int b0 = st.block_begin, nbk = std::max(st.block_end - st.block_begin, 0);
    std::vector<int> slots((size_t)std::max(nbk, 1));
    std::vector<unsigned> fl((size_t)std::max(nbk, 1) * 64);
    if (nbk > 0) {
        HIPCHECK(hipMemcpy(slots.data(), c->block_slot + b0, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(fl.data(), c->g_flags + (size_t)b0 * 64, (size_t)nbk * 64 * sizeof(unsigned), hipMemcpyDeviceToHost));
    }
    std::fill(out, out + total, 0u);
    for (int b = 0; b < nbk; ++b)
        for (int l = 0; l < 64; ++l) {
            int64_t o;
            if (dense_index(c, slots[(size_t)b], l, o)) out[o] = fl[(size_t)b * 64 + l];
        }
    return MPM_OK;
}

int mpm_read_block_flags(mpm_handle c, int64_t cap_blocks, int32_t* block_origin, uint32_t* flags,
                         int32_t* nblocks) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_read_block_flags: null handle");
    mpm::SceneState st{};
    int rc = read_scene(c, 0, st, nullptr);
    if (rc) return rc;
    const int D = c->D, B = block_edge(D), nb = c->table[0].nb;
    const int b0 = st.block_begin, nbk = std::max(st.block_end - st.block_begin, 0);
    if (nblocks) *nblocks = nbk;
    if (!block_origin && !flags) return MPM_OK;
    if (cap_blocks < nbk) return mpm_set_error(MPM_E_CAPACITY, "mpm_read_block_flags: cap_blocks too small");
    if (nbk == 0) return MPM_OK;
    std::vector<int> slots((size_t)nbk);
    HIPCHECK(hipMemcpy(slots.data(), c->block_slot + b0, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
    if (flags)
        HIPCHECK(hipMemcpy(flags, c->g_flags + (size_t)b0 * 64, (size_t)nbk * 64 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (block_origin)
        for (int b = 0; b < nbk; ++b) {
            int s = slots[(size_t)b] - c->table[0].slot_base;
            int bb[3];
            for (int a = D - 1; a >= 0; --a) { bb[a] = s % nb; s /= nb; }
            for (int a = 0; a < D; ++a) block_origin[(int64_t)b * D + a] = B * bb[a] - B;
        }
    return MPM_OK;
}

int mpm_read_bins(mpm_handle c, int64_t cap_blocks, int32_t* block_coords, int32_t* block_particle_counts,
                  int64_t cap_particles, int32_t* particle_block, int32_t* particle_ids, int32_t* nblocks,
                  int64_t* nparticles) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_read_bins: null handle");
    mpm::SceneState st{};
    int rc = read_scene(c, 0, st, nullptr);
    if (rc) return rc;
    const int D = c->D;
    const int b0 = st.block_begin, nbk = std::max(st.block_end - st.block_begin, 0);
    const int64_t p0 = st.pbegin, np = std::max(st.pend - st.pbegin, 0);
    if (nblocks) *nblocks = nbk;
    if (nparticles) *nparticles = np;
    std::vector<int> slots((size_t)std::max(nbk, 1)), pstart((size_t)nbk + 1);
    if (nbk > 0) {
        HIPCHECK(hipMemcpy(slots.data(), c->block_slot + b0, (size_t)nbk * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(pstart.data(), c->block_pstart + b0, ((size_t)nbk + 1) * sizeof(int), hipMemcpyDeviceToHost));
    } else {
        pstart[0] = (int)p0;
    }
    const int nb = c->table[0].nb;
    if (block_coords || block_particle_counts) {
        if (cap_blocks < nbk) return mpm_set_error(MPM_E_CAPACITY, "mpm_read_bins: cap_blocks too small");
        for (int b = 0; b < nbk; ++b) {
            int s = slots[(size_t)b] - c->table[0].slot_base;
            int bb[3];
            for (int a = D - 1; a >= 0; --a) { bb[a] = s % nb; s /= nb; }
            if (block_coords) for (int a = 0; a < D; ++a) block_coords[(int64_t)b * D + a] = bb[a];
            if (block_particle_counts) block_particle_counts[b] = pstart[(size_t)b + 1] - pstart[(size_t)b];
        }
    }
    if (particle_block || particle_ids) {
        if (cap_particles < np) return mpm_set_error(MPM_E_CAPACITY, "mpm_read_bins: cap_particles too small");
        if (particle_ids && np > 0)
            HIPCHECK(hipMemcpy(particle_ids, c->idbuf[c->cur] + p0, (size_t)np * sizeof(int), hipMemcpyDeviceToHost));
        if (particle_block)
            for (int b = 0; b < nbk; ++b)
                for (int i = pstart[(size_t)b]; i < pstart[(size_t)b + 1]; ++i) particle_block[i - p0] = b;
    }
    return MPM_OK;
}

int mpm_cg_stats(mpm_handle c, mpm_step_stats* stats) {
    if (!c || !stats) return mpm_set_error(MPM_E_INVALID, "mpm_cg_stats: null argument");
    return fill_stats(c, 0, stats);
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
    if (!c->stepped) return mpm_set_error(MPM_E_NOT_READY, "mpm_time_cg_operator: run a substep first");
    if (!c->records_fresh)
        return mpm_set_error(MPM_E_NOT_READY, "mpm_time_cg_operator: the operator records are stale (run a substep)");
    // the operator kernel skips finished solves: reopen them for the timed
    // launches (same dt, same search direction) and restore the state after
    const size_t n = (size_t)c->nscenes;
    std::vector<mpm::CGState> saved(n), open(n);
    mpm::Counters hc{};
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(saved.data(), c->cg, sizeof(mpm::CGState) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&hc, c->ctr, sizeof(hc), hipMemcpyDeviceToHost));
    open = saved;
    for (auto& g : open) g.done = 0;
    mpm::Counters hopen = hc;
    hopen.cg_ndone = 0;
    HIPCHECK(hipMemcpy(c->cg, open.data(), sizeof(mpm::CGState) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->ctr, &hopen, sizeof(hopen), hipMemcpyHostToDevice));
    std::vector<float> dts(n);
    for (size_t s = 0; s < n; ++s) dts[s] = c->sc[s].last_dt;
    set_step(c, dts.data(), 0.0f, false, false);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    mpm::KP p = make_kp(c);
    HIPCHECK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) {
        if (c->D == 3) launch_operator<3>(c, p, false, 1, c->nscenes == 1);
        else launch_operator<2>(c, p, false, 1, c->nscenes == 1);
    }
    HIPCHECK(hipEventRecord(e1, c->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    HIPCHECK(hipMemcpy(c->cg, saved.data(), sizeof(mpm::CGState) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->ctr, &hc, sizeof(hc), hipMemcpyHostToDevice));
    int rc = check_launch("mpm_time_cg_operator");
    if (rc) return rc;
    *avg_ms = ms / (float)reps;
    return MPM_OK;
}

int mpm_time_stage(mpm_handle c, int32_t stage, int32_t reps, float* avg_ms) {
    if (!c || !avg_ms || reps <= 0 ||
        (stage != MPM_STAGE_P2G && stage != MPM_STAGE_CONSTITUTIVE && stage != MPM_STAGE_G2P))
        return mpm_set_error(MPM_E_INVALID, "mpm_time_stage: bad argument");
    if (!c->stepped) return mpm_set_error(MPM_E_NOT_READY, "mpm_time_stage: run a substep first");
    // the particle fields and keys go to the other buffer (unused between
    // substeps) and come back after the timed launches
    const size_t n = (size_t)c->nscenes;
    const size_t pbytes = (size_t)nfields(c->D) * (size_t)c->pcap * sizeof(float);
    const int cur = c->cur;
    HIPCHECK(hipStreamSynchronize(c->stream));
    // G2P's reduction rewrites each scene's max speed, removal count and
    // non-finite flag: they come back with the particles
    std::vector<mpm::SceneState> sst_saved(n);
    HIPCHECK(hipMemcpy(sst_saved.data(), c->sst, sizeof(mpm::SceneState) * n, hipMemcpyDeviceToHost));
    std::vector<float> dts(n);
    for (size_t s = 0; s < n; ++s) dts[s] = c->sc[s].last_dt;
    set_step(c, dts.data(), 0.0f, false, false);
    HIPCHECK(hipMemcpyAsync(c->pbuf[cur ^ 1], c->pbuf[cur], pbytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->keybuf[cur ^ 1], c->keybuf[cur], (size_t)c->pcap * sizeof(int), hipMemcpyDeviceToDevice,
                            c->stream));
    // from here on the saved state is restored on every path
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    mpm::KP p = make_kp(c);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        if (stage == MPM_STAGE_P2G) {
            if (c->D == 3) mpm::launch_p2g<3>(p, c->stream, c->grid_block);
            else mpm::launch_p2g<2>(p, c->stream, c->grid_block);
        } else if (stage == MPM_STAGE_CONSTITUTIVE) {  // the substep's fused constitutive update and force
            // the records are not part of the restored state
            c->records_fresh = false;
            if (c->D == 3) mpm::launch_particle_force<3>(p, c->stream, c->grid_block, true, true);
            else mpm::launch_particle_force<2>(p, c->stream, c->grid_block, true, true);
        } else {
            if (c->D == 3) mpm::launch_g2p<3>(p, c->stream, c->grid_block);
            else mpm::launch_g2p<2>(p, c->stream, c->grid_block);
        }
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    HIPCHECK(hipMemcpyAsync(c->pbuf[cur], c->pbuf[cur ^ 1], pbytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->keybuf[cur], c->keybuf[cur ^ 1], (size_t)c->pcap * sizeof(int), hipMemcpyDeviceToDevice,
                            c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(c->sst, sst_saved.data(), sizeof(mpm::SceneState) * n, hipMemcpyHostToDevice));
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string("mpm_time_stage: ") + hipGetErrorString(e));
    *avg_ms = ms / (float)reps;
    return MPM_OK;
}

int mpm_measure_copy_bandwidth(int64_t bytes, int32_t reps, double* gbs) {
    if (!gbs || reps <= 0 || bytes < 16 || bytes % 16 != 0)
        return mpm_set_error(MPM_E_INVALID, "mpm_measure_copy_bandwidth: bad argument");
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *src = nullptr, *dst = nullptr;
    float ms = 0.0f;
    const int64_t n4 = bytes / 16;
    if ((n4 + mpm::WG - 1) / mpm::WG > INT32_MAX)
        return mpm_set_error(MPM_E_INVALID, "mpm_measure_copy_bandwidth: bytes too large");
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMalloc(&src, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, (size_t)bytes);
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, (size_t)bytes, s);
    if (e == hipSuccess) {
        mpm::launch_copy((const float4*)src, (float4*)dst, n4, s);  // untimed first launch
        e = hipEventRecord(e0, s);
    }
    if (e == hipSuccess) {
        for (int r = 0; r < reps; ++r) mpm::launch_copy((const float4*)src, (float4*)dst, n4, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (s) (void)hipStreamSynchronize(s);
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) return mpm_set_error(MPM_E_HIP, std::string("mpm_measure_copy_bandwidth: ") + hipGetErrorString(e));
    if (!(ms > 0.0f)) return mpm_set_error(MPM_E_HIP, "mpm_measure_copy_bandwidth: no elapsed time");
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return MPM_OK;
}

int mpm_get_launch_config(mpm_handle c, mpm_launch_config* cfg) {
    if (!c || !cfg) return mpm_set_error(MPM_E_INVALID, "mpm_get_launch_config: null argument");
    *cfg = current_config(c);
    return MPM_OK;
}

int mpm_set_launch_config(mpm_handle c, const mpm_launch_config* cfg) {
    if (!c || !cfg) return mpm_set_error(MPM_E_INVALID, "mpm_set_launch_config: null argument");
    int rc = check_config(c, *cfg);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    apply_config(c, *cfg);
    return MPM_OK;
}

int mpm_autotune(mpm_handle c, const float* dt, int32_t flags, mpm_launch_config* out) {
    if (!c || !dt) return mpm_set_error(MPM_E_INVALID, "mpm_autotune: null argument");
    if (!c->initialized) return mpm_set_error(MPM_E_NOT_READY, "mpm_autotune: call mpm_initialize first");
    const TuneKey key = tune_key(c);
    mpm_launch_config best{};
    bool cached = false;
    if (flags & MPM_TUNE_CACHED) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            best = it->second;
            cached = true;
        }
    }
    if (!cached) {
        std::vector<float> dts((size_t)c->nscenes, dt[0]);
        if (c->nscenes > 1) std::copy(dt, dt + c->nscenes, dts.begin());
        int rc = c->D == 3 ? autotune<3>(c, dts.data(), best) : autotune<2>(c, dts.data(), best);
        if (rc) return rc;
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = best;
    }
    int rc = check_config(c, best);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    apply_config(c, best);
    best.from_cache = cached ? 1 : 0;
    if (out) *out = best;
    return MPM_OK;
}

int mpm_set_graph(mpm_handle c, int32_t on) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_set_graph: null handle");
    if (on && max_cg_iterations(c) > MAX_GRAPH_CG_ITERATIONS)
        return mpm_set_error(MPM_E_INVALID, "mpm_set_graph: captured substeps need CG caps of at most 256 iterations");
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (!on) invalidate_graphs(c);
    c->graph = on != 0;
    return MPM_OK;
}

const char* mpm_kernel_class_name(int32_t cls) { return (cls >= 0 && cls < KC_COUNT) ? kClassNames[cls] : ""; }
int32_t mpm_kernel_class_count(void) { return KC_COUNT; }

// ---- batches ----------------------------------------------------------------
int mpm_batch_create(const mpm_scene_desc* descs, const int64_t* capacities, int32_t nscenes, mpm_handle* out) {
    return create_batch(descs, capacities, nscenes, out);
}

int32_t mpm_batch_size(mpm_handle c) { return c ? c->nscenes : 0; }

int mpm_batch_upload(mpm_handle c, int32_t scene, int64_t n, const float* x, const float* v, const float* F,
                     const float* Jp, const float* mass, const float* volume) {
    return mpm_batch_upload_ex(c, scene, n, x, v, F, Jp, mass, volume, nullptr);
}

int mpm_batch_upload_ex(mpm_handle c, int32_t scene, int64_t n, const float* x, const float* v, const float* F,
                        const float* Jp, const float* mass, const float* volume, const int32_t* material) {
    if (!c || (n > 0 && (!x || !v || !mass))) return mpm_set_error(MPM_E_INVALID, "mpm_batch_upload: null argument");
    if (scene < 0 || scene >= c->nscenes) return mpm_set_error(MPM_E_INVALID, "mpm_batch_upload: no such scene");
    if (c->ever_binned) return mpm_set_error(MPM_E_NOT_READY, "mpm_batch_upload: the batch is already initialized");
    int rc = upload_scene(c, scene, n, x, v, F, Jp, mass, volume, material);
    if (rc) return rc;
    return upload_max_speed(c);
}

int mpm_batch_substep(mpm_handle c, const float* dt) {
    if (!c || !dt) return mpm_set_error(MPM_E_INVALID, "mpm_batch_substep: null argument");
    if (!c->initialized) return mpm_set_error(MPM_E_NOT_READY, "mpm_batch_substep: call mpm_initialize first");
    for (int s = 0; s < c->nscenes; ++s)
        if (!std::isfinite(dt[s])) return mpm_set_error(MPM_E_INVALID, "mpm_batch_substep: dt must be finite");
    return do_substep(c, dt);
}

int mpm_batch_max_velocity(mpm_handle c, float* out) {
    if (!c || !out) return mpm_set_error(MPM_E_INVALID, "mpm_batch_max_velocity: null argument");
    return max_velocities(c, out);
}

int mpm_batch_stats(mpm_handle c, int32_t scene, mpm_step_stats* stats) {
    if (!c || !stats) return mpm_set_error(MPM_E_INVALID, "mpm_batch_stats: null argument");
    if (scene < 0 || scene >= c->nscenes) return mpm_set_error(MPM_E_INVALID, "mpm_batch_stats: no such scene");
    return fill_stats(c, scene, stats);
}

int mpm_batch_download(mpm_handle c, int32_t scene, int64_t cap, float* x, float* v, float* F, float* Jp,
                       float* mass, float* volume, int32_t* ids, int64_t* n_out) {
    return mpm_batch_download_ex(c, scene, cap, x, v, F, Jp, mass, volume, ids, nullptr, n_out);
}

int mpm_batch_download_ex(mpm_handle c, int32_t scene, int64_t cap, float* x, float* v, float* F, float* Jp,
                          float* mass, float* volume, int32_t* ids, int32_t* material, int64_t* n_out) {
    if (!c) return mpm_set_error(MPM_E_INVALID, "mpm_batch_download: null handle");
    if (scene < 0 || scene >= c->nscenes) return mpm_set_error(MPM_E_INVALID, "mpm_batch_download: no such scene");
    return download_scene(c, scene, cap, x, v, F, Jp, mass, volume, ids, n_out, material);
}

int mpm_batch_advance_to_time(mpm_handle c, const double* time, const double* target, double* time_out,
                              int32_t* substeps_out) {
    if (!c || !time || !target) return mpm_set_error(MPM_E_INVALID, "mpm_batch_advance_to_time: null argument");
    if (!c->initialized) return mpm_set_error(MPM_E_NOT_READY, "mpm_batch_advance_to_time: call mpm_initialize first");
    const size_t n = (size_t)c->nscenes;
    std::vector<float> t(n), tgt(n), vmax(n), dts(n);
    std::vector<int> steps(n, 0);
    std::vector<char> done(n, 0), last(n, 0);
    for (size_t s = 0; s < n; ++s) {
        t[s] = (float)time[s];
        tgt[s] = (float)target[s];
        done[s] = !(tgt[s] > t[s]);
    }
    if (c->graph && !c->emitted) {  // the device frame loop
        int rc = device_frames(c, t.data(), tgt.data(), t.data(), steps.data());
        if (rc) return rc;
        std::fill(done.begin(), done.end(), 1);
    }
    for (;;) {
        bool any = false;
        int rc = max_velocities(c, vmax.data());
        if (rc) return rc;
        for (size_t s = 0; s < n; ++s) {
            dts[s] = 0.0f;
            if (done[s]) continue;
            bool l = false;
            const float dt = next_dt(c->sc[s].desc, vmax[s], t[s], tgt[s], l);
            if (!(dt > 0.0f)) { done[s] = 1; continue; }
            dts[s] = dt;
            last[s] = l ? 1 : 0;
            any = true;
        }
        if (!any) break;
        rc = do_substep(c, dts.data());
        if (rc) return rc;
        for (size_t s = 0; s < n; ++s)
            if (dts[s] > 0.0f) {
                t[s] += dts[s];
                ++steps[s];
                if (last[s]) done[s] = 1;
            }
    }
    for (size_t s = 0; s < n; ++s) {
        if (time_out) time_out[s] = (double)t[s];
        if (substeps_out) substeps_out[s] = steps[s];
    }
    return MPM_OK;
}

}  // extern "C"

// ---- allocation-free substep ------------------------------------------------
namespace {

// Scratch of mpm_substep_noalloc: every device buffer of a one-scene context
// (sorted particle copies, block tables, grid channels, tiles, CG partials,
// counters), carved afresh by each call from this array.  Each device that
// runs the entry point holds its own copy until the process ends.  In host
// code the array's name is a host-side stand-in and never a kernel argument:
// its device address comes from hipGetSymbolAddress, once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)40 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static_assert(MPM_NOALLOC_MAX_BLOCKS <= 2048 && MPM_NOALLOC_MAX_COUNTS <= 64 && MPM_NOALLOC_MAX_PARTICLES <= 65536,
              "the scratch size is computed for these limits");

struct NoallocDevice {
    unsigned char* base = nullptr;  // device address of `scratch` on this device
    mpm_context* c = nullptr;       // the host view every call on this device reuses
};
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mutex;  // held for the host part of a call
NoallocDevice g_noalloc[NOALLOC_MAX_DEVICES];

// Set-up on the first call on a device (with that device current): the
// scratch's device address and the reusable host view.
int noalloc_setup(NoallocDevice& nd) {
    if (nd.base) return MPM_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        return mpm_set_error(MPM_E_HIP, std::string("mpm_substep_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    mpm_context* c = new mpm_context;
    c->nscenes = 1;
    c->sc.resize(1);
    c->table.resize(1);
    nd.c = c;
    nd.base = static_cast<unsigned char*>(ptr);
    return MPM_OK;
}

template <int D>
int noalloc_substep(NoallocDevice& nd, const mpm_scene_desc& d, int64_t n, float* x, float* v, float* F, float* Jp,
                    const float* mass, float* volume, int32_t* valid, float dt, hipStream_t stream,
                    mpm_step_stats* stats) {
    mpm_context* c = nd.c;
    c->D = D;
    c->nscenes = 1;
    c->pcap = n;
    c->nslots = slots_of_desc(d);
    c->bcap = std::min(c->nslots, MPM_NOALLOC_MAX_BLOCKS);
    size_context(c);
    c->stream = stream;
    c->sc[0] = SceneHost{};
    c->sc[0].desc = d;
    c->sc[0].capacity = n;
    c->sc[0].n_uploaded = n;
    c->sc[0].uploaded = true;
    c->cur = 0;
    c->solo = -1;
    c->pool = false;
    c->graph = false;
    c->timing = false;
    c->emitted = true;  // every call runs Process_Waiting_Particles for particles of volume 0
    c->ever_binned = true;
    size_t off = 0;
    bool fits = true;
    layout_buffers(c, [&](auto** ptr, size_t count) {
        using T = std::remove_reference_t<decltype(**ptr)>;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        off = (off + 255) & ~(size_t)255;
        if (off + bytes > NOALLOC_SCRATCH_BYTES) {
            fits = false;
            *ptr = nullptr;
            return;
        }
        *ptr = reinterpret_cast<T*>(nd.base + off);
        off += bytes;
    });
    if (!fits) return mpm_set_error(MPM_E_CAPACITY, "mpm_substep_noalloc: the scene does not fit the scratch");
    mpm::SceneParams sp = scene_params(d, 0, c->nslots);
    sp.waiting = 1;
    mpm::NoallocIn in{x, v, F, Jp, mass, volume, valid, c->d_dt, c->slot_scene, (int)n, dt};
    mpm::launch_noalloc_load<D>(make_kp(c), stream, c->grid_particle, sp, in);
    int rc = substep<D>(c, nullptr, true);  // dt from d_dt, CG of a fixed count (no host polls)
    if (rc) return rc;
    const mpm::NoallocOut o{x, v, F, Jp, volume, valid};
    mpm::launch_noalloc_store<D>(make_kp(c), stream, c->grid_particle, o);
    rc = check_launch("mpm_substep_noalloc");
    if (rc) return rc;
    if (stats) return fill_stats(c, 0, stats);
    return MPM_OK;
}

}  // namespace

extern "C" int mpm_substep_noalloc(const mpm_scene_desc* desc, int64_t n, float* x, float* v, float* F, float* Jp,
                                   const float* mass, float* volume, int32_t* valid, float dt, void* stream,
                                   mpm_step_stats* stats) {
    if (!desc || !x || !v || !mass || !volume || !valid)
        return mpm_set_error(MPM_E_INVALID, "mpm_substep_noalloc: null argument");
    int rc = check_desc(*desc, "mpm_substep_noalloc");
    if (rc) return rc;
    if (n < 1 || n > MPM_NOALLOC_MAX_PARTICLES)
        return mpm_set_error(MPM_E_CAPACITY, "mpm_substep_noalloc: n must be in [1, " +
                                                 std::to_string(MPM_NOALLOC_MAX_PARTICLES) + "]");
    if (desc->counts > MPM_NOALLOC_MAX_COUNTS)
        return mpm_set_error(MPM_E_CAPACITY, "mpm_substep_noalloc: counts above " + std::to_string(MPM_NOALLOC_MAX_COUNTS));
    if (desc->cg_max_iterations < 1 || desc->cg_max_iterations > MAX_GRAPH_CG_ITERATIONS)
        return mpm_set_error(MPM_E_INVALID, "mpm_substep_noalloc: the CG cap must be in [1, 256]");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return mpm_set_error(MPM_E_INVALID, "mpm_substep_noalloc: dt must be positive");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return mpm_set_error(MPM_E_INVALID, "mpm_substep_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    NoallocDevice& nd = g_noalloc[dev];
    rc = noalloc_setup(nd);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    return desc->dim == 3 ? noalloc_substep<3>(nd, *desc, n, x, v, F, Jp, mass, volume, valid, dt, st, stats)
                          : noalloc_substep<2>(nd, *desc, n, x, v, F, Jp, mass, volume, valid, dt, st, stats);
}
