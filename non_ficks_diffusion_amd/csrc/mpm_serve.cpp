// Request server: continuous batching of snow scenes on one device arena.
//
// The arena is one batch context whose scene entries come and go (see
// mpm_context.h): an admitted request attaches its scene to a free entry, a
// contiguous range of block slots and particle space, and its particle,
// grid and CG state stay on the device until it finishes.  One tick admits
// queued requests in arrival order while the arena has room, then runs
// batched substeps until every scene resident at its start has completed a
// frame, each scene with its own dt from the reference's rule (MPM_Driver::Advance_To_Target_Time,
// MPM_Driver.cpp:55-80) toward its own next frame end.  A scene whose frames
// are short goes on with its next frame instead of waiting for the slowest
// one, so every resident particle advances in every substep.  Between
// substeps, finished requests leave the arena and queued ones take the room
// they free.  No reduction mixes scenes, so a scene evolves in the mix as it
// does alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/mpm_gpu.h"
#include "mpm_context.h"
#include "mpm_internal.h"

namespace {

struct Request {
    mpm_scene_desc desc{};
    int64_t n = 0;
    std::vector<float> x, v, m;  // host copy until admitted
    std::vector<int32_t> mat;    // material entries (empty: all 0)
    int frames_total = 0, frames_done = 0;
    int state = MPM_REQ_QUEUED;
    int entry = -1;
    int slots = 0;
    int64_t blocks = 0, bytes = 0;  // admission estimate
    float time = 0.0f;              // simulated time
    int substeps = 0;
    double submitted = -1.0, admitted = -1.0, finished = -1.0;
    std::vector<double> frame_t;    // completion time of each frame (server clock)
    // final particles (keep_results)
    std::vector<float> rx, rv, rF, rJp, rm, rvol;
    std::vector<int32_t> rids;
    int64_t rn = 0;
};

}  // namespace

struct mpm_server {
    mpm_serve_config cfg{};
    mpm_context* c = nullptr;
    mpm_host::UnitBytes ub{};
    int64_t pcap = 0, nslots = 0, bcap = 0, arena_bytes = 0;
    int64_t particles_in_use = 0, blocks_in_use = 0, bytes_in_use = 0;
    int64_t tuned_at = 0;  // resident particles at the last launch-shape tune
    std::map<int64_t, Request> req;
    std::deque<int64_t> queue;
    std::vector<int64_t> by_entry;  // arena entry -> request id, -1 if free
    int64_t next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Blocks of storage a scene of n particles may need: home blocks plus
// their halo at low particle density, at most every slot of its grid.
int64_t block_estimate(int64_t n, int slots) { return std::min<int64_t>(slots, n / 8 + 64); }

bool fits(const mpm_server* s, const Request& r) {
    return s->particles_in_use + r.n <= s->pcap && s->blocks_in_use + r.blocks <= s->bcap &&
           s->bytes_in_use + r.bytes <= s->cfg.hbm_budget_bytes && mpm_host::pool_largest_free_slots(s->c) >= r.slots &&
           std::find(s->by_entry.begin(), s->by_entry.end(), (int64_t)-1) != s->by_entry.end();
}

int finish(mpm_server* s, int64_t id, Request& r) {
    if (s->cfg.keep_results) {
        const int D = s->cfg.dim;
        const int64_t cap = r.n;
        r.rx.resize((size_t)cap * D);
        r.rv.resize((size_t)cap * D);
        r.rF.resize((size_t)cap * D * D);
        r.rJp.resize((size_t)cap);
        r.rm.resize((size_t)cap);
        r.rvol.resize((size_t)cap);
        r.rids.resize((size_t)cap);
        int rc = mpm_host::download(s->c, r.entry, cap, r.rx.data(), r.rv.data(), r.rF.data(), r.rJp.data(),
                                    r.rm.data(), r.rvol.data(), r.rids.data(), &r.rn);
        if (rc) return rc;
    }
    int rc = mpm_host::pool_detach(s->c, r.entry);
    if (rc) return rc;
    s->by_entry[(size_t)r.entry] = -1;
    s->particles_in_use -= r.n;
    s->blocks_in_use -= r.blocks;
    s->bytes_in_use -= r.bytes;
    r.entry = -1;
    r.state = MPM_REQ_DONE;
    r.finished = s->now();
    (void)id;
    return MPM_OK;
}

// Admits queued requests in arrival order while the arena has room.
int admit(mpm_server* s, mpm_tick_stats& st) {
    while (!s->queue.empty()) {
        Request& r = s->req[s->queue.front()];
        if (!fits(s, r)) break;
        int e = -1;
        int rc = mpm_host::pool_attach(s->c, r.desc, r.n, r.x.data(), r.v.data(), r.m.data(),
                                       r.mat.empty() ? nullptr : r.mat.data(), &e);
        if (rc == MPM_E_CAPACITY) break;  // fragmented slots or dead particles not yet compacted
        if (rc) return rc;
        r.entry = e;
        r.state = MPM_REQ_RUNNING;
        r.admitted = s->now();
        std::vector<float>().swap(r.x);
        std::vector<float>().swap(r.v);
        std::vector<float>().swap(r.m);
        std::vector<int32_t>().swap(r.mat);
        s->by_entry[(size_t)e] = s->queue.front();
        s->particles_in_use += r.n;
        s->blocks_in_use += r.blocks;
        s->bytes_in_use += r.bytes;
        s->queue.pop_front();
        ++st.admitted;
    }
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_serve_create(const mpm_serve_config* cfg, mpm_server_handle* out) {
    if (!cfg || !out) return mpm_set_error(MPM_E_INVALID, "mpm_serve_create: null argument");
    const mpm_serve_config& g = *cfg;
    if (g.dim != 2 && g.dim != 3) return mpm_set_error(MPM_E_INVALID, "mpm_serve_create: dim must be 2 or 3");
    if (g.max_scenes < 1 || g.max_counts < 4 || g.max_counts > 4096 || g.hbm_budget_bytes <= 0 ||
        g.cg_max_iterations < 1 || g.cg_max_iterations > 256)
        return mpm_set_error(MPM_E_INVALID, "mpm_serve_create: bad configuration");
    mpm_server* s = new mpm_server;
    s->cfg = g;
    s->ub = mpm_host::unit_bytes(g.dim);
    // Arena: the slots of max_scenes largest grids (block keys are 32-bit),
    // the rest of the budget split between particles and blocks of storage
    // as a scene at 8 particles per cell uses them (n/8 blocks per n).
    mpm_scene_desc big{};
    big.dim = g.dim;
    big.counts = g.max_counts;
    const int64_t key_slots = (((int64_t)1 << 31) - 1) / 64;
    s->nslots = std::min<int64_t>(key_slots, (int64_t)g.max_scenes * mpm_host::slots_of(big));
    s->nslots = std::min<int64_t>(s->nslots, std::max<int64_t>(1, g.hbm_budget_bytes / 4 / s->ub.slot));
    const int64_t rest = g.hbm_budget_bytes - s->nslots * s->ub.slot;
    const double per_particle = (double)s->ub.particle + (double)s->ub.block / 8.0;
    const int64_t np = (int64_t)((double)rest / per_particle);
    s->pcap = std::max<int64_t>(1, std::min<int64_t>(np, ((int64_t)1 << 31) - 1));
    // block indices * 65 stay below 2^31
    s->bcap = std::max<int64_t>(1, std::min<int64_t>({np / 8, s->nslots, (((int64_t)1 << 31) - 1) / 65}));
    s->arena_bytes = s->pcap * s->ub.particle + s->nslots * s->ub.slot + s->bcap * s->ub.block;
    if (rest <= 0 || np < 1)
        return delete s, mpm_set_error(MPM_E_INVALID, "mpm_serve_create: budget too small for the slot table");
    int rc = mpm_host::pool_create(g.dim, g.max_scenes, s->pcap, s->nslots, s->bcap, g.cg_max_iterations, &s->c);
    if (rc) {
        delete s;
        return rc;
    }
    if (g.graph) {
        rc = mpm_set_graph(s->c, 1);
        if (rc) {
            mpm_destroy(s->c);
            delete s;
            return rc;
        }
    }
    s->by_entry.assign((size_t)g.max_scenes, -1);
    *out = s;
    return MPM_OK;
}

int mpm_serve_destroy(mpm_server_handle s) {
    if (!s) return MPM_OK;
    mpm_destroy(s->c);
    delete s;
    return MPM_OK;
}

int mpm_serve_arena(mpm_server_handle s, int64_t* particles, int64_t* slots, int64_t* blocks, int64_t* bytes) {
    if (!s) return mpm_set_error(MPM_E_INVALID, "mpm_serve_arena: null handle");
    if (particles) *particles = s->pcap;
    if (slots) *slots = s->nslots;
    if (blocks) *blocks = s->bcap;
    if (bytes) *bytes = s->arena_bytes;
    return MPM_OK;
}

int mpm_serve_submit(mpm_server_handle s, const mpm_scene_desc* desc, int64_t n, const float* x, const float* v,
                     const float* mass, int32_t frames, int64_t* id) {
    return mpm_serve_submit_ex(s, desc, n, x, v, mass, nullptr, frames, id);
}

int mpm_serve_submit_ex(mpm_server_handle s, const mpm_scene_desc* desc, int64_t n, const float* x, const float* v,
                        const float* mass, const int32_t* material, int32_t frames, int64_t* id) {
    if (!s || !desc || !x || !v || !mass || !id) return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: null argument");
    if (desc->dim != s->cfg.dim) return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: dimension differs from the server's");
    if (desc->counts < 4 || desc->counts > s->cfg.max_counts)
        return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: grid larger than the server's max_counts");
    if (desc->cg_max_iterations < 1 || desc->cg_max_iterations > s->cfg.cg_max_iterations)
        return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: CG cap above the server's");
    if (n < 1 || frames < 1) return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: no particles or frames");
    {
        const int rc = mpm_host::check_particles("mpm_serve_submit", desc->dim, n, x, v, nullptr, nullptr, mass, nullptr);
        if (rc) return rc;
    }
    if (!(desc->frame_rate > 0.0f)) return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: frame rate must be positive");
    Request r;
    r.desc = *desc;
    r.n = n;
    r.frames_total = frames;
    r.slots = mpm_host::slots_of(*desc);
    r.blocks = block_estimate(n, r.slots);
    r.bytes = n * s->ub.particle + (int64_t)r.slots * s->ub.slot + r.blocks * s->ub.block;
    if (r.n > s->pcap || r.blocks > s->bcap || r.slots > s->nslots || r.bytes > s->cfg.hbm_budget_bytes)
        return mpm_set_error(MPM_E_CAPACITY, "mpm_serve_submit: the request exceeds the whole arena");
    const int D = desc->dim;
    r.x.assign(x, x + n * D);
    r.v.assign(v, v + n * D);
    r.m.assign(mass, mass + n);
    if (material) {
        const int nmat = std::max(1, (int)desc->nmaterials);
        for (int64_t i = 0; i < n; ++i)
            if (material[i] < 0 || material[i] >= nmat)
                return mpm_set_error(MPM_E_INVALID, "mpm_serve_submit: particle material outside the scene's table");
        r.mat.assign(material, material + n);
    }
    r.submitted = s->now();
    const int64_t k = s->next_id++;
    s->req.emplace(k, std::move(r));
    s->queue.push_back(k);
    *id = k;
    return MPM_OK;
}

int mpm_serve_tick(mpm_server_handle s, mpm_tick_stats* out) {
    if (!s) return mpm_set_error(MPM_E_INVALID, "mpm_serve_tick: null handle");
    const double t_start = s->now();
    mpm_tick_stats st{};
    int rc = admit(s, st);
    if (rc) return rc;
    const int ne = s->cfg.max_scenes;
    // every scene resident at the start owes one frame to this tick
    std::vector<char> owe((size_t)ne, 0), last((size_t)ne, 0), completed((size_t)ne, 0);
    for (int e = 0; e < ne; ++e) {
        if (s->by_entry[(size_t)e] < 0) continue;
        owe[(size_t)e] = 1;
        ++st.resident;
    }
    std::vector<float> vmax((size_t)ne), dts((size_t)ne);
    bool any_completed = false;
    for (;;) {
        rc = mpm_host::max_velocities(s->c, vmax.data());  // waits for the last substep
        if (rc) return rc;
        if (any_completed) {
            // frames ended by the last substep: time stamps, retirement, admission
            const double t = s->now();
            bool freed = false;
            for (int e = 0; e < ne; ++e) {
                if (!completed[(size_t)e]) continue;
                completed[(size_t)e] = 0;
                const int64_t id = s->by_entry[(size_t)e];
                Request& r = s->req[id];
                r.frame_t.push_back(t);
                if (r.frames_done >= r.frames_total) {
                    rc = finish(s, id, r);
                    if (rc) return rc;
                    ++st.finished;
                    owe[(size_t)e] = 0;
                    freed = true;
                }
            }
            any_completed = false;
            if (freed && !s->queue.empty()) {
                const int before = st.admitted;
                rc = admit(s, st);
                if (rc) return rc;
                if (st.admitted > before) {  // their max |V| for the first dt
                    rc = mpm_host::max_velocities(s->c, vmax.data());
                    if (rc) return rc;
                }
            }
        }
        bool owing = false;
        for (int e = 0; e < ne; ++e) owing = owing || owe[(size_t)e];
        if (!owing || (s->cfg.max_tick_substeps > 0 && st.substeps >= s->cfg.max_tick_substeps)) break;
        bool any = false;
        int64_t work = 0;
        for (int e = 0; e < ne; ++e) {
            dts[(size_t)e] = 0.0f;
            const int64_t id = s->by_entry[(size_t)e];
            if (id < 0) continue;
            Request& r = s->req[id];
            const float target = (float)(r.frames_done + 1) / r.desc.frame_rate;
            bool l = false;
            const float dt = mpm_host::next_dt(r.desc, vmax[(size_t)e], r.time, target, l);
            if (!(dt > 0.0f)) {  // already at the frame end
                ++r.frames_done;
                ++st.frames;
                owe[(size_t)e] = 0;
                completed[(size_t)e] = 1;
                any_completed = true;
                continue;
            }
            dts[(size_t)e] = dt;
            last[(size_t)e] = l ? 1 : 0;
            work += r.n;
            any = true;
        }
        if (!any) continue;
        if (s->cfg.tune && s->particles_in_use >= std::max<int64_t>(1000000, 2 * s->tuned_at)) {
            rc = mpm_autotune(s->c, dts.data(), 0, nullptr);  // times shapes; advances nothing
            if (rc) return rc;
            s->tuned_at = s->particles_in_use;
        }
        rc = mpm_host::substep(s->c, dts.data());
        if (rc) return rc;
        ++st.substeps;
        st.particle_substeps += work;
        for (int e = 0; e < ne; ++e) {
            if (!(dts[(size_t)e] > 0.0f)) continue;
            Request& r = s->req[s->by_entry[(size_t)e]];
            r.time += dts[(size_t)e];
            ++r.substeps;
            if (last[(size_t)e]) {
                ++r.frames_done;
                ++st.frames;
                owe[(size_t)e] = 0;
                completed[(size_t)e] = 1;
                any_completed = true;
            }
        }
    }
    int64_t n_live = 0;
    rc = mpm_host::live_particles(s->c, &n_live);  // reports a block-capacity overflow
    if (rc) return rc;
    st.queued = (int32_t)s->queue.size();
    st.bytes_in_use = s->bytes_in_use;
    st.ms = (s->now() - t_start) * 1e3;
    if (out) *out = st;
    return MPM_OK;
}

int mpm_serve_frame_times(mpm_server_handle s, int64_t id, int32_t cap, double* times, int32_t* n) {
    if (!s) return mpm_set_error(MPM_E_INVALID, "mpm_serve_frame_times: null handle");
    auto it = s->req.find(id);
    if (it == s->req.end()) return mpm_set_error(MPM_E_INVALID, "mpm_serve_frame_times: unknown request");
    const std::vector<double>& t = it->second.frame_t;
    if (n) *n = (int32_t)t.size();
    if (times) std::copy(t.begin(), t.begin() + std::min<size_t>(t.size(), (size_t)std::max(cap, 0)), times);
    return MPM_OK;
}

int mpm_serve_poll(mpm_server_handle s, int64_t id, mpm_request_status* out) {
    if (!s || !out) return mpm_set_error(MPM_E_INVALID, "mpm_serve_poll: null argument");
    auto it = s->req.find(id);
    if (it == s->req.end()) return mpm_set_error(MPM_E_INVALID, "mpm_serve_poll: unknown request");
    const Request& r = it->second;
    out->state = r.state;
    out->frames_done = r.frames_done;
    out->frames_total = r.frames_total;
    out->substeps = r.substeps;
    out->particles = r.n;
    out->bytes = r.bytes;
    out->sim_time = r.time;
    out->submitted = r.submitted;
    out->admitted = r.admitted;
    out->finished = r.finished;
    return MPM_OK;
}

int mpm_serve_download(mpm_server_handle s, int64_t id, int64_t cap, float* x, float* v, float* F, float* Jp,
                       float* mass, float* volume, int32_t* ids, int64_t* n_out) {
    if (!s) return mpm_set_error(MPM_E_INVALID, "mpm_serve_download: null handle");
    auto it = s->req.find(id);
    if (it == s->req.end()) return mpm_set_error(MPM_E_INVALID, "mpm_serve_download: unknown request");
    Request& r = it->second;
    if (r.state == MPM_REQ_RUNNING)
        return mpm_host::download(s->c, r.entry, cap, x, v, F, Jp, mass, volume, ids, n_out);
    if (r.state != MPM_REQ_DONE || !s->cfg.keep_results)
        return mpm_set_error(MPM_E_NOT_READY, "mpm_serve_download: no particles kept for this request");
    if (n_out) *n_out = r.rn;
    if (r.rn > cap) return mpm_set_error(MPM_E_CAPACITY, "mpm_serve_download: cap too small");
    const int D = s->cfg.dim;
    const size_t n = (size_t)r.rn;
    if (x) std::memcpy(x, r.rx.data(), n * D * sizeof(float));
    if (v) std::memcpy(v, r.rv.data(), n * D * sizeof(float));
    if (F) std::memcpy(F, r.rF.data(), n * D * D * sizeof(float));
    if (Jp) std::memcpy(Jp, r.rJp.data(), n * sizeof(float));
    if (mass) std::memcpy(mass, r.rm.data(), n * sizeof(float));
    if (volume) std::memcpy(volume, r.rvol.data(), n * sizeof(float));
    if (ids) std::memcpy(ids, r.rids.data(), n * sizeof(int32_t));
    return MPM_OK;
}

int mpm_serve_release(mpm_server_handle s, int64_t id) {
    if (!s) return mpm_set_error(MPM_E_INVALID, "mpm_serve_release: null handle");
    auto it = s->req.find(id);
    if (it == s->req.end()) return mpm_set_error(MPM_E_INVALID, "mpm_serve_release: unknown request");
    if (it->second.state != MPM_REQ_DONE && it->second.state != MPM_REQ_CANCELLED)
        return mpm_set_error(MPM_E_NOT_READY, "mpm_serve_release: request not finished");
    s->req.erase(it);
    return MPM_OK;
}

int mpm_serve_cancel(mpm_server_handle s, int64_t id) {
    if (!s) return mpm_set_error(MPM_E_INVALID, "mpm_serve_cancel: null handle");
    auto it = s->req.find(id);
    if (it == s->req.end()) return mpm_set_error(MPM_E_INVALID, "mpm_serve_cancel: unknown request");
    Request& r = it->second;
    if (r.state == MPM_REQ_DONE || r.state == MPM_REQ_CANCELLED)
        return mpm_set_error(MPM_E_INVALID, "mpm_serve_cancel: request already finished");
    if (r.state == MPM_REQ_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
        std::vector<float>().swap(r.x);
        std::vector<float>().swap(r.v);
        std::vector<float>().swap(r.m);
        std::vector<int32_t>().swap(r.mat);
    } else {
        // leaves the arena at once: its particles, entry and slot range are freed
        const int rc = mpm_host::pool_detach(s->c, r.entry);
        if (rc) return rc;
        s->by_entry[(size_t)r.entry] = -1;
        s->particles_in_use -= r.n;
        s->blocks_in_use -= r.blocks;
        s->bytes_in_use -= r.bytes;
        r.entry = -1;
    }
    r.state = MPM_REQ_CANCELLED;
    r.finished = s->now();
    return MPM_OK;
}

}  // extern "C"
