// Host-side context of libmpm_gpu (a batch of scenes; a single scene is a
// batch of one, a serving arena a batch whose scene entries come and go) and
// the internal entry points the request server uses.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/mpm_gpu.h"
#include "mpm_kernels.h"

enum KClass {
    KC_BIN = 0, KC_SCAN, KC_SETUP, KC_REORDER, KC_P2G, KC_NORMALIZE, KC_INIT_VOLUME, KC_PFORCE, KC_GRID,
    KC_CG_FORCE, KC_CG_NODE, KC_G2P, KC_MAXSPEED, KC_COUNT
};
// Host view of one scene of a batch.
struct SceneHost {
    mpm_scene_desc desc{};
    int64_t capacity = 0;  // particle entries reserved for the scene
    int64_t pbase = 0;     // its range in the particle buffer before the first binning
    int64_t n_uploaded = 0;
    bool uploaded = false;
    float last_dt = 0.0f;
};

struct mpm_context {
    int D = 3;
    int nscenes = 0;
    std::vector<SceneHost> sc;
    std::vector<mpm::SceneParams> table;  // host copy of the scene table
    int64_t pcap = 0;
    int bcap = 0, nslots = 0, nparts = 0;
    int64_t gcap = 0;
    hipStream_t stream = nullptr;
    float* pbuf[2] = {nullptr, nullptr};
    int* idbuf[2] = {nullptr, nullptr};
    int* keybuf[2] = {nullptr, nullptr};
    int* scbuf[2] = {nullptr, nullptr};
    int cur = 0;
    int solo = -1;  // batch or arena substep with one active scene: it takes the single-scene CG
    int* prank = nullptr;
    int* cell_order = nullptr;  // per particle: index scattered by binning rank (stable in-cell order)
    float* rcache = nullptr;
    float* pop = nullptr;
    int *slot_count = nullptr, *slot_touched = nullptr, *slot_active = nullptr;
    int *block_slot = nullptr, *block_scene = nullptr, *block_pstart = nullptr, *block_nbr = nullptr;
    int* block_tnbr = nullptr;
    int* block_org = nullptr;
    int* block_order = nullptr;  // per XCD range: the active blocks by descending particle count, then the
                                 // same by descending fullest cell (2 bcap ints)
    int* block_full = nullptr;   // per active block, its fullest cell's particle count
    int *scan_part = nullptr, *cell_count = nullptr, *cell_start = nullptr;
    float *g_m = nullptr, *g_v = nullptr, *g_vs = nullptr, *g_q = nullptr, *g_s = nullptr, *g_r = nullptr;
    unsigned* g_flags = nullptr;
    float* tiles = nullptr;
    mpm::Counters* ctr = nullptr;
    mpm::SceneParams* scn = nullptr;
    mpm::SceneState* sst = nullptr;
    mpm::CGState* cg = nullptr;
    int* slot_scene = nullptr;   // per slot: owning scene, -1 if free
    double *bsum = nullptr, *bmax = nullptr, *bsq = nullptr;
    unsigned* bu = nullptr;
    int cg_predict = 0;    // the next solve's first launch batch: the last solve's iterations plus one
    int* h_done = nullptr;  // pinned: per poll slot (8 ints), Counters::overflow .. cg_iters
    int* d_done = nullptr;  // the device's address of h_done
    bool overflow_seen = false;  // the last substep's poll saw a block-capacity overflow
    hipEvent_t done_ev[2] = {nullptr, nullptr};
    // done polls that a solve running to its last chunk left unread; they are
    // read once the rest of the substep is queued (finish_cg_poll)
    bool poll_open = false;
    bool poll_pending[2] = {false, false};
    int poll_boundary[2] = {0, 0};
    int poll_older = 0, poll_target = 1, poll_max_it = 0;
    int grid_particle = 0, grid_block = 0, grid_node = 0;
    int op_mode = 0, grid_op = 0;  // CG operator: 0 four waves per block, 1 one wave per block
    size_t nred = 0;           // entries of the reduction partial arrays
    // captured substeps: one graph per particle-buffer parity
    bool graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    float* d_dt = nullptr;     // per-scene dt read by a replayed graph
    float* h_dt = nullptr;     // pinned staging of d_dt
    hipEvent_t dt_ev = nullptr;
    bool dt_pending = false;
    // device frame loop: per-scene FrameState, graphs of FRAME_K substeps per
    // buffer parity, pinned read-back of the frame-done count
    mpm::FrameState* frame = nullptr;
    hipGraphExec_t fexec[2] = {nullptr, nullptr};
    int* h_fdone = nullptr;
    hipEvent_t fdone_ev = nullptr;
    // serving arena: scenes attach to free entries, slot ranges and particle space
    bool pool = false;
    int pool_cg_max = 0;                            // CG cap of every arena scene
    std::vector<std::pair<int, int>> free_slots;    // (first slot, count), sorted by first slot
    std::vector<char> entry_used;
    bool initialized = false;
    bool binned = false;       // particles sit in the block ranges of the last binning
    bool ever_binned = false;  // particle ranges are the binned ones (not the upload ranges)
    bool stepped = false;      // a substep has run (operator timing needs its state)
    // the operator records (pop) hold the constitutive update of the particles
    // as the last binning placed them: set by the constitutive kernel on binned
    // particles, cleared by binning (which reorders the particles but not the
    // records) and by stage timing of the constitutive kernel
    bool records_fresh = false;
    bool emitted = false;      // particles added since the last substep wait for their volumes
    std::vector<void*> allocations;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_class;
    size_t ev_used = 0;
    double ms[KC_COUNT] = {};
    int64_t launches[KC_COUNT] = {};
};

namespace mpm_host {

// Bytes of device memory per particle, block slot and block of storage of a
// context (3D or 2D): the admission estimate of the request server.
struct UnitBytes {
    int64_t particle, slot, block;
};
UnitBytes unit_bytes(int D);

// An arena of max_scenes free scene entries with room for pcap particles,
// nslots block slots and bcap blocks of storage; every attached scene must
// cap its CG at cg_max iterations.
int pool_create(int D, int max_scenes, int64_t pcap, int64_t nslots, int64_t bcap, int cg_max, mpm_context** out);
// Attaches a scene (its particles waiting for Process_Waiting_Particles) to
// a free entry; MPM_E_CAPACITY if the entry, slot range or particle space is
// not available.
int pool_attach(mpm_context* c, const mpm_scene_desc& d, int64_t n, const float* x, const float* v,
                const float* mass, const int32_t* material, int* entry);
// Finite particle inputs with a positive mass (F, Jp, volume may be NULL);
// MPM_E_INVALID otherwise.
int check_particles(const char* who, int D, int64_t n, const float* x, const float* v, const float* F,
                    const float* Jp, const float* mass, const float* volume);
// Removes a scene's particles and frees its entry and slot range.
int pool_detach(mpm_context* c, int entry);
// Largest free slot range.
int pool_largest_free_slots(const mpm_context* c);
int slots_of(const mpm_scene_desc& d);

int substep(mpm_context* c, const float* dts);
int max_velocities(mpm_context* c, float* out);
int download(mpm_context* c, int scene, int64_t cap, float* x, float* v, float* F, float* Jp, float* mass,
             float* volume, int32_t* ids, int64_t* n_out);
int stats(mpm_context* c, int scene, mpm_step_stats* out);
int live_particles(mpm_context* c, int64_t* n);
// The reference's dt rule (MPM_Driver.cpp:55-80) for one scene, in float.
float next_dt(const mpm_scene_desc& d, float max_v, float time, float target, bool& last);

}  // namespace mpm_host
