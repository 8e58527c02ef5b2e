// Kernel parameter block and launch entry points of the gfx950 substep.
//
// One launch advances a *batch* of independent scenes (a single scene is a
// batch of one).  Scenes are concatenated: their block slots occupy
// consecutive ranges of one slot table, their particles share one buffer
// (grouped by scene after binning, because slots are scene-major), and
// every block or particle looks its scene's parameters up in a scene table.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "mpm_device.h"

namespace mpm {

constexpr int WG = 256;          // work-group size of every kernel
// floats per particle of the CG force operator record (see OpRec)
constexpr int op_record_floats(int D) { return D == 3 ? 16 : 12; }
constexpr int SCAN_PER_THREAD = 4;  // one int4 per array and thread
constexpr int SCAN_TILE = WG * SCAN_PER_THREAD;
constexpr int MAX_PLANES = 8;
constexpr unsigned MPM_FLAG_INTERIOR_DEV = 1u;  // Cell_Type_Interior (Flag_Setup_Helper.h:32)

// Read-only parameters of one scene during a substep.
struct SceneParams {
    int counts, nb;            // cells per axis; block slots per axis (padded)
    int slot_base, nslots;     // the scene's slots: [slot_base, slot_base + nslots)
    float dx, inv_dx;
    Material mats[MAX_MATERIALS];  // the scene's material table (particles hold an entry)
    int nmat;
    float gravity[3];
    float flip;
    int nplanes;
    float pn[MAX_PLANES][3], ps[MAX_PLANES][3], pmu[MAX_PLANES];
    int crop_ghost_x, project_mode;
    double cg_tol;
    int cg_min, cg_max;
    float dt;                  // dt of this substep
    int active;                // 0: the scene keeps its state this substep
    int waiting;               // particles still need Process_Waiting_Particles
    int pad;
};

// Per-scene results written by the kernels.
struct SceneState {
    int block_begin, block_end;  // active blocks of the scene after binning
    int pbegin, pend;            // its particles after binning
    unsigned max_v2;             // bits of max |V|^2 (non-negative float) after G2P
    int n_removed;               // particles removed by the last G2P
    int nonfinite;               // NaN/Inf particle met in binning
    int pad;
};

// Batch counters.
struct Counters {
    int nblocks;        // active blocks after binning (all scenes)
    int n_live;         // particles in the buffer after the last binning / upload
    int n_prev;         // particles in the buffer before the last binning
    int overflow;       // block capacity exceeded
    int cg_ndone;       // scenes whose CG solve is finished
    int slot_hw;        // slots [0, slot_hw) hold every scene's range (the binning scans them)
    int frame_ndone;    // device frame loop: scenes that reached their frame end
    int cg_iters;       // largest iteration count of the solves finished in this substep
};

// Device frame loop of one scene (MPM_Driver::Advance_To_Target_Time,
// MPM_Driver.cpp:55-80): simulated time toward the frame end, the dt rule's
// parameters (host-rounded exactly as the host rule rounds them), and the
// substeps taken.
struct FrameState {
    float time, target;
    float cfl, dx, min_dt, max_dt;
    float last_dt;
    int done, steps, pad;
};

// CG scalars of one scene.  rho of iteration k lives in rho_hist[k & 1].
struct CGState {
    double rho_hist[2], alpha, beta, conv, rr, sq;
    int iter, done, converged, multiplies;
};

struct KP {
    // scenes
    SceneParams* scn;
    SceneState* sst;
    CGState* cg;
    const int* slot_scene;  // per global slot: the scene that owns it, -1 if none
    int nscenes;
    int solo;        // the scene of the single-scene CG kernels (0 for a one-scene context)
    int cg_target;   // finished solves that end the CG: 1 on the single-scene path, else nscenes
    // particles: field-major arrays of capacity pcap
    float* pcur;
    float* palt;
    int* idcur;
    int* idalt;
    int* keycur;     // global key slot*64 + cell per particle, -1 = removed or empty
    int* keyalt;
    int* scur;       // scene of each particle
    int* salt;
    int* prank;
    int* cell_order;  // per particle slot of the binned layout: the index it is filled from
    float* rcache;   // rotation R per particle (sorted order), field-major
    float* pop;      // per particle (sorted order) op_record_floats(D) floats:
                     // the stencil offsets and the linearised CG force operator
    int64_t pcap;
    // slots and blocks
    int nslots;        // all scenes
    int* slot_count;
    int* slot_touched;
    int* slot_active;
    int* block_slot;
    int* block_scene;
    int* block_pstart;
    int* block_nbr;    // per active block, the 3^D neighbours' active indices (-1: none)
    int* block_tnbr;   // the same, -1 also for neighbours without particles (no tile)
    int* block_org;    // per active block, its D coordinates in its scene's slot grid (3 ints)
    int* block_order;  // each XCD's eighth of the active blocks by descending particle count ([0, bcap))
                       // and by descending fullest cell ([bcap, 2 bcap), for the CG operator)
    int* block_full;   // per active block, the particle count of its fullest cell
    int* cell_count;   // per slot, 64 cell counters (zero between substeps)
    int* cell_start;   // per active block, 65 particle offsets (cell ranges)
    int bcap;
    int* scan_part;    // 2 * nparts
    int nparts;
    // grid channels, capacity gcap nodes each; vector channels hold D arrays
    float* g_m;
    float* g_v;
    float* g_vs;
    float* g_q;      // explicit force, then the CG q
    float* g_s;
    float* g_r;
    unsigned* g_flags;
    int64_t gcap;
    float* tiles;    // per active block, (D+1) * (B+2)^D floats
    // control
    Counters* ctr;
    double* bsum;    // partial sums of r.r, per active block (per work-group for one scene)
    double* bmax;    // partial max |r|^2, same layout
    double* bsq;     // partial sums of s.q, same layout
    unsigned* bu;    // 2 * bcap: per block max |V|^2 bits | removals (G2P)
};

// Per-substep scene settings, passed by value (kernel arguments are copied
// at launch, so consecutive substeps never race on a host buffer).
constexpr int STEP_ARGS_MAX = 480;
struct StepArgs {
    int first, n;          // scenes [first, first + n)
    int force_active;      // 1: every scene active (grid build, stage-by-stage calls)
    int clear_waiting;     // 1: volumes are set, clear the waiting flags
    float dt[STEP_ARGS_MAX];
};

// The caller's particle-major arrays of an allocation-free substep (device
// memory): inputs, and the outputs written back at each particle's index.
struct NoallocIn {
    const float* x;
    const float* v;
    const float* F;       // may be null: identity
    const float* Jp;      // may be null: 1 (an EOS material: its density)
    const float* mass;
    const float* volume;  // 0: the particle waits for Process_Waiting_Particles
    int* valid;
    float* d_dt;          // the scene's dt in device memory (read by k_set_step_dev)
    int* slot_scene;      // the context's slot owners (all slots: scene 0)
    int n;
    float dt;
};
struct NoallocOut {
    float* x;
    float* v;
    float* F;   // may be null
    float* Jp;  // may be null
    float* volume;
    int* valid;  // 1: binned and inside the domain after G2P; 0: removed
};

// launchers (host)
void launch_set_step(const KP& p, hipStream_t s, const StepArgs& a);
void launch_set_step_dev(const KP& p, hipStream_t s, const float* dt);
void launch_frame_dt(const KP& p, hipStream_t s, FrameState* f, float* dt);  // nscenes dts in device memory
void launch_clear_waiting(const KP& p, hipStream_t s);
void launch_remove_scene(const KP& p, hipStream_t s, int grid, int scene);
template <int D> void launch_bin(const KP& p, hipStream_t s, int grid);
void launch_scan(const KP& p, hipStream_t s);
void launch_copy(const float4* src, float4* dst, int64_t n4, hipStream_t s);
template <int D> void launch_block_setup(const KP& p, hipStream_t s, int grid);
// block_order: every XCD range of the active blocks sorted by descending particle count
// block order, scene ranges and the in-cell order scatter (cell_groups work-groups)
void launch_setup_order(const KP& p, hipStream_t s, int cell_groups);
void launch_poll(const KP& p, hipStream_t s, int* dst);
template <int D> void launch_reorder(const KP& p, hipStream_t s, int grid);
template <int D> void launch_p2g(const KP& p, hipStream_t s, int grid);
template <int D> void launch_normalize(const KP& p, hipStream_t s, int grid);
template <int D> void launch_init_volume(const KP& p, hipStream_t s, int grid);
template <int D> void launch_particle_force(const KP& p, hipStream_t s, int grid, bool constitutive, bool force);
template <int D> void launch_grid_update(const KP& p, hipStream_t s, int grid, bool force, bool collision, bool cg_init, bool write_f);
template <int D> void launch_cg_force(const KP& p, hipStream_t s, int grid, bool restart, int k, bool one);
// the same with one wave per block and a lane's cell accumulated in registers
template <int D> void launch_cg_force_cell(const KP& p, hipStream_t s, int grid, bool restart, int k, bool one);
// `one`: a batch of one scene (per-work-group partials, see k_cg_update_s)
template <int D> void launch_cg_restart_node(const KP& p, hipStream_t s, int grid, bool one);
template <int D> void launch_cg_update_s(const KP& p, hipStream_t s, int grid, int k);
template <int D> void launch_cg_node1(const KP& p, hipStream_t s, int grid, int k, bool one);
template <int D> void launch_cg_node2(const KP& p, hipStream_t s, int grid, int k, bool one);
void launch_cg_reduce_sq(const KP& p, hipStream_t s, int k);
void launch_cg_reduce_rr(const KP& p, hipStream_t s, int k);
template <int D> void launch_g2p(const KP& p, hipStream_t s, int grid);
template <int D> void launch_max_speed(const KP& p, hipStream_t s, int grid);
template <int D> void launch_noalloc_load(const KP& p, hipStream_t s, int grid, const SceneParams& sp, const NoallocIn& a);
template <int D> void launch_noalloc_store(const KP& p, hipStream_t s, int grid, const NoallocOut& o);

}  // namespace mpm
