// Kernel parameter block and launch entry points of the gfx950 substep.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "mpm_device.h"

namespace mpm {

constexpr int WG = 256;          // work-group size of every kernel
// floats per particle of the CG force operator record (see OpRec)
constexpr int op_record_floats(int D) { return D == 3 ? 24 : 12; }
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_TILE = WG * SCAN_PER_THREAD;
constexpr int MAX_PLANES = 8;
constexpr unsigned MPM_FLAG_INTERIOR_DEV = 1u;  // Cell_Type_Interior (Flag_Setup_Helper.h:32)

// Device counters of one scene.
struct Counters {
    int nblocks;        // active blocks after binning
    int n_live;         // live particles (after the last binning / upload)
    int n_prev;         // particles in the buffer before the last binning
    int overflow;       // block capacity exceeded
    int nonfinite;      // NaN/Inf or out-of-domain particle met in binning
    unsigned max_v2;    // bits of max |V|^2 (non-negative float) after G2P
    int n_removed;      // particles removed by the last G2P
    int pad;
};

// CG scalars.  rho of iteration k lives in rho_hist[k & 1] so a kernel never
// reads a slot that another work-group of the same kernel writes.
struct CGState {
    double rho_hist[2], alpha, beta, conv, rr, sq;
    int iter, done, converged, multiplies;
};

struct KP {
    // scene
    int counts, nb, nslots;
    float dx, inv_dx;
    Material mat;
    float gravity[3];
    float flip;
    int nplanes;
    float pn[MAX_PLANES][3], ps[MAX_PLANES][3], pmu[MAX_PLANES];
    int crop_ghost_x, project_mode;
    double cg_tol;
    int cg_max, cg_min;
    // particles: field-major arrays of capacity pcap
    float* pcur;
    float* palt;
    int* idcur;
    int* idalt;
    int* keycur;     // home slot per particle, -1 = removed
    int* keyalt;
    int* prank;
    float* rcache;   // rotation R per particle (sorted order), field-major
    float* pop;      // per particle (sorted order) op_record_floats(D) floats:
                     // X and the linearised CG force operator
    int64_t pcap;
    // slots and blocks
    int* slot_count;
    int* slot_touched;
    int* slot_active;
    int* block_slot;
    int* block_pstart;
    int* block_nbr;
    int* cell_count;   // per slot, 64 cell counters (zero between substeps)
    int* cell_start;   // per active block, 65 particle offsets (cell ranges)
    int bcap;
    int* scan_part;  // 2 * nparts
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
    float* tiles;     // per active block, (D+1) * (B+2)^D floats
    // control
    Counters* ctr;
    CGState* cg;
    double* red;      // 3 * red_stride partials: s.q | r.r | max |r|^2
    int red_stride;   // node-kernel work-groups
    int red_count;
    unsigned* wg_part;  // 2 * wg_part_count: per-work-group max |v|^2 | removals of G2P
    int wg_part_count;
    float dt;
};

// launchers (host)
template <int D> void launch_bin(const KP& p, hipStream_t s, int grid);
template <int D> void launch_scan(const KP& p, hipStream_t s);
template <int D> void launch_block_setup(const KP& p, hipStream_t s, int grid);
template <int D> void launch_reorder(const KP& p, hipStream_t s, int grid);
template <int D> void launch_p2g(const KP& p, hipStream_t s, int grid);
template <int D> void launch_normalize(const KP& p, hipStream_t s, int grid);
template <int D> void launch_init_volume(const KP& p, hipStream_t s, int grid);
template <int D> void launch_particle_force(const KP& p, hipStream_t s, int grid, bool constitutive, bool force);
template <int D> void launch_grid_update(const KP& p, hipStream_t s, int grid, bool force, bool collision, bool cg_init, bool write_f);
template <int D> void launch_cg_force(const KP& p, hipStream_t s, int grid, bool restart);
template <int D> void launch_cg_restart_node(const KP& p, hipStream_t s, int grid);
template <int D> void launch_cg_update_s(const KP& p, hipStream_t s, int grid, int k);
template <int D> void launch_cg_node1(const KP& p, hipStream_t s, int grid);
template <int D> void launch_cg_node2(const KP& p, hipStream_t s, int grid, int k);
void launch_cg_finalize(const KP& p, hipStream_t s, int k_end);
template <int D> void launch_g2p(const KP& p, hipStream_t s, int grid);
template <int D> void launch_max_speed(const KP& p, hipStream_t s, int grid);

}  // namespace mpm
