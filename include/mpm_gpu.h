/*
 * mpm_gpu.h -- C ABI of the MI355X (gfx950) MPM snow substep.
 *
 * The library replaces the CPU substep behind the reference's
 * MPM_Example / MPM_Driver surface.  Every entry point below names the
 * reference interface it stands in for (paths are relative to the
 * reference repository SoldierDown/Non_Ficks_Diffusion).
 *
 * Conventions
 *   - extern "C", plain pointers and sizes; no C++ or torch types.
 *   - Every function returns an int status: MPM_OK (0) on success, a
 *     negative MPM_E_* code otherwise; mpm_last_error() gives the text.
 *     (The reference returns void everywhere and ignores failures, e.g.
 *     the CG result at MPM_Example.cpp:711.)
 *   - Host buffers are owned by the caller; device memory is owned by the
 *     library.  One HIP stream per handle.  A handle is not thread safe:
 *     one host thread drives one handle, as one thread drives
 *     MPM_Driver in the reference (mpm_3d/main.cpp:16-24).
 *   - Particle arrays on the host are particle-major: x[i*dim+a],
 *     F[(i*dim+a)*dim+b] (row a, column b of particle i's F).
 *   - Domain is [0,1]^dim with counts[0] cells per axis (cubic cells,
 *     dx = 1/counts[0]); grid nodes sit at cell centres
 *     (mpm_3d/Standard_Tests/Standard_Tests.h:44, MPM_Particle.h:81-83).
 */
#ifndef MPM_GPU_H
#define MPM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_E_INVALID (-1)       /* bad argument */
#define MPM_E_HIP (-2)           /* HIP runtime error */
#define MPM_E_CAPACITY (-3)      /* particle or block capacity exceeded */
#define MPM_E_NOT_READY (-4)     /* call order violated (e.g. substep before initialize) */
#define MPM_E_NONFINITE (-5)     /* NaN/Inf detected in particle state */
#define MPM_E_BUDGET (-6)        /* serving admission: over the HBM budget */

#define MPM_MAX_PLANES 8
#define MPM_MAX_MATERIALS 8

/* Node flag bits (the reference uses Nova's SPGrid flags,
 * Flag_Setup_Helper.h:32 sets Cell_Type_Interior where m > 0). */
#define MPM_FLAG_INTERIOR 1u

/* MPM_Plane_Barrier (MPM_Plane_Barrier.h:13-32): half-space collider with
 * Coulomb friction mu; the node is inside when (x - surface).normal < 0. */
typedef struct mpm_plane {
    float normal[3];
    float surface[3];
    float mu;
    int32_t sticky; /* unused, as in the reference (MPM_Plane_Barrier.h:23) */
} mpm_plane;

/* A particle material: the constitutive model a reference particle carries
 * (MPM_Constitutive_Model.h:78-93; mu0/lambda0 from Compute_Lame_Parameters,
 * MPM_Constitutive_Model.cpp:22-27), or an EOS fluid (MPM_Particle.h:33-37,
 * 51-55).  A scene holds a small table of them and every particle names
 * its entry. */
typedef struct mpm_material {
    float mu0;
    float lambda0;
    float stretching_yield;  /* 1 + theta_s */
    float compression_yield; /* 1 - theta_c */
    float hardening_factor;  /* xi */
    float hardening_max;     /* FLT_MAX in the reference */
    int32_t plastic;
    /* EOS fluid particle: no constitutive model; explicit force
     * (V0/rho) K rho^(gamma-1) grad w (MPM_Example.cpp:774-782), CG force
     * V0 K gamma rho^(gamma-1) (grad w . v) (MPM_CG_System.h:169-174,
     * 205-207), density divided by 1 + dt tr(grad v) in G2P
     * (MPM_Example.cpp:646).  The particle's density rides in its Jp slot. */
    int32_t eos;
    float bulk_modulus;      /* K; the reference's default is 1e7 */
    float gamma;             /* 7 */
    float density;           /* initial density, 1 */
} mpm_material;

/* Scene description: what MPM_Example::Parse_Options and the scene hook
 * Initialize_Particles set up (MPM_Example.cpp:73-119, 925-935;
 * mpm_3d/Standard_Tests/Standard_Tests.h:39-154), with the reference's
 * hard-coded constants turned into fields. */
typedef struct mpm_scene_desc {
    int32_t dim;                /* 2 or 3 */
    int32_t counts;             /* cells per axis (-size) */
    float gravity[3];           /* (0,-2,0) hard-coded at MPM_Example.cpp:81 */
    float flip;                 /* 0.9 (MPM_Example.cpp:82) */
    float cfl;                  /* 0.1 (-cfl default, MPM_Example.cpp:862) */
    float min_dt, max_dt;       /* 1e-6, 5e-3 (MPM_Driver.cpp:58-59) */
    float frame_rate;           /* frames per second of Time_At_Frame */
    float cg_tolerance;         /* 1e-7 in the reference (MPM_Example.cpp:78) */
    int32_t cg_max_iterations;  /* 10000 in the reference (MPM_Example.cpp:79) */
    int32_t cg_min_iterations;  /* 0 (MPM_Example.cpp:711) */
    int32_t nplanes;
    mpm_plane planes[MPM_MAX_PLANES];
    mpm_material material;
    /* 1: particle->grid scatters skip nodes whose x index lies outside
     * [0, counts-1], which is what the reference's x-slab crop does
     * (Tools/Cropped_Influence_Iterator.h:31-42 with the slab bounds of
     * MPM_Example.cpp:468-474).  0: every ghost node receives scatters. */
    int32_t crop_ghost_x;
    /* CG Project (Implicit_Force_Helper/Project_Helper.h:30-45).  The
     * reference's predicate compares Nova's Vector with a scalar
     * (Project_Helper.h:39), whose meaning Nova fixes:
     * 1 = no-op (the predicate is false for every plane the reference and
     *     the scene presets define; the presets' default),
     * 0 = zero the CG vector at every node inside a plane
     *     ((x_i - s).n < 0). */
    int32_t project_mode;
    /* Material table: entry 0 is `material`; nmaterials (0 or 1: only
     * `material`) up to MPM_MAX_MATERIALS, entries 1.. in extra_materials.
     * Particles uploaded without ids use entry 0. */
    int32_t nmaterials;
    mpm_material extra_materials[MPM_MAX_MATERIALS - 1];
} mpm_scene_desc;

typedef struct mpm_step_stats {
    int32_t cg_iterations;      /* completed CG updates (I) */
    int32_t cg_multiplies;      /* operator applications (M = I + 1) */
    int32_t cg_converged;
    int32_t active_blocks;      /* 4^3 (3D) / 8^2 (2D) blocks with storage */
    int64_t live_particles;     /* after removal */
    double cg_residual;         /* convergence norm at exit (L-inf of |r|) */
    double max_speed;           /* max |V| over particles after the substep */
} mpm_step_stats;

typedef struct mpm_context* mpm_handle;

/* Launch shapes of a handle's kernels (no counterpart in the reference,
 * whose loops run on CPU threads).  Every count is a multiple of 8 (the
 * XCDs of an MI355X). */
typedef struct mpm_launch_config {
    int32_t block_groups;     /* work-groups of the per-block kernels (P2G, forces, CG operator, G2P);
                                 each takes about active_blocks / block_groups 4^3 blocks */
    int32_t node_groups;      /* work-groups of the per-node kernels (normalize, grid update, CG vector
                                 updates), at most 2048 */
    int32_t particle_groups;  /* work-groups of the per-particle kernels (binning, reorder); particles
                                 per thread = live particles / (256 particle_groups) */
    int32_t operator_mode;    /* CG operator: 0 = four waves per block, a cell's particles split over
                                 them; 1 = one wave per block, a lane runs every particle of its cell */
    int32_t operator_groups;  /* work-groups of the CG operator (four waves each) */
    int32_t from_cache;       /* mpm_autotune: 1 if taken from the size-bucket cache */
    float block_ms, node_ms, particle_ms, operator_ms;  /* mpm_autotune: the chosen shapes' timing sequences */
} mpm_launch_config;

/* ---- library ---------------------------------------------------------- */
const char* mpm_last_error(void);
/* Writes "<device name>;<CUs>;<arch>" into buf. */
int mpm_device_info(char* buf, int32_t len);

/* ---- scene lifecycle (MPM_Example constructor / destructor,
 *      MPM_Example.cpp:73-135) ------------------------------------------ */
/* Creates a scene with room for `capacity` particles. */
int mpm_create(const mpm_scene_desc* desc, int64_t capacity, mpm_handle* out);
int mpm_destroy(mpm_handle h);

/* Replaces MPM_Example::particles being filled by Initialize_Particles
 * (MPM_Example.h:306, mpm_3d/Standard_Tests/Standard_Tests.h:47-154).
 * volume may be NULL: the particles then wait for
 * Process_Waiting_Particles (MPM_Example.cpp:827-840) in mpm_initialize.
 * F and Jp may be NULL (identity and 1). */
int mpm_upload_particles(mpm_handle h, int64_t n, const float* x, const float* v,
                         const float* F, const float* Jp, const float* mass,
                         const float* volume);
/* The same with each particle's entry in the scene's material table
 * (MPM_Particle's per-particle constitutive model and eos flag,
 * MPM_Particle.h:33-37; NULL: all 0).  An EOS particle uploaded with
 * Jp == NULL starts at its material's density. */
int mpm_upload_particles_ex(mpm_handle h, int64_t n, const float* x, const float* v, const float* F,
                            const float* Jp, const float* mass, const float* volume,
                            const int32_t* material);

/* Copies live particles back in upload order (ids ascending); removed
 * particles are skipped, their ids missing from `ids`.  Any output
 * pointer may be NULL.  Replaces reading MPM_Example::particles
 * (MPM_Example.h:271) and Write_Output_Files (MPM_Example.cpp:1049-1063).
 * For an EOS particle the Jp output holds its density. */
int mpm_download_particles(mpm_handle h, int64_t cap, float* x, float* v, float* F,
                           float* Jp, float* mass, float* volume, int32_t* ids,
                           int64_t* n_out);
int mpm_download_particles_ex(mpm_handle h, int64_t cap, float* x, float* v, float* F,
                              float* Jp, float* mass, float* volume, int32_t* ids,
                              int32_t* material, int64_t* n_out);

/* MPM_Example::Add_Fluid_Source with Allocate_Particle
 * (MPM_Example.cpp:925-1027; the call sits, commented out, at the start of
 * Advance_Step, MPM_Driver.cpp:108): after mpm_initialize, n particles join
 * scene `scene` (0 for a single scene) between substeps as waiting
 * particles -- F = I, Jp = 1 (an EOS material: its density), volume from
 * the grid mass in the next substep's Process_Waiting_Particles -- with the
 * scene's next ids.  They are binned, and so visible to downloads and
 * stats, from the next substep on.  material may be NULL (entry 0).
 * MPM_E_CAPACITY if the handle's particle capacity is exhausted. */
int mpm_add_particles(mpm_handle h, int32_t scene, int64_t n, const float* x, const float* v,
                      const float* mass, const int32_t* material);

/* MPM_Driver::Initialize (MPM_Driver.cpp:18-32): one grid build, P2G and
 * Process_Waiting_Particles for particles uploaded without volume.
 * MPM_E_CAPACITY if the particles need more active blocks than the handle
 * stores (nothing is built; upload a set that fits). */
int mpm_initialize(mpm_handle h);

/* MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116): one substep with the
 * given dt.  Asynchronous unless stats != NULL (then it waits and fills
 * stats).  A substep whose particles need more active blocks than the
 * handle stores changes nothing and returns MPM_E_CAPACITY (seen when the
 * CG's done poll or the stats read the device counters). */
int mpm_substep(mpm_handle h, float dt, mpm_step_stats* stats);

/* MPM_Driver::Advance_To_Target_Time (MPM_Driver.cpp:55-80): substeps from
 * `time` until `target_time`, each with dt = clamp(cfl*dx/max(vmax,1e-2),
 * min_dt, max_dt) from the device-reduced max speed, shortened or halved
 * near the target exactly as the reference does (float arithmetic).
 * Returns the substep count and the final time; fills stats of the last
 * substep when stats != NULL. */
int mpm_advance_to_time(mpm_handle h, double time, double target_time,
                        int32_t* substeps_out, double* time_out, mpm_step_stats* stats);

/* MPM_Example::Max_Particle_Velocity (MPM_Example.cpp:375-387). */
int mpm_max_particle_velocity(mpm_handle h, float* out);

int mpm_synchronize(mpm_handle h);

/* ---- per-stage entry points for parity (MPM_Driver.cpp:103-116) ------ */
enum mpm_stage {
    MPM_STAGE_BIN = 0,          /* Initialize_SPGrid + Update_Particle_Weights + Group_Particles
                                   + Rasterize_Voxels (MPM_Example.cpp:325-331, 418-427, 464-485, 291-311) */
    MPM_STAGE_P2G = 1,          /* Rasterize incl. flags and v/=m (MPM_Example.cpp:520-547) */
    MPM_STAGE_WAITING = 2,      /* Process_Waiting_Particles (MPM_Example.cpp:827-840) */
    MPM_STAGE_CONSTITUTIVE = 3, /* Update_Constitutive_Model_State (MPM_Example.cpp:563-571) */
    MPM_STAGE_EXPLICIT_FORCE = 4, /* Apply_Explicit_Force (MPM_Example.cpp:757-790) */
    MPM_STAGE_COLLISION = 5,    /* Grid_Based_Collision (MPM_Example.cpp:802-806) */
    MPM_STAGE_CG = 6,           /* RHS + Reset_Solver_Channels + CG solve (MPM_Example.cpp:700-711) */
    MPM_STAGE_G2P = 7           /* G2P, F and X update, removal (MPM_Example.cpp:631-660) */
};
/* Runs one stage.  Every stage but BIN and CONSTITUTIVE needs the particles
 * as the last BIN placed them (MPM_E_NOT_READY after G2P), and CG also needs
 * CONSTITUTIVE to have run after that BIN: the CG operator reads the
 * per-particle operator records the constitutive update writes, in the
 * order of the binning (MPM_E_NOT_READY otherwise). */
int mpm_run_stage(mpm_handle h, int32_t stage, float dt);

/* Grid channels readable after a stage (names follow MPM_Data.h:12-46). */
enum mpm_channel {
    MPM_CH_MASS = 0,            /* ch0 */
    MPM_CH_VELOCITY = 1,        /* ch1..3 (momentum after scatter, velocity after P2G) */
    MPM_CH_VELOCITY_STAR = 2,   /* ch4..6 (also the CG solution) */
    MPM_CH_FORCE = 3,           /* ch7..9 */
    MPM_CH_CG_R = 4,            /* ch19..21 */
    MPM_CH_CG_S = 5,            /* ch16..18 */
    MPM_CH_CG_Q = 6             /* ch13..15 */
};
/* Dense copy of a grid channel over node indices [-1, counts]^dim, i.e.
 * (counts+2)^dim nodes, x slowest; vector channels give dim values per
 * node.  Nodes outside active blocks read 0. */
int mpm_read_grid(mpm_handle h, int32_t channel, float* out, int64_t out_len);
/* Dense node flags over the same index range. */
int mpm_read_flags(mpm_handle h, uint32_t* out, int64_t out_len);
/* Sparse flags for the viewer's hierarchy files (MPM_Example::
 * Write_Output_Files, MPM_Example.cpp:1061-1062, read back by
 * MPM_Renderable.h:125-168): for every active block its first node index
 * (dim ints, node index as in mpm_read_flags) and its flags, 4^3 (2D: 8^2)
 * per block with the last axis fastest.  Pointers may be NULL. */
int mpm_read_block_flags(mpm_handle h, int64_t cap_blocks, int32_t* block_origin, uint32_t* flags,
                         int32_t* nblocks);
/* Binning result: active block coordinates (dim ints each, block index =
 * (node index + B) / B with B = 4 in 3D, 8 in 2D), particles homed in each
 * block, and for every live particle its home block's ordinal and its id.
 * Any pointer may be NULL; counts are returned through nblocks/nparticles. */
int mpm_read_bins(mpm_handle h, int64_t cap_blocks, int32_t* block_coords,
                  int32_t* block_particle_counts, int64_t cap_particles,
                  int32_t* particle_block, int32_t* particle_ids, int32_t* nblocks,
                  int64_t* nparticles);
/* CG diagnostics of the last solve. */
int mpm_cg_stats(mpm_handle h, mpm_step_stats* stats);

/* ---- timing (HIP events on the handle's stream) ------------------------ */
/* Enables per-kernel-class event timing; mpm_kernel_times returns the
 * accumulated milliseconds and launch counts per class since the last
 * reset.  Class ids: see mpm_kernel_class_name. */
int mpm_enable_timing(mpm_handle h, int32_t on);
int mpm_kernel_times(mpm_handle h, int32_t nclasses, double* ms, int64_t* launches);
const char* mpm_kernel_class_name(int32_t cls);
int32_t mpm_kernel_class_count(void);
/* Benchmark support (no counterpart in the reference): launches the CG
 * operator application (MPM_CG_System::Force, MPM_CG_System.h:150-215) `reps`
 * times back to back on the current particle state and search direction,
 * bracketed by two HIP events on the handle's stream, and returns the
 * average milliseconds per launch.  Needs a completed substep whose operator
 * records are current (not after mpm_time_stage of MPM_STAGE_CONSTITUTIVE,
 * MPM_E_NOT_READY then); leaves the state of that substep unchanged except
 * the operator's tile buffer. */
int mpm_time_cg_operator(mpm_handle h, int32_t reps, float* avg_ms);

/* Benchmark support (no counterpart in the reference): launches one
 * stage kernel of the last substep `reps` times back to back on its state,
 * bracketed by two HIP events on the handle's stream, and returns the
 * average milliseconds per launch.  stage: MPM_STAGE_P2G (the scatter
 * kernel), MPM_STAGE_CONSTITUTIVE (the substep's fused constitutive update
 * and explicit force) or MPM_STAGE_G2P.  Needs a completed substep; the
 * particle state is saved before the launches and restored after them, the
 * grid scratch they write is not (the next substep rebuilds it). */
int mpm_time_stage(mpm_handle h, int32_t stage, int32_t reps, float* avg_ms);

/* Benchmark support (no counterpart in the reference): the measured copy
 * bandwidth roof of the current device.  A streaming copy kernel moves
 * `bytes` (a multiple of 16) from one device buffer to another `reps` times
 * back to back, after one untimed launch, between two HIP events on a stream
 * of its own; *gbs = 2 * bytes * reps / elapsed (read plus write).  The call
 * allocates and frees the two buffers. */
int mpm_measure_copy_bandwidth(int64_t bytes, int32_t reps, double* gbs);

/* ---- one substep without allocation ------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116) of one
 * scene on particle arrays in device memory that the caller owns, with no
 * handle and no allocation.  Arrays are particle-major as on the host (x[i*dim+a],
 * F row-major per particle); x, v, F, Jp and volume are read and then
 * overwritten with the state after the substep at each particle's index; F
 * and Jp may be NULL (identity and 1, not written).  A particle of volume 0
 * gets its volume from the grid mass first (Process_Waiting_Particles,
 * MPM_Example.cpp:827-840), so a first call on fresh particles is
 * MPM_Driver::Initialize plus one substep.  valid[i] = 1 for particles inside
 * the domain after the substep, 0 for removed ones (MPM_Example.cpp:650-660).
 * Every particle takes material entry 0; the CG runs a fixed count of
 * cg_max_iterations iterations (those after convergence return at once).
 * stream: a hipStream_t, or NULL for the null stream.  With stats == NULL the
 * call is asynchronous on the stream and calls no synchronizing HIP routine;
 * with stats it waits and returns MPM_E_CAPACITY when the particles needed
 * more than MPM_NOALLOC_MAX_BLOCKS active blocks (the substep was then a
 * no-op) and MPM_E_NONFINITE as mpm_substep does.
 * Every working buffer lives in one static device array per GPU (40 MiB,
 * held until the process ends) that every call on that GPU shares and that
 * keeps the last call's contents: calls that overlap on one GPU race.  Sizes
 * beyond the limits are refused with MPM_E_CAPACITY before any GPU work;
 * CG caps above 256 with MPM_E_INVALID. */
#define MPM_NOALLOC_MAX_PARTICLES 65536
#define MPM_NOALLOC_MAX_COUNTS 64
#define MPM_NOALLOC_MAX_BLOCKS 2048
int mpm_substep_noalloc(const mpm_scene_desc* desc, int64_t n, float* x, float* v, float* F, float* Jp,
                        const float* mass, float* volume, int32_t* valid, float dt, void* stream,
                        mpm_step_stats* stats);

/* ---- launch shapes, autotuner, captured substeps ------------------------ */
/* (No counterpart in the reference.)  Current launch shapes of a handle. */
int mpm_get_launch_config(mpm_handle h, mpm_launch_config* cfg);
/* Sets the launch shapes (results do not depend on them). */
int mpm_set_launch_config(mpm_handle h, const mpm_launch_config* cfg);
/* Autotuner: times candidate work-group counts for the per-block,
 * per-node and per-particle kernels on the handle's current particles
 * (re-binned, with the grid and CG set up for dt; nscenes dts, or one dt for
 * a single scene) and keeps the fastest of each.  Particle state is not
 * advanced; grid and CG buffers are left as scratch, so call it between
 * substeps.  flags & MPM_TUNE_CACHED: reuse the result of an earlier tune of
 * a handle in the same size bucket (dimension, one scene or a batch,
 * log2 of particle capacity and of block slots) without timing.  Results
 * go to *out (may be NULL). */
#define MPM_TUNE_CACHED 1
int mpm_autotune(mpm_handle h, const float* dt, int32_t flags, mpm_launch_config* out);
/* Captured substeps: with on != 0 every substep replays a HIP graph of its
 * whole launch sequence (captured at the first such substep, one graph per
 * particle-buffer parity), the CG with a fixed count of max(cg_max_iterations)
 * iterations whose kernels return at once for finished solves, and dts read
 * from device memory.  Results equal the launched substep's.  Needs CG caps
 * of at most 256 iterations. */
int mpm_set_graph(mpm_handle h, int32_t on);

/* ---- batches of scenes ---------------------------------------------------
 * No counterpart in the reference, which runs one MPM_Example per process.
 * A batch is many independent scenes of one dimension that advance together:
 * every kernel of a substep covers all of them (their block slots and
 * particles are concatenated with per-scene offsets, and each block reads
 * its scene's parameters from a scene table), and every scene runs its own
 * CG solve with its own dot products.  A handle from mpm_create is a batch
 * of one; every single-scene call above acts on scene 0. */
int mpm_batch_create(const mpm_scene_desc* descs, const int64_t* capacities, int32_t nscenes,
                     mpm_handle* out);
int32_t mpm_batch_size(mpm_handle h);
/* Particles of one scene (as mpm_upload_particles); every scene is uploaded
 * before mpm_initialize, which initialises the whole batch. */
int mpm_batch_upload(mpm_handle h, int32_t scene, int64_t n, const float* x, const float* v,
                     const float* F, const float* Jp, const float* mass, const float* volume);
int mpm_batch_upload_ex(mpm_handle h, int32_t scene, int64_t n, const float* x, const float* v,
                        const float* F, const float* Jp, const float* mass, const float* volume,
                        const int32_t* material);
/* One substep of every scene s with dt[s]; dt[s] <= 0 leaves scene s
 * unchanged (MPM_Driver::Advance_Step per scene). */
int mpm_batch_substep(mpm_handle h, const float* dt);
/* Max |V| per scene (nscenes values) after the last substep. */
int mpm_batch_max_velocity(mpm_handle h, float* out);
int mpm_batch_stats(mpm_handle h, int32_t scene, mpm_step_stats* stats);
int mpm_batch_download(mpm_handle h, int32_t scene, int64_t cap, float* x, float* v, float* F,
                       float* Jp, float* mass, float* volume, int32_t* ids, int64_t* n_out);
int mpm_batch_download_ex(mpm_handle h, int32_t scene, int64_t cap, float* x, float* v, float* F,
                          float* Jp, float* mass, float* volume, int32_t* ids, int32_t* material,
                          int64_t* n_out);
/* MPM_Driver::Advance_To_Target_Time (MPM_Driver.cpp:55-80) for every scene:
 * each scene takes the reference's dt rule from its own max speed, and
 * scenes that reach their target idle while the others finish.  time_out
 * and substeps_out (nscenes entries each) may be NULL. */
int mpm_batch_advance_to_time(mpm_handle h, const double* time, const double* target,
                              double* time_out, int32_t* substeps_out);

/* ---- serving: continuous batching on one device arena ---------------- */
/* (No counterpart in the reference, which runs one scene per process.)  A
 * server owns an arena sized from an HBM budget.  Requests are scenes with
 * a frame count; mpm_serve_tick admits queued requests in arrival order
 * while the arena has room (particle space, a free slot range, blocks of
 * storage by estimate, the byte budget), then runs batched substeps until
 * every scene resident at its start has completed a frame
 * (MPM_Driver::Advance_To_Target_Time, MPM_Driver.cpp:55-80, each scene with
 * its own dt toward its own next frame end).  Scenes with shorter frames go
 * on with their next ones meanwhile; between substeps, finished requests
 * leave the arena and queued ones are admitted into the room they free.  A
 * handle is not thread-safe. */
typedef struct mpm_server* mpm_server_handle;

typedef struct mpm_serve_config {
    int32_t dim;                 /* 3 (or 2) */
    int32_t max_scenes;          /* resident scenes at most */
    int32_t max_counts;          /* largest grid of a request, cells per axis */
    int32_t cg_max_iterations;   /* CG cap every request must stay within (<= 256) */
    int64_t hbm_budget_bytes;    /* the arena is sized from it; admission stays within it */
    int32_t graph;               /* replay captured substeps */
    int32_t keep_results;        /* keep finished requests' particles for mpm_serve_download */
    int32_t max_tick_substeps;   /* > 0: a tick also ends after this many substeps */
    int32_t tune;                /* run mpm_autotune on the resident scenes whenever their particle
                                    count has doubled since the last tune (from 1M particles on) */
} mpm_serve_config;

enum { MPM_REQ_QUEUED = 0, MPM_REQ_RUNNING = 1, MPM_REQ_DONE = 2, MPM_REQ_CANCELLED = 3 };

typedef struct mpm_request_status {
    int32_t state;               /* MPM_REQ_* */
    int32_t frames_done, frames_total;
    int32_t substeps;
    int64_t particles;
    int64_t bytes;               /* admission estimate */
    double sim_time;
    double submitted, admitted, finished;  /* server clock in seconds, -1 until then */
} mpm_request_status;

typedef struct mpm_tick_stats {
    int32_t admitted, finished, resident, queued;
    int32_t substeps;            /* batched substeps of the tick */
    int32_t frames;              /* frames completed */
    int64_t particle_substeps;   /* particles of the advancing scenes, summed over substeps */
    int64_t bytes_in_use;        /* admission estimates of the resident scenes */
    double ms;                   /* wall time of the tick, device work included */
} mpm_tick_stats;

int mpm_serve_create(const mpm_serve_config* cfg, mpm_server_handle* out);
int mpm_serve_destroy(mpm_server_handle s);
/* Arena capacities and device bytes. */
int mpm_serve_arena(mpm_server_handle s, int64_t* particles, int64_t* slots, int64_t* blocks, int64_t* bytes);
/* Queues a request (particles copied; F = I, Jp = 1, volumes from the grid
 * at admission).  MPM_E_CAPACITY if it could never fit the arena. */
int mpm_serve_submit(mpm_server_handle s, const mpm_scene_desc* desc, int64_t n, const float* x,
                     const float* v, const float* mass, int32_t frames, int64_t* request_id);
/* The same with each particle's material entry (NULL: all 0). */
int mpm_serve_submit_ex(mpm_server_handle s, const mpm_scene_desc* desc, int64_t n, const float* x,
                        const float* v, const float* mass, const int32_t* material, int32_t frames,
                        int64_t* request_id);
int mpm_serve_tick(mpm_server_handle s, mpm_tick_stats* out);
int mpm_serve_poll(mpm_server_handle s, int64_t request_id, mpm_request_status* out);
/* Particles of a running request (current state) or of a finished one
 * (final state, with keep_results), sorted by id as mpm_download_particles. */
int mpm_serve_download(mpm_server_handle s, int64_t request_id, int64_t cap, float* x, float* v,
                       float* F, float* Jp, float* mass, float* volume, int32_t* ids,
                       int64_t* n_out);
/* Withdraws a queued request, or a running one (its particles leave the
 * arena at once, freeing room for admissions); the request is then
 * MPM_REQ_CANCELLED and can be released.  MPM_E_INVALID if it already
 * finished.  (No reference counterpart: the reference runs one scene per
 * process.) */
int mpm_serve_cancel(mpm_server_handle s, int64_t request_id);
/* Completion times (server clock, seconds) of a request's frames so far;
 * *n = frames completed, at most cap times written. */
int mpm_serve_frame_times(mpm_server_handle s, int64_t request_id, int32_t cap, double* times,
                          int32_t* n);
/* Forgets a finished request. */
int mpm_serve_release(mpm_server_handle s, int64_t request_id);

/* ---- scenes (scene hook Initialize_Particles) ------------------------- */
/* Fills desc with the preset of a named scene and returns the particle
 * count it generates.  Known names: "disk2d", "two_balls", "avalanche",
 * "blob", "slab", and the reference's Standard_Tests scenes "ref2d_test1",
 * "ref2d_test15", "ref2d_test16", "ref2d_test17", "ref2d_test18",
 * "ref2d_test19", "ref2d_test20" (no particles: the fluid source is the host's,
 * through mpm_add_particles), "ref3d_test18", "ref3d_test19". */
int mpm_scene_preset(const char* name, int32_t counts, uint64_t seed, mpm_scene_desc* desc,
                     int64_t* n_particles);
/* Generates the particles of a named scene (same arguments as the preset);
 * arrays must hold n_particles entries. */
int mpm_scene_generate(const char* name, int32_t counts, uint64_t seed, int64_t cap, float* x,
                       float* v, float* mass, int64_t* n_out);
/* Sized variants for batched and served scenes (no reference counterpart:
 * the reference's scenes fix their own particle counts).  target_n > 0
 * scales the preset's shapes toward target_n particles at 8 per cell of the
 * given grid, clamped to the domain, and draws exactly target_n uniform
 * samples; target_n == 0 is the plain preset.  Presets with a fixed count
 * (the test 18 scenes) reject target_n > 0. */
int mpm_scene_preset_sized(const char* name, int32_t counts, uint64_t seed, int64_t target_n,
                           mpm_scene_desc* desc, int64_t* n_particles);
int mpm_scene_generate_sized(const char* name, int32_t counts, uint64_t seed, int64_t target_n,
                             int64_t cap, float* x, float* v, float* mass, int64_t* n_out);
/* Material entry of every particle of a (sized) scene, in generation order
 * (the scenes that mix materials: the reference's 2D test 16). */
int mpm_scene_materials(const char* name, int32_t counts, uint64_t seed, int64_t target_n,
                        int64_t cap, int32_t* material, int64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* MPM_GPU_H */
