/*
 * mpm_oracle.h -- CPU oracle for the MPM snow substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a C++/OpenMP restatement of the
 * reference's algorithm (SoldierDown/Non_Ficks_Diffusion, which cannot be
 * built here: its Nova dependency is not vendored).  The GPU path's parity
 * tests compare against it and bench.py times it as the CPU baseline; the
 * product never links, loads or calls it.
 *
 * Every stage follows the reference stage for stage: AoS particles,
 * x-slab binning with `threads` slabs, a dense grid over the particle
 * bounding box rebuilt every substep, slab-owned cropped scatters, and
 * Nova's unpreconditioned CG with the MPM_CG_System callbacks.  Pieces
 * that live in Nova are restated from mathematics (see mpm_oracle.cpp).
 *
 * Particle arrays use the same layout as include/mpm_gpu.h (particle
 * major; F row-major per particle).
 *
 * How each part is pinned (the reference ships no golden vectors, and it
 * cannot be built without Nova; tests/test_oracle_pins.py holds the checks):
 *   weights          closed forms of MPM_Particle.h:90-100, partition of
 *                    unity, linear reproduction, sum of gradients = 0
 *   SVD              F = U S V^T, orthogonality, det U = det V = +1,
 *                    ordering, sign on the smallest value, special cases,
 *                    against numpy's SVD
 *   plasticity       diagonal-F closed form, Jp update identity
 *   P, dP            P = dPsi/dF and dP = directional derivative of P by
 *                    finite differences; dP symmetric
 *   force map        symmetric, block structure in the SVD frame
 *   explicit force   snow: sum_i f_i = g sum_p m_p and the first moment
 *                    sum_i f_i x_i^T = g (sum_p m_p x_p)^T - sum_p V0 P F^T
 *                    with P from the textbook closed form
 *   CG operator      M-symmetric; solve matches a dense numpy solve; rigid
 *                    translation maps to itself; K u equals the central
 *                    difference of a numpy restatement of the elastic force
 *                    with F_p -> (I + e grad u_p) F_p (fixes V0, dt^2 and
 *                    the transposes)
 *   waiting volume   isolated particle V0 = dx^d / sum_i w_ip^2; a cluster
 *                    against a brute-force textbook-weight sum
 *   collision        closed form of Grid_Based_Collision_Helper.h:72-80
 *   P2G / G2P        mass and momentum conservation; G2P reproduces a linear
 *                    velocity field; free fall without elasticity
 *   binning          every stencil node has exactly one slab owner for any
 *                    thread count
 *   driver dt        sequence of MPM_Driver.cpp:55-80 by hand
 *   EOS particles    explicit force: first moment sum_i f_i x_i^T =
 *                    (sum_p V0/rho K rho^(gamma-1)) I; CG force on a linear
 *                    field: tr(G) (sum_p V0 K gamma rho^(gamma-1)) I; G2P
 *                    density rho / (1 + dt tr B) with F unchanged
 *   material table   a particle of entry k evolves as in a scene whose only
 *                    material is entry k
 * Parity unpinned (no reference behaviour to check against; the readings
 * taken are the scene options of mpm_gpu.h):
 *   - Project_Helper's predicate (Project_Helper.h:39, Nova operator<):
 *     project_mode 1 (no-op) by default, 0 = zero inside planes;
 *   - Nova CG internals beyond the solution: the stop-test order, restart
 *     at iteration 0, s.q <= 0 handling and the Multiply count;
 *   - SPGrid handling of stencil writes at ghost index -1 (crop_ghost_x);
 *   - Time_At_Frame's frame rate (24 fps taken);
 *   - Nova's matrix-plus-scalar in the EOS CG force (MPM_CG_System.h:172),
 *     taken as adding to the diagonal.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H
#include <stdint.h>
#include "../include/mpm_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* option bits */
#define ORACLE_DOT_FP64 1    /* CG inner products and norm accumulate in double */
#define ORACLE_SVD_FP64 2    /* 3x3 / 2x2 SVD computed in double */
#define ORACLE_JP_SCALAR 4   /* carry det(Fp) as a scalar instead of the Fp matrix */

typedef struct oracle_ctx* oracle_handle;

/* precision: 4 = float state (the reference build), 8 = double (FP64 oracle) */
oracle_handle oracle_create(const mpm_scene_desc* desc, int32_t precision, int32_t threads,
                            int32_t options);
void oracle_destroy(oracle_handle h);
/* volume may be NULL (particles then wait for Process_Waiting_Particles);
 * F, Jp may be NULL (identity, 1). */
void oracle_set_particles(oracle_handle h, int64_t n, const float* x, const float* v,
                          const float* F, const float* Jp, const float* mass, const float* volume);
/* The same with each particle's entry in the scene's material table (NULL:
 * all 0); an EOS particle's Jp is its density (NULL: the material's). */
void oracle_set_particles_ex(oracle_handle h, int64_t n, const float* x, const float* v,
                             const float* F, const float* Jp, const float* mass,
                             const float* volume, const int32_t* material);
/* New waiting particles between substeps (Add_Fluid_Source, MPM_Example.cpp:977-1027). */
void oracle_add_particles(oracle_handle h, int64_t n, const float* x, const float* v,
                          const float* mass, const int32_t* material);
int64_t oracle_num_particles(oracle_handle h);
int64_t oracle_num_simulated(oracle_handle h);
/* All particles in creation order; valid[i] = 0 for removed ones. */
void oracle_get_particles(oracle_handle h, double* x, double* v, double* F, double* Jp,
                          double* mass, double* volume, uint8_t* valid);

/* MPM_Driver::Initialize (MPM_Driver.cpp:18-32) */
void oracle_initialize(oracle_handle h);
/* MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116) */
void oracle_advance_step(oracle_handle h, double dt);
/* One stage of Advance_Step; stage ids are enum mpm_stage of mpm_gpu.h. */
void oracle_run_stage(oracle_handle h, int32_t stage, double dt);
/* Populate_Simulated_Particles (MPM_Example.cpp:347-355) */
void oracle_populate_simulated(oracle_handle h);
/* MPM_Example::Max_Particle_Velocity (MPM_Example.cpp:375-387) */
double oracle_max_particle_velocity(oracle_handle h);
/* MPM_Driver::Advance_To_Target_Time (MPM_Driver.cpp:55-80); returns the new time. */
double oracle_advance_to_target_time(oracle_handle h, double time, double target_time,
                                     int32_t* substeps);

/* Dense grid channel over node indices [-1, counts]^dim (x slowest);
 * vector channels give dim values per node (enum mpm_channel). */
void oracle_get_grid(oracle_handle h, int32_t channel, double* out);
void oracle_get_flags(oracle_handle h, uint32_t* out);
/* Applies the CG operator (MPM_CG_System::Multiply followed by Project)
 * to a dense grid vector x (same layout as oracle_get_grid of a vector
 * channel) using the current grid and particle state; writes q. */
void oracle_apply_operator(oracle_handle h, double dt, const double* x, double* q);
/* Overwrites a grid channel (same dense layout as oracle_get_grid) on the
 * nodes of the current grid; for tests that prescribe a grid field. */
void oracle_set_grid(oracle_handle h, int32_t channel, const double* in);
/* Interior-masked M inner product of two dense vectors (Inner_Product_Helper). */
double oracle_inner_product(oracle_handle h, const double* a, const double* b);
void oracle_cg_stats(oracle_handle h, int32_t* iterations, int32_t* multiplies, int32_t* converged,
                     double* residual);
/* Bounding-box cell range of the last grid build (Rasterize_Voxels). */
void oracle_bbox_cells(oracle_handle h, int32_t* lo, int32_t* hi);

/* ---- per-particle maths, for closed-form and invariant tests --------- */
/* Quadratic B-spline weights of a particle at X (MPM_Particle.h:79-100):
 * closest cell, w[3*dim] (row k = node c-1+k), dw[3*dim]. */
void oracle_weights(int32_t dim, int32_t counts, const double* X, int32_t* cell, double* w,
                    double* dw);
/* 3x3 (dim 3) or 2x2 SVD, row-major matrices: F = U diag(S) V^T. */
void oracle_svd(int32_t dim, const double* F, double* U, double* S, double* V);
/* MPM_Constitutive_Model::Precompute + P + Psi for one particle in double:
 * F, Fp-determinant in; F, Jp out, R, P, psi out. */
void oracle_constitutive(int32_t dim, const mpm_material* mat, const double* F_in, double Jp_in,
                         double* F_out, double* Jp_out, double* R, double* P, double* psi);
/* MPM_Constitutive_Model::Times_dP_dF at state (F, R, lambda, mu). */
void oracle_dP(int32_t dim, const double* F, const double* R, double lambda, double mu,
               const double* dF, double* dP);

#ifdef __cplusplus
}
#endif
#endif
