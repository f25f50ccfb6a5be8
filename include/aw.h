/*
 * aw.h — C ABI of the MI355X MSBG FLIP step (anitowave_adaptive_amd).
 *
 * Plain pointers and sizes only.  Every pointer argument named d_* or described
 * as "device" is HIP device memory; work is enqueued on the context's stream
 * and is asynchronous unless the function says it synchronises.  Handles are
 * not thread-safe: one host thread per context, as the reference's classes.
 * The caller owns particle buffers; the library owns grid and solver memory.
 * Destroying a handle that others were made from (a context with live grids
 * or aw_malloc buffers, a grid with live solvers, a solver with live tapes)
 * only closes it; it is released together with the last of them.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference checkout):
 *   aw_grid_create / aw_grid_destroy     MSBGManager ctor / cleanup          include/MSBG.cuh:506-597
 *   aw_grid_build_topology               initRefinementMap, markActiveBlocks, enforceGrading,
 *                                        countBlockElements, exclusive_scan, allocateChannels
 *                                                                            src/main.cu:156-302, include/MSBG.cuh:249-376,599-659
 *   aw_grid_build_topology_radius        the same, with markActiveBlocks's particle-radius levels
 *                                                                            src/main.cu:156-196, include/MSBG.cuh:271-351
 *   aw_grid_layout                       MSBG POD fields                     include/MSBG.cuh:71-106
 *   aw_p2g                               P2GTransfer::run                    include/P2G.cuh:173-196
 *   aw_divergence, aw_project_velocity   (no counterpart: projection is missing upstream)
 *   aw_p2g_particles                     P2GTransfer::run (caller's order kept) include/P2G.cuh:173-196
 *   aw_g2p_advect                        G2PManager::performG2P + advect     include/G2P.cuh:198-227
 *   aw_g2p, aw_advect                    G2PManager::performG2P, ::advect    include/G2P.cuh:192-227
 *   aw_solver_create / _destroy          VCycleSolver ctor / init / dtor     include/VCycle.cuh:161-249
 *   aw_solver_vcycle                     VCycleSolver::runCycle              include/VCycle.cuh:251-304
 *   aw_solver_smooth                     AdaptiveSmoother::solveLevel        include/Smoother.cuh:249-291
 *   aw_solver_residual                   computeExplicitResidual             include/VCycle.cuh:81-131
 *   aw_solver_restrict / _prolong        MultigridTransfer::restrictResiduals / prolongateAndCorrect
 *                                                                            include/MultigridTransfer.cuh:107-149
 *   aw_solver_solve                      main.cu's fixed V-cycle loop        src/main.cu:314-317
 *   aw_step                              the main.cu step sequence, completed with the projection
 *   aw_resample_cells                    sampleMSBG                          include/vtx_visualization.cuh:8-62
 *   aw_remap_cells                       (no counterpart: warm start across topology changes)
 *   aw_remap_cells_noalloc               (no counterpart: aw_remap_cells for callers that cannot allocate)
 *   aw_write_vtk_*, aw_export_*_vtk      exportUniformPressure_VTK, exportMultigrid_VTK, exportParticles_VTK
 *                                                                            include/vtx_visualization.cuh:64-216
 *   aw_*_adjoint, aw_tape_*, aw_rollout_* (no counterpart: the reference has no gradient code; these
 *                                        are the reverse-mode derivative of aw_step and of a rollout)
 */
#ifndef ANITOWAVE_AW_H
#define ANITOWAVE_AW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AW_OK = 0,
  AW_ERR_INVALID = 1,
  AW_ERR_OOM = 2,
  AW_ERR_DEVICE = 3,
  AW_ERR_NOT_CONVERGED = 4,
  AW_ERR_NAN = 5,
  AW_ERR_UNSUPPORTED = 6
} aw_status;

#define AW_MAX_LEVELS 4     /* native levels: cells per block edge 8, 4, 2, 1 */
#define AW_MAX_MG_LEVELS 12

typedef struct aw_ctx aw_ctx;
typedef struct aw_grid aw_grid;
typedef struct aw_solver aw_solver;

/* Grid: dims[] blocks of base_res^3 cells at native level 0 (world size
 * base_res*h), num_levels native levels, distance bands band[l] in blocks
 * (0..254). */
typedef struct {
  int dims[3];
  int base_res;        /* must be 8 */
  int num_levels;      /* 1..AW_MAX_LEVELS */
  float h;             /* finest cell size; a power of two */
  float domain_min[3]; /* must be 0,0,0 */
  int band[AW_MAX_LEVELS];
} aw_grid_desc;

typedef struct {
  uint32_t num_blocks;
  uint32_t total_cells;
  uint32_t total_faces[3];
  int num_mg_levels;
  const uint8_t* d_level;       /* native level per block */
  const uint32_t* d_cell_off;   /* num_blocks+1 */
  const uint32_t* d_face_off[3];
  const uint32_t* d_particle_count; /* particles per block from the last binning */
  const uint32_t* d_particle_off;   /* num_blocks+1 */
  float* d_pressure;            /* solver units: p * dt / (rho h) */
  float* d_rhs;                 /* -divergence of u* (h^2 velocity) */
  uint8_t* d_liquid;            /* 1 = liquid cell */
  float* d_beta[3];             /* face coefficients 1 / rho (0 on the tank walls): the uniform 1, or from
                                   aw_step_params.d_density; written by P2G.  A caller writing them
                                   directly calls aw_grid_beta_changed before the next solve */
  float* d_u_old[3];            /* P2G velocity */
  float* d_u_star[3];           /* + body force */
  float* d_u_new[3];            /* after projection */
  float* d_mass[3];             /* folded particle weight per face */
  float* d_fine;                /* padded fine lattice [block][field 0 new, 1 delta][axis][900]: each
                                   block's 8^3 nodes plus a one-node halo; filled for the blocks within
                                   one block of a block holding particles, others not current */
} aw_layout;

typedef struct {
  int method;          /* 0 = conjugate gradients preconditioned by one V-cycle; 1 = stationary V-cycles */
  int pre_sweeps;      /* V(pre, post) */
  int post_sweeps;
  int coarse_sweeps;   /* symmetric sweeps on the coarsest level */
  int local_sweeps;    /* in-block red-black sweeps per block phase (3 in the
                          Python and C++ wrappers: the sweeps run in LDS, and
                          each one more saves V-cycles; the reference sweeps
                          a block until its residual falls by 0.33, at most 10) */
  float omega;         /* Gauss-Seidel weight (reference 0.85714) */
  float prolong_scale; /* coarse-correction scale (1 for PCG, 0.75 recommended for stationary) */
  int precision;       /* 0 = FP32; 1 = FP32 V-cycles inside FP64 iterative refinement */
  /* The reference's adaptive smoothing (AdaptiveSmoother, Smoother.cuh:105-291), method 1 only (it
   * makes the V-cycle nonlinear, which conjugate gradients cannot take): every smoothing call starts
   * with all blocks active; an active block runs up to max_block_sweeps in-block sweeps and stops
   * after the first whose largest |r| is below block_reduction times the first sweep's; a block still
   * above theta * max|b| (the reference's absolute 1e-4, made relative) activates itself and its six
   * face neighbours for the next iteration.  The flags stay on the device (no host readback). */
  int adaptive;        /* 0 = plain block red-black sweeps (local_sweeps each); 1 = adaptive */
  float theta;         /* activation threshold relative to max |b| of the level (reference 1e-4) */
  int max_block_sweeps; /* reference MAX_BLOCK_ITER 10 */
  float block_reduction; /* reference TARGET_RESIDUAL_REDUCTION 0.33 */
} aw_solver_params;

/* Convergence contract of every solve: AW_OK only when the true relative
 * residual ||b - A p||_2 / ||b||_2, recomputed from the returned p, is at or
 * below rel_tol; otherwise AW_ERR_NOT_CONVERGED with these statistics
 * filled and p left at the last iterate.  With precision 0 the FP32 pressure
 * bounds the attainable true residual (about eps32 |A| |p| / |b|, 1e-5 to
 * 2e-5 relative on free-surface scenes at 256^3): PCG then stops when its
 * recurrence residual has met rel_tol twice (one restart from the true
 * residual in between) and reports the true residual above it as
 * AW_ERR_NOT_CONVERGED.  Tolerances below that floor need precision 1. */
typedef struct {
  int iterations;      /* PCG iterations or V-cycles (outer+inner for mixed precision) */
  int outer_iterations;
  double rel_residual; /* true ||b - A p||_2 / ||b||_2 of the returned p (the recurrence residual only when
                          the iteration cap ended an FP32 PCG solve) */
  double b_norm;
  double recurrence_rel_residual; /* FP32 PCG: the last recurrence residual; otherwise rel_residual */
} aw_solve_stats;

typedef struct {
  float dt;
  float gravity[3];
  float flip_alpha;
  double rel_tol;
  int max_iters;
  int rebuild_topology; /* 1 = levels from the distance bands every step; 2 = from the particle
                           radii d_radius every step (aw_grid_build_topology_radius's rule) */
  int warm_start;       /* 1 = start the pressure solve from the previous step's pressure, remapped
                           cell by cell when the topology changed (no counterpart in the reference,
                           whose p = 0 start main.cu:300 is the default 0); ignored by rollouts */
  const float* d_radius; /* rebuild_topology 2: radius of each particle, indexed by particle id;
                            the step needs d_ids (ids in [0, n); another id fails the step with
                            AW_ERR_INVALID) to find them as it reorders.  Read during aw_step only;
                            aw_rollout_forward copies the radii into its tape for the backward pass */
  const float* d_density; /* two-phase / variable density: relative density of each particle (the
                            reference density is 1), indexed by particle id (needs d_ids), or NULL for
                            beta = 1.  P2G also deposits it, beta = 1 / rho_face on faces with mass
                            (rho_face = sum w rho / sum w; 1 on faces without), and the solver coarsens
                            beta through its levels (the reference's BETA_COEFF / MASS_DENSITY channels
                            and galerkinCoarsen, MSBG.cuh:21-34,401-498).  Not differentiated: the
                            adjoint entry points and rollouts reject it (AW_ERR_UNSUPPORTED) */
} aw_step_params;

typedef struct {
  int topology_changed;
  aw_solve_stats solve;
  uint32_t num_blocks_by_level[AW_MAX_LEVELS];
  uint32_t total_cells;
  uint32_t liquid_cells;
} aw_step_stats;

typedef struct {
  int dims[3];
  uint32_t num_blocks;
  uint32_t total_cells;
  uint32_t total_faces[3];
  int aggregated;          /* 1 if built from the previous level by 2x2x2 block aggregation */
  const uint8_t* d_e;      /* log2 cell edge in fine cells, per block */
  const uint32_t* d_cell_off;
  float* d_x;
  float* d_b;
  float* d_r;
  uint8_t* d_liquid;
  const double* d_x64;     /* level 0 with precision 1: the FP64 solution of the last solve */
  const uint8_t* d_block_sweeps; /* adaptive smoothing: in-block sweeps each block ran in the last
                                    iteration of the last smoothing call (0 = inactive), or NULL */
  const uint8_t* d_active;       /* adaptive smoothing: the blocks that call left active, or NULL */
  uint32_t smoothed_blocks;      /* blocks the plain smoother's launches cover: those holding a
                                    liquid cell at the last setup or solve (all-air blocks are left
                                    out), num_blocks where the level keeps the full lists (tail and
                                    uniform levels, adaptive smoothing) */
} aw_level_info;

const char* aw_status_str(int status);
const char* aw_last_error(void);     /* thread-local detail of the last failure */
int aw_version(void);

int aw_ctx_create(int device, void* stream, aw_ctx** out);
int aw_ctx_destroy(aw_ctx* ctx);
int aw_ctx_synchronize(aw_ctx* ctx);
/* Plumbing for hosts without a device allocator of their own (no reference
 * counterpart): device allocation and a synchronous copy in any direction. */
int aw_malloc(aw_ctx* ctx, uint64_t bytes, void** out);
int aw_free(aw_ctx* ctx, void* p);
int aw_copy(aw_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* d[0..n) = value on the context's stream (asynchronous; a fill kernel, no
 * memset node, so that it can sit in a captured graph). */
int aw_fill_f32(aw_ctx* ctx, float* d, uint64_t n, float value);
/* Device memory the library holds in this process (grids, solvers, tapes,
 * work buffers and aw_malloc buffers, over all contexts) and its high-water
 * mark since load or the last reset (no reference counterpart).  reset_peak
 * sets the mark to the current figure after reading it. */
int aw_memory_stats(uint64_t* current_bytes, uint64_t* peak_bytes, int reset_peak);
/* Cap on that memory (no reference counterpart; 0 = none, the default): an
 * allocation that would take the library past it fails with AW_ERR_OOM, as
 * a failed hipMalloc does, and the object it was for stays destroyable.  For
 * a caller that shares the GPU under a budget. */
int aw_set_memory_limit(uint64_t bytes);
/* Kernel timing with HIP events on the context's stream (no reference
 * counterpart).  on = 1: an event pair around every launch; on = 2: one pair
 * around each run of consecutive launches of one kernel group (no events
 * between back-to-back launches, so short kernels are not stretched by them).
 * enable clears earlier records; report synchronises and writes JSON
 * {"kernel": {"launches": n, "ms": total, "bytes": algorithmic bytes}}. */
int aw_profile_enable(aw_ctx* ctx, int on);
int aw_profile_report(aw_ctx* ctx, char* buf, uint64_t len);

int aw_grid_create(aw_ctx* ctx, const aw_grid_desc* desc, aw_grid** out);
int aw_grid_destroy(aw_grid* grid);
/* Bins d_pos (AoS float3, n) to blocks, marks levels (distance bands + 2:1
 * grading), scans offsets and (re)allocates the channels.  Synchronises.
 * *changed (may be NULL) is set to 1 when the level map differs from the
 * previous one. */
int aw_grid_build_topology(aw_grid* grid, const float* d_pos, int64_t n, int* changed);
/* The reference's level rule instead of the distance bands: every block
 * starts at the coarsest level; each particle lowers the blocks within
 * `support` blocks of its own to plev = floor(max(0, log2(h / radius))),
 * clamped to num_levels - 1, with support = ceil(radius / res) at that level
 * (res = 8 >> plev; one block for every radius up to res world units);
 * then 2:1 grading to its fixed point.  d_radius is float[n] in d_pos's
 * order; every radius must be positive and finite (else AW_ERR_INVALID and
 * the topology is left as it was).  Otherwise as aw_grid_build_topology.
 * aw_step with rebuild_topology = 1 rebuilds with the distance bands: keep
 * rebuild_topology = 0 to step on this topology, as the reference does,
 * whose topology is fixed at start-up, or set rebuild_topology = 2 with
 * aw_step_params.d_radius to re-mark it from the radii every step. */
int aw_grid_build_topology_radius(aw_grid* grid, const float* d_pos, const float* d_radius, int64_t n,
                                  int* changed);
/* Marks the grid's face coefficients (aw_layout.d_beta) as written by the
 * caller: the next solver setup coarsens them through the multigrid levels
 * (uniform-coefficient fast paths off) until P2G rewrites them.  No reference
 * counterpart beyond runGalerkin (MSBG.cuh:661-676). */
int aw_grid_beta_changed(aw_grid* grid);
/* Explicit native levels (host array, num_blocks) instead of the band rule. */
int aw_grid_set_levels(aw_grid* grid, const uint8_t* h_levels);
int aw_grid_layout(aw_grid* grid, aw_layout* out);
/* Counting sort of particles by block into the caller's buffers (positions,
 * velocities and optional ids permuted together).  d_scratch_* are library
 * buffers; the sorted order is written back to d_pos/d_vel/d_ids. */
int aw_bin_particles(aw_grid* grid, float* d_pos, float* d_vel, int32_t* d_ids, int64_t n);

/* P2G of block-sorted particles: virtual-fine deposit, Pi^T, fold, normalise
 * (u_old), body force (u_star), liquid mask.  Particles must be sorted by
 * aw_bin_particles for the current topology. */
int aw_p2g(aw_grid* grid, const float* d_pos, const float* d_vel, int64_t n, float dt, const float gravity[3]);
int aw_divergence(aw_grid* grid);                 /* rhs = -D u_star on liquid cells */
int aw_project_velocity(aw_grid* grid);           /* u_new = u_star - beta grad p; fine lattice */
int aw_g2p_advect(aw_grid* grid, float* d_pos, float* d_vel, int64_t n, float alpha, float dt);
/* P2G of particles in any order (P2GTransfer::run, include/P2G.cuh:173-196):
 * bins and sorts them into the library's own buffers, so the caller's arrays
 * keep their order and contents, then deposits as aw_p2g.  Needs a topology
 * (aw_grid_build_topology); asynchronous. */
int aw_p2g_particles(aw_grid* grid, const float* d_pos, const float* d_vel, int64_t n, float dt,
                     const float gravity[3]);
/* The two halves of aw_g2p_advect on particles in any order, in place:
 * aw_g2p the FLIP/PIC velocity update at each particle's position
 * (G2PManager::performG2P, include/G2P.cuh:192-214; positions untouched),
 * aw_advect the RK3 advection through u_new (G2PManager::advect,
 * include/G2P.cuh:216-227; velocities untouched).  Both read the fine lattice
 * aw_project_velocity leaves; the advection does not read the particle
 * velocities, so either order gives aw_g2p_advect's result.  Asynchronous. */
int aw_g2p(aw_grid* grid, const float* d_pos, float* d_vel, int64_t n, float alpha);
int aw_advect(aw_grid* grid, float* d_pos, int64_t n, float dt);

int aw_solver_create(aw_grid* grid, const aw_solver_params* params, aw_solver** out);
int aw_solver_destroy(aw_solver* s);
/* Builds the multigrid hierarchy for the grid's current topology and liquid
 * mask (coarse masks, offsets, neighbour tables, work lists). */
int aw_solver_setup(aw_solver* s);
int aw_solver_level_info(aw_solver* s, int level, aw_level_info* out);
int aw_solver_smooth(aw_solver* s, int level, int iterations, int reverse);
int aw_solver_residual(aw_solver* s, int level);  /* r = b - A x on the level */
int aw_solver_apply(aw_solver* s, int level, const float* d_x, float* d_y);
int aw_solver_restrict(aw_solver* s, int level);  /* b[level+1] = R r[level] */
int aw_solver_prolong(aw_solver* s, int level);   /* x[level] += scale * P x[level+1] */
int aw_solver_vcycle(aw_solver* s, int level);    /* one V-cycle rooted at level on its x, b */
/* Solves A p = rhs on level 0 into the grid's pressure (initial guess: its
 * current content).  Synchronises. */
int aw_solver_solve(aw_solver* s, double rel_tol, int max_iters, aw_solve_stats* stats);
/* Replays each PCG iteration (V-cycle preconditioner and vector updates
 * between two scalar readbacks) from a captured HIP graph when on != 0; the
 * default (0) launches its kernels one by one, which measures faster on
 * ROCm 7 for this launch sequence.  The graph is captured on first use after
 * each hierarchy build.  No reference counterpart (the reference launches
 * every kernel from the host, VCycle.cuh:251-304). */
int aw_solver_set_graph(aw_solver* s, int on);

/* One FLIP step on caller-owned AoS float3 particles (positions, velocities,
 * optional int32 ids permuted with them).  Synchronises. */
int aw_step(aw_grid* grid, aw_solver* s, float* d_pos, float* d_vel, int32_t* d_ids, int64_t n,
            const aw_step_params* params, aw_step_stats* stats);

/* ---- Adjoint (reverse mode) of the step ----
 * Derivative of aw_step with the topology, the liquid mask and the mass
 * thresholds held at their forward values (they are piecewise constant in
 * the particle state).  The stage entry points run on the state the last
 * forward step left in the grid (which must have been binned with particle
 * ids); particle adjoints d_xbar, d_vbar are float[3 n] indexed by particle
 * id, not by sorted slot. */
typedef struct {
  float* d_ubar_new[3];   /* dL/du_new per face slot */
  float* d_ubar_old[3];   /* dL/du_old through G2P's FLIP term */
  float* d_ubar_star[3];  /* dL/du_star (projection and divergence) */
  float* d_Mbar[3];       /* dL/d(momentum accumulator) per slot */
  float* d_mbar[3];       /* dL/d(mass accumulator) per slot */
  float* d_pbar;          /* dL/dp on liquid cells */
  float* d_rhsbar;        /* A^{-1} pbar */
  float* d_fine_bar;      /* fine lattice [block][field][axis][512]: field 0 Fnew-bar + Fdel-bar, field 1 Fdel-bar */
  float* d_fine_mbar;     /* fine lattice: field 0 Pi Mbar, field 1 Pi mbar; current for the blocks
                             within one block of a block holding particles */
  const int32_t* d_sorted_ids; /* particle id of each sorted slot of the last binning */
  int64_t num_sorted;
} aw_adjoint_layout;

int aw_grid_adjoint_layout(aw_grid* grid, aw_adjoint_layout* out);

/* G2P + RK3 adjoint: d_xbar, d_vbar hold dL/dx', dL/dv' on entry and the
 * direct partials dL/dx, dL/dv on return; fills ubar_new and ubar_old. */
int aw_g2p_advect_adjoint(aw_grid* grid, float alpha, float dt, float* d_xbar, float* d_vbar);
/* Projection adjoint: pbar from ubar_new, the adjoint Poisson solve
 * A rhsbar = pbar with the forward operator and hierarchy (A is symmetric),
 * then ubar_star.  Overwrites the grid's pressure and rhs.  Synchronises. */
int aw_project_adjoint(aw_solver* s, double rel_tol, int max_iters, aw_solve_stats* stats);
/* P2G adjoint: fold adjoint from ubar_old + ubar_star, then adds the deposit
 * partials to d_xbar, d_vbar. */
int aw_p2g_adjoint(aw_grid* grid, float* d_xbar, float* d_vbar);
/* The three stages: dL/d(x', v') -> dL/d(x, v) of the last aw_step.  Synchronises. */
int aw_step_adjoint(aw_grid* grid, aw_solver* s, const aw_step_params* params, float* d_xbar, float* d_vbar,
                    aw_solve_stats* stats);

/* Checkpointed rollout.  The forward pass keeps (x, v, ids) every
 * checkpoint_every steps; the backward pass re-runs each segment from its
 * checkpoint, keeps the per-step state the step adjoint needs, and walks the
 * segment backwards.  keep_masks = 1 also records every step's native levels
 * and liquid mask on the host (aw_tape_masks). */
typedef struct aw_tape aw_tape;

typedef struct {
  int steps;
  double forward_ms;        /* wall time of aw_rollout_forward */
  double recompute_ms;      /* segment re-runs inside aw_rollout_backward */
  double backward_ms;       /* step adjoints inside aw_rollout_backward */
  int forward_iterations;   /* solver iterations summed over the forward steps */
  int adjoint_iterations;   /* adjoint solver iterations summed over the steps */
  double forward_rel_residual;  /* worst forward solve */
  double adjoint_rel_residual;  /* worst adjoint solve */
  /* Replay check of aw_rollout_backward's recomputation against the forward
   * pass (the gradient follows the recomputed trajectory):
   * replay_mismatched_steps: recomputed steps whose native levels or liquid
   * mask differ from the forward step's; replay_max_state_diff: largest
   * |position| or |velocity| difference, particle by particle id, between a
   * recomputed segment's end and the checkpoint the forward pass stored
   * there.  0 and 0.0 when the replay is bit-identical; the particle sort and
   * the P2G deposit add floats in arrival order, so round-off can remain. */
  int replay_mismatched_steps;
  double replay_max_state_diff;
} aw_rollout_stats;

int aw_tape_create(aw_grid* grid, aw_solver* s, int64_t n, int num_steps, int checkpoint_every, int keep_masks,
                   aw_tape** out);
int aw_tape_destroy(aw_tape* tape);
/* num_steps steps of aw_step on the particles (ids required).  Synchronises. */
int aw_rollout_forward(aw_tape* tape, float* d_pos, float* d_vel, int32_t* d_ids, const aw_step_params* params,
                       aw_rollout_stats* stats);
/* d_xbar, d_vbar: dL/d(x_K, v_K) on entry, dL/d(x_0, v_0) on return, indexed
 * by particle id.  Leaves the grid on the state of step 0.  Synchronises. */
int aw_rollout_backward(aw_tape* tape, float* d_xbar, float* d_vbar, aw_rollout_stats* stats);
/* Native levels (num_blocks bytes) and liquid mask (*ncells bytes) that
 * forward step `step` used; h_liquid may be NULL to query *ncells. */
int aw_tape_masks(aw_tape* tape, int step, uint8_t* h_levels, uint8_t* h_liquid, uint32_t liquid_cap,
                  uint32_t* ncells);

/* Remap of a cell field across a topology change of this grid (no reference
 * counterpart; the reference builds its topology once and starts every solve
 * from p = 0, src/main.cu:300): d_src holds a field of the layout whose native
 * levels were d_level_prev (device uint8[num_blocks], e.g. copied aside before
 * aw_grid_build_topology), d_dst (float[total_cells]) receives it on the
 * current layout, each cell taking the value of the previous cell that
 * contains its centre (a centre on a face of the previous cells takes the
 * cell above).  The warm start of aw_step remaps the pressure this way.
 * Asynchronous; d_src and d_dst must not alias. */
int aw_remap_cells(aw_grid* grid, const uint8_t* d_level_prev, const float* d_src, float* d_dst);
/* aw_remap_cells without any device allocation (no reference counterpart):
 * for callers that cannot allocate, such as code under HIP graph capture or
 * tight serving loops.  Its scratch (the previous layout's cell offsets and
 * their scan's chunk sums) is a fixed array in device memory, one per GPU,
 * held until the process ends (about 1 MB); grids of more than 262,144
 * blocks (2^18, the 512^3 grid's count) are rejected with AW_ERR_INVALID.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams or threads) race.  Same result as aw_remap_cells. */
int aw_remap_cells_noalloc(aw_grid* grid, const uint8_t* d_level_prev, const float* d_src, float* d_dst);

/* ---- Uniform resample and VTK export (include/vtx_visualization.cuh) ----
 * aw_resample_cells: nearest-cell resample of a cell field (float[total_cells]
 * of the current topology, e.g. aw_layout.d_pressure) onto an out_dims[0] x
 * out_dims[1] x out_dims[2] grid of voxel centres over the domain, x fastest,
 * into d_out (sampleMSBG, vtx_visualization.cuh:8-62).  A centre on a cell
 * face samples the cell above it (exact integer arithmetic).  Each size
 * 1..65535 with (2 size - 1) * 8 * dims below 2^32.  Asynchronous.
 * The writers are host-only and write legacy VTK with BINARY (big-endian)
 * sections: aw_write_vtk_structured a STRUCTURED_POINTS scalar field (the
 * reference's exportUniformPressure_VTK, vtx_visualization.cuh:64-126, wrote
 * ASCII), aw_write_vtk_blocks one hexahedron per block with its level as
 * refineLevel (exportMultigrid_VTK, :128-190), aw_write_vtk_particles
 * positions as POLYDATA with one vertex cell each and optional radius and
 * velocity (exportParticles_VTK, :192-216; h_radius, h_vel may be NULL).
 * aw_export_cells_vtk resamples and writes (ORIGIN at the first voxel
 * centre, where the reference put 0); aw_export_blocks_vtk writes the grid's
 * blocks.  Synchronous; AW_ERR_INVALID if a file cannot be written. */
int aw_resample_cells(aw_grid* grid, const float* d_field, const int out_dims[3], float* d_out);
int aw_write_vtk_structured(const char* path, const char* name, const float* h_data, const int dims[3],
                            const float origin[3], const float spacing[3]);
int aw_write_vtk_blocks(const char* path, const int dims[3], float block_size, const uint8_t* h_levels);
int aw_write_vtk_particles(const char* path, const float* h_pos, const float* h_radius, const float* h_vel,
                           int64_t n);
int aw_export_cells_vtk(aw_grid* grid, const float* d_field, const int out_dims[3], const char* name,
                        const char* path);
int aw_export_blocks_vtk(aw_grid* grid, const char* path);

#ifdef __cplusplus
}
#endif

#endif /* ANITOWAVE_AW_H */
