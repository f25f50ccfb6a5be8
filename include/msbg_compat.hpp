// msbg_compat.hpp — the reference's host classes (MSBGManager, P2GTransfer,
// G2PManager, AdaptiveSmoother, MultigridTransfer, VCycleSolver) over the C
// ABI in aw.h, so that a caller shaped like the reference's src/main.cu keeps
// its structure.  Header-only C++17; needs only aw.h and libanitowave.so (no
// HIP headers: device buffers come from the caller's allocator or from
// aw_malloc through msbg::DeviceBuffer).
//
// Reference surface and what each method does here:
//   MSBGManager(dims, blockSize, domainMin)     include/MSBG.cuh:502-597   aw_ctx_create + aw_grid_create
//     .allocateChannels()                        MSBG.cuh:599-659           no-op: channels follow the topology
//     .buildTopology(d_pos, n)                   main.cu:156-275            aw_grid_build_topology
//     .runGalerkin()                             MSBG.cuh:661-676           no-op: VCycleSolver::init coarsens
//     .cleanup()                                 MSBG.cuh:544-597           aw_grid_destroy + aw_ctx_destroy
//     .grid()                                    MSBG.cuh:71-106            aw_grid_layout
//   P2GTransfer::run(mg, d_pos, d_vel, n)        include/P2G.cuh:173-196    aw_p2g_particles + aw_divergence
//   G2PManager::init / backupGridVelocity / performG2P / advect / cleanup
//                                                include/G2P.cuh:175-228    aw_project_velocity + aw_g2p, aw_advect
//   AdaptiveSmoother::solveLevel(solver, level, iterations)
//                                                include/Smoother.cuh:249-291  aw_solver_smooth
//   MultigridTransfer::restrictResiduals / prolongateAndCorrect(solver, level)
//                                                include/MultigridTransfer.cuh:107-149  aw_solver_restrict / _prolong
//   VCycleSolver(levels, nBlocks).init(mg).runCycle(mg, level)
//                                                include/VCycle.cuh:140-305  aw_solver_create/setup, aw_solver_vcycle
//
// Differences a caller sees:
//   * Errors throw msbg::Error (status and aw_last_error() detail); the
//     reference drops every CUDA status.  VCycleSolver::solve returns the
//     status instead, so that a caller can accept AW_ERR_NOT_CONVERGED.
//   * Particle buffers are AoS float3 (any struct of three floats, or float*
//     with 3 n values), as in the reference, and keep their order: P2G bins
//     them into the library's buffers.
//   * P2GTransfer::run also writes the pressure right-hand side -D u* (the
//     reference put particle mass into its DIVERGENCE channel, P2G.cuh:91-92)
//     and takes dt and gravity for the body force.  performG2P applies the
//     pressure projection u_new = u* - beta grad p before the FLIP update
//     (the reference has no projection).
//   * Calls are asynchronous on the manager's stream; synchronize() waits.
#ifndef ANITOWAVE_MSBG_COMPAT_HPP
#define ANITOWAVE_MSBG_COMPAT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>

#include "aw.h"

namespace msbg {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& where)
      : std::runtime_error(where + ": " + aw_last_error() + " (" + aw_status_str(status) + ")"), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* where) {
  if (status != AW_OK) throw Error(status, where);
}

template <class P>
inline const float* fptr(const P* p) {
  static_assert(sizeof(P) == 3 * sizeof(float) || sizeof(P) == sizeof(float), "AoS float3 or float");
  return reinterpret_cast<const float*>(p);
}
template <class P>
inline float* fptr(P* p) {
  static_assert(sizeof(P) == 3 * sizeof(float) || sizeof(P) == sizeof(float), "AoS float3 or float");
  return reinterpret_cast<float*>(p);
}

// Native level bands of the distance rule (in blocks); the defaults are the
// 256^3 pool's: level 0 within 1 block of the surface, level 1 within 4.
struct Bands {
  int num_levels = 3;
  int band[AW_MAX_LEVELS] = {1, 4, 16, 0};
};

class MSBGManager {
 public:
  // dims: blocks per axis (any type with x, y, z ints, e.g. int3); blockSize:
  // world edge of a block of 8^3 finest cells (h = blockSize / 8, a power of
  // two); domainMin must be the origin.
  template <class I3, class F3>
  MSBGManager(I3 dims, float blockSize, F3 domainMin, Bands bands = Bands(), int device = 0,
              void* stream = nullptr) {
    check(aw_ctx_create(device, stream, &ctx_), "aw_ctx_create");
    aw_grid_desc d{};
    d.dims[0] = dims.x;
    d.dims[1] = dims.y;
    d.dims[2] = dims.z;
    d.base_res = 8;
    d.num_levels = bands.num_levels;
    d.h = blockSize / 8.0f;
    d.domain_min[0] = domainMin.x;
    d.domain_min[1] = domainMin.y;
    d.domain_min[2] = domainMin.z;
    for (int l = 0; l < AW_MAX_LEVELS; ++l) d.band[l] = bands.band[l];
    const int st = aw_grid_create(ctx_, &d, &grid_);
    if (st != AW_OK) {
      Error e(st, "aw_grid_create");
      aw_ctx_destroy(ctx_);
      ctx_ = nullptr;
      throw e;
    }
    num_blocks_ = (uint32_t)dims.x * dims.y * dims.z;
  }
  MSBGManager(const MSBGManager&) = delete;
  MSBGManager& operator=(const MSBGManager&) = delete;
  ~MSBGManager() { cleanup(); }

  void allocateChannels() {}
  template <class P>
  bool buildTopology(const P* d_pos, uint32_t n) {
    int changed = 0;
    check(aw_grid_build_topology(grid_, fptr(d_pos), n, &changed), "aw_grid_build_topology");
    return changed != 0;
  }
  template <class P>
  bool buildTopologyRadius(const P* d_pos, const float* d_radius, uint32_t n) {
    int changed = 0;
    check(aw_grid_build_topology_radius(grid_, fptr(d_pos), d_radius, n, &changed), "aw_grid_build_topology_radius");
    return changed != 0;
  }
  void runGalerkin() {}
  void cleanup() {
    if (grid_) aw_grid_destroy(grid_);
    if (ctx_) aw_ctx_destroy(ctx_);
    grid_ = nullptr;
    ctx_ = nullptr;
  }
  aw_layout grid() const {
    aw_layout l;
    check(aw_grid_layout(grid_, &l), "aw_grid_layout");
    return l;
  }
  void synchronize() const { check(aw_ctx_synchronize(ctx_), "aw_ctx_synchronize"); }
  uint32_t numBlocks() const { return num_blocks_; }
  aw_grid* handle() const { return grid_; }
  aw_ctx* context() const { return ctx_; }

 private:
  aw_ctx* ctx_ = nullptr;
  aw_grid* grid_ = nullptr;
  uint32_t num_blocks_ = 0;
};

// Device memory from the library's allocator, for hosts without one.
class DeviceBuffer {
 public:
  DeviceBuffer(MSBGManager& mg, uint64_t bytes) : ctx_(mg.context()), bytes_(bytes) {
    check(aw_malloc(ctx_, bytes, &p_), "aw_malloc");
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() {
    if (p_) aw_free(ctx_, p_);
  }
  void upload(const void* h, uint64_t bytes) { check(aw_copy(ctx_, p_, h, bytes), "aw_copy"); }
  void download(void* h, uint64_t bytes) const { check(aw_copy(ctx_, h, p_, bytes), "aw_copy"); }
  template <class T>
  T* as() const {
    return static_cast<T*>(p_);
  }
  uint64_t size() const { return bytes_; }

 private:
  aw_ctx* ctx_;
  void* p_ = nullptr;
  uint64_t bytes_;
};

struct P2GTransfer {
  // Deposit (trilinear MAC P2G), normalise, body force, liquid mask and the
  // pressure right-hand side.  The grid must have a topology.
  template <class P>
  static void run(MSBGManager& mg, const P* d_pos, const P* d_vel, uint32_t n, float dt = 0.0f,
                  const float gravity[3] = nullptr) {
    const float zero[3] = {0.f, 0.f, 0.f};
    check(aw_p2g_particles(mg.handle(), fptr(d_pos), fptr(d_vel), n, dt, gravity ? gravity : zero),
          "aw_p2g_particles");
    check(aw_divergence(mg.handle()), "aw_divergence");
  }
};

class VCycleSolver {
 public:
  // levels and nBlocks are the reference's arguments; the hierarchy's depth
  // here follows the grid (blocks coarsen until every block is one cell,
  // then 2x2x2 aggregation), so they are not used.
  VCycleSolver(int levels, uint32_t nBlocks, aw_solver_params params = defaults()) : prm_(params) {
    (void)levels;
    (void)nBlocks;
  }
  VCycleSolver(const VCycleSolver&) = delete;
  VCycleSolver& operator=(const VCycleSolver&) = delete;
  ~VCycleSolver() {
    if (s_) aw_solver_destroy(s_);
  }
  static aw_solver_params defaults() {
    aw_solver_params p{};
    p.method = 0;
    p.pre_sweeps = 2;
    p.post_sweeps = 2;
    p.coarse_sweeps = 16;
    p.local_sweeps = 3;
    p.omega = 0.85714f;
    p.prolong_scale = 1.0f;
    p.precision = 0;
    p.adaptive = 0;
    p.theta = 1e-4f;
    p.max_block_sweeps = 10;
    p.block_reduction = 0.33f;
    return p;
  }
  // builds the hierarchy for the grid's current topology and liquid mask
  void init(MSBGManager& mg) {
    if (!s_) check(aw_solver_create(mg.handle(), &prm_, &s_), "aw_solver_create");
    check(aw_solver_setup(s_), "aw_solver_setup");
  }
  // one V-cycle rooted at `level` on its x, b (level 0: the grid's pressure and right-hand side)
  void runCycle(MSBGManager& mg, int level = 0) {
    if (!s_) init(mg);
    check(aw_solver_vcycle(s_, level), "aw_solver_vcycle");
  }
  // the pressure initial guess p = 0 (main.cu:300)
  static void zeroPressure(MSBGManager& mg) {
    const aw_layout l = mg.grid();
    check(aw_fill_f32(mg.context(), l.d_pressure, l.total_cells, 0.0f), "aw_fill_f32");
  }
  // the convergence loop the reference's main.cu replaces by a fixed 10
  // cycles; from p = 0 unless zero_guess is false (then from the grid's pressure)
  int solve(MSBGManager& mg, double rel_tol, int max_iters, aw_solve_stats* stats = nullptr, bool zero_guess = true) {
    if (!s_) init(mg);
    if (zero_guess) zeroPressure(mg);
    aw_solve_stats st{};
    const int rc = aw_solver_solve(s_, rel_tol, max_iters, &st);
    if (stats) *stats = st;
    if (rc != AW_OK && rc != AW_ERR_NOT_CONVERGED) throw Error(rc, "aw_solver_solve");
    return rc;
  }
  aw_solver* handle() const { return s_; }

 private:
  aw_solver_params prm_;
  aw_solver* s_ = nullptr;
};

struct AdaptiveSmoother {
  static void solveLevel(VCycleSolver& vc, int mgLevel, int iterations, bool reverse = false) {
    check(aw_solver_smooth(vc.handle(), mgLevel, iterations, reverse ? 1 : 0), "aw_solver_smooth");
  }
};

struct MultigridTransfer {
  static void restrictResiduals(VCycleSolver& vc, int mgLevel) {
    check(aw_solver_residual(vc.handle(), mgLevel), "aw_solver_residual");
    check(aw_solver_restrict(vc.handle(), mgLevel), "aw_solver_restrict");
  }
  static void prolongateAndCorrect(VCycleSolver& vc, int mgLevel) {
    check(aw_solver_prolong(vc.handle(), mgLevel), "aw_solver_prolong");
  }
};

class G2PManager {
 public:
  void init(MSBGManager& mg) { grid_ = mg.handle(); }
  // the pre-step grid velocity is P2G's u_old, kept by the grid: nothing to copy
  void backupGridVelocity(MSBGManager& mg) { grid_ = mg.handle(); }
  // u_new = u* - beta grad p from the grid's current pressure, then the FLIP /
  // PIC update of the particle velocities in place (positions untouched)
  template <class P>
  void performG2P(MSBGManager& mg, P* d_vel, const P* d_pos, uint32_t n, float alpha = 0.97f) {
    grid_ = mg.handle();
    check(aw_project_velocity(grid_), "aw_project_velocity");
    check(aw_g2p(grid_, fptr(d_pos), fptr(d_vel), n, alpha), "aw_g2p");
  }
  // RK3 advection through u_new in place (after performG2P)
  template <class P>
  void advect(MSBGManager& mg, P* d_pos, uint32_t n, float dt) {
    grid_ = mg.handle();
    check(aw_advect(grid_, fptr(d_pos), n, dt), "aw_advect");
  }
  void cleanup() { grid_ = nullptr; }

 private:
  aw_grid* grid_ = nullptr;
};

}  // namespace msbg

#endif  // ANITOWAVE_MSBG_COMPAT_HPP
