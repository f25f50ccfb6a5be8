// Internal declarations shared by the HIP kernels (aw_kernels.hip) and the
// host runtime (aw_runtime.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace aw {

// Events for the kernel launches that follow (the per-launch profiler):
// start recorded by the next launch's dispatch, stop by every launch's until
// taken back.
struct LaunchEvents {
  hipEvent_t start, stop;
};
void set_launch_events(hipEvent_t start, hipEvent_t stop);
LaunchEvents take_launch_events();

constexpr int BASE = 8;        // cells per block edge at native level 0
constexpr int LOGB = 3;
constexpr int PAD = 900;       // padded fine nodes per axis per block (9 x 10 x 10)
constexpr int SLAB = 2 * 3 * PAD;  // floats per block in the P2G slab / fine lattice
constexpr int FINE = 512;      // own fine nodes per axis per block
constexpr int ADJ = 6 * FINE;  // floats per block of a two-field fine-lattice adjoint buffer

// One multigrid level as the kernels see it.  Level 0 aliases the grid.
struct LevelDev {
  int D[3];
  int nb;
  int bs;                      // block edge in fine cells (8, or 8 << k after k aggregations)
  const uint8_t* e;            // log2 of the cell edge in fine cells, per block
  const uint32_t* coff;        // cell offsets, nb+1
  const uint32_t* foff[3];     // face-slot offsets per axis, nb+1
  const int32_t* nbr;          // 6 face neighbours per block (-x,+x,-y,+y,-z,+z), -1 = tank wall
  const uint8_t* liquid;       // per cell
  const float* beta[3];        // per face slot
  const uint32_t* bd = nullptr;  // block descriptors in work-list order (BD_WORDS each), or null
};

// Block descriptor of one work-list entry, so that a stencil kernel reaches
// every offset it needs with one dependent load:
// [0] block, [1] cell offset, [2] face-slot offset, [3..8] neighbour blocks
// (-1 = tank wall), [9..14] neighbour cell offsets, [15..20] neighbour
// face-slot offsets, [21] e | ne0 << 8 | ne1 << 16 | ne2 << 24,
// [22] ne3 | ne4 << 8 | ne5 << 16 (e = log2 cell edge), [23] unused.
constexpr int BD_WORDS = 24;

// The native grid as the transfer kernels see it.
struct GridDev {
  int D[3];
  int N[3];                    // fine cells per axis
  int nb;
  float h, inv_h;
  const uint8_t* level;
  const uint32_t* coff;
  const uint32_t* foff[3];
  const int32_t* nbr;
  const uint8_t* halo;         // per block: the padded fine lattice's halo is filled
  const uint32_t* count;       // per block: binned particles (their neighbourhoods have halos)
  const float* un[3];          // u_new and u_old faces (the padded lattice's sources)
  const float* uo[3];
};

// Work-group of a stencil launch: `count` blocks of resolution 1 << lr
// starting at work-list entry `start`.  A group holds at most 512 cells.
struct WG {
  int start, count, lr, pad;
};

// Work-groups of one launch, decoded arithmetically from blockIdx (no table
// load on the kernels' critical path).  The work list holds runs of blocks of
// one resolution (per colour, resolutions from coarse cells to fine); run k
// covers work-list entries [s0[k], s0[k] + cnt[k]) packed 512 >> 3 lr[k]
// blocks per group, and its groups start at launch index g0[k].  Unused runs
// have g0 = INT_MAX.
constexpr int WGMAP_RUNS = 8;
struct WGMap {
  int n;                     // work-groups in the launch
  int g0[WGMAP_RUNS];
  int s0[WGMAP_RUNS];
  int cnt[WGMAP_RUNS];
  int lr[WGMAP_RUNS];
};

// Restriction fused into the residual pass (coarse level with the same block
// dims): b_c = sum of children, x_c = 0.
struct Restr {
  LevelDev C;
  float* bc;
  float* xc;
  int on;
};

// Trailing multigrid levels run by one work-group with x, b, r in LDS.
constexpr int TAIL_T = 1024;
constexpr int TAIL_MAX_LEVELS = 8;
constexpr size_t TAIL_LDS_BYTES = 64 * 1024;
// one work-group evaluates every cell of a tail level per phase: beyond about
// a thousand cells of a mixed-resolution level separate launches are faster;
// a level of one resolution (arithmetic stencil, no table lookups) stays
// cheaper in the tail up to a few thousand cells
constexpr uint32_t TAIL_CELLS_MAX = 1024;
constexpr uint32_t TAIL_UNIFORM_CELLS_MAX = 4096;

struct TailLevel {
  LevelDev L;
  const uint32_t* cblk;  // block of each cell
  int ncells;
  int agg;               // built from the previous tail level by 2x2x2 aggregation
  int off;               // LDS offset (floats) of x; b at off + ncells, r at off + 2 ncells
  int loff;              // byte offset of the liquid mask after the float arrays
  int ulr;               // log2 block resolution when every block has it (arithmetic stencil), else -1
  int ue;                // log2 cell edge of such a uniform level
  int sh[3];             // log2 dims when the level is a lattice of 1^3 blocks with power-of-two dims, else -1
};

struct TailArgs {
  TailLevel lv[TAIL_MAX_LEVELS];
  int n;
  int fl;                // floats before the liquid masks
  float omega, scale;
  int pre, post, coarse_sweeps, local_sweeps;
  float* x0;
  const float* b0;
  const double* cinv;    // FP64 inverse of the coarsest level, or null (sweeps)
};

constexpr int DIRECT_MAX = 64;  // coarsest levels up to this many cells are solved exactly

// PCG scalars in device memory: SC_SLOTS accumulators each (a reduction's
// work-groups spread their atomics over the slots; readers add them).
enum Scalar { SC_RZ = 0, SC_RZOLD = 1, SC_PQ = 2, SC_RR = 3, SC_BB = 4, SC_N = 8 };
constexpr int SC_SLOTS = 32;

// ---- launchers (aw_kernels.hip) ----
void k_fill_f32(hipStream_t s, float* p, size_t n, float v);
void k_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);
void k_fill_f64(hipStream_t s, double* p, size_t n, double v);

// topology
void k_bin(hipStream_t s, const GridDev& g, const float* pos, int64_t n, int32_t* key, uint32_t* count);
void k_levels(hipStream_t s, const int* D, int L, const int* band, const uint32_t* count, uint8_t* surf,
              uint8_t* level, uint8_t* tmp);
// levels from particle radii (rad[ids[i]] for particle slot i, rad[i] if ids
// is null; rmap: 2 num_blocks + 1 words, the direct and per-block marks and
// the invalid-radius flag), graded; result in level, tmp is scratch
void k_radius_levels(hipStream_t s, const GridDev& g, int logh, int L, const float* pos, const float* rad,
                     const int32_t* ids, int64_t n, uint32_t* rmap, uint8_t* level, uint8_t* tmp);
// nearest-cell resample of a cell field onto an o[0] x o[1] x o[2] voxel grid
// (x fastest); o[1], o[2] <= 65535 and (2 o - 1) N < 2^32 per axis
void k_resample(hipStream_t s, const GridDev& g, const float* field, const int* o, float* out);
// exclusive scans, out[n] = total; `scratch` holds scan_scratch_words(n) words
size_t scan_scratch_words(int n);
void k_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, int n, uint32_t* scratch);
void k_cell_offsets(hipStream_t s, const uint8_t* level, int nb, uint32_t* coff, uint32_t* scratch);
void k_layout_offsets(hipStream_t s, const uint8_t* level, int nb, uint32_t* coff, uint32_t* fx, uint32_t* fy,
                      uint32_t* fz, uint32_t* scratch);
// dens (needs ids): per-particle density by particle id into sdens (sorted order), or null
// Largest-|value| bounds (the P2G deposit's fixed-point scale) are VB_SLOTS
// words of float bits each, the largest slot being the bound; NaN bits mark a
// value that is not finite.
constexpr int VB_SLOTS = 64;
// the binning sort's copy; also the bounds of |vel| (vbound[0 .. VB_SLOTS))
// and |dens| (vbound[VB_SLOTS .. 2 VB_SLOTS), when dens)
void k_scatter(hipStream_t s, const float* pos, const float* vel, const int32_t* ids, const int32_t* key,
               uint32_t* cursor, int64_t n, float* spos, float* svel, int32_t* sids, const float* dens,
               float* sdens, uint32_t* vbound);
// the bound of |v[i]|, i < n, into out[0 .. VB_SLOTS)
void k_vbound(hipStream_t s, const float* v, int64_t n, uint32_t* out);
void k_init_beta(hipStream_t s, const LevelDev& L, float* bx, float* by, float* bz);

// transfers
// sdens (sorted): deposit the particles' density instead of their velocity (the
// slab's first field then holds the density-weighted weights)
// vbound: the bounds of the deposited particles' |velocity| (and, with
// sdens, |density|), which set the fixed-point scale
void k_p2g_deposit(hipStream_t s, const GridDev& g, const float* spos, const float* svel, const uint32_t* pcount,
                   const uint32_t* poff, float* slab, uint8_t* liquid, const uint32_t* vbound,
                   const float* sdens = nullptr);
// beta = m / D folded (two-phase coefficients), 0 on walls
void k_beta_fold(hipStream_t s, const GridDev& g, float* const* macc, float* const* Dacc, float* const* beta);
void k_p2g_gather(hipStream_t s, const GridDev& g, const uint32_t* pcount, const float* slab, float* const* Macc,
                  float* const* macc);
void k_fold(hipStream_t s, const GridDev& g, float* const* Macc, float* const* macc, float* const* uold,
            float* const* ustar, float* const* mass, float dt, const float* grav);
void k_divergence(hipStream_t s, const GridDev& g, const uint8_t* liquid, float* const* ustar, float* rhs);
void k_project(hipStream_t s, const GridDev& g, const float* p, float* const* ustar, float* const* beta,
               float* const* unew);
// padded fine lattice of the blocks within one block of a particle, every
// node from its owner's faces (count: particles per block; halo: per-block
// flag written)
void k_to_fine(hipStream_t s, const GridDev& g, float* fine, const uint32_t* count, uint8_t* halo);
// p (current layout) from p_prev of the previous layout (its per-block levels and cell offsets)
void k_pressure_remap(hipStream_t s, const GridDev& g, const uint8_t* level_prev, const uint32_t* coff_prev,
                      const float* p_prev, float* p);
// list: n + 1 words of scratch for the particles whose stage points leave the
// blocks with fine-lattice halos (taken in a second pass).  mode 0: FLIP blend
// and RK3 advection; 1: the blend only (vel_out); 2: the advection only (pos_out)
void k_g2p_advect(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const float* svel,
                  int64_t n, float alpha, float dt, float* pos_out, float* vel_out, uint32_t* list, int mode = 0);
// the step's G2P + advection of block-sorted particles (block B's particles
// at poff[B] .. poff[B] + pcount[B]), each block's lattice read into LDS once
void k_g2p_advect_sorted(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const float* svel,
                         const uint32_t* pcount, const uint32_t* poff, int64_t n, float alpha, float dt,
                         float* pos_out, float* vel_out, uint32_t* list);

// adjoint (fine-lattice adjoint buffers: [block][field][axis][512], stride ADJ)
// G2P adjoint: stage pass over the n sorted particles (partials, and one
// scatter record of adj_record_bytes() per particle), then the scatter pass,
// one work-group per block over its particles [poff[b], poff[b] + pcount[b])
size_t adj_record_bytes();
// pcount, poff: the block-sorted particles' per-block ranges (one work-group
// per block, the forward lattice in LDS), or null (one thread per particle)
void k_g2p_adjoint_stages(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const int32_t* sids,
                          int64_t n, float alpha, float dt, float* xbar, float* vbar, void* rec, uint32_t* list,
                          const uint32_t* pcount = nullptr, const uint32_t* poff = nullptr);
// touched (nb bytes): set for every block whose fine-lattice adjoint the
// scatter writes; the transpose writes 0 for the others without reading
// them, and k_fbar_clear zeroes the touched blocks' adjoint and flags
void k_g2p_adjoint_scatter(hipStream_t s, const GridDev& g, const float* spos, const uint32_t* pcount,
                           const uint32_t* poff, const void* rec, float* fbar, uint8_t* touched);
void k_fbar_clear(hipStream_t s, int nb, float* fbar, uint8_t* touched);
void k_pi_transpose(hipStream_t s, const GridDev& g, const float* fbar, const uint8_t* touched, float* const* unb,
                    float* const* uob);
void k_project_adjoint(hipStream_t s, const GridDev& g, float* const* beta, float* const* unb, float* const* usb,
                       float* pbar);
void k_divergence_adjoint(hipStream_t s, const GridDev& g, const uint8_t* liquid, const float* rhsbar,
                          float* const* usb);
void k_fold_adjoint(hipStream_t s, const GridDev& g, float* const* uold, float* const* mass, float* const* uob,
                    float* const* usb, float* const* Mb, float* const* mb);
void k_pi_forward2(hipStream_t s, const GridDev& g, float* const* A, float* const* Bf, float* out);
// pcount, poff: the block-sorted particles' per-block ranges (one work-group
// per block, the lattice's padded copy in LDS), or null (one thread per particle)
void k_p2g_adjoint(hipStream_t s, const GridDev& g, const float* fm, const float* spos, const float* svel,
                   const int32_t* sids, int64_t n, float* xbar, float* vbar, const uint32_t* pcount = nullptr,
                   const uint32_t* poff = nullptr);
void k_masked_copy(hipStream_t s, float* dst, const float* src, const uint8_t* liquid, int n);
void k_ids_range(hipStream_t s, const int32_t* ids, int64_t n, int* out);  // out: {min, max}, preset by the caller

// solver
void k_smooth(hipStream_t s, const LevelDev& L, float* x, const float* b, const WGMap& wg,
              const uint32_t* blist, float omega, int local_sweeps, int reverse);
// the reference's adaptive smoothing of one block colour: blocks with act set,
// up to max_sweeps in-block sweeps with the early exit at `reduction`, then
// act_next set on blocks above theta * (*bmax) and their face neighbours;
// sweeps[block] = sweeps run (0 = inactive)
void k_smooth_adaptive(hipStream_t s, const LevelDev& L, float* x, const float* b, const WGMap& wg, float omega,
                       int reverse, const uint8_t* act, uint8_t* act_next, uint8_t* sweeps, const float* bmax,
                       float theta, int max_sweeps, float reduction);
void k_absmax(hipStream_t s, const float* v, const uint8_t* liquid, int n, float* out);  // *out = max |v| on liquid
void k_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v);
// live[B] = block B holds a liquid cell; out[j] = descriptor idx[j] of bd
void k_block_live(hipStream_t s, const LevelDev& L, uint8_t* live);
// order-independent 64-bit fingerprint of a liquid mask
void k_mask_hash(hipStream_t s, const uint8_t* m, uint32_t n, unsigned long long* out);
// out[0] = max |x - x_ref|, |v - v_ref| (float bits) over particles matched by
// id, out[1] = 1 if an id of rids lies outside [0, n); inv: n ints of scratch
void k_state_diff(hipStream_t s, const float* pos, const float* vel, const int32_t* ids, const float* rpos,
                  const float* rvel, const int32_t* rids, int64_t n, int32_t* inv, uint32_t* out);
void k_gather_bd(hipStream_t s, const uint32_t* bd, const uint32_t* idx, int n, uint32_t* out);
// the same sweep on a uniform coarse level with power-of-two block dims sh
// (log2), block resolution 1 << lr and cell edge 1 << e, colour col
void k_smooth_uniform(hipStream_t s, const int* sh, int lr, int e, const uint8_t* liquid, float* x, const float* b,
                      int col, float omega, int local_sweeps, int reverse);
void k_residual(hipStream_t s, const LevelDev& L, const float* x, const float* b, float* r, const WGMap& wg,
                const uint32_t* blist, const Restr& rs);
void k_apply_dot(hipStream_t s, const LevelDev& L, const float* x, float* y, const WGMap& wg,
                 const uint32_t* blist, double* acc);  // acc: SC_SLOTS accumulators, or null
void k_restrict(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, const float* rf, float* bc,
                float* xc, const WGMap& wgc, const uint32_t* blistc);
void k_prolong(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, float* xf, const float* xc,
               float scale, const WGMap& wgf, const uint32_t* blistf);
void k_coarsen_liquid(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, uint8_t* liquid_c,
                      const WGMap& wgc, const uint32_t* blistc);
// beta of level C from level F (area means of the finer side's faces)
void k_coarsen_beta(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, float* bx, float* by,
                    float* bz);
void k_tail(hipStream_t s, const TailArgs& args, size_t lds_bytes);
void k_coarse_factor(hipStream_t s, const LevelDev& L, int n, double* inv);

// PCG vector work (scalars live in device memory, see Scalar)
void k_dot(hipStream_t s, const float* a, const float* b, const uint8_t* liquid, int n, double* acc);
void k_scalar_shift(hipStream_t s, double* sc);
void k_pcg_p(hipStream_t s, const float* z, float* p, int n, const double* sc, int first);
void k_pcg_xr(hipStream_t s, float* x, float* r, const float* p, const float* q, const uint8_t* liquid, int n,
              double* sc);
void k_mask(hipStream_t s, float* x, const uint8_t* liquid, int n);
// *out = number of nonzero bytes of m[0..n) (cleared by the launcher)
void k_count_nonzero(hipStream_t s, const uint8_t* m, uint32_t n, uint32_t* out);
// FP64 iterative refinement
void k_residual64(hipStream_t s, const LevelDev& L, const double* x, const float* b, float* r32, const WGMap& wg,
                  const uint32_t* blist, double* rr);
void k_axpy64(hipStream_t s, double* x, const float* e, const uint8_t* liquid, int n);
void k_to64(hipStream_t s, double* x, const float* p, const uint8_t* liquid, int n);
void k_to32(hipStream_t s, float* p, const double* x, int n);
void k_add(hipStream_t s, float* x, const float* z, const uint8_t* liquid, int n);

}  // namespace aw
