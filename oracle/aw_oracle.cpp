// TEST INFRASTRUCTURE ONLY — the FP64 CPU oracle of the MSBG FLIP step.
//
// This file is the checker: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load the library built from it.  The product (the HIP library
// under anitowave_adaptive_amd/) never links, loads or calls it.
//
// It is a plain FP64 C++/OpenMP restatement of the discrete step that the HIP
// path implements.  Where the reference fixes a rule it is followed and cited
// (paths are relative to the reference checkout); where the reference is
// incomplete or defective the corrected rule is stated in place.
//
// Independence from the product: the Poisson operator here is assembled
// geometrically — every cell probes the fine-resolution points just across each
// of its faces, finds the cell that contains them and accumulates the shared
// area — and is then applied as an explicit CSR matrix.  The HIP path instead
// evaluates a structured stencil from block-neighbour tables.  Multigrid
// parents are likewise found by point containment, not by the halving /
// aggregation case analysis the HIP transfer kernels use.
//
// Parity pins (tests/test_oracle_*.py): closed-form layout and Morton values,
// brute-force topology on tiny grids (distance bands; particle radii against
// exact rational arithmetic), A = A^T and A SPD, a dense solve of the
// single-level box problem, partition of unity / linear-field reproduction of
// the transfers, mass conservation of P2G, hydrostatic column, divergence of
// the projected field equal to the solver residual, RK3 exact translation;
// the adjoint: Pi / Pi^T transpose identity, whole-step and rollout
// directional derivatives against central finite differences of the FP64
// forward with frozen topology (tests/test_oracle_adjoint.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr int B = 8;       // cells per block edge at native level 0
constexpr int LOGB = 3;

inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Morton interleave, 10 bits per axis (MSBG.cuh:139-149).
inline uint32_t expand_bits(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Face slot index inside a block of resolution r: the normal axis carries r+1
// slots, the others r (MSBG.cuh:169-179).
inline int face_index(int a, int x, int y, int z, int r) {
  if (a == 0) return x + (r + 1) * (y + r * z);
  if (a == 1) return x + r * (y + (r + 1) * z);
  return x + r * (y + r * z);
}
inline int cell_index(int x, int y, int z, int r) { return x + r * (y + r * z); }  // MSBG.cuh:165-167
inline int faces_per_block(int r) { return (r + 1) * r * r; }

struct Level {              // one multigrid level
  int D[3] = {0, 0, 0};
  int nb = 0;
  int bs = B;               // block edge in fine cells
  std::vector<int> e;       // log2 of the cell edge (fine units) per block
  std::vector<uint32_t> coff;
  int ncells = 0;
  std::vector<int> cblock, clocal;
  std::vector<uint8_t> liquid;
  std::vector<int> rowptr, col;
  std::vector<double> val, diag;
  std::vector<int> parent;
  std::vector<double> x, b, r;
  std::vector<int> bcol[2];
  std::vector<uint8_t> act, sweeps;   // adaptive smoothing: flags left active, sweeps run per block
};

struct OGrid {
  int D[3];
  int N[3];
  int nb;
  int L;                    // native levels
  double h;
  int band[4];
  std::vector<uint8_t> level;
  std::vector<uint32_t> coff, foff[3];
  int64_t ncells = 0, nfaces[3] = {0, 0, 0};
  // native fields
  std::vector<double> Macc[3], macc[3], uold[3], ustar[3], unew[3], mt[3], beta[3];
  std::vector<double> Dacc[3];          // density-weighted deposit weight per slot (two-phase beta)
  std::vector<double> p, rhs;
  std::vector<uint8_t> liquid;
  std::vector<uint32_t> pcount;
  // virtual fine lattice: [block][axis][512]
  std::vector<double> Fm, FM, Fnew, Fdel;
  std::vector<double> FD;               // fine lattice: density-weighted deposit weight
  // adjoint fields (bar = derivative of the loss with respect to the field)
  std::vector<double> Fbar_nd, Fbar_d;   // fine lattice: Fnew-bar + Fdel-bar, Fdel-bar
  std::vector<double> FMbar, Fmbar;      // fine lattice: deposit accumulators
  std::vector<double> ubar_new[3], ubar_old[3], ubar_star[3], Mbar[3], mbar[3];
  std::vector<double> pbar, rhsbar;
  // multigrid
  std::vector<Level> mg;
  double omega = 0.85714f;
  int local_sweeps = 3, pre = 2, post = 2, coarse_sweeps = 16;
  double prolong_scale = 1.0;
  // adaptive smoothing (AdaptiveSmoother, Smoother.cuh:105-291)
  int adaptive = 0;
  double theta = 1e-4, block_reduction = 0.33;
  int max_block_sweeps = 10;
};

inline int block_res(const OGrid& g, int b) { return B >> g.level[b]; }
inline void block_coords(const int* D, int b, int* c) {
  c[0] = b % D[0];
  c[1] = (b / D[0]) % D[1];
  c[2] = b / (D[0] * D[1]);
}
inline int block_id(const int* D, int x, int y, int z) { return x + D[0] * (y + D[1] * z); }

// ---------------------------------------------------------------- binning
// Fine cell of a position: g = floor(x / h), clamped to the domain.  h is a
// power of two, so x * (1/h) is exact in FP32 and the result is bit-identical
// on every implementation (worldToBlockCoord, MSBG.cuh:151-159, rewritten on
// exact fine-cell integers).
inline void fine_cell(const OGrid& g, const float* x, int* gc) {
  const float inv_h = (float)(1.0 / g.h);
  for (int a = 0; a < 3; ++a) {
    float u = x[a] * inv_h;
    int v = (int)std::floor(u);
    gc[a] = iclamp(v, 0, g.N[a] - 1);
  }
}
// FP64 positions (the finite-difference rollouts): x / h is exact as well.
inline void fine_cell(const OGrid& g, const double* x, int* gc) {
  for (int a = 0; a < 3; ++a) gc[a] = iclamp((int)std::floor(x[a] / g.h), 0, g.N[a] - 1);
}

// ------------------------------------------------------- native topology
void set_topology(OGrid& g, const uint8_t* level) {
  g.level.assign(level, level + g.nb);
  g.coff.assign(g.nb + 1, 0);
  for (int a = 0; a < 3; ++a) g.foff[a].assign(g.nb + 1, 0);
  for (int b = 0; b < g.nb; ++b) {          // countBlockElements + exclusive_scan (MSBG.cuh:353-376, main.cu:270-273)
    int r = block_res(g, b);
    g.coff[b + 1] = g.coff[b] + r * r * r;
    for (int a = 0; a < 3; ++a) g.foff[a][b + 1] = g.foff[a][b] + faces_per_block(r);
  }
  g.ncells = g.coff[g.nb];
  for (int a = 0; a < 3; ++a) {
    g.nfaces[a] = g.foff[a][g.nb];
    for (auto* v : {&g.Macc[a], &g.macc[a], &g.uold[a], &g.ustar[a], &g.unew[a], &g.mt[a], &g.ubar_new[a],
                    &g.ubar_old[a], &g.ubar_star[a], &g.Mbar[a], &g.mbar[a]})
      v->assign(g.nfaces[a], 0.0);
    // beta = 1/rho on every interior face, 0 on the tank walls (Neumann); the
    // reference sets 1 everywhere and treats the outside as Dirichlet air
    // (main.cu:294-296, VCycle.cuh:43-44).
    g.beta[a].assign(g.nfaces[a], 1.0);
  }
  for (int b = 0; b < g.nb; ++b) {
    int c[3];
    block_coords(g.D, b, c);
    int r = block_res(g, b);
    for (int a = 0; a < 3; ++a) {
      int n[3] = {r, r, r};
      n[a] = r + 1;
      for (int z = 0; z < n[2]; ++z)
        for (int y = 0; y < n[1]; ++y)
          for (int x = 0; x < n[0]; ++x) {
            int sl[3] = {x, y, z};
            bool wall = (sl[a] == 0 && c[a] == 0) || (sl[a] == r && c[a] == g.D[a] - 1);
            if (wall) g.beta[a][g.foff[a][b] + face_index(a, x, y, z, r)] = 0.0;
          }
    }
  }
  g.p.assign(g.ncells, 0.0);
  g.rhs.assign(g.ncells, 0.0);
  g.pbar.assign(g.ncells, 0.0);
  g.rhsbar.assign(g.ncells, 0.0);
  g.liquid.assign(g.ncells, 0);
  g.pcount.assign(g.ncells, 0);
  g.mg.clear();
}

// 2:1 grading over the 26-neighbourhood, double-buffered, to the fixed point
// (L - 1 passes: the result is min over q of l_q + Chebyshev distance, and
// terms at distance >= L - 1 cannot go below any level) (enforceGrading,
// MSBG.cuh:316-351, ran 3 racy in-place passes).
void grade(const OGrid& g, std::vector<uint8_t>& lv) {
  std::vector<uint8_t> tmp(g.nb);
  for (int it = 0; it < g.L - 1; ++it) {
    for (int b = 0; b < g.nb; ++b) {
      int c[3];
      block_coords(g.D, b, c);
      int l = lv[b];
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
            if (x < 0 || y < 0 || z < 0 || x >= g.D[0] || y >= g.D[1] || z >= g.D[2]) continue;
            l = std::min(l, (int)lv[block_id(g.D, x, y, z)] + 1);
          }
      tmp[b] = (uint8_t)l;
    }
    lv.swap(tmp);
  }
}

// Level rule (replaces the particle-radius rule of markActiveBlocks,
// MSBG.cuh:271-314, by the distance bands the workload specifies):
//   occupied  = block holds >= 1 particle
//   surface   = occupied block with an unoccupied in-domain 26-neighbour
//   dist      = Chebyshev block distance to the nearest surface block
//   level     = first l with dist <= band[l], else L-1
// followed by 2:1 grading over the 26-neighbourhood to a fixed point with a
// double buffer (enforceGrading, MSBG.cuh:316-351, ran 3 racy in-place passes).
void levels_from_counts(const OGrid& g, const uint32_t* count, uint8_t* out) {
  std::vector<uint8_t> surf(g.nb, 0);
  for (int b = 0; b < g.nb; ++b) {
    if (!count[b]) continue;
    int c[3];
    block_coords(g.D, b, c);
    bool s = false;
    for (int dz = -1; dz <= 1 && !s; ++dz)
      for (int dy = -1; dy <= 1 && !s; ++dy)
        for (int dx = -1; dx <= 1 && !s; ++dx) {
          int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
          if (x < 0 || y < 0 || z < 0 || x >= g.D[0] || y >= g.D[1] || z >= g.D[2]) continue;
          if (!count[block_id(g.D, x, y, z)]) s = true;
        }
    surf[b] = s;
  }
  int bmax = 0;
  for (int l = 0; l < g.L - 1; ++l) bmax = std::max(bmax, g.band[l]);
  std::vector<uint8_t> lv(g.nb);
  for (int b = 0; b < g.nb; ++b) {
    int c[3];
    block_coords(g.D, b, c);
    int dist = 1 << 20;
    for (int z = std::max(0, c[2] - bmax); z <= std::min(g.D[2] - 1, c[2] + bmax); ++z)
      for (int y = std::max(0, c[1] - bmax); y <= std::min(g.D[1] - 1, c[1] + bmax); ++y)
        for (int x = std::max(0, c[0] - bmax); x <= std::min(g.D[0] - 1, c[0] + bmax); ++x)
          if (surf[block_id(g.D, x, y, z)]) {
            int d = std::max(std::abs(x - c[0]), std::max(std::abs(y - c[1]), std::abs(z - c[2])));
            dist = std::min(dist, d);
          }
    int l = g.L - 1;
    for (int k = 0; k < g.L - 1; ++k)
      if (dist <= g.band[k]) { l = k; break; }
    lv[b] = (uint8_t)l;
  }
  grade(g, lv);
  std::copy(lv.begin(), lv.end(), out);
}

// The reference's own level rule (markActiveBlocks, MSBG.cuh:271-314, after
// initRefinementMap to MAX_LEVELS - 1, main.cu:158-162): every block starts
// at the coarsest level; each particle lowers the blocks within `support`
// blocks of its bin to its level
//   plev    = floor(max(0, log2(h / radius))), clamped to [0, L-1]
//   support = ceil(radius * cellSize / blockSize) = ceil(radius / res(plev))
// then the same graded fixed point.  plev is the largest k <= L-1 with
// radius 2^k <= h, found with exact power-of-two scaling in FP64.  Returns 1
// (out untouched) if a radius is not positive and finite.
int levels_from_radii(const OGrid& g, const float* pos, const float* radius, int64_t n, uint8_t* out) {
  for (int64_t i = 0; i < n; ++i)
    if (!(radius[i] > 0.f) || std::isinf(radius[i])) return 1;
  std::vector<uint8_t> lv(g.nb, (uint8_t)(g.L - 1));
  const int smax = std::max(g.D[0], std::max(g.D[1], g.D[2]));
  for (int64_t i = 0; i < n; ++i) {
    const double r = radius[i];
    int plev = 0;
    while (plev < g.L - 1 && std::ldexp(r, plev + 1) <= g.h) ++plev;
    const int res = B >> plev;
    const int support = (int)std::min<double>(std::ceil(r / res), smax);
    int gc[3];
    fine_cell(g, pos + 3 * i, gc);
    const int c[3] = {gc[0] >> LOGB, gc[1] >> LOGB, gc[2] >> LOGB};
    for (int z = std::max(0, c[2] - support); z <= std::min(g.D[2] - 1, c[2] + support); ++z)
      for (int y = std::max(0, c[1] - support); y <= std::min(g.D[1] - 1, c[1] + support); ++y)
        for (int x = std::max(0, c[0] - support); x <= std::min(g.D[0] - 1, c[0] + support); ++x) {
          uint8_t& l = lv[block_id(g.D, x, y, z)];
          l = std::min<uint8_t>(l, (uint8_t)plev);
        }
  }
  grade(g, lv);
  std::copy(lv.begin(), lv.end(), out);
  return 0;
}

// ----------------------------------------------------- multigrid levels
// Level k of the block part: every block's cell edge is
//   e_b(k) = min(max(l_b, l_min + k), 3)
// i.e. the finest blocks coarsen first and a block never vanishes (the
// reference halves every block at every level and drops blocks whose
// resolution reaches 0, MSBG.cuh:689-692, MultigridTransfer.cuh:26).  Once all
// blocks are 1^3 the grid is a dense D^3 array and aggregation levels halve the
// dims while every dim stays >= 4.
inline int cell_of(const Level& L, const int* gc) {
  int Bc[3], l[3];
  for (int a = 0; a < 3; ++a) Bc[a] = gc[a] / L.bs;
  int b = block_id(L.D, Bc[0], Bc[1], Bc[2]);
  int e = L.e[b];
  int r = L.bs >> e;
  for (int a = 0; a < 3; ++a) l[a] = (gc[a] - Bc[a] * L.bs) >> e;
  return (int)L.coff[b] + cell_index(l[0], l[1], l[2], r);
}

inline void cell_geom(const Level& L, int c, int* g0, int& s) {
  int b = L.cblock[c];
  int e = L.e[b];
  int r = L.bs >> e;
  s = 1 << e;
  int Bc[3];
  block_coords(L.D, b, Bc);
  int li = L.clocal[c];
  int l[3] = {li % r, (li / r) % r, li / (r * r)};
  for (int a = 0; a < 3; ++a) g0[a] = Bc[a] * L.bs + l[a] * s;
}

void finish_level_cells(Level& L) {
  L.nb = L.D[0] * L.D[1] * L.D[2];
  L.coff.assign(L.nb + 1, 0);
  for (int b = 0; b < L.nb; ++b) {
    int r = L.bs >> L.e[b];
    L.coff[b + 1] = L.coff[b] + r * r * r;
  }
  L.ncells = L.coff[L.nb];
  L.cblock.resize(L.ncells);
  L.clocal.resize(L.ncells);
  for (int b = 0; b < L.nb; ++b)
    for (uint32_t c = L.coff[b]; c < L.coff[b + 1]; ++c) {
      L.cblock[c] = b;
      L.clocal[c] = c - L.coff[b];
    }
  L.bcol[0].clear();
  L.bcol[1].clear();
  for (int b = 0; b < L.nb; ++b) {
    int c[3];
    block_coords(L.D, b, c);
    L.bcol[(c[0] + c[1] + c[2]) & 1].push_back(b);
  }
  L.x.assign(L.ncells, 0.0);
  L.b.assign(L.ncells, 0.0);
  L.r.assign(L.ncells, 0.0);
  L.liquid.assign(L.ncells, 0);
}

// beta of the native face that holds the unit fine sub-face on the plane
// between fine cells P - 1 and P along axis a, at fine transverse coordinates
// q[t1], q[t2]: the slot of the finer of the two native cells there (a level
// change's coarse side is not an unknown; equal cells hold equal duplicates).
double beta_native(const OGrid& g, int a, int P, const int* q) {
  int lo[3] = {q[0], q[1], q[2]}, hi[3] = {q[0], q[1], q[2]};
  lo[a] = P - 1;
  hi[a] = P;
  const int bl = block_id(g.D, lo[0] >> LOGB, lo[1] >> LOGB, lo[2] >> LOGB);
  const int bu = block_id(g.D, hi[0] >> LOGB, hi[1] >> LOGB, hi[2] >> LOGB);
  const int X = g.level[bu] < g.level[bl] ? bu : bl;
  const int e = g.level[X], r = B >> e;
  int Xc[3];
  block_coords(g.D, X, Xc);
  int l[3];
  for (int d = 0; d < 3; ++d) l[d] = ((d == a ? P : q[d]) - B * Xc[d]) >> e;
  return g.beta[a][g.foff[a][X] + face_index(a, l[0], l[1], l[2], r)];
}

// Geometric assembly.  For a cell pair (c, c') sharing area A (fine units)
// the face coefficient is  w = beta A / d,  d = (s_c + s_c') / 2 the centre
// distance along the axis — h and 2h for equal cells, 1.5 h_f across a
// coarse–fine (T-junction) interface — where beta A is the sum of the native
// beta over the unit fine sub-faces of the shared area: on every level the
// coefficient is the area integral of the finest beta (the reference's
// galerkinCoarsen averages the four fine faces under each coarse face,
// MSBG.cuh:401-498).  A = sum over faces of w (e_c - e_c') (e_c - e_c')^T is
// symmetric by construction, unlike the reference's point-sampled ghosts
// (Smoother.cuh:70-91, VCycle.cuh:56-75).  Air cells are Dirichlet p = 0, tank
// walls have beta = 0.  Rows exist for liquid cells.
void assemble(const OGrid& g, Level& L) {
  L.rowptr.assign(L.ncells + 1, 0);
  L.diag.assign(L.ncells, 0.0);
  std::vector<std::vector<std::pair<int, double>>> rows(L.ncells);
#pragma omp parallel for schedule(dynamic, 256)
  for (int c = 0; c < L.ncells; ++c) {
    if (!L.liquid[c]) continue;
    int g0[3], s;
    cell_geom(L, c, g0, s);
    std::vector<std::pair<int, double>> hits;   // (neighbour, sum of beta over its fine sub-faces)
    double dsum = 0.0;
    for (int a = 0; a < 3; ++a) {
      int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
      for (int side = 0; side < 2; ++side) {
        int ga = side ? g0[a] + s : g0[a] - 1;
        if (ga < 0 || ga >= g.N[a]) continue;           // tank wall: beta = 0
        const int P = side ? g0[a] + s : g0[a];         // the plane of the face
        hits.clear();
        for (int j = 0; j < s; ++j)
          for (int k = 0; k < s; ++k) {
            int q[3];
            q[a] = ga;
            q[t1] = g0[t1] + j;
            q[t2] = g0[t2] + k;
            int cn = cell_of(L, q);
            const double bt = beta_native(g, a, P, q);
            bool found = false;
            for (auto& hcnt : hits)
              if (hcnt.first == cn) { hcnt.second += bt; found = true; break; }
            if (!found) hits.push_back({cn, bt});
          }
        for (auto& hcnt : hits) {
          int gn[3], sn;
          cell_geom(L, hcnt.first, gn, sn);
          double w = hcnt.second / (0.5 * (s + sn));
          dsum += w;
          if (L.liquid[hcnt.first]) rows[c].push_back({hcnt.first, -w});
        }
      }
    }
    L.diag[c] = dsum;
  }
  for (int c = 0; c < L.ncells; ++c) L.rowptr[c + 1] = L.rowptr[c] + (int)rows[c].size();
  L.col.resize(L.rowptr[L.ncells]);
  L.val.resize(L.rowptr[L.ncells]);
  for (int c = 0; c < L.ncells; ++c)
    for (size_t k = 0; k < rows[c].size(); ++k) {
      L.col[L.rowptr[c] + k] = rows[c][k].first;
      L.val[L.rowptr[c] + k] = rows[c][k].second;
    }
}

void mg_build(OGrid& g) {
  g.mg.clear();
  int lmin = 3;
  for (int b = 0; b < g.nb; ++b) lmin = std::min(lmin, (int)g.level[b]);
  for (int k = 0;; ++k) {
    Level L;
    for (int a = 0; a < 3; ++a) L.D[a] = g.D[a];
    L.bs = B;
    L.e.resize(g.nb);
    bool all1 = true;
    for (int b = 0; b < g.nb; ++b) {
      L.e[b] = std::min(std::max((int)g.level[b], lmin + k), LOGB);
      all1 = all1 && L.e[b] == LOGB;
    }
    finish_level_cells(L);
    g.mg.push_back(std::move(L));
    if (all1) break;
  }
  for (;;) {
    const Level& P = g.mg.back();
    int mn = std::min(P.D[0], std::min(P.D[1], P.D[2]));
    if (mn < 4 || (P.D[0] & 1) || (P.D[1] & 1) || (P.D[2] & 1)) break;
    Level L;
    for (int a = 0; a < 3; ++a) L.D[a] = P.D[a] / 2;
    L.bs = P.bs * 2;
    int eb = P.e[0] + 1;
    L.e.assign(L.D[0] * L.D[1] * L.D[2], eb);
    finish_level_cells(L);
    g.mg.push_back(std::move(L));
  }
  // parents by containment of the child's fine origin
  for (size_t k = 0; k + 1 < g.mg.size(); ++k) {
    Level& F = g.mg[k];
    Level& C = g.mg[k + 1];
    F.parent.resize(F.ncells);
    for (int c = 0; c < F.ncells; ++c) {
      int g0[3], s;
      cell_geom(F, c, g0, s);
      F.parent[c] = cell_of(C, g0);
    }
  }
}

// Liquid mask per level: level 0 is the native mask (cell holds >= 1
// particle); a coarse cell is liquid only if every child is liquid, so the
// Dirichlet (air) label dominates and every coarse level keeps its free
// surface.  Then assemble each level.
void mg_refresh(OGrid& g) {
  Level& L0 = g.mg[0];
  for (int c = 0; c < L0.ncells; ++c) L0.liquid[c] = g.liquid[c];
  for (size_t k = 0; k + 1 < g.mg.size(); ++k) {
    Level& F = g.mg[k];
    Level& C = g.mg[k + 1];
    std::fill(C.liquid.begin(), C.liquid.end(), 1);
    for (int c = 0; c < F.ncells; ++c)
      if (!F.liquid[c]) C.liquid[F.parent[c]] = 0;
  }
  for (auto& L : g.mg) assemble(g, L);
}

// --------------------------------------------------------------- solver
// Weighted Gauss–Seidel, block red-black over block parity (Smoother.cuh:123)
// times cell red-black over cell parity (Smoother.cuh:150), x += w r / diag
// (Smoother.cuh:160-175), `local_sweeps` in-block sweeps per block phase
// (MAX_BLOCK_ITER, Smoother.cuh:138, without the early exit).  The reverse
// order (block colour 1 then 0, cell colour 1 then 0) makes the post-smoother
// the adjoint of the pre-smoother, so the V-cycle is a symmetric
// preconditioner.
void smooth_adaptive(OGrid& g, Level& L, int iters, bool reverse);

// plain = true: the plain sweeps even with adaptive smoothing on (the
// coarsest level's symmetric sweeps)
void smooth(OGrid& g, Level& L, int iters, bool reverse, bool plain = false) {
  if (g.adaptive && !plain) {
    smooth_adaptive(g, L, iters, reverse);
    return;
  }
  for (int it = 0; it < iters; ++it)
    for (int ph = 0; ph < 2; ++ph) {
      int bc = reverse ? 1 - ph : ph;
      const std::vector<int>& blocks = L.bcol[bc];
#pragma omp parallel for schedule(dynamic, 16)
      for (size_t bi = 0; bi < blocks.size(); ++bi) {
        int b = blocks[bi];
        int r = L.bs >> L.e[b];
        for (int ls = 0; ls < g.local_sweeps; ++ls)
          for (int cp = 0; cp < 2; ++cp) {
            int cc = reverse ? 1 - cp : cp;
            for (uint32_t c = L.coff[b]; c < L.coff[b + 1]; ++c) {
              int li = c - L.coff[b];
              int lx = li % r, ly = (li / r) % r, lz = li / (r * r);
              if (((lx + ly + lz) & 1) != cc || !L.liquid[c]) continue;
              double ax = L.diag[c] * L.x[c];
              for (int k = L.rowptr[c]; k < L.rowptr[c + 1]; ++k) ax += L.val[k] * L.x[L.col[k]];
              double res = L.b[c] - ax;
              if (L.diag[c] > 0) L.x[c] += g.omega * res / L.diag[c];
            }
          }
      }
    }
}

// The reference's adaptive smoothing (smoothActiveBlocks + solveLevel,
// Smoother.cuh:105-291), its divergent early exit (A12) decided per block:
// every call starts with all blocks active; an active block of the phase's
// colour runs up to max_block_sweeps in-block sweeps (both cell colours, in
// the phase order) and stops after the first sweep whose largest |r| over
// its updates is below block_reduction times the first sweep's (when that is
// above 1e-9); if the last sweep's largest |r| exceeds theta max|b| (the
// reference's absolute 1e-4 made relative, A22) the block and its six face
// neighbours are active in the next iteration.
void smooth_adaptive(OGrid& g, Level& L, int iters, bool reverse) {
  std::vector<uint8_t> act(L.nb, 1), nxt(L.nb, 0);
  L.sweeps.assign(L.nb, 0);
  double bmax = 0.0;
  for (int c = 0; c < L.ncells; ++c)
    if (L.liquid[c]) bmax = std::max(bmax, std::fabs(L.b[c]));
  const double theta = g.theta * bmax;
  for (int it = 0; it < iters; ++it) {
    std::fill(nxt.begin(), nxt.end(), 0);
    for (int ph = 0; ph < 2; ++ph) {
      int bc = reverse ? 1 - ph : ph;
      const std::vector<int>& blocks = L.bcol[bc];
#pragma omp parallel for schedule(dynamic, 16)
      for (size_t bi = 0; bi < blocks.size(); ++bi) {
        int b = blocks[bi];
        if (!act[b]) {
          L.sweeps[b] = 0;
          continue;
        }
        int r = L.bs >> L.e[b];
        double first = 0.0, last = 0.0;
        int ran = 0;
        for (int ls = 0; ls < g.max_block_sweeps; ++ls) {
          double m = 0.0;
          for (int cp = 0; cp < 2; ++cp) {
            int cc = reverse ? 1 - cp : cp;
            for (uint32_t c = L.coff[b]; c < L.coff[b + 1]; ++c) {
              int li = c - L.coff[b];
              int lx = li % r, ly = (li / r) % r, lz = li / (r * r);
              if (((lx + ly + lz) & 1) != cc || !L.liquid[c] || !(L.diag[c] > 0)) continue;
              double ax = L.diag[c] * L.x[c];
              for (int k = L.rowptr[c]; k < L.rowptr[c + 1]; ++k) ax += L.val[k] * L.x[L.col[k]];
              double res = L.b[c] - ax;
              L.x[c] += g.omega * res / L.diag[c];
              m = std::max(m, std::fabs(res));
            }
          }
          ran = ls + 1;
          if (ls == 0) first = m;
          last = m;
          if (first > 1e-9 && m < first * g.block_reduction) break;
        }
        L.sweeps[b] = (uint8_t)ran;
        if (last > theta) {
          int c[3];
          block_coords(L.D, b, c);
#pragma omp atomic write
          nxt[b] = 1;
          for (int a = 0; a < 3; ++a)
            for (int sd = -1; sd <= 1; sd += 2) {
              int q[3] = {c[0], c[1], c[2]};
              q[a] += sd;
              if (q[a] < 0 || q[a] >= L.D[a]) continue;
#pragma omp atomic write
              nxt[block_id(L.D, q[0], q[1], q[2])] = 1;
            }
        }
      }
    }
    act.swap(nxt);
  }
  L.act = act;
}

void apply(const Level& L, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < L.ncells; ++c) {
    if (!L.liquid[c]) { y[c] = 0.0; continue; }
    double ax = L.diag[c] * x[c];
    for (int k = L.rowptr[c]; k < L.rowptr[c + 1]; ++k) ax += L.val[k] * x[L.col[k]];
    y[c] = ax;
  }
}

void residual(Level& L) {   // computeExplicitResidual (VCycle.cuh:81-131), r = 0 at air cells
  apply(L, L.x.data(), L.r.data());
  for (int c = 0; c < L.ncells; ++c) L.r[c] = L.liquid[c] ? L.b[c] - L.r[c] : 0.0;
}

void restrict_to(OGrid& g, int k) {   // restrictResidualsKernel: sum of children (MultigridTransfer.cuh:9-59)
  Level& F = g.mg[k];
  Level& C = g.mg[k + 1];
  std::fill(C.b.begin(), C.b.end(), 0.0);
  for (int c = 0; c < F.ncells; ++c) C.b[F.parent[c]] += F.r[c];
}

void prolong_from(OGrid& g, int k) {  // prolongateCorrection: injection-add (MultigridTransfer.cuh:63-102)
  Level& F = g.mg[k];
  Level& C = g.mg[k + 1];
  for (int c = 0; c < F.ncells; ++c)
    if (F.liquid[c]) F.x[c] += g.prolong_scale * C.x[F.parent[c]];
}

// Exact solve of a small coarsest level (<= 64 cells) restricted to its
// liquid cells: Gauss-Jordan elimination in FP64 without pivot search (the
// matrix is SPD when the level touches air); a pivot below 1e-12 of the
// largest diagonal drops its unknown (x = 0), which keeps an all-Neumann
// component solvable.
constexpr int DIRECT_MAX = 64;

void coarse_direct(Level& L) {
  const int n = L.ncells;
  std::vector<double> A((size_t)n * n, 0.0), I((size_t)n * n, 0.0);
  double dmax = 0.0;
  for (int c = 0; c < n; ++c) {
    I[(size_t)c * n + c] = 1.0;
    if (!L.liquid[c]) {
      A[(size_t)c * n + c] = 1.0;
      continue;
    }
    A[(size_t)c * n + c] = L.diag[c];
    dmax = std::max(dmax, L.diag[c]);
    for (int k = L.rowptr[c]; k < L.rowptr[c + 1]; ++k) A[(size_t)c * n + L.col[k]] = L.val[k];
  }
  std::vector<char> dead(n, 0);
  for (int p = 0; p < n; ++p) {
    double piv = A[(size_t)p * n + p];
    if (!(std::fabs(piv) > 1e-12 * dmax)) {
      dead[p] = 1;
      continue;
    }
    const double inv = 1.0 / piv;
    for (int j = 0; j < n; ++j) {
      A[(size_t)p * n + j] *= inv;
      I[(size_t)p * n + j] *= inv;
    }
    for (int i = 0; i < n; ++i) {
      if (i == p) continue;
      const double f = A[(size_t)i * n + p];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        A[(size_t)i * n + j] -= f * A[(size_t)p * n + j];
        I[(size_t)i * n + j] -= f * I[(size_t)p * n + j];
      }
    }
  }
  std::vector<double> x(n, 0.0);
  for (int i = 0; i < n; ++i) {
    if (!L.liquid[i] || dead[i]) continue;
    double s = 0.0;
    for (int j = 0; j < n; ++j)
      if (L.liquid[j] && !dead[j]) s += I[(size_t)i * n + j] * L.b[j];
    x[i] = s;
  }
  L.x = x;
}

// Recursive V-cycle (VCycleSolver::runCycle, VCycle.cuh:251-304): pre-smooth,
// residual, restrict, zero the coarse guess, recurse, prolongate, post-smooth.
// The coarsest level is solved exactly when it has at most 64 cells and
// otherwise gets `coarse_sweeps` symmetric sweeps (the reference runs 50
// one-sided sweeps, VCycle.cuh:252-256).
void vcycle(OGrid& g, int k) {
  Level& L = g.mg[k];
  if (k + 1 == (int)g.mg.size()) {
    if (L.ncells <= DIRECT_MAX) {
      coarse_direct(L);
      return;
    }
    for (int i = 0; i < g.coarse_sweeps; ++i) {
      smooth(g, L, 1, false, true);
      smooth(g, L, 1, true, true);
    }
    return;
  }
  smooth(g, L, g.pre, false);
  residual(L);
  restrict_to(g, k);
  std::fill(g.mg[k + 1].x.begin(), g.mg[k + 1].x.end(), 0.0);
  vcycle(g, k + 1);
  prolong_from(g, k);
  smooth(g, L, g.post, true);
}

double dot(const std::vector<double>& a, const std::vector<double>& b, const std::vector<uint8_t>& m) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i)
    if (m[i]) s += a[i] * b[i];
  return s;
}

// Pressure solve on level 0.  method 0: conjugate gradients preconditioned by
// one V-cycle (zero initial guess on every level); method 1: stationary
// V-cycles x += V(b - A x) as the reference's main loop (main.cu:314-317).
// Stops when ||b - A x||_2 <= tol ||b||_2.
int solve(OGrid& g, int method, double tol, int max_iters, int* iters_out, double* rel_out) {
  Level& L = g.mg[0];
  const int n = L.ncells;
  std::vector<double> x(g.p), bvec(g.rhs);
  for (int c = 0; c < n; ++c)
    if (!L.liquid[c]) { x[c] = 0.0; bvec[c] = 0.0; }
  double bn = std::sqrt(dot(bvec, bvec, L.liquid));
  std::vector<double> r(n), z(n), pv(n), q(n);
  apply(L, x.data(), q.data());
  for (int c = 0; c < n; ++c) r[c] = L.liquid[c] ? bvec[c] - q[c] : 0.0;
  double rn = std::sqrt(dot(r, r, L.liquid));
  int it = 0;
  if (bn == 0.0) {
    g.p = x;
    *iters_out = 0;
    *rel_out = 0.0;
    return 0;
  }
  auto precond = [&](const std::vector<double>& rin, std::vector<double>& zout) {
    L.b = rin;
    std::fill(L.x.begin(), L.x.end(), 0.0);
    vcycle(g, 0);
    zout = L.x;
  };
  if (method == 1) {
    while (rn > tol * bn && it < max_iters) {
      precond(r, z);
      for (int c = 0; c < n; ++c) x[c] += z[c];
      apply(L, x.data(), q.data());
      for (int c = 0; c < n; ++c) r[c] = L.liquid[c] ? bvec[c] - q[c] : 0.0;
      rn = std::sqrt(dot(r, r, L.liquid));
      ++it;
    }
  } else {
    precond(r, z);
    pv = z;
    double rz = dot(r, z, L.liquid);
    while (rn > tol * bn && it < max_iters) {
      apply(L, pv.data(), q.data());
      double alpha = rz / dot(pv, q, L.liquid);
      for (int c = 0; c < n; ++c) {
        x[c] += alpha * pv[c];
        r[c] -= alpha * q[c];
      }
      rn = std::sqrt(dot(r, r, L.liquid));
      ++it;
      if (rn <= tol * bn) break;
      precond(r, z);
      double rz2 = dot(r, z, L.liquid);
      double bt = rz2 / rz;
      rz = rz2;
      for (int c = 0; c < n; ++c) pv[c] = z[c] + bt * pv[c];
    }
  }
  g.p = x;
  *iters_out = it;
  *rel_out = rn / bn;
  return rn <= tol * bn ? 0 : 4;
}

// ---------------------------------------------------- particle transfers
// Velocities live on the native MAC faces of each block (r+1 slots on the
// normal axis, MSBG.cuh:118-120).  Particles talk to a virtual fine MAC
// lattice at spacing h: X nodes at (i, j+1/2, k+1/2) h, i in [0, N), node i = N
// (the +x wall) is 0; transverse indices clamp to [0, N-1].  A block of native
// level e maps its slots onto the fine nodes it covers with the operator Pi:
// normal coordinate xi = l / s, transverse eta = (l + 1/2) / s - 1/2 clamped
// into [0, r-1], trilinear in (xi, eta, zeta).  For e = 0 Pi is the identity.
// G2P samples Pi(u) trilinearly (the reference's interpolateComponent,
// G2P.cuh:11-72, with its MAC offsets fixed: A5/A6); P2G deposits with the
// same weights (transpose), then applies Pi^T, so G2P and P2G are exact
// transposes and every weight set sums to one.
struct PiW { int slot[8]; double w[8]; int n; };

inline void pi_weights(int a, int e, const int* l, PiW& o) {
  int s = 1 << e, r = B >> e;
  double xi[3];
  int i0[3];
  double t[3];
  for (int d = 0; d < 3; ++d) {
    if (d == a) {
      xi[d] = (double)l[d] / s;
      i0[d] = l[d] >> e;
      t[d] = xi[d] - i0[d];
    } else {
      double eta = (l[d] + 0.5) / s - 0.5;
      eta = std::min(std::max(eta, 0.0), (double)(r - 1));
      int j = (int)std::floor(eta);
      if (r >= 2) j = std::min(j, r - 2); else j = 0;
      i0[d] = j;
      t[d] = eta - j;
    }
  }
  o.n = 0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        int dd[3] = {dx, dy, dz};
        double w = 1.0;
        int c[3];
        for (int d = 0; d < 3; ++d) {
          w *= dd[d] ? t[d] : 1.0 - t[d];
          c[d] = i0[d] + dd[d];
        }
        if (w == 0.0) continue;
        o.slot[o.n] = face_index(a, c[0], c[1], c[2], r);
        o.w[o.n] = w;
        o.n++;
      }
}

// fine lattice address; -1 where the node is the +wall / outside (value 0)
inline int64_t fine_addr(const OGrid& g, int a, const int* gn) {
  int q[3];
  for (int d = 0; d < 3; ++d) {
    if (d == a) {
      if (gn[d] < 0 || gn[d] >= g.N[d]) return -1;
      q[d] = gn[d];
    } else {
      q[d] = iclamp(gn[d], 0, g.N[d] - 1);
    }
  }
  int b = block_id(g.D, q[0] >> LOGB, q[1] >> LOGB, q[2] >> LOGB);
  int l = (q[0] & 7) + 8 * ((q[1] & 7) + 8 * (q[2] & 7));
  return ((int64_t)b * 3 + a) * 512 + l;
}

// Trilinear stencil of a point (fine units u) for component a: the 8 nodes
// and weights.  Normal axis nodes at integer u, transverse at u - 1/2.
inline void stencil8(const OGrid& g, int a, const double* u, int64_t* addr, double* w) {
  int i0[3];
  double t[3];
  for (int d = 0; d < 3; ++d) {
    double c = (d == a) ? u[d] : u[d] - 0.5;
    i0[d] = (int)std::floor(c);
    t[d] = c - i0[d];
  }
  int k = 0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx, ++k) {
        int dd[3] = {dx, dy, dz};
        int gn[3];
        double ww = 1.0;
        for (int d = 0; d < 3; ++d) {
          gn[d] = i0[d] + dd[d];
          ww *= dd[d] ? t[d] : 1.0 - t[d];
        }
        addr[k] = fine_addr(g, a, gn);
        w[k] = ww;
      }
}

// FP32 position -> fine units, clamped inside [0, N - 2^-10] exactly as the
// HIP path does it.
inline void to_units(const OGrid& g, const float* x, double* u) {
  const float inv_h = (float)(1.0 / g.h);
  for (int a = 0; a < 3; ++a) {
    float v = x[a] * inv_h;
    v = std::fmin(std::fmax(v, 0.0f), (float)g.N[a] - 0.0009765625f);
    u[a] = v;
  }
}
inline void to_units(const OGrid& g, const double* x, double* u) {
  for (int a = 0; a < 3; ++a) u[a] = std::min(std::max(x[a] / g.h, 0.0), (double)g.N[a] - 0.0009765625);
}

// dens (may be null): relative density of each particle (input order).  With
// it the fold also sets beta = 1 / rho_face, rho_face = sum w rho / sum w (the
// density channels the reference reserves, MSBG.cuh:21-34, and its commented
// beta deposit, P2G.cuh:93-97).
template <class T>
void p2g(OGrid& g, const T* pos, const T* vel, int64_t n, double dt, const double* grav, const double* dens = nullptr) {
  const size_t fl = (size_t)g.nb * 3 * 512;
  g.FM.assign(fl, 0.0);
  g.Fm.assign(fl, 0.0);
  g.FD.assign(dens ? fl : 0, 0.0);
  std::fill(g.pcount.begin(), g.pcount.end(), 0u);
  // bin by block; deposit in 8 parity classes so blocks of one class touch
  // disjoint nodes (a block's stencils reach one node past its faces)
  std::vector<int> key(n);
  std::vector<std::vector<int64_t>> lists(g.nb);
  for (int64_t p = 0; p < n; ++p) {
    int gc[3];
    fine_cell(g, pos + 3 * p, gc);
    int b = block_id(g.D, gc[0] >> LOGB, gc[1] >> LOGB, gc[2] >> LOGB);
    key[p] = b;
    lists[b].push_back(p);
    int e = g.level[b], r = B >> e;
    int l[3];
    for (int a = 0; a < 3; ++a) l[a] = (gc[a] & 7) >> e;
    g.pcount[g.coff[b] + cell_index(l[0], l[1], l[2], r)]++;
  }
  for (int cls = 0; cls < 8; ++cls) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int b = 0; b < g.nb; ++b) {
      int c[3];
      block_coords(g.D, b, c);
      if (((c[0] & 1) | ((c[1] & 1) << 1) | ((c[2] & 1) << 2)) != cls) continue;
      for (int64_t p : lists[b]) {
        double u[3];
        to_units(g, pos + 3 * p, u);
        for (int a = 0; a < 3; ++a) {
          int64_t ad[8];
          double w[8];
          stencil8(g, a, u, ad, w);
          for (int k = 0; k < 8; ++k) {
            if (ad[k] < 0) continue;
            g.FM[ad[k]] += w[k] * (double)vel[3 * p + a];
            g.Fm[ad[k]] += w[k];
            if (dens) g.FD[ad[k]] += w[k] * dens[p];
          }
        }
      }
    }
  }
  // Pi^T per block into the native accumulators
  for (int a = 0; a < 3; ++a) {
    std::fill(g.Macc[a].begin(), g.Macc[a].end(), 0.0);
    std::fill(g.macc[a].begin(), g.macc[a].end(), 0.0);
    g.Dacc[a].assign(dens ? g.nfaces[a] : 0, 0.0);
  }
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int e = g.level[b];
    for (int a = 0; a < 3; ++a)
      for (int li = 0; li < 512; ++li) {
        int64_t ad = ((int64_t)b * 3 + a) * 512 + li;
        double M = g.FM[ad], m = g.Fm[ad];
        if (M == 0.0 && m == 0.0) continue;
        int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
        PiW pw;
        pi_weights(a, e, l, pw);
        for (int k = 0; k < pw.n; ++k) {
          g.Macc[a][g.foff[a][b] + pw.slot[k]] += pw.w[k] * M;
          g.macc[a][g.foff[a][b] + pw.slot[k]] += pw.w[k] * m;
          if (dens) g.Dacc[a][g.foff[a][b] + pw.slot[k]] += pw.w[k] * g.FD[ad];
        }
      }
  }
  for (int64_t c = 0; c < g.ncells; ++c) g.liquid[c] = g.pcount[c] > 0;
  // Fold duplicate and coarse–fine slots, normalise u = M / m (P2G.cuh:139-171,
  // threshold 1e-6), add the body force on faces that carry mass.
  // Boundary slot ownership: equal-level neighbours hold duplicate copies of
  // one face (summed); across a level change the fine side's slots are the
  // unknowns and the coarse side's slot is their average, so its deposit is
  // split 1/4 to each fine sub-face; wall slots are 0.
  const double eps = 1e-6;
  auto norm = [&](double M, double m) { return m > eps ? M / m : 0.0; };
  // beta = 1 / rho_face = m / D on a face with mass, the reference density's
  // 1 on one without; folded like the velocity (duplicates summed, the coarse
  // side of a level change the mean of its four fine sub-faces); walls 0
  auto binv = [&](double D, double m) { return m > eps ? m / D : 1.0; };
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int c[3];
    block_coords(g.D, b, c);
    int r = block_res(g, b);
    for (int a = 0; a < 3; ++a) {
      int nn[3] = {r, r, r};
      nn[a] = r + 1;
      for (int z = 0; z < nn[2]; ++z)
        for (int y = 0; y < nn[1]; ++y)
          for (int x = 0; x < nn[0]; ++x) {
            int sl[3] = {x, y, z};
            int64_t id = g.foff[a][b] + face_index(a, x, y, z, r);
            double M = g.Macc[a][id], m = g.macc[a][id];
            double uo, us, mtot;
            bool bnd = sl[a] == 0 || sl[a] == r;
            int nbb = -1;
            if (bnd) {
              int cc[3] = {c[0], c[1], c[2]};
              cc[a] += sl[a] == 0 ? -1 : 1;
              if (cc[a] >= 0 && cc[a] < g.D[a]) nbb = block_id(g.D, cc[0], cc[1], cc[2]);
            }
            const double D = dens ? g.Dacc[a][id] : 0.0;
            double bt = 1.0;
            if (bnd && nbb < 0) {
              uo = us = 0.0;
              mtot = 0.0;
              bt = 0.0;
            } else if (!bnd) {
              mtot = m;
              uo = norm(M, m);
              us = uo + (m > eps ? dt * grav[a] : 0.0);
              if (dens) bt = binv(D, m);
            } else {
              int en = g.level[nbb], rn = B >> en, e = g.level[b];
              int pa = sl[a] == 0 ? rn : 0;    // partner slot along the axis
              int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
              if (en == e) {
                int ps[3] = {x, y, z};
                ps[a] = pa;
                int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
                mtot = m + g.macc[a][pid];
                uo = norm(M + g.Macc[a][pid], mtot);
                us = uo + (mtot > eps ? dt * grav[a] : 0.0);
                if (dens) bt = binv(D + g.Dacc[a][pid], mtot);
              } else if (en == e + 1) {
                int ps[3];
                ps[a] = pa;
                ps[t1] = sl[t1] >> 1;
                ps[t2] = sl[t2] >> 1;
                int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
                mtot = m + 0.25 * g.macc[a][pid];
                uo = norm(M + 0.25 * g.Macc[a][pid], mtot);
                us = uo + (mtot > eps ? dt * grav[a] : 0.0);
                if (dens) bt = binv(D + 0.25 * g.Dacc[a][pid], mtot);
              } else {
                uo = us = mtot = 0.0;
                bt = 0.0;
                for (int d1 = 0; d1 < 2; ++d1)
                  for (int d2 = 0; d2 < 2; ++d2) {
                    int ps[3];
                    ps[a] = pa;
                    ps[t1] = 2 * sl[t1] + d1;
                    ps[t2] = 2 * sl[t2] + d2;
                    int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
                    double ms = g.macc[a][pid] + 0.25 * m;
                    double u0 = norm(g.Macc[a][pid] + 0.25 * M, ms);
                    uo += 0.25 * u0;
                    us += 0.25 * (u0 + (ms > eps ? dt * grav[a] : 0.0));
                    mtot += ms;
                    if (dens) bt += 0.25 * binv(g.Dacc[a][pid] + 0.25 * D, ms);
                  }
                if (!dens) bt = 1.0;
              }
            }
            g.uold[a][id] = uo;
            g.ustar[a][id] = us;
            g.mt[a][id] = mtot;
            if (dens) g.beta[a][id] = bt;
          }
    }
  }
}

// Volume-integrated divergence in units of h^2 times velocity:
// D_c = s^2 sum_a (u[+] - u[-]); rhs b = -D_c on liquid cells.
void divergence(OGrid& g) {
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int r = block_res(g, b);
    double s = (double)(1 << g.level[b]);
    for (int z = 0; z < r; ++z)
      for (int y = 0; y < r; ++y)
        for (int x = 0; x < r; ++x) {
          int64_t c = g.coff[b] + cell_index(x, y, z, r);
          double d = 0.0;
          d += g.ustar[0][g.foff[0][b] + face_index(0, x + 1, y, z, r)] - g.ustar[0][g.foff[0][b] + face_index(0, x, y, z, r)];
          d += g.ustar[1][g.foff[1][b] + face_index(1, x, y + 1, z, r)] - g.ustar[1][g.foff[1][b] + face_index(1, x, y, z, r)];
          d += g.ustar[2][g.foff[2][b] + face_index(2, x, y, z + 1, r)] - g.ustar[2][g.foff[2][b] + face_index(2, x, y, z, r)];
          g.rhs[c] = g.liquid[c] ? -s * s * d : 0.0;
        }
  }
}

// Pressure-gradient update u = u* - beta (p+ - p-) / d on every slot, with
// d = s between equal cells and 1.5 s_fine across a level change; air cells
// carry p = 0; the coarse side's slot of a level change is the average of
// the four fine sub-faces.
void project(OGrid& g) {
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int c[3];
    block_coords(g.D, b, c);
    int e = g.level[b], r = B >> e;
    double s = (double)(1 << e);
    for (int a = 0; a < 3; ++a) {
      int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
      int nn[3] = {r, r, r};
      nn[a] = r + 1;
      for (int z = 0; z < nn[2]; ++z)
        for (int y = 0; y < nn[1]; ++y)
          for (int x = 0; x < nn[0]; ++x) {
            int sl[3] = {x, y, z};
            int64_t id = g.foff[a][b] + face_index(a, x, y, z, r);
            double us = g.ustar[a][id], bt = g.beta[a][id];
            double out;
            if (sl[a] > 0 && sl[a] < r) {
              int cm[3] = {x, y, z}, cp[3] = {x, y, z};
              cm[a] -= 1;
              double pp = g.p[g.coff[b] + cell_index(cp[0], cp[1], cp[2], r)];
              double pm = g.p[g.coff[b] + cell_index(cm[0], cm[1], cm[2], r)];
              out = us - bt * (pp - pm) / s;
            } else {
              bool lo = sl[a] == 0;
              int cc[3] = {c[0], c[1], c[2]};
              cc[a] += lo ? -1 : 1;
              if (cc[a] < 0 || cc[a] >= g.D[a]) {
                out = 0.0;
              } else {
                int nbb = block_id(g.D, cc[0], cc[1], cc[2]);
                int en = g.level[nbb], rn = B >> en;
                int own[3] = {x, y, z};
                own[a] = lo ? 0 : r - 1;
                double po = g.p[g.coff[b] + cell_index(own[0], own[1], own[2], r)];
                if (en == e) {
                  int q[3] = {x, y, z};
                  q[a] = lo ? rn - 1 : 0;
                  double pn = g.p[g.coff[nbb] + cell_index(q[0], q[1], q[2], rn)];
                  out = us - bt * (lo ? po - pn : pn - po) / s;
                } else if (en == e + 1) {
                  int q[3];
                  q[a] = lo ? rn - 1 : 0;
                  q[t1] = sl[t1] >> 1;
                  q[t2] = sl[t2] >> 1;
                  double pn = g.p[g.coff[nbb] + cell_index(q[0], q[1], q[2], rn)];
                  out = us - bt * (lo ? po - pn : pn - po) / (1.5 * s);
                } else {
                  out = 0.0;
                  double sf = 0.5 * s;
                  for (int d1 = 0; d1 < 2; ++d1)
                    for (int d2 = 0; d2 < 2; ++d2) {
                      int q[3], ps[3];
                      q[a] = lo ? rn - 1 : 0;
                      ps[a] = lo ? rn : 0;
                      q[t1] = ps[t1] = 2 * sl[t1] + d1;
                      q[t2] = ps[t2] = 2 * sl[t2] + d2;
                      int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
                      double pn = g.p[g.coff[nbb] + cell_index(q[0], q[1], q[2], rn)];
                      out += 0.25 * (g.ustar[a][pid] - g.beta[a][pid] * (lo ? po - pn : pn - po) / (1.5 * sf));
                    }
                }
              }
            }
            g.unew[a][id] = out;
          }
    }
  }
}

// Pi of the new field and of the change (u_new - u_old) on the fine lattice.
void to_fine(OGrid& g) {
  const size_t fl = (size_t)g.nb * 3 * 512;
  g.Fnew.assign(fl, 0.0);
  g.Fdel.assign(fl, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int e = g.level[b];
    for (int a = 0; a < 3; ++a)
      for (int li = 0; li < 512; ++li) {
        int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
        PiW pw;
        pi_weights(a, e, l, pw);
        double vn = 0.0, vd = 0.0;
        for (int k = 0; k < pw.n; ++k) {
          int64_t id = g.foff[a][b] + pw.slot[k];
          vn += pw.w[k] * g.unew[a][id];
          vd += pw.w[k] * (g.unew[a][id] - g.uold[a][id]);
        }
        int64_t ad = ((int64_t)b * 3 + a) * 512 + li;
        g.Fnew[ad] = vn;
        g.Fdel[ad] = vd;
      }
  }
}

inline void sample(const OGrid& g, const std::vector<double>& F, const double* u, double* v) {
  for (int a = 0; a < 3; ++a) {
    int64_t ad[8];
    double w[8];
    stencil8(g, a, u, ad, w);
    double s = 0.0;
    for (int k = 0; k < 8; ++k)
      if (ad[k] >= 0) s += w[k] * F[ad[k]];
    v[a] = s;
  }
}

// FLIP/PIC blend (g2pUpdateKernel, G2P.cuh:87-128) then Ralston RK3 through the
// new field (advectParticlesKernel, G2P.cuh:131-173) with the wall clamp in
// cells: [h/4, extent - h/4] (the reference clamps 0.1 world units, A9).
template <class T>
void g2p_advect(const OGrid& g, const T* pos, const T* vel, int64_t n, double alpha, double dt,
                double* pos_out, double* vel_out) {
  const double h = g.h;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    double u[3], vn[3], vd[3];
    to_units(g, pos + 3 * p, u);
    sample(g, g.Fnew, u, vn);
    sample(g, g.Fdel, u, vd);
    for (int a = 0; a < 3; ++a)
      vel_out[3 * p + a] = alpha * ((double)vel[3 * p + a] + vd[a]) + (1.0 - alpha) * vn[a];
    double x[3], x2[3], x3[3], v2[3], v3[3], uu[3];
    for (int a = 0; a < 3; ++a) x[a] = pos[3 * p + a];
    auto samp = [&](const double* xx, double* out) {
      for (int a = 0; a < 3; ++a) {
        double t = xx[a] / h;
        uu[a] = std::min(std::max(t, -1.0), (double)g.N[a] + 1.0);
      }
      sample(g, g.Fnew, uu, out);
    };
    for (int a = 0; a < 3; ++a) x2[a] = x[a] + 0.5 * dt * vn[a];
    samp(x2, v2);
    for (int a = 0; a < 3; ++a) x3[a] = x[a] + 0.75 * dt * v2[a];
    samp(x3, v3);
    for (int a = 0; a < 3; ++a) {
      double xn = x[a] + dt * ((2.0 / 9.0) * vn[a] + (3.0 / 9.0) * v2[a] + (4.0 / 9.0) * v3[a]);
      double lo = 0.25 * h, hi = g.N[a] * h - 0.25 * h;
      pos_out[3 * p + a] = std::min(std::max(xn, lo), hi);
    }
  }
}

// ------------------------------------------------------------ adjoint
// Reverse-mode derivative of one step with the topology, the liquid mask and
// the mass thresholds held at their forward values (they are piecewise
// constant in the particle state).  Every linear stage is transposed slot by
// slot from its forward formula above; the nonlinear ones (u = M / m, the
// trilinear weights as functions of position, the position clamps) are
// differentiated at the forward point.  Accumulations into shared fields use
// atomics so the transposes read exactly like the forward loops.

// Trilinear stencil with d(weight)/d(fine units) along each axis.
inline void stencil8_grad(const OGrid& g, int a, const double* u, int64_t* addr, double* w, double (*dw)[3]) {
  int i0[3];
  double t[3];
  for (int d = 0; d < 3; ++d) {
    double c = (d == a) ? u[d] : u[d] - 0.5;
    i0[d] = (int)std::floor(c);
    t[d] = c - i0[d];
  }
  int k = 0;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx, ++k) {
        int dd[3] = {dx, dy, dz};
        int gn[3];
        double f[3], df[3];
        for (int d = 0; d < 3; ++d) {
          gn[d] = i0[d] + dd[d];
          f[d] = dd[d] ? t[d] : 1.0 - t[d];
          df[d] = dd[d] ? 1.0 : -1.0;
        }
        addr[k] = fine_addr(g, a, gn);
        w[k] = f[0] * f[1] * f[2];
        dw[k][0] = df[0] * f[1] * f[2];
        dw[k][1] = f[0] * df[1] * f[2];
        dw[k][2] = f[0] * f[1] * df[2];
      }
}

// grad[d] += sum_a wbar[a] d S_a(F)(u) / du_d
inline void sample_vjp(const OGrid& g, const std::vector<double>& F, const double* u, const double* wbar, double* grad) {
  for (int a = 0; a < 3; ++a) {
    if (wbar[a] == 0.0) continue;
    int64_t ad[8];
    double w[8], dw[8][3];
    stencil8_grad(g, a, u, ad, w, dw);
    for (int k = 0; k < 8; ++k) {
      if (ad[k] < 0) continue;
      for (int d = 0; d < 3; ++d) grad[d] += wbar[a] * F[ad[k]] * dw[k][d];
    }
  }
}

// F[node] += w(u) * val[a] on every component lattice (the transpose of sample)
inline void deposit_atomic(const OGrid& g, std::vector<double>& F, const double* u, const double* val) {
  for (int a = 0; a < 3; ++a) {
    if (val[a] == 0.0) continue;
    int64_t ad[8];
    double w[8];
    stencil8(g, a, u, ad, w);
    for (int k = 0; k < 8; ++k) {
      if (ad[k] < 0) continue;
#pragma omp atomic
      F[ad[k]] += w[k] * val[a];
    }
  }
}

// Pi^T of a fine-lattice field into per-slot accumulators, scaled by sgn.
void pi_transpose(const OGrid& g, const std::vector<double>& F, std::vector<double>* out, double sgn) {
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int e = g.level[b];
    for (int a = 0; a < 3; ++a)
      for (int li = 0; li < 512; ++li) {
        double v = F[((int64_t)b * 3 + a) * 512 + li];
        if (v == 0.0) continue;
        int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
        PiW pw;
        pi_weights(a, e, l, pw);
        for (int k = 0; k < pw.n; ++k) out[a][g.foff[a][b] + pw.slot[k]] += sgn * pw.w[k] * v;
      }
  }
}

// Pi of per-slot fields onto the fine lattice.
void pi_forward(const OGrid& g, const std::vector<double>* S, std::vector<double>& F) {
  F.assign((size_t)g.nb * 3 * 512, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int e = g.level[b];
    for (int a = 0; a < 3; ++a)
      for (int li = 0; li < 512; ++li) {
        int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
        PiW pw;
        pi_weights(a, e, l, pw);
        double v = 0.0;
        for (int k = 0; k < pw.n; ++k) v += pw.w[k] * S[a][g.foff[a][b] + pw.slot[k]];
        F[((int64_t)b * 3 + a) * 512 + li] = v;
      }
  }
}

// Adjoint of g2p_advect: given xbo = dL/dx' and vbo = dL/dv', write the
// particles' direct partials xbi, vbi and accumulate Fbar_nd (= Fnew-bar +
// Fdel-bar) and Fbar_d on the fine lattice; then ubar_new = Pi^T Fbar_nd and
// ubar_old = -Pi^T Fbar_d.
template <class T>
void g2p_advect_adjoint(OGrid& g, const T* pos, const T* vel, int64_t n, double alpha, double dt, const double* xbo,
                        const double* vbo, double* xbi, double* vbi) {
  const size_t fl = (size_t)g.nb * 3 * 512;
  g.Fbar_nd.assign(fl, 0.0);
  g.Fbar_d.assign(fl, 0.0);
  const double h = g.h;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    double u[3], m1[3], vn[3], vd[3], x[3];
    to_units(g, pos + 3 * p, u);
    for (int a = 0; a < 3; ++a) {
      x[a] = (double)pos[3 * p + a];
      double raw = x[a] / h;
      m1[a] = (raw > 0.0 && raw < (double)g.N[a] - 0.0009765625) ? 1.0 : 0.0;
    }
    sample(g, g.Fnew, u, vn);
    sample(g, g.Fdel, u, vd);
    auto units = [&](const double* xx, double* uu, double* mm) {
      for (int a = 0; a < 3; ++a) {
        double t = xx[a] / h;
        mm[a] = (t > -1.0 && t < (double)g.N[a] + 1.0) ? 1.0 : 0.0;
        uu[a] = std::min(std::max(t, -1.0), (double)g.N[a] + 1.0);
      }
    };
    double x2[3], u2[3], m2[3], v2[3], x3[3], u3[3], m3[3], v3[3];
    for (int a = 0; a < 3; ++a) x2[a] = x[a] + 0.5 * dt * vn[a];
    units(x2, u2, m2);
    sample(g, g.Fnew, u2, v2);
    for (int a = 0; a < 3; ++a) x3[a] = x[a] + 0.75 * dt * v2[a];
    units(x3, u3, m3);
    sample(g, g.Fnew, u3, v3);
    double xb[3], k1[3], k2[3], k3[3], xacc[3];
    for (int a = 0; a < 3; ++a) {
      double xn = x[a] + dt * ((2.0 / 9.0) * vn[a] + (3.0 / 9.0) * v2[a] + (4.0 / 9.0) * v3[a]);
      double lo = 0.25 * h, hi = g.N[a] * h - 0.25 * h;
      xb[a] = (xn > lo && xn < hi) ? xbo[3 * p + a] : 0.0;
      k1[a] = dt * (2.0 / 9.0) * xb[a];
      k2[a] = dt * (3.0 / 9.0) * xb[a];
      k3[a] = dt * (4.0 / 9.0) * xb[a];
      xacc[a] = xb[a];
    }
    // stage 3: v3 = S(Fnew)(x3 / h), x3 = x + 3/4 dt v2
    double gr[3] = {0, 0, 0};
    deposit_atomic(g, g.Fbar_nd, u3, k3);
    sample_vjp(g, g.Fnew, u3, k3, gr);
    for (int d = 0; d < 3; ++d) {
      double x3b = gr[d] / h * m3[d];
      xacc[d] += x3b;
      k2[d] += 0.75 * dt * x3b;
    }
    // stage 2: v2 = S(Fnew)(x2 / h), x2 = x + 1/2 dt vn
    gr[0] = gr[1] = gr[2] = 0.0;
    deposit_atomic(g, g.Fbar_nd, u2, k2);
    sample_vjp(g, g.Fnew, u2, k2, gr);
    for (int d = 0; d < 3; ++d) {
      double x2b = gr[d] / h * m2[d];
      xacc[d] += x2b;
      k1[d] += 0.5 * dt * x2b;
    }
    // G2P blend and stage 1 at x
    double vnb[3], vdb[3], both[3];
    for (int a = 0; a < 3; ++a) {
      vnb[a] = k1[a] + (1.0 - alpha) * vbo[3 * p + a];
      vdb[a] = alpha * vbo[3 * p + a];
      both[a] = vnb[a] + vdb[a];
    }
    deposit_atomic(g, g.Fbar_nd, u, both);
    deposit_atomic(g, g.Fbar_d, u, vdb);
    gr[0] = gr[1] = gr[2] = 0.0;
    sample_vjp(g, g.Fnew, u, vnb, gr);
    sample_vjp(g, g.Fdel, u, vdb, gr);
    for (int d = 0; d < 3; ++d) {
      xbi[3 * p + d] = xacc[d] + gr[d] / h * m1[d];
      vbi[3 * p + d] = alpha * vbo[3 * p + d];
    }
  }
  for (int a = 0; a < 3; ++a) {
    std::fill(g.ubar_new[a].begin(), g.ubar_new[a].end(), 0.0);
    std::fill(g.ubar_old[a].begin(), g.ubar_old[a].end(), 0.0);
  }
  pi_transpose(g, g.Fbar_nd, g.ubar_new, 1.0);
  pi_transpose(g, g.Fbar_d, g.ubar_old, -1.0);
}

inline void atomic_add(double& t, double v) {
#pragma omp atomic
  t += v;
}

// Adjoint of project(): ubar_new -> ubar_star (the direct u* terms) and pbar
// (the gradient terms), pbar kept on liquid cells only (air cells are not
// unknowns).
void project_adjoint(OGrid& g) {
  for (int a = 0; a < 3; ++a) std::fill(g.ubar_star[a].begin(), g.ubar_star[a].end(), 0.0);
  std::fill(g.pbar.begin(), g.pbar.end(), 0.0);
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int c[3];
    block_coords(g.D, b, c);
    int e = g.level[b], r = B >> e;
    double s = (double)(1 << e);
    for (int a = 0; a < 3; ++a) {
      int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
      int nn[3] = {r, r, r};
      nn[a] = r + 1;
      for (int z = 0; z < nn[2]; ++z)
        for (int y = 0; y < nn[1]; ++y)
          for (int x = 0; x < nn[0]; ++x) {
            int sl[3] = {x, y, z};
            int64_t id = g.foff[a][b] + face_index(a, x, y, z, r);
            double ub = g.ubar_new[a][id], bt = g.beta[a][id];
            if (ub == 0.0) continue;
            if (sl[a] > 0 && sl[a] < r) {
              int cm[3] = {x, y, z};
              cm[a] -= 1;
              atomic_add(g.ubar_star[a][id], ub);
              atomic_add(g.pbar[g.coff[b] + cell_index(x, y, z, r)], -bt / s * ub);
              atomic_add(g.pbar[g.coff[b] + cell_index(cm[0], cm[1], cm[2], r)], bt / s * ub);
              continue;
            }
            bool lo = sl[a] == 0;
            double sg = lo ? 1.0 : -1.0;  // out = us - bt * sg * (po - pn) / d
            int cc[3] = {c[0], c[1], c[2]};
            cc[a] += lo ? -1 : 1;
            if (cc[a] < 0 || cc[a] >= g.D[a]) continue;
            int nbb = block_id(g.D, cc[0], cc[1], cc[2]);
            int en = g.level[nbb], rn = B >> en;
            int own[3] = {x, y, z};
            own[a] = lo ? 0 : r - 1;
            int64_t co = g.coff[b] + cell_index(own[0], own[1], own[2], r);
            if (en == e || en == e + 1) {
              int q[3] = {x, y, z};
              q[a] = lo ? rn - 1 : 0;
              double d = s;
              if (en == e + 1) {
                q[t1] = sl[t1] >> 1;
                q[t2] = sl[t2] >> 1;
                d = 1.5 * s;
              }
              int64_t cn = g.coff[nbb] + cell_index(q[0], q[1], q[2], rn);
              atomic_add(g.ubar_star[a][id], ub);
              atomic_add(g.pbar[co], -bt * sg / d * ub);
              atomic_add(g.pbar[cn], bt * sg / d * ub);
            } else {
              double sf = 0.5 * s;
              for (int d1 = 0; d1 < 2; ++d1)
                for (int d2 = 0; d2 < 2; ++d2) {
                  int q[3], ps[3];
                  q[a] = lo ? rn - 1 : 0;
                  ps[a] = lo ? rn : 0;
                  q[t1] = ps[t1] = 2 * sl[t1] + d1;
                  q[t2] = ps[t2] = 2 * sl[t2] + d2;
                  int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
                  int64_t cn = g.coff[nbb] + cell_index(q[0], q[1], q[2], rn);
                  double k = 0.25 * g.beta[a][pid] * sg / (1.5 * sf);
                  atomic_add(g.ubar_star[a][pid], 0.25 * ub);
                  atomic_add(g.pbar[co], -k * ub);
                  atomic_add(g.pbar[cn], k * ub);
                }
            }
          }
    }
  }
  for (int64_t c = 0; c < g.ncells; ++c)
    if (!g.liquid[c]) g.pbar[c] = 0.0;
}

// Adjoint of divergence(): ubar_star += D^T rhsbar with rhs = -s^2 sum(u+ - u-).
void divergence_adjoint(OGrid& g) {
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int r = block_res(g, b);
    double s = (double)(1 << g.level[b]);
    for (int z = 0; z < r; ++z)
      for (int y = 0; y < r; ++y)
        for (int x = 0; x < r; ++x) {
          int64_t c = g.coff[b] + cell_index(x, y, z, r);
          if (!g.liquid[c]) continue;
          double k = s * s * g.rhsbar[c];
          g.ubar_star[0][g.foff[0][b] + face_index(0, x + 1, y, z, r)] -= k;
          g.ubar_star[0][g.foff[0][b] + face_index(0, x, y, z, r)] += k;
          g.ubar_star[1][g.foff[1][b] + face_index(1, x, y + 1, z, r)] -= k;
          g.ubar_star[1][g.foff[1][b] + face_index(1, x, y, z, r)] += k;
          g.ubar_star[2][g.foff[2][b] + face_index(2, x, y, z + 1, r)] -= k;
          g.ubar_star[2][g.foff[2][b] + face_index(2, x, y, z, r)] += k;
        }
  }
}

// Adjoint pressure solve: A is symmetric, so rhsbar = A^{-1} pbar with the
// forward hierarchy.  The forward p and rhs are kept.
int solve_adjoint(OGrid& g, double tol, int max_iters, int* iters, double* rel) {
  std::vector<double> p0 = g.p, r0 = g.rhs;
  for (int64_t c = 0; c < g.ncells; ++c) {
    g.rhs[c] = g.liquid[c] ? g.pbar[c] : 0.0;
    g.p[c] = 0.0;
  }
  int st = solve(g, 0, tol, max_iters, iters, rel);
  g.rhsbar = g.p;
  g.p = p0;
  g.rhs = r0;
  return st;
}

// Adjoint of the fold (u_old, u_star as functions of Macc, macc): with
// ub = ubar_old + ubar_star per slot, every slot formula u = M' / m' with
// M' = sum_j c_j Macc[j], m' = sum_j c_j macc[j] sends c_j ub / m' to
// Mbar[j] and -c_j ub u / m' to mbar[j] (m' > 1e-6; frozen branch).
void fold_adjoint(OGrid& g) {
  const double eps = 1e-6;
  for (int a = 0; a < 3; ++a) {
    std::fill(g.Mbar[a].begin(), g.Mbar[a].end(), 0.0);
    std::fill(g.mbar[a].begin(), g.mbar[a].end(), 0.0);
  }
#pragma omp parallel for schedule(dynamic, 8)
  for (int b = 0; b < g.nb; ++b) {
    int c[3];
    block_coords(g.D, b, c);
    int r = block_res(g, b), e = g.level[b];
    for (int a = 0; a < 3; ++a) {
      int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
      int nn[3] = {r, r, r};
      nn[a] = r + 1;
      for (int z = 0; z < nn[2]; ++z)
        for (int y = 0; y < nn[1]; ++y)
          for (int x = 0; x < nn[0]; ++x) {
            int sl[3] = {x, y, z};
            int64_t id = g.foff[a][b] + face_index(a, x, y, z, r);
            double ub = g.ubar_old[a][id] + g.ubar_star[a][id];
            if (ub == 0.0) continue;
            double M = g.Macc[a][id], m = g.macc[a][id];
            bool bnd = sl[a] == 0 || sl[a] == r;
            if (!bnd) {
              if (m > eps) {
                atomic_add(g.Mbar[a][id], ub / m);
                atomic_add(g.mbar[a][id], -ub * (M / m) / m);
              }
              continue;
            }
            int cc[3] = {c[0], c[1], c[2]};
            cc[a] += sl[a] == 0 ? -1 : 1;
            if (cc[a] < 0 || cc[a] >= g.D[a]) continue;
            int nbb = block_id(g.D, cc[0], cc[1], cc[2]);
            int en = g.level[nbb], rn = B >> en;
            int pa = sl[a] == 0 ? rn : 0;
            if (en == e || en == e + 1) {
              int ps[3] = {x, y, z};
              ps[a] = pa;
              double cp = 1.0;
              if (en == e + 1) {
                ps[t1] = sl[t1] >> 1;
                ps[t2] = sl[t2] >> 1;
                cp = 0.25;
              }
              int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
              double mt = m + cp * g.macc[a][pid];
              if (mt <= eps) continue;
              double u = (M + cp * g.Macc[a][pid]) / mt;
              atomic_add(g.Mbar[a][id], ub / mt);
              atomic_add(g.mbar[a][id], -ub * u / mt);
              atomic_add(g.Mbar[a][pid], cp * ub / mt);
              atomic_add(g.mbar[a][pid], -cp * ub * u / mt);
            } else {
              for (int d1 = 0; d1 < 2; ++d1)
                for (int d2 = 0; d2 < 2; ++d2) {
                  int ps[3];
                  ps[a] = pa;
                  ps[t1] = 2 * sl[t1] + d1;
                  ps[t2] = 2 * sl[t2] + d2;
                  int64_t pid = g.foff[a][nbb] + face_index(a, ps[0], ps[1], ps[2], rn);
                  double ms = g.macc[a][pid] + 0.25 * m;
                  if (ms <= eps) continue;
                  double u0 = (g.Macc[a][pid] + 0.25 * M) / ms;
                  double uk = 0.25 * ub;
                  atomic_add(g.Mbar[a][pid], uk / ms);
                  atomic_add(g.mbar[a][pid], -uk * u0 / ms);
                  atomic_add(g.Mbar[a][id], 0.25 * uk / ms);
                  atomic_add(g.mbar[a][id], -0.25 * uk * u0 / ms);
                }
            }
          }
    }
  }
}

// Adjoint of the deposit: FMbar = Pi Mbar, Fmbar = Pi mbar, then per particle
// vbar += S(FMbar)(x / h) and xbar += (v . dS(FMbar) + dS(Fmbar)) / h.
template <class T>
void p2g_adjoint(OGrid& g, const T* pos, const T* vel, int64_t n, double* xbar, double* vbar) {
  fold_adjoint(g);
  pi_forward(g, g.Mbar, g.FMbar);
  pi_forward(g, g.mbar, g.Fmbar);
  const double h = g.h;
  const double ones[3] = {1.0, 1.0, 1.0};
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    double u[3], m1[3], v[3], sv[3], gr[3] = {0, 0, 0};
    to_units(g, pos + 3 * p, u);
    for (int a = 0; a < 3; ++a) {
      double raw = (double)pos[3 * p + a] / h;
      m1[a] = (raw > 0.0 && raw < (double)g.N[a] - 0.0009765625) ? 1.0 : 0.0;
      v[a] = (double)vel[3 * p + a];
    }
    sample(g, g.FMbar, u, sv);
    sample_vjp(g, g.FMbar, u, v, gr);
    sample_vjp(g, g.Fmbar, u, ones, gr);
    for (int a = 0; a < 3; ++a) {
      vbar[3 * p + a] += sv[a];
      xbar[3 * p + a] += gr[a] / h * m1[a];
    }
  }
}

}  // namespace

extern "C" {

void* og_create(int dx, int dy, int dz, int num_levels, double h, const int* band) {
  if (dx <= 0 || dy <= 0 || dz <= 0 || num_levels < 1 || num_levels > 4) return nullptr;
  OGrid* g = new OGrid();
  g->D[0] = dx; g->D[1] = dy; g->D[2] = dz;
  for (int a = 0; a < 3; ++a) g->N[a] = g->D[a] * B;
  g->nb = dx * dy * dz;
  g->L = num_levels;
  g->h = h;
  for (int l = 0; l < 4; ++l) g->band[l] = band ? band[l] : 0;
  std::vector<uint8_t> lv(g->nb, 0);
  set_topology(*g, lv.data());
  return g;
}

void og_destroy(void* p) { delete static_cast<OGrid*>(p); }

uint32_t og_morton3d(uint32_t x, uint32_t y, uint32_t z) { return expand_bits(x) | (expand_bits(y) << 1) | (expand_bits(z) << 2); }

int og_bin(void* p, const float* pos, int64_t n, int32_t* key, uint32_t* count) {
  OGrid& g = *static_cast<OGrid*>(p);
  std::fill(count, count + g.nb, 0u);
  for (int64_t i = 0; i < n; ++i) {
    int gc[3];
    fine_cell(g, pos + 3 * i, gc);
    int b = block_id(g.D, gc[0] >> LOGB, gc[1] >> LOGB, gc[2] >> LOGB);
    key[i] = b;
    count[b]++;
  }
  return 0;
}

int og_levels_from_counts(void* p, const uint32_t* count, uint8_t* out) {
  levels_from_counts(*static_cast<OGrid*>(p), count, out);
  return 0;
}

int og_levels_from_radii(void* p, const float* pos, const float* radius, int64_t n, uint8_t* out) {
  return levels_from_radii(*static_cast<OGrid*>(p), pos, radius, n, out);
}

// Nearest-cell resample onto an ox x oy x oz grid of voxel centres, x fastest
// (sampleMSBG, vtx_visualization.cuh:8-62): the centre of voxel q along an
// axis of N fine cells lies at (q + 1/2) N / o fine cells, in fine cell
// floor((2q + 1) N / (2o)) (exact; a centre on a face goes to the cell above),
// which names the block and the native cell holding it.
int og_resample(void* p, const double* field, int ox, int oy, int oz, double* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  const int o[3] = {ox, oy, oz};
  for (int a = 0; a < 3; ++a)
    if (o[a] < 1) return 1;
  for (int z = 0; z < oz; ++z)
    for (int y = 0; y < oy; ++y)
      for (int x = 0; x < ox; ++x) {
        const int q[3] = {x, y, z};
        int gc[3];
        for (int a = 0; a < 3; ++a) gc[a] = (int)(((2 * (int64_t)q[a] + 1) * g.N[a]) / (2 * (int64_t)o[a]));
        const int b = block_id(g.D, gc[0] >> LOGB, gc[1] >> LOGB, gc[2] >> LOGB);
        const int e = g.level[b], r = B >> e;
        const int l[3] = {(gc[0] & (B - 1)) >> e, (gc[1] & (B - 1)) >> e, (gc[2] & (B - 1)) >> e};
        out[x + (size_t)ox * (y + (size_t)oy * z)] = field[g.coff[b] + cell_index(l[0], l[1], l[2], r)];
      }
  return 0;
}

// Remap of a cell field from a previous layout (native levels level_prev, same
// blocks) onto the current one: every current cell takes the value of the
// previous layout's cell that contains its centre.  The centre of a cell of
// edge s at local index l lies at fine coordinate 8 B + s l + s / 2; the
// point is located by containment in a level built for the previous levels
// (cell_of), at the fine cell just above the centre, so a centre that falls on
// a face of the previous cells takes the cell above it (the warm start of the
// pressure solve across topology changes; the reference has no counterpart:
// main.cu:300 starts every solve from p = 0).
int og_remap_cells(void* p, const uint8_t* level_prev, const double* src, double* dst) {
  OGrid& g = *static_cast<OGrid*>(p);
  Level P;
  for (int a = 0; a < 3; ++a) P.D[a] = g.D[a];
  P.bs = B;
  P.e.resize(g.nb);
  for (int b = 0; b < g.nb; ++b) {
    if (level_prev[b] > LOGB) return 1;
    P.e[b] = level_prev[b];
  }
  finish_level_cells(P);
  for (int b = 0; b < g.nb; ++b) {
    int Bc[3];
    block_coords(g.D, b, Bc);
    const int e = g.level[b], r = B >> e, s = 1 << e;
    for (int li = 0; li < r * r * r; ++li) {
      const int l[3] = {li % r, (li / r) % r, li / (r * r)};
      int q[3];
      for (int a = 0; a < 3; ++a) q[a] = Bc[a] * B + l[a] * s + s / 2;   // s = 1: the cell itself
      dst[g.coff[b] + li] = src[cell_of(P, q)];
    }
  }
  return 0;
}

int og_set_levels(void* p, const uint8_t* level) {
  OGrid& g = *static_cast<OGrid*>(p);
  for (int b = 0; b < g.nb; ++b)
    if (level[b] >= g.L) return 1;
  set_topology(g, level);
  return 0;
}

int og_get_levels(void* p, uint8_t* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  std::copy(g.level.begin(), g.level.end(), out);
  return 0;
}

int og_get_offsets(void* p, uint32_t* c, uint32_t* fx, uint32_t* fy, uint32_t* fz) {
  OGrid& g = *static_cast<OGrid*>(p);
  std::copy(g.coff.begin(), g.coff.end(), c);
  std::copy(g.foff[0].begin(), g.foff[0].end(), fx);
  std::copy(g.foff[1].begin(), g.foff[1].end(), fy);
  std::copy(g.foff[2].begin(), g.foff[2].end(), fz);
  return 0;
}

int og_totals(void* p, int64_t* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  out[0] = g.ncells;
  out[1] = g.nfaces[0];
  out[2] = g.nfaces[1];
  out[3] = g.nfaces[2];
  return 0;
}

int og_p2g(void* p, const float* pos, const float* vel, int64_t n, double dt, double gx, double gy, double gz,
           const double* density) {
  double gr[3] = {gx, gy, gz};
  p2g(*static_cast<OGrid*>(p), pos, vel, n, dt, gr, density);
  return 0;
}

int og_divergence(void* p) {
  divergence(*static_cast<OGrid*>(p));
  return 0;
}

int og_set_params(void* p, double omega, int local_sweeps, int pre, int post, int coarse_sweeps, double prolong_scale) {
  OGrid& g = *static_cast<OGrid*>(p);
  g.prolong_scale = prolong_scale;
  g.omega = omega;
  g.local_sweeps = local_sweeps;
  g.pre = pre;
  g.post = post;
  g.coarse_sweeps = coarse_sweeps;
  return 0;
}

int og_set_adaptive(void* p, int on, double theta, int max_block_sweeps, double block_reduction) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (max_block_sweeps < 1 || !(block_reduction > 0.0 && block_reduction < 1.0) || theta < 0.0) return 1;
  g.adaptive = on;
  g.theta = theta;
  g.max_block_sweeps = max_block_sweeps;
  g.block_reduction = block_reduction;
  return 0;
}

// per-block sweeps run and flags left active by the last adaptive smoothing of level k
int og_mg_adaptive(void* p, int k, uint8_t* sweeps, uint8_t* act) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (k < 0 || k >= (int)g.mg.size()) return 1;
  const Level& L = g.mg[k];
  if ((int)L.sweeps.size() != L.nb) return 1;
  std::copy(L.sweeps.begin(), L.sweeps.end(), sweeps);
  std::copy(L.act.begin(), L.act.end(), act);
  return 0;
}

int og_mg_build(void* p) {
  OGrid& g = *static_cast<OGrid*>(p);
  mg_build(g);
  mg_refresh(g);
  return 0;
}

int og_num_mg_levels(void* p) { return (int)static_cast<OGrid*>(p)->mg.size(); }

int og_mg_level_info(void* p, int k, int* info) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (k < 0 || k >= (int)g.mg.size()) return 1;
  const Level& L = g.mg[k];
  info[0] = L.D[0];
  info[1] = L.D[1];
  info[2] = L.D[2];
  info[3] = L.nb;
  info[4] = L.ncells;
  info[5] = L.bs;
  info[6] = (int)L.col.size();
  return 0;
}

int og_mg_level_e(void* p, int k, uint8_t* e, uint32_t* coff) {
  OGrid& g = *static_cast<OGrid*>(p);
  const Level& L = g.mg[k];
  for (int b = 0; b < L.nb; ++b) e[b] = (uint8_t)L.e[b];
  std::copy(L.coff.begin(), L.coff.end(), coff);
  return 0;
}

static std::vector<double>* mg_field(Level& L, int which) {
  switch (which) {
    case 0: return &L.x;
    case 1: return &L.b;
    case 2: return &L.r;
    case 4: return &L.diag;
  }
  return nullptr;
}

int og_mg_get(void* p, int k, int which, double* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  Level& L = g.mg[k];
  if (which == 3) {
    for (int c = 0; c < L.ncells; ++c) out[c] = L.liquid[c];
    return 0;
  }
  if (which == 5) {
    for (int c = 0; c < L.ncells; ++c) out[c] = k + 1 < (int)g.mg.size() ? L.parent[c] : -1;
    return 0;
  }
  auto* v = mg_field(L, which);
  if (!v) return 1;
  std::copy(v->begin(), v->end(), out);
  return 0;
}

int og_mg_set(void* p, int k, int which, const double* in) {
  OGrid& g = *static_cast<OGrid*>(p);
  Level& L = g.mg[k];
  auto* v = mg_field(L, which);
  if (!v || which == 4) return 1;
  std::copy(in, in + L.ncells, v->begin());
  return 0;
}

int og_smooth(void* p, int k, int iters, int reverse) {
  OGrid& g = *static_cast<OGrid*>(p);
  smooth(g, g.mg[k], iters, reverse != 0);
  return 0;
}

int og_residual(void* p, int k) {
  OGrid& g = *static_cast<OGrid*>(p);
  residual(g.mg[k]);
  return 0;
}

int og_apply(void* p, int k, const double* x, double* y) {
  OGrid& g = *static_cast<OGrid*>(p);
  apply(g.mg[k], x, y);
  return 0;
}

int og_restrict(void* p, int k) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (k + 1 >= (int)g.mg.size()) return 1;
  restrict_to(g, k);
  return 0;
}

int og_prolong(void* p, int k) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (k + 1 >= (int)g.mg.size()) return 1;
  prolong_from(g, k);
  return 0;
}

int og_vcycle(void* p, int k) {
  OGrid& g = *static_cast<OGrid*>(p);
  vcycle(g, k);
  return 0;
}

int og_export_matrix(void* p, int k, int* rowptr, int* col, double* val, double* diag) {
  OGrid& g = *static_cast<OGrid*>(p);
  const Level& L = g.mg[k];
  std::copy(L.rowptr.begin(), L.rowptr.end(), rowptr);
  std::copy(L.col.begin(), L.col.end(), col);
  std::copy(L.val.begin(), L.val.end(), val);
  std::copy(L.diag.begin(), L.diag.end(), diag);
  return 0;
}

int og_solve(void* p, int method, double tol, int max_iters, int* iters, double* rel) {
  return solve(*static_cast<OGrid*>(p), method, tol, max_iters, iters, rel);
}

int og_project(void* p) {
  OGrid& g = *static_cast<OGrid*>(p);
  project(g);
  to_fine(g);
  return 0;
}

int og_g2p_advect(void* p, const float* pos, const float* vel, int64_t n, double alpha, double dt, double* pos_out,
                  double* vel_out) {
  g2p_advect(*static_cast<OGrid*>(p), pos, vel, n, alpha, dt, pos_out, vel_out);
  return 0;
}

int og_sample(void* p, int field, const double* units, int64_t n, double* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  const std::vector<double>& F = field == 0 ? g.Fnew : g.Fdel;
  if (F.empty()) return 1;
  for (int64_t i = 0; i < n; ++i) sample(g, F, units + 3 * i, out + 3 * i);
  return 0;
}

static std::vector<double>* face_field(OGrid& g, int which, int a) {
  switch (which) {
    case 0: return &g.Macc[a];
    case 1: return &g.macc[a];
    case 2: return &g.uold[a];
    case 3: return &g.ustar[a];
    case 4: return &g.unew[a];
    case 5: return &g.mt[a];
    case 6: return &g.beta[a];
    case 7: return &g.ubar_new[a];
    case 8: return &g.ubar_old[a];
    case 9: return &g.ubar_star[a];
    case 10: return &g.Mbar[a];
    case 11: return &g.mbar[a];
  }
  return nullptr;
}

int og_get_face(void* p, int which, int a, double* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  auto* v = face_field(g, which, a);
  if (!v) return 1;
  std::copy(v->begin(), v->end(), out);
  return 0;
}

int og_set_face(void* p, int which, int a, const double* in) {
  OGrid& g = *static_cast<OGrid*>(p);
  auto* v = face_field(g, which, a);
  if (!v) return 1;
  std::copy(in, in + v->size(), v->begin());
  return 0;
}

int og_get_cell(void* p, int which, double* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  for (int64_t c = 0; c < g.ncells; ++c) {
    switch (which) {
      case 0: out[c] = g.p[c]; break;
      case 1: out[c] = g.rhs[c]; break;
      case 2: out[c] = g.liquid[c]; break;
      case 3: out[c] = g.pcount[c]; break;
      case 4: out[c] = g.pbar[c]; break;
      case 5: out[c] = g.rhsbar[c]; break;
      default: return 1;
    }
  }
  return 0;
}

int og_set_cell(void* p, int which, const double* in) {
  OGrid& g = *static_cast<OGrid*>(p);
  for (int64_t c = 0; c < g.ncells; ++c) {
    switch (which) {
      case 0: g.p[c] = in[c]; break;
      case 1: g.rhs[c] = in[c]; break;
      case 2: g.liquid[c] = in[c] != 0.0; break;
      case 4: g.pbar[c] = in[c]; break;
      case 5: g.rhsbar[c] = in[c]; break;
      default: return 1;
    }
  }
  return 0;
}

// The whole FLIP step on FP32 particles: topology from the particles, P2G with
// the body force, divergence, pressure solve, projection, G2P + RK3.  Outputs
// are FP64; the caller rounds them to FP32 for the next step.  density (may
// be null): relative density per particle (input order), beta = 1 / rho.
int og_step(void* p, const float* pos, const float* vel, int64_t n, double dt, double gx, double gy, double gz,
            double alpha, int method, double tol, int max_iters, double* pos_out, double* vel_out, int* iters,
            double* rel, double* stage_ms, const double* density) {
  OGrid& g = *static_cast<OGrid*>(p);
  auto now = [] { return omp_get_wtime(); };
  double t0 = now();
  std::vector<int32_t> key(n);
  std::vector<uint32_t> cnt(g.nb);
  og_bin(p, pos, n, key.data(), cnt.data());
  std::vector<uint8_t> lv(g.nb);
  levels_from_counts(g, cnt.data(), lv.data());
  set_topology(g, lv.data());
  double t1 = now();
  double gr[3] = {gx, gy, gz};
  p2g(g, pos, vel, n, dt, gr, density);
  divergence(g);
  double t2 = now();
  mg_build(g);
  mg_refresh(g);
  std::fill(g.p.begin(), g.p.end(), 0.0);
  int st = solve(g, method, tol, max_iters, iters, rel);
  double t3 = now();
  project(g);
  to_fine(g);
  g2p_advect(g, pos, vel, n, alpha, dt, pos_out, vel_out);
  double t4 = now();
  if (stage_ms) {
    stage_ms[0] = (t1 - t0) * 1e3;
    stage_ms[1] = (t2 - t1) * 1e3;
    stage_ms[2] = (t3 - t2) * 1e3;
    stage_ms[3] = (t4 - t3) * 1e3;
  }
  return st;
}

int og_num_threads(void) { return omp_get_max_threads(); }
int og_set_threads(int n) {
  if (n < 1) return 1;
  omp_set_num_threads(n);
  return 0;
}

static std::vector<double>* fine_field(OGrid& g, int which) {
  switch (which) {
    case 0: return &g.Fnew;
    case 1: return &g.Fdel;
    case 2: return &g.Fbar_nd;
    case 3: return &g.Fbar_d;
    case 4: return &g.FMbar;
    case 5: return &g.Fmbar;
    case 6: return &g.FM;
    case 7: return &g.Fm;
  }
  return nullptr;
}

int og_get_fine(void* p, int which, double* out) {
  OGrid& g = *static_cast<OGrid*>(p);
  auto* v = fine_field(g, which);
  if (!v) return 1;
  std::fill(out, out + (size_t)g.nb * 3 * 512, 0.0);
  std::copy(v->begin(), v->end(), out);
  return 0;
}

int og_set_fine(void* p, int which, const double* in) {
  OGrid& g = *static_cast<OGrid*>(p);
  auto* v = fine_field(g, which);
  if (!v) return 1;
  v->assign(in, in + (size_t)g.nb * 3 * 512);
  return 0;
}

// Pi and Pi^T between per-slot fields (which: face field index) and fine
// lattice fields, for transpose tests.
int og_pi(void* p, int face_which, int fine_which) {
  OGrid& g = *static_cast<OGrid*>(p);
  std::vector<double> S[3];
  for (int a = 0; a < 3; ++a) S[a] = *face_field(g, face_which, a);
  pi_forward(g, S, *fine_field(g, fine_which));
  return 0;
}

int og_pi_transpose(void* p, int fine_which, int face_which) {
  OGrid& g = *static_cast<OGrid*>(p);
  std::vector<double> S[3];
  for (int a = 0; a < 3; ++a) S[a].assign(g.nfaces[a], 0.0);
  pi_transpose(g, *fine_field(g, fine_which), S, 1.0);
  for (int a = 0; a < 3; ++a) *face_field(g, face_which, a) = S[a];
  return 0;
}

// One FP64 step on FP64 particles (the finite-difference and adjoint
// rollouts).  levels_in / liquid_in, when given, replace the topology and the
// liquid mask the particles would produce (frozen topology); levels_out /
// liquid_out receive the ones used.  Intermediates stay in the grid for
// og_step_adjoint.
int og_step64(void* p, const double* pos, const double* vel, int64_t n, double dt, double gx, double gy, double gz,
              double alpha, int method, double tol, int max_iters, const uint8_t* levels_in,
              const uint8_t* liquid_in, uint8_t* levels_out, uint8_t* liquid_out, double* pos_out, double* vel_out,
              int* iters, double* rel) {
  OGrid& g = *static_cast<OGrid*>(p);
  std::vector<uint8_t> lv(g.nb);
  if (levels_in) {
    std::copy(levels_in, levels_in + g.nb, lv.begin());
  } else {
    std::vector<uint32_t> cnt(g.nb, 0);
    for (int64_t i = 0; i < n; ++i) {
      int gc[3];
      fine_cell(g, pos + 3 * i, gc);
      cnt[block_id(g.D, gc[0] >> LOGB, gc[1] >> LOGB, gc[2] >> LOGB)]++;
    }
    levels_from_counts(g, cnt.data(), lv.data());
  }
  set_topology(g, lv.data());
  double gr[3] = {gx, gy, gz};
  p2g(g, pos, vel, n, dt, gr);
  if (liquid_in)
    for (int64_t c = 0; c < g.ncells; ++c) g.liquid[c] = liquid_in[c] != 0;
  divergence(g);
  mg_build(g);
  mg_refresh(g);
  std::fill(g.p.begin(), g.p.end(), 0.0);
  int st = solve(g, method, tol, max_iters, iters, rel);
  project(g);
  to_fine(g);
  g2p_advect(g, pos, vel, n, alpha, dt, pos_out, vel_out);
  if (levels_out) std::copy(g.level.begin(), g.level.end(), levels_out);
  if (liquid_out) std::copy(g.liquid.begin(), g.liquid.end(), liquid_out);
  return st;
}

// Adjoint stages.  fp64 selects the particle type (0: float, 1: double).
int og_adj_g2p(void* p, int fp64, const void* pos, const void* vel, int64_t n, double alpha, double dt,
               const double* xbo, const double* vbo, double* xbi, double* vbi) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (g.Fnew.empty()) return 1;
  if (fp64)
    g2p_advect_adjoint(g, (const double*)pos, (const double*)vel, n, alpha, dt, xbo, vbo, xbi, vbi);
  else
    g2p_advect_adjoint(g, (const float*)pos, (const float*)vel, n, alpha, dt, xbo, vbo, xbi, vbi);
  return 0;
}

int og_adj_project(void* p, double tol, int max_iters, int* iters, double* rel) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (g.mg.empty()) return 1;
  project_adjoint(g);
  int st = solve_adjoint(g, tol, max_iters, iters, rel);
  divergence_adjoint(g);
  return st;
}

int og_adj_p2g(void* p, int fp64, const void* pos, const void* vel, int64_t n, double* xbar, double* vbar) {
  OGrid& g = *static_cast<OGrid*>(p);
  if (fp64)
    p2g_adjoint(g, (const double*)pos, (const double*)vel, n, xbar, vbar);
  else
    p2g_adjoint(g, (const float*)pos, (const float*)vel, n, xbar, vbar);
  return 0;
}

// Whole-step adjoint on the intermediates of the last og_step / og_step64
// with these particles: xbar, vbar hold dL/d(x', v') on entry and dL/d(x, v)
// on return.
int og_step_adjoint(void* p, int fp64, const void* pos, const void* vel, int64_t n, double alpha, double dt,
                    double tol, int max_iters, double* xbar, double* vbar, int* iters, double* rel) {
  std::vector<double> xo(xbar, xbar + 3 * n), vo(vbar, vbar + 3 * n);
  int st = og_adj_g2p(p, fp64, pos, vel, n, alpha, dt, xo.data(), vo.data(), xbar, vbar);
  if (st) return st;
  st = og_adj_project(p, tol, max_iters, iters, rel);
  og_adj_p2g(p, fp64, pos, vel, n, xbar, vbar);
  return st;
}

}  // extern "C"
