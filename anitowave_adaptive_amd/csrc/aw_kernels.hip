// gfx950 kernels of the MSBG FLIP step.
//
// Layout (per native block b, resolution r = 8 >> level[b]):
//   cells   [coff[b], coff[b+1])   x + r (y + r z)                     (MSBG.cuh:165-167)
//   faces   [foff[b], foff[b+1])   normal axis r+1 slots, others r     (MSBG.cuh:169-179)
// The three face-offset arrays coincide because every axis holds (r+1) r^2
// slots per block.  Fine-lattice buffers are block-major: [block][field][axis][512].
//
// All kernels are HBM-bound gathers/scatters and stencils; none uses MFMA.

#include "aw_internal.h"

namespace aw {
namespace {

constexpr int TPB = 256;

__device__ __forceinline__ int face_idx(int a, int x, int y, int z, int r) {
  return a == 0 ? x + (r + 1) * (y + r * z) : (a == 1 ? x + r * (y + (r + 1) * z) : x + r * (y + r * z));
}
__device__ __forceinline__ int face_idx(int a, const int* c, int r) { return face_idx(a, c[0], c[1], c[2], r); }
__device__ __forceinline__ int cell_idx(const int* c, int r) { return c[0] + r * (c[1] + r * c[2]); }
__device__ __forceinline__ void face_decode(int a, int k, int r, int* c) {
  if (a == 0) {
    c[0] = k % (r + 1);
    c[1] = (k / (r + 1)) % r;
    c[2] = k / ((r + 1) * r);
  } else if (a == 1) {
    c[0] = k % r;
    c[1] = (k / r) % (r + 1);
    c[2] = k / (r * (r + 1));
  } else {
    c[0] = k % r;
    c[1] = (k / r) % r;
    c[2] = k / (r * r);
  }
}
__device__ __forceinline__ void block_coords(const int* D, int b, int* c) {
  c[0] = b % D[0];
  c[1] = (b / D[0]) % D[1];
  c[2] = b / (D[0] * D[1]);
}
__device__ __forceinline__ int block_id(const int* D, const int* c) { return c[0] + D[0] * (c[1] + D[1] * c[2]); }
__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// padded fine lattice of one block for axis a: normal index [0, 8], transverse [-1, 8]
__device__ __forceinline__ int pad_idx(int a, const int* l) {
  const int n0 = a == 0 ? 9 : 10, n1 = a == 1 ? 9 : 10;
  const int o0 = a == 0 ? 0 : 1, o1 = a == 1 ? 0 : 1, o2 = a == 2 ? 0 : 1;
  return (l[0] + o0) + n0 * ((l[1] + o1) + n1 * (l[2] + o2));
}

__device__ __forceinline__ float norm_u(float M, float m) { return m > 1e-6f ? M / m : 0.f; }

// Pi: the native slots (and weights) of a block of level e that define its
// own fine node l on axis a.  Normal coordinate l / s, transverse
// (l + 1/2) / s - 1/2 clamped to [0, r-1]; trilinear.
__device__ __forceinline__ int pi_weights(int a, int e, const int* l, int* slot, float* w) {
  const int s = 1 << e, r = BASE >> e;
  int i0[3];
  float t[3];
  const float inv_s = 1.0f / (float)s;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d == a) {
      i0[d] = l[d] >> e;
      t[d] = (float)(l[d] & (s - 1)) * inv_s;
    } else {
      float eta = ((float)l[d] + 0.5f) * inv_s - 0.5f;
      eta = fminf(fmaxf(eta, 0.f), (float)(r - 1));
      int j = (int)floorf(eta);
      j = r >= 2 ? min(j, r - 2) : 0;
      i0[d] = j;
      t[d] = eta - (float)j;
    }
  }
  int n = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
    const float ww = (dx ? t[0] : 1.f - t[0]) * (dy ? t[1] : 1.f - t[1]) * (dz ? t[2] : 1.f - t[2]);
    if (ww == 0.f) continue;
    slot[n] = face_idx(a, i0[0] + dx, i0[1] + dy, i0[2] + dz, r);
    w[n] = ww;
    ++n;
  }
  return n;
}

template <typename T>
__device__ __forceinline__ T block_reduce_sum(T v) {
  __shared__ T red[TPB / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < TPB / 64; ++i) s += red[i];
  return s;
}

// ------------------------------------------------------------- fills
template <typename T>
__global__ void fill_kernel(T* p, size_t n, T v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

inline int grid_for(size_t n) {
  size_t g = (n + TPB - 1) / TPB;
  return (int)(g < 65536 ? (g ? g : 1) : 65536);
}

// ------------------------------------------------------------- topology
// Block of each particle from its exact fine cell (worldToBlockCoord,
// MSBG.cuh:151-159) and the per-block particle count.
__global__ void bin_kernel(GridDev g, const float* __restrict__ pos, int64_t n, int32_t* __restrict__ key,
                           uint32_t* __restrict__ count) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float u = fminf(fmaxf(pos[3 * i + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
    c[a] = ((int)u) >> LOGB;
  }
  int b = block_id(g.D, c);
  key[i] = b;
  atomicAdd(&count[b], 1u);
}

// occupied block with an unoccupied in-domain 26-neighbour
__global__ void surface_kernel(int D0, int D1, int D2, const uint32_t* __restrict__ count, uint8_t* __restrict__ surf) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int D[3] = {D0, D1, D2};
  if (b >= D0 * D1 * D2) return;
  uint8_t s = 0;
  if (count[b]) {
    int c[3];
    block_coords(D, b, c);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
          if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= D0 || q[1] >= D1 || q[2] >= D2) continue;
          if (!count[block_id(D, q)]) s = 1;
        }
  }
  surf[b] = s;
}

struct Bands {
  int v[4];
};

// first level whose band (Chebyshev block distance to a surface block) holds the block
__global__ void band_kernel(int D0, int D1, int D2, int L, Bands band, int bmax, const uint8_t* __restrict__ surf,
                            uint8_t* __restrict__ level) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int D[3] = {D0, D1, D2};
  if (b >= D0 * D1 * D2) return;
  int c[3];
  block_coords(D, b, c);
  int dist = 1 << 20;
  for (int z = max(0, c[2] - bmax); z <= min(D2 - 1, c[2] + bmax); ++z)
    for (int y = max(0, c[1] - bmax); y <= min(D1 - 1, c[1] + bmax); ++y)
      for (int x = max(0, c[0] - bmax); x <= min(D0 - 1, c[0] + bmax); ++x) {
        int q[3] = {x, y, z};
        if (surf[block_id(D, q)]) dist = min(dist, max(abs(x - c[0]), max(abs(y - c[1]), abs(z - c[2]))));
      }
  int l = L - 1;
  for (int k = 0; k < L - 1; ++k)
    if (dist <= band.v[k]) {
      l = k;
      break;
    }
  level[b] = (uint8_t)l;
}

// one Jacobi pass of 2:1 grading: l = min(l, min_n l_n + 1) over 26 neighbours
// (enforceGrading, MSBG.cuh:316-351, double-buffered)
__global__ void grade_kernel(int D0, int D1, int D2, const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int D[3] = {D0, D1, D2};
  if (b >= D0 * D1 * D2) return;
  int c[3];
  block_coords(D, b, c);
  int l = in[b];
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= D0 || q[1] >= D1 || q[2] >= D2) continue;
        l = min(l, (int)in[block_id(D, q)] + 1);
      }
  out[b] = (uint8_t)l;
}

// Single-work-group exclusive scan (the reference's thrust::exclusive_scan,
// main.cu:270-273).  mode 0: values; 1: cells r^3 from levels; 2: face
// slots (r+1) r^2 from levels.  out[n] receives the total.
constexpr int SCAN_T = 1024;
__global__ __launch_bounds__(SCAN_T) void scan_kernel(const uint32_t* __restrict__ in, const uint8_t* __restrict__ lv,
                                                      int mode, uint32_t* __restrict__ out, int n) {
  __shared__ uint32_t sh[SCAN_T];
  uint32_t carry = 0;
  const int t = threadIdx.x;
  for (int base = 0; base < n; base += SCAN_T * 4) {
    uint32_t v[4];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int i = base + t * 4 + k;
      uint32_t x = 0;
      if (i < n) {
        if (mode == 0) {
          x = in[i];
        } else {
          int r = BASE >> lv[i];
          x = mode == 1 ? (uint32_t)(r * r * r) : (uint32_t)((r + 1) * r * r);
        }
      }
      v[k] = x;
      local += x;
    }
    sh[t] = local;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
      uint32_t y = t >= off ? sh[t - off] : 0;
      __syncthreads();
      sh[t] += y;
      __syncthreads();
    }
    uint32_t excl = carry + sh[t] - local;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int i = base + t * 4 + k;
      if (i < n) out[i] = excl;
      excl += v[k];
    }
    carry += sh[SCAN_T - 1];
    __syncthreads();
  }
  if (t == 0) out[n] = carry;
}

__global__ void scatter_kernel(const float* __restrict__ pos, const float* __restrict__ vel,
                               const int32_t* __restrict__ ids, const int32_t* __restrict__ key,
                               uint32_t* __restrict__ cursor, int64_t n, float* __restrict__ spos,
                               float* __restrict__ svel, int32_t* __restrict__ sids) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t s = atomicAdd(&cursor[key[i]], 1u);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    spos[3 * (int64_t)s + a] = pos[3 * i + a];
    svel[3 * (int64_t)s + a] = vel[3 * i + a];
  }
  if (ids) sids[s] = ids[i];
}

// beta = 1 on interior slots, 0 on tank-wall slots
__global__ void init_beta_kernel(LevelDev L, float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz) {
  const int B = blockIdx.x;
  const int r = L.bs >> L.e[B];
  int Bc[3];
  block_coords(L.D, B, Bc);
  const int nf = (r + 1) * r * r;
  float* dst[3] = {bx, by, bz};
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    int c[3];
    face_decode(a, f, r, c);
    const bool wall = (c[a] == 0 && Bc[a] == 0) || (c[a] == r && Bc[a] == L.D[a] - 1);
    dst[a][L.foff[a][B] + f] = wall ? 0.f : 1.f;
  }
}

// ------------------------------------------------------------- P2G
// One work-group per block: deposit the block's particles on the padded
// virtual fine lattice in LDS (trilinear MAC weights; the reference's
// per-particle global atomics, P2G.cuh:16-135, become LDS atomics), count
// particles per native cell, then store the padded lattice to the block's
// slab with plain stores.  No global atomics, no cross-block races.
__global__ __launch_bounds__(TPB) void p2g_deposit_kernel(GridDev g, const float* __restrict__ spos,
                                                          const float* __restrict__ svel,
                                                          const uint32_t* __restrict__ pcount,
                                                          const uint32_t* __restrict__ poff, float* __restrict__ slab,
                                                          uint8_t* __restrict__ liquid) {
  __shared__ float acc[SLAB];
  __shared__ uint32_t ccnt[512];
  const int B = blockIdx.x, tid = threadIdx.x;
  const uint32_t cnt = pcount[B];
  const int e = g.level[B], r = BASE >> e, nc = r * r * r;
  const uint32_t co = g.coff[B];
  if (cnt == 0) {
    for (int k = tid; k < nc; k += TPB) liquid[co + k] = 0;
    return;
  }
  for (int k = tid; k < SLAB; k += TPB) acc[k] = 0.f;
  for (int k = tid; k < 512; k += TPB) ccnt[k] = 0;
  __syncthreads();
  int Bc[3];
  block_coords(g.D, B, Bc);
  const uint32_t p0 = poff[B];
  for (uint32_t p = p0 + tid; p < p0 + cnt; p += TPB) {
    float u[3], v[3];
    int gc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      u[a] = fminf(fmaxf(spos[3 * (size_t)p + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
      v[a] = svel[3 * (size_t)p + a];
      gc[a] = (int)u[a];
    }
    // a particle outside this block (unsorted input) is skipped: it would
    // index outside the block's padded lattice
    if ((gc[0] >> LOGB) != Bc[0] || (gc[1] >> LOGB) != Bc[1] || (gc[2] >> LOGB) != Bc[2]) continue;
    int lc[3] = {(gc[0] - BASE * Bc[0]) >> e, (gc[1] - BASE * Bc[1]) >> e, (gc[2] - BASE * Bc[2]) >> e};
    atomicAdd(&ccnt[cell_idx(lc, r)], 1u);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      int i0[3];
      float t[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float c = d == a ? u[d] : u[d] - 0.5f;
        const float fl = floorf(c);
        i0[d] = (int)fl;
        t[d] = c - fl;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int dd[3] = {k & 1, (k >> 1) & 1, k >> 2};
        int l[3];
        bool skip = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          int gn = i0[d] + dd[d];
          if (d == a) {
            if (gn >= g.N[d]) skip = true;
          } else {
            gn = iclamp(gn, 0, g.N[d] - 1);
          }
          l[d] = gn - BASE * Bc[d];
        }
        if (skip) continue;
        const float w = (dd[0] ? t[0] : 1.f - t[0]) * (dd[1] ? t[1] : 1.f - t[1]) * (dd[2] ? t[2] : 1.f - t[2]);
        const int pi = a * PAD + pad_idx(a, l);
        atomicAdd(&acc[pi], w * v[a]);
        atomicAdd(&acc[3 * PAD + pi], w);
      }
    }
  }
  __syncthreads();
  float* dst = slab + (size_t)B * SLAB;
  for (int k = tid; k < SLAB; k += TPB) dst[k] = acc[k];
  for (int k = tid; k < nc; k += TPB) liquid[co + k] = ccnt[k] > 0;
}

// One work-group per block: collect the block's own fine nodes from its slab
// and the halos of the occupied neighbours' slabs, then apply Pi^T into the
// native accumulators (M, m) of every slot of the block.
__global__ __launch_bounds__(TPB) void p2g_gather_kernel(GridDev g, const uint32_t* __restrict__ pcount,
                                                         const float* __restrict__ slab, float* Mx, float* My,
                                                         float* Mz, float* mx, float* my, float* mz) {
  __shared__ float fM[3 * FINE], fm[3 * FINE];
  __shared__ float nM[3 * 576], nm[3 * 576];
  const int B = blockIdx.x, tid = threadIdx.x;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  int Bc[3];
  block_coords(g.D, B, Bc);
  for (int k = tid; k < 3 * FINE; k += TPB) {
    const int a = k >> 9, li = k & 511;
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = (l[d] == 0 && Bc[d] > 0) ? -1 : 0;
      hi[d] = (d != a && l[d] == 7 && Bc[d] < g.D[d] - 1) ? 1 : 0;
    }
    float sM = 0.f, sm = 0.f;
    for (int dz = lo[2]; dz <= hi[2]; ++dz)
      for (int dy = lo[1]; dy <= hi[1]; ++dy)
        for (int dx = lo[0]; dx <= hi[0]; ++dx) {
          const int q[3] = {Bc[0] + dx, Bc[1] + dy, Bc[2] + dz};
          const int N = block_id(g.D, q);
          if (!pcount[N]) continue;
          const int lp[3] = {l[0] - BASE * dx, l[1] - BASE * dy, l[2] - BASE * dz};
          const float* src = slab + (size_t)N * SLAB + a * PAD + pad_idx(a, lp);
          sM += src[0];
          sm += src[3 * PAD];
        }
    fM[k] = sM;
    fm[k] = sm;
  }
  for (int k = tid; k < 3 * 576; k += TPB) {
    nM[k] = 0.f;
    nm[k] = 0.f;
  }
  __syncthreads();
  for (int k = tid; k < 3 * FINE; k += TPB) {
    const float M = fM[k], m = fm[k];
    if (M == 0.f && m == 0.f) continue;
    const int a = k >> 9, li = k & 511;
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    int slot[8];
    float w[8];
    const int n = pi_weights(a, e, l, slot, w);
    for (int j = 0; j < n; ++j) {
      atomicAdd(&nM[a * 576 + slot[j]], w[j] * M);
      atomicAdd(&nm[a * 576 + slot[j]], w[j] * m);
    }
  }
  __syncthreads();
  float* dM[3] = {Mx, My, Mz};
  float* dm[3] = {mx, my, mz};
  for (int k = tid; k < 3 * nf; k += TPB) {
    const int a = k / nf, f = k - a * nf;
    dM[a][g.foff[a][B] + f] = nM[a * 576 + f];
    dm[a][g.foff[a][B] + f] = nm[a * 576 + f];
  }
}

// Fold the duplicate / coarse-fine boundary slots, normalise u = M / m
// (normalizeVelocity, P2G.cuh:139-171), add the body force on faces with mass.
struct FacePtrs {
  float* p[3];
};

__global__ __launch_bounds__(TPB) void fold_kernel(GridDev g, FacePtrs Macc, FacePtrs macc, FacePtrs uold,
                                                   FacePtrs ustar, FacePtrs mass, float dt, float g0, float g1,
                                                   float g2) {
  const int B = blockIdx.x;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  const float grav[3] = {g0, g1, g2};
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int sl[3];
    face_decode(a, f, r, sl);
    const size_t id = g.foff[a][B] + f;
    const float M = Macc.p[a][id], m = macc.p[a][id];
    float uo, us, mt;
    if (sl[a] > 0 && sl[a] < r) {
      mt = m;
      uo = norm_u(M, m);
      us = uo + (m > 1e-6f ? dt * grav[a] : 0.f);
    } else {
      const int side = sl[a] == r;
      const int N = g.nbr[B * 6 + 2 * a + side];
      if (N < 0) {
        uo = us = mt = 0.f;
      } else {
        const int en = g.level[N], rn = BASE >> en;
        int ps[3] = {sl[0], sl[1], sl[2]};
        ps[a] = side ? 0 : rn;
        if (en == e) {
          const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
          mt = m + macc.p[a][pid];
          uo = norm_u(M + Macc.p[a][pid], mt);
          us = uo + (mt > 1e-6f ? dt * grav[a] : 0.f);
        } else if (en == e + 1) {
          ps[t1] = sl[t1] >> 1;
          ps[t2] = sl[t2] >> 1;
          const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
          mt = m + 0.25f * macc.p[a][pid];
          uo = norm_u(M + 0.25f * Macc.p[a][pid], mt);
          us = uo + (mt > 1e-6f ? dt * grav[a] : 0.f);
        } else {
          uo = us = mt = 0.f;
          for (int d1 = 0; d1 < 2; ++d1)
            for (int d2 = 0; d2 < 2; ++d2) {
              ps[t1] = 2 * sl[t1] + d1;
              ps[t2] = 2 * sl[t2] + d2;
              const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
              const float ms = macc.p[a][pid] + 0.25f * m;
              const float u0 = norm_u(Macc.p[a][pid] + 0.25f * M, ms);
              uo += 0.25f * u0;
              us += 0.25f * (u0 + (ms > 1e-6f ? dt * grav[a] : 0.f));
              mt += ms;
            }
        }
      }
    }
    uold.p[a][id] = uo;
    ustar.p[a][id] = us;
    mass.p[a][id] = mt;
  }
}

// rhs = -s^2 sum_a (u[+] - u[-]) on liquid cells
__global__ __launch_bounds__(TPB) void divergence_kernel(GridDev g, const uint8_t* __restrict__ liquid, FacePtrs u,
                                                         float* __restrict__ rhs) {
  const int B = blockIdx.x;
  const int e = g.level[B], r = BASE >> e, nc = r * r * r;
  const float s = (float)(1 << e);
  for (int k = threadIdx.x; k < nc; k += blockDim.x) {
    const int c[3] = {k % r, (k / r) % r, k / (r * r)};
    float d = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      int cp[3] = {c[0], c[1], c[2]};
      cp[a] += 1;
      d += u.p[a][g.foff[a][B] + face_idx(a, cp, r)] - u.p[a][g.foff[a][B] + face_idx(a, c, r)];
    }
    const size_t id = g.coff[B] + k;
    rhs[id] = liquid[id] ? -(s * s) * d : 0.f;
  }
}

// u_new = u* - beta (p+ - p-) / d on every slot of the block
__global__ __launch_bounds__(TPB) void project_kernel(GridDev g, const float* __restrict__ p, FacePtrs ustar,
                                                      FacePtrs beta, FacePtrs unew) {
  const int B = blockIdx.x;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  const float s = (float)(1 << e);
  const uint32_t co = g.coff[B];
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int sl[3];
    face_decode(a, f, r, sl);
    const size_t id = g.foff[a][B] + f;
    const float us = ustar.p[a][id], bt = beta.p[a][id];
    float out;
    if (sl[a] > 0 && sl[a] < r) {
      int cm[3] = {sl[0], sl[1], sl[2]};
      cm[a] -= 1;
      out = us - bt * (p[co + cell_idx(sl, r)] - p[co + cell_idx(cm, r)]) / s;
    } else {
      const int side = sl[a] == r;
      const int N = g.nbr[B * 6 + 2 * a + side];
      if (N < 0) {
        out = 0.f;
      } else {
        const int en = g.level[N], rn = BASE >> en;
        int own[3] = {sl[0], sl[1], sl[2]};
        own[a] = side ? r - 1 : 0;
        const float po = p[co + cell_idx(own, r)];
        int q[3] = {sl[0], sl[1], sl[2]};
        q[a] = side ? 0 : rn - 1;
        if (en == e) {
          const float pn = p[g.coff[N] + cell_idx(q, rn)];
          out = us - bt * (side ? pn - po : po - pn) / s;
        } else if (en == e + 1) {
          q[t1] = sl[t1] >> 1;
          q[t2] = sl[t2] >> 1;
          const float pn = p[g.coff[N] + cell_idx(q, rn)];
          out = us - bt * (side ? pn - po : po - pn) / (1.5f * s);
        } else {
          const float sf = 0.5f * s;
          out = 0.f;
          int ps[3];
          ps[a] = side ? 0 : rn;
          for (int d1 = 0; d1 < 2; ++d1)
            for (int d2 = 0; d2 < 2; ++d2) {
              q[t1] = ps[t1] = 2 * sl[t1] + d1;
              q[t2] = ps[t2] = 2 * sl[t2] + d2;
              const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
              const float pn = p[g.coff[N] + cell_idx(q, rn)];
              out += 0.25f * (ustar.p[a][pid] - beta.p[a][pid] * (side ? pn - po : po - pn) / (1.5f * sf));
            }
        }
      }
    }
    unew.p[a][id] = out;
  }
}

// Pi of u_new and of (u_new - u_old) on the block's own fine nodes
__global__ __launch_bounds__(TPB) void to_fine_kernel(GridDev g, FacePtrs unew, FacePtrs uold, float* __restrict__ fine) {
  const int B = blockIdx.x;
  const int e = g.level[B];
  float* dst = fine + (size_t)B * SLAB;
  for (int k = threadIdx.x; k < 3 * FINE; k += blockDim.x) {
    const int a = k >> 9, li = k & 511;
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    int slot[8];
    float w[8];
    const int n = pi_weights(a, e, l, slot, w);
    float vn = 0.f, vd = 0.f;
    for (int j = 0; j < n; ++j) {
      const size_t id = g.foff[a][B] + slot[j];
      const float un = unew.p[a][id];
      vn += w[j] * un;
      vd += w[j] * (un - uold.p[a][id]);
    }
    dst[k] = vn;
    dst[3 * FINE + k] = vd;
  }
}

__device__ __forceinline__ float fine_at(const GridDev& g, const float* __restrict__ F, int field, int a, int* q) {
  if (q[a] < 0 || q[a] >= g.N[a]) return 0.f;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d != a) q[d] = iclamp(q[d], 0, g.N[d] - 1);
  const int B = (q[0] >> LOGB) + g.D[0] * ((q[1] >> LOGB) + g.D[1] * (q[2] >> LOGB));
  const int l = (q[0] & 7) + 8 * ((q[1] & 7) + 8 * (q[2] & 7));
  return F[(size_t)B * SLAB + field * 3 * FINE + a * FINE + l];
}

__device__ __forceinline__ float sample_comp(const GridDev& g, const float* __restrict__ F, int field, int a,
                                             const float* u) {
  int i0[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = d == a ? u[d] : u[d] - 0.5f;
    const float fl = floorf(c);
    i0[d] = (int)fl;
    t[d] = c - fl;
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int dd[3] = {k & 1, (k >> 1) & 1, k >> 2};
    int q[3] = {i0[0] + dd[0], i0[1] + dd[1], i0[2] + dd[2]};
    const float w = (dd[0] ? t[0] : 1.f - t[0]) * (dd[1] ? t[1] : 1.f - t[1]) * (dd[2] ? t[2] : 1.f - t[2]);
    s += w * fine_at(g, F, field, a, q);
  }
  return s;
}

// FLIP/PIC blend (G2P.cuh:87-128) and Ralston RK3 (G2P.cuh:131-173) per particle
__global__ __launch_bounds__(TPB) void g2p_advect_kernel(GridDev g, const float* __restrict__ F,
                                                         const float* __restrict__ spos,
                                                         const float* __restrict__ svel, int64_t n, float alpha,
                                                         float dt, float* __restrict__ pos_out,
                                                         float* __restrict__ vel_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[3], u[3], v1[3], vd[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = spos[3 * i + a];
    u[a] = fminf(fmaxf(x[a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    v1[a] = sample_comp(g, F, 0, a, u);
    vd[a] = sample_comp(g, F, 1, a, u);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) vel_out[3 * i + a] = alpha * (svel[3 * i + a] + vd[a]) + (1.f - alpha) * v1[a];
  float xx[3], v2[3], v3[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    xx[a] = x[a] + 0.5f * dt * v1[a];
    u[a] = fminf(fmaxf(xx[a] * g.inv_h, -1.f), (float)g.N[a] + 1.f);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) v2[a] = sample_comp(g, F, 0, a, u);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    xx[a] = x[a] + 0.75f * dt * v2[a];
    u[a] = fminf(fmaxf(xx[a] * g.inv_h, -1.f), (float)g.N[a] + 1.f);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) v3[a] = sample_comp(g, F, 0, a, u);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float xn = x[a] + dt * ((2.f / 9.f) * v1[a] + (3.f / 9.f) * v2[a] + (4.f / 9.f) * v3[a]);
    const float lo = 0.25f * g.h, hi = (float)g.N[a] * g.h - 0.25f * g.h;
    pos_out[3 * i + a] = fminf(fmaxf(xn, lo), hi);
  }
}

// ------------------------------------------------------------- solver
struct CellRef {
  int B, r, lr, q, li;
  int l[3];
  uint32_t c;
};

__device__ __forceinline__ bool decode(const LevelDev& L, const WG& w, const uint32_t* __restrict__ blist, int i,
                                       CellRef& o) {
  const int sh = 3 * w.lr;
  if (i >= (w.count << sh)) return false;
  o.q = i >> sh;
  o.li = i & ((1 << sh) - 1);
  o.B = blist[w.start + o.q];
  o.lr = w.lr;
  o.r = 1 << w.lr;
  o.l[0] = o.li & (o.r - 1);
  o.l[1] = (o.li >> w.lr) & (o.r - 1);
  o.l[2] = o.li >> (2 * w.lr);
  o.c = L.coff[o.B] + o.li;
  return true;
}

// 7-point adaptive Poisson stencil of one cell: diag = sum of face
// coefficients, off = sum of coefficient * neighbour.  Coefficient beta * s
// between equal cells (s = cell edge in fine units), beta * s_f / 1.5 across a
// coarse-fine interface (one term per fine sub-face); tank walls carry beta = 0.
// inb(local index) reads the own block, outb(cell) any other block.
template <class In, class Out>
__device__ __forceinline__ void stencil(const LevelDev& L, const CellRef& cr, In inb, Out outb, float& diag,
                                        float& off) {
  const int B = cr.B, r = cr.r;
  const int e = L.e[B];
  const float s = (float)(1 << e);
  diag = 0.f;
  off = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const uint32_t fo = L.foff[a][B];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int nl = cr.l[a] + (side ? 1 : -1);
      int sl[3] = {cr.l[0], cr.l[1], cr.l[2]};
      sl[a] += side;
      if (nl >= 0 && nl < r) {
        const float w = L.beta[a][fo + face_idx(a, sl, r)] * s;
        int nc[3] = {cr.l[0], cr.l[1], cr.l[2]};
        nc[a] = nl;
        diag += w;
        off += w * inb(cell_idx(nc, r));
        continue;
      }
      const int N = L.nbr[B * 6 + 2 * a + side];
      if (N < 0) continue;
      const int en = L.e[N], rn = L.bs >> en;
      int nc[3] = {cr.l[0], cr.l[1], cr.l[2]};
      nc[a] = side ? 0 : rn - 1;
      if (en == e) {
        const float w = L.beta[a][fo + face_idx(a, sl, r)] * s;
        diag += w;
        off += w * outb(L.coff[N] + cell_idx(nc, rn));
      } else if (en > e) {
        const float w = L.beta[a][fo + face_idx(a, sl, r)] * s / 1.5f;
        nc[t1] = cr.l[t1] >> 1;
        nc[t2] = cr.l[t2] >> 1;
        diag += w;
        off += w * outb(L.coff[N] + cell_idx(nc, rn));
      } else {
        int ps[3];
        ps[a] = side ? 0 : rn;
        const uint32_t fon = L.foff[a][N], con = L.coff[N];
        const float sf = 0.5f * s;
#pragma unroll
        for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
          for (int d2 = 0; d2 < 2; ++d2) {
            nc[t1] = ps[t1] = 2 * cr.l[t1] + d1;
            nc[t2] = ps[t2] = 2 * cr.l[t2] + d2;
            const float w = L.beta[a][fon + face_idx(a, ps, rn)] * sf / 1.5f;
            diag += w;
            off += w * outb(con + cell_idx(nc, rn));
          }
      }
    }
  }
}

// Block red-black x cell red-black weighted Gauss-Seidel on one block colour
// (smoothActiveBlocks, Smoother.cuh:105-202): the work-group's blocks sit in
// LDS, neighbours of the other block colour are frozen in HBM.  `reverse`
// runs the cell colours in the opposite order (post-smoother).
__global__ __launch_bounds__(TPB) void smooth_kernel(LevelDev L, float* __restrict__ x, const float* __restrict__ b,
                                                     const WG* __restrict__ wgt, const uint32_t* __restrict__ blist,
                                                     float omega, int local_sweeps, int reverse) {
  __shared__ float sx[512];
  const WG w = wgt[blockIdx.x];
  const int sh = 3 * w.lr;
  CellRef cr[2];
  bool ok[2];
  float bb[2];
  bool liq[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int i = threadIdx.x + k * TPB;
    ok[k] = decode(L, w, blist, i, cr[k]);
    liq[k] = false;
    if (ok[k]) {
      sx[i] = x[cr[k].c];
      bb[k] = b[cr[k].c];
      liq[k] = L.liquid[cr[k].c] != 0;
    }
  }
  __syncthreads();
  for (int ls = 0; ls < local_sweeps; ++ls)
    for (int ph = 0; ph < 2; ++ph) {
      const int col = reverse ? 1 - ph : ph;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!ok[k] || !liq[k]) continue;
        const CellRef& c = cr[k];
        if (((c.l[0] + c.l[1] + c.l[2]) & 1) != col) continue;
        const int base = c.q << sh;
        float diag, off;
        stencil(
            L, c, [&](int li) { return sx[base + li]; }, [&](uint32_t cc) { return x[cc]; }, diag, off);
        const int i = threadIdx.x + k * TPB;
        const float res = bb[k] - (diag * sx[i] - off);
        if (diag > 0.f) sx[i] += omega * res / diag;
      }
      __syncthreads();
    }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (ok[k] && liq[k]) x[cr[k].c] = sx[threadIdx.x + k * TPB];
}

// r = b - A x (computeExplicitResidual, VCycle.cuh:81-131); 0 on air cells
__global__ __launch_bounds__(TPB) void residual_kernel(LevelDev L, const float* __restrict__ x,
                                                       const float* __restrict__ b, float* __restrict__ r,
                                                       const WG* __restrict__ wgt,
                                                       const uint32_t* __restrict__ blist) {
  const WG w = wgt[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(L, w, blist, threadIdx.x + k * TPB, c)) continue;
    if (!L.liquid[c.c]) {
      r[c.c] = 0.f;
      continue;
    }
    const uint32_t co = L.coff[c.B];
    float diag, off;
    stencil(
        L, c, [&](int li) { return x[co + li]; }, [&](uint32_t cc) { return x[cc]; }, diag, off);
    r[c.c] = b[c.c] - (diag * x[c.c] - off);
  }
}

// y = A x with acc += x . y
__global__ __launch_bounds__(TPB) void apply_dot_kernel(LevelDev L, const float* __restrict__ x, float* __restrict__ y,
                                                        const WG* __restrict__ wgt,
                                                        const uint32_t* __restrict__ blist, double* acc) {
  const WG w = wgt[blockIdx.x];
  double part = 0.0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(L, w, blist, threadIdx.x + k * TPB, c)) continue;
    if (!L.liquid[c.c]) {
      y[c.c] = 0.f;
      continue;
    }
    const uint32_t co = L.coff[c.B];
    float diag, off;
    stencil(
        L, c, [&](int li) { return x[co + li]; }, [&](uint32_t cc) { return x[cc]; }, diag, off);
    const float v = diag * x[c.c] - off;
    y[c.c] = v;
    part += (double)x[c.c] * (double)v;
  }
  if (acc) {
    const double s = block_reduce_sum(part);
    if (threadIdx.x == 0) atomicAdd(acc, s);
  }
}

// b_c = sum of the children's residuals (restrictResidualsKernel,
// MultigridTransfer.cuh:9-59): 8 in-block children where the block halves, the
// same cell where it keeps its resolution, 8 child blocks after aggregation.
__global__ __launch_bounds__(TPB) void restrict_kernel(LevelDev F, LevelDev C, int aggregated,
                                                       const float* __restrict__ rf, float* __restrict__ bc,
                                                       const WG* __restrict__ wgt,
                                                       const uint32_t* __restrict__ blist) {
  const WG w = wgt[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(C, w, blist, threadIdx.x + k * TPB, c)) continue;
    float s = 0.f;
    if (aggregated) {
      int Bc[3];
      block_coords(C.D, c.B, Bc);
      for (int d = 0; d < 8; ++d) {
        const int q[3] = {2 * Bc[0] + (d & 1), 2 * Bc[1] + ((d >> 1) & 1), 2 * Bc[2] + (d >> 2)};
        s += rf[F.coff[block_id(F.D, q)]];
      }
    } else if (C.e[c.B] == F.e[c.B]) {
      s = rf[F.coff[c.B] + c.li];
    } else {
      const int rF = 2 * c.r;
      const uint32_t fo = F.coff[c.B];
      for (int d = 0; d < 8; ++d) {
        const int q[3] = {2 * c.l[0] + (d & 1), 2 * c.l[1] + ((d >> 1) & 1), 2 * c.l[2] + (d >> 2)};
        s += rf[fo + cell_idx(q, rF)];
      }
    }
    bc[c.c] = s;
  }
}

// liquid_c = all children liquid (Dirichlet air dominates on coarse levels)
__global__ __launch_bounds__(TPB) void coarsen_liquid_kernel(LevelDev F, LevelDev C, int aggregated,
                                                             uint8_t* __restrict__ liq, const WG* __restrict__ wgt,
                                                             const uint32_t* __restrict__ blist) {
  const WG w = wgt[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(C, w, blist, threadIdx.x + k * TPB, c)) continue;
    uint8_t all = 1;
    if (aggregated) {
      int Bc[3];
      block_coords(C.D, c.B, Bc);
      for (int d = 0; d < 8; ++d) {
        const int q[3] = {2 * Bc[0] + (d & 1), 2 * Bc[1] + ((d >> 1) & 1), 2 * Bc[2] + (d >> 2)};
        all &= F.liquid[F.coff[block_id(F.D, q)]];
      }
    } else if (C.e[c.B] == F.e[c.B]) {
      all = F.liquid[F.coff[c.B] + c.li];
    } else {
      const int rF = 2 * c.r;
      const uint32_t fo = F.coff[c.B];
      for (int d = 0; d < 8; ++d) {
        const int q[3] = {2 * c.l[0] + (d & 1), 2 * c.l[1] + ((d >> 1) & 1), 2 * c.l[2] + (d >> 2)};
        all &= F.liquid[fo + cell_idx(q, rF)];
      }
    }
    liq[c.c] = all ? 1 : 0;
  }
}

// x_f += scale * x_c[parent] on liquid cells (prolongateCorrection,
// MultigridTransfer.cuh:63-102)
__global__ __launch_bounds__(TPB) void prolong_kernel(LevelDev F, LevelDev C, int aggregated, float* __restrict__ xf,
                                                      const float* __restrict__ xc, float scale,
                                                      const WG* __restrict__ wgt,
                                                      const uint32_t* __restrict__ blist) {
  const WG w = wgt[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(F, w, blist, threadIdx.x + k * TPB, c)) continue;
    if (!F.liquid[c.c]) continue;
    uint32_t pc;
    if (aggregated) {
      int Bc[3];
      block_coords(F.D, c.B, Bc);
      const int q[3] = {Bc[0] >> 1, Bc[1] >> 1, Bc[2] >> 1};
      pc = C.coff[block_id(C.D, q)];
    } else if (C.e[c.B] == F.e[c.B]) {
      pc = C.coff[c.B] + c.li;
    } else {
      const int q[3] = {c.l[0] >> 1, c.l[1] >> 1, c.l[2] >> 1};
      pc = C.coff[c.B] + cell_idx(q, c.r >> 1);
    }
    xf[c.c] += scale * xc[pc];
  }
}

// Whole coarsest level in one work-group's LDS: `sweeps` symmetric sweeps
// (forward block/cell colour order, then reverse).
__global__ __launch_bounds__(TPB) void coarse_solve_kernel(LevelDev L, float* __restrict__ x,
                                                           const float* __restrict__ b, int ncells, int sweeps,
                                                           float omega, int local_sweeps) {
  __shared__ float sx[512];
  CellRef cr[2];
  bool ok[2], liq[2];
  float bb[2];
  int bcol[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int i = threadIdx.x + k * TPB;
    ok[k] = i < ncells;
    liq[k] = false;
    if (ok[k]) {
      // locate the cell's block by scanning offsets (the coarsest level is tiny)
      int B = 0;
      while (L.coff[B + 1] <= (uint32_t)i) ++B;
      CellRef& c = cr[k];
      c.B = B;
      c.r = L.bs >> L.e[B];
      c.lr = 31 - __clz(c.r);
      c.li = i - L.coff[B];
      c.q = 0;
      c.l[0] = c.li & (c.r - 1);
      c.l[1] = (c.li >> c.lr) & (c.r - 1);
      c.l[2] = c.li >> (2 * c.lr);
      c.c = i;
      int Bc[3];
      block_coords(L.D, B, Bc);
      bcol[k] = (Bc[0] + Bc[1] + Bc[2]) & 1;
      sx[i] = x[i];
      bb[k] = b[i];
      liq[k] = L.liquid[i] != 0;
    }
  }
  __syncthreads();
  for (int sw = 0; sw < sweeps; ++sw)
    for (int dir = 0; dir < 2; ++dir)
      for (int ph = 0; ph < 2; ++ph) {
        const int bc = dir ? 1 - ph : ph;
        for (int ls = 0; ls < local_sweeps; ++ls)
          for (int cp = 0; cp < 2; ++cp) {
            const int cc = dir ? 1 - cp : cp;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
              if (!ok[k] || !liq[k] || bcol[k] != bc) continue;
              const CellRef& c = cr[k];
              if (((c.l[0] + c.l[1] + c.l[2]) & 1) != cc) continue;
              const uint32_t co = L.coff[c.B];
              float diag, off;
              stencil(
                  L, c, [&](int li) { return sx[co + li]; }, [&](uint32_t q) { return sx[q]; }, diag, off);
              const int i = c.c;
              const float res = bb[k] - (diag * sx[i] - off);
              if (diag > 0.f) sx[i] += omega * res / diag;
            }
            __syncthreads();
          }
      }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (ok[k] && liq[k]) x[cr[k].c] = sx[cr[k].c];
}

// ------------------------------------------------------------- vectors
__global__ __launch_bounds__(TPB) void dot_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                  const uint8_t* __restrict__ liq, int n, double* acc) {
  double s = 0.0;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (liq[i]) s += (double)a[i] * (double)b[i];
  s = block_reduce_sum(s);
  if (threadIdx.x == 0) atomicAdd(acc, s);
}

__global__ void scalar_shift_kernel(double* sc) {
  sc[SC_RZOLD] = sc[SC_RZ];
  sc[SC_RZ] = 0.0;
  sc[SC_PQ] = 0.0;
  sc[SC_RR] = 0.0;
}

__global__ __launch_bounds__(TPB) void pcg_p_kernel(const float* __restrict__ z, float* __restrict__ p, int n,
                                                    const double* __restrict__ sc, int first) {
  const float beta = first ? 0.f : (float)(sc[SC_RZ] / sc[SC_RZOLD]);
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(TPB) void pcg_xr_kernel(float* __restrict__ x, float* __restrict__ r,
                                                     const float* __restrict__ p, const float* __restrict__ q,
                                                     const uint8_t* __restrict__ liq, int n, double* sc) {
  const double pq = sc[SC_PQ];
  const float alpha = pq > 0.0 ? (float)(sc[SC_RZ] / pq) : 0.f;
  double s = 0.0;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
    x[i] += alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    if (liq[i]) s += (double)ri * (double)ri;
  }
  s = block_reduce_sum(s);
  if (threadIdx.x == 0) atomicAdd(&sc[SC_RR], s);
}

__global__ void mask_kernel(float* __restrict__ x, const uint8_t* __restrict__ liq, int n) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (!liq[i]) x[i] = 0.f;
}

__global__ void add_kernel(float* __restrict__ x, const float* __restrict__ z, const uint8_t* __restrict__ liq, int n) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (liq[i]) x[i] += z[i];
}

inline int red_grid(int n) {
  int g = (n + TPB - 1) / TPB;
  return g < 1 ? 1 : (g > 1024 ? 1024 : g);
}

}  // namespace

// ------------------------------------------------------------- launchers
void k_fill_f32(hipStream_t s, float* p, size_t n, float v) {
  if (n) fill_kernel<float><<<grid_for(n), TPB, 0, s>>>(p, n, v);
}
void k_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
  if (n) fill_kernel<uint32_t><<<grid_for(n), TPB, 0, s>>>(p, n, v);
}
void k_fill_f64(hipStream_t s, double* p, size_t n, double v) {
  if (n) fill_kernel<double><<<grid_for(n), TPB, 0, s>>>(p, n, v);
}

void k_bin(hipStream_t s, const GridDev& g, const float* pos, int64_t n, int32_t* key, uint32_t* count) {
  if (n > 0) bin_kernel<<<(unsigned)((n + TPB - 1) / TPB), TPB, 0, s>>>(g, pos, n, key, count);
}

void k_levels(hipStream_t s, const int* D, int L, const int* band, const uint32_t* count, uint8_t* surf,
              uint8_t* level, uint8_t* tmp) {
  const int nb = D[0] * D[1] * D[2];
  const int gr = (nb + TPB - 1) / TPB;
  surface_kernel<<<gr, TPB, 0, s>>>(D[0], D[1], D[2], count, surf);
  Bands bd;
  int bmax = 0;
  for (int i = 0; i < 4; ++i) bd.v[i] = band[i];
  for (int i = 0; i < L - 1; ++i) bmax = band[i] > bmax ? band[i] : bmax;
  band_kernel<<<gr, TPB, 0, s>>>(D[0], D[1], D[2], L, bd, bmax, surf, level);
  for (int it = 0; it < L - 1; ++it) {
    grade_kernel<<<gr, TPB, 0, s>>>(D[0], D[1], D[2], level, tmp);
    hipMemcpyAsync(level, tmp, nb, hipMemcpyDeviceToDevice, s);
  }
}

void k_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, int n) {
  scan_kernel<<<1, SCAN_T, 0, s>>>(in, nullptr, 0, out, n);
}

void k_layout_offsets(hipStream_t s, const uint8_t* level, int nb, uint32_t* coff, uint32_t* fx, uint32_t* fy,
                      uint32_t* fz) {
  scan_kernel<<<1, SCAN_T, 0, s>>>(nullptr, level, 1, coff, nb);
  scan_kernel<<<1, SCAN_T, 0, s>>>(nullptr, level, 2, fx, nb);
  if (fy != fx) hipMemcpyAsync(fy, fx, (nb + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
  if (fz != fx) hipMemcpyAsync(fz, fx, (nb + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
}

void k_scatter(hipStream_t s, const float* pos, const float* vel, const int32_t* ids, const int32_t* key,
               uint32_t* cursor, int64_t n, float* spos, float* svel, int32_t* sids) {
  if (n > 0)
    scatter_kernel<<<(unsigned)((n + TPB - 1) / TPB), TPB, 0, s>>>(pos, vel, ids, key, cursor, n, spos, svel, sids);
}

void k_init_beta(hipStream_t s, const LevelDev& L, float* bx, float* by, float* bz) {
  init_beta_kernel<<<L.nb, TPB, 0, s>>>(L, bx, by, bz);
}

void k_p2g_deposit(hipStream_t s, const GridDev& g, const float* spos, const float* svel, const uint32_t* pcount,
                   const uint32_t* poff, float* slab, uint8_t* liquid) {
  p2g_deposit_kernel<<<g.nb, TPB, 0, s>>>(g, spos, svel, pcount, poff, slab, liquid);
}

void k_p2g_gather(hipStream_t s, const GridDev& g, const uint32_t* pcount, const float* slab, float* const* Macc,
                  float* const* macc) {
  p2g_gather_kernel<<<g.nb, TPB, 0, s>>>(g, pcount, slab, Macc[0], Macc[1], Macc[2], macc[0], macc[1], macc[2]);
}

static FacePtrs fp(float* const* p) { return FacePtrs{{p[0], p[1], p[2]}}; }

void k_fold(hipStream_t s, const GridDev& g, float* const* Macc, float* const* macc, float* const* uold,
            float* const* ustar, float* const* mass, float dt, const float* grav) {
  fold_kernel<<<g.nb, TPB, 0, s>>>(g, fp(Macc), fp(macc), fp(uold), fp(ustar), fp(mass), dt, grav[0], grav[1],
                                    grav[2]);
}

void k_divergence(hipStream_t s, const GridDev& g, const uint8_t* liquid, float* const* ustar, float* rhs) {
  divergence_kernel<<<g.nb, TPB, 0, s>>>(g, liquid, fp(ustar), rhs);
}

void k_project(hipStream_t s, const GridDev& g, const float* p, float* const* ustar, float* const* beta,
               float* const* unew) {
  project_kernel<<<g.nb, TPB, 0, s>>>(g, p, fp(ustar), fp(beta), fp(unew));
}

void k_to_fine(hipStream_t s, const GridDev& g, float* const* unew, float* const* uold, float* fine) {
  to_fine_kernel<<<g.nb, TPB, 0, s>>>(g, fp(unew), fp(uold), fine);
}

void k_g2p_advect(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const float* svel,
                  int64_t n, float alpha, float dt, float* pos_out, float* vel_out) {
  if (n > 0)
    g2p_advect_kernel<<<(unsigned)((n + TPB - 1) / TPB), TPB, 0, s>>>(g, fine, spos, svel, n, alpha, dt, pos_out,
                                                                     vel_out);
}

void k_smooth(hipStream_t s, const LevelDev& L, float* x, const float* b, const WG* wg, int nwg,
              const uint32_t* blist, float omega, int local_sweeps, int reverse) {
  if (nwg > 0) smooth_kernel<<<nwg, TPB, 0, s>>>(L, x, b, wg, blist, omega, local_sweeps, reverse);
}

void k_residual(hipStream_t s, const LevelDev& L, const float* x, const float* b, float* r, const WG* wg, int nwg,
                const uint32_t* blist) {
  if (nwg > 0) residual_kernel<<<nwg, TPB, 0, s>>>(L, x, b, r, wg, blist);
}

void k_apply_dot(hipStream_t s, const LevelDev& L, const float* x, float* y, const WG* wg, int nwg,
                 const uint32_t* blist, double* acc) {
  if (nwg > 0) apply_dot_kernel<<<nwg, TPB, 0, s>>>(L, x, y, wg, blist, acc);
}

void k_restrict(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, const float* rf, float* bc,
                const WG* wgc, int nwgc, const uint32_t* blistc) {
  if (nwgc > 0) restrict_kernel<<<nwgc, TPB, 0, s>>>(F, C, aggregated, rf, bc, wgc, blistc);
}

void k_prolong(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, float* xf, const float* xc,
               float scale, const WG* wgf, int nwgf, const uint32_t* blistf) {
  if (nwgf > 0) prolong_kernel<<<nwgf, TPB, 0, s>>>(F, C, aggregated, xf, xc, scale, wgf, blistf);
}

void k_coarsen_liquid(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, uint8_t* liquid_c,
                      const WG* wgc, int nwgc, const uint32_t* blistc) {
  if (nwgc > 0) coarsen_liquid_kernel<<<nwgc, TPB, 0, s>>>(F, C, aggregated, liquid_c, wgc, blistc);
}

void k_coarse_solve(hipStream_t s, const LevelDev& L, float* x, const float* b, int ncells, int sweeps, float omega,
                    int local_sweeps) {
  coarse_solve_kernel<<<1, TPB, 0, s>>>(L, x, b, ncells, sweeps, omega, local_sweeps);
}

void k_dot(hipStream_t s, const float* a, const float* b, const uint8_t* liquid, int n, double* acc) {
  dot_kernel<<<red_grid(n), TPB, 0, s>>>(a, b, liquid, n, acc);
}

void k_scalar_shift(hipStream_t s, double* sc) { scalar_shift_kernel<<<1, 1, 0, s>>>(sc); }

void k_pcg_p(hipStream_t s, const float* z, float* p, int n, const double* sc, int first) {
  pcg_p_kernel<<<red_grid(n), TPB, 0, s>>>(z, p, n, sc, first);
}

void k_pcg_xr(hipStream_t s, float* x, float* r, const float* p, const float* q, const uint8_t* liquid, int n,
              double* sc) {
  pcg_xr_kernel<<<red_grid(n), TPB, 0, s>>>(x, r, p, q, liquid, n, sc);
}

void k_mask(hipStream_t s, float* x, const uint8_t* liquid, int n) {
  mask_kernel<<<red_grid(n), TPB, 0, s>>>(x, liquid, n);
}

void k_add(hipStream_t s, float* x, const float* z, const uint8_t* liquid, int n) {
  add_kernel<<<red_grid(n), TPB, 0, s>>>(x, z, liquid, n);
}

}  // namespace aw
