// gfx950 kernels of the MSBG FLIP step.
//
// Layout (per native block b, resolution r = 8 >> level[b]):
//   cells   [coff[b], coff[b+1])   x + r (y + r z)                     (MSBG.cuh:165-167)
//   faces   [foff[b], foff[b+1])   normal axis r+1 slots, others r     (MSBG.cuh:169-179)
// The three face-offset arrays coincide because every axis holds (r+1) r^2
// slots per block.  Fine-lattice buffers are block-major: [block][field][axis][512].
//
// All kernels are HBM-bound gathers/scatters and stencils; none uses MFMA.

#include <hip/hip_ext.h>

#include "aw_internal.h"

namespace aw {

// Timing events for the launches that follow (the per-launch profiler): the
// dispatch packets themselves carry the timestamps (hipExtLaunchKernel), so
// timing puts no marker packet, and no wait, between two kernels.
namespace {
thread_local LaunchEvents g_launch_events;
}
void set_launch_events(hipEvent_t start, hipEvent_t stop) { g_launch_events = {start, stop}; }
LaunchEvents take_launch_events() {
  const LaunchEvents e = g_launch_events;
  g_launch_events = {nullptr, nullptr};
  return e;
}

namespace {

// Every kernel launch: plain, or with the profiler's events (the start event
// on the first launch after they are set, the stop event on every launch, so
// a run of launches spans the first start to the last end).
template <typename K, typename... A>
void launch(K kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t s, A... args) {
  if (g_launch_events.stop) {
    hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)shmem, s, g_launch_events.start, g_launch_events.stop, 0,
                          args...);
    g_launch_events.start = nullptr;
  } else {
    kernel<<<grid, block, shmem, s>>>(args...);
  }
}

// Device-to-device copy in stream order.  Like the kernel launches above it
// reports nothing itself: the runtime's launch check (hipGetLastError, which
// holds the last error of any runtime call) sees a failure.
inline void copy_async(hipStream_t s, void* dst, const void* src, size_t bytes) {
  (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
}

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

// LDS float add through compare-and-swap.  On gfx950 the LDS float atomic
// (ds_add_f32) retires about 0.33 lanes per clock per CU, while a read plus
// ds_cmpst_rtn_b32 loop sustains about 8x that when lanes rarely hit the same
// word; with many lanes on one word the retries make it far slower, so it is
// used only where addresses are spread.
__device__ __forceinline__ void lds_add_spread(float* p, float v) {
  unsigned* a = reinterpret_cast<unsigned*>(p);
  unsigned old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (true) {
    const unsigned got = atomicCAS(a, old, __float_as_uint(__uint_as_float(old) + v));
    if (got == old) break;
    old = got;
  }
}

// (x, y) += (dx, dy) on an 8-byte-aligned LDS float pair through one 64-bit
// compare-and-swap (ds_cmpst_rtn_b64): two sums for the price of one retry
// loop, same use conditions as lds_add_spread.
__device__ __forceinline__ void lds_add2_spread(float2* p, float dx, float dy) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(p);
  unsigned long long old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (true) {
    float2 o;
    __builtin_memcpy(&o, &old, 8);
    const float2 nv = make_float2(o.x + dx, o.y + dy);
    unsigned long long want;
    __builtin_memcpy(&want, &nv, 8);
    const unsigned long long got = atomicCAS(a, old, want);
    if (got == old) break;
    old = got;
  }
}

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

// One axis factor of pi_weights: fine index l on the normal (face) axis or a
// transverse axis of a level-e block weighs 1 - t on slot i0 and t on i0 + 1.
__device__ __forceinline__ float pi_w1(int e, bool normal, int l, int I) {
  const int s = 1 << e, r = BASE >> e;
  const float inv_s = 1.0f / (float)s;
  int i0;
  float t;
  if (normal) {
    i0 = l >> e;
    t = (float)(l & (s - 1)) * inv_s;
  } else {
    float eta = ((float)l + 0.5f) * inv_s - 0.5f;
    eta = fminf(fmaxf(eta, 0.f), (float)(r - 1));
    int j = (int)floorf(eta);
    j = r >= 2 ? min(j, r - 2) : 0;
    i0 = j;
    t = eta - (float)j;
  }
  return I == i0 ? 1.f - t : (I == i0 + 1 ? t : 0.f);
}

// Pi^T of a block of level e > 0 for two fields: slot (I0, I1, I2) of axis a
// (face_idx order) = sum over the block's own fine nodes l of the product of
// the three axis factors times F_a(l).  The product separates, so it runs as
// three 1-D passes (x, then y, then z; at most 8 terms per output) through
// LDS scratch instead of scattering each node's 8 weights with atomics (the
// LDS float atomic is slow on gfx950 and neighbouring nodes share slots).
// in(f, a, li) reads field f, axis a, own node li = l0 + 8 (l1 + 8 l2);
// out(a, slot, v0, v1) receives both fields of a slot.  tA*: 3 * 576 floats,
// tB*: 3 * 512 floats; the caller synchronises before reusing them.
template <class In, class Out>
__device__ __forceinline__ void pi_transpose_sep(int e, In in, float* tA0, float* tA1, float* tB0, float* tB1,
                                                 Out out) {
  const int r = BASE >> e, tid = threadIdx.x;
  const int cA = (3 * r + 1) * 64;   // pass x outputs: n0(a) x 8 x 8 over the three axes
  for (int q = tid; q < cA; q += blockDim.x) {
    const int a = q < (r + 1) * 64 ? 0 : (q < (2 * r + 1) * 64 ? 1 : 2);
    const int qa = q - (a == 0 ? 0 : (a == 1 ? (r + 1) * 64 : (2 * r + 1) * 64));
    const int n0 = a == 0 ? r + 1 : r;
    const int I0 = qa % n0, l12 = qa / n0;
    float v0 = 0.f, v1 = 0.f;
    for (int l0 = 0; l0 < BASE; ++l0) {
      const float w = pi_w1(e, a == 0, l0, I0);
      if (w == 0.f) continue;
      v0 += w * in(0, a, l0 + BASE * l12);
      v1 += w * in(1, a, l0 + BASE * l12);
    }
    tA0[a * 576 + qa] = v0;
    tA1[a * 576 + qa] = v1;
  }
  __syncthreads();
  // pass y: n0 x n1 x 8 per axis
  const int cB = ((r + 1) * r + r * (r + 1) + r * r) * 8;
  for (int q = tid; q < cB; q += blockDim.x) {
    const int c0 = (r + 1) * r * 8, c1 = c0 + r * (r + 1) * 8;
    const int a = q < c0 ? 0 : (q < c1 ? 1 : 2);
    const int qa = q - (a == 0 ? 0 : (a == 1 ? c0 : c1));
    const int n0 = a == 0 ? r + 1 : r, n1 = a == 1 ? r + 1 : r;
    const int I0 = qa % n0, I1 = (qa / n0) % n1, l2 = qa / (n0 * n1);
    float v0 = 0.f, v1 = 0.f;
    for (int l1 = 0; l1 < BASE; ++l1) {
      const float w = pi_w1(e, a == 1, l1, I1);
      if (w == 0.f) continue;
      v0 += w * tA0[a * 576 + I0 + n0 * (l1 + BASE * l2)];
      v1 += w * tA1[a * 576 + I0 + n0 * (l1 + BASE * l2)];
    }
    tB0[a * 512 + qa] = v0;
    tB1[a * 512 + qa] = v1;
  }
  __syncthreads();
  // pass z: the slots
  const int nfa = (r + 1) * r * r;
  for (int q = tid; q < 3 * nfa; q += blockDim.x) {
    const int a = q / nfa, f = q - a * nfa;
    const int n0 = a == 0 ? r + 1 : r, n1 = a == 1 ? r + 1 : r, n2 = a == 2 ? r + 1 : r;
    const int I0 = f % n0, I1 = (f / n0) % n1, I2 = f / (n0 * n1);
    (void)n2;
    float v0 = 0.f, v1 = 0.f;
    for (int l2 = 0; l2 < BASE; ++l2) {
      const float w = pi_w1(e, a == 2, l2, I2);
      if (w == 0.f) continue;
      v0 += w * tB0[a * 512 + I0 + n0 * (I1 + n1 * l2)];
      v1 += w * tB1[a * 512 + I0 + n0 * (I1 + n1 * l2)];
    }
    out(a, f, v0, v1);
  }
}

template <typename T, int NT = TPB>
__device__ __forceinline__ T block_reduce_sum(T v) {
  __shared__ T red[NT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  return s;
}

// Grid-wide sums go to SC_SLOTS accumulators per scalar (work-group b adds
// into slot b % SC_SLOTS), so that thousands of work-groups do not serialise
// on one address; readers add the slots.  Call from every thread.
template <int NT = TPB>
__device__ __forceinline__ void grid_sum(double v, double* acc) {
  const double s = block_reduce_sum<double, NT>(v);
  if (threadIdx.x == 0) atomicAdd(acc + (blockIdx.x & (SC_SLOTS - 1)), s);
}

// value of a slotted scalar, for every thread of the work-group
__device__ __forceinline__ double sc_get(const double* sc, int idx) {
  __shared__ double v;
  if (threadIdx.x < 64) {
    double t = threadIdx.x < SC_SLOTS ? sc[idx * SC_SLOTS + threadIdx.x] : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (threadIdx.x == 0) v = t;
  }
  __syncthreads();
  const double r = v;
  __syncthreads();
  return r;
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
// Work-group aggregation of per-block atomics for the binning and sorting
// passes.  Particles arrive mostly grouped by block (each step leaves them
// sorted), so the lanes of a wave form runs of one key (one atomic per run),
// and the runs of a work-group's AGG_PPT * TPB particles fall into a few
// keys: each key takes an LDS slot (open addressing), the runs add their
// lengths there, and one global atomic per key and work-group reserves the
// work-group's range.  Without the LDS step every wave of a block's segment
// hits the same global counter, which serialises.  A key that finds no free
// slot within AGG_PROBE probes takes the global atomic directly.
constexpr int AGG_SLOTS = 512;
constexpr int AGG_PROBE = 16;
constexpr int AGG_PPT = 8;

__device__ __forceinline__ int agg_slot(int* hk, int key) {
  const uint32_t h = ((uint32_t)key * 2654435761u) >> 23;   // 9 bits
#pragma unroll 1
  for (int p = 0; p < AGG_PROBE; ++p) {
    const int s = (int)((h + p) & (AGG_SLOTS - 1));
    const int cur = __hip_atomic_load(&hk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == key) return s;
    if (cur == -1) {
      const int old = atomicCAS(&hk[s], -1, key);
      if (old == -1 || old == key) return s;
    }
  }
  return -1;
}

// lane's run within its wave: head flag, first lane and length of the run
__device__ __forceinline__ void wave_run(int key, bool active, bool& head, int& start, int& len) {
  const int lane = threadIdx.x & 63;
  const int k = active ? key : -1;
  const int prev = __shfl_up(k, 1, 64);
  head = active && (lane == 0 || prev != k);
  const unsigned long long heads = __ballot(head);
  const unsigned long long act = __ballot(active);
  const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const unsigned long long my_heads = heads & le;
  start = my_heads ? 63 - __clzll(my_heads) : lane;
  const unsigned long long stop = (heads | ~act) & ~le;
  const int end = stop ? __ffsll(stop) - 1 : 64;
  len = end - start;
}

// Block of each particle from its exact fine cell (worldToBlockCoord,
// MSBG.cuh:151-159) and the per-block particle count.
__global__ __launch_bounds__(TPB) void bin_kernel(GridDev g, const float* __restrict__ pos, int64_t n,
                                                  int32_t* __restrict__ key, uint32_t* __restrict__ count) {
  __shared__ int hk[AGG_SLOTS];
  __shared__ uint32_t hc[AGG_SLOTS];
  for (int s = threadIdx.x; s < AGG_SLOTS; s += TPB) {
    hk[s] = -1;
    hc[s] = 0;
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * (TPB * AGG_PPT);
#pragma unroll
  for (int j = 0; j < AGG_PPT; ++j) {
    const int64_t i = base + j * TPB + threadIdx.x;
    const bool active = i < n;
    int b = 0;
    if (active) {
      int c[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float u = fminf(fmaxf(pos[3 * i + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
        c[a] = ((int)u) >> LOGB;
      }
      b = block_id(g.D, c);
      key[i] = b;
    }
    bool head;
    int start, len;
    wave_run(b, active, head, start, len);
    if (head) {
      const int sl = agg_slot(hk, b);
      if (sl >= 0)
        atomicAdd(&hc[sl], (uint32_t)len);
      else
        atomicAdd(&count[b], (uint32_t)len);
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < AGG_SLOTS; s += TPB)
    if (hk[s] >= 0) atomicAdd(&count[hk[s]], hc[s]);
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

// Level of each block: the first level whose band holds the block's
// Chebyshev (L-infinity) block distance to the nearest surface block.  The
// distance within the largest band bmax is separable: one pass per axis,
// out(b) = min over |k| <= bmax of max(|k|, in(b + k e_axis)), the first
// pass seeded with 0 at surface blocks and CHEB_FAR elsewhere (distances
// beyond bmax stay CHEB_FAR and fall to the last level).  The last pass
// writes the level instead of the distance.  3 (2 bmax + 1) reads per block
// instead of (2 bmax + 1)^3.
constexpr int CHEB_FAR = 255;

__global__ void cheb_pass_kernel(int D0, int D1, int D2, int axis, int bmax, int seed, int L, Bands band,
                                 const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int D[3] = {D0, D1, D2};
  if (b >= D0 * D1 * D2) return;
  int c[3];
  block_coords(D, b, c);
  const int stride = axis == 0 ? 1 : (axis == 1 ? D0 : D0 * D1);
  const int lo = max(0, c[axis] - bmax), hi = min(D[axis] - 1, c[axis] + bmax);
  int dist = CHEB_FAR;
  for (int q = lo; q <= hi; ++q) {
    const int v = in[b + (q - c[axis]) * stride];
    const int d = seed ? (v ? 0 : CHEB_FAR) : v;
    dist = min(dist, max(abs(q - c[axis]), d));
  }
  if (L == 0) {
    out[b] = (uint8_t)dist;
    return;
  }
  int l = L - 1;
  for (int k = 0; k < L - 1; ++k)
    if (dist <= band.v[k]) {
      l = k;
      break;
    }
  out[b] = (uint8_t)l;
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

// ---- alternative level rule: particle radii (markActiveBlocks, MSBG.cuh:271-314)
// Each particle marks its block's neighbourhood of `support` blocks with its
// level plev = floor(max(0, log2(h / radius))) clamped to [0, L-1]
// (MSBG.cuh:286-291); the map starts at the coarsest level everywhere
// (initRefinementMap with MAX_LEVELS - 1, main.cu:158-162), so every block
// exists.  plev is evaluated exactly, not through log2f: with radius = m 2^E
// (m in [0.5, 1)) and h = 2^H, log2(h / radius) = H - E + log2(1/m), and
// log2(1/m) lies in (0, 1], reaching 1 only for m = 0.5.
__device__ __forceinline__ int radius_level(float r, int logh, int L) {
  int E;
  const float m = frexpf(r, &E);
  const int k = logh - E + (m == 0.5f ? 1 : 0);
  return min(max(k, 0), L - 1);
}

// support = ceil(radius * cellSize / blockSize) = ceil(radius / res) at the
// particle's level (MSBG.cuh:294-296, one for every radius up to res world
// units); capped at the largest dimension, beyond which the box covers the
// domain anyway.  The block is the particle's bin (clamped into the domain).
// A radius that is not positive and finite sets bit 0 of *bad, a particle id
// outside [0, n) bit 1.
// Marking is split in two so that the same addresses are not hit 27 times
// per wave: a particle of support 1 lowers only its own block's minimum
// own[b] (radius_dilate_kernel then spreads each minimum over the 26
// neighbours); a wider support marks its whole box in direct[] at once.
// Lanes that repeat the previous lane's (block, level, support), the common
// case for block-sorted particles of one radius, skip their atomics.
constexpr int RM_PPT = 4;  // particles per thread, TPB apart (coalesced)

__global__ __launch_bounds__(TPB) void radius_mark_kernel(GridDev g, int logh, int L, int smax,
                                                          const float* __restrict__ pos,
                                                          const float* __restrict__ rad,
                                                          const int32_t* __restrict__ ids, int64_t n,
                                                          uint32_t* __restrict__ direct, uint32_t* __restrict__ own,
                                                          uint32_t* __restrict__ bad) {
  const int64_t base = (int64_t)blockIdx.x * (TPB * RM_PPT);
#pragma unroll 1
  for (int it = 0; it < RM_PPT; ++it) {
    const int64_t i = base + it * TPB + threadIdx.x;
    long long key = -1;
    int c[3] = {0, 0, 0}, lev = 0, sup = 0, b = 0;
    if (i < n) {
      // radii by particle id: an id outside [0, n) is an invalid input (bit 1)
      const int32_t id = ids ? ids[i] : (int32_t)i;
      const bool id_ok = id >= 0 && (int64_t)id < n;
      const float r = id_ok ? rad[id] : 0.f;
      if (!id_ok) {
        atomicOr(bad, 2u);
      } else if (!(r > 0.f) || isinf(r)) {
        atomicOr(bad, 1u);
      } else {
        lev = radius_level(r, logh, L);
        sup = (int)fminf(ceilf(r / (float)(BASE >> lev)), (float)smax);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float u = fminf(fmaxf(pos[3 * i + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
          c[a] = ((int)u) >> LOGB;
        }
        b = block_id(g.D, c);
        key = ((long long)b << 32) | ((long long)lev << 24) | sup;
      }
    }
    const long long prev = __shfl_up(key, 1, 64);
    if (key < 0 || ((threadIdx.x & 63) != 0 && prev == key)) continue;
    if (sup == 1) {
      atomicMin(&own[b], (uint32_t)lev);
      continue;
    }
    const int lo0 = max(0, c[0] - sup), hi0 = min(g.D[0] - 1, c[0] + sup);
    const int lo1 = max(0, c[1] - sup), hi1 = min(g.D[1] - 1, c[1] + sup);
    const int lo2 = max(0, c[2] - sup), hi2 = min(g.D[2] - 1, c[2] + sup);
    for (int z = lo2; z <= hi2; ++z)
      for (int y = lo1; y <= hi1; ++y)
        for (int x = lo0; x <= hi0; ++x) atomicMin(&direct[x + g.D[0] * (y + g.D[1] * z)], (uint32_t)lev);
  }
}

// level = min(direct[b], min over b and its 26 neighbours q of own[q])
__global__ void radius_dilate_kernel(int D0, int D1, int D2, const uint32_t* __restrict__ direct,
                                     const uint32_t* __restrict__ own, uint8_t* __restrict__ level) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  int D[3] = {D0, D1, D2};
  if (b >= D0 * D1 * D2) return;
  int c[3];
  block_coords(D, b, c);
  uint32_t l = direct[b];
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= D0 || q[1] >= D1 || q[2] >= D2) continue;
        l = min(l, own[block_id(D, q)]);
      }
  level[b] = (uint8_t)l;
}

// ---- uniform resample of a cell field (sampleMSBG, vtx_visualization.cuh:8-62)
// Voxel (x, y, z) of an o0 x o1 x o2 grid over the domain samples the cell
// that contains its centre.  Along each axis that centre lies in fine cell
// floor((2 q + 1) N / (2 o)), evaluated in exact integer arithmetic (the
// reference's FP32 position can land on either side of a cell face; the
// host keeps (2 o - 1) N below 2^32); its block and the native cell holding
// it follow from the block's level.  A thread writes RS_V consecutive voxels
// of one row (one float4 store when rows are multiples of 4), stepping the
// fine cell by the quotient and remainder of 2N / 2o instead of dividing
// again; a work-group covers RS_Y rows, blockIdx.z is z.
constexpr int RS_X = 128, RS_Y = 2, RS_V = 4;

__global__ __launch_bounds__(RS_X * RS_Y) void resample_kernel(GridDev g, const float* __restrict__ field, int o0,
                                                               int o1, int o2, float* __restrict__ out) {
  const int tx = threadIdx.x % RS_X, ty = threadIdx.x / RS_X;
  const uint32_t x0 = ((uint32_t)blockIdx.x * RS_X + tx) * RS_V;
  const uint32_t y = blockIdx.y * RS_Y + ty, z = blockIdx.z;
  if (x0 >= (uint32_t)o0 || y >= (uint32_t)o1) return;
  const uint32_t fy = ((2u * y + 1u) * (uint32_t)g.N[1]) / (2u * (uint32_t)o1);
  const uint32_t fz = ((2u * z + 1u) * (uint32_t)g.N[2]) / (2u * (uint32_t)o2);
  const uint32_t d = 2u * (uint32_t)o0, step = 2u * (uint32_t)g.N[0];
  const uint32_t kq = step / d, kr = step % d;
  const uint32_t num = (2u * x0 + 1u) * (uint32_t)g.N[0];
  uint32_t fx = num / d, rem = num % d;
  float v[RS_V];
#pragma unroll
  for (int j = 0; j < RS_V; ++j) {
    if (j) {
      fx += kq;
      rem += kr;
      if (rem >= d) {
        rem -= d;
        ++fx;
      }
    }
    v[j] = 0.f;
    if (x0 + j < (uint32_t)o0) {
      const int c[3] = {(int)(fx >> LOGB), (int)(fy >> LOGB), (int)(fz >> LOGB)};
      const int B = block_id(g.D, c);
      const int e = g.level[B];
      const int l[3] = {(int)(fx & (BASE - 1)) >> e, (int)(fy & (BASE - 1)) >> e, (int)(fz & (BASE - 1)) >> e};
      v[j] = field[g.coff[B] + cell_idx(l, BASE >> e)];
    }
  }
  float* row = out + (size_t)o0 * (y + (size_t)o1 * z);
  if ((o0 & 3) == 0) {
    *reinterpret_cast<float4*>(row + x0) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int j = 0; j < RS_V; ++j)
      if (x0 + j < (uint32_t)o0) row[x0 + j] = v[j];
  }
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

// Many-work-group form of the same scan for long inputs: chunk totals, an
// exclusive scan of the totals (scan_kernel), then each chunk rescans itself
// from its base.  Chunks of SCAN_CH values, SCAN_VPT consecutive per thread.
constexpr int SCAN_VPT = 16;
constexpr int SCAN_CH = TPB * SCAN_VPT;

__device__ __forceinline__ uint32_t scan_value(const uint32_t* in, const uint8_t* lv, int mode, int i) {
  if (mode == 0) return in[i];
  const int r = BASE >> lv[i];
  return mode == 1 ? (uint32_t)(r * r * r) : (uint32_t)((r + 1) * r * r);
}

// sum over a work-group of one value per thread, returned to every thread,
// and the exclusive prefix of the thread's value
__device__ __forceinline__ uint32_t wg_exclusive(uint32_t v, uint32_t* wsum, uint32_t& total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t base = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < TPB / 64; ++k) {
    base += k < w ? wsum[k] : 0u;
    total += wsum[k];
  }
  return base + incl - v;
}

__global__ __launch_bounds__(TPB) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                          const uint8_t* __restrict__ lv, int mode, int n,
                                                          uint32_t* __restrict__ sums) {
  __shared__ uint32_t wsum[TPB / 64];
  const int base = blockIdx.x * SCAN_CH;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_VPT; ++k) {
    const int i = base + k * TPB + threadIdx.x;
    if (i < n) s += scan_value(in, lv, mode, i);
  }
  uint32_t total;
  wg_exclusive(s, wsum, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(TPB) void scan_final_kernel(const uint32_t* __restrict__ in,
                                                         const uint8_t* __restrict__ lv, int mode, int n,
                                                         const uint32_t* __restrict__ cbase, int nchunks,
                                                         uint32_t* __restrict__ out) {
  __shared__ uint32_t wsum[TPB / 64];
  const int i0 = blockIdx.x * SCAN_CH + threadIdx.x * SCAN_VPT;
  uint32_t v[SCAN_VPT], s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_VPT; ++k) {
    v[k] = i0 + k < n ? scan_value(in, lv, mode, i0 + k) : 0u;
    s += v[k];
  }
  uint32_t total;
  uint32_t e = cbase[blockIdx.x] + wg_exclusive(s, wsum, total);
#pragma unroll
  for (int k = 0; k < SCAN_VPT; ++k) {
    if (i0 + k < n) out[i0 + k] = e;
    e += v[k];
  }
  if (blockIdx.x == nchunks - 1 && threadIdx.x == 0) out[n] = cbase[nchunks];
}

// dens (with ids): per-particle density by particle id, gathered into sdens
// in the sorted order
// A largest-|v| bound is kept in VB_SLOTS words (float bits, work-group b
// into slot b mod VB_SLOTS), so that tens of thousands of work-groups do not
// serialise on one address; vbound_get takes the largest slot.  Call
// vbound_add from every thread of the work-group.
__device__ __forceinline__ void vbound_add(uint32_t* slots, uint32_t bits) {
  __shared__ uint32_t wmax;
  if (threadIdx.x == 0) wmax = 0u;
  __syncthreads();
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) bits = max(bits, (uint32_t)__shfl_down((int)bits, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&wmax, bits);
  __syncthreads();
  if (threadIdx.x == 0 && wmax) atomicMax(&slots[blockIdx.x % VB_SLOTS], wmax);
}

__global__ __launch_bounds__(TPB) void scatter_kernel(const float* __restrict__ pos, const float* __restrict__ vel,
                                                      const int32_t* __restrict__ ids,
                                                      const int32_t* __restrict__ key, uint32_t* __restrict__ cursor,
                                                      int64_t n, float* __restrict__ spos, float* __restrict__ svel,
                                                      int32_t* __restrict__ sids, const float* __restrict__ dens,
                                                      float* __restrict__ sdens, uint32_t* __restrict__ vbound) {
  __shared__ int hk[AGG_SLOTS];
  __shared__ uint32_t hc[AGG_SLOTS];
  for (int s = threadIdx.x; s < AGG_SLOTS; s += TPB) {
    hk[s] = -1;
    hc[s] = 0;
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * (TPB * AGG_PPT);
  const int lane = threadIdx.x & 63;
  int slot[AGG_PPT];
  uint32_t off[AGG_PPT];   // offset within the slot's range, or the global position (slot -1)
#pragma unroll
  for (int j = 0; j < AGG_PPT; ++j) {
    const int64_t i = base + j * TPB + threadIdx.x;
    const bool active = i < n;
    const int k = active ? key[i] : 0;
    bool head;
    int start, len;
    wave_run(k, active, head, start, len);
    int sl = -1;
    uint32_t b = 0;
    if (head) {
      sl = agg_slot(hk, k);
      b = sl >= 0 ? atomicAdd(&hc[sl], (uint32_t)len) : atomicAdd(&cursor[k], (uint32_t)len);
    }
    slot[j] = __shfl(sl, start, 64);
    off[j] = __shfl(b, start, 64) + (uint32_t)(lane - start);
  }
  __syncthreads();
  // one global reservation per key; hc becomes the key's first position
  for (int s = threadIdx.x; s < AGG_SLOTS; s += TPB)
    if (hk[s] >= 0) hc[s] = atomicAdd(&cursor[hk[s]], hc[s]);
  __syncthreads();
  // the largest |v| and |density| on the way (the P2G deposit's fixed-point
  // scale); a non-finite value gives NaN bits, which compare above every
  // finite value's
  float vm = 0.f, dm = 0.f;
  bool bad = false, dbad = false;
#pragma unroll
  for (int j = 0; j < AGG_PPT; ++j) {
    const int64_t i = base + j * TPB + threadIdx.x;
    if (i >= n) break;
    const int64_t s = slot[j] >= 0 ? (int64_t)hc[slot[j]] + off[j] : (int64_t)off[j];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      spos[3 * s + a] = pos[3 * i + a];
      const float v = vel[3 * i + a];
      svel[3 * s + a] = v;
      bad = bad || !(fabsf(v) <= 3.402823466e38f);
      vm = fmaxf(vm, fabsf(v));
    }
    if (ids) sids[s] = ids[i];
    if (dens) {
      const float d = dens[ids[i]];
      sdens[s] = d;
      dbad = dbad || !(fabsf(d) <= 3.402823466e38f);
      dm = fmaxf(dm, fabsf(d));
    }
  }
  vbound_add(vbound, bad ? 0x7fc00000u : __float_as_uint(vm));
  if (dens) vbound_add(vbound + VB_SLOTS, dbad ? 0x7fc00000u : __float_as_uint(dm));
}

// the largest |value| of v[0..n) into *out (float bits, NaN bits for a
// non-finite value; *out cleared by the launcher)
__global__ __launch_bounds__(TPB) void vbound_kernel(const float* __restrict__ v, int64_t n, uint32_t* __restrict__ out) {
  float m = 0.f;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    const float a = fabsf(v[i]);
    bad = bad || !(a <= 3.402823466e38f);
    m = fmaxf(m, a);
  }
  vbound_add(out, bad ? 0x7fc00000u : __float_as_uint(m));
}

// beta = 1 on interior slots, 0 on tank-wall slots
// beta = 1 on every face slot, 0 on the tank's walls.  One wave per block,
// the axes in turn (no division by the slot count); four blocks per
// work-group, so that the many small coarse blocks do not each take one.
__global__ __launch_bounds__(TPB) void init_beta_kernel(LevelDev L, float* __restrict__ bx, float* __restrict__ by,
                                                        float* __restrict__ bz) {
  const int B = blockIdx.x * (TPB / 64) + (int)(threadIdx.x >> 6);
  if (B >= L.nb) return;
  const int lane = threadIdx.x & 63;
  const int r = L.bs >> L.e[B];
  int Bc[3];
  block_coords(L.D, B, Bc);
  const int nf = (r + 1) * r * r;
  float* dst[3] = {bx, by, bz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float* d = dst[a] + L.foff[a][B];
    for (int f = lane; f < nf; f += 64) {
      int c[3];
      face_decode(a, f, r, c);
      const bool wall = (c[a] == 0 && Bc[a] == 0) || (c[a] == r && Bc[a] == L.D[a] - 1);
      d[f] = wall ? 0.f : 1.f;
    }
  }
}

// Work-groups are dealt to the 8 XCDs round-robin (group w on XCD w mod 8),
// and each XCD has its own L2.  xcd_group gives XCD x the consecutive groups
// [x q + min(x, r), ...) of a launch of q * 8 + r groups.  Particle kernels
// sample the padded lattice of the particle's block, and a block's particles
// are consecutive after the binning sort: with the plain mapping a block's
// groups run on all 8 XCDs and each XCD's L2 fetches the block's lattice.
// Kernels of one work-group per block that read their neighbours' data take
// their block the same way, so that neighbours in x and y (block ids 1 and
// D_x apart) mostly run on the XCD whose L2 already holds those lines.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int64_t xcd_group(int64_t w, int64_t nwg) {
  const int64_t q = nwg / NUM_XCD, r = nwg % NUM_XCD, x = w % NUM_XCD, j = w / NUM_XCD;
  return x * q + (x < r ? x : r) + j;
}

// ------------------------------------------------------------- P2G
// One work-group per block: the block's particles (contiguous after the
// binning sort) scatter their trilinear weights into the block's padded fine
// nodes held in LDS, and the node sums then go to the block's slab with
// plain stores, so P2G issues no global float atomics (P2G.cuh:16-135
// scatters with global float atomics instead).
//
// The sums are fixed point: each weight w and each w v is rounded to a
// 64-bit integer at a per-block scale 2^S and added with LDS integer adds,
// which commute, so the deposit does not depend on the order in which
// threads add (the same bits on every run).  With n the block's particles
// and |v| < 2^E (E from vbound, the largest |v| of the launch, which the
// binning sort records as it copies the particles), every node sum is below
// n 2^E, so S = 62 - ceil(log2 n) - E keeps it, rounding included, inside
// int64: the resolution 2^-S, about 1e-13 of max |v| at 8^3 blocks, is far
// below the rounding of the FP32 product w v each term starts from.  (A
// 32-bit sum, 2^-30 of even a tighter bound, lost up to 3e-4 of max |u| on
// nodes of small mass.)  A non-finite velocity makes the sums NaN, as float
// sums would.  On gfx950 a no-return LDS u64 add retires about as many
// pair-adds as a float compare-and-swap pair (1.05 against 0.96 T
// pair-adds/s measured), and LDS float atomics (ds_add_f32) a tenth.
constexpr int DEP_T = 256;            // threads per work-group
constexpr int DEP_NODES = 3 * PAD;    // padded nodes of all three axes

// DENS: the deposited value is the particle's density sdens[p] on every
// axis instead of its velocity (the density-weighted weight of the
// two-phase coefficients), with the same weights
template <bool DENS>
__global__ __launch_bounds__(DEP_T) void p2g_deposit_kernel(GridDev g, const float* __restrict__ spos,
                                                          const float* __restrict__ svel,
                                                          const uint32_t* __restrict__ pcount,
                                                          const uint32_t* __restrict__ poff, float* __restrict__ slab,
                                                          uint8_t* __restrict__ liquid,
                                                          const uint32_t* __restrict__ vbound) {
  __shared__ unsigned long long accM[DEP_NODES];   // sum of w v per padded node (int64 bits, scale 2^sM)
  __shared__ unsigned long long accm[DEP_NODES];   // sum of w (scale 2^sm)
  __shared__ uint32_t ccnt[512];                   // particles per fine cell
  const int B = blockIdx.x, tid = threadIdx.x;
  const uint32_t cnt = pcount[B];
  const int e = g.level[B], r = BASE >> e, nc = r * r * r;
  const uint32_t co = g.coff[B];
  if (cnt == 0) {
    for (int k = tid; k < nc; k += DEP_T) liquid[co + k] = 0;
    return;
  }
  int Bc[3];
  block_coords(g.D, B, Bc);
  for (int k = tid; k < DEP_NODES; k += DEP_T) accM[k] = accm[k] = 0ull;
  for (int k = tid; k < 512; k += DEP_T) ccnt[k] = 0;
  __shared__ uint32_t svb;
  if (tid < 64) {
    uint32_t m = tid < VB_SLOTS ? vbound[(DENS ? VB_SLOTS : 0) + tid] : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_down((int)m, o, 64));
    if (tid == 0) svb = m;
  }
  __syncthreads();
  const uint32_t vb = svb;
  const bool bad = vb >= 0x7f800000u;              // a non-finite |v|
  int ev, lc = 0;
  (void)frexpf(bad ? 1.f : __uint_as_float(vb), &ev);   // |v| < 2^ev
  while (lc < 32 && (1u << lc) < cnt) ++lc;              // cnt <= 2^lc
  const int sM = min(100, 62 - lc - ev), sm = 62 - lc;
  const float scM = ldexpf(1.f, sM), scm = ldexpf(1.f, sm);
  __syncthreads();
  const uint32_t p0 = poff[B];
  // windows of DEP_WIN consecutive particles; in a window lane l of
  // wave v takes particles DEP_S l + DEP_SPW v + (0 .. DEP_SPW-1).  Lanes sit
  // DEP_S particles apart (particles of one cell are often neighbours in the
  // sorted order, and lanes of one wave on one node serialise), and a
  // work-group's cache lines are those of one window, which L2 holds until
  // the window is done (one contiguous run per thread over the whole block
  // re-read the particle arrays from HBM about nine times over).
  constexpr uint32_t DEP_S = 16, DEP_WIN = 64 * DEP_S, DEP_SPW = DEP_WIN / DEP_T;
  const uint32_t lane = tid & 63, wv = tid >> 6;
  const uint32_t nit = (cnt + DEP_WIN - 1) / DEP_WIN * DEP_SPW;
  for (uint32_t it = 0; it < nit; ++it) {
    const uint32_t i = (it / DEP_SPW) * DEP_WIN + lane * DEP_S + wv * DEP_SPW + it % DEP_SPW;
    if (i >= cnt) continue;
    const size_t p = p0 + i;
    float u[3], v[3];
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float ua = fminf(fmaxf(spos[3 * p + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
      u[a] = ua - (float)(BASE * Bc[a]);
      v[a] = DENS ? svel[p] : svel[3 * p + a];
      inside = inside && ((int)ua >> LOGB) == Bc[a];
    }
    // a particle outside this block (unsorted input) is skipped
    if (!inside) continue;
    atomicAdd(&ccnt[(int)u[0] + 8 * ((int)u[1] + 8 * (int)u[2])], 1u);
    // per axis: lower node and the two hat weights on the face-normal lattice
    // (nodes at integer u, padded index = node) and on the transverse one
    // (nodes at u - 1/2, padded index = node + 1).  A transverse node outside
    // the tank gives its weight to the wall node; a normal node on the top
    // wall face gets none.
    int ln[3], lt[3];
    float wn[3][2], wt[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float fun = floorf(u[d]);
      const float tn = u[d] - fun;
      ln[d] = (int)fun;
      wn[d][0] = 1.f - tn;
      wn[d][1] = (BASE * Bc[d] + ln[d] + 1 >= g.N[d]) ? 0.f : tn;
      const float ut = u[d] - 0.5f, fut = floorf(ut);
      const float tt = ut - fut;
      lt[d] = (int)fut + 1;
      const int gl = BASE * Bc[d] + lt[d] - 1;   // global index of the lower transverse node
      wt[d][0] = gl < 0 ? 0.f : (gl + 1 >= g.N[d] ? 1.f : 1.f - tt);
      wt[d][1] = gl < 0 ? 1.f : (gl + 1 >= g.N[d] ? 0.f : tt);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int nx = a == 0 ? 9 : 10, ny = a == 1 ? 9 : 10;
      const int lx = a == 0 ? ln[0] : lt[0], ly = a == 1 ? ln[1] : lt[1], lz = a == 2 ? ln[2] : lt[2];
      const float* wx = a == 0 ? wn[0] : wt[0];
      const float* wy = a == 1 ? wn[1] : wt[1];
      const float* wz = a == 2 ? wn[2] : wt[2];
      const int base = a * PAD + lx + nx * (ly + ny * lz);
      const float vs = v[a] * scM;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        const float w = wx[dx] * wy[dy] * wz[dz];
        const int k = base + dx + nx * (dy + ny * dz);
        atomicAdd(&accM[k], (unsigned long long)__float2ll_rn(w * vs));
        atomicAdd(&accm[k], (unsigned long long)__float2ll_rn(w * scm));
      }
    }
  }
  __syncthreads();
  float* dst = slab + (size_t)B * SLAB;
  const float iM = bad ? __uint_as_float(0x7fc00000u) : ldexpf(1.f, -sM), im = ldexpf(1.f, -sm);
  for (int k = tid; k < DEP_NODES; k += DEP_T) {
    dst[k] = (float)(long long)accM[k] * iM;
    dst[DEP_NODES + k] = (float)(long long)accm[k] * im;
  }
  // native-cell liquid mask from the fine-cell counts
  for (int k = tid; k < nc; k += DEP_T) {
    const int x = k % r, y = (k / r) % r, z = k / (r * r);
    const int s = 1 << e;
    uint32_t t = 0;
    for (int dz = 0; dz < s; ++dz)
      for (int dy = 0; dy < s; ++dy)
        for (int dx = 0; dx < s; ++dx) t += ccnt[(x * s + dx) + 8 * ((y * s + dy) + 8 * (z * s + dz))];
    liquid[co + k] = t > 0;
  }
}

// One work-group per block: collect the block's own fine nodes from its slab
// and the halos of the occupied neighbours' slabs, then apply Pi^T into the
// native accumulators (M, m) of every slot of the block.
__global__ __launch_bounds__(TPB) void p2g_gather_kernel(GridDev g, const uint32_t* __restrict__ pcount,
                                                         const float* __restrict__ slab, float* Mx, float* My,
                                                         float* Mz, float* mx, float* my, float* mz) {
  __shared__ float fM[3 * FINE], fm[3 * FINE];
  __shared__ float nM[3 * 576], nm[3 * 576];
  __shared__ int occ[27];   // occupied 3x3x3 neighbours, (dx + 1) + 3 (dy + 1) + 9 (dz + 1)
  __shared__ int any_occ;
  const int B = (int)xcd_group(blockIdx.x, gridDim.x), tid = threadIdx.x;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  int Bc[3];
  block_coords(g.D, B, Bc);
  if (tid == 0) any_occ = 0;
  __syncthreads();
  if (tid < 27) {
    const int q[3] = {Bc[0] + tid % 3 - 1, Bc[1] + (tid / 3) % 3 - 1, Bc[2] + tid / 9 - 1};
    const bool in = q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < g.D[0] && q[1] < g.D[1] && q[2] < g.D[2];
    occ[tid] = in && pcount[block_id(g.D, q)] > 0;
    if (occ[tid]) any_occ = 1;
  }
  __syncthreads();
  float* dM[3] = {Mx, My, Mz};
  float* dm[3] = {mx, my, mz};
  if (!any_occ) {
    // no particle within one block: every slot of the block is empty
    for (int k = tid; k < 3 * nf; k += TPB) {
      const int a = k / nf, f = k - a * nf;
      dM[a][g.foff[a][B] + f] = 0.f;
      dm[a][g.foff[a][B] + f] = 0.f;
    }
    return;
  }
  // the block's own nodes first, every load of the thread in flight together
  // (a slab is written only for a block holding particles); then the sums
  // over the owners of each node, in the same order as the node's loop
  constexpr int NPT = 3 * FINE / TPB;
  static_assert(NPT * TPB == 3 * FINE, "nodes per thread");
  const bool own = occ[13] != 0;
  float oM[NPT], om[NPT];
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    const int k = tid + j * TPB, a = k >> 9, li = k & 511;
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    const float* src = slab + (size_t)B * SLAB + a * PAD + pad_idx(a, l);
    oM[j] = own ? src[0] : 0.f;
    om[j] = own ? src[3 * PAD] : 0.f;
  }
#pragma unroll
  for (int j = 0; j < NPT; ++j) {
    const int k = tid + j * TPB, a = k >> 9, li = k & 511;
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
          if (!occ[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)]) continue;
          if ((dx | dy | dz) == 0) {
            sM += oM[j];
            sm += om[j];
            continue;
          }
          const int q[3] = {Bc[0] + dx, Bc[1] + dy, Bc[2] + dz};
          const int N = block_id(g.D, q);
          const int lp[3] = {l[0] - BASE * dx, l[1] - BASE * dy, l[2] - BASE * dz};
          const float* src = slab + (size_t)N * SLAB + a * PAD + pad_idx(a, lp);
          sM += src[0];
          sm += src[3 * PAD];
        }
    fM[k] = sM;
    fm[k] = sm;
  }
  __syncthreads();
  if (e > 0) {
    pi_transpose_sep(
        e, [&](int f, int a, int li) { return f ? fm[a * FINE + li] : fM[a * FINE + li]; }, nM, nm, fM, fm,
        [&](int a, int f, float v0, float v1) {
          dM[a][g.foff[a][B] + f] = v0;
          dm[a][g.foff[a][B] + f] = v1;
        });
    return;
  }
  // a level-0 block's fine nodes are its native slots, one to one
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
    (void)pi_weights(a, e, l, slot, w);
    nM[a * 576 + slot[0]] += w[0] * M;
    nm[a * 576 + slot[0]] += w[0] * m;
  }
  __syncthreads();
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
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
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

// Two-phase coefficients: beta = 1 / rho_face = m / D on a face with mass
// (m > 1e-6), the reference density's 1 on one without, 0 on the tank walls;
// duplicates and coarse-fine slots folded as fold_kernel folds the velocity
// (the coarse side of a level change, not an unknown of the operator, gets
// the mean of its four fine sub-faces).  The reference reserves BETA = 1/rho
// and MASS_DENSITY channels for this (MSBG.cuh:21-34, P2G.cuh:93-97).
__device__ __forceinline__ float beta_of(float D, float m) { return m > 1e-6f ? m / D : 1.f; }

__global__ __launch_bounds__(TPB) void beta_fold_kernel(GridDev g, FacePtrs macc, FacePtrs Dacc, FacePtrs beta) {
  const int B = blockIdx.x;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int sl[3];
    face_decode(a, f, r, sl);
    const size_t id = g.foff[a][B] + f;
    const float m = macc.p[a][id], D = Dacc.p[a][id];
    float bt;
    if (sl[a] > 0 && sl[a] < r) {
      bt = beta_of(D, m);
    } else {
      const int side = sl[a] == r;
      const int N = g.nbr[B * 6 + 2 * a + side];
      if (N < 0) {
        bt = 0.f;
      } else {
        const int en = g.level[N], rn = BASE >> en;
        int ps[3] = {sl[0], sl[1], sl[2]};
        ps[a] = side ? 0 : rn;
        if (en == e) {
          const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
          bt = beta_of(D + Dacc.p[a][pid], m + macc.p[a][pid]);
        } else if (en == e + 1) {
          ps[t1] = sl[t1] >> 1;
          ps[t2] = sl[t2] >> 1;
          const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
          bt = beta_of(D + 0.25f * Dacc.p[a][pid], m + 0.25f * macc.p[a][pid]);
        } else {
          bt = 0.f;
          for (int d1 = 0; d1 < 2; ++d1)
            for (int d2 = 0; d2 < 2; ++d2) {
              ps[t1] = 2 * sl[t1] + d1;
              ps[t2] = 2 * sl[t2] + d2;
              const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
              bt += 0.25f * beta_of(Dacc.p[a][pid] + 0.25f * D, macc.p[a][pid] + 0.25f * m);
            }
        }
      }
    }
    beta.p[a][id] = bt;
  }
}

// rhs = -s^2 sum_a (u[+] - u[-]) on liquid cells
__global__ __launch_bounds__(TPB) void divergence_kernel(GridDev g, const uint8_t* __restrict__ liquid, FacePtrs u,
                                                         float* __restrict__ rhs) {
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
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
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
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

// The forward fine lattice is padded per block: for component a, the
// block's 8^3 nodes and a one-node halo, so that the 2x2x2 nodes of any
// trilinear sample at a point inside the block lie in that block's array
// (face-normal axis nodes 0..8, transverse axes -1..8, index pad_idx, PAD
// floats per component, 3 PAD per field; field 0 = u_new, 1 = u_new - u_old).
// Halo nodes hold their owner block's values, or the sampling rule's value
// past the tank: 0 across a face-normal wall, the nearest node inside
// transversally.  The lattice is filled only for the blocks flagged `halo`
// (within one block of a block holding particles: where RK3 stage points
// fall); a sample elsewhere takes the owners' nodes one by one, from the
// owner's lattice where it is filled and from its faces otherwise.

// Warm start of the pressure solve: each cell of the current layout takes the
// previous step's pressure of the previous layout's cell (same block, that
// block's previous resolution) holding its centre.  Blocks do not move, so a
// block whose level did not change copies its cells one to one.
__global__ __launch_bounds__(TPB) void pressure_remap_kernel(GridDev g, const uint8_t* __restrict__ level_prev,
                                                             const uint32_t* __restrict__ coff_prev,
                                                             const float* __restrict__ p_prev,
                                                             float* __restrict__ p) {
  const int B = blockIdx.x;
  const int e = g.level[B], r = BASE >> e, s = 1 << e;
  const int eo = level_prev[B], ro = BASE >> eo;
  const uint32_t co = g.coff[B], cp = coff_prev[B];
  for (int li = threadIdx.x; li < r * r * r; li += TPB) {
    const int l[3] = {li & (r - 1), (li / r) & (r - 1), li / (r * r)};
    int lo[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) lo[d] = min((l[d] * s + (s >> 1)) >> eo, ro - 1);
    p[co + li] = p_prev[cp + cell_idx(lo, ro)];
  }
}

// halo[B] = a block of the 3x3x3 neighbourhood of B holds particles
__global__ __launch_bounds__(TPB) void halo_flags_kernel(GridDev g, const uint32_t* __restrict__ count,
                                                         uint8_t* __restrict__ halo) {
  const int B = blockIdx.x * TPB + threadIdx.x;
  if (B >= g.nb) return;
  int Bc[3];
  block_coords(g.D, B, Bc);
  bool any = false;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int q[3] = {Bc[0] + dx, Bc[1] + dy, Bc[2] + dz};
        if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= g.D[0] || q[1] >= g.D[1] || q[2] >= g.D[2]) continue;
        any = any || count[q[0] + g.D[0] * (q[1] + g.D[1] * q[2])] > 0;
      }
  halo[B] = any ? 1 : 0;
}

// Pi of u_new (field 0) or of u_new - u_old (field 1) at own fine node l of
// block B on component a: the weighted sum of the block's native slots
__device__ __forceinline__ float node_from_faces(const GridDev& g, int field, int a, int B, const int* l) {
  int slot[8];
  float w[8];
  const int n = pi_weights(a, g.level[B], l, slot, w);
  float vn = 0.f, vd = 0.f;
  for (int j = 0; j < n; ++j) {
    const size_t id = g.foff[a][B] + slot[j];
    const float un = g.un[a][id];
    vn += w[j] * un;
    vd += w[j] * (un - g.uo[a][id]);
  }
  return field ? vd : vn;
}

// Fine node l (block-local fine coordinates) on axis a of a block of level
// e, from that block's face slots on axis a (un_a, uo_a, in face_idx order):
// Pi(u_new) and Pi(u_new - u_old).  The node takes its 2x2x2 slot products
// in pi_weights's order, loading only the slots of nonzero weight, so a
// node has the same sums in every lattice that holds it.
__device__ __forceinline__ void fine_node_from(int a, int e, const float* un_a, const float* uo_a, const int* l,
                                               float& vn, float& vd) {
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
  float un[8], uo[8], ww[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
    ww[k] = (dx ? t[0] : 1.f - t[0]) * (dy ? t[1] : 1.f - t[1]) * (dz ? t[2] : 1.f - t[2]);
    un[k] = uo[k] = 0.f;
    if (ww[k] != 0.f) {
      const int id = face_idx(a, i0[0] + dx, i0[1] + dy, i0[2] + dz, r);
      un[k] = un_a[id];
      uo[k] = uo_a[id];
    }
  }
  vn = 0.f;
  vd = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (ww[k] == 0.f) continue;
    vn += ww[k] * un[k];
    vd += ww[k] * (un[k] - uo[k]);
  }
}

// The padded lattice of every flagged block, each node from the faces of
// the block that owns it.  Its own nodes (block-local fine coordinates in
// [0, 8)^3, 512 per axis) come from the block's faces, staged in LDS with
// coalesced loads; the 388 halo nodes per axis (the normal layer at 8 and
// the transverse rings at -1 and 8) from their owners' faces in global
// memory, two per thread, the owners' levels and face offsets (the block
// and its 26 neighbours) read into LDS first.
constexpr int TF_HALO = PAD - FINE;
__global__ __launch_bounds__(TPB) void to_fine_kernel(GridDev g, float* __restrict__ fine) {
  __shared__ int nlev[27];
  __shared__ uint32_t nfo[27];
  __shared__ float sun[3 * 576], suo[3 * 576];   // the block's face slots, (r + 1) r^2 <= 576 per axis
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
  if (!g.halo[B]) return;
  int Bc[3];
  block_coords(g.D, B, Bc);
  if (threadIdx.x < 27) {
    const int t = threadIdx.x;
    const int q[3] = {Bc[0] + t % 3 - 1, Bc[1] + (t / 3) % 3 - 1, Bc[2] + t / 9 - 1};
    const bool in = q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < g.D[0] && q[1] < g.D[1] && q[2] < g.D[2];
    const int Bo = in ? block_id(g.D, q) : B;
    nlev[t] = g.level[Bo];
    nfo[t] = g.foff[0][Bo];   // the three axes' face offsets coincide
  }
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  const uint32_t fo = g.foff[0][B];
  for (int k = threadIdx.x; k < 3 * nf; k += TPB) {
    const int a = k / nf, i = k - a * nf;
    sun[a * 576 + i] = g.un[a][fo + i];
    suo[a * 576 + i] = g.uo[a][fo + i];
  }
  __syncthreads();
  float* dst = fine + (size_t)B * SLAB;
  for (int k = threadIdx.x; k < 3 * FINE; k += TPB) {
    const int a = k >> 9, li = k & (FINE - 1);
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    float vn, vd;
    fine_node_from(a, e, sun + a * 576, suo + a * 576, l, vn, vd);
    const int pi = a * PAD + pad_idx(a, l);
    dst[pi] = vn;
    dst[3 * PAD + pi] = vd;
  }
  for (int k0 = threadIdx.x; k0 < 3 * TF_HALO; k0 += 2 * TPB) {
    float vn[2], vd[2];
    int pi[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = k0 + h * TPB;
      vn[h] = vd[h] = 0.f;
      pi[h] = -1;
      if (k >= 3 * TF_HALO) continue;
      const int a = k / TF_HALO, hh = k - a * TF_HALO;
      // padded coordinates: normal tn in [0, 9), transverse t1, t2 in [0, 10)
      // (the lower axis first); the core tn < 8, 1 <= t1, t2 <= 8 is own
      int tn, t1, t2;
      if (hh < 8 * 36) {
        tn = hh / 36;
        const int ri = hh - tn * 36;
        if (ri < 10) {
          t1 = ri;
          t2 = 0;
        } else if (ri < 20) {
          t1 = ri - 10;
          t2 = 9;
        } else if (ri < 28) {
          t1 = 0;
          t2 = ri - 19;
        } else {
          t1 = 9;
          t2 = ri - 27;
        }
      } else {
        tn = 8;
        t1 = (hh - 8 * 36) % 10;
        t2 = (hh - 8 * 36) / 10;
      }
      const int d1 = a == 0 ? 1 : 0, d2 = a == 2 ? 1 : 2;
      int lp[3];
      lp[a] = tn;
      lp[d1] = t1;
      lp[d2] = t2;
      const int n0 = a == 0 ? 9 : 10, n1 = a == 1 ? 9 : 10;
      pi[h] = a * PAD + lp[0] + n0 * (lp[1] + n1 * lp[2]);
      int q[3];
      bool outside = false;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        q[d] = BASE * Bc[d] + lp[d] - (d == a ? 0 : 1);
        if (d == a)
          outside = outside || q[d] >= g.N[d];
        else
          q[d] = iclamp(q[d], 0, g.N[d] - 1);
      }
      if (outside) continue;
      const int nb = ((q[0] >> LOGB) - Bc[0] + 1) + 3 * ((q[1] >> LOGB) - Bc[1] + 1) + 9 * ((q[2] >> LOGB) - Bc[2] + 1);
      const int l[3] = {q[0] & 7, q[1] & 7, q[2] & 7};
      fine_node_from(a, nlev[nb], g.un[a] + nfo[nb], g.uo[a] + nfo[nb], l, vn[h], vd[h]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (pi[h] < 0) continue;
      dst[pi[h]] = vn[h];
      dst[3 * PAD + pi[h]] = vd[h];
    }
  }
}

// node l of owner block Bo from the padded lattice where it is filled, else
// from the owner's faces
__device__ __forceinline__ float own_node(const GridDev& g, const float* __restrict__ F, int field, int a, int Bo,
                                          const int* l) {
  if (g.halo[Bo]) return F[(size_t)Bo * SLAB + field * 3 * PAD + a * PAD + pad_idx(a, l)];
  return node_from_faces(g, field, a, Bo, l);
}

// Trilinear sample of component A of a field of the padded fine lattice at u
// (fine units): the 2x2x2 nodes lie in one block's padded array, at fixed
// offsets from the first.  Node rule: a face-normal index outside [0, N) is
// 0 (a tank wall or beyond), a transverse index clamps to [0, N - 1].  The
// coordinate is clamped first, which reproduces that rule up to rounding (a
// point past a wall normal to A is only sampled for u_new, 0 on the wall).
// The block of a binned particle and whether it holds binned particles: then
// every block of its 3x3x3 neighbourhood has its halo filled, known without
// reading the flags.
struct Near {
  int bc[3];
  bool own;
};

__device__ __forceinline__ Near near_of(const GridDev& g, const float* u) {
  Near nr;
#pragma unroll
  for (int d = 0; d < 3; ++d) nr.bc[d] = (int)u[d] >> LOGB;
  nr.own = g.count[nr.bc[0] + g.D[0] * (nr.bc[1] + g.D[1] * nr.bc[2])] > 0;
  return nr;
}

// SLOW: a sample whose block has no halo reads each node from its owner
// block; otherwise such a sample clears ok (and its value is not used).
// Trilinear interpolation of the 2x2x2 values v[k] (k = dx + 2 dy + 4 dz)
// at fractions t: seven blends (1 - t) a + t b, each one multiply and one
// fused multiply-add, exact at t = 0 and t = 1.
__device__ __forceinline__ float trilerp(const float* v, const float* t) {
  const float s0 = 1.f - t[0], s1 = 1.f - t[1], s2 = 1.f - t[2];
  const float c00 = fmaf(t[0], v[1], s0 * v[0]), c10 = fmaf(t[0], v[3], s0 * v[2]);
  const float c01 = fmaf(t[0], v[5], s0 * v[4]), c11 = fmaf(t[0], v[7], s0 * v[6]);
  const float c0 = fmaf(t[1], c10, s1 * c00), c1 = fmaf(t[1], c11, s1 * c01);
  return fmaf(t[2], c1, s2 * c0);
}

// trilerp of the nodes at fixed offsets from f in an n0 x n1 x . array
template <int n0, int n1>
__device__ __forceinline__ float trilerp_at(const float* f, const float* t) {
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = f[(k & 1) + n0 * (((k >> 1) & 1) + n1 * (k >> 2))];
  return trilerp(v, t);
}

template <int A, bool SLOW>
__device__ __forceinline__ float sample_comp(const GridDev& g, const float* __restrict__ F, int field,
                                             const float* u, const Near& nr, bool& ok) {
  constexpr int n0 = A == 0 ? 9 : 10, n1 = A == 1 ? 9 : 10;
  int bc[3], lp[3];
  float t[3];
  bool zero = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int i;
    if (d == A) {
      zero = zero || u[d] < 0.f || u[d] > (float)g.N[d];
      const float c = fminf(fmaxf(u[d], 0.f), (float)g.N[d]);
      i = min((int)floorf(c), g.N[d] - 1);
      t[d] = c - (float)i;
    } else {
      const float c = fminf(fmaxf(u[d] - 0.5f, 0.f), (float)(g.N[d] - 1));
      i = (int)floorf(c);
      t[d] = c - (float)i;
    }
    bc[d] = i >> LOGB;
    lp[d] = (i & 7) + (d == A ? 0 : 1);
  }
  const int B = bc[0] + g.D[0] * (bc[1] + g.D[1] * bc[2]);
  bool fast = nr.own;
#pragma unroll
  for (int d = 0; d < 3; ++d) fast = fast && abs(bc[d] - nr.bc[d]) <= 1;
  if (!fast) fast = g.halo[B] != 0;
  float s = 0.f;
  if (!SLOW && !fast) ok = false;
  if (!SLOW || fast) {
    const float* f = F + (size_t)B * SLAB + field * 3 * PAD + A * PAD + lp[0] + n0 * (lp[1] + n1 * lp[2]);
    s = trilerp_at<n0, n1>(f, t);
  } else if (SLOW) {
    // no halo: each node from its owner block's own nodes
    float vk[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int dd[3] = {k & 1, (k >> 1) & 1, k >> 2};
      int q[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) q[d] = BASE * bc[d] + lp[d] - (d == A ? 0 : 1) + dd[d];
      float v = 0.f;
      if (q[A] < g.N[A]) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (d != A) q[d] = iclamp(q[d], 0, g.N[d] - 1);
        const int Bo = (q[0] >> LOGB) + g.D[0] * ((q[1] >> LOGB) + g.D[1] * (q[2] >> LOGB));
        const int l[3] = {q[0] & 7, q[1] & 7, q[2] & 7};
        v = own_node(g, F, field, A, Bo, l);
      }
      vk[k] = v;
    }
    s = trilerp(vk, t);
  }
  return zero ? 0.f : s;
}

template <bool SLOW>
__device__ __forceinline__ float sample_comp(const GridDev& g, const float* __restrict__ F, int field, int a,
                                             const float* u, const Near& nr, bool& ok) {
  return a == 0 ? sample_comp<0, SLOW>(g, F, field, u, nr, ok)
                : (a == 1 ? sample_comp<1, SLOW>(g, F, field, u, nr, ok)
                          : sample_comp<2, SLOW>(g, F, field, u, nr, ok));
}

// the per-node path of sample_comp, out of line
template <int A>
__device__ __attribute__((noinline)) float sample_comp_slow(const GridDev& g, const float* __restrict__ F, int field,
                                                            float u0, float u1, float u2) {
  const float u[3] = {u0, u1, u2};
  const Near far = {{-2, -2, -2}, false};
  bool ok = true;
  return sample_comp<A, true>(g, F, field, u, far, ok);
}

template <int A>
__device__ __forceinline__ float sample_comp_any(const GridDev& g, const float* __restrict__ F, int field,
                                                 const float* u, const Near& nr) {
  bool ok = true;
  const float v = sample_comp<A, false>(g, F, field, u, nr, ok);
  return ok ? v : sample_comp_slow<A>(g, F, field, u[0], u[1], u[2]);
}

// sample_comp<A, true> with the gradient of the interpolant with respect to
// the sampled coordinate (fine units): G = 0 along an axis where the
// coordinate was clamped (the sample is flat there), as for the node rule of
// sample_comp; the value is computed exactly as sample_comp computes it.
struct SG {
  float s, g0, g1, g2;
};

// trilerp of v at t (the same value, bit for bit) with its derivatives in
// t0, t1, t2: the x blends and x differences of the four (y, z) node pairs,
// then blends of those in y and z
__device__ __forceinline__ SG trilerp_grad(const float* v, const float* t) {
  const float s0 = 1.f - t[0], s1 = 1.f - t[1], s2 = 1.f - t[2];
  const float c00 = fmaf(t[0], v[1], s0 * v[0]), c10 = fmaf(t[0], v[3], s0 * v[2]);
  const float c01 = fmaf(t[0], v[5], s0 * v[4]), c11 = fmaf(t[0], v[7], s0 * v[6]);
  const float d00 = v[1] - v[0], d10 = v[3] - v[2], d01 = v[5] - v[4], d11 = v[7] - v[6];
  const float c0 = fmaf(t[1], c10, s1 * c00), c1 = fmaf(t[1], c11, s1 * c01);
  SG r;
  r.s = fmaf(t[2], c1, s2 * c0);
  r.g0 = fmaf(t[2], fmaf(t[1], d11, s1 * d01), s2 * fmaf(t[1], d10, s1 * d00));
  r.g1 = fmaf(t[2], c11 - c01, s2 * (c10 - c00));
  r.g2 = c1 - c0;
  return r;
}

template <int A, bool FAST>
__device__ __forceinline__ SG sample_grad_impl(const GridDev& g, const float* __restrict__ F, int field,
                                               const float* u) {
  constexpr int n0 = A == 0 ? 9 : 10, n1 = A == 1 ? 9 : 10;
  int bc[3], lp[3];
  float t[3], dm[3];
  bool zero = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int i;
    if (d == A) {
      zero = zero || u[d] < 0.f || u[d] > (float)g.N[d];
      const float c = fminf(fmaxf(u[d], 0.f), (float)g.N[d]);
      i = min((int)floorf(c), g.N[d] - 1);
      t[d] = c - (float)i;
      dm[d] = 1.f;
    } else {
      const float cr = u[d] - 0.5f;
      const float c = fminf(fmaxf(cr, 0.f), (float)(g.N[d] - 1));
      i = (int)floorf(c);
      t[d] = c - (float)i;
      dm[d] = (cr >= 0.f && cr < (float)(g.N[d] - 1)) ? 1.f : 0.f;
    }
    bc[d] = i >> LOGB;
    lp[d] = (i & 7) + (d == A ? 0 : 1);
  }
  const int B = bc[0] + g.D[0] * (bc[1] + g.D[1] * bc[2]);
  const float* p = F + (size_t)B * SLAB + field * 3 * PAD + A * PAD + lp[0] + n0 * (lp[1] + n1 * lp[2]);
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int dd[3] = {k & 1, (k >> 1) & 1, k >> 2};
    float f;
    if (FAST) {
      f = p[dd[0] + n0 * (dd[1] + n1 * dd[2])];
    } else {
      int q[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) q[d] = BASE * bc[d] + lp[d] - (d == A ? 0 : 1) + dd[d];
      f = 0.f;
      if (q[A] < g.N[A]) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (d != A) q[d] = iclamp(q[d], 0, g.N[d] - 1);
        const int Bo = (q[0] >> LOGB) + g.D[0] * ((q[1] >> LOGB) + g.D[1] * (q[2] >> LOGB));
        const int l[3] = {q[0] & 7, q[1] & 7, q[2] & 7};
        f = own_node(g, F, field, A, Bo, l);
      }
    }
    v[k] = f;
  }
  SG r = trilerp_grad(v, t);
  if (zero) return SG{0.f, 0.f, 0.f, 0.f};
  r.g0 *= dm[0];
  r.g1 *= dm[1];
  r.g2 *= dm[2];
  return r;
}

// the per-node path, out of line so that its registers do not weigh on the
// callers' fast path
template <int A>
__device__ __attribute__((noinline)) SG sample_grad_slow(const GridDev& g, const float* __restrict__ F, int field,
                                                         float u0, float u1, float u2) {
  const float u[3] = {u0, u1, u2};
  return sample_grad_impl<A, false>(g, F, field, u);
}

template <int A>
__device__ __forceinline__ float sample_grad(const GridDev& g, const float* __restrict__ F, int field,
                                             const float* u, const Near& nr, float* G) {
  int bc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = d == A ? fminf(fmaxf(u[d], 0.f), (float)g.N[d]) : fminf(fmaxf(u[d] - 0.5f, 0.f), (float)(g.N[d] - 1));
    const int i = d == A ? min((int)floorf(c), g.N[d] - 1) : (int)floorf(c);
    bc[d] = i >> LOGB;
  }
  bool fast = nr.own;
#pragma unroll
  for (int d = 0; d < 3; ++d) fast = fast && abs(bc[d] - nr.bc[d]) <= 1;
  if (!fast) fast = g.halo[bc[0] + g.D[0] * (bc[1] + g.D[1] * bc[2])] != 0;
  const SG r = fast ? sample_grad_impl<A, true>(g, F, field, u) : sample_grad_slow<A>(g, F, field, u[0], u[1], u[2]);
  G[0] = r.g0;
  G[1] = r.g1;
  G[2] = r.g2;
  return r.s;
}

__device__ __forceinline__ float sample_grad(const GridDev& g, const float* __restrict__ F, int field, int a,
                                             const float* u, const Near& nr, float* G) {
  return a == 0 ? sample_grad<0>(g, F, field, u, nr, G)
                : (a == 1 ? sample_grad<1>(g, F, field, u, nr, G) : sample_grad<2>(g, F, field, u, nr, G));
}

// sample_grad on the halo path only: ok = false (and nothing sampled) where
// the point needs the per-node path
template <int A>
__device__ __forceinline__ float sample_grad_fast(const GridDev& g, const float* __restrict__ F, int field,
                                                  const float* u, const Near& nr, float* G, bool& ok) {
  int bc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = d == A ? fminf(fmaxf(u[d], 0.f), (float)g.N[d]) : fminf(fmaxf(u[d] - 0.5f, 0.f), (float)(g.N[d] - 1));
    const int i = d == A ? min((int)floorf(c), g.N[d] - 1) : (int)floorf(c);
    bc[d] = i >> LOGB;
  }
  bool fast = nr.own;
#pragma unroll
  for (int d = 0; d < 3; ++d) fast = fast && abs(bc[d] - nr.bc[d]) <= 1;
  if (!fast) fast = g.halo[bc[0] + g.D[0] * (bc[1] + g.D[1] * bc[2])] != 0;
  if (!fast) {
    ok = false;
    G[0] = G[1] = G[2] = 0.f;
    return 0.f;
  }
  const SG r = sample_grad_impl<A, true>(g, F, field, u);
  G[0] = r.g0;
  G[1] = r.g1;
  G[2] = r.g2;
  return r.s;
}

template <bool SLOW>
__device__ __forceinline__ float sample_grad_sel(const GridDev& g, const float* __restrict__ F, int field, int a,
                                                 const float* u, const Near& nr, float* G, bool& ok) {
  if (SLOW) return sample_grad(g, F, field, a, u, nr, G);
  return a == 0 ? sample_grad_fast<0>(g, F, field, u, nr, G, ok)
                : (a == 1 ? sample_grad_fast<1>(g, F, field, u, nr, G, ok)
                          : sample_grad_fast<2>(g, F, field, u, nr, G, ok));
}

// FLIP/PIC blend (G2P.cuh:87-128) and Ralston RK3 (G2P.cuh:131-173) of
// particle i.  Without SLOW a particle with a stage point in a block without
// halo is appended to list (list[0] counts) and nothing is written for it.
// MODE: G2P_BOTH the blend and the advection (the step's fused pass);
// G2P_VEL the blend only (G2PManager::performG2P, G2P.cuh:192-214), positions
// untouched; G2P_POS the advection only (G2PManager::advect, G2P.cuh:216-227),
// velocities untouched.  The advection samples u_new only, so the two halves
// in either order give the fused pass's result.
enum { G2P_BOTH = 0, G2P_VEL = 1, G2P_POS = 2 };

template <bool SLOW, int MODE>
__device__ __forceinline__ void g2p_advect_one(const GridDev& g, const float* __restrict__ F,
                                               const float* __restrict__ spos, const float* __restrict__ svel,
                                               int64_t i, float alpha, float dt, float* __restrict__ pos_out,
                                               float* __restrict__ vel_out, uint32_t* __restrict__ list) {
  bool ok = true;
  float x[3], u[3], v1[3], vd[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = spos[3 * i + a];
    u[a] = fminf(fmaxf(x[a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
  }
  const Near nr = near_of(g, u);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    v1[a] = sample_comp<SLOW>(g, F, 0, a, u, nr, ok);
    vd[a] = MODE == G2P_POS ? 0.f : sample_comp<SLOW>(g, F, 1, a, u, nr, ok);
  }
  float vn[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) vn[a] = MODE == G2P_POS ? 0.f : alpha * (svel[3 * i + a] + vd[a]) + (1.f - alpha) * v1[a];
  if (MODE == G2P_VEL) {
    if (!SLOW && !ok) {
      list[1 + atomicAdd(&list[0], 1u)] = (uint32_t)i;
      return;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) vel_out[3 * i + a] = vn[a];
    return;
  }
  float xx[3], v2[3], v3[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    xx[a] = x[a] + 0.5f * dt * v1[a];
    u[a] = fminf(fmaxf(xx[a] * g.inv_h, -1.f), (float)g.N[a] + 1.f);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) v2[a] = sample_comp<SLOW>(g, F, 0, a, u, nr, ok);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    xx[a] = x[a] + 0.75f * dt * v2[a];
    u[a] = fminf(fmaxf(xx[a] * g.inv_h, -1.f), (float)g.N[a] + 1.f);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) v3[a] = sample_comp<SLOW>(g, F, 0, a, u, nr, ok);
  if (!SLOW && !ok) {
    list[1 + atomicAdd(&list[0], 1u)] = (uint32_t)i;
    return;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float xn = x[a] + dt * ((2.f / 9.f) * v1[a] + (3.f / 9.f) * v2[a] + (4.f / 9.f) * v3[a]);
    const float lo = 0.25f * g.h, hi = (float)g.N[a] * g.h - 0.25f * g.h;
    if (MODE == G2P_BOTH) vel_out[3 * i + a] = vn[a];
    pos_out[3 * i + a] = fminf(fmaxf(xn, lo), hi);
  }
}


// FIXUP = false: every particle; FIXUP = true: the particles listed by the
// first pass (a grid-stride loop over list[0] entries, so the host needs no
// count)
template <bool FIXUP, int MODE>
__global__ __launch_bounds__(TPB) void g2p_advect_kernel(GridDev g, const float* __restrict__ F,
                                                         const float* __restrict__ spos,
                                                         const float* __restrict__ svel, int64_t n, float alpha,
                                                         float dt, float* __restrict__ pos_out,
                                                         float* __restrict__ vel_out, uint32_t* __restrict__ list) {
  const int64_t i = (FIXUP ? (int64_t)blockIdx.x : xcd_group(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
  if (FIXUP) {
    const int64_t m = (int64_t)list[0];
    for (int64_t j = i; j < m; j += (int64_t)gridDim.x * blockDim.x)
      g2p_advect_one<true, MODE>(g, F, spos, svel, (int64_t)list[1 + j], alpha, dt, pos_out, vel_out, list);
    return;
  }
  if (i >= n) return;
  g2p_advect_one<false, MODE>(g, F, spos, svel, i, alpha, dt, pos_out, vel_out, list);
}

// ---- G2P of the step's block-sorted particles with the block's lattice in LDS ----
// Trilinear sample of component A of `field` from one block's padded lattice
// staged in LDS (lat: the block's SLAB floats; o: the block's first fine
// cell per axis), with sample_comp's node rule and sums.  Returns false, and
// reads nothing, when the sample's 2x2x2 nodes are not all in that lattice
// (a sample point outside the block).
template <int A>
__device__ __forceinline__ bool sample_lds(const GridDev& g, const float* lat, int field, const int* o, const float* u,
                                           float& out) {
  constexpr int n0 = A == 0 ? 9 : 10, n1 = A == 1 ? 9 : 10, n2 = A == 2 ? 9 : 10;
  int lp[3];
  float t[3];
  bool zero = false, in = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int i;
    if (d == A) {
      zero = zero || u[d] < 0.f || u[d] > (float)g.N[d];
      const float c = fminf(fmaxf(u[d], 0.f), (float)g.N[d]);
      i = min((int)floorf(c), g.N[d] - 1);
      t[d] = c - (float)i;
    } else {
      const float c = fminf(fmaxf(u[d] - 0.5f, 0.f), (float)(g.N[d] - 1));
      i = (int)floorf(c);
      t[d] = c - (float)i;
    }
    lp[d] = i - o[d] + (d == A ? 0 : 1);
    const int nd = d == 0 ? n0 : (d == 1 ? n1 : n2);
    in = in && lp[d] >= 0 && lp[d] <= nd - 2;
  }
  if (!in) return false;
  const float s = trilerp_at<n0, n1>(lat + field * 3 * PAD + A * PAD + lp[0] + n0 * (lp[1] + n1 * lp[2]), t);
  out = zero ? 0.f : s;
  return true;
}

// component a of field at u: from the LDS lattice, else through the global
// lattice's halo path (sample_comp; a sample it cannot serve clears ok)
template <int A>
__device__ __forceinline__ float sample_blk(const GridDev& g, const float* lat, const float* __restrict__ F, int field,
                                            const int* o, const float* u, const Near& nr, bool& ok) {
  float v;
  if (!sample_lds<A>(g, lat, field, o, u, v)) v = sample_comp<A, false>(g, F, field, u, nr, ok);
  return v;
}

// One work-group per block: the block's padded lattice (both fields, 21.6 KB)
// is read into LDS once, and the block's particles (consecutive after the
// binning sort, poff[B] .. poff[B] + pcount[B]) sample it there; a stage
// point outside the block samples the global lattice, and a particle none of
// whose paths can serve goes to the list that g2p_advect_kernel<true>
// finishes.  Same arithmetic as g2p_advect_one<false, G2P_BOTH>, so the
// same results.  (One thread per particle over the global lattice, the 96
// gathers of a particle go through the vector L1/L2 and the kernel ran at
// 1.64 ms for 40.9 M particles.)
constexpr int G2B_T = 256;
__global__ __launch_bounds__(G2B_T) void g2p_block_kernel(GridDev g, const float* __restrict__ F,
                                                          const float* __restrict__ spos,
                                                          const float* __restrict__ svel,
                                                          const uint32_t* __restrict__ pcount,
                                                          const uint32_t* __restrict__ poff, float alpha, float dt,
                                                          float* __restrict__ pos_out, float* __restrict__ vel_out,
                                                          uint32_t* __restrict__ list) {
  __shared__ float lat[SLAB];
  const int B = blockIdx.x;
  const uint32_t cnt = pcount[B];
  if (cnt == 0) return;
  {
    // SLAB floats per block: a multiple of 4, and every block's lattice 16-byte aligned
    const float4* src = reinterpret_cast<const float4*>(F + (size_t)B * SLAB);
    float4* dst = reinterpret_cast<float4*>(lat);
    for (int k = threadIdx.x; k < SLAB / 4; k += G2B_T) dst[k] = src[k];
  }
  Near nr;
  block_coords(g.D, B, nr.bc);
  nr.own = true;
  const int o[3] = {BASE * nr.bc[0], BASE * nr.bc[1], BASE * nr.bc[2]};
  __syncthreads();
  const uint32_t p0 = poff[B];
  for (uint32_t j = threadIdx.x; j < cnt; j += G2B_T) {
    const int64_t i = (int64_t)p0 + j;
    bool ok = true;
    float x[3], u[3], v1[3], vd[3], vp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      x[a] = spos[3 * i + a];
      vp[a] = svel[3 * i + a];
      u[a] = fminf(fmaxf(x[a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
    }
    v1[0] = sample_blk<0>(g, lat, F, 0, o, u, nr, ok);
    vd[0] = sample_blk<0>(g, lat, F, 1, o, u, nr, ok);
    v1[1] = sample_blk<1>(g, lat, F, 0, o, u, nr, ok);
    vd[1] = sample_blk<1>(g, lat, F, 1, o, u, nr, ok);
    v1[2] = sample_blk<2>(g, lat, F, 0, o, u, nr, ok);
    vd[2] = sample_blk<2>(g, lat, F, 1, o, u, nr, ok);
    float vn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) vn[a] = alpha * (vp[a] + vd[a]) + (1.f - alpha) * v1[a];
    float xx[3], v2[3], v3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xx[a] = x[a] + 0.5f * dt * v1[a];
      u[a] = fminf(fmaxf(xx[a] * g.inv_h, -1.f), (float)g.N[a] + 1.f);
    }
    v2[0] = sample_blk<0>(g, lat, F, 0, o, u, nr, ok);
    v2[1] = sample_blk<1>(g, lat, F, 0, o, u, nr, ok);
    v2[2] = sample_blk<2>(g, lat, F, 0, o, u, nr, ok);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xx[a] = x[a] + 0.75f * dt * v2[a];
      u[a] = fminf(fmaxf(xx[a] * g.inv_h, -1.f), (float)g.N[a] + 1.f);
    }
    v3[0] = sample_blk<0>(g, lat, F, 0, o, u, nr, ok);
    v3[1] = sample_blk<1>(g, lat, F, 0, o, u, nr, ok);
    v3[2] = sample_blk<2>(g, lat, F, 0, o, u, nr, ok);
    if (!ok) {
      list[1 + atomicAdd(&list[0], 1u)] = (uint32_t)i;
      continue;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float xn = x[a] + dt * ((2.f / 9.f) * v1[a] + (3.f / 9.f) * v2[a] + (4.f / 9.f) * v3[a]);
      const float lo = 0.25f * g.h, hi = (float)g.N[a] * g.h - 0.25f * g.h;
      vel_out[3 * i + a] = vn[a];
      pos_out[3 * i + a] = fminf(fmaxf(xn, lo), hi);
    }
  }
}

// ------------------------------------------------------------- adjoint
// Reverse-mode derivative of the step with the topology, the liquid mask and
// the mass thresholds held at their forward values.  Fine-lattice adjoint
// buffers are block-major [block][field][axis][512] with two fields and a
// per-block stride of ADJ floats.  Scatters into the fine lattice use global
// float atomics (no return); every other stage is a gather.

// address of fine node q of component a (transverse indices clamp to the
// domain), or -1 where the node is the +wall or outside; stride SLAB is the
// padded forward lattice (the node in its owner block's array), any other
// stride a buffer of the blocks' own nodes
__device__ __forceinline__ int64_t fine_node(const GridDev& g, int stride, int field, int a, int* q) {
  if (q[a] < 0 || q[a] >= g.N[a]) return -1;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (d != a) q[d] = iclamp(q[d], 0, g.N[d] - 1);
  const int B = (q[0] >> LOGB) + g.D[0] * ((q[1] >> LOGB) + g.D[1] * (q[2] >> LOGB));
  if (stride == SLAB) {
    const int l[3] = {q[0] & 7, q[1] & 7, q[2] & 7};
    return (int64_t)B * SLAB + field * 3 * PAD + a * PAD + pad_idx(a, l);
  }
  const int l = (q[0] & 7) + 8 * ((q[1] & 7) + 8 * (q[2] & 7));
  return (int64_t)B * stride + field * 3 * FINE + a * FINE + l;
}

// Component a of the field at u (fine units); adds wb * d(value)/du_d to grad.
__device__ __forceinline__ float sample_vjp(const GridDev& g, const float* __restrict__ F, int stride, int field,
                                            int a, const float* u, float wb, float* grad) {
  int i0[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = d == a ? u[d] : u[d] - 0.5f;
    const float fl = floorf(c);
    i0[d] = (int)fl;
    t[d] = c - fl;
  }
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int dd[3] = {k & 1, (k >> 1) & 1, k >> 2};
    int q[3] = {i0[0] + dd[0], i0[1] + dd[1], i0[2] + dd[2]};
    const int64_t ad = fine_node(g, stride, field, a, q);
    v[k] = ad >= 0 ? F[ad] : 0.f;
  }
  const SG r = trilerp_grad(v, t);
  grad[0] = fmaf(wb, r.g0, grad[0]);
  grad[1] = fmaf(wb, r.g1, grad[1]);
  grad[2] = fmaf(wb, r.g2, grad[2]);
  return r.s;
}

// LDS tile of the fine-lattice adjoint around one block: nodes at block-local
// coordinates [-AH, 8 + AH) on every axis, per component, the two fields of
// a node side by side (float2: field 0, field 1).
constexpr int AH = 2, AT = BASE + 2 * AH, AT3 = AT * AT * AT;

// F[node] += w(u) * val (field 0; PAIR: and w(u) * val1 on field 1) on
// component a, into the block's LDS tile where the node lies in it and
// straight into the global buffer (stride ADJ) otherwise.  Node indices are
// resolved as in sampling: normal index outside [0, N) is a dead node,
// transverse indices clamp to [0, N - 1].  The two fields at one point share
// the weights, so PAIR adds both with one 64-bit compare-and-swap.
template <bool PAIR>
__device__ __forceinline__ void tile_scatter(const GridDev& g, float2* tile, float* fbar, uint8_t* touched,
                                             const int* o, int a, const float* u, float val, float val1 = 0.f) {
  if (val == 0.f && (!PAIR || val1 == 0.f)) return;
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
    int q[3] = {i0[0] + dd[0], i0[1] + dd[1], i0[2] + dd[2]};
    if (q[a] < 0 || q[a] >= g.N[a]) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d != a) q[d] = iclamp(q[d], 0, g.N[d] - 1);
    const float w = (dd[0] ? t[0] : 1.f - t[0]) * (dd[1] ? t[1] : 1.f - t[1]) * (dd[2] ? t[2] : 1.f - t[2]);
    const int t0 = q[0] - o[0], t1 = q[1] - o[1], t2 = q[2] - o[2];
    if ((unsigned)t0 < (unsigned)AT && (unsigned)t1 < (unsigned)AT && (unsigned)t2 < (unsigned)AT) {
      float2* cell = &tile[a * AT3 + t0 + AT * (t1 + AT * t2)];
      if (PAIR)
        lds_add2_spread(cell, w * val, w * val1);
      else
        lds_add_spread(&cell->x, w * val);
    } else {
      const int B = (q[0] >> LOGB) + g.D[0] * ((q[1] >> LOGB) + g.D[1] * (q[2] >> LOGB));
      const int l = (q[0] & 7) + 8 * ((q[1] & 7) + 8 * (q[2] & 7));
      if (val != 0.f) unsafeAtomicAdd(fbar + (int64_t)B * ADJ + a * FINE + l, w * val);
      if (PAIR && val1 != 0.f) unsafeAtomicAdd(fbar + (int64_t)B * ADJ + 3 * FINE + a * FINE + l, w * val1);
      touched[B] = 1;
    }
  }
}

// Adjoint of g2p_advect_kernel in two passes.
//
// Stage pass, one thread per sorted particle (the forward kernel's shape):
// recomputes the forward stages, reads dL/dx', dL/dv' of the particle (by
// id), samples the stage fields' gradients, writes its direct partials dL/dx,
// dL/dv, and leaves the field-adjoint scatters it owes as a record: the two
// RK stage positions and the weights of the twelve trilinear scatters.
//
// Scatter pass, one work-group per block over the block's records: the
// scatters go into an LDS tile around the block (field 0 = Fnew-bar +
// Fdel-bar, field 1 = Fdel-bar), flushed once per node at the end.
//
// In one kernel the gradient sampling ran at the occupancy the LDS tile
// allows, each thread walking its particles' serial RK chains one after the
// other; split, the sampling runs at the forward kernel's parallelism.
struct AdjRec {
  float4 q0;   // u2.xyz, u3.x
  float4 q1;   // u3.yz, k3.xy
  float4 q2;   // k3.z, k2.xyz
  float4 q3;   // s1.xyz (field 0 at u), d1.x
  float4 q4;   // d1.yz (field 1 at u), unused
};

// Stage pass of particle i.  Without SLOW only the halo path samples: a
// particle with a stage point in a block without halo goes to `list`
// (list[0] counts) and nothing is written for it; the fix-up pass (SLOW)
// redoes the listed particles with the per-node path.
// Samplers of the stage pass.  GlobalSmp: the padded lattice in global
// memory (SLOW: any point, through the per-node path where a block has no
// lattice; otherwise such a point clears ok).  BlockSmp: one block's lattice
// staged in LDS, points outside it through the global halo path.
template <bool SLOW>
struct GlobalSmp {
  const GridDev& g;
  const float* __restrict__ F;
  Near nr;
  __device__ float comp(int field, int a, const float* u, bool& ok) const {
    return sample_comp<SLOW>(g, F, field, a, u, nr, ok);
  }
  __device__ void grad(int field, int a, const float* u, float* G, bool& ok) const {
    (void)sample_grad_sel<SLOW>(g, F, field, a, u, nr, G, ok);
  }
  __device__ float grad_v(int field, int a, const float* u, float* G, bool& ok) const {
    return sample_grad_sel<SLOW>(g, F, field, a, u, nr, G, ok);
  }
};

// sample_grad_impl<A, true> on one block's padded lattice staged in LDS
// (lat; o: the block's first fine cell per axis), the same sums; false, and
// nothing read, when the 2x2x2 nodes are not all in it
template <int A>
__device__ __forceinline__ bool sample_grad_lds(const GridDev& g, const float* lat, int field, const int* o,
                                                const float* u, SG& r) {
  constexpr int n0 = A == 0 ? 9 : 10, n1 = A == 1 ? 9 : 10, n2 = A == 2 ? 9 : 10;
  int lp[3];
  float t[3], dm[3];
  bool zero = false, in = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int i;
    if (d == A) {
      zero = zero || u[d] < 0.f || u[d] > (float)g.N[d];
      const float c = fminf(fmaxf(u[d], 0.f), (float)g.N[d]);
      i = min((int)floorf(c), g.N[d] - 1);
      t[d] = c - (float)i;
      dm[d] = 1.f;
    } else {
      const float cr = u[d] - 0.5f;
      const float c = fminf(fmaxf(cr, 0.f), (float)(g.N[d] - 1));
      i = (int)floorf(c);
      t[d] = c - (float)i;
      dm[d] = (cr >= 0.f && cr < (float)(g.N[d] - 1)) ? 1.f : 0.f;
    }
    lp[d] = i - o[d] + (d == A ? 0 : 1);
    const int nd = d == 0 ? n0 : (d == 1 ? n1 : n2);
    in = in && lp[d] >= 0 && lp[d] <= nd - 2;
  }
  if (!in) return false;
  const float* p = lat + field * 3 * PAD + A * PAD + lp[0] + n0 * (lp[1] + n1 * lp[2]);
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = p[(k & 1) + n0 * (((k >> 1) & 1) + n1 * (k >> 2))];
  r = trilerp_grad(v, t);
  if (zero) {
    r = SG{0.f, 0.f, 0.f, 0.f};
    return true;
  }
  r.g0 *= dm[0];
  r.g1 *= dm[1];
  r.g2 *= dm[2];
  return true;
}

template <int A>
__device__ __forceinline__ float sample_grad_blk(const GridDev& g, const float* lat, const float* __restrict__ F,
                                                 int field, const int* o, const float* u, const Near& nr, float* G,
                                                 bool& ok) {
  SG r;
  if (sample_grad_lds<A>(g, lat, field, o, u, r)) {
    G[0] = r.g0;
    G[1] = r.g1;
    G[2] = r.g2;
    return r.s;
  }
  return sample_grad_fast<A>(g, F, field, u, nr, G, ok);
}

struct BlockSmp {
  const GridDev& g;
  const float* __restrict__ F;
  const float* lat;
  const int* o;
  Near nr;
  __device__ float comp(int field, int a, const float* u, bool& ok) const {
    return a == 0 ? sample_blk<0>(g, lat, F, field, o, u, nr, ok)
                  : (a == 1 ? sample_blk<1>(g, lat, F, field, o, u, nr, ok) : sample_blk<2>(g, lat, F, field, o, u, nr, ok));
  }
  __device__ float grad_v(int field, int a, const float* u, float* G, bool& ok) const {
    return a == 0 ? sample_grad_blk<0>(g, lat, F, field, o, u, nr, G, ok)
                  : (a == 1 ? sample_grad_blk<1>(g, lat, F, field, o, u, nr, G, ok)
                            : sample_grad_blk<2>(g, lat, F, field, o, u, nr, G, ok));
  }
  __device__ void grad(int field, int a, const float* u, float* G, bool& ok) const { (void)grad_v(field, a, u, G, ok); }
};

// S samples the forward lattice (GlobalSmp or BlockSmp); without SLOW a
// particle with a point S cannot serve goes to the fix-up list.
template <bool SLOW, class Smp>
__device__ __forceinline__ void g2p_adjoint_one(const GridDev& g, const Smp& S, const float* __restrict__ spos,
                                                const int32_t* __restrict__ sids, int64_t i, float alpha, float dt,
                                                float* __restrict__ xbar, float* __restrict__ vbar,
                                                AdjRec* __restrict__ rec, uint32_t* __restrict__ list) {
  const float umax[3] = {(float)g.N[0] - 0.0009765625f, (float)g.N[1] - 0.0009765625f,
                         (float)g.N[2] - 0.0009765625f};
  const int64_t id = sids[i];
  bool ok = true;
  // live ranges kept short (the masks follow from the clamped coordinates,
  // the RK sum accumulates as the stages finish, the record is written as
  // its parts become final): the kernel is held to 128 VGPRs
  float x[3], u[3], u2[3], u3[3], xn[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    x[a] = spos[3 * i + a];
    u[a] = fminf(fmaxf(x[a] * g.inv_h, 0.f), umax[a]);
  }
  // forward stages: x2 = x + 1/2 dt v1, x3 = x + 3/4 dt v2, x' = x + dt (2 v1 + 3 v2 + 4 v3) / 9
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float v1 = S.comp(0, a, u, ok);
    u2[a] = fminf(fmaxf((x[a] + 0.5f * dt * v1) * g.inv_h, -1.f), (float)g.N[a] + 1.f);
    xn[a] = (2.f / 9.f) * v1;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float v2 = S.comp(0, a, u2, ok);
    u3[a] = fminf(fmaxf((x[a] + 0.75f * dt * v2) * g.inv_h, -1.f), (float)g.N[a] + 1.f);
    xn[a] += (3.f / 9.f) * v2;
  }
  // stage 3 gradient G[a][d] = d v3_a / d u3_d
  float G[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float v3 = S.grad_v(0, a, u3, G[a], ok);
    xn[a] = x[a] + dt * (xn[a] + (4.f / 9.f) * v3);
  }
  float k1[3], k2[3], k3[3], xacc[3], vo[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lo = 0.25f * g.h, hi = (float)g.N[a] * g.h - 0.25f * g.h;
    const float xb = (xn[a] > lo && xn[a] < hi) ? xbar[3 * id + a] : 0.f;
    vo[a] = vbar[3 * id + a];
    k1[a] = dt * (2.f / 9.f) * xb;
    k2[a] = dt * (3.f / 9.f) * xb;
    k3[a] = dt * (4.f / 9.f) * xb;
    xacc[a] = xb;
  }
  // stage 3 (the coordinate mask: the unclamped stage coordinate lies inside
  // (-1, N + 1), i.e. the clamped one is off both bounds)
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float m3 = (u3[d] > -1.f && u3[d] < (float)g.N[d] + 1.f) ? 1.f : 0.f;
    const float b = (k3[0] * G[0][d] + k3[1] * G[1][d] + k3[2] * G[2][d]) * g.inv_h * m3;
    xacc[d] += b;
    k2[d] += 0.75f * dt * b;
  }
  rec[i].q0 = make_float4(u2[0], u2[1], u2[2], u3[0]);
  rec[i].q1 = make_float4(u3[1], u3[2], k3[0], k3[1]);
  rec[i].q2 = make_float4(k3[2], k2[0], k2[1], k2[2]);
  // stage 2
#pragma unroll
  for (int a = 0; a < 3; ++a) S.grad(0, a, u2, G[a], ok);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float m2 = (u2[d] > -1.f && u2[d] < (float)g.N[d] + 1.f) ? 1.f : 0.f;
    const float b = (k2[0] * G[0][d] + k2[1] * G[1][d] + k2[2] * G[2][d]) * g.inv_h * m2;
    xacc[d] += b;
    k1[d] += 0.5f * dt * b;
  }
  // FLIP/PIC blend and stage 1 at x
  float gr[3] = {0.f, 0.f, 0.f}, s1[3], d1[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float vnb = k1[a] + (1.f - alpha) * vo[a], vdb = alpha * vo[a];
    s1[a] = vnb + vdb;
    d1[a] = vdb;
    float Gn[3], Gd[3];
    S.grad(0, a, u, Gn, ok);
    S.grad(1, a, u, Gd, ok);
#pragma unroll
    for (int d = 0; d < 3; ++d) gr[d] += vnb * Gn[d] + vdb * Gd[d];
  }
  if (!SLOW && !ok) {
    // a stage point without halo: the fix-up pass redoes the particle
    list[1 + atomicAdd(&list[0], 1u)] = (uint32_t)i;
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float m1 = (u[d] > 0.f && u[d] < umax[d]) ? 1.f : 0.f;
    xbar[3 * id + d] = xacc[d] + gr[d] * g.inv_h * m1;
    vbar[3 * id + d] = d1[d];
  }
  rec[i].q3 = make_float4(s1[0], s1[1], s1[2], d1[0]);
  rec[i].q4 = make_float4(d1[1], d1[2], 0.f, 0.f);
}

template <bool FIXUP>
__global__ __launch_bounds__(TPB) void g2p_adjoint_stage_kernel(GridDev g, const float* __restrict__ F,
                                                                const float* __restrict__ spos,
                                                                const int32_t* __restrict__ sids, int64_t n,
                                                                float alpha, float dt, float* __restrict__ xbar,
                                                                float* __restrict__ vbar, AdjRec* __restrict__ rec,
                                                                uint32_t* __restrict__ list) {
  const int64_t i = (FIXUP ? (int64_t)blockIdx.x : xcd_group(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
  if (FIXUP) {
    const int64_t m = (int64_t)list[0];
    for (int64_t j = i; j < m; j += (int64_t)gridDim.x * blockDim.x) {
      const int64_t q = (int64_t)list[1 + j];
      float u[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) u[a] = fminf(fmaxf(spos[3 * q + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
      const GlobalSmp<true> S{g, F, near_of(g, u)};
      g2p_adjoint_one<true>(g, S, spos, sids, q, alpha, dt, xbar, vbar, rec, list);
    }
    return;
  }
  if (i >= n) return;
  float u[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) u[a] = fminf(fmaxf(spos[3 * i + a] * g.inv_h, 0.f), (float)g.N[a] - 0.0009765625f);
  const GlobalSmp<false> S{g, F, near_of(g, u)};
  g2p_adjoint_one<false>(g, S, spos, sids, i, alpha, dt, xbar, vbar, rec, list);
}

// The stage pass of the step's block-sorted particles, one work-group per
// block with the block's forward lattice (both fields) in LDS, as the
// forward's g2p_block_kernel; the same arithmetic as g2p_adjoint_stage_kernel.
__global__ __launch_bounds__(G2B_T) __attribute__((amdgpu_waves_per_eu(4))) void g2p_adjoint_block_kernel(GridDev g, const float* __restrict__ F,
                                                                  const float* __restrict__ spos,
                                                                  const int32_t* __restrict__ sids,
                                                                  const uint32_t* __restrict__ pcount,
                                                                  const uint32_t* __restrict__ poff, float alpha,
                                                                  float dt, float* __restrict__ xbar,
                                                                  float* __restrict__ vbar, AdjRec* __restrict__ rec,
                                                                  uint32_t* __restrict__ list) {
  __shared__ float lat[SLAB];
  const int B = blockIdx.x;
  const uint32_t cnt = pcount[B];
  if (cnt == 0) return;
  {
    const float4* src = reinterpret_cast<const float4*>(F + (size_t)B * SLAB);
    float4* dst = reinterpret_cast<float4*>(lat);
    for (int k = threadIdx.x; k < SLAB / 4; k += G2B_T) dst[k] = src[k];
  }
  Near nr;
  block_coords(g.D, B, nr.bc);
  nr.own = true;
  const int o[3] = {BASE * nr.bc[0], BASE * nr.bc[1], BASE * nr.bc[2]};
  __syncthreads();
  const BlockSmp S{g, F, lat, o, nr};
  const uint32_t p0 = poff[B];
  for (uint32_t j = threadIdx.x; j < cnt; j += G2B_T)
    g2p_adjoint_one<false>(g, S, spos, sids, (int64_t)p0 + j, alpha, dt, xbar, vbar, rec, list);
}

// 512 threads per work-group (the LDS tile leaves room for three per CU, 24
// waves to cover the latency of the compare-and-swap adds): measured at
// 512^3, 16.8 ms against 24.5 ms with 256 threads and 20.4 ms with 1024.
// (Both fields of a node are held side by side: the samples at the particle
// position add both with one 64-bit compare-and-swap, the two other stages,
// field 0 only, with a 32-bit one on the pair's first word.)
constexpr int ADJ_T = 512;

__global__ __launch_bounds__(ADJ_T) void g2p_adjoint_scatter_kernel(GridDev g, const float* __restrict__ spos,
                                                                  const uint32_t* __restrict__ pcount,
                                                                  const uint32_t* __restrict__ poff,
                                                                  const AdjRec* __restrict__ rec, float* fbar,
                                                                  uint8_t* touched) {
  __shared__ float2 tile[3 * AT3];
  const int B = blockIdx.x;
  const uint32_t cnt = pcount[B];
  if (cnt == 0) return;
  for (int k = threadIdx.x; k < 3 * AT3; k += ADJ_T) tile[k] = make_float2(0.f, 0.f);
  int o[3];
  block_coords(g.D, B, o);
#pragma unroll
  for (int d = 0; d < 3; ++d) o[d] = o[d] * BASE - AH;
  __syncthreads();
  const float umax[3] = {(float)g.N[0] - 0.0009765625f, (float)g.N[1] - 0.0009765625f,
                         (float)g.N[2] - 0.0009765625f};
  // particle order as in the deposit: windows of ADJ_WIN particles, lane l of
  // wave v taking ADJ_S l + ADJ_SPW v + (0 .. ADJ_SPW-1), so the lanes of a
  // wave hold particles far apart in the sorted order (particles of one cell
  // are often neighbours there, and their tile adds would meet on a word)
  // while a work-group's records stay within one window (15.7 ms against
  // 16.8 ms at 512^3 for one contiguous run per thread)
  constexpr uint32_t ADJ_S = 16, ADJ_WIN = 64 * ADJ_S, ADJ_SPW = ADJ_WIN / ADJ_T;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t nit = (cnt + ADJ_WIN - 1) / ADJ_WIN * ADJ_SPW;
  for (uint32_t it = 0; it < nit; ++it) {
    const uint32_t j = (it / ADJ_SPW) * ADJ_WIN + lane * ADJ_S + wv * ADJ_SPW + it % ADJ_SPW;
    if (j >= cnt) continue;
    const int64_t i = (int64_t)poff[B] + j;
    float u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = fminf(fmaxf(spos[3 * i + a] * g.inv_h, 0.f), umax[a]);
    const AdjRec r = rec[i];
    const float u2[3] = {r.q0.x, r.q0.y, r.q0.z}, u3[3] = {r.q0.w, r.q1.x, r.q1.y};
    const float k3[3] = {r.q1.z, r.q1.w, r.q2.x}, k2[3] = {r.q2.y, r.q2.z, r.q2.w};
    const float s1[3] = {r.q3.x, r.q3.y, r.q3.z}, d1[3] = {r.q3.w, r.q4.x, r.q4.y};
#pragma unroll
    for (int a = 0; a < 3; ++a) tile_scatter<false>(g, tile, fbar, touched, o, a, u3, k3[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) tile_scatter<false>(g, tile, fbar, touched, o, a, u2, k2[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) tile_scatter<true>(g, tile, fbar, touched, o, a, u, s1[a], d1[a]);
  }
  __syncthreads();
  // flush: tile nodes are in-domain (clamped) by construction except the
  // normal-axis ones past the +wall, which were never written
  for (int k = threadIdx.x; k < 3 * AT3; k += ADJ_T) {
    const float2 v = tile[k];
    if (v.x == 0.f && v.y == 0.f) continue;
    const int a = k / AT3, r = k - a * AT3;
    const int q[3] = {o[0] + r % AT, o[1] + (r / AT) % AT, o[2] + r / (AT * AT)};
    const int Bq = (q[0] >> LOGB) + g.D[0] * ((q[1] >> LOGB) + g.D[1] * (q[2] >> LOGB));
    const int l = (q[0] & 7) + 8 * ((q[1] & 7) + 8 * (q[2] & 7));
    float* dst = fbar + (int64_t)Bq * ADJ + a * FINE + l;
    touched[Bq] = 1;
    if (v.x != 0.f) unsafeAtomicAdd(dst, v.x);
    if (v.y != 0.f) unsafeAtomicAdd(dst + 3 * FINE, v.y);
  }
}

// Zero the fine-lattice adjoint of the blocks the last scatter touched and
// clear their flags (one wave per block; every other block's adjoint is
// already 0), so that the buffer starts each G2P adjoint at zero without a
// pass over every block's nodes.
__global__ __launch_bounds__(TPB) void fbar_clear_kernel(int nb, float* __restrict__ fbar, uint8_t* touched) {
  const int B = blockIdx.x * (TPB / 64) + (int)(threadIdx.x >> 6);
  if (B >= nb || !touched[B]) return;
  float4* p = reinterpret_cast<float4*>(fbar + (size_t)B * ADJ);   // ADJ floats, a multiple of 4
  for (int k = threadIdx.x & 63; k < ADJ / 4; k += 64) p[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  if ((threadIdx.x & 63) == 0) touched[B] = 0;
}

// Pi^T of the two fine-lattice adjoint fields onto the block's own slots:
// unb = Pi^T field 0, uob = -Pi^T field 1.
__global__ __launch_bounds__(TPB) void pi_transpose_kernel(GridDev g, const float* __restrict__ fbar,
                                                           const uint8_t* __restrict__ touched, FacePtrs unb,
                                                           FacePtrs uob) {
  __shared__ float nA[3 * 576], nB[3 * 576];
  __shared__ float tA[3 * 512], tB[3 * 512];
  const int B = (int)xcd_group(blockIdx.x, gridDim.x), tid = threadIdx.x;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  if (!touched[B]) {
    // nothing was scattered into this block: its fine-lattice adjoint is 0
    for (int k = tid; k < 3 * nf; k += TPB) {
      const int a = k / nf, f = k - a * nf;
      unb.p[a][g.foff[a][B] + f] = 0.f;
      uob.p[a][g.foff[a][B] + f] = 0.f;
    }
    return;
  }
  const float* src = fbar + (size_t)B * ADJ;
  if (e > 0) {
    pi_transpose_sep(
        e, [&](int f, int a, int li) { return src[f * 3 * FINE + a * FINE + li]; }, nA, nB, tA, tB,
        [&](int a, int f, float v0, float v1) {
          unb.p[a][g.foff[a][B] + f] = v0;
          uob.p[a][g.foff[a][B] + f] = -v1;
        });
    return;
  }
  // level 0: own fine nodes are the native slots, one to one
  for (int k = tid; k < 3 * 576; k += TPB) {
    nA[k] = 0.f;
    nB[k] = 0.f;
  }
  __syncthreads();
  for (int k = tid; k < 3 * FINE; k += TPB) {
    const float fa = src[k], fb = src[3 * FINE + k];
    if (fa == 0.f && fb == 0.f) continue;
    const int a = k >> 9, li = k & 511;
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    int slot[8];
    float w[8];
    (void)pi_weights(a, e, l, slot, w);
    nA[a * 576 + slot[0]] += w[0] * fa;
    nB[a * 576 + slot[0]] += w[0] * fb;
  }
  __syncthreads();
  for (int k = tid; k < 3 * nf; k += TPB) {
    const int a = k / nf, f = k - a * nf;
    unb.p[a][g.foff[a][B] + f] = nA[a * 576 + f];
    uob.p[a][g.foff[a][B] + f] = -nB[a * 576 + f];
  }
}

// Adjoint of project_kernel.  Per slot: the u* terms are gathered (own
// formula, plus 1/4 of the coarse neighbour's average slot on the fine side
// of a level change) into usb; the pressure terms are scattered into pbar.
__global__ __launch_bounds__(TPB) void project_adjoint_kernel(GridDev g, FacePtrs beta, FacePtrs unb, FacePtrs usb,
                                                              float* pbar) {
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  const float s = (float)(1 << e);
  const uint32_t co = g.coff[B];
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int sl[3];
    face_decode(a, f, r, sl);
    const size_t id = g.foff[a][B] + f;
    const float ub = unb.p[a][id], bt = beta.p[a][id];
    float direct = 0.f;
    if (sl[a] > 0 && sl[a] < r) {
      direct = ub;
      if (ub != 0.f) {
        int cm[3] = {sl[0], sl[1], sl[2]};
        cm[a] -= 1;
        unsafeAtomicAdd(pbar + co + cell_idx(sl, r), -bt / s * ub);
        unsafeAtomicAdd(pbar + co + cell_idx(cm, r), bt / s * ub);
      }
    } else {
      const int side = sl[a] == r;
      const int N = g.nbr[B * 6 + 2 * a + side];
      if (N >= 0) {
        const int en = g.level[N], rn = BASE >> en;
        const float sg = side ? -1.f : 1.f;  // u = u* - bt sg (p_own - p_nbr) / d
        int own[3] = {sl[0], sl[1], sl[2]};
        own[a] = side ? r - 1 : 0;
        const uint32_t po = co + cell_idx(own, r);
        int q[3] = {sl[0], sl[1], sl[2]};
        q[a] = side ? 0 : rn - 1;
        if (en == e || en == e + 1) {
          float d = s;
          direct = ub;
          if (en == e + 1) {
            q[t1] = sl[t1] >> 1;
            q[t2] = sl[t2] >> 1;
            d = 1.5f * s;
            int ps[3] = {sl[0], sl[1], sl[2]};
            ps[a] = side ? 0 : rn;
            ps[t1] = q[t1];
            ps[t2] = q[t2];
            direct += 0.25f * unb.p[a][g.foff[a][N] + face_idx(a, ps, rn)];
          }
          if (ub != 0.f) {
            const float c = bt * sg / d * ub;
            unsafeAtomicAdd(pbar + po, -c);
            unsafeAtomicAdd(pbar + g.coff[N] + cell_idx(q, rn), c);
          }
        } else if (ub != 0.f) {
          const float sf = 0.5f * s;
          int ps[3];
          ps[a] = side ? 0 : rn;
          for (int d1 = 0; d1 < 2; ++d1)
            for (int d2 = 0; d2 < 2; ++d2) {
              q[t1] = ps[t1] = 2 * sl[t1] + d1;
              q[t2] = ps[t2] = 2 * sl[t2] + d2;
              const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
              const float c = 0.25f * beta.p[a][pid] * sg / (1.5f * sf) * ub;
              unsafeAtomicAdd(pbar + po, -c);
              unsafeAtomicAdd(pbar + g.coff[N] + cell_idx(q, rn), c);
            }
        }
      }
    }
    usb.p[a][id] = direct;
  }
}

// usb += D^T rhsbar: rhs_c = -s^2 sum_a (u[c+] - u[c-]) on liquid cells
__global__ __launch_bounds__(TPB) void divergence_adjoint_kernel(GridDev g, const uint8_t* __restrict__ liquid,
                                                                 const float* __restrict__ rhsbar, FacePtrs usb) {
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  const float s2 = (float)(1 << e) * (float)(1 << e);
  const uint32_t co = g.coff[B];
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    int sl[3];
    face_decode(a, f, r, sl);
    float acc = 0.f;
    if (sl[a] > 0) {
      int cm[3] = {sl[0], sl[1], sl[2]};
      cm[a] -= 1;
      const uint32_t c = co + cell_idx(cm, r);
      if (liquid[c]) acc -= s2 * rhsbar[c];
    }
    if (sl[a] < r) {
      const uint32_t c = co + cell_idx(sl, r);
      if (liquid[c]) acc += s2 * rhsbar[c];
    }
    usb.p[a][g.foff[a][B] + f] += acc;
  }
}

// Adjoint of fold_kernel.  Each unknown face value u = M'/m' (m' > 1e-6)
// collects the adjoints of every slot that copies it (lambda); the slots
// whose accumulators enter M', m' with weight c receive c lambda / m' and
// -c lambda u / m'.  u and m' are the forward's u_old and mass.
__global__ __launch_bounds__(TPB) void fold_adjoint_kernel(GridDev g, FacePtrs uold, FacePtrs mass, FacePtrs uob,
                                                           FacePtrs usb, FacePtrs Mb, FacePtrs mb) {
  const int B = (int)xcd_group(blockIdx.x, gridDim.x);
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, f = k - a * nf;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int sl[3];
    face_decode(a, f, r, sl);
    const size_t id = g.foff[a][B] + f;
    const float ub = uob.p[a][id] + usb.p[a][id];
    float oM = 0.f, om = 0.f;
    if (sl[a] > 0 && sl[a] < r) {
      const float mt = mass.p[a][id];
      if (mt > 1e-6f) {
        oM = ub / mt;
        om = -ub * uold.p[a][id] / mt;
      }
    } else {
      const int side = sl[a] == r;
      const int N = g.nbr[B * 6 + 2 * a + side];
      if (N >= 0) {
        const int en = g.level[N], rn = BASE >> en;
        int ps[3] = {sl[0], sl[1], sl[2]};
        ps[a] = side ? 0 : rn;
        if (en == e || en == e + 1) {
          float cp = 1.f;
          if (en == e + 1) {
            ps[t1] = sl[t1] >> 1;
            ps[t2] = sl[t2] >> 1;
            cp = 0.25f;
          }
          const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
          const float lam = ub + cp * (uob.p[a][pid] + usb.p[a][pid]);
          const float mt = mass.p[a][id];
          if (mt > 1e-6f) {
            oM = lam / mt;
            om = -lam * uold.p[a][id] / mt;
          }
        } else {
          for (int d1 = 0; d1 < 2; ++d1)
            for (int d2 = 0; d2 < 2; ++d2) {
              ps[t1] = 2 * sl[t1] + d1;
              ps[t2] = 2 * sl[t2] + d2;
              const size_t pid = g.foff[a][N] + face_idx(a, ps, rn);
              const float mt = mass.p[a][pid];
              if (mt <= 1e-6f) continue;
              const float lam = uob.p[a][pid] + usb.p[a][pid] + 0.25f * ub;
              oM += 0.25f * lam / mt;
              om -= 0.25f * lam * uold.p[a][pid] / mt;
            }
        }
      }
    }
    Mb.p[a][id] = oM;
    mb.p[a][id] = om;
  }
}

// Pi of two per-slot fields onto the block's own fine nodes (fields 0, 1),
// for the blocks flagged `halo` (a particle within one block): the particle
// samples that read the result (the deposit adjoint) take their nodes from
// the particle's block and its neighbours only
__global__ __launch_bounds__(TPB) void pi_forward2_kernel(GridDev g, FacePtrs A, FacePtrs Bf, float* __restrict__ out) {
  // the block's slots of both fields, read once with coalesced loads (the
  // nodes' weighted sums then read LDS, in pi_weights's order as before)
  __shared__ float sA[3 * 576], sB[3 * 576];   // (r + 1) r^2 <= 576 slots per axis
  const int B = blockIdx.x;
  if (!g.halo[B]) return;
  const int e = g.level[B], r = BASE >> e, nf = (r + 1) * r * r;
  for (int k = threadIdx.x; k < 3 * nf; k += blockDim.x) {
    const int a = k / nf, i = k - a * nf;
    sA[a * 576 + i] = A.p[a][g.foff[a][B] + i];
    sB[a * 576 + i] = Bf.p[a][g.foff[a][B] + i];
  }
  __syncthreads();
  float* dst = out + (size_t)B * ADJ;
  for (int k = threadIdx.x; k < 3 * FINE; k += blockDim.x) {
    const int a = k >> 9, li = k & 511;
    const int l[3] = {li & 7, (li >> 3) & 7, li >> 6};
    int slot[8];
    float w[8];
    const int nw = pi_weights(a, e, l, slot, w);
    float va = 0.f, vb = 0.f;
    for (int j = 0; j < nw; ++j) {
      va += w[j] * sA[a * 576 + slot[j]];
      vb += w[j] * sB[a * 576 + slot[j]];
    }
    dst[k] = va;
    dst[3 * FINE + k] = vb;
  }
}

// Adjoint of the deposit per particle: vbar += S(FMbar)(x / h),
// xbar += (v . dS(FMbar) + dS(Fmbar)) / h  (fm field 0 = FMbar, 1 = Fmbar)
__global__ __launch_bounds__(TPB) void p2g_adjoint_kernel(GridDev g, const float* __restrict__ fm,
                                                          const float* __restrict__ spos,
                                                          const float* __restrict__ svel,
                                                          const int32_t* __restrict__ sids, int64_t n,
                                                          float* __restrict__ xbar, float* __restrict__ vbar) {
  const int64_t i = xcd_group(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t id = sids[i];
  float u[3], m1[3], gr[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float umax = (float)g.N[a] - 0.0009765625f;
    const float raw = spos[3 * i + a] * g.inv_h;
    u[a] = fminf(fmaxf(raw, 0.f), umax);
    m1[a] = (raw > 0.f && raw < umax) ? 1.f : 0.f;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float sv = sample_vjp(g, fm, ADJ, 0, a, u, svel[3 * i + a], gr);
    sample_vjp(g, fm, ADJ, 1, a, u, 1.f, gr);
    vbar[3 * id + a] += sv;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) xbar[3 * id + d] += gr[d] * g.inv_h * m1[d];
}

// sample_vjp on one block's padded copy of the two-field fine lattice in LDS
// (lat: field f, component a at f 3 PAD + a PAD + pad_idx, every node already
// resolved by fine_node's rule; o: the block's first fine cell per axis), the
// same sums; false, and nothing read, when a node is not in it.
template <int A>
__device__ __forceinline__ bool sample_vjp_lds(const float* lat, const int* o, int field, const float* u, float wb,
                                               float* grad, float& out) {
  constexpr int n0 = A == 0 ? 9 : 10, n1 = A == 1 ? 9 : 10, n2 = A == 2 ? 9 : 10;
  int lp[3];
  float t[3];
  bool in = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = d == A ? u[d] : u[d] - 0.5f;
    const float fl = floorf(c);
    lp[d] = (int)fl - o[d] + (d == A ? 0 : 1);
    t[d] = c - fl;
    const int nd = d == 0 ? n0 : (d == 1 ? n1 : n2);
    in = in && lp[d] >= 0 && lp[d] <= nd - 2;
  }
  if (!in) return false;
  const float* p = lat + field * 3 * PAD + A * PAD + lp[0] + n0 * (lp[1] + n1 * lp[2]);
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = p[(k & 1) + n0 * (((k >> 1) & 1) + n1 * (k >> 2))];
  const SG r = trilerp_grad(v, t);
  grad[0] = fmaf(wb, r.g0, grad[0]);
  grad[1] = fmaf(wb, r.g1, grad[1]);
  grad[2] = fmaf(wb, r.g2, grad[2]);
  out = r.s;
  return true;
}

template <int A>
__device__ __forceinline__ float sample_vjp_blk(const GridDev& g, const float* lat, const float* __restrict__ fm,
                                                const int* o, int field, const float* u, float wb, float* grad) {
  float s;
  if (sample_vjp_lds<A>(lat, o, field, u, wb, grad, s)) return s;
  return sample_vjp(g, fm, ADJ, field, A, u, wb, grad);
}

// p2g_adjoint_kernel for the block-sorted particles, one work-group per block:
// the block's padded copy of the two fields (its own nodes and a one-node
// halo from the neighbours, 21.6 KB) is gathered into LDS once, and the
// block's particles sample it there (the same sums; a particle outside the
// block samples the global lattice).
__global__ __launch_bounds__(G2B_T) void p2g_adjoint_block_kernel(GridDev g, const float* __restrict__ fm,
                                                                  const float* __restrict__ spos,
                                                                  const float* __restrict__ svel,
                                                                  const int32_t* __restrict__ sids,
                                                                  const uint32_t* __restrict__ pcount,
                                                                  const uint32_t* __restrict__ poff,
                                                                  float* __restrict__ xbar, float* __restrict__ vbar) {
  __shared__ float lat[SLAB];
  const int B = blockIdx.x;
  const uint32_t cnt = pcount[B];
  if (cnt == 0) return;
  int o[3];
  block_coords(g.D, B, o);
#pragma unroll
  for (int d = 0; d < 3; ++d) o[d] *= BASE;
  for (int k = threadIdx.x; k < SLAB; k += G2B_T) {
    const int f = k / (3 * PAD), r = k - f * 3 * PAD;
    const int a = r / PAD, pi = r - a * PAD;
    const int n0 = a == 0 ? 9 : 10, n1 = a == 1 ? 9 : 10;
    int q[3] = {o[0] + pi % n0 - (a == 0 ? 0 : 1), o[1] + (pi / n0) % n1 - (a == 1 ? 0 : 1),
                o[2] + pi / (n0 * n1) - (a == 2 ? 0 : 1)};
    const int64_t ad = fine_node(g, ADJ, f, a, q);
    lat[k] = ad >= 0 ? fm[ad] : 0.f;
  }
  __syncthreads();
  const uint32_t p0 = poff[B];
  for (uint32_t j = threadIdx.x; j < cnt; j += G2B_T) {
    const int64_t i = (int64_t)p0 + j;
    const int64_t id = sids[i];
    float u[3], m1[3], gr[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float umax = (float)g.N[a] - 0.0009765625f;
      const float raw = spos[3 * i + a] * g.inv_h;
      u[a] = fminf(fmaxf(raw, 0.f), umax);
      m1[a] = (raw > 0.f && raw < umax) ? 1.f : 0.f;
    }
    float sv[3];
    sv[0] = sample_vjp_blk<0>(g, lat, fm, o, 0, u, svel[3 * i + 0], gr);
    sample_vjp_blk<0>(g, lat, fm, o, 1, u, 1.f, gr);
    sv[1] = sample_vjp_blk<1>(g, lat, fm, o, 0, u, svel[3 * i + 1], gr);
    sample_vjp_blk<1>(g, lat, fm, o, 1, u, 1.f, gr);
    sv[2] = sample_vjp_blk<2>(g, lat, fm, o, 0, u, svel[3 * i + 2], gr);
    sample_vjp_blk<2>(g, lat, fm, o, 1, u, 1.f, gr);
#pragma unroll
    for (int a = 0; a < 3; ++a) vbar[3 * id + a] += sv[a];
#pragma unroll
    for (int d = 0; d < 3; ++d) xbar[3 * id + d] += gr[d] * g.inv_h * m1[d];
  }
}

// out[0] = min id, out[1] = max id (out preset to INT_MAX, INT_MIN)
__global__ __launch_bounds__(TPB) void ids_range_kernel(const int32_t* __restrict__ ids, int64_t n, int* out) {
  int lo = 0x7FFFFFFF, hi = -0x7FFFFFFF - 1;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    lo = min(lo, ids[i]);
    hi = max(hi, ids[i]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_down(lo, off, 64));
    hi = max(hi, __shfl_down(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(out, lo);
    atomicMax(out + 1, hi);
  }
}

__global__ void masked_copy_kernel(float* __restrict__ dst, const float* __restrict__ src,
                                   const uint8_t* __restrict__ liq, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = liq[i] ? src[i] : 0.f;
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

__device__ __forceinline__ void cell_ref(const LevelDev& L, int B, int li, CellRef& o) {
  o.B = B;
  o.r = L.bs >> L.e[B];
  o.lr = 31 - __clz(o.r);
  o.q = 0;
  o.li = li;
  o.l[0] = li & (o.r - 1);
  o.l[1] = (li >> o.lr) & (o.r - 1);
  o.l[2] = li >> (2 * o.lr);
  o.c = L.coff[B] + li;
}

// 7-point adaptive Poisson stencil of one cell, split into its parts:
// diag = sum of all face coefficients; ext = sum of coefficient * value over
// neighbours in other blocks, read through outb(N, local coords, res, cell);
// w[2a+side] = coefficient of the in-block neighbour on that side (0: none).
// Coefficients: beta * s between equal cells (s = cell edge in fine cells),
// beta * s_f / 1.5 across a coarse-fine interface, one term per fine
// sub-face; tank walls carry beta = 0.
template <class Out, typename T>
__device__ __forceinline__ void stencil_split(const LevelDev& L, const CellRef& cr, Out outb, T& diag, T& ext, T* w) {
  const int B = cr.B, r = cr.r;
  const int e = L.e[B];
  const T s = (T)(1 << e);
  diag = 0;
  ext = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const uint32_t fo = L.foff[a][B];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int d = 2 * a + side;
      w[d] = 0;
      const int nl = cr.l[a] + (side ? 1 : -1);
      int sl[3] = {cr.l[0], cr.l[1], cr.l[2]};
      sl[a] += side;
      if (nl >= 0 && nl < r) {
        const T c = (T)L.beta[a][fo + face_idx(a, sl, r)] * s;
        w[d] = c;
        diag += c;
        continue;
      }
      const int N = L.nbr[B * 6 + d];
      if (N < 0) continue;
      const int en = L.e[N], rn = L.bs >> en;
      int nc[3] = {cr.l[0], cr.l[1], cr.l[2]};
      nc[a] = side ? 0 : rn - 1;
      if (en == e) {
        const T c = (T)L.beta[a][fo + face_idx(a, sl, r)] * s;
        diag += c;
        ext += c * outb(N, nc, rn, L.coff[N] + cell_idx(nc, rn));
      } else if (en > e) {
        const T c = (T)L.beta[a][fo + face_idx(a, sl, r)] * s / (T)1.5;
        nc[t1] = cr.l[t1] >> 1;
        nc[t2] = cr.l[t2] >> 1;
        diag += c;
        ext += c * outb(N, nc, rn, L.coff[N] + cell_idx(nc, rn));
      } else {
        int ps[3];
        ps[a] = side ? 0 : rn;
        const uint32_t fon = L.foff[a][N], con = L.coff[N];
        const T sf = (T)0.5 * s;
#pragma unroll
        for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
          for (int d2 = 0; d2 < 2; ++d2) {
            nc[t1] = ps[t1] = 2 * cr.l[t1] + d1;
            nc[t2] = ps[t2] = 2 * cr.l[t2] + d2;
            const T c = (T)L.beta[a][fon + face_idx(a, ps, rn)] * sf / (T)1.5;
            diag += c;
            ext += c * outb(N, nc, rn, con + cell_idx(nc, rn));
          }
      }
    }
  }
}

// group g of a launch from its run table (scalar arithmetic only)
__device__ __forceinline__ WG wg_at(const WGMap& m, int g) {
  int k = 0;
#pragma unroll
  for (int j = 1; j < WGMAP_RUNS; ++j) k += g >= m.g0[j] ? 1 : 0;
  const int lr = m.lr[k], per = 512 >> (3 * lr);
  const int q0 = m.s0[k] + (g - m.g0[k]) * per;
  WG w;
  w.start = q0;
  w.count = min(per, m.s0[k] + m.cnt[k] - q0);
  w.lr = lr;
  w.pad = 0;
  return w;
}

// Descriptor of one work-list entry, held in registers: read with six
// 16-byte loads at the start of a stencil kernel, so that every data load of
// the cell (its value, mask, right-hand side, face coefficients and the
// neighbour values) depends on that one load and all of them are in flight
// together.  Read lazily instead (a load per field, behind the branches that
// use it), the smoother's level-0 phase waited on about a dozen dependent
// loads per thread.
struct BDR {
  uint32_t B, coff, foff, m0, m1;
  int nbr[6];
  uint32_t ncoff[6], nfoff[6];
};

typedef const __attribute__((address_space(4))) uint32_t* ConstU32;

// UNIFORM: p is the same for every thread; the table is read through the
// constant address space (it does not change during a launch), so the
// descriptor is loaded with scalar loads into scalar registers.
template <bool UNIFORM>
__device__ __forceinline__ void bd_read(const uint32_t* p, BDR& d) {
  uint32_t v[BD_WORDS];
#pragma unroll
  for (int k = 0; k < BD_WORDS / 4; ++k) {
    if (UNIFORM) {
      const ConstU32 c = (ConstU32)p + 4 * k;
      v[4 * k] = c[0];
      v[4 * k + 1] = c[1];
      v[4 * k + 2] = c[2];
      v[4 * k + 3] = c[3];
    } else {
      const uint4 t = reinterpret_cast<const uint4*>(p)[k];
      v[4 * k] = t.x;
      v[4 * k + 1] = t.y;
      v[4 * k + 2] = t.z;
      v[4 * k + 3] = t.w;
    }
  }
  d.B = v[0];
  d.coff = v[1];
  d.foff = v[2];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    d.nbr[j] = (int)v[3 + j];
    d.ncoff[j] = v[9 + j];
    d.nfoff[j] = v[15 + j];
  }
  d.m0 = v[21];
  d.m1 = v[22];
}

// Work-list cell i of group w and its block's descriptor.  ONE: the group is
// a single block (resolution 8), so the descriptor address is the same for
// every thread and the descriptor lands in scalar registers.  A thread past
// the group's cells gets false and the cell of the same local index in the
// group's first block (in bounds), so that its loads need no branch.
template <bool ONE>
__device__ __forceinline__ bool decode_r(const LevelDev& L, const WG& w, int i, CellRef& o, BDR& d) {
  const int sh = 3 * w.lr;
  const bool ok = i < (w.count << sh);
  o.q = ONE ? 0 : (ok ? i >> sh : 0);
  o.li = i & ((1 << sh) - 1);
  bd_read<ONE>(L.bd + (size_t)(w.start + o.q) * BD_WORDS, d);
  o.B = (int)d.B;
  o.lr = w.lr;
  o.r = 1 << w.lr;
  o.l[0] = o.li & (o.r - 1);
  o.l[1] = (o.li >> w.lr) & (o.r - 1);
  o.l[2] = o.li >> (2 * w.lr);
  o.c = d.coff + o.li;
  return ok;
}

// log2 cell edge of face neighbour d from the packed descriptor words
__device__ __forceinline__ int bd_ne(const BDR& bd, int d) {
  return (int)((d < 3 ? bd.m0 >> (8 * (d + 1)) : bd.m1 >> (8 * (d - 3))) & 255u);
}

// The stencil of stencil_split from a register descriptor, with every face's
// loads issued before any is used: the coefficient from the block's own face
// slot, and the neighbour value of each boundary face with a single
// neighbour (a load under the lane's own predicate, no branch on loaded
// data).  Only a face with four finer neighbours loads again.  rd(cell) reads
// the value of a cell.  Sums run in face order as in stencil_split.
template <typename T, class Rd>
__device__ __forceinline__ void stencil_r(const LevelDev& L, const CellRef& cr, const BDR& bd, Rd rd, T& diag,
                                          T& ext, T* w) {
  const int r = cr.r;
  const int e = (int)(bd.m0 & 255u);
  const T s = (T)(1 << e);
  T bt[6], xn[6];
  int kind[6];  // 0 in-block, 1 wall, 2 equal neighbour, 3 coarser neighbour, 4 finer neighbours
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const int a = d >> 1, side = d & 1;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const int nl = cr.l[a] + (side ? 1 : -1);
    const bool inb = nl >= 0 && nl < r;
    int sl[3] = {cr.l[0], cr.l[1], cr.l[2]};
    sl[a] += side;
    bt[d] = (T)L.beta[a][bd.foff + face_idx(a, sl, r)];
    const int N = bd.nbr[d];
    const int en = bd_ne(bd, d);
    const int rn = L.bs >> en;
    const bool coarser = en > e;
    int nc[3] = {cr.l[0], cr.l[1], cr.l[2]};
    nc[a] = side ? 0 : rn - 1;
    nc[t1] = coarser ? cr.l[t1] >> 1 : cr.l[t1];
    nc[t2] = coarser ? cr.l[t2] >> 1 : cr.l[t2];
    const bool single = !inb && N >= 0 && en >= e;
    xn[d] = (T)0;
    if (single) xn[d] = rd(bd.ncoff[d] + cell_idx(nc, rn));
    kind[d] = inb ? 0 : (N < 0 ? 1 : (en == e ? 2 : (coarser ? 3 : 4)));
  }
  diag = 0;
  ext = 0;
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    w[d] = 0;
    if (kind[d] == 0) {
      const T c = bt[d] * s;
      w[d] = c;
      diag += c;
    } else if (kind[d] == 2) {
      const T c = bt[d] * s;
      diag += c;
      ext += c * xn[d];
    } else if (kind[d] == 3) {
      const T c = bt[d] * s / (T)1.5;
      diag += c;
      ext += c * xn[d];
    } else if (kind[d] == 4) {
      const int a = d >> 1, side = d & 1;
      const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
      const int rn = L.bs >> bd_ne(bd, d);
      const uint32_t con = bd.ncoff[d], fon = bd.nfoff[d];
      int nc[3], ps[3];
      nc[a] = side ? 0 : rn - 1;
      ps[a] = side ? 0 : rn;
      const T sf = (T)0.5 * s;
#pragma unroll
      for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
        for (int d2 = 0; d2 < 2; ++d2) {
          nc[t1] = ps[t1] = 2 * cr.l[t1] + d1;
          nc[t2] = ps[t2] = 2 * cr.l[t2] + d2;
          const T c = (T)L.beta[a][fon + face_idx(a, ps, rn)] * sf / (T)1.5;
          diag += c;
          ext += c * rd(con + cell_idx(nc, rn));
        }
    }
  }
}

// in-block part: sum of w[d] * value of the in-block neighbour d
template <class In>
__device__ __forceinline__ float inblock(const float* w, int li, int r, In inb) {
  float s = 0.f;
  const int st[3] = {1, r, r * r};
#pragma unroll
  for (int d = 0; d < 6; ++d)
    if (w[d] != 0.f) s += w[d] * inb(li + ((d & 1) ? st[d >> 1] : -st[d >> 1]));
  return s;
}

// inblock without branches: every face reads a value (the cell's own where
// the face has no in-block neighbour, weighed by 0), the same sum
template <class In>
__device__ __forceinline__ float inblock_all(const float* w, int li, int r, In inb) {
  float s = 0.f;
  const int st[3] = {1, r, r * r};
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const int nl = li + ((d & 1) ? st[d >> 1] : -st[d >> 1]);
    s += w[d] * inb(w[d] != 0.f ? nl : li);
  }
  return s;
}

// parent cell on the next coarser level: same block halving (local / 2),
// same block keeping its resolution, or the aggregated block (coords / 2)
__device__ __forceinline__ uint32_t parent_cell(const LevelDev& F, const LevelDev& C, int aggregated, int B,
                                                const int* l, int r, int li) {
  if (aggregated) {
    int Bc[3];
    block_coords(F.D, B, Bc);
    const int q[3] = {Bc[0] >> 1, Bc[1] >> 1, Bc[2] >> 1};
    return C.coff[block_id(C.D, q)];
  }
  if (C.e[B] == F.e[B]) return C.coff[B] + li;
  const int q[3] = {l[0] >> 1, l[1] >> 1, l[2] >> 1};
  return C.coff[B] + cell_idx(q, r >> 1);
}


// Block red-black x cell red-black weighted Gauss-Seidel on one block colour
// (smoothActiveBlocks, Smoother.cuh:105-202).  Load phase: own x into LDS,
// b, diag and the contribution of the frozen other-colour neighbour blocks;
// then `local_sweeps` x (cell colour 0, 1) sweeps entirely in LDS; store.
// `reverse` runs the cell colours in the opposite order (post-smoother).
// The stencil kernels are latency-bound (one round of work-groups): they are
// held to 64 VGPRs for eight waves per SIMD.
#define AW_STENCIL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(8, 8)))
// A stencil work-group covers 512 cells with ST_T threads, ST_CPT cells each.
// One cell per thread: each thread's chain of dependent loads is as short as
// it gets (measured at 256^3: level-1 smoother phase 11.7 us against 15.9 us
// with 256 threads x 2 cells, level 0 even; 4 cells per thread: 33 us).
constexpr int ST_T = 512, ST_CPT = 512 / ST_T;

static_assert(ST_CPT == 1, "the stencil kernels below give each thread one cell");

template <bool ONE>
__device__ __forceinline__ void smooth_group(const LevelDev& L, float* __restrict__ x, const float* __restrict__ b,
                                             const WG& w, float omega, int local_sweeps, int reverse, float* sx) {
  const int i = threadIdx.x;
  CellRef cr;
  BDR bd;
  const bool ok = decode_r<ONE>(L, w, i, cr, bd);
  float dg, ex, cw[6];
  const float xv = x[cr.c];
  const bool act = ok & (L.liquid[cr.c] != 0);
  const float bb = b[cr.c];
  stencil_r(L, cr, bd, [&](uint32_t cc) { return x[cc]; }, dg, ex, cw);
  sx[i] = xv;
  // the relaxation weight omega / diag once per cell (not a division per update)
  const float wod = dg > 0.f ? omega / dg : 0.f;
  const bool mine = act && dg > 0.f;
  const int cc = (cr.l[0] + cr.l[1] + cr.l[2]) & 1;
  __syncthreads();
  const int base = i - cr.li;
  for (int ls = 0; ls < local_sweeps; ++ls)
    for (int ph = 0; ph < 2; ++ph) {
      const int col = reverse ? 1 - ph : ph;
      if (mine && cc == col) {
        const float off = ex + inblock_all(cw, cr.li, cr.r, [&](int li) { return sx[base + li]; });
        const float res = bb - (dg * sx[i] - off);
        sx[i] += wod * res;
      }
      __syncthreads();
    }
  if (act) x[cr.c] = sx[i];
}

__global__ __launch_bounds__(ST_T) AW_STENCIL_OCCUPANCY void smooth_kernel(
    LevelDev L, float* __restrict__ x, const float* __restrict__ b, WGMap wgm,
    const uint32_t* __restrict__ blist, float omega, int local_sweeps, int reverse) {
  __shared__ float sx[512];
  const WG w = wg_at(wgm, blockIdx.x);
  if (w.lr == 3)
    smooth_group<true>(L, x, b, w, omega, local_sweeps, reverse, sx);
  else
    smooth_group<false>(L, x, b, w, omega, local_sweeps, reverse, sx);
}

// live[B] = 1 when block B of the level holds a liquid cell: one wave per
// block, its lanes strided over the block's cells.
constexpr int LIVE_T = 256;
__global__ __launch_bounds__(LIVE_T) void block_live_kernel(LevelDev L, uint8_t* __restrict__ live) {
  const int B = blockIdx.x * (LIVE_T / 64) + (int)(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (B >= L.nb) return;  // the whole wave leaves together
  const uint32_t c0 = L.coff[B], c1 = L.coff[B + 1];
  bool any = false;
  for (uint32_t c = c0 + lane; c < c1; c += 64) any = any || L.liquid[c] != 0;
  const bool wave_any = __ballot(any) != 0;
  if (lane == 0) live[B] = wave_any ? 1 : 0;
}

// out[j] = descriptor idx[j] of bd, one word per thread
__global__ __launch_bounds__(TPB) void gather_bd_kernel(const uint32_t* __restrict__ bd,
                                                        const uint32_t* __restrict__ idx, int n,
                                                        uint32_t* __restrict__ out) {
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= n * BD_WORDS) return;
  const int j = t / BD_WORDS, w = t - j * BD_WORDS;
  out[t] = bd[(size_t)idx[j] * BD_WORDS + w];
}

// The reference's adaptive smoothing of one block colour (smoothActiveBlocks,
// Smoother.cuh:105-202, with its defects A12 fixed: the exit is decided per
// block in LDS, so every thread of the work-group leaves the loop together).
// A block runs only if act[B] is set.  It runs up to max_sweeps in-block
// sweeps (both cell colours each), tracking the largest |r| met by its
// updates in each sweep; it stops after the first sweep whose maximum is
// below `reduction` times the first sweep's (when that is above 1e-9).  A
// block whose last sweep's maximum exceeds theta (absolute; the caller scales
// the reference's 1e-4 by max |b|, A22) sets act_next on itself and its six
// face neighbours (activateBlock, Smoother.cuh:97-102,188-201; plain stores of
// 1, so the order does not matter).  sweeps[B] receives the sweeps run (0:
// inactive).  The loads and the update are smooth_kernel's.
__global__ __launch_bounds__(ST_T) AW_STENCIL_OCCUPANCY void adaptive_smooth_kernel(
    LevelDev L, float* __restrict__ x, const float* __restrict__ b, WGMap wgm, float omega, int reverse,
    const uint8_t* __restrict__ act, uint8_t* __restrict__ act_next, uint8_t* __restrict__ sweeps,
    const float* __restrict__ theta_p, float theta_rel, int max_sweeps, float reduction) {
  __shared__ float sx[512];
  __shared__ uint32_t smax[512];   // per block of the group: this sweep's max |r| (float bits)
  __shared__ float sfirst[512];
  __shared__ uint8_t sdone[512];
  const WG w = wg_at(wgm, blockIdx.x);
  CellRef cr;
  BDR bd;
  const int i = threadIdx.x;
  const bool ok = decode_r<false>(L, w, i, cr, bd);
  bool on = false, liq = false;
  float bb = 0.f, dg = 0.f, ex = 0.f, cw[6];
  if (ok) {
    sx[i] = x[cr.c];
    on = act[cr.B] != 0;
    liq = L.liquid[cr.c] != 0;
    bb = b[cr.c];
    stencil_r(L, cr, bd, [&](uint32_t cc) { return x[cc]; }, dg, ex, cw);
  }
  const bool live = on && ok && liq;
  if (!__syncthreads_or(on ? 1 : 0)) {   // no active block in the group
    if (ok && cr.li == 0) sweeps[cr.B] = 0;
    return;
  }
  // block q of the group is active: one thread per block slot
  const int nblk = w.count;
  if (i < nblk) {
    const uint32_t* bq = L.bd + (size_t)(w.start + i) * BD_WORDS;
    sdone[i] = act[bq[0]] ? 0 : 1;
    sfirst[i] = 0.f;
  }
  const float theta = theta_rel * *theta_p;
  int ran = 0;
  for (int ls = 0; ls < max_sweeps; ++ls) {
    if (i < nblk) smax[i] = 0u;
    __syncthreads();
    const bool go = ok && !sdone[cr.q];
    for (int ph = 0; ph < 2; ++ph) {
      const int col = reverse ? 1 - ph : ph;
      if (go && live && ((cr.l[0] + cr.l[1] + cr.l[2]) & 1) == col) {
        const int base = i - cr.li;
        const float off = ex + inblock(cw, cr.li, cr.r, [&](int li) { return sx[base + li]; });
        const float res = bb - (dg * sx[i] - off);
        if (dg > 0.f) {
          sx[i] += omega * res / dg;
          atomicMax(&smax[cr.q], __float_as_uint(fabsf(res)));
        }
      }
      __syncthreads();
    }
    if (go) ran = ls + 1;
    int more = 0;
    if (i < nblk && !sdone[i]) {
      const float m = __uint_as_float(smax[i]);
      if (ls == 0) sfirst[i] = m;
      if ((sfirst[i] > 1e-9f && m < sfirst[i] * reduction) || ls + 1 == max_sweeps) {
        sdone[i] = 1;
        // activation from the last sweep's maximum
        if (m > theta) {
          const uint32_t* bq = L.bd + (size_t)(w.start + i) * BD_WORDS;
          act_next[bq[0]] = 1;
#pragma unroll
          for (int d = 0; d < 6; ++d) {
            const int N = (int)bq[3 + d];
            if (N >= 0) act_next[N] = 1;
          }
        }
      } else {
        more = 1;
      }
    }
    if (!__syncthreads_or(more)) break;
  }
  if (ok) {
    if (on) x[cr.c] = sx[i];
    if (cr.li == 0) sweeps[cr.B] = (uint8_t)(on ? ran : 0);
  }
}

// max |v| over liquid cells into *out (float bits; *out cleared by the launcher)
__global__ __launch_bounds__(TPB) void absmax_kernel(const float* __restrict__ v, const uint8_t* __restrict__ liq, int n,
                                                     uint32_t* __restrict__ out) {
  float m = 0.f;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (liq[i]) m = fmaxf(m, fabsf(v[i]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(out, __float_as_uint(m));
}

// smooth_kernel for a uniform coarse level (every block of resolution
// 1 << lr, beta 1 inside the tank and 0 on its walls, power-of-two block
// dims): the blocks of one colour are enumerated and every neighbour address
// and coefficient comes from the cell's coordinates, so the only loads are
// x, b and the liquid mask.  Same sums, in the same order, as smooth_kernel.
struct ULevel {
  int sh[3];   // log2 block dims
  int lr, e;   // log2 block resolution, log2 cell edge
};

__global__ __launch_bounds__(ST_T) AW_STENCIL_OCCUPANCY void uniform_smooth_kernel(
    ULevel U, const uint8_t* __restrict__ liquid, float* __restrict__ x, const float* __restrict__ b, int col,
    float omega, int local_sweeps, int reverse) {
  __shared__ float sx[512];
  const int lr = U.lr, r = 1 << lr, sh3 = 3 * lr;
  const int D[3] = {1 << U.sh[0], 1 << U.sh[1], 1 << U.sh[2]};
  const int ncol = (D[0] * D[1] * D[2]) >> 1;       // blocks of one colour
  const int per = 512 >> sh3;                        // blocks per work-group
  const float s = (float)(1 << U.e);
  int c[ST_CPT], lo[ST_CPT][3];
  bool act[ST_CPT];
  float bb[ST_CPT], dg[ST_CPT], ex[ST_CPT];
  uint8_t inb[ST_CPT];   // bit d: the face-d neighbour is in the block
#pragma unroll
  for (int k = 0; k < ST_CPT; ++k) {
    const int i = threadIdx.x + k * ST_T;
    act[k] = false;
    c[k] = -1;
    const int kb = blockIdx.x * per + (i >> sh3);
    if (kb >= ncol) continue;
    const int hs = U.sh[0] - 1;
    const int row = kb >> hs;
    const int by = row & (D[1] - 1), bz = row >> U.sh[1];
    const int bx = 2 * (kb & ((1 << hs) - 1)) + ((col + by + bz) & 1);
    const int B = bx + (row << U.sh[0]);
    const int li = i & ((1 << sh3) - 1);
    const int l[3] = {li & (r - 1), (li >> lr) & (r - 1), li >> (2 * lr)};
    const int Bc[3] = {bx, by, bz};
    c[k] = (B << sh3) + li;
    // every load of the cell issued before any is used (the neighbour
    // addresses come from coordinates): value, mask, right-hand side and
    // the out-of-block neighbours
    sx[i] = x[c[k]];
    act[k] = liquid[c[k]] != 0;
    bb[k] = b[c[k]];
    float dgv = 0.f, exv = 0.f;
    uint8_t m = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      const int a = d >> 1, side = d & 1;
      int G[3] = {Bc[0] * r + l[0], Bc[1] * r + l[1], Bc[2] * r + l[2]};
      G[a] += side ? 1 : -1;
      if (G[a] < 0 || G[a] >= D[a] * r) continue;   // tank wall: beta 0
      dgv += s;
      const int nl = l[a] + (side ? 1 : -1);
      if (nl >= 0 && nl < r) {
        m |= (uint8_t)(1 << d);
        continue;
      }
      const int q[3] = {G[0] >> lr, G[1] >> lr, G[2] >> lr};
      const int lq[3] = {G[0] & (r - 1), G[1] & (r - 1), G[2] & (r - 1)};
      exv += s * x[((q[0] + D[0] * (q[1] + D[1] * q[2])) << sh3) + cell_idx(lq, r)];
    }
    dg[k] = dgv;
    ex[k] = exv;
    inb[k] = m;
    lo[k][0] = l[0];
    lo[k][1] = l[1];
    lo[k][2] = l[2];
  }
  // the relaxation weight omega / diag once per cell; every face reads a
  // value (the cell's own, weighed by 0, where it has no in-block neighbour)
  float wod[ST_CPT];
#pragma unroll
  for (int k = 0; k < ST_CPT; ++k) {
    wod[k] = act[k] && dg[k] > 0.f ? omega / dg[k] : 0.f;
    act[k] = act[k] && dg[k] > 0.f;
  }
  __syncthreads();
  const int st[3] = {1, r, r * r};
  for (int ls = 0; ls < local_sweeps; ++ls)
    for (int ph = 0; ph < 2; ++ph) {
      const int cc = reverse ? 1 - ph : ph;
#pragma unroll
      for (int k = 0; k < ST_CPT; ++k) {
        if (!act[k]) continue;
        if (((lo[k][0] + lo[k][1] + lo[k][2]) & 1) != cc) continue;
        const int i = threadIdx.x + k * ST_T;
        float in = 0.f;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          const bool has = inb[k] & (1 << d);
          in += (has ? s : 0.f) * sx[has ? i + ((d & 1) ? st[d >> 1] : -st[d >> 1]) : i];
        }
        const float off = ex[k] + in;
        const float res = bb[k] - (dg[k] * sx[i] - off);
        sx[i] += wod[k] * res;
      }
      __syncthreads();
    }
#pragma unroll
  for (int k = 0; k < ST_CPT; ++k)
    if (act[k]) x[c[k]] = sx[threadIdx.x + k * ST_T];
}

// r = b - A x (computeExplicitResidual, VCycle.cuh:81-131), 0 on air cells;
// optionally fused with the restriction b_c = sum of children
// (restrictResidualsKernel, MultigridTransfer.cuh:9-59) and x_c = 0.
template <bool ONE>
__device__ __forceinline__ void residual_group(const LevelDev& L, const float* __restrict__ x,
                                               const float* __restrict__ b, float* __restrict__ r, const WG& w,
                                               const Restr& rs, float* sx, float* sr, uint32_t* sco, int* shalf) {
  const int i = threadIdx.x;
  CellRef cr;
  BDR bd;
  const bool ok = decode_r<ONE>(L, w, i, cr, bd);
  float dg, ex, cw[6];
  const float xv = x[cr.c];
  const bool liq = L.liquid[cr.c] != 0;
  const float bb = b[cr.c];
  stencil_r(L, cr, bd, [&](uint32_t cc) { return x[cc]; }, dg, ex, cw);
  // the fused restriction's coarse cell offset of each block of the group
  // (and whether the coarse level halves the group's blocks), in LDS
  if (rs.on && ok && cr.li == 0) {
    sco[cr.q] = rs.C.coff[cr.B];
    if (cr.q == 0) *shalf = (int)rs.C.e[cr.B] != (int)(bd.m0 & 255u);
  }
  sx[i] = xv;
  __syncthreads();
  if (ok) {
    float rv = 0.f;
    if (liq) {
      const int base = i - cr.li;
      const float off = ex + inblock_all(cw, cr.li, cr.r, [&](int li) { return sx[base + li]; });
      rv = bb - (dg * sx[i] - off);
    }
    sr[i] = rv;
    if (r) r[cr.c] = rv;
  }
  if (!rs.on) return;
  __syncthreads();
  // coarse cells of this group's blocks (all of one resolution)
  const int rF = 1 << w.lr;
  const bool halves = *shalf != 0;
  const int rC = halves ? rF >> 1 : rF;
  const int cpbC = rC * rC * rC, cpbF = rF * rF * rF;
  for (int j = threadIdx.x; j < w.count * cpbC; j += ST_T) {
    const int q = j / cpbC, lj = j - q * cpbC;
    const int l[3] = {lj % rC, (lj / rC) % rC, lj / (rC * rC)};
    float s = 0.f;
    if (halves) {
      for (int d = 0; d < 8; ++d) {
        const int f[3] = {2 * l[0] + (d & 1), 2 * l[1] + ((d >> 1) & 1), 2 * l[2] + (d >> 2)};
        s += sr[q * cpbF + cell_idx(f, rF)];
      }
    } else {
      s = sr[q * cpbF + lj];
    }
    const uint32_t cc = sco[q] + lj;
    rs.bc[cc] = s;
    rs.xc[cc] = 0.f;
  }
}

__global__ __launch_bounds__(ST_T) AW_STENCIL_OCCUPANCY void residual_kernel(
    LevelDev L, const float* __restrict__ x, const float* __restrict__ b, float* __restrict__ r,
    WGMap wgm, const uint32_t* __restrict__ blist, Restr rs) {
  __shared__ float sx[512];
  __shared__ float sr[512];
  __shared__ uint32_t sco[512];
  __shared__ int shalf;
  const WG w = wg_at(wgm, blockIdx.x);
  if (w.lr == 3)
    residual_group<true>(L, x, b, r, w, rs, sx, sr, sco, &shalf);
  else
    residual_group<false>(L, x, b, r, w, rs, sx, sr, sco, &shalf);
}

// y = A x with sum.acc += x . y
template <bool ONE>
__device__ __forceinline__ double apply_group(const LevelDev& L, const float* __restrict__ x, float* __restrict__ y,
                                              const WG& w, float* sx) {
  const int i = threadIdx.x;
  CellRef cr;
  BDR bd;
  const bool ok = decode_r<ONE>(L, w, i, cr, bd);
  float dg, ex, cw[6];
  const float xv = x[cr.c];
  const bool liq = L.liquid[cr.c] != 0;
  stencil_r(L, cr, bd, [&](uint32_t cc) { return x[cc]; }, dg, ex, cw);
  sx[i] = xv;
  __syncthreads();
  if (!ok) return 0.0;
  if (!liq) {
    y[cr.c] = 0.f;
    return 0.0;
  }
  const int base = i - cr.li;
  const float off = ex + inblock_all(cw, cr.li, cr.r, [&](int li) { return sx[base + li]; });
  const float v = dg * sx[i] - off;
  y[cr.c] = v;
  return (double)sx[i] * (double)v;
}

__global__ __launch_bounds__(ST_T) AW_STENCIL_OCCUPANCY void apply_dot_kernel(
    LevelDev L, const float* __restrict__ x, float* __restrict__ y, WGMap wgm,
    const uint32_t* __restrict__ blist, double* acc) {
  __shared__ float sx[512];
  const WG w = wg_at(wgm, blockIdx.x);
  const double part = w.lr == 3 ? apply_group<true>(L, x, y, w, sx) : apply_group<false>(L, x, y, w, sx);
  if (acc) grid_sum<ST_T>(part, acc);
}

// children sum of one coarse cell (halving / keeping block, or aggregation)
template <class Rd>
__device__ __forceinline__ float children_sum(const LevelDev& F, const LevelDev& C, int aggregated, const CellRef& c,
                                              Rd rd) {
  float s = 0.f;
  if (aggregated) {
    int Bc[3];
    block_coords(C.D, c.B, Bc);
    for (int d = 0; d < 8; ++d) {
      const int q[3] = {2 * Bc[0] + (d & 1), 2 * Bc[1] + ((d >> 1) & 1), 2 * Bc[2] + (d >> 2)};
      s += rd(F.coff[block_id(F.D, q)]);
    }
  } else if (C.e[c.B] == F.e[c.B]) {
    s = rd(F.coff[c.B] + c.li);
  } else {
    const int rF = 2 * c.r;
    const uint32_t fo = F.coff[c.B];
    for (int d = 0; d < 8; ++d) {
      const int q[3] = {2 * c.l[0] + (d & 1), 2 * c.l[1] + ((d >> 1) & 1), 2 * c.l[2] + (d >> 2)};
      s += rd(fo + cell_idx(q, rF));
    }
  }
  return s;
}

// standalone restriction (used across aggregation levels): b_c = sum of
// children, x_c = 0
__global__ __launch_bounds__(TPB) void restrict_kernel(LevelDev F, LevelDev C, int aggregated,
                                                       const float* __restrict__ rf, float* __restrict__ bc,
                                                       float* __restrict__ xc, WGMap wgm,
                                                       const uint32_t* __restrict__ blist) {
  const WG w = wg_at(wgm, blockIdx.x);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(C, w, blist, threadIdx.x + k * TPB, c)) continue;
    bc[c.c] = children_sum(F, C, aggregated, c, [&](uint32_t i) { return rf[i]; });
    if (xc) xc[c.c] = 0.f;
  }
}

// Coefficient coarsening (galerkinCoarsen / runGalerkin, MSBG.cuh:401-498,
// 661-676): every face slot of the coarser level C is the area mean of the
// finer level F's beta over the slot's square, taken from the side that holds
// the finer faces there (the block itself, or across a block face its
// neighbour when that was finer on F: a level change's coarse-side slot is
// not an unknown and is never read).  A block that kept its resolution copies
// its slots.  Across a 2x2x2 aggregation a slot is the mean of the four fine
// blocks' slots on its side.  One work-group per block of C.
__global__ __launch_bounds__(TPB) void coarsen_beta_kernel(LevelDev F, LevelDev C, int aggregated, float* __restrict__ bx,
                                                           float* __restrict__ by, float* __restrict__ bz) {
  const int B = blockIdx.x;
  float* out[3] = {bx, by, bz};
  const int eC = C.e[B], rC = C.bs >> eC, nf = (rC + 1) * rC * rC;
  int Bc[3];
  block_coords(C.D, B, Bc);
  for (int k = threadIdx.x; k < 3 * nf; k += TPB) {
    const int a = k / nf, f = k - a * nf;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    int sl[3];
    face_decode(a, f, rC, sl);
    float v = 0.f;
    if (aggregated) {
      // one-cell blocks: slot sl[a] (0 or 1) is the face on that side
#pragma unroll
      for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
        for (int d2 = 0; d2 < 2; ++d2) {
          int q[3];
          q[a] = 2 * Bc[a] + sl[a];
          q[t1] = 2 * Bc[t1] + d1;
          q[t2] = 2 * Bc[t2] + d2;
          const int X = block_id(F.D, q);
          v += F.beta[a][F.foff[a][X] + sl[a]];
        }
      v *= 0.25f;
    } else {
      int X = B;
      bool wall = false;
      if (sl[a] == 0 || sl[a] == rC) {
        const int N = C.nbr[B * 6 + 2 * a + (sl[a] == rC ? 1 : 0)];
        if (N < 0)
          wall = true;
        else if (F.e[N] < F.e[B])
          X = N;
      }
      if (!wall) {
        const int S = 1 << eC;                        // the slot's edge in fine cells
        const int P = C.bs * Bc[a] + sl[a] * S;       // its plane along a
        const int eX = F.e[X], rX = F.bs >> eX, m = S >> eX;
        int Xc[3];
        block_coords(F.D, X, Xc);
        int l[3];
        l[a] = (P - F.bs * Xc[a]) >> eX;
        const int o1 = (C.bs * Bc[t1] + sl[t1] * S - F.bs * Xc[t1]) >> eX;
        const int o2 = (C.bs * Bc[t2] + sl[t2] * S - F.bs * Xc[t2]) >> eX;
        const uint32_t fo = F.foff[a][X];
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) {
            l[t1] = o1 + i;
            l[t2] = o2 + j;
            v += F.beta[a][fo + face_idx(a, l, rX)];
          }
        v /= (float)(m * m);
      }
    }
    out[a][C.foff[a][B] + f] = v;
  }
}

// liquid_c = all children liquid (Dirichlet air dominates on coarse levels)
__global__ __launch_bounds__(TPB) void coarsen_liquid_kernel(LevelDev F, LevelDev C, int aggregated,
                                                             uint8_t* __restrict__ liq, WGMap wgm,
                                                             const uint32_t* __restrict__ blist) {
  const WG w = wg_at(wgm, blockIdx.x);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    if (!decode(C, w, blist, threadIdx.x + k * TPB, c)) continue;
    const float all = children_sum(F, C, aggregated, c, [&](uint32_t i) { return F.liquid[i] ? 0.f : 1.f; });
    liq[c.c] = all == 0.f ? 1 : 0;
  }
}

// x_f += scale * x_c[parent] on liquid cells (prolongateCorrection,
// MultigridTransfer.cuh:63-102).  The fine block's descriptor in registers
// (scalar when the group is one block), then the cell's mask, value and
// parent offset together, then the parent's value.
template <bool ONE>
__device__ __forceinline__ void prolong_group(const LevelDev& F, const LevelDev& C, int aggregated,
                                              float* __restrict__ xf, const float* __restrict__ xc, float scale,
                                              const WG& w) {
  CellRef cr;
  BDR bd;
  if (!decode_r<ONE>(F, w, threadIdx.x, cr, bd)) return;
  const bool liq = F.liquid[cr.c] != 0;
  const float xv = xf[cr.c];
  const uint32_t pc = parent_cell(F, C, aggregated, cr.B, cr.l, cr.r, cr.li);
  if (liq) xf[cr.c] = xv + scale * xc[pc];
}

__global__ __launch_bounds__(ST_T) void prolong_kernel(LevelDev F, LevelDev C, int aggregated, float* __restrict__ xf,
                                                       const float* __restrict__ xc, float scale,
                                                       WGMap wgm,
                                                       const uint32_t* __restrict__ blist) {
  const WG w = wg_at(wgm, blockIdx.x);
  if (w.lr == 3)
    prolong_group<true>(F, C, aggregated, xf, xc, scale, w);
  else
    prolong_group<false>(F, C, aggregated, xf, xc, scale, w);
}

// ---- single-work-group V-cycle over the small trailing levels, in LDS ----
// Levels with one cell size everywhere (every level after the first
// aggregation, and block levels of uniform resolution) evaluate the stencil
// arithmetically from cell coordinates: coefficient s to every in-domain
// neighbour (coarse levels carry beta = 1 inside the tank), no metadata loads.

struct UStencil {
  int diagn;     // number of in-domain neighbours
  int nb[6];     // neighbour cell indices, -1 = tank wall
  int cls;       // block colour * 2 + cell colour
};

__device__ __forceinline__ void uniform_stencil(const TailLevel& T, int i, UStencil& u) {
  const int lr = T.ulr, r = 1 << lr, sh = 3 * lr;
  const int B = i >> sh, li = i & ((1 << sh) - 1);
  int Bc[3];
  block_coords(T.L.D, B, Bc);
  const int l[3] = {li & (r - 1), (li >> lr) & (r - 1), li >> (2 * lr)};
  u.cls = (((Bc[0] + Bc[1] + Bc[2]) & 1) << 1) | ((l[0] + l[1] + l[2]) & 1);
  u.diagn = 0;
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const int a = d >> 1;
    int G[3] = {Bc[0] * r + l[0], Bc[1] * r + l[1], Bc[2] * r + l[2]};
    G[a] += (d & 1) ? 1 : -1;
    if (G[a] < 0 || G[a] >= T.L.D[a] * r) {
      u.nb[d] = -1;
      continue;
    }
    const int q[3] = {G[0] >> lr, G[1] >> lr, G[2] >> lr};
    const int lq[3] = {G[0] & (r - 1), G[1] & (r - 1), G[2] & (r - 1)};
    u.nb[d] = (block_id(T.L.D, q) << sh) + cell_idx(lq, r);
    ++u.diagn;
  }
}

// class of a cell of a generic level (block colour * 2 + cell colour)
__device__ __forceinline__ int generic_class(const TailLevel& T, int i, CellRef& c) {
  const int B = T.cblk[i];
  cell_ref(T.L, B, i - T.L.coff[B], c);
  int Bc[3];
  block_coords(T.L.D, B, Bc);
  return (((Bc[0] + Bc[1] + Bc[2]) & 1) << 1) | ((c.l[0] + c.l[1] + c.l[2]) & 1);
}

// Lattice levels (one-cell blocks, power-of-two dims): the uniform stencil
// from the cell's coordinates by shifts, neighbours in the order of
// uniform_stencil (same sums).
__device__ __forceinline__ float lattice_ax(const TailLevel& T, const float* xs, int i, int x, int y, int z) {
  const int D0 = 1 << T.sh[0], D1 = 1 << T.sh[1], D2 = 1 << T.sh[2];
  const int sy = D0, sz = D0 * D1;
  const float s = (float)(1 << T.ue);
  float off = 0.f;
  int n = 0;
  if (x > 0) { off += s * xs[i - 1]; ++n; }
  if (x + 1 < D0) { off += s * xs[i + 1]; ++n; }
  if (y > 0) { off += s * xs[i - sy]; ++n; }
  if (y + 1 < D1) { off += s * xs[i + sy]; ++n; }
  if (z > 0) { off += s * xs[i - sz]; ++n; }
  if (z + 1 < D2) { off += s * xs[i + sz]; ++n; }
  return s * (float)n * xs[i] - off;
}

__device__ __forceinline__ void lattice_xyz(const TailLevel& T, int i, int& x, int& y, int& z) {
  x = i & ((1 << T.sh[0]) - 1);
  y = (i >> T.sh[0]) & ((1 << T.sh[1]) - 1);
  z = i >> (T.sh[0] + T.sh[1]);
}

// A x (own row) of cell i on a tail level, from LDS values xs
__device__ __forceinline__ float tail_ax(const TailLevel& T, const float* xs, int i, int* cls_out) {
  if (T.sh[0] >= 0 && !cls_out) {
    int x, y, z;
    lattice_xyz(T, i, x, y, z);
    return lattice_ax(T, xs, i, x, y, z);
  }
  if (T.ulr >= 0) {
    UStencil u;
    uniform_stencil(T, i, u);
    const float s = (float)(1 << T.ue);
    float off = 0.f;
#pragma unroll
    for (int d = 0; d < 6; ++d)
      if (u.nb[d] >= 0) off += s * xs[u.nb[d]];
    if (cls_out) *cls_out = u.cls;
    return s * (float)u.diagn * xs[i] - off;
  }
  CellRef c;
  const int cls = generic_class(T, i, c);
  if (cls_out) *cls_out = cls;
  float dg, ex, cw[6];
  stencil_split(
      T.L, c, [&](int, const int*, int, uint32_t q) { return xs[q]; }, dg, ex, cw);
  const int base = i - c.li;
  return dg * xs[i] - (ex + inblock(cw, c.li, c.r, [&](int li) { return xs[base + li]; }));
}

// colour class of cell i of a uniform level, without the stencil
__device__ __forceinline__ int uniform_class(const TailLevel& T, int i) {
  const int lr = T.ulr, r = 1 << lr, sh = 3 * lr;
  const int B = i >> sh, li = i & ((1 << sh) - 1);
  int Bc[3];
  block_coords(T.L.D, B, Bc);
  return (((Bc[0] + Bc[1] + Bc[2]) & 1) << 1) | (((li & (r - 1)) + ((li >> lr) & (r - 1)) + (li >> (2 * lr))) & 1);
}

// One weighted Gauss-Seidel update of the cells of class `want`: the class is
// decided first, and the stencil (row and diagonal) evaluated once, only for
// the cells that update.
__device__ __forceinline__ void tail_update(const TailLevel& T, float* xs, const float* bs, const uint8_t* liq,
                                            int want, float omega) {
  if (T.sh[0] >= 0) {
    // lattice: the cells of block colour want >> 1 directly, two per row
    // pair (x even or odd with the row's parity); one-cell blocks have only
    // cell colour 0
    if (want & 1) return;
    const int bcol = want >> 1;
    const int hs = T.sh[0] - 1;
    for (int t = threadIdx.x; t < (T.ncells >> 1); t += TAIL_T) {
      const int row = t >> hs;
      const int y = row & ((1 << T.sh[1]) - 1), z = row >> T.sh[1];
      const int x = 2 * (t & ((1 << hs) - 1)) + ((bcol + y + z) & 1);
      const int i = x + (row << T.sh[0]);
      if (!liq[i]) continue;
      const float s = (float)(1 << T.ue);
      const int n = (x > 0) + (x + 1 < (1 << T.sh[0])) + (y > 0) + (y + 1 < (1 << T.sh[1])) + (z > 0) +
                    (z + 1 < (1 << T.sh[2]));
      const float dg = s * (float)n;
      const float ax = lattice_ax(T, xs, i, x, y, z);
      if (dg > 0.f) xs[i] += omega * (bs[i] - ax) / dg;
    }
    return;
  }
  for (int i = threadIdx.x; i < T.ncells; i += TAIL_T) {
    if (!liq[i]) continue;
    if (T.ulr >= 0) {
      if (uniform_class(T, i) != want) continue;
      UStencil u;
      uniform_stencil(T, i, u);
      const float s = (float)(1 << T.ue);
      float off = 0.f;
#pragma unroll
      for (int d = 0; d < 6; ++d)
        if (u.nb[d] >= 0) off += s * xs[u.nb[d]];
      const float dg = s * (float)u.diagn;
      const float ax = dg * xs[i] - off;
      if (dg > 0.f) xs[i] += omega * (bs[i] - ax) / dg;
    } else {
      CellRef c;
      if (generic_class(T, i, c) != want) continue;
      float dg, ex, cw[6];
      stencil_split(
          T.L, c, [&](int, const int*, int, uint32_t q) { return xs[q]; }, dg, ex, cw);
      const int base = i - c.li;
      const float ax = dg * xs[i] - (ex + inblock(cw, c.li, c.r, [&](int li) { return xs[base + li]; }));
      if (dg > 0.f) xs[i] += omega * (bs[i] - ax) / dg;
    }
  }
}

__device__ __forceinline__ void tail_smooth(const TailLevel& T, float* sm, const uint8_t* liq, int iters,
                                            bool reverse, float omega, int local_sweeps) {
  float* xs = sm + T.off;
  const float* bs = sm + T.off + T.ncells;
  // Blocks of one cell have a single cell colour, and their in-block sweeps
  // relax the same cell against frozen neighbours: k updates with weight w
  // leave (1 - w)^k of its residual, so they are one update with weight
  // 1 - (1 - w)^k (equal up to rounding), one barrier instead of k.
  const bool one_cell = T.ulr == 0;
  float w = omega;
  int ls_n = local_sweeps;
  if (one_cell) {
    float keep = 1.f;
    for (int ls = 0; ls < local_sweeps; ++ls) keep *= 1.f - omega;
    w = 1.f - keep;
    ls_n = 1;
  }
  for (int it = 0; it < iters; ++it)
    for (int ph = 0; ph < 2; ++ph) {
      const int bc = reverse ? 1 - ph : ph;
      for (int ls = 0; ls < ls_n; ++ls)
        for (int cp = 0; cp < 2; ++cp) {
          const int cc = reverse ? 1 - cp : cp;
          if (one_cell && cc == 1) continue;
          tail_update(T, xs, bs, liq, (bc << 1) | cc, w);
          __syncthreads();
        }
    }
}

__global__ __launch_bounds__(TAIL_T) void tail_kernel(TailArgs A) {
  extern __shared__ float sm[];
  // liquid masks of every tail level, as bytes after the float arrays
  uint8_t* lq = reinterpret_cast<uint8_t*>(sm + A.fl);
  for (int t = 0; t < A.n; ++t)
    for (int i = threadIdx.x; i < A.lv[t].ncells; i += TAIL_T) lq[A.lv[t].loff + i] = A.lv[t].L.liquid[i];
  const TailLevel& T0 = A.lv[0];
  for (int i = threadIdx.x; i < T0.ncells; i += TAIL_T) {
    sm[T0.off + i] = 0.f;
    sm[T0.off + T0.ncells + i] = A.b0[i];
  }
  __syncthreads();
  for (int t = 0; t + 1 < A.n; ++t) {
    const TailLevel& F = A.lv[t];
    const TailLevel& C = A.lv[t + 1];
    const uint8_t* fl = lq + F.loff;
    tail_smooth(F, sm, fl, A.pre, false, A.omega, A.local_sweeps);
    float* xs = sm + F.off;
    const float* bs = sm + F.off + F.ncells;
    float* rs = sm + F.off + 2 * F.ncells;
    for (int i = threadIdx.x; i < F.ncells; i += TAIL_T) rs[i] = fl[i] ? bs[i] - tail_ax(F, xs, i, nullptr) : 0.f;
    __syncthreads();
    const bool lat = C.agg && F.sh[0] >= 0 && C.sh[0] >= 0;
    for (int i = threadIdx.x; i < C.ncells; i += TAIL_T) {
      if (lat) {
        // children of lattice cell (X, Y, Z): (2X + dx, 2Y + dy, 2Z + dz), in children_sum's order
        int X, Y, Z;
        lattice_xyz(C, i, X, Y, Z);
        float s = 0.f;
        for (int d = 0; d < 8; ++d) {
          const int q = (2 * X + (d & 1)) + ((2 * Y + ((d >> 1) & 1)) << F.sh[0]) +
                        ((2 * Z + (d >> 2)) << (F.sh[0] + F.sh[1]));
          s += rs[q];
        }
        sm[C.off + C.ncells + i] = s;
        sm[C.off + i] = 0.f;
        continue;
      }
      const int B = C.cblk[i];
      CellRef c;
      cell_ref(C.L, B, i - C.L.coff[B], c);
      sm[C.off + C.ncells + i] = children_sum(F.L, C.L, C.agg, c, [&](uint32_t q) { return rs[q]; });
      sm[C.off + i] = 0.f;
    }
    __syncthreads();
  }
  const TailLevel& TC = A.lv[A.n - 1];
  if (A.cinv) {
    // exact coarsest solve: x = A^-1 b with the FP64 inverse (rows and
    // columns of air or dropped unknowns are zero)
    const int n = TC.ncells;
    const float* bs = sm + TC.off + n;
    for (int i = threadIdx.x; i < n; i += TAIL_T) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += A.cinv[(size_t)i * n + j] * (double)bs[j];
      sm[TC.off + i] = (float)s;
    }
    __syncthreads();
  } else {
    for (int sw = 0; sw < A.coarse_sweeps; ++sw) {
      tail_smooth(TC, sm, lq + TC.loff, 1, false, A.omega, A.local_sweeps);
      tail_smooth(TC, sm, lq + TC.loff, 1, true, A.omega, A.local_sweeps);
    }
  }
  for (int t = A.n - 2; t >= 0; --t) {
    const TailLevel& F = A.lv[t];
    const TailLevel& C = A.lv[t + 1];
    const uint8_t* fl = lq + F.loff;
    const bool lat = C.agg && F.sh[0] >= 0 && C.sh[0] >= 0;
    for (int i = threadIdx.x; i < F.ncells; i += TAIL_T) {
      if (!fl[i]) continue;
      if (lat) {
        int x, y, z;
        lattice_xyz(F, i, x, y, z);
        sm[F.off + i] += A.scale * sm[C.off + (x >> 1) + ((y >> 1) << C.sh[0]) + ((z >> 1) << (C.sh[0] + C.sh[1]))];
        continue;
      }
      const int B = F.cblk[i];
      CellRef c;
      cell_ref(F.L, B, i - F.L.coff[B], c);
      sm[F.off + i] += A.scale * sm[C.off + parent_cell(F.L, C.L, C.agg, c.B, c.l, c.r, c.li)];
    }
    __syncthreads();
    tail_smooth(F, sm, fl, A.post, true, A.omega, A.local_sweeps);
  }
  for (int i = threadIdx.x; i < T0.ncells; i += TAIL_T)
    if (lq[T0.loff + i]) A.x0[i] = sm[T0.off + i];
}

// FP64 inverse of the coarsest level restricted to its liquid cells
// (Gauss-Jordan in LDS; the same rule as the exact coarsest solve of the
// oracle: a pivot below 1e-12 of the largest diagonal drops its unknown).

__global__ __launch_bounds__(TPB) void coarse_factor_kernel(LevelDev L, int n, double* __restrict__ inv) {
  __shared__ double Am[DIRECT_MAX * DIRECT_MAX];
  __shared__ double Im[DIRECT_MAX * DIRECT_MAX];
  __shared__ double col[DIRECT_MAX];
  __shared__ int dead[DIRECT_MAX];
  __shared__ double dmax;
  for (int k = threadIdx.x; k < n * n; k += TPB) {
    Am[k] = 0.0;
    Im[k] = (k / n == k % n) ? 1.0 : 0.0;
  }
  if (threadIdx.x == 0) dmax = 0.0;
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += TPB) {
    dead[c] = 0;
    if (!L.liquid[c]) {
      Am[c * n + c] = 1.0;
      dead[c] = 1;
      continue;
    }
    int B = 0;
    while (L.coff[B + 1] <= (uint32_t)c) ++B;
    CellRef cr;
    cell_ref(L, B, c - L.coff[B], cr);
    float dg, ex, cw[6];
    // cross-block neighbours: the coefficient is recovered by reading a
    // unit value at one neighbour at a time
    stencil_split(
        L, cr,
        [&](int, const int*, int, uint32_t q) {
          return 0.f;
        },
        dg, ex, cw);
    Am[c * n + c] = dg;
    const int st[3] = {1, cr.r, cr.r * cr.r};
    for (int d = 0; d < 6; ++d)
      if (cw[d] != 0.f) {
        const int q = c + ((d & 1) ? st[d >> 1] : -st[d >> 1]);
        if (L.liquid[q]) Am[c * n + q] = -cw[d];
      }
  }
  __syncthreads();
  // cross-block couplings: every cell visits its cross-block neighbours
  for (int c = threadIdx.x; c < n; c += TPB) {
    if (!L.liquid[c]) continue;
    int B = 0;
    while (L.coff[B + 1] <= (uint32_t)c) ++B;
    CellRef cr;
    cell_ref(L, B, c - L.coff[B], cr);
    // coefficient of neighbour q = diag - (diag with q's term removed): use
    // the linearity of ext in the neighbour values
    uint32_t qs[24];
    int nq = 0;
    float dg, ex, cw[6];
    stencil_split(
        L, cr,
        [&](int, const int*, int, uint32_t q) {
          if (nq < 24) qs[nq++] = q;
          return 0.f;
        },
        dg, ex, cw);
    for (int k = 0; k < nq; ++k) {
      float e1;
      stencil_split(
          L, cr, [&](int, const int*, int, uint32_t q) { return q == qs[k] ? 1.f : 0.f; }, dg, e1, cw);
      if (L.liquid[qs[k]]) Am[c * n + qs[k]] -= (double)e1;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += TPB)
    if (!dead[c]) atomicMax((unsigned long long*)&dmax, __double_as_longlong(Am[c * n + c]));
  __syncthreads();
  for (int p = 0; p < n; ++p) {
    const double piv = Am[p * n + p];
    const bool drop = dead[p] || !(fabs(piv) > 1e-12 * dmax);
    __syncthreads();
    if (drop) {
      if (threadIdx.x == 0) dead[p] = 1;
      __syncthreads();
      continue;
    }
    const double ip = 1.0 / piv;
    for (int j = threadIdx.x; j < n; j += TPB) {
      Am[p * n + j] *= ip;
      Im[p * n + j] *= ip;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += TPB) col[i] = Am[i * n + p];
    __syncthreads();
    for (int k = threadIdx.x; k < n * n; k += TPB) {
      const int i = k / n, j = k - i * n;
      if (i == p || col[i] == 0.0) continue;
      Am[k] -= col[i] * Am[p * n + j];
      Im[k] -= col[i] * Im[p * n + j];
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < n * n; k += TPB) {
    const int i = k / n, j = k - i * n;
    inv[k] = (dead[i] || dead[j]) ? 0.0 : Im[k];
  }
}


// ---- FP64 iterative refinement on level 0 ----
// r = b - A x with x and the stencil in FP64 (b is the FP32 right-hand side);
// r is handed to the FP32 inner solve, ||r||^2 accumulates in FP64.
__global__ __launch_bounds__(TPB) void residual64_kernel(LevelDev L, const double* __restrict__ x,
                                                         const float* __restrict__ b, float* __restrict__ r32,
                                                         WGMap wgm,
                                                         const uint32_t* __restrict__ blist, double* rr) {
  const WG w = wg_at(wgm, blockIdx.x);
  double part = 0.0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    CellRef c;
    BDR bd;
    if (!decode_r<false>(L, w, threadIdx.x + k * TPB, c, bd)) continue;
    if (!L.liquid[c.c]) {
      r32[c.c] = 0.f;
      continue;
    }
    double dg, ex, cw[6];
    stencil_r(L, c, bd, [&](uint32_t cc) { return x[cc]; }, dg, ex, cw);
    const uint32_t base = c.c - c.li;
    double off = ex;
    const int st[3] = {1, c.r, c.r * c.r};
#pragma unroll
    for (int d = 0; d < 6; ++d)
      if (cw[d] != 0.0) off += cw[d] * x[base + c.li + ((d & 1) ? st[d >> 1] : -st[d >> 1])];
    const double rv = (double)b[c.c] - (dg * x[c.c] - off);
    r32[c.c] = (float)rv;
    part += rv * rv;
  }
  grid_sum(part, rr);
}

__global__ void axpy64_kernel(double* __restrict__ x, const float* __restrict__ e, const uint8_t* __restrict__ liq,
                              int n) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (liq[i]) x[i] += (double)e[i];
}

__global__ void to64_kernel(double* __restrict__ x, const float* __restrict__ p, const uint8_t* __restrict__ liq,
                            int n) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) x[i] = liq[i] ? (double)p[i] : 0.0;
}

__global__ void to32_kernel(float* __restrict__ p, const double* __restrict__ x, int n) {
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) p[i] = (float)x[i];
}

// ------------------------------------------------------------- vectors
__global__ __launch_bounds__(TPB) void dot_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                  const uint8_t* __restrict__ liq, int n, double* acc) {
  double s = 0.0;
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (liq[i]) s += (double)a[i] * (double)b[i];
  grid_sum(s, acc);
}

// RZOLD = RZ (summed into slot 0), then RZ, PQ, RR cleared; one wave
__global__ void scalar_shift_kernel(double* sc) {
  const int i = threadIdx.x;
  double t = i < SC_SLOTS ? sc[SC_RZ * SC_SLOTS + i] : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  if (i < SC_SLOTS) {
    sc[SC_RZOLD * SC_SLOTS + i] = i == 0 ? t : 0.0;
    sc[SC_RZ * SC_SLOTS + i] = 0.0;
    sc[SC_PQ * SC_SLOTS + i] = 0.0;
    sc[SC_RR * SC_SLOTS + i] = 0.0;
  }
}

__global__ __launch_bounds__(TPB) void pcg_p_kernel(const float* __restrict__ z, float* __restrict__ p, int n,
                                                    const double* __restrict__ sc, int first) {
  // the first direction is z itself: p may hold anything (a fresh buffer), and 0 * NaN is NaN
  if (first) {
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) p[i] = z[i];
    return;
  }
  const float beta = (float)(sc_get(sc, SC_RZ) / sc_get(sc, SC_RZOLD));
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(TPB) void pcg_xr_kernel(float* __restrict__ x, float* __restrict__ r,
                                                     const float* __restrict__ p, const float* __restrict__ q,
                                                     const uint8_t* __restrict__ liq, int n, double* sc) {
  const double pq = sc_get(sc, SC_PQ);
  const float alpha = pq > 0.0 ? (float)(sc_get(sc, SC_RZ) / pq) : 0.f;
  double s = 0.0;
  // liquid cells only: q = A p may be left unwritten elsewhere (p is 0 there)
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
    if (!liq[i]) continue;
    x[i] += alpha * p[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += (double)ri * (double)ri;
  }
  grid_sum(s, sc + SC_RR * SC_SLOTS);
}

__global__ __launch_bounds__(TPB) void count_nonzero_kernel(const uint8_t* __restrict__ m, uint32_t n,
                                                            uint32_t* __restrict__ out) {
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) c += m[i] != 0;
  const uint32_t s = block_reduce_sum<uint32_t>(c);
  if (threadIdx.x == 0 && s) atomicAdd(out, s);
}

// Order-independent fingerprint of a liquid mask: the 64-bit wrapping sum of
// a mixed hash of every liquid cell's index (splitmix64's finaliser).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__global__ __launch_bounds__(TPB) void mask_hash_kernel(const uint8_t* __restrict__ m, uint32_t n,
                                                        unsigned long long* __restrict__ out) {
  unsigned long long h = 0;
  for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
    if (m[i]) h += mix64((unsigned long long)i + 0x9e3779b97f4a7c15ull);
  const unsigned long long s = block_reduce_sum<unsigned long long>(h);
  if (threadIdx.x == 0 && s) atomicAdd(out, s);
}

// inv[ids[i]] = i for a permutation ids of [0, n); an id outside marks bad
__global__ void inverse_ids_kernel(const int32_t* __restrict__ ids, int64_t n, int32_t* __restrict__ inv,
                                   uint32_t* __restrict__ bad) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    const int32_t id = ids[i];
    if (id < 0 || id >= n)
      atomicOr(bad, 1u);
    else
      inv[id] = (int32_t)i;
  }
}

// Largest |difference| of positions and velocities between two particle sets
// matched by id (ref slot inv[id]); NaN counts as infinite.  out holds the
// float's bits (non-negative floats order as unsigned integers).
__global__ __launch_bounds__(TPB) void state_diff_kernel(const float* __restrict__ pos, const float* __restrict__ vel,
                                                         const int32_t* __restrict__ ids,
                                                         const float* __restrict__ rpos,
                                                         const float* __restrict__ rvel,
                                                         const int32_t* __restrict__ inv, int64_t n,
                                                         uint32_t* __restrict__ out) {
  float d = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    const int32_t id = ids[i];
    if (id < 0 || id >= n) {
      d = INFINITY;
      continue;
    }
    const int64_t j = inv[id];
    if (j < 0) {  // an id the reference set does not hold
      d = INFINITY;
      continue;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float dp = fabsf(pos[3 * i + a] - rpos[3 * j + a]);
      const float dv = fabsf(vel[3 * i + a] - rvel[3 * j + a]);
      d = fmaxf(d, dp != dp ? INFINITY : dp);
      d = fmaxf(d, dv != dv ? INFINITY : dv);
    }
  }
  __shared__ float red[TPB / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) d = fmaxf(d, __shfl_down(d, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int w = 0; w < TPB / 64; ++w) m = fmaxf(m, red[w]);
    if (m > 0.f) atomicMax(out, __float_as_uint(m));
  }
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
  if (n) launch(fill_kernel<float>, grid_for(n), TPB, 0, s, p, n, v);
}
void k_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
  if (n) launch(fill_kernel<uint32_t>, grid_for(n), TPB, 0, s, p, n, v);
}
void k_fill_f64(hipStream_t s, double* p, size_t n, double v) {
  if (n) launch(fill_kernel<double>, grid_for(n), TPB, 0, s, p, n, v);
}

void k_bin(hipStream_t s, const GridDev& g, const float* pos, int64_t n, int32_t* key, uint32_t* count) {
  if (n > 0)
    launch(bin_kernel, (unsigned)((n + TPB * AGG_PPT - 1) / (TPB * AGG_PPT)), TPB, 0, s, g, pos, n, key, count);
}

// 2:1 grading to its fixed point: level l_b ends as min over blocks q of
// l_q + (Chebyshev distance b..q), and terms with distance >= L - 1 cannot
// go below l_b <= L - 1, so L - 1 Jacobi passes reach it.
static void grade_passes(hipStream_t s, const int* D, int L, uint8_t* level, uint8_t* tmp) {
  const int nb = D[0] * D[1] * D[2];
  const int gr = (nb + TPB - 1) / TPB;
  for (int it = 0; it < L - 1; ++it) {
    launch(grade_kernel, gr, TPB, 0, s, D[0], D[1], D[2], level, tmp);
    copy_async(s, level, tmp, nb);
  }
}

void k_levels(hipStream_t s, const int* D, int L, const int* band, const uint32_t* count, uint8_t* surf,
              uint8_t* level, uint8_t* tmp) {
  const int nb = D[0] * D[1] * D[2];
  const int gr = (nb + TPB - 1) / TPB;
  launch(surface_kernel, gr, TPB, 0, s, D[0], D[1], D[2], count, surf);
  Bands bd;
  int bmax = 0;
  for (int i = 0; i < 4; ++i) bd.v[i] = band[i];
  for (int i = 0; i < L - 1; ++i) bmax = band[i] > bmax ? band[i] : bmax;
  // z, y, x passes: surf -> tmp -> level -> tmp (levels), then back into level
  launch(cheb_pass_kernel, gr, TPB, 0, s, D[0], D[1], D[2], 2, bmax, 1, 0, bd, surf, tmp);
  launch(cheb_pass_kernel, gr, TPB, 0, s, D[0], D[1], D[2], 1, bmax, 0, 0, bd, tmp, level);
  launch(cheb_pass_kernel, gr, TPB, 0, s, D[0], D[1], D[2], 0, bmax, 0, L, bd, level, tmp);
  copy_async(s, level, tmp, nb);
  grade_passes(s, D, L, level, tmp);
}

void k_resample(hipStream_t s, const GridDev& g, const float* field, const int* o, float* out) {
  launch(resample_kernel, dim3((o[0] + RS_X * RS_V - 1) / (RS_X * RS_V), (o[1] + RS_Y - 1) / RS_Y, o[2]), RS_X * RS_Y, 0, s, g, field, o[0], o[1], o[2], out);
}

void k_radius_levels(hipStream_t s, const GridDev& g, int logh, int L, const float* pos, const float* rad,
                     const int32_t* ids, int64_t n, uint32_t* rmap, uint8_t* level, uint8_t* tmp) {
  const int nb = g.nb;
  uint32_t *direct = rmap, *own = rmap + nb, *bad = rmap + 2 * (size_t)nb;
  k_fill_u32(s, rmap, 2 * (size_t)nb, (uint32_t)(L - 1));
  k_fill_u32(s, bad, 1, 0u);
  const int smax = max(g.D[0], max(g.D[1], g.D[2]));
  if (n > 0)
    launch(radius_mark_kernel, (unsigned)((n + TPB * RM_PPT - 1) / (TPB * RM_PPT)), TPB, 0, s, g, logh, L, smax, pos, rad,
                                                                                         ids, n, direct, own, bad);
  launch(radius_dilate_kernel, (nb + TPB - 1) / TPB, TPB, 0, s, g.D[0], g.D[1], g.D[2], direct, own, level);
  grade_passes(s, g.D, L, level, tmp);
}

size_t scan_scratch_words(int n) { return 2 * (size_t)(n / SCAN_CH + 2); }

// exclusive scan of n values (in, or cells / face slots from levels) into
// out[0..n], out[n] = total; inputs beyond one chunk use `scratch`
// (scan_scratch_words(n) words)
static void scan_any(hipStream_t s, const uint32_t* in, const uint8_t* lv, int mode, uint32_t* out, int n,
                     uint32_t* scratch) {
  if (n <= SCAN_CH) {
    launch(scan_kernel, 1, SCAN_T, 0, s, in, lv, mode, out, n);
    return;
  }
  const int nch = (n + SCAN_CH - 1) / SCAN_CH;
  uint32_t* sums = scratch;
  uint32_t* cbase = scratch + nch;
  launch(scan_reduce_kernel, nch, TPB, 0, s, in, lv, mode, n, sums);
  launch(scan_kernel, 1, SCAN_T, 0, s, sums, nullptr, 0, cbase, nch);
  launch(scan_final_kernel, nch, TPB, 0, s, in, lv, mode, n, cbase, nch, out);
}

void k_scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, int n, uint32_t* scratch) {
  scan_any(s, in, nullptr, 0, out, n, scratch);
}

void k_cell_offsets(hipStream_t s, const uint8_t* level, int nb, uint32_t* coff, uint32_t* scratch) {
  scan_any(s, nullptr, level, 1, coff, nb, scratch);
}

void k_layout_offsets(hipStream_t s, const uint8_t* level, int nb, uint32_t* coff, uint32_t* fx, uint32_t* fy,
                      uint32_t* fz, uint32_t* scratch) {
  scan_any(s, nullptr, level, 1, coff, nb, scratch);
  scan_any(s, nullptr, level, 2, fx, nb, scratch);
  if (fy != fx) copy_async(s, fy, fx, (nb + 1) * sizeof(uint32_t));
  if (fz != fx) copy_async(s, fz, fx, (nb + 1) * sizeof(uint32_t));
}

void k_scatter(hipStream_t s, const float* pos, const float* vel, const int32_t* ids, const int32_t* key,
               uint32_t* cursor, int64_t n, float* spos, float* svel, int32_t* sids, const float* dens, float* sdens,
               uint32_t* vbound) {
  launch(fill_kernel<uint32_t>, 1, 64, 0, s, vbound, 2 * VB_SLOTS, 0u);
  if (n > 0)
    launch(scatter_kernel, (unsigned)((n + TPB * AGG_PPT - 1) / (TPB * AGG_PPT)), TPB, 0, s, pos, vel, ids, key,
           cursor, n, spos, svel, sids, dens, sdens, vbound);
}

void k_vbound(hipStream_t s, const float* v, int64_t n, uint32_t* out) {
  launch(fill_kernel<uint32_t>, 1, 64, 0, s, out, VB_SLOTS, 0u);
  const int64_t wg = (n + TPB - 1) / TPB;
  if (n > 0) launch(vbound_kernel, (unsigned)(wg < 2048 ? wg : 2048), TPB, 0, s, v, n, out);
}

void k_init_beta(hipStream_t s, const LevelDev& L, float* bx, float* by, float* bz) {
  launch(init_beta_kernel, (L.nb + TPB / 64 - 1) / (TPB / 64), TPB, 0, s, L, bx, by, bz);
}

void k_p2g_deposit(hipStream_t s, const GridDev& g, const float* spos, const float* svel, const uint32_t* pcount,
                   const uint32_t* poff, float* slab, uint8_t* liquid, const uint32_t* vbound, const float* sdens) {
  if (sdens)
    launch(p2g_deposit_kernel<true>, g.nb, DEP_T, 0, s, g, spos, sdens, pcount, poff, slab, liquid, vbound);
  else
    launch(p2g_deposit_kernel<false>, g.nb, DEP_T, 0, s, g, spos, svel, pcount, poff, slab, liquid, vbound);
}

void k_p2g_gather(hipStream_t s, const GridDev& g, const uint32_t* pcount, const float* slab, float* const* Macc,
                  float* const* macc) {
  launch(p2g_gather_kernel, g.nb, TPB, 0, s, g, pcount, slab, Macc[0], Macc[1], Macc[2], macc[0], macc[1], macc[2]);
}

static FacePtrs fp(float* const* p) { return FacePtrs{{p[0], p[1], p[2]}}; }

void k_fold(hipStream_t s, const GridDev& g, float* const* Macc, float* const* macc, float* const* uold,
            float* const* ustar, float* const* mass, float dt, const float* grav) {
  launch(fold_kernel, g.nb, TPB, 0, s, g, fp(Macc), fp(macc), fp(uold), fp(ustar), fp(mass), dt, grav[0], grav[1],
                                    grav[2]);
}

void k_beta_fold(hipStream_t s, const GridDev& g, float* const* macc, float* const* Dacc, float* const* beta) {
  launch(beta_fold_kernel, g.nb, TPB, 0, s, g, fp(macc), fp(Dacc), fp(beta));
}

void k_divergence(hipStream_t s, const GridDev& g, const uint8_t* liquid, float* const* ustar, float* rhs) {
  launch(divergence_kernel, g.nb, TPB, 0, s, g, liquid, fp(ustar), rhs);
}

void k_project(hipStream_t s, const GridDev& g, const float* p, float* const* ustar, float* const* beta,
               float* const* unew) {
  launch(project_kernel, g.nb, TPB, 0, s, g, p, fp(ustar), fp(beta), fp(unew));
}

void k_pressure_remap(hipStream_t s, const GridDev& g, const uint8_t* level_prev, const uint32_t* coff_prev,
                      const float* p_prev, float* p) {
  launch(pressure_remap_kernel, g.nb, TPB, 0, s, g, level_prev, coff_prev, p_prev, p);
}

void k_to_fine(hipStream_t s, const GridDev& g, float* fine, const uint32_t* count, uint8_t* halo) {
  launch(halo_flags_kernel, (g.nb + TPB - 1) / TPB, TPB, 0, s, g, count, halo);
  launch(to_fine_kernel, g.nb, TPB, 0, s, g, fine);
}

template <int MODE>
static void g2p_launch(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const float* svel,
                       int64_t n, float alpha, float dt, float* pos_out, float* vel_out, uint32_t* list) {
  launch(fill_kernel<uint32_t>, 1, 64, 0, s, list, 1, 0u);
  launch(g2p_advect_kernel<false, MODE>, (unsigned)((n + TPB - 1) / TPB), TPB, 0, s, g, fine, spos, svel, n, alpha, dt,
                                                                                  pos_out, vel_out, list);
  const int64_t nwg = (n + TPB - 1) / TPB;
  const unsigned fix = (unsigned)(nwg < 1024 ? nwg : 1024);
  launch(g2p_advect_kernel<true, MODE>, fix, TPB, 0, s, g, fine, spos, svel, n, alpha, dt, pos_out, vel_out, list);
}

void k_g2p_advect(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const float* svel,
                  int64_t n, float alpha, float dt, float* pos_out, float* vel_out, uint32_t* list, int mode) {
  if (n <= 0) return;
  if (mode == G2P_VEL)
    g2p_launch<G2P_VEL>(s, g, fine, spos, svel, n, alpha, dt, pos_out, vel_out, list);
  else if (mode == G2P_POS)
    g2p_launch<G2P_POS>(s, g, fine, spos, svel, n, alpha, dt, pos_out, vel_out, list);
  else
    g2p_launch<G2P_BOTH>(s, g, fine, spos, svel, n, alpha, dt, pos_out, vel_out, list);
}

void k_g2p_advect_sorted(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const float* svel,
                         const uint32_t* pcount, const uint32_t* poff, int64_t n, float alpha, float dt,
                         float* pos_out, float* vel_out, uint32_t* list) {
  if (n <= 0) return;
  launch(fill_kernel<uint32_t>, 1, 64, 0, s, list, 1, 0u);
  launch(g2p_block_kernel, g.nb, G2B_T, 0, s, g, fine, spos, svel, pcount, poff, alpha, dt, pos_out, vel_out, list);
  const int64_t nwg = (n + TPB - 1) / TPB;
  const unsigned fix = (unsigned)(nwg < 1024 ? nwg : 1024);
  launch(g2p_advect_kernel<true, G2P_BOTH>, fix, TPB, 0, s, g, fine, spos, svel, n, alpha, dt, pos_out, vel_out,
         list);
}

size_t adj_record_bytes() { return sizeof(AdjRec); }

void k_g2p_adjoint_stages(hipStream_t s, const GridDev& g, const float* fine, const float* spos, const int32_t* sids,
                          int64_t n, float alpha, float dt, float* xbar, float* vbar, void* rec, uint32_t* list,
                          const uint32_t* pcount, const uint32_t* poff) {
  if (n <= 0) return;
  AdjRec* r = static_cast<AdjRec*>(rec);
  launch(fill_kernel<uint32_t>, 1, 64, 0, s, list, 1, 0u);
  const int64_t nwg = (n + TPB - 1) / TPB;
  if (pcount)
    launch(g2p_adjoint_block_kernel, g.nb, G2B_T, 0, s, g, fine, spos, sids, pcount, poff, alpha, dt, xbar, vbar, r,
           list);
  else
    launch(g2p_adjoint_stage_kernel<false>, (unsigned)nwg, TPB, 0, s, g, fine, spos, sids, n, alpha, dt, xbar, vbar,
           r, list);
  const unsigned fix = (unsigned)(nwg < 1024 ? nwg : 1024);
  launch(g2p_adjoint_stage_kernel<true>, fix, TPB, 0, s, g, fine, spos, sids, n, alpha, dt, xbar, vbar, r, list);
}

void k_g2p_adjoint_scatter(hipStream_t s, const GridDev& g, const float* spos, const uint32_t* pcount,
                           const uint32_t* poff, const void* rec, float* fbar, uint8_t* touched) {
  launch(g2p_adjoint_scatter_kernel, g.nb, ADJ_T, 0, s, g, spos, pcount, poff, static_cast<const AdjRec*>(rec), fbar,
         touched);
}
void k_fbar_clear(hipStream_t s, int nb, float* fbar, uint8_t* touched) {
  const int per = TPB / 64;
  if (nb > 0) launch(fbar_clear_kernel, (nb + per - 1) / per, TPB, 0, s, nb, fbar, touched);
}
void k_pi_transpose(hipStream_t s, const GridDev& g, const float* fbar, const uint8_t* touched, float* const* unb,
                    float* const* uob) {
  launch(pi_transpose_kernel, g.nb, TPB, 0, s, g, fbar, touched, fp(unb), fp(uob));
}
void k_project_adjoint(hipStream_t s, const GridDev& g, float* const* beta, float* const* unb, float* const* usb,
                       float* pbar) {
  launch(project_adjoint_kernel, g.nb, TPB, 0, s, g, fp(beta), fp(unb), fp(usb), pbar);
}
void k_divergence_adjoint(hipStream_t s, const GridDev& g, const uint8_t* liquid, const float* rhsbar,
                          float* const* usb) {
  launch(divergence_adjoint_kernel, g.nb, TPB, 0, s, g, liquid, rhsbar, fp(usb));
}
void k_fold_adjoint(hipStream_t s, const GridDev& g, float* const* uold, float* const* mass, float* const* uob,
                    float* const* usb, float* const* Mb, float* const* mb) {
  launch(fold_adjoint_kernel, g.nb, TPB, 0, s, g, fp(uold), fp(mass), fp(uob), fp(usb), fp(Mb), fp(mb));
}
void k_pi_forward2(hipStream_t s, const GridDev& g, float* const* A, float* const* Bf, float* out) {
  launch(pi_forward2_kernel, g.nb, TPB, 0, s, g, fp(A), fp(Bf), out);
}
void k_p2g_adjoint(hipStream_t s, const GridDev& g, const float* fm, const float* spos, const float* svel,
                   const int32_t* sids, int64_t n, float* xbar, float* vbar, const uint32_t* pcount,
                   const uint32_t* poff) {
  if (n <= 0) return;
  if (pcount)
    launch(p2g_adjoint_block_kernel, g.nb, G2B_T, 0, s, g, fm, spos, svel, sids, pcount, poff, xbar, vbar);
  else
    launch(p2g_adjoint_kernel, (unsigned)((n + TPB - 1) / TPB), TPB, 0, s, g, fm, spos, svel, sids, n, xbar, vbar);
}
void k_ids_range(hipStream_t s, const int32_t* ids, int64_t n, int* out) {
  int64_t blocks = (n + TPB - 1) / TPB;
  if (blocks > 2048) blocks = 2048;
  if (n > 0) launch(ids_range_kernel, (unsigned)blocks, TPB, 0, s, ids, n, out);
}
void k_masked_copy(hipStream_t s, float* dst, const float* src, const uint8_t* liquid, int n) {
  if (n > 0) launch(masked_copy_kernel, (n + TPB - 1) / TPB, TPB, 0, s, dst, src, liquid, n);
}

void k_smooth(hipStream_t s, const LevelDev& L, float* x, const float* b, const WGMap& wg,
              const uint32_t* blist, float omega, int local_sweeps, int reverse) {
  if (wg.n > 0) launch(smooth_kernel, wg.n, ST_T, 0, s, L, x, b, wg, blist, omega, local_sweeps, reverse);
}

void k_block_live(hipStream_t s, const LevelDev& L, uint8_t* live) {
  const int per = LIVE_T / 64;
  if (L.nb > 0) launch(block_live_kernel, (L.nb + per - 1) / per, LIVE_T, 0, s, L, live);
}

void k_gather_bd(hipStream_t s, const uint32_t* bd, const uint32_t* idx, int n, uint32_t* out) {
  const int words = n * BD_WORDS;
  if (n > 0) launch(gather_bd_kernel, (words + TPB - 1) / TPB, TPB, 0, s, bd, idx, n, out);
}

void k_smooth_adaptive(hipStream_t s, const LevelDev& L, float* x, const float* b, const WGMap& wg, float omega,
                       int reverse, const uint8_t* act, uint8_t* act_next, uint8_t* sweeps, const float* bmax,
                       float theta, int max_sweeps, float reduction) {
  if (wg.n > 0)
    launch(adaptive_smooth_kernel, wg.n, ST_T, 0, s, L, x, b, wg, omega, reverse, act, act_next, sweeps, bmax, theta,
                                                 max_sweeps, reduction);
}

void k_absmax(hipStream_t s, const float* v, const uint8_t* liquid, int n, float* out) {
  k_fill_u32(s, reinterpret_cast<uint32_t*>(out), 1, 0u);
  if (n > 0) launch(absmax_kernel, grid_for(n) < 1024 ? grid_for(n) : 1024, TPB, 0, s, v, liquid, n, reinterpret_cast<uint32_t*>(out));
}

void k_fill_u8(hipStream_t s, uint8_t* p, size_t n, uint8_t v) {
  if (n) launch(fill_kernel<uint8_t>, grid_for(n), TPB, 0, s, p, n, v);
}

void k_smooth_uniform(hipStream_t s, const int* sh, int lr, int e, const uint8_t* liquid, float* x, const float* b,
                      int col, float omega, int local_sweeps, int reverse) {
  ULevel U;
  for (int d = 0; d < 3; ++d) U.sh[d] = sh[d];
  U.lr = lr;
  U.e = e;
  const int ncol = (1 << (sh[0] + sh[1] + sh[2])) >> 1;
  const int per = 512 >> (3 * lr);
  const int nwg = (ncol + per - 1) / per;
  if (nwg > 0)
    launch(uniform_smooth_kernel, nwg, ST_T, 0, s, U, liquid, x, b, col, omega, local_sweeps, reverse);
}

void k_residual(hipStream_t s, const LevelDev& L, const float* x, const float* b, float* r, const WGMap& wg,
                const uint32_t* blist, const Restr& rs) {
  if (wg.n > 0) launch(residual_kernel, wg.n, ST_T, 0, s, L, x, b, r, wg, blist, rs);
}

void k_apply_dot(hipStream_t s, const LevelDev& L, const float* x, float* y, const WGMap& wg,
                 const uint32_t* blist, double* acc) {
  if (wg.n > 0) launch(apply_dot_kernel, wg.n, ST_T, 0, s, L, x, y, wg, blist, acc);
}

void k_restrict(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, const float* rf, float* bc,
                float* xc, const WGMap& wgc, const uint32_t* blistc) {
  if (wgc.n > 0) launch(restrict_kernel, wgc.n, TPB, 0, s, F, C, aggregated, rf, bc, xc, wgc, blistc);
}

void k_prolong(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, float* xf, const float* xc,
               float scale, const WGMap& wgf, const uint32_t* blistf) {
  if (wgf.n > 0) launch(prolong_kernel, wgf.n, ST_T, 0, s, F, C, aggregated, xf, xc, scale, wgf, blistf);
}

void k_coarsen_liquid(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, uint8_t* liquid_c,
                      const WGMap& wgc, const uint32_t* blistc) {
  if (wgc.n > 0) launch(coarsen_liquid_kernel, wgc.n, TPB, 0, s, F, C, aggregated, liquid_c, wgc, blistc);
}

void k_coarsen_beta(hipStream_t s, const LevelDev& F, const LevelDev& C, int aggregated, float* bx, float* by,
                    float* bz) {
  launch(coarsen_beta_kernel, C.nb, TPB, 0, s, F, C, aggregated, bx, by, bz);
}

void k_tail(hipStream_t s, const TailArgs& args, size_t lds_bytes) {
  launch(tail_kernel, 1, TAIL_T, lds_bytes, s, args);
}

void k_coarse_factor(hipStream_t s, const LevelDev& L, int n, double* inv) {
  launch(coarse_factor_kernel, 1, TPB, 0, s, L, n, inv);
}

void k_residual64(hipStream_t s, const LevelDev& L, const double* x, const float* b, float* r32, const WGMap& wg,
                  const uint32_t* blist, double* rr) {
  if (wg.n > 0) launch(residual64_kernel, wg.n, TPB, 0, s, L, x, b, r32, wg, blist, rr);
}

void k_axpy64(hipStream_t s, double* x, const float* e, const uint8_t* liquid, int n) {
  launch(axpy64_kernel, red_grid(n), TPB, 0, s, x, e, liquid, n);
}

void k_to64(hipStream_t s, double* x, const float* p, const uint8_t* liquid, int n) {
  launch(to64_kernel, red_grid(n), TPB, 0, s, x, p, liquid, n);
}

void k_to32(hipStream_t s, float* p, const double* x, int n) { launch(to32_kernel, red_grid(n), TPB, 0, s, p, x, n); }

void k_dot(hipStream_t s, const float* a, const float* b, const uint8_t* liquid, int n, double* acc) {
  launch(dot_kernel, red_grid(n), TPB, 0, s, a, b, liquid, n, acc);
}

void k_scalar_shift(hipStream_t s, double* sc) { launch(scalar_shift_kernel, 1, 64, 0, s, sc); }

void k_pcg_p(hipStream_t s, const float* z, float* p, int n, const double* sc, int first) {
  launch(pcg_p_kernel, red_grid(n), TPB, 0, s, z, p, n, sc, first);
}

void k_pcg_xr(hipStream_t s, float* x, float* r, const float* p, const float* q, const uint8_t* liquid, int n,
              double* sc) {
  launch(pcg_xr_kernel, red_grid(n), TPB, 0, s, x, r, p, q, liquid, n, sc);
}

void k_mask_hash(hipStream_t s, const uint8_t* m, uint32_t n, unsigned long long* out) {
  k_fill_u32(s, reinterpret_cast<uint32_t*>(out), 2, 0u);
  if (n > 0) launch(mask_hash_kernel, grid_for(n) < 1024 ? grid_for(n) : 1024, TPB, 0, s, m, n, out);
}

void k_state_diff(hipStream_t s, const float* pos, const float* vel, const int32_t* ids, const float* rpos,
                  const float* rvel, const int32_t* rids, int64_t n, int32_t* inv, uint32_t* out) {
  // out[0]: max difference (float bits), out[1]: an id outside [0, n)
  k_fill_u32(s, out, 2, 0u);
  if (n <= 0) return;
  const unsigned grid = grid_for((size_t)n) < 2048 ? grid_for((size_t)n) : 2048;
  k_fill_u32(s, reinterpret_cast<uint32_t*>(inv), (size_t)n, 0xFFFFFFFFu);
  launch(inverse_ids_kernel, grid, TPB, 0, s, rids, n, inv, out + 1);
  launch(state_diff_kernel, grid, TPB, 0, s, pos, vel, ids, rpos, rvel, inv, n, out);
}

void k_count_nonzero(hipStream_t s, const uint8_t* m, uint32_t n, uint32_t* out) {
  k_fill_u32(s, out, 1, 0u);
  if (n > 0) launch(count_nonzero_kernel, grid_for(n) < 1024 ? grid_for(n) : 1024, TPB, 0, s, m, n, out);
}

void k_mask(hipStream_t s, float* x, const uint8_t* liquid, int n) {
  launch(mask_kernel, red_grid(n), TPB, 0, s, x, liquid, n);
}

void k_add(hipStream_t s, float* x, const float* z, const uint8_t* liquid, int n) {
  launch(add_kernel, red_grid(n), TPB, 0, s, x, z, liquid, n);
}

}  // namespace aw
