// Host runtime of the MI355X MSBG FLIP step: contexts, grids, the topology
// pipeline, the multigrid hierarchy and the solver loops behind the C ABI in
// include/aw.h.  Device work is enqueued on the context's stream; host
// synchronisation happens only where the ABI says so.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <atomic>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/aw.h"
#include "aw_internal.h"

using namespace aw;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define AW_HIP(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(AW_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define AW_TRY(expr)         \
  do {                       \
    int s_ = (expr);         \
    if (s_ != AW_OK) return s_; \
  } while (0)

// device memory the library holds (all contexts), and its high-water mark
std::atomic<uint64_t> g_mem_cur{0}, g_mem_peak{0};
void mem_add(uint64_t b) {
  const uint64_t now = g_mem_cur.fetch_add(b) + b;
  uint64_t pk = g_mem_peak.load();
  while (now > pk && !g_mem_peak.compare_exchange_weak(pk, now)) {
  }
}
void mem_sub(uint64_t b) { g_mem_cur.fetch_sub(b); }
// cap on the device memory the library holds (aw_set_memory_limit), 0: none
std::atomic<uint64_t> g_mem_limit{0};
int over_limit(uint64_t want) {
  const uint64_t lim = g_mem_limit.load(), cur = g_mem_cur.load();
  if (lim == 0 || cur + want <= lim) return AW_OK;
  return fail(AW_ERR_OOM, "allocation of %llu bytes would pass the memory limit (%llu of %llu bytes held)",
              (unsigned long long)want, (unsigned long long)cur, (unsigned long long)lim);
}
std::mutex g_user_mu;
std::unordered_map<void*, uint64_t> g_user_alloc;  // aw_malloc'd buffers and their sizes

// Scratch of the entry points that allocate nothing (aw_remap_cells_noalloc):
// the previous layout's cell offsets (num_blocks + 1 words) and the chunk
// sums of their scan, for grids of up to NOALLOC_MAX_BLOCKS blocks.  Each GPU
// holds its own copy until the process ends; host code uses only the device
// address resolved per device (scratch_dev), never the array's name.
constexpr int NOALLOC_MAX_BLOCKS = 1 << 18;
constexpr size_t NOALLOC_SCRATCH_WORDS = (size_t)NOALLOC_MAX_BLOCKS + 1024;
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device `device` (looked up once per device)
int scratch_address(int device, uint32_t** out) {
  *out = nullptr;
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return AW_OK;  // that device has no no-allocation entry points
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return fail(AW_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch_dev[device] = static_cast<uint32_t*>(ptr);
  }
  *out = g_scratch_dev[device];
  return AW_OK;
}

// grow-only device allocation
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) {
      (void)hipFree(p);
      mem_sub(cap);
    }
  }
  int ensure(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (bytes <= cap) return AW_OK;
    if (p) {
      (void)hipFree(p);
      mem_sub(cap);
    }
    // a buffer that has to grow gets 1/8 headroom, so that sizes that creep
    // up over topology changes do not reallocate every time
    const size_t want = cap ? bytes + bytes / 8 : bytes;
    p = nullptr;
    cap = 0;
    AW_TRY(over_limit(want));
    if (hipMalloc(&p, want) != hipSuccess) {
      (void)hipGetLastError();
      return fail(AW_ERR_OOM, "hipMalloc of %zu bytes failed", want);
    }
    cap = want;
    mem_add(cap);
    return AW_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  // exchange the allocations (no copy; kernels already enqueued keep the
  // pointers they were launched with)
  void swap(DevBuf& o) noexcept {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
};

int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(AW_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  return AW_OK;
}

void neighbour_table(const int* D, std::vector<int32_t>& nbr) {
  const int nb = D[0] * D[1] * D[2];
  nbr.assign((size_t)nb * 6, -1);
  const int st[3] = {1, D[0], D[0] * D[1]};
  int b = 0;
  for (int z = 0; z < D[2]; ++z)
    for (int y = 0; y < D[1]; ++y)
      for (int x = 0; x < D[0]; ++x, ++b) {
        const int c[3] = {x, y, z};
        int32_t* n = &nbr[(size_t)b * 6];
        for (int a = 0; a < 3; ++a) {
          if (c[a] > 0) n[2 * a] = b - st[a];
          if (c[a] + 1 < D[a]) n[2 * a + 1] = b + st[a];
        }
      }
}

}  // namespace

// Optional per-launch timing with HIP events on the context's stream.
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
  double bytes;
  long launches;
};

// Handles are reference-counted by the handles made from them (grids hold
// their context, solvers their grid, tapes their solver, aw_malloc buffers
// their context): destroying a handle that others still use only marks it
// closed, and it is freed with its last user, whatever order callers (or a
// garbage collector) destroy them in.
struct aw_ctx {
  int refs = 0;
  bool closed = false;
  int device = 0;
  uint32_t* scratch_dev = nullptr;  // this device's copy of the no-allocation scratch
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool prof = false;
  bool prof_runs = false;   // consecutive launches of one name share an event pair
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t event() {
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      (void)hipEventCreate(&e);
      pool.push_back(e);
    }
    return pool[used++];
  }
};

namespace {
// RAII: times the kernel launches in its scope with an event pair that the
// launches' own dispatches record (start by the first, stop by each), with
// their algorithmic bytes; no marker packets go between kernels.
struct Prof {
  aw_ctx* c;
  long idx = -1;
  Prof(aw_ctx* ctx, const char* name, double bytes) : c(ctx) {
    if (!c->prof) return;
    if (c->prof_runs && !c->recs.empty() && std::strcmp(c->recs.back().name, name) == 0) {
      // extends the run: this scope's launches move its stop event on
      idx = (long)c->recs.size() - 1;
      c->recs[idx].bytes += bytes;
      c->recs[idx].launches++;
      set_launch_events(nullptr, c->recs[idx].b);
      return;
    }
    ProfRec r{name, c->event(), c->event(), bytes, 1};
    c->recs.push_back(r);
    idx = (long)c->recs.size() - 1;
    set_launch_events(r.a, r.b);
  }
  ~Prof() {
    if (idx < 0) return;
    const LaunchEvents e = take_launch_events();
    if (e.start) {
      // nothing was launched in a new record's scope: an empty interval here
      (void)hipEventRecord(c->recs[idx].a, c->stream);
      (void)hipEventRecord(c->recs[idx].b, c->stream);
    }
  }
};
}  // namespace

struct aw_grid {
  int refs = 0;
  bool closed = false;
  aw_ctx* ctx = nullptr;
  aw_grid_desc desc{};
  int D[3]{}, N[3]{};
  int nb = 0;
  float h = 0.f, inv_h = 0.f;
  DevBuf level, surf, ltmp, count, poff, cursor, coff, foff, nbr, key, halo, g2p_list, scan_tmp, rmap;
  DevBuf spos, svel, sids;
  // largest |vel| and |density| of the particles in spos / svel, sdens (float
  // bits), recorded by the binning sort; vbound_of is the velocity array
  // they describe (another array gets its own pass before a deposit)
  DevBuf vbound;
  const void* vbound_of = nullptr;
  const void* vbound_dens_of = nullptr;
  DevBuf liquid, p, rhs, slab;
  DevBuf beta[3], Macc[3], macc[3], uold[3], ustar[3], unew[3], mass[3];
  // adjoint fields: fine-lattice adjoints (two fields each, stride ADJ), per-slot and per-cell adjoints
  DevBuf fbar, fmbar, pbar, rhsbar, idrange, adjrec;
  // blocks whose fbar the last G2P adjoint scattered into (the rest of fbar
  // is 0), and the fbar allocation those flags describe
  DevBuf fbar_touched;
  const void* fbar_zeroed_for = nullptr;
  DevBuf liq_count;                  // liquid cells of the last step, read back with the step's final sync
  // two-phase coefficients: beta = 1 / rho from a per-particle density (the
  // last step deposited one), its density-weighted deposit and sorted densities
  bool beta_variable = false;
  DevBuf Dacc[3], sdens;
  uint32_t* h_liq_count = nullptr;   // pinned host copy
  uint8_t* h_level_read = nullptr;   // pinned host copy of the level map (nb bytes), read back per topology
  hipEvent_t level_read_done = nullptr;
  DevBuf ubn[3], ubo[3], ubs[3], Mb[3], mb[3];
  bool sorted_ids = false;  // the last binning carried particle ids
  std::vector<uint8_t> h_level;
  std::vector<uint32_t> h_coff;
  bool have_topology = false;
  int topo_version = 0;
  // warm start: the pressure of the last step's solve, its layout version,
  // and that layout's levels and cell offsets once the topology moves on
  bool p_valid = false;
  int p_version = -1;
  DevBuf p_prev, level_prev, coff_prev, remap_coff;
  uint32_t ncells = 0, nfaces = 0;
  int64_t n_sorted = -1;
  double occupied = 0.0;   // blocks holding particles (for byte accounting only)
  double lattice_blocks = 0.0, lattice_faces = 0.0;   // blocks with a filled fine lattice, their face slots (same)

  GridDev view() const {
    GridDev g;
    for (int a = 0; a < 3; ++a) {
      g.D[a] = D[a];
      g.N[a] = N[a];
      g.foff[a] = foff.as<uint32_t>();
      g.un[a] = unew[a].as<float>();
      g.uo[a] = uold[a].as<float>();
    }
    g.nb = nb;
    g.h = h;
    g.inv_h = inv_h;
    g.level = level.as<uint8_t>();
    g.coff = coff.as<uint32_t>();
    g.nbr = nbr.as<int32_t>();
    g.halo = halo.as<uint8_t>();
    g.count = count.as<uint32_t>();
    return g;
  }
  float* face(DevBuf* f, int a) const { return f[a].as<float>(); }
};

namespace {

struct MGLevel {
  int D[3]{};
  int nb = 0, bs = BASE, aggregated = 0;
  uint32_t ncells = 0, nfaces = 0;
  std::vector<uint8_t> h_e;
  std::vector<uint32_t> h_coff;
  // host sources of the uploaded tables, kept for the level's lifetime so no
  // asynchronous copy can outlive its source
  std::vector<uint32_t> h_blist, h_bd, h_cblk, h_foff;
  std::vector<int32_t> h_nbr;
  int nbr_dims[3] = {0, 0, 0};   // dims h_nbr was built for
  bool valid = false;            // tables and device buffers built for h_e
  // a coarse level with one block resolution and power-of-two dims: log2
  // dims (ush[0] < 0 otherwise), log2 resolution and cell edge
  int ush[3] = {-1, -1, -1};
  int ulr = 0, ue = 0;
  DevBuf e, coff, foff, nbr, liquid, beta[3], x, b, r, blist, cblk, bdesc;
  DevBuf act[2], sweeps;           // adaptive smoothing: active flags (current, next), sweeps run per block
  int act_cur = 0;
  bool act_valid = false;
  WGMap wgm[3]{};  // stencil launches over block colour 0, colour 1, both
  // algorithmic bytes of one stencil pass over colour 0, colour 1 (x read +
  // x write + b read = 12 B per cell, beta 3 x 4 B per face slot)
  double sweep_bytes[2] = {0.0, 0.0};
  int tables_version = 0;  // bumped with every upload of the work list and descriptors
  // The smoother's launches without the all-air blocks (it writes nothing on
  // air cells, so leaving them out changes no value): per block colour the
  // runs of wgm over the work-list entries whose block holds a liquid cell,
  // in work-list order, their descriptors gathered into bdesc_live.  Set by
  // refresh_live from the liquid masks; live_on false: the full lists.
  bool live_on = false;
  int live_tables_version = -1;
  WGMap wgm_live[3]{};  // colour 0, colour 1, both (the runs of colour 0, then those of colour 1)
  DevBuf bdesc_live, live_idx;
  std::vector<uint32_t> h_live_idx;  // work-list entries of bdesc_live (host source of live_idx)
  uint32_t live_blocks = 0;
  double live_bytes[2] = {0.0, 0.0};
  double live_cells = 0.0, live_ccells = 0.0;  // cells of those blocks, and their cells on the next level
  // device views (level 0 aliases the grid)
  const uint8_t* d_e = nullptr;
  const uint32_t* d_coff = nullptr;
  const uint32_t* d_foff = nullptr;
  const int32_t* d_nbr = nullptr;
  uint8_t* d_liq = nullptr;
  float* d_beta[3] = {nullptr, nullptr, nullptr};
  float* d_x = nullptr;
  float* d_b = nullptr;
  float* d_r = nullptr;

  LevelDev view() const {
    LevelDev L;
    for (int a = 0; a < 3; ++a) {
      L.D[a] = D[a];
      L.foff[a] = d_foff;
      L.beta[a] = d_beta[a];
    }
    L.nb = nb;
    L.bs = bs;
    L.e = d_e;
    L.coff = d_coff;
    L.nbr = d_nbr;
    L.liquid = d_liq;
    L.bd = bdesc.as<uint32_t>();
    return L;
  }
  LevelDev live_view() const {
    LevelDev L = view();
    L.bd = bdesc_live.as<uint32_t>();
    return L;
  }
  const uint32_t* bl() const { return blist.as<uint32_t>(); }
};

}  // namespace

struct aw_solver {
  int refs = 0;
  bool closed = false;
  aw_grid* g = nullptr;
  aw_solver_params prm{};
  std::vector<std::unique_ptr<MGLevel>> lv;
  int built_version = -1;
  int tail = 0;          // first level run by the single-work-group tail kernel
  size_t tail_lds = 0;
  DevBuf cinv;           // FP64 inverse of a small coarsest level
  bool direct = false;
  DevBuf x64, e32, r32;  // FP64 iterative refinement: solution, FP32 correction, FP32 residual
  DevBuf r0, z, pv, q, sc;
  DevBuf bmax;           // adaptive smoothing: max |b| of the level being smoothed
  DevBuf live_flags;     // per block of the list-smoothed levels: holds a liquid cell
  uint8_t* h_live = nullptr;  // pinned host copy of live_flags
  size_t h_live_cap = 0;
  double* h_sc = nullptr;       // host scalars (slots added), points at h_sc_sum
  double h_sc_sum[SC_N] = {};
  double* h_slots = nullptr;    // pinned host copy of the slotted device scalars
  // The PCG iteration body (V-cycle preconditioner and vector updates, all
  // between two scalar readbacks) replayed from a captured HIP graph:
  // [0] continuing iteration, [1] first iteration after a (re)start.  The
  // kernels' arguments are baked in, so a graph belongs to one hierarchy
  // build (graph_epoch) and one solution buffer.
  struct Captured {
    hipGraphExec_t exec = nullptr;
    int epoch = -1;
    const void* x = nullptr;
  };
  Captured pcg_body[2];
  int graph_epoch = 0;
  hipEvent_t upload_event = nullptr;  // after the last hierarchy build's table uploads
  bool upload_event_set = false;
  bool use_graph = false;  // aw_solver_set_graph
  // the coarse levels' beta arrays hold coarsened variable coefficients (they
  // are coarsened again, to the uniform 1 / 0, once the grid's beta is uniform)
  bool coarse_beta_variable = false;
  bool capture_ok = true;  // false once capture failed (e.g. on the legacy null stream): run eagerly
};

namespace {

int apply_levels(aw_grid* g) {
  hipStream_t s = g->ctx->stream;
  const int nb = g->nb;
  uint64_t nc = 0, nf = 0;
  g->h_coff.assign(nb + 1, 0);
  for (int b = 0; b < nb; ++b) {
    const int r = BASE >> g->h_level[b];
    g->h_coff[b + 1] = g->h_coff[b] + r * r * r;
    nc += (uint64_t)r * r * r;
    nf += (uint64_t)(r + 1) * r * r;
  }
  if (nf > 0xFFFFFFFFull) return fail(AW_ERR_INVALID, "grid too large for 32-bit offsets");
  g->ncells = (uint32_t)nc;
  g->nfaces = (uint32_t)nf;
  AW_TRY(g->coff.ensure((nb + 1) * sizeof(uint32_t)));
  AW_TRY(g->foff.ensure((nb + 1) * sizeof(uint32_t)));
  // countBlockElements + exclusive_scan on the device (MSBG.cuh:353-376, main.cu:270-273)
  AW_TRY(g->scan_tmp.ensure(scan_scratch_words(nb) * sizeof(uint32_t)));
  k_layout_offsets(s, g->level.as<uint8_t>(), nb, g->coff.as<uint32_t>(), g->foff.as<uint32_t>(),
                   g->foff.as<uint32_t>(), g->foff.as<uint32_t>(), g->scan_tmp.as<uint32_t>());
  AW_TRY(g->liquid.ensure(nc));
  AW_TRY(g->p.ensure(nc * sizeof(float)));
  AW_TRY(g->rhs.ensure(nc * sizeof(float)));
  for (int a = 0; a < 3; ++a)
    for (DevBuf* f : {g->beta, g->Macc, g->macc, g->uold, g->ustar, g->unew, g->mass})
      AW_TRY(f[a].ensure(nf * sizeof(float)));
  AW_TRY(g->slab.ensure((size_t)nb * SLAB * sizeof(float)));
  AW_TRY(g->halo.ensure((size_t)nb));
  LevelDev L;
  for (int a = 0; a < 3; ++a) {
    L.D[a] = g->D[a];
    L.foff[a] = g->foff.as<uint32_t>();
    L.beta[a] = g->beta[a].as<float>();
  }
  L.nb = nb;
  L.bs = BASE;
  L.e = g->level.as<uint8_t>();
  L.coff = g->coff.as<uint32_t>();
  L.nbr = g->nbr.as<int32_t>();
  L.liquid = g->liquid.as<uint8_t>();
  k_init_beta(s, L, g->beta[0].as<float>(), g->beta[1].as<float>(), g->beta[2].as<float>());
  g->beta_variable = false;
  k_fill_f32(s, g->p.as<float>(), nc, 0.f);
  k_fill_f32(s, g->rhs.as<float>(), nc, 0.f);
  AW_HIP(hipMemsetAsync(g->liquid.p, 0, nc, s));
  g->have_topology = true;
  g->topo_version++;
  g->n_sorted = -1;
  return launch_status("apply_levels");
}

// radius_rule: levels from the particle radii d_radius (by particle id
// through d_rids when given, else by slot) instead of the distance bands.
// meanwhile (distance bands only): GPU work that needs the binning but not
// the levels, enqueued before the host waits for the level map, so that the
// host's compare and layout work runs while the GPU does it.
int bin_and_levels(aw_grid* g, const float* d_pos, int64_t n, bool mark_levels, int* changed,
                   bool radius_rule = false, const float* d_radius = nullptr, const int32_t* d_rids = nullptr,
                   const std::function<int()>& meanwhile = nullptr) {
  hipStream_t s = g->ctx->stream;
  // the counts below are rewritten for these positions: the sorted particles
  // of an earlier binning no longer match them, whatever happens next (a
  // failure below leaves no sorted state the adjoint could pair with them)
  g->n_sorted = -1;
  g->sorted_ids = false;
  AW_TRY(g->key.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)));
  k_fill_u32(s, g->count.as<uint32_t>(), g->nb, 0u);
  GridDev gv = g->view();
  {
    Prof pf(g->ctx, "bin", 16.0 * n);
    k_bin(s, gv, d_pos, n, g->key.as<int32_t>(), g->count.as<uint32_t>());
  }
  if (changed) *changed = 0;
  if (g->ctx->prof) {  // byte accounting of the transfer kernels needs the occupied-block count
    std::vector<uint32_t> cnt(g->nb);
    AW_HIP(hipMemcpyAsync(cnt.data(), g->count.p, g->nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    AW_HIP(hipStreamSynchronize(s));
    g->occupied = (double)std::count_if(cnt.begin(), cnt.end(), [](uint32_t c) { return c > 0; });
    // blocks whose padded lattice is filled (a particle within one block) and their face slots
    g->lattice_blocks = 0.0;
    g->lattice_faces = 0.0;
    for (int b = 0; b < g->nb; ++b) {
      const int x = b % g->D[0], y = (b / g->D[0]) % g->D[1], z = b / (g->D[0] * g->D[1]);
      bool any = false;
      for (int dz = -1; dz <= 1 && !any; ++dz)
        for (int dy = -1; dy <= 1 && !any; ++dy)
          for (int dx = -1; dx <= 1 && !any; ++dx) {
            const int qx = x + dx, qy = y + dy, qz = z + dz;
            if (qx < 0 || qy < 0 || qz < 0 || qx >= g->D[0] || qy >= g->D[1] || qz >= g->D[2]) continue;
            any = cnt[qx + g->D[0] * (qy + g->D[1] * qz)] > 0;
          }
      if (!any || !g->have_topology) continue;
      const int r = BASE >> g->h_level[b];
      g->lattice_blocks += 1.0;
      g->lattice_faces += 3.0 * (r + 1) * r * r;
    }
  }
  if (!mark_levels) return launch_status("bin");
  std::vector<uint8_t> nl(g->nb);
  if (radius_rule) {
    // graded into surf (scratch here), checked, then made the level map
    AW_TRY(g->rmap.ensure((size_t)(2 * g->nb + 1) * sizeof(uint32_t)));
    uint32_t bad = 0;
    {
      Prof pf(g->ctx, "levels", 16.0 * n + 12.0 * g->nb);
      k_radius_levels(s, gv, std::ilogb(g->h), g->desc.num_levels, d_pos, d_radius, d_rids, n,
                      g->rmap.as<uint32_t>(), g->surf.as<uint8_t>(), g->ltmp.as<uint8_t>());
    }
    AW_HIP(hipMemcpyAsync(&bad, g->rmap.as<uint32_t>() + 2 * (size_t)g->nb, sizeof bad, hipMemcpyDeviceToHost, s));
    AW_HIP(hipMemcpyAsync(nl.data(), g->surf.p, g->nb, hipMemcpyDeviceToHost, s));
    AW_HIP(hipStreamSynchronize(s));
    AW_TRY(launch_status("levels"));
    if (bad & 2u) return fail(AW_ERR_INVALID, "particle ids outside [0, %lld)", (long long)n);
    if (bad) return fail(AW_ERR_INVALID, "particle radius not positive and finite");
    AW_HIP(hipMemcpyAsync(g->level.p, g->surf.p, g->nb, hipMemcpyDeviceToDevice, s));
  } else {
    {
      Prof pf(g->ctx, "levels", 8.0 * g->nb);
      k_levels(s, g->D, g->desc.num_levels, g->desc.band, g->count.as<uint32_t>(), g->surf.as<uint8_t>(),
               g->level.as<uint8_t>(), g->ltmp.as<uint8_t>());
    }
    AW_HIP(hipMemcpyAsync(g->h_level_read, g->level.p, g->nb, hipMemcpyDeviceToHost, s));
    AW_HIP(hipEventRecord(g->level_read_done, s));
    if (meanwhile) AW_TRY(meanwhile());
    AW_HIP(hipEventSynchronize(g->level_read_done));
    AW_TRY(launch_status("levels"));
    nl.assign(g->h_level_read, g->h_level_read + g->nb);
  }
  bool ch = !g->have_topology || nl != g->h_level;
  if (ch) {
    g->h_level.swap(nl);
    AW_TRY(apply_levels(g));
  } else {
    // the device map was rewritten with identical values; nothing else changes
  }
  if (changed) *changed = ch ? 1 : 0;
  return AW_OK;
}

int sort_particles(aw_grid* g, const float* d_pos, const float* d_vel, const int32_t* d_ids, int64_t n,
                   const float* d_density = nullptr) {
  hipStream_t s = g->ctx->stream;
  AW_TRY(g->spos.ensure((size_t)std::max<int64_t>(n, 1) * 3 * sizeof(float)));
  AW_TRY(g->svel.ensure((size_t)std::max<int64_t>(n, 1) * 3 * sizeof(float)));
  if (d_ids) AW_TRY(g->sids.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t)));
  if (d_density) AW_TRY(g->sdens.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(float)));
  Prof pf(g->ctx, "sort_particles", 52.0 * n);
  AW_TRY(g->scan_tmp.ensure(scan_scratch_words(g->nb) * sizeof(uint32_t)));
  k_scan_u32(s, g->count.as<uint32_t>(), g->poff.as<uint32_t>(), g->nb, g->scan_tmp.as<uint32_t>());
  AW_HIP(hipMemcpyAsync(g->cursor.p, g->poff.p, (size_t)g->nb * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  AW_TRY(g->vbound.ensure(2 * VB_SLOTS * sizeof(uint32_t)));
  k_scatter(s, d_pos, d_vel, d_ids, g->key.as<int32_t>(), g->cursor.as<uint32_t>(), n, g->spos.as<float>(),
            g->svel.as<float>(), d_ids ? g->sids.as<int32_t>() : nullptr, d_density,
            d_density ? g->sdens.as<float>() : nullptr, g->vbound.as<uint32_t>());
  g->vbound_of = g->svel.p;
  g->vbound_dens_of = d_density ? g->sdens.p : nullptr;
  g->n_sorted = n;
  g->sorted_ids = d_ids != nullptr;
  return launch_status("sort_particles");
}

// sdens (sorted densities, or null): also deposit the density-weighted
// weights and set beta = 1 / rho_face (two-phase); without it beta keeps the
// uniform 1 (reset here if the last step had variable beta)
int p2g_sorted(aw_grid* g, const float* spos, const float* svel, float dt, const float* grav,
               const float* sdens = nullptr) {
  hipStream_t s = g->ctx->stream;
  GridDev gv = g->view();
  const double nf3 = 3.0 * g->nfaces;
  float* mm[3] = {g->macc[0].as<float>(), g->macc[1].as<float>(), g->macc[2].as<float>()};
  // the deposit's fixed-point scale needs the largest |v| (and |density|):
  // recorded by the binning sort for its own copy, measured here otherwise
  AW_TRY(g->vbound.ensure(2 * VB_SLOTS * sizeof(uint32_t)));
  uint32_t* vb = g->vbound.as<uint32_t>();
  if (svel != g->vbound_of) k_vbound(s, svel, 3 * g->n_sorted, vb);
  if (sdens && sdens != g->vbound_dens_of) k_vbound(s, sdens, g->n_sorted, vb + VB_SLOTS);
  if (sdens) {
    for (int a = 0; a < 3; ++a) AW_TRY(g->Dacc[a].ensure((size_t)g->nfaces * sizeof(float)));
    float* D[3] = {g->Dacc[0].as<float>(), g->Dacc[1].as<float>(), g->Dacc[2].as<float>()};
    {
      Prof pf(g->ctx, "p2g_deposit_density", 8.0 * g->n_sorted + 4.0 * SLAB * g->occupied + g->ncells);
      k_p2g_deposit(s, gv, spos, nullptr, g->count.as<uint32_t>(), g->poff.as<uint32_t>(), g->slab.as<float>(),
                    g->liquid.as<uint8_t>(), vb, sdens);
    }
    {
      // the weights go to macc here and are deposited again with the velocities below
      Prof pf(g->ctx, "p2g_gather_density", 4.0 * SLAB * g->occupied + 8.0 * nf3);
      k_p2g_gather(s, gv, g->count.as<uint32_t>(), g->slab.as<float>(), D, mm);
    }
    Prof pf(g->ctx, "beta_fold", 8.0 * nf3 + 4.0 * nf3);
    float* bt[3] = {g->beta[0].as<float>(), g->beta[1].as<float>(), g->beta[2].as<float>()};
    k_beta_fold(s, gv, mm, D, bt);
    g->beta_variable = true;
  } else if (g->beta_variable) {
    LevelDev L;
    for (int a = 0; a < 3; ++a) {
      L.D[a] = g->D[a];
      L.foff[a] = g->foff.as<uint32_t>();
      L.beta[a] = g->beta[a].as<float>();
    }
    L.nb = g->nb;
    L.bs = BASE;
    L.e = g->level.as<uint8_t>();
    L.coff = g->coff.as<uint32_t>();
    L.nbr = g->nbr.as<int32_t>();
    k_init_beta(s, L, g->beta[0].as<float>(), g->beta[1].as<float>(), g->beta[2].as<float>());
    g->beta_variable = false;
  }
  {
    Prof pf(g->ctx, "p2g_deposit", 24.0 * g->n_sorted + 4.0 * SLAB * g->occupied + g->ncells);
    k_p2g_deposit(s, gv, spos, svel, g->count.as<uint32_t>(), g->poff.as<uint32_t>(), g->slab.as<float>(),
                  g->liquid.as<uint8_t>(), vb);
  }
  float* M[3] = {g->Macc[0].as<float>(), g->Macc[1].as<float>(), g->Macc[2].as<float>()};
  float* m[3] = {g->macc[0].as<float>(), g->macc[1].as<float>(), g->macc[2].as<float>()};
  float* uo[3] = {g->uold[0].as<float>(), g->uold[1].as<float>(), g->uold[2].as<float>()};
  float* us[3] = {g->ustar[0].as<float>(), g->ustar[1].as<float>(), g->ustar[2].as<float>()};
  float* ms[3] = {g->mass[0].as<float>(), g->mass[1].as<float>(), g->mass[2].as<float>()};
  {
    Prof pf(g->ctx, "p2g_gather", 4.0 * SLAB * g->occupied + 8.0 * nf3);
    k_p2g_gather(s, gv, g->count.as<uint32_t>(), g->slab.as<float>(), M, m);
  }
  {
    Prof pf(g->ctx, "p2g_fold", 8.0 * nf3 + 12.0 * nf3);
    k_fold(s, gv, M, m, uo, us, ms, dt, grav);
  }
  return launch_status("p2g");
}

int divergence(aw_grid* g) {
  float* us[3] = {g->ustar[0].as<float>(), g->ustar[1].as<float>(), g->ustar[2].as<float>()};
  Prof pf(g->ctx, "divergence", 12.0 * g->nfaces + 5.0 * g->ncells);
  k_divergence(g->ctx->stream, g->view(), g->liquid.as<uint8_t>(), us, g->rhs.as<float>());
  return launch_status("divergence");
}

int project(aw_grid* g) {
  hipStream_t s = g->ctx->stream;
  GridDev gv = g->view();
  float* us[3] = {g->ustar[0].as<float>(), g->ustar[1].as<float>(), g->ustar[2].as<float>()};
  float* bt[3] = {g->beta[0].as<float>(), g->beta[1].as<float>(), g->beta[2].as<float>()};
  float* un[3] = {g->unew[0].as<float>(), g->unew[1].as<float>(), g->unew[2].as<float>()};
  {
    Prof pf(g->ctx, "project", 36.0 * g->nfaces + 4.0 * g->ncells);
    k_project(s, gv, g->p.as<float>(), us, bt, un);
  }
  {
    // u_new, u_old of the flagged blocks' faces, the padded lattice written, flags
    Prof pf(g->ctx, "to_fine", 8.0 * g->lattice_faces + 4.0 * SLAB * g->lattice_blocks + 5.0 * g->nb);
    k_to_fine(s, gv, g->slab.as<float>(), g->count.as<uint32_t>(), g->halo.as<uint8_t>());
  }
  return launch_status("project");
}

// ---------------------------------------------------------------- multigrid
// Work list: per block colour, the blocks of each resolution (8^3 cells
// first), 512 cells per work-group; the launch maps of colour 0, colour 1
// and both describe its runs.
void build_wg(MGLevel& L, std::vector<uint32_t>& blist) {
  blist.clear();
  for (WGMap& m : L.wgm) {
    m.n = 0;
    for (int k = 0; k < WGMAP_RUNS; ++k) {
      m.g0[k] = INT_MAX;
      m.s0[k] = m.cnt[k] = m.lr[k] = 0;
    }
  }
  // one pass: bucket (colour, resolution) of every block, then a stable
  // counting sort into the work list
  int lbs = 0;
  while ((1 << lbs) < L.bs) ++lbs;
  std::vector<uint8_t> key((size_t)L.nb);
  int count[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  {
    int b = 0;
    for (int z = 0; z < L.D[2]; ++z)
      for (int y = 0; y < L.D[1]; ++y)
        for (int x = 0; x < L.D[0]; ++x, ++b) {
          const int lr = lbs - (int)L.h_e[b];
          const int k = ((x + y + z) & 1) * 4 + (3 - lr);
          key[b] = (uint8_t)k;
          ++count[k];
        }
  }
  int first[8];
  for (int k = 0, acc = 0; k < 8; ++k) {
    first[k] = acc;
    acc += count[k];
  }
  blist.assign((size_t)L.nb, 0u);
  {
    int pos[8];
    std::copy(first, first + 8, pos);
    for (int b = 0; b < L.nb; ++b) blist[pos[key[b]]++] = (uint32_t)b;
  }
  int nrun[3] = {0, 0, 0};
  for (int col = 0; col < 2; ++col) {
    for (int lr = 3; lr >= 0; --lr) {
      const int per = 512 >> (3 * lr);
      const int start = first[col * 4 + 3 - lr];
      const int cnt = count[col * 4 + 3 - lr];
      if (cnt == 0) continue;
      for (int mi : {col, 2}) {
        WGMap& m = L.wgm[mi];
        const int k = nrun[mi]++;
        m.g0[k] = m.n;
        m.s0[k] = start;
        m.cnt[k] = cnt;
        m.lr[k] = lr;
        m.n += (cnt + per - 1) / per;
      }
    }
  }
}

int upload_level(hipStream_t s, MGLevel& L, bool own_topology) {
  L.nb = L.D[0] * L.D[1] * L.D[2];
  L.h_coff.assign(L.nb + 1, 0);
  std::vector<uint32_t>& foff = L.h_foff;
  foff.assign(L.nb + 1, 0);
  for (int b = 0; b < L.nb; ++b) {
    const int r = L.bs >> L.h_e[b];
    L.h_coff[b + 1] = L.h_coff[b] + r * r * r;
    foff[b + 1] = foff[b] + (r + 1) * r * r;
  }
  L.ncells = L.h_coff[L.nb];
  L.nfaces = foff[L.nb];
  L.sweep_bytes[0] = L.sweep_bytes[1] = 0.0;
  {
    int b = 0;
    for (int z = 0; z < L.D[2]; ++z)
      for (int y = 0; y < L.D[1]; ++y)
        for (int x = 0; x < L.D[0]; ++x, ++b) {
          const int r = L.bs >> L.h_e[b];
          L.sweep_bytes[(x + y + z) & 1] += 12.0 * r * r * r + 12.0 * (r + 1) * r * r;
        }
  }
  std::vector<uint32_t>& blist = L.h_blist;
  build_wg(L, blist);
  {
    // one descriptor per work-list entry: the block's and its face neighbours' offsets and levels
    std::vector<int32_t>& nb6 = L.h_nbr;
    if (nb6.size() != (size_t)L.nb * 6 || L.nbr_dims[0] != L.D[0] || L.nbr_dims[1] != L.D[1] ||
        L.nbr_dims[2] != L.D[2]) {
      neighbour_table(L.D, nb6);
      for (int a = 0; a < 3; ++a) L.nbr_dims[a] = L.D[a];
    }
    std::vector<uint32_t>& bd = L.h_bd;
    bd.assign(blist.size() * BD_WORDS, 0u);
    for (size_t k = 0; k < blist.size(); ++k) {
      const uint32_t B = blist[k];
      uint32_t* d = &bd[k * BD_WORDS];
      d[0] = B;
      d[1] = L.h_coff[B];
      d[2] = foff[B];
      uint32_t ne[6] = {0, 0, 0, 0, 0, 0};
      for (int j = 0; j < 6; ++j) {
        const int32_t N = nb6[(size_t)B * 6 + j];
        d[3 + j] = (uint32_t)N;
        if (N >= 0) {
          d[9 + j] = L.h_coff[N];
          d[15 + j] = foff[N];
          ne[j] = L.h_e[N];
        }
      }
      d[21] = (uint32_t)L.h_e[B] | (ne[0] << 8) | (ne[1] << 16) | (ne[2] << 24);
      d[22] = ne[3] | (ne[4] << 8) | (ne[5] << 16);
    }
    AW_TRY(L.bdesc.ensure(bd.size() * sizeof(uint32_t)));
    AW_HIP(hipMemcpyAsync(L.bdesc.p, bd.data(), bd.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    L.tables_version++;
    L.live_on = false;
  }
  AW_TRY(L.blist.ensure(blist.size() * sizeof(uint32_t)));
  AW_HIP(hipMemcpyAsync(L.blist.p, blist.data(), blist.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  AW_TRY(L.r.ensure((size_t)L.ncells * sizeof(float)));
  L.d_r = L.r.as<float>();
  if ((size_t)L.ncells * 13 <= TAIL_LDS_BYTES) {
    std::vector<uint32_t>& cb = L.h_cblk;
    cb.assign(L.ncells, 0u);
    for (int b = 0; b < L.nb; ++b)
      for (uint32_t c = L.h_coff[b]; c < L.h_coff[b + 1]; ++c) cb[c] = (uint32_t)b;
    AW_TRY(L.cblk.ensure(cb.size() * sizeof(uint32_t)));
    AW_HIP(hipMemcpyAsync(L.cblk.p, cb.data(), cb.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  }
  L.ush[0] = L.ush[1] = L.ush[2] = -1;
  if (own_topology) {
    bool uni = true, p2 = true;
    for (uint8_t e : L.h_e) uni = uni && e == L.h_e[0];
    int sh[3];
    for (int d = 0; d < 3; ++d) {
      sh[d] = 0;
      while ((1 << sh[d]) < L.D[d]) ++sh[d];
      p2 = p2 && (1 << sh[d]) == L.D[d];
    }
    if (uni && p2 && sh[0] >= 1) {
      int r = L.bs >> L.h_e[0], lr = 0;
      while ((1 << lr) < r) ++lr;
      for (int d = 0; d < 3; ++d) L.ush[d] = sh[d];
      L.ulr = lr;
      L.ue = L.h_e[0];
    }
  }
  if (!own_topology) return AW_OK;
  const std::vector<int32_t>& nbr = L.h_nbr;
  AW_TRY(L.e.ensure(L.nb));
  AW_TRY(L.coff.ensure((L.nb + 1) * sizeof(uint32_t)));
  AW_TRY(L.foff.ensure((L.nb + 1) * sizeof(uint32_t)));
  AW_TRY(L.nbr.ensure(nbr.size() * sizeof(int32_t)));
  AW_TRY(L.liquid.ensure(L.ncells));
  AW_TRY(L.x.ensure((size_t)L.ncells * sizeof(float)));
  AW_TRY(L.b.ensure((size_t)L.ncells * sizeof(float)));
  for (int a = 0; a < 3; ++a) AW_TRY(L.beta[a].ensure((size_t)L.nfaces * sizeof(float)));
  AW_HIP(hipMemcpyAsync(L.e.p, L.h_e.data(), L.nb, hipMemcpyHostToDevice, s));
  AW_HIP(hipMemcpyAsync(L.coff.p, L.h_coff.data(), (L.nb + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  AW_HIP(hipMemcpyAsync(L.foff.p, foff.data(), (L.nb + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  AW_HIP(hipMemcpyAsync(L.nbr.p, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  L.d_e = L.e.as<uint8_t>();
  L.d_coff = L.coff.as<uint32_t>();
  L.d_foff = L.foff.as<uint32_t>();
  L.d_nbr = L.nbr.as<int32_t>();
  L.d_liq = L.liquid.as<uint8_t>();
  for (int a = 0; a < 3; ++a) L.d_beta[a] = L.beta[a].as<float>();
  L.d_x = L.x.as<float>();
  L.d_b = L.b.as<float>();
  k_init_beta(s, L.view(), L.d_beta[0], L.d_beta[1], L.d_beta[2]);
  k_fill_f32(s, L.d_x, L.ncells, 0.f);
  return AW_OK;
}

// Hierarchy (replaces VCycleSolver::init, VCycle.cuh:177-249): level 0 is the
// native grid; block level k has cell edge e_b = min(max(l_b, l_min + k), 3)
// (the finest blocks coarsen first, no block vanishes); once every block is
// 1^3, 2x2x2 aggregation halves the dims while every dim is even and >= 4.
int build_hierarchy(aw_solver* S) {
  aw_grid* g = S->g;
  hipStream_t s = g->ctx->stream;
  std::vector<std::unique_ptr<MGLevel>> lv;
  // The previous build's levels are reused in order: their device buffers
  // (grow-only) are not freed and allocated again on every topology change,
  // and a level whose block resolutions did not change keeps its tables as
  // they are.  Host tables are rewritten only after the previous build's
  // copies from them have completed.
  std::vector<std::unique_ptr<MGLevel>> pool;
  pool.swap(S->lv);
  if (S->upload_event_set) AW_HIP(hipEventSynchronize(S->upload_event));
  size_t reuse = 0;
  auto level = [&]() {
    return reuse < pool.size() ? std::move(pool[reuse++]) : std::make_unique<MGLevel>();
  };
  // (re)builds L for dims D, block size bs and resolutions he unless it holds them already
  auto prepare = [&](MGLevel& L, const int* D, int bs, int aggregated, std::vector<uint8_t>& he) -> int {
    if (L.valid && L.bs == bs && L.aggregated == aggregated && L.D[0] == D[0] && L.D[1] == D[1] &&
        L.D[2] == D[2] && L.h_e == he)
      return AW_OK;
    L.valid = false;
    for (int a = 0; a < 3; ++a) L.D[a] = D[a];
    L.bs = bs;
    L.aggregated = aggregated;
    L.h_e.swap(he);
    AW_TRY(upload_level(s, L, true));
    L.valid = true;
    return AW_OK;
  };
  {
    auto L0 = level();
    L0->valid = false;
    L0->aggregated = 0;
    for (int a = 0; a < 3; ++a) L0->D[a] = g->D[a];
    L0->bs = BASE;
    L0->h_e = g->h_level;
    AW_TRY(upload_level(s, *L0, false));
    L0->d_e = g->level.as<uint8_t>();
    L0->d_coff = g->coff.as<uint32_t>();
    L0->d_foff = g->foff.as<uint32_t>();
    L0->d_nbr = g->nbr.as<int32_t>();
    L0->d_liq = g->liquid.as<uint8_t>();
    for (int a = 0; a < 3; ++a) L0->d_beta[a] = g->beta[a].as<float>();
    L0->d_x = g->p.as<float>();
    L0->d_b = g->rhs.as<float>();
    lv.push_back(std::move(L0));
  }
  int lmin = 3;
  for (uint8_t l : g->h_level) lmin = std::min(lmin, (int)l);
  bool all1 = true;
  for (uint8_t l : g->h_level) all1 = all1 && l == LOGB;
  for (int k = 1; !all1; ++k) {
    auto L = level();
    std::vector<uint8_t> he((size_t)g->nb);
    all1 = true;
    for (int b = 0; b < g->nb; ++b) {
      he[b] = (uint8_t)std::min(std::max((int)g->h_level[b], lmin + k), LOGB);
      all1 = all1 && he[b] == LOGB;
    }
    AW_TRY(prepare(*L, g->D, BASE, 0, he));
    lv.push_back(std::move(L));
  }
  for (;;) {
    const MGLevel& P = *lv.back();
    const int mn = std::min(P.D[0], std::min(P.D[1], P.D[2]));
    if (mn < 4 || (P.D[0] & 1) || (P.D[1] & 1) || (P.D[2] & 1)) break;
    if ((int)lv.size() >= AW_MAX_MG_LEVELS) break;
    auto L = level();
    const int D[3] = {P.D[0] / 2, P.D[1] / 2, P.D[2] / 2};
    std::vector<uint8_t> he((size_t)D[0] * D[1] * D[2], (uint8_t)(P.h_e[0] + 1));
    AW_TRY(prepare(*L, D, P.bs * 2, 1, he));
    lv.push_back(std::move(L));
  }
  S->lv.swap(lv);
  S->built_version = g->topo_version;
  S->graph_epoch++;
  if (!S->upload_event) AW_HIP(hipEventCreateWithFlags(&S->upload_event, hipEventDisableTiming));
  AW_HIP(hipEventRecord(S->upload_event, s));
  S->upload_event_set = true;
  // tail: the longest suffix of levels whose x, b, r fit the LDS budget
  {
    size_t bytes = 0;
    int t = (int)S->lv.size();
    auto fits = [](const MGLevel& L) {
      if (L.ncells <= TAIL_CELLS_MAX) return true;
      bool uni = true;
      for (uint8_t e : L.h_e) uni = uni && e == L.h_e[0];
      return uni && L.ncells <= TAIL_UNIFORM_CELLS_MAX;
    };
    while (t > 0 && (int)S->lv.size() - t < TAIL_MAX_LEVELS && fits(*S->lv[t - 1]) &&
           bytes + (size_t)S->lv[t - 1]->ncells * 13 <= TAIL_LDS_BYTES) {
      bytes += (size_t)S->lv[t - 1]->ncells * 13;
      --t;
    }
    if (t == (int)S->lv.size()) return fail(AW_ERR_UNSUPPORTED, "coarsest level too large for the tail solver");
    S->tail = t;
    S->tail_lds = bytes;
    const uint32_t nc = S->lv.back()->ncells;
    S->direct = nc <= (uint32_t)DIRECT_MAX;
    if (S->direct) AW_TRY(S->cinv.ensure((size_t)nc * nc * sizeof(double)));
  }
  const uint32_t n0 = S->lv[0]->ncells;
  AW_TRY(S->r0.ensure((size_t)n0 * sizeof(float)));
  AW_TRY(S->z.ensure((size_t)n0 * sizeof(float)));
  AW_TRY(S->pv.ensure((size_t)n0 * sizeof(float)));
  AW_TRY(S->q.ensure((size_t)n0 * sizeof(float)));
  return launch_status("build_hierarchy");
}

void coarsen_masks(aw_solver* S) {
  hipStream_t s = S->g->ctx->stream;
  for (size_t k = 0; k + 1 < S->lv.size(); ++k) {
    const MGLevel& F = *S->lv[k];
    MGLevel& C = *S->lv[k + 1];
    k_coarsen_liquid(s, F.view(), C.view(), C.aggregated, C.d_liq, C.wgm[2], C.bl());
  }
}

// coefficient coarsening level by level (runGalerkin, MSBG.cuh:661-676), when
// the grid's beta is variable or the coarse levels still hold variable values
void coarsen_betas(aw_solver* S) {
  aw_grid* g = S->g;
  if (!g->beta_variable && !S->coarse_beta_variable) return;
  hipStream_t s = g->ctx->stream;
  for (size_t k = 0; k + 1 < S->lv.size(); ++k) {
    const MGLevel& F = *S->lv[k];
    MGLevel& C = *S->lv[k + 1];
    Prof pf(g->ctx, "coarsen_beta", 12.0 * F.nfaces + 12.0 * C.nfaces);
    k_coarsen_beta(s, F.view(), C.view(), C.aggregated, C.d_beta[0], C.d_beta[1], C.d_beta[2]);
  }
  S->coarse_beta_variable = g->beta_variable;
}

// Level k's plain smoother runs the arithmetic stencil of a uniform level
// (beta = 1 inside the tank) rather than the work-list kernel.
bool uniform_smoother(const aw_solver* S, int k) {
  return k > 0 && S->lv[k]->ush[0] >= 0 && !S->g->beta_variable;
}

// The smoother's work lists without the all-air blocks (MGLevel::live_on),
// from the liquid masks as they are now: every level below the tail that the
// plain work-list smoother runs.  One flag per block is read back (one
// synchronisation); a list whose entries and tables are unchanged is kept,
// any other is uploaded, its descriptors gathered on the device, and the
// captured PCG iteration (whose launches hold the lists) captured again.
int refresh_live(aw_solver* S) {
  hipStream_t s = S->g->ctx->stream;
  const int nl = S->prm.adaptive ? 0 : S->tail;
  std::vector<size_t> off((size_t)nl + 1, 0);
  size_t tot = 0;
  for (int k = 0; k < (int)S->lv.size(); ++k) S->lv[k]->live_on = false;
  for (int k = 0; k < nl; ++k) {
    off[k] = tot;
    if (!uniform_smoother(S, k)) tot += (size_t)S->lv[k]->nb;
  }
  if (tot == 0) return AW_OK;
  AW_TRY(S->live_flags.ensure(tot));
  uint8_t* flags = S->live_flags.as<uint8_t>();
  for (int k = 0; k < nl; ++k)
    if (!uniform_smoother(S, k)) k_block_live(s, S->lv[k]->view(), flags + off[k]);
  if (S->h_live_cap < tot) {
    if (S->h_live) (void)hipHostFree(S->h_live);
    S->h_live = nullptr;
    S->h_live_cap = 0;
    AW_HIP(hipHostMalloc((void**)&S->h_live, tot, hipHostMallocDefault));
    S->h_live_cap = tot;
  }
  AW_HIP(hipMemcpyAsync(S->h_live, flags, tot, hipMemcpyDeviceToHost, s));
  AW_HIP(hipStreamSynchronize(s));
  AW_TRY(launch_status("liquid blocks"));
  // every earlier copy from a level's h_live_idx has completed (synchronised
  // above).  This host pass runs with the GPU idle, so its inner loop only
  // filters the work list: a run's blocks share one resolution, so its bytes
  // follow from its count; the next level's cells under the live blocks
  // (byte accounting of the fused restriction and the prolongation) are
  // counted only while the per-launch profiler is on.
  const bool prof = S->g->ctx->prof;
  std::vector<uint32_t> idx;
  for (int k = 0; k < nl; ++k) {
    if (uniform_smoother(S, k)) continue;
    MGLevel& L = *S->lv[k];
    const uint8_t* f = S->h_live + off[k];
    idx.clear();
    idx.reserve(L.h_blist.size());
    WGMap m[2]{};
    double by[2] = {0.0, 0.0}, cells = 0.0, ccells = 0.0;
    const MGLevel* C = k + 1 < (int)S->lv.size() ? S->lv[k + 1].get() : nullptr;
    for (int col = 0; col < 2; ++col) {
      for (int j = 0; j < WGMAP_RUNS; ++j) {
        m[col].g0[j] = INT_MAX;
        m[col].s0[j] = m[col].cnt[j] = m[col].lr[j] = 0;
      }
      const WGMap& full = L.wgm[col];
      int nrun = 0;
      for (int j = 0; j < WGMAP_RUNS; ++j) {
        if (full.cnt[j] == 0) continue;
        const int start = (int)idx.size();
        const uint32_t* bl = L.h_blist.data();
        for (int q = full.s0[j]; q < full.s0[j] + full.cnt[j]; ++q)
          if (f[bl[q]]) idx.push_back((uint32_t)q);
        const int cnt = (int)idx.size() - start;
        if (cnt == 0) continue;
        const double r = (double)(1 << full.lr[j]);
        by[col] += cnt * (12.0 * r * r * r + 12.0 * (r + 1.0) * r * r);
        cells += cnt * r * r * r;
        if (prof && C && !C->aggregated)
          for (int q = start; q < start + cnt; ++q) {
            const double rc = (double)(C->bs >> C->h_e[bl[idx[q]]]);
            ccells += rc * rc * rc;
          }
        const int per = 512 >> (3 * full.lr[j]);
        const int t = nrun++;
        m[col].g0[t] = m[col].n;
        m[col].s0[t] = start;
        m[col].cnt[t] = cnt;
        m[col].lr[t] = full.lr[j];
        m[col].n += (cnt + per - 1) / per;
      }
    }
    if (L.live_tables_version != L.tables_version || idx != L.h_live_idx) {
      L.h_live_idx.swap(idx);
      const size_t n = L.h_live_idx.size();
      AW_TRY(L.live_idx.ensure(std::max<size_t>(n, 1) * sizeof(uint32_t)));
      AW_TRY(L.bdesc_live.ensure(std::max<size_t>(n, 1) * BD_WORDS * sizeof(uint32_t)));
      if (n > 0) {
        AW_HIP(hipMemcpyAsync(L.live_idx.p, L.h_live_idx.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        k_gather_bd(s, L.bdesc.as<uint32_t>(), L.live_idx.as<uint32_t>(), (int)n, L.bdesc_live.as<uint32_t>());
      }
      L.live_tables_version = L.tables_version;
      S->graph_epoch++;
    }
    L.wgm_live[0] = m[0];
    L.wgm_live[1] = m[1];
    {
      // both colours in one launch: colour 1's entries follow colour 0's in
      // bdesc_live, so its runs keep their starts and shift their groups
      WGMap& u = L.wgm_live[2];
      u = m[0];
      int t = 0;
      while (t < WGMAP_RUNS && u.cnt[t] > 0) ++t;
      for (int j = 0; j < WGMAP_RUNS && m[1].cnt[j] > 0; ++j, ++t) {
        if (t >= WGMAP_RUNS) return fail(AW_ERR_UNSUPPORTED, "smoother lists: more than WGMAP_RUNS runs");
        u.g0[t] = m[1].g0[j] + m[0].n;
        u.s0[t] = m[1].s0[j];
        u.cnt[t] = m[1].cnt[j];
        u.lr[t] = m[1].lr[j];
      }
      u.n = m[0].n + m[1].n;
    }
    L.live_bytes[0] = by[0];
    L.live_bytes[1] = by[1];
    L.live_cells = cells;
    L.live_ccells = ccells;
    L.live_blocks = (uint32_t)L.h_live_idx.size();
    L.live_on = true;
    // The V-cycle's fused residual and restriction then run over these lists
    // too, and the restriction zeroes the coarse guess only under the
    // blocks holding liquid: the rest of the next level's guess (air cells,
    // which neighbours read as 0 and nothing else writes) is cleared here.
    if (k + 1 < (int)S->lv.size()) k_fill_f32(s, S->lv[k + 1]->d_x, S->lv[k + 1]->ncells, 0.f);
  }
  return launch_status("smoother lists");
}

int ensure_setup(aw_solver* S) {
  if (!S->g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  if (S->built_version != S->g->topo_version) AW_TRY(build_hierarchy(S));
  coarsen_betas(S);
  coarsen_masks(S);
  AW_TRY(refresh_live(S));
  if (S->direct) {
    const MGLevel& C = *S->lv.back();
    Prof pf(S->g->ctx, "coarse_factor", 8.0 * C.ncells * C.ncells);
    k_coarse_factor(S->g->ctx->stream, C.view(), (int)C.ncells, S->cinv.as<double>());
  }
  return launch_status("solver setup");
}

// AdaptiveSmoother::solveLevel (Smoother.cuh:249-291) with device-side flags:
// every block active at the start (initActiveBlocks), `iters` iterations of
// both block colours over the active blocks, the flags swapped in between.
int smooth_level_adaptive(aw_solver* S, int k, float* x, const float* b, int iters, bool reverse) {
  MGLevel& L = *S->lv[k];
  hipStream_t s = S->g->ctx->stream;
  const LevelDev V = L.view();
  for (DevBuf* f : {&L.act[0], &L.act[1], &L.sweeps}) AW_TRY(f->ensure((size_t)L.nb));
  AW_TRY(S->bmax.ensure(sizeof(float)));
  int cur = 0;
  k_fill_u8(s, L.act[cur].as<uint8_t>(), L.nb, 1);
  k_fill_u8(s, L.sweeps.as<uint8_t>(), L.nb, 0);
  k_absmax(s, b, L.d_liq, (int)L.ncells, S->bmax.as<float>());
  for (int it = 0; it < iters; ++it) {
    k_fill_u8(s, L.act[1 - cur].as<uint8_t>(), L.nb, 0);
    for (int ph = 0; ph < 2; ++ph) {
      const int col = reverse ? 1 - ph : ph;
      Prof pf(S->g->ctx, k == 0 ? "smooth_adaptive_l0" : "smooth_adaptive_coarse", L.sweep_bytes[col]);
      k_smooth_adaptive(s, V, x, b, L.wgm[col], S->prm.omega, reverse ? 1 : 0, L.act[cur].as<uint8_t>(),
                        L.act[1 - cur].as<uint8_t>(), L.sweeps.as<uint8_t>(), S->bmax.as<float>(), S->prm.theta,
                        S->prm.max_block_sweeps, S->prm.block_reduction);
    }
    cur = 1 - cur;
  }
  L.act_cur = cur;
  L.act_valid = true;
  return launch_status("adaptive smoothing");
}

void smooth_level(aw_solver* S, int k, float* x, const float* b, int iters, bool reverse) {
  if (S->prm.adaptive) {
    (void)smooth_level_adaptive(S, k, x, b, iters, reverse);
    return;
  }
  const MGLevel& L = *S->lv[k];
  hipStream_t s = S->g->ctx->stream;
  // the arithmetic stencil of uniform levels assumes beta = 1 inside the tank
  const bool uni = uniform_smoother(S, k);
  const bool live = !uni && L.live_on;
  const LevelDev V = live ? L.live_view() : L.view();
  for (int it = 0; it < iters; ++it)
    for (int ph = 0; ph < 2; ++ph) {
      const int col = reverse ? 1 - ph : ph;
      Prof pf(S->g->ctx, k == 0 ? "smooth_l0" : (uni ? "smooth_uniform" : "smooth_coarse"),
              live ? L.live_bytes[col] : L.sweep_bytes[col]);
      if (uni)
        k_smooth_uniform(s, L.ush, L.ulr, L.ue, L.d_liq, x, b, col, S->prm.omega, S->prm.local_sweeps,
                         reverse ? 1 : 0);
      else
        k_smooth(s, V, x, b, live ? L.wgm_live[col] : L.wgm[col], L.bl(), S->prm.omega, S->prm.local_sweeps,
                 reverse ? 1 : 0);
    }
}

// V-cycle of levels t0.. (all in the tail) in one work-group
void tail_vcycle(aw_solver* S, int t0, float* x, const float* b) {
  TailArgs A;
  std::memset(&A, 0, sizeof A);
  int off = 0;
  A.n = (int)S->lv.size() - t0;
  size_t cells = 0;
  for (int i = 0; i < A.n; ++i) {
    const MGLevel& L = *S->lv[t0 + i];
    TailLevel& T = A.lv[i];
    T.L = L.view();
    T.cblk = L.cblk.as<uint32_t>();
    T.ncells = (int)L.ncells;
    T.agg = i > 0 ? L.aggregated : 0;
    T.off = off;
    off += 3 * (int)L.ncells;
    cells += L.ncells;
    // uniform coarse levels use the arithmetic stencil (their beta is 1 inside
    // the tank); level 0 keeps the general one
    T.ulr = -1;
    T.ue = 0;
    if (t0 + i > 0 && !S->g->beta_variable) {
      bool uni = true;
      for (uint8_t e : L.h_e) uni = uni && e == L.h_e[0];
      if (uni) {
        int r = L.bs >> L.h_e[0], lr = 0;
        while ((1 << lr) < r) ++lr;
        T.ulr = lr;
        T.ue = L.h_e[0];
      }
    }
    // a lattice of one-cell blocks with power-of-two dims (cell index = block
    // index): neighbours, colours and parents by shifts
    for (int d = 0; d < 3; ++d) T.sh[d] = -1;
    if (T.ulr == 0) {
      int sh[3];
      bool p2 = true;
      for (int d = 0; d < 3; ++d) {
        sh[d] = 0;
        while ((1 << sh[d]) < L.D[d]) ++sh[d];
        p2 = p2 && (1 << sh[d]) == L.D[d];
      }
      if (p2 && sh[0] >= 1)
        for (int d = 0; d < 3; ++d) T.sh[d] = sh[d];
    }
  }
  A.fl = off;
  int loff = 0;
  for (int i = 0; i < A.n; ++i) {
    A.lv[i].loff = loff;
    loff += A.lv[i].ncells;
  }
  A.cinv = S->direct ? S->cinv.as<double>() : nullptr;
  A.omega = S->prm.omega;
  A.scale = S->prm.prolong_scale;
  A.pre = S->prm.pre_sweeps;
  A.post = S->prm.post_sweeps;
  A.coarse_sweeps = S->prm.coarse_sweeps;
  A.local_sweeps = S->prm.local_sweeps;
  A.x0 = x;
  A.b0 = b;
  Prof pf(S->g->ctx, "vcycle_tail", 12.0 * cells);
  k_tail(S->g->ctx->stream, A, (size_t)off * sizeof(float) + (size_t)loff);
}

void residual_level(aw_solver* S, int k, const float* x, const float* b, float* r, bool fuse_restrict = false) {
  const MGLevel& L = *S->lv[k];
  Restr rs;
  std::memset(&rs, 0, sizeof rs);
  double bytes = L.sweep_bytes[0] + L.sweep_bytes[1];
  if (fuse_restrict) {
    const MGLevel& C = *S->lv[k + 1];
    rs.C = C.view();
    rs.bc = C.d_b;
    rs.xc = C.d_x;
    rs.on = 1;
    bytes += 8.0 * C.ncells - 4.0 * L.ncells;  // r is not stored; b_c and x_c are
  }
  // fused with the restriction, only the blocks holding liquid matter: the
  // fine residual is not stored, and the coarse right-hand side of air cells
  // is not read (their guess stays 0)
  const bool live = fuse_restrict && L.live_on;
  if (live) bytes = L.live_bytes[0] + L.live_bytes[1] + 8.0 * L.live_ccells - 4.0 * L.live_cells;
  Prof pf(S->g->ctx,
          k == 0 ? (fuse_restrict ? "residual_restrict_l0" : "residual_l0")
                 : (fuse_restrict ? "residual_restrict_coarse" : "residual_coarse"),
          bytes);
  k_residual(S->g->ctx->stream, live ? L.live_view() : L.view(), x, b, fuse_restrict ? nullptr : r,
             live ? L.wgm_live[2] : L.wgm[2], L.bl(), rs);
}

void restrict_level(aw_solver* S, int k, const float* r, float* bc, float* xc = nullptr) {
  const MGLevel& F = *S->lv[k];
  const MGLevel& C = *S->lv[k + 1];
  Prof pf(S->g->ctx, "restrict", 4.0 * F.ncells + 4.0 * C.ncells);
  k_restrict(S->g->ctx->stream, F.view(), C.view(), C.aggregated, r, bc, xc, C.wgm[2], C.bl());
}

void prolong_level(aw_solver* S, int k, float* x, const float* xc) {
  const MGLevel& F = *S->lv[k];
  const MGLevel& C = *S->lv[k + 1];
  // air cells take no correction: the blocks holding liquid are enough
  const bool live = F.live_on;
  Prof pf(S->g->ctx, "prolong", live ? 8.0 * F.live_cells + 4.0 * F.live_ccells : 8.0 * F.ncells + 4.0 * C.ncells);
  k_prolong(S->g->ctx->stream, live ? F.live_view() : F.view(), C.view(), C.aggregated, x, xc,
            S->prm.prolong_scale, live ? F.wgm_live[2] : F.wgm[2], F.bl());
}

// recursive V-cycle (VCycleSolver::runCycle, VCycle.cuh:251-304): the
// restriction is fused into the residual pass (across an aggregation the
// restriction kernel zeroes the coarse guess itself), the prolongation is a
// pass of its own before post-smoothing (fusing it into the first smoothing
// phase lengthens that phase's dependent loads more than the pass costs), and
// the trailing small levels run in one work-group.
void vcycle(aw_solver* S, int k, float* x, const float* b) {
  if (k >= S->tail) {
    tail_vcycle(S, k, x, b);
    return;
  }
  MGLevel& L = *S->lv[k];
  MGLevel& C = *S->lv[k + 1];
  smooth_level(S, k, x, b, S->prm.pre_sweeps, false);
  if (!C.aggregated) {
    residual_level(S, k, x, b, L.d_r, true);
  } else {
    residual_level(S, k, x, b, L.d_r);
    restrict_level(S, k, L.d_r, C.d_b, C.d_x);
  }
  vcycle(S, k + 1, C.d_x, C.d_b);
  prolong_level(S, k, x, C.d_x);
  smooth_level(S, k, x, b, S->prm.post_sweeps, true);
}

// slot group of scalar idx
double* scp(aw_solver* S, int idx) { return S->sc.as<double>() + idx * SC_SLOTS; }

// copies the slotted scalars to the host and adds each one's slots into h_sc[idx]
int read_scalars(aw_solver* S) {
  hipStream_t s = S->g->ctx->stream;
  AW_HIP(hipMemcpyAsync(S->h_slots, S->sc.p, SC_N * SC_SLOTS * sizeof(double), hipMemcpyDeviceToHost, s));
  AW_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < SC_N; ++i) {
    double t = 0.0;
    for (int k = 0; k < SC_SLOTS; ++k) t += S->h_slots[i * SC_SLOTS + k];
    S->h_sc[i] = t;
  }
  return launch_status("solver");
}

// One PCG iteration between scalar readbacks: z = V-cycle(r), beta from
// (r.z)/(r.z)_old, p = z + beta p, q = A p with p.q, x += alpha p, r -= alpha q.
void pcg_body(aw_solver* S, float* x, int first) {
  aw_grid* g = S->g;
  hipStream_t s = g->ctx->stream;
  MGLevel& L0 = *S->lv[0];
  const int n = (int)L0.ncells;
  const uint8_t* liq = g->liquid.as<uint8_t>();
  float* r = S->r0.as<float>();
  float* z = S->z.as<float>();
  float* pv = S->pv.as<float>();
  float* q = S->q.as<float>();
  double* sc = S->sc.as<double>();
  k_scalar_shift(s, sc);
  k_fill_f32(s, z, n, 0.f);
  vcycle(S, 0, z, r);
  {
    Prof pf(g->ctx, "pcg_dot_p", 21.0 * n);
    k_dot(s, r, z, liq, n, scp(S, SC_RZ));
    k_pcg_p(s, z, pv, n, sc, first);
  }
  {
    // q = A p is read on liquid cells only (p.q here, r -= alpha q below):
    // the blocks holding liquid are enough
    const bool live = L0.live_on;
    Prof pf(g->ctx, "apply_l0",
            live ? L0.live_bytes[0] + L0.live_bytes[1] - 4.0 * L0.live_cells
                 : L0.sweep_bytes[0] + L0.sweep_bytes[1] - 4.0 * n);
    k_apply_dot(s, live ? L0.live_view() : L0.view(), pv, q, live ? L0.wgm_live[2] : L0.wgm[2], L0.bl(),
                scp(S, SC_PQ));
  }
  {
    Prof pf(g->ctx, "pcg_update", 28.0 * n);
    k_pcg_xr(s, x, r, pv, q, liq, n, sc);
  }
}

// pcg_body through its captured graph (captured on first use after a
// hierarchy build); eager when the per-launch profiler is on or capture is
// unavailable on the context's stream.
int run_pcg_body(aw_solver* S, float* x, int first) {
  hipStream_t s = S->g->ctx->stream;
  if (S->g->ctx->prof || !S->use_graph || !S->capture_ok) {
    pcg_body(S, x, first);
    return launch_status("pcg iteration");
  }
  aw_solver::Captured& c = S->pcg_body[first ? 1 : 0];
  if (!c.exec || c.epoch != S->graph_epoch || c.x != x) {
    if (c.exec) {
      (void)hipGraphExecDestroy(c.exec);
      c.exec = nullptr;
    }
    AW_TRY(launch_status("before capture"));
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      S->capture_ok = false;
      pcg_body(S, x, first);
      return launch_status("pcg iteration");
    }
    pcg_body(S, x, first);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (ec != hipSuccess || !graph) {
      (void)hipGetLastError();
      if (graph) (void)hipGraphDestroy(graph);
      return fail(AW_ERR_DEVICE, "capture of the PCG iteration failed: %s", hipGetErrorString(ec));
    }
    const hipError_t ei = hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      (void)hipGetLastError();
      c.exec = nullptr;
      return fail(AW_ERR_DEVICE, "instantiation of the PCG iteration graph failed: %s", hipGetErrorString(ei));
    }
    c.epoch = S->graph_epoch;
    c.x = x;
  }
  AW_HIP(hipGraphLaunch(c.exec, s));
  return AW_OK;
}

// FP32 conjugate gradients (or stationary V-cycles) on level 0 for A x = b.
int solve_fp32(aw_solver* S, float* x, const float* b, double tol, int max_iters, aw_solve_stats* st) {
  aw_grid* g = S->g;
  hipStream_t s = g->ctx->stream;
  MGLevel& L0 = *S->lv[0];
  const int n = (int)L0.ncells;
  const uint8_t* liq = g->liquid.as<uint8_t>();
  float* r = S->r0.as<float>();
  float* z = S->z.as<float>();
  double* sc = S->sc.as<double>();
  k_mask(s, x, liq, n);
  k_fill_f64(s, sc, SC_N * SC_SLOTS, 0.0);
  residual_level(S, 0, x, b, r);
  k_dot(s, b, b, liq, n, scp(S, SC_BB));
  k_dot(s, r, r, liq, n, scp(S, SC_RR));
  AW_TRY(read_scalars(S));
  const double bn = std::sqrt(S->h_sc[SC_BB]);
  double rn = std::sqrt(S->h_sc[SC_RR]);
  int it = 0;
  if (st) {
    st->b_norm = bn;
    st->outer_iterations = 1;
    st->recurrence_rel_residual = 0.0;
  }
  // finiteness first: a NaN norm fails every comparison, the zero test too
  if (!std::isfinite(bn) || !std::isfinite(rn)) return fail(AW_ERR_NAN, "non-finite right-hand side");
  if (bn == 0.0) {
    k_fill_f32(s, x, n, 0.f);
    if (st) {
      st->iterations = 0;
      st->rel_residual = 0.0;
    }
    return launch_status("solve");
  }
  if (S->prm.method == 1) {
    while (rn > tol * bn && it < max_iters) {
      k_fill_f32(s, z, n, 0.f);
      vcycle(S, 0, z, r);
      k_add(s, x, z, liq, n);
      residual_level(S, 0, x, b, r);
      k_scalar_shift(s, sc);
      k_dot(s, r, r, liq, n, scp(S, SC_RR));
      AW_TRY(read_scalars(S));
      rn = std::sqrt(S->h_sc[SC_RR]);
      ++it;
      if (!std::isfinite(rn)) return fail(AW_ERR_NAN, "V-cycle iteration diverged to non-finite values");
    }
    if (st) st->recurrence_rel_residual = rn / bn;
  } else {
    // Conjugate gradients preconditioned by one V-cycle.  The recursively
    // updated FP32 residual drifts from b - A x; when it reports convergence
    // the true residual is recomputed and, if it is still above the
    // tolerance, the iteration restarts once from it.  The FP32 pressure
    // itself bounds the attainable true residual (rounding p to FP32 leaves
    // a residual of order eps32 |A| |p|): when the recurrence converges a
    // second time and the true residual is still above the tolerance, the
    // solve stops there and reports AW_ERR_NOT_CONVERGED with both
    // residuals (tolerances below that floor need precision 1).
    int restarts = 0;
    int first = 1;
    double rec = rn;
    while (it < max_iters) {
      AW_TRY(run_pcg_body(S, x, first));
      first = 0;
      AW_TRY(read_scalars(S));
      rn = std::sqrt(S->h_sc[SC_RR]);
      rec = rn;
      ++it;
      if (!std::isfinite(rn)) return fail(AW_ERR_NAN, "PCG produced non-finite residual");
      if (rn > tol * bn) continue;
      residual_level(S, 0, x, b, r);
      k_scalar_shift(s, sc);
      k_dot(s, r, r, liq, n, scp(S, SC_RR));
      AW_TRY(read_scalars(S));
      rn = std::sqrt(S->h_sc[SC_RR]);
      if (rn <= tol * bn || restarts == 1) break;
      ++restarts;
      first = 1;
    }
    if (st) st->recurrence_rel_residual = rec / bn;
    if (rec <= tol * bn && rn > tol * bn) {
      if (st) {
        st->iterations = it;
        st->rel_residual = rn / bn;
      }
      AW_TRY(launch_status("solve"));
      return fail(AW_ERR_NOT_CONVERGED,
                  "FP32 floor: true relative residual %.3e above %.1e after %d iterations (recurrence %.3e)", rn / bn,
                  tol, it, rec / bn);
    }
  }
  if (st) {
    st->iterations = it;
    st->rel_residual = rn / bn;
  }
  AW_TRY(launch_status("solve"));
  return rn <= tol * bn ? AW_OK : fail(AW_ERR_NOT_CONVERGED, "relative residual %.3e after %d iterations", rn / bn, it);
}

// FP32 V-cycle PCG inside FP64 iterative refinement: the fine-level
// residual and the accumulated solution are FP64, every correction is an FP32
// solve to relative accuracy 1e-4, the coarsest level is solved in FP64; the
// FP32 pressure used by the projection is the rounded FP64 solution.
int solve_mixed(aw_solver* S, double tol, int max_iters, aw_solve_stats* st) {
  aw_grid* g = S->g;
  hipStream_t s = g->ctx->stream;
  MGLevel& L0 = *S->lv[0];
  const int n = (int)L0.ncells;
  const uint8_t* liq = g->liquid.as<uint8_t>();
  double* sc = S->sc.as<double>();
  AW_TRY(S->x64.ensure((size_t)n * sizeof(double)));
  AW_TRY(S->e32.ensure((size_t)n * sizeof(float)));
  AW_TRY(S->r32.ensure((size_t)n * sizeof(float)));
  double* x64 = S->x64.as<double>();
  float* e = S->e32.as<float>();
  float* r = S->r32.as<float>();
  k_to64(s, x64, g->p.as<float>(), liq, n);
  k_fill_f64(s, sc, SC_N * SC_SLOTS, 0.0);
  k_dot(s, g->rhs.as<float>(), g->rhs.as<float>(), liq, n, scp(S, SC_BB));
  AW_TRY(read_scalars(S));
  const double bn = std::sqrt(S->h_sc[SC_BB]);
  if (st) {
    st->b_norm = bn;
    st->iterations = 0;
    st->outer_iterations = 0;
    st->rel_residual = 0.0;
    st->recurrence_rel_residual = 0.0;
  }
  // finiteness first: a NaN norm fails every comparison, the zero test too
  if (!std::isfinite(bn)) return fail(AW_ERR_NAN, "non-finite right-hand side");
  if (bn == 0.0) {
    k_fill_f32(s, g->p.as<float>(), n, 0.f);
    k_fill_f64(s, x64, n, 0.0);
    return launch_status("solve");
  }
  int inner_total = 0, outer = 0;
  double rn = 0.0;
  for (;;) {
    k_scalar_shift(s, sc);
    {
      Prof pf(g->ctx, "residual64_l0", 8.0 * n + 4.0 * n + 4.0 * n + 12.0 * L0.nfaces);
      k_residual64(s, L0.view(), x64, g->rhs.as<float>(), r, L0.wgm[2], L0.bl(), scp(S, SC_RR));
    }
    AW_TRY(read_scalars(S));
    rn = std::sqrt(S->h_sc[SC_RR]);
    if (!std::isfinite(rn)) return fail(AW_ERR_NAN, "iterative refinement diverged");
    if (rn <= tol * bn || inner_total >= max_iters || outer >= 30) break;
    k_fill_f32(s, e, n, 0.f);
    aw_solve_stats ist{};
    int sti = solve_fp32(S, e, r, 1e-4, max_iters - inner_total, &ist);
    if (sti != AW_OK && sti != AW_ERR_NOT_CONVERGED) return sti;
    inner_total += std::max(ist.iterations, 1);
    k_axpy64(s, x64, e, liq, n);
    ++outer;
  }
  k_to32(s, g->p.as<float>(), x64, n);
  if (st) {
    st->iterations = inner_total;
    st->outer_iterations = outer;
    st->rel_residual = rn / bn;
    st->recurrence_rel_residual = rn / bn;  // the FP64 residual is evaluated directly every refinement
  }
  AW_TRY(launch_status("solve"));
  return rn <= tol * bn ? AW_OK
                        : fail(AW_ERR_NOT_CONVERGED, "relative residual %.3e after %d refinements", rn / bn, outer);
}

int solve(aw_solver* S, double tol, int max_iters, aw_solve_stats* st) {
  if (S->prm.precision == 1) return solve_mixed(S, tol, max_iters, st);
  return solve_fp32(S, S->g->p.as<float>(), S->g->rhs.as<float>(), tol, max_iters, st);
}

// ---------------------------------------------------------------- adjoint
// Reverse-mode derivative of aw_step with the topology, the liquid mask and
// the mass thresholds held at their forward values.  It runs on the state the
// last forward step left in the grid: the sorted particles (spos, svel, sids),
// the fine lattice of Pi(u_new), Pi(u_new - u_old), u_old and the folded mass.
// Particle adjoints are indexed by particle id.
int ensure_adjoint(aw_grid* g) {
  const size_t nf = (size_t)g->nfaces * sizeof(float);
  AW_TRY(g->fbar.ensure((size_t)g->nb * ADJ * sizeof(float)));
  AW_TRY(g->fmbar.ensure((size_t)g->nb * ADJ * sizeof(float)));
  AW_TRY(g->pbar.ensure((size_t)g->ncells * sizeof(float)));
  AW_TRY(g->rhsbar.ensure((size_t)g->ncells * sizeof(float)));
  for (int a = 0; a < 3; ++a)
    for (DevBuf* f : {g->ubn, g->ubo, g->ubs, g->Mb, g->mb}) AW_TRY(f[a].ensure(nf));
  return AW_OK;
}

struct Faces {
  float* p[3];
  explicit Faces(DevBuf* f) {
    for (int a = 0; a < 3; ++a) p[a] = f[a].as<float>();
  }
};

// Particle adjoints are indexed by id: every id must lie in [0, n).
int validate_ids(aw_grid* g, const int32_t* ids, int64_t n) {
  if (n <= 0) return AW_OK;
  hipStream_t s = g->ctx->stream;
  AW_TRY(g->idrange.ensure(2 * sizeof(int)));
  const int init[2] = {0x7FFFFFFF, -0x7FFFFFFF - 1};
  int mm[2];
  AW_HIP(hipMemcpyAsync(g->idrange.p, init, sizeof init, hipMemcpyHostToDevice, s));
  k_ids_range(s, ids, n, g->idrange.as<int>());
  AW_HIP(hipMemcpyAsync(mm, g->idrange.p, sizeof mm, hipMemcpyDeviceToHost, s));
  AW_HIP(hipStreamSynchronize(s));
  if (mm[0] < 0 || mm[1] >= n) return fail(AW_ERR_INVALID, "particle ids span [%d, %d], outside [0, %lld)", mm[0], mm[1], (long long)n);
  return AW_OK;
}

int check_adjoint_state(aw_grid* g) {
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  if (g->beta_variable)
    return fail(AW_ERR_UNSUPPORTED, "the adjoint holds beta = 1: the last step had variable density");
  if (g->n_sorted < 0 || !g->sorted_ids)
    return fail(AW_ERR_INVALID, "the adjoint needs a forward step binned with particle ids");
  return ensure_adjoint(g);
}

// dL/d(x', v') -> direct particle partials; ubar_new = Pi^T (Fnew-bar + Fdel-bar), ubar_old = -Pi^T Fdel-bar
int g2p_adjoint(aw_grid* g, float alpha, float dt, float* xbar, float* vbar) {
  AW_TRY(check_adjoint_state(g));
  hipStream_t s = g->ctx->stream;
  const GridDev gv = g->view();
  const int64_t n = g->n_sorted;
  AW_TRY(g->adjrec.ensure((size_t)std::max<int64_t>(n, 1) * adj_record_bytes()));
  AW_TRY(g->g2p_list.ensure((size_t)(n + 1) * sizeof(uint32_t)));
  const size_t flag_words = ((size_t)g->nb + 3) / 4;
  AW_TRY(g->fbar_touched.ensure(flag_words * sizeof(uint32_t)));
  if (g->fbar_zeroed_for != g->fbar.p) {
    // a new allocation: zero it all once
    k_fill_f32(s, g->fbar.as<float>(), (size_t)g->nb * ADJ, 0.f);
    k_fill_u32(s, g->fbar_touched.as<uint32_t>(), flag_words, 0u);
    g->fbar_zeroed_for = g->fbar.p;
  } else {
    k_fbar_clear(s, g->nb, g->fbar.as<float>(), g->fbar_touched.as<uint8_t>());
  }
  {
    // particle position and id 16 B + adjoints read and written 48 B + record; forward fine lattice per block
    Prof pf(g->ctx, "g2p_adjoint_stages", (64.0 + adj_record_bytes()) * n + 2.0 * 4.0 * 3 * FINE * g->nb);
    k_g2p_adjoint_stages(s, gv, g->slab.as<float>(), g->spos.as<float>(), g->sids.as<int32_t>(), n, alpha, dt, xbar,
                         vbar, g->adjrec.p, g->g2p_list.as<uint32_t>(), g->count.as<uint32_t>(),
                         g->poff.as<uint32_t>());
  }
  {
    // position 12 B + record per particle; the fine-lattice adjoint per block
    Prof pf(g->ctx, "g2p_adjoint_scatter", (12.0 + adj_record_bytes()) * n + 4.0 * ADJ * g->nb);
    k_g2p_adjoint_scatter(s, gv, g->spos.as<float>(), g->count.as<uint32_t>(), g->poff.as<uint32_t>(), g->adjrec.p,
                          g->fbar.as<float>(), g->fbar_touched.as<uint8_t>());
  }
  {
    Prof pf(g->ctx, "pi_transpose", 4.0 * ADJ * g->nb + 8.0 * 3 * g->nfaces);
    Faces un(g->ubn), uo(g->ubo);
    k_pi_transpose(s, gv, g->fbar.as<float>(), g->fbar_touched.as<uint8_t>(), un.p, uo.p);
  }
  return launch_status("g2p_advect_adjoint");
}

// ubar_new -> pbar -> rhsbar = A^{-1} pbar (same operator and hierarchy) -> ubar_star
int project_adjoint(aw_solver* S, double tol, int max_iters, aw_solve_stats* st) {
  aw_grid* g = S->g;
  AW_TRY(ensure_adjoint(g));
  hipStream_t s = g->ctx->stream;
  const GridDev gv = g->view();
  Faces bt(g->beta), un(g->ubn), us(g->ubs);
  k_fill_f32(s, g->pbar.as<float>(), g->ncells, 0.f);
  {
    Prof pf(g->ctx, "project_adjoint", 12.0 * 3 * g->nfaces + 8.0 * g->ncells);
    k_project_adjoint(s, gv, bt.p, un.p, us.p, g->pbar.as<float>());
  }
  k_masked_copy(s, g->rhs.as<float>(), g->pbar.as<float>(), g->liquid.as<uint8_t>(), (int)g->ncells);
  k_fill_f32(s, g->p.as<float>(), g->ncells, 0.f);
  AW_TRY(ensure_setup(S));
  int rc = solve(S, tol, max_iters, st);
  if (rc != AW_OK && rc != AW_ERR_NOT_CONVERGED) return rc;
  std::string keep = g_err;
  AW_HIP(hipMemcpyAsync(g->rhsbar.p, g->p.p, (size_t)g->ncells * sizeof(float), hipMemcpyDeviceToDevice, s));
  {
    Prof pf(g->ctx, "divergence_adjoint", 8.0 * 3 * g->nfaces + 5.0 * g->ncells);
    k_divergence_adjoint(s, gv, g->liquid.as<uint8_t>(), g->rhsbar.as<float>(), us.p);
  }
  AW_TRY(launch_status("project_adjoint"));
  if (rc != AW_OK) g_err = keep;
  return rc;
}

// (ubar_old + ubar_star) -> Mbar, mbar -> Pi -> particle partials (accumulated)
int p2g_adjoint(aw_grid* g, float* xbar, float* vbar) {
  AW_TRY(check_adjoint_state(g));
  hipStream_t s = g->ctx->stream;
  const GridDev gv = g->view();
  Faces uo(g->uold), ms(g->mass), ubo(g->ubo), ubs(g->ubs), Mb(g->Mb), mb(g->mb);
  {
    Prof pf(g->ctx, "fold_adjoint", 24.0 * 3 * g->nfaces);
    k_fold_adjoint(s, gv, uo.p, ms.p, ubo.p, ubs.p, Mb.p, mb.p);
  }
  {
    Prof pf(g->ctx, "pi_forward", 8.0 * g->lattice_faces + 4.0 * ADJ * g->lattice_blocks);
    k_pi_forward2(s, gv, Mb.p, mb.p, g->fmbar.as<float>());
  }
  {
    Prof pf(g->ctx, "p2g_adjoint", 76.0 * g->n_sorted + 4.0 * ADJ * g->nb);
    k_p2g_adjoint(s, gv, g->fmbar.as<float>(), g->spos.as<float>(), g->svel.as<float>(), g->sids.as<int32_t>(),
                  g->n_sorted, xbar, vbar, g->count.as<uint32_t>(), g->poff.as<uint32_t>());
  }
  return launch_status("p2g_adjoint");
}

int step_adjoint(aw_grid* g, aw_solver* S, const aw_step_params* prm, float* xbar, float* vbar,
                 aw_solve_stats* st) {
  AW_TRY(g2p_adjoint(g, prm->flip_alpha, prm->dt, xbar, vbar));
  int rc = project_adjoint(S, prm->rel_tol, prm->max_iters, st);
  if (rc != AW_OK && rc != AW_ERR_NOT_CONVERGED) return rc;
  std::string keep = g_err;
  AW_TRY(p2g_adjoint(g, xbar, vbar));
  if (rc != AW_OK) g_err = keep;
  return rc;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// Checkpointed rollout: (x, v, ids) every `every` steps in the forward pass;
// the reverse pass re-runs each segment from its checkpoint and keeps what
// the step adjoint needs per step (topology, liquid mask, u_old, mass, u_new
// and the sorted particles), then walks the segment backwards.
struct aw_tape {
  struct Step {
    std::vector<uint8_t> level;
    uint32_t ncells = 0, nfaces = 0;
    DevBuf liquid, uold[3], mass[3], unew[3], spos, svel, sids, count, poff;
  };
  struct Ckpt {
    DevBuf pos, vel, ids;
  };
  aw_grid* g = nullptr;
  aw_solver* S = nullptr;
  int64_t n = 0;
  int steps = 0, every = 1;
  bool keep_masks = false;
  bool have_forward = false;
  aw_step_params prm{};
  std::vector<std::unique_ptr<Ckpt>> ck;
  std::vector<std::unique_ptr<Step>> rec;
  DevBuf wpos, wvel, wids;
  DevBuf radius;  // rebuild_topology 2: the tape's own copy of the radii (the recompute re-reads them)
  std::vector<std::vector<uint8_t>> mask_level, mask_liquid;
  // replay check: every forward step's native levels and liquid-mask
  // fingerprint; scratch of the fingerprint and of the state comparison
  std::vector<std::vector<uint8_t>> fwd_level;
  std::vector<unsigned long long> fwd_hash;
  DevBuf hash, diff, inv;
};

namespace {

int copy_d2d(hipStream_t s, DevBuf& dst, const void* src, size_t bytes) {
  AW_TRY(dst.ensure(bytes));
  if (bytes) AW_HIP(hipMemcpyAsync(dst.p, src, bytes, hipMemcpyDeviceToDevice, s));
  return AW_OK;
}

int tape_record(aw_tape* t, aw_tape::Step& r) {
  aw_grid* g = t->g;
  hipStream_t s = g->ctx->stream;
  const size_t nf = (size_t)g->nfaces * sizeof(float);
  r.level = g->h_level;
  r.ncells = g->ncells;
  r.nfaces = g->nfaces;
  AW_TRY(copy_d2d(s, r.liquid, g->liquid.p, g->ncells));
  for (int a = 0; a < 3; ++a) {
    AW_TRY(copy_d2d(s, r.uold[a], g->uold[a].p, nf));
    AW_TRY(copy_d2d(s, r.mass[a], g->mass[a].p, nf));
    AW_TRY(copy_d2d(s, r.unew[a], g->unew[a].p, nf));
  }
  // The sorted particle arrays move into the record instead of being
  // copied: the grid takes the record's previous buffers, whose contents the
  // next step's sort rewrites entry by entry (and which are grown here if the
  // record's were smaller).  The bound the sort took of the velocities is
  // of a buffer the grid no longer holds.
  r.spos.swap(g->spos);
  r.svel.swap(g->svel);
  r.sids.swap(g->sids);
  const size_t np = (size_t)std::max<int64_t>(t->n, 1);
  AW_TRY(g->spos.ensure(np * 3 * sizeof(float)));
  AW_TRY(g->svel.ensure(np * 3 * sizeof(float)));
  AW_TRY(g->sids.ensure(np * sizeof(int32_t)));
  g->vbound_of = g->vbound_dens_of = nullptr;
  AW_TRY(copy_d2d(s, r.count, g->count.p, (size_t)g->nb * sizeof(uint32_t)));
  AW_TRY(copy_d2d(s, r.poff, g->poff.p, (size_t)(g->nb + 1) * sizeof(uint32_t)));
  return AW_OK;
}

// fingerprint of the grid's liquid mask after a step (synchronises)
int liquid_fingerprint(aw_tape* t, unsigned long long* out) {
  aw_grid* g = t->g;
  hipStream_t s = g->ctx->stream;
  AW_TRY(t->hash.ensure(sizeof(unsigned long long)));
  k_mask_hash(s, g->liquid.as<uint8_t>(), g->ncells, t->hash.as<unsigned long long>());
  AW_HIP(hipMemcpyAsync(out, t->hash.p, sizeof *out, hipMemcpyDeviceToHost, s));
  AW_HIP(hipStreamSynchronize(s));
  return launch_status("liquid fingerprint");
}

int tape_restore(aw_tape* t, aw_tape::Step& r) {
  aw_grid* g = t->g;
  hipStream_t s = g->ctx->stream;
  if (g->h_level != r.level) {
    g->h_level = r.level;
    AW_HIP(hipMemcpyAsync(g->level.p, g->h_level.data(), g->nb, hipMemcpyHostToDevice, s));
    AW_TRY(apply_levels(g));
  }
  const size_t nf = (size_t)g->nfaces * sizeof(float);
  AW_HIP(hipMemcpyAsync(g->liquid.p, r.liquid.p, g->ncells, hipMemcpyDeviceToDevice, s));
  for (int a = 0; a < 3; ++a) {
    AW_HIP(hipMemcpyAsync(g->uold[a].p, r.uold[a].p, nf, hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(g->mass[a].p, r.mass[a].p, nf, hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(g->unew[a].p, r.unew[a].p, nf, hipMemcpyDeviceToDevice, s));
  }
  // the record's sorted particle arrays move back into the grid (the record
  // keeps the grid's buffers, rewritten when it next records)
  g->spos.swap(r.spos);
  g->svel.swap(r.svel);
  g->sids.swap(r.sids);
  g->vbound_of = g->vbound_dens_of = nullptr;   // the sorted arrays no longer hold what the bound was taken of
  AW_HIP(hipMemcpyAsync(g->count.p, r.count.p, (size_t)g->nb * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  AW_HIP(hipMemcpyAsync(g->poff.p, r.poff.p, (size_t)(g->nb + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  g->n_sorted = t->n;
  g->sorted_ids = true;
  Faces un(g->unew), uo(g->uold);
  {
    Prof pf(g->ctx, "to_fine", 8.0 * g->lattice_faces + 4.0 * SLAB * g->lattice_blocks + 5.0 * g->nb);
    k_to_fine(s, g->view(), g->slab.as<float>(), g->count.as<uint32_t>(), g->halo.as<uint8_t>());
  }
  return launch_status("tape restore");
}

}  // namespace

namespace {

void ctx_free(aw_ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  delete c;
}
void ctx_release(aw_ctx* c) {
  if (--c->refs == 0 && c->closed) ctx_free(c);
}
void grid_free(aw_grid* g) {
  aw_ctx* c = g->ctx;
  (void)hipStreamSynchronize(c->stream);
  if (g->h_liq_count) (void)hipHostFree(g->h_liq_count);
  if (g->h_level_read) (void)hipHostFree(g->h_level_read);
  if (g->level_read_done) (void)hipEventDestroy(g->level_read_done);
  delete g;
  ctx_release(c);
}
void grid_release(aw_grid* g) {
  if (--g->refs == 0 && g->closed) grid_free(g);
}
void solver_free(aw_solver* S) {
  aw_grid* g = S->g;
  (void)hipStreamSynchronize(g->ctx->stream);
  for (auto& c : S->pcg_body)
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
  if (S->upload_event) (void)hipEventDestroy(S->upload_event);
  if (S->h_slots) (void)hipHostFree(S->h_slots);
  if (S->h_live) (void)hipHostFree(S->h_live);
  delete S;
  grid_release(g);
}
void solver_release(aw_solver* S) {
  if (--S->refs == 0 && S->closed) solver_free(S);
}

bool valid_params(const aw_solver_params* p) {
  if (!p) return false;
  const bool base = (p->method == 0 || p->method == 1) && p->pre_sweeps >= 0 && p->post_sweeps >= 0 &&
                    p->coarse_sweeps >= 0 && p->local_sweeps >= 1 && p->omega > 0.f && p->omega < 2.f &&
                    (p->precision == 0 || p->precision == 1) && (p->adaptive == 0 || p->adaptive == 1);
  if (!base || !p->adaptive) return base;
  // adaptive smoothing is nonlinear: the stationary method only
  return p->method == 1 && p->theta >= 0.f && p->max_block_sweeps >= 1 && p->max_block_sweeps <= 255 &&
         p->block_reduction > 0.f && p->block_reduction < 1.f;
}

}  // namespace

extern "C" {

const char* aw_status_str(int status) {
  switch (status) {
    case AW_OK: return "ok";
    case AW_ERR_INVALID: return "invalid argument";
    case AW_ERR_OOM: return "out of device memory";
    case AW_ERR_DEVICE: return "device error";
    case AW_ERR_NOT_CONVERGED: return "not converged";
    case AW_ERR_NAN: return "non-finite values";
    case AW_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char* aw_last_error(void) { return g_err.c_str(); }

int aw_version(void) { return 1; }

int aw_ctx_create(int device, void* stream, aw_ctx** out) {
  if (!out) return fail(AW_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(AW_ERR_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= ndev) return fail(AW_ERR_INVALID, "device %d out of range [0,%d)", device, ndev);
  AW_HIP(hipSetDevice(device));
  uint32_t* scr = nullptr;
  AW_TRY(scratch_address(device, &scr));
  auto* c = new aw_ctx();
  c->device = device;
  c->scratch_dev = scr;
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return fail(AW_ERR_DEVICE, "hipStreamCreate failed");
    }
    c->own_stream = true;
  }
  *out = c;
  return AW_OK;
}

int aw_ctx_destroy(aw_ctx* c) {
  if (!c) return fail(AW_ERR_INVALID, "ctx is NULL");
  c->closed = true;
  if (c->refs == 0) ctx_free(c);
  return AW_OK;
}

int aw_profile_enable(aw_ctx* c, int on) {
  if (!c) return fail(AW_ERR_INVALID, "ctx is NULL");
  AW_HIP(hipStreamSynchronize(c->stream));
  c->recs.clear();
  c->used = 0;
  c->prof = on != 0;
  c->prof_runs = on == 2;
  return AW_OK;
}

int aw_profile_report(aw_ctx* c, char* buf, uint64_t len) {
  if (!c || !buf || len == 0) return fail(AW_ERR_INVALID, "bad arguments");
  AW_HIP(hipStreamSynchronize(c->stream));
  struct Agg {
    std::string name;
    long n = 0;
    double ms = 0.0, bytes = 0.0;
  };
  std::vector<Agg> agg;
  for (const ProfRec& r : c->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    Agg* a = nullptr;
    for (Agg& x : agg)
      if (x.name == r.name) a = &x;
    if (!a) {
      agg.push_back(Agg{r.name});
      a = &agg.back();
    }
    a->n += r.launches;
    a->ms += ms;
    a->bytes += r.bytes;
  }
  std::string out = "{";
  for (size_t i = 0; i < agg.size(); ++i) {
    char tmp[256];
    snprintf(tmp, sizeof tmp, "%s\"%s\": {\"launches\": %ld, \"ms\": %.6f, \"bytes\": %.0f}", i ? ", " : "",
             agg[i].name.c_str(), agg[i].n, agg[i].ms, agg[i].bytes);
    out += tmp;
  }
  out += "}";
  if (out.size() + 1 > len) return fail(AW_ERR_INVALID, "report buffer too small (%zu bytes needed)", out.size() + 1);
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return AW_OK;
}

int aw_ctx_synchronize(aw_ctx* c) {
  if (!c) return fail(AW_ERR_INVALID, "ctx is NULL");
  AW_HIP(hipStreamSynchronize(c->stream));
  return launch_status("synchronize");
}

int aw_memory_stats(uint64_t* current_bytes, uint64_t* peak_bytes, int reset_peak) {
  if (current_bytes) *current_bytes = g_mem_cur.load();
  if (peak_bytes) *peak_bytes = g_mem_peak.load();
  if (reset_peak) g_mem_peak.store(g_mem_cur.load());
  return AW_OK;
}

int aw_set_memory_limit(uint64_t bytes) {
  g_mem_limit.store(bytes);
  return AW_OK;
}

int aw_malloc(aw_ctx* c, uint64_t bytes, void** out) {
  if (!c || !out) return fail(AW_ERR_INVALID, "NULL argument");
  *out = nullptr;
  AW_TRY(over_limit(bytes ? bytes : 1));
  AW_HIP(hipSetDevice(c->device));
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return fail(AW_ERR_OOM, "hipMalloc of %llu bytes failed", (unsigned long long)bytes);
  }
  c->refs++;
  {
    std::lock_guard<std::mutex> lk(g_user_mu);
    g_user_alloc[*out] = bytes ? bytes : 1;
  }
  mem_add(bytes ? bytes : 1);
  return AW_OK;
}

int aw_free(aw_ctx* c, void* p) {
  if (!c) return fail(AW_ERR_INVALID, "ctx is NULL");
  if (p) {
    (void)hipStreamSynchronize(c->stream);
    {
      std::lock_guard<std::mutex> lk(g_user_mu);
      auto it = g_user_alloc.find(p);
      if (it != g_user_alloc.end()) {
        mem_sub(it->second);
        g_user_alloc.erase(it);
      }
    }
    const hipError_t e = hipFree(p);
    ctx_release(c);
    if (e != hipSuccess) return fail(AW_ERR_DEVICE, "hipFree: %s", hipGetErrorString(e));
  }
  return AW_OK;
}

int aw_copy(aw_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return fail(AW_ERR_INVALID, "NULL argument");
  if (!bytes) return AW_OK;
  AW_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c->stream));
  AW_HIP(hipStreamSynchronize(c->stream));
  return launch_status("copy");
}

int aw_fill_f32(aw_ctx* c, float* d, uint64_t n, float value) {
  if (!c || (n && !d)) return fail(AW_ERR_INVALID, "NULL argument");
  if (n) k_fill_f32(c->stream, d, n, value);
  return launch_status("fill");
}

int aw_grid_create(aw_ctx* ctx, const aw_grid_desc* d, aw_grid** out) {
  if (!ctx || !d || !out) return fail(AW_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (d->base_res != BASE) return fail(AW_ERR_UNSUPPORTED, "base_res %d (only 8 is supported)", d->base_res);
  if (d->num_levels < 1 || d->num_levels > AW_MAX_LEVELS)
    return fail(AW_ERR_INVALID, "num_levels %d outside [1,%d]", d->num_levels, AW_MAX_LEVELS);
  for (int a = 0; a < 3; ++a) {
    if (d->dims[a] < 1 || d->dims[a] > 1024) return fail(AW_ERR_INVALID, "dims[%d] = %d", a, d->dims[a]);
    if (d->domain_min[a] != 0.f) return fail(AW_ERR_UNSUPPORTED, "domain_min must be the origin");
  }
  int ex = 0;
  const float fr = std::frexp(d->h, &ex);
  if (!(d->h > 0.f) || fr != 0.5f) return fail(AW_ERR_INVALID, "h must be a positive power of two");
  for (int l = 0; l + 1 < d->num_levels; ++l)
    if (d->band[l] < 0 || d->band[l] > 254) return fail(AW_ERR_INVALID, "band[%d] = %d outside 0..254", l, d->band[l]);
  const int64_t nb = (int64_t)d->dims[0] * d->dims[1] * d->dims[2];
  if (nb > (1 << 24)) return fail(AW_ERR_INVALID, "too many blocks");
  AW_HIP(hipSetDevice(ctx->device));
  auto g = std::make_unique<aw_grid>();
  g->ctx = ctx;
  g->desc = *d;
  g->nb = (int)nb;
  for (int a = 0; a < 3; ++a) {
    g->D[a] = d->dims[a];
    g->N[a] = d->dims[a] * BASE;
  }
  g->h = d->h;
  g->inv_h = 1.0f / d->h;
  hipStream_t s = ctx->stream;
  AW_TRY(g->level.ensure(nb));
  AW_TRY(g->surf.ensure(nb));
  AW_TRY(g->ltmp.ensure(nb));
  AW_TRY(g->count.ensure(nb * sizeof(uint32_t)));
  AW_TRY(g->poff.ensure((nb + 1) * sizeof(uint32_t)));
  AW_TRY(g->cursor.ensure(nb * sizeof(uint32_t)));
  AW_TRY(g->liq_count.ensure(sizeof(uint32_t)));
  AW_HIP(hipHostMalloc((void**)&g->h_liq_count, sizeof(uint32_t), hipHostMallocDefault));
  *g->h_liq_count = 0;
  AW_HIP(hipHostMalloc((void**)&g->h_level_read, nb, hipHostMallocDefault));
  AW_HIP(hipEventCreateWithFlags(&g->level_read_done, hipEventDisableTiming));
  std::vector<int32_t> nbr;
  neighbour_table(g->D, nbr);
  AW_TRY(g->nbr.ensure(nbr.size() * sizeof(int32_t)));
  AW_HIP(hipMemcpyAsync(g->nbr.p, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  AW_HIP(hipStreamSynchronize(s));
  ctx->refs++;
  *out = g.release();
  return AW_OK;
}

int aw_grid_destroy(aw_grid* g) {
  if (!g) return fail(AW_ERR_INVALID, "grid is NULL");
  g->closed = true;
  if (g->refs == 0) grid_free(g);
  return AW_OK;
}

int aw_grid_build_topology(aw_grid* g, const float* d_pos, int64_t n, int* changed) {
  if (!g || (n > 0 && !d_pos) || n < 0) return fail(AW_ERR_INVALID, "bad arguments");
  if (n > 0x7FFFFFFF) return fail(AW_ERR_INVALID, "too many particles");
  AW_TRY(bin_and_levels(g, d_pos, n, true, changed));
  AW_HIP(hipStreamSynchronize(g->ctx->stream));
  return launch_status("build_topology");
}

int aw_grid_build_topology_radius(aw_grid* g, const float* d_pos, const float* d_radius, int64_t n, int* changed) {
  if (!g || (n > 0 && (!d_pos || !d_radius)) || n < 0) return fail(AW_ERR_INVALID, "bad arguments");
  if (n > 0x7FFFFFFF) return fail(AW_ERR_INVALID, "too many particles");
  AW_TRY(bin_and_levels(g, d_pos, n, true, changed, true, d_radius));
  AW_HIP(hipStreamSynchronize(g->ctx->stream));
  return launch_status("build_topology_radius");
}

int aw_grid_beta_changed(aw_grid* g) {
  if (!g) return fail(AW_ERR_INVALID, "grid is NULL");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  g->beta_variable = true;
  return AW_OK;
}

int aw_grid_set_levels(aw_grid* g, const uint8_t* h_levels) {
  if (!g || !h_levels) return fail(AW_ERR_INVALID, "NULL argument");
  for (int b = 0; b < g->nb; ++b)
    if (h_levels[b] >= g->desc.num_levels) return fail(AW_ERR_INVALID, "level %d at block %d", h_levels[b], b);
  // 2:1 grading over the 26-neighbourhood is a precondition of the stencil
  for (int b = 0; b < g->nb; ++b) {
    int c[3] = {b % g->D[0], (b / g->D[0]) % g->D[1], b / (g->D[0] * g->D[1])};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
          if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= g->D[0] || q[1] >= g->D[1] || q[2] >= g->D[2]) continue;
          int nb = q[0] + g->D[0] * (q[1] + g->D[1] * q[2]);
          if (std::abs((int)h_levels[b] - (int)h_levels[nb]) > 1)
            return fail(AW_ERR_INVALID, "levels violate 2:1 grading at block %d", b);
        }
  }
  hipStream_t s = g->ctx->stream;
  g->h_level.assign(h_levels, h_levels + g->nb);
  AW_HIP(hipMemcpyAsync(g->level.p, h_levels, g->nb, hipMemcpyHostToDevice, s));
  AW_TRY(apply_levels(g));
  AW_HIP(hipStreamSynchronize(s));
  return launch_status("set_levels");
}

int aw_grid_layout(aw_grid* g, aw_layout* o) {
  if (!g || !o) return fail(AW_ERR_INVALID, "NULL argument");
  std::memset(o, 0, sizeof *o);
  o->num_blocks = (uint32_t)g->nb;
  o->total_cells = g->ncells;
  for (int a = 0; a < 3; ++a) {
    o->total_faces[a] = g->nfaces;
    o->d_face_off[a] = g->foff.as<uint32_t>();
    o->d_beta[a] = g->beta[a].as<float>();
    o->d_u_old[a] = g->uold[a].as<float>();
    o->d_u_star[a] = g->ustar[a].as<float>();
    o->d_u_new[a] = g->unew[a].as<float>();
    o->d_mass[a] = g->mass[a].as<float>();
  }
  o->d_level = g->level.as<uint8_t>();
  o->d_cell_off = g->coff.as<uint32_t>();
  o->d_particle_count = g->count.as<uint32_t>();
  o->d_particle_off = g->poff.as<uint32_t>();
  o->d_pressure = g->p.as<float>();
  o->d_rhs = g->rhs.as<float>();
  o->d_liquid = g->liquid.as<uint8_t>();
  o->d_fine = g->slab.as<float>();
  return AW_OK;
}

int aw_bin_particles(aw_grid* g, float* d_pos, float* d_vel, int32_t* d_ids, int64_t n) {
  if (!g || n < 0 || (n > 0 && (!d_pos || !d_vel))) return fail(AW_ERR_INVALID, "bad arguments");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  hipStream_t s = g->ctx->stream;
  AW_TRY(bin_and_levels(g, d_pos, n, false, nullptr));
  AW_TRY(sort_particles(g, d_pos, d_vel, d_ids, n));
  if (n > 0) {
    AW_HIP(hipMemcpyAsync(d_pos, g->spos.p, n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(d_vel, g->svel.p, n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (d_ids) AW_HIP(hipMemcpyAsync(d_ids, g->sids.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  }
  return launch_status("bin_particles");
}

int aw_p2g(aw_grid* g, const float* d_pos, const float* d_vel, int64_t n, float dt, const float gravity[3]) {
  if (!g || n < 0 || (n > 0 && (!d_pos || !d_vel))) return fail(AW_ERR_INVALID, "bad arguments");
  if (!g->have_topology || g->n_sorted != n) return fail(AW_ERR_INVALID, "particles not binned for this topology");
  const float zero[3] = {0.f, 0.f, 0.f};
  return p2g_sorted(g, d_pos, d_vel, dt, gravity ? gravity : zero);
}

int aw_divergence(aw_grid* g) {
  if (!g || !g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  return divergence(g);
}

int aw_project_velocity(aw_grid* g) {
  if (!g || !g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  return project(g);
}

int aw_g2p_advect(aw_grid* g, float* d_pos, float* d_vel, int64_t n, float alpha, float dt) {
  if (!g || n < 0 || (n > 0 && (!d_pos || !d_vel))) return fail(AW_ERR_INVALID, "bad arguments");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  AW_TRY(g->g2p_list.ensure((size_t)(n + 1) * sizeof(uint32_t)));
  k_g2p_advect(g->ctx->stream, g->view(), g->slab.as<float>(), d_pos, d_vel, n, alpha, dt, d_pos, d_vel,
               g->g2p_list.as<uint32_t>());
  return launch_status("g2p_advect");
}

int aw_p2g_particles(aw_grid* g, const float* d_pos, const float* d_vel, int64_t n, float dt,
                     const float gravity[3]) {
  if (!g || n < 0 || (n > 0 && (!d_pos || !d_vel))) return fail(AW_ERR_INVALID, "bad arguments");
  if (n > 0x7FFFFFFF) return fail(AW_ERR_INVALID, "too many particles");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  // binned and sorted into the library's own buffers: the caller's arrays keep their order
  AW_TRY(bin_and_levels(g, d_pos, n, false, nullptr));
  AW_TRY(sort_particles(g, d_pos, d_vel, nullptr, n));
  const float zero[3] = {0.f, 0.f, 0.f};
  return p2g_sorted(g, g->spos.as<float>(), g->svel.as<float>(), dt, gravity ? gravity : zero);
}

int aw_g2p(aw_grid* g, const float* d_pos, float* d_vel, int64_t n, float alpha) {
  if (!g || n < 0 || (n > 0 && (!d_pos || !d_vel))) return fail(AW_ERR_INVALID, "bad arguments");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  AW_TRY(g->g2p_list.ensure((size_t)(n + 1) * sizeof(uint32_t)));
  Prof pf(g->ctx, "g2p", 36.0 * n);
  k_g2p_advect(g->ctx->stream, g->view(), g->slab.as<float>(), d_pos, d_vel, n, alpha, 0.f, nullptr, d_vel,
               g->g2p_list.as<uint32_t>(), 1);
  return launch_status("g2p");
}

int aw_advect(aw_grid* g, float* d_pos, int64_t n, float dt) {
  if (!g || n < 0 || (n > 0 && !d_pos)) return fail(AW_ERR_INVALID, "bad arguments");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  AW_TRY(g->g2p_list.ensure((size_t)(n + 1) * sizeof(uint32_t)));
  Prof pf(g->ctx, "advect", 24.0 * n);
  k_g2p_advect(g->ctx->stream, g->view(), g->slab.as<float>(), d_pos, nullptr, n, 0.f, dt, d_pos, nullptr,
               g->g2p_list.as<uint32_t>(), 2);
  return launch_status("advect");
}

int aw_solver_create(aw_grid* g, const aw_solver_params* p, aw_solver** out) {
  if (!g || !out) return fail(AW_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (!valid_params(p)) return fail(AW_ERR_INVALID, "invalid solver parameters");
  auto S = std::make_unique<aw_solver>();
  S->g = g;
  S->prm = *p;
  AW_TRY(S->sc.ensure(SC_N * SC_SLOTS * sizeof(double)));
  AW_HIP(hipHostMalloc((void**)&S->h_slots, SC_N * SC_SLOTS * sizeof(double), hipHostMallocDefault));
  S->h_sc = S->h_sc_sum;
  g->refs++;
  *out = S.release();
  return AW_OK;
}

int aw_solver_destroy(aw_solver* S) {
  if (!S) return fail(AW_ERR_INVALID, "solver is NULL");
  S->closed = true;
  if (S->refs == 0) solver_free(S);
  return AW_OK;
}

int aw_solver_set_graph(aw_solver* S, int on) {
  if (!S) return fail(AW_ERR_INVALID, "solver is NULL");
  S->use_graph = on != 0;
  return AW_OK;
}

int aw_solver_setup(aw_solver* S) {
  if (!S) return fail(AW_ERR_INVALID, "solver is NULL");
  return ensure_setup(S);
}

int aw_solver_level_info(aw_solver* S, int k, aw_level_info* o) {
  if (!S || !o) return fail(AW_ERR_INVALID, "NULL argument");
  if (k < 0 || k >= (int)S->lv.size()) return fail(AW_ERR_INVALID, "level %d not built", k);
  const MGLevel& L = *S->lv[k];
  std::memset(o, 0, sizeof *o);
  for (int a = 0; a < 3; ++a) {
    o->dims[a] = L.D[a];
    o->total_faces[a] = L.nfaces;
  }
  o->num_blocks = (uint32_t)L.nb;
  o->total_cells = L.ncells;
  o->aggregated = L.aggregated;
  o->d_e = L.d_e;
  o->d_cell_off = L.d_coff;
  o->d_x = L.d_x;
  o->d_b = L.d_b;
  o->d_r = L.d_r;
  o->d_liquid = L.d_liq;
  o->d_x64 = (k == 0 && S->prm.precision == 1 && S->x64.p) ? S->x64.as<double>() : nullptr;
  o->d_block_sweeps = L.act_valid ? L.sweeps.as<uint8_t>() : nullptr;
  o->d_active = L.act_valid ? L.act[L.act_cur].as<uint8_t>() : nullptr;
  o->smoothed_blocks = L.live_on ? L.live_blocks : (uint32_t)L.nb;
  return AW_OK;
}

static int check_level(aw_solver* S, int k, bool need_next) {
  if (!S) return fail(AW_ERR_INVALID, "solver is NULL");
  if (S->built_version != S->g->topo_version) return fail(AW_ERR_INVALID, "solver not set up for this topology");
  if (k < 0 || k >= (int)S->lv.size() || (need_next && k + 1 >= (int)S->lv.size()))
    return fail(AW_ERR_INVALID, "level %d out of range", k);
  return AW_OK;
}

int aw_solver_smooth(aw_solver* S, int k, int iters, int reverse) {
  AW_TRY(check_level(S, k, false));
  if (iters < 0) return fail(AW_ERR_INVALID, "iterations < 0");
  MGLevel& L = *S->lv[k];
  smooth_level(S, k, L.d_x, L.d_b, iters, reverse != 0);
  return launch_status("smooth");
}

int aw_solver_residual(aw_solver* S, int k) {
  AW_TRY(check_level(S, k, false));
  MGLevel& L = *S->lv[k];
  residual_level(S, k, L.d_x, L.d_b, L.d_r, false);
  return launch_status("residual");
}

int aw_solver_apply(aw_solver* S, int k, const float* d_x, float* d_y) {
  AW_TRY(check_level(S, k, false));
  if (!d_x || !d_y) return fail(AW_ERR_INVALID, "NULL vector");
  MGLevel& L = *S->lv[k];
  k_apply_dot(S->g->ctx->stream, L.view(), d_x, d_y, L.wgm[2], L.bl(), nullptr);
  return launch_status("apply");
}

int aw_solver_restrict(aw_solver* S, int k) {
  AW_TRY(check_level(S, k, true));
  restrict_level(S, k, S->lv[k]->d_r, S->lv[k + 1]->d_b);
  return launch_status("restrict");
}

int aw_solver_prolong(aw_solver* S, int k) {
  AW_TRY(check_level(S, k, true));
  prolong_level(S, k, S->lv[k]->d_x, S->lv[k + 1]->d_x);
  return launch_status("prolong");
}

int aw_solver_vcycle(aw_solver* S, int k) {
  AW_TRY(check_level(S, k, false));
  MGLevel& L = *S->lv[k];
  vcycle(S, k, L.d_x, L.d_b);
  return launch_status("vcycle");
}

int aw_solver_solve(aw_solver* S, double tol, int max_iters, aw_solve_stats* st) {
  if (!S) return fail(AW_ERR_INVALID, "solver is NULL");
  if (!(tol > 0.0) || max_iters < 0) return fail(AW_ERR_INVALID, "bad tolerance or iteration cap");
  AW_TRY(ensure_setup(S));
  return solve(S, tol, max_iters, st);
}

int aw_step(aw_grid* g, aw_solver* S, float* d_pos, float* d_vel, int32_t* d_ids, int64_t n,
            const aw_step_params* prm, aw_step_stats* stats) {
  if (!g || !S || !prm || S->g != g || n < 0 || (n > 0 && (!d_pos || !d_vel)))
    return fail(AW_ERR_INVALID, "bad arguments");
  if (n > 0x7FFFFFFF) return fail(AW_ERR_INVALID, "too many particles");
  if (prm->rebuild_topology < 0 || prm->rebuild_topology > 2) return fail(AW_ERR_INVALID, "rebuild_topology 0, 1 or 2");
  if (prm->rebuild_topology == 2 && (!prm->d_radius || (n > 0 && !d_ids)))
    return fail(AW_ERR_INVALID, "rebuild_topology 2 needs d_radius and d_ids");
  if (prm->d_density && n > 0 && !d_ids) return fail(AW_ERR_INVALID, "d_density needs d_ids");
  // (rebuild_topology 2 reads the radii through the ids: the marking kernel
  // range-checks each id and flags one outside [0, n) as an invalid input)
  hipStream_t s = g->ctx->stream;
  int changed = 0;
  // warm start from the last step's pressure, if the layout it was solved in
  // is still the grid's (nothing else rebuilt the topology since)
  const bool warm = prm->warm_start && g->p_valid && g->have_topology && g->p_version == g->topo_version;
  if (warm) {
    AW_TRY(g->p_prev.ensure((size_t)g->ncells * sizeof(float)));
    AW_TRY(g->level_prev.ensure(g->nb));
    AW_TRY(g->coff_prev.ensure((size_t)(g->nb + 1) * sizeof(uint32_t)));
    AW_HIP(hipMemcpyAsync(g->p_prev.p, g->p.p, (size_t)g->ncells * sizeof(float), hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(g->level_prev.p, g->level.p, g->nb, hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(g->coff_prev.p, g->coff.p, (size_t)(g->nb + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                          s));
  }
  g->p_valid = false;
  const float* dens = n > 0 ? prm->d_density : nullptr;
  bool sorted = false;
  if (prm->rebuild_topology == 2) {
    AW_TRY(bin_and_levels(g, d_pos, n, true, &changed, true, prm->d_radius, d_ids));
  } else if (prm->rebuild_topology || !g->have_topology) {
    // the sort needs the binning only: it runs while the host handles the
    // level map (a topology change resets the sorted state, restored below)
    AW_TRY(bin_and_levels(g, d_pos, n, true, &changed, false, nullptr, nullptr, [&]() {
      return sort_particles(g, d_pos, d_vel, d_ids, n, dens);
    }));
    sorted = true;
  } else {
    AW_TRY(bin_and_levels(g, d_pos, n, false, nullptr));
  }
  if (sorted) {
    g->n_sorted = n;
    g->sorted_ids = d_ids != nullptr;
  } else {
    AW_TRY(sort_particles(g, d_pos, d_vel, d_ids, n, dens));
  }
  AW_TRY(p2g_sorted(g, g->spos.as<float>(), g->svel.as<float>(), prm->dt, prm->gravity,
                    dens ? g->sdens.as<float>() : nullptr));
  AW_TRY(divergence(g));
  AW_TRY(ensure_setup(S));
  if (warm) {
    Prof pf(g->ctx, "pressure_remap", 8.0 * g->ncells + 9.0 * g->nb);
    k_pressure_remap(s, g->view(), g->level_prev.as<uint8_t>(), g->coff_prev.as<uint32_t>(), g->p_prev.as<float>(),
                     g->p.as<float>());
  } else {
    k_fill_f32(s, g->p.as<float>(), g->ncells, 0.f);
  }
  aw_solve_stats sst{};
  int st = solve(S, prm->rel_tol, prm->max_iters, &sst);
  if (st != AW_OK && st != AW_ERR_NOT_CONVERGED) return st;
  g->p_valid = true;
  g->p_version = g->topo_version;
  std::string keep = g_err;
  AW_TRY(project(g));
  {
    Prof pf(g->ctx, "g2p_advect", 48.0 * n + 2.0 * 4.0 * 3 * FINE * g->occupied);
    AW_TRY(g->g2p_list.ensure((size_t)(n + 1) * sizeof(uint32_t)));
    k_g2p_advect_sorted(s, g->view(), g->slab.as<float>(), g->spos.as<float>(), g->svel.as<float>(),
                        g->count.as<uint32_t>(), g->poff.as<uint32_t>(), n, prm->flip_alpha, prm->dt, d_pos, d_vel,
                        g->g2p_list.as<uint32_t>());
  }
  if (d_ids && n > 0) AW_HIP(hipMemcpyAsync(d_ids, g->sids.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  if (stats) {
    k_count_nonzero(s, g->liquid.as<uint8_t>(), g->ncells, g->liq_count.as<uint32_t>());
    AW_HIP(hipMemcpyAsync(g->h_liq_count, g->liq_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  }
  AW_HIP(hipStreamSynchronize(s));
  AW_TRY(launch_status("step"));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->topology_changed = changed;
    stats->solve = sst;
    for (uint8_t l : g->h_level) stats->num_blocks_by_level[l]++;
    stats->total_cells = g->ncells;
    stats->liquid_cells = *g->h_liq_count;
  }
  if (st != AW_OK) g_err = keep;
  return st;
}

int aw_grid_adjoint_layout(aw_grid* g, aw_adjoint_layout* o) {
  if (!g || !o) return fail(AW_ERR_INVALID, "NULL argument");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "grid has no topology");
  AW_TRY(ensure_adjoint(g));
  std::memset(o, 0, sizeof *o);
  for (int a = 0; a < 3; ++a) {
    o->d_ubar_new[a] = g->ubn[a].as<float>();
    o->d_ubar_old[a] = g->ubo[a].as<float>();
    o->d_ubar_star[a] = g->ubs[a].as<float>();
    o->d_Mbar[a] = g->Mb[a].as<float>();
    o->d_mbar[a] = g->mb[a].as<float>();
  }
  o->d_pbar = g->pbar.as<float>();
  o->d_rhsbar = g->rhsbar.as<float>();
  o->d_fine_bar = g->fbar.as<float>();
  o->d_fine_mbar = g->fmbar.as<float>();
  o->d_sorted_ids = g->sids.as<int32_t>();
  o->num_sorted = g->n_sorted;
  return AW_OK;
}

int aw_g2p_advect_adjoint(aw_grid* g, float alpha, float dt, float* d_xbar, float* d_vbar) {
  if (!g || !d_xbar || !d_vbar) return fail(AW_ERR_INVALID, "NULL argument");
  AW_TRY(check_adjoint_state(g));
  AW_TRY(validate_ids(g, g->sids.as<int32_t>(), g->n_sorted));
  return g2p_adjoint(g, alpha, dt, d_xbar, d_vbar);
}

int aw_project_adjoint(aw_solver* S, double rel_tol, int max_iters, aw_solve_stats* stats) {
  if (!S || !(rel_tol > 0.0) || max_iters < 1) return fail(AW_ERR_INVALID, "bad arguments");
  AW_TRY(check_adjoint_state(S->g));
  aw_solve_stats st{};
  int rc = project_adjoint(S, rel_tol, max_iters, &st);
  AW_HIP(hipStreamSynchronize(S->g->ctx->stream));
  if (stats) *stats = st;
  return rc;
}

int aw_p2g_adjoint(aw_grid* g, float* d_xbar, float* d_vbar) {
  if (!g || !d_xbar || !d_vbar) return fail(AW_ERR_INVALID, "NULL argument");
  AW_TRY(check_adjoint_state(g));
  AW_TRY(validate_ids(g, g->sids.as<int32_t>(), g->n_sorted));
  return p2g_adjoint(g, d_xbar, d_vbar);
}

int aw_step_adjoint(aw_grid* g, aw_solver* S, const aw_step_params* prm, float* d_xbar, float* d_vbar,
                    aw_solve_stats* stats) {
  if (!g || !S || S->g != g || !prm || !d_xbar || !d_vbar) return fail(AW_ERR_INVALID, "bad arguments");
  AW_TRY(check_adjoint_state(g));
  AW_TRY(validate_ids(g, g->sids.as<int32_t>(), g->n_sorted));
  aw_solve_stats st{};
  int rc = step_adjoint(g, S, prm, d_xbar, d_vbar, &st);
  AW_HIP(hipStreamSynchronize(g->ctx->stream));
  if (stats) *stats = st;
  return rc;
}

int aw_tape_create(aw_grid* g, aw_solver* S, int64_t n, int num_steps, int checkpoint_every, int keep_masks,
                   aw_tape** out) {
  if (!out) return fail(AW_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (!g || !S || S->g != g || n < 1 || n > 0x7FFFFFFF || num_steps < 1 || checkpoint_every < 1)
    return fail(AW_ERR_INVALID, "bad arguments");
  auto t = std::make_unique<aw_tape>();
  t->g = g;
  t->S = S;
  t->n = n;
  t->steps = num_steps;
  t->every = std::min(checkpoint_every, num_steps);
  t->keep_masks = keep_masks != 0;
  const int nck = (num_steps + t->every - 1) / t->every;
  for (int i = 0; i < nck; ++i) t->ck.push_back(std::make_unique<aw_tape::Ckpt>());
  for (int i = 0; i < t->every; ++i) t->rec.push_back(std::make_unique<aw_tape::Step>());
  S->refs++;
  *out = t.release();
  return AW_OK;
}

int aw_tape_destroy(aw_tape* t) {
  if (!t) return fail(AW_ERR_INVALID, "tape is NULL");
  aw_solver* S = t->S;
  (void)hipStreamSynchronize(t->g->ctx->stream);
  delete t;
  solver_release(S);
  return AW_OK;
}

int aw_rollout_forward(aw_tape* t, float* d_pos, float* d_vel, int32_t* d_ids, const aw_step_params* prm,
                       aw_rollout_stats* stats) {
  if (!t || !d_pos || !d_vel || !d_ids || !prm) return fail(AW_ERR_INVALID, "NULL argument");
  if (prm->d_density) return fail(AW_ERR_UNSUPPORTED, "rollouts differentiate with beta = 1 (no d_density)");
  aw_grid* g = t->g;
  hipStream_t s = g->ctx->stream;
  const int64_t n = t->n;
  t->prm = *prm;
  t->prm.warm_start = 0;   // every step of a rollout starts its solve from zero, as its recomputation does
  if (prm->rebuild_topology == 2 && prm->d_radius) {
    // the backward pass re-marks the levels of every recomputed step: keep the
    // radii as they were, whatever the caller does with its buffer meanwhile
    AW_TRY(copy_d2d(s, t->radius, prm->d_radius, (size_t)n * sizeof(float)));
    t->prm.d_radius = t->radius.as<float>();
  }
  t->have_forward = false;
  t->mask_level.clear();
  t->mask_liquid.clear();
  t->fwd_level.clear();
  t->fwd_hash.clear();
  AW_TRY(validate_ids(g, d_ids, n));
  aw_rollout_stats rs{};
  AW_HIP(hipStreamSynchronize(s));
  const double t0 = now_ms();
  int worst = AW_OK;
  for (int k = 0; k < t->steps; ++k) {
    if (k % t->every == 0) {
      aw_tape::Ckpt& c = *t->ck[k / t->every];
      AW_TRY(copy_d2d(s, c.pos, d_pos, (size_t)n * 3 * sizeof(float)));
      AW_TRY(copy_d2d(s, c.vel, d_vel, (size_t)n * 3 * sizeof(float)));
      AW_TRY(copy_d2d(s, c.ids, d_ids, (size_t)n * sizeof(int32_t)));
    }
    aw_step_stats ss{};
    int rc = aw_step(g, t->S, d_pos, d_vel, d_ids, n, &t->prm, &ss);
    if (rc != AW_OK && rc != AW_ERR_NOT_CONVERGED) return rc;
    if (rc != AW_OK) worst = rc;
    rs.forward_iterations += ss.solve.iterations;
    rs.forward_rel_residual = std::max(rs.forward_rel_residual, ss.solve.rel_residual);
    unsigned long long h = 0;
    AW_TRY(liquid_fingerprint(t, &h));
    t->fwd_level.push_back(g->h_level);
    t->fwd_hash.push_back(h);
    if (t->keep_masks) {
      t->mask_level.push_back(g->h_level);
      std::vector<uint8_t> lq(g->ncells);
      AW_HIP(hipMemcpy(lq.data(), g->liquid.p, g->ncells, hipMemcpyDeviceToHost));
      t->mask_liquid.push_back(std::move(lq));
    }
  }
  AW_HIP(hipStreamSynchronize(s));
  rs.forward_ms = now_ms() - t0;
  rs.steps = t->steps;
  t->have_forward = true;
  if (stats) *stats = rs;
  return worst == AW_OK ? AW_OK : fail(worst, "a forward solve ended above its tolerance");
}

int aw_rollout_backward(aw_tape* t, float* d_xbar, float* d_vbar, aw_rollout_stats* stats) {
  if (!t || !d_xbar || !d_vbar) return fail(AW_ERR_INVALID, "NULL argument");
  if (!t->have_forward) return fail(AW_ERR_INVALID, "no forward rollout on this tape");
  aw_grid* g = t->g;
  hipStream_t s = g->ctx->stream;
  const int64_t n = t->n;
  AW_TRY(t->wpos.ensure((size_t)n * 3 * sizeof(float)));
  AW_TRY(t->wvel.ensure((size_t)n * 3 * sizeof(float)));
  AW_TRY(t->wids.ensure((size_t)n * sizeof(int32_t)));
  aw_rollout_stats rs{};
  if (stats) rs = *stats;
  rs.recompute_ms = rs.backward_ms = 0.0;
  rs.adjoint_iterations = 0;
  rs.adjoint_rel_residual = 0.0;
  rs.replay_mismatched_steps = 0;
  rs.replay_max_state_diff = 0.0;
  AW_TRY(t->diff.ensure(2 * sizeof(uint32_t)));
  AW_TRY(t->inv.ensure((size_t)n * sizeof(int32_t)));
  int worst = AW_OK;
  const int nseg = (int)t->ck.size();
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int k0 = seg * t->every, k1 = std::min(t->steps, k0 + t->every);
    AW_HIP(hipStreamSynchronize(s));
    double ta = now_ms();
    aw_tape::Ckpt& c = *t->ck[seg];
    AW_HIP(hipMemcpyAsync(t->wpos.p, c.pos.p, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(t->wvel.p, c.vel.p, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    AW_HIP(hipMemcpyAsync(t->wids.p, c.ids.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    for (int k = k0; k < k1; ++k) {
      aw_step_stats ss{};
      int rc = aw_step(g, t->S, t->wpos.as<float>(), t->wvel.as<float>(), t->wids.as<int32_t>(), n, &t->prm, &ss);
      if (rc != AW_OK && rc != AW_ERR_NOT_CONVERGED) return rc;
      unsigned long long h = 0;
      AW_TRY(liquid_fingerprint(t, &h));
      if (g->h_level != t->fwd_level[k] || h != t->fwd_hash[k]) rs.replay_mismatched_steps++;
      if (k < k1 - 1) AW_TRY(tape_record(t, *t->rec[k - k0]));
    }
    if (seg + 1 < nseg) {
      // the segment's end is the next checkpoint's state
      const aw_tape::Ckpt& nx = *t->ck[seg + 1];
      uint32_t hd[2] = {0u, 0u};
      k_state_diff(s, t->wpos.as<float>(), t->wvel.as<float>(), t->wids.as<int32_t>(), nx.pos.as<float>(),
                   nx.vel.as<float>(), nx.ids.as<int32_t>(), n, t->inv.as<int32_t>(), t->diff.as<uint32_t>());
      AW_HIP(hipMemcpyAsync(hd, t->diff.p, sizeof hd, hipMemcpyDeviceToHost, s));
      AW_HIP(hipStreamSynchronize(s));
      AW_TRY(launch_status("replay check"));
      float d;
      std::memcpy(&d, &hd[0], sizeof d);
      if (hd[1]) d = INFINITY;
      rs.replay_max_state_diff = std::max(rs.replay_max_state_diff, (double)d);
    }
    AW_HIP(hipStreamSynchronize(s));
    double tb = now_ms();
    rs.recompute_ms += tb - ta;
    for (int k = k1 - 1; k >= k0; --k) {
      if (k < k1 - 1) AW_TRY(tape_restore(t, *t->rec[k - k0]));
      aw_solve_stats st{};
      int rc = step_adjoint(g, t->S, &t->prm, d_xbar, d_vbar, &st);
      if (rc != AW_OK && rc != AW_ERR_NOT_CONVERGED) return rc;
      if (rc != AW_OK) worst = rc;
      rs.adjoint_iterations += st.iterations;
      rs.adjoint_rel_residual = std::max(rs.adjoint_rel_residual, st.rel_residual);
    }
    AW_HIP(hipStreamSynchronize(s));
    rs.backward_ms += now_ms() - tb;
  }
  AW_TRY(launch_status("rollout_backward"));
  rs.steps = t->steps;
  if (stats) *stats = rs;
  return worst == AW_OK ? AW_OK : fail(worst, "an adjoint solve ended above its tolerance");
}

int aw_tape_masks(aw_tape* t, int step, uint8_t* h_levels, uint8_t* h_liquid, uint32_t liquid_cap,
                  uint32_t* ncells) {
  if (!t || !h_levels || !ncells) return fail(AW_ERR_INVALID, "NULL argument");
  if (!t->keep_masks || step < 0 || step >= (int)t->mask_level.size())
    return fail(AW_ERR_INVALID, "no masks recorded for step %d", step);
  std::copy(t->mask_level[step].begin(), t->mask_level[step].end(), h_levels);
  const auto& lq = t->mask_liquid[step];
  *ncells = (uint32_t)lq.size();
  if (h_liquid) {
    if (liquid_cap < lq.size()) return fail(AW_ERR_INVALID, "liquid buffer holds %u of %zu cells", liquid_cap, lq.size());
    std::copy(lq.begin(), lq.end(), h_liquid);
  }
  return AW_OK;
}

// ---- uniform resample and VTK export (vtx_visualization.cuh) ----

int aw_remap_cells(aw_grid* g, const uint8_t* d_level_prev, const float* d_src, float* d_dst) {
  if (!g || !d_level_prev || !d_src || !d_dst) return fail(AW_ERR_INVALID, "NULL argument");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "no topology");
  if (d_src == d_dst) return fail(AW_ERR_INVALID, "d_src and d_dst must not alias");
  hipStream_t s = g->ctx->stream;
  // the previous layout's cell offsets from its levels (an exclusive scan of r^3)
  AW_TRY(g->remap_coff.ensure((size_t)(g->nb + 1) * sizeof(uint32_t)));
  AW_TRY(g->scan_tmp.ensure(scan_scratch_words(g->nb) * sizeof(uint32_t)));
  k_cell_offsets(s, d_level_prev, g->nb, g->remap_coff.as<uint32_t>(), g->scan_tmp.as<uint32_t>());
  Prof pf(g->ctx, "pressure_remap", 8.0 * g->ncells + 9.0 * g->nb);
  k_pressure_remap(s, g->view(), d_level_prev, g->remap_coff.as<uint32_t>(), d_src, d_dst);
  return launch_status("remap_cells");
}

int aw_remap_cells_noalloc(aw_grid* g, const uint8_t* d_level_prev, const float* d_src, float* d_dst) {
  if (!g || !d_level_prev || !d_src || !d_dst) return fail(AW_ERR_INVALID, "NULL argument");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "no topology");
  if (d_src == d_dst) return fail(AW_ERR_INVALID, "d_src and d_dst must not alias");
  if (g->nb > NOALLOC_MAX_BLOCKS)
    return fail(AW_ERR_INVALID, "grid of %d blocks: aw_remap_cells_noalloc handles at most %d", g->nb,
                NOALLOC_MAX_BLOCKS);
  uint32_t* scr = g->ctx->scratch_dev;
  if (!scr) return fail(AW_ERR_UNSUPPORTED, "no scratch on device %d", g->ctx->device);
  static_assert(NOALLOC_MAX_BLOCKS + 1 + 2 * (NOALLOC_MAX_BLOCKS / 4096 + 2) <= (int)NOALLOC_SCRATCH_WORDS,
                "scratch holds the offsets and the scan's chunk sums");
  if ((size_t)g->nb + 1 + scan_scratch_words(g->nb) > NOALLOC_SCRATCH_WORDS)
    return fail(AW_ERR_INVALID, "scratch too small for %d blocks", g->nb);
  // the scan writes every word it reads (no clearing needed); no profiler
  // events (they could allocate)
  hipStream_t s = g->ctx->stream;
  uint32_t* coff_prev = scr;
  uint32_t* scan_tmp = scr + g->nb + 1;
  k_cell_offsets(s, d_level_prev, g->nb, coff_prev, scan_tmp);
  k_pressure_remap(s, g->view(), d_level_prev, coff_prev, d_src, d_dst);
  return launch_status("remap_cells_noalloc");
}

int aw_resample_cells(aw_grid* g, const float* d_field, const int out_dims[3], float* d_out) {
  if (!g || !d_field || !out_dims || !d_out) return fail(AW_ERR_INVALID, "NULL argument");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "no topology");
  for (int a = 0; a < 3; ++a) {
    const int o = out_dims[a];
    if (o < 1 || o > 65535 || (2ull * o - 1) * (uint64_t)(BASE * g->D[a]) >= (1ull << 32))
      return fail(AW_ERR_INVALID, "output size %d along axis %d out of range", o, a);
  }
  // algorithmic bytes: the output once and the field once
  Prof pf(g->ctx, "resample", 4.0 * out_dims[0] * out_dims[1] * out_dims[2] + 4.0 * g->ncells);
  k_resample(g->ctx->stream, g->view(), d_field, out_dims, d_out);
  return launch_status("resample");
}

namespace {

bool one_token(const char* s) {
  if (!s || !*s || std::strlen(s) > 200) return false;
  for (const char* c = s; *c; ++c)
    if (*c == ' ' || *c == '\n' || *c == '\r' || *c == '\t') return false;
  return true;
}

// legacy VTK binary sections are big-endian 32-bit words
void put_be32(FILE* f, const void* src, size_t words) {
  const uint32_t* w = static_cast<const uint32_t*>(src);
  std::vector<uint32_t> tmp(std::min<size_t>(std::max<size_t>(words, 1), 1 << 16));
  for (size_t i = 0; i < words; i += tmp.size()) {
    const size_t k = std::min(tmp.size(), words - i);
    for (size_t j = 0; j < k; ++j) tmp[j] = __builtin_bswap32(w[i + j]);
    std::fwrite(tmp.data(), 4, k, f);
  }
}

struct OutFile {
  FILE* f = nullptr;
  explicit OutFile(const char* path) : f(std::fopen(path, "wb")) {}
  ~OutFile() {
    if (f) std::fclose(f);
  }
  int close(const char* path) {
    const bool bad = std::ferror(f) != 0;
    const int rc = std::fclose(f);
    f = nullptr;
    return bad || rc ? fail(AW_ERR_INVALID, "writing %s failed", path) : AW_OK;
  }
};

}  // namespace

int aw_write_vtk_structured(const char* path, const char* name, const float* h_data, const int dims[3],
                            const float origin[3], const float spacing[3]) {
  if (!path || !h_data || !dims || !origin || !spacing) return fail(AW_ERR_INVALID, "NULL argument");
  if (!one_token(name)) return fail(AW_ERR_INVALID, "field name must be one word");
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return fail(AW_ERR_INVALID, "bad dimensions");
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  OutFile o(path);
  if (!o.f) return fail(AW_ERR_INVALID, "cannot open %s", path);
  std::fprintf(o.f, "# vtk DataFile Version 3.0\nMSBG %s resample\nBINARY\nDATASET STRUCTURED_POINTS\n", name);
  std::fprintf(o.f, "DIMENSIONS %d %d %d\nORIGIN %.9g %.9g %.9g\nSPACING %.9g %.9g %.9g\n", dims[0], dims[1], dims[2],
               origin[0], origin[1], origin[2], spacing[0], spacing[1], spacing[2]);
  std::fprintf(o.f, "POINT_DATA %zu\nSCALARS %s float 1\nLOOKUP_TABLE default\n", n, name);
  put_be32(o.f, h_data, n);
  std::fputc('\n', o.f);
  return o.close(path);
}

int aw_write_vtk_blocks(const char* path, const int dims[3], float block_size, const uint8_t* h_levels) {
  if (!path || !dims || !h_levels) return fail(AW_ERR_INVALID, "NULL argument");
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || (int64_t)dims[0] * dims[1] * dims[2] > (1 << 26))
    return fail(AW_ERR_INVALID, "bad dimensions");
  if (!(block_size > 0.f)) return fail(AW_ERR_INVALID, "block size must be positive");
  const uint32_t nb = (uint32_t)dims[0] * dims[1] * dims[2];
  OutFile o(path);
  if (!o.f) return fail(AW_ERR_INVALID, "cannot open %s", path);
  std::fprintf(o.f, "# vtk DataFile Version 3.0\nMSBG blocks\nBINARY\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(o.f, "POINTS %u float\n", 8 * nb);
  // the eight corners of each block in VTK hexahedron order, blocks x fastest
  static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<float> pts((size_t)24 * nb);
  for (uint32_t b = 0; b < nb; ++b) {
    const uint32_t c[3] = {b % dims[0], (b / dims[0]) % dims[1], b / (dims[0] * dims[1])};
    for (int k = 0; k < 8; ++k)
      for (int a = 0; a < 3; ++a) pts[24 * (size_t)b + 3 * k + a] = (float)(c[a] + corner[k][a]) * block_size;
  }
  put_be32(o.f, pts.data(), pts.size());
  std::fprintf(o.f, "\nCELLS %u %u\n", nb, 9 * nb);
  std::vector<int32_t> cells((size_t)9 * nb);
  for (uint32_t b = 0; b < nb; ++b) {
    cells[9 * (size_t)b] = 8;
    for (int k = 0; k < 8; ++k) cells[9 * (size_t)b + 1 + k] = (int32_t)(8 * b + k);
  }
  put_be32(o.f, cells.data(), cells.size());
  std::fprintf(o.f, "\nCELL_TYPES %u\n", nb);
  std::vector<int32_t> v(nb, 12);  // VTK_HEXAHEDRON
  put_be32(o.f, v.data(), nb);
  std::fprintf(o.f, "\nCELL_DATA %u\nSCALARS refineLevel int 1\nLOOKUP_TABLE default\n", nb);
  for (uint32_t b = 0; b < nb; ++b) v[b] = h_levels[b];
  put_be32(o.f, v.data(), nb);
  std::fputc('\n', o.f);
  return o.close(path);
}

int aw_write_vtk_particles(const char* path, const float* h_pos, const float* h_radius, const float* h_vel,
                           int64_t n) {
  if (!path || (n > 0 && !h_pos) || n < 0) return fail(AW_ERR_INVALID, "bad arguments");
  if (n > (1 << 30)) return fail(AW_ERR_INVALID, "too many particles for one legacy VTK file");
  OutFile o(path);
  if (!o.f) return fail(AW_ERR_INVALID, "cannot open %s", path);
  std::fprintf(o.f, "# vtk DataFile Version 3.0\nParticle Cloud\nBINARY\nDATASET POLYDATA\nPOINTS %lld float\n",
               (long long)n);
  put_be32(o.f, h_pos, 3 * (size_t)n);
  // one vertex cell per particle, so that the points render
  std::fprintf(o.f, "\nVERTICES %lld %lld\n", (long long)n, 2 * (long long)n);
  std::vector<int32_t> vc(2 * (size_t)std::min<int64_t>(n, 1 << 16));
  for (int64_t i = 0; i < n; i += (int64_t)vc.size() / 2) {
    const int64_t k = std::min<int64_t>((int64_t)vc.size() / 2, n - i);
    for (int64_t j = 0; j < k; ++j) {
      vc[2 * j] = 1;
      vc[2 * j + 1] = (int32_t)(i + j);
    }
    put_be32(o.f, vc.data(), 2 * (size_t)k);
  }
  if (h_radius || h_vel) std::fprintf(o.f, "\nPOINT_DATA %lld", (long long)n);
  if (h_radius) {
    std::fprintf(o.f, "\nSCALARS radius float 1\nLOOKUP_TABLE default\n");
    put_be32(o.f, h_radius, (size_t)n);
  }
  if (h_vel) {
    std::fprintf(o.f, "\nVECTORS velocity float\n");
    put_be32(o.f, h_vel, 3 * (size_t)n);
  }
  std::fputc('\n', o.f);
  return o.close(path);
}

int aw_export_cells_vtk(aw_grid* g, const float* d_field, const int out_dims[3], const char* name, const char* path) {
  if (!g || !d_field || !out_dims || !path) return fail(AW_ERR_INVALID, "NULL argument");
  if (!one_token(name)) return fail(AW_ERR_INVALID, "field name must be one word");
  const size_t n = (size_t)std::max(out_dims[0], 0) * std::max(out_dims[1], 0) * std::max(out_dims[2], 0);
  DevBuf tmp;
  AW_TRY(tmp.ensure(std::max<size_t>(n, 1) * sizeof(float)));
  AW_TRY(aw_resample_cells(g, d_field, out_dims, tmp.as<float>()));
  std::vector<float> h(n);
  AW_HIP(hipMemcpyAsync(h.data(), tmp.p, n * sizeof(float), hipMemcpyDeviceToHost, g->ctx->stream));
  AW_HIP(hipStreamSynchronize(g->ctx->stream));
  float spacing[3], origin[3];
  for (int a = 0; a < 3; ++a) {
    // samples sit at voxel centres, half a spacing in from the domain's corner
    spacing[a] = (float)((double)BASE * g->D[a] * g->h / out_dims[a]);
    origin[a] = 0.5f * spacing[a];
  }
  return aw_write_vtk_structured(path, name, h.data(), out_dims, origin, spacing);
}

int aw_export_blocks_vtk(aw_grid* g, const char* path) {
  if (!g || !path) return fail(AW_ERR_INVALID, "NULL argument");
  if (!g->have_topology) return fail(AW_ERR_INVALID, "no topology");
  return aw_write_vtk_blocks(path, g->D, (float)(BASE * g->h), g->h_level.data());
}

}  // extern "C"
