// TEST INFRASTRUCTURE ONLY — drives the FP64 CPU oracle (aw_oracle.cpp)
// through its entry points on a small scene, for sanitizer builds of the
// oracle (`make -C oracle sanitize`): AddressSanitizer and
// UndefinedBehaviorSanitizer with OpenMP threads.  It exercises topology,
// P2G with and without densities, both solver methods, adaptive smoothing,
// the FP64 step and its adjoint, the pressure remap and the resample, and
// prints one line of summary figures; any sanitizer report aborts it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
void* og_create(int dx, int dy, int dz, int num_levels, double h, const int* band);
void og_destroy(void* p);
int og_bin(void* p, const float* pos, int64_t n, int32_t* key, uint32_t* count);
int og_get_levels(void* p, uint8_t* out);
int og_totals(void* p, int64_t* out);
int og_set_params(void* p, double omega, int local_sweeps, int pre, int post, int coarse_sweeps, double prolong_scale);
int og_set_adaptive(void* p, int on, double theta, int max_block_sweeps, double block_reduction);
int og_step(void* p, const float* pos, const float* vel, int64_t n, double dt, double gx, double gy, double gz,
            double alpha, int method, double tol, int max_iters, double* pos_out, double* vel_out, int* iters,
            double* rel, double* stage_ms, const double* density);
int og_step64(void* p, const double* pos, const double* vel, int64_t n, double dt, double gx, double gy, double gz,
              double alpha, int method, double tol, int max_iters, const uint8_t* levels_in,
              const uint8_t* liquid_in, uint8_t* levels_out, uint8_t* liquid_out, double* pos_out, double* vel_out,
              int* iters, double* rel);
int og_step_adjoint(void* p, int fp64, const void* pos, const void* vel, int64_t n, double alpha, double dt,
                    double tol, int max_iters, double* xbar, double* vbar, int* iters, double* rel);
int og_remap_cells(void* p, const uint8_t* level_prev, const double* src, double* dst);
int og_resample(void* p, const double* field, int ox, int oy, int oz, double* out);
int og_get_cell(void* p, int which, double* out);
}

namespace {

// a dam of particles in the lower corner of the unit tank, 2 per cell per
// axis with a fixed jitter, resting velocity plus a swirl
void scene(int n, std::vector<float>& pos, std::vector<float>& vel) {
  uint32_t s = 12345u;
  auto rnd = [&s] {
    s = s * 1664525u + 1013904223u;
    return (float)((s >> 8) & 0xFFFF) / 65536.0f;
  };
  for (int z = 0; z < n / 2; ++z)
    for (int y = 0; y < n / 2; ++y)
      for (int x = 0; x < n / 2; ++x)
        for (int k = 0; k < 8; ++k) {
          const float p[3] = {(x + 0.25f + 0.5f * (k & 1) + 0.1f * (rnd() - 0.5f)) / n,
                              (y + 0.25f + 0.5f * ((k >> 1) & 1) + 0.1f * (rnd() - 0.5f)) / n,
                              (z + 0.25f + 0.5f * (k >> 2) + 0.1f * (rnd() - 0.5f)) / n};
          for (int a = 0; a < 3; ++a) pos.push_back(p[a]);
          vel.push_back(0.5f * (p[1] - 0.25f));
          vel.push_back(-0.3f);
          vel.push_back(-0.5f * (p[0] - 0.25f));
        }
}

int fails = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++fails;
  }
}

}  // namespace

int main() {
  const int n = 32, D = n / 8;
  const int band[4] = {0, 1, 0, 0};
  std::vector<float> pos, vel;
  scene(n, pos, vel);
  const int64_t np = (int64_t)pos.size() / 3;
  const double dt = 1.0 / 240.0, alpha = 0.95;
  void* g = og_create(D, D, D, 3, 1.0 / n, band);
  expect(g != nullptr, "og_create");
  og_set_params(g, 0.85714, 2, 2, 2, 16, 1.0);

  std::vector<double> po(3 * np), vo(3 * np), stage(4);
  int it = 0;
  double rel = 0.0;
  expect(og_step(g, pos.data(), vel.data(), np, dt, 0, -9.81, 0, alpha, 0, 1e-8, 200, po.data(), vo.data(), &it,
                 &rel, stage.data(), nullptr) == 0, "PCG step");
  const int it_pcg = it;

  std::vector<double> dens(np);
  for (int64_t i = 0; i < np; ++i) dens[i] = pos[3 * i + 1] < 0.125f ? 2.0 : 0.5;
  expect(og_step(g, pos.data(), vel.data(), np, dt, 0, -9.81, 0, alpha, 0, 1e-8, 200, po.data(), vo.data(), &it,
                 &rel, stage.data(), dens.data()) == 0, "two-phase step");

  og_set_params(g, 0.85714, 2, 2, 2, 16, 0.75);
  og_set_adaptive(g, 1, 1e-4, 10, 0.33);
  expect(og_step(g, pos.data(), vel.data(), np, dt, 0, -9.81, 0, alpha, 1, 1e-8, 400, po.data(), vo.data(), &it,
                 &rel, stage.data(), nullptr) == 0, "adaptive stationary step");
  const int it_adaptive = it;
  og_set_adaptive(g, 0, 1e-4, 10, 0.33);
  og_set_params(g, 0.85714, 2, 2, 2, 16, 1.0);

  // FP64 step and its adjoint
  std::vector<double> p64(pos.begin(), pos.end()), v64(vel.begin(), vel.end());
  std::vector<uint8_t> lv(D * D * D), lq(D * D * D * 512);
  expect(og_step64(g, p64.data(), v64.data(), np, dt, 0, -9.81, 0, alpha, 0, 1e-10, 400, nullptr, nullptr, lv.data(),
                   lq.data(), po.data(), vo.data(), &it, &rel) == 0, "FP64 step");
  std::vector<double> xbar(3 * np, 0.0), vbar(3 * np, 0.0);
  for (int64_t i = 0; i < np; ++i) xbar[3 * i + 1] = 1.0 / (double)np;
  expect(og_step_adjoint(g, 1, p64.data(), v64.data(), np, alpha, dt, 1e-10, 400, xbar.data(), vbar.data(), &it,
                         &rel) == 0, "step adjoint");
  double gnorm = 0.0;
  for (double v : vbar) gnorm += v * v;

  // pressure remap onto the topology of a coarser band, and a resample
  int64_t tot[4];
  og_totals(g, tot);
  std::vector<double> p(tot[0]);
  og_get_cell(g, 0, p.data());
  std::vector<uint8_t> prev(D * D * D);
  og_get_levels(g, prev.data());
  const int band2[4] = {0, 0, 0, 0};
  void* g2 = og_create(D, D, D, 3, 1.0 / n, band2);
  int64_t tot2[4];
  og_totals(g2, tot2);
  std::vector<double> p2(tot2[0]);
  expect(og_remap_cells(g2, prev.data(), p.data(), p2.data()) == 0, "remap");
  std::vector<double> uni(16 * 16 * 16);
  expect(og_resample(g, p.data(), 16, 16, 16, uni.data()) == 0, "resample");
  og_destroy(g2);
  og_destroy(g);
  std::printf("sanitize_driver: %lld particles, PCG %d iterations, adaptive %d V-cycles, |dL/dv| %.3e, %s\n",
              (long long)np, it_pcg, it_adaptive, std::sqrt(gnorm), fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}
