// A caller shaped like the reference's src/main.cu (topology -> P2G ->
// V-cycle hierarchy -> pressure solve -> G2P -> advection), written against
// the reference's class names in include/msbg_compat.hpp and linked only to
// libanitowave.so.  Particles come from and go back to a file, so that the
// same run can be compared with other callers of the library.
//
//   msbg_main in.bin out.bin DX DY DZ LEVELS BAND0 BAND1 STEPS TOL [CYCLES]
//
// in.bin / out.bin: int64 n, then float pos[3 n], float vel[3 n] (AoS,
// little-endian).  The grid is DX x DY x DZ blocks of 8^3 finest cells with
// h = 1 / (8 DX); every step rebuilds the topology from the distance bands
// (BAND0, BAND1 blocks), deposits with dt = 1/240 and gravity -9.81, solves
// the pressure to relative residual TOL (PCG preconditioned by one V-cycle,
// from p = 0; with CYCLES > 0 instead CYCLES stationary V-cycles through
// runCycle, as the reference's main loop runs 10), projects, and moves the
// particles with FLIP 0.95 and RK3.  Exit status 0 on success (a solve at the
// FP32 floor counts as success and is reported), 1 on an error.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "msbg_compat.hpp"

namespace {

struct I3 {
  int x, y, z;
};
struct F3 {
  float x, y, z;
};

bool read_particles(const char* path, std::vector<F3>& pos, std::vector<F3>& vel) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  int64_t n = 0;
  bool ok = std::fread(&n, sizeof n, 1, f) == 1 && n >= 0;
  if (ok) {
    pos.resize((size_t)n);
    vel.resize((size_t)n);
    ok = std::fread(pos.data(), sizeof(F3), (size_t)n, f) == (size_t)n &&
         std::fread(vel.data(), sizeof(F3), (size_t)n, f) == (size_t)n;
  }
  std::fclose(f);
  return ok;
}

bool write_particles(const char* path, const std::vector<F3>& pos, const std::vector<F3>& vel) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const int64_t n = (int64_t)pos.size();
  bool ok = std::fwrite(&n, sizeof n, 1, f) == 1 && std::fwrite(pos.data(), sizeof(F3), pos.size(), f) == pos.size() &&
            std::fwrite(vel.data(), sizeof(F3), vel.size(), f) == vel.size();
  return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 11) {
    std::fprintf(stderr, "usage: %s in.bin out.bin DX DY DZ LEVELS BAND0 BAND1 STEPS TOL [CYCLES]\n", argv[0]);
    return 1;
  }
  const I3 dims = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
  msbg::Bands bands;
  bands.num_levels = std::atoi(argv[6]);
  bands.band[0] = std::atoi(argv[7]);
  bands.band[1] = std::atoi(argv[8]);
  const int steps = std::atoi(argv[9]);
  const double tol = std::atof(argv[10]);
  const int cycles = argc > 11 ? std::atoi(argv[11]) : 0;
  std::vector<F3> pos, vel;
  if (!read_particles(argv[1], pos, vel)) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  const uint32_t n = (uint32_t)pos.size();
  const float h = 1.0f / (8.0f * dims.x);
  const float dt = 1.0f / 240.0f, gravity[3] = {0.0f, -9.81f, 0.0f};
  try {
    // MSBGManager mg(dims, blockSize, domainMin)  (main.cu:99-106)
    msbg::MSBGManager mg(dims, 8.0f * h, F3{0.f, 0.f, 0.f}, bands);
    mg.allocateChannels();
    msbg::DeviceBuffer d_pos(mg, sizeof(F3) * (uint64_t)n), d_vel(mg, sizeof(F3) * (uint64_t)n);
    d_pos.upload(pos.data(), sizeof(F3) * (uint64_t)n);
    d_vel.upload(vel.data(), sizeof(F3) * (uint64_t)n);
    msbg::VCycleSolver vc(3, mg.numBlocks());
    msbg::G2PManager g2p;
    g2p.init(mg);
    for (int k = 0; k < steps; ++k) {
      mg.buildTopology(d_pos.as<F3>(), n);                           // main.cu:156-275
      msbg::P2GTransfer::run(mg, d_pos.as<F3>(), d_vel.as<F3>(), n, dt, gravity);  // main.cu:278-302
      mg.runGalerkin();
      vc.init(mg);                                                    // VCycle.cuh:177-249
      if (cycles > 0) {
        msbg::VCycleSolver::zeroPressure(mg);
        for (int c = 0; c < cycles; ++c) vc.runCycle(mg, 0);          // main.cu:314-317
      } else {
        aw_solve_stats st{};
        const int rc = vc.solve(mg, tol, 200, &st);
        std::printf("step %d: %d iterations, true rel. residual %.3e%s\n", k, st.iterations, st.rel_residual,
                    rc == AW_OK ? "" : " (FP32 floor: not converged)");
      }
      g2p.backupGridVelocity(mg);
      g2p.performG2P(mg, d_vel.as<F3>(), d_pos.as<F3>(), n, 0.95f);  // G2P.cuh:192-214
      g2p.advect(mg, d_pos.as<F3>(), n, dt);                          // G2P.cuh:216-227
    }
    mg.synchronize();
    d_pos.download(pos.data(), sizeof(F3) * (uint64_t)n);
    d_vel.download(vel.data(), sizeof(F3) * (uint64_t)n);
    g2p.cleanup();
  } catch (const msbg::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  if (!write_particles(argv[2], pos, vel)) {
    std::fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  std::printf("%u particles, %d steps\n", n, steps);
  return 0;
}
