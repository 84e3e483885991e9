// CPU build of the forward projector: fwd_body and reduce_body from
// pytomography_amd/csrc/forward_body.hpp compiled with the host stand-in in
// tests/cpu_build/hip, one std::thread per GPU lane.  Blocks run one at a
// time; dynamic LDS and the chunk partials start as NaN so that a read of an
// entry the kernel never wrote shows in the output.
//
// With attenuation the field or suffix table is made first by the march
// bodies (cpu_attenuation.hpp) and the body runs in the matching mode.
//
//   fwd_cpu <dir> [trb]  (problem files as in cpu_runtime.hpp, plus f.bin
//                         float32 [N][N][Nz]); writes <dir>/proj.bin float32
//                         [n_sel][N][Nz] and prints "H .. blocks .. threads ..
//                         nan_partial ..".  With "trb" the body runs its
//                         B^T mode: the transpose of the reference back
//                         projection's gather replaces the forward gather.
#include "cpu_attenuation.hpp"

using namespace spect;

int main(int argc, char **argv) {
  if (argc < 2 || argc > 3 || (argc == 3 && std::string(argv[2]) != "trb")) {
    std::fprintf(stderr, "usage: fwd_cpu <dir> [trb]\n");
    return 2;
  }
  const std::string dir = argv[1];
  const bool trb = argc == 3;
  const cpu_build::Problem p = cpu_build::load_problem(dir);
  const Geo<float> &g = p.g;
  const int N = p.N, Nz = p.Nz, n = p.n, H = p.H;
  const auto f = cpu_build::read_file<float>(dir + "/f.bin", (size_t)N * N * Nz);
  const int JO = kTJ + 2 * H;
  std::vector<float> partial((size_t)p.C * n * g.ntiles * JO * Nz, std::numeric_limits<float>::quiet_NaN());
  const cpu_build::Attenuation att = cpu_build::build_attenuation(p);
  const FwdJob<float> J{g, p.tables(), f.data(), att.mu, att.att, att.att ? 1 : 0, att.attf, p.sel.data(), n,
                        partial.data()};

  // launch_fwd_h: block size, phase-A coverage, dynamic LDS
  constexpr int RZ = FwdTile<float>::RZ, QS = FwdTile<float>::QS;
  const int bt = cpu_build::block_threads(Nz, RZ, QS);
  if (bt > 512 || !phase_a_fits<float, QS>(bt, Nz)) {
    std::fprintf(stderr, "launch configuration rejected\n");
    return 2;
  }
  const int runs = (Nz + RZ - 1) / RZ;
  std::vector<unsigned char> smem(
      (2 * (size_t)kTJ * (runs * RZ + 2 * H + 1) + (size_t)p.L * 2 * (2 * H + 1)) * sizeof(float));
  auto run = [&](auto h_c) {
    constexpr int HH = decltype(h_c)::value;
    cpu_build::run_grid(n, g.ntiles, p.C, bt, smem, [&](int al, int jt, int c, unsigned char *s) {
      if (trb) {
        if (att.attf) fwd_body<float, RZ, QS, HH, true, true>(J, al, jt, c, s);
        else fwd_body<float, RZ, QS, HH, true, false>(J, al, jt, c, s);
      } else {
        if (att.attf) fwd_body<float, RZ, QS, HH, false, true>(J, al, jt, c, s);
        else fwd_body<float, RZ, QS, HH, false, false>(J, al, jt, c, s);
      }
    });
  };
  switch (H) {
    case 0: run(std::integral_constant<int, 0>{}); break;
    case 2: run(std::integral_constant<int, 2>{}); break;
    case 4: run(std::integral_constant<int, 4>{}); break;
    case 6: run(std::integral_constant<int, 6>{}); break;
    case 8: run(std::integral_constant<int, 8>{}); break;
    default: std::fprintf(stderr, "PSF half width %d not built here\n", H); return 2;
  }
  const size_t nan_partial = cpu_build::count_nan(partial);

  ReduceJob<float> RJ{g, partial.data(), n, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0f};
  std::vector<float> proj((size_t)n * N * Nz);
  RJ.proj = proj.data();
  for (size_t i = 0; i < proj.size(); ++i) reduce_body<float>(RJ, i);
  cpu_build::write_file(dir + "/proj.bin", proj);
  std::printf("H %d blocks %d threads %d nan_partial %zu\n", H, n * g.ntiles * p.C, bt, nan_partial);
  return 0;
}
