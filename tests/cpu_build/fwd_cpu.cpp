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
  const std::string mode = argc >= 3 ? std::string(argv[2]) : std::string();
  if (argc < 2 || argc > 4 || (argc >= 3 && mode != "trb" && mode != "sched")) {
    std::fprintf(stderr, "usage: fwd_cpu <dir> [trb | sched [blocks]]\n");
    return 2;
  }
  const std::string dir = argv[1];
  const bool trb = mode == "trb", sched_mode = mode == "sched";
  const int sched_blocks = argc == 4 ? std::atoi(argv[3]) : 7;
  const cpu_build::Problem p = cpu_build::load_problem(dir);
  const Geo<float> &g = p.g;
  const int N = p.N, Nz = p.Nz, n = p.n, H = p.H;
  const auto f = cpu_build::read_file<float>(dir + "/f.bin", (size_t)N * N * Nz);
  const int JO = kTJ + 2 * H;
  const int npairs = n * g.ntiles;
  const int grid = sched_grid_blocks(n, g.ntiles, g.P, sched_blocks);
  const size_t slots = sched_mode ? (size_t)grid : (size_t)p.C * npairs;
  std::vector<float> partial(slots * JO * Nz, std::numeric_limits<float>::quiet_NaN());
  const cpu_build::Attenuation att = cpu_build::build_attenuation(p);
  if (sched_mode && att.att) {
    std::fprintf(stderr, "the balanced schedule runs without attenuation or with the field\n");
    return 2;
  }
  const FwdJob<float> J{g, p.tables(), f.data(), att.mu, att.att, att.att ? 1 : 0, att.attf, p.sel.data(), n,
                        partial.data()};

  // launch_fwd_h: block size, phase-A coverage, dynamic LDS
  constexpr int RZ = FwdTile<float>::RZ, QS = FwdTile<float>::QS;
  const int bt = cpu_build::block_threads(Nz, RZ, QS);
  if (bt > 512 || !phase_a_fits<float, QS>(bt, Nz)) {
    std::fprintf(stderr, "launch configuration rejected\n");
    return 2;
  }
  std::vector<int> sched(2 + (size_t)npairs + 1, -1);
  std::vector<unsigned char> smem(std::max(fwd_smem_bytes<float, RZ, 8>(Nz, p.L), (size_t)64));
  auto run = [&](auto h_c) {
    constexpr int HH = decltype(h_c)::value;
    if (sched_mode) {
      cpu_build::run_grid(grid, 1, 1, bt, smem, [&](int b, int, int, unsigned char *s) {
        const int piece = xcd_piece(b, grid);  // fwd_sched_kernel's block -> piece order
        if (att.attf) fwd_sched_body<float, RZ, QS, HH, true>(J, p.tables().ranges, sched.data(), sched_blocks, piece, s);
        else fwd_sched_body<float, RZ, QS, HH, false>(J, p.tables().ranges, sched.data(), sched_blocks,
                                                        piece, s);
      });
      return;
    }
    cpu_build::run_grid(n, g.ntiles, p.C, bt, smem, [&](int al, int jt, int c, unsigned char *s) {
      const FwdPiece pc = chunk_piece(g, n, al, jt, c);
      if (trb) {
        if (att.attf) fwd_body<float, RZ, QS, HH, true, true>(J, al, jt, pc, s);
        else fwd_body<float, RZ, QS, HH, true, false>(J, al, jt, pc, s);
      } else {
        if (att.attf) fwd_body<float, RZ, QS, HH, false, true>(J, al, jt, pc, s);
        else fwd_body<float, RZ, QS, HH, false, false>(J, al, jt, pc, s);
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
  // every slot the reduce reads must have been written
  const size_t used = sched_mode ? (size_t)sched[0] : slots;
  const size_t nan_partial =
      cpu_build::count_nan(std::vector<float>(partial.begin(), partial.begin() + used * JO * Nz));

  ReduceJob<float> RJ{g, partial.data(), n, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0f};
  std::vector<float> proj((size_t)n * N * Nz);
  RJ.proj = proj.data();
  for (size_t i = 0; i < proj.size(); ++i) {
    if (sched_mode) reduce_sched_body<float>(RJ, sched.data(), i);
    else reduce_body<float>(RJ, i);
  }
  cpu_build::write_file(dir + "/proj.bin", proj);
  std::printf("H %d blocks %d threads %d nan_partial %zu slots %zu\n", H,
              sched_mode ? grid : n * g.ntiles * p.C, bt, nan_partial, used);
  return 0;
}
