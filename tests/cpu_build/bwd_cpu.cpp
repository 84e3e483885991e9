// CPU build of the back projection: field_body and gather_body from
// pytomography_amd/csrc/backward_body.hpp compiled with the host stand-in in
// tests/cpu_build/hip, one std::thread per GPU lane for the field blocks.
// The field W starts as NaN, as do the LDS buffers, so that a read of an entry
// the field kernel never wrote -- also one taken with weight zero -- shows in
// the output.  With attenuation the field or suffix table is made first by the
// march bodies (cpu_attenuation.hpp).
//
//   bwd_cpu <dir>     (problem files as in cpu_runtime.hpp, plus p.bin
//                      float32 [n_sel][N][Nz], the projections of the
//                      selected angles); writes <dir>/bp_ref.bin (reference
//                      back projection, gather mode 0) and <dir>/bp_adj.bin
//                      (exact adjoint, mode 1), float32 [N][N][Nz], and
//                      prints "H .. w_written .. w_total .."
//   bwd_cpu <dir> grad  (plus f.bin and mu.bin float32 [N][N][Nz]): the
//                      differentiable projector's VJP of u = p.bin as
//                      spect_forward_vjp runs it -- FP64 attenuation rows of the
//                      call's angles, the field with the mu-gradient parts
//                      (GRAD), the prefix of chunk totals, gather mode 2 --
//                      writing <dir>/grad_f.bin and <dir>/grad_mu.bin
//   bwd_cpu <dir> sched [blocks]  the field from the balanced schedule sized
//                      for that many resident blocks (one piece per block,
//                      each block locating its own)
#include "backward_body.hpp"
#include "cpu_attenuation.hpp"

using namespace spect;

namespace {

// grad_impl (capi.hip) with the mu gradient requested.
int run_grad(const std::string &dir, const cpu_build::Problem &p, const std::vector<float> &u,
             std::vector<float> &W) {
  const Geo<float> &g = p.g;
  const int N = p.N, Nz = p.Nz, n = p.n, H = p.H, C = p.C;
  if (p.mu.empty()) {
    std::fprintf(stderr, "grad needs mu.bin (att_mode > 0)\n");
    return 2;
  }
  const auto f = cpu_build::read_file<float>(dir + "/f.bin", (size_t)N * N * Nz);
  const Tables<float> tb = p.tables();
  const size_t n_jext = g.je1 - g.je0, plane = n_jext * Nz;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  // FP64 suffix rows of the call's angles (launch_att_table over grid (n, ntiles))
  std::vector<double> rows((size_t)n * C * plane, nan);
  {
    constexpr int RZ = FwdTile<float>::RZ;
    const int bt = (kTJ * ((Nz + RZ - 1) / RZ) + 63) / 64 * 64;
    std::vector<unsigned char> none(16);
    cpu_build::run_grid(n, g.ntiles, 1, bt, none, [&](int al, int jt, int, unsigned char *) {
      att_table_body<float, RZ>(g, tb, p.mu.data(), p.sel.data(), rows.data(), al, jt);
    });
  }
  std::vector<double> Wm((size_t)n * p.P * plane, nan), Dsum((size_t)n * C * plane, nan);
  const FieldJob<float> FJ{g, tb, u.data(), p.mu.data(), rows.data(), 0, nullptr, p.sel.data(), n,
                           W.data(), f.data(), Wm.data(), Dsum.data()};
  constexpr int RZ = BpTile<float>::RZ, QS = BpTile<float>::QS;
  const int bt = cpu_build::block_threads(Nz, RZ, QS);
  const int runs = (Nz + RZ - 1) / RZ;
  std::vector<unsigned char> smem(
      ((size_t)kTJ * (runs * RZ + 2 * H + 1) + (size_t)p.L * 2 * (2 * H + 1)) * sizeof(float));
  auto run = [&](auto h_c) {
    constexpr int HH = decltype(h_c)::value;
    cpu_build::run_grid(n, g.ntiles, C, bt, smem, [&](int al, int jt, int c, unsigned char *s) {
      field_body<float, RZ, QS, HH, true, false>(FJ, al, jt, c, s);
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
  // chunk_prefix_kernel: Dsum[a][c] <- sum over c' <= c
  for (int a = 0; a < n; ++a)
    for (size_t r = 0; r < plane; ++r) {
      double acc = 0.0;
      for (int c = 0; c < C; ++c) {
        double &v = Dsum[((size_t)a * C + c) * plane + r];
        acc += v;
        v = acc;
      }
    }
  constexpr int ZV = gather_zv<float>();  // as launch_bp_gather
  std::vector<float> grad_f((size_t)N * N * Nz, std::numeric_limits<float>::quiet_NaN());
  std::vector<float> grad_mu(grad_f.size(), std::numeric_limits<float>::quiet_NaN());
  const GatherJob<float> GJ{g, tb, W.data(), p.sel.data(), n, SPECT_OUT_STORE, grad_f.data(), nullptr,
                            nullptr, 0.0f, 0.0f, Wm.data(), Dsum.data(), grad_mu.data()};
  // as bp_gather_adj_kernel: 256-lane blocks sharing their columns' lists
  const int zq = (Nz + ZV - 1) / ZV;
  const size_t lanes = (size_t)N * N * zq;
  std::vector<unsigned char> asmem(adj_block_smem<float>(zq));
  cpu_build::run_grid((int)((lanes + 255) / 256), 1, 1, 256, asmem,
                      [&](int b, int, int, unsigned char *s) { gather_adj_block<float, ZV, 2>(GJ, b, s); });
  cpu_build::write_file(dir + "/grad_f.bin", grad_f);
  cpu_build::write_file(dir + "/grad_mu.bin", grad_mu);
  std::printf("H %d grad 1\n", H);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string mode = argc >= 3 ? std::string(argv[2]) : std::string();
  if (argc < 2 || argc > 4 || (argc >= 3 && mode != "grad" && mode != "sched") ||
      (argc == 4 && mode != "sched")) {
    std::fprintf(stderr, "usage: bwd_cpu <dir> [grad | sched [blocks]]\n");
    return 2;
  }
  const std::string dir = argv[1];
  const bool sched_mode = mode == "sched";
  const int sched_blocks = argc == 4 ? std::atoi(argv[3]) : 7;
  const cpu_build::Problem p = cpu_build::load_problem(dir);
  const Geo<float> &g = p.g;
  const int N = p.N, Nz = p.Nz, n = p.n, H = p.H;
  const auto proj = cpu_build::read_file<float>(dir + "/p.bin", (size_t)n * N * Nz);
  const size_t n_jext = g.je1 - g.je0;
  std::vector<float> W((size_t)n * p.P * n_jext * Nz, std::numeric_limits<float>::quiet_NaN());
  if (mode == "grad") return run_grad(dir, p, proj, W);
  const cpu_build::Attenuation att = cpu_build::build_attenuation(p);
  if (sched_mode && att.att) {
    std::fprintf(stderr, "the balanced schedule runs without attenuation or with the field\n");
    return 2;
  }
  const FieldJob<float> FJ{g, p.tables(), proj.data(), att.mu, att.att, att.att ? 1 : 0, att.attf,
                           p.sel.data(), n, W.data(), nullptr, nullptr, nullptr};

  // launch_field_h: block size and dynamic LDS
  constexpr int RZ = BpTile<float>::RZ, QS = BpTile<float>::QS;
  const int bt = cpu_build::block_threads(Nz, RZ, QS);
  if (bt > 1024) {
    std::fprintf(stderr, "launch configuration rejected\n");
    return 2;
  }
  const int runs = (Nz + RZ - 1) / RZ;
  // a scheduled piece holds up to kPieceMax planes, a chunk L
  const size_t planes = sched_mode ? std::max(p.L, kPieceMax) : p.L;
  std::vector<unsigned char> smem(
      ((size_t)kTJ * (runs * RZ + 2 * H + 1) + planes * 2 * (2 * H + 1)) * sizeof(float));
  const int npairs = n * g.ntiles;
  const int grid = sched_grid_blocks(n, g.ntiles, g.P, sched_blocks);
  int pieces = npairs * p.C;
  if (sched_mode) {
    // the schedule as block 0 of the forward writes it (for the printout)
    std::vector<int> sched(2 + (size_t)npairs + 1, -1);
    std::vector<unsigned char> none(64);
    cpu_build::run_grid(1, 1, 1, bt, none, [&](int, int, int, unsigned char *) {
      sched_locate<float>(g, p.sel.data(), p.tables().ranges, n, sched_blocks, 0, sched.data());
    });
    pieces = sched[0];
  }
  auto run = [&](auto h_c) {
    constexpr int HH = decltype(h_c)::value;
    if (sched_mode) {
      cpu_build::run_grid(grid, 1, 1, bt, smem, [&](int b, int, int, unsigned char *s) {
        const int piece = xcd_piece(b, grid);  // field_sched_kernel's block -> piece order
        if (att.attf) field_sched_body<float, RZ, QS, HH, true>(FJ, sched_blocks, piece, s);
        else field_sched_body<float, RZ, QS, HH, false>(FJ, sched_blocks, piece, s);
      });
      return;
    }
    cpu_build::run_grid(n, g.ntiles, p.C, bt, smem, [&](int al, int jt, int c, unsigned char *s) {
      if (att.attf) field_body<float, RZ, QS, HH, false, true>(FJ, al, jt, c, s);
      else field_body<float, RZ, QS, HH, false, false>(FJ, al, jt, c, s);
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
  const size_t w_written = W.size() - cpu_build::count_nan(W);

  // launch_bp_gather: one lane per (x, y, ZV z); the reference mode has no
  // barriers, the exact adjoint runs as 256-lane blocks sharing their
  // columns' contributor lists (bp_gather_adj_kernel)
  constexpr int ZV = gather_zv<float>();  // as launch_bp_gather
  const int zq = (Nz + ZV - 1) / ZV;
  const size_t lanes = (size_t)N * N * zq;
  for (int mode = 0; mode < 2; ++mode) {
    std::vector<float> out((size_t)N * N * Nz, std::numeric_limits<float>::quiet_NaN());
    const GatherJob<float> GJ{g, p.tables(), W.data(), p.sel.data(), n, SPECT_OUT_STORE, out.data(),
                              nullptr, nullptr, 0.0f, 0.0f, nullptr, nullptr, nullptr};
    if (mode == 0) {
      for (size_t gid = 0; gid < lanes; ++gid) gather_body<float, ZV, 0>(GJ, gid);
    } else if (lanes <= ((size_t)1 << 16)) {
      std::vector<unsigned char> asmem(adj_block_smem<float>(zq));
      const int gblocks = (int)((lanes + 255) / 256);
      cpu_build::run_grid(gblocks, 1, 1, 256, asmem, [&](int b, int, int, unsigned char *s) {
        gather_adj_block<float, ZV, 1>(GJ, b, s);
      });
    } else {
      // large problems: the per-lane search of gather_body (the same sums;
      // the block form's barriers between 256 threads per angle would take
      // hours here)
      for (size_t gid = 0; gid < lanes; ++gid) gather_body<float, ZV, 1>(GJ, gid);
    }
    cpu_build::write_file(dir + (mode == 0 ? "/bp_ref.bin" : "/bp_adj.bin"), out);
  }
  if (H == 0 && p.mu.empty()) {
    // the direct plain back projection (backward_impl without a field): the
    // projection itself as a field whose rows all alias one row, gather mode 3
    std::vector<float> out((size_t)N * N * Nz, std::numeric_limits<float>::quiet_NaN());
    const size_t nout = (size_t)(g.jo1 - g.jo0);
    const GatherJob<float> GJ{g, p.tables(), proj.data(), p.sel.data(), n, SPECT_OUT_STORE, out.data(),
                              nullptr, nullptr, 0.0f, 0.0f, nullptr, nullptr, nullptr, nout * Nz, 0};
    for (size_t gid = 0; gid < lanes; ++gid) gather_body<float, ZV, 3>(GJ, gid);
    cpu_build::write_file(dir + "/bp_direct.bin", out);
  }
  std::printf("H %d w_written %zu w_total %zu blocks %d pieces %d\n", H, w_written, W.size(),
              sched_mode ? grid : npairs * p.C, pieces);
  return 0;
}
