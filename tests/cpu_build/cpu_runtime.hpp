// Definitions behind the host stand-in for <hip/hip_runtime.h> (lane index,
// block barrier, __syncthreads_or) and the plan-side inputs of the kernel
// bodies as spect_plan_create builds them (capi.hip): geometry, rotation and
// linspace tables, PSF taps widened to K = 2H+1.  Included once by each CPU
// build executable.
//
// Problem directory: meta.txt "N Nz A L kr kz n_radii n_sel att_mode dx"
// (att_mode 0: none, 1: attenuation field, 2: FP64 suffix table; dx in cm),
// angles.bin float32 [A], sel.bin int32 [n_sel]; with a PSF ridx.bin int32
// [A], taps_r.bin float32 [n_radii][P][kr], taps_z.bin float32
// [n_radii][P][kz]; with attenuation mu.bin float32 [N][N][Nz].
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "activity.hpp"
#include "tiers.hpp"

thread_local cpu_dim3 threadIdx;
cpu_dim3 blockDim;
std::barrier<> *cpu_block_barrier;

namespace {
std::atomic<int> g_or_slot[2];
thread_local unsigned g_or_calls = 0;
}  // namespace

int __syncthreads_or(int pred) {
  const int s = g_or_calls++ & 1;
  if (pred) g_or_slot[s].fetch_or(1);
  cpu_block_barrier->arrive_and_wait();
  if (threadIdx.x == 0) g_or_slot[s ^ 1].store(0);  // the previous call's slot
  const int r = g_or_slot[s].load();
  cpu_block_barrier->arrive_and_wait();
  return r;
}

namespace {
int g_shfl_slot[1024];
}  // namespace

int __builtin_amdgcn_update_dpp(int old, int src, int dpp_ctrl, int row_mask, int bank_mask,
                                bool bound_ctrl) {
  if (row_mask != 0xF || bank_mask != 0xF) {
    std::fprintf(stderr, "DPP row/bank masks other than 0xF are not emulated\n");
    std::abort();
  }
  const unsigned t = threadIdx.x, l = t & 15u, row0 = t & ~15u;
  int from = -1;  // source lane, -1: outside the row
  const int n = dpp_ctrl & 0xF;
  if (dpp_ctrl < 0x100) from = (int)((t & ~3u) | ((unsigned)(dpp_ctrl >> (2 * (t & 3u))) & 3u));
  else if (dpp_ctrl >= 0x101 && dpp_ctrl <= 0x10F) from = (int)l + n <= 15 ? (int)(t + n) : -1;
  else if (dpp_ctrl >= 0x111 && dpp_ctrl <= 0x11F) from = (int)l >= n ? (int)(t - n) : -1;
  else if (dpp_ctrl >= 0x121 && dpp_ctrl <= 0x12F) from = (int)(row0 + ((l - n) & 15u));
  else {
    std::fprintf(stderr, "DPP control 0x%x is not emulated\n", dpp_ctrl);
    std::abort();
  }
  g_shfl_slot[t] = src;
  cpu_block_barrier->arrive_and_wait();
  const int r = from >= 0 ? g_shfl_slot[from] : (bound_ctrl ? 0 : old);
  cpu_block_barrier->arrive_and_wait();
  return r;
}

int __builtin_amdgcn_readlane(int v, int lane) {
  const unsigned t = threadIdx.x;
  g_shfl_slot[t] = v;
  cpu_block_barrier->arrive_and_wait();
  const int r = g_shfl_slot[(t & ~63u) + (unsigned)lane];
  cpu_block_barrier->arrive_and_wait();
  return r;
}

unsigned long long __ballot(int pred) {
  const unsigned t = threadIdx.x, w0 = t & ~63u;
  g_shfl_slot[t] = pred ? 1 : 0;
  cpu_block_barrier->arrive_and_wait();
  unsigned long long m = 0;
  for (unsigned l = 0; l < 64 && w0 + l < blockDim.x; ++l)
    if (g_shfl_slot[w0 + l]) m |= 1ull << l;
  cpu_block_barrier->arrive_and_wait();
  return m;
}

namespace cpu_build {

template <typename T> std::vector<T> read_file(const std::string &path, size_t count) {
  std::vector<T> v(count);
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp || std::fread(v.data(), sizeof(T), count, fp) != count) {
    std::fprintf(stderr, "cannot read %zu items from %s\n", count, path.c_str());
    std::exit(2);
  }
  std::fclose(fp);
  return v;
}

template <typename T> void write_file(const std::string &path, const std::vector<T> &v) {
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp || std::fwrite(v.data(), sizeof(T), v.size(), fp) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(fp);
}

// tile_block_threads (forward.hip)
inline int block_threads(int Nz, int run_len, int splits) {
  const int runs = (Nz + run_len - 1) / run_len;
  const int nzc = (Nz + 63) / 64 * 64;
  int t = spect::kTJ * runs;
  if (splits * nzc > t) t = splits * nzc;
  return (t + 63) / 64 * 64;
}

struct Problem {
  int N, Nz, A, L, kr, kz, n_radii, n, att_mode, pad, P, C, H;
  float dx;
  spect::Geo<float> g;
  std::vector<float2> trig;
  std::vector<float> lin, taps_r, taps_z, mu;
  std::vector<int32_t> ridx, sel, tier;
  std::vector<uint64_t> act;
  std::vector<int2> ranges;
  spect::Tables<float> tables() const {
    return {trig.data(), lin.data(), ridx.data(), H > 0 ? taps_r.data() : nullptr,
            H > 0 ? taps_z.data() : nullptr, H > 0 ? tier.data() : nullptr,
            reinterpret_cast<const unsigned long long *>(act.data()), ranges.data()};
  }
};

inline std::vector<float> widen(const std::vector<float> &src, int rows, int k, int K, bool unit_if_empty) {
  std::vector<float> out((size_t)rows * K, 0.0f);
  const int off = (K - k) / 2;
  for (int r = 0; r < rows; ++r) {
    if (k > 0)
      for (int t = 0; t < k; ++t) out[(size_t)r * K + off + t] = src[(size_t)r * k + t];
    else if (unit_if_empty)
      out[(size_t)r * K + K / 2] = 1.0f;
  }
  return out;
}

inline Problem load_problem(const std::string &dir) {
  Problem p;
  FILE *fm = std::fopen((dir + "/meta.txt").c_str(), "r");
  if (!fm || std::fscanf(fm, "%d %d %d %d %d %d %d %d %d %f", &p.N, &p.Nz, &p.A, &p.L, &p.kr, &p.kz,
                         &p.n_radii, &p.n, &p.att_mode, &p.dx) != 10) {
    std::fprintf(stderr, "bad meta.txt\n");
    std::exit(2);
  }
  std::fclose(fm);
  p.pad = (int)std::ceil((std::sqrt(2.0) * p.N - p.N) / 2.0);
  p.P = p.N + 2 * p.pad;
  p.C = (p.P + p.L - 1) / p.L;
  p.H = 0;
  if (p.kr > 0) {
    p.H = (std::max(p.kr, p.kz) / 2 + 1) / 2 * 2;  // even half widths are compiled
    if (p.H == 0) p.H = 2;
  }
  spect::Geo<float> &g = p.g;
  g.N = p.N; g.Nz = p.Nz; g.P = p.P; g.pad = p.pad; g.A = p.A; g.H = p.H;
  g.jo0 = p.pad; g.jo1 = p.pad + p.N;
  g.je0 = std::max(0, g.jo0 - p.H); g.je1 = std::min(p.P, g.jo1 + p.H);
  g.L = p.L; g.C = p.C;
  g.ntiles = (g.je1 - g.je0 + spect::kTJ - 1) / spect::kTJ;
  g.half = float(p.P - 1) / 2.0f;
  g.dx = p.dx;
  g.lo = float(p.pad) - 2.0f;
  g.hi = float(p.pad + p.N) + 1.0f;
  const auto angles = read_file<float>(dir + "/angles.bin", p.A);
  if (p.att_mode != 0) p.mu = read_file<float>(dir + "/mu.bin", (size_t)p.N * p.N * p.Nz);
  p.sel = read_file<int32_t>(dir + "/sel.bin", p.n);
  p.ridx.assign(p.A, 0);
  if (p.H > 0) {
    const int K = 2 * p.H + 1, rows = p.n_radii * p.P;
    p.ridx = read_file<int32_t>(dir + "/ridx.bin", p.A);
    p.taps_r = widen(read_file<float>(dir + "/taps_r.bin", (size_t)rows * p.kr), rows, p.kr, K, false);
    std::vector<float> tz;
    if (p.kz > 0) tz = read_file<float>(dir + "/taps_z.bin", (size_t)rows * p.kz);
    p.taps_z = widen(tz, rows, p.kz, K, true);
    p.tier = spect::tap_tiers<float>(p.taps_r.data(), p.taps_z.data(), rows, p.H, spect::kTierEps32);
  }
  // theta_f = -(270 - beta) in float32, radians as float32 deg2rad
  p.trig.resize(p.A);
  const float pi32 = (float)M_PI;
  for (int a = 0; a < p.A; ++a) {
    const float th = -(270.0f - angles[a]);
    const float rad = (th * pi32) / 180.0f;
    p.trig[a] = make_float2((float)std::cos((double)rad), (float)std::sin((double)rad));
  }
  // torch.linspace(-1, 1, P) as ATen computes it in float32
  p.lin.resize(p.P);
  const float step = 2.0f / float(p.P - 1);
  const int halfway = p.P / 2;
  for (int k = 0; k < p.P; ++k)
    p.lin[k] = k < halfway ? -1.0f + step * float(k) : 1.0f - step * float(p.P - k - 1);
  // planes each forward block visits (as spect_plan_create tabulates them)
  {
    std::vector<double> cv(p.A), sv(p.A), l64(p.P);
    for (int a = 0; a < p.A; ++a) {
      const double r64 = -(270.0 - (double)angles[a]) * M_PI / 180.0;
      cv[a] = std::cos(r64);
      sv[a] = std::sin(r64);
    }
    const double st64 = 2.0 / double(p.P - 1);
    for (int k = 0; k < p.P; ++k)
      l64[k] = k < halfway ? -1.0 + st64 * double(k) : 1.0 - st64 * double(p.P - k - 1);
    std::vector<int> rng;
    p.act = spect::forward_activity(cv.data(), sv.data(), p.A, l64.data(), p.P, g.je0, g.je1, p.L,
                                    p.C, spect::kTJ, (double)g.lo, (double)g.hi, &rng);
    p.ranges.resize(rng.size() / 2);
    for (size_t k = 0; k < p.ranges.size(); ++k) p.ranges[k] = int2{rng[2 * k], rng[2 * k + 1]};
  }
  return p;
}

// Runs body(al, jt, c) for every block of a (n, ntiles, C) grid with bt
// lanes, one block at a time; dynamic LDS is refilled with NaN per block.
template <typename Body>
void run_grid(int n, int ntiles, int C, int bt, std::vector<unsigned char> &smem, Body body) {
  std::barrier<> bar(bt);
  cpu_block_barrier = &bar;
  blockDim.x = bt;
  g_or_slot[0].store(0);  // fresh lanes start their __syncthreads_or count at 0
  g_or_slot[1].store(0);
  const int nblocks = n * ntiles * C;
  std::vector<std::thread> lanes;
  for (int t = 0; t < bt; ++t)
    lanes.emplace_back([&, t] {
      threadIdx.x = t;
      for (int b = 0; b < nblocks; ++b) {
        if (t == 0) {
          float *s = reinterpret_cast<float *>(smem.data());
          for (size_t q = 0; q < smem.size() / sizeof(float); ++q) s[q] = std::numeric_limits<float>::quiet_NaN();
        }
        bar.arrive_and_wait();
        const int al = b / (ntiles * C), jt = (b / C) % ntiles, c = b % C;
        body(al, jt, c, smem.data());
        bar.arrive_and_wait();
      }
    });
  for (auto &l : lanes) l.join();
}

inline size_t count_nan(const std::vector<float> &v) {
  size_t k = 0;
  for (float x : v) k += std::isnan(x) ? 1 : 0;
  return k;
}

}  // namespace cpu_build
