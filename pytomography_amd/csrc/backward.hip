// backward.hip -- back projection kernels.
//
// SPECTSystemMatrix.backward (spect_system_matrix.py:159-201): per angle the
// padded projection is smeared along x (:180,191), blurred by the PSF
// (psf.py:216-231), multiplied by the attenuation factor (attenuation.py:
// 94-115), rotated back with kornia rotate(270 - beta) (:196) and summed over
// the angles, then cropped (:200).
//
// bp_field_kernel: one block = (angle, tile of kTJ rotated-frame columns,
//   chunk of L planes); writes W(i,j,z) = a(i,j,z) * (G_i * p)(j,z).  The y
//   blur reads a window of the projection that each thread keeps in
//   registers for all planes; the z blur uses a register window with an LDS
//   halo; a is the FP64 reverse-cumsum attenuation as in the forward kernel.
//   Planes whose tile lies outside the (1.5-pixel widened) footprint are
//   skipped: the gathers below never read them.
// bp_gather_kernel<MODE 0>: W gathered at theta_b = 270 - beta, summed over
//   the angles in registers, cropped, then stored / added / used for the OSEM
//   update.
// bp_gather_kernel<MODE 3>: MODE 0 on a field constant along the rows (the
//   plain back projection of the projection itself).
// bp_gather_kernel<MODE 1>: the exact transpose of the forward's bilinear
//   gather, written as a gather over the <= 16 rotated-frame candidates of
//   each voxel (race-free and deterministic); MODE 2 adds the mu gradient.
// The block bodies (field_body, gather_body) are in backward_body.hpp.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "backward_body.hpp"

namespace spect {
namespace {

template <typename R, int RZ, int QS, int H, bool GRAD, bool AF>
__global__ __launch_bounds__(1024) void bp_field_kernel(FieldJob<R> J) {
  extern __shared__ __align__(16) unsigned char smem[];
  field_body<R, RZ, QS, H, GRAD, AF>(J, blockIdx.x, blockIdx.y, blockIdx.z, smem);
}

template <typename R, int RZ, int QS, int H, bool AF>
__global__ __launch_bounds__(1024) void field_batch_kernel(const FieldJob<R> *__restrict__ jobs,
                                                           const int *__restrict__ starts,
                                                           int njobs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const int j = find_job(starts, njobs, b);
  const FieldJob<R> J = ld_uniform_struct(jobs + j);
  const int local = b - ld_uniform(starts + j);
  const int c = local % J.g.C, r = local / J.g.C;
  field_body<R, RZ, QS, H, false, AF>(J, r / J.g.ntiles, r % J.g.ntiles, c, smem);
}

template <typename R, int RZ, int QS, int H, bool AF>
__global__ __launch_bounds__(1024) void field_sched_kernel(FieldJob<R> J, int nblocks) {
  extern __shared__ __align__(16) unsigned char smem[];
  field_sched_body<R, RZ, QS, H, AF>(J, nblocks, xcd_piece(blockIdx.x, gridDim.x), smem);
}

// In place: Dsum[a][c] <- sum_{c' <= c} Dsum[a][c'] (prefix over chunks, FP64).
__global__ __launch_bounds__(256) void chunk_prefix_kernel(double *__restrict__ Dsum, int n, int C,
                                                           size_t plane) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)n * plane) return;
  const size_t a = idx / plane, r = idx - a * plane;
  double s = 0.0;
  for (int c = 0; c < C; ++c) {
    double &v = Dsum[(a * C + c) * plane + r];
    s += v;
    v = s;
  }
}

template <typename R, int ZV, int MODE>
__global__ __launch_bounds__(256) void bp_gather_kernel(GatherJob<R> J) {
  gather_body<R, ZV, MODE>(J, (size_t)blockIdx.x * 256 + threadIdx.x);
}

// Exact-adjoint modes with the block's contributor lists in LDS
// (gather_adj_block; dynamic LDS adj_block_smem).
template <typename R, int ZV, int MODE>
__global__ __launch_bounds__(256) void bp_gather_adj_kernel(GatherJob<R> J) {
  extern __shared__ __align__(16) unsigned char smem[];
  gather_adj_block<R, ZV, MODE>(J, blockIdx.x, smem);
}

template <typename R, int ZV, int MODE>
__global__ __launch_bounds__(256) void gather_batch_kernel(const GatherJob<R> *__restrict__ jobs,
                                                           const int *__restrict__ starts,
                                                           int njobs) {
  const int b = blockIdx.x;
  const int j = find_job(starts, njobs, b);
  const GatherJob<R> J = ld_uniform_struct(jobs + j);
  gather_body<R, ZV, MODE>(J, (size_t)(b - ld_uniform(starts + j)) * 256 + threadIdx.x);
}

template <typename R, int H>
hipError_t launch_field_h(const Geo<R> &g, const Tables<R> &tb, const R *proj, const R *mu,
                          const double *att, bool att_global, const R *attf, const int32_t *idx,
                          int n, R *W, const R *f, double *Wm, double *Dsum, hipStream_t st) {
  constexpr int RZ = BpTile<R>::RZ, QS = BpTile<R>::QS;
  const int bt = tile_block_threads(g.Nz, RZ, QS);
  if (bt > 1024) return hipErrorInvalidConfiguration;  // __launch_bounds__(1024)
  const int runs = (g.Nz + RZ - 1) / RZ;
  const size_t smem = ((size_t)kTJ * (runs * RZ + 2 * H + 1) + (size_t)g.L * 2 * (2 * H + 1)) * sizeof(R);
  dim3 grid(n, g.ntiles, g.C);
  if (Wm) {
    if (attf) return hipErrorInvalidValue;  // the gradient field gathers mu itself
    auto kern = bp_field_kernel<R, RZ, QS, H, true, false>;
    if (smem > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return e;
    }
    kern<<<grid, bt, smem, st>>>(
        FieldJob<R>{g, tb, proj, mu, att, att_global ? 1 : 0, nullptr, idx, n, W, f, Wm, Dsum});
  } else {
    auto kern = attf ? bp_field_kernel<R, RZ, QS, H, false, true> : bp_field_kernel<R, RZ, QS, H, false, false>;
    if (smem > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kern,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return e;
    }
    kern<<<grid, bt, smem, st>>>(
        FieldJob<R>{g, tb, proj, mu, att, att_global ? 1 : 0, attf, idx, n, W, nullptr, nullptr, nullptr});
  }
  return hipGetLastError();
}

// Dynamic LDS of the field kernels: U ([kTJ][stride]) and the taps of up
// to max(L, kPieceMax) planes.
template <typename R, int RZ, int H> size_t field_smem_bytes(int Nz, int L) {
  const int runs = (Nz + RZ - 1) / RZ;
  const int planes = L > kPieceMax ? L : kPieceMax;
  return ((size_t)kTJ * (runs * RZ + 2 * H + 1) + (size_t)planes * 2 * (2 * H + 1)) * sizeof(R);
}

template <typename R, int H, bool AF>
int field_resident_blocks(int Nz, int L) {
  constexpr int RZ = BpTile<R>::RZ, QS = BpTile<R>::QS;
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({dev, Nz, L});
  if (it != cache.end()) return it->second;
  const int bt = tile_block_threads(Nz, RZ, QS);
  const size_t smem = field_smem_bytes<R, RZ, H>(Nz, L);
  auto kern = field_sched_kernel<R, RZ, QS, H, AF>;
  if (smem > 64 * 1024 &&
      hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)smem) != hipSuccess)
    return 0;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, bt, smem) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const int nb = per_cu > 0 && cus > 0 ? per_cu * cus : 0;
  cache[{dev, Nz, L}] = nb;
  return nb;
}

template <typename R, int H>
hipError_t launch_field_sched_h(const Geo<R> &g, const Tables<R> &tb, const R *proj, const R *attf,
                                const int32_t *idx, int n, R *W, int nb, hipStream_t st) {
  constexpr int RZ = BpTile<R>::RZ, QS = BpTile<R>::QS;
  const int bt = tile_block_threads(g.Nz, RZ, QS);
  if (bt > 1024) return hipErrorInvalidConfiguration;  // __launch_bounds__(1024)
  if (nb <= 0) return hipErrorInvalidConfiguration;
  const size_t smem = field_smem_bytes<R, RZ, H>(g.Nz, g.L);
  auto kern = attf ? field_sched_kernel<R, RZ, QS, H, true> : field_sched_kernel<R, RZ, QS, H, false>;
  if (smem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  const int grid = sched_grid_blocks(n, g.ntiles, g.P, nb);
  kern<<<grid, bt, smem, st>>>(
      FieldJob<R>{g, tb, proj, nullptr, nullptr, 0, attf, idx, n, W, nullptr, nullptr, nullptr}, nb);
  return hipGetLastError();
}

}  // namespace

template <typename R>
int field_sched_blocks(const Geo<R> &g) {
#define SPECT_FSB_CASE(HH)                                                                   \
  case HH:                                                                                    \
    return std::max(field_resident_blocks<R, HH, false>(g.Nz, g.L),                         \
                    field_resident_blocks<R, HH, true>(g.Nz, g.L));
  switch (g.H) {
    SPECT_FSB_CASE(0) SPECT_FSB_CASE(2) SPECT_FSB_CASE(4) SPECT_FSB_CASE(6) SPECT_FSB_CASE(8)
    SPECT_FSB_CASE(10) SPECT_FSB_CASE(12) SPECT_FSB_CASE(14) SPECT_FSB_CASE(16)
    SPECT_FSB_CASE(18) SPECT_FSB_CASE(20)
    default: return 0;
  }
#undef SPECT_FSB_CASE
}

template <typename R>
hipError_t launch_bp_field_sched(const Geo<R> &g, const Tables<R> &tb, const R *proj,
                                 const R *attf, const int32_t *idx, int n, R *W, int blocks,
                                 hipStream_t st) {
#define SPECT_FSL_CASE(HH) \
  case HH: return launch_field_sched_h<R, HH>(g, tb, proj, attf, idx, n, W, blocks, st);
  switch (g.H) {
    SPECT_FSL_CASE(0) SPECT_FSL_CASE(2) SPECT_FSL_CASE(4) SPECT_FSL_CASE(6) SPECT_FSL_CASE(8)
    SPECT_FSL_CASE(10) SPECT_FSL_CASE(12) SPECT_FSL_CASE(14) SPECT_FSL_CASE(16)
    SPECT_FSL_CASE(18) SPECT_FSL_CASE(20)
    default: return hipErrorInvalidValue;
  }
#undef SPECT_FSL_CASE
}

template <typename R>
hipError_t launch_bp_field(const Geo<R> &g, const Tables<R> &tb, const R *proj, const R *mu,
                           const double *att, bool att_global, const R *attf, const int32_t *idx,
                           int n, R *W, const R *f, double *Wm, double *Dsum, hipStream_t st) {
#define SPECT_FIELD_CASE(HH) \
  case HH:                   \
    return launch_field_h<R, HH>(g, tb, proj, mu, att, att_global, attf, idx, n, W, f, Wm, Dsum, st);
  switch (g.H) {
    SPECT_FIELD_CASE(0) SPECT_FIELD_CASE(2) SPECT_FIELD_CASE(4) SPECT_FIELD_CASE(6)
    SPECT_FIELD_CASE(8) SPECT_FIELD_CASE(10) SPECT_FIELD_CASE(12) SPECT_FIELD_CASE(14)
    SPECT_FIELD_CASE(16) SPECT_FIELD_CASE(18) SPECT_FIELD_CASE(20)
    default: return hipErrorInvalidValue;
  }
#undef SPECT_FIELD_CASE
}

template <typename R>
hipError_t launch_chunk_prefix(const Geo<R> &g, double *Dsum, int n, hipStream_t st) {
  const size_t plane = (size_t)(g.je1 - g.je0) * g.Nz;
  const size_t total = (size_t)n * plane;
  chunk_prefix_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(Dsum, n, g.C, plane);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_bp_gather(const Geo<R> &g, const Tables<R> &tb, const R *W,
                            const int32_t *idx, int n, int mode, int out_mode, R *out, const R *f,
                            const R *nrm, R delta, R floor_value, const double *Wm,
                            const double *Pref, R *out_mu, hipStream_t st, size_t w_astride,
                            size_t w_rstride) {
  constexpr int ZV = gather_zv<R>();
  const unsigned blocks = (unsigned)gather_blocks(g);
  const GatherJob<R> J{g,  tb,      W,      idx,     n,          out_mode, out, f, nrm,
                       delta, floor_value, Wm, Pref, out_mu, w_astride, w_rstride};
  if (mode == SPECT_BP_REFERENCE && w_astride != 0 && w_rstride == 0)
    bp_gather_kernel<R, ZV, 3><<<blocks, 256, 0, st>>>(J);  // rows alias one projection row
  else if (mode == SPECT_BP_REFERENCE)
    bp_gather_kernel<R, ZV, 0><<<blocks, 256, 0, st>>>(J);
  else {
    const size_t smem = adj_block_smem<R>((g.Nz + ZV - 1) / ZV);  // <= 50 KB (Nz <= ZV)
    if (mode == SPECT_BP_ADJOINT)
      bp_gather_adj_kernel<R, ZV, 1><<<blocks, 256, smem, st>>>(J);
    else
      bp_gather_adj_kernel<R, ZV, 2><<<blocks, 256, smem, st>>>(J);
  }
  return hipGetLastError();
}

template <typename R, int H>
hipError_t launch_field_batch_h(const FieldJob<R> *jobs, const int *starts, int njobs, int blocks,
                                int Nz, int max_L, bool af, hipStream_t st) {
  constexpr int RZ = BpTile<R>::RZ, QS = BpTile<R>::QS;
  const int bt = tile_block_threads(Nz, RZ, QS);
  if (bt > 1024) return hipErrorInvalidConfiguration;  // __launch_bounds__(1024)
  const int runs = (Nz + RZ - 1) / RZ;
  const size_t smem = ((size_t)kTJ * (runs * RZ + 2 * H + 1) + (size_t)max_L * 2 * (2 * H + 1)) * sizeof(R);
  auto kern = af ? field_batch_kernel<R, RZ, QS, H, true> : field_batch_kernel<R, RZ, QS, H, false>;
  if (smem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  kern<<<blocks, bt, smem, st>>>(jobs, starts, njobs);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_field_batch(const FieldJob<R> *jobs, const int *starts, int njobs, int blocks,
                              int H, int Nz, int max_L, bool af, hipStream_t st) {
#define SPECT_FLB_CASE(HH) \
  case HH: return launch_field_batch_h<R, HH>(jobs, starts, njobs, blocks, Nz, max_L, af, st);
  switch (H) {
    SPECT_FLB_CASE(0) SPECT_FLB_CASE(2) SPECT_FLB_CASE(4) SPECT_FLB_CASE(6) SPECT_FLB_CASE(8)
    SPECT_FLB_CASE(10) SPECT_FLB_CASE(12) SPECT_FLB_CASE(14) SPECT_FLB_CASE(16)
    SPECT_FLB_CASE(18) SPECT_FLB_CASE(20)
    default: return hipErrorInvalidValue;
  }
#undef SPECT_FLB_CASE
}

template <typename R>
hipError_t launch_gather_batch(const GatherJob<R> *jobs, const int *starts, int njobs, int blocks,
                               int mode, hipStream_t st) {
  constexpr int ZV = gather_zv<R>();
  if (mode == SPECT_BP_REFERENCE)
    gather_batch_kernel<R, ZV, 0><<<blocks, 256, 0, st>>>(jobs, starts, njobs);
  else
    gather_batch_kernel<R, ZV, 1><<<blocks, 256, 0, st>>>(jobs, starts, njobs);
  return hipGetLastError();
}

#define SPECT_INST(R)                                                                          \
  template hipError_t launch_bp_field<R>(const Geo<R> &, const Tables<R> &, const R *,         \
                                         const R *, const double *, bool, const R *,            \
                                         const int32_t *, int, R *, const R *, double *,        \
                                         double *, hipStream_t);                                \
  template hipError_t launch_chunk_prefix<R>(const Geo<R> &, double *, int, hipStream_t);      \
  template int field_sched_blocks<R>(const Geo<R> &);                                          \
  template hipError_t launch_bp_field_sched<R>(const Geo<R> &, const Tables<R> &, const R *,    \
                                               const R *, const int32_t *, int, R *, int,       \
                                               hipStream_t);                                   \
  template hipError_t launch_field_batch<R>(const FieldJob<R> *, const int *, int, int, int,    \
                                            int, int, bool, hipStream_t);                      \
  template hipError_t launch_gather_batch<R>(const GatherJob<R> *, const int *, int, int, int,  \
                                             hipStream_t);                                     \
  template hipError_t launch_bp_gather<R>(const Geo<R> &, const Tables<R> &, const R *,        \
                                          const int32_t *, int, int, int, R *, const R *,      \
                                          const R *, R, R, const double *, const double *, R *, \
                                          hipStream_t, size_t, size_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
