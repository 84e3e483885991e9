// forward.hip -- attenuation table and forward projection kernels.
//
// SPECTSystemMatrix.forward (spect_system_matrix.py:120-157) with
// SPECTAttenuationTransform.forward (attenuation.py:72-92) and
// SPECTPSFTransform.forward (psf.py:199-214, Seperable1DBlurNet psf.py:27-41).
//
// One block = (angle, tile of kTJ rotated-frame columns, chunk of L planes).
// The block walks its planes from the detector side (high i) down.  Per plane:
//   A  gather: V(j,z) = a(i,j,z) * f_rot(i,j,z) for the tile's columns; each
//      thread owns one column and a run of RZ z values in registers; the
//      attenuation a comes from an FP64 running sum of mu_rot*dx that starts at
//      the chunk's suffix sum (att table) -- the reverse cumsum of
//      utils/misc.py:18 in the order a sequential scan would add.
//   B  z blur with the plane's taps: register window (own run + halo read from
//      LDS).
//   C  y blur written as a scatter: each thread owns one z and half of the
//      tile's kTJ + 2H output columns and accumulates, in registers across all
//      planes, sum_t w[t] * Vz(j + t - H); the ray sum is that accumulation.
// Planes where no column of the tile samples the object are skipped.  Each
// block writes its kTJ + 2H output columns; a reduce kernel adds chunks and
// overlapping tiles in a fixed order (deterministic, no atomics).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "forward_body.hpp"

namespace spect {

int tile_block_threads(int Nz, int run_len, int splits) {
  const int runs = (Nz + run_len - 1) / run_len;
  const int nzc = (Nz + 63) / 64 * 64;
  int t = kTJ * runs;
  if (splits * nzc > t) t = splits * nzc;
  return (t + 63) / 64 * 64;
}

namespace {

template <typename R, int RZ>
__global__ __launch_bounds__(1024) void att_table_kernel(Geo<R> g, Tables<R> tb,
                                                        const R *__restrict__ mu,
                                                        const int32_t *__restrict__ angle_idx,
                                                        double *__restrict__ table) {
  att_table_body<R, RZ>(g, tb, mu, angle_idx, table, blockIdx.x, blockIdx.y);
}

template <typename R, int RZ>
__global__ __launch_bounds__(1024) void att_field_kernel(Geo<R> g, Tables<R> tb,
                                                        const R *__restrict__ mu,
                                                        const int32_t *__restrict__ angle_idx,
                                                        R *__restrict__ field) {
  att_field_body<R, RZ>(g, tb, mu, angle_idx, field, blockIdx.x, blockIdx.y);
}

template <typename R, int RZ, int QS, int H, bool TRB, bool AF>
__global__ __launch_bounds__(512) void fwd_kernel(FwdJob<R> J) {
  extern __shared__ __align__(16) unsigned char smem[];
  // tiles and chunks are taken centre-out: the heaviest blocks are dispatched first
  const int jt = centre_out(blockIdx.y, J.g.ntiles);
  fwd_body<R, RZ, QS, H, TRB, AF>(J, blockIdx.x, jt,
                                  chunk_piece(J.g, J.n, blockIdx.x, jt, centre_out(blockIdx.z, J.g.C)),
                                  smem);
}

// The forward blocks of the balanced schedule (fwd_sched_body), each locating
// its piece for nblocks resident blocks; block 0 writes the schedule to sched.
template <typename R, int RZ, int QS, int H, bool AF>
__global__ __launch_bounds__(512) void fwd_sched_kernel(FwdJob<R> J, int *__restrict__ sched, int nblocks) {
  extern __shared__ __align__(16) unsigned char smem[];
  fwd_sched_body<R, RZ, QS, H, AF>(J, J.tb.ranges, sched, nblocks, xcd_piece(blockIdx.x, gridDim.x), smem);
}

template <typename R>
__global__ __launch_bounds__(256) void reduce_sched_kernel(ReduceJob<R> J, const int *__restrict__ sched) {
  reduce_sched_body<R>(J, sched, (size_t)blockIdx.x * 256 + threadIdx.x);
}

// Batched: block b belongs to job find_job(b); within a job the blocks run
// over (angle, tile) fastest and then over the chunks centre-out.
template <typename R, int RZ, int QS, int H, bool AF>
__global__ __launch_bounds__(512) void fwd_batch_kernel(const FwdJob<R> *__restrict__ jobs,
                                                        const int *__restrict__ starts, int njobs) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const int j = find_job(starts, njobs, b);
  const FwdJob<R> J = ld_uniform_struct(jobs + j);
  const int local = b - ld_uniform(starts + j);
  const int per_chunk = J.n * J.g.ntiles;
  const int c = centre_out(local / per_chunk, J.g.C), r = local % per_chunk;
  const int al = r % J.n, jt = r / J.n;
  fwd_body<R, RZ, QS, H, false, AF>(J, al, jt, chunk_piece(J.g, J.n, al, jt, c), smem);
}

template <typename R>
__global__ __launch_bounds__(256) void fwd_reduce_kernel(ReduceJob<R> J) {
  reduce_body<R>(J, (size_t)blockIdx.x * 256 + threadIdx.x);
}

template <typename R>
__global__ __launch_bounds__(256) void reduce_batch_kernel(const ReduceJob<R> *__restrict__ jobs,
                                                           const int *__restrict__ starts,
                                                           int njobs) {
  const int b = blockIdx.x;
  const int j = find_job(starts, njobs, b);
  const ReduceJob<R> J = ld_uniform_struct(jobs + j);
  reduce_body<R>(J, (size_t)(b - ld_uniform(starts + j)) * 256 + threadIdx.x);
}

template <typename R, int H>
hipError_t launch_fwd_h(const Geo<R> &g, const Tables<R> &tb, const R *f, const R *mu,
                        const double *att, bool att_global, const R *attf, const int32_t *idx,
                        int n, R *partial, bool trb, hipStream_t st) {
  constexpr int RZ = FwdTile<R>::RZ, QS = FwdTile<R>::QS;
  const int bt = tile_block_threads(g.Nz, RZ, QS);
  if (bt > 512) return hipErrorInvalidConfiguration;  // __launch_bounds__(512)
  if (!phase_a_fits<R, QS>(bt, g.Nz)) return hipErrorInvalidConfiguration;
  const size_t smem = fwd_smem_bytes<R, RZ, H>(g.Nz, g.L);
  auto kern = attf ? (trb ? fwd_kernel<R, RZ, QS, H, true, true> : fwd_kernel<R, RZ, QS, H, false, true>)
                   : (trb ? fwd_kernel<R, RZ, QS, H, true, false> : fwd_kernel<R, RZ, QS, H, false, false>);
  if (smem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  FwdJob<R> J{g, tb, f, mu, att, att_global ? 1 : 0, attf, idx, n, partial};
  dim3 grid(n, g.ntiles, g.C);
  kern<<<grid, bt, smem, st>>>(J);
  return hipGetLastError();
}

template <typename R, int H>
hipError_t launch_fwd_batch_h(const FwdJob<R> *jobs, const int *starts, int njobs, int blocks,
                              int Nz, int max_L, bool af, hipStream_t st) {
  constexpr int RZ = FwdTile<R>::RZ, QS = FwdTile<R>::QS;
  const int bt = tile_block_threads(Nz, RZ, QS);
  if (bt > 512) return hipErrorInvalidConfiguration;  // __launch_bounds__(512)
  if (!phase_a_fits<R, QS>(bt, Nz)) return hipErrorInvalidConfiguration;
  const size_t smem = fwd_smem_bytes<R, RZ, H>(Nz, max_L);
  auto kern = af ? fwd_batch_kernel<R, RZ, QS, H, true> : fwd_batch_kernel<R, RZ, QS, H, false>;
  if (smem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  kern<<<blocks, bt, smem, st>>>(jobs, starts, njobs);
  return hipGetLastError();
}

// Blocks of fwd_sched_kernel resident on the device at once (occupancy x
// CUs), cached per kernel instance, z extent and device.
template <typename R, int H, bool AF>
int sched_resident_blocks(int Nz, int L) {
  constexpr int RZ = FwdTile<R>::RZ, QS = FwdTile<R>::QS;
  static std::mutex mu;
  static std::map<std::tuple<int, int, int>, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({dev, Nz, L});
  if (it != cache.end()) return it->second;
  const int bt = tile_block_threads(Nz, RZ, QS);
  const size_t smem = fwd_smem_bytes<R, RZ, H>(Nz, L);
  auto kern = fwd_sched_kernel<R, RZ, QS, H, AF>;
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
int sched_blocks_h(int Nz, int L, bool af) {
  return af ? sched_resident_blocks<R, H, true>(Nz, L) : sched_resident_blocks<R, H, false>(Nz, L);
}

template <typename R, int H>
hipError_t launch_fwd_sched_h(const Geo<R> &g, const Tables<R> &tb, const R *f, const R *attf,
                              const int32_t *idx, int n, R *partial, int *sched, int nb,
                              hipStream_t st) {
  constexpr int RZ = FwdTile<R>::RZ, QS = FwdTile<R>::QS;
  const int bt = tile_block_threads(g.Nz, RZ, QS);
  if (bt > 512) return hipErrorInvalidConfiguration;  // __launch_bounds__(512)
  if (!phase_a_fits<R, QS>(bt, g.Nz)) return hipErrorInvalidConfiguration;
  if (nb <= 0) return hipErrorInvalidConfiguration;
  const int grid = sched_grid_blocks(n, g.ntiles, g.P, nb);
  const size_t smem = fwd_smem_bytes<R, RZ, H>(g.Nz, g.L);
  auto kern = attf ? fwd_sched_kernel<R, RZ, QS, H, true> : fwd_sched_kernel<R, RZ, QS, H, false>;
  if (smem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  FwdJob<R> J{g, tb, f, nullptr, nullptr, 0, attf, idx, n, partial};
  kern<<<grid, bt, smem, st>>>(J, sched, nb);
  return hipGetLastError();
}

}  // namespace

template <typename R>
int forward_sched_blocks(const Geo<R> &g) {
#define SPECT_FS_CASE(HH) \
  case HH: return std::max(sched_blocks_h<R, HH>(g.Nz, g.L, false), sched_blocks_h<R, HH>(g.Nz, g.L, true));
  switch (g.H) {
    SPECT_FS_CASE(0) SPECT_FS_CASE(2) SPECT_FS_CASE(4) SPECT_FS_CASE(6) SPECT_FS_CASE(8)
    SPECT_FS_CASE(10) SPECT_FS_CASE(12) SPECT_FS_CASE(14) SPECT_FS_CASE(16) SPECT_FS_CASE(18)
    SPECT_FS_CASE(20)
    default: return 0;
  }
#undef SPECT_FS_CASE
}

template <typename R>
hipError_t launch_forward_sched(const Geo<R> &g, const Tables<R> &tb, const R *f, const R *attf,
                                const int32_t *idx, int n, R *partial, int *sched, int blocks,
                                hipStream_t st) {
#define SPECT_FSL_CASE(HH) \
  case HH: return launch_fwd_sched_h<R, HH>(g, tb, f, attf, idx, n, partial, sched, blocks, st);
  switch (g.H) {
    SPECT_FSL_CASE(0) SPECT_FSL_CASE(2) SPECT_FSL_CASE(4) SPECT_FSL_CASE(6) SPECT_FSL_CASE(8)
    SPECT_FSL_CASE(10) SPECT_FSL_CASE(12) SPECT_FSL_CASE(14) SPECT_FSL_CASE(16)
    SPECT_FSL_CASE(18) SPECT_FSL_CASE(20)
    default: return hipErrorInvalidValue;
  }
#undef SPECT_FSL_CASE
}

template <typename R>
hipError_t launch_reduce_sched(const Geo<R> &g, const R *partial, int n, const int *sched, R *proj,
                               const R *gm, const R *s, R *pred, R *ratio, R delta, hipStream_t st) {
  ReduceJob<R> J{g, partial, n, proj, gm, s, pred, ratio, delta};
  reduce_sched_kernel<R><<<reduce_blocks(g, n), 256, 0, st>>>(J, sched);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_att_table(const Geo<R> &g, const Tables<R> &tb, const R *mu,
                            const int32_t *angle_idx, int n, double *table, hipStream_t st) {
  constexpr int RZ = FwdTile<R>::RZ;
  const int runs = (g.Nz + RZ - 1) / RZ;
  const int bt = (kTJ * runs + 63) / 64 * 64;
  dim3 grid(n, g.ntiles);
  att_table_kernel<R, RZ><<<grid, bt, 0, st>>>(g, tb, mu, angle_idx, table);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_att_field(const Geo<R> &g, const Tables<R> &tb, const R *mu,
                            const int32_t *angle_idx, int n, R *field, hipStream_t st) {
  constexpr int RZ = FwdTile<R>::RZ;
  const int runs = (g.Nz + RZ - 1) / RZ;
  const int bt = (kTJ * runs + 63) / 64 * 64;
  dim3 grid(n, g.ntiles);
  att_field_kernel<R, RZ><<<grid, bt, 0, st>>>(g, tb, mu, angle_idx, field);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_forward(const Geo<R> &g, const Tables<R> &tb, const R *f, const R *mu,
                          const double *att, bool att_global, const R *attf, const int32_t *idx,
                          int n, R *partial, bool trb, hipStream_t st) {
#define SPECT_F_CASE(HH) \
  case HH: return launch_fwd_h<R, HH>(g, tb, f, mu, att, att_global, attf, idx, n, partial, trb, st);
  switch (g.H) {
    SPECT_F_CASE(0) SPECT_F_CASE(2) SPECT_F_CASE(4) SPECT_F_CASE(6) SPECT_F_CASE(8)
    SPECT_F_CASE(10) SPECT_F_CASE(12) SPECT_F_CASE(14) SPECT_F_CASE(16) SPECT_F_CASE(18)
    SPECT_F_CASE(20)
    default: return hipErrorInvalidValue;
  }
#undef SPECT_F_CASE
}

template <typename R>
hipError_t launch_forward_reduce(const Geo<R> &g, const R *partial, int n, R *proj, const R *gm,
                                 const R *s, R *pred, R *ratio, R delta, hipStream_t st) {
  ReduceJob<R> J{g, partial, n, proj, gm, s, pred, ratio, delta};
  fwd_reduce_kernel<R><<<reduce_blocks(g, n), 256, 0, st>>>(J);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_forward_batch(const FwdJob<R> *jobs, const int *starts, int njobs, int blocks,
                                int H, int Nz, int max_L, bool af, hipStream_t st) {
#define SPECT_FB_CASE(HH) \
  case HH: return launch_fwd_batch_h<R, HH>(jobs, starts, njobs, blocks, Nz, max_L, af, st);
  switch (H) {
    SPECT_FB_CASE(0) SPECT_FB_CASE(2) SPECT_FB_CASE(4) SPECT_FB_CASE(6) SPECT_FB_CASE(8)
    SPECT_FB_CASE(10) SPECT_FB_CASE(12) SPECT_FB_CASE(14) SPECT_FB_CASE(16) SPECT_FB_CASE(18)
    SPECT_FB_CASE(20)
    default: return hipErrorInvalidValue;
  }
#undef SPECT_FB_CASE
}

template <typename R>
hipError_t launch_reduce_batch(const ReduceJob<R> *jobs, const int *starts, int njobs, int blocks,
                               hipStream_t st) {
  reduce_batch_kernel<R><<<blocks, 256, 0, st>>>(jobs, starts, njobs);
  return hipGetLastError();
}

#define SPECT_INST(R)                                                                          \
  template hipError_t launch_att_table<R>(const Geo<R> &, const Tables<R> &, const R *,        \
                                          const int32_t *, int, double *, hipStream_t);        \
  template hipError_t launch_att_field<R>(const Geo<R> &, const Tables<R> &, const R *,        \
                                          const int32_t *, int, R *, hipStream_t);             \
  template hipError_t launch_forward<R>(const Geo<R> &, const Tables<R> &, const R *,          \
                                        const R *, const double *, bool, const R *,            \
                                        const int32_t *, int, R *, bool, hipStream_t);         \
  template hipError_t launch_forward_reduce<R>(const Geo<R> &, const R *, int, R *, const R *, \
                                               const R *, R *, R *, R, hipStream_t);           \
  template hipError_t launch_forward_batch<R>(const FwdJob<R> *, const int *, int, int, int,    \
                                              int, int, bool, hipStream_t);                    \
  template hipError_t launch_reduce_batch<R>(const ReduceJob<R> *, const int *, int, int,       \
                                             hipStream_t);                                     \
  template int forward_sched_blocks<R>(const Geo<R> &);                                        \
  template hipError_t launch_forward_sched<R>(const Geo<R> &, const Tables<R> &, const R *,     \
                                              const R *, const int32_t *, int, R *, int *, int, \
                                              hipStream_t);                                    \
  template hipError_t launch_reduce_sched<R>(const Geo<R> &, const R *, int, const int *, R *,  \
                                             const R *, const R *, R *, R *, R, hipStream_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
