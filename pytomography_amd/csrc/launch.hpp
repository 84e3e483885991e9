// launch.hpp -- host-side launchers of the SPECT kernels (one per translation unit).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.hpp"

namespace spect {

// Work descriptors of one problem for each kernel.  A single-problem launch
// passes one by value; a batched launch reads an array of them from device
// memory (one per problem) and maps its blocks onto them.
template <typename R> struct FwdJob {
  Geo<R> g;
  Tables<R> tb;
  const R *f;
  const R *mu;
  const double *att;
  int att_global;
  const R *attf;  // precomputed attenuation factors [A][P][je1-je0][Nz] (or null)
  const int32_t *angle_idx;
  int n;
  R *partial;
};

template <typename R> struct ReduceJob {
  Geo<R> g;
  const R *partial;
  int n;
  R *proj;
  const R *gm;
  const R *s;
  R *pred;
  R *ratio;
  R delta;
};

template <typename R> struct FieldJob {
  Geo<R> g;
  Tables<R> tb;
  const R *proj;
  const R *mu;
  const double *att;
  int att_global;
  const R *attf;  // precomputed attenuation factors [A][P][je1-je0][Nz] (or null)
  const int32_t *angle_idx;
  int n;
  R *W;
  const R *f;
  double *Wm;
  double *Dsum;
};

template <typename R> struct GatherJob {
  Geo<R> g;
  Tables<R> tb;
  const R *W;
  const int32_t *angle_idx;
  int n;
  int out_mode;
  R *out;
  const R *f;
  const R *nrm;
  R delta;
  R floor_value;
  const double *Wm;
  const double *Pref;
  R *out_mu;
  // Strides (elements) of W between angles and between rows; 0 selects the
  // field layout [n][P][je1-je0][Nz].  Without attenuation and PSF the field
  // is the projection repeated along the rows, so the gather can read the
  // projection [n][jo1-jo0][Nz] itself: angle stride (jo1-jo0)*Nz, row stride 0.
  size_t w_astride;
  size_t w_rstride;
};

// Blocks per job of each kernel (the batched launchers sum these).
template <typename R> int fwd_blocks(const Geo<R> &g, int n) { return n * g.ntiles * g.C; }
template <typename R> int reduce_blocks(const Geo<R> &g, int n) {
  return (int)(((size_t)n * (g.jo1 - g.jo0) * g.Nz + 255) / 256);
}
// z values per lane of the gather kernels: 32 bytes (two 16-byte loads per
// bilinear tap), so each angle's coordinate and tap arithmetic is shared by
// eight float (four double) outputs.
template <typename R> constexpr int gather_zv() { return 32 / (int)sizeof(R); }
template <typename R> int gather_blocks(const Geo<R> &g) {
  constexpr int ZV = gather_zv<R>();
  return (int)(((size_t)g.N * g.N * ((g.Nz + ZV - 1) / ZV) + 255) / 256);
}

// Batched launches: jobs[] and starts[] (first block of each job, ascending)
// are device arrays; all jobs of one launch share the PSF half width H and Nz.
template <typename R>
hipError_t launch_forward_batch(const FwdJob<R> *jobs, const int *starts, int njobs, int blocks,
                                int H, int Nz, int max_L, bool af, hipStream_t st);
template <typename R>
hipError_t launch_reduce_batch(const ReduceJob<R> *jobs, const int *starts, int njobs, int blocks,
                               hipStream_t st);
template <typename R>
hipError_t launch_field_batch(const FieldJob<R> *jobs, const int *starts, int njobs, int blocks,
                              int H, int Nz, int max_L, bool af, hipStream_t st);
template <typename R>
hipError_t launch_gather_batch(const GatherJob<R> *jobs, const int *starts, int njobs, int blocks,
                               int mode, hipStream_t st);

// Threads per block of the tile kernels for a z extent Nz.
int tile_block_threads(int Nz, int run_len, int splits);

// Suffix sums of mu_rot*dx (FP64) at every chunk boundary, for the angles in
// angle_idx; row of angle a is (by_global ? angle_idx[a] : a).
// table layout: [rows][C][je1-je0][Nz] doubles.
template <typename R>
hipError_t launch_att_table(const Geo<R> &g, const Tables<R> &tb, const R *mu,
                            const int32_t *angle_idx, int n, double *table, hipStream_t st);

// Attenuation factors of every visited plane: field[n][P][je1-je0][Nz] (R).
template <typename R>
hipError_t launch_att_field(const Geo<R> &g, const Tables<R> &tb, const R *mu,
                            const int32_t *angle_idx, int n, R *field, hipStream_t st);

// Forward projection partials: partial[C][n][ntiles][kTJ + 2H][Nz].  With trb
// the rotation is the transpose of the reference back projection's gather, so
// the result is B^T f (B = spect_backward REFERENCE).
template <typename R>
hipError_t launch_forward(const Geo<R> &g, const Tables<R> &tb, const R *f, const R *mu,
                          const double *att, bool att_global, const R *attf,
                          const int32_t *angle_idx, int n, R *partial, bool trb, hipStream_t st);

// Balanced-schedule forward (no attenuation, or the attenuation field; no
// B^T): sched_grid_blocks(n, ntiles, P, blocks) blocks, each locating its
// piece of the schedule for `blocks` resident blocks (a multiple of
// forward_sched_blocks(g), the blocks resident at once; 0: the device cannot
// run the kernel), write partial[slot][kTJ + 2H][Nz], one slot each; block 0
// writes the schedule into sched (ints: 2 + n * ntiles + 1) for the reduce.
template <typename R> int forward_sched_blocks(const Geo<R> &g);
template <typename R>
hipError_t launch_forward_sched(const Geo<R> &g, const Tables<R> &tb, const R *f, const R *attf,
                                const int32_t *angle_idx, int n, R *partial, int *sched,
                                int blocks, hipStream_t st);
// Back-projection field over the balanced schedule (attenuation field or no
// attenuation): sched_grid_blocks(n, ntiles, P, blocks) blocks, one piece each;
// field_sched_blocks(g) = its blocks resident at once (0: cannot run).
template <typename R> int field_sched_blocks(const Geo<R> &g);
template <typename R>
hipError_t launch_bp_field_sched(const Geo<R> &g, const Tables<R> &tb, const R *proj,
                                 const R *attf, const int32_t *angle_idx, int n, R *W, int blocks,
                                 hipStream_t st);
template <typename R>
hipError_t launch_reduce_sched(const Geo<R> &g, const R *partial, int n, const int *sched, R *proj,
                               const R *g_meas, const R *s, R *pred, R *ratio, R delta,
                               hipStream_t st);

// proj[n][jo1-jo0][Nz] = sum of the partials; if g_meas != nullptr also
// pred = proj + s (s may be null) and ratio = g_meas / (pred + delta).
template <typename R>
hipError_t launch_forward_reduce(const Geo<R> &g, const R *partial, int n, R *proj,
                                 const R *g_meas, const R *s, R *pred, R *ratio, R delta,
                                 hipStream_t st);

// Back projection field W[n][P][je1-je0][Nz] = a * (G_i * p).  With Wm non-null
// also the attenuation-gradient field (needs f): Wm[n][P][je1-je0][Nz] and the
// chunk totals Dsum[n][C][je1-je0][Nz], both FP64.
template <typename R>
hipError_t launch_bp_field(const Geo<R> &g, const Tables<R> &tb, const R *proj, const R *mu,
                           const double *att, bool att_global, const R *attf,
                           const int32_t *angle_idx, int n, R *W, const R *f, double *Wm,
                           double *Dsum, hipStream_t st);

// Dsum <- prefix over chunks (in place).
template <typename R>
hipError_t launch_chunk_prefix(const Geo<R> &g, double *Dsum, int n, hipStream_t st);

// Gather W back into the object: mode 0 reference rotation, 1 exact adjoint,
// 2 exact adjoint of W -> out and of dx*(Wm - Pref) -> out_mu; out_mode 0
// store, 1 add, 2 OSEM update.
template <typename R>
hipError_t launch_bp_gather(const Geo<R> &g, const Tables<R> &tb, const R *W,
                            const int32_t *angle_idx, int n, int mode, int out_mode, R *out,
                            const R *f, const R *nrm, R delta, R floor_value, const double *Wm,
                            const double *Pref, R *out_mu, hipStream_t st,
                            size_t w_astride = 0, size_t w_rstride = 0);

// Elementwise and stencil kernels (misc.hip).
template <typename R>
hipError_t launch_ratio(const R *g, const R *hf, const R *s, R *pred, R *ratio, int64_t count,
                        R delta, hipStream_t st);
template <typename R>
hipError_t launch_rdp(const R *f, R *out, int nx, int ny, int nz, double dx, double dy, double dz,
                      R beta_scale, R gamma, R delta, hipStream_t st);
template <typename R>
hipError_t launch_rdp_hvp(const R *f, const R *u, R *out, int nx, int ny, int nz, double dx,
                          double dy, double dz, R c, R gamma, R delta, hipStream_t st);
template <typename R>
hipError_t launch_rdp_jvp(const R *f, const R *u, R *out, int nx, int ny, int nz, double dx,
                          double dy, double dz, R c, R gamma, R delta, hipStream_t st);
template <typename R>
hipError_t launch_bsrem_update(R *f, const R *b, const R *nrm, const R *nsum, const R *prior,
                               int64_t count, R relax, R weight, R delta, R floor_value,
                               hipStream_t st);

// Energy-window scatter combination and one-axis smoothing (scatter.hip).
template <typename R>
hipError_t launch_ew_scatter(const R *lower, const R *upper, R *est, R *var, int64_t count, R wl,
                             R wu, R wp, R kl, R ku, R c1, R c2, hipStream_t st);
template <typename R>
hipError_t launch_smooth_axis(const R *in, R *out, int n0, int n1, int n2, int axis, const R *taps,
                              int K, int mode, hipStream_t st);

// Kernel-EM basis and DVF warp (shared.hip).
template <typename R>
hipError_t launch_kem(const R *K, const R *f, R *out, int nx, int ny, int nz, int m,
                      hipStream_t st);
template <typename R>
hipError_t launch_dvf(const R *f, const R *dvf, R *out, int nx, int ny, int nz, hipStream_t st);

// Whole rotated-frame volumes for the per-angle projector path (volume.hip):
// out[n][P][P][Nz] = the object sampled on the rotated frame of each angle;
// the PSF blur of such volumes (g.H > 0), out != vol.
template <typename R>
hipError_t launch_rotate(const Geo<R> &g, const Tables<R> &tb, const R *f, const int32_t *angle_idx,
                         int n, R *out, hipStream_t st);
template <typename R> size_t psf_volume_smem_bytes(const Geo<R> &g);
template <typename R>
hipError_t launch_psf_volume(const Geo<R> &g, const Tables<R> &tb, const R *vol,
                             const int32_t *angle_idx, int n, R *out, hipStream_t st);

}  // namespace spect
