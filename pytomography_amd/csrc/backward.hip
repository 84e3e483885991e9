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
// bp_ref_kernel: W gathered at theta_b = 270 - beta, summed over the angles
//   in registers, cropped, then stored / added / used for the OSEM update.
// bp_adj_kernel: the exact transpose of the forward's bilinear gather,
//   written as a gather over the <= 16 rotated-frame candidates of each voxel
//   (race-free and deterministic).
#include <type_traits>

#include "../../include/spect_hip.h"
#include "launch.hpp"

namespace spect {
namespace {

// GRAD: also the attenuation-map gradient field of the differentiable
// projector: with D(i) = W(i) * f_rot(i) along a ray, dL/dM(k) =
// sum_{i>k} D(i) + D(k)/2 - sum_i D(i)  (M = mu_rot*dx).  The block writes the
// chunk-local part Wm(k) = sum_{i>k, i in chunk} D(i) + D(k)/2 (FP64) and the
// chunk total Dsum (FP64); the gather subtracts the prefix of chunk totals.
template <typename R, int RZ, int QS, int H, bool GRAD, bool AF>
__device__ __forceinline__ void field_body(const FieldJob<R> &J, int al, int jt, int c,
                                           unsigned char *smem) {
  const Geo<R> &g = J.g;
  const Tables<R> &tb = J.tb;
  const R *__restrict__ proj = J.proj;
  const R *__restrict__ mu = J.mu;
  const double *__restrict__ att = J.att;
  const int att_global = J.att_global;
  const R *__restrict__ attf = J.attf;  // AF: precomputed attenuation factors
  const int32_t *__restrict__ angle_idx = J.angle_idx;
  R *__restrict__ W = J.W;
  const R *__restrict__ f = J.f;
  double *__restrict__ Wm = J.Wm;
  double *__restrict__ Dsum = J.Dsum;
  constexpr int K = 2 * H + 1;
  constexpr int HALF_ROWS = kTJ / QS;  // tile rows per y-phase thread
  __shared__ Src<R> src[2][kTJ];
  const int ang = ld_uniform(angle_idx + al);
  if ((unsigned)ang >= (unsigned)g.A) return;
  const int tid = threadIdx.x;
  const int Nz = g.Nz;
  const int runs = (Nz + RZ - 1) / RZ;
  const int ldu = runs * RZ + 2 * H + 1;  // H zeros | data | H zeros | pad (odd stride)
  R *U = reinterpret_cast<R *>(smem);     // [kTJ][ldu] after the y blur, data at +H
  for (int q = tid; q < kTJ * ldu; q += blockDim.x) U[q] = R(0);
  const int j0 = g.je0 + jt * kTJ;
  const int tj = min(kTJ, g.je1 - j0);
  const int n_jext = g.je1 - g.je0;
  const int nout = g.jo1 - g.jo0;
  // phase A ownership: one z, 1/QS of the tile rows (wave-uniform split)
  const int nzc = (Nz + 63) / 64 * 64;
  const int half = tid / nzc, za = tid - half * nzc;
  const bool aact = half < QS && za < Nz;
  // phase B ownership: column jl, z run [z0, z0 + RZ)
  const int jl = tid % kTJ, run = tid / kTJ, z0 = run * RZ;
  const bool bact = run < runs;
  const bool mine = bact && jl < tj;
  const auto cs = ld_uniform(tb.trig + ang);
  const bool has_mu = mu != nullptr;

  // projection window for the y blur: tile rows [half*16 - H, half*16 + 16 + H)
  R pw[HALF_ROWS + 2 * H];
#pragma unroll
  for (int m = 0; m < HALF_ROWS + 2 * H; ++m) {
    pw[m] = R(0);
    if (aact) {
      const int j = j0 + half * HALF_ROWS - H + m;
      if (j >= g.jo0 && j < g.jo1) pw[m] = proj[((size_t)al * nout + (j - g.jo0)) * Nz + za];
    }
  }
  double S[AF ? 1 : RZ];
#pragma unroll
  for (int m = 0; m < (AF ? 1 : RZ); ++m) {
    S[m] = 0.0;
    if (!AF && has_mu && mine && z0 + m < Nz) {
      const int rowa = att_global ? ang : al;
      S[m] = att[(((size_t)rowa * g.C + c) * n_jext + (j0 - g.je0 + jl)) * Nz + z0 + m];
    }
  }
  double SD[GRAD ? RZ : 1];
#pragma unroll
  for (int m = 0; m < (GRAD ? RZ : 1); ++m) SD[m] = 0.0;
  const int rix = ld_uniform(tb.radius_index + ang);
  const int i_lo = c * g.L, i_hi = min(g.P, i_lo + g.L);
  // the chunk's PSF taps (z then y, per plane) staged in LDS: the blur phases
  // read them as LDS broadcasts instead of waiting on scalar loads
  R *Wt = U + (size_t)kTJ * ldu;  // [i_hi - i_lo][2K]
  if constexpr (H > 0) {
    const int nt = (i_hi - i_lo) * 2 * K;
    for (int q = tid; q < nt; q += blockDim.x) {
      const int l = q / (2 * K), r = q - l * (2 * K);
      const size_t row = ((size_t)rix * g.P + i_lo + l) * K;
      Wt[q] = r < K ? tb.taps_z[row + r] : tb.taps_r[row + r - K];
    }
  }
  int buf = 0;
  for (int i = i_hi - 1; i >= i_lo; --i) {
    int inside = 0;
    if (tid < tj) {
      R ix, iy;
      rot_coord<R>(cs.x, cs.y, tb.lin[j0 + tid], ld_uniform(tb.lin + i), g.half, ix, iy);
      src[buf][tid] = make_src<R>(ix, iy, g.pad);
      inside = in_footprint<R>(ix, iy, g);
    }
    if (!__syncthreads_or(inside)) continue;
    R bz[RZ];
    // AF: this plane's attenuation factors are requested before the z blur
    // (after the y-blur barrier, which would otherwise wait for them) so that
    // their HBM latency overlaps the blur arithmetic
    R a[AF ? RZ : 1];
    auto load_att = [&]() {
      if constexpr (AF) {
        if (mine)
          load_run<R, RZ>(attf + (((size_t)ang * g.P + i) * n_jext + (j0 - g.je0 + jl)) * Nz, z0,
                          Nz, a);
      }
    };
    if constexpr (H > 0) {
      // ---- A: y blur of the projection with this plane's taps -> U
      if (aact) {
        const R *wy = Wt + (size_t)(i - i_lo) * 2 * K + K;
        R u[HALF_ROWS];
#pragma unroll
        for (int r = 0; r < HALF_ROWS; ++r) u[r] = R(0);
#pragma unroll
        for (int t = 0; t < K; ++t) {
          const R w = wy[t];
#pragma unroll
          for (int r = 0; r < HALF_ROWS; ++r) u[r] += w * pw[r + t];
        }
#pragma unroll
        for (int r = 0; r < HALF_ROWS; ++r) U[(size_t)(half * HALF_ROWS + r) * ldu + H + za] = u[r];
      }
      __syncthreads();
      load_att();
      // ---- B: z blur of the own run
      if (bact) {
        const R *wz = Wt + (size_t)(i - i_lo) * 2 * K;
        const R *ur = U + (size_t)jl * ldu + z0;  // ur[m] = U(z0 - H + m)
        R win[RZ + 2 * H];
#pragma unroll
        for (int m = 0; m < RZ + 2 * H; ++m) win[m] = ur[m];
#pragma unroll
        for (int m = 0; m < RZ; ++m) bz[m] = R(0);
#pragma unroll
        for (int t = 0; t < K; ++t) {
          const R w = wz[t];
#pragma unroll
          for (int m = 0; m < RZ; ++m) bz[m] += w * win[m + t];
        }
      }
    } else {
      load_att();
      // no PSF: the smeared projection itself
      if (bact) {
        const int j = j0 + jl;
#pragma unroll
        for (int m = 0; m < RZ; ++m) {
          const int z = z0 + m;
          bz[m] = (j >= g.jo0 && j < g.jo1 && z < Nz) ? proj[((size_t)al * nout + (j - g.jo0)) * Nz + z]
                                                       : R(0);
        }
      }
    }
    // ---- attenuation and store
    if (mine) {
      Taps4<R> t = taps4<R>(src[buf][jl], g.N, Nz);
      if constexpr (AF) {
#pragma unroll
        for (int m = 0; m < RZ; ++m) bz[m] = bz[m] * a[m];
      } else if (has_mu) {
        R x[RZ];
        gather_run<R, RZ>(mu, t, z0, Nz, x);
#pragma unroll
        for (int m = 0; m < RZ; ++m) {
          const R xm = x[m] * g.dx;
          S[m] += (double)xm;
          bz[m] = bz[m] * att_factor<R>(S[m], xm);
        }
      }
      if constexpr (GRAD) {
        R fv[RZ];
        gather_run<R, RZ>(f, t, z0, Nz, fv);
        double *dm = Wm + (((size_t)al * g.P + i) * n_jext + (j0 - g.je0 + jl)) * Nz + z0;
#pragma unroll
        for (int m = 0; m < RZ; ++m) {
          const double d = (double)bz[m] * (double)fv[m];
          if (z0 + m < Nz) dm[m] = SD[m] + 0.5 * d;
          SD[m] += d;
        }
      }
      R *dst = W + (((size_t)al * g.P + i) * n_jext + (j0 - g.je0 + jl)) * Nz + z0;
      if (z0 + RZ <= Nz && (Nz % 4) == 0) {
        if constexpr (sizeof(R) == 4) {
#pragma unroll
          for (int m = 0; m < RZ; m += 4)
            *reinterpret_cast<float4 *>(dst + m) = make_float4(bz[m], bz[m + 1], bz[m + 2], bz[m + 3]);
        } else {
#pragma unroll
          for (int m = 0; m < RZ; m += 2)
            *reinterpret_cast<double2 *>(dst + m) = make_double2(bz[m], bz[m + 1]);
        }
      } else {
#pragma unroll
        for (int m = 0; m < RZ; ++m)
          if (z0 + m < Nz) dst[m] = bz[m];
      }
    }
    buf ^= 1;
  }
  if constexpr (GRAD) {
    if (mine) {
#pragma unroll
      for (int m = 0; m < RZ; ++m)
        if (z0 + m < Nz)
          Dsum[(((size_t)al * g.C + c) * n_jext + (j0 - g.je0 + jl)) * Nz + z0 + m] = SD[m];
    }
  }
}

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

// Output stage of a back projection for one voxel value b.
template <typename R>
__device__ __forceinline__ void bp_store(int out_mode, R *__restrict__ out, size_t o, R b,
                                         const R *__restrict__ f, const R *__restrict__ nrm,
                                         R delta, R floor_value) {
#pragma clang fp contract(off)
  if (out_mode == SPECT_OUT_STORE) {
    out[o] = b;
  } else if (out_mode == SPECT_OUT_ADD) {
    out[o] = out[o] + b;
  } else {
    // preconditioned_gradient_ascent.py:121-122 with OSEM's factor (:329)
    const R n = nrm[o];
    const R fo = f[o];
    const R pre = fo * (R(1) / (n + delta));
    const R grad = b - n;
    R v = fo + pre * (grad - R(0));
    if (v <= floor_value) v = floor_value;
    out[o] = v;
  }
}

// Vector of ZV consecutive z values.
template <typename R, int ZV> struct ZVec { R v[ZV]; };

template <typename R, int ZV>
__device__ __forceinline__ ZVec<R, ZV> load_zvec(const R *p) {
  ZVec<R, ZV> r;
  if constexpr (sizeof(R) == 4 && ZV == 4) {
    const float4 x = *reinterpret_cast<const float4 *>(p);
    r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
  } else if constexpr (sizeof(R) == 8 && ZV == 2) {
    const double2 x = *reinterpret_cast<const double2 *>(p);
    r.v[0] = x.x; r.v[1] = x.y;
  } else {
#pragma unroll
    for (int m = 0; m < ZV; ++m) r.v[m] = p[m];
  }
  return r;
}

template <typename R, int ZV>
__device__ __forceinline__ void store_zvec(R *p, const R (&v)[ZV]) {
  if constexpr (sizeof(R) == 4 && ZV == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (sizeof(R) == 8 && ZV == 2) {
    *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
  } else {
#pragma unroll
    for (int m = 0; m < ZV; ++m) p[m] = v[m];
  }
}

// VEC: all ZV values valid and 16-byte aligned; the operands of the update are
// read as vectors (one request each instead of one per value).
template <typename R, int ZV, bool VEC = false>
__device__ __forceinline__ void bp_store_vec(int out_mode, R *__restrict__ out, size_t o,
                                             const R (&b)[ZV], int nvalid,
                                             const R *__restrict__ f, const R *__restrict__ nrm,
                                             R delta, R floor_value) {
#pragma clang fp contract(off)
  if constexpr (VEC) {
    R r[ZV];
    if (out_mode == SPECT_OUT_STORE) {
#pragma unroll
      for (int m = 0; m < ZV; ++m) r[m] = b[m];
    } else if (out_mode == SPECT_OUT_ADD) {
      const ZVec<R, ZV> ov = load_zvec<R, ZV>(out + o);
#pragma unroll
      for (int m = 0; m < ZV; ++m) r[m] = ov.v[m] + b[m];
    } else {
      const ZVec<R, ZV> nv = load_zvec<R, ZV>(nrm + o);
      const ZVec<R, ZV> fv = load_zvec<R, ZV>(f + o);
#pragma unroll
      for (int m = 0; m < ZV; ++m) {
        // the same expression as bp_store
        const R n = nv.v[m];
        const R fo = fv.v[m];
        const R pre = fo * (R(1) / (n + delta));
        const R grad = b[m] - n;
        R v = fo + pre * (grad - R(0));
        if (v <= floor_value) v = floor_value;
        r[m] = v;
      }
    }
    store_zvec<R, ZV>(out + o, r);
  } else {
#pragma unroll
    for (int m = 0; m < ZV; ++m)
      if (m < nvalid) bp_store<R>(out_mode, out, o + m, b[m], f, nrm, delta, floor_value);
  }
}

// Each thread owns one object column (x, y) and ZV consecutive z; the wave's
// lanes cover consecutive z (then the next columns), so W reads coalesce.
// mode 0: reference gather at theta_b = -theta_f (spect_system_matrix.py:196);
// mode 1: exact transpose of the forward gather.
// mode 2: exact transpose as mode 1 for W -> out, and for the attenuation
// gradient field dx * (Wm - prefix of chunk totals) -> out_mu (FP64 sums).
template <typename R, int ZV, int MODE>
__device__ __forceinline__ void gather_body(const GatherJob<R> &J, size_t gid) {
  const Geo<R> &g = J.g;
  const Tables<R> &tb = J.tb;
  const R *__restrict__ W = J.W;
  const int32_t *__restrict__ angle_idx = J.angle_idx;
  const int n_call = J.n;
  const int out_mode = J.out_mode;
  R *__restrict__ out = J.out;
  const R *__restrict__ f = J.f;
  const R *__restrict__ nrm = J.nrm;
  const R delta = J.delta, floor_value = J.floor_value;
  const double *__restrict__ Wm = J.Wm;
  const double *__restrict__ Pref = J.Pref;
  R *__restrict__ out_mu = J.out_mu;
  const int Nz = g.Nz, N = g.N;
  const int zq = (Nz + ZV - 1) / ZV;  // z groups per column
  if (gid >= (size_t)N * N * zq) return;
  const int col = (int)(gid / zq);
  const int z0 = (int)(gid - (size_t)col * zq) * ZV;
  const int nvalid = min(ZV, Nz - z0);
  const bool vec = (nvalid == ZV) && (Nz % 4 == 0);
  const int x = col / N, y = col - x * N;
  const int X = x + g.pad, Y = y + g.pad;
  const int n_jext = g.je1 - g.je0;
  const size_t plane = (size_t)n_jext * Nz;
  R acc[ZV];
  double accm[MODE == 2 ? ZV : 1];
#pragma unroll
  for (int m = 0; m < ZV; ++m) acc[m] = R(0);
#pragma unroll
  for (int m = 0; m < (MODE == 2 ? ZV : 1); ++m) accm[m] = 0.0;
  // The vector / element choice is made once per thread (VEC is a
  // compile-time constant inside run): a branch around every load would make
  // the compiler wait for all outstanding loads at each merge.
  const R linY = tb.lin[Y], linX = tb.lin[X];  // angle-independent
  auto run = [&](auto vec_c) {
  constexpr bool VEC = decltype(vec_c)::value;
  auto fetch = [&](const R *p, R (&v)[ZV]) {
    if constexpr (VEC) {
      const ZVec<R, ZV> t = load_zvec<R, ZV>(p);
#pragma unroll
      for (int m = 0; m < ZV; ++m) v[m] = t.v[m];
    } else {
#pragma unroll
      for (int m = 0; m < ZV; ++m) v[m] = m < nvalid ? p[m] : R(0);
    }
  };
  for (int al = 0; al < n_call; ++al) {
    const int ang = ld_uniform(angle_idx + al);
    if ((unsigned)ang >= (unsigned)g.A) continue;
    const auto cs = ld_uniform(tb.trig + ang);
    const R *Wa = W + (size_t)al * g.P * plane;
    if constexpr (MODE == 0) {
      R ix, iy;
      rot_coord<R>(cs.x, -cs.y, linY, linX, g.half, ix, iy);  // theta_b = -theta_f
      const R fx = floor(ix), fy = floor(iy);
      const int c0 = int(fx), r0 = int(fy);
      const R tx = ix - fx, ty = iy - fy;
      const R ex = R(1) - tx, ny = R(1) - ty;
      const int rr[2] = {r0, r0 + 1}, cc[2] = {c0, c0 + 1};
      const R wr[2] = {ny, ty}, wc[2] = {ex, tx};
      R v[ZV];
#pragma unroll
      for (int m = 0; m < ZV; ++m) v[m] = R(0);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const bool ok = (unsigned)rr[a] < (unsigned)g.P && cc[b] >= g.je0 && cc[b] < g.je1;
          const int r = min(max(rr[a], 0), g.P - 1);
          const int cl = min(max(cc[b], g.je0), g.je1 - 1);
          const R w = ok ? wr[a] * wc[b] : R(0);
          R t[ZV];
          fetch(Wa + (size_t)r * plane + (size_t)(cl - g.je0) * Nz + z0, t);
          // ((nw + ne) + sw) + se
#pragma unroll
          for (int m = 0; m < ZV; ++m) v[m] = (a == 0 && b == 0) ? t[m] * w : v[m] + t[m] * w;
        }
      }
#pragma unroll
      for (int m = 0; m < ZV; ++m) acc[m] = acc[m] + v[m];
    } else {
      // rotated-frame position whose forward sample lands on (row X, col Y)
      R jc, ic;
      rot_coord<R>(cs.x, -cs.y, linY, linX, g.half, jc, ic);
      const int ib = int(floor(ic - R(1.5))), jb = int(floor(jc - R(1.5)));
      R v[ZV];
#pragma unroll
      for (int m = 0; m < ZV; ++m) v[m] = R(0);
      for (int di = 0; di < 4; ++di) {
        const int i = ib + di;
        if ((unsigned)i >= (unsigned)g.P) continue;
        for (int dj = 0; dj < 4; ++dj) {
          const int j = jb + dj;
          if (j < g.je0 || j >= g.je1) continue;
          R ix, iy;
          rot_coord<R>(cs.x, cs.y, tb.lin[j], tb.lin[i], g.half, ix, iy);
          const R fx = floor(ix), fy = floor(iy);
          const int c0 = int(fx), r0 = int(fy);
          const R tx = ix - fx, ty = iy - fy;
          R wr, wc;
          if (r0 == X) wr = R(1) - ty;
          else if (r0 + 1 == X) wr = ty;
          else continue;
          if (c0 == Y) wc = R(1) - tx;
          else if (c0 + 1 == Y) wc = tx;
          else continue;
          const R w = wr * wc;
          R t[ZV];
          fetch(Wa + (size_t)i * plane + (size_t)(j - g.je0) * Nz + z0, t);
#pragma unroll
          for (int m = 0; m < ZV; ++m) v[m] = v[m] + t[m] * w;
          if constexpr (MODE == 2) {
            const size_t o = (size_t)i * plane + (size_t)(j - g.je0) * Nz + z0;
            const double *wm = Wm + (size_t)al * g.P * plane + o;
            const double *pr = Pref + ((size_t)al * g.C + i / g.L) * plane + (size_t)(j - g.je0) * Nz + z0;
#pragma unroll
            for (int m = 0; m < ZV; ++m)
              if (m < nvalid) accm[m] += (double)w * (wm[m] - pr[m]);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < ZV; ++m) acc[m] = acc[m] + v[m];
    }
  }
  bp_store_vec<R, ZV, VEC>(out_mode, out, ((size_t)x * N + y) * Nz + z0, acc, nvalid, f, nrm,
                           delta, floor_value);
  };
  if (vec) run(std::true_type{});
  else run(std::false_type{});
  if constexpr (MODE == 2) {
    R gm[ZV];
#pragma unroll
    for (int m = 0; m < ZV; ++m) gm[m] = R((double)g.dx * accm[m]);
    bp_store_vec<R, ZV>(out_mode, out_mu, ((size_t)x * N + y) * Nz + z0, gm, nvalid, f, nrm, delta,
                        floor_value);
  }
}

template <typename R, int ZV, int MODE>
__global__ __launch_bounds__(256) void bp_gather_kernel(GatherJob<R> J) {
  gather_body<R, ZV, MODE>(J, (size_t)blockIdx.x * 256 + threadIdx.x);
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

}  // namespace

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
                            const double *Pref, R *out_mu, hipStream_t st) {
  constexpr int ZV = sizeof(R) == 4 ? 4 : 2;
  const unsigned blocks = (unsigned)gather_blocks(g);
  const GatherJob<R> J{g, tb, W, idx, n, out_mode, out, f, nrm, delta, floor_value, Wm, Pref, out_mu};
  if (mode == SPECT_BP_REFERENCE)
    bp_gather_kernel<R, ZV, 0><<<blocks, 256, 0, st>>>(J);
  else if (mode == SPECT_BP_ADJOINT)
    bp_gather_kernel<R, ZV, 1><<<blocks, 256, 0, st>>>(J);
  else
    bp_gather_kernel<R, ZV, 2><<<blocks, 256, 0, st>>>(J);
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
  constexpr int ZV = sizeof(R) == 4 ? 4 : 2;
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
  template hipError_t launch_field_batch<R>(const FieldJob<R> *, const int *, int, int, int,    \
                                            int, int, bool, hipStream_t);                      \
  template hipError_t launch_gather_batch<R>(const GatherJob<R> *, const int *, int, int, int,  \
                                             hipStream_t);                                     \
  template hipError_t launch_bp_gather<R>(const Geo<R> &, const Tables<R> &, const R *,        \
                                          const int32_t *, int, int, int, R *, const R *,      \
                                          const R *, R, R, const double *, const double *, R *, \
                                          hipStream_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
