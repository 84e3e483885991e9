// backward_body.hpp -- block bodies of the back projection: the field
// W = a * (G_i * p) of one (angle, tile, chunk) block (field_body) and the
// rotation back with the output stage of one thread (gather_body), shared by
// the single-problem and batched kernels of backward.hip.  Plain device
// functions of the launch descriptors and the block / thread indices, so the
// host build in tests/cpu_build executes the same source.
#pragma once

#include <type_traits>

#include "../../include/spect_hip.h"
#include "launch.hpp"

namespace spect {

// GRAD: also the attenuation-map gradient field of the differentiable
// projector: with D(i) = W(i) * f_rot(i) along a ray, dL/dM(k) =
// sum_{i>k} D(i) + D(k)/2 - sum_i D(i)  (M = mu_rot*dx).  The block writes the
// chunk-local part Wm(k) = sum_{i>k, i in chunk} D(i) + D(k)/2 (FP64) and the
// chunk total Dsum (FP64); the gather subtracts the prefix of chunk totals.
// Planes [i_lo, i_hi) of (angle al, tile jt); c is the chunk whose FP64
// suffix-table rows start the attenuation march (suffix-table mode only).
template <typename R, int RZ, int QS, int H, bool GRAD, bool AF, bool MARCH = !AF>
__device__ __forceinline__ void field_body_range(const FieldJob<R> &J, int al, int jt, int c,
                                                 int i_lo, int i_hi, unsigned char *smem) {
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
  // MARCH: mu may be given and is marched here (FP64 suffix sums); a scheduled
  // piece runs with the attenuation field or without attenuation, never both
  const bool has_mu = MARCH && mu != nullptr;

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
  double S[MARCH ? RZ : 1];
#pragma unroll
  for (int m = 0; m < (MARCH ? RZ : 1); ++m) {
    S[m] = 0.0;
    if (MARCH && has_mu && mine && z0 + m < Nz) {
      const int rowa = att_global ? ang : al;
      S[m] = att[(((size_t)rowa * g.C + c) * n_jext + (j0 - g.je0 + jl)) * Nz + z0 + m];
    }
  }
  double SD[GRAD ? RZ : 1];
#pragma unroll
  for (int m = 0; m < (GRAD ? RZ : 1); ++m) SD[m] = 0.0;
  const int rix = ld_uniform(tb.radius_index + ang);
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
    // AF: this plane's attenuation factors are requested after the y-blur
    // barrier, so their HBM latency overlaps the z blur (requested before the
    // y blur as well, they measured 10 us slower per 15-angle launch).
    R a[AF ? RZ : 1];
    auto load_att = [&]() {
      if constexpr (AF) {
        if (mine)
          load_run<R, RZ>(attf + (((size_t)ang * g.P + i) * n_jext + (j0 - g.je0 + jl)) * Nz, z0,
                          Nz, a);
      }
    };
    // A scheduled piece (c < 0) covers planes of the activity table's range
    // only, so the plane is taken without the block-wide test (a plane of
    // that superset with no sample inside gets W written all the same).
    if (c < 0) __syncthreads();
    else if (!__syncthreads_or(inside)) continue;
    R bz[RZ];
    if constexpr (H > 0) {
      // ---- A: y blur of the projection with this plane's taps -> U, over
      // the plane's tier of taps [H - HE, H + HE]
      const int he = ld_uniform(tb.tier + (size_t)rix * g.P + i);
      if (aact) {
        const R *wy = Wt + (size_t)(i - i_lo) * 2 * K + K;
        with_tier<H>(he, [&](auto he_c) {
          constexpr int HE = decltype(he_c)::value;
          R u[HALF_ROWS];
#pragma unroll
          for (int r = 0; r < HALF_ROWS; ++r) u[r] = R(0);
#pragma unroll
          for (int t = H - HE; t <= H + HE; ++t) {
            const R w = wy[t];
#pragma unroll
            for (int r = 0; r < HALF_ROWS; ++r) u[r] += w * pw[r + t];
          }
#pragma unroll
          for (int r = 0; r < HALF_ROWS; ++r) U[(size_t)(half * HALF_ROWS + r) * ldu + H + za] = u[r];
        });
      }
      __syncthreads();
      load_att();
      // ---- B: z blur of the own run
      if (bact) {
        const R *wz = Wt + (size_t)(i - i_lo) * 2 * K;
        with_tier<H>(he, [&](auto he_c) {
          constexpr int HE = decltype(he_c)::value;
          const R *ur = U + (size_t)jl * ldu + z0 + (H - HE);  // ur[m] = U(z0 - HE + m)
          R win[RZ + 2 * HE];
#pragma unroll
          for (int m = 0; m < RZ + 2 * HE; ++m) win[m] = ur[m];
#pragma unroll
          for (int m = 0; m < RZ; ++m) bz[m] = R(0);
#pragma unroll
          for (int t = 0; t < 2 * HE + 1; ++t) {
            const R w = wz[H - HE + t];
#pragma unroll
            for (int m = 0; m < RZ; ++m) bz[m] += w * win[m + t];
          }
        });
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
      } else if constexpr (MARCH) {
        if (has_mu) {
          R x[RZ];
          gather_run<R, RZ>(mu, t, z0, Nz, x);
#pragma unroll
          for (int m = 0; m < RZ; ++m) {
            const R xm = x[m] * g.dx;
            S[m] += (double)xm;
            bz[m] = bz[m] * att_factor<R>(S[m], xm);
          }
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

// Chunk c of the uniform schedule: planes [c L, c L + L).
template <typename R, int RZ, int QS, int H, bool GRAD, bool AF>
__device__ __forceinline__ void field_body(const FieldJob<R> &J, int al, int jt, int c,
                                           unsigned char *smem) {
  const int i_lo = c * J.g.L;
  field_body_range<R, RZ, QS, H, GRAD, AF>(J, al, jt, c, i_lo, min(J.g.P, i_lo + J.g.L), smem);
}

// Block b of the balanced field schedule (sched_locate in forward_body.hpp:
// the forward's pairs and plane ranges, split for this kernel's resident
// blocks): piece b, planes walked from the top of its range.
template <typename R, int RZ, int QS, int H, bool AF>
__device__ __forceinline__ void field_sched_body(const FieldJob<R> &J, int nblocks, int b,
                                                 unsigned char *smem) {
  const Geo<R> &g = J.g;
  const SchedLoc s = sched_locate<R>(g, J.angle_idx, J.tb.ranges, J.n, nblocks, b, nullptr);
  if (s.p < 0) return;
  const int al = s.p / g.ntiles, jt = s.p - al * g.ntiles;
  const int ang = ld_uniform(J.angle_idx + al);
  const int2 r = J.tb.ranges[(size_t)ang * g.ntiles + jt];
  int i_lo, i_hi;
  sched_piece_planes(r, s.k, s.c, i_lo, i_hi);
  field_body_range<R, RZ, QS, H, false, AF, false>(J, al, jt, -1, i_lo, i_hi, smem);
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

// 16-byte vector accesses when ZV values fill whole 16-byte words.
template <typename R, int ZV>
__device__ __forceinline__ ZVec<R, ZV> load_zvec(const R *p) {
  ZVec<R, ZV> r;
  if constexpr (sizeof(R) == 4 && ZV % 4 == 0) {
#pragma unroll
    for (int q = 0; q < ZV; q += 4) {
      const float4 x = *reinterpret_cast<const float4 *>(p + q);
      r.v[q] = x.x; r.v[q + 1] = x.y; r.v[q + 2] = x.z; r.v[q + 3] = x.w;
    }
  } else if constexpr (sizeof(R) == 8 && ZV % 2 == 0) {
#pragma unroll
    for (int q = 0; q < ZV; q += 2) {
      const double2 x = *reinterpret_cast<const double2 *>(p + q);
      r.v[q] = x.x; r.v[q + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int m = 0; m < ZV; ++m) r.v[m] = p[m];
  }
  return r;
}

template <typename R, int ZV>
__device__ __forceinline__ void store_zvec(R *p, const R (&v)[ZV]) {
  if constexpr (sizeof(R) == 4 && ZV % 4 == 0) {
#pragma unroll
    for (int q = 0; q < ZV; q += 4)
      *reinterpret_cast<float4 *>(p + q) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
  } else if constexpr (sizeof(R) == 8 && ZV % 2 == 0) {
#pragma unroll
    for (int q = 0; q < ZV; q += 2) *reinterpret_cast<double2 *>(p + q) = make_double2(v[q], v[q + 1]);
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
  const size_t astride = J.w_astride ? J.w_astride : (size_t)g.P * plane;
  const size_t rstride = J.w_astride ? J.w_rstride : plane;
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
    const R *Wa = W + (size_t)al * astride;
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
          fetch(Wa + (size_t)r * rstride + (size_t)(cl - g.je0) * Nz + z0, t);
          // ((nw + ne) + sw) + se
#pragma unroll
          for (int m = 0; m < ZV; ++m) v[m] = (a == 0 && b == 0) ? t[m] * w : v[m] + t[m] * w;
        }
      }
#pragma unroll
      for (int m = 0; m < ZV; ++m) acc[m] = acc[m] + v[m];
    } else if constexpr (MODE == 3) {
      // MODE 0 on a field whose rows all hold the same values (the plain
      // back projection reads the projection itself): the two row weights
      // of a column add up, two loads per sample instead of four
      R ix, iy;
      rot_coord<R>(cs.x, -cs.y, linY, linX, g.half, ix, iy);  // theta_b = -theta_f
      const R fx = floor(ix), fy = floor(iy);
      const int c0 = int(fx), r0 = int(fy);
      const R tx = ix - fx, ty = iy - fy;
      const R ex = R(1) - tx, ny = R(1) - ty;
      const R wrow = ((unsigned)r0 < (unsigned)g.P ? ny : R(0)) +
                     ((unsigned)(r0 + 1) < (unsigned)g.P ? ty : R(0));
      const int cc[2] = {c0, c0 + 1};
      const R wc[2] = {ex, tx};
      R v[ZV];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const bool ok = cc[b] >= g.je0 && cc[b] < g.je1;
        const int cl = min(max(cc[b], g.je0), g.je1 - 1);
        const R w = ok ? wrow * wc[b] : R(0);
        R t[ZV];
        fetch(Wa + (size_t)(cl - g.je0) * Nz + z0, t);
#pragma unroll
        for (int m = 0; m < ZV; ++m) v[m] = b == 0 ? t[m] * w : v[m] + t[m] * w;
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
          fetch(Wa + (size_t)i * rstride + (size_t)(j - g.je0) * Nz + z0, t);
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

// Dynamic LDS of gather_adj_block: per column of the block a contributor
// count and 16 (row << 16 | column, weight) entries.
template <typename R> __host__ __device__ inline int adj_block_cols(int zq) { return 256 / zq + 2; }
template <typename R> __host__ __device__ inline size_t adj_block_smem(int zq) {
  return (size_t)adj_block_cols<R>(zq) * (sizeof(int) + 16 * (sizeof(uint32_t) + sizeof(R)));
}

// Exact-adjoint gather (MODE 1; MODE 2 adds the mu-gradient sums) of block b
// of a 256-lane launch over the lanes of gather_body.  Per angle, the block
// first builds the contributor list of each of its object columns once -- the
// rotated-frame pixels (i, j) whose forward bilinear sample holds the
// column's pixel, tested over the 4 x 4 candidates of gather_body's search by
// one 16-lane group per column and packed by a wave ballot in candidate order
// -- and every z lane of the column then sums its list in that order: the
// same sums as gather_body, without each of a column's z lanes repeating the
// search.  Every thread of the block calls (barriers).
template <typename R, int ZV, int MODE>
__device__ __forceinline__ void gather_adj_block(const GatherJob<R> &J, int b, unsigned char *smem) {
  static_assert(MODE == 1 || MODE == 2, "exact-adjoint modes");
  const Geo<R> &g = J.g;
  const Tables<R> &tb = J.tb;
  const R *__restrict__ W = J.W;
  const int Nz = g.Nz, N = g.N;
  const int zq = (Nz + ZV - 1) / ZV;
  const size_t total = (size_t)N * N * zq;
  const size_t gid0 = (size_t)b * 256;
  const int tid = threadIdx.x;
  const size_t gid = gid0 + tid;
  const bool lane_ok = gid < total;
  const int col_lo = (int)(gid0 / zq);
  const size_t last = gid0 + 255 < total - 1 ? gid0 + 255 : total - 1;
  const int ncols = (int)(last / zq) - col_lo + 1;
  const int cols_max = adj_block_cols<R>(zq);
  int *cnt = reinterpret_cast<int *>(smem);                    // [cols_max]
  uint32_t *ent_ij = reinterpret_cast<uint32_t *>(cnt + cols_max);  // [cols_max][16]
  R *ent_w = reinterpret_cast<R *>(ent_ij + (size_t)cols_max * 16);  // [cols_max][16]
  const int col = lane_ok ? (int)(gid / zq) : col_lo;
  const int lc = col - col_lo;
  const int z0 = lane_ok ? (int)(gid - (size_t)col * zq) * ZV : 0;
  const int nvalid = min(ZV, Nz - z0);
  const bool vec = (nvalid == ZV) && (Nz % 4 == 0);
  const int x = col / N, y = col - x * N;
  const int n_jext = g.je1 - g.je0;
  const size_t plane = (size_t)n_jext * Nz;
  const size_t astride = J.w_astride ? J.w_astride : (size_t)g.P * plane;
  const size_t rstride = J.w_astride ? J.w_rstride : plane;
  R acc[ZV];
  double accm[MODE == 2 ? ZV : 1];
#pragma unroll
  for (int m = 0; m < ZV; ++m) acc[m] = R(0);
#pragma unroll
  for (int m = 0; m < (MODE == 2 ? ZV : 1); ++m) accm[m] = 0.0;
  const unsigned lane = tid & 63u;
  for (int al = 0; al < J.n; ++al) {
    const int ang = ld_uniform(J.angle_idx + al);
    if ((unsigned)ang >= (unsigned)g.A) continue;  // block-uniform
    const auto cs = ld_uniform(tb.trig + ang);
    // ---- the block's columns' lists (item q: column q / 16, candidate q % 16)
    for (int q0 = 0; q0 < ncols * 16; q0 += 256) {
      const int q = q0 + tid;
      const int qc = q >> 4, k = q & 15;
      bool use = false;
      uint32_t pij = 0;
      R w = R(0);
      if (q < ncols * 16) {
        const int c = col_lo + qc;
        const int cx = c / N, cy = c - cx * N;
        const int X = cx + g.pad, Y = cy + g.pad;
        // rotated-frame position whose forward sample lands on (row X, col Y)
        R jc, ic;
        rot_coord<R>(cs.x, -cs.y, tb.lin[Y], tb.lin[X], g.half, jc, ic);
        const int i = int(floor(ic - R(1.5))) + (k >> 2), j = int(floor(jc - R(1.5))) + (k & 3);
        if ((unsigned)i < (unsigned)g.P && j >= g.je0 && j < g.je1) {
          R ix, iy;
          rot_coord<R>(cs.x, cs.y, tb.lin[j], tb.lin[i], g.half, ix, iy);
          const R fx = floor(ix), fy = floor(iy);
          const int c0 = int(fx), r0 = int(fy);
          const R tx = ix - fx, ty = iy - fy;
          if ((r0 == X || r0 + 1 == X) && (c0 == Y || c0 + 1 == Y)) {
            const R wr = r0 == X ? R(1) - ty : ty;
            const R wc = c0 == Y ? R(1) - tx : tx;
            use = true;
            pij = ((uint32_t)i << 16) | (uint32_t)(j - g.je0);
            w = wr * wc;
          }
        }
      }
      const unsigned grp = (unsigned)(__ballot(use) >> (lane & 48u)) & 0xFFFFu;
      if (use) {
        const int slot = __popc(grp & ((1u << (lane & 15u)) - 1u));
        ent_ij[(size_t)qc * 16 + slot] = pij;
        ent_w[(size_t)qc * 16 + slot] = w;
      }
      if (q < ncols * 16 && k == 0) cnt[qc] = __popc(grp);
    }
    __syncthreads();
    // ---- the z lanes sum their column's list
    if (lane_ok) {
      const R *Wa = W + (size_t)al * astride;
      R v[ZV];
#pragma unroll
      for (int m = 0; m < ZV; ++m) v[m] = R(0);
      const int nc = cnt[lc];
      auto sum = [&](auto vec_c) {
        constexpr bool VEC = decltype(vec_c)::value;
        for (int e = 0; e < nc; ++e) {
          const uint32_t pij = ent_ij[(size_t)lc * 16 + e];
          const R w = ent_w[(size_t)lc * 16 + e];
          const int i = (int)(pij >> 16), jr = (int)(pij & 0xFFFFu);
          const R *p = Wa + (size_t)i * rstride + (size_t)jr * Nz + z0;
          R t[ZV];
          if constexpr (VEC) {
            const ZVec<R, ZV> tv = load_zvec<R, ZV>(p);
#pragma unroll
            for (int m = 0; m < ZV; ++m) t[m] = tv.v[m];
          } else {
#pragma unroll
            for (int m = 0; m < ZV; ++m) t[m] = m < nvalid ? p[m] : R(0);
          }
#pragma unroll
          for (int m = 0; m < ZV; ++m) v[m] = v[m] + t[m] * w;
          if constexpr (MODE == 2) {
            const size_t o = (size_t)i * plane + (size_t)jr * Nz + z0;
            const double *wm = J.Wm + (size_t)al * g.P * plane + o;
            const double *pr = J.Pref + ((size_t)al * g.C + i / g.L) * plane + (size_t)jr * Nz + z0;
#pragma unroll
            for (int m = 0; m < ZV; ++m)
              if (m < nvalid) accm[m] += (double)w * (wm[m] - pr[m]);
          }
        }
      };
      if (vec) sum(std::true_type{});
      else sum(std::false_type{});
#pragma unroll
      for (int m = 0; m < ZV; ++m) acc[m] = acc[m] + v[m];
    }
    __syncthreads();  // the lists are rebuilt for the next angle
  }
  if (!lane_ok) return;
  const size_t o = ((size_t)x * N + y) * Nz + z0;
  if (vec) bp_store_vec<R, ZV, true>(J.out_mode, J.out, o, acc, nvalid, J.f, J.nrm, J.delta, J.floor_value);
  else bp_store_vec<R, ZV, false>(J.out_mode, J.out, o, acc, nvalid, J.f, J.nrm, J.delta, J.floor_value);
  if constexpr (MODE == 2) {
    R gm[ZV];
#pragma unroll
    for (int m = 0; m < ZV; ++m) gm[m] = R((double)g.dx * accm[m]);
    bp_store_vec<R, ZV>(J.out_mode, J.out_mu, o, gm, nvalid, J.f, J.nrm, J.delta, J.floor_value);
  }
}

}  // namespace spect
