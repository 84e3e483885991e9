// common.hpp -- device helpers and launch descriptors shared by the SPECT kernels.
//
// Frames.  The object f[x][y][z] (z fastest) sits in a zero-padded P x P
// transaxial grid, P = N + 2*pad (utils/spatial.py:36-81).  For detector angle
// beta the "rotated frame" V(i, j, z) is the object resampled by kornia's
// rotate(theta_f = beta - 270) (spect_system_matrix.py:150,
// transforms/shared/spatial.py:53): rotated-frame pixel (row i, column j)
// samples the padded object at
//     col = (c*xn_j - s*yn_i + 1) * (P-1)/2,  row = (s*xn_j + c*yn_i + 1) * (P-1)/2
// with xn = linspace(-1, 1, P) and (c, s) = (cos, sin)(theta_f), bilinear
// weights and zeros outside (grid_sample, align_corners=True).  The detector
// is at +i; a projection is the sum over i (spect_system_matrix.py:154).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace spect {

constexpr int kTJ = 32;  // rotated-frame columns per tile (forward and back projection)

// Most planes in one piece of the forward's balanced schedule.
constexpr int kPieceMax = 32;

// XCDs of the device; blocks are dispatched to them round-robin (block b
// runs on XCD b % kXcds).
constexpr int kXcds = 8;

// Blocks a scheduled launch runs for n angles: the resident blocks, or more
// when even kPieceMax-plane pieces outnumber them (large calls), rounded up to
// a multiple of kXcds.
__host__ __device__ inline int sched_grid_blocks(int n, int ntiles, int P, int resident) {
  const int most = n * ntiles * ((P + kPieceMax - 1) / kPieceMax);
  const int g = most > resident ? most : resident;
  return (g + kXcds - 1) / kXcds * kXcds;
}

// The piece block b of a scheduled launch of `grid` blocks (a multiple of
// kXcds) takes: each XCD works through one contiguous run of pieces -- whole
// (angle, tile) pairs, whose object and attenuation samples then stay in that
// XCD's L2 instead of being fetched by all eight.  A permutation of [0, grid).
__host__ __device__ inline int xcd_piece(int b, int grid) {
  return (b % kXcds) * (grid / kXcds) + b / kXcds;
}

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

// Thread mapping of the tile kernels: in the gather / z-blur phases a thread
// owns one rotated-frame column and RZ consecutive z values; in the y phase it
// owns one z value and 1/QS of the tile's rows (back projection) or output
// columns (forward).
// Forward blocks are 32 * ceil(Nz / RZ) threads under a 512-thread launch
// bound: RZ = 16 keeps Nz <= 256 launchable in both precisions.
template <typename R> struct FwdTile {
  static constexpr int RZ = 16;
  static constexpr int QS = 2;
};
template <typename R> struct BpTile {
  static constexpr int RZ = 8;
  static constexpr int QS = 4;
};

// Device tables of a plan in one precision.
template <typename R> struct Tables {
  const typename V2<R>::T *trig;  // [A] (cos, sin) of theta_f = beta - 270
  const R *lin;                   // [P] linspace(-1, 1, P)
  const int32_t *radius_index;    // [A] -> PSF table row block
  const R *taps_r;                // [n_radii][P][K] taps along r (y), K = 2H+1
  const R *taps_z;                // [n_radii][P][K] taps along z (a unit tap when '1D')
  const int32_t *tier;            // [n_radii][P] half width each plane's blur runs with
  // [A][ntiles][C] planes of each forward block that can touch the object
  // (bit l: plane c * L + l), or null: the forward kernel tests them itself
  const unsigned long long *act;
  // [A][ntiles] (i_min, i_max) of the planes each (angle, tile) visits
  // (i_max < i_min: none), the balanced schedule's input
  const int2 *ranges;
};

// Geometry of one call.
template <typename R> struct Geo {
  int N, Nz, P, pad, A;
  int H;          // PSF half width (0: no PSF)
  int je0, je1;   // rotated-frame columns that can be non-zero after the y blur
  int jo0, jo1;   // detector columns kept by unpad_proj
  int L, C;       // planes per chunk, chunks
  int ntiles;     // ceil((je1 - je0) / kTJ)
  R half;         // (P - 1) / 2
  R dx;           // attenuation step (cm)
  R lo, hi;       // footprint test: a sample matters iff lo < col, row < hi
};

// Bilinear source of one rotated-frame pixel: rows/cols in unpadded object
// coordinates (may be out of range) plus fractional parts.
template <typename R> struct Src {
  int r0, c0;
  R tx, ty;
};

// affine_grid with theta = [[c, -s, 0], [s, c, 0]] followed by grid_sample's
// unnormalisation (g + 1) * (size - 1) / 2: column ix and row iy (padded px).
//
// No floating-point contraction here or in make_src: every coordinate is a
// rounded value of R, so the fractional part ix - floor(ix) is exact and lies
// in [0, 1).  If the product (gx + 1) * half were fused into that subtraction
// (an FMA of the unrounded product minus floor of the rounded one), a sample
// that lands just below an integer -- as at the quarter-turn angles -- would
// get a fractional part down to about -6e-6 and a negative bilinear weight,
// and a non-negative object could project to slightly negative values.
template <typename R>
__device__ __forceinline__ void rot_coord(R c, R s, R xn, R yn, R half, R &ix, R &iy) {
#pragma clang fp contract(off)
  const R gx = c * xn - s * yn;
  const R gy = s * xn + c * yn;
  ix = (gx + R(1)) * half;
  iy = (gy + R(1)) * half;
}

template <typename R>
__device__ __forceinline__ Src<R> make_src(R ix, R iy, int pad) {
#pragma clang fp contract(off)
  const R fx = floor(ix), fy = floor(iy);
  Src<R> s;
  s.c0 = int(fx) - pad;
  s.r0 = int(fy) - pad;
  s.tx = ix - fx;
  s.ty = iy - fy;
  return s;
}

// Load through the constant address space: a wave-uniform address becomes a
// scalar (s_load) load.  Only for tables the kernels never write.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T *p) {
  return *(const __attribute__((address_space(4))) T *)(p);
}
// Whole descriptor struct (plain data) through scalar loads.
template <typename T>
__device__ __forceinline__ T ld_uniform_struct(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size must be a multiple of 4");
  union {
    T t;
    uint32_t w[sizeof(T) / 4];
  } u;
  const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) u.w[k] = ld_uniform(q + k);
  return u.t;
}

// Index of grid slice z taken centre-out (mid, mid+1, mid-1, ...): the
// chunks through the middle of the padded frame, which hold the most planes
// inside the object, and the central column tiles are dispatched first and
// the light edge ones last, so the tail of a launch is made of short blocks.
// A bijection of [0, C).
__device__ __forceinline__ int centre_out(int z, int C) {
  const int mid = (C - 1) / 2;
  return (z & 1) ? mid + (z + 1) / 2 : mid - z / 2;
}

// Index of the job owning block b: starts[] ascending, starts[0] == 0.
__device__ __forceinline__ int find_job(const int *starts, int njobs, int b) {
  int j = 0;
  for (int k = 1; k < njobs; ++k)
    if (ld_uniform(starts + k) <= b) j = k;
  return j;
}

__device__ __forceinline__ float2 ld_uniform(const float2 *p) {
  const float *q = reinterpret_cast<const float *>(p);
  return make_float2(ld_uniform(q), ld_uniform(q + 1));
}
__device__ __forceinline__ double2 ld_uniform(const double2 *p) {
  const double *q = reinterpret_cast<const double *>(p);
  return make_double2(ld_uniform(q), ld_uniform(q + 1));
}

// Bilinear taps in grid_sample's CPU order: nw=(1-ty)(1-tx), ne=(1-ty)tx,
// sw=ty(1-tx), se=ty*tx.  Taps outside the object get weight zero and an
// offset clamped onto a valid voxel, so the loads need no branches.
template <typename R> struct Taps4 {
  uint32_t o[4];  // element offsets of (row, col, 0)
  R w[4];
};

template <typename R>
__device__ __forceinline__ Taps4<R> taps4(const Src<R> &s, int N, int Nz) {
  Taps4<R> t;
  const R ex = R(1) - s.tx, ny = R(1) - s.ty;
  const bool r0 = (unsigned)s.r0 < (unsigned)N, r1 = (unsigned)(s.r0 + 1) < (unsigned)N;
  const bool c0 = (unsigned)s.c0 < (unsigned)N, c1 = (unsigned)(s.c0 + 1) < (unsigned)N;
  const int rr0 = min(max(s.r0, 0), N - 1), rr1 = min(max(s.r0 + 1, 0), N - 1);
  const int cc0 = min(max(s.c0, 0), N - 1), cc1 = min(max(s.c0 + 1, 0), N - 1);
  t.o[0] = (uint32_t)(rr0 * N + cc0) * (uint32_t)Nz;
  t.o[1] = (uint32_t)(rr0 * N + cc1) * (uint32_t)Nz;
  t.o[2] = (uint32_t)(rr1 * N + cc0) * (uint32_t)Nz;
  t.o[3] = (uint32_t)(rr1 * N + cc1) * (uint32_t)Nz;
  t.w[0] = (r0 && c0) ? ny * ex : R(0);
  t.w[1] = (r0 && c1) ? ny * s.tx : R(0);
  t.w[2] = (r1 && c0) ? s.ty * ex : R(0);
  t.w[3] = (r1 && c1) ? s.ty * s.tx : R(0);
  return t;
}

template <typename R>
__device__ __forceinline__ R interp(const R *__restrict__ vol, const Taps4<R> &t, int z) {
  R acc = vol[t.o[0] + z] * t.w[0];
  acc = acc + vol[t.o[1] + z] * t.w[1];
  acc = acc + vol[t.o[2] + z] * t.w[2];
  acc = acc + vol[t.o[3] + z] * t.w[3];
  return acc;
}

// Gather RZ consecutive z of one rotated-frame pixel (bilinear); 16-byte
// vector loads when the run is complete and Nz is a multiple of 4.
template <typename R, int RZ>
__device__ __forceinline__ void gather_run(const R *__restrict__ vol, const Taps4<R> &t, int z0,
                                           int Nz, R (&out)[RZ]) {
  if (z0 + RZ <= Nz && (Nz % 4) == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const R *p = vol + t.o[q] + z0;
      R v[RZ];
      if constexpr (sizeof(R) == 4) {
#pragma unroll
        for (int m = 0; m < RZ; m += 4) {
          const float4 x = *reinterpret_cast<const float4 *>(p + m);
          v[m] = x.x; v[m + 1] = x.y; v[m + 2] = x.z; v[m + 3] = x.w;
        }
      } else {
#pragma unroll
        for (int m = 0; m < RZ; m += 2) {
          const double2 x = *reinterpret_cast<const double2 *>(p + m);
          v[m] = x.x; v[m + 1] = x.y;
        }
      }
      // ((nw + ne) + sw) + se, as interp()
#pragma unroll
      for (int m = 0; m < RZ; ++m) out[m] = q == 0 ? v[m] * t.w[0] : out[m] + v[m] * t.w[q];
    }
  } else {
#pragma unroll
    for (int m = 0; m < RZ; ++m) out[m] = (z0 + m < Nz) ? interp<R>(vol, t, z0 + m) : R(0);
  }
}

// ZV consecutive z of one object pixel column starting at z0 (one 16-byte
// load when Nz is a multiple of ZV, else element loads within Nz).
template <typename R, int ZV>
__device__ __forceinline__ void load_zv(const R *__restrict__ p, int z0, int Nz, bool vec,
                                        R (&v)[ZV]) {
  if (vec) {
    if constexpr (sizeof(R) == 4) {
      const float4 x = *reinterpret_cast<const float4 *>(p + z0);
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    } else {
      const double2 x = *reinterpret_cast<const double2 *>(p + z0);
      v[0] = x.x; v[1] = x.y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < ZV; ++e) v[e] = z0 + e < Nz ? p[z0 + e] : R(0);
  }
}

// RZ consecutive z of one row starting at z0 (16-byte loads when the run is
// complete and Nz is a multiple of 4, else element loads within Nz).
template <typename R, int RZ>
__device__ __forceinline__ void load_run(const R *__restrict__ p, int z0, int Nz, R (&v)[RZ]) {
  constexpr int ZV = 16 / sizeof(R);
  if (z0 + RZ <= Nz && (Nz % 4) == 0) {
#pragma unroll
    for (int m = 0; m < RZ; m += ZV) {
      R t[ZV];
      load_zv<R, ZV>(p, z0 + m, Nz, true, t);
#pragma unroll
      for (int e = 0; e < ZV; ++e) v[m + e] = t[e];
    }
  } else {
#pragma unroll
    for (int m = 0; m < RZ; ++m) v[m] = z0 + m < Nz ? p[z0 + m] : R(0);
  }
}

// Bilinear sample of ZV consecutive z, in interp()'s order.
template <typename R, int ZV>
__device__ __forceinline__ void interp_zv(const R *__restrict__ vol, const Taps4<R> &t, int z0,
                                          int Nz, bool vec, R (&out)[ZV]) {
  R v[4][ZV];
#pragma unroll
  for (int q = 0; q < 4; ++q) load_zv<R, ZV>(vol + t.o[q], z0, Nz, vec, v[q]);
#pragma unroll
  for (int e = 0; e < ZV; ++e) {
    R a = v[0][e] * t.w[0];
    a = a + v[1][e] * t.w[1];
    a = a + v[2][e] * t.w[2];
    a = a + v[3][e] * t.w[3];
    out[e] = a;
  }
}

// The forward kernel's phase A covers kTJ columns in NPA passes of
// blockDim / ceil(Nz / ZV) columns; blocks have >= QS * Nz threads, so
// NPA = kTJ / (QS * ZV) passes suffice (checked at launch).
template <typename R, int QS>
__host__ __device__ constexpr int phase_a_passes() {
  return kTJ / (QS * (16 / (int)sizeof(R))) > 0 ? kTJ / (QS * (16 / (int)sizeof(R))) : 1;
}
template <typename R, int QS>
inline bool phase_a_fits(int block_threads, int Nz) {
  constexpr int ZV = 16 / sizeof(R);
  const int nq = (Nz + ZV - 1) / ZV;
  return (block_threads / nq) * phase_a_passes<R, QS>() >= kTJ;
}

__device__ __forceinline__ float att_exp(float v) { return __expf(v); }
__device__ __forceinline__ double att_exp(double v) { return exp(v); }

// Reverse-cumsum attenuation factor (utils/misc.py:18, attenuation.py:21):
// x = mu_rot*dx in R; S (FP64) already includes x; a = exp(-(R(S) - x/2)).
template <typename R>
__device__ __forceinline__ R att_factor(double S, R x) {
  const R c = R(S);
  const R r = c - x / R(2);
  return att_exp(-r);
}

// Runs body(std::integral_constant<int, T>) with the blur tier T for a plane
// whose taps reach h from the centre, in a kernel compiled for half width H:
// the smallest compiled tier >= h among 2, H - 4k (k >= 1) and H (tiers.hpp);
// taps beyond T are skipped.  The comparisons are block-uniform.
template <int H, int T, typename F>
__device__ __forceinline__ void tier_dispatch(int h, F &body) {
  if constexpr (T >= H) {
    body(std::integral_constant<int, H>{});
  } else {
    if constexpr (T == 2 || (H - T) % 4 == 0) {
      if (h <= T) {
        body(std::integral_constant<int, T>{});
        return;
      }
    }
    tier_dispatch<H, T + 2>(h, body);
  }
}
template <int H, typename F>
__device__ __forceinline__ void with_tier(int h, F &&body) {
  if constexpr (H == 0) body(std::integral_constant<int, 0>{});
  else tier_dispatch<H, 2>(h, body);
}

// Does rotated-frame pixel (i, j) sample inside the footprint square?
template <typename R>
__device__ __forceinline__ bool in_footprint(R ix, R iy, const Geo<R> &g) {
  return ix > g.lo && ix < g.hi && iy > g.lo && iy < g.hi;
}

// ---------------------------------------------------------------------------
// Balanced schedule (used by the forward and the back-projection field):
// pair p = al * ntiles + jt visits the n_p planes of its range [i_min, i_max]
// in the plan's activity table; it is split into c_p = ceil(n_p / Q) pieces,
// Q the smallest plane count <= kPieceMax for which all pieces fit the
// resident blocks at once.

// Wave-level sums over the 64 lanes by DPP moves (VALU operations; no LDS
// round trips).  Called from control flow that every lane of the wave
// reaches.  dpp_move<C>: the value of the lane that DPP control C selects;
// 0 where that lane lies outside the row of 16.
template <int C>
__device__ __forceinline__ int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(0, v, C, 0xF, 0xF, false);
}

// Sum over the 64 lanes of a wave (every lane gets it): two butterflies
// within quads (quad_perm [1,0,3,2], [2,3,0,1]) and row rotations by 4 and 8
// leave every lane with its row's total; then the four row totals.
__device__ __forceinline__ int wave_sum(int v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x124>(v);
  v += dpp_move<0x128>(v);
  return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) +
         __builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48);
}

// Inclusive prefix sum over the lanes of a wave: Hillis-Steele within each row
// (row_shr by 1, 2, 4, 8), then the totals of the rows before.
__device__ __forceinline__ int wave_scan(int v) {
  v += dpp_move<0x111>(v);
  v += dpp_move<0x112>(v);
  v += dpp_move<0x114>(v);
  v += dpp_move<0x118>(v);
  const int lane = threadIdx.x & 63;
  const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31),
            r2 = __builtin_amdgcn_readlane(v, 47);
  return v + (lane >= 16 ? r0 : 0) + (lane >= 32 ? r1 : 0) + (lane >= 48 ? r2 : 0);
}

// Piece of one block: pair p (< 0: the block has none), its index k within
// the pair and the pair's piece count c.
struct SchedLoc {
  int p, k, c;
};

// The schedule for nblocks resident blocks, evaluated by the whole block
// (blockDim.x a multiple of 64, at most 1024; every thread calls), located at
// piece b.  With `out`, the block of piece 0 writes the schedule there.
template <typename R>
__device__ __forceinline__ SchedLoc sched_locate(const Geo<R> &g, const int32_t *__restrict__ angle_idx,
                                                 const int2 *__restrict__ ranges, int n, int nblocks,
                                                 int b, int *__restrict__ out) {
  __shared__ int red[16 * 8];  // per-wave partial sums
  __shared__ int loc[3];
  const int tid = threadIdx.x, bd = blockDim.x;
  const int w = tid >> 6, lane = tid & 63, nw = bd >> 6;
  const int npairs = n * g.ntiles;
  // thread tid owns the contiguous pairs [p0, p1)
  const int per = (npairs + bd - 1) / bd;
  const int p0 = min(npairs, tid * per), p1 = min(npairs, p0 + per);
  auto plen = [&](int p) -> int {
    const int al = p / g.ntiles, jt = p - al * g.ntiles;
    const int ang = angle_idx[al];
    if ((unsigned)ang >= (unsigned)g.A) return 0;
    const int2 r = ranges[(size_t)ang * g.ntiles + jt];
    return r.y >= r.x ? r.y - r.x + 1 : 0;
  };
  // the first kKeep lengths stay in registers; threads (whole waves, when
  // the pairs are fewer than the threads) without pairs skip every pass
  constexpr int kKeep = 4;
  int keep[kKeep];
#pragma unroll
  for (int i = 0; i < kKeep; ++i) keep[i] = p0 + i < p1 ? plen(p0 + i) : 0;
  auto each = [&](auto fn) {
#pragma unroll
    for (int i = 0; i < kKeep; ++i)
      if (p0 + i < p1) fn(p0 + i, keep[i]);
    for (int p = p0 + kKeep; p < p1; ++p) fn(p, plen(p));
  };
  // ceil(len / q) = floor((a + 0.5) / q), a = len + q - 1, in float: the
  // quotient lies at least 0.5 / q from an integer and the two roundings
  // (1 / q, the product) move it by under (a / q + 1) 2^-23 -- exact for
  // a < 2^20 (len <= P).  No integer division on the hot start of a block.
  auto pieces = [](int len, int q, float inv_q) {
    return (int)(((float)(len + q - 1) + 0.5f) * inv_q);
  };

  // T = sum of n_p
  int t = 0;
  each([&](int, int len) { t += len; });
  t = wave_sum(t);
  if (lane == 0) red[w] = t;
  if (tid == 0) loc[0] = -1;
  __syncthreads();
  int T = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (i < nw) T += red[i];
  __syncthreads();
  // smallest Q in [ceil(T / nblocks), kPieceMax] whose pieces fit nblocks,
  // eight candidates per pass (nearly always one pass)
  int q = kPieceMax;
  for (int q0 = max(1, min(kPieceMax, (T + nblocks - 1) / nblocks)); q0 <= kPieceMax; q0 += 8) {
    int s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] = 0;
    if (p0 < p1) {
      float inv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) inv[k] = 1.0f / (float)min(q0 + k, kPieceMax);
      each([&](int, int len) {
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += pieces(len, min(q0 + k, kPieceMax), inv[k]);
      });
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      s[k] = wave_sum(s[k]);
      if (lane == 0) red[w * 8 + k] = s[k];
    }
    __syncthreads();
    int found = 0;  // 1 + the first fitting candidate
#pragma unroll
    for (int k = 7; k >= 0; --k) {
      int S = 0;
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i < nw) S += red[i * 8 + k];
      if (S <= nblocks) found = k + 1;
    }
    __syncthreads();
    if (found) {
      q = min(q0 + found - 1, kPieceMax);
      break;
    }
  }
  // pieces before each thread's pairs, then the owner of piece b
  const float inv_q = 1.0f / (float)q;
  int cnt = 0;
  each([&](int, int len) { cnt += pieces(len, q, inv_q); });
  const int incl = wave_scan(cnt);
  if (lane == 63) red[w] = incl;
  __syncthreads();
  int pre = 0, total = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int v = i < nw ? red[i] : 0;
    if (i < w) pre += v;
    total += v;
  }
  int b0 = pre + incl - cnt;
  const bool write = out != nullptr && b == 0;
  if (write || (b >= b0 && b < b0 + cnt)) {
    each([&](int p, int len) {
      if (p >= p1) return;
      const int c = pieces(len, q, inv_q);
      if (write) out[2 + p] = b0;
      if (b >= b0 && b < b0 + c) {
        loc[0] = p;
        loc[1] = b - b0;
        loc[2] = c;
      }
      b0 += c;
    });
  }
  if (write && tid == 0) {
    out[0] = total;
    out[1] = q;
    out[2 + npairs] = total;
  }
  __syncthreads();
  return {__builtin_amdgcn_readfirstlane(loc[0]), __builtin_amdgcn_readfirstlane(loc[1]),
          __builtin_amdgcn_readfirstlane(loc[2])};
}

// Planes of piece k of c over the range r (walked from r.y down).
__device__ __forceinline__ void sched_piece_planes(int2 r, int k, int c, int &i_lo, int &i_hi) {
  const int np = r.y - r.x + 1;
  const int a = (int)((long long)k * np / c), e = (int)((long long)(k + 1) * np / c);
  i_hi = r.y - a + 1;
  i_lo = r.y - (e - 1);
}

}  // namespace spect
