// spect_kernels.hip -- gfx950 kernels and C ABI of libspect_hip.so.
//
// The SPECT rotate+sum projector of PyTomography 3.3.2
// (src/pytomography/projectors/SPECT/spect_system_matrix.py:120-201) done as
// a few fused, stream-ordered kernels instead of ~40 ATen launches per angle.
//
// Frames.  The object f[x][y][z] (z fastest) sits in a zero-padded P x P
// transaxial grid, P = N + 2*pad (utils/spatial.py:36-81).  For detector angle
// beta the "rotated frame" V(i, j, z) is the object resampled by kornia's
// rotate(theta_f = beta - 270): the rotated-frame pixel (row i, column j)
// samples the padded object at
//     col = (c*xn_j - s*yn_i + 1) * (P-1)/2,  row = (s*xn_j + c*yn_i + 1) * (P-1)/2
// (xn = linspace(-1, 1, P); c, s = cos, sin of theta_f in radians) with
// bilinear weights and zeros outside (grid_sample, align_corners=True).  The
// detector is at +i; a projection sums over i (spect_system_matrix.py:154).
//
// Kernels (R = real type, float or double):
//   att_sums_kernel  : per (angle, column tile, plane chunk) the FP64 sum of
//                      mu_rot*dx over the chunk -- the starting value of the
//                      reverse cumsum of utils/misc.py:18 for the chunk below.
//   fwd_kernel       : per (angle, column tile, plane chunk) walk planes from
//                      the detector side: gather f and mu, FP64 running
//                      attenuation sum, exp, y-blur then z-blur with the
//                      plane's taps (psf.py:27-41) staged in LDS, accumulate
//                      the ray-sum in registers; write one partial per chunk.
//   fwd_reduce_kernel: sum the chunk partials in a fixed order (deterministic).
//   bww_kernel       : back projection field W(i,j,z) = a(i,j,z) * (G_i * p)(j,z)
//                      (psf.py:216-231 then attenuation.py:94-115) per angle.
//   bw_ref_kernel    : gather W at 270-beta (spect_system_matrix.py:196), sum
//                      over the angles, crop, then store / add / OSEM update.
//   bw_adj_kernel    : exact transpose of the forward bilinear gather, written
//                      as a gather (race-free, deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spect_hip.h"

#define SPECT_VERSION 1

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(SPECT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

constexpr int kThreads = 256;

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

// Bilinear source of one rotated-frame pixel: rows/cols in unpadded object
// coordinates (may be out of range) plus fractional parts.
template <typename R> struct Src {
  int r0, c0;
  R tx, ty;
};

// Kornia/grid_sample coordinate of rotated-frame pixel (i, j) for a rotation
// with cosine c and sine s: returns column (ix) and row (iy) in padded pixels.
template <typename R>
__device__ __forceinline__ void rot_coord(R c, R s, R xn, R yn, R half, R &ix, R &iy) {
  // affine_grid: grid = [xn, yn, 1] @ theta^T with theta = [[c, -s, 0], [s, c, 0]]
  R gx = c * xn - s * yn;
  R gy = s * xn + c * yn;
  // grid_sample unnormalize (align_corners=True): (g + 1) * (size - 1) / 2
  ix = (gx + R(1)) * half;
  iy = (gy + R(1)) * half;
}

template <typename R>
__device__ __forceinline__ Src<R> make_src(R ix, R iy, int pad) {
  R fx = floor(ix), fy = floor(iy);
  Src<R> s;
  s.c0 = int(fx) - pad;
  s.r0 = int(fy) - pad;
  s.tx = ix - fx;
  s.ty = iy - fy;
  return s;
}

// Bilinear sample of a [N][N][Nz] volume at Src for slice z; zero outside.
// Weights and summation order follow grid_sample's CPU kernel:
// nw=(1-ty)(1-tx), ne=(1-ty)tx, sw=ty(1-tx), se=ty*tx, ((nw+ne)+sw)+se.
template <typename R>
__device__ __forceinline__ R bilinear(const R *__restrict__ vol, const Src<R> &s, int z, int N,
                                      int Nz) {
  const R ex = R(1) - s.tx, ny = R(1) - s.ty;
  const bool r0ok = (unsigned)s.r0 < (unsigned)N, r1ok = (unsigned)(s.r0 + 1) < (unsigned)N;
  const bool c0ok = (unsigned)s.c0 < (unsigned)N, c1ok = (unsigned)(s.c0 + 1) < (unsigned)N;
  const size_t base = ((size_t)s.r0 * N + s.c0) * Nz + z;
  R vnw = (r0ok && c0ok) ? vol[base] : R(0);
  R vne = (r0ok && c1ok) ? vol[base + Nz] : R(0);
  R vsw = (r1ok && c0ok) ? vol[base + (size_t)N * Nz] : R(0);
  R vse = (r1ok && c1ok) ? vol[base + (size_t)N * Nz + Nz] : R(0);
  R acc = vnw * (ny * ex);
  acc = acc + vne * (ny * s.tx);
  acc = acc + vsw * (s.ty * ex);
  acc = acc + vse * (s.ty * s.tx);
  return acc;
}

// exp for the attenuation factor in the run's precision.
__device__ __forceinline__ float att_exp(float v) { return expf(v); }
__device__ __forceinline__ double att_exp(double v) { return exp(v); }

// Reverse-cumsum attenuation factor (utils/misc.py:18, attenuation.py:21):
// x = mu_rot*dx (in R), S += x in FP64, a = exp(-(R(S) - x/2)).
template <typename R>
__device__ __forceinline__ R att_factor(double S, R x) {
  R c = R(S);
  R r = c - x / R(2);
  return att_exp(-r);
}

template <typename R> struct Tables {
  const typename V2<R>::T *trig;  // [A] (cos, sin) of theta_f = beta - 270
  const R *lin;                   // [P] linspace(-1, 1, P)
  const int32_t *radius_index;    // [A]
  const R *taps_r;                // [n_radii][P][kr]
  const R *taps_z;                // [n_radii][P][kz]
};

template <typename R> struct Shape {
  int N, Nz, P, pad;
  int A;               // number of angles of the plan (valid angle indices are [0, A))
  int kr, kz, hy, hz;  // taps and half widths (kr == 0: no PSF; kz == 0: no z blur)
  R half;              // (P - 1) / 2
  R dx;                // attenuation step (cm)
};

// ---------------------------------------------------------------------------
// Attenuation chunk sums: T[a][c][jx][z] = sum_{i in chunk c} R(mu_rot(i,j,z)*dx)
// accumulated in FP64, for columns j = je0 + jx in [je0, je1).
// ---------------------------------------------------------------------------
template <typename R, int NI>
__global__ __launch_bounds__(kThreads) void att_sums_kernel(
    Shape<R> sh, Tables<R> tb, const R *__restrict__ mu, const int32_t *__restrict__ angle_idx,
    int je0, int je1, int TJ, int L, int n_chunks, double *__restrict__ sums) {
  __shared__ Src<R> src[64];
  const int al = blockIdx.x, jt = blockIdx.y, c = blockIdx.z;
  const int ang = angle_idx[al];
  if ((unsigned)ang >= (unsigned)sh.A) return;  // invalid index: whole block exits
  const int j0 = je0 + jt * TJ;
  const int tj = min(TJ, je1 - j0);
  const int n_items = tj * sh.Nz;
  const auto cs = tb.trig[ang];
  double S[NI];
#pragma unroll
  for (int k = 0; k < NI; ++k) S[k] = 0.0;
  const int i_lo = c * L, i_hi = min(sh.P, i_lo + L);
  for (int i = i_hi - 1; i >= i_lo; --i) {
    for (int q = threadIdx.x; q < tj; q += kThreads) {
      R ix, iy;
      rot_coord<R>(cs.x, cs.y, tb.lin[j0 + q], tb.lin[i], sh.half, ix, iy);
      src[q] = make_src<R>(ix, iy, sh.pad);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NI; ++k) {
      const int idx = threadIdx.x + k * kThreads;
      if (idx < n_items) {
        const int jl = idx / sh.Nz, z = idx - jl * sh.Nz;
        const R m = bilinear<R>(mu, src[jl], z, sh.N, sh.Nz);
        S[k] += (double)(m * sh.dx);
      }
    }
    __syncthreads();
  }
  const int n_jext = je1 - je0;
#pragma unroll
  for (int k = 0; k < NI; ++k) {
    const int idx = threadIdx.x + k * kThreads;
    if (idx < n_items) {
      const int jl = idx / sh.Nz, z = idx - jl * sh.Nz;
      sums[(((size_t)al * n_chunks + c) * n_jext + (j0 - je0 + jl)) * sh.Nz + z] = S[k];
    }
  }
}

// Starting FP64 suffix sum for chunk c: sum of chunk sums above it, top down.
__device__ __forceinline__ double att_start(const double *__restrict__ sums, int al, int c,
                                            int n_chunks, int n_jext, int jx, int z, int Nz) {
  double s = 0.0;
  for (int cc = n_chunks - 1; cc > c; --cc)
    s += sums[(((size_t)al * n_chunks + cc) * n_jext + jx) * Nz + z];
  return s;
}

// ---------------------------------------------------------------------------
// Forward projection, one (angle, output column tile, plane chunk) per block.
// LDS: V[(TJ+2hy)][Nz] attenuated rotated object, U[TJ][Nz] after the y blur.
// ---------------------------------------------------------------------------
template <typename R, int NA, int NC>
__global__ __launch_bounds__(kThreads) void fwd_kernel(
    Shape<R> sh, Tables<R> tb, const R *__restrict__ f, const R *__restrict__ mu,
    const int32_t *__restrict__ angle_idx, int jo0, int jo1, int TJ, int L, int n_chunks,
    int je0, int je1, const double *__restrict__ sums, R *__restrict__ partial, int n_call) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int al = blockIdx.x, jt = blockIdx.y, c = blockIdx.z;
  const int ang = angle_idx[al];
  if ((unsigned)ang >= (unsigned)sh.A) return;  // invalid index: whole block exits
  const int Nz = sh.Nz, hy = sh.hy;
  const int j0 = jo0 + jt * TJ;
  const int tj = min(TJ, jo1 - j0);
  const int wj = tj + 2 * hy;
  R *V = reinterpret_cast<R *>(smem);                  // [(TJ + 2hy)][Nz]
  R *U = V + (size_t)(TJ + 2 * hy) * Nz;               // [TJ][Nz]
  Src<R> *src = reinterpret_cast<Src<R> *>(U + (size_t)TJ * Nz);  // [TJ + 2hy]
  const auto cs = tb.trig[ang];
  const int nA = wj * Nz, nC = tj * Nz;
  const int n_jext = je1 - je0;
  const bool has_mu = (mu != nullptr);

  double S[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    S[k] = 0.0;
    const int idx = threadIdx.x + k * kThreads;
    if (has_mu && idx < nA) {
      const int jl = idx / Nz, z = idx - jl * Nz;
      const int j = j0 - hy + jl;
      if (j >= je0 && j < je1) S[k] = att_start(sums, al, c, n_chunks, n_jext, j - je0, z, Nz);
    }
  }
  R acc[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k] = R(0);

  const int rix = tb.radius_index[ang];
  const int i_lo = c * L, i_hi = min(sh.P, i_lo + L);
  for (int i = i_hi - 1; i >= i_lo; --i) {
    // A0: bilinear sources of the wj columns of plane i
    for (int q = threadIdx.x; q < wj; q += kThreads) {
      const int j = j0 - hy + q;
      Src<R> s;
      if (j >= 0 && j < sh.P) {
        R ix, iy;
        rot_coord<R>(cs.x, cs.y, tb.lin[j], tb.lin[i], sh.half, ix, iy);
        s = make_src<R>(ix, iy, sh.pad);
      } else {  // column outside the padded grid: never sampled
        s.r0 = -2; s.c0 = -2; s.tx = R(0); s.ty = R(0);
      }
      src[q] = s;
    }
    __syncthreads();
    // A: V = a * F
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      const int idx = threadIdx.x + k * kThreads;
      if (idx < nA) {
        const int jl = idx / Nz, z = idx - jl * Nz;
        const Src<R> s = src[jl];
        R v = bilinear<R>(f, s, z, sh.N, Nz);
        if (has_mu) {
          const R m = bilinear<R>(mu, s, z, sh.N, Nz) * sh.dx;
          S[k] += (double)m;
          v = v * att_factor<R>(S[k], m);
        }
        V[idx] = v;
      }
    }
    __syncthreads();
    if (sh.kr > 0) {
      // B: y blur (correlation, zero padded) with this plane's taps
      const R *wy = tb.taps_r + ((size_t)rix * sh.P + i) * sh.kr;
      for (int idx = threadIdx.x; idx < nC; idx += kThreads) {
        const int jl = idx / Nz, z = idx - jl * Nz;
        const R *col = V + (size_t)jl * Nz + z;
        R u = R(0);
        for (int t = 0; t < sh.kr; ++t) u += wy[t] * col[(size_t)t * Nz];
        U[idx] = u;
      }
      __syncthreads();
      // C: z blur, accumulate the ray sum
      if (sh.kz > 0) {
        const R *wz = tb.taps_z + ((size_t)rix * sh.P + i) * sh.kz;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          const int idx = threadIdx.x + k * kThreads;
          if (idx < nC) {
            const int jl = idx / Nz, z = idx - jl * Nz;
            const R *row = U + (size_t)jl * Nz;
            R u = R(0);
            const int t0 = max(0, sh.hz - z), t1 = min(sh.kz, Nz + sh.hz - z);
            for (int t = t0; t < t1; ++t) u += wz[t] * row[z + t - sh.hz];
            acc[k] += u;
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          const int idx = threadIdx.x + k * kThreads;
          if (idx < nC) acc[k] += U[idx];
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const int idx = threadIdx.x + k * kThreads;
        if (idx < nC) acc[k] += V[idx];
      }
    }
    // no barrier needed here: the next plane writes src/V only after every
    // thread has passed this plane's second barrier, and U after its first.
  }
  const int n_out = jo1 - jo0;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int idx = threadIdx.x + k * kThreads;
    if (idx < nC) {
      const int jl = idx / Nz, z = idx - jl * Nz;
      partial[(((size_t)c * n_call + al) * n_out + (j0 - jo0 + jl)) * Nz + z] = acc[k];
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kThreads) void fwd_reduce_kernel(const R *__restrict__ partial,
                                                               R *__restrict__ proj, size_t count,
                                                               int n_chunks) {
  const size_t idx = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= count) return;
  R s = partial[idx];
  for (int c = 1; c < n_chunks; ++c) s += partial[(size_t)c * count + idx];
  proj[idx] = s;
}

// ---------------------------------------------------------------------------
// Back projection field W[a][i][jx][z] = a(i,j,z) * (G_i * p)(j,z) for
// columns j = je0 + jx; p is the unpadded projection [n][N][Nz] (columns
// jo0..jo1 of the padded frame).
// ---------------------------------------------------------------------------
template <typename R, int NC>
__global__ __launch_bounds__(kThreads) void bww_kernel(
    Shape<R> sh, Tables<R> tb, const R *__restrict__ proj, const R *__restrict__ mu,
    const int32_t *__restrict__ angle_idx, int jo0, int jo1, int je0, int je1, int TJ, int L,
    int n_chunks, const double *__restrict__ sums, R *__restrict__ W) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int al = blockIdx.x, jt = blockIdx.y, c = blockIdx.z;
  const int ang = angle_idx[al];
  if ((unsigned)ang >= (unsigned)sh.A) return;  // invalid index: whole block exits
  const int Nz = sh.Nz, hy = sh.hy;
  const int j0 = je0 + jt * TJ;
  const int tj = min(TJ, je1 - j0);
  const int wj = tj + 2 * hy;
  R *Rp = reinterpret_cast<R *>(smem);                 // [(TJ + 2hy)][Nz]
  R *U = Rp + (size_t)(TJ + 2 * hy) * Nz;              // [TJ][Nz]
  Src<R> *src = reinterpret_cast<Src<R> *>(U + (size_t)TJ * Nz);  // [TJ]
  const auto cs = tb.trig[ang];
  const int nC = tj * Nz;
  const int n_jext = je1 - je0;
  const int n_out = jo1 - jo0;
  const bool has_mu = (mu != nullptr);

  // stage projection rows j0-hy .. j0+tj+hy (zero outside the detector)
  for (int idx = threadIdx.x; idx < wj * Nz; idx += kThreads) {
    const int jl = idx / Nz, z = idx - jl * Nz;
    const int j = j0 - hy + jl;
    Rp[idx] = (j >= jo0 && j < jo1) ? proj[((size_t)al * n_out + (j - jo0)) * Nz + z] : R(0);
  }
  double S[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    S[k] = 0.0;
    const int idx = threadIdx.x + k * kThreads;
    if (has_mu && idx < nC) {
      const int jl = idx / Nz, z = idx - jl * Nz;
      S[k] = att_start(sums, al, c, n_chunks, n_jext, j0 - je0 + jl, z, Nz);
    }
  }
  __syncthreads();
  const int rix = tb.radius_index[ang];
  const int i_lo = c * L, i_hi = min(sh.P, i_lo + L);
  for (int i = i_hi - 1; i >= i_lo; --i) {
    if (has_mu) {
      for (int q = threadIdx.x; q < tj; q += kThreads) {
        R ix, iy;
        rot_coord<R>(cs.x, cs.y, tb.lin[j0 + q], tb.lin[i], sh.half, ix, iy);
        src[q] = make_src<R>(ix, iy, sh.pad);
      }
    }
    const R *Bsrc = Rp + (size_t)hy * Nz;  // rows aligned with the tile when no PSF
    if (sh.kr > 0) {
      const R *wy = tb.taps_r + ((size_t)rix * sh.P + i) * sh.kr;
      for (int idx = threadIdx.x; idx < nC; idx += kThreads) {
        const int jl = idx / Nz, z = idx - jl * Nz;
        const R *col = Rp + (size_t)jl * Nz + z;
        R u = R(0);
        for (int t = 0; t < sh.kr; ++t) u += wy[t] * col[(size_t)t * Nz];
        U[idx] = u;
      }
      Bsrc = U;
    }
    __syncthreads();
    const R *wz = (sh.kr > 0 && sh.kz > 0) ? tb.taps_z + ((size_t)rix * sh.P + i) * sh.kz : nullptr;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int idx = threadIdx.x + k * kThreads;
      if (idx < nC) {
        const int jl = idx / Nz, z = idx - jl * Nz;
        const R *row = Bsrc + (size_t)jl * Nz;
        R b;
        if (wz) {
          b = R(0);
          const int t0 = max(0, sh.hz - z), t1 = min(sh.kz, Nz + sh.hz - z);
          for (int t = t0; t < t1; ++t) b += wz[t] * row[z + t - sh.hz];
        } else {
          b = row[z];
        }
        if (has_mu) {
          const R m = bilinear<R>(mu, src[jl], z, sh.N, Nz) * sh.dx;
          S[k] += (double)m;
          b = b * att_factor<R>(S[k], m);
        }
        W[(((size_t)al * sh.P + i) * n_jext + (j0 - je0 + jl)) * Nz + z] = b;
      }
    }
    __syncthreads();
  }
}

// Output stage of a back projection for one voxel value b.
template <typename R>
__device__ __forceinline__ void bp_store(int out_mode, R *__restrict__ out, size_t o, R b,
                                         const R *__restrict__ f, const R *__restrict__ nrm,
                                         R delta, R floor_value) {
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

// ---------------------------------------------------------------------------
// Reference back projection gather: rotation by theta_b = 270 - beta, i.e.
// (cos, sin) = (c_f, -s_f).  One wave per object column (x, y); lanes run
// along z.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kThreads) void bw_ref_kernel(
    Shape<R> sh, Tables<R> tb, const R *__restrict__ W, const int32_t *__restrict__ angle_idx,
    int n_call, int je0, int je1, int out_mode, R *__restrict__ out, const R *__restrict__ f,
    const R *__restrict__ nrm, R delta, R floor_value) {
  const int wave = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int Nz = sh.Nz, N = sh.N;
  const int zblocks = (Nz + 63) >> 6;
  const int col = wave / zblocks;
  if (col >= N * N) return;
  const int z = (wave - col * zblocks) * 64 + lane;
  if (z >= Nz) return;
  const int x = col / N, y = col - x * N;
  const int X = x + sh.pad, Y = y + sh.pad;
  const int n_jext = je1 - je0;
  const R xn = tb.lin[Y], yn = tb.lin[X];
  R acc = R(0);
  for (int al = 0; al < n_call; ++al) {
    const int ang = angle_idx[al];
    if ((unsigned)ang >= (unsigned)sh.A) continue;
    const auto cs = tb.trig[ang];
    R ix, iy;
    rot_coord<R>(cs.x, -cs.y, xn, yn, sh.half, ix, iy);
    const R fx = floor(ix), fy = floor(iy);
    const int c0 = int(fx), r0 = int(fy);
    const R tx = ix - fx, ty = iy - fy;
    const R ex = R(1) - tx, ny = R(1) - ty;
    const R *Wa = W + (size_t)al * sh.P * n_jext * Nz;
    auto tap = [&](int r, int cc) -> R {
      if ((unsigned)r >= (unsigned)sh.P || cc < je0 || cc >= je1) return R(0);
      return Wa[((size_t)r * n_jext + (cc - je0)) * Nz + z];
    };
    R v = tap(r0, c0) * (ny * ex);
    v = v + tap(r0, c0 + 1) * (ny * tx);
    v = v + tap(r0 + 1, c0) * (ty * ex);
    v = v + tap(r0 + 1, c0 + 1) * (ty * tx);
    acc = acc + v;
  }
  bp_store<R>(out_mode, out, ((size_t)x * N + y) * Nz + z, acc, f, nrm, delta, floor_value);
}

// ---------------------------------------------------------------------------
// Exact adjoint gather: transpose of the forward bilinear gather.  For object
// voxel (X, Y) (padded) enumerate the rotated-frame pixels (i, j) whose
// forward sample touches it, recomputing their coordinates with the forward's
// own arithmetic.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kThreads) void bw_adj_kernel(
    Shape<R> sh, Tables<R> tb, const R *__restrict__ W, const int32_t *__restrict__ angle_idx,
    int n_call, int je0, int je1, int out_mode, R *__restrict__ out, const R *__restrict__ f,
    const R *__restrict__ nrm, R delta, R floor_value) {
  const int wave = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int Nz = sh.Nz, N = sh.N;
  const int zblocks = (Nz + 63) >> 6;
  const int col = wave / zblocks;
  if (col >= N * N) return;
  const int z = (wave - col * zblocks) * 64 + lane;
  if (z >= Nz) return;
  const int x = col / N, y = col - x * N;
  const int X = x + sh.pad, Y = y + sh.pad;
  const int n_jext = je1 - je0;
  R acc = R(0);
  for (int al = 0; al < n_call; ++al) {
    const int ang = angle_idx[al];
    if ((unsigned)ang >= (unsigned)sh.A) continue;
    const auto cs = tb.trig[ang];
    // inverse rotation of (row X, col Y): rotation by -theta_f
    R jc, ic;
    rot_coord<R>(cs.x, -cs.y, tb.lin[Y], tb.lin[X], sh.half, jc, ic);
    const int ib = int(floor(ic - R(1.5))), jb = int(floor(jc - R(1.5)));
    const R *Wa = W + (size_t)al * sh.P * n_jext * Nz;
    R v = R(0);
    for (int di = 0; di < 4; ++di) {
      const int i = ib + di;
      if ((unsigned)i >= (unsigned)sh.P) continue;
      for (int dj = 0; dj < 4; ++dj) {
        const int j = jb + dj;
        if (j < je0 || j >= je1) continue;
        R ix, iy;
        rot_coord<R>(cs.x, cs.y, tb.lin[j], tb.lin[i], sh.half, ix, iy);
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
        v = v + Wa[((size_t)i * n_jext + (j - je0)) * Nz + z] * (wr * wc);
      }
    }
    acc = acc + v;
  }
  bp_store<R>(out_mode, out, ((size_t)x * N + y) * Nz + z, acc, f, nrm, delta, floor_value);
}

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kThreads) void ratio_kernel(const R *__restrict__ g,
                                                         const R *__restrict__ hf,
                                                         const R *__restrict__ s,
                                                         R *__restrict__ pred, R *__restrict__ ratio,
                                                         int64_t count, R delta) {
  const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= count) return;
  R p = hf[idx];
  if (s) p = p + s[idx];
  if (pred) pred[idx] = p;
  ratio[idx] = g[idx] / (p + delta);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void rdp_kernel(const R *__restrict__ f, R *__restrict__ out,
                                                       int nx, int ny, int nz, R dx, R dy, R dz,
                                                       R beta_scale, R gamma, R delta) {
  const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  const int64_t count = (int64_t)nx * ny * nz;
  if (idx >= count) return;
  const int z = idx % nz;
  const int y = (idx / nz) % ny;
  const int x = idx / ((int64_t)nz * ny);
  const R fr = f[idx];
  R acc = R(0);
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        // EuclideanNeighbourWeight (nearest_neighbour.py:219-220), in float64 like numpy
        const double w = (double)dx / sqrt((double)dx * i * (double)dx * i +
                                           (double)dy * j * (double)dy * j +
                                           (double)dz * k * (double)dz * k);
        const int xs = min(max(x + i, 0), nx - 1);
        const int ys = min(max(y + j, 0), ny - 1);
        const int zs = min(max(z + k, 0), nz - 1);
        const R fs = f[((int64_t)xs * ny + ys) * nz + zs];
        const R d = fr - fs;
        const R ad = fabs(d);
        const R num = d * (gamma * ad + R(3) * fs + fr + R(2) * delta);
        const R den = fr + fs + gamma * ad + delta;
        acc = acc + (num / (den * den)) * R(w);
      }
  out[idx] = beta_scale * acc;
}

template <typename R>
__global__ __launch_bounds__(kThreads) void bsrem_update_kernel(
    R *__restrict__ f, const R *__restrict__ b, const R *__restrict__ nrm,
    const R *__restrict__ nsum, const R *__restrict__ prior, int64_t count, R relax, R weight,
    R delta, R floor_value) {
  const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= count) return;
  const R fo = f[idx];
  const R d = relax / (weight * nsum[idx] + delta);
  const R pre = fo * d;
  const R grad = b[idx] - nrm[idx];
  const R pg = prior ? prior[idx] : R(0);
  R v = fo + pre * (grad - pg);
  if (v <= floor_value) v = floor_value;
  f[idx] = v;
}

}  // namespace

// ===========================================================================
// Plan and host side
// ===========================================================================
struct spect_plan {
  int N, Nz, A, P, pad;
  double dx, dy, dz;
  int kr, kz, n_radii;
  bool has64;
  int device;
  float2 *trig32 = nullptr;
  double2 *trig64 = nullptr;
  float *lin32 = nullptr;
  double *lin64 = nullptr;
  int32_t *radius_index = nullptr;
  float *taps_r32 = nullptr, *taps_z32 = nullptr;
  double *taps_r64 = nullptr, *taps_z64 = nullptr;
};

namespace {

int host_pad_size(int width) {
  return (int)std::ceil((std::sqrt(2.0) * width - width) / 2.0);
}

// torch.linspace(-1, 1, P) as computed by ATen's CPU range kernel: a float
// step, ascending from -1 for the first half, descending from +1 after.
template <typename R> std::vector<R> host_linspace(int P) {
  std::vector<R> v(P);
  if (P == 1) {
    v[0] = R(-1);
    return v;
  }
  const R step = (R(1) - R(-1)) / R(P - 1);
  const int halfway = P / 2;
  for (int k = 0; k < P; ++k)
    v[k] = k < halfway ? R(-1) + step * R(k) : R(1) - step * R(P - k - 1);
  return v;
}

template <typename T> int upload(T **dst, const T *src, size_t count) {
  HIP_TRY(hipMalloc((void **)dst, count * sizeof(T)));
  HIP_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return SPECT_OK;
}

struct Launch {
  int TJ_f, TJ_b, L, n_chunks;
  int jo0, jo1, je0, je1;
  int NA, NC_f, NC_b;
};

// Tile / chunk choice for n angles.
Launch choose_launch(const spect_plan *p, int n, size_t real_size) {
  Launch L{};
  const int hy = p->kr > 0 ? p->kr / 2 : 0;
  L.jo0 = p->pad;
  L.jo1 = p->pad + p->N;
  L.je0 = std::max(0, L.jo0 - hy);
  L.je1 = std::min(p->P, L.jo1 + hy);
  L.TJ_f = real_size == 4 ? 16 : 8;
  L.TJ_b = real_size == 4 ? 16 : 8;
  const int jt_f = (p->N + L.TJ_f - 1) / L.TJ_f;
  // aim for >= 1024 blocks, chunks of >= 8 planes
  int chunks = std::max(1, (1024 + n * jt_f - 1) / (n * jt_f));
  chunks = std::min(chunks, std::max(1, p->P / 8));
  L.L = (p->P + chunks - 1) / chunks;
  L.n_chunks = (p->P + L.L - 1) / L.L;
  const int nA = (L.TJ_f + 2 * hy) * p->Nz;
  L.NA = (nA + kThreads - 1) / kThreads;
  L.NC_f = (L.TJ_f * p->Nz + kThreads - 1) / kThreads;
  L.NC_b = (L.TJ_b * p->Nz + kThreads - 1) / kThreads;
  return L;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct WsLayout {
  size_t sums_off, sums_bytes;
  size_t part_off, part_bytes;
  size_t w_off, w_bytes;
  size_t total;
};

WsLayout ws_layout(const spect_plan *p, int dtype, int n) {
  const size_t rs = dtype == SPECT_F64 ? 8 : 4;
  Launch L = choose_launch(p, n, rs);
  WsLayout w{};
  const size_t n_jext = L.je1 - L.je0;
  w.sums_bytes = align256((size_t)n * L.n_chunks * n_jext * p->Nz * sizeof(double));
  w.part_bytes = align256((size_t)L.n_chunks * n * p->N * p->Nz * rs);
  w.w_bytes = align256((size_t)n * p->P * n_jext * p->Nz * rs);
  w.sums_off = 0;
  w.part_off = w.sums_off + w.sums_bytes;
  w.w_off = w.sums_off + w.sums_bytes;  // forward partials and W never coexist
  w.total = w.sums_bytes + std::max(w.part_bytes, w.w_bytes);
  return w;
}

template <typename R> Tables<R> tables(const spect_plan *p);
template <> Tables<float> tables<float>(const spect_plan *p) {
  return {p->trig32, p->lin32, p->radius_index, p->taps_r32, p->taps_z32};
}
template <> Tables<double> tables<double>(const spect_plan *p) {
  return {p->trig64, p->lin64, p->radius_index, p->taps_r64, p->taps_z64};
}

template <typename R> Shape<R> shape(const spect_plan *p) {
  Shape<R> s;
  s.N = p->N;
  s.Nz = p->Nz;
  s.P = p->P;
  s.pad = p->pad;
  s.A = p->A;
  s.kr = p->kr;
  s.kz = p->kr > 0 ? p->kz : 0;
  s.hy = p->kr > 0 ? p->kr / 2 : 0;
  s.hz = s.kz > 0 ? s.kz / 2 : 0;
  s.half = R(p->P - 1) / R(2);
  s.dx = R(p->dx);
  return s;
}

// Dispatch on a runtime item count to a compiled register-array size.
// SWITCH_A: per-thread items of a halo tile; SWITCH_C: items of an output tile.
#define SPECT_SWITCH_A(n, NAME, ...)                                                  \
  [&]() -> int {                                                                      \
    if ((n) <= 8) { constexpr int NAME = 8; __VA_ARGS__; return SPECT_OK; }           \
    if ((n) <= 16) { constexpr int NAME = 16; __VA_ARGS__; return SPECT_OK; }         \
    if ((n) <= 24) { constexpr int NAME = 24; __VA_ARGS__; return SPECT_OK; }         \
    if ((n) <= 32) { constexpr int NAME = 32; __VA_ARGS__; return SPECT_OK; }         \
    if ((n) <= 48) { constexpr int NAME = 48; __VA_ARGS__; return SPECT_OK; }         \
    return fail(SPECT_E_ARG, "tile too large for the register budget");              \
  }()
#define SPECT_SWITCH_C(n, NAME, ...)                                                  \
  [&]() -> int {                                                                      \
    if ((n) <= 4) { constexpr int NAME = 4; __VA_ARGS__; return SPECT_OK; }           \
    if ((n) <= 8) { constexpr int NAME = 8; __VA_ARGS__; return SPECT_OK; }           \
    if ((n) <= 16) { constexpr int NAME = 16; __VA_ARGS__; return SPECT_OK; }         \
    return fail(SPECT_E_ARG, "tile too large for the register budget");              \
  }()

template <typename R>
int att_sums(const spect_plan *p, const Launch &L, const R *mu, const int32_t *idx, int n,
             double *sums, hipStream_t st) {
  const Shape<R> sh = shape<R>(p);
  const Tables<R> tb = tables<R>(p);
  const int TJ = 16;
  const int n_jext = L.je1 - L.je0;
  dim3 grid(n, (n_jext + TJ - 1) / TJ, L.n_chunks);
  const int ni = (TJ * p->Nz + kThreads - 1) / kThreads;
  int rc = SPECT_SWITCH_C(ni, NI, {
    att_sums_kernel<R, NI><<<grid, kThreads, 0, st>>>(sh, tb, mu, idx, L.je0, L.je1, TJ, L.L,
                                                      L.n_chunks, sums);
  });
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return SPECT_OK;
}

template <typename R>
int forward_impl(const spect_plan *p, const R *f, const R *mu, const int32_t *idx, int n, R *proj,
                 unsigned char *ws, hipStream_t st) {
  const Launch L = choose_launch(p, n, sizeof(R));
  const WsLayout w = ws_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, n);
  double *sums = reinterpret_cast<double *>(ws + w.sums_off);
  R *part = reinterpret_cast<R *>(ws + w.part_off);
  if (mu) {
    int rc = att_sums<R>(p, L, mu, idx, n, sums, st);
    if (rc) return rc;
  }
  const Shape<R> sh = shape<R>(p);
  const Tables<R> tb = tables<R>(p);
  const int jt = (p->N + L.TJ_f - 1) / L.TJ_f;
  dim3 grid(n, jt, L.n_chunks);
  const size_t smem = (size_t)(2 * L.TJ_f + 2 * sh.hy) * p->Nz * sizeof(R) +
                      (size_t)(L.TJ_f + 2 * sh.hy) * sizeof(Src<R>);
  if (smem > 160 * 1024) return fail(SPECT_E_ARG, "forward tile exceeds LDS");
  int rc = SPECT_SWITCH_A(L.NA, NA, {
    return SPECT_SWITCH_C(L.NC_f, NC, {
      auto kern = fwd_kernel<R, NA, NC>;
      if (smem > 64 * 1024)
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)smem);
      kern<<<grid, kThreads, smem, st>>>(sh, tb, f, mu, idx, L.jo0, L.jo1, L.TJ_f, L.L,
                                         L.n_chunks, L.je0, L.je1, sums, part, n);
    });
  });
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  const size_t count = (size_t)n * p->N * p->Nz;
  fwd_reduce_kernel<R><<<(unsigned)((count + kThreads - 1) / kThreads), kThreads, 0, st>>>(
      part, proj, count, L.n_chunks);
  HIP_TRY(hipGetLastError());
  return SPECT_OK;
}

template <typename R>
int backward_impl(const spect_plan *p, int mode, int out_mode, const R *proj, const R *mu,
                  const int32_t *idx, int n, R *out, const R *f, const R *nrm, double delta,
                  double floor_value, unsigned char *ws, hipStream_t st) {
  const Launch L = choose_launch(p, n, sizeof(R));
  const WsLayout w = ws_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, n);
  double *sums = reinterpret_cast<double *>(ws + w.sums_off);
  R *W = reinterpret_cast<R *>(ws + w.w_off);
  if (mu) {
    int rc = att_sums<R>(p, L, mu, idx, n, sums, st);
    if (rc) return rc;
  }
  const Shape<R> sh = shape<R>(p);
  const Tables<R> tb = tables<R>(p);
  const int n_jext = L.je1 - L.je0;
  dim3 grid(n, (n_jext + L.TJ_b - 1) / L.TJ_b, L.n_chunks);
  const size_t smem = (size_t)(2 * L.TJ_b + 2 * sh.hy) * p->Nz * sizeof(R) +
                      (size_t)L.TJ_b * sizeof(Src<R>);
  if (smem > 160 * 1024) return fail(SPECT_E_ARG, "backward tile exceeds LDS");
  int rc = SPECT_SWITCH_C(L.NC_b, NC, {
    auto kern = bww_kernel<R, NC>;
    if (smem > 64 * 1024)
      (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)smem);
    kern<<<grid, kThreads, smem, st>>>(sh, tb, proj, mu, idx, L.jo0, L.jo1, L.je0, L.je1, L.TJ_b,
                                       L.L, L.n_chunks, sums, W);
  });
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  const int zblocks = (p->Nz + 63) / 64;
  const size_t waves = (size_t)p->N * p->N * zblocks;
  const unsigned blocks = (unsigned)((waves * 64 + kThreads - 1) / kThreads);
  if (mode == SPECT_BP_REFERENCE)
    bw_ref_kernel<R><<<blocks, kThreads, 0, st>>>(sh, tb, W, idx, n, L.je0, L.je1, out_mode, out,
                                                 f, nrm, R(delta), R(floor_value));
  else
    bw_adj_kernel<R><<<blocks, kThreads, 0, st>>>(sh, tb, W, idx, n, L.je0, L.je1, out_mode, out,
                                                 f, nrm, R(delta), R(floor_value));
  HIP_TRY(hipGetLastError());
  return SPECT_OK;
}

int check_plan_call(const spect_plan *p, int dtype, int n, const void *ws, size_t ws_bytes) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  if (dtype == SPECT_F64 && !p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
  if (n <= 0 || n > p->A) return fail(SPECT_E_ARG, "angle count out of range");
  const size_t need = ws_layout(p, dtype, n).total;
  if (!ws || ws_bytes < need)
    return fail(SPECT_E_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  return SPECT_OK;
}

}  // namespace

extern "C" {

const char *spect_last_error(void) { return g_err.c_str(); }

int spect_version(void) { return SPECT_VERSION; }

int spect_plan_create(const spect_geom *g, spect_plan **out) {
  if (!g || !out) return fail(SPECT_E_ARG, "null argument");
  *out = nullptr;
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return fail(SPECT_E_ARG, "object shape must be positive");
  if (g->nx != g->ny) return fail(SPECT_E_ARG, "nx must equal ny (square transaxial grid)");
  if (g->n_angles <= 0 || !g->angles_deg) return fail(SPECT_E_ARG, "angles required");
  if (!(g->dx > 0) || !(g->dz > 0)) return fail(SPECT_E_ARG, "voxel sizes must be positive");
  if (g->psf_kr < 0 || g->psf_kz < 0 || (g->psf_kr > 0 && g->psf_kr % 2 == 0) ||
      (g->psf_kz > 0 && g->psf_kz % 2 == 0))
    return fail(SPECT_E_ARG, "PSF tap counts must be odd");
  if (g->psf_kr > 0 && (g->n_radii <= 0 || !g->radius_index || !g->taps_r ||
                        (g->psf_kz > 0 && !g->taps_z)))
    return fail(SPECT_E_ARG, "PSF tables missing");
  auto *p = new spect_plan();
  p->N = g->nx;
  p->Nz = g->nz;
  p->A = g->n_angles;
  p->pad = host_pad_size(g->ny);
  p->P = g->nx + 2 * p->pad;
  p->dx = g->dx;
  p->dy = g->dy;
  p->dz = g->dz;
  p->kr = g->psf_kr;
  p->kz = g->psf_kr > 0 ? g->psf_kz : 0;
  p->n_radii = g->psf_kr > 0 ? g->n_radii : 1;
  p->has64 = g->angles_deg64 != nullptr && (g->psf_kr == 0 || g->taps_r64 != nullptr);
  const int hy = p->kr / 2;
  if ((16 + 2 * hy) * p->Nz > 48 * kThreads) {
    delete p;
    return fail(SPECT_E_ARG, "volume too deep along z for the forward tile");
  }
  (void)hipGetDevice(&p->device);
  auto cleanup = [&](int rc) {
    spect_plan_destroy(p);
    return rc;
  };
  // theta_f = -(270 - beta) in float32; radians as kornia's deg2rad in float32
  std::vector<float2> t32(p->A);
  std::vector<double2> t64(p->A);
  const float pi32 = (float)M_PI;
  for (int a = 0; a < p->A; ++a) {
    const float th = -(270.0f - g->angles_deg[a]);
    const float rad = (th * pi32) / 180.0f;
    t32[a] = make_float2((float)std::cos((double)rad), (float)std::sin((double)rad));
    if (g->angles_deg64) {
      const double th64 = -(270.0 - g->angles_deg64[a]);
      const double r64 = th64 * M_PI / 180.0;
      t64[a] = make_double2(std::cos(r64), std::sin(r64));
    }
  }
  std::vector<int32_t> ridx(p->A, 0);
  if (p->kr > 0) {
    for (int a = 0; a < p->A; ++a) {
      if (g->radius_index[a] < 0 || g->radius_index[a] >= p->n_radii) {
        delete p;
        return fail(SPECT_E_ARG, "radius_index out of range");
      }
      ridx[a] = g->radius_index[a];
    }
  }
  int rc;
  if ((rc = upload(&p->trig32, t32.data(), t32.size()))) return cleanup(rc);
  if ((rc = upload(&p->trig64, t64.data(), t64.size()))) return cleanup(rc);
  auto l32 = host_linspace<float>(p->P);
  auto l64 = host_linspace<double>(p->P);
  if ((rc = upload(&p->lin32, l32.data(), l32.size()))) return cleanup(rc);
  if ((rc = upload(&p->lin64, l64.data(), l64.size()))) return cleanup(rc);
  if ((rc = upload(&p->radius_index, ridx.data(), ridx.size()))) return cleanup(rc);
  if (p->kr > 0) {
    const size_t nr = (size_t)p->n_radii * p->P;
    if ((rc = upload(&p->taps_r32, g->taps_r, nr * p->kr))) return cleanup(rc);
    if (p->kz > 0 && (rc = upload(&p->taps_z32, g->taps_z, nr * p->kz))) return cleanup(rc);
    if (p->has64) {
      if ((rc = upload(&p->taps_r64, g->taps_r64, nr * p->kr))) return cleanup(rc);
      if (p->kz > 0 && (rc = upload(&p->taps_z64, g->taps_z64, nr * p->kz))) return cleanup(rc);
    }
  }
  *out = p;
  return SPECT_OK;
}

int spect_plan_destroy(spect_plan *p) {
  if (!p) return SPECT_OK;
  (void)hipFree(p->trig32);
  (void)hipFree(p->trig64);
  (void)hipFree(p->lin32);
  (void)hipFree(p->lin64);
  (void)hipFree(p->radius_index);
  (void)hipFree(p->taps_r32);
  (void)hipFree(p->taps_z32);
  (void)hipFree(p->taps_r64);
  (void)hipFree(p->taps_z64);
  delete p;
  return SPECT_OK;
}

int spect_plan_dims(const spect_plan *p, int32_t *padded, int32_t *pad) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (padded) *padded = p->P;
  if (pad) *pad = p->pad;
  return SPECT_OK;
}

size_t spect_workspace_bytes(const spect_plan *p, int dtype, int n) {
  if (!p || n <= 0) return 0;
  return ws_layout(p, dtype, n).total;
}

int spect_forward(const spect_plan *p, int dtype, const void *f, const void *mu,
                  const int32_t *angle_idx, int n, void *proj, void *workspace,
                  size_t workspace_bytes, void *stream) {
  int rc = check_plan_call(p, dtype, n, workspace, workspace_bytes);
  if (rc) return rc;
  if (!f || !proj || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  if (dtype == SPECT_F32)
    return forward_impl<float>(p, (const float *)f, (const float *)mu, angle_idx, n, (float *)proj,
                               ws, st);
  return forward_impl<double>(p, (const double *)f, (const double *)mu, angle_idx, n,
                              (double *)proj, ws, st);
}

int spect_backward(const spect_plan *p, int dtype, int mode, int out_mode, const void *proj,
                   const void *mu, const int32_t *angle_idx, int n, void *out, const void *f,
                   const void *norm, double delta, double floor_value, void *workspace,
                   size_t workspace_bytes, void *stream) {
  int rc = check_plan_call(p, dtype, n, workspace, workspace_bytes);
  if (rc) return rc;
  if (!proj || !out || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  if (mode != SPECT_BP_REFERENCE && mode != SPECT_BP_ADJOINT) return fail(SPECT_E_ARG, "bad mode");
  if (out_mode < SPECT_OUT_STORE || out_mode > SPECT_OUT_OSEM) return fail(SPECT_E_ARG, "bad out_mode");
  if (out_mode == SPECT_OUT_OSEM && (!f || !norm)) return fail(SPECT_E_ARG, "OSEM update needs f and norm");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  if (dtype == SPECT_F32)
    return backward_impl<float>(p, mode, out_mode, (const float *)proj, (const float *)mu,
                                angle_idx, n, (float *)out, (const float *)f, (const float *)norm,
                                delta, floor_value, ws, st);
  return backward_impl<double>(p, mode, out_mode, (const double *)proj, (const double *)mu,
                               angle_idx, n, (double *)out, (const double *)f,
                               (const double *)norm, delta, floor_value, ws, st);
}

int spect_poisson_ratio(int dtype, const void *g, const void *hf, const void *s, void *pred,
                        void *ratio, int64_t count, double delta, void *stream) {
  if (!g || !hf || !ratio || count < 0) return fail(SPECT_E_ARG, "bad ratio arguments");
  if (count == 0) return SPECT_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned blocks = (unsigned)((count + kThreads - 1) / kThreads);
  if (dtype == SPECT_F32)
    ratio_kernel<float><<<blocks, kThreads, 0, st>>>((const float *)g, (const float *)hf,
                                                     (const float *)s, (float *)pred,
                                                     (float *)ratio, count, (float)delta);
  else if (dtype == SPECT_F64)
    ratio_kernel<double><<<blocks, kThreads, 0, st>>>((const double *)g, (const double *)hf,
                                                      (const double *)s, (double *)pred,
                                                      (double *)ratio, count, delta);
  else
    return fail(SPECT_E_ARG, "bad dtype");
  HIP_TRY(hipGetLastError());
  return SPECT_OK;
}

int spect_rdp_gradient(int dtype, const void *f, void *out, int32_t nx, int32_t ny, int32_t nz,
                       double dx, double dy, double dz, double beta, double gamma, double delta,
                       double scale, void *stream) {
  if (!f || !out || nx <= 0 || ny <= 0 || nz <= 0) return fail(SPECT_E_ARG, "bad prior arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int64_t count = (int64_t)nx * ny * nz;
  const unsigned blocks = (unsigned)((count + kThreads - 1) / kThreads);
  const double bs = beta * scale;
  if (dtype == SPECT_F32)
    rdp_kernel<float><<<blocks, kThreads, 0, st>>>((const float *)f, (float *)out, nx, ny, nz,
                                                   (float)dx, (float)dy, (float)dz, (float)bs,
                                                   (float)gamma, (float)delta);
  else if (dtype == SPECT_F64)
    rdp_kernel<double><<<blocks, kThreads, 0, st>>>((const double *)f, (double *)out, nx, ny, nz,
                                                    dx, dy, dz, bs, gamma, delta);
  else
    return fail(SPECT_E_ARG, "bad dtype");
  HIP_TRY(hipGetLastError());
  return SPECT_OK;
}

int spect_bsrem_update(int dtype, void *f, const void *b, const void *norm, const void *norm_sum,
                       const void *prior, int64_t count, double relax, double weight, double delta,
                       double floor_value, void *stream) {
  if (!f || !b || !norm || !norm_sum || count < 0) return fail(SPECT_E_ARG, "bad update arguments");
  if (count == 0) return SPECT_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned blocks = (unsigned)((count + kThreads - 1) / kThreads);
  if (dtype == SPECT_F32)
    bsrem_update_kernel<float><<<blocks, kThreads, 0, st>>>(
        (float *)f, (const float *)b, (const float *)norm, (const float *)norm_sum,
        (const float *)prior, count, (float)relax, (float)weight, (float)delta,
        (float)floor_value);
  else if (dtype == SPECT_F64)
    bsrem_update_kernel<double><<<blocks, kThreads, 0, st>>>(
        (double *)f, (const double *)b, (const double *)norm, (const double *)norm_sum,
        (const double *)prior, count, relax, weight, delta, floor_value);
  else
    return fail(SPECT_E_ARG, "bad dtype");
  HIP_TRY(hipGetLastError());
  return SPECT_OK;
}

}  // extern "C"
