// shared.hip -- object-space transforms that wrap the projector:
// the kernel-EM basis K (transforms/shared/kem.py:113-146) and the
// deformation-vector-field warp (transforms/shared/motion.py:54-110).
#include "../../include/spect_hip.h"
#include "launch.hpp"

namespace spect {
namespace {

// out(v) = sum_{i,j,k} K[i][j][k](v) * f(clamp(v + (i, j, k))), i, j, k in
// [-m, m] ascending, accumulated in that order from 0 (KEMTransform.forward,
// kem.py:113-131, with get_object_nearest_neighbour's replicate clamp,
// utils/misc.py:40-57).  K is [S^3][nx*ny*nz]: each tap plane is read once,
// coalesced along v; f is re-read through L2.
// MT > 0: the half-width m = MT is a compile-time constant and the tap loop is
// unrolled, so the size^3 independent tap-plane loads are all in flight at
// once; MT = 0 is the generic loop.
template <typename R, int MT>
__global__ __launch_bounds__(256) void kem_kernel(const R *__restrict__ K, const R *__restrict__ f,
                                                  R *__restrict__ out, int nx, int ny, int nz,
                                                  int m_rt) {
#pragma clang fp contract(off)
  const int m = MT > 0 ? MT : m_rt;
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  const int z = v % nz;
  const int y = (v / nz) % ny;
  const int x = v / ((int64_t)nz * ny);
  R acc = R(0);
  int t = 0;
#pragma unroll
  for (int i = -m; i <= m; ++i) {
    const int xs = min(max(x + i, 0), nx - 1);
#pragma unroll
    for (int j = -m; j <= m; ++j) {
      const int ys = min(max(y + j, 0), ny - 1);
      const R *row = f + ((int64_t)xs * ny + ys) * nz;
#pragma unroll
      for (int k = -m; k <= m; ++k, ++t) {
        const int zs = min(max(z + k, 0), nz - 1);
        acc = acc + K[(int64_t)t * n + v] * row[zs];
      }
    }
  }
  out[v] = acc;
}

// DVFMotionTransform._apply_dvf (motion.py:54-78): sample f at
// v + dvf(v) with torch grid_sample(trilinear, zeros, align_corners=True).
// The normalised coordinate is 2/(L-1) * p - 1 in R (motion.py:50-52) and
// grid_sample maps it back with ((c + 1) / 2) * (L - 1); corners outside the
// volume contribute 0.  The dvf is [nx][ny][nz][3] (x, y, z components).
template <typename R>
__global__ __launch_bounds__(256) void dvf_kernel(const R *__restrict__ f,
                                                  const R *__restrict__ dvf, R *__restrict__ out,
                                                  int nx, int ny, int nz) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= n) return;
  const int z = v % nz;
  const int y = (v / nz) % ny;
  const int x = v / ((int64_t)nz * ny);
  const R sx = R(2) / R(nx - 1), sy = R(2) / R(ny - 1), sz = R(2) / R(nz - 1);
  const R px = R(x) + dvf[3 * v + 0];
  const R py = R(y) + dvf[3 * v + 1];
  const R pz = R(z) + dvf[3 * v + 2];
  const R gx = sx * px - R(1), gy = sy * py - R(1), gz = sz * pz - R(1);
  // grid_sample's (W, H, D) = (z, y, x) after the reference's flip
  const R iw = ((gz + R(1)) / R(2)) * R(nz - 1);
  const R ih = ((gy + R(1)) / R(2)) * R(ny - 1);
  const R id = ((gx + R(1)) / R(2)) * R(nx - 1);
  const R fw = floor(iw), fh = floor(ih), fd = floor(id);
  const int w0 = (int)fw, h0 = (int)fh, d0 = (int)fd;
  const R w1f = fw + R(1), h1f = fh + R(1), d1f = fd + R(1);
  // corner weights in grid_sample's order (t/b = depth, n/s = height, w/e = width)
  const R tnw = (w1f - iw) * (h1f - ih) * (d1f - id);
  const R tne = (iw - fw) * (h1f - ih) * (d1f - id);
  const R tsw = (w1f - iw) * (ih - fh) * (d1f - id);
  const R tse = (iw - fw) * (ih - fh) * (d1f - id);
  const R bnw = (w1f - iw) * (h1f - ih) * (id - fd);
  const R bne = (iw - fw) * (h1f - ih) * (id - fd);
  const R bsw = (w1f - iw) * (ih - fh) * (id - fd);
  const R bse = (iw - fw) * (ih - fh) * (id - fd);
  auto at = [&](int d, int h, int w) -> R {
    if ((unsigned)d >= (unsigned)nx || (unsigned)h >= (unsigned)ny || (unsigned)w >= (unsigned)nz)
      return R(0);
    return f[((int64_t)d * ny + h) * nz + w];
  };
  auto inb = [&](int d, int h, int w) {
    return (unsigned)d < (unsigned)nx && (unsigned)h < (unsigned)ny && (unsigned)w < (unsigned)nz;
  };
  R acc = R(0);
  if (inb(d0, h0, w0)) acc = acc + at(d0, h0, w0) * tnw;
  if (inb(d0, h0, w0 + 1)) acc = acc + at(d0, h0, w0 + 1) * tne;
  if (inb(d0, h0 + 1, w0)) acc = acc + at(d0, h0 + 1, w0) * tsw;
  if (inb(d0, h0 + 1, w0 + 1)) acc = acc + at(d0, h0 + 1, w0 + 1) * tse;
  if (inb(d0 + 1, h0, w0)) acc = acc + at(d0 + 1, h0, w0) * bnw;
  if (inb(d0 + 1, h0, w0 + 1)) acc = acc + at(d0 + 1, h0, w0 + 1) * bne;
  if (inb(d0 + 1, h0 + 1, w0)) acc = acc + at(d0 + 1, h0 + 1, w0) * bsw;
  if (inb(d0 + 1, h0 + 1, w0 + 1)) acc = acc + at(d0 + 1, h0 + 1, w0 + 1) * bse;
  out[v] = acc;
}

}  // namespace

template <typename R>
hipError_t launch_kem(const R *K, const R *f, R *out, int nx, int ny, int nz, int m,
                      hipStream_t st) {
  const int64_t n = (int64_t)nx * ny * nz;
  const dim3 grid((unsigned)((n + 255) / 256));
  switch (m) {
    case 1: kem_kernel<R, 1><<<grid, 256, 0, st>>>(K, f, out, nx, ny, nz, m); break;
    case 2: kem_kernel<R, 2><<<grid, 256, 0, st>>>(K, f, out, nx, ny, nz, m); break;
    default: kem_kernel<R, 0><<<grid, 256, 0, st>>>(K, f, out, nx, ny, nz, m); break;
  }
  return hipGetLastError();
}

template <typename R>
hipError_t launch_dvf(const R *f, const R *dvf, R *out, int nx, int ny, int nz, hipStream_t st) {
  const int64_t n = (int64_t)nx * ny * nz;
  dvf_kernel<R><<<dim3((unsigned)((n + 255) / 256)), 256, 0, st>>>(f, dvf, out, nx, ny, nz);
  return hipGetLastError();
}

#define SPECT_INST(R)                                                                          \
  template hipError_t launch_kem<R>(const R *, const R *, R *, int, int, int, int, hipStream_t); \
  template hipError_t launch_dvf<R>(const R *, const R *, R *, int, int, int, hipStream_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
