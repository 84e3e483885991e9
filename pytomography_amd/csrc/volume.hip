// volume.hip -- whole rotated-frame volumes, one per angle, for the per-angle
// path of SPECTSystemMatrix (spect_system_matrix.py:147-156, 187-198), which
// runs when a caller's transforms are not the fused pair (attenuation, then
// PSF): the padded object rotated to an angle, and the PSF blur of such a
// volume.  The back rotation and the attenuation factors reuse the gather and
// attenuation-field bodies over the full frame width (capi.hip).
#include "../../include/spect_hip.h"
#include "launch.hpp"

namespace spect {
namespace {

// out[al][i][j][z]: the bilinear sample of the object (zero outside it) at
// rotated-frame voxel (i, j) of angle angle_idx[al] -- the sample the fused
// forward takes there (rot_coord, make_src, taps4, interp).  Grid (z-groups
// of the plane, P planes, n angles).
template <typename R>
__global__ __launch_bounds__(256) void rotate_kernel(Geo<R> g, Tables<R> tb, const R *__restrict__ f,
                                                     const int32_t *__restrict__ angle_idx,
                                                     R *__restrict__ out) {
  const int al = blockIdx.z, i = blockIdx.y;
  const int e = blockIdx.x * 256 + threadIdx.x;  // j * Nz + z within plane i
  if (e >= g.P * g.Nz) return;
  const int j = e / g.Nz, z = e - j * g.Nz;
  R *dst = out + ((size_t)al * g.P + i) * (size_t)g.P * g.Nz + e;
  const int ang = ld_uniform(angle_idx + al);
  if ((unsigned)ang >= (unsigned)g.A) {
    *dst = R(0);
    return;
  }
  const auto cs = ld_uniform(tb.trig + ang);
  R ix, iy;
  rot_coord<R>(cs.x, cs.y, tb.lin[j], ld_uniform(tb.lin + i), g.half, ix, iy);
  const Taps4<R> t = taps4<R>(make_src<R>(ix, iy, g.pad), g.N, g.Nz);
  *dst = interp<R>(f, t, z);
}

constexpr int kPsfTJ = 16;  // columns per block of the volume blur

// PSF blur of one plane's kPsfTJ columns: along j with the plane's taps_r,
// then along z with its taps_z, as Seperable1DBlurNet applies layer_r then
// layer_z (psf.py:27-40), each a 'same' cross-correlation with zeros outside
// the frame.  LDS: the columns with H halo columns each side (all z), then the
// j-blurred columns with H zeros each side of z.
template <typename R>
__global__ __launch_bounds__(256) void psf_volume_kernel(Geo<R> g, Tables<R> tb, const R *__restrict__ vol,
                                                         const int32_t *__restrict__ angle_idx,
                                                         R *__restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int al = blockIdx.z, i = blockIdx.y, j0 = blockIdx.x * kPsfTJ;
  const int H = g.H, K = 2 * H + 1, Nz = g.Nz, P = g.P;
  const int tj = min(kPsfTJ, P - j0);
  const size_t base = ((size_t)al * P + i) * (size_t)P * Nz;
  const int ang = ld_uniform(angle_idx + al);
  if ((unsigned)ang >= (unsigned)g.A) {
    for (int q = threadIdx.x; q < tj * Nz; q += blockDim.x) out[base + (size_t)j0 * Nz + q] = R(0);
    return;
  }
  const int rix = ld_uniform(tb.radius_index + ang);
  const R *wr = tb.taps_r + ((size_t)rix * P + i) * K;
  const R *wz = tb.taps_z + ((size_t)rix * P + i) * K;
  const int ldu = Nz + 2 * H;
  R *V = reinterpret_cast<R *>(smem);      // [kPsfTJ + 2H][Nz]
  R *U = V + (size_t)(kPsfTJ + 2 * H) * Nz;  // [kPsfTJ][Nz + 2H]
  for (int q = threadIdx.x; q < (tj + 2 * H) * Nz; q += blockDim.x) {
    const int r = q / Nz, z = q - r * Nz;
    const int j = j0 - H + r;
    V[q] = (j >= 0 && j < P) ? vol[base + (size_t)j * Nz + z] : R(0);
  }
  for (int q = threadIdx.x; q < tj * ldu; q += blockDim.x) U[q] = R(0);
  __syncthreads();
  for (int q = threadIdx.x; q < tj * Nz; q += blockDim.x) {
    const int c = q / Nz, z = q - c * Nz;
    R acc = R(0);
    for (int t = 0; t < K; ++t) acc += wr[t] * V[(c + t) * Nz + z];
    U[c * ldu + H + z] = acc;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < tj * Nz; q += blockDim.x) {
    const int c = q / Nz, z = q - c * Nz;
    R acc = R(0);
    for (int t = 0; t < K; ++t) acc += wz[t] * U[c * ldu + z + t];
    out[base + (size_t)(j0 + c) * Nz + z] = acc;
  }
}

}  // namespace

template <typename R>
hipError_t launch_rotate(const Geo<R> &g, const Tables<R> &tb, const R *f, const int32_t *angle_idx,
                         int n, R *out, hipStream_t st) {
  if (n > 65535 || g.P > 65535) return hipErrorInvalidConfiguration;
  dim3 grid((g.P * g.Nz + 255) / 256, g.P, n);
  rotate_kernel<R><<<grid, 256, 0, st>>>(g, tb, f, angle_idx, out);
  return hipGetLastError();
}

template <typename R>
size_t psf_volume_smem_bytes(const Geo<R> &g) {
  return ((size_t)(kPsfTJ + 2 * g.H) * g.Nz + (size_t)kPsfTJ * (g.Nz + 2 * g.H)) * sizeof(R);
}

template <typename R>
hipError_t launch_psf_volume(const Geo<R> &g, const Tables<R> &tb, const R *vol,
                             const int32_t *angle_idx, int n, R *out, hipStream_t st) {
  if (n > 65535 || g.P > 65535 || g.H <= 0) return hipErrorInvalidConfiguration;
  const size_t smem = psf_volume_smem_bytes(g);
  if (smem > 160 * 1024) return hipErrorInvalidConfiguration;
  if (smem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)psf_volume_kernel<R>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  dim3 grid((g.P + kPsfTJ - 1) / kPsfTJ, g.P, n);
  psf_volume_kernel<R><<<grid, 256, smem, st>>>(g, tb, vol, angle_idx, out);
  return hipGetLastError();
}

#define SPECT_INST(R)                                                                            \
  template hipError_t launch_rotate<R>(const Geo<R> &, const Tables<R> &, const R *,           \
                                       const int32_t *, int, R *, hipStream_t);                \
  template size_t psf_volume_smem_bytes<R>(const Geo<R> &);                                     \
  template hipError_t launch_psf_volume<R>(const Geo<R> &, const Tables<R> &, const R *,       \
                                           const int32_t *, int, R *, hipStream_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
