// misc.hip -- elementwise / stencil kernels of the likelihood, prior and updates.
#include "../../include/spect_hip.h"
#include "launch.hpp"

#include <cmath>

namespace spect {
namespace {

// EuclideanNeighbourWeight (priors/nearest_neighbour.py:207-220) of the 26
// neighbours, dx / |(i dx, j dy, k dz)| evaluated in double on the host and
// rounded to R once; index (i+1)*9 + (j+1)*3 + (k+1), centre unused.  Passed
// by value, so the unrolled stencils read them from the kernel arguments.
template <typename R> struct RdpWeights {
  R w[27];
};

template <typename R>
RdpWeights<R> rdp_weights(double dx, double dy, double dz) {
  RdpWeights<R> W{};
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        const double w = dx / std::sqrt((dx * i) * (dx * i) + (dy * j) * (dy * j) + (dz * k) * (dz * k));
        W.w[(i + 1) * 9 + (j + 1) * 3 + (k + 1)] = R(w);
      }
  return W;
}

// likelihoods/poisson_log_likelihood.py:34-36
template <typename R>
__global__ __launch_bounds__(256) void ratio_kernel(const R *__restrict__ g,
                                                    const R *__restrict__ hf,
                                                    const R *__restrict__ s, R *__restrict__ pred,
                                                    R *__restrict__ ratio, int64_t count, R delta) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= count) return;
  R p = hf[idx];
  if (s) p = p + s[idx];
  if (pred) pred[idx] = p;
  ratio[idx] = g[idx] / (p + delta);
}

// RelativeDifferencePrior.phi1 (priors/nearest_neighbour.py:176-177) summed
// over the 26 replicate-clamped neighbours (utils/misc.py:40-57) with
// EuclideanNeighbourWeight (:207-220), in the reference's loop order.
template <typename R>
__global__ __launch_bounds__(256) void rdp_kernel(const R *__restrict__ f, R *__restrict__ out,
                                                  int nx, int ny, int nz, RdpWeights<R> W,
                                                  R beta_scale, R gamma, R delta) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t count = (int64_t)nx * ny * nz;
  if (idx >= count) return;
  const int z = idx % nz;
  const int y = (idx / nz) % ny;
  const int x = idx / ((int64_t)nz * ny);
  const R fr = f[idx];
  R acc = R(0);
#pragma unroll
  for (int i = -1; i <= 1; ++i)
#pragma unroll
    for (int j = -1; j <= 1; ++j)
#pragma unroll
      for (int k = -1; k <= 1; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        const R w = W.w[(i + 1) * 9 + (j + 1) * 3 + (k + 1)];
        const int xs = min(max(x + i, 0), nx - 1);
        const int ys = min(max(y + j, 0), ny - 1);
        const int zs = min(max(z + k, 0), nz - 1);
        const R fs = f[((int64_t)xs * ny + ys) * nz + zs];
        const R d = fr - fs;
        const R ad = fabs(d);
        const R num = d * (gamma * ad + R(3) * fs + fr + R(2) * delta);
        const R den = fr + fs + gamma * ad + delta;
        acc = acc + (num / (den * den)) * w;
      }
  out[idx] = beta_scale * acc;
}

// Transposed Jacobian of the RDP gradient times u: with r_r = c sum_s w_s
// phi1(f_r, f_nb(r,s)) (nb = replicate-clamped neighbour),
//   (J^T u)_q = c sum_s w_s [ u_q phi2_1(f_q, f_nb(q,s))
//                             + sum_{r: nb(r,s) = q} u_r phi2_2(f_r, f_q) ]
// with phi2_1 = d phi1 / d f_r and phi2_2 = d phi1 / d f_s as in
// RelativeDifferencePrior.phi2_1 / phi2_2 (priors/nearest_neighbour.py:179-183).
// The second sum is a gather: per axis, the r with clamp(r + d) = q.
template <typename R>
__global__ __launch_bounds__(256) void rdp_hvp_kernel(const R *__restrict__ f,
                                                      const R *__restrict__ u,
                                                      R *__restrict__ out, int nx, int ny, int nz,
                                                      RdpWeights<R> W, R c, R gamma, R delta) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t count = (int64_t)nx * ny * nz;
  if (idx >= count) return;
  const int z = idx % nz;
  const int y = (idx / nz) % ny;
  const int x = idx / ((int64_t)nz * ny);
  const int n3[3] = {nx, ny, nz};
  const int q3[3] = {x, y, z};
  const R fq = f[idx], uq = u[idx];
  auto phi2_1 = [&](R fr, R fs) {
    const R den = gamma * fabs(fr - fs) + fr + fs + delta;
    const R a = R(2) * fs + delta;
    return (R(2) * a * a) / (den * den * den);
  };
  auto phi2_2 = [&](R fr, R fs) {
    const R den = gamma * fabs(fr - fs) + fr + fs + delta;
    return -(R(2) * (R(2) * fr + delta) * (R(2) * fs + delta)) / (den * den * den);
  };
  R acc = R(0);
#pragma unroll
  for (int i = -1; i <= 1; ++i)
#pragma unroll
    for (int j = -1; j <= 1; ++j)
#pragma unroll
      for (int k = -1; k <= 1; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        const R w = W.w[(i + 1) * 9 + (j + 1) * 3 + (k + 1)];
        const int d3[3] = {i, j, k};
        // own term: neighbour of q
        const int xs = min(max(x + i, 0), nx - 1), ys = min(max(y + j, 0), ny - 1),
                  zs = min(max(z + k, 0), nz - 1);
        R term = uq * phi2_1(fq, f[((int64_t)xs * ny + ys) * nz + zs]);
        // voxels r whose clamped neighbour along (i, j, k) is q
        int cand[3][2], nc[3];
        for (int a = 0; a < 3; ++a) {
          nc[a] = 0;
          const int d = d3[a], qa = q3[a], na = n3[a];
          if (d == 0) {
            cand[a][nc[a]++] = qa;
          } else if (d == 1) {
            if (qa >= 1) cand[a][nc[a]++] = qa - 1;
            if (qa == na - 1) cand[a][nc[a]++] = na - 1;
          } else {
            if (qa <= na - 2) cand[a][nc[a]++] = qa + 1;
            if (qa == 0) cand[a][nc[a]++] = 0;
          }
        }
        for (int a = 0; a < nc[0]; ++a)
          for (int b = 0; b < nc[1]; ++b)
            for (int e = 0; e < nc[2]; ++e) {
              const int64_t r = ((int64_t)cand[0][a] * ny + cand[1][b]) * nz + cand[2][e];
              term += u[r] * phi2_2(f[r], fq);
            }
        acc += term * w;
      }
  out[idx] = c * acc;
}

// The prior's derivative of order 2 applied to u, as NearestNeighbourPrior.__call__(2)
// (priors/nearest_neighbour.py:105-108): out = c * [u_r sum_s w_s phi2_1(f_r, f_s)]
// + c * [sum_s w_s phi2_2(f_r, f_s) u_s], s over the replicate-clamped neighbours
// (the Jacobian of the prior gradient times u; rdp_hvp_kernel is its transpose).
template <typename R>
__global__ __launch_bounds__(256) void rdp_jvp_kernel(const R *__restrict__ f,
                                                      const R *__restrict__ u,
                                                      R *__restrict__ out, int nx, int ny, int nz,
                                                      RdpWeights<R> W, R c, R gamma, R delta) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t count = (int64_t)nx * ny * nz;
  if (idx >= count) return;
  const int z = idx % nz;
  const int y = (idx / nz) % ny;
  const int x = idx / ((int64_t)nz * ny);
  const R fr = f[idx];
  R diag = R(0), blur = R(0);
#pragma unroll
  for (int i = -1; i <= 1; ++i)
#pragma unroll
    for (int j = -1; j <= 1; ++j)
#pragma unroll
      for (int k = -1; k <= 1; ++k) {
        if (i == 0 && j == 0 && k == 0) continue;
        const R w = W.w[(i + 1) * 9 + (j + 1) * 3 + (k + 1)];
        const int64_t s = ((int64_t)min(max(x + i, 0), nx - 1) * ny + min(max(y + j, 0), ny - 1)) * nz +
                          min(max(z + k, 0), nz - 1);
        const R fs = f[s];
        const R den = gamma * fabs(fr - fs) + fr + fs + delta;
        const R den3 = den * den * den;
        const R a = R(2) * fs + delta;
        const R p21 = (R(2) * (a * a)) / den3;
        const R p22 = -((R(2) * (R(2) * fr + delta)) * (R(2) * fs + delta)) / den3;
        diag = diag + p21 * w;
        blur = blur + (p22 * u[s]) * w;
      }
  out[idx] = (c * diag) * u[idx] + c * blur;
}

// PreconditionedGradientAscentAlgorithm.__call__ (preconditioned_gradient_ascent.py:119-122)
// with BSREM's preconditioner (:461-474).
template <typename R>
__global__ __launch_bounds__(256) void bsrem_update_kernel(R *__restrict__ f, const R *__restrict__ b,
                                                           const R *__restrict__ nrm,
                                                           const R *__restrict__ nsum,
                                                           const R *__restrict__ prior,
                                                           int64_t count, R relax, R weight,
                                                           R delta, R floor_value) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
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

template <typename R>
hipError_t launch_ratio(const R *g, const R *hf, const R *s, R *pred, R *ratio, int64_t count,
                        R delta, hipStream_t st) {
  ratio_kernel<R><<<(unsigned)((count + 255) / 256), 256, 0, st>>>(g, hf, s, pred, ratio, count,
                                                                   delta);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_rdp(const R *f, R *out, int nx, int ny, int nz, double dx, double dy, double dz,
                      R beta_scale, R gamma, R delta, hipStream_t st) {
  const int64_t count = (int64_t)nx * ny * nz;
  rdp_kernel<R><<<(unsigned)((count + 255) / 256), 256, 0, st>>>(
      f, out, nx, ny, nz, rdp_weights<R>(dx, dy, dz), beta_scale, gamma, delta);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_rdp_hvp(const R *f, const R *u, R *out, int nx, int ny, int nz, double dx,
                          double dy, double dz, R c, R gamma, R delta, hipStream_t st) {
  const int64_t count = (int64_t)nx * ny * nz;
  rdp_hvp_kernel<R><<<(unsigned)((count + 255) / 256), 256, 0, st>>>(
      f, u, out, nx, ny, nz, rdp_weights<R>(dx, dy, dz), c, gamma, delta);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_rdp_jvp(const R *f, const R *u, R *out, int nx, int ny, int nz, double dx,
                          double dy, double dz, R c, R gamma, R delta, hipStream_t st) {
  const int64_t count = (int64_t)nx * ny * nz;
  rdp_jvp_kernel<R><<<(unsigned)((count + 255) / 256), 256, 0, st>>>(
      f, u, out, nx, ny, nz, rdp_weights<R>(dx, dy, dz), c, gamma, delta);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_bsrem_update(R *f, const R *b, const R *nrm, const R *nsum, const R *prior,
                               int64_t count, R relax, R weight, R delta, R floor_value,
                               hipStream_t st) {
  bsrem_update_kernel<R><<<(unsigned)((count + 255) / 256), 256, 0, st>>>(
      f, b, nrm, nsum, prior, count, relax, weight, delta, floor_value);
  return hipGetLastError();
}

#define SPECT_INST(R)                                                                          \
  template hipError_t launch_ratio<R>(const R *, const R *, const R *, R *, R *, int64_t, R,   \
                                      hipStream_t);                                            \
  template hipError_t launch_rdp<R>(const R *, R *, int, int, int, double, double, double, R,  \
                                    R, R, hipStream_t);                                        \
  template hipError_t launch_rdp_hvp<R>(const R *, const R *, R *, int, int, int, double,      \
                                        double, double, R, R, R, hipStream_t);                 \
  template hipError_t launch_rdp_jvp<R>(const R *, const R *, R *, int, int, int, double,      \
                                        double, double, R, R, R, hipStream_t);                 \
  template hipError_t launch_bsrem_update<R>(R *, const R *, const R *, const R *, const R *,  \
                                             int64_t, R, R, R, R, hipStream_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
