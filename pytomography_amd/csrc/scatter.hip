// scatter.hip -- energy-window scatter estimate of the additive term
// (utils/scatter.py:10-99): the window combination and the Gaussian smoothing
// along one projection axis.  Both are HBM streams over [A, Lr, Lz]
// projections; the smoothing reads each input row through L2 (K taps of a
// short row), so one launch costs ~2 x the projection bytes.
#include "../../include/spect_hip.h"
#include "launch.hpp"

namespace spect {
namespace {

// compute_EW_scatter (utils/scatter.py:52-99), evaluated op by op in R as the
// reference's tensor-scalar expressions are:
//   estimate = (lower / wl * kl + upper / wu * ku) * wp
//   variance = c1 * lower + c2 * upper,   c1 = (wp / wl * kl)^2, c2 likewise
// An absent upper window contributes an exact 0 to both.
template <typename R>
__global__ __launch_bounds__(256) void ew_scatter_kernel(const R *__restrict__ lower,
                                                         const R *__restrict__ upper,
                                                         R *__restrict__ est, R *__restrict__ var,
                                                         int64_t count, R wl, R wu, R wp, R kl,
                                                         R ku, R c1, R c2) {
#pragma clang fp contract(off)
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= count) return;
  const R lo = lower[idx];
  R a = lo / wl;
  a = a * kl;
  R up = R(0);
  if (upper) {
    up = upper[idx];
    R b = up / wu;
    b = b * ku;
    a = a + b;
  }
  est[idx] = a * wp;
  if (var) {
    R v = c1 * lo;
    if (upper) v = v + c2 * up;
    var[idx] = v;
  }
}

// One torch Conv1d(padding='same', padding_mode=mode) with a symmetric
// kernel along `axis` of an [n0, n1, n2] array (get_smoothed_scatter,
// utils/scatter.py:41-49): out[.., n, ..] = sum_t w[t] in[.., n + t - H, ..]
// with the index wrapped (circular), clamped (replicate) or the term dropped
// (zeros).
template <typename R>
__global__ __launch_bounds__(256) void smooth_axis_kernel(const R *__restrict__ in,
                                                          R *__restrict__ out, int n0, int n1,
                                                          int n2, int axis,
                                                          const R *__restrict__ taps, int K,
                                                          int mode) {
#pragma clang fp contract(off)
  const int64_t count = (int64_t)n0 * n1 * n2;
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= count) return;
  const int i2 = idx % n2;
  const int i1 = (idx / n2) % n1;
  const int i0 = idx / ((int64_t)n2 * n1);
  const int len = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  const int pos = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
  const int64_t stride = axis == 0 ? (int64_t)n1 * n2 : (axis == 1 ? n2 : 1);
  const R *base = in + idx - (int64_t)pos * stride;
  const int H = K / 2;
  R acc = R(0);
  for (int t = 0; t < K; ++t) {
    int q = pos + t - H;
    if (q < 0 || q >= len) {
      if (mode == SPECT_PAD_ZEROS) continue;
      if (mode == SPECT_PAD_REPLICATE) {
        q = q < 0 ? 0 : len - 1;
      } else {
        q %= len;
        if (q < 0) q += len;
      }
    }
    acc = acc + taps[t] * base[(int64_t)q * stride];
  }
  out[idx] = acc;
}

}  // namespace

template <typename R>
hipError_t launch_ew_scatter(const R *lower, const R *upper, R *est, R *var, int64_t count, R wl,
                             R wu, R wp, R kl, R ku, R c1, R c2, hipStream_t st) {
  const int64_t blocks = (count + 255) / 256;
  ew_scatter_kernel<R><<<dim3((unsigned)blocks), 256, 0, st>>>(lower, upper, est, var, count, wl,
                                                               wu, wp, kl, ku, c1, c2);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_smooth_axis(const R *in, R *out, int n0, int n1, int n2, int axis, const R *taps,
                              int K, int mode, hipStream_t st) {
  const int64_t count = (int64_t)n0 * n1 * n2;
  const int64_t blocks = (count + 255) / 256;
  smooth_axis_kernel<R><<<dim3((unsigned)blocks), 256, 0, st>>>(in, out, n0, n1, n2, axis, taps,
                                                                K, mode);
  return hipGetLastError();
}

#define SPECT_INST(R)                                                                          \
  template hipError_t launch_ew_scatter<R>(const R *, const R *, R *, R *, int64_t, R, R, R, R, \
                                           R, R, R, hipStream_t);                              \
  template hipError_t launch_smooth_axis<R>(const R *, R *, int, int, int, int, const R *, int, \
                                            int, hipStream_t);
SPECT_INST(float)
SPECT_INST(double)
#undef SPECT_INST

}  // namespace spect
