// tiers.hpp -- per-plane effective widths of the PSF taps (host code).
//
// SPECTPSFTransform fixes one kernel size K = 2H+1 for all planes of a radius
// (transforms/SPECT/psf.py:170-178), sized for the widest plane.  Planes
// closer to the detector have a much narrower Gaussian: at 128^3 with
// sigma(d) = 0.03 d + 0.1 cm the half width needed to carry all but 2^-26 of
// a plane's central tap falls from 14 taps at the far side to 1 behind the
// detector.  The blur kernels therefore run each plane with a compile-time
// half width ("tier") taken from a short list, and skip the taps outside it;
// what they drop is below the stated fraction of the plane's central tap.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace spect {

// Half widths a kernel compiled for half width H instantiates: H, H - 4,
// H - 8, ... down to 2 or 4, and 2.
inline bool tier_available(int h, int H) { return h == H || h == 2 || (h >= 2 && (H - h) % 4 == 0); }

inline int tier_round_up(int h, int H) {
  int t = h < 2 ? 2 : h;
  while (t < H && !tier_available(t, H)) ++t;
  return t > H ? H : t;
}

// Smallest half width h such that the taps farther than h from the centre of
// one plane's row (K = 2H+1 entries) add up to at most eps times the central
// tap, for the y row and the z row alike.
template <typename T>
int plane_half_width(const T *wr, const T *wz, int H, double eps) {
  for (int h = 0; h < H; ++h) {
    double tr = 0.0, tz = 0.0;
    for (int d = h + 1; d <= H; ++d) {
      tr += std::fabs((double)wr[H - d]) + std::fabs((double)wr[H + d]);
      tz += std::fabs((double)wz[H - d]) + std::fabs((double)wz[H + d]);
    }
    if (tr <= eps * std::fabs((double)wr[H]) && tz <= eps * std::fabs((double)wz[H])) return h;
  }
  return H;
}

// Tier of every row of the widened tap tables [rows][2H+1].
template <typename T>
std::vector<int32_t> tap_tiers(const T *taps_r, const T *taps_z, int rows, int H, double eps) {
  std::vector<int32_t> out((size_t)rows, H);
  const int K = 2 * H + 1;
  for (int r = 0; r < rows; ++r)
    out[r] = tier_round_up(plane_half_width(taps_r + (size_t)r * K, taps_z + (size_t)r * K, H, eps), H);
  return out;
}

// Fractions of the central tap the float32 and float64 kernels may drop.
constexpr double kTierEps32 = 1.0 / (1 << 26);
constexpr double kTierEps64 = 1.0 / 72057594037927936.0;  // 2^-56

}  // namespace spect
