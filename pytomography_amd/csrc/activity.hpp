// activity.hpp -- which planes of a forward work block can touch the object
// (host code).
//
// A forward block covers one angle, one tile of kTJ rotated-frame columns and
// one chunk of L planes.  Its planes whose samples all fall outside the
// footprint square (lo, hi)^2 of the padded object (Geo::lo = pad - 2,
// Geo::hi = pad + N + 1: at least one pixel beyond the object on every side,
// so those samples have no bilinear tap inside the object) contribute
// nothing and are skipped.  The plan tabulates that set once per
// (angle, tile, chunk) as a bit mask, so a block reads one word instead of
// resampling its whole chunk.  The host evaluates the sample positions in
// double precision and widens the square by a quarter pixel, so the mask is a
// superset of what the float or double kernels would find: a plane it keeps
// needlessly only costs time; a plane it drops has no tap inside the object.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace spect {

// mask[(a * ntiles + jt) * C + c], bit l: plane c * L + l is visited.
// (cosv, sinv)[a] of theta_f for angle a; lin = linspace(-1, 1, P).  With
// ranges, also (i_min, i_max) per (a, jt) of all visited planes (the set is
// an interval: rows of a convex square cut by a segment of columns);
// i_max < i_min when there are none.
inline std::vector<uint64_t> forward_activity(const double *cosv, const double *sinv, int A,
                                              const double *lin, int P, int je0, int je1, int L,
                                              int C, int tile, double lo, double hi,
                                              std::vector<int> *ranges = nullptr) {
  const int ntiles = (je1 - je0 + tile - 1) / tile;
  const double half = (P - 1) / 2.0;
  const double lo_w = lo - 0.25, hi_w = hi + 0.25;
  std::vector<uint64_t> mask((size_t)A * ntiles * C, 0);
  if (ranges) ranges->assign((size_t)A * ntiles * 2, 0);
  for (int a = 0; a < A; ++a)
    for (int jt = 0; jt < ntiles; ++jt) {
      const int j0 = je0 + jt * tile, j1 = std::min(je1, j0 + tile);
      int imin = P, imax = -1;
      for (int i = 0; i < P; ++i) {
        bool hit = false;
        for (int j = j0; j < j1 && !hit; ++j) {
          const double ix = (cosv[a] * lin[j] - sinv[a] * lin[i] + 1.0) * half;
          const double iy = (sinv[a] * lin[j] + cosv[a] * lin[i] + 1.0) * half;
          hit = ix > lo_w && ix < hi_w && iy > lo_w && iy < hi_w;
        }
        if (hit) {
          mask[((size_t)a * ntiles + jt) * C + i / L] |= 1ull << (i % L);
          imin = std::min(imin, i);
          imax = std::max(imax, i);
        }
      }
      if (ranges) {
        // the interval covers every visited plane (planes between two visited
        // ones are added, which only costs time)
        (*ranges)[((size_t)a * ntiles + jt) * 2] = imax >= 0 ? imin : 0;
        (*ranges)[((size_t)a * ntiles + jt) * 2 + 1] = imax;
      }
    }
  return mask;
}

}  // namespace spect
