// Attenuation inputs of the projector bodies for the CPU builds, made by the
// attenuation march bodies of forward_body.hpp over all angles, as
// spect_att_field_build / spect_att_table_build make them (grid (A, ntiles)).
// Both start as NaN, so a read of an entry the march never wrote shows.
#pragma once
#include <numeric>

#include "cpu_runtime.hpp"
#include "forward_body.hpp"

namespace cpu_build {

struct Attenuation {
  std::vector<float> field;   // att_mode 1: [A][P][je1-je0][Nz]
  std::vector<double> table;  // att_mode 2: [A][C][je1-je0][Nz]
  const float *mu = nullptr, *attf = nullptr;
  const double *att = nullptr;
};

inline Attenuation build_attenuation(const Problem &p) {
  Attenuation a;
  if (p.att_mode == 0) return a;
  constexpr int RZ = spect::FwdTile<float>::RZ;
  const int runs = (p.Nz + RZ - 1) / RZ;
  const int bt = (spect::kTJ * runs + 63) / 64 * 64;  // launch_att_field / launch_att_table
  std::vector<int32_t> all(p.A);
  std::iota(all.begin(), all.end(), 0);
  const spect::Tables<float> tb = p.tables();
  const size_t n_jext = p.g.je1 - p.g.je0;
  std::vector<unsigned char> smem(16);  // the march uses static LDS only
  if (p.att_mode == 1) {
    a.field.assign((size_t)p.A * p.P * n_jext * p.Nz, std::numeric_limits<float>::quiet_NaN());
    run_grid(p.A, p.g.ntiles, 1, bt, smem, [&](int al, int jt, int, unsigned char *) {
      spect::att_field_body<float, RZ>(p.g, tb, p.mu.data(), all.data(), a.field.data(), al, jt);
    });
    a.attf = a.field.data();
  } else {
    a.table.assign((size_t)p.A * p.C * n_jext * p.Nz, std::numeric_limits<double>::quiet_NaN());
    run_grid(p.A, p.g.ntiles, 1, bt, smem, [&](int al, int jt, int, unsigned char *) {
      spect::att_table_body<float, RZ>(p.g, tb, p.mu.data(), all.data(), a.table.data(), al, jt);
    });
    a.att = a.table.data();
  }
  a.mu = p.mu.data();
  return a;
}

}  // namespace cpu_build
