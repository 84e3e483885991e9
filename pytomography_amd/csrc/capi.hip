// capi.hip -- plan management and the extern "C" entry points of libspect_hip.so
// (declared in include/spect_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spect_hip.h"
#include "launch.hpp"
#include "activity.hpp"
#include "tiers.hpp"

#define SPECT_VERSION 2

using namespace spect;

struct spect_plan {
  int N, Nz, A, P, pad;
  double dx, dy, dz;
  int H;                 // PSF half width used by the kernels (0: none)
  int L, C;              // planes per chunk, chunk count
  int rounds;            // rounds of the balanced schedule (0: automatic, -1: uniform chunks)
  int pieces_max = 0;    // most pieces of kPieceMax planes one angle's tiles need
  int planes_max = 0;    // most visited planes one angle's tiles hold
  int n_radii;
  bool has64;
  int device;
  float2 *trig32 = nullptr;
  double2 *trig64 = nullptr;
  float *lin32 = nullptr;
  double *lin64 = nullptr;
  int32_t *radius_index = nullptr;
  int32_t *all_angles = nullptr;  // 0..A-1
  float *taps_r32 = nullptr, *taps_z32 = nullptr;
  double *taps_r64 = nullptr, *taps_z64 = nullptr;
  int32_t *tier32 = nullptr, *tier64 = nullptr;  // [n_radii][P] blur tier per plane
  unsigned long long *act = nullptr;  // [A][ntiles][C] visited planes per forward block
  int2 *ranges = nullptr;             // [A][ntiles] visited plane range per (angle, tile)
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(SPECT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                          \
  do {                                                         \
    hipError_t e_ = (expr);                                    \
    if (e_ != hipSuccess) return hip_fail(e_, #expr);          \
  } while (0)

int host_pad_size(int width) { return (int)std::ceil((std::sqrt(2.0) * width - width) / 2.0); }

// torch.linspace(-1, 1, P) as ATen's CPU range kernel computes it: a step in
// the output dtype, ascending from -1 for the first half, descending from +1.
template <typename R> std::vector<R> host_linspace(int P) {
  std::vector<R> v(P);
  if (P == 1) {
    v[0] = R(-1);
    return v;
  }
  const R step = (R(1) - R(-1)) / R(P - 1);
  const int halfway = P / 2;
  for (int k = 0; k < P; ++k) v[k] = k < halfway ? R(-1) + step * R(k) : R(1) - step * R(P - k - 1);
  return v;
}

template <typename T> int upload(T **dst, const T *src, size_t count) {
  HIP_TRY(hipMalloc((void **)dst, count * sizeof(T)));
  HIP_TRY(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return SPECT_OK;
}

// Re-centre per-plane taps of width k into width K = 2H+1 (zero taps outside):
// a zero tap adds an exact 0 to the correlation.
template <typename T>
std::vector<T> widen_taps(const T *src, int rows, int k, int K, bool unit_if_null) {
  std::vector<T> out((size_t)rows * K, T(0));
  const int off = (K - k) / 2;
  for (int r = 0; r < rows; ++r) {
    if (src) {
      for (int t = 0; t < k; ++t) out[(size_t)r * K + off + t] = src[(size_t)r * k + t];
    } else if (unit_if_null) {
      out[(size_t)r * K + K / 2] = T(1);
    }
  }
  return out;
}

template <typename R> Tables<R> tables(const spect_plan *p);
template <> Tables<float> tables<float>(const spect_plan *p) {
  return {p->trig32, p->lin32, p->radius_index, p->taps_r32, p->taps_z32, p->tier32, p->act,
          p->ranges};
}
template <> Tables<double> tables<double>(const spect_plan *p) {
  return {p->trig64, p->lin64, p->radius_index, p->taps_r64, p->taps_z64, p->tier64, p->act,
          p->ranges};
}

template <typename R> Geo<R> geo(const spect_plan *p) {
  Geo<R> g;
  g.N = p->N;
  g.Nz = p->Nz;
  g.P = p->P;
  g.pad = p->pad;
  g.A = p->A;
  g.H = p->H;
  g.jo0 = p->pad;
  g.jo1 = p->pad + p->N;
  g.je0 = std::max(0, g.jo0 - p->H);
  g.je1 = std::min(p->P, g.jo1 + p->H);
  g.L = p->L;
  g.C = p->C;
  g.ntiles = (g.je1 - g.je0 + kTJ - 1) / kTJ;
  g.half = R(p->P - 1) / R(2);
  g.dx = R(p->dx);
  g.lo = R(p->pad) - R(2);
  g.hi = R(p->pad + p->N) + R(1);
  return g;
}

struct Ws {
  size_t att_off, part_off, w_off, ratio_off, total;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// The forward's partials for n angles: room for the chunk layout
// [C][n][ntiles] and for the balanced schedule's slots (one per block) of
// (kTJ + 2H) x Nz values each, followed by the schedule's ints.
struct FwdPart {
  size_t sched_off, total;  // relative to the start of the partials
  int sched_blocks;         // blocks of the scheduled forward (0: not runnable)
};

// Automatic schedule choice (sched_rounds 0): the balanced schedule when the
// pieces of an n-angle call fit one round of `blocks` resident blocks even at
// the longest piece (kPieceMax planes); otherwise -- large problems, where the
// schedule would run several rounds of kPieceMax-plane pieces with a ragged
// last round -- uniform chunks, which the hardware dispatcher balances.  With
// max_q > 0 the call's planes must also average at most max_q per block: one
// round of long pieces cannot even out their different tap tiers.
bool sched_auto_ok(const spect_plan *p, int n, int blocks, int max_q = 0) {
  if (blocks <= 0) return false;
  if (p->rounds != 0) return true;
  if ((long long)n * p->pieces_max > (long long)blocks) return false;
  return max_q <= 0 || (long long)n * p->planes_max <= (long long)max_q * blocks;
}

// The forward's bound on planes per block for the automatic choice: measured
// with the balanced schedule ahead at up to 15.9 planes per block (160^3, 8
// angles) and behind at 22.2 (192^3, 8 angles).
constexpr int kFwdAutoPlanesPerBlock = 19;

FwdPart fwd_part_layout(const spect_plan *p, int dtype, int n) {
  const size_t rs = dtype == SPECT_F64 ? 8 : 4;
  const Geo<float> g = geo<float>(p);
  const size_t JO = kTJ + 2 * p->H;
  const size_t npairs = (size_t)n * g.ntiles;
  FwdPart l;
  const int res = p->rounds < 0 ? 0
                  : std::max(1, p->rounds) *
                        (dtype == SPECT_F64 ? forward_sched_blocks<double>(geo<double>(p))
                                            : forward_sched_blocks<float>(g));
  l.sched_blocks = sched_auto_ok(p, n, res, kFwdAutoPlanesPerBlock) ? res : 0;
  const size_t chunk = (size_t)p->C * npairs * JO * p->Nz * rs;
  const size_t sched = (size_t)sched_grid_blocks(n, g.ntiles, p->P, res) * JO * p->Nz * rs;
  l.sched_off = align256(std::max(chunk, sched));
  l.total = l.sched_off + align256((2 + npairs + 1) * sizeof(int));
  return l;
}

Ws ws_layout(const spect_plan *p, int dtype, int n) {
  const size_t rs = dtype == SPECT_F64 ? 8 : 4;
  const Geo<float> g = geo<float>(p);
  const size_t n_jext = g.je1 - g.je0;
  const size_t att = align256((size_t)n * p->C * n_jext * p->Nz * sizeof(double));
  const size_t part = fwd_part_layout(p, dtype, n).total;
  const size_t w = align256((size_t)n * p->P * n_jext * p->Nz * rs);
  const size_t ratio = align256((size_t)n * p->N * p->Nz * rs);
  Ws l;
  l.att_off = 0;
  l.part_off = att;
  l.w_off = att;  // forward partials are consumed before the field is written
  l.ratio_off = att + std::max(part, w);
  l.total = l.ratio_off + ratio;
  return l;
}

// Gradient calls: attenuation rows + W + Wm (FP64) + chunk totals (FP64).
struct GradWs {
  size_t att_off, w_off, wm_off, dsum_off, total;
};

GradWs grad_ws_layout(const spect_plan *p, int dtype, int n) {
  const size_t rs = dtype == SPECT_F64 ? 8 : 4;
  const Geo<float> g = geo<float>(p);
  const size_t n_jext = g.je1 - g.je0;
  GradWs l;
  l.att_off = 0;
  l.w_off = align256((size_t)n * p->C * n_jext * p->Nz * sizeof(double));
  l.wm_off = l.w_off + align256((size_t)n * p->P * n_jext * p->Nz * rs);
  l.dsum_off = l.wm_off + align256((size_t)n * p->P * n_jext * p->Nz * sizeof(double));
  l.total = l.dsum_off + align256((size_t)n * p->C * n_jext * p->Nz * sizeof(double));
  return l;
}

bool aligned16(const void *ptr) { return ((uintptr_t)ptr & 15) == 0; }

int check_call(const spect_plan *p, int dtype, int n, const void *ws, size_t ws_bytes) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  if (dtype == SPECT_F64 && !p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
  if (n <= 0 || n > p->A) return fail(SPECT_E_ARG, "angle count out of range");
  const size_t need = ws_layout(p, dtype, n).total;
  if (!ws || ws_bytes < need)
    return fail(SPECT_E_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if (!aligned16(ws)) return fail(SPECT_E_ARG, "workspace must be 16-byte aligned");
  return SPECT_OK;
}

int check_arrays(std::initializer_list<const void *> arrays) {
  for (const void *a : arrays)
    if (a && !aligned16(a)) return fail(SPECT_E_ARG, "arrays must be 16-byte aligned");
  return SPECT_OK;
}

// Attenuation rows for the call: the caller's table (indexed by global angle)
// or rows built now in the workspace (indexed by position in angle_idx).
template <typename R>
int att_rows(const spect_plan *p, const Geo<R> &g, const R *mu, const void *table, bool field,
             const int32_t *idx, int n, double *rows, hipStream_t st, const double **att,
             bool *global, const R **attf) {
  *att = nullptr;
  *global = false;
  *attf = nullptr;
  if (field) {  // precomputed attenuation factors of every angle
    if (!table) return fail(SPECT_E_ARG, "SPECT_ATT_FIELD needs the attenuation field");
    *attf = reinterpret_cast<const R *>(table);
    return SPECT_OK;
  }
  if (!mu) return SPECT_OK;
  if (table) {
    *att = reinterpret_cast<const double *>(table);
    *global = true;
    return SPECT_OK;
  }
  HIP_TRY(launch_att_table<R>(g, tables<R>(p), mu, idx, n, rows, st));
  *att = rows;
  return SPECT_OK;
}

// Forward projection with its scratch given explicitly: rows (FP64
// attenuation rows, used only with mu and no table) and part (the chunk
// partials).
template <typename R>
int forward_core(const spect_plan *p, const R *f, const R *mu, const void *table, bool field,
                 const int32_t *idx, int n, R *proj, const R *gm, const R *s, R *pred, R *ratio,
                 R delta, double *rows, R *part, hipStream_t st, bool trb) {
  const Geo<R> g = geo<R>(p);
  const double *att;
  bool global;
  const R *attf;
  int rc = att_rows<R>(p, g, mu, table, field, idx, n, rows, st, &att, &global, &attf);
  if (rc) return rc;
  const FwdPart pl = fwd_part_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, n);
  if (!trb && (attf || !mu) && pl.sched_blocks > 0) {
    // the balanced schedule: every block carries the same number of planes
    int *sched = reinterpret_cast<int *>(reinterpret_cast<unsigned char *>(part) + pl.sched_off);
    HIP_TRY(launch_forward_sched<R>(g, tables<R>(p), f, attf, idx, n, part, sched,
                                    pl.sched_blocks, st));
    HIP_TRY(launch_reduce_sched<R>(g, part, n, sched, proj, gm, s, pred, ratio, delta, st));
    return SPECT_OK;
  }
  HIP_TRY(launch_forward<R>(g, tables<R>(p), f, mu, att, global, attf, idx, n, part, trb, st));
  HIP_TRY(launch_forward_reduce<R>(g, part, n, proj, gm, s, pred, ratio, delta, st));
  return SPECT_OK;
}

template <typename R>
int forward_impl(const spect_plan *p, const R *f, const R *mu, const void *table, bool field,
                 const int32_t *idx, int n, R *proj, const R *gm, const R *s, R *pred, R *ratio,
                 R delta, unsigned char *ws, hipStream_t st, bool trb = false) {
  const Ws l = ws_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, n);
  return forward_core<R>(p, f, mu, table, field, idx, n, proj, gm, s, pred, ratio, delta,
                         reinterpret_cast<double *>(ws + l.att_off),
                         reinterpret_cast<R *>(ws + l.part_off), st, trb);
}

template <typename R>
int backward_impl(const spect_plan *p, int mode, int out_mode, const R *proj, const R *mu,
                  const void *table, bool field, const int32_t *idx, int n, R *out, const R *f,
                  const R *nrm, double delta, double floor_value, unsigned char *ws,
                  hipStream_t st) {
  const Geo<R> g = geo<R>(p);
  const Ws l = ws_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, n);
  const double *att;
  bool global;
  const R *attf;
  int rc = att_rows<R>(p, g, mu, table, field, idx, n, reinterpret_cast<double *>(ws + l.att_off),
                       st, &att, &global, &attf);
  if (rc) return rc;
  if (!mu && p->H == 0) {
    // Without attenuation and PSF the field is the projection repeated along
    // the rows (pad_object 'back_project', utils/spatial.py:76-79): the
    // gather reads the projection itself (same values, no field pass).
    const size_t nout = (size_t)(g.jo1 - g.jo0);
    HIP_TRY(launch_bp_gather<R>(g, tables<R>(p), proj, idx, n, mode, out_mode, out, f, nrm,
                                R(delta), R(floor_value), nullptr, nullptr, nullptr, st,
                                nout * g.Nz, 0));
    return SPECT_OK;
  }
  R *W = reinterpret_cast<R *>(ws + l.w_off);
  const int fb = (attf || !mu) && p->rounds >= 0 ? std::max(1, p->rounds) * field_sched_blocks<R>(g)
                                                  : 0;
  if (sched_auto_ok(p, n, fb)) {
    // the balanced schedule (as the forward's, for this kernel's blocks)
    HIP_TRY(launch_bp_field_sched<R>(g, tables<R>(p), proj, attf, idx, n, W, fb, st));
  } else {
    HIP_TRY(launch_bp_field<R>(g, tables<R>(p), proj, mu, att, global, attf, idx, n, W, nullptr,
                               nullptr, nullptr, st));
  }
  HIP_TRY(launch_bp_gather<R>(g, tables<R>(p), W, idx, n, mode, out_mode, out, f, nrm, R(delta),
                              R(floor_value), nullptr, nullptr, nullptr, st));
  return SPECT_OK;
}

template <typename R>
int grad_impl(const spect_plan *p, const R *u, const R *f, const R *mu, const int32_t *idx, int n,
              R *grad_f, R *grad_mu, unsigned char *ws, hipStream_t st) {
  const Geo<R> g = geo<R>(p);
  const GradWs l = grad_ws_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, n);
  R *W = reinterpret_cast<R *>(ws + l.w_off);
  const double *att = nullptr;
  if (mu) {
    double *rows = reinterpret_cast<double *>(ws + l.att_off);
    HIP_TRY(launch_att_table<R>(g, tables<R>(p), mu, idx, n, rows, st));
    att = rows;
  }
  if (grad_mu) {
    double *Wm = reinterpret_cast<double *>(ws + l.wm_off);
    double *Dsum = reinterpret_cast<double *>(ws + l.dsum_off);
    HIP_TRY(launch_bp_field<R>(g, tables<R>(p), u, mu, att, false, nullptr, idx, n, W, f, Wm, Dsum,
                               st));
    HIP_TRY(launch_chunk_prefix<R>(g, Dsum, n, st));
    HIP_TRY(launch_bp_gather<R>(g, tables<R>(p), W, idx, n, 2, SPECT_OUT_STORE, grad_f, nullptr,
                                nullptr, R(0), R(0), Wm, Dsum, grad_mu, st));
  } else {
    HIP_TRY(launch_bp_field<R>(g, tables<R>(p), u, mu, att, false, nullptr, idx, n, W, nullptr,
                               nullptr, nullptr, st));
    HIP_TRY(launch_bp_gather<R>(g, tables<R>(p), W, idx, n, SPECT_BP_ADJOINT, SPECT_OUT_STORE,
                                grad_f, nullptr, nullptr, R(0), R(0), nullptr, nullptr, nullptr, st));
  }
  return SPECT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched OSEM
// ---------------------------------------------------------------------------
struct spect_batch {
  int dtype;
  int n_subsets;
  struct Group {
    int H, Nz;
    bool af;   // members pass precomputed attenuation factors
    int nprob;
    int maxL;  // largest chunk length among the members (LDS tap staging)
    // per subset: device offsets (bytes) into `tables` and block counts
    std::vector<size_t> fwd_off, red_off, fld_off, gat_off, starts_off[4];
    std::vector<int> blocks[4];
  };
  std::vector<Group> groups;
  void *tables = nullptr;  // device: job arrays and block starts
  std::vector<std::pair<int, hipGraphExec_t>> graphs;  // (n_iters, exec)
  hipStream_t capture_stream = nullptr;  // private stream the schedule is captured on
  // one stream and one join event per group: in a captured schedule the
  // groups are parallel branches (their problems share nothing)
  std::vector<hipStream_t> group_streams;
  std::vector<hipEvent_t> group_events;
  hipEvent_t fork_event = nullptr;
};

namespace {

template <typename R>
int batch_build(spect_batch *b, const spect_batch_problem *probs, int n, double delta,
                double floor_value) {
  // group problems by (H, Nz): a launch compiles one PSF width and one block shape
  std::vector<std::vector<int>> members;
  std::vector<std::pair<int, int>> keys;
  for (int i = 0; i < n; ++i) {
    // the key packs (H, attenuation-field flag) and Nz
    const std::pair<int, int> key{probs[i].plan->H * 2 + (probs[i].att_field ? 1 : 0),
                                  probs[i].plan->Nz};
    size_t k = 0;
    while (k < keys.size() && keys[k] != key) ++k;
    if (k == keys.size()) {
      keys.push_back(key);
      members.emplace_back();
    }
    members[k].push_back(i);
  }
  // Within a group the largest problems come first in the block order, so a
  // launch ends on the blocks of the small ones (results do not depend on it).
  for (auto &m : members)
    std::stable_sort(m.begin(), m.end(), [&](int a, int c) {
      const spect_plan *pa = probs[a].plan, *pc = probs[c].plan;
      return (double)pa->N * pa->N * pa->Nz * pa->A > (double)pc->N * pc->N * pc->Nz * pc->A;
    });
  const int S = b->n_subsets;
  std::vector<unsigned char> host;
  auto put = [&](const void *src, size_t bytes) {
    const size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + bytes);
    std::memcpy(host.data() + off, src, bytes);
    return off;
  };
  for (size_t k = 0; k < keys.size(); ++k) {
    spect_batch::Group G;
    G.H = keys[k].first / 2;
    G.af = (keys[k].first & 1) != 0;
    G.Nz = keys[k].second;
    G.nprob = (int)members[k].size();
    G.maxL = 0;
    for (int i : members[k]) G.maxL = std::max(G.maxL, probs[i].plan->L);
    for (int sub = 0; sub < S; ++sub) {
      std::vector<FwdJob<R>> fj;
      std::vector<ReduceJob<R>> rj;
      std::vector<FieldJob<R>> lj;
      std::vector<GatherJob<R>> gj;
      std::vector<int> st[4];
      int tot[4] = {0, 0, 0, 0};
      for (int i : members[k]) {
        const spect_batch_problem &pb = probs[i];
        const spect_plan *p = pb.plan;
        const Geo<R> g = geo<R>(p);
        const Tables<R> tb = tables<R>(p);
        const int cnt = pb.counts[sub];
        const Ws l = ws_layout(p, sizeof(R) == 8 ? SPECT_F64 : SPECT_F32, cnt);
        auto *ws = reinterpret_cast<unsigned char *>(pb.workspace);
        R *part = reinterpret_cast<R *>(ws + l.part_off);
        R *W = reinterpret_cast<R *>(ws + l.w_off);
        R *ratio = reinterpret_cast<R *>(ws + l.ratio_off);
        const R *mu = reinterpret_cast<const R *>(pb.mu);
        const double *att = pb.att_field ? nullptr : reinterpret_cast<const double *>(pb.att_table);
        const R *attf = pb.att_field ? reinterpret_cast<const R *>(pb.att_table) : nullptr;
        const int32_t *idx = pb.angle_idx[sub];
        R *f = reinterpret_cast<R *>(pb.f);
        fj.push_back(FwdJob<R>{g, tb, f, mu, att, 1, attf, idx, cnt, part});
        rj.push_back(ReduceJob<R>{g, part, cnt, nullptr, reinterpret_cast<const R *>(pb.g[sub]),
                                  pb.s ? reinterpret_cast<const R *>(pb.s[sub]) : nullptr,
                                  pb.pred ? reinterpret_cast<R *>(pb.pred[sub]) : nullptr, ratio,
                                  R(delta)});
        lj.push_back(FieldJob<R>{g, tb, ratio, mu, att, 1, attf, idx, cnt, W, nullptr, nullptr, nullptr});
        gj.push_back(GatherJob<R>{g, tb, W, idx, cnt, SPECT_OUT_OSEM, f, f,
                                  reinterpret_cast<const R *>(pb.norm[sub]), R(delta),
                                  R(floor_value), nullptr, nullptr, nullptr});
        const int nb[4] = {fwd_blocks(g, cnt), reduce_blocks(g, cnt), fwd_blocks(g, cnt),
                           gather_blocks(g)};
        for (int q = 0; q < 4; ++q) {
          st[q].push_back(tot[q]);
          tot[q] += nb[q];
        }
      }
      G.fwd_off.push_back(put(fj.data(), fj.size() * sizeof(fj[0])));
      G.red_off.push_back(put(rj.data(), rj.size() * sizeof(rj[0])));
      G.fld_off.push_back(put(lj.data(), lj.size() * sizeof(lj[0])));
      G.gat_off.push_back(put(gj.data(), gj.size() * sizeof(gj[0])));
      for (int q = 0; q < 4; ++q) {
        G.starts_off[q].push_back(put(st[q].data(), st[q].size() * sizeof(int)));
        G.blocks[q].push_back(tot[q]);
      }
    }
    b->groups.push_back(G);
  }
  HIP_TRY(hipMalloc(&b->tables, host.size()));
  HIP_TRY(hipMemcpy(b->tables, host.data(), host.size(), hipMemcpyHostToDevice));
  return SPECT_OK;
}

// Every sub-iteration of one group's problems: forward, reduce (Poisson
// ratio), back-projection field, gather with the OSEM update -- one launch
// each for all members.
template <typename R>
int batch_launch_group(spect_batch *b, const spect_batch::Group &G, int n_iters, hipStream_t st) {
  auto *base = reinterpret_cast<unsigned char *>(b->tables);
  for (int it = 0; it < n_iters; ++it) {
    for (int sub = 0; sub < b->n_subsets; ++sub) {
      HIP_TRY(launch_forward_batch<R>(reinterpret_cast<const FwdJob<R> *>(base + G.fwd_off[sub]),
                                      reinterpret_cast<const int *>(base + G.starts_off[0][sub]),
                                      G.nprob, G.blocks[0][sub], G.H, G.Nz, G.maxL, G.af, st));
      HIP_TRY(launch_reduce_batch<R>(reinterpret_cast<const ReduceJob<R> *>(base + G.red_off[sub]),
                                     reinterpret_cast<const int *>(base + G.starts_off[1][sub]),
                                     G.nprob, G.blocks[1][sub], st));
      HIP_TRY(launch_field_batch<R>(reinterpret_cast<const FieldJob<R> *>(base + G.fld_off[sub]),
                                    reinterpret_cast<const int *>(base + G.starts_off[2][sub]),
                                    G.nprob, G.blocks[2][sub], G.H, G.Nz, G.maxL, G.af, st));
      HIP_TRY(launch_gather_batch<R>(reinterpret_cast<const GatherJob<R> *>(base + G.gat_off[sub]),
                                     reinterpret_cast<const int *>(base + G.starts_off[3][sub]),
                                     G.nprob, G.blocks[3][sub], SPECT_BP_REFERENCE, st));
    }
  }
  return SPECT_OK;
}

// The whole schedule on one stream, group after group.
template <typename R>
int batch_launch(spect_batch *b, int n_iters, hipStream_t st) {
  for (const auto &G : b->groups)
    if (int rc = batch_launch_group<R>(b, G, n_iters, st)) return rc;
  return SPECT_OK;
}

// A stream and a join event per group, and the fork event (created before
// any capture begins).
int batch_branch_streams(spect_batch *b) {
  while (b->group_streams.size() < b->groups.size()) {
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b->group_streams.push_back(s);
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->group_events.push_back(e);
  }
  if (!b->fork_event) HIP_TRY(hipEventCreateWithFlags(&b->fork_event, hipEventDisableTiming));
  return SPECT_OK;
}

// The schedule with every group on its own stream, forked from and joined
// back into st (under capture: parallel branches of the graph).
template <typename R>
int batch_launch_branches(spect_batch *b, int n_iters, hipStream_t st) {
  const size_t ng = b->groups.size();
  HIP_TRY(hipEventRecord(b->fork_event, st));
  for (size_t k = 0; k < ng; ++k) {
    HIP_TRY(hipStreamWaitEvent(b->group_streams[k], b->fork_event, 0));
    if (int rc = batch_launch_group<R>(b, b->groups[k], n_iters, b->group_streams[k])) return rc;
    HIP_TRY(hipEventRecord(b->group_events[k], b->group_streams[k]));
    HIP_TRY(hipStreamWaitEvent(st, b->group_events[k], 0));
  }
  return SPECT_OK;
}

}  // namespace

extern "C" {

int spect_batch_create(const spect_batch_problem *probs, int n, int dtype, double delta,
                       double floor_value, spect_batch **out) {
  if (!probs || n <= 0 || !out) return fail(SPECT_E_ARG, "bad batch arguments");
  *out = nullptr;
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  const int S = probs[0].n_subsets;
  for (int i = 0; i < n; ++i) {
    const spect_batch_problem &pb = probs[i];
    if (!pb.plan || !pb.f || !pb.angle_idx || !pb.counts || !pb.g || !pb.norm || !pb.workspace)
      return fail(SPECT_E_ARG, "batch problem " + std::to_string(i) + ": null argument");
    if (pb.n_subsets != S || S <= 0) return fail(SPECT_E_ARG, "all problems need the same subset count");
    if (dtype == SPECT_F64 && !pb.plan->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
    if (pb.mu && !pb.att_table) return fail(SPECT_E_ARG, "batched problems with mu need an attenuation table");
    int cmax = 0;
    for (int k = 0; k < S; ++k) {
      if (pb.counts[k] <= 0 || pb.counts[k] > pb.plan->A || !pb.angle_idx[k] || !pb.g[k] || !pb.norm[k])
        return fail(SPECT_E_ARG, "batch problem " + std::to_string(i) + ": bad subset " + std::to_string(k));
      cmax = std::max(cmax, (int)pb.counts[k]);
    }
    if (pb.workspace_bytes < ws_layout(pb.plan, dtype, cmax).total)
      return fail(SPECT_E_WORKSPACE, "batch problem " + std::to_string(i) + ": workspace too small");
  }
  // The jobs of one launch run concurrently: a workspace (partials, ratio,
  // field) or an iterate shared by two problems would be written by both.
  {
    const size_t rs = dtype == SPECT_F64 ? 8 : 4;
    struct Span { uintptr_t lo, hi; int prob; const char *what; };
    std::vector<Span> spans;
    for (int i = 0; i < n; ++i) {
      const spect_batch_problem &pb = probs[i];
      int cmax = 0;
      for (int k = 0; k < S; ++k) cmax = std::max(cmax, (int)pb.counts[k]);
      const uintptr_t w = (uintptr_t)pb.workspace, f = (uintptr_t)pb.f;
      spans.push_back({w, w + ws_layout(pb.plan, dtype, cmax).total, i, "workspace"});
      spans.push_back({f, f + (size_t)pb.plan->N * pb.plan->N * pb.plan->Nz * rs, i, "iterate f"});
    }
    for (size_t a = 0; a < spans.size(); ++a)
      for (size_t c = a + 1; c < spans.size(); ++c)
        if (spans[a].lo < spans[c].hi && spans[c].lo < spans[a].hi)
          return fail(SPECT_E_ARG, std::string("batch problems ") + std::to_string(spans[a].prob) +
                                       " (" + spans[a].what + ") and " +
                                       std::to_string(spans[c].prob) + " (" + spans[c].what +
                                       ") overlap in device memory; each problem needs its own "
                                       "workspace and iterate");
  }
  auto *b = new spect_batch();
  b->dtype = dtype;
  b->n_subsets = S;
  const int rc = dtype == SPECT_F32 ? batch_build<float>(b, probs, n, delta, floor_value)
                                    : batch_build<double>(b, probs, n, delta, floor_value);
  if (rc) {
    spect_batch_destroy(b);
    return rc;
  }
  *out = b;
  return SPECT_OK;
}

int spect_batch_run(spect_batch *b, int n_iters, int use_graph, void *stream) {
  if (!b || n_iters <= 0) return fail(SPECT_E_ARG, "bad batch run arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!use_graph)
    return b->dtype == SPECT_F32 ? batch_launch<float>(b, n_iters, st) : batch_launch<double>(b, n_iters, st);
  for (auto &e : b->graphs)
    if (e.first == n_iters) {
      HIP_TRY(hipGraphLaunch(e.second, st));
      return SPECT_OK;
    }
  // Capture on a private stream (the caller's may be the null stream, which
  // cannot be captured); the graph is then launched on the caller's stream.
  if (!b->capture_stream) HIP_TRY(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
  if (int rc = batch_branch_streams(b)) return rc;
  hipStream_t cs = b->capture_stream;
  HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  const int rc = b->dtype == SPECT_F32 ? batch_launch_branches<float>(b, n_iters, cs)
                                       : batch_launch_branches<double>(b, n_iters, cs);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(cs, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
  b->graphs.emplace_back(n_iters, exec);
  HIP_TRY(hipGraphLaunch(exec, st));
  return SPECT_OK;
}

int spect_batch_destroy(spect_batch *b) {
  if (!b) return SPECT_OK;
  for (auto &e : b->graphs) (void)hipGraphExecDestroy(e.second);
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (hipStream_t s : b->group_streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->group_events) (void)hipEventDestroy(e);
  if (b->fork_event) (void)hipEventDestroy(b->fork_event);
  (void)hipFree(b->tables);
  delete b;
  return SPECT_OK;
}

}  // extern "C"

extern "C" {

const char *spect_last_error(void) { return g_err.c_str(); }

int spect_version(void) { return SPECT_VERSION; }

int spect_plan_create(const spect_geom *g, spect_plan **out) {
  if (!g || !out) return fail(SPECT_E_ARG, "null argument");
  *out = nullptr;
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return fail(SPECT_E_ARG, "object shape must be positive");
  if (g->nx != g->ny) return fail(SPECT_E_ARG, "nx must equal ny (square transaxial grid)");
  if (g->nz > 256) return fail(SPECT_E_ARG, "nz > 256 is not supported by the tile kernels");
  if (g->nx > 16384) return fail(SPECT_E_ARG, "nx > 16384 is not supported");
  if (g->chunk_planes < 0 || g->chunk_planes > 64)
    return fail(SPECT_E_ARG, "chunk_planes must be 0 (default) or 1..64");
  if (g->sched_rounds < -1 || g->sched_rounds > 16)
    return fail(SPECT_E_ARG, "sched_rounds must be -1, 0 (default) or 1..16");
  if (g->n_angles <= 0 || !g->angles_deg) return fail(SPECT_E_ARG, "angles required");
  if (!(g->dx > 0) || !(g->dz > 0)) return fail(SPECT_E_ARG, "voxel sizes must be positive");
  if (g->psf_kr < 0 || g->psf_kz < 0 || (g->psf_kr > 0 && g->psf_kr % 2 == 0) ||
      (g->psf_kz > 0 && g->psf_kz % 2 == 0))
    return fail(SPECT_E_ARG, "PSF tap counts must be odd");
  if (g->psf_kr > 0 && (g->n_radii <= 0 || !g->radius_index || !g->taps_r ||
                        (g->psf_kz > 0 && !g->taps_z)))
    return fail(SPECT_E_ARG, "PSF tables missing");
  if (g->psf_kr > 0 && g->psf_kz > 0 && g->taps_r64 && !g->taps_z64)
    return fail(SPECT_E_ARG, "float64 PSF tables: taps_z64 is required with taps_r64 when psf_kz > 0");
  const int hmax = std::max(g->psf_kr, g->psf_kz) / 2;
  int H = g->psf_kr > 0 ? (hmax + 1) / 2 * 2 : 0;  // even half widths are compiled
  if (g->psf_kr > 0 && H == 0) H = 2;
  if (H > 20) return fail(SPECT_E_ARG, "PSF wider than 41 taps is not supported");

  auto *p = new spect_plan();
  p->N = g->nx;
  p->Nz = g->nz;
  p->A = g->n_angles;
  p->pad = host_pad_size(g->ny);
  p->P = g->nx + 2 * p->pad;
  p->dx = g->dx;
  p->dy = g->dy;
  p->dz = g->dz;
  p->H = H;
  p->L = g->chunk_planes > 0 ? g->chunk_planes : 16;
  p->C = (p->P + p->L - 1) / p->L;
  p->rounds = g->sched_rounds;
  p->n_radii = g->psf_kr > 0 ? g->n_radii : 1;
  p->has64 = g->angles_deg64 != nullptr && (g->psf_kr == 0 || g->taps_r64 != nullptr);
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
    const double b64 = g->angles_deg64 ? g->angles_deg64[a] : (double)g->angles_deg[a];
    const double r64 = -(270.0 - b64) * M_PI / 180.0;
    t64[a] = make_double2(std::cos(r64), std::sin(r64));
  }
  std::vector<int32_t> ridx(p->A, 0), all(p->A);
  for (int a = 0; a < p->A; ++a) {
    all[a] = a;
    if (g->psf_kr > 0) {
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
  {
    // planes each forward block visits (activity.hpp), from the float64 geometry
    const Geo<double> g = geo<double>(p);
    std::vector<double> cv(p->A), sv(p->A);
    for (int a = 0; a < p->A; ++a) {
      cv[a] = t64[a].x;
      sv[a] = t64[a].y;
    }
    std::vector<int> rng;
    auto act = forward_activity(cv.data(), sv.data(), p->A, l64.data(), p->P, g.je0, g.je1, p->L,
                                p->C, kTJ, (double)g.lo, (double)g.hi, &rng);
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "64-bit masks");
    if ((rc = upload(&p->act, reinterpret_cast<const unsigned long long *>(act.data()), act.size())))
      return cleanup(rc);
    std::vector<int2> r2(rng.size() / 2);
    for (size_t k = 0; k < r2.size(); ++k) r2[k] = make_int2(rng[2 * k], rng[2 * k + 1]);
    // the most pieces of kPieceMax planes the tiles of one angle split into:
    // n times this bounds the pieces of any n-angle call at the longest piece
    const size_t nt = p->A > 0 ? r2.size() / (size_t)p->A : 0;
    for (int a = 0; a < p->A; ++a) {
      int pc = 0, pl = 0;
      for (size_t t = 0; t < nt; ++t) {
        const int2 r = r2[(size_t)a * nt + t];
        if (r.y >= r.x) {
          pc += (r.y - r.x + kPieceMax) / kPieceMax;
          pl += r.y - r.x + 1;
        }
      }
      p->pieces_max = std::max(p->pieces_max, pc);
      p->planes_max = std::max(p->planes_max, pl);
    }
    if ((rc = upload(&p->ranges, r2.data(), r2.size()))) return cleanup(rc);
  }
  if ((rc = upload(&p->radius_index, ridx.data(), ridx.size()))) return cleanup(rc);
  if ((rc = upload(&p->all_angles, all.data(), all.size()))) return cleanup(rc);
  if (H > 0) {
    const int K = 2 * H + 1;
    const int rows = p->n_radii * p->P;
    auto r32 = widen_taps<float>(g->taps_r, rows, g->psf_kr, K, false);
    auto z32 = widen_taps<float>(g->psf_kz > 0 ? g->taps_z : nullptr, rows, g->psf_kz, K, true);
    if ((rc = upload(&p->taps_r32, r32.data(), r32.size()))) return cleanup(rc);
    if ((rc = upload(&p->taps_z32, z32.data(), z32.size()))) return cleanup(rc);
    auto t32v = tap_tiers<float>(r32.data(), z32.data(), rows, H, kTierEps32);
    if ((rc = upload(&p->tier32, t32v.data(), t32v.size()))) return cleanup(rc);
    if (p->has64) {
      auto r64 = widen_taps<double>(g->taps_r64, rows, g->psf_kr, K, false);
      auto z64 = widen_taps<double>(g->psf_kz > 0 ? g->taps_z64 : nullptr, rows, g->psf_kz, K, true);
      if ((rc = upload(&p->taps_r64, r64.data(), r64.size()))) return cleanup(rc);
      if ((rc = upload(&p->taps_z64, z64.data(), z64.size()))) return cleanup(rc);
      auto t64v = tap_tiers<double>(r64.data(), z64.data(), rows, H, kTierEps64);
      if ((rc = upload(&p->tier64, t64v.data(), t64v.size()))) return cleanup(rc);
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
  (void)hipFree(p->all_angles);
  (void)hipFree(p->taps_r32);
  (void)hipFree(p->taps_z32);
  (void)hipFree(p->taps_r64);
  (void)hipFree(p->taps_z64);
  (void)hipFree(p->tier32);
  (void)hipFree(p->tier64);
  (void)hipFree(p->act);
  (void)hipFree(p->ranges);
  delete p;
  return SPECT_OK;
}

int spect_plan_dims(const spect_plan *p, int32_t *padded, int32_t *pad, int32_t *psf_half,
                    int32_t *chunks) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (padded) *padded = p->P;
  if (pad) *pad = p->pad;
  if (psf_half) *psf_half = p->H;
  if (chunks) *chunks = p->C;
  return SPECT_OK;
}

size_t spect_workspace_bytes(const spect_plan *p, int dtype, int n) {
  if (!p || n <= 0) return 0;
  return ws_layout(p, dtype, n).total;
}

int spect_schedule_info(const spect_plan *p, int dtype, int n, int32_t *forward, int32_t *field) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (!forward || !field) return fail(SPECT_E_ARG, "null output argument");
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  if (n <= 0 || n > p->A) return fail(SPECT_E_ARG, "angle count out of range");
  // the same choices as forward_core and backward_impl
  *forward = fwd_part_layout(p, dtype, n).sched_blocks > 0 ? 1 : 0;
  const int per = dtype == SPECT_F64 ? field_sched_blocks<double>(geo<double>(p))
                                     : field_sched_blocks<float>(geo<float>(p));
  const int fb = p->rounds >= 0 ? std::max(1, p->rounds) * per : 0;
  *field = sched_auto_ok(p, n, fb) ? 1 : 0;
  return SPECT_OK;
}

size_t spect_att_table_bytes(const spect_plan *p) {
  if (!p) return 0;
  const Geo<float> g = geo<float>(p);
  return (size_t)p->A * p->C * (g.je1 - g.je0) * p->Nz * sizeof(double);
}

int spect_att_table_build(const spect_plan *p, int dtype, const void *mu, void *table,
                          void *stream) {
  if (!p || !mu || !table) return fail(SPECT_E_ARG, "null argument");
  if (int rc = check_arrays({mu, table})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32) {
    HIP_TRY(launch_att_table<float>(geo<float>(p), tables<float>(p), (const float *)mu,
                                    p->all_angles, p->A, (double *)table, st));
  } else if (dtype == SPECT_F64) {
    if (!p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
    HIP_TRY(launch_att_table<double>(geo<double>(p), tables<double>(p), (const double *)mu,
                                     p->all_angles, p->A, (double *)table, st));
  } else {
    return fail(SPECT_E_ARG, "bad dtype");
  }
  return SPECT_OK;
}

size_t spect_att_field_bytes(const spect_plan *p, int dtype) {
  if (!p || (dtype != SPECT_F32 && dtype != SPECT_F64)) return 0;
  const Geo<float> g = geo<float>(p);
  return (size_t)p->A * p->P * (g.je1 - g.je0) * p->Nz * (dtype == SPECT_F64 ? 8 : 4);
}

int spect_att_field_build(const spect_plan *p, int dtype, const void *mu, void *field,
                          void *stream) {
  if (!p || !mu || !field) return fail(SPECT_E_ARG, "null argument");
  if (int rc = check_arrays({mu, field})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32) {
    HIP_TRY(launch_att_field<float>(geo<float>(p), tables<float>(p), (const float *)mu,
                                    p->all_angles, p->A, (float *)field, st));
  } else if (dtype == SPECT_F64) {
    if (!p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
    HIP_TRY(launch_att_field<double>(geo<double>(p), tables<double>(p), (const double *)mu,
                                     p->all_angles, p->A, (double *)field, st));
  } else {
    return fail(SPECT_E_ARG, "bad dtype");
  }
  return SPECT_OK;
}

int spect_forward(const spect_plan *p, int dtype, const void *f, const void *mu,
                  const void *att_table, const int32_t *angle_idx, int n, void *proj,
                  void *workspace, size_t workspace_bytes, void *stream) {
  const bool field = (dtype & SPECT_ATT_FIELD) != 0;
  dtype &= ~SPECT_ATT_FIELD;
  if (int rc = check_call(p, dtype, n, workspace, workspace_bytes)) return rc;
  if (!f || !proj || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  if (int rc = check_arrays({f, mu, att_table, proj})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  if (dtype == SPECT_F32)
    return forward_impl<float>(p, (const float *)f, (const float *)mu, att_table, field, angle_idx,
                               n, (float *)proj, nullptr, nullptr, nullptr, nullptr, 0.f, ws, st);
  return forward_impl<double>(p, (const double *)f, (const double *)mu, att_table, field, angle_idx,
                              n, (double *)proj, nullptr, nullptr, nullptr, nullptr, 0.0, ws, st);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free forward projection
// ---------------------------------------------------------------------------
// The scratch of spect_forward_noalloc: every GPU the library runs on holds
// its own copy of this array until the process ends.  In host code the name
// is only a stand-in for the device symbol; the device address comes from
// hipGetSymbolAddress and is what the launches receive.
static __device__ __align__(256) unsigned char scratch[SPECT_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of the scratch array on `device`, which must be current;
// looked up on the first call per device and kept.
int scratch_address(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices) return fail(SPECT_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(SPECT_E_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[device] = ptr;
  }
  *out = reinterpret_cast<unsigned char *>(g_scratch_ptr[device]);
  return SPECT_OK;
}

}  // namespace

extern "C" {

int spect_forward_noalloc(const spect_plan *p, int dtype, const void *f, const void *mu,
                          const void *att_table, const int32_t *angle_idx, int n, void *proj,
                          void *stream) {
  const bool field = (dtype & SPECT_ATT_FIELD) != 0;
  dtype &= ~SPECT_ATT_FIELD;
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  if (dtype == SPECT_F64 && !p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
  if (n <= 0 || n > p->A) return fail(SPECT_E_ARG, "angle count out of range");
  if (!f || !proj || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  if (field && !att_table) return fail(SPECT_E_ARG, "SPECT_ATT_FIELD needs the attenuation field");
  if (int rc = check_arrays({f, mu, att_table, proj})) return rc;
  // scratch: FP64 attenuation rows (only when mu comes without a table), then
  // the chunk partials; every entry is written before it is read
  const size_t rs = dtype == SPECT_F64 ? 8 : 4;
  const Geo<float> g = geo<float>(p);
  const size_t n_jext = g.je1 - g.je0;
  const size_t rows_b = (mu && !att_table)
                            ? align256((size_t)n * p->C * n_jext * p->Nz * sizeof(double)) : 0;
  const size_t part_b = fwd_part_layout(p, dtype, n).total;
  if (rows_b + part_b > SPECT_NOALLOC_SCRATCH_BYTES)
    return fail(SPECT_E_WORKSPACE,
                "spect_forward_noalloc: " + std::to_string(n) + " angles need " +
                    std::to_string(rows_b + part_b) + " bytes of scratch, above the limit of " +
                    std::to_string((size_t)SPECT_NOALLOC_SCRATCH_BYTES) +
                    " (SPECT_NOALLOC_SCRATCH_BYTES): project fewer angles per call");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != p->device)
    return fail(SPECT_E_ARG, "the current device is not the device the plan was created on");
  unsigned char *base = nullptr;
  if (int rc = scratch_address(dev, &base)) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  double *rows = reinterpret_cast<double *>(base);
  if (dtype == SPECT_F32)
    return forward_core<float>(p, (const float *)f, (const float *)mu, att_table, field, angle_idx,
                               n, (float *)proj, nullptr, nullptr, nullptr, nullptr, 0.f, rows,
                               reinterpret_cast<float *>(base + rows_b), st, false);
  return forward_core<double>(p, (const double *)f, (const double *)mu, att_table, field, angle_idx,
                              n, (double *)proj, nullptr, nullptr, nullptr, nullptr, 0.0, rows,
                              reinterpret_cast<double *>(base + rows_b), st, false);
}

int spect_backward_transpose(const spect_plan *p, int dtype, const void *v, const void *mu,
                             const void *att_table, const int32_t *angle_idx, int n, void *proj,
                             void *workspace, size_t workspace_bytes, void *stream) {
  const bool field = (dtype & SPECT_ATT_FIELD) != 0;
  dtype &= ~SPECT_ATT_FIELD;
  if (int rc = check_call(p, dtype, n, workspace, workspace_bytes)) return rc;
  if (!v || !proj || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  if (int rc = check_arrays({v, mu, att_table, proj})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  if (dtype == SPECT_F32)
    return forward_impl<float>(p, (const float *)v, (const float *)mu, att_table, field, angle_idx,
                               n, (float *)proj, nullptr, nullptr, nullptr, nullptr, 0.f, ws, st,
                               true);
  return forward_impl<double>(p, (const double *)v, (const double *)mu, att_table, field, angle_idx,
                              n, (double *)proj, nullptr, nullptr, nullptr, nullptr, 0.0, ws, st,
                              true);
}

int spect_backward(const spect_plan *p, int dtype, int mode, int out_mode, const void *proj,
                   const void *mu, const void *att_table, const int32_t *angle_idx, int n,
                   void *out, const void *f, const void *norm, double delta, double floor_value,
                   void *workspace, size_t workspace_bytes, void *stream) {
  const bool field = (dtype & SPECT_ATT_FIELD) != 0;
  dtype &= ~SPECT_ATT_FIELD;
  if (int rc = check_call(p, dtype, n, workspace, workspace_bytes)) return rc;
  if (!proj || !out || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  if (mode != SPECT_BP_REFERENCE && mode != SPECT_BP_ADJOINT) return fail(SPECT_E_ARG, "bad mode");
  if (out_mode < SPECT_OUT_STORE || out_mode > SPECT_OUT_OSEM) return fail(SPECT_E_ARG, "bad out_mode");
  if (out_mode == SPECT_OUT_OSEM && (!f || !norm)) return fail(SPECT_E_ARG, "OSEM update needs f and norm");
  if (int rc = check_arrays({proj, mu, att_table, out, f, norm})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  if (dtype == SPECT_F32)
    return backward_impl<float>(p, mode, out_mode, (const float *)proj, (const float *)mu,
                                att_table, field, angle_idx, n, (float *)out, (const float *)f,
                                (const float *)norm, delta, floor_value, ws, st);
  return backward_impl<double>(p, mode, out_mode, (const double *)proj, (const double *)mu,
                               att_table, field, angle_idx, n, (double *)out, (const double *)f,
                               (const double *)norm, delta, floor_value, ws, st);
}

int spect_osem_subiteration(const spect_plan *p, int dtype, void *f, const void *mu,
                            const void *att_table, const int32_t *angle_idx, int n, const void *g,
                            const void *s, const void *norm, void *pred, double delta,
                            double floor_value, void *workspace, size_t workspace_bytes,
                            void *stream) {
  const bool field = (dtype & SPECT_ATT_FIELD) != 0;
  dtype &= ~SPECT_ATT_FIELD;
  if (int rc = check_call(p, dtype, n, workspace, workspace_bytes)) return rc;
  if (!f || !g || !norm || !angle_idx) return fail(SPECT_E_ARG, "null array argument");
  if (int rc = check_arrays({f, mu, att_table, g, s, norm, pred})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  const Ws l = ws_layout(p, dtype, n);
  if (dtype == SPECT_F32) {
    float *ratio = reinterpret_cast<float *>(ws + l.ratio_off);
    int rc = forward_impl<float>(p, (const float *)f, (const float *)mu, att_table, field,
                                 angle_idx, n, nullptr, (const float *)g, (const float *)s,
                                 (float *)pred, ratio, (float)delta, ws, st);
    if (rc) return rc;
    return backward_impl<float>(p, SPECT_BP_REFERENCE, SPECT_OUT_OSEM, ratio, (const float *)mu,
                                att_table, field, angle_idx, n, (float *)f, (const float *)f,
                                (const float *)norm, delta, floor_value, ws, st);
  }
  double *ratio = reinterpret_cast<double *>(ws + l.ratio_off);
  int rc = forward_impl<double>(p, (const double *)f, (const double *)mu, att_table, field,
                                angle_idx, n, nullptr, (const double *)g, (const double *)s,
                                (double *)pred, ratio, delta, ws, st);
  if (rc) return rc;
  return backward_impl<double>(p, SPECT_BP_REFERENCE, SPECT_OUT_OSEM, ratio, (const double *)mu,
                               att_table, field, angle_idx, n, (double *)f, (const double *)f,
                               (const double *)norm, delta, floor_value, ws, st);
}

size_t spect_grad_workspace_bytes(const spect_plan *p, int dtype, int n) {
  if (!p || n <= 0) return 0;
  return grad_ws_layout(p, dtype, n).total;
}

int spect_forward_vjp(const spect_plan *p, int dtype, const void *u, const void *f, const void *mu,
                      const int32_t *angle_idx, int n, void *grad_f, void *grad_mu,
                      void *workspace, size_t workspace_bytes, void *stream) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  if (dtype == SPECT_F64 && !p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
  if (n <= 0 || n > p->A) return fail(SPECT_E_ARG, "angle count out of range");
  if (!u || !angle_idx || !grad_f) return fail(SPECT_E_ARG, "null array argument");
  if (grad_mu && (!mu || !f)) return fail(SPECT_E_ARG, "the attenuation gradient needs mu and f");
  const size_t need = grad_ws_layout(p, dtype, n).total;
  if (!workspace || workspace_bytes < need)
    return fail(SPECT_E_WORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
  if (int rc = check_arrays({u, f, mu, grad_f, grad_mu, workspace})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  auto *ws = reinterpret_cast<unsigned char *>(workspace);
  if (dtype == SPECT_F32)
    return grad_impl<float>(p, (const float *)u, (const float *)f, (const float *)mu, angle_idx, n,
                            (float *)grad_f, (float *)grad_mu, ws, st);
  return grad_impl<double>(p, (const double *)u, (const double *)f, (const double *)mu, angle_idx, n,
                           (double *)grad_f, (double *)grad_mu, ws, st);
}

int spect_poisson_ratio(int dtype, const void *g, const void *hf, const void *s, void *pred,
                        void *ratio, int64_t count, double delta, void *stream) {
  if (!g || !hf || !ratio || count < 0) return fail(SPECT_E_ARG, "bad ratio arguments");
  if (count == 0) return SPECT_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32)
    HIP_TRY(launch_ratio<float>((const float *)g, (const float *)hf, (const float *)s,
                                (float *)pred, (float *)ratio, count, (float)delta, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_ratio<double>((const double *)g, (const double *)hf, (const double *)s,
                                 (double *)pred, (double *)ratio, count, delta, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_rdp_gradient(int dtype, const void *f, void *out, int32_t nx, int32_t ny, int32_t nz,
                       double dx, double dy, double dz, double beta, double gamma, double delta,
                       double scale, void *stream) {
  if (!f || !out || nx <= 0 || ny <= 0 || nz <= 0) return fail(SPECT_E_ARG, "bad prior arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double bs = beta * scale;
  if (dtype == SPECT_F32)
    HIP_TRY(launch_rdp<float>((const float *)f, (float *)out, nx, ny, nz, dx, dy, dz, (float)bs,
                              (float)gamma, (float)delta, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_rdp<double>((const double *)f, (double *)out, nx, ny, nz, dx, dy, dz, bs, gamma,
                               delta, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_rdp_hessian_vjp(int dtype, const void *f, const void *u, void *out, int32_t nx,
                          int32_t ny, int32_t nz, double dx, double dy, double dz, double beta,
                          double gamma, double delta, double scale, void *stream) {
  if (!f || !u || !out || nx <= 0 || ny <= 0 || nz <= 0) return fail(SPECT_E_ARG, "bad prior arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double bs = beta * scale;
  if (dtype == SPECT_F32)
    HIP_TRY(launch_rdp_hvp<float>((const float *)f, (const float *)u, (float *)out, nx, ny, nz, dx,
                                  dy, dz, (float)bs, (float)gamma, (float)delta, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_rdp_hvp<double>((const double *)f, (const double *)u, (double *)out, nx, ny, nz,
                                   dx, dy, dz, bs, gamma, delta, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_rdp_hessian_jvp(int dtype, const void *f, const void *u, void *out, int32_t nx,
                          int32_t ny, int32_t nz, double dx, double dy, double dz, double beta,
                          double gamma, double delta, double scale, void *stream) {
  if (!f || !u || !out || nx <= 0 || ny <= 0 || nz <= 0) return fail(SPECT_E_ARG, "bad prior arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const double bs = beta * scale;
  if (dtype == SPECT_F32)
    HIP_TRY(launch_rdp_jvp<float>((const float *)f, (const float *)u, (float *)out, nx, ny, nz, dx,
                                  dy, dz, (float)bs, (float)gamma, (float)delta, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_rdp_jvp<double>((const double *)f, (const double *)u, (double *)out, nx, ny, nz,
                                   dx, dy, dz, bs, gamma, delta, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_bsrem_update(int dtype, void *f, const void *b, const void *norm, const void *norm_sum,
                       const void *prior, int64_t count, double relax, double weight, double delta,
                       double floor_value, void *stream) {
  if (!f || !b || !norm || !norm_sum || count < 0) return fail(SPECT_E_ARG, "bad update arguments");
  if (count == 0) return SPECT_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32)
    HIP_TRY(launch_bsrem_update<float>((float *)f, (const float *)b, (const float *)norm,
                                       (const float *)norm_sum, (const float *)prior, count,
                                       (float)relax, (float)weight, (float)delta,
                                       (float)floor_value, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_bsrem_update<double>((double *)f, (const double *)b, (const double *)norm,
                                        (const double *)norm_sum, (const double *)prior, count,
                                        relax, weight, delta, floor_value, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_ew_scatter(int dtype, const void *lower, const void *upper, void *estimate,
                     void *variance_diag, int64_t count, double width_lower, double width_upper,
                     double width_peak, double weighting_lower, double weighting_upper,
                     void *stream) {
  if (!lower || !estimate || count < 0) return fail(SPECT_E_ARG, "bad scatter arguments");
  if (width_lower == 0.0 || (upper && width_upper == 0.0))
    return fail(SPECT_E_ARG, "energy window widths must be non-zero");
  if (count == 0) return SPECT_OK;
  if (count > (int64_t)256 * 0x7fffffff) return fail(SPECT_E_ARG, "array too large");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // the variance coefficients are Python float (double) expressions (scatter.py:86)
  const double a1 = width_peak / width_lower * weighting_lower;
  const double a2 = upper ? width_peak / width_upper * weighting_upper : 0.0;
  const double c1 = a1 * a1, c2 = a2 * a2;
  if (dtype == SPECT_F32)
    HIP_TRY(launch_ew_scatter<float>((const float *)lower, (const float *)upper, (float *)estimate,
                                     (float *)variance_diag, count, (float)width_lower,
                                     (float)width_upper, (float)width_peak, (float)weighting_lower,
                                     (float)weighting_upper, (float)c1, (float)c2, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_ew_scatter<double>((const double *)lower, (const double *)upper,
                                      (double *)estimate, (double *)variance_diag, count,
                                      width_lower, width_upper, width_peak, weighting_lower,
                                      weighting_upper, c1, c2, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_smooth_axis(int dtype, const void *in, void *out, int32_t n0, int32_t n1, int32_t n2,
                      int32_t axis, const void *taps, int32_t K, int32_t mode, void *stream) {
  if (!in || !out || !taps || n0 <= 0 || n1 <= 0 || n2 <= 0)
    return fail(SPECT_E_ARG, "bad smoothing arguments");
  if (in == out) return fail(SPECT_E_ARG, "smoothing input and output must differ");
  if (axis < 0 || axis > 2) return fail(SPECT_E_ARG, "axis must be 0, 1 or 2");
  if (K < 1 || (K % 2) == 0) return fail(SPECT_E_ARG, "kernel size must be odd and positive");
  if (mode != SPECT_PAD_CIRCULAR && mode != SPECT_PAD_REPLICATE && mode != SPECT_PAD_ZEROS)
    return fail(SPECT_E_ARG, "bad padding mode");
  if ((int64_t)n0 * n1 * n2 > (int64_t)256 * 0x7fffffff)
    return fail(SPECT_E_ARG, "array too large");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32)
    HIP_TRY(launch_smooth_axis<float>((const float *)in, (float *)out, n0, n1, n2, axis,
                                      (const float *)taps, K, mode, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_smooth_axis<double>((const double *)in, (double *)out, n0, n1, n2, axis,
                                       (const double *)taps, K, mode, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

int spect_kem_apply(int dtype, const void *kernel, const void *f, void *out, int32_t nx,
                    int32_t ny, int32_t nz, int32_t size, void *stream) {
  if (!kernel || !f || !out || nx <= 0 || ny <= 0 || nz <= 0)
    return fail(SPECT_E_ARG, "bad KEM arguments");
  if (size < 1 || (size % 2) == 0) return fail(SPECT_E_ARG, "KEM kernel size must be odd");
  if (f == out) return fail(SPECT_E_ARG, "KEM input and output must differ");
  if ((int64_t)nx * ny * nz > (int64_t)256 * 0x7fffffff) return fail(SPECT_E_ARG, "object too large");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const int m = (size - 1) / 2;
  if (dtype == SPECT_F32)
    HIP_TRY(launch_kem<float>((const float *)kernel, (const float *)f, (float *)out, nx, ny, nz, m, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_kem<double>((const double *)kernel, (const double *)f, (double *)out, nx, ny, nz,
                               m, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

}  // extern "C"

namespace {

// The plan's geometry over the full frame width: every rotated-frame column
// [0, P), not only those the fused projectors visit.
template <typename R> Geo<R> geo_full(const spect_plan *p) {
  Geo<R> g = geo<R>(p);
  g.je0 = 0;
  g.je1 = p->P;
  g.ntiles = (p->P + kTJ - 1) / kTJ;
  return g;
}

int check_volume_call(const spect_plan *p, int dtype, const int32_t *angle_idx, int n,
                      std::initializer_list<const void *> arrays) {
  if (!p) return fail(SPECT_E_ARG, "null plan");
  if (dtype != SPECT_F32 && dtype != SPECT_F64) return fail(SPECT_E_ARG, "dtype must be F32 or F64");
  if (dtype == SPECT_F64 && !p->has64) return fail(SPECT_E_ARG, "plan has no float64 tables");
  if (!angle_idx) return fail(SPECT_E_ARG, "null angle_idx");
  if (n <= 0 || n > 65535) return fail(SPECT_E_ARG, "angle count must be 1..65535");
  for (const void *a : arrays)
    if (!a) return fail(SPECT_E_ARG, "null array argument");
  return check_arrays(arrays);
}

}  // namespace

extern "C" {

int spect_rotate_object(const spect_plan *p, int dtype, const void *f, const int32_t *angle_idx,
                        int n, void *out, void *stream) {
  if (int rc = check_volume_call(p, dtype, angle_idx, n, {f, out})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32)
    HIP_TRY(launch_rotate<float>(geo_full<float>(p), tables<float>(p), (const float *)f, angle_idx, n,
                                 (float *)out, st));
  else
    HIP_TRY(launch_rotate<double>(geo_full<double>(p), tables<double>(p), (const double *)f, angle_idx,
                                  n, (double *)out, st));
  return SPECT_OK;
}

int spect_rotate_back(const spect_plan *p, int dtype, const void *vol, const int32_t *angle_idx, int n,
                      int out_mode, void *out, void *stream) {
  if (int rc = check_volume_call(p, dtype, angle_idx, n, {vol, out})) return rc;
  if (out_mode != SPECT_OUT_STORE && out_mode != SPECT_OUT_ADD)
    return fail(SPECT_E_ARG, "out_mode must be SPECT_OUT_STORE or SPECT_OUT_ADD");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // the reference back projection's gather (mode 0) over full-width frames
  if (dtype == SPECT_F32)
    HIP_TRY(launch_bp_gather<float>(geo_full<float>(p), tables<float>(p), (const float *)vol, angle_idx,
                                    n, 0, out_mode, (float *)out, nullptr, nullptr, 0.0f, 0.0f,
                                    nullptr, nullptr, nullptr, st));
  else
    HIP_TRY(launch_bp_gather<double>(geo_full<double>(p), tables<double>(p), (const double *)vol,
                                     angle_idx, n, 0, out_mode, (double *)out, nullptr, nullptr, 0.0,
                                     0.0, nullptr, nullptr, nullptr, st));
  return SPECT_OK;
}

int spect_att_volume(const spect_plan *p, int dtype, const void *mu, const int32_t *angle_idx, int n,
                     void *out, void *stream) {
  if (int rc = check_volume_call(p, dtype, angle_idx, n, {mu, out})) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  // the attenuation-field body over full-width frames
  if (dtype == SPECT_F32)
    HIP_TRY(launch_att_field<float>(geo_full<float>(p), tables<float>(p), (const float *)mu, angle_idx,
                                    n, (float *)out, st));
  else
    HIP_TRY(launch_att_field<double>(geo_full<double>(p), tables<double>(p), (const double *)mu,
                                     angle_idx, n, (double *)out, st));
  return SPECT_OK;
}

int spect_psf_volume(const spect_plan *p, int dtype, const void *vol, const int32_t *angle_idx, int n,
                     void *out, void *stream) {
  if (int rc = check_volume_call(p, dtype, angle_idx, n, {vol, out})) return rc;
  if (p->H <= 0) return fail(SPECT_E_ARG, "the plan has no PSF");
  if (vol == out) return fail(SPECT_E_ARG, "vol and out must differ");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32)
    HIP_TRY(launch_psf_volume<float>(geo_full<float>(p), tables<float>(p), (const float *)vol, angle_idx,
                                     n, (float *)out, st));
  else
    HIP_TRY(launch_psf_volume<double>(geo_full<double>(p), tables<double>(p), (const double *)vol,
                                      angle_idx, n, (double *)out, st));
  return SPECT_OK;
}

int spect_dvf_warp(int dtype, const void *f, const void *dvf, void *out, int32_t nx, int32_t ny,
                   int32_t nz, void *stream) {
  if (!f || !dvf || !out) return fail(SPECT_E_ARG, "bad DVF arguments");
  if (nx < 2 || ny < 2 || nz < 2) return fail(SPECT_E_ARG, "DVF warp needs every side >= 2");
  if (f == out) return fail(SPECT_E_ARG, "DVF input and output must differ");
  if ((int64_t)nx * ny * nz > (int64_t)256 * 0x7fffffff) return fail(SPECT_E_ARG, "object too large");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (dtype == SPECT_F32)
    HIP_TRY(launch_dvf<float>((const float *)f, (const float *)dvf, (float *)out, nx, ny, nz, st));
  else if (dtype == SPECT_F64)
    HIP_TRY(launch_dvf<double>((const double *)f, (const double *)dvf, (double *)out, nx, ny, nz, st));
  else
    return fail(SPECT_E_ARG, "bad dtype");
  return SPECT_OK;
}

}  // extern "C"
