// forward_body.hpp -- block bodies of the forward projector (fwd_body) and of
// its chunk reduce (reduce_body), shared by the single-problem and batched
// kernels of forward.hip.  They are plain device functions of the launch
// descriptors and the block / thread indices, so a host build that runs one
// std::thread per lane (tests/cpu_build) executes the same source.
#pragma once

#include <type_traits>

#include "launch.hpp"

namespace spect {

// Attenuation march of one (angle al, 32-column tile jt) block, grid (n, ntiles).
template <typename R, int RZ>
__device__ __forceinline__ void att_table_body(const Geo<R> &g, const Tables<R> &tb,
                                               const R *__restrict__ mu,
                                               const int32_t *__restrict__ angle_idx,
                                               double *__restrict__ table, int al, int jt) {
  __shared__ Src<R> src[kTJ];
  const int ang = ld_uniform(angle_idx + al);
  if ((unsigned)ang >= (unsigned)g.A) return;
  const int tid = threadIdx.x;
  const int j0 = g.je0 + jt * kTJ;
  const int tj = min(kTJ, g.je1 - j0);
  const int runs = (g.Nz + RZ - 1) / RZ;
  const int jl = tid % kTJ, run = tid / kTJ, z0 = run * RZ;
  const bool mine = run < runs && jl < tj;
  const int n_jext = g.je1 - g.je0;
  const auto cs = ld_uniform(tb.trig + ang);
  double S[RZ];
#pragma unroll
  for (int m = 0; m < RZ; ++m) S[m] = 0.0;
  auto row = [&](int c, int m) -> size_t {
    return (((size_t)al * g.C + c) * n_jext + (j0 - g.je0 + jl)) * g.Nz + z0 + m;
  };
  if (mine) {
#pragma unroll
    for (int m = 0; m < RZ; ++m)
      if (z0 + m < g.Nz) table[row(g.C - 1, m)] = 0.0;
  }
  for (int i = g.P - 1; i >= 0; --i) {
    int inside = 0;
    if (tid < tj) {
      R ix, iy;
      rot_coord<R>(cs.x, cs.y, tb.lin[j0 + tid], ld_uniform(tb.lin + i), g.half, ix, iy);
      src[tid] = make_src<R>(ix, iy, g.pad);
      inside = in_footprint<R>(ix, iy, g);
    }
    if (__syncthreads_or(inside)) {
      if (mine) {
        const Taps4<R> t = taps4<R>(src[jl], g.N, g.Nz);
        R x[RZ];
        gather_run<R, RZ>(mu, t, z0, g.Nz, x);
#pragma unroll
        for (int m = 0; m < RZ; ++m) S[m] += (double)(x[m] * g.dx);
      }
      __syncthreads();  // src is rewritten by the next plane
    }
    if (mine && i > 0 && i % g.L == 0) {
      const int c = i / g.L - 1;
#pragma unroll
      for (int m = 0; m < RZ; ++m)
        if (z0 + m < g.Nz) table[row(c, m)] = S[m];
    }
  }
}

// The attenuation factor itself, a(i, j, z) = exp(-(S(i) - x(i)/2)) in R, for
// every plane whose tile footprint is non-empty (the planes the projection
// kernels visit), from the same FP64 march as att_table_kernel: the forward
// and field kernels then read it instead of gathering mu (bitwise the same
// factors).  field layout: [angle][P][je1-je0][Nz].
template <typename R, int RZ>
__device__ __forceinline__ void att_field_body(const Geo<R> &g, const Tables<R> &tb,
                                               const R *__restrict__ mu,
                                               const int32_t *__restrict__ angle_idx,
                                               R *__restrict__ field, int al, int jt) {
  __shared__ Src<R> src[kTJ];
  const int ang = ld_uniform(angle_idx + al);
  if ((unsigned)ang >= (unsigned)g.A) return;
  const int tid = threadIdx.x;
  const int j0 = g.je0 + jt * kTJ;
  const int tj = min(kTJ, g.je1 - j0);
  const int runs = (g.Nz + RZ - 1) / RZ;
  const int jl = tid % kTJ, run = tid / kTJ, z0 = run * RZ;
  const bool mine = run < runs && jl < tj;
  const int n_jext = g.je1 - g.je0;
  const auto cs = ld_uniform(tb.trig + ang);
  double S[RZ];
#pragma unroll
  for (int m = 0; m < RZ; ++m) S[m] = 0.0;
  for (int i = g.P - 1; i >= 0; --i) {
    int inside = 0;
    if (tid < tj) {
      R ix, iy;
      rot_coord<R>(cs.x, cs.y, tb.lin[j0 + tid], ld_uniform(tb.lin + i), g.half, ix, iy);
      src[tid] = make_src<R>(ix, iy, g.pad);
      inside = in_footprint<R>(ix, iy, g);
    }
    // Every plane gets its factors, also where the tile samples no object
    // (mu is zero there, so S carries over): the projectors may visit a
    // plane a conservative activity test keeps.
    const bool any = __syncthreads_or(inside);
    if (mine) {
      R x[RZ];
      if (any) {
        const Taps4<R> t = taps4<R>(src[jl], g.N, g.Nz);
        gather_run<R, RZ>(mu, t, z0, g.Nz, x);
      } else {
#pragma unroll
        for (int m = 0; m < RZ; ++m) x[m] = R(0);
      }
      R *dst = field + (((size_t)al * g.P + i) * n_jext + (j0 - g.je0 + jl)) * g.Nz + z0;
#pragma unroll
      for (int m = 0; m < RZ; ++m) {
        const R xm = x[m] * g.dx;
        S[m] += (double)xm;
        if (z0 + m < g.Nz) dst[m] = att_factor<R>(S[m], xm);
      }
    }
    __syncthreads();  // src is rewritten by the next plane
  }
}

// LDS rows: H zeros | runs*RZ values | H zeros | 1 pad (odd stride).
template <int RZ, int H> __device__ __forceinline__ int lds_stride(int runs) {
  return runs * RZ + 2 * H + 1;
}

// Dynamic LDS of fwd_body: V and Vz ([kTJ][stride] each) and the taps of the
// piece's planes ([planes][2][2H+1], planes = the larger of the chunk length
// and kPieceMax).
template <typename R, int RZ, int H> inline size_t fwd_smem_bytes(int Nz, int L) {
  const int runs = (Nz + RZ - 1) / RZ;
  const int planes = L > kPieceMax ? L : kPieceMax;
  return (2 * (size_t)kTJ * (runs * RZ + 2 * H + 1) + (size_t)planes * 2 * (2 * H + 1)) * sizeof(R);
}

// TRB: instead of the forward bilinear gather of f, the transpose of the
// reference back projection's gather at theta_b = -theta_f
// (spect_system_matrix.py:196): rotated-frame pixel (i, j) collects every
// object pixel whose theta_b sample has (i, j) among its bilinear taps, with
// that tap's weight -- the body then computes B^T v (attenuation, PSF, sum).
// Contributors per rotated-frame pixel: the search below visits a 4 x 4 block
// of object pixels, so a list of 16 can never overflow (<= 9 occur in exact
// arithmetic: lattice points of a side-2 square).
constexpr int kTrbSearch = 4;
constexpr int kTrbMax = kTrbSearch * kTrbSearch;

// The planes one call of fwd_body covers and where its result goes.
struct FwdPiece {
  int i_lo, i_hi;  // planes [i_lo, i_hi) (at most 64), walked from i_hi - 1 down
  int c;           // chunk of the FP64 suffix-table rows the march starts from
  int ranged;      // 1: every plane of the range is visited; 0: the chunk's
                   // activity mask (table or in-kernel test) selects them
  size_t slot;     // partial + slot * (kTJ + 2H) * Nz receives the tile's sums
};

// Chunk c of the uniform schedule: planes [c L, c L + L), partial slot
// (c, al, jt) of the [C][n][ntiles] layout.
template <typename R>
__device__ __forceinline__ FwdPiece chunk_piece(const Geo<R> &g, int n, int al, int jt, int c) {
  FwdPiece pc;
  pc.i_lo = c * g.L;
  pc.i_hi = min(g.P, pc.i_lo + g.L);
  pc.c = c;
  pc.ranged = 0;
  pc.slot = ((size_t)c * n + al) * g.ntiles + jt;
  return pc;
}

// MARCH: mu may be given without the attenuation field and is marched here
// (FP64 suffix sums per lane); a scheduled piece never does.
template <typename R, int RZ, int QS, int H, bool TRB, bool AF, bool MARCH = !AF>
__device__ __forceinline__ void fwd_body(const FwdJob<R> &J, int al, int jt, const FwdPiece &pc,
                                         unsigned char *smem) {
  constexpr int K = 2 * H + 1;
  constexpr int JO = kTJ + 2 * H;      // output columns touched by the tile
  constexpr int JH = (JO + QS - 1) / QS;  // per split
  // bilinear taps per tile column, double buffered: the next plane's are
  // built while the current plane is gathered
  __shared__ Taps4<R> tp[2][kTJ];
  __shared__ uint32_t trb_off[TRB ? 2 : 1][TRB ? kTJ : 1][kTrbMax];
  __shared__ R trb_w[TRB ? 2 : 1][TRB ? kTJ : 1][kTrbMax];
  __shared__ int trb_n[TRB ? 2 : 1][TRB ? kTJ : 1];
  __shared__ unsigned long long act_s;  // planes of the chunk with a non-empty footprint
  const Geo<R> &g = J.g;
  const Tables<R> &tb = J.tb;
  const R *__restrict__ f = J.f;
  const R *__restrict__ mu = J.mu;
  const double *__restrict__ att = J.att;
  const int att_global = J.att_global;
  const R *__restrict__ attf = J.attf;  // AF: precomputed attenuation factors
  const int32_t *__restrict__ angle_idx = J.angle_idx;
  R *__restrict__ partial = J.partial;
  const int c = pc.c;
  const int ang = ld_uniform(angle_idx + al);
  if ((unsigned)ang >= (unsigned)g.A) return;
  const int tid = threadIdx.x;
  const int Nz = g.Nz;
  const int runs = (Nz + RZ - 1) / RZ;
  const int ldv = lds_stride<RZ, H>(runs);
  R *V = reinterpret_cast<R *>(smem);         // [kTJ][ldv], data at +H
  R *Vz = V + (size_t)kTJ * ldv;              // [kTJ][ldv], data at +0
  const int j0 = g.je0 + jt * kTJ;
  const int tj = min(kTJ, g.je1 - j0);
  const int n_jext = g.je1 - g.je0;
  // phase B ownership: column jl, z run [z0, z0 + RZ)
  const int jl = tid % kTJ, run = tid / kTJ, z0 = run * RZ;
  const bool ab = run < runs;
  // phase A ownership: ZV consecutive z (one 16-byte load) per lane with the
  // lanes along z, so every wave load covers whole 64-byte sectors; the
  // thread covers columns aco + p * cpp, p < NPA (cpp * NPA >= kTJ is
  // checked at launch)
  constexpr int ZV = 16 / sizeof(R);
  constexpr int NPA = phase_a_passes<R, QS>();
  const int nq = (Nz + ZV - 1) / ZV;
  const int cpp = blockDim.x / nq;
  const int aco = tid / nq, za = (tid - aco * nq) * ZV;
  const bool aact = aco < cpp;
  const bool zvec = (Nz % ZV) == 0;
  // phase C ownership: one z, 1/QS of the output columns (wave-uniform split)
  const int nzc = (Nz + 63) / 64 * 64;
  const int half = tid / nzc, zc = tid - half * nzc;
  const bool cact = half < QS && zc < Nz;
  const auto cs = ld_uniform(tb.trig + ang);
  const bool has_mu = MARCH && mu != nullptr;

  // The LDS entries the phases read but never write start at zero: V's
  // halo pads, V's z slots past Nz, and whole rows of both buffers past the
  // tile's last column (phase A writes columns < tj, z < Nz; phase B writes
  // Vz's data slots of every row)
  {
    const int pads = 2 * H + 1;  // per row: [0, H) and [H + runs*RZ, ldv)
    for (int q = tid; q < kTJ * pads; q += blockDim.x) {
      const int row = q / pads, k = q - row * pads;
      V[(size_t)row * ldv + (k < H ? k : runs * RZ + k)] = R(0);
    }
    const int tailz = runs * RZ - Nz;
    if (tailz > 0)
      for (int q = tid; q < kTJ * tailz; q += blockDim.x) {
        const int row = q / tailz, k = q - row * tailz;
        V[(size_t)row * ldv + H + Nz + k] = R(0);
        if constexpr (H == 0) Vz[(size_t)row * ldv + Nz + k] = R(0);
      }
    if (tj < kTJ)
      for (int q = tid; q < (kTJ - tj) * ldv; q += blockDim.x) {
        V[(size_t)tj * ldv + q] = R(0);
        Vz[(size_t)tj * ldv + q] = R(0);
      }
  }
  double S[MARCH ? NPA : 1][ZV];
#pragma unroll
  for (int p = 0; p < (MARCH ? NPA : 1); ++p) {
    const int jj = aco + p * cpp;
#pragma unroll
    for (int e = 0; e < ZV; ++e) {
      S[p][e] = 0.0;
      if (MARCH && has_mu && aact && jj < tj && za + e < Nz) {
        const int rowa = att_global ? ang : al;
        S[p][e] = att[(((size_t)rowa * g.C + c) * n_jext + (j0 - g.je0 + jj)) * Nz + za + e];
      }
    }
  }
  R acc[JH];
#pragma unroll
  for (int q = 0; q < JH; ++q) acc[q] = R(0);
  // RAY (no PSF, forward gather): the ray sums are accumulated by the
  // gathering lanes themselves, in registers, in the same plane order -- no
  // LDS round trip and one barrier per plane (for the taps' double buffer)
  constexpr bool RAY = H == 0 && !TRB;
  R accA[RAY ? NPA : 1][ZV];
#pragma unroll
  for (int p = 0; p < (RAY ? NPA : 1); ++p)
#pragma unroll
    for (int e = 0; e < ZV; ++e) accA[p][e] = R(0);

  const int rix = ld_uniform(tb.radius_index + ang);
  const int i_lo = pc.i_lo, i_hi = pc.i_hi;
  // the piece's PSF taps (z then y, per plane) staged in LDS: the blur phases
  // read them as LDS broadcasts instead of waiting on scalar loads
  R *Wt = Vz + (size_t)kTJ * ldv;  // [i_hi - i_lo][2K]
  const R lin_j = tid < tj ? tb.lin[j0 + tid] : R(0);
  if constexpr (H > 0) {
    // all of a thread's tap loads are issued before the first LDS store
    constexpr int TB = 8;
    const int nt = (i_hi - i_lo) * 2 * K;
    for (int q0 = tid; q0 < nt; q0 += TB * (int)blockDim.x) {
      R w[TB];
#pragma unroll
      for (int u = 0; u < TB; ++u) {
        const int q = min(q0 + u * (int)blockDim.x, nt - 1);
        const int l = q / (2 * K), r = q - l * (2 * K);
        const size_t row = ((size_t)rix * g.P + i_lo + l) * K;
        w[u] = r < K ? tb.taps_z[row + r] : tb.taps_r[row + r - K];
      }
#pragma unroll
      for (int u = 0; u < TB; ++u)
        if (q0 + u * (int)blockDim.x < nt) Wt[q0 + u * blockDim.x] = w[u];
    }
  }
  // which planes of the chunk can touch the object (bit l: plane i_lo + l);
  // the skipped planes sample outside the object, where f and mu vanish
  if (pc.ranged) {
    // every plane of the range
    if (tid == 0) act_s = (i_hi - i_lo >= 64) ? ~0ull : ((1ull << (i_hi - i_lo)) - 1);
  } else if (tb.act) {
    if (tid == 0) act_s = ld_uniform(tb.act + ((size_t)ang * g.ntiles + jt) * g.C + c);
  } else {
    if (tid == 0) act_s = 0ull;
    __syncthreads();
    for (int q = tid; q < (i_hi - i_lo) * kTJ; q += blockDim.x) {
      const int l = q / kTJ, jj = q - l * kTJ;
      if (jj < tj) {
        R ix, iy;
        rot_coord<R>(cs.x, cs.y, tb.lin[j0 + jj], tb.lin[i_lo + l], g.half, ix, iy);
        if (in_footprint<R>(ix, iy, g)) atomicOr(&act_s, 1ull << l);
      }
    }
  }
  // bilinear taps of plane i into buffer b; run by tid < tj
  auto setup = [&](int i, int b) {
    R ix, iy;
    rot_coord<R>(cs.x, cs.y, lin_j, ld_uniform(tb.lin + i), g.half, ix, iy);
    tp[b][tid] = taps4<R>(make_src<R>(ix, iy, g.pad), g.N, Nz);
  };
  // TRB: the contributor lists of plane i into buffer b, by the whole block
  // (every thread calls, i block-uniform).  Item q = column * 16 + k tests
  // object pixel (Xb + k / 4, Yb + k % 4) near the forward sample position of
  // (i, column); each column's 16 candidates sit in 16 lanes of one wave, and
  // a ballot packs its contributors in candidate order (X, then Y ascending)
  // -- at most one entry per candidate, so cnt <= kTrbMax.
  static_assert(kTrbMax == 16, "one 16-lane group per column");
  auto trb_lists = [&](int i, int b) {
    if constexpr (TRB) {
      const int nitems = tj * kTrbMax;
      const R lin_i = ld_uniform(tb.lin + i);
      const unsigned lane = tid & 63u;
      for (int q0 = 0; q0 < nitems; q0 += blockDim.x) {
        const int q = q0 + tid;
        const int col = q >> 4, k = q & 15;
        bool use = false;
        uint32_t off = 0;
        R w = R(0);
        if (q < nitems) {
          R ix, iy;
          rot_coord<R>(cs.x, cs.y, tb.lin[j0 + col], lin_i, g.half, ix, iy);
          const int X = int(floor(iy - R(1.5))) + (k >> 2), Y = int(floor(ix - R(1.5))) + (k & 3);
          if ((unsigned)(X - g.pad) < (unsigned)g.N && (unsigned)(Y - g.pad) < (unsigned)g.N) {
            R jc, ic;
            rot_coord<R>(cs.x, -cs.y, tb.lin[Y], tb.lin[X], g.half, jc, ic);
            const R fjc = floor(jc), fic = floor(ic);
            const int c0 = int(fjc), r0 = int(fic);
            const R tx = jc - fjc, ty = ic - fic;
            const int j = j0 + col;
            const bool rok = r0 == i || r0 + 1 == i, cok = c0 == j || c0 + 1 == j;
            if (rok && cok) {
              const R wr = r0 == i ? R(1) - ty : ty;
              const R wc = c0 == j ? R(1) - tx : tx;
              use = true;
              off = (uint32_t)((X - g.pad) * g.N + (Y - g.pad)) * (uint32_t)Nz;
              w = wr * wc;
            }
          }
        }
        const unsigned grp = (unsigned)(__ballot(use) >> (lane & 48u)) & 0xFFFFu;
        if (use) {
          const int slot = __popc(grp & ((1u << (lane & 15u)) - 1u));
          trb_off[b][col][slot] = off;
          trb_w[b][col][slot] = w;
        }
        // a contributor's forward sample lies within sqrt(2) px of the object,
        // inside the footprint square, so the plane is in act
        if (q < nitems && k == 0) trb_n[b][col] = __popc(grp);
      }
    }
  };
  __syncthreads();
  unsigned long long rem = act_s;
  int i = rem ? i_lo + 63 - __clzll(rem) : -1;
  if (i >= 0) {
    rem &= ~(1ull << (i - i_lo));
    if (tid < tj) setup(i, 0);
    trb_lists(i, 0);
  }
  __syncthreads();
  // PF: the attenuation factors of the next plane (a coalesced stream from
  // HBM) are loaded into registers right after the phase-A barrier, so their
  // latency hides behind the blur phases of the current plane
  constexpr bool PF = AF;
  R Ar[PF ? NPA : 1][ZV];
  // The 16-byte / element-wise choice of load_zv is made once per phase
  // (VEC is a compile-time constant inside the phase bodies): a per-load
  // runtime branch would make the compiler wait for every outstanding load at
  // each branch merge and serialise the independent gathers.
  // (a block-uniform 64-bit plane base plus 32-bit lane offsets: fewer
  // vector registers than a 64-bit address per pass)
  const uint32_t af_col0 = (uint32_t)((j0 - g.je0 + aco) * Nz + za);
  auto issue_v = [&](int ip, auto vec_c) {
    constexpr bool VEC = decltype(vec_c)::value;
    const R *base = attf + ((size_t)ang * g.P + ip) * ((size_t)n_jext * Nz);
#pragma unroll
    for (int p = 0; p < NPA; ++p) {
      const int jj = aco + p * cpp;
      if (jj < tj)
        load_zv<R, ZV>(base + (af_col0 + (uint32_t)(p * cpp * Nz)), 0, Nz - za, VEC,
                       Ar[PF ? p : 0]);
    }
  };
  auto issue = [&](int ip, int bb) {
    if (!aact) return;
    if (zvec) issue_v(ip, std::true_type{});
    else issue_v(ip, std::false_type{});
  };
  if constexpr (PF) {
    if (i >= 0) issue(i, 0);
  }
  int b = 0;
  while (i >= 0) {
    const int inext = rem ? i_lo + 63 - __clzll(rem) : -1;
    if (inext >= 0) rem &= ~(1ull << (inext - i_lo));
    // the plane's blur tier (a scalar load, requested before the gather)
    const int he = H > 0 ? ld_uniform(tb.tier + (size_t)rix * g.P + i) : 0;
    // ---- A: V = a * f_rot, ZV z per lane, NPA columns per thread
    auto phase_a = [&](auto vec_c) {
      constexpr bool VEC = decltype(vec_c)::value;
      R *dstA = (H == 0 ? Vz : V + H) + za;
      if constexpr (VEC && !TRB) {
        // the 4 x GA tap loads of GA passes are issued before the first is
        // used (all NPA passes for the float attenuation-field kernel, one
        // pass otherwise, to stay within the register budget); columns past the
        // tile load a valid column's taps and are not stored, so no divergent
        // branch separates the loads
        constexpr int GA = (sizeof(R) == 4 && !MARCH) ? NPA : 1;
#pragma unroll
        for (int p0 = 0; p0 < NPA; p0 += GA) {
        R raw[GA][4][ZV];
        Taps4<R> tt[GA];
#pragma unroll
        for (int g = 0; g < GA; ++g) {
          const int jc = min(aco + (p0 + g) * cpp, tj - 1);
          tt[g] = tp[b][jc];
#pragma unroll
          for (int q = 0; q < 4; ++q) load_zv<R, ZV>(f + tt[g].o[q], za, Nz, true, raw[g][q]);
        }
#pragma unroll
        for (int g = 0; g < GA; ++g) {
          const int p = p0 + g;
          const int jj = aco + p * cpp;
          R val[ZV];
#pragma unroll
          for (int e = 0; e < ZV; ++e) {
            R a = raw[g][0][e] * tt[g].w[0];
            a = a + raw[g][1][e] * tt[g].w[1];
            a = a + raw[g][2][e] * tt[g].w[2];
            a = a + raw[g][3][e] * tt[g].w[3];
            val[e] = a;
          }
          if constexpr (PF) {
#pragma unroll
            for (int e = 0; e < ZV; ++e) val[e] = val[e] * Ar[p][e];
          } else if constexpr (MARCH) {
            if (has_mu) {
              R x[ZV];
              interp_zv<R, ZV>(mu, tt[g], za, Nz, true, x);
#pragma unroll
              for (int e = 0; e < ZV; ++e) {
                const R xm = x[e] * J.g.dx;
                S[AF ? 0 : p][e] += (double)xm;
                val[e] = val[e] * att_factor<R>(S[AF ? 0 : p][e], xm);
              }
            }
          }
          if (jj < tj) {
            if constexpr (RAY) {
#pragma unroll
              for (int e = 0; e < ZV; ++e) accA[p][e] += val[e];
            } else {
#pragma unroll
              for (int e = 0; e < ZV; ++e) dstA[(size_t)jj * ldv + e] = val[e];
            }
          }
        }
        }
        return;
      }
#pragma unroll
      for (int p = 0; p < NPA; ++p) {
        const int jj = aco + p * cpp;
        if (jj < tj) {
          const Taps4<R> t = tp[b][jj];
          R val[ZV];
#pragma unroll
          for (int e = 0; e < ZV; ++e) val[e] = R(0);
          if constexpr (TRB) {
            // contributors in groups of 4 whose loads are issued together;
            // slots past cnt re-read contributor 0 with weight 0 (no branch
            // between the loads); the sum keeps the contributor order
            const int cnt = trb_n[b][jj];
            for (int q0 = 0; q0 < cnt; q0 += 4) {
              R v[4][ZV];
              R w[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const bool use = q0 + u < cnt;
                const int q = use ? q0 + u : 0;
                w[u] = use ? trb_w[b][jj][q] : R(0);
                load_zv<R, ZV>(f + trb_off[b][jj][q], za, Nz, VEC, v[u]);
              }
#pragma unroll
              for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int e = 0; e < ZV; ++e) val[e] += v[u][e] * w[u];
            }
          } else {
            interp_zv<R, ZV>(f, t, za, Nz, VEC, val);
          }
          if constexpr (PF) {
#pragma unroll
            for (int e = 0; e < ZV; ++e) val[e] = val[e] * Ar[p][e];
          } else if constexpr (AF) {
            R a[ZV];
            load_zv<R, ZV>(attf + (((size_t)ang * g.P + i) * n_jext + (j0 - g.je0 + jj)) * Nz, za,
                           Nz, VEC, a);
#pragma unroll
            for (int e = 0; e < ZV; ++e) val[e] = val[e] * a[e];
          } else if constexpr (MARCH) {
            if (has_mu) {
              R x[ZV];
              interp_zv<R, ZV>(mu, t, za, Nz, VEC, x);
#pragma unroll
              for (int e = 0; e < ZV; ++e) {
                const R xm = x[e] * g.dx;
                S[AF ? 0 : p][e] += (double)xm;
                val[e] = val[e] * att_factor<R>(S[AF ? 0 : p][e], xm);
              }
            }
          }
          if constexpr (RAY) {
#pragma unroll
            for (int e = 0; e < ZV; ++e) accA[p][e] += val[e];
          } else {
#pragma unroll
            for (int e = 0; e < ZV; ++e)
              if (VEC || za + e < Nz) dstA[(size_t)jj * ldv + e] = val[e];
          }
        }
      }
    };
    if (aact) {
      if (zvec) phase_a(std::true_type{});
      else phase_a(std::false_type{});
    }
    if (inext >= 0) {
      if (tid < tj) setup(inext, b ^ 1);
      trb_lists(inext, b ^ 1);
    }
    if constexpr (RAY) {
      // the next plane's taps are complete; this plane's may be rewritten
      __syncthreads();
      if constexpr (PF) {
        if (inext >= 0) issue(inext, b ^ 1);
      }
    } else if constexpr (H == 0) {
      __syncthreads();
      if constexpr (PF) {
        if (inext >= 0) issue(inext, b ^ 1);
      }
      if (cact) {
        // no PSF: the ray sum of column jl is the plain sum over planes
#pragma unroll
        for (int q = 0; q < JH; ++q) {
          const int jo = half * JH + q;
          if (jo < JO) acc[q] += Vz[(size_t)jo * ldv + zc];
        }
      }
      __syncthreads();
    } else {
      __syncthreads();
      if constexpr (PF) {
        if (inext >= 0) issue(inext, b ^ 1);
      }
      // ---- B: z blur (correlation, zero padded) of the own run, over the
      // plane's tier of taps [H - HE, H + HE]
      const R *wz = Wt + (size_t)(i - i_lo) * 2 * K;
      if (ab) {
        R win[RZ + 2 * H];
        const R *vr = V + (size_t)jl * ldv + z0;  // win[m] = V(z0 - H + m)
#pragma unroll
        for (int m = 0; m < RZ + 2 * H; ++m) win[m] = vr[m];
        R out[RZ];
#pragma unroll
        for (int m = 0; m < RZ; ++m) out[m] = R(0);
        with_tier<H>(he, [&](auto he_c) {
          constexpr int HE = decltype(he_c)::value;
#pragma unroll
          for (int t = H - HE; t <= H + HE; ++t) {
            const R w = wz[t];
#pragma unroll
            for (int m = 0; m < RZ; ++m) out[m] += w * win[m + t];
          }
        });
#pragma unroll
        for (int m = 0; m < RZ; ++m) Vz[(size_t)jl * ldv + z0 + m] = out[m];
      }
      __syncthreads();
      // ---- C: y blur as a scatter into the tile's output columns
      if (cact) {
        const R *wy = Wt + (size_t)(i - i_lo) * 2 * K + K;
        auto scatter = [&](auto half_c) {
          constexpr int HC = decltype(half_c)::value;
          constexpr int jlo = HC * JH;
          constexpr int jhi = (jlo + JH < JO) ? jlo + JH : JO;
          constexpr int jv0 = (jlo - 2 * H > 0) ? jlo - 2 * H : 0;
          constexpr int jv1 = jhi < kTJ ? jhi : kTJ;
          R v[kTJ];
#pragma unroll
          for (int jv = jv0; jv < jv1; ++jv) v[jv] = Vz[(size_t)jv * ldv + zc];
          with_tier<H>(he, [&](auto he_c) {
            constexpr int HE = decltype(he_c)::value;
            // out[jo] += w[t] * in[jo + t - 2H] (tile-local); t ascending is jv ascending
#pragma unroll
            for (int t = H - HE; t <= H + HE; ++t) {
              const R w = wy[t];
#pragma unroll
              for (int q = 0; q < JH; ++q) {
                const int jo = jlo + q;
                const int jv = jo + t - 2 * H;
                if (jo < jhi && jv >= jv0 && jv < jv1) acc[q] += w * v[jv];
              }
            }
          });
        };
        switch (half) {
          case 0: scatter(std::integral_constant<int, 0>{}); break;
          case 1: if constexpr (QS > 1) scatter(std::integral_constant<int, 1>{}); break;
          case 2: if constexpr (QS > 2) scatter(std::integral_constant<int, 2>{}); break;
          case 3: if constexpr (QS > 3) scatter(std::integral_constant<int, 3>{}); break;
          default: break;
        }
      }
    }
    i = inext;
    b ^= 1;
  }
  if constexpr (RAY) {
    // every column of the tile (zeros past tj), z < Nz
    if (aact) {
      R *dst = partial + pc.slot * JO * Nz;
#pragma unroll
      for (int p = 0; p < NPA; ++p) {
        const int jj = aco + p * cpp;
        if (jj < kTJ)
#pragma unroll
          for (int e = 0; e < ZV; ++e)
            if (za + e < Nz) dst[(size_t)jj * Nz + za + e] = accA[p][e];
      }
    }
  } else if (cact) {
    R *dst = partial + pc.slot * JO * Nz;
#pragma unroll
    for (int q = 0; q < JH; ++q) {
      const int jo = half * JH + q;
      if (jo < JO) dst[(size_t)jo * Nz + zc] = acc[q];
    }
  }
}

// ---------------------------------------------------------------------------
// Balanced schedule.  Every (angle, tile) pair p = al * ntiles + jt visits a
// contiguous range of planes [i_min, i_max] (the plan's activity table), n_p
// planes.  The schedule splits each pair into c_p = ceil(n_p / Q) pieces of
// nearly equal length, with the smallest Q <= kPieceMax for which all pieces
// fit in the resident blocks at once (Q >= ceil(T / blocks), T = sum n_p), so
// one launch runs as one round of blocks of at most Q planes each -- instead
// of uniform chunks, whose central ones are full and edge ones nearly empty.
// Block b takes piece b (pieces ordered by pair, then position): piece k of
// pair p covers positions [k n_p / c_p, (k + 1) n_p / c_p) of the pair's
// planes walked from i_max down, and writes partial slot b.
//
// Every block of a scheduled launch computes the schedule itself (a few
// hundred pairs: two loads per pair, wave reductions and four barriers), so
// no separate launch precedes it; block 0 also writes it out for the reduce.
// sched layout (ints): [0] total pieces, [1] Q, then base[0..npairs]
// (base[p] = pieces of the pairs before p).

// The block of the scheduled forward that takes piece b: it writes partial
// slot b; piece 0's block writes the schedule to `sched` for the reduce.
template <typename R, int RZ, int QS, int H, bool AF>
__device__ __forceinline__ void fwd_sched_body(const FwdJob<R> &J, const int2 *__restrict__ ranges,
                                               int *__restrict__ sched, int nblocks, int b,
                                               unsigned char *smem) {
  const Geo<R> &g = J.g;
  const SchedLoc s = sched_locate<R>(g, J.angle_idx, ranges, J.n, nblocks, b, sched);
  if (s.p < 0) return;
  const int al = s.p / g.ntiles, jt = s.p - al * g.ntiles;
  const int ang = ld_uniform(J.angle_idx + al);
  const int2 r = ranges[(size_t)ang * g.ntiles + jt];
  FwdPiece pc;
  sched_piece_planes(r, s.k, s.c, pc.i_lo, pc.i_hi);
  pc.c = -1;
  pc.ranged = 1;
  pc.slot = (size_t)b;
  fwd_body<R, RZ, QS, H, false, AF, false>(J, al, jt, pc, smem);
}

// Reduce of the scheduled forward: the pieces of the <= 3 tiles a column
// lies in, tiles ascending, pieces ascending (a fixed order).
template <typename R>
__device__ __forceinline__ void reduce_sched_body(const ReduceJob<R> &J, const int *__restrict__ sched,
                                                  size_t idx) {
  const Geo<R> &g = J.g;
  const R *__restrict__ partial = J.partial;
  const int n = J.n;
  const int nout = g.jo1 - g.jo0;
  const size_t count = (size_t)n * nout * g.Nz;
  if (idx >= count) return;
  const int z = idx % g.Nz;
  const int jx = (idx / g.Nz) % nout;
  const int al = idx / ((size_t)g.Nz * nout);
  const int JO = kTJ + 2 * g.H;
  const int rel = g.jo0 + jx - g.je0 + g.H;  // column within tile t: rel - t*kTJ
  int t_hi = rel / kTJ;
  if (t_hi > g.ntiles - 1) t_hi = g.ntiles - 1;
  const int t_lo = (rel - JO) < 0 ? 0 : (rel - JO) / kTJ + 1;
  const int *base = sched + 2;
  const size_t plane = (size_t)JO * g.Nz;
  R acc = R(0);
  for (int t = t_lo; t <= t_hi; ++t) {
    const int p = al * g.ntiles + t;
    const int k0 = base[p], k1 = base[p + 1];
    const R *src = partial + (size_t)(rel - t * kTJ) * g.Nz + z;
    // 8 loads in flight; slots past k1 re-read a valid piece and are not added
    constexpr int kRB = 8;
    for (int k = k0; k < k1; k += kRB) {
      R v[kRB];
#pragma unroll
      for (int u = 0; u < kRB; ++u) v[u] = src[(size_t)min(k + u, k1 - 1) * plane];
#pragma unroll
      for (int u = 0; u < kRB; ++u)
        if (k + u < k1) acc += v[u];
    }
  }
  if (J.proj) J.proj[idx] = acc;
  if (J.gm) {
    R p = acc;
    if (J.s) p = p + J.s[idx];
    if (J.pred) J.pred[idx] = p;
    J.ratio[idx] = J.gm[idx] / (p + J.delta);
  }
}

template <typename R>
__device__ __forceinline__ void reduce_body(const ReduceJob<R> &J, size_t idx) {
  const Geo<R> &g = J.g;
  const R *__restrict__ partial = J.partial;
  const int n = J.n;
  R *__restrict__ proj = J.proj;
  const R *__restrict__ gm = J.gm;
  const R *__restrict__ s = J.s;
  R *__restrict__ pred = J.pred;
  R *__restrict__ ratio = J.ratio;
  const R delta = J.delta;
  const int nout = g.jo1 - g.jo0;
  const size_t count = (size_t)n * nout * g.Nz;
  if (idx >= count) return;
  const int z = idx % g.Nz;
  const int jx = (idx / g.Nz) % nout;
  const int al = idx / ((size_t)g.Nz * nout);
  const int JO = kTJ + 2 * g.H;
  const int j = g.jo0 + jx;
  const int rel = j - g.je0 + g.H;  // column within tile t: rel - t*kTJ
  int t_hi = rel / kTJ;
  if (t_hi > g.ntiles - 1) t_hi = g.ntiles - 1;
  int t_lo = (rel - JO) < 0 ? 0 : (rel - JO) / kTJ + 1;
  // The partials are added in the order (chunk ascending, tile ascending).
  // kRB chunks x kRT tiles of loads are issued before the first add; a slot
  // past the chunk count or the column's last tile re-reads a valid entry and
  // is not added, so no branch separates the loads.  A column lies in at most
  // ceil((kTJ + 2H) / kTJ) <= 3 tiles (H <= 20).
  constexpr int kRB = 8, kRT = 3;
  const size_t plane = (size_t)JO * g.Nz;
  const size_t base0 = (size_t)al * g.ntiles * plane + (size_t)rel * g.Nz + z;
  R acc = R(0);
  for (int c0 = 0; c0 < g.C; c0 += kRB) {
    R v[kRB][kRT];
#pragma unroll
    for (int u = 0; u < kRB; ++u) {
      const int c = min(c0 + u, g.C - 1);
#pragma unroll
      for (int k = 0; k < kRT; ++k) {
        const int t = min(t_lo + k, t_hi);
        // entry (c, al, t, jo = rel - t*kTJ, z)
        v[u][k] = partial[((size_t)c * n * g.ntiles + t) * plane + base0 - (size_t)t * kTJ * g.Nz];
      }
    }
#pragma unroll
    for (int u = 0; u < kRB; ++u)
#pragma unroll
      for (int k = 0; k < kRT; ++k)
        if (c0 + u < g.C && t_lo + k <= t_hi) acc += v[u][k];
  }
  if (proj) proj[idx] = acc;
  if (gm) {
    R p = acc;
    if (s) p = p + s[idx];
    if (pred) pred[idx] = p;
    ratio[idx] = gm[idx] / (p + delta);
  }
}

}  // namespace spect
