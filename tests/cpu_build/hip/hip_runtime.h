// Host stand-in for <hip/hip_runtime.h> used by the CPU build of the forward
// projector's block bodies (tests/cpu_build/fwd_cpu.cpp): the device
// qualifiers compile away, __shared__ arrays become function statics (one
// block runs at a time), every GPU lane is a std::thread and __syncthreads
// is a std::barrier over the block's threads.
#pragma once
#include <algorithm>
#include <barrier>
#include <cmath>
#include <cstddef>
#include <cstdint>

using std::max;
using std::min;

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
#define __align__(x) alignas(x)

struct cpu_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
extern thread_local cpu_dim3 threadIdx;
extern cpu_dim3 blockDim;
extern std::barrier<> *cpu_block_barrier;

inline void __syncthreads() { cpu_block_barrier->arrive_and_wait(); }
inline unsigned long long atomicOr(unsigned long long *p, unsigned long long v) {
  return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
inline int __clzll(unsigned long long v) { return __builtin_clzll(v); }
// Block-wide OR of pred: two barriers, the partial ORs double-buffered so that
// consecutive calls never share a slot.
int __syncthreads_or(int pred);
inline float __expf(float v) { return std::exp(v); }
// Cross-lane operations of 64-lane waves, emulated block-synchronously (two
// barriers each, like __syncthreads_or): the bodies call them from control
// flow that every lane of the block reaches.  DPP controls: quad_perm
// (0x00-0xFF), row_shl (0x101-0x10F), row_shr (0x111-0x11F), row_ror
// (0x121-0x12F); a source lane outside the row gives `old` (or 0 with
// bound_ctrl).
int __builtin_amdgcn_update_dpp(int old, int src, int dpp_ctrl, int row_mask, int bank_mask,
                                bool bound_ctrl);
int __builtin_amdgcn_readlane(int v, int lane);
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
// Bit l of the result: pred of lane l of the caller's wave (lanes past
// blockDim.x read as 0).
unsigned long long __ballot(int pred);
inline int __popc(unsigned v) { return __builtin_popcount(v); }

struct float2 {
  float x, y;
};
struct int2 {
  int x, y;
};
inline int2 make_int2(int a, int b) { return {a, b}; }
struct double2 {
  double x, y;
};
struct alignas(16) float4 {
  float x, y, z, w;
};
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline double2 make_double2(double a, double b) { return {a, b}; }

typedef int hipError_t;
typedef void *hipStream_t;
