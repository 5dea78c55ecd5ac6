// sf_common.hpp -- shared definitions of the sfhip device code (gfx950).
//
// Launch shapes, the supported (n, q) list, the workgroup -> XCD remap and
// small device helpers used by every kernel family.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sf {

// Supported (n = p + 1, q) pairs: p = 1..9 and q = n-1 .. n+3 (q >= 1).
// This covers the reference's own choices (q = p for quadrature degree 2p-1
// or 2(p-1), test/benchmark_v3.jl:36-40, test/Demo.jl:22-23) and q = p + 2
// used by the perturbed-mesh workloads.
#define SF_FOR_EACH_NQ(X)                                                      \
  X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5)                                      \
  X(3, 2) X(3, 3) X(3, 4) X(3, 5) X(3, 6)                                      \
  X(4, 3) X(4, 4) X(4, 5) X(4, 6) X(4, 7)                                      \
  X(5, 4) X(5, 5) X(5, 6) X(5, 7) X(5, 8)                                      \
  X(6, 5) X(6, 6) X(6, 7) X(6, 8) X(6, 9)                                      \
  X(7, 6) X(7, 7) X(7, 8) X(7, 9) X(7, 10)                                     \
  X(8, 7) X(8, 8) X(8, 9) X(8, 10) X(8, 11)                                    \
  X(9, 8) X(9, 9) X(9, 10) X(9, 11) X(9, 12)                                   \
  X(10, 9) X(10, 10) X(10, 11) X(10, 12) X(10, 13)

constexpr int kMaxN = 10;
constexpr int kMaxQ = 13;
constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int round_up(int a, int m) { return (a + m - 1) / m * m; }

// Workgroups are dealt to the 8 XCDs round-robin (id % 8).  Remap so that
// each XCD works on one contiguous range of cell batches: neighbouring
// cells share face dofs, so the gather of x and the scatter into y then hit
// the same XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int per = nblk / kNumXcd;
  const int rem = nblk - per * kNumXcd;
  const int xcd = bid % kNumXcd;
  const int j = bid / kNumXcd;
  return xcd * per + (xcd < rem ? xcd : rem) + j;
}

// The 1D tables are read with uniform scalar loads.  A fully unrolled
// contraction uses Q*N entries of each table; left alone the compiler hoists
// all of them into SGPRs and spills the excess into VGPR lanes.  table_row()
// hands each unrolled output row an opaque copy of the table pointer (kept
// in the constant address space, so loads stay scalar): the loads of a row
// cannot move above it, and only about one row of each table is live.
template <typename P>
using cptr = const __attribute__((address_space(4))) P*;

template <typename P>
__device__ __forceinline__ cptr<P> table_row(const P* p) {
  cptr<P> c = (cptr<P>)p;
  asm volatile("" : "+s"(c));
  return c;
}

// Rows of the B and D tables for an unrolled contraction, re-fetched every
// TK rows: at(r) hands out a fresh opaque pointer pair when r % TK == 0, so
// the scalar loads of TK rows are issued together and waited for once
// (one wait per row leaves the wave stalled on the scalar cache between
// short bursts of FMAs; all rows at once overflow the SGPRs).
template <typename T, int TK>
struct TableRows {
  const T* tb;
  const T* td;
  cptr<T> b, d;
  __device__ __forceinline__ TableRows(const T* tb_, const T* td_) : tb(tb_), td(td_) {}
  __device__ __forceinline__ void at(int r) {
    if (r % TK == 0) {
      b = table_row(tb);
      d = table_row(td);
    }
  }
};

// bf16 storage for the geometric factor (round-to-nearest-even on store).
struct bf16s {
  uint16_t bits;
};

__host__ __device__ inline uint16_t f32_to_bf16_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// Streamed, read-once load of the geometric factor (non-temporal hint so
// the G stream does not evict the reused x / y lines from L2).
template <typename TG, typename T>
__device__ __forceinline__ T load_g(const TG* p);

template <>
__device__ __forceinline__ double load_g<double, double>(const double* p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ float load_g<float, float>(const float* p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ float load_g<bf16s, float>(const bf16s* p) {
  const uint16_t h = __builtin_nontemporal_load(&p->bits);
  return bf16_bits_to_f32(h);
}

// LDS (workgroup-scope) atomic add; the result is unused, so the
// no-return form ds_add_f64 / ds_add_f32 is emitted.
// Two adjacent reduced-precision G entries in one load.
template <typename TG, typename T>
__device__ __forceinline__ void load_g_pair(const TG* p, T* lo, T* hi);

template <>
__device__ __forceinline__ void load_g_pair<float, float>(const float* p,
                                                         float* lo, float* hi) {
  typedef float fv2 __attribute__((ext_vector_type(2)));
  const fv2 v = __builtin_nontemporal_load((const fv2*)p);
  *lo = v.x;
  *hi = v.y;
}
template <>
__device__ __forceinline__ void load_g_pair<bf16s, float>(const bf16s* p,
                                                         float* lo, float* hi) {
  const uint32_t v = __builtin_nontemporal_load((const uint32_t*)p);
  *lo = __uint_as_float(v << 16);
  *hi = __uint_as_float(v & 0xffff0000u);
}

__device__ __forceinline__ void lds_atomic_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_atomic_add(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void atomic_add(double* p, double v) {
  unsafeAtomicAdd(p, v);
}
__device__ __forceinline__ void atomic_add(float* p, float v) {
  unsafeAtomicAdd(p, v);
}

// Sum over the 64 lanes of a wave (valid in lane 0).
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

}  // namespace sf
