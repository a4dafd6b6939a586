// Shared helpers for the augur HIP library: status codes, the thread-local
// last-error string, launch checking and small device utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <stdarg.h>
#include <stdio.h>
#include <cmath>

#include "../../include/augur.h"

namespace augur {

void set_error(const char* fmt, ...);

#define AUGUR_REQUIRE(cond, code, ...)        \
  do {                                        \
    if (!(cond)) {                            \
      ::augur::set_error(__VA_ARGS__);        \
      return (code);                          \
    }                                         \
  } while (0)

#define AUGUR_HIP(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::augur::set_error("%s failed: %s", #call, hipGetErrorString(e_));       \
      return AUGUR_EHIP;                                                       \
    }                                                                          \
  } while (0)

#define AUGUR_LAUNCHED(name)                                                   \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      ::augur::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
      return AUGUR_EHIP;                                                       \
    }                                                                          \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

inline int ilog2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Carves aligned sub-buffers out of a caller-provided workspace.
struct Carver {
  char* base;
  size_t cap, used = 0;
  Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <class T>
  T* take(size_t count) {
    size_t off = align_up(used);
    used = off + count * sizeof(T);
    return (base && used <= cap) ? reinterpret_cast<T*>(base + off) : nullptr;
  }
  bool ok() const { return base != nullptr && used <= cap; }
};

// Grid-stride bound for 1D launches: enough blocks to fill 256 CUs several
// times over without exceeding the work.
inline unsigned blocks_for(int64_t n, int threads, int64_t cap = 65536) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace augur

// ---------------------------------------------------------------- device side
namespace augur_dev {

__device__ __forceinline__ int wrap(int i, int G) {  // G is a power of two
  return i & (G - 1);
}

// CIC base coordinate u = (p + off)/h - 0.5 with the reference's rounding
// sequence: rounded add, IEEE divide, rounded subtract, no contraction.
__device__ __forceinline__ float cic_u(float p, float off, float h) {
  float t = __fadd_rn(p, off);
  t = __fdiv_rn(t, h);
  return __fsub_rn(t, 0.5f);
}

// The same value when h is a power of two and inv_h = 1/h exactly: both
// operations then round the same real number p/h once, and p + 0 only turns
// -0 into +0, which the subtraction of 0.5 maps to the same result.
__device__ __forceinline__ float cic_u_pow2(float p, float inv_h) { return __fsub_rn(__fmul_rn(p, inv_h), 0.5f); }

// Fixed-point exponent of the order-independent deposits: the largest e with
// n * max|v| * 2^e <= 2^52.  A contribution v counts rint(v * 2^e) units, and
// any sum of n of them is an integer of magnitude below 2^53: exact in int64
// and in FP64 alike, so every summation order gives the same bits.
__device__ __forceinline__ int fixed_exponent(unsigned maxbits, int64_t n) {
  const float m = __uint_as_float(maxbits);
  if (!(m > 0.0f)) return 0;
  int e = (int)floor(52.0 - log2((double)m * (double)(n > 0 ? n : 1)));
  return e > 1000 ? 1000 : (e < -1000 ? -1000 : e);
}

// rint(v * 2^e) as a 64-bit integer, exact whenever |v * 2^e| < 2^62 (which
// fixed_exponent guarantees for every contribution).  All FP32: the scaling by
// 2^e is exact, rintf rounds half to even, and the magnitude splits exactly
// into two 32-bit halves (a - hi * 2^32 is a multiple of ulp(a) below 2^32).
__device__ __forceinline__ long long fixed_round(float v, int e) {
  const float r = rintf(ldexpf(v, e));
  const float a = fabsf(r);
  const float hi = floorf(a * 2.3283064365386963e-10f);
  const float lo = fmaf(hi, -4294967296.0f, a);
  const unsigned long long q = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  return r < 0.0f ? -(long long)q : (long long)q;
}

// acc[key] += q for every active lane of the wave, with runs of equal keys in
// adjacent lanes summed first (a segmented scan over the wave) so that only
// the run's last lane issues the atomic.  Cell-sorted particles put a dense
// cell's visits in adjacent lanes, where same-address atomics would
// serialise; the integer sums are exact, so merging changes no bits.  Every
// lane of the wave must call it (inactive lanes with active = false).
__device__ __forceinline__ void wave_merged_add(unsigned long long* __restrict__ acc, int64_t key, long long q,
                                                bool active) {
  const int lane = (int)(threadIdx.x & 63);
  const int64_t k = active ? key : (int64_t)(-1 - lane);  // inactive lanes match nothing
  const int64_t kprev = __shfl_up(k, 1);
  const bool head = lane == 0 || kprev != k;
  if (__ballot(head) == ~0ull) {  // no runs in this wave: one atomic per lane
    if (active && q) atomicAdd(&acc[key], (unsigned long long)q);
    return;
  }
  long long v = active ? q : 0;
  int f = head ? 1 : 0;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const long long vo = __shfl_up(v, off);
    const int fo = __shfl_up(f, off);
    if (lane >= off && !f) {
      v += vo;
      f |= fo;
    }
  }
  const int64_t knext = __shfl_down(k, 1);
  const bool tail = lane == 63 || knext != k;
  if (active && tail && v) atomicAdd(&acc[key], (unsigned long long)v);
}

// h = L / G in FP32 is a power of two (then cic_u_pow2 applies)
inline bool pow2_float(float h) {
  int ex;
  return h > 0.0f && std::frexp(h, &ex) == 0.5f;
}

}  // namespace augur_dev
