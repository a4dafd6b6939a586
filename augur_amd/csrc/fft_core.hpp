// In-LDS radix-4/2 Stockham FFT core shared by the FFT and PM kernels.
#pragma once
#include "internal.hpp"

namespace augur {

constexpr int FFT_THREADS = 256;
constexpr int FFT_TILE = 4096;  // complex elements per workgroup

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

template <int N>
__device__ __forceinline__ void make_twiddles(float2* tw) {
  // tw[t] = exp(-2 pi i t / N); sincospif is accurate to ~1 ulp
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    float s, c;
    sincospif(2.0f * (float)t / (float)N, &s, &c);
    tw[t] = make_float2(c, -s);
  }
}

// Transforms `nl` lines of length N stored at a + line*ls (and b as scratch).
// Returns the buffer holding the result.  Caller must __syncthreads() before.
template <int N, bool INV>
__device__ float2* lds_fft(float2* a, float2* b, const float2* tw, int nl, int ls) {
  float2* src = a;
  float2* dst = b;
  int Ns = 1;
  constexpr int LOGN = (N <= 1) ? 0 : 31 - __builtin_clz(N);
#pragma unroll 1
  for (int st = 0; st < LOGN / 2; ++st) {
    constexpr int Q = N / 4;
    const int nb = nl * Q;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
      const int line = t / Q, j = t - line * Q;
      const float2* x = src + line * ls;
      float2* y = dst + line * ls;
      float2 v0 = x[j], v1 = x[j + Q], v2 = x[j + 2 * Q], v3 = x[j + 3 * Q];
      const int k = j & (Ns - 1);
      if (k) {
        const int ts = k * (N / (Ns * 4));
        float2 w1 = tw[ts], w2 = tw[2 * ts], w3 = tw[3 * ts];
        if (INV) {
          w1.y = -w1.y;
          w2.y = -w2.y;
          w3.y = -w3.y;
        }
        v1 = cmul(v1, w1);
        v2 = cmul(v2, w2);
        v3 = cmul(v3, w3);
      }
      const float2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3);
      float2 t3 = csub(v1, v3);
      // forward: t3 * (-i); inverse: t3 * (+i)
      t3 = INV ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);
      const int d = (j - k) * 4 + k;
      y[d] = cadd(t0, t2);
      y[d + Ns] = cadd(t1, t3);
      y[d + 2 * Ns] = csub(t0, t2);
      y[d + 3 * Ns] = csub(t1, t3);
    }
    __syncthreads();
    float2* tmp = src;
    src = dst;
    dst = tmp;
    Ns *= 4;
  }
  if (LOGN & 1) {
    constexpr int Q = N / 2;
    const int nb = nl * Q;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
      const int line = t / Q, j = t - line * Q;
      const float2* x = src + line * ls;
      float2* y = dst + line * ls;
      float2 v0 = x[j], v1 = x[j + Q];
      const int k = j & (Ns - 1);
      if (k) {
        float2 w = tw[k * (N / (Ns * 2))];
        if (INV) w.y = -w.y;
        v1 = cmul(v1, w);
      }
      const int d = (j - k) * 2 + k;
      y[d] = cadd(v0, v1);
      y[d + Ns] = csub(v0, v1);
    }
    __syncthreads();
    float2* tmp = src;
    src = dst;
    dst = tmp;
  }
  return src;
}


}  // namespace augur
