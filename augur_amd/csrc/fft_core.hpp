// Register-resident Stockham FFT core shared by the FFT and PM kernels.
//
// A line of N points is transformed in P passes of radix R_p (N = prod R_p):
// each thread holds one radix-R butterfly in registers (a DFT-R done with
// compile-time twiddles), multiplies by the inter-pass twiddles from a
// per-workgroup table, and the passes exchange data through one padded LDS
// buffer (a barrier separates the loads and the stores of a pass, so the
// buffer is reused in place).  The first pass reads and the last pass writes
// through caller functors, so loads from HBM feed registers directly and
// results go straight back out: a 256-point line costs one LDS round trip.
#pragma once
#include "internal.hpp"

namespace augur {

constexpr int FFT_THREADS = 256;

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

// ------------------------------------------------------ in-register DFTs
// exp(-+ 2 pi i k / 16) for k = 0..7 (forward sign); cos and sin of k*22.5 deg
__device__ constexpr float C16[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                                     0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f};
__device__ constexpr float S16[8] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                                     1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f};

template <bool INV>
__device__ __forceinline__ float2 w16(int k) {  // W_16^k for k in [0, 8)
  return make_float2(C16[k], INV ? S16[k] : -S16[k]);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  const float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
  float2 t3 = csub(a1, a3);
  t3 = INV ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);  // * -+i
  a0 = cadd(t0, t2);
  a1 = cadd(t1, t3);
  a2 = csub(t0, t2);
  a3 = csub(t1, t3);
}

// natural-order DFT of R points held in v[0..R)
template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<INV>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(v[0], v[1], v[2], v[3]);
  } else {
    constexpr int H = R / 2;
    float2 e[H], o[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      e[k] = v[2 * k];
      o[k] = v[2 * k + 1];
    }
    dft<H, INV>(e);
    dft<H, INV>(o);
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const float2 t = cmul(o[k], w16<INV>(k * (16 / R)));
      v[k] = cadd(e[k], t);
      v[k + H] = csub(e[k], t);
    }
  }
}

// ----------------------------------------------------------- pass plans
template <int N>
struct FftPlan;
template <> struct FftPlan<4> { static constexpr int P = 1, R0 = 4, R1 = 1, R2 = 1; };
template <> struct FftPlan<8> { static constexpr int P = 1, R0 = 8, R1 = 1, R2 = 1; };
template <> struct FftPlan<16> { static constexpr int P = 1, R0 = 16, R1 = 1, R2 = 1; };
template <> struct FftPlan<32> { static constexpr int P = 2, R0 = 8, R1 = 4, R2 = 1; };
template <> struct FftPlan<64> { static constexpr int P = 2, R0 = 8, R1 = 8, R2 = 1; };
template <> struct FftPlan<128> { static constexpr int P = 2, R0 = 16, R1 = 8, R2 = 1; };
template <> struct FftPlan<256> { static constexpr int P = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct FftPlan<512> { static constexpr int P = 3, R0 = 8, R1 = 8, R2 = 8; };

// lines per workgroup so that every pass has at most FFT_THREADS butterflies
template <int N, int T = FFT_THREADS>
struct FftLines {
  static constexpr int RMIN = FftPlan<N>::P == 1 ? FftPlan<N>::R0
                              : (FftPlan<N>::P == 2 ? (FftPlan<N>::R0 < FftPlan<N>::R1 ? FftPlan<N>::R0 : FftPlan<N>::R1)
                                                     : 8);
  static constexpr int value = T / (N / RMIN);
};

// padded LDS index: one spare complex every 16 spreads strided accesses over banks
__host__ __device__ constexpr int fft_pad(int i) { return i + (i >> 4); }
template <int N>
struct FftStride {
  static constexpr int value = fft_pad(N) + 1;
};

// One pass: butterfly t of the workgroup.  LF (line-fastest) maps
// consecutive threads to consecutive lines (adjacent columns in memory for
// the strided-axis kernels); otherwise consecutive threads take consecutive
// butterflies of one line (contiguous rows).
template <int N, int R, bool INV, bool LF, class Src>
__device__ __forceinline__ bool pass_load(int nl, int Ns, const float2* tw, Src src, float2* v, int& line, int& j) {
  constexpr int Q = N / R;
  const int t = threadIdx.x;
  if (t >= nl * Q) return false;
  if (LF) {
    line = t % nl;
    j = t / nl;
  } else {
    line = t / Q;
    j = t - line * Q;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src(line, j + r * Q);
  const int k = j & (Ns - 1);
  if (Ns > 1 && k) {
    const int step = k * (N / (Ns * R));
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[(step * r) & (N - 1)];
      if (INV) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  dft<R, INV>(v);
  return true;
}

template <int R, class Dst>
__device__ __forceinline__ void pass_store(int Ns, Dst dst, const float2* v, int line, int j) {
  const int k = j & (Ns - 1);
  const int base = (j - k) * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) dst(line, base + r * Ns, v[r]);
}

// Transforms nl lines: src(line, i) gives input point i, dst(line, i, v)
// receives output point i (natural order).  lds: nl * FftStride<N> complex.
// SRC_LDS / DST_LDS: src reads, or dst writes, the lds buffer itself (an
// in-place transform); a barrier then separates the first pass's loads from
// its stores (and the last pass's loads from its stores).
// All threads of the block must call it (it contains barriers).
template <int N, bool INV, bool LF = false, bool SRC_LDS = false, bool DST_LDS = false, class Src, class Dst>
__device__ void fft_lines(float2* lds, const float2* tw, int nl, Src src, Dst dst) {
  using PL = FftPlan<N>;
  constexpr int LS = FftStride<N>::value;
  auto lds_src = [&](int line, int i) { return lds[line * LS + fft_pad(i)]; };
  auto lds_dst = [&](int line, int i, float2 v) { lds[line * LS + fft_pad(i)] = v; };
  int line = 0, j = 0;
  if constexpr (PL::P == 1) {
    float2 v[PL::R0];
    const bool ok = pass_load<N, PL::R0, INV, LF>(nl, 1, tw, src, v, line, j);
    if constexpr (SRC_LDS || DST_LDS) __syncthreads();
    if (ok) pass_store<PL::R0>(1, dst, v, line, j);
  } else if constexpr (PL::P == 2) {
    {
      float2 v[PL::R0];
      const bool ok = pass_load<N, PL::R0, INV, LF>(nl, 1, tw, src, v, line, j);
      if constexpr (SRC_LDS) __syncthreads();
      if (ok) pass_store<PL::R0>(1, lds_dst, v, line, j);
    }
    __syncthreads();
    {
      float2 v[PL::R1];
      const bool ok = pass_load<N, PL::R1, INV, LF>(nl, PL::R0, tw, lds_src, v, line, j);
      if constexpr (DST_LDS) __syncthreads();
      if (ok) pass_store<PL::R1>(PL::R0, dst, v, line, j);
    }
  } else {
    {
      float2 v[PL::R0];
      const bool ok = pass_load<N, PL::R0, INV, LF>(nl, 1, tw, src, v, line, j);
      if constexpr (SRC_LDS) __syncthreads();
      if (ok) pass_store<PL::R0>(1, lds_dst, v, line, j);
    }
    __syncthreads();
    {
      float2 v[PL::R1];
      const bool ok = pass_load<N, PL::R1, INV, LF>(nl, PL::R0, tw, lds_src, v, line, j);
      __syncthreads();
      if (ok) pass_store<PL::R1>(PL::R0, lds_dst, v, line, j);
    }
    __syncthreads();
    {
      float2 v[PL::R2];
      const bool ok = pass_load<N, PL::R2, INV, LF>(nl, PL::R0 * PL::R1, tw, lds_src, v, line, j);
      if constexpr (DST_LDS) __syncthreads();
      if (ok) pass_store<PL::R2>(PL::R0 * PL::R1, dst, v, line, j);
    }
  }
  __syncthreads();  // the buffer may be reused by the caller
}

}  // namespace augur
