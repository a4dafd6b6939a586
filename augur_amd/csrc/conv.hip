// BF16 MFMA 3x3x3 periodic convolution (forward, data-gradient, weight-
// gradient) for the scatter-field surrogate on gfx950, FP32 accumulation.
//
// Activations are channels-last BF16: x[b][z][y][x][c].  Weights are stored
// tap-major [27][32][CIN] (tap = (dz*3 + dy)*3 + dx, offsets d - 1).
// One workgroup (4 waves) computes a 32 x 4 x 2 output tile for all 32 output
// channels: the tile's input with a one-voxel periodic halo and the whole
// weight tensor are staged in LDS, and each wave runs
// v_mfma_f32_32x32x16_bf16 over 27 taps x CIN/16 K-steps for two 32-voxel
// rows, sharing each weight fragment between its two rows.
#include <type_traits>

#include "internal.hpp"

using namespace augur;

namespace augur {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
typedef __attribute__((ext_vector_type(2))) uint32_t u32x2;

constexpr int CT_X = 32, CT_Y = 4, CT_Z = 2;
constexpr int CONV_THREADS = 256;
constexpr int HX = CT_X + 2, HY = CT_Y + 2, HZ = CT_Z + 2;

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// round-to-nearest-even FP32 -> BF16 (the hardware conversion)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
// two values in one packed conversion: lo in bits 0..15, hi in bits 16..31
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// GELU in its tanh form, 0.5 z (1 + tanh(u)) with u = sqrt(2/pi) (z + 0.044715 z^3),
// written as z / (1 + e^(-2u)) on the hardware exp2 and reciprocal (relative
// error ~1e-6, far below the 2^-8 of the BF16 activations)
constexpr float GELU_K = -2.3022082f;           // -2 sqrt(2/pi) log2(e)
constexpr float GELU_C = 0.044715f;
constexpr float GELU_D = 1.5957691216057308f;   // 2 sqrt(2/pi)
__device__ __forceinline__ float gelu(float z) {
  const float e = __builtin_amdgcn_exp2f(z * fmaf(GELU_K * GELU_C, z * z, GELU_K));
  return z * __builtin_amdgcn_rcpf(1.0f + e);
}
// d/dz of the same function: s + z s (1 - s) 2u'(z), s = 1 / (1 + e^(-2u))
__device__ __forceinline__ float gelu_grad(float z) {
  const float e = __builtin_amdgcn_exp2f(z * fmaf(GELU_K * GELU_C, z * z, GELU_K));
  const float sg = __builtin_amdgcn_rcpf(1.0f + e);
  return fmaf(z * sg * (1.0f - sg), GELU_D * fmaf(3.0f * GELU_C, z * z, 1.0f), sg);
}

enum ConvMode : int {
  CONV_OUT_F32_CF = 0,   // y[b][co][vox] = conv + bias, co < cout (FP32, channels-first)
  CONV_GELU = 1,         // z = conv + bias -> z_out; h_out = gelu(z)            (BF16 channels-last)
  CONV_RES_GELU = 2,     // z -> z_out; h_out = x_center + gelu(z)               (BF16, CIN = 32)
  CONV_OUT_F32_CL = 3,   // y[vox][co] = conv (+ addend[vox][co])                (FP32 channels-last, no bias)
  CONV_OUT_F32_CL_DGELU = 4,  // as 3 (32 -> 32), and dz[vox][co] = BF16(y * gelu'(z_in[vox][co])): the next
                              // layer's GELU backward fused into this data gradient (z_out: z_in, h_out: dz)
  CONV_OUT_F32_CF_INTO = 5,   // y[b][co][vox] = alpha (conv + bias) (+ y[b][co][vox] when accumulating), FP32
                              // channels-first in place; the launch does nothing while *skip != 0
};

// ------------------------------------------------------- z-marching forward
// One workgroup computes a 32 x 16 (x, y) column of output voxels over ZM_ZC
// consecutive z planes, all 32 output channels.  Each wave keeps the whole
// weight tensor in registers (27 taps x CIN/16 MFMA A fragments) and owns RPW
// output rows (4 rows, four waves; for CIN = 16, 2 rows and eight waves, so two
// waves share a SIMD); the input planes z-1, z, z+1 (with a one-voxel
// periodic halo) sit in a 3-plane LDS ring.  An input row fragment is loaded
// once per (dz, dx, K-step) and feeds every output row it touches (dy = 0..2),
// two MFMAs per LDS read on average.  The next plane is fetched into
// registers while the current plane is computed, then written over the
// plane that is no longer needed.
constexpr int ZM_X = 32, ZM_Y = 16, ZM_ZC = 16, ZM_THREADS = 256, ZM_RPW = 4;
constexpr int ZM_PX = ZM_X + 2, ZM_PY = ZM_Y + 2;
constexpr int ZS_THREADS = 512;  // the eight-wave kernels (split outputs, narrow outputs)

template <int CIN, int MODE, int RPW = ZM_RPW>
__global__ __launch_bounds__(64 * (ZM_Y / RPW)) void k_conv3_zm(const uint16_t* __restrict__ x,
                                                         const uint16_t* __restrict__ w,
                                                         const float* __restrict__ bias, int G, int zc, int cout,
                                                         uint16_t* __restrict__ z_out, uint16_t* __restrict__ h_out,
                                                         float* __restrict__ y_out, const float* __restrict__ addend) {
  constexpr int CP = CIN + 8;  // padded channel stride (conflict-free 16-byte LDS reads)
  constexpr int NH = CIN / 16;
  constexpr int PLANE = ZM_PY * ZM_PX * CP;
  constexpr int CHUNKS = ZM_PY * ZM_PX * (CIN / 8);
  constexpr int NT = 64 * (ZM_Y / RPW);  // threads: one wave per RPW rows
  constexpr int PER = (CHUNKS + NT - 1) / NT;
  __shared__ __attribute__((aligned(16))) uint16_t ring[3 * PLANE];
  __shared__ __attribute__((aligned(16))) float sbias[32];
  const int tid = threadIdx.x;
  const int ntx = G / ZM_X, nty = G / ZM_Y, ntz = G / zc;
  int t = blockIdx.x;
  const int bx = t % ntx;
  t /= ntx;
  const int by = t % nty;
  t /= nty;
  const int bz = t % ntz;
  const int b = t / ntz;
  const int x0 = bx * ZM_X, y0 = by * ZM_Y, zs = bz * zc;
  const int64_t G3 = (int64_t)G * G * G;
  const uint16_t* xb = x + (int64_t)b * G3 * CIN;
  const int lane = tid & 63, wv = tid >> 6;
  const int i = lane & 31, kh = lane >> 5;
  if (tid < 32) sbias[tid] = (MODE == CONV_OUT_F32_CL || tid >= cout) ? 0.0f : bias[tid];

  // A operand: w[tap][co = i][k-slice of lane]
  bf16x8 wr[27][NH];
#pragma unroll
  for (int tap = 0; tap < 27; ++tap)
#pragma unroll
    for (int hh = 0; hh < NH; ++hh)
      wr[tap][hh] = *reinterpret_cast<const bf16x8*>(&w[((int64_t)tap * 32 + i) * CIN + hh * 16 + kh * 8]);

  // the plane's 16-byte chunks travel global -> registers -> LDS as native
  // vectors (plain loads and stores, so the arrays stay in registers)
  auto load_plane = [&](int zz, u32x4 (&r)[PER]) __attribute__((always_inline)) {
    const int gz = zz & (G - 1);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = tid + k * NT;
      if (c < CHUNKS) {
        const int ch = c % (CIN / 8), v = c / (CIN / 8);
        const int tx = v % ZM_PX, ty = v / ZM_PX;
        const int gx = (x0 + tx - 1) & (G - 1), gy = (y0 + ty - 1) & (G - 1);
        r[k] = *reinterpret_cast<const u32x4*>(&xb[(((int64_t)gz * G + gy) * G + gx) * CIN + ch * 8]);
      }
    }
  };
  auto store_plane = [&](int slot, const u32x4 (&r)[PER]) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = tid + k * NT;
      if (c < CHUNKS) {
        const int ch = c % (CIN / 8), v = c / (CIN / 8);
        *reinterpret_cast<u32x4*>(&ring[slot * PLANE + v * CP + ch * 8]) = r[k];
      }
    }
  };
  {  // planes zs - 1, zs, zs + 1 -> slots 0, 1, 2
    u32x4 r0[PER], r1[PER];
    load_plane(zs - 1, r0);
    load_plane(zs, r1);
    store_plane(0, r0);
    load_plane(zs + 1, r0);
    store_plane(1, r1);
    store_plane(2, r0);
  }
  __syncthreads();

#pragma unroll 1
  for (int zi = 0; zi < zc; ++zi) {
    const int z = zs + zi;
    u32x4 pre[PER];
    if (zi + 1 < zc) load_plane(z + 2, pre);  // in flight during the MFMAs below
    f32x16 acc[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[r][q] = 0.0f;
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      const uint16_t* P = ring + ((zi + dz) % 3) * PLANE;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int yr = 0; yr < RPW + 2; ++yr)
#pragma unroll
          for (int hh = 0; hh < NH; ++hh) {
            // B operand: input row voxel i + dx - 1, k-slice of lane
            const bf16x8 av =
                *reinterpret_cast<const bf16x8*>(&P[((wv * RPW + yr) * ZM_PX + i + dx) * CP + hh * 16 + kh * 8]);
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
              const int dy = yr - r;
              if (dy < 0 || dy > 2) continue;
              acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wr[(dz * 3 + dy) * 3 + dx][hh], av, acc[r], 0, 0, 0);
            }
          }
    }
    // epilogue: acc[r][4 j + e] holds output channel co = 8 j + 4 kh + e of
    // voxel (x0 + i, y0 + row, z): four consecutive channels per j
    const uint16_t* C = ring + ((zi + 1) % 3) * PLANE;  // input plane z (residual centre)
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int ly = wv * RPW + r;
      const int64_t sv = ((int64_t)z * G + y0 + ly) * G + x0 + i;  // voxel index within the scene
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co0 = 8 * j + 4 * kh;
        const float4 bj = *reinterpret_cast<const float4*>(&sbias[co0]);
        float v[4] = {acc[r][4 * j] + bj.x, acc[r][4 * j + 1] + bj.y, acc[r][4 * j + 2] + bj.z,
                      acc[r][4 * j + 3] + bj.w};
        if (MODE == CONV_OUT_F32_CF) {
          float* yb = y_out + (int64_t)b * cout * G3 + sv;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (co0 + e < cout) yb[(int64_t)(co0 + e) * G3] = v[e];
        } else if (MODE == CONV_OUT_F32_CL) {
          if (cout == 32) {
            const int64_t o = ((int64_t)b * G3 + sv) * 32 + co0;
            float4 val = {v[0], v[1], v[2], v[3]};
            if (addend) {
              const float4 ad = *reinterpret_cast<const float4*>(&addend[o]);
              val.x += ad.x;
              val.y += ad.y;
              val.z += ad.z;
              val.w += ad.w;
            }
            *reinterpret_cast<float4*>(&y_out[o]) = val;
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (co0 + e < cout) {
                const int64_t o = ((int64_t)b * G3 + sv) * cout + co0 + e;
                y_out[o] = addend ? v[e] + addend[o] : v[e];
              }
          }
        } else {
          const int64_t o = ((int64_t)b * G3 + sv) * 32 + co0;
          if (z_out) {
            uint2 zp;
            zp.x = pack_bf16x2(v[0], v[1]);
            zp.y = pack_bf16x2(v[2], v[3]);
            *reinterpret_cast<uint2*>(&z_out[o]) = zp;
          }
          float hv[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) hv[e] = gelu(v[e]);
          if (MODE == CONV_RES_GELU) {
            const uint2 xr = *reinterpret_cast<const uint2*>(&C[((ly + 1) * ZM_PX + 1 + i) * CP + co0]);
            hv[0] += __uint_as_float(xr.x << 16);
            hv[1] += __uint_as_float(xr.x & 0xffff0000u);
            hv[2] += __uint_as_float(xr.y << 16);
            hv[3] += __uint_as_float(xr.y & 0xffff0000u);
          }
          uint2 hp;
          hp.x = pack_bf16x2(hv[0], hv[1]);
          hp.y = pack_bf16x2(hv[2], hv[3]);
          *reinterpret_cast<uint2*>(&h_out[o]) = hp;
        }
      }
    }
    __syncthreads();  // every wave is done with the slot of plane z - 1
    if (zi + 1 < zc) {
      store_plane(zi % 3, pre);  // plane z + 2 replaces plane z - 1
      __syncthreads();
    }
  }
}

// ------------------------------------------- 4-channel input forward
// The network's input layer has 4 real channels.  With 16 padded channels a
// 32x32x16 MFMA spends 12 of its 16 K slots on zeros; here K packs 4 taps x
// 4 channels instead (K index 4 j + ci for tap 4 g + j of K-group g), so a row
// takes 7 MFMAs (28 tap slots, the last one empty) instead of 27.  Input:
// BF16 channels-last [B][G^3][4] (8 bytes per voxel); weights [27][32][4].
// Lane (i, kh) of the A operand holds output channel i for taps 4 g + 2 kh
// and 4 g + 2 kh + 1; of the B operand, voxel i of the output row for the
// same two taps, each an 8-byte read of the shifted input voxel.  Same
// z-marching column and epilogue as k_conv3_zm (modes 1: GELU, z optional).
constexpr int IN4_RPW = 2;                   // output rows per wave (eight waves)
constexpr int IN4_THREADS = 64 * (ZM_Y / IN4_RPW);
constexpr int IN4_GROUPS = 7;                // ceil(27 / 4) K-groups of 4 taps

template <int MODE>
__global__ __launch_bounds__(IN4_THREADS) void k_conv3_in4(const uint16_t* __restrict__ x,
                                                         const uint16_t* __restrict__ w,
                                                         const float* __restrict__ bias, int G, int zc,
                                                         uint16_t* __restrict__ z_out,
                                                         uint16_t* __restrict__ h_out) {
  constexpr int PLANE = ZM_PY * ZM_PX;  // voxels per plane; 4 channels (8 bytes) each
  constexpr int PER = (PLANE + IN4_THREADS - 1) / IN4_THREADS;
  __shared__ __attribute__((aligned(16))) uint2 ring[3 * PLANE];
  __shared__ __attribute__((aligned(16))) float sbias[32];
  const int tid = threadIdx.x;
  const int ntx = G / ZM_X, nty = G / ZM_Y, ntz = G / zc;
  int t = blockIdx.x;
  const int bx = t % ntx;
  t /= ntx;
  const int by = t % nty;
  t /= nty;
  const int bz = t % ntz;
  const int b = t / ntz;
  const int x0 = bx * ZM_X, y0 = by * ZM_Y, zs = bz * zc;
  const int64_t G3 = (int64_t)G * G * G;
  const uint2* xb = reinterpret_cast<const uint2*>(x) + (int64_t)b * G3;
  const int lane = tid & 63, wv = tid >> 6;
  const int i = lane & 31, kh = lane >> 5;
  if (tid < 32) sbias[tid] = bias[tid];

  // A operand of K-group g: w[tap][co = i][0..3] for taps 4 g + 2 kh, + 1 (zero past tap 26)
  bf16x8 wr[IN4_GROUPS];
#pragma unroll
  for (int g = 0; g < IN4_GROUPS; ++g) {
    u32x4 u = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int tap = 4 * g + 2 * kh + jj;
      if (tap < 27) {
        const uint2 q = *reinterpret_cast<const uint2*>(&w[((int64_t)tap * 32 + i) * 4]);
        u[2 * jj] = q.x;
        u[2 * jj + 1] = q.y;
      }
    }
    wr[g] = __builtin_bit_cast(bf16x8, u);
  }

  auto load_plane = [&](int zz, uint2 (&r)[PER]) __attribute__((always_inline)) {
    const int gz = zz & (G - 1);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int v = tid + k * IN4_THREADS;
      if (v < PLANE) {
        const int tx = v % ZM_PX, ty = v / ZM_PX;
        const int gx = (x0 + tx - 1) & (G - 1), gy = (y0 + ty - 1) & (G - 1);
        r[k] = xb[((int64_t)gz * G + gy) * G + gx];
      }
    }
  };
  auto store_plane = [&](int slot, const uint2 (&r)[PER]) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int v = tid + k * IN4_THREADS;
      if (v < PLANE) ring[slot * PLANE + v] = r[k];
    }
  };
  {  // planes zs - 1, zs, zs + 1 -> slots 0, 1, 2
    uint2 r0[PER];
#pragma unroll 1
    for (int sl = 0; sl < 3; ++sl) {
      load_plane(zs - 1 + sl, r0);
      store_plane(sl, r0);
    }
  }
  __syncthreads();

#pragma unroll 1
  for (int zi = 0; zi < zc; ++zi) {
    const int z = zs + zi;
    uint2 pre[PER];
    if (zi + 1 < zc) load_plane(z + 2, pre);  // in flight during the MFMAs below
    f32x16 acc[IN4_RPW];
#pragma unroll
    for (int r = 0; r < IN4_RPW; ++r)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[r][q] = 0.0f;
#pragma unroll
    for (int r = 0; r < IN4_RPW; ++r) {
      const int ly = wv * IN4_RPW + r;  // output row within the column
#pragma unroll
      for (int g = 0; g < IN4_GROUPS; ++g) {
        u32x4 u = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int tap = 4 * g + 2 * kh + jj;
          if (tap < 27) {
            const int dzo = tap / 9, dyo = (tap / 3) % 3, dxo = tap % 3;
            const uint2 q = ring[((zi + dzo) % 3) * PLANE + (ly + dyo) * ZM_PX + i + dxo];
            u[2 * jj] = q.x;
            u[2 * jj + 1] = q.y;
          }
        }
        acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wr[g], __builtin_bit_cast(bf16x8, u), acc[r], 0, 0, 0);
      }
    }
    // epilogue: acc[r][4 j + e] holds output channel co = 8 j + 4 kh + e of voxel (x0 + i, y0 + row, z)
#pragma unroll
    for (int r = 0; r < IN4_RPW; ++r) {
      const int ly = wv * IN4_RPW + r;
      const int64_t sv = ((int64_t)z * G + y0 + ly) * G + x0 + i;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co0 = 8 * j + 4 * kh;
        const float4 bj = *reinterpret_cast<const float4*>(&sbias[co0]);
        const float v0 = acc[r][4 * j] + bj.x, v1 = acc[r][4 * j + 1] + bj.y, v2 = acc[r][4 * j + 2] + bj.z,
                    v3 = acc[r][4 * j + 3] + bj.w;
        const int64_t o = ((int64_t)b * G3 + sv) * 32 + co0;
        if (z_out) {
          uint2 zp;
          zp.x = pack_bf16x2(v0, v1);
          zp.y = pack_bf16x2(v2, v3);
          *reinterpret_cast<uint2*>(&z_out[o]) = zp;
        }
        uint2 hp;
        hp.x = pack_bf16x2(gelu(v0), gelu(v1));
        hp.y = pack_bf16x2(gelu(v2), gelu(v3));
        *reinterpret_cast<uint2*>(&h_out[o]) = hp;
      }
    }
    __syncthreads();  // every wave is done with the slot of plane z - 1
    if (zi + 1 < zc) {
      store_plane(zi % 3, pre);  // plane z + 2 replaces plane z - 1
      __syncthreads();
    }
  }
}

// ------------------------------------------------- narrow output forward
// The same z-marching column for a layer with at most 16 output channels and
// CIN = 32 (the network's output convolution): v_mfma_f32_16x16x32_bf16 with
// A = weights (16 output channels x 32 input channels of one tap) and B = 16
// voxels of an input row, so no matrix work is spent on padded channels.
// Eight waves: each keeps the 27 weight fragments in registers and owns one
// 16-voxel half of 4 output rows (waves 0-3 the first halves, 4-7 the second),
// so two waves share a SIMD.
typedef __attribute__((ext_vector_type(4))) float f32x4;

template <int MODE>
__global__ __launch_bounds__(ZS_THREADS) void k_conv3_zm_narrow(const uint16_t* __restrict__ x,
                                                                const uint16_t* __restrict__ w,
                                                                const float* __restrict__ bias, int G, int zc, int cout,
                                                                float* __restrict__ y_out,
                                                                const float* __restrict__ addend, float alpha,
                                                                int accumulate, const uint32_t* __restrict__ skip) {
  constexpr int CIN = 32, CP = CIN + 8;
  if (MODE == CONV_OUT_F32_CF_INTO && skip && *skip != 0u) return;  // uniform: the whole grid exits
  constexpr int PLANE = ZM_PY * ZM_PX * CP;
  constexpr int CHUNKS = ZM_PY * ZM_PX * (CIN / 8);
  constexpr int PER = (CHUNKS + ZS_THREADS - 1) / ZS_THREADS;
  __shared__ __attribute__((aligned(16))) uint16_t ring[3 * PLANE];
  __shared__ __attribute__((aligned(16))) float sbias[32];
  const int tid = threadIdx.x;
  const int ntx = G / ZM_X, nty = G / ZM_Y, ntz = G / zc;
  int t = blockIdx.x;
  const int bx = t % ntx;
  t /= ntx;
  const int by = t % nty;
  t /= nty;
  const int bz = t % ntz;
  const int b = t / ntz;
  const int x0 = bx * ZM_X, y0 = by * ZM_Y, zs = bz * zc;
  const int64_t G3 = (int64_t)G * G * G;
  const uint16_t* xb = x + (int64_t)b * G3 * CIN;
  const int lane = tid & 63, wv = tid >> 6;
  const int hf = wv >> 2, wrow = wv & 3;  // voxels 2 n16 + hf (even or odd) of rows 4 wrow .. + 3
  const int n16 = lane & 15, kq = lane >> 4;
  if (tid < 32) sbias[tid] = (MODE == CONV_OUT_F32_CL || tid >= cout) ? 0.0f : bias[tid];

  // A operand: w[tap][co = n16][ci = 8 kq .. 8 kq + 7]
  bf16x8 wr[27];
#pragma unroll
  for (int tap = 0; tap < 27; ++tap)
    wr[tap] = *reinterpret_cast<const bf16x8*>(&w[((int64_t)tap * 32 + n16) * CIN + kq * 8]);

  auto load_plane = [&](int zz, u32x4 (&r)[PER]) {
    const int gz = zz & (G - 1);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = tid + k * ZS_THREADS;
      if (c < CHUNKS) {
        const int ch = c % (CIN / 8), v = c / (CIN / 8);
        const int tx = v % ZM_PX, ty = v / ZM_PX;
        const int gx = (x0 + tx - 1) & (G - 1), gy = (y0 + ty - 1) & (G - 1);
        r[k] = *reinterpret_cast<const u32x4*>(&xb[(((int64_t)gz * G + gy) * G + gx) * CIN + ch * 8]);
      }
    }
  };
  auto store_plane = [&](int slot, const u32x4 (&r)[PER]) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = tid + k * ZS_THREADS;
      if (c < CHUNKS) {
        const int ch = c % (CIN / 8), v = c / (CIN / 8);
        *reinterpret_cast<u32x4*>(&ring[slot * PLANE + v * CP + ch * 8]) = r[k];
      }
    }
  };
  // the plane ring as in k_conv3_zm_split: one barrier per plane, after the
  // dz = 0 MFMAs, then plane z + 2 goes into the slot of plane z - 1
  u32x4 pre[PER];
  {
    u32x4 r[PER];
#pragma unroll 1
    for (int s = 0; s < 3; ++s) {
      load_plane(zs - 1 + s, r);
      store_plane(s, r);
    }
  }
  load_plane(zs + 2, pre);  // zc >= 2
  __syncthreads();

  // the into-field epilogue reads and writes this scene's planes through a
  // raw buffer view: a lane whose channel is past cout gets an offset past
  // the view (its load gives 0, its store is dropped), so the loads and
  // stores carry no branches and no wait on earlier stores
  constexpr bool CF = MODE == CONV_OUT_F32_CF_INTO || MODE == CONV_OUT_F32_CF;
  const __amdgpu_buffer_rsrc_t rfield = __builtin_amdgcn_make_buffer_rsrc(
      CF ? (void*)(y_out + (int64_t)b * cout * G3) : (void*)y_out, 0, CF ? (int)((int64_t)cout * G3 * 4) : 0,
      0x00020000);
  auto field_off = [&](int co, int64_t sv) -> uint32_t {
    return co < cout ? (uint32_t)(((int64_t)co * G3 + sv) * 4) : 0xFFFFFFF0u;
  };

#pragma unroll 1
  for (int zi = 0; zi < zc; ++zi) {
    const int z = zs + zi;
    // the field values this thread updates, fetched before the MFMAs (in place, accumulating)
    float old[ZM_RPW][4];
    if (MODE == CONV_OUT_F32_CF_INTO) {
#pragma unroll
      for (int r = 0; r < ZM_RPW; ++r) {
        const int64_t sv = ((int64_t)z * G + y0 + wrow * ZM_RPW + r) * G + x0 + 2 * n16 + hf;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          old[r][q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rfield, field_off(4 * kq + q, sv), 0, 0));
      }
    }
    f32x4 acc[ZM_RPW];
#pragma unroll
    for (int r = 0; r < ZM_RPW; ++r) acc[r] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      if (dz == 1) {
        __syncthreads();
        if (zi + 1 < zc) store_plane(zi % 3, pre);
        if (zi + 2 < zc) load_plane(z + 3, pre);
      }
      const uint16_t* P = ring + ((zi + dz) % 3) * PLANE;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int yr = 0; yr < ZM_RPW + 2; ++yr) {
          // B operand: voxel 2 n16 + hf + dx - 1 of the row, ci = 8 kq .. 8 kq + 7 (40-dword lane stride)
          const bf16x8 av =
              *reinterpret_cast<const bf16x8*>(&P[((wrow * ZM_RPW + yr) * ZM_PX + 2 * n16 + hf + dx) * CP + kq * 8]);
#pragma unroll
          for (int r = 0; r < ZM_RPW; ++r) {
            const int dy = yr - r;
            if (dy < 0 || dy > 2) continue;
            acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[(dz * 3 + dy) * 3 + dx], av, acc[r], 0, 0, 0);
          }
        }
    }
    // acc[r][q] holds output channel 4 kq + q of voxel (x0 + 2 n16 + hf, y0 + row, z)
#pragma unroll
    for (int r = 0; r < ZM_RPW; ++r) {
      const int ly = wrow * ZM_RPW + r;
      const int64_t sv = ((int64_t)z * G + y0 + ly) * G + x0 + 2 * n16 + hf;
      if (CF) {  // channel-first planes of this scene, through the buffer view
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int co = 4 * kq + q;
          const float val = acc[r][q] + sbias[co < 32 ? co : 0];
          const float out = MODE == CONV_OUT_F32_CF ? val
                            : (accumulate ? fmaf(alpha, val, old[r][q]) : alpha * val);  // uniform select
          __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(out), rfield, field_off(co, sv), 0, 0);
        }
        continue;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int co = 4 * kq + q;
        if (co < cout) {
          const float val = acc[r][q] + sbias[co];
          {
            const int64_t o = ((int64_t)b * G3 + sv) * cout + co;
            y_out[o] = addend ? val + addend[o] : val;
          }
        }
      }
    }
  }
}

// --------------------------------------- 32 -> 32 GELU layers, split outputs
// The same z-marching column with eight waves: waves 0-3 compute output
// channels 0-15 and waves 4-7 channels 16-31 of the same 16 rows, each with
// v_mfma_f32_16x16x32_bf16 (K = all 32 input channels per instruction) and
// only its half of the weights in registers (27 taps x 4 VGPRs).  Two waves
// per SIMD: one wave's epilogue (bias, GELU, residual, BF16 stores) runs
// while the other's MFMAs do.  Input fragments are read by both halves.  Also
// the FP32 channels-last 32 -> 32 layer of the backward pass.
template <int MODE>
__global__ __launch_bounds__(ZS_THREADS) void k_conv3_zm_split(const uint16_t* __restrict__ x,
                                                               const uint16_t* __restrict__ w,
                                                               const float* __restrict__ bias, int G, int zc,
                                                               uint16_t* __restrict__ z_out,
                                                               uint16_t* __restrict__ h_out,
                                                               float* __restrict__ y_out,
                                                               const float* __restrict__ addend) {
  constexpr int CIN = 32, CP = CIN + 8;
  constexpr int PLANE = ZM_PY * ZM_PX * CP;
  constexpr int CHUNKS = ZM_PY * ZM_PX * (CIN / 8);
  constexpr int PER = (CHUNKS + ZS_THREADS - 1) / ZS_THREADS;
  __shared__ __attribute__((aligned(16))) uint16_t ring[3 * PLANE];
  __shared__ __attribute__((aligned(16))) float sbias[32];
  const int tid = threadIdx.x;
  const int ntx = G / ZM_X, nty = G / ZM_Y, ntz = G / zc;
  int t = blockIdx.x;
  const int bx = t % ntx;
  t /= ntx;
  const int by = t % nty;
  t /= nty;
  const int bz = t % ntz;
  const int b = t / ntz;
  const int x0 = bx * ZM_X, y0 = by * ZM_Y, zs = bz * zc;
  const int64_t G3 = (int64_t)G * G * G;
  const uint16_t* xb = x + (int64_t)b * G3 * CIN;
  const int lane = tid & 63, wv = tid >> 6;
  const int half = wv >> 2, wrow = wv & 3;  // output channels 16 half .. + 15, rows 4 wrow .. + 3
  const int n16 = lane & 15, kq = lane >> 4;
  if (tid < 32) sbias[tid] = (MODE == CONV_OUT_F32_CL || MODE == CONV_OUT_F32_CL_DGELU) ? 0.0f : bias[tid];

  // A operand: w[tap][co = 16 half + n16][ci = 8 kq .. 8 kq + 7]
  bf16x8 wr[27];
#pragma unroll
  for (int tap = 0; tap < 27; ++tap)
    wr[tap] = *reinterpret_cast<const bf16x8*>(&w[((int64_t)tap * 32 + 16 * half + n16) * CIN + kq * 8]);

  auto load_plane = [&](int zz, u32x4 (&r)[PER]) {
    const int gz = zz & (G - 1);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = tid + k * ZS_THREADS;
      if (c < CHUNKS) {
        const int ch = c % (CIN / 8), v = c / (CIN / 8);
        const int tx = v % ZM_PX, ty = v / ZM_PX;
        const int gx = (x0 + tx - 1) & (G - 1), gy = (y0 + ty - 1) & (G - 1);
        r[k] = *reinterpret_cast<const u32x4*>(&xb[(((int64_t)gz * G + gy) * G + gx) * CIN + ch * 8]);
      }
    }
  };
  auto store_plane = [&](int slot, const u32x4 (&r)[PER]) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = tid + k * ZS_THREADS;
      if (c < CHUNKS) {
        const int ch = c % (CIN / 8), v = c / (CIN / 8);
        *reinterpret_cast<u32x4*>(&ring[slot * PLANE + v * CP + ch * 8]) = r[k];
      }
    }
  };
  // slot zi % 3 holds plane z - 1 at step zi (z = zs + zi), slot (zi + 1) % 3
  // plane z and slot (zi + 2) % 3 plane z + 1.  One barrier per plane, after
  // the dz = 0 MFMAs: every wave is then done with plane z - 1, so plane z + 2
  // (fetched into registers a step earlier) goes into its slot, and the next
  // barrier publishes it before step zi + 1 reads it as dz = 2.  A wave's
  // epilogue runs while the other waves start the next plane.
  u32x4 pre[PER];
  {
    u32x4 r[PER];
#pragma unroll 1
    for (int s = 0; s < 3; ++s) {
      load_plane(zs - 1 + s, r);
      store_plane(s, r);
    }
  }
  load_plane(zs + 2, pre);  // zc >= 2
  __syncthreads();

#pragma unroll 1
  for (int zi = 0; zi < zc; ++zi) {
    const int z = zs + zi;
    f32x4 acc[ZM_RPW][2];
#pragma unroll
    for (int r = 0; r < ZM_RPW; ++r)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) acc[r][hf] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      if (dz == 1) {
        __syncthreads();
        if (zi + 1 < zc) store_plane(zi % 3, pre);
        if (zi + 2 < zc) load_plane(z + 3, pre);
      }
      const uint16_t* P = ring + ((zi + dz) % 3) * PLANE;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx)
#pragma unroll
        for (int yr = 0; yr < ZM_RPW + 2; ++yr)
#pragma unroll
          for (int hf = 0; hf < 2; ++hf) {
            // B operand: voxel 2 n16 + hf + dx - 1 of the row, ci = 8 kq .. 8 kq + 7 (the even
            // or odd voxels: a 40-dword lane stride, so the 32 lanes served together hit
            // distinct banks)
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(
                &P[((wrow * ZM_RPW + yr) * ZM_PX + 2 * n16 + hf + dx) * CP + kq * 8]);
#pragma unroll
            for (int r = 0; r < ZM_RPW; ++r) {
              const int dy = yr - r;
              if (dy < 0 || dy > 2) continue;
              acc[r][hf] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[(dz * 3 + dy) * 3 + dx], av, acc[r][hf], 0, 0, 0);
            }
          }
    }
    // acc[r][hf][q] holds output channel 16 half + 4 kq + q of voxel (x0 + 2 n16 + hf, y0 + row, z)
    const uint16_t* C = ring + ((zi + 1) % 3) * PLANE;  // input plane z (residual centre)
    const int co0 = 16 * half + 4 * kq;
    const float4 bj = *reinterpret_cast<const float4*>(&sbias[co0]);
    // plane z of each channels-last tensor through a raw buffer view (byte
    // offsets below G^2 x 32 x 4 B <= 2^27); a null tensor gets an empty view,
    // so its loads give 0 and its stores are dropped.  All of the plane's reads
    // are issued before any is used: no branch or wait per row.
    const int64_t pe = ((int64_t)b * G + z) * G * G * 32;  // first element of plane z
    const uint32_t nb16 = (uint32_t)G * G * 64, nb32 = 2 * nb16;
    const auto view = [](const void* p, uint32_t bytes) {
      return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, p ? bytes : 0, 0x00020000);
    };
    constexpr bool CL = MODE == CONV_OUT_F32_CL || MODE == CONV_OUT_F32_CL_DGELU;
    const __amdgpu_buffer_rsrc_t rz = view(z_out ? z_out + pe : nullptr, nb16);  // z_out, or z_in (mode 4)
    const __amdgpu_buffer_rsrc_t rh = view(h_out ? h_out + pe : nullptr, nb16);  // h_out, or dz (mode 4)
    const __amdgpu_buffer_rsrc_t ry = view(CL && y_out ? y_out + pe : nullptr, nb32);
    const __amdgpu_buffer_rsrc_t ra = view(CL && addend ? addend + pe : nullptr, nb32);
    // element offset of the lane's first channel at (row r, hf): e0 + 32 hf + r G 32.
    // Loads take the row part as the scalar offset; stores take the whole
    // offset in the vector register (a store with a scalar offset whose data
    // registers are rewritten by the next instruction lost data on this GPU)
    const uint32_t e0 = ((uint32_t)(y0 + wrow * ZM_RPW) * (uint32_t)G + (uint32_t)(x0 + 2 * n16)) * 32u + (uint32_t)co0;
    const uint32_t rowe = (uint32_t)G * 32u;
    // two rows at a time: their reads are issued together, then used
#pragma unroll
    for (int rp = 0; rp < ZM_RPW; rp += 2) {
      u32x4 adr[2][2];   // addend (modes 3, 4)
      uint2 side[2][2];  // residual centre (mode 2) or z_in (mode 4)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
          const int r = rp + i, ly = wrow * ZM_RPW + r, lx = 2 * n16 + hf;
          if (CL) adr[i][hf] = __builtin_amdgcn_raw_buffer_load_b128(ra, (e0 + 32u * hf) * 4u, r * rowe * 4u, 0);
          if (MODE == CONV_OUT_F32_CL_DGELU) {
            const u32x2 zr = __builtin_amdgcn_raw_buffer_load_b64(rz, (e0 + 32u * hf) * 2u, r * rowe * 2u, 0);
            side[i][hf] = make_uint2(zr[0], zr[1]);
          }
          if (MODE == CONV_RES_GELU)
            side[i][hf] = *reinterpret_cast<const uint2*>(&C[((ly + 1) * ZM_PX + 1 + lx) * CP + co0]);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
          const int r = rp + i;
          const uint32_t ev = e0 + 32u * hf + r * rowe;
          const float v0 = acc[r][hf][0] + bj.x, v1 = acc[r][hf][1] + bj.y, v2 = acc[r][hf][2] + bj.z,
                      v3 = acc[r][hf][3] + bj.w;
          if (CL) {  // FP32 channels-last, 32 channels (+ addend: uniform select, no branch)
            const u32x4 ad = adr[i][hf];
            float4 val = {v0, v1, v2, v3};
            if (addend) {
              val.x += __uint_as_float(ad[0]);
              val.y += __uint_as_float(ad[1]);
              val.z += __uint_as_float(ad[2]);
              val.w += __uint_as_float(ad[3]);
            }
            const u32x4 yv = {__float_as_uint(val.x), __float_as_uint(val.y), __float_as_uint(val.z),
                              __float_as_uint(val.w)};
            __builtin_amdgcn_raw_buffer_store_b128(yv, ry, ev * 4u, 0, 0);
            if (MODE == CONV_OUT_F32_CL_DGELU) {  // dz = y * gelu'(z_in), BF16 (k_gelu_bwd's arithmetic)
              const uint2 zr = side[i][hf];
              const u32x2 dp = {(uint32_t)f2bf(val.x * gelu_grad(__uint_as_float(zr.x << 16))) |
                                    ((uint32_t)f2bf(val.y * gelu_grad(__uint_as_float(zr.x & 0xffff0000u))) << 16),
                                (uint32_t)f2bf(val.z * gelu_grad(__uint_as_float(zr.y << 16))) |
                                    ((uint32_t)f2bf(val.w * gelu_grad(__uint_as_float(zr.y & 0xffff0000u))) << 16)};
              __builtin_amdgcn_raw_buffer_store_b64(dp, rh, ev * 2u, 0, 0);
            }
            continue;
          }
          const u32x2 zp = {pack_bf16x2(v0, v1), pack_bf16x2(v2, v3)};
          __builtin_amdgcn_raw_buffer_store_b64(zp, rz, ev * 2u, 0, 0);
          float h0 = gelu(v0), h1 = gelu(v1), h2 = gelu(v2), h3 = gelu(v3);
          if (MODE == CONV_RES_GELU) {
            const uint2 xr = side[i][hf];
            h0 += __uint_as_float(xr.x << 16);
            h1 += __uint_as_float(xr.x & 0xffff0000u);
            h2 += __uint_as_float(xr.y << 16);
            h3 += __uint_as_float(xr.y & 0xffff0000u);
          }
          const u32x2 hp = {pack_bf16x2(h0, h1), pack_bf16x2(h2, h3)};
          __builtin_amdgcn_raw_buffer_store_b64(hp, rh, ev * 2u, 0, 0);
        }
    }
  }
}

// ------------------------------------------------------------ weight grad
// dW[tap][co][ci] = sum_v dz[v][co] * x[v + o_tap][ci]; db[co] = sum_v dz[v][co].
// Persistent workgroups walk 32x4x2-voxel tiles in a fixed order and write
// one partial per workgroup; k_sum_partials adds them in workgroup order
// (deterministic).  In LDS the tiles are channel-major, so eight consecutive
// voxels of one channel form a 16-byte MFMA fragment: A = dz^T (co x voxels),
// B = x shifted by the tap (voxels x ci).  x is staged twice, the second copy
// one voxel to the left, so the B fragment of a dx = 0 or dx = 1 tap is one
// aligned 16-byte read and that of a dx = 2 tap four 4-byte reads.  Channel
// strides are padded to 36 (mod 64) banks, so the 16 lanes of a 16-byte read
// hit distinct banks.  Wave w accumulates taps w, w+4, ... (compile-time
// tap lists per wave).
constexpr int DW_TAPS_PER_WAVE = 7;  // ceil(27 / 4)
constexpr int XS = HX + 6;           // x stride of a row of the channel-major input copies (40)
constexpr int DS = CT_X + 8;         // x stride of a row of the channel-major dz tile (40)
constexpr int XCS = HZ * HY * XS + 8;    // per-channel stride of an input copy (968 = 484 dwords)
constexpr int DCS = CT_Z * CT_Y * DS + 8;  // per-channel stride of the dz tile (328 = 164 dwords)
static_assert((XCS / 2) % 64 == 36 && (DCS / 2) % 64 == 36, "conflict-free 16-byte fragment reads");
static_assert(HX % 2 == 0, "x halo rows load in voxel pairs");

// 16 input channels: one 32-column MFMA serves two taps of the same dx (lanes
// 0-15 the first tap's channels, 16-31 the second's), so the 27 taps are 12
// pairs and 3 singles, 15 MFMA chains (waves take chains w, w + 4, ...).
// Chain c: taps wg16_tap(c, 0) and wg16_tap(c, 1) (-1: none).
constexpr int WG16_CHAINS = 15, WG16_PER_WAVE = 4;
__host__ __device__ constexpr int wg16_tap(int c, int second) {
  // taps of one dx in (dz, dy) order: m -> 9 (m / 3) + 3 (m % 3) + dx
  return c < 12 ? 9 * ((2 * (c % 4) + second) / 3) + 3 * ((2 * (c % 4) + second) % 3) + c / 4
                : (second ? -1 : 9 * (8 / 3) + 3 * (8 % 3) + (c - 12));
}

template <int CIN>
__global__ __launch_bounds__(CONV_THREADS) void k_conv3_wgrad(const uint16_t* __restrict__ x,
                                                              const uint16_t* __restrict__ dz, int B, int G,
                                                              int64_t ntiles, float* __restrict__ part) {
  constexpr int NB = CIN / 32 > 0 ? CIN / 32 : 1;
  constexpr int XCH = HZ * HY * (HX / 2) * (CIN / 8);      // voxel pairs x 16-byte chunks of the x tile
  constexpr int XPER = (XCH + CONV_THREADS - 1) / CONV_THREADS;
  constexpr int DCH = CT_Z * CT_Y * CT_X * 4;              // 16-byte chunks of the dz tile
  constexpr int DPER = DCH / CONV_THREADS;
  __shared__ __attribute__((aligned(16))) uint16_t xl[2][CIN * XCS];  // [0]: voxel tx at tx; [1]: at tx - 1
  __shared__ __attribute__((aligned(16))) uint16_t dl[32 * DCS];
  __shared__ float bred[4][32];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i = lane & 31, kh = lane >> 5;
  const int ntx = G / CT_X, nty = G / CT_Y, ntz = G / CT_Z;
  const int64_t G3 = (int64_t)G * G * G;
  constexpr int NCH = CIN == 16 ? WG16_PER_WAVE : DW_TAPS_PER_WAVE;  // MFMA chains per wave
  f32x16 acc[NCH][NB];
#pragma unroll
  for (int k = 0; k < NCH; ++k)
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[k][n][q] = 0.0f;
  // bias: db[co] = sum_v dz[v][co], as MFMAs of dz^T with a ones operand,
  // a quarter of the K steps of each tile per wave
  f32x16 accb;
#pragma unroll
  for (int q = 0; q < 16; ++q) accb[q] = 0.0f;
  bf16x8 ones;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;

  // the next tile's chunks travel global -> registers (in flight during the
  // current tile's MFMAs) -> LDS, transposed to channel-major; x in pairs of
  // voxels adjacent in x, so the unshifted copy takes 4-byte stores
  u32x4 rx[XPER][2], rd[DPER];
  auto load_tile = [&](int64_t tile) {
    int64_t t = tile;
    const int bx = (int)(t % ntx);
    t /= ntx;
    const int by = (int)(t % nty);
    t /= nty;
    const int bz = (int)(t % ntz);
    const int b = (int)(t / ntz);
    const int x0 = bx * CT_X, y0 = by * CT_Y, z0 = bz * CT_Z;
#pragma unroll
    for (int k = 0; k < XPER; ++k) {
      const int c = tid + k * CONV_THREADS;
      if (c < XCH) {
        const int ch = c % (CIN / 8);
        int v = c / (CIN / 8);
        const int tp = v % (HX / 2);
        v /= HX / 2;
        const int ty = v % HY;
        const int tz = v / HY;
        const int gy = (y0 + ty - 1) & (G - 1), gz = (z0 + tz - 1) & (G - 1);
        const int64_t row = (int64_t)b * G3 + ((int64_t)gz * G + gy) * G;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int gx = (x0 + 2 * tp + u - 1) & (G - 1);
          rx[k][u] = *reinterpret_cast<const u32x4*>(&x[(row + gx) * CIN + ch * 8]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < DPER; ++k) {
      const int c = tid + k * CONV_THREADS;
      const int ch = c % 4;
      int v = c / 4;
      const int tx = v % CT_X;
      v /= CT_X;
      const int ty = v % CT_Y;
      const int tz = v / CT_Y;
      const int64_t gv = ((int64_t)(z0 + tz) * G + (y0 + ty)) * G + x0 + tx;
      rd[k] = *reinterpret_cast<const u32x4*>(&dz[((int64_t)b * G3 + gv) * 32 + ch * 8]);
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int k = 0; k < XPER; ++k) {
      const int c = tid + k * CONV_THREADS;
      if (c < XCH) {
        const int ch = c % (CIN / 8);
        int v = c / (CIN / 8);
        const int tp = v % (HX / 2);
        v /= HX / 2;
        const int ty = v % HY;
        const int tz = v / HY;
        const int o = (tz * HY + ty) * XS + 2 * tp;  // voxel 2 tp of the row (4-byte aligned)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t a0 = rx[k][0][e], a1 = rx[k][1][e];  // channels 2e (low half), 2e + 1 (high half)
          const int c0 = (ch * 8 + 2 * e) * XCS, c1 = c0 + XCS;
          *reinterpret_cast<uint32_t*>(&xl[0][c0 + o]) = (a0 & 0xffffu) | (a1 << 16);
          *reinterpret_cast<uint32_t*>(&xl[0][c1 + o]) = (a0 >> 16) | (a1 & 0xffff0000u);
          // copy 1 holds voxel tx at tx - 1: voxel 2 tp + 1 at 2 tp, voxel 2 tp at 2 tp - 1
          xl[1][c0 + o] = (uint16_t)(a1 & 0xffffu);
          xl[1][c1 + o] = (uint16_t)(a1 >> 16);
          if (tp > 0) {
            xl[1][c0 + o - 1] = (uint16_t)(a0 & 0xffffu);
            xl[1][c1 + o - 1] = (uint16_t)(a0 >> 16);
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < DPER; ++k) {
      const int c = tid + k * CONV_THREADS;
      const int ch = c % 4;
      int v = c / 4;
      const int tx = v % CT_X;
      v /= CT_X;
      const int ty = v % CT_Y;
      const int tz = v / CT_Y;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        dl[(ch * 8 + 2 * e) * DCS + (tz * CT_Y + ty) * DS + tx] = (uint16_t)(rd[k][e] & 0xffffu);
        dl[(ch * 8 + 2 * e + 1) * DCS + (tz * CT_Y + ty) * DS + tx] = (uint16_t)(rd[k][e] >> 16);
      }
    }
  };

  // K loop of one tile for wave WV: 8 rows x 32 voxels = 16 steps of 16
  // voxels (two per row); each step's dz fragment feeds the wave's 7 taps,
  // 7 independent MFMA chains
  auto mfma_tile = [&](auto wvc) __attribute__((always_inline)) {
    constexpr int WV = decltype(wvc)::value;
#pragma unroll 2
    for (int st = 0; st < 2 * CT_Z * CT_Y; ++st) {
      const int r = st >> 1, half = st & 1;
      const int rz = r / CT_Y, ry = r % CT_Y;
      const int vb = half * 16 + kh * 8;  // first of this lane's 8 voxels
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(&dl[i * DCS + (rz * CT_Y + ry) * DS + vb]);
      if ((st & 3) == WV) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, ones, accb, 0, 0, 0);
      if constexpr (CIN == 16) {
#pragma unroll
        for (int k = 0; k < WG16_PER_WAVE; ++k) {
          const int c = WV + 4 * k;
          if (c < WG16_CHAINS) {
            const int dxo = c < 12 ? c / 4 : c - 12;  // shared by the chain's two taps
            const int tp = (i < 16) ? wg16_tap(c, 0) : wg16_tap(c, 1);
            const int ci = i & 15;
            bf16x8 bv;
            if (tp >= 0) {
              const int dzo = tp / 9, dyo = (tp / 3) % 3;
              const int row = ((rz + dzo) * HY + ry + dyo) * XS + vb;
              if (dxo == 2) {
                const uint32_t* p = reinterpret_cast<const uint32_t*>(&xl[0][ci * XCS + row + 2]);
                const u32x4 u = {p[0], p[1], p[2], p[3]};
                bv = __builtin_bit_cast(bf16x8, u);
              } else {
                bv = *reinterpret_cast<const bf16x8*>(&xl[dxo][ci * XCS + row]);
              }
            } else {
#pragma unroll
              for (int e = 0; e < 8; ++e) bv[e] = (__bf16)0.0f;
            }
            acc[k][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[k][0], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < (CIN == 16 ? 0 : DW_TAPS_PER_WAVE); ++k) {
        const int tap = WV + 4 * k;
        if (tap < 27) {
          const int dzo = tap / 9, dyo = (tap / 3) % 3, dxo = tap % 3;
          const int row = ((rz + dzo) * HY + ry + dyo) * XS + vb;
#pragma unroll
          for (int n = 0; n < NB; ++n) {
            const int ci = n * 32 + i;
            bf16x8 bv;
            if (ci < CIN) {
              if (dxo == 2) {  // voxels vb + 2 .. vb + 9 of the unshifted copy: 4-byte aligned
                const uint32_t* p = reinterpret_cast<const uint32_t*>(&xl[0][ci * XCS + row + 2]);
                const u32x4 u = {p[0], p[1], p[2], p[3]};
                bv = __builtin_bit_cast(bf16x8, u);
              } else {
                bv = *reinterpret_cast<const bf16x8*>(&xl[dxo][ci * XCS + row]);
              }
            } else {
#pragma unroll
              for (int e = 0; e < 8; ++e) bv[e] = (__bf16)0.0f;
            }
            acc[k][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[k][n], 0, 0, 0);
          }
        }
      }
    }
  };

  int64_t tile = blockIdx.x;
  if (tile < ntiles) load_tile(tile);
  for (; tile < ntiles; tile += gridDim.x) {
    __syncthreads();  // the previous tile's MFMAs are done with LDS
    store_tile();
    __syncthreads();
    if (tile + gridDim.x < ntiles) load_tile(tile + gridDim.x);
    switch (wv) {  // wave-uniform: each wave's tap list is compile-time
      case 0: mfma_tile(std::integral_constant<int, 0>{}); break;
      case 1: mfma_tile(std::integral_constant<int, 1>{}); break;
      case 2: mfma_tile(std::integral_constant<int, 2>{}); break;
      default: mfma_tile(std::integral_constant<int, 3>{}); break;
    }
  }
  // bias partial of the wave: accb[q] holds (co = 8 (q/4) + 4 kh + q%4, any column): lanes 0 and 32
  if (i == 0)
#pragma unroll
    for (int q = 0; q < 16; ++q) bred[wv][8 * (q / 4) + 4 * kh + (q % 4)] = accb[q];
  __syncthreads();
  // partial layout: [wg][27][32 co][CIN] then [wg] bias block of 32
  float* pw = part + (int64_t)blockIdx.x * (27 * 32 * CIN + 32);
  if constexpr (CIN == 16) {
#pragma unroll
    for (int k = 0; k < WG16_PER_WAVE; ++k) {
      const int c = wv + 4 * k;
      const int tp = c < WG16_CHAINS ? ((i < 16) ? wg16_tap(c, 0) : wg16_tap(c, 1)) : -1;
      if (tp >= 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int co = 8 * (q / 4) + 4 * kh + (q % 4);
          pw[(tp * 32 + co) * CIN + (i & 15)] = acc[k][0][q];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < (CIN == 16 ? 0 : DW_TAPS_PER_WAVE); ++k) {
    const int tap = wv + 4 * k;
    if (tap >= 27) break;
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      const int ci = n * 32 + i;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int co = 8 * (q / 4) + 4 * kh + (q % 4);
        if (ci < CIN) pw[(tap * 32 + co) * CIN + ci] = acc[k][n][q];
      }
    }
  }
  if (tid < 32) pw[27 * 32 * CIN + tid] = ((bred[0][tid] + bred[1][tid]) + bred[2][tid]) + bred[3][tid];
}

// Weight gradient of the 32-channel layers, streamed along z.  A task is a
// column of 32x2 voxels (x, y) over WZ_ZC planes of one scene; the workgroup
// walks it plane by plane: x planes (with their one-row y halo) pass through a
// three-plane LDS ring, so each is staged once per task instead of once per
// tap row, and dz through a one-plane buffer.  A plane is 4 K-steps of 16
// voxels; each step's dz fragment feeds the wave's taps.  Every fragment is
// one aligned 16-byte read: x is held twice (copy 0 at voxel x0 + p - 1,
// copy 1 at x0 + p), dz twice (copy 0 at x0 + p, copy 1 at x0 + p - 1).  A
// dx = 0 or 1 tap pairs dz copy 0 with x copy 0 or 1; a dx = 2 tap sums its
// column over voxels x0 - 1 .. x0 + 30 instead (every voxel is still summed
// once over the periodic grid) and pairs dz copy 1 with x copy 1.  60 KB of
// LDS: two workgroups per CU.  Persistent workgroups take tasks s, s + nwg,
// ... (s groups each XCD's workgroups on adjacent columns) and write one
// partial each, summed in workgroup order by k_sum_partials.
constexpr int WZ_X = 32, WZ_Y = 2;
constexpr int WZ_HY = WZ_Y + 2;
constexpr int WZ_XS = WZ_X;             // row stride of a staged x plane (32 voxels)
constexpr int WZ_SLOT = WZ_HY * WZ_XS;  // one plane of one channel (128)
constexpr int WZ_XCS = 392;             // channel stride of an x copy: 3 planes + pad (196 dwords)
constexpr int WZ_DS = WZ_X;             // row stride of the dz plane (32)
constexpr int WZ_DCS = 72;              // channel stride of a dz copy: 2 rows + pad (36 dwords)
constexpr int WZ_ZC = 64;               // planes per task (fewer when G is smaller)
// 32 lanes (consecutive channels) of a 16-byte read are served together: a
// channel stride of 4 x odd dwords (mod 128) puts them on distinct banks
static_assert(WZ_XCS >= 3 * WZ_SLOT && (WZ_XCS / 2) % 8 == 4 && (WZ_DCS / 2) % 8 == 4 && WZ_DCS >= WZ_Y * WZ_DS,
              "16-byte fragment reads of 32 consecutive channels hit distinct banks");

__global__ __launch_bounds__(CONV_THREADS) void k_conv3_wgrad_z(const uint16_t* __restrict__ x,
                                                                const uint16_t* __restrict__ dz, int B, int G,
                                                                int zc, int64_t ntasks, float* __restrict__ part) {
  static_assert(CONV_THREADS == 256 && WZ_HY == 4 && WZ_Y == 2 && WZ_X == 32,
                "staging threads: 16 channel pairs x 4 runs of 8 voxels x 4 x rows (2 dz rows)");
  static_assert(WZ_XCS % 8 == 0 && WZ_SLOT % 8 == 0 && WZ_XS % 8 == 0 && WZ_DCS % 8 == 0 && WZ_DS % 8 == 0,
                "16-byte aligned runs");
  __shared__ __attribute__((aligned(16))) uint16_t xl[2][32 * WZ_XCS];  // [0]: voxel x0 + p - 1 at p; [1]: x0 + p
  __shared__ __attribute__((aligned(16))) uint16_t dl[2][32 * WZ_DCS];  // [0]: voxel x0 + p at p; [1]: x0 + p - 1
  __shared__ float bred[4][32];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i = lane & 31, kh = lane >> 5;
  const int ntx = G / WZ_X, nty = G / WZ_Y, ntz = G / zc;
  const int64_t G3 = (int64_t)G * G * G;
  f32x16 acc[DW_TAPS_PER_WAVE];
#pragma unroll
  for (int k = 0; k < DW_TAPS_PER_WAVE; ++k)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[k][q] = 0.0f;
  f32x16 accb;
#pragma unroll
  for (int q = 0; q < 16; ++q) accb[q] = 0.0f;
  bf16x8 ones;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;

  // one x plane (rows y0-1 .. y0+2, voxels x0-1 .. x0+31) and one dz plane
  // (rows y0, y0+1, voxels x0-1 .. x0+31) travel global -> registers -> LDS
  // (channel-major).  Thread (cp, vg, ty) holds channels 2cp, 2cp+1 of voxels
  // x0 + 8vg - 1 .. x0 + 8vg + 7 as 4-byte words; byte permutes pair
  // neighbouring voxels, so each copy is written as 16-byte runs of 8 voxels
  // of one channel at position 8vg.  Lanes cp = 0..7 of a 16-byte store land
  // on distinct banks.  Threads 0..127 stage the dz plane the same way (row r
  // for ty = r).
  uint32_t rx[9], rd[9];
  int cb = 0, cx0 = 0, cy0 = 0;  // the task's scene and column origin
  const int xcp = tid & 15, xvg = (tid >> 4) & 3, xty = tid >> 6;
  auto load_x = [&](int zp) {
    const int gz = zp & (G - 1);
    const int gy = (cy0 + xty - 1) & (G - 1);
    const uint32_t* row =
        reinterpret_cast<const uint32_t*>(x) + ((int64_t)cb * G3 + ((int64_t)gz * G + gy) * G) * 16 + xcp;
#pragma unroll
    for (int u = 0; u < 9; ++u) {
      const int gx = (cx0 + 8 * xvg + u - 1) & (G - 1);
      rx[u] = row[(int64_t)gx * 16];
    }
  };
  // 16-byte runs of channels 2cp (lo) and 2cp + 1 (hi): words w[j0 .. j0 + 8)
  auto run_lo = [](const uint32_t (&w)[9], int j0) {
    u32x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (w[j0 + 2 * j] & 0xffffu) | (w[j0 + 2 * j + 1] << 16);
    return v;
  };
  auto run_hi = [](const uint32_t (&w)[9], int j0) {
    u32x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (w[j0 + 2 * j] >> 16) | (w[j0 + 2 * j + 1] & 0xffff0000u);
    return v;
  };
  auto store_x = [&](int slot) {
    const int o = 2 * xcp * WZ_XCS + slot * WZ_SLOT + xty * WZ_XS + 8 * xvg;
    *reinterpret_cast<u32x4*>(&xl[0][o]) = run_lo(rx, 0);
    *reinterpret_cast<u32x4*>(&xl[0][o + WZ_XCS]) = run_hi(rx, 0);
    *reinterpret_cast<u32x4*>(&xl[1][o]) = run_lo(rx, 1);
    *reinterpret_cast<u32x4*>(&xl[1][o + WZ_XCS]) = run_hi(rx, 1);
  };
  const int dr = (tid >> 6) & 1;
  auto load_dz = [&](int zp) {
    if (tid < 128) {
      const uint32_t* row =
          reinterpret_cast<const uint32_t*>(dz) + ((int64_t)cb * G3 + ((int64_t)zp * G + (cy0 + dr)) * G) * 16 + xcp;
#pragma unroll
      for (int u = 0; u < 9; ++u) {
        const int gx = (cx0 + 8 * xvg + u - 1) & (G - 1);
        rd[u] = row[(int64_t)gx * 16];
      }
    }
  };
  auto store_dz = [&]() {
    if (tid < 128) {
      const int o = 2 * xcp * WZ_DCS + dr * WZ_DS + 8 * xvg;
      *reinterpret_cast<u32x4*>(&dl[0][o]) = run_lo(rd, 1);
      *reinterpret_cast<u32x4*>(&dl[0][o + WZ_DCS]) = run_hi(rd, 1);
      *reinterpret_cast<u32x4*>(&dl[1][o]) = run_lo(rd, 0);
      *reinterpret_cast<u32x4*>(&dl[1][o + WZ_DCS]) = run_hi(rd, 0);
    }
  };
  // the MFMAs of one plane for wave WV: x planes z-1, z, z+1 in ring slots
  // (s0, s0+1, s0+2) mod 3; 4 K-steps (2 rows x 2 halves of 16 voxels)
  auto mfma_plane = [&](auto wvc, int s0) __attribute__((always_inline)) {
    constexpr int WV = decltype(wvc)::value;
#pragma unroll
    for (int st = 0; st < 2 * WZ_Y; ++st) {
      const int r = st >> 1, half = st & 1;
      const int vb = half * 16 + kh * 8;
      const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&dl[0][i * WZ_DCS + r * WZ_DS + vb]);
      const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&dl[1][i * WZ_DCS + r * WZ_DS + vb]);
      if ((st & 3) == WV) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, ones, accb, 0, 0, 0);
#pragma unroll
      for (int k = 0; k < DW_TAPS_PER_WAVE; ++k) {
        const int tap = WV + 4 * k;
        if (tap < 27) {
          const int dzo = tap / 9, dyo = (tap / 3) % 3, dxo = tap % 3;
          int sl = s0 + dzo;
          sl = sl >= 3 ? sl - 3 : sl;
          const int o = i * WZ_XCS + sl * WZ_SLOT + (r + dyo) * WZ_XS + vb;
          const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&xl[dxo == 0 ? 0 : 1][o]);
          acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dxo == 2 ? a1 : a0, bv, acc[k], 0, 0, 0);
        }
      }
    }
  };

  const int64_t nwg = gridDim.x;
  const int64_t s = (nwg % 8 == 0) ? (int64_t)(blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8 : (int64_t)blockIdx.x;
  for (int64_t task = s; task < ntasks; task += nwg) {
    int64_t t = task;
    const int bx = (int)(t % ntx);
    t /= ntx;
    const int by = (int)(t % nty);
    t /= nty;
    const int bz = (int)(t % ntz);
    cb = (int)(t / ntz);
    cx0 = bx * WZ_X;
    cy0 = by * WZ_Y;
    const int z0 = bz * zc;
    // prologue: x planes z0 - 1 and z0 into slots 0 and 1
    load_x(z0 - 1);
    __syncthreads();  // the previous task's MFMAs are done with LDS
    store_x(0);
    load_x(z0);
    store_x(1);
    load_x(z0 + 1);
    load_dz(z0);
    for (int zi = 0; zi < zc; ++zi) {
      __syncthreads();  // the previous plane's MFMAs are done with slot (zi + 2) % 3 and the dz plane
      store_x((zi + 2) % 3);
      store_dz();
      __syncthreads();
      if (zi + 1 < zc) {
        load_x(z0 + zi + 2);
        load_dz(z0 + zi + 1);
      }
      const int s0 = zi % 3;
      switch (wv) {  // wave-uniform: each wave's tap list is compile-time
        case 0: mfma_plane(std::integral_constant<int, 0>{}, s0); break;
        case 1: mfma_plane(std::integral_constant<int, 1>{}, s0); break;
        case 2: mfma_plane(std::integral_constant<int, 2>{}, s0); break;
        default: mfma_plane(std::integral_constant<int, 3>{}, s0); break;
      }
    }
  }
  if (i == 0)
#pragma unroll
    for (int q = 0; q < 16; ++q) bred[wv][8 * (q / 4) + 4 * kh + (q % 4)] = accb[q];
  __syncthreads();
  float* pw = part + (int64_t)blockIdx.x * (27 * 32 * 32 + 32);
#pragma unroll
  for (int k = 0; k < DW_TAPS_PER_WAVE; ++k) {
    const int tap = wv + 4 * k;
    if (tap >= 27) break;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int co = 8 * (q / 4) + 4 * kh + (q % 4);
      pw[(tap * 32 + co) * 32 + i] = acc[k][q];
    }
  }
  if (tid < 32) pw[27 * 32 * 32 + tid] = ((bred[0][tid] + bred[1][tid]) + bred[2][tid]) + bred[3][tid];
}

// Weight gradient of the 4-channel input layer: dw[tap][co][ci], ci < 4.  The
// MFMA's 32 columns are (tap, channel) pairs, 8 taps x 4 channels, so the 27
// taps take 5 MFMAs per K-step of 16 voxels: three for the 18 taps with
// dx = 0 or 1 (dz copy 0 against x copy dx) and two for the 9 taps with
// dx = 2 (dz copy 1 against x copy 1, the column of k_conv3_wgrad_z shifted
// by one voxel); one more against a ones operand sums the bias.  Columns past
// a group's taps read a zero fragment.  A task is a column of 32x8 voxels
// over WZ_ZC planes of one scene; wave w takes rows 2w and 2w + 1 of each
// plane (4 K-steps) and writes a partial of its own.  The input is 8 bytes a
// voxel and dz 64: the kernel streams dz.
constexpr int W4_Y = 8;                    // dz rows per task
constexpr int W4_HY = W4_Y + 2;            // x rows (one-row halo each side)
constexpr int W4_SLOT = W4_HY * 32;        // one x plane of one channel (320)
constexpr int W4_XCS = 968;                // channel stride of an x copy: 3 planes + pad (484 dwords)
constexpr int W4_DCS = 264;                // channel stride of a dz copy: 8 rows + pad (132 dwords)
constexpr int W4_PART = 27 * 32 * 4 + 32;  // floats of one partial
static_assert(W4_XCS >= 3 * W4_SLOT && (W4_XCS / 2) % 8 == 4 && (W4_DCS / 2) % 8 == 4 && W4_DCS >= W4_Y * 32,
              "16-byte fragment reads of 32 consecutive channels hit distinct banks");

__global__ __launch_bounds__(CONV_THREADS) void k_conv3_wgrad_in4(const uint16_t* __restrict__ x,
                                                                  const uint16_t* __restrict__ dz, int B, int G,
                                                                  int zc, int64_t ntasks, float* __restrict__ part) {
  static_assert(CONV_THREADS == 256, "staging threads: x 2 channel pairs x 4 runs x 10 rows; dz 16 x 4 x 4 (2 rows)");
  __shared__ __attribute__((aligned(16))) uint16_t xl[2][4 * W4_XCS];   // [0]: voxel x0 + p - 1 at p; [1]: x0 + p
  __shared__ __attribute__((aligned(16))) uint16_t dl[2][32 * W4_DCS];  // [0]: voxel x0 + p at p; [1]: x0 + p - 1
  __shared__ __attribute__((aligned(16))) uint16_t zero8[8];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int i = lane & 31, kh = lane >> 5;
  const int ntx = G / 32, nty = G / W4_Y, ntz = G / zc;
  const int64_t G3 = (int64_t)G * G * G;
  if (tid < 8) zero8[tid] = 0;
  f32x16 acc[5], accb;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    accb[q] = 0.0f;
#pragma unroll
    for (int g = 0; g < 5; ++g) acc[g][q] = 0.0f;
  }
  bf16x8 ones;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;
  // this lane's column in group g: channel ci = i & 3 of tap (dzo, dyo, dx)
  const int ci = i & 3, tc = i >> 2;
  bool cval[5];
  int ccopy[5], cdzo[5], cdyo[5], ctap[5];
#pragma unroll
  for (int g = 0; g < 5; ++g) {
    if (g < 3) {
      const int t = 8 * g + tc;  // the 18 taps with dx = 0, 1 in (dz, dy, dx) order
      cval[g] = t < 18;
      cdzo[g] = t / 6;
      cdyo[g] = (t / 2) % 3;
      ccopy[g] = t % 2;
      ctap[g] = cdzo[g] * 9 + cdyo[g] * 3 + ccopy[g];
    } else {
      const int t = 8 * (g - 3) + tc;  // the 9 taps with dx = 2
      cval[g] = t < 9;
      cdzo[g] = t / 3;
      cdyo[g] = t % 3;
      ccopy[g] = 1;
      ctap[g] = cdzo[g] * 9 + cdyo[g] * 3 + 2;
    }
  }

  // x plane: rows y0-1 .. y0+8, voxels x0-1 .. x0+31, 4 channels; thread
  // (cp, vg, ty) < 80 holds channels 2cp, 2cp+1 of voxels x0+8vg-1 .. x0+8vg+7.
  // dz plane: rows y0 .. y0+7, thread (cp, vg, q) rows 2q, 2q+1, channels 2cp,
  // 2cp+1.  Both written as 16-byte channel runs (byte permutes of voxel pairs).
  uint32_t rx[9], rd[2][9];
  int cb = 0, cx0 = 0, cy0 = 0;
  const int xcp = tid & 1, xvg = (tid >> 1) & 3, xty = tid >> 3;
  const int dcp = tid & 15, dvg = (tid >> 4) & 3, dq = tid >> 6;
  auto run_lo = [](const uint32_t (&w)[9], int j0) {
    u32x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (w[j0 + 2 * j] & 0xffffu) | (w[j0 + 2 * j + 1] << 16);
    return v;
  };
  auto run_hi = [](const uint32_t (&w)[9], int j0) {
    u32x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (w[j0 + 2 * j] >> 16) | (w[j0 + 2 * j + 1] & 0xffff0000u);
    return v;
  };
  auto load_x = [&](int zp) {
    if (tid < 2 * 4 * W4_HY) {
      const int gz = zp & (G - 1);
      const int gy = (cy0 + xty - 1) & (G - 1);
      const uint32_t* row =
          reinterpret_cast<const uint32_t*>(x) + ((int64_t)cb * G3 + ((int64_t)gz * G + gy) * G) * 2 + xcp;
#pragma unroll
      for (int u = 0; u < 9; ++u) {
        const int gx = (cx0 + 8 * xvg + u - 1) & (G - 1);
        rx[u] = row[(int64_t)gx * 2];
      }
    }
  };
  auto store_x = [&](int slot) {
    if (tid < 2 * 4 * W4_HY) {
      const int o = 2 * xcp * W4_XCS + slot * W4_SLOT + xty * 32 + 8 * xvg;
      *reinterpret_cast<u32x4*>(&xl[0][o]) = run_lo(rx, 0);
      *reinterpret_cast<u32x4*>(&xl[0][o + W4_XCS]) = run_hi(rx, 0);
      *reinterpret_cast<u32x4*>(&xl[1][o]) = run_lo(rx, 1);
      *reinterpret_cast<u32x4*>(&xl[1][o + W4_XCS]) = run_hi(rx, 1);
    }
  };
  auto load_dz = [&](int zp) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint32_t* row = reinterpret_cast<const uint32_t*>(dz) +
                            ((int64_t)cb * G3 + ((int64_t)zp * G + (cy0 + 2 * dq + k)) * G) * 16 + dcp;
#pragma unroll
      for (int u = 0; u < 9; ++u) {
        const int gx = (cx0 + 8 * dvg + u - 1) & (G - 1);
        rd[k][u] = row[(int64_t)gx * 16];
      }
    }
  };
  auto store_dz = [&]() {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int o = 2 * dcp * W4_DCS + (2 * dq + k) * 32 + 8 * dvg;
      *reinterpret_cast<u32x4*>(&dl[0][o]) = run_lo(rd[k], 1);
      *reinterpret_cast<u32x4*>(&dl[0][o + W4_DCS]) = run_hi(rd[k], 1);
      *reinterpret_cast<u32x4*>(&dl[1][o]) = run_lo(rd[k], 0);
      *reinterpret_cast<u32x4*>(&dl[1][o + W4_DCS]) = run_hi(rd[k], 0);
    }
  };
  // the MFMAs of one plane for this wave: x planes z-1, z, z+1 in ring slots
  // (s0, s0+1, s0+2) mod 3; rows 2 wv and 2 wv + 1, two halves of 16 voxels
  auto mfma_plane = [&](int s0) __attribute__((always_inline)) {
    int xo[5];  // element offset of each group's fragment without the row and voxel terms
#pragma unroll
    for (int g = 0; g < 5; ++g) {
      int sl = s0 + cdzo[g];
      sl = sl >= 3 ? sl - 3 : sl;
      xo[g] = ci * W4_XCS + sl * W4_SLOT + cdyo[g] * 32;
    }
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int r = 2 * wv + (st >> 1), vb = (st & 1) * 16 + kh * 8;
      const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(&dl[0][i * W4_DCS + r * 32 + vb]);
      const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(&dl[1][i * W4_DCS + r * 32 + vb]);
      accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, ones, accb, 0, 0, 0);
#pragma unroll
      for (int g = 0; g < 5; ++g) {
        const uint16_t* src = cval[g] ? &xl[ccopy[g]][xo[g] + r * 32 + vb] : zero8;
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(src);
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(g < 3 ? a0 : a1, bv, acc[g], 0, 0, 0);
      }
    }
  };

  const int64_t nwg = gridDim.x;
  const int64_t s = (nwg % 8 == 0) ? (int64_t)(blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8 : (int64_t)blockIdx.x;
  for (int64_t task = s; task < ntasks; task += nwg) {
    int64_t t = task;
    const int bx = (int)(t % ntx);
    t /= ntx;
    const int by = (int)(t % nty);
    t /= nty;
    const int bz = (int)(t % ntz);
    cb = (int)(t / ntz);
    cx0 = bx * 32;
    cy0 = by * W4_Y;
    const int z0 = bz * zc;
    load_x(z0 - 1);
    __syncthreads();  // the previous task's MFMAs are done with LDS (and zero8 is written)
    store_x(0);
    load_x(z0);
    store_x(1);
    load_x(z0 + 1);
    load_dz(z0);
    for (int zi = 0; zi < zc; ++zi) {
      __syncthreads();  // the previous plane's MFMAs are done with slot (zi + 2) % 3 and the dz plane
      store_x((zi + 2) % 3);
      store_dz();
      __syncthreads();
      if (zi + 1 < zc) {
        load_x(z0 + zi + 2);
        load_dz(z0 + zi + 1);
      }
      mfma_plane(zi % 3);
    }
  }
  float* pw = part + ((int64_t)blockIdx.x * 4 + wv) * W4_PART;
#pragma unroll
  for (int g = 0; g < 5; ++g) {
    if (cval[g])
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int co = 8 * (q / 4) + 4 * kh + (q % 4);
        pw[(ctap[g] * 32 + co) * 4 + ci] = acc[g][q];
      }
  }
  if (i == 0)
#pragma unroll
    for (int q = 0; q < 16; ++q) pw[27 * 32 * 4 + 8 * (q / 4) + 4 * kh + (q % 4)] = accb[q];
}

#ifndef AUGUR_WGRAD_Z
#define AUGUR_WGRAD_Z 1
#endif

// out_w[j] (j < mw) and out_b[j - mw] = sum over partials in workgroup order
__global__ void k_sum_partials(const float* __restrict__ part, int nparts, int64_t m, int64_t mw,
                               float* __restrict__ out_w, float* __restrict__ out_b) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < m; j += (int64_t)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int p = 0; p < nparts; ++p) s += part[(int64_t)p * m + j];
    if (j < mw)
      out_w[j] = s;
    else
      out_b[j - mw] = s;
  }
}

// dz = dh * gelu'(z)  (BF16 out), dh FP32 channels-last, z BF16
__global__ void k_gelu_bwd(const float* __restrict__ dh, const uint16_t* __restrict__ z, uint16_t* __restrict__ dz,
                           int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dz[i] = f2bf(dh[i] * gelu_grad(bf2f(z[i])));
}

// Data-gradient weights: W'[t'][a][c] = W[26 - t'][c][a] for c < cout_fwd
// (else 0), a < 32 rows (rows a >= cin_fwd are 0); W is [27][32][cin_fwd],
// W' is [27][32][cinp].
__global__ void k_flip_weights(const uint16_t* __restrict__ w, int cin_fwd, int cout_fwd, int cinp,
                               uint16_t* __restrict__ wt) {
  const int n = 27 * 32 * cinp;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const int c = j % cinp, a = (j / cinp) % 32, tp = j / (32 * cinp);
    wt[j] = (c < cout_fwd && a < cin_fwd) ? w[((26 - tp) * 32 + c) * cin_fwd + a] : (uint16_t)0;
  }
}

// FP32 channels-first [B][C][G^3] * scale -> BF16 channels-last [B][G^3][CP] (zero padded)
// one voxel per thread: C strided FP32 channel reads (coalesced across the
// voxels), CP BF16 channels written as 16-byte chunks (CP a multiple of 8)
__global__ void k_pack_cl8(const float* __restrict__ in, int C, int64_t B, int64_t G3, int CP, float scale,
                           uint16_t* __restrict__ out) {
  const int64_t nv = B * G3;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nv; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / G3, v = t - b * G3;
    const float* src = in + b * C * G3 + v;
    for (int g = 0; g < CP / 8; ++g) {
      float f[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int c = g * 8 + k;
        f[k] = c < C ? src[(int64_t)c * G3] * scale : 0.0f;
      }
      u32x4 w;
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = pack_bf16x2(f[2 * k], f[2 * k + 1]);
      *reinterpret_cast<u32x4*>(&out[t * CP + g * 8]) = w;
    }
  }
}

// the same for a 4-channel image: one voxel per thread, 8 bytes written
__global__ void k_pack_cl4(const float* __restrict__ in, int C, int64_t B, int64_t G3, float scale,
                           uint16_t* __restrict__ out) {
  const int64_t nv = B * G3;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nv; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / G3, v = t - b * G3;
    const float* src = in + b * C * G3 + v;
    float f[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) f[c] = c < C ? src[(int64_t)c * G3] * scale : 0.0f;
    *reinterpret_cast<uint2*>(&out[t * 4]) = uint2{pack_bf16x2(f[0], f[1]), pack_bf16x2(f[2], f[3])};
  }
}

// the correction net's input straight from the PM grids: channel 0 =
// rho * rho_mul + rho_add (the scaled density contrast), channels 1..3 =
// force planes * force_mul, channels 4..CP-1 zero; BF16 channels-last
__global__ void k_pack_features(const float* __restrict__ rho, const float* __restrict__ force, int64_t B,
                                int64_t G3, float rho_mul, float rho_add, float force_mul, int CP,
                                uint16_t* __restrict__ out) {
  const int64_t nv = B * G3;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nv; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / G3, v = t - b * G3;
    const float* fb = force + b * 3 * G3 + v;
    u32x4 w;
    w[0] = pack_bf16x2(fmaf(rho[t], rho_mul, rho_add), fb[0] * force_mul);
    w[1] = pack_bf16x2(fb[G3] * force_mul, fb[2 * G3] * force_mul);
    w[2] = 0u;
    w[3] = 0u;
    if (CP == 4) {  // the 4-channel input image: 8 bytes per voxel
      *reinterpret_cast<uint2*>(&out[t * 4]) = uint2{w[0], w[1]};
      continue;
    }
    *reinterpret_cast<u32x4*>(&out[t * CP]) = w;
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (int g = 1; g < CP / 8; ++g) *reinterpret_cast<u32x4*>(&out[t * CP + g * 8]) = z;
  }
}

__global__ void k_pack_cl(const float* __restrict__ in, int C, int64_t B, int64_t G3, int CP, float scale,
                          uint16_t* __restrict__ out) {
  const int64_t n = B * G3 * CP;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(j % CP);
    const int64_t v = (j / CP) % G3, b = j / (CP * G3);
    out[j] = c < C ? f2bf(in[(b * C + c) * G3 + v] * scale) : (uint16_t)0;
  }
}

// Adam (decoupled bias correction), FP32 master weights, optional BF16 copy.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2, float gscale,
                       uint16_t* __restrict__ p_bf16) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const float gj = g[j] * gscale;
    const float mj = b1 * m[j] + (1.0f - b1) * gj;
    const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float pj = p[j] - lr * (mj / c1) / (sqrtf(vj / c2) + eps);
    p[j] = pj;
    if (p_bf16) p_bf16[j] = f2bf(pj);
  }
}

// Adam gated on the device: the step counter and the skip flag live in
// opt[0] (steps taken) and opt[2] (1: this step skipped); the gradient scale
// (the clip factor) in *gscale.  Nothing is written when the step is skipped.
__global__ void k_adam_gated(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                             float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                             const int* __restrict__ opt, const float* __restrict__ gscale,
                             uint16_t* __restrict__ p_bf16) {
  if (opt[2] != 0) return;
  const int step = opt[0];
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  const float gs = *gscale;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const float gj = g[j] * gs;
    const float mj = b1 * m[j] + (1.0f - b1) * gj;
    const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float pj = p[j] - lr * (mj / c1) / (sqrtf(vj / c2) + eps);
    p[j] = pj;
    if (p_bf16) p_bf16[j] = f2bf(pj);
  }
}

// acc[0] += sum g^2 over n values (FP64, one workgroup, fixed order)
constexpr int GN_T = 1024;
__global__ __launch_bounds__(GN_T) void k_grad_sumsq(const float* __restrict__ g, int64_t n, double* __restrict__ acc) {
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += GN_T) s += (double)g[j] * (double)g[j];
  __shared__ double red[GN_T / 64];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < GN_T / 64; ++w) t += red[w];
    acc[0] += t;
  }
}

// clip_grad_norm_(max_norm) and the non-finite skip: norm = sqrt(acc[0]);
// skip when the norm or the loss is not finite; scale = min(1, max_norm /
// (norm + 1e-6)); the step counter advances only on a step taken.  acc[0] is
// cleared for the next step; acc[1] keeps the last norm.
__global__ void k_grad_clip_decide(double* __restrict__ acc, const float* __restrict__ loss, float max_norm,
                                   float* __restrict__ gscale, int* __restrict__ opt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double norm = sqrt(acc[0]);
  const bool ok = isfinite(norm) && (loss == nullptr || isfinite(*loss));
  float sc = 1.0f;
  if (max_norm > 0.0f) {
    const float c = max_norm / ((float)norm + 1e-6f);
    sc = c < 1.0f ? c : 1.0f;
  }
  *gscale = sc;
  acc[1] = norm;
  acc[0] = 0.0;
  if (ok) {
    opt[0] += 1;
    opt[2] = 0;
  } else {
    opt[1] += 1;
    opt[2] = 1;
  }
}

int conv_tiles_ok(int64_t G) { return is_pow2(G) && G >= 32 && G <= 1024; }

#ifndef AUGUR_CONV_SPLIT
#define AUGUR_CONV_SPLIT 1
#endif
// Kernel families of the forward convolution (the tile variants a caller or
// an autotuner may pick; AUGUR_CONV_TILE_* in augur.h):
//   ZM     k_conv3_zm: 32x32x16 MFMAs, four waves of 4 rows (CIN 16: eight waves of 2 rows)
//   SPLIT  k_conv3_zm_split: eight waves, each half the output channels (CIN 32, 32 outputs)
//   NARROW k_conv3_zm_narrow: 16x16x32 MFMAs over at most 16 output channels (CIN 32)
// Every family computes each output voxel with the same per-voxel MFMA chain
// for any z planes per workgroup, so results do not depend on zc.
enum ConvTile : int { TILE_DEFAULT = 0, TILE_ZM = 1, TILE_SPLIT = 2, TILE_NARROW = 3, TILE_IN4 = 4 };

int conv_variant_mask(int cin, int mode, int cout) {
  if (cin == 4) return (mode == CONV_GELU && cout == 32) ? 1 << TILE_IN4 : 0;  // k_conv3_in4 only
  int m = 1 << TILE_ZM;
  if (cin == 32 && (mode == CONV_GELU || mode == CONV_RES_GELU || (mode == CONV_OUT_F32_CL && cout == 32)))
    m |= 1 << TILE_SPLIT;
  if (cin == 32 && cout <= 16 && (mode == CONV_OUT_F32_CF || mode == CONV_OUT_F32_CL)) m |= 1 << TILE_NARROW;
  if (mode == CONV_OUT_F32_CL_DGELU) m = (cin == 32 && cout == 32) ? 1 << TILE_SPLIT : 0;  // split kernel only
  if (mode == CONV_OUT_F32_CF_INTO) m = (cin == 32 && cout <= 16) ? 1 << TILE_NARROW : 0;  // narrow kernel only
  return m;
}

int conv_default_variant(int mask) {
  if (mask & (1 << TILE_IN4)) return TILE_IN4;
  if (mask & (1 << TILE_NARROW)) return TILE_NARROW;
  if (AUGUR_CONV_SPLIT && (mask & (1 << TILE_SPLIT))) return TILE_SPLIT;
  return TILE_ZM;
}

template <int CIN, int MODE>
int launch_fwd(const uint16_t* x, const uint16_t* w, const float* bias, int64_t B, int64_t G, int cout, uint16_t* z,
               uint16_t* h, float* y, const float* addend, int variant, int zc, hipStream_t s, const char* name,
               float alpha = 1.0f, int accumulate = 0, const uint32_t* skip = nullptr) {
  const int64_t blocks = B * (G / ZM_X) * (G / ZM_Y) * (G / zc);
  AUGUR_REQUIRE(blocks < ((int64_t)1 << 31), AUGUR_EINVAL, "conv: grid too large");
  TimedScope ts_(name, s);
  if (variant == TILE_SPLIT) {
    if constexpr (CIN == 32 && MODE != CONV_OUT_F32_CF && MODE != CONV_OUT_F32_CF_INTO) {
      hipLaunchKernelGGL((k_conv3_zm_split<MODE>), dim3((unsigned)blocks), dim3(ZS_THREADS), 0, s, x, w, bias,
                         (int)G, zc, z, h, y, addend);
      AUGUR_LAUNCHED(name);
      return AUGUR_OK;
    }
  } else if (variant == TILE_NARROW) {
    if constexpr (CIN == 32 && (MODE == CONV_OUT_F32_CF || MODE == CONV_OUT_F32_CL || MODE == CONV_OUT_F32_CF_INTO)) {
      hipLaunchKernelGGL((k_conv3_zm_narrow<MODE>), dim3((unsigned)blocks), dim3(ZS_THREADS), 0, s, x, w, bias,
                         (int)G, zc, cout, y, addend, alpha, accumulate, skip);
      AUGUR_LAUNCHED(name);
      return AUGUR_OK;
    }
  } else if (variant == TILE_ZM) {
    if constexpr (MODE != CONV_OUT_F32_CL_DGELU && MODE != CONV_OUT_F32_CF_INTO) {
      if constexpr (CIN == 16) {  // 108 weight VGPRs: eight waves of 2 rows fit two waves per SIMD
        hipLaunchKernelGGL((k_conv3_zm<CIN, MODE, 2>), dim3((unsigned)blocks), dim3(64 * (ZM_Y / 2)), 0, s, x, w,
                           bias, (int)G, zc, cout, z, h, y, addend);
      } else {
        hipLaunchKernelGGL((k_conv3_zm<CIN, MODE>), dim3((unsigned)blocks), dim3(ZM_THREADS), 0, s, x, w, bias,
                           (int)G, zc, cout, z, h, y, addend);
      }
      AUGUR_LAUNCHED(name);
      return AUGUR_OK;
    }
  }
  set_error("conv: tile variant %d has no kernel for cin %d, mode %d", variant, CIN, MODE);
  return AUGUR_EUNSUPPORTED;
}

}  // namespace augur

extern "C" {

int augur_conv3_tile_variants(int cin, int mode, int cout) {
  if ((cin != 4 && cin != 16 && cin != 32) || mode < CONV_OUT_F32_CF || mode > CONV_OUT_F32_CF_INTO || cout < 1 || cout > 32)
    return 0;
  if (mode == CONV_RES_GELU && cin != 32) return 0;
  return conv_variant_mask(cin, mode, cout);
}

int augur_conv3_bf16_fwd_tile(const uint16_t* x, int cin, const uint16_t* w, const float* bias, int64_t B, int64_t G,
                              int mode, int cout, uint16_t* z_out, uint16_t* h_out, float* y_out, const float* addend,
                              int variant, int zc, void* stream) {
  AUGUR_REQUIRE(conv_tiles_ok(G), AUGUR_EUNSUPPORTED, "conv: G must be a power of two in [32, 1024]");
  AUGUR_REQUIRE(B >= 1 && x && w, AUGUR_EINVAL, "conv: bad arguments");
  AUGUR_REQUIRE(cin == 4 || cin == 16 || cin == 32, AUGUR_EUNSUPPORTED,
                "conv: cin must be 4, 16 or 32 (pad channels)");
  AUGUR_REQUIRE(cout >= 1 && cout <= 32, AUGUR_EINVAL, "conv: cout must be 1..32");
  AUGUR_REQUIRE(is_pow2(zc) && zc >= 2 && zc <= G, AUGUR_EINVAL, "conv: zc must be a power of two in [2, G]");
  AUGUR_REQUIRE(cin != 4 || (mode == CONV_GELU && cout == 32 && h_out && bias), AUGUR_EUNSUPPORTED,
                "conv: cin 4 (the input layer) supports mode 1 with 32 outputs only");
  AUGUR_REQUIRE(mode != CONV_OUT_F32_CL_DGELU || (cin == 32 && cout == 32), AUGUR_EINVAL,
                "conv: mode 4 needs z_out (read), h_out (dz), cin = cout = 32");
  const int mask = conv_variant_mask(cin, mode, cout);
  if (variant == TILE_DEFAULT) variant = conv_default_variant(mask);
  AUGUR_REQUIRE(variant >= TILE_ZM && variant <= TILE_IN4 && ((mask >> variant) & 1), AUGUR_EUNSUPPORTED,
                "conv: tile variant cannot run this layer");
  // the narrow channel-first epilogue addresses one scene's planes through a
  // 32-bit buffer view
  if (variant == TILE_NARROW && mode == CONV_OUT_F32_CF && (int64_t)cout * G * G * G * 4 >= ((int64_t)1 << 31))
    variant = TILE_ZM;
  hipStream_t s = as_stream(stream);
  if (cin == 4) {
    const int64_t blocks = B * (G / ZM_X) * (G / ZM_Y) * (G / zc);
    AUGUR_REQUIRE(blocks < ((int64_t)1 << 31), AUGUR_EINVAL, "conv: grid too large");
    TimedScope ts_("conv_fwd_gelu", s);
    hipLaunchKernelGGL((k_conv3_in4<CONV_GELU>), dim3((unsigned)blocks), dim3(IN4_THREADS), 0, s, x, w, bias, (int)G,
                       zc, z_out, h_out);
    AUGUR_LAUNCHED("conv_fwd_gelu");
    return AUGUR_OK;
  }
  switch (mode) {
    case CONV_OUT_F32_CF:
      AUGUR_REQUIRE(y_out && bias, AUGUR_EINVAL, "conv: mode 0 needs y_out and bias");
      return cin == 16 ? launch_fwd<16, CONV_OUT_F32_CF>(x, w, bias, B, G, cout, nullptr, nullptr, y_out, nullptr,
                                                         variant, zc, s, "conv_fwd_out")
                       : launch_fwd<32, CONV_OUT_F32_CF>(x, w, bias, B, G, cout, nullptr, nullptr, y_out, nullptr,
                                                         variant, zc, s, "conv_fwd_out");
    case CONV_GELU:
      AUGUR_REQUIRE(h_out && bias && cout == 32, AUGUR_EINVAL, "conv: mode 1 needs h_out, bias, cout 32");
      return cin == 16 ? launch_fwd<16, CONV_GELU>(x, w, bias, B, G, 32, z_out, h_out, nullptr, nullptr, variant, zc,
                                                   s, "conv_fwd_gelu")
                       : launch_fwd<32, CONV_GELU>(x, w, bias, B, G, 32, z_out, h_out, nullptr, nullptr, variant, zc,
                                                   s, "conv_fwd_gelu");
    case CONV_RES_GELU:
      AUGUR_REQUIRE(h_out && bias && cout == 32 && cin == 32, AUGUR_EINVAL, "conv: mode 2 needs cin = cout = 32");
      return launch_fwd<32, CONV_RES_GELU>(x, w, bias, B, G, 32, z_out, h_out, nullptr, nullptr, variant, zc, s,
                                           "conv_fwd_res");
    case CONV_OUT_F32_CL:
      AUGUR_REQUIRE(y_out, AUGUR_EINVAL, "conv: mode 3 needs y_out");
      return cin == 16 ? launch_fwd<16, CONV_OUT_F32_CL>(x, w, nullptr, B, G, cout, nullptr, nullptr, y_out, addend,
                                                         variant, zc, s, "conv_bwd_data")
                       : launch_fwd<32, CONV_OUT_F32_CL>(x, w, nullptr, B, G, cout, nullptr, nullptr, y_out, addend,
                                                         variant, zc, s, "conv_bwd_data");
    case CONV_OUT_F32_CL_DGELU:
      AUGUR_REQUIRE(z_out && h_out && cin == 32 && cout == 32, AUGUR_EINVAL,
                    "conv: mode 4 needs z_out (read), h_out (dz), cin = cout = 32");
      return launch_fwd<32, CONV_OUT_F32_CL_DGELU>(x, w, nullptr, B, G, 32, z_out, h_out, y_out, addend, variant,
                                                   zc, s, "conv_bwd_data");
    default:
      set_error("conv: unknown mode %d", mode);
      return AUGUR_EINVAL;
  }
}

int augur_conv3_bf16_fwd_into(const uint16_t* x, int cin, const uint16_t* w, const float* bias, int64_t B,
                              int64_t G, int cout, float alpha, int accumulate, float* field, const uint32_t* skip,
                              int zc, void* stream) {
  AUGUR_REQUIRE(conv_tiles_ok(G), AUGUR_EUNSUPPORTED, "conv: G must be a power of two in [32, 1024]");
  AUGUR_REQUIRE(B >= 1 && x && w && bias && field, AUGUR_EINVAL, "conv_into: bad arguments");
  AUGUR_REQUIRE(cin == 32 && cout >= 1 && cout <= 16, AUGUR_EUNSUPPORTED, "conv_into: needs cin 32, cout 1..16");
  AUGUR_REQUIRE((int64_t)cout * G * G * G * 4 < ((int64_t)1 << 31), AUGUR_EUNSUPPORTED,
                "conv_into: one scene's field (cout G^3 FP32) must stay below 2 GiB");
  if (zc == 0) zc = (int)(G < ZM_ZC ? G : ZM_ZC);
  AUGUR_REQUIRE(is_pow2(zc) && zc >= 2 && zc <= G, AUGUR_EINVAL, "conv: zc must be a power of two in [2, G]");
  return launch_fwd<32, CONV_OUT_F32_CF_INTO>(x, w, bias, B, G, cout, nullptr, nullptr, field, nullptr, TILE_NARROW,
                                               zc, as_stream(stream), "conv_fwd_out", alpha, accumulate ? 1 : 0,
                                               skip);
}

int augur_conv3_bf16_fwd(const uint16_t* x, int cin, const uint16_t* w, const float* bias, int64_t B, int64_t G,
                         int mode, int cout, uint16_t* z_out, uint16_t* h_out, float* y_out, const float* addend,
                         void* stream) {
  return augur_conv3_bf16_fwd_tile(x, cin, w, bias, B, G, mode, cout, z_out, h_out, y_out, addend, TILE_DEFAULT,
                                   ZM_ZC, stream);
}

size_t augur_conv3_wgrad_workspace_bytes(int cin, int64_t B, int64_t G, int nwg) {
  (void)B;
  (void)G;
  const size_t parts = cin == 4 ? (size_t)nwg * 4 : (size_t)nwg;  // cin 4: one partial per wave
  return align_up(parts * (27 * 32 * cin + 32) * sizeof(float)) + 256;
}

int augur_conv3_bf16_wgrad(const uint16_t* x, int cin, const uint16_t* dz, int64_t B, int64_t G, float* dw, float* db,
                           int nwg, void* ws, size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(conv_tiles_ok(G), AUGUR_EUNSUPPORTED, "conv: G must be a power of two in [32, 1024]");
  AUGUR_REQUIRE(cin == 4 || cin == 16 || cin == 32, AUGUR_EUNSUPPORTED, "conv wgrad: cin must be 4, 16 or 32");
  AUGUR_REQUIRE(B >= 1 && x && dz && dw && db && nwg >= 1 && nwg <= 65536, AUGUR_EINVAL, "conv wgrad: bad arguments");
  const int64_t m = 27 * 32 * cin + 32;
  const int nparts = cin == 4 ? 4 * nwg : nwg;
  AUGUR_REQUIRE(ws && ws_bytes >= (size_t)nparts * m * sizeof(float), AUGUR_EWORKSPACE,
                "conv wgrad: workspace too small");
  hipStream_t s = as_stream(stream);
  const int64_t ntiles = B * (G / CT_X) * (G / CT_Y) * (G / CT_Z);
  {
    TimedScope ts_("conv_wgrad", s);
    if (cin == 4) {
      const int zc = (int)(G < WZ_ZC ? G : WZ_ZC);
      const int64_t ntasks = B * (G / 32) * (G / W4_Y) * (G / zc);
      hipLaunchKernelGGL(k_conv3_wgrad_in4, dim3(nwg), dim3(CONV_THREADS), 0, s, x, dz, (int)B, (int)G, zc, ntasks,
                         (float*)ws);
    } else if (cin == 16)
      hipLaunchKernelGGL((k_conv3_wgrad<16>), dim3(nwg), dim3(CONV_THREADS), 0, s, x, dz, (int)B, (int)G, ntiles,
                         (float*)ws);
    else if (AUGUR_WGRAD_Z) {
      const int zc = (int)(G < WZ_ZC ? G : WZ_ZC);
      const int64_t ntasks = B * (G / WZ_X) * (G / WZ_Y) * (G / zc);
      hipLaunchKernelGGL(k_conv3_wgrad_z, dim3(nwg), dim3(CONV_THREADS), 0, s, x, dz, (int)B, (int)G, zc, ntasks,
                         (float*)ws);
    } else
      hipLaunchKernelGGL((k_conv3_wgrad<32>), dim3(nwg), dim3(CONV_THREADS), 0, s, x, dz, (int)B, (int)G, ntiles,
                         (float*)ws);
    AUGUR_LAUNCHED("conv_wgrad");
  }
  hipLaunchKernelGGL(k_sum_partials, dim3(blocks_for(m, 256)), dim3(256), 0, s, (const float*)ws, nparts, m,
                     (int64_t)27 * 32 * cin, dw, db);
  AUGUR_LAUNCHED("sum_partials");
  return AUGUR_OK;
}

int augur_conv3_flip_weights(const uint16_t* w, int cin_fwd, int cout_fwd, int cinp, uint16_t* wt, void* stream) {
  AUGUR_REQUIRE(w && wt && cin_fwd >= 1 && cin_fwd <= 32 && cout_fwd >= 1 && cout_fwd <= cinp &&
                    (cinp == 16 || cinp == 32),
                AUGUR_EINVAL, "conv flip: bad arguments");
  hipLaunchKernelGGL(k_flip_weights, dim3(blocks_for(27 * 32 * cinp, 256)), dim3(256), 0, as_stream(stream), w,
                     cin_fwd, cout_fwd, cinp, wt);
  AUGUR_LAUNCHED("flip_weights");
  return AUGUR_OK;
}

int augur_gelu_bwd(const float* dh, const uint16_t* z, int64_t n, uint16_t* dz, void* stream) {
  AUGUR_REQUIRE(n >= 0 && (n == 0 || (dh && z && dz)), AUGUR_EINVAL, "gelu_bwd: bad arguments");
  if (n == 0) return AUGUR_OK;
  hipStream_t s = as_stream(stream);
  TimedScope ts_("gelu_bwd", s);
  hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, s, dh, z, dz, n);
  AUGUR_LAUNCHED("gelu_bwd");
  return AUGUR_OK;
}

int augur_pack_cl_bf16(const float* in, int C, int64_t B, int64_t G, int CP, float scale, uint16_t* out,
                       void* stream) {
  AUGUR_REQUIRE(in && out && C >= 1 && C <= CP && B >= 1 && G >= 1, AUGUR_EINVAL, "pack_cl: bad arguments");
  const int64_t G3 = G * G * G;
  TimedScope ts_("pack_cl", as_stream(stream));
  if (CP % 8 == 0)
    hipLaunchKernelGGL(k_pack_cl8, dim3(blocks_for(B * G3, 256)), dim3(256), 0, as_stream(stream), in, C, B, G3, CP,
                       scale, out);
  else if (CP == 4)
    hipLaunchKernelGGL(k_pack_cl4, dim3(blocks_for(B * G3, 256)), dim3(256), 0, as_stream(stream), in, C, B, G3,
                       scale, out);
  else
    hipLaunchKernelGGL(k_pack_cl, dim3(blocks_for(B * G3 * CP, 256)), dim3(256), 0, as_stream(stream), in, C, B, G3,
                       CP, scale, out);
  AUGUR_LAUNCHED("pack_cl");
  return AUGUR_OK;
}

int augur_pack_features_bf16(const float* rho, const float* force, int64_t B, int64_t G, float rho_mul, float rho_add,
                             float force_mul, int CP, uint16_t* out, void* stream) {
  AUGUR_REQUIRE(rho && force && out && B >= 1 && G >= 1 && (CP == 4 || (CP >= 8 && CP % 8 == 0)), AUGUR_EINVAL,
                "pack_features: bad arguments");
  const int64_t G3 = G * G * G;
  TimedScope ts_("pack_features", as_stream(stream));
  hipLaunchKernelGGL(k_pack_features, dim3(blocks_for(B * G3, 256)), dim3(256), 0, as_stream(stream), rho, force, B,
                     G3, rho_mul, rho_add, force_mul, CP, out);
  AUGUR_LAUNCHED("pack_features");
  return AUGUR_OK;
}

// y += alpha x over n floats (float4 where aligned); nothing while *skip != 0
__global__ __launch_bounds__(256) void k_field_add(float* __restrict__ y, const float* __restrict__ x, float alpha,
                                                   int64_t n, const uint32_t* __restrict__ skip) {
  if (skip && *skip != 0u) return;
  const int64_t n4 = n / 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += stride) {
    float4 a = reinterpret_cast<float4*>(y)[i];
    const float4 b = reinterpret_cast<const float4*>(x)[i];
    a.x = fmaf(alpha, b.x, a.x);
    a.y = fmaf(alpha, b.y, a.y);
    a.z = fmaf(alpha, b.z, a.z);
    a.w = fmaf(alpha, b.w, a.w);
    reinterpret_cast<float4*>(y)[i] = a;
  }
  for (int64_t i = 4 * n4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = fmaf(alpha, x[i], y[i]);
}

int augur_field_add(float* y, const float* x, float alpha, int64_t n, const uint32_t* skip, void* stream) {
  AUGUR_REQUIRE(y && x && n >= 0, AUGUR_EINVAL, "field_add: bad arguments");
  AUGUR_REQUIRE((reinterpret_cast<uintptr_t>(y) & 15) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0,
                AUGUR_EINVAL, "field_add: y and x must be 16-byte aligned");
  if (n == 0) return AUGUR_OK;
  hipStream_t s = as_stream(stream);
  TimedScope ts_("field_add", s);
  hipLaunchKernelGGL(k_field_add, dim3(blocks_for((n + 3) / 4, 256, 4096)), dim3(256), 0, s, y, x, alpha, n, skip);
  AUGUR_LAUNCHED("field_add");
  return AUGUR_OK;
}

int augur_adam_update(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                      float eps, int step, float gscale, uint16_t* p_bf16, void* stream) {
  AUGUR_REQUIRE(p && g && m && v && n >= 0 && step >= 1, AUGUR_EINVAL, "adam: bad arguments");
  if (n == 0) return AUGUR_OK;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), p, g, m, v, n, lr, b1, b2, eps,
                     c1, c2, gscale, p_bf16);
  AUGUR_LAUNCHED("adam");
  return AUGUR_OK;
}

int augur_grad_sumsq(const float* g, int64_t n, double* acc, void* stream) {
  AUGUR_REQUIRE(g && acc && n >= 0, AUGUR_EINVAL, "grad_sumsq: bad arguments");
  if (n == 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_grad_sumsq, dim3(1), dim3(GN_T), 0, as_stream(stream), g, n, acc);
  AUGUR_LAUNCHED("grad_sumsq");
  return AUGUR_OK;
}

int augur_grad_clip_decide(double* acc, const float* loss, float max_norm, float* gscale, int* opt, void* stream) {
  AUGUR_REQUIRE(acc && gscale && opt, AUGUR_EINVAL, "grad_clip_decide: bad arguments");
  hipLaunchKernelGGL(k_grad_clip_decide, dim3(1), dim3(64), 0, as_stream(stream), acc, loss, max_norm, gscale, opt);
  AUGUR_LAUNCHED("grad_clip_decide");
  return AUGUR_OK;
}

int augur_adam_update_gated(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                            float eps, const int* opt, const float* gscale, uint16_t* p_bf16, void* stream) {
  AUGUR_REQUIRE(p && g && m && v && opt && gscale && n >= 0, AUGUR_EINVAL, "adam_gated: bad arguments");
  if (n == 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_adam_gated, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), p, g, m, v, n, lr, b1,
                     b2, eps, opt, gscale, p_bf16);
  AUGUR_LAUNCHED("adam");
  return AUGUR_OK;
}

}  // extern "C"
