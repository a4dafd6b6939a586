// Large-N driver with compact particle state: positions and velocities kept
// in 16- or 32-bit encodings in HBM and decoded to FP32 per particle inside
// the kernels, so a step needs no FP32 copies of the state and the launches
// can run over chunks of the particles.  The state is in the merged KDK form
// M = v + (dt/2) a of the reference's billion-body loop
// (scripts/scatter_bh_1b.py:1-9,107-146): one step drifts x += dt M and
// deposits (one pass), solves the grid, then kicks M += dt (a - abar) at the
// new positions (a second pass), abar being the FP64 mass-weighted mean
// acceleration taken from the grid (:119-120), which removes the mean force
// FP32 round-off leaves (momentum conservation).
//
// Encodings (per axis): positions F32, FIX32 (x = q L / 2^32), FIX16
// (x = q L / 2^16) -- fixed point wraps the periodic box for free -- and
// velocities F32, F16, BF16.  The deposit is the fixed-point CIC of the PM
// stepper (the same units, int64 atomics with runs of equal cells merged in
// the wave), so the density is bitwise independent of order and of the chunk
// split.
#include <cstring>

#include "internal.hpp"

using namespace augur;
using namespace augur_dev;

namespace {

enum PosEnc : int { POS_F32 = 0, POS_FIX32 = 1, POS_FIX16 = 2 };
enum VelEnc : int { VEL_F32 = 0, VEL_F16 = 1, VEL_BF16 = 2 };

__device__ __forceinline__ float load_pos(const void* p, int enc, int64_t k, float L) {
  if (enc == POS_FIX32) return (float)((double)static_cast<const uint32_t*>(p)[k] * (double)L * 0x1p-32);
  if (enc == POS_FIX16) return (float)static_cast<const uint16_t*>(p)[k] * (L * 0x1p-16f);
  return static_cast<const float*>(p)[k];
}

// x (any real) wrapped into [0, L) and stored
__device__ __forceinline__ void store_pos(void* p, int enc, int64_t k, float x, float L) {
  if (enc == POS_FIX32 || enc == POS_FIX16) {
    const int bits = enc == POS_FIX32 ? 32 : 16;
    double f = (double)x / (double)L;
    f -= floor(f);
    const double q = rint(ldexp(f, bits));      // 2^bits wraps to 0
    if (enc == POS_FIX32)
      static_cast<uint32_t*>(p)[k] = (uint32_t)(unsigned long long)q;
    else
      static_cast<uint16_t*>(p)[k] = (uint16_t)(unsigned)q;
    return;
  }
  float y = x;
  if (y >= L || y < 0.0f) {
    y = y - L * floorf(y / L);
    if (y >= L) y = 0.0f;
  }
  static_cast<float*>(p)[k] = y;
}

__device__ __forceinline__ float load_vel(const void* p, int enc, int64_t k) {
  if (enc == VEL_F16) return (float)static_cast<const _Float16*>(p)[k];
  if (enc == VEL_BF16) return __uint_as_float((uint32_t) static_cast<const uint16_t*>(p)[k] << 16);
  return static_cast<const float*>(p)[k];
}

__device__ __forceinline__ void store_vel(void* p, int enc, int64_t k, float v) {
  if (enc == VEL_F16) {
    static_cast<_Float16*>(p)[k] = (_Float16)v;   // round to nearest even
  } else if (enc == VEL_BF16) {  // round to nearest even
    const uint32_t u = __float_as_uint(v);
    static_cast<uint16_t*>(p)[k] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
  } else {
    static_cast<float*>(p)[k] = v;
  }
}

// CIC base cell and weights of a position in [0, L) (the stepper's rounding)
__device__ __forceinline__ void cic_base3(const float x[3], int G, float h, float inv_h, bool p2, int c[3], float w[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float u = p2 ? cic_u_pow2(x[d], inv_h) : cic_u(x[d], 0.0f, h);
    const float f = floorf(u);
    c[d] = (int)f;
    w[d] = __fsub_rn(u, f);
  }
}

// drift (dt != 0: x += dt M, stored back) and the fixed-point deposit of the
// particles [lo, hi), mass_value each, units rint(m scale w 2^E)
__global__ __launch_bounds__(256) void k_lp_drift_deposit(void* __restrict__ pos, int penc,
                                                          const void* __restrict__ vel, int venc, int64_t lo,
                                                          int64_t hi, int64_t n_total, float dt, int G, float L,
                                                          float h, float inv_h, int p2, float mass_scale,
                                                          unsigned maxbits, unsigned long long* __restrict__ img) {
  const int E = fixed_exponent(maxbits, n_total);
  const float valE = ldexpf(mass_scale, E);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = lo + (int64_t)blockIdx.x * blockDim.x; base < hi; base += stride) {
    const int64_t i = base + threadIdx.x;
    const bool live = i < hi;
    const int64_t k = live ? i : hi - 1;
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = load_pos(pos, penc, 3 * k + d, L);
    if (dt != 0.0f) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float y = fmaf(dt, load_vel(vel, venc, 3 * k + d), x[d]);
        if (live) store_pos(pos, penc, 3 * k + d, y, L);
      }
      // re-read what was stored: the deposit sees the encoded position
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = live ? load_pos(pos, penc, 3 * k + d, L) : x[d];
    }
    int c[3];
    float w[3];
    cic_base3(x, G, h, inv_h, p2 != 0, c, w);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      const float wx = ox ? w[0] : __fsub_rn(1.0f, w[0]);
      const float wy = oy ? w[1] : __fsub_rn(1.0f, w[1]);
      const float wz = oz ? w[2] : __fsub_rn(1.0f, w[2]);
      const int64_t cell = ((int64_t)wrap(c[2] + oz, G) * G + wrap(c[1] + oy, G)) * G + wrap(c[0] + ox, G);
      const long long q = fixed_round(__fmul_rn(valE, __fmul_rn(__fmul_rn(wx, wy), wz)), 0);
      wave_merged_add(img, cell, q, live);
    }
  }
}

// rho = units 2^-E (image cleared for the next step)
__global__ void k_lp_finish(unsigned long long* __restrict__ img, int64_t G3, int64_t n_total, unsigned maxbits,
                            float* __restrict__ rho) {
  const int E = fixed_exponent(maxbits, n_total);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < G3; i += (int64_t)gridDim.x * blockDim.x) {
    rho[i] = (float)ldexp((double)(long long)img[i], -E);
    img[i] = 0ull;
  }
}

// abar = sum_c rho_c a_c h^3 / sum_c rho_c h^3 (FP64): the mean acceleration
// the particles see (the gather is the transpose of the deposit); per-block
// partials, then one block sums them in order into out[0..2] (out[3] = mass)
constexpr int GM_BLOCKS = 1024, GM_T = 256;

__global__ __launch_bounds__(GM_T) void k_grid_mean_partial(const float* __restrict__ rho,
                                                            const float* __restrict__ force, int64_t G3,
                                                            double* __restrict__ part) {
  const int64_t per = (G3 + GM_BLOCKS - 1) / GM_BLOCKS;
  const int64_t lo = blockIdx.x * per, hi = lo + per < G3 ? lo + per : G3;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += GM_T) {
    const double r = rho[i];
    s[0] += r * force[i];
    s[1] += r * force[G3 + i];
    s[2] += r * force[2 * G3 + i];
    s[3] += r;
  }
  __shared__ double red[GM_T / 64][4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_down(s[k], o);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 4; ++k) red[threadIdx.x >> 6][k] = s[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < GM_T / 64; ++w) t += red[w][threadIdx.x];
    part[4 * blockIdx.x + threadIdx.x] = t;
  }
}

__global__ void k_grid_mean_final(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int b = 0; b < GM_BLOCKS; ++b) t += part[4 * b + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// M += kick (a(x) - abar) for particles [lo, hi) (abar NULL: zero); with
// sums, per-block FP64 [sum v (3), sum v^2 / 2] of the kicked velocities
__global__ __launch_bounds__(256) void k_lp_kick(const void* __restrict__ pos, int penc, void* __restrict__ vel,
                                                 int venc, int64_t lo, int64_t hi, float kick,
                                                 const float* __restrict__ force, const double* __restrict__ abar,
                                                 int G, float L, float h, float inv_h, int p2) {
  const int64_t G3 = (int64_t)G * G * G;
  float ab[3] = {0.0f, 0.0f, 0.0f};
  if (abar) {
    const double m = abar[3] > 0.0 ? abar[3] : 1.0;
    for (int d = 0; d < 3; ++d) ab[d] = (float)(abar[d] / m);
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += stride) {
    float x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = load_pos(pos, penc, 3 * i + d, L);
    int c[3];
    float w[3];
    cic_base3(x, G, h, inv_h, p2 != 0, c, w);
    float a[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      const float wt = __fmul_rn(__fmul_rn(ox ? w[0] : __fsub_rn(1.0f, w[0]), oy ? w[1] : __fsub_rn(1.0f, w[1])),
                                 oz ? w[2] : __fsub_rn(1.0f, w[2]));
      const int64_t cell = ((int64_t)wrap(c[2] + oz, G) * G + wrap(c[1] + oy, G)) * G + wrap(c[0] + ox, G);
#pragma unroll
      for (int d = 0; d < 3; ++d) a[d] = fmaf(wt, force[d * G3 + cell], a[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
      store_vel(vel, venc, 3 * i + d, fmaf(kick, a[d] - ab[d], load_vel(vel, venc, 3 * i + d)));
  }
}

// decode to FP32 (pos in [0, L), vel) for the caller's diagnostics
__global__ void k_lp_decode(const void* __restrict__ pos, int penc, const void* __restrict__ vel, int venc,
                            int64_t n3, float L, float* __restrict__ pout, float* __restrict__ vout) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x) {
    if (pout) pout[k] = load_pos(pos, penc, k, L);
    if (vout) vout[k] = load_vel(vel, venc, k);
  }
}

__global__ void k_lp_encode(void* __restrict__ pos, int penc, void* __restrict__ vel, int venc, int64_t n3, float L,
                            const float* __restrict__ pin, const float* __restrict__ vin) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n3; k += (int64_t)gridDim.x * blockDim.x) {
    if (pin) store_pos(pos, penc, k, pin[k], L);
    if (vin) store_vel(vel, venc, k, vin[k]);
  }
}

unsigned float_bits(float f) {
  unsigned u;
  std::memcpy(&u, &f, sizeof u);
  return u;
}

bool enc_ok(int penc, int venc) { return penc >= POS_F32 && penc <= POS_FIX16 && venc >= VEL_F32 && venc <= VEL_BF16; }

}  // namespace

extern "C" {

int augur_lp_encode(void* pos, int penc, void* vel, int venc, int64_t n, float L, const float* pin, const float* vin,
                    void* stream) {
  AUGUR_REQUIRE(enc_ok(penc, venc) && n >= 0 && L > 0.0f, AUGUR_EINVAL, "lp_encode: bad arguments");
  AUGUR_REQUIRE((!pin || pos) && (!vin || vel), AUGUR_EINVAL, "lp_encode: null state");
  if (n == 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_lp_encode, dim3(blocks_for(3 * n, 256, 8192)), dim3(256), 0, as_stream(stream), pos, penc, vel,
                     venc, 3 * n, L, pin, vin);
  AUGUR_LAUNCHED("lp_encode");
  return AUGUR_OK;
}

int augur_lp_decode(const void* pos, int penc, const void* vel, int venc, int64_t n, float L, float* pout,
                    float* vout, void* stream) {
  AUGUR_REQUIRE(enc_ok(penc, venc) && n >= 0 && L > 0.0f, AUGUR_EINVAL, "lp_decode: bad arguments");
  AUGUR_REQUIRE((!pout || pos) && (!vout || vel), AUGUR_EINVAL, "lp_decode: null state");
  if (n == 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_lp_decode, dim3(blocks_for(3 * n, 256, 8192)), dim3(256), 0, as_stream(stream), pos, penc, vel,
                     venc, 3 * n, L, pout, vout);
  AUGUR_LAUNCHED("lp_decode");
  return AUGUR_OK;
}

int augur_lp_drift_deposit(void* pos, int penc, const void* vel, int venc, int64_t lo, int64_t hi, int64_t n_total,
                           float mass_value, float dt, int64_t G, float L, unsigned long long* img, void* stream) {
  AUGUR_REQUIRE(enc_ok(penc, venc) && pos && img && 0 <= lo && lo <= hi && hi <= n_total && L > 0.0f,
                AUGUR_EINVAL, "lp_drift_deposit: bad arguments");
  AUGUR_REQUIRE(dt == 0.0f || vel, AUGUR_EINVAL, "lp_drift_deposit: drift needs vel");
  AUGUR_REQUIRE(is_pow2(G) && G >= 2 && G <= 2048, AUGUR_EUNSUPPORTED, "lp_drift_deposit: G must be a power of two");
  if (hi == lo) return AUGUR_OK;
  const float h = (float)((double)L / (double)G);
  const float scale = (float)(1.0 / ((double)h * (double)h * (double)h));
  const float ms = mass_value * scale;   // IEEE single product, as the stepper's bound
  const bool p2 = pow2_float(h);
  TimedScope ts_("lp_drift_deposit", as_stream(stream));
  hipLaunchKernelGGL(k_lp_drift_deposit, dim3(blocks_for(hi - lo, 256, 16384)), dim3(256), 0, as_stream(stream), pos,
                     penc, vel, venc, lo, hi, n_total, dt, (int)G, L, h, p2 ? 1.0f / h : 0.0f, p2 ? 1 : 0, ms,
                     float_bits(fabsf(ms)), img);
  AUGUR_LAUNCHED("lp_drift_deposit");
  return AUGUR_OK;
}

int augur_lp_finish(unsigned long long* img, int64_t G, int64_t n_total, float mass_value, float L, float* rho,
                    void* stream) {
  AUGUR_REQUIRE(img && rho && G >= 2 && n_total >= 1 && L > 0.0f, AUGUR_EINVAL, "lp_finish: bad arguments");
  const float h = (float)((double)L / (double)G);
  const float ms = mass_value * (float)(1.0 / ((double)h * (double)h * (double)h));
  TimedScope ts_("lp_finish", as_stream(stream));
  hipLaunchKernelGGL(k_lp_finish, dim3(blocks_for(G * G * G, 256, 16384)), dim3(256), 0, as_stream(stream), img,
                     G * G * G, n_total, float_bits(fabsf(ms)), rho);
  AUGUR_LAUNCHED("lp_finish");
  return AUGUR_OK;
}

size_t augur_grid_mean_workspace_bytes(void) { return GM_BLOCKS * 4 * sizeof(double); }

int augur_grid_mean_acc_f64(const float* rho, const float* force, int64_t G, double* out, void* ws, size_t ws_bytes,
                            void* stream) {
  AUGUR_REQUIRE(rho && force && out && G >= 1, AUGUR_EINVAL, "grid_mean_acc: bad arguments");
  AUGUR_REQUIRE(ws && ws_bytes >= GM_BLOCKS * 4 * sizeof(double), AUGUR_EWORKSPACE, "grid_mean_acc: workspace too small");
  hipStream_t s = as_stream(stream);
  TimedScope ts_("grid_mean_acc", s);
  hipLaunchKernelGGL(k_grid_mean_partial, dim3(GM_BLOCKS), dim3(GM_T), 0, s, rho, force, G * G * G, (double*)ws);
  AUGUR_LAUNCHED("grid_mean_partial");
  hipLaunchKernelGGL(k_grid_mean_final, dim3(1), dim3(64), 0, s, (const double*)ws, out);
  AUGUR_LAUNCHED("grid_mean_final");
  return AUGUR_OK;
}

int augur_lp_kick(const void* pos, int penc, void* vel, int venc, int64_t lo, int64_t hi, float kick,
                  const float* force, const double* abar, int64_t G, float L, void* stream) {
  AUGUR_REQUIRE(enc_ok(penc, venc) && pos && vel && force && 0 <= lo && lo <= hi && L > 0.0f, AUGUR_EINVAL,
                "lp_kick: bad arguments");
  AUGUR_REQUIRE(is_pow2(G) && G >= 2 && G <= 2048, AUGUR_EUNSUPPORTED, "lp_kick: G must be a power of two");
  if (hi == lo) return AUGUR_OK;
  const float h = (float)((double)L / (double)G);
  const bool p2 = pow2_float(h);
  TimedScope ts_("lp_kick", as_stream(stream));
  hipLaunchKernelGGL(k_lp_kick, dim3(blocks_for(hi - lo, 256, 16384)), dim3(256), 0, as_stream(stream), pos, penc,
                     vel, venc, lo, hi, kick, force, abar, (int)G, L, h, p2 ? 1.0f / h : 0.0f, p2 ? 1 : 0);
  AUGUR_LAUNCHED("lp_kick");
  return AUGUR_OK;
}

}  // extern "C"
