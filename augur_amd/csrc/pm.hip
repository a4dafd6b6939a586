// Periodic PM stepper: particles held sorted by CIC base cell; one step is
// deposit -> rfftn -> Green's function + spectral gradient -> irfftn ->
// fused CIC gather + kick (+ drift of the next step).
#include "internal.hpp"

using namespace augur;
using namespace augur_dev;

namespace augur {

struct PmBufs {
  unsigned long long* fixed;  // G^3 fixed-point deposit accumulator
  unsigned* maxbits;          // deposit scale bound
  float* rho;                 // G^3
  float* spec;                // G*G*(G/2+1) complex
  float* aspec;               // 3 * G*G*(G/2+1) complex
  float* force;               // 3 * G^3: [a_x; a_y; a_z]
};

static bool carve_pm(Carver& c, int64_t G, PmBufs& b) {
  const int64_t G3 = G * G * G, nc = G * G * (G / 2 + 1);
  b.fixed = c.take<unsigned long long>(G3);
  b.maxbits = c.take<unsigned>(4);
  b.rho = c.take<float>(G3);
  b.spec = c.take<float>(2 * nc);
  b.aspec = c.take<float>(6 * nc);
  b.force = c.take<float>(3 * G3);
  return c.ok();
}

__global__ void k_kick_drift(float* __restrict__ pos, float* __restrict__ vel, const float* __restrict__ acc, int64_t n,
                             float hdt, float dt, float L) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float v = vel[3 * i + d] + hdt * acc[3 * i + d];
      float x = pos[3 * i + d] + dt * v;
      if (x >= L) x -= L;
      if (x < 0.0f) x += L;
      vel[3 * i + d] = v;
      pos[3 * i + d] = x;
    }
  }
}

// a = CIC gather of the three force grids at pos; acc = a; v += dt/2 a; and,
// unless this is the last step, the next step's v += dt/2 a; x += dt v.
// With do_kick == 0 only acc is written (force evaluation).
__global__ __launch_bounds__(256) void k_gather_kdk(float* __restrict__ pos, float* __restrict__ vel,
                                                    float* __restrict__ acc, const float* __restrict__ force, int64_t n,
                                                    int G, float h, float hdt, float dt, float L, int do_kick,
                                                    int drift) {
  const int64_t G3 = (int64_t)G * G * G;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float p[3];
    int i0[3];
    float w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      p[d] = pos[3 * i + d];
      const float u = cic_u(p[d], 0.0f, h);
      const float f = floorf(u);
      i0[d] = (int)f;
      w[d] = __fsub_rn(u, f);
    }
    float a[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int oz = 0; oz < 2; ++oz) {
          const float wx = ox ? w[0] : __fsub_rn(1.0f, w[0]);
          const float wy = oy ? w[1] : __fsub_rn(1.0f, w[1]);
          const float wz = oz ? w[2] : __fsub_rn(1.0f, w[2]);
          const float wt = __fmul_rn(__fmul_rn(wx, wy), wz);
          const int64_t cell =
              ((int64_t)wrap(i0[2] + oz, G) * G + wrap(i0[1] + oy, G)) * G + wrap(i0[0] + ox, G);
          a[0] += force[cell] * wt;
          a[1] += force[G3 + cell] * wt;
          a[2] += force[2 * G3 + cell] * wt;
        }
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[3 * i + d] = a[d];
    if (do_kick) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float v = vel[3 * i + d] + hdt * a[d];
        if (drift) {
          v = v + hdt * a[d];
          float x = p[d] + dt * v;
          if (x >= L) x -= L;
          if (x < 0.0f) x += L;
          pos[3 * i + d] = x;
        }
        vel[3 * i + d] = v;
      }
    }
  }
}

__global__ void k_base_keys(const float* __restrict__ pos, int64_t n, int G, float h, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = wrap((int)floorf(cic_u(pos[3 * i + d], 0.0f, h)), G);
    keys[i] = ((uint32_t)c[2] * G + c[1]) * G + c[0];
    vals[i] = (uint32_t)i;
  }
}

template <class T, int K>
__global__ void k_permute(const T* __restrict__ src, T* __restrict__ dst, const uint32_t* __restrict__ perm, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = perm[i];
#pragma unroll
    for (int k = 0; k < K; ++k) dst[i * K + k] = src[j * K + k];
  }
}

// rho = deposit(pos) -> force grids (3, G^3)
static int pm_force_grids(const float* pos, const float* mass, float mass_value, int64_t n, int64_t G, float L,
                          float Gconst, PmBufs& b, hipStream_t s) {
  const float h = (float)((double)L / (double)G);
  const float scale = (float)(1.0 / ((double)h * (double)h * (double)h));
  const size_t dep_ws = augur_cic_deposit_workspace_bytes(1, 3, G, 1);
  // The deposit's own workspace layout is (fixed grid, maxbits): b.fixed and
  // b.maxbits are carved consecutively with the same alignment.
  int rc = augur_cic_deposit(pos, nullptr, mass, mass_value, nullptr, 1, n, 3, G, 0.0f, h, 1, scale, b.rho, b.fixed,
                             dep_ws, s);
  if (rc) return rc;
  const int64_t shape[3] = {G, G, G};
  rc = augur_fft_r2c(b.rho, nullptr, b.spec, 3, shape, 1, nullptr, 0, s);
  if (rc) return rc;
  const float inv = (float)(1.0 / ((double)G * (double)G * (double)G));
  rc = augur_poisson_green_grad(b.spec, 1, G, L, Gconst, inv, b.aspec, nullptr, s);
  if (rc) return rc;
  return augur_fft_c2r(b.aspec, b.force, nullptr, 3, shape, 3, 1.0f, nullptr, 0, s);
}

static int check_pm(int64_t n, int64_t G, float L) {
  AUGUR_REQUIRE(n >= 1 && n < ((int64_t)1 << 31), AUGUR_EINVAL, "pm: need 1 <= n < 2^31");
  AUGUR_REQUIRE(is_pow2(G) && G >= 8 && G <= 512, AUGUR_EUNSUPPORTED, "pm: G must be a power of two in [8, 512]");
  AUGUR_REQUIRE(L > 0.0f, AUGUR_EINVAL, "pm: L must be positive");
  return AUGUR_OK;
}

}  // namespace augur

extern "C" {

size_t augur_pm_workspace_bytes(int64_t n, int64_t G) {
  const int64_t G3 = G * G * G, nc = G * G * (G / 2 + 1);
  size_t grid = align_up(G3 * 8) + align_up(16) + align_up(G3 * 4) + align_up(nc * 8) + align_up(3 * nc * 8) +
                align_up(3 * G3 * 4);
  size_t sort = 2 * align_up(n * 4) + radix_workspace_bytes(n) + align_up(n * 3 * 4);
  return grid + sort + 4096;
}

int augur_pm_prepare(float* pos, float* vel, float* acc, float* mass, int32_t* ids, int64_t n, int64_t G, float L,
                     void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos, AUGUR_EINVAL, "pm_prepare: null positions");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  carve_pm(c, G, b);
  uint32_t* keys = c.take<uint32_t>(n);
  uint32_t* vals = c.take<uint32_t>(n);
  void* rws = c.take<char>(radix_workspace_bytes(n));
  float* tmp = c.take<float>(n * 3);
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "pm_prepare: workspace too small (%zu < %zu)", ws_bytes, c.used);
  const float h = (float)((double)L / (double)G);
  hipLaunchKernelGGL(k_base_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, n, (int)G, h, keys, vals);
  AUGUR_LAUNCHED("base_keys");
  uint32_t *ks, *perm;
  rc = radix_sort_pairs(keys, vals, n, 3 * ilog2(G), rws, radix_workspace_bytes(n), &ks, &perm, s);
  if (rc) return rc;
  const dim3 g(blocks_for(n, 256)), bl(256);
  float* arrs3[3] = {pos, vel, acc};
  for (float* a : arrs3) {
    if (!a) continue;
    hipLaunchKernelGGL((k_permute<float, 3>), g, bl, 0, s, a, tmp, perm, n);
    AUGUR_LAUNCHED("permute3");
    AUGUR_HIP(hipMemcpyAsync(a, tmp, n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  if (mass) {
    hipLaunchKernelGGL((k_permute<float, 1>), g, bl, 0, s, mass, tmp, perm, n);
    AUGUR_LAUNCHED("permute1");
    AUGUR_HIP(hipMemcpyAsync(mass, tmp, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  if (ids) {
    hipLaunchKernelGGL((k_permute<int32_t, 1>), g, bl, 0, s, ids, (int32_t*)tmp, perm, n);
    AUGUR_LAUNCHED("permute_ids");
    AUGUR_HIP(hipMemcpyAsync(ids, tmp, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  }
  return AUGUR_OK;
}

int augur_pm_accel(const float* pos, float* acc, const float* mass, float mass_value, int64_t n, int64_t G, float L,
                   float Gconst, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && acc, AUGUR_EINVAL, "pm_accel: null pointer");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, G, b), AUGUR_EWORKSPACE, "pm_accel: workspace too small (%zu < %zu)", ws_bytes, c.used);
  rc = pm_force_grids(pos, mass, mass_value, n, G, L, Gconst, b, s);
  if (rc) return rc;
  const float h = (float)((double)L / (double)G);
  TimedScope ts_("gather", s);
  hipLaunchKernelGGL(k_gather_kdk, dim3(blocks_for(n, 256)), dim3(256), 0, s, (float*)pos, (float*)nullptr, acc,
                     b.force, n, (int)G, h, 0.0f, 0.0f, L, 0, 0);
  AUGUR_LAUNCHED("gather_kdk");
  return AUGUR_OK;
}

int augur_pm_kdk(float* pos, float* vel, float* acc, const float* mass, float mass_value, int64_t n, int64_t G,
                 float L, float Gconst, float dt, int nsteps, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && vel && acc && nsteps >= 0, AUGUR_EINVAL, "pm_kdk: bad arguments");
  if (nsteps == 0) return AUGUR_OK;
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, G, b), AUGUR_EWORKSPACE, "pm_kdk: workspace too small (%zu < %zu)", ws_bytes, c.used);
  const float h = (float)((double)L / (double)G);
  const float hdt = 0.5f * dt;
  {
    TimedScope ts_("kick_drift", s);
    hipLaunchKernelGGL(k_kick_drift, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, vel, acc, n, hdt, dt, L);
    AUGUR_LAUNCHED("kick_drift");
  }
  for (int st = 0; st < nsteps; ++st) {
    rc = pm_force_grids(pos, mass, mass_value, n, G, L, Gconst, b, s);
    if (rc) return rc;
    const int last = st == nsteps - 1;
    TimedScope ts_("gather_kdk", s);
    hipLaunchKernelGGL(k_gather_kdk, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, vel, acc, b.force, n, (int)G, h,
                       hdt, dt, L, 1, last ? 0 : 1);
    AUGUR_LAUNCHED("gather_kdk");
  }
  return AUGUR_OK;
}

}  // extern "C"
