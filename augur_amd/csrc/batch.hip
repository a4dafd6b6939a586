// Batched stepping of many scenes of uneven particle counts: one persistent
// launch applies the CIC gather of every scene's force planes, the KDK half
// kicks and the drift to the whole ragged batch.
//
// A ragged batch concatenates the scenes' particles: scene b is rows
// starts[b] .. starts[b+1]-1 of pos/vel/acc (FP32 (total, 3)), and its force
// field is force[b] = three FP32 planes a_x, a_y, a_z of G^3 cells (the
// layout of the batched PM solve, (B, 3, G, G, G)).  All scenes share the
// periodic box [0, L)^3 and the grid.  Workgroups stride over the particles
// in order (coalesced 12-byte rows); a particle's scene is a binary search of
// the offsets staged in LDS.  The arithmetic per particle is that of the
// single-scene stepper: reference rounding of the CIC base coordinate, corner
// weights (wx wy) wz, and every kick and drift one explicit FMA, so a scene
// stepped in a batch and the same scene stepped alone give the same bits.
#include "internal.hpp"

using namespace augur;
using namespace augur_dev;

namespace augur {

constexpr int RG_THREADS = 256;
constexpr int RG_MAX_SCENES = 1024;  // scene offsets in LDS: 8 KB

enum RaggedPhase : int {
  PHASE_GATHER = 0,      // acc = F(x)
  PHASE_OPEN = 1,        // v += dt/2 acc; x += dt v                  (no gather)
  PHASE_CLOSE = 2,       // a = F(x); v += dt/2 a; acc = a
  PHASE_CLOSE_OPEN = 3,  // a = F(x); v += dt/2 a; v += dt/2 a; x += dt v (acc not stored)
};

__device__ __forceinline__ float wrap_box(float x, float L) {
  if (x >= L) x -= L;
  if (x < 0.0f) x += L;
  if (x >= L) x = 0.0f;  // -tiny + L rounds to L
  return x;
}

template <bool P2>
__global__ __launch_bounds__(RG_THREADS) void k_step_ragged(float* __restrict__ pos, float* __restrict__ vel,
                                                            float* __restrict__ acc, const float* __restrict__ force,
                                                            const int64_t* __restrict__ starts, int B, int64_t total,
                                                            int G, float h, float inv_h, float hdt, float dt, float L,
                                                            int phase, uint32_t* __restrict__ vmax_bits) {
  __shared__ int64_t s_start[RG_MAX_SCENES + 1];
  float vmax = 0.0f;  // largest |v| component of this thread's drifts
  for (int b = threadIdx.x; b <= B; b += RG_THREADS) s_start[b] = starts[b];
  __syncthreads();
  const int64_t G3 = (int64_t)G * G * G;
  const int64_t stride = (int64_t)gridDim.x * RG_THREADS;
  for (int64_t i = blockIdx.x * (int64_t)RG_THREADS + threadIdx.x; i < total; i += stride) {
    float p[3], v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      p[d] = pos[3 * i + d];
      v[d] = phase == PHASE_GATHER ? 0.0f : vel[3 * i + d];
    }
    if (phase == PHASE_OPEN) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float vv = __fmaf_rn(hdt, acc[3 * i + d], v[d]);
        vel[3 * i + d] = vv;
        pos[3 * i + d] = wrap_box(__fmaf_rn(dt, vv, p[d]), L);
        vmax = fmaxf(vmax, fabsf(vv));
      }
      continue;
    }
    // scene of particle i: the largest b with starts[b] <= i (empty scenes
    // share their start with the next scene, which then wins)
    int lo = 0, hi = B - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_start[mid] <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
    const float* fb = force + (int64_t)lo * 3 * G3;
    int ic[3];
    float w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float uu = P2 ? cic_u_pow2(p[d], inv_h) : cic_u(p[d], 0.0f, h);
      const float f = floorf(uu);
      ic[d] = (int)f;
      w[d] = __fsub_rn(uu, f);
    }
    const int x0 = wrap(ic[0], G), x1 = wrap(ic[0] + 1, G);
    const int y0 = wrap(ic[1], G), y1 = wrap(ic[1] + 1, G);
    const int z0 = wrap(ic[2], G), z1 = wrap(ic[2] + 1, G);
    float a[3] = {0.0f, 0.0f, 0.0f};
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
          const int64_t cell = ((int64_t)(oz ? z1 : z0) * G + (oy ? y1 : y0)) * G + (ox ? x1 : x0);
          a[0] += fb[cell] * wt;
          a[1] += fb[G3 + cell] * wt;
          a[2] += fb[2 * G3 + cell] * wt;
        }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (phase == PHASE_GATHER) {
        acc[3 * i + d] = a[d];
        continue;
      }
      float vv = __fmaf_rn(hdt, a[d], v[d]);
      if (phase == PHASE_CLOSE) {
        acc[3 * i + d] = a[d];
      } else {
        vv = __fmaf_rn(hdt, a[d], vv);
        pos[3 * i + d] = wrap_box(__fmaf_rn(dt, vv, p[d]), L);
        vmax = fmaxf(vmax, fabsf(vv));
      }
      vel[3 * i + d] = vv;
    }
  }
  // the drift bound of the tracked caller (a planned ragged deposit): one
  // atomic max per wave on the float bits (non-negative, so ordered as uints)
  if (vmax_bits) {
    for (int o = 32; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
    if ((threadIdx.x & 63) == 0 && vmax > 0.0f) atomicMax(vmax_bits, __float_as_uint(vmax));
  }
}

}  // namespace augur

extern "C" {

static int persistent_step(float* pos, float* vel, float* acc, const int64_t* starts, int64_t B, int64_t total,
                           const float* force, int64_t G, float L, float dt, int phase, uint32_t* vmax_bits,
                           void* stream) {
  AUGUR_REQUIRE(B >= 1 && B <= RG_MAX_SCENES, AUGUR_EINVAL, "persistent_step: B must be in [1, 1024]");
  AUGUR_REQUIRE(total >= 0 && starts && pos, AUGUR_EINVAL, "persistent_step: bad arguments");
  AUGUR_REQUIRE(is_pow2(G) && G >= 2 && G <= 1024, AUGUR_EUNSUPPORTED,
                "persistent_step: G must be a power of two in [2, 1024]");
  AUGUR_REQUIRE(L > 0.0f && dt == dt, AUGUR_EINVAL, "persistent_step: L must be positive and dt finite");
  AUGUR_REQUIRE(phase >= PHASE_GATHER && phase <= PHASE_CLOSE_OPEN, AUGUR_EINVAL,
                "persistent_step: unknown phase %d", phase);
  AUGUR_REQUIRE(acc, AUGUR_EINVAL, "persistent_step: acc is required");
  AUGUR_REQUIRE(phase == PHASE_GATHER || vel, AUGUR_EINVAL, "persistent_step: vel is required for kicks");
  AUGUR_REQUIRE(phase == PHASE_OPEN || force, AUGUR_EINVAL, "persistent_step: force is required to gather");
  if (total == 0) return AUGUR_OK;
  hipStream_t s = as_stream(stream);
  const float h = (float)((double)L / (double)G);
  const bool p2 = pow2_float(h);
  const float inv_h = p2 ? 1.0f / h : 0.0f;
  // persistent grid: eight workgroups per CU, each striding over the batch
  const int64_t cap = (int64_t)device_cu_count() * 8;
  const unsigned blocks = blocks_for(total, RG_THREADS, cap);
  TimedScope ts_("step_ragged", s);
  if (p2)
    hipLaunchKernelGGL((k_step_ragged<true>), dim3(blocks), dim3(RG_THREADS), 0, s, pos, vel, acc, force, starts,
                       (int)B, total, (int)G, h, inv_h, 0.5f * dt, dt, L, phase, vmax_bits);
  else
    hipLaunchKernelGGL((k_step_ragged<false>), dim3(blocks), dim3(RG_THREADS), 0, s, pos, vel, acc, force, starts,
                       (int)B, total, (int)G, h, inv_h, 0.5f * dt, dt, L, phase, vmax_bits);
  AUGUR_LAUNCHED("step_ragged");
  return AUGUR_OK;
}

int augur_persistent_step(float* pos, float* vel, float* acc, const int64_t* starts, int64_t B, int64_t total,
                          const float* force, int64_t G, float L, float dt, int phase, void* stream) {
  return persistent_step(pos, vel, acc, starts, B, total, force, G, L, dt, phase, nullptr, stream);
}

int augur_persistent_step_tracked(float* pos, float* vel, float* acc, const int64_t* starts, int64_t B,
                                  int64_t total, const float* force, int64_t G, float L, float dt, int phase,
                                  uint32_t* drift_word, void* stream) {
  AUGUR_REQUIRE(drift_word, AUGUR_EINVAL, "persistent_step_tracked: null drift word");
  return persistent_step(pos, vel, acc, starts, B, total, force, G, L, dt, phase, drift_word, stream);
}

}  // extern "C"
