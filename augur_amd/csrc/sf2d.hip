// The reference's 2D scatter-field step on gfx950: the grid net's 3x3
// convolutions (UNet and LocalNet, FP32, fused bias + GELU, with the 2x2
// average pool, the nearest x2 upsampling and the skip concatenation folded
// into the input loads), the learned kernel grid of the FFT force, the
// finite-difference gradient (zero and "ext" padding, with the kernel
// acceleration added), the cell-list kNN pair term, the momentum fix and the
// Verlet / plain step updates.  FP32 arithmetic with erf GELU, as the
// reference's FP32 modules (scripts/scatter_field.py:15-67,263-289,291-359,
// 551-555,591-688).
#include "internal.hpp"

using namespace augur;

namespace {

__device__ __forceinline__ float gelu_erf(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

// The two-hidden-layer MLP nn.Sequential(Linear(2, 64), GELU, Linear(64, 64),
// GELU, Linear(64, 1)) of the reference (kmlp and ppmlp, scripts/scatter_field.py:
// 120-121,128-130) at one input (a, b).  Parameters packed in the modules'
// order: W0 [64][2], b0 [64], W1 [64][64], b1 [64], W2 [64], b2 [1]; the
// indices are wave-uniform, so the weights come through the scalar cache.
constexpr int MLP_H = 64;
constexpr int MLP_FLOATS = MLP_H * 2 + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;

__device__ float mlp2(const float* __restrict__ p, float a, float b) {
  const float* W0 = p;
  const float* b0 = W0 + 2 * MLP_H;
  const float* W1 = b0 + MLP_H;
  const float* b1 = W1 + MLP_H * MLP_H;
  const float* W2 = b1 + MLP_H;
  const float b2 = W2[MLP_H];
  float h1[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; ++i) h1[i] = gelu_erf(fmaf(W0[2 * i], a, fmaf(W0[2 * i + 1], b, b0[i])));
  float out = b2;
#pragma unroll 4
  for (int i = 0; i < MLP_H; ++i) {
    float s = b1[i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) s = fmaf(W1[i * MLP_H + j], h1[j], s);
    out = fmaf(W2[i], gelu_erf(s), out);
  }
  return out;
}

// ------------------------------------------------------------------ conv
// out[b][co][y][x] = act(bias[co] + sum_{ci, dy, dx} w[co][ci][dy][dx] in[b][ci][y+dy-1][x+dx-1])
// with zero padding.  Input channels ci < ca come from `a` through amode
// (0: same size; 1: a is 2H x 2W, 2x2 average; 2: a is H/2 x W/2, nearest),
// channels ca.. from `bsrc` (same size).  act: 0 none, 1 GELU, 2 res + GELU.
// A workgroup computes a 16 x 16 pixel tile for 8 output channels; input
// channels are staged through LDS 8 at a time.
constexpr int CV_T = 16, CV_CO = 8, CV_CI = 8, CV_P = CV_T + 2;

__global__ __launch_bounds__(256) void k_conv2d(const float* __restrict__ a, int ca, int amode,
                                                const float* __restrict__ bsrc, int cb,
                                                const float* __restrict__ w, const float* __restrict__ bias, int H,
                                                int W, int cout, int act, const float* __restrict__ res,
                                                float* __restrict__ out) {
  __shared__ float tile[CV_CI][CV_P][CV_P];
  const int cin = ca + cb;
  const int ntx = (W + CV_T - 1) / CV_T;
  const int tx0 = (blockIdx.x % ntx) * CV_T, ty0 = (blockIdx.x / ntx) * CV_T;
  const int co0 = blockIdx.y * CV_CO;
  const int b = blockIdx.z;
  const int tid = threadIdx.x, lx = tid % CV_T, ly = tid / CV_T;
  const int64_t HW = (int64_t)H * W;
  const int Ha = amode == 1 ? 2 * H : (amode == 2 ? H / 2 : H);
  const int Wa = amode == 1 ? 2 * W : (amode == 2 ? W / 2 : W);
  const float* ab = a + (int64_t)b * ca * Ha * Wa;
  const float* bb = bsrc ? bsrc + (int64_t)b * cb * HW : nullptr;
  float acc[CV_CO];
#pragma unroll
  for (int k = 0; k < CV_CO; ++k) acc[k] = 0.0f;
  for (int c0 = 0; c0 < cin; c0 += CV_CI) {
    for (int e = tid; e < CV_CI * CV_P * CV_P; e += 256) {
      const int c = c0 + e / (CV_P * CV_P);
      const int r = e % (CV_P * CV_P);
      const int y = ty0 + r / CV_P - 1, x = tx0 + r % CV_P - 1;
      float v = 0.0f;
      if (c < cin && y >= 0 && y < H && x >= 0 && x < W) {
        if (c < ca) {
          const float* ac = ab + (int64_t)c * Ha * Wa;
          if (amode == 1) {
            const float* q = ac + (int64_t)(2 * y) * Wa + 2 * x;
            v = (((q[0] + q[1]) + q[Wa]) + q[Wa + 1]) * 0.25f;
          } else if (amode == 2) {
            v = ac[(int64_t)(y >> 1) * Wa + (x >> 1)];
          } else {
            v = ac[(int64_t)y * Wa + x];
          }
        } else {
          v = bb[(int64_t)(c - ca) * HW + (int64_t)y * W + x];
        }
      }
      tile[e / (CV_P * CV_P)][r / CV_P][r % CV_P] = v;
    }
    __syncthreads();
    const int nc = cin - c0 < CV_CI ? cin - c0 : CV_CI;
    for (int ci = 0; ci < nc; ++ci) {
      float in[9];
#pragma unroll
      for (int t = 0; t < 9; ++t) in[t] = tile[ci][ly + t / 3][lx + t % 3];
#pragma unroll
      for (int k = 0; k < CV_CO; ++k) {
        const int co = co0 + k < cout ? co0 + k : cout - 1;   // uniform clamp; results past cout are dropped
        const float* wk = w + ((int64_t)co * cin + c0 + ci) * 9;
#pragma unroll
        for (int t = 0; t < 9; ++t) acc[k] = fmaf(wk[t], in[t], acc[k]);
      }
    }
    __syncthreads();
  }
  const int y = ty0 + ly, x = tx0 + lx;
  if (y >= H || x >= W) return;
#pragma unroll
  for (int k = 0; k < CV_CO; ++k) {
    const int co = co0 + k;
    if (co >= cout) break;
    float v = acc[k] + (bias ? bias[co] : 0.0f);
    const int64_t o = ((int64_t)b * cout + co) * HW + (int64_t)y * W + x;
    if (act == 1) v = gelu_erf(v);
    if (act == 2) v = res[o] + gelu_erf(v);
    out[o] = v;
  }
}

// ------------------------------------------------------- kernel grid
// K(r) at min-image distances of the (2G)^2 grid: r = sqrt(d_y^2 + d_x^2),
// d = min(i, 2G - i) h; features [r, log(r + h)]; times the split taper
// 1 - (1 - min(r/pp, 1)^2)^2 when pp > 0 (scripts/scatter_field.py:263-272).
__global__ void k_kernel_grid(const float* __restrict__ p, int G, float h, float pp, float* __restrict__ K) {
  const int n2 = 2 * G;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)n2 * n2) return;
  const int iy = (int)(i / n2), ix = (int)(i % n2);
  const float dy = (float)(iy < n2 - iy ? iy : n2 - iy) * h;
  const float dx = (float)(ix < n2 - ix ? ix : n2 - ix) * h;
  const float r = sqrtf(__fadd_rn(__fmul_rn(dy, dy), __fmul_rn(dx, dx)));
  float k = mlp2(p, r, logf(r + h));
  if (pp > 0.0f) {
    const float t = fminf(r / pp, 1.0f);
    const float u = 1.0f - t * t;
    k = k * (1.0f - u * u);
  }
  K[i] = k;
}

// ------------------------------------------------------ gradient
// -(phi[i+1] - phi[i-1]) / (2h) per axis into out (B, 2, G, G) (+ addend).
// ext = 0: phi (B, G, G), zero outside; ext = 1: phi (B, 2G, 2G), the full
// zero-padded convolution, read at rows/columns -1 (= 2G-1) .. G.
__global__ void k_neg_grad(const float* __restrict__ phi, int ext, int64_t B, int G, float h,
                           const float* __restrict__ addend, float* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t GG = (int64_t)G * G;
  if (i >= B * GG) return;
  const int64_t b = i / GG;
  const int y = (int)((i % GG) / G), x = (int)(i % G);
  const int P = ext ? 2 * G : G;
  const float* pb = phi + b * (int64_t)P * P;
  auto at = [&](int yy, int xx) -> float {
    if (ext) return pb[(int64_t)((yy + P) % P) * P + (xx + P) % P];
    return (yy < 0 || yy >= G || xx < 0 || xx >= G) ? 0.0f : pb[(int64_t)yy * G + xx];
  };
  const float inv = 2.0f * h;
  float gx = -((at(y, x + 1) - at(y, x - 1)) / inv);
  float gy = -((at(y + 1, x) - at(y - 1, x)) / inv);
  const int64_t o = b * 2 * GG + (int64_t)y * G + x;
  if (addend) {
    gx += addend[o];
    gy += addend[o + GG];
  }
  out[o] = gx;
  out[o + GG] = gy;
}

// ------------------------------------------------------ pair term
// For every unmasked particle q: the (at most KMAX) nearest unmasked others
// with r = sqrt(|d|^2 + 1e-8) < pp among the 3x3 cells of side pp around
// its cell (clamped cells, as the keys of ScatterField._cell_keys), sorted
// by r; the first min(knn, .) of them add
// ppmlp([r, log(r + 0.05)]) (1 - (r/pp)^2)^2 m_j d / r to acc[q].  Every
// pair with r < pp lies in those cells, so this is the dense pair term.
constexpr int KMAX = 16;

__global__ __launch_bounds__(128) void k_pair_knn(const float* __restrict__ pos, const float* __restrict__ mass,
                                                  const float* __restrict__ mask, int64_t B, int64_t N,
                                                  const int64_t* __restrict__ perm,
                                                  const int64_t* __restrict__ starts, int C, float off, float cell,
                                                  float pp, int knn, const float* __restrict__ mlp,
                                                  float* __restrict__ acc) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= B * N) return;
  if (mask && mask[q] <= 0.0f) return;
  const int64_t b = q / N;
  const float px = pos[2 * q], py = pos[2 * q + 1];
  int cx = (int)floorf(__fdiv_rn(__fadd_rn(px, off), cell));
  int cy = (int)floorf(__fdiv_rn(__fadd_rn(py, off), cell));
  cx = cx < 0 ? 0 : (cx > C - 1 ? C - 1 : cx);
  cy = cy < 0 ? 0 : (cy > C - 1 ? C - 1 : cy);
  float rk[KMAX];
  int64_t jk[KMAX];
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    rk[t] = INFINITY;
    jk[t] = -1;
  }
  for (int ny = cy - 1; ny <= cy + 1; ++ny) {
    if (ny < 0 || ny >= C) continue;
    for (int nx = cx - 1; nx <= cx + 1; ++nx) {
      if (nx < 0 || nx >= C) continue;
      const int64_t c = b * C * C + (int64_t)ny * C + nx;
      for (int64_t s = starts[c], e = starts[c + 1]; s < e; ++s) {
        const int64_t j = perm[s];
        if (j == q) continue;
        const float dx = pos[2 * j] - px, dy = pos[2 * j + 1] - py;
        const float r = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), 1e-8f));
        if (!(r < pp) || !(r < rk[KMAX - 1])) continue;
        rk[KMAX - 1] = r;
        jk[KMAX - 1] = j;
#pragma unroll
        for (int t = KMAX - 1; t > 0; --t) {
          if (rk[t] < rk[t - 1]) {
            const float tr = rk[t];
            rk[t] = rk[t - 1];
            rk[t - 1] = tr;
            const int64_t tj = jk[t];
            jk[t] = jk[t - 1];
            jk[t - 1] = tj;
          }
        }
      }
    }
  }
  float ax = 0.0f, ay = 0.0f;
#pragma unroll
  for (int t = 0; t < KMAX; ++t) {
    if (t >= knn || jk[t] < 0) break;
    const int64_t j = jk[t];
    const float r = rk[t];
    const float g = mlp2(mlp, r, logf(r + 0.05f));
    const float u = fminf(r / pp, 1.0f);
    const float win = (1.0f - u * u) * (1.0f - u * u);
    const float mj = mass ? mass[j] * (mask ? mask[j] : 1.0f) : (mask ? mask[j] : 1.0f);
    const float wgt = g * win * mj;
    ax += (wgt * (pos[2 * j] - px)) / r;
    ay += (wgt * (pos[2 * j + 1] - py)) / r;
  }
  acc[2 * q] += ax;
  acc[2 * q + 1] += ay;
}

// ------------------------------------------------------ momentum fix
// a -= sum(m a) / max(sum m, 1e-9) per scene (m = mass * mask, FP64 sums in
// a fixed order), then a *= mask (scripts/scatter_field.py:612-615).
constexpr int MF_T = 256;

__global__ __launch_bounds__(MF_T) void k_momfix(float* __restrict__ a, const float* __restrict__ mass,
                                                 const float* __restrict__ mask, int64_t N, int D, int fix) {
  const int64_t b = blockIdx.x;
  float* ab = a + b * N * D;
  const float* mb = mass ? mass + b * N : nullptr;
  const float* kb = mask ? mask + b * N : nullptr;
  __shared__ double red[MF_T / 64][4];
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (fix) {
    for (int64_t i = threadIdx.x; i < N; i += MF_T) {
      const double m = (double)(mb ? mb[i] : 1.0f) * (double)(kb ? kb[i] : 1.0f);
      for (int d = 0; d < D; ++d) s[d] += m * (double)ab[i * D + d];
      s[3] += m;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_down(s[k], o);
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < 4; ++k) red[threadIdx.x >> 6][k] = s[k];
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
      double t = 0.0;
      for (int w = 0; w < MF_T / 64; ++w) t += red[w][k];
      s[k] = t;
    }
  }
  const double msum = s[3] > 1e-9 ? s[3] : 1e-9;
  float mean[3];
  for (int d = 0; d < D; ++d) mean[d] = fix ? (float)(s[d] / msum) : 0.0f;
  for (int64_t i = threadIdx.x; i < N; i += MF_T) {
    const float k = kb ? kb[i] : 1.0f;
    for (int d = 0; d < D; ++d) ab[i * D + d] = (ab[i * D + d] - mean[d]) * k;
  }
}

// ------------------------------------------------------ step updates
// phase 0 (Verlet drift): pos += dt v + (dt^2 / 2) a_prev; vh = v + dt a_prev
// phase 1 (Verlet close): v += (dt / 2) (a_prev + a_new); dv = (dt / 2) (a_prev + a_new)
// phase 2 (plain):        dv = dt a; pos += dt (v + dv / 2); v += dv
// in the reference's operation order (scripts/scatter_field.py:677-688), no contraction.
__global__ void k_sf_update(float* __restrict__ pos, float* __restrict__ vel, const float* __restrict__ a0,
                            const float* __restrict__ a1, float* __restrict__ aux, int64_t n, float dt, float c2,
                            int phase) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (phase == 0) {
    const float p = __fadd_rn(__fadd_rn(pos[i], __fmul_rn(dt, vel[i])), __fmul_rn(c2, a0[i]));
    aux[i] = __fadd_rn(vel[i], __fmul_rn(dt, a0[i]));
    pos[i] = p;
  } else if (phase == 1) {
    const float dv = __fmul_rn(0.5f * dt, __fadd_rn(a0[i], a1[i]));
    if (aux) aux[i] = dv;
    vel[i] = __fadd_rn(vel[i], dv);
  } else {
    const float dv = __fmul_rn(a0[i], dt);
    pos[i] = __fadd_rn(pos[i], __fmul_rn(dt, __fadd_rn(vel[i], __fmul_rn(0.5f, dv))));
    vel[i] = __fadd_rn(vel[i], dv);
    if (aux) aux[i] = dv;
  }
}

}  // namespace

extern "C" {

int augur_conv2d_f32(const float* a, int ca, int amode, const float* b, int cb, const float* w, const float* bias,
                     int64_t B, int64_t H, int64_t W, int cout, int act, const float* res, float* out, void* stream) {
  AUGUR_REQUIRE(a && w && out && ca >= 1 && cb >= 0 && cout >= 1 && B >= 1 && H >= 1 && W >= 1, AUGUR_EINVAL,
                "conv2d: bad arguments");
  AUGUR_REQUIRE(cb == 0 || b, AUGUR_EINVAL, "conv2d: cb > 0 needs b");
  AUGUR_REQUIRE(amode >= 0 && amode <= 2 && act >= 0 && act <= 2 && (act != 2 || res), AUGUR_EINVAL,
                "conv2d: bad amode/act (act 2 needs res)");
  AUGUR_REQUIRE(amode != 2 || (H % 2 == 0 && W % 2 == 0), AUGUR_EINVAL, "conv2d: upsampling needs even H, W");
  AUGUR_REQUIRE(B < 65536 && (cout + CV_CO - 1) / CV_CO < 65536, AUGUR_EINVAL, "conv2d: grid too large");
  hipStream_t s = as_stream(stream);
  const int64_t tiles = ((W + CV_T - 1) / CV_T) * ((H + CV_T - 1) / CV_T);
  AUGUR_REQUIRE(tiles < ((int64_t)1 << 31), AUGUR_EINVAL, "conv2d: grid too large");
  TimedScope ts_("conv2d", s);
  hipLaunchKernelGGL(k_conv2d, dim3((unsigned)tiles, (unsigned)((cout + CV_CO - 1) / CV_CO), (unsigned)B), dim3(256),
                     0, s, a, ca, amode, b, cb, w, bias, (int)H, (int)W, cout, act, res, out);
  AUGUR_LAUNCHED("conv2d");
  return AUGUR_OK;
}

int augur_sf_kernel_grid(const float* mlp, int64_t G, float h, float pp, float* K, void* stream) {
  AUGUR_REQUIRE(mlp && K && G >= 1 && G <= 8192 && h > 0.0f, AUGUR_EINVAL, "sf_kernel_grid: bad arguments");
  hipStream_t s = as_stream(stream);
  const int64_t n = 4 * G * G;
  TimedScope ts_("sf_kernel_grid", s);
  hipLaunchKernelGGL(k_kernel_grid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mlp, (int)G, h, pp, K);
  AUGUR_LAUNCHED("sf_kernel_grid");
  return AUGUR_OK;
}

int augur_sf_neg_grad(const float* phi, int ext, int64_t B, int64_t G, float h, const float* addend, float* out,
                      void* stream) {
  AUGUR_REQUIRE(phi && out && B >= 1 && G >= 1 && h > 0.0f && (ext == 0 || ext == 1), AUGUR_EINVAL,
                "sf_neg_grad: bad arguments");
  hipStream_t s = as_stream(stream);
  const int64_t n = B * G * G;
  TimedScope ts_("sf_neg_grad", s);
  hipLaunchKernelGGL(k_neg_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phi, ext, B, (int)G, h, addend,
                     out);
  AUGUR_LAUNCHED("sf_neg_grad");
  return AUGUR_OK;
}

int augur_pair_knn_f32(const float* pos, const float* mass, const float* mask, int64_t B, int64_t N,
                       const int64_t* perm, const int64_t* starts, int C, float off, float cell, float pp, int knn,
                       const float* mlp, float* acc, void* stream) {
  AUGUR_REQUIRE(pos && perm && starts && mlp && acc && B >= 1 && N >= 1 && C >= 1 && cell > 0.0f && pp > 0.0f,
                AUGUR_EINVAL, "pair_knn: bad arguments");
  AUGUR_REQUIRE(knn >= 1 && knn <= KMAX, AUGUR_EUNSUPPORTED, "pair_knn: knn must be in [1, %d]", KMAX);
  hipStream_t s = as_stream(stream);
  TimedScope ts_("pair_knn", s);
  hipLaunchKernelGGL(k_pair_knn, dim3((unsigned)((B * N + 127) / 128)), dim3(128), 0, s, pos, mass, mask, B, N, perm,
                     starts, C, off, cell, pp, knn, mlp, acc);
  AUGUR_LAUNCHED("pair_knn");
  return AUGUR_OK;
}

int augur_sf_momfix(float* acc, const float* mass, const float* mask, int64_t B, int64_t N, int D, int fix,
                    void* stream) {
  AUGUR_REQUIRE(acc && B >= 1 && B < 65536 && N >= 1 && D >= 1 && D <= 3, AUGUR_EINVAL, "sf_momfix: bad arguments");
  hipStream_t s = as_stream(stream);
  TimedScope ts_("sf_momfix", s);
  hipLaunchKernelGGL(k_momfix, dim3((unsigned)B), dim3(MF_T), 0, s, acc, mass, mask, N, D, fix ? 1 : 0);
  AUGUR_LAUNCHED("sf_momfix");
  return AUGUR_OK;
}

int augur_sf_update(float* pos, float* vel, const float* a0, const float* a1, float* aux, int64_t n, float dt,
                    int phase, void* stream) {
  AUGUR_REQUIRE(pos && vel && a0 && n >= 0 && phase >= 0 && phase <= 2, AUGUR_EINVAL, "sf_update: bad arguments");
  AUGUR_REQUIRE(phase != 0 || aux, AUGUR_EINVAL, "sf_update: phase 0 writes aux (the half-step velocity)");
  AUGUR_REQUIRE(phase != 1 || a1, AUGUR_EINVAL, "sf_update: phase 1 needs a1");
  if (n == 0) return AUGUR_OK;
  hipStream_t s = as_stream(stream);
  const float c2 = (float)(0.5 * (double)dt * (double)dt);
  TimedScope ts_("sf_update", s);
  hipLaunchKernelGGL(k_sf_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos, vel, a0, a1, aux, n, dt,
                     c2, phase);
  AUGUR_LAUNCHED("sf_update");
  return AUGUR_OK;
}

}  // extern "C"
