// FP64 oracles on the GPU: the softened direct sum (isolated) and the Ewald
// sum (periodic cube, mean density removed).  Work is split over
// (target block, source slice); each block writes a partial sum and a second
// pass adds the slices in a fixed order, so results are run-to-run identical.
#include <mutex>

#include "common.hpp"

using namespace augur;

namespace augur {

constexpr int DS_THREADS = 256;
constexpr int DS_TILE = 512;  // sources staged in LDS per step

inline int ds_splits(int64_t t, int64_t n) {
  const int64_t tb = (t + DS_THREADS - 1) / DS_THREADS;
  int64_t s = 2048 / (tb > 0 ? tb : 1);
  const int64_t max_by_n = (n + DS_TILE - 1) / DS_TILE;
  if (s > max_by_n) s = max_by_n;
  if (s > 256) s = 256;
  if (s < 1) s = 1;
  return (int)s;
}

// POT = false: accelerations (t, D); POT = true: potentials
// phi_t = sum_{j != t} m_j (|x_j - x_t|^2 + eps^2)^-1/2 (t, 1).
template <int D, bool POT>
__global__ __launch_bounds__(DS_THREADS) void k_direct(const double* __restrict__ p64, const float* __restrict__ p32,
                                                       const double* __restrict__ mass, int64_t n,
                                                       const int64_t* __restrict__ targets, int64_t t, double eps2,
                                                       int splits, double* __restrict__ part) {
  __shared__ double sp[DS_TILE * 4];
  const int64_t ti = (int64_t)blockIdx.x * DS_THREADS + threadIdx.x;
  const int split = blockIdx.y;
  const int64_t per = (n + splits - 1) / splits;
  const int64_t lo = split * per, hi = min(n, lo + per);
  int64_t self = -1;
  double xi[3] = {0, 0, 0};
  if (ti < t) {
    self = targets ? targets[ti] : ti;
    for (int d = 0; d < D; ++d) xi[d] = p64 ? p64[self * D + d] : (double)p32[self * D + d];
  }
  double a[3] = {0, 0, 0};
  for (int64_t j0 = lo; j0 < hi; j0 += DS_TILE) {
    const int cnt = (int)min((int64_t)DS_TILE, hi - j0);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt; e += DS_THREADS) {
      const int64_t j = j0 + e;
      for (int d = 0; d < D; ++d) sp[e * 4 + d] = p64 ? p64[j * D + d] : (double)p32[j * D + d];
      sp[e * 4 + 3] = mass ? mass[j] : 1.0;
    }
    __syncthreads();
    if (ti < t) {
      for (int e = 0; e < cnt; ++e) {
        const double dx = sp[e * 4] - xi[0], dy = sp[e * 4 + 1] - xi[1];
        const double dz = D == 3 ? sp[e * 4 + 2] - xi[2] : 0.0;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double rs = 1.0 / sqrt(r2);
        if (POT) {
          if (j0 + e != self) a[0] += sp[e * 4 + 3] * rs;
          continue;
        }
        double w = sp[e * 4 + 3] * rs * rs * rs;
        if (j0 + e == self) w = 0.0;
        a[0] += dx * w;
        a[1] += dy * w;
        a[2] += dz * w;
      }
    }
  }
  constexpr int K = POT ? 1 : D;
  if (ti < t)
    for (int d = 0; d < K; ++d) part[((int64_t)split * t + ti) * K + d] = a[d];
}

__global__ void k_sum_splits(const double* __restrict__ part, int splits, int64_t m, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < splits; ++k) s += part[(int64_t)k * m + i];
    out[i] = s;
  }
}

// ------------------------------------------------------------------ Ewald
constexpr double PI = 3.14159265358979323846;

// (erf(s) - 2 s exp(-s^2)/sqrt(pi)) / s^3, finite at s = 0
__device__ double ewald_g_over_s3(double s) {
  const double c = 2.0 / sqrt(PI);
  if (s < 0.5) {
    // sum_{n>=1} (-1)^(n+1) 2n s^(2n-2) / (n! (2n+1))
    double s2 = s * s, term = 1.0, sum = 0.0, fact = 1.0;
    for (int k = 1; k <= 12; ++k) {
      fact *= k;
      const double cterm = (double)(2 * k) / (fact * (2 * k + 1));
      sum += ((k & 1) ? 1.0 : -1.0) * cterm * term;
      term *= s2;
    }
    return c * sum;
  }
  return (erf(s) - c * s * exp(-s * s)) / (s * s * s);
}

__global__ __launch_bounds__(DS_THREADS) void k_ewald_real(const float* __restrict__ p32,
                                                           const double* __restrict__ mass, double mval, int64_t n,
                                                           const int64_t* __restrict__ targets, int64_t t, double L,
                                                           double eps2, double alpha, int splits,
                                                           double* __restrict__ part) {
  __shared__ double sp[DS_TILE * 4];
  const int64_t ti = (int64_t)blockIdx.x * DS_THREADS + threadIdx.x;
  const int split = blockIdx.y;
  const int64_t per = (n + splits - 1) / splits;
  const int64_t lo = split * per, hi = min(n, lo + per);
  int64_t self = -1;
  double xi[3] = {0, 0, 0};
  if (ti < t) {
    self = targets ? targets[ti] : ti;
    for (int d = 0; d < 3; ++d) xi[d] = (double)p32[self * 3 + d];
  }
  const double a3 = alpha * alpha * alpha, invL = 1.0 / L;
  double a[3] = {0, 0, 0};
  for (int64_t j0 = lo; j0 < hi; j0 += DS_TILE) {
    const int cnt = (int)min((int64_t)DS_TILE, hi - j0);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt; e += DS_THREADS) {
      const int64_t j = j0 + e;
      for (int d = 0; d < 3; ++d) sp[e * 4 + d] = (double)p32[j * 3 + d];
      sp[e * 4 + 3] = mass ? mass[j] : mval;
    }
    __syncthreads();
    if (ti < t) {
      for (int e = 0; e < cnt; ++e) {
        if (j0 + e == self) continue;
        // r0 = x_i - x_j, minimum image
        double r[3];
        for (int d = 0; d < 3; ++d) {
          double v = xi[d] - sp[e * 4 + d];
          v -= L * rint(v * invL);
          r[d] = v;
        }
        const double rr2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        const double rr = sqrt(rr2);
        const double q = rr2 + eps2;
        const double soft = 1.0 / (q * sqrt(q));
        // F = -r / (r^2 + eps^2)^1.5 + r * alpha^3 g(s)/s^3
        const double f = -soft + a3 * ewald_g_over_s3(alpha * rr);
        const double m = sp[e * 4 + 3];
        a[0] += m * f * r[0];
        a[1] += m * f * r[1];
        a[2] += m * f * r[2];
      }
    }
  }
  if (ti < t)
    for (int d = 0; d < 3; ++d) part[((int64_t)split * t + ti) * 3 + d] = a[d];
}

// Half-space k vectors: index -> (nx, ny, nz) with (nz > 0) or (nz = 0, ny > 0)
// or (nz = ny = 0, nx > 0).
__device__ __forceinline__ void kvec(int idx, int kmax, int* nv) {
  const int side = 2 * kmax + 1;
  // enumerate the full cube in order and take the second half (after the origin)
  const int full = idx + (side * side * side + 1) / 2;
  nv[0] = full % side - kmax;
  nv[1] = (full / side) % side - kmax;
  nv[2] = full / (side * side) - kmax;
}

__global__ __launch_bounds__(DS_THREADS) void k_ewald_sk(const float* __restrict__ p32, const double* __restrict__ mass,
                                                         double mval, int64_t n, int nk, int kmax, double kf,
                                                         int splits, double* __restrict__ part) {
  __shared__ double sp[DS_TILE * 4];
  const int ki = blockIdx.x * DS_THREADS + threadIdx.x;
  const int split = blockIdx.y;
  const int64_t per = (n + splits - 1) / splits;
  const int64_t lo = split * per, hi = min(n, lo + per);
  double kv[3] = {0, 0, 0};
  if (ki < nk) {
    int nv[3];
    kvec(ki, kmax, nv);
    for (int d = 0; d < 3; ++d) kv[d] = kf * nv[d];
  }
  double C = 0.0, S = 0.0;
  for (int64_t j0 = lo; j0 < hi; j0 += DS_TILE) {
    const int cnt = (int)min((int64_t)DS_TILE, hi - j0);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt; e += DS_THREADS) {
      const int64_t j = j0 + e;
      for (int d = 0; d < 3; ++d) sp[e * 4 + d] = (double)p32[j * 3 + d];
      sp[e * 4 + 3] = mass ? mass[j] : mval;
    }
    __syncthreads();
    if (ki < nk) {
      for (int e = 0; e < cnt; ++e) {
        const double ph = kv[0] * sp[e * 4] + kv[1] * sp[e * 4 + 1] + kv[2] * sp[e * 4 + 2];
        double sn, cs;
        sincos(ph, &sn, &cs);
        C += sp[e * 4 + 3] * cs;
        S += sp[e * 4 + 3] * sn;
      }
    }
  }
  if (ki < nk) {
    part[((int64_t)split * nk + ki) * 2] = C;
    part[((int64_t)split * nk + ki) * 2 + 1] = S;
  }
}

__global__ __launch_bounds__(DS_THREADS) void k_ewald_recip(const float* __restrict__ p32,
                                                            const int64_t* __restrict__ targets, int64_t t,
                                                            const double* __restrict__ sk, int nk, int kmax, double kf,
                                                            double alpha, double V, double* __restrict__ out) {
  const int64_t ti = (int64_t)blockIdx.x * DS_THREADS + threadIdx.x;
  if (ti >= t) return;
  const int64_t self = targets ? targets[ti] : ti;
  double xi[3];
  for (int d = 0; d < 3; ++d) xi[d] = (double)p32[self * 3 + d];
  double a[3] = {0, 0, 0};
  const double pref = -2.0 * 4.0 * PI / V;  // factor 2: the half space stands for +-k
  for (int ki = 0; ki < nk; ++ki) {
    int nv[3];
    kvec(ki, kmax, nv);
    const double k0 = kf * nv[0], k1 = kf * nv[1], k2v = kf * nv[2];
    const double k2 = k0 * k0 + k1 * k1 + k2v * k2v;
    const double fac = exp(-k2 / (4.0 * alpha * alpha)) / k2;
    double sn, cs;
    sincos(k0 * xi[0] + k1 * xi[1] + k2v * xi[2], &sn, &cs);
    const double w = pref * fac * (sn * sk[2 * ki] - cs * sk[2 * ki + 1]);
    a[0] += w * k0;
    a[1] += w * k1;
    a[2] += w * k2v;
  }
  for (int d = 0; d < 3; ++d) out[ti * 3 + d] += a[d];
}

}  // namespace augur

extern "C" {

size_t augur_direct_sum_workspace_bytes(int64_t n, int64_t t, int D) {
  return align_up((size_t)ds_splits(t, n) * t * D * sizeof(double)) + 256;
}

static int direct_impl(bool pot, const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                       const int64_t* targets, int64_t t, double eps, double* out, void* ws, size_t ws_bytes,
                       void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "direct_sum: D must be 2 or 3");
  AUGUR_REQUIRE(n >= 1 && t >= 0, AUGUR_EINVAL, "direct_sum: bad sizes");
  AUGUR_REQUIRE((pos64 != nullptr) != (pos32 != nullptr), AUGUR_EINVAL, "direct_sum: pass exactly one of pos64/pos32");
  if (t == 0) return AUGUR_OK;
  AUGUR_REQUIRE(out, AUGUR_EINVAL, "direct_sum: null output");
  const int splits = ds_splits(t, n);
  const int K = pot ? 1 : D;
  AUGUR_REQUIRE(ws && ws_bytes >= (size_t)splits * t * D * sizeof(double), AUGUR_EWORKSPACE,
                "direct_sum: workspace too small");
  hipStream_t s = as_stream(stream);
  dim3 g((unsigned)((t + DS_THREADS - 1) / DS_THREADS), (unsigned)splits);
  const double e2 = eps * eps;
  double* part = (double*)ws;
  if (D == 2 && !pot)
    hipLaunchKernelGGL((k_direct<2, false>), g, dim3(DS_THREADS), 0, s, pos64, pos32, mass, n, targets, t, e2, splits, part);
  else if (D == 3 && !pot)
    hipLaunchKernelGGL((k_direct<3, false>), g, dim3(DS_THREADS), 0, s, pos64, pos32, mass, n, targets, t, e2, splits, part);
  else if (D == 2)
    hipLaunchKernelGGL((k_direct<2, true>), g, dim3(DS_THREADS), 0, s, pos64, pos32, mass, n, targets, t, e2, splits, part);
  else
    hipLaunchKernelGGL((k_direct<3, true>), g, dim3(DS_THREADS), 0, s, pos64, pos32, mass, n, targets, t, e2, splits, part);
  AUGUR_LAUNCHED("direct");
  hipLaunchKernelGGL(k_sum_splits, dim3(blocks_for(t * K, 256)), dim3(256), 0, s, (const double*)ws, splits, t * K,
                     out);
  AUGUR_LAUNCHED("sum_splits");
  return AUGUR_OK;
}

int augur_direct_sum_f64(const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                         const int64_t* targets, int64_t t, double eps, double* out, void* ws, size_t ws_bytes,
                         void* stream) {
  return direct_impl(false, pos64, pos32, mass, n, D, targets, t, eps, out, ws, ws_bytes, stream);
}

// ------------------------------------------------ no-allocation variant
// The partial sums of augur_direct_sum_f64 in a static device array instead
// of a caller workspace, for callers that cannot allocate.  Every call on a
// device shares that device's copy; each partial it reads is written earlier
// in the same call, so nothing needs clearing, but calls that overlap on one
// device race.
constexpr size_t DS_SCRATCH_DOUBLES = (size_t)1 << 21;  // 16 MiB per device
static __device__ __align__(256) double scratch[DS_SCRATCH_DOUBLES];

constexpr int DS_MAX_DEVICES = 64;
static void* g_scratch_ptr[DS_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// the device address of this device's `scratch` (looked up once per device,
// with that device current); host code never uses the array's name itself
static int scratch_address(double** out) {
  int dev = -1;
  AUGUR_HIP(hipGetDevice(&dev));
  AUGUR_REQUIRE(dev >= 0 && dev < DS_MAX_DEVICES, AUGUR_EUNSUPPORTED, "direct_sum_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("direct_sum_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return AUGUR_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<double*>(g_scratch_ptr[dev]);
  return AUGUR_OK;
}

int augur_direct_sum_f64_noalloc(const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                                 const int64_t* targets, int64_t t, double eps, double* out, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "direct_sum_noalloc: D must be 2 or 3");
  AUGUR_REQUIRE(n >= 1 && t >= 0, AUGUR_EINVAL, "direct_sum_noalloc: bad sizes");
  if (t == 0) return AUGUR_OK;
  const size_t need = (size_t)ds_splits(t, n) * (size_t)t * (size_t)D;
  AUGUR_REQUIRE(need <= DS_SCRATCH_DOUBLES, AUGUR_EINVAL,
                "direct_sum_noalloc: %zu partial sums exceed the static scratch of %zu doubles (at most 699050 "
                "targets in 3D); use augur_direct_sum_f64 with a workspace",
                need, DS_SCRATCH_DOUBLES);
  double* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  return direct_impl(false, pos64, pos32, mass, n, D, targets, t, eps, out, ws,
                     DS_SCRATCH_DOUBLES * sizeof(double), stream);
}

int augur_direct_potential_f64(const double* pos64, const float* pos32, const double* mass, int64_t n, int D,
                               const int64_t* targets, int64_t t, double eps, double* out, void* ws, size_t ws_bytes,
                               void* stream) {
  return direct_impl(true, pos64, pos32, mass, n, D, targets, t, eps, out, ws, ws_bytes, stream);
}

static int ewald_nk(int kmax) {
  const int side = 2 * kmax + 1;
  return (side * side * side - 1) / 2;
}

size_t augur_ewald_workspace_bytes(int64_t n, int64_t t, int kmax) {
  const int nk = ewald_nk(kmax);
  const int s_real = ds_splits(t, n), s_k = ds_splits(nk, n);
  return align_up((size_t)s_real * t * 3 * sizeof(double)) + align_up((size_t)s_k * nk * 2 * sizeof(double)) +
         align_up((size_t)nk * 2 * sizeof(double)) + 1024;
}

int augur_ewald_f64(const float* pos32, const double* mass, double mass_value, int64_t n, const int64_t* targets,
                    int64_t t, double L, double eps, double alpha_l, int kmax, double* out, void* ws, size_t ws_bytes,
                    void* stream) {
  AUGUR_REQUIRE(n >= 1 && t >= 0 && L > 0 && alpha_l > 0 && kmax >= 1 && kmax <= 64, AUGUR_EINVAL,
                "ewald: bad arguments");
  if (t == 0) return AUGUR_OK;
  AUGUR_REQUIRE(pos32 && out, AUGUR_EINVAL, "ewald: null pointer");
  const int nk = ewald_nk(kmax);
  const int s_real = ds_splits(t, n), s_k = ds_splits(nk, n);
  Carver c(ws, ws_bytes);
  double* preal = c.take<double>((size_t)s_real * t * 3);
  double* pk = c.take<double>((size_t)s_k * nk * 2);
  double* sk = c.take<double>((size_t)nk * 2);
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "ewald: workspace too small (%zu < %zu)", ws_bytes, c.used);
  hipStream_t s = as_stream(stream);
  const double alpha = alpha_l / L, kf = 2.0 * PI / L;
  hipLaunchKernelGGL(k_ewald_real, dim3((unsigned)((t + DS_THREADS - 1) / DS_THREADS), (unsigned)s_real),
                     dim3(DS_THREADS), 0, s, pos32, mass, mass_value, n, targets, t, L, eps * eps, alpha, s_real,
                     preal);
  AUGUR_LAUNCHED("ewald_real");
  hipLaunchKernelGGL(k_sum_splits, dim3(blocks_for(t * 3, 256)), dim3(256), 0, s, (const double*)preal, s_real, t * 3,
                     out);
  AUGUR_LAUNCHED("sum_splits");
  hipLaunchKernelGGL(k_ewald_sk, dim3((unsigned)((nk + DS_THREADS - 1) / DS_THREADS), (unsigned)s_k),
                     dim3(DS_THREADS), 0, s, pos32, mass, mass_value, n, nk, kmax, kf, s_k, pk);
  AUGUR_LAUNCHED("ewald_sk");
  hipLaunchKernelGGL(k_sum_splits, dim3(blocks_for((int64_t)nk * 2, 256)), dim3(256), 0, s, (const double*)pk, s_k,
                     (int64_t)nk * 2, sk);
  AUGUR_LAUNCHED("sum_splits");
  hipLaunchKernelGGL(k_ewald_recip, dim3((unsigned)((t + DS_THREADS - 1) / DS_THREADS)), dim3(DS_THREADS), 0, s,
                     pos32, targets, t, (const double*)sk, nk, kmax, kf, alpha, L * L * L, out);
  AUGUR_LAUNCHED("ewald_recip");
  return AUGUR_OK;
}

}  // extern "C"
