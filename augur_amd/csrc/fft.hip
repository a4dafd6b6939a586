// LDS-staged radix-4/2 Stockham FFT for gfx950, batched over lines of a 2D or
// 3D grid (x fastest), with the periodic Green's-function Poisson solve,
// spectral gradient, open-boundary softened kernel and FD gradient.
//
// One workgroup transforms a tile of lines held in LDS:
//   rows pass (x axis):  L = 4096/N contiguous lines, loaded row by row;
//   cols pass (y or z):  W adjacent columns (consecutive x) of N elements,
//                        loaded so that each LDS row is one coalesced segment.
// The in-LDS transform is Stockham autosort (no bit reversal) with a
// per-workgroup twiddle table (fft_core.hpp).
#include "fft_core.hpp"

using namespace augur;

namespace augur {

// ---------------------------------------------------------------- rows R2C
// Real rows of the FFT shape; line l = (b, outer...) with `rows_per_batch`
// rows per batch entry.  A row exists in the input when its outer coordinates
// are inside in_shape; columns beyond in_nx read as zero.
struct RowMap {
  int64_t rows_per_batch;  // prod(shape[:-1])
  int64_t n_rows;          // batch * rows_per_batch
  int ny, nz;              // FFT outer extents (nz = 1 for 2D)
  int in_ny, in_nz, in_nx; // input (or cropped output) extents
  int64_t pitch = 0;       // complex row stride of the half spectrum (0: N/2 + 1)
};

template <int N>
struct RowCfg {
  static constexpr int T = FFT_THREADS;
  static constexpr int NL = FftLines<N, T>::value;
  static constexpr int LS = FftStride<N>::value;
};

template <int N>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows_r2c(const float* __restrict__ in, float2* __restrict__ out,
                                                              RowMap m) {
  using C = RowCfg<N>;
  __shared__ float2 lds[C::NL * C::LS];
  __shared__ float2 tw[N];
  make_twiddles<N>(tw);
  __syncthreads();
  const int64_t row0 = (int64_t)blockIdx.x * C::NL;
  constexpr int nxc = N / 2 + 1;
  auto src = [&](int l, int x) -> float2 {
    const int64_t row = row0 + l;
    float v = 0.0f;
    if (row < m.n_rows && x < m.in_nx) {
      const int64_t b = row / m.rows_per_batch;
      const int64_t r = row - b * m.rows_per_batch;
      const int y = (int)(r % m.ny), z = (int)(r / m.ny);
      if (y < m.in_ny && z < m.in_nz) v = in[((b * m.in_nz + z) * m.in_ny + y) * (int64_t)m.in_nx + x];
    }
    return make_float2(v, 0.0f);
  };
  const int64_t pitch = m.pitch ? m.pitch : nxc;
  auto dst = [&](int l, int k, float2 v) {
    const int64_t row = row0 + l;
    if (row < m.n_rows && k < nxc) out[row * pitch + k] = v;
  };
  fft_lines<N, false>(lds, tw, C::NL, src, dst);
}

// ---------------------------------------------------------------- rows C2R
template <int N>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows_c2r(const float2* __restrict__ in, float* __restrict__ out,
                                                              RowMap m, float scale) {
  // m.in_* here describe the cropped OUTPUT extents.
  using C = RowCfg<N>;
  __shared__ float2 lds[C::NL * C::LS];
  __shared__ float2 tw[N];
  make_twiddles<N>(tw);
  __syncthreads();
  const int64_t row0 = (int64_t)blockIdx.x * C::NL;
  constexpr int nxc = N / 2 + 1;
  auto src = [&](int l, int k) -> float2 {
    const int64_t row = row0 + l;
    float2 v = make_float2(0.f, 0.f);
    if (row < m.n_rows) {
      v = in[row * nxc + (k < nxc ? k : N - k)];
      if (k >= nxc) v.y = -v.y;
      if (k == 0 || k == N / 2) v.y = 0.0f;  // C2R ignores these imaginary parts
    }
    return v;
  };
  auto dst = [&](int l, int x, float2 v) {
    const int64_t row = row0 + l;
    if (row >= m.n_rows || x >= m.in_nx) return;
    const int64_t b = row / m.rows_per_batch;
    const int64_t rr = row - b * m.rows_per_batch;
    const int y = (int)(rr % m.ny), z = (int)(rr / m.ny);
    if (y < m.in_ny && z < m.in_nz) out[((b * m.in_nz + z) * m.in_ny + y) * (int64_t)m.in_nx + x] = v.x * scale;
  };
  fft_lines<N, true>(lds, tw, C::NL, src, dst);
}

// ------------------------------------------------------ strided columns C2C
// Lines along a strided axis: for outer index o in [0, n_outer) the line base
// is (o / q) * s1 + (o % q) * s2, element stride es; columns c in [0, ncols)
// are adjacent (stride 1) and one workgroup takes NL of them.  In-place.
struct ColMap {
  int64_t n_outer, q, s1, s2, es;
  int ncols;
};

template <int N>
struct ColCfg {
  static constexpr int T = N >= 512 ? 1024 : FFT_THREADS;  // keep >= 16 adjacent columns per workgroup
  static constexpr int NL = FftLines<N, T>::value > 64 ? 64 : FftLines<N, T>::value;
  static constexpr int LS = FftStride<N>::value;
};

// Column tiles of NL adjacent columns of one outer row.  For a half spectrum
// whose last column (the Nyquist column) would be a tile of its own, that
// column instead goes to the first ceil(n_outer / COL_NQ) workgroups, COL_NQ
// rows each (lines past COL_NQ empty), and the rest hold NL full columns.
constexpr int COL_NQ = 2;

template <int N, bool INV, bool NYQ>
__device__ __forceinline__ void fft_cols_tile(float2* __restrict__ data, const ColMap& m, float2* lds,
                                              const float2* tw, int64_t b) {
  using C = ColCfg<N>;
  if (NYQ) {  // Nyquist tile: line c is the last column of row b * COL_NQ + c
    const int64_t o0 = b * COL_NQ;
    const int col = m.ncols - 1;
    auto base_of = [&](int c) -> int64_t {
      const int64_t o = o0 + c;
      return (o / m.q) * m.s1 + (o % m.q) * m.s2 + col;
    };
    auto used = [&](int c) { return c < COL_NQ && o0 + c < m.n_outer; };
    auto src = [&](int c, int k) -> float2 {
      return used(c) ? data[base_of(c) + k * m.es] : make_float2(0.f, 0.f);
    };
    auto dst = [&](int c, int k, float2 v) {
      if (used(c)) data[base_of(c) + k * m.es] = v;
    };
    fft_lines<N, INV, true>(lds, tw, C::NL, src, dst);
  } else {
    const int nblk = NYQ ? 0 : (m.ncols - 1) / C::NL;  // full tiles per row (Nyquist column elsewhere)
    const int64_t o = b / nblk;
    const int c0 = (int)(b - o * nblk) * C::NL;
    const int64_t base = (o / m.q) * m.s1 + (o % m.q) * m.s2 + c0;
    auto src = [&](int c, int k) -> float2 { return data[base + k * m.es + c]; };
    auto dst = [&](int c, int k, float2 v) { data[base + k * m.es + c] = v; };
    fft_lines<N, INV, true>(lds, tw, C::NL, src, dst);
  }
}

// nnyq > 0: Nyquist tiles first (nnyq of them), then full tiles; nnyq == 0:
// every row in ceil(ncols / NL) tiles, the last one partly empty
template <int N, bool INV>
__global__ __launch_bounds__(ColCfg<N>::T) void k_fft_cols(float2* __restrict__ data, ColMap m, int64_t nnyq) {
  using C = ColCfg<N>;
  __shared__ float2 lds[C::NL * C::LS];
  __shared__ float2 tw[N];
  make_twiddles<N>(tw);
  __syncthreads();
  if (nnyq > 0) {
    if ((int64_t)blockIdx.x < nnyq)
      fft_cols_tile<N, INV, true>(data, m, lds, tw, blockIdx.x);
    else
      fft_cols_tile<N, INV, false>(data, m, lds, tw, (int64_t)blockIdx.x - nnyq);
    return;
  }
  const int nblk = (m.ncols + C::NL - 1) / C::NL;
  const int64_t o = blockIdx.x / nblk;
  const int c0 = (int)(blockIdx.x - o * nblk) * C::NL;
  const int64_t base = (o / m.q) * m.s1 + (o % m.q) * m.s2 + c0;
  auto src = [&](int c, int k) -> float2 {
    return (c0 + c < m.ncols) ? data[base + k * m.es + c] : make_float2(0.f, 0.f);
  };
  auto dst = [&](int c, int k, float2 v) {
    if (c0 + c < m.ncols) data[base + k * m.es + c] = v;
  };
  fft_lines<N, INV, true>(lds, tw, C::NL, src, dst);
}

template <bool INV>
int launch_cols(int n, float2* data, const ColMap& m, hipStream_t s) {
  TimedScope ts_(INV ? "fft_cols_inv" : "fft_cols_fwd", s);
  switch (n) {
#define CASE(NN)                                                                                  \
  case NN: {                                                                                      \
    constexpr int NL = ColCfg<NN>::NL;                                                            \
    const bool nyq = m.ncols > NL && (m.ncols - 1) % NL == 0;                                     \
    const int64_t nnyq = nyq ? (m.n_outer + COL_NQ - 1) / COL_NQ : 0;                             \
    const int64_t blocks = nyq ? nnyq + m.n_outer * ((m.ncols - 1) / NL)                          \
                               : m.n_outer * ((m.ncols + NL - 1) / NL);                           \
    AUGUR_REQUIRE(blocks < ((int64_t)1 << 31), AUGUR_EINVAL, "fft: too many lines");              \
    hipLaunchKernelGGL((k_fft_cols<NN, INV>), dim3((unsigned)blocks), dim3(ColCfg<NN>::T), 0, s, data, m, nnyq); \
    break;                                                                                        \
  }
    CASE(4) CASE(8) CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512)
#undef CASE
    default: set_error("fft: unsupported length %d", n); return AUGUR_EUNSUPPORTED;
  }
  AUGUR_LAUNCHED("fft_cols");
  return AUGUR_OK;
}

int launch_rows_r2c(int n, const float* in, float2* out, const RowMap& m, hipStream_t s) {
  TimedScope ts_("fft_rows_r2c", s);
  switch (n) {
#define CASE(NN)                                                                                         \
  case NN:                                                                                               \
    hipLaunchKernelGGL((k_fft_rows_r2c<NN>), dim3((unsigned)((m.n_rows + RowCfg<NN>::NL - 1) / RowCfg<NN>::NL)), \
                       dim3(FFT_THREADS), 0, s, in, out, m);                                             \
    break;
    CASE(4) CASE(8) CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512)
#undef CASE
    default: set_error("fft: unsupported length %d", n); return AUGUR_EUNSUPPORTED;
  }
  AUGUR_LAUNCHED("fft_rows_r2c");
  return AUGUR_OK;
}

int launch_rows_c2r(int n, const float2* in, float* out, const RowMap& m, float scale, hipStream_t s) {
  TimedScope ts_("fft_rows_c2r", s);
  switch (n) {
#define CASE(NN)                                                                                         \
  case NN:                                                                                               \
    hipLaunchKernelGGL((k_fft_rows_c2r<NN>), dim3((unsigned)((m.n_rows + RowCfg<NN>::NL - 1) / RowCfg<NN>::NL)), \
                       dim3(FFT_THREADS), 0, s, in, out, m, scale);                                      \
    break;
    CASE(4) CASE(8) CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512)
#undef CASE
    default: set_error("fft: unsupported length %d", n); return AUGUR_EUNSUPPORTED;
  }
  AUGUR_LAUNCHED("fft_rows_c2r");
  return AUGUR_OK;
}

// x-row R2C and y-column forward transforms of one real G^3 grid (the z
// transform is left to the fused Poisson pass of the PM stepper).  The half
// spectrum's rows are `pitch` complex apart (>= G/2 + 1).
int fft_r2c_xy(const float* in, float2* out, int64_t G, int64_t pitch, hipStream_t s) {
  RowMap m;
  m.ny = (int)G;
  m.nz = (int)G;
  m.rows_per_batch = G * G;
  m.n_rows = G * G;
  m.in_nx = m.in_ny = m.in_nz = (int)G;
  m.pitch = pitch;
  int rc = launch_rows_r2c((int)G, in, out, m, s);
  if (rc) return rc;
  const int nxc = (int)(G / 2 + 1);
  ColMap my{G, 1, G * pitch, 0, pitch, nxc};
  return launch_cols<false>((int)G, out, my, s);
}

// y-column forward transforms of a (G, G, pitch) half spectrum (G/2 + 1 columns)
int fft_cols_y(float2* data, int64_t G, int64_t pitch, hipStream_t s, int64_t nb) {
  // nb grids of G planes each, consecutive: every plane is one outer index
  ColMap my{nb * G, 1, G * pitch, 0, pitch, (int)(G / 2 + 1)};
  return launch_cols<false>((int)G, data, my, s);
}

static int check_shape(int ndim, const int64_t* shape) {
  AUGUR_REQUIRE(ndim == 2 || ndim == 3, AUGUR_EINVAL, "fft: ndim must be 2 or 3");
  AUGUR_REQUIRE(shape != nullptr, AUGUR_EINVAL, "fft: null shape");
  for (int d = 0; d < ndim; ++d)
    AUGUR_REQUIRE(is_pow2(shape[d]) && shape[d] >= 4 && shape[d] <= 512, AUGUR_EUNSUPPORTED,
                  "fft: extents must be powers of two in [4, 512] (got %lld)", (long long)shape[d]);
  return AUGUR_OK;
}

// Forward/inverse C2C over the strided axes of a (batch, [nz,] ny, nxc) spectrum.
template <bool INV>
int spectrum_cols(float2* spec, int ndim, const int64_t* shape, int64_t batch, hipStream_t s) {
  const int nxc = (int)(shape[ndim - 1] / 2 + 1);
  const int64_t ny = shape[ndim - 2];
  const int64_t nz = ndim == 3 ? shape[0] : 1;
  // y axis: outer = (b, z), base = o * ny * nxc, element stride nxc
  ColMap my{batch * nz, 1, ny * nxc, 0, nxc, nxc};
  // z axis: outer = (b, y), base = b * nz*ny*nxc + y * nxc, element stride ny*nxc
  ColMap mz{batch * ny, ny, nz * ny * nxc, nxc, ny * nxc, nxc};
  int rc;
  if (!INV) {
    rc = launch_cols<INV>((int)ny, spec, my, s);
    if (rc) return rc;
    if (ndim == 3) rc = launch_cols<INV>((int)nz, spec, mz, s);
  } else {
    if (ndim == 3) {
      rc = launch_cols<INV>((int)nz, spec, mz, s);
      if (rc) return rc;
    }
    rc = launch_cols<INV>((int)ny, spec, my, s);
  }
  return rc;
}

// ----------------------------------------------------------- Poisson solve
__global__ void k_poisson(const float2* __restrict__ rho, int64_t B, int G, float kf, float green, float scale,
                          float2* __restrict__ acc, float2* __restrict__ phi) {
  const int nxc = G / 2 + 1;
  const int64_t per = (int64_t)G * G * nxc;
  const int64_t tot = B * per;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / per;
    const int64_t e = i - b * per;
    const int kx = (int)(e % nxc);
    const int ky = (int)((e / nxc) % G);
    const int kz = (int)(e / ((int64_t)nxc * G));
    const int ny = ky <= G / 2 ? ky : ky - G;
    const int nz = kz <= G / 2 ? kz : kz - G;
    const float fx = kf * kx, fy = kf * ny, fz = kf * nz;
    const float k2 = fx * fx + fy * fy + fz * fz;
    float2 p = make_float2(0.f, 0.f);
    if (e != 0) {
      const float g = green * scale / k2;  // -4 pi G / k^2 times the output scale
      const float2 r = rho[i];
      p = make_float2(r.x * g, r.y * g);
    }
    if (phi) phi[i] = p;
    // a_d = -i k_d phi = (k_d * phi.im, -k_d * phi.re); Nyquist of axis d zeroed
    const float ax = (kx == G / 2) ? 0.f : fx;
    const float ay = (ky == G / 2) ? 0.f : fy;
    const float az = (kz == G / 2) ? 0.f : fz;
    float2* out = acc + b * 3 * per + e;
    out[0] = make_float2(ax * p.y, -ax * p.x);
    out[per] = make_float2(ay * p.y, -ay * p.x);
    out[2 * per] = make_float2(az * p.y, -az * p.x);
  }
}

__global__ void k_complex_mul(const float2* __restrict__ a, const float2* __restrict__ b, float2* __restrict__ out,
                              int64_t n, int64_t batch, float scale) {
  const int64_t tot = n * batch;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    float2 p = cmul(a[i], b[i % n]);
    out[i] = make_float2(p.x * scale, p.y * scale);
  }
}

__global__ void k_open_kernel_2d(int G, float h, float eps2, float* out) {
  const int n = 2 * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n * n; i += gridDim.x * blockDim.x) {
    const int iy = i / n, ix = i % n;
    // reference: d = min(idx, 2G - idx) * h in the tensor dtype (FP32)
    const float dy = __fmul_rn((float)min(iy, n - iy), h), dx = __fmul_rn((float)min(ix, n - ix), h);
    const float r2 = __fadd_rn(__fadd_rn(__fmul_rn(dy, dy), __fmul_rn(dx, dx)), eps2);
    out[i] = __fdiv_rn(-1.0f, __fsqrt_rn(r2));
  }
}

__global__ void k_neg_grad_fd(const float* __restrict__ phi, int64_t B, int G, float twoh, float* __restrict__ out) {
  const int64_t per = (int64_t)G * G;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < B * per; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / per;
    const int e = (int)(i - b * per);
    const int y = e / G, x = e % G;
    const float* p = phi + b * per;
    const float xp = x + 1 < G ? p[e + 1] : 0.f, xm = x > 0 ? p[e - 1] : 0.f;
    const float yp = y + 1 < G ? p[e + G] : 0.f, ym = y > 0 ? p[e - G] : 0.f;
    // reference: -(p[i+1] - p[i-1]) / (2h), IEEE divide
    out[b * 2 * per + e] = -__fdiv_rn(__fsub_rn(xp, xm), twoh);
    out[b * 2 * per + per + e] = -__fdiv_rn(__fsub_rn(yp, ym), twoh);
  }
}

}  // namespace augur

extern "C" {

size_t augur_fft_workspace_bytes(int ndim, const int64_t* shape, int64_t batch) {
  (void)ndim;
  (void)shape;
  (void)batch;
  return 0;  // the LDS kernels need no global scratch
}

int augur_fft_r2c(const float* in, const int64_t* in_shape, float* out, int ndim, const int64_t* shape,
                  int64_t batch, void* ws, size_t ws_bytes, void* stream) {
  (void)ws;
  (void)ws_bytes;
  int rc = check_shape(ndim, shape);
  if (rc) return rc;
  AUGUR_REQUIRE(batch >= 1 && in && out, AUGUR_EINVAL, "fft_r2c: bad arguments");
  const int64_t* is = in_shape ? in_shape : shape;
  for (int d = 0; d < ndim; ++d)
    AUGUR_REQUIRE(is[d] >= 1 && is[d] <= shape[d], AUGUR_EINVAL, "fft_r2c: in_shape must be within shape");
  hipStream_t s = as_stream(stream);
  RowMap m;
  m.ny = (int)shape[ndim - 2];
  m.nz = ndim == 3 ? (int)shape[0] : 1;
  m.rows_per_batch = (int64_t)m.ny * m.nz;
  m.n_rows = batch * m.rows_per_batch;
  m.in_nx = (int)is[ndim - 1];
  m.in_ny = (int)is[ndim - 2];
  m.in_nz = ndim == 3 ? (int)is[0] : 1;
  rc = launch_rows_r2c((int)shape[ndim - 1], in, (float2*)out, m, s);
  if (rc) return rc;
  return spectrum_cols<false>((float2*)out, ndim, shape, batch, s);
}

int augur_fft_c2r(float* in, float* out, const int64_t* out_shape, int ndim, const int64_t* shape, int64_t batch,
                  float scale, void* ws, size_t ws_bytes, void* stream) {
  (void)ws;
  (void)ws_bytes;
  int rc = check_shape(ndim, shape);
  if (rc) return rc;
  AUGUR_REQUIRE(batch >= 1 && in && out, AUGUR_EINVAL, "fft_c2r: bad arguments");
  const int64_t* os = out_shape ? out_shape : shape;
  for (int d = 0; d < ndim; ++d)
    AUGUR_REQUIRE(os[d] >= 1 && os[d] <= shape[d], AUGUR_EINVAL, "fft_c2r: out_shape must be within shape");
  hipStream_t s = as_stream(stream);
  rc = spectrum_cols<true>((float2*)in, ndim, shape, batch, s);
  if (rc) return rc;
  RowMap m;
  m.ny = (int)shape[ndim - 2];
  m.nz = ndim == 3 ? (int)shape[0] : 1;
  m.rows_per_batch = (int64_t)m.ny * m.nz;
  m.n_rows = batch * m.rows_per_batch;
  m.in_nx = (int)os[ndim - 1];
  m.in_ny = (int)os[ndim - 2];
  m.in_nz = ndim == 3 ? (int)os[0] : 1;
  return launch_rows_c2r((int)shape[ndim - 1], (const float2*)in, out, m, scale, s);
}

int augur_complex_mul(const float* a, const float* b, float* out, int64_t n, int64_t batch, float scale, void* stream) {
  AUGUR_REQUIRE(n >= 0 && batch >= 1, AUGUR_EINVAL, "complex_mul: bad sizes");
  if (n == 0) return AUGUR_OK;
  AUGUR_REQUIRE(a && b && out, AUGUR_EINVAL, "complex_mul: null pointer");
  hipLaunchKernelGGL(k_complex_mul, dim3(blocks_for(n * batch, 256)), dim3(256), 0, as_stream(stream),
                     (const float2*)a, (const float2*)b, (float2*)out, n, batch, scale);
  AUGUR_LAUNCHED("complex_mul");
  return AUGUR_OK;
}

int augur_poisson_green_grad(const float* rho_hat, int64_t B, int64_t G, float L, float Gconst, float scale,
                             float* acc_hat, float* phi_hat, void* stream) {
  AUGUR_REQUIRE(B >= 1 && is_pow2(G) && G >= 4 && G <= 512 && L > 0.0f, AUGUR_EINVAL,
                "poisson: need B >= 1, power-of-two G in [4, 512], L > 0");
  AUGUR_REQUIRE(rho_hat && acc_hat, AUGUR_EINVAL, "poisson: null pointer");
  const float kf = (float)(2.0 * 3.14159265358979323846 / (double)L);
  const float green = (float)(-4.0 * 3.14159265358979323846 * (double)Gconst);
  const int64_t tot = B * G * G * (G / 2 + 1);
  TimedScope ts_("poisson", as_stream(stream));
  hipLaunchKernelGGL(k_poisson, dim3(blocks_for(tot, 256)), dim3(256), 0, as_stream(stream), (const float2*)rho_hat, B,
                     (int)G, kf, green, scale, (float2*)acc_hat, (float2*)phi_hat);
  AUGUR_LAUNCHED("poisson");
  return AUGUR_OK;
}

int augur_open_kernel_2d(int64_t G, float h, float eps, float* out, void* stream) {
  AUGUR_REQUIRE(G >= 1 && G <= 4096 && out, AUGUR_EINVAL, "open_kernel_2d: bad arguments");
    // eps^2 as the reference forms it: a Python float squared, then cast to FP32
  const float eps2 = (float)((double)eps * (double)eps);
  hipLaunchKernelGGL(k_open_kernel_2d, dim3(blocks_for(4 * G * G, 256)), dim3(256), 0, as_stream(stream), (int)G, h,
                     eps2, out);
  AUGUR_LAUNCHED("open_kernel_2d");
  return AUGUR_OK;
}

int augur_neg_grad_fd(const float* phi, int64_t B, int64_t G, float h, float* out, void* stream) {
  AUGUR_REQUIRE(B >= 1 && G >= 1 && h > 0.0f && phi && out, AUGUR_EINVAL, "neg_grad_fd: bad arguments");
  hipLaunchKernelGGL(k_neg_grad_fd, dim3(blocks_for(B * G * G, 256)), dim3(256), 0, as_stream(stream), phi, B, (int)G,
                     (float)(2.0 * (double)h), out);
  AUGUR_LAUNCHED("neg_grad_fd");
  return AUGUR_OK;
}

}  // extern "C"
