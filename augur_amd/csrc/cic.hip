// Cell keys, stable sort-by-cell, deterministic CIC deposit, CIC gather and
// FP64 reductions for the PM path (gfx950).
//
// Deposit determinism: every CIC contribution c (computed in FP32 with the
// reference's op order) is converted to an exact fixed-point integer
// rint(c * 2^e), with one power-of-two exponent e per channel chosen from an
// order-independent bound (n * max|value| * 2^e <= 2^52), and accumulated with
// 64-bit integer adds.  Integer addition is associative, so the grid is
// bitwise the same for any particle order, thread count or launch shape.
#include "internal.hpp"

namespace augur {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

}  // namespace augur

extern "C" const char* augur_last_error(void) { return augur::g_err; }
extern "C" int augur_abi_version(void) { return 1; }

using namespace augur;
using namespace augur_dev;

namespace augur {

// ------------------------------------------------------------------ fills
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

int fill_u64(void* p, int64_t n, unsigned long long v, hipStream_t s) {
  TimedScope ts_("fill_u64", s);
  if (n <= 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, (unsigned long long*)p, n, v);
  AUGUR_LAUNCHED("fill_u64");
  return AUGUR_OK;
}
int device_cu_count() {
  // per-device cache; a device index beyond the table is queried each call
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (dev >= 0 && dev < 64) cached[dev] = cus;
  return cus;
}

int fill_u32(void* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, s, (uint32_t*)p, n, v);
  AUGUR_LAUNCHED("fill_u32");
  return AUGUR_OK;
}

// ------------------------------------------------------------------- scan
// Exclusive scan of uint32 values, 4096 elements per block (256 threads x 16).
constexpr int SC_THREADS = 256, SC_ITEMS = 16, SC_TILE = SC_THREADS * SC_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan_u32(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[SC_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int i = 0; i < SC_THREADS / 64; ++i) {
    uint32_t s = wsum[i];
    if (i < w) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_block(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* sums,
                                                           const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  const int64_t base = (int64_t)blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
  uint32_t v[SC_ITEMS];
  uint32_t local = 0;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0u;
    local += v[j];
  }
  uint32_t tot;
  uint32_t pre = block_exclusive_scan_u32(local, &tot);
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    if (base + j < n) out[base + j] = pre;
    pre += v[j];
  }
  if (threadIdx.x == 0 && sums) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_add(uint32_t* out, int64_t n, const uint32_t* offs,
                                                         const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  const int64_t base = (int64_t)blockIdx.x * SC_TILE;
  const uint32_t o = offs[blockIdx.x];
  for (int j = threadIdx.x; j < SC_TILE; j += SC_THREADS)
    if (base + j < n) out[base + j] += o;
}

// For a short list of block sums: workgroup b adds sum(sums[0 .. b)) to its
// tile itself (a block reduction of at most SC_FUSE values), which saves the
// separate scan of the block sums.
constexpr int64_t SC_FUSE = 1024;
__global__ __launch_bounds__(SC_THREADS) void k_scan_add_prefix(uint32_t* out, int64_t n, const uint32_t* sums,
                                                                const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t part[SC_THREADS / 64];
  const int b = blockIdx.x;
  uint32_t v = 0;
  for (int j = threadIdx.x; j < b; j += SC_THREADS) v += sums[j];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t o = 0;
  for (int w = 0; w < SC_THREADS / 64; ++w) o += part[w];
  const int64_t base = (int64_t)b * SC_TILE;
  for (int j = threadIdx.x; j < SC_TILE; j += SC_THREADS)
    if (base + j < n) out[base + j] += o;
}

size_t scan_workspace_bytes(int64_t n) {
  size_t b = 0;
  while (n > SC_TILE) {
    n = (n + SC_TILE - 1) / SC_TILE;
    b += align_up(n * sizeof(uint32_t)) * 2;
  }
  return b + 256;
}

// out may alias in.
// one workgroup for a short array, staged in LDS: coalesced load, each
// thread scans a contiguous chunk, chunk totals scanned across the block,
// coalesced store (in place allowed)
constexpr int64_t SC_SMALL = 16384;  // one launch of k_scan_small up to here
__device__ __forceinline__ void scan_small_body(const uint32_t* in, uint32_t* out, int64_t n) {
  __shared__ uint32_t buf[SC_SMALL];
  for (int64_t i = threadIdx.x; i < n; i += SC_THREADS) buf[i] = in[i];
  __syncthreads();
  const int64_t chunk = (n + SC_THREADS - 1) / SC_THREADS;
  const int64_t lo = (int64_t)threadIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  uint32_t local = 0;
  for (int64_t i = lo; i < hi; ++i) local += buf[i];
  uint32_t tot;
  uint32_t pre = block_exclusive_scan_u32(local, &tot);
  for (int64_t i = lo; i < hi; ++i) {
    const uint32_t v = buf[i];
    buf[i] = pre;
    pre += v;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += SC_THREADS) out[i] = buf[i];
}
__global__ __launch_bounds__(SC_THREADS) void k_scan_small(const uint32_t* in, uint32_t* out, int64_t n,
                                                           const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  scan_small_body(in, out, n);
}
// two short arrays in one launch, one workgroup each
__global__ __launch_bounds__(SC_THREADS) void k_scan_small2(const uint32_t* in0, uint32_t* out0, int64_t n0,
                                                            const uint32_t* in1, uint32_t* out1, int64_t n1,
                                                            const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  if (blockIdx.x == 0)
    scan_small_body(in0, out0, n0);
  else
    scan_small_body(in1, out1, n1);
}

// two exclusive scans; one launch when both arrays are short
int exclusive_scan2_u32(const uint32_t* in0, uint32_t* out0, int64_t n0, const uint32_t* in1, uint32_t* out1,
                        int64_t n1, void* ws, hipStream_t s, const uint32_t* gate) {
  if (n0 > 0 && n1 > 0 && n0 <= SC_SMALL && n1 <= SC_SMALL) {
    hipLaunchKernelGGL(k_scan_small2, dim3(2), dim3(SC_THREADS), 0, s, in0, out0, n0, in1, out1, n1, gate);
    AUGUR_LAUNCHED("scan_small2");
    return AUGUR_OK;
  }
  int rc = exclusive_scan_u32(in0, out0, n0, ws, s, gate);
  return rc ? rc : exclusive_scan_u32(in1, out1, n1, ws, s, gate);
}

// First level only, when the block sums are short enough for the consumer to
// add their prefix itself (at most SC_FUSE of them): *bsums / *nbs name them,
// and out[i] + (exclusive prefix of bsums)[i / SC_TILE] is the scan.  Else
// (*bsums = nullptr) the scan is complete.
static int exclusive_scan_u32_split(const uint32_t* in, uint32_t* out, int64_t n, void* ws, hipStream_t s,
                                    const uint32_t* gate, const uint32_t** bsums, int* nbs) {
  *bsums = nullptr;
  *nbs = 0;
  const int64_t nb = (n + SC_TILE - 1) / SC_TILE;
  if (n <= SC_SMALL || nb > SC_FUSE) return exclusive_scan_u32(in, out, n, ws, s, gate);
  uint32_t* sums = (uint32_t*)ws;
  hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(SC_THREADS), 0, s, in, out, n, sums, gate);
  AUGUR_LAUNCHED("scan_block");
  *bsums = sums;
  *nbs = (int)nb;
  return AUGUR_OK;
}

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* ws, hipStream_t s, const uint32_t* gate) {
  if (n <= 0) return AUGUR_OK;
  int64_t nb = (n + SC_TILE - 1) / SC_TILE;
  if (nb > 1 && n <= SC_SMALL) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SC_THREADS), 0, s, in, out, n, gate);
    AUGUR_LAUNCHED("scan_small");
    return AUGUR_OK;
  }
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(SC_THREADS), 0, s, in, out, n, (uint32_t*)nullptr, gate);
    AUGUR_LAUNCHED("scan_block");
    return AUGUR_OK;
  }
  uint32_t* sums = (uint32_t*)ws;
  uint32_t* offs = (uint32_t*)((char*)ws + align_up(nb * sizeof(uint32_t)));
  void* rest = (char*)offs + align_up(nb * sizeof(uint32_t));
  hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(SC_THREADS), 0, s, in, out, n, sums, gate);
  AUGUR_LAUNCHED("scan_block");
  if (nb <= SC_FUSE) {
    hipLaunchKernelGGL(k_scan_add_prefix, dim3((unsigned)nb), dim3(SC_THREADS), 0, s, out, n, (const uint32_t*)sums,
                       gate);
    AUGUR_LAUNCHED("scan_add_prefix");
    return AUGUR_OK;
  }
  int rc = exclusive_scan_u32(sums, offs, nb, rest, s, gate);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SC_THREADS), 0, s, out, n, offs, gate);
  AUGUR_LAUNCHED("scan_add");
  return AUGUR_OK;
}

// -------------------------------------------------------------- radix sort
// Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits.
// Per pass: block digit histograms -> exclusive scan in digit-major order ->
// stable scatter.  Stability inside a block: each wave takes a contiguous run
// of 1024 elements in 64-element batches; lanes with equal digits are found
// with 8 ballots and ranked by lane order, and per-wave digit counters are
// combined in wave order.
constexpr int RS_THREADS = 256, RS_TILE = 4096, RS_WAVE_RUN = 1024;

// Tiles of RS_TILE keys; a workgroup takes tiles blockIdx.x, + gridDim.x, ...
// The launch caps its grid: a gated launch that does nothing costs less with
// fewer workgroups (2^24 keys, stepper gate closed: 40.6 -> 37.9 us for the
// whole gated sort at 2048, 36.6 at 1024), and an open sort more (at 1024 the
// scatter runs at half its occupancy; 2048 keeps it).
constexpr int64_t RS_MAX_GRID = 2048;
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t* keys, int64_t n, int shift, uint32_t* hist,
                                                        int nblocks, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t h[256];
  for (int tile = blockIdx.x; tile < nblocks; tile += gridDim.x) {
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)tile * RS_TILE;
    for (int i = threadIdx.x; i < RS_TILE; i += RS_THREADS) {
      int64_t idx = base + i;
      if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + tile] = h[threadIdx.x];
  }
}

// offs: the scanned histograms, or their first scan level with the block sums
// bsums (nbs <= SC_FUSE of them) whose prefix each workgroup adds itself
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout,
                                                           uint32_t* vout, int64_t n, int shift, const uint32_t* offs,
                                                           const uint32_t* bsums, int nbs, int nblocks,
                                                           const uint32_t* gate) {
  static_assert(RS_THREADS == SC_THREADS && SC_FUSE == 4 * SC_THREADS, "block-sum prefix layout");
  if (gate && *gate == 0u) return;
  __shared__ uint32_t cnt[RS_THREADS / 64][256];
  __shared__ uint32_t bpre[SC_FUSE];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (bsums) {  // exclusive prefix of the block sums: four per thread
    uint32_t v[4], loc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = 4 * threadIdx.x + j;
      v[j] = i < nbs ? bsums[i] : 0u;
      loc += v[j];
    }
    uint32_t tot;
    uint32_t pre = block_exclusive_scan_u32(loc, &tot);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bpre[4 * threadIdx.x + j] = pre;
      pre += v[j];
    }
  }
  for (int tile = blockIdx.x; tile < nblocks; tile += gridDim.x) {
    for (int i = threadIdx.x; i < (RS_THREADS / 64) * 256; i += RS_THREADS) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)tile * RS_TILE + (int64_t)w * RS_WAVE_RUN;
    constexpr int NB = RS_WAVE_RUN / 64;
    uint32_t key[NB], val[NB], rk[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int64_t idx = base + j * 64 + lane;
      const bool valid = idx < n;
      const uint32_t k = valid ? kin[idx] : 0u;
      const uint32_t d = (k >> shift) & 255u;
      uint64_t m = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
      }
      const int leader = __ffsll((unsigned long long)m) - 1;
      const uint32_t below = __popcll(m & __lanemask_lt());
      uint32_t old = 0;
      if (valid && lane == leader) {
        old = cnt[w][d];
        cnt[w][d] = old + (uint32_t)__popcll(m);
      }
      old = __shfl(old, leader < 0 ? 0 : leader, 64);
      key[j] = k;
      val[j] = valid ? vin[idx] : 0u;
      rk[j] = old + below;
    }
    __syncthreads();
    {
      const int d = threadIdx.x;  // 256 threads, one digit each
      const int64_t at = (int64_t)d * nblocks + tile;
      uint32_t s = offs[at] + (bsums ? bpre[at / SC_TILE] : 0u);
      for (int q = 0; q < RS_THREADS / 64; ++q) {
        uint32_t c = cnt[q][d];
        cnt[q][d] = s;
        s += c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int64_t idx = base + j * 64 + lane;
      if (idx < n) {
        const uint32_t dest = cnt[w][(key[j] >> shift) & 255u] + rk[j];
        kout[dest] = key[j];
        vout[dest] = val[j];
      }
    }
    __syncthreads();  // cnt is cleared for the next tile
  }
}

__global__ void k_iota_u32(uint32_t* v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (uint32_t)i;
}

size_t radix_workspace_bytes(int64_t n) {
  const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
  return 4 * align_up(n * sizeof(uint32_t)) + align_up(nb * 256 * sizeof(uint32_t)) +
         scan_workspace_bytes(nb * 256) + 1024;
}

// Sorts (keys, vals) by the low `bits` bits of keys, stably.  On return
// *keys_out / *vals_out point at the sorted arrays (inside ws or the inputs).
// keys/vals are clobbered.  If vals == nullptr they are initialised to iota.
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int bits, void* ws, size_t ws_bytes,
                     uint32_t** keys_out, uint32_t** vals_out, hipStream_t s, const uint32_t* gate) {
  const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
  Carver c(ws, ws_bytes);
  uint32_t* k2 = c.take<uint32_t>(n);
  uint32_t* v2 = c.take<uint32_t>(n);
  uint32_t* hist = c.take<uint32_t>(nb * 256);
  void* sws = c.take<char>(scan_workspace_bytes(nb * 256));
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "radix sort: workspace too small (%zu < %zu)", ws_bytes, c.used);
  if (nb > (int64_t)1 << 24) {
    set_error("radix sort: n too large");
    return AUGUR_EINVAL;
  }
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
  for (int shift = 0; shift < bits; shift += 8) {
    const unsigned rgrid = (unsigned)(nb < RS_MAX_GRID ? nb : RS_MAX_GRID);
    hipLaunchKernelGGL(k_rs_hist, dim3(rgrid), dim3(RS_THREADS), 0, s, ka, n, shift, hist, (int)nb, gate);
    AUGUR_LAUNCHED("rs_hist");
    const uint32_t* bsums;
    int nbs;
    int rc = exclusive_scan_u32_split(hist, hist, nb * 256, sws, s, gate, &bsums, &nbs);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rs_scatter, dim3(rgrid), dim3(RS_THREADS), 0, s, ka, va, kb, vb, n, shift, hist,
                       bsums, nbs, (int)nb, gate);
    AUGUR_LAUNCHED("rs_scatter");
    uint32_t* t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
  }
  *keys_out = ka;
  *vals_out = va;
  return AUGUR_OK;
}

// ------------------------------------------------------------------- keys
__global__ void k_cell_keys(const float* __restrict__ pos, const float* __restrict__ mask, int64_t B, int64_t N,
                            int D, float off, float cell, int64_t C, int64_t* __restrict__ keys,
                            unsigned long long* __restrict__ counts) {
  const int64_t tot = B * N;
  int64_t CD = 1;
  for (int d = 0; d < D; ++d) CD *= C;
  const float cmax = (float)(C - 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / N;
    int64_t lin = b * CD, mul = 1;
    for (int d = 0; d < D; ++d) {
      float t = __fdiv_rn(__fadd_rn(pos[i * D + d], off), cell);
      t = floorf(t);
      t = fminf(fmaxf(t, 0.0f), cmax);
      lin += (int64_t)t * mul;
      mul *= C;
    }
    const bool ok = mask == nullptr || mask[i] > 0.0f;
    const int64_t key = ok ? lin : B * CD;
    keys[i] = key;
    if (counts) atomicAdd(&counts[key], 1ull);
  }
}

__global__ void k_cic_base(const float* __restrict__ pos, int64_t n, int D, float off, float h, int32_t* __restrict__ i0,
                           float* __restrict__ w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * D; i += (int64_t)gridDim.x * blockDim.x) {
    const float u = cic_u(pos[i], off, h);
    const float f = floorf(u);
    i0[i] = (int32_t)f;
    w[i] = __fsub_rn(u, f);
  }
}

// --------------------------------------------------------------- deposit
__device__ __forceinline__ float chan_value(int c, float mm, const float* vel, int64_t i, int D, float scale) {
  // reference: stack([mm, mm*v0, mm*v1]) * (in_scale / h^D)  (scatter_field.py:241)
  const float base = (c == 0) ? mm : __fmul_rn(mm, vel[i * D + (c - 1)]);
  return __fmul_rn(base, scale);
}

__global__ void k_dep_absmax(const float* __restrict__ vel, const float* __restrict__ mass, float mass_value,
                             const float* __restrict__ mask, int64_t tot, int D, int C, float scale,
                             unsigned* __restrict__ maxbits) {
  float mx[4] = {0.f, 0.f, 0.f, 0.f};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    float mm = mass ? mass[i] : mass_value;
    if (mask) mm = __fmul_rn(mm, mask[i]);
    for (int c = 0; c < C; ++c) mx[c] = fmaxf(mx[c], fabsf(chan_value(c, mm, vel, i, D, scale)));
  }
  for (int c = 0; c < C; ++c) {
    float v = mx[c];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&maxbits[c], __float_as_uint(v));
  }
}

// scene of particle i in a ragged batch: the last b with starts[b] <= i
__device__ __forceinline__ int64_t scene_of(const int64_t* __restrict__ starts, int64_t B, int64_t i) {
  int64_t lo = 0, hi = B;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (starts[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// starts (may be NULL): ragged batch, scene b holds particles [starts[b], starts[b+1]);
// otherwise B scenes of N particles each.  The loop runs in whole waves (a
// lane past the end is inactive) so that wave_merged_add sees every lane.
template <int D>
__global__ __launch_bounds__(256) void k_dep_fixed(const float* __restrict__ pos, const float* __restrict__ vel,
                                                   const float* __restrict__ mass, float mass_value,
                                                   const float* __restrict__ mask, int64_t B, int64_t N,
                                                   const int64_t* __restrict__ starts, int64_t tot, int G,
                                                   float off, float h, int periodic, float scale, int C,
                                                   const unsigned* __restrict__ maxbits,
                                                   unsigned long long* __restrict__ acc) {
  int E[4];
  for (int c = 0; c < C; ++c) E[c] = fixed_exponent(maxbits[c], tot);
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < tot; base += stride) {
    const int64_t ii = base + threadIdx.x;
    const bool live = ii < tot;
    const int64_t i = live ? ii : tot - 1;  // a dead lane computes on a valid row and adds nothing
    const int64_t b = starts ? scene_of(starts, B, i) : i / N;
    float mm = mass ? mass[i] : mass_value;
    if (mask) mm = __fmul_rn(mm, mask[i]);
    float valE[4];
    for (int c = 0; c < C; ++c) valE[c] = ldexpf(chan_value(c, mm, vel, i, D, scale), E[c]);
    int i0[D];
    float w[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float u = cic_u(pos[i * D + d], off, h);
      const float f = floorf(u);
      i0[d] = (int)f;
      w[d] = __fsub_rn(u, f);
    }
#pragma unroll
    for (int o = 0; o < (1 << D); ++o) {
      float wt = 1.0f;
      bool ok = live;
      int64_t flat = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int od = (o >> d) & 1;
        const float wd = od ? w[d] : __fsub_rn(1.0f, w[d]);
        wt = (d == 0) ? wd : __fmul_rn(wt, wd);
        int ix = i0[d] + od;
        if (periodic) {
          ix = wrap(ix, G);
        } else {
          ok = ok && ix >= 0 && ix < G;
        }
        flat += (int64_t)ix * mul;
        mul *= G;
      }
      for (int c = 0; c < C; ++c) {
        // rint(val * wt * 2^e) with the exact power-of-two scaling applied to val
        const long long q = fixed_round(__fmul_rn(valE[c], wt), 0);
        wave_merged_add(acc, ((b * C) + c) * GD + (ok ? flat : 0), q, ok);
      }
    }
  }
}

__global__ void k_dep_convert(const unsigned long long* __restrict__ acc, float* __restrict__ out, int64_t B, int C,
                              int64_t GD, int64_t tot_particles, const unsigned* __restrict__ maxbits) {
  const int64_t n = B * C * GD;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)((i / GD) % C);
    out[i] = (float)ldexp((double)(long long)acc[i], -fixed_exponent(maxbits[c], tot_particles));
  }
}

// ---------------------------------------------------------------- gather
template <int D>
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ field, const float* __restrict__ pos,
                                                int64_t B, int64_t N, int G, int C, float off, float h, int periodic,
                                                float* __restrict__ out) {
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  const int64_t tot = B * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / N;
    int i0[D];
    float w[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float u = cic_u(pos[i * D + d], off, h);
      const float f = floorf(u);
      i0[d] = (int)f;
      w[d] = __fsub_rn(u, f);
    }
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // corner order of the reference's _corner_iter: x offset outermost
#pragma unroll
    for (int ox = 0; ox < 2; ++ox)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int oz = 0; oz < (D == 3 ? 2 : 1); ++oz) {
          const int od[3] = {ox, oy, oz};
          float wt = 1.0f;
          bool ok = true;
          int64_t flat = 0, mul = 1;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const float wd = od[d] ? w[d] : __fsub_rn(1.0f, w[d]);
            wt = (d == 0) ? wd : __fmul_rn(wt, wd);
            int ix = i0[d] + od[d];
            if (periodic) {
              ix = wrap(ix, G);
            } else {
              ok = ok && ix >= 0 && ix < G;
              ix = min(max(ix, 0), G - 1);
            }
            flat += (int64_t)ix * mul;
            mul *= G;
          }
          if (!ok) wt = 0.0f;
          for (int c = 0; c < C && c < 8; ++c) acc[c] += field[(b * C + c) * GD + flat] * wt;
        }
    for (int c = 0; c < C && c < 8; ++c) out[i * C + c] = acc[c];
  }
}

// ------------------------------------------------- TSC ("quad") stencil
// ScatterField._quad_axis (scripts/scatter_field.py:184-188): i0 = round(u)
// (half to even), d = u - i0, weights 0.5 (0.5 - d)^2, 0.75 - d^2,
// 0.5 (0.5 + d)^2 at offsets -1, 0, +1; the corner weight is the product
// over axes in axis order; off-grid corners weigh 0 (or wrap when periodic).
__device__ __forceinline__ void tsc_axis(float u, int& i0, float w[3]) {
  const float r = rintf(u);
  i0 = (int)r;
  const float d = __fsub_rn(u, r);
  const float a = __fsub_rn(0.5f, d), b = __fadd_rn(0.5f, d);
  w[0] = __fmul_rn(0.5f, __fmul_rn(a, a));
  w[1] = __fsub_rn(0.75f, __fmul_rn(d, d));
  w[2] = __fmul_rn(0.5f, __fmul_rn(b, b));
}

template <int D>
__global__ __launch_bounds__(256) void k_dep_tsc(const float* __restrict__ pos, const float* __restrict__ vel,
                                                 const float* __restrict__ mass, float mass_value,
                                                 const float* __restrict__ mask, int64_t B, int64_t N, int64_t tot,
                                                 int G, float off, float h, int periodic, float scale, int C,
                                                 const unsigned* __restrict__ maxbits,
                                                 unsigned long long* __restrict__ acc) {
  int E[4];
  for (int c = 0; c < C; ++c) E[c] = fixed_exponent(maxbits[c], tot);
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  constexpr int NC = D == 2 ? 9 : 27;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < tot; base += stride) {
    const int64_t ii = base + threadIdx.x;
    const bool live = ii < tot;
    const int64_t i = live ? ii : tot - 1;
    const int64_t b = i / N;
    float mm = mass ? mass[i] : mass_value;
    if (mask) mm = __fmul_rn(mm, mask[i]);
    float valE[4];
    for (int c = 0; c < C; ++c) valE[c] = ldexpf(chan_value(c, mm, vel, i, D, scale), E[c]);
    int i0[D];
    float w[D][3];
#pragma unroll
    for (int d = 0; d < D; ++d) tsc_axis(cic_u(pos[i * D + d], off, h), i0[d], w[d]);
#pragma unroll
    for (int o = 0; o < NC; ++o) {
      float wt = 1.0f;
      bool ok = live;
      int64_t flat = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int od = (d == 0 ? o : (d == 1 ? o / 3 : o / 9)) % 3;
        wt = (d == 0) ? w[d][od] : __fmul_rn(wt, w[d][od]);
        int ix = i0[d] + od - 1;
        if (periodic) {
          ix = wrap(ix, G);
        } else {
          ok = ok && ix >= 0 && ix < G;
        }
        flat += (int64_t)ix * mul;
        mul *= G;
      }
      for (int c = 0; c < C; ++c) {
        const long long q = fixed_round(__fmul_rn(valE[c], wt), 0);
        wave_merged_add(acc, ((b * C) + c) * GD + (ok ? flat : 0), q, ok);
      }
    }
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_gather_tsc(const float* __restrict__ field, const float* __restrict__ pos,
                                                    int64_t B, int64_t N, int G, int C, float off, float h,
                                                    int periodic, float* __restrict__ out) {
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  constexpr int NC = D == 2 ? 9 : 27;
  const int64_t tot = B * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / N;
    int i0[D];
    float w[D][3];
#pragma unroll
    for (int d = 0; d < D; ++d) tsc_axis(cic_u(pos[i * D + d], off, h), i0[d], w[d]);
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // the reference's corner order: x offset outermost (scripts/scatter_field.py:193-201)
#pragma unroll
    for (int o = 0; o < NC; ++o) {
      int od[3];
      od[0] = o / (NC / 3);
      od[1] = (o / (D == 3 ? 3 : 1)) % 3;
      od[2] = o % 3;
      float wt = 1.0f;
      bool ok = true;
      int64_t flat = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        wt = (d == 0) ? w[d][od[d]] : __fmul_rn(wt, w[d][od[d]]);
        int ix = i0[d] + od[d] - 1;
        if (periodic) {
          ix = wrap(ix, G);
        } else {
          ok = ok && ix >= 0 && ix < G;
          ix = min(max(ix, 0), G - 1);
        }
        flat += (int64_t)ix * mul;
        mul *= G;
      }
      if (!ok) wt = 0.0f;
      for (int c = 0; c < C && c < 8; ++c) acc[c] += field[(b * C + c) * GD + flat] * wt;
    }
    for (int c = 0; c < C && c < 8; ++c) out[i * C + c] = acc[c];
  }
}

// ------------------------------------------------------- FP64 reductions
constexpr int RED_BLOCKS = 1024, RED_THREADS = 256;

__device__ __forceinline__ double block_sum_f64(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;  // valid in thread 0
}

// Each block reduces a fixed contiguous slice in a fixed order, so the
// result depends only on n (not on timing).
__global__ __launch_bounds__(RED_THREADS) void k_red_partial(const float* __restrict__ vel,
                                                             const float* __restrict__ mass, float mass_value,
                                                             const float* __restrict__ phi, int64_t n,
                                                             double* __restrict__ part) {
  __shared__ double sh[RED_THREADS / 64];
  const int64_t per = (n + RED_BLOCKS - 1) / RED_BLOCKS;
  const int64_t lo = blockIdx.x * per, hi = min(n, lo + per);
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = lo + threadIdx.x; i < hi; i += RED_THREADS) {
    const double m = mass ? (double)mass[i] : (double)mass_value;
    const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
    a[0] += m * vx;
    a[1] += m * vy;
    a[2] += m * vz;
    a[3] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    if (phi) a[4] += 0.5 * m * (double)phi[i];
    a[5] += m;
  }
  for (int k = 0; k < 6; ++k) {
    double t = block_sum_f64(a[k], sh);
    if (threadIdx.x == 0) part[blockIdx.x * 6 + k] = t;
  }
}

__global__ __launch_bounds__(RED_THREADS) void k_red_final(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[RED_THREADS / 64];
  for (int k = 0; k < 6; ++k) {
    double a = 0.0;
    for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS) a += part[b * 6 + k];
    double t = block_sum_f64(a, sh);
    if (threadIdx.x == 0) out[k] = t;
  }
}

}  // namespace augur

// =================================================================== C ABI
extern "C" {

int augur_cell_keys(const float* pos, const float* mask, int64_t B, int64_t N, int D, float off, float cell,
                    int64_t C, int64_t* keys, int64_t* counts, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "cell_keys: D must be 2 or 3 (got %d)", D);
  AUGUR_REQUIRE(B >= 0 && N >= 0 && C >= 1 && cell > 0.0f, AUGUR_EINVAL, "cell_keys: bad sizes");
  AUGUR_REQUIRE(B * N == 0 || (pos && keys), AUGUR_EINVAL, "cell_keys: null pointer");
  hipStream_t s = as_stream(stream);
  int64_t CD = 1;
  for (int d = 0; d < D; ++d) CD *= C;
  if (counts) {
    int rc = fill_u64(counts, B * CD + 1, 0ull, s);
    if (rc) return rc;
  }
  if (B * N == 0) return AUGUR_OK;
  hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, pos, mask, B, N, D, off, cell, C,
                     keys, (unsigned long long*)counts);
  AUGUR_LAUNCHED("cell_keys");
  return AUGUR_OK;
}

int augur_cic_base(const float* pos, int64_t n, int D, float off, float h, int32_t* i0, float* w, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "cic_base: D must be 2 or 3");
  AUGUR_REQUIRE(n >= 0 && h > 0.0f, AUGUR_EINVAL, "cic_base: bad sizes");
  if (n == 0) return AUGUR_OK;
  AUGUR_REQUIRE(pos && i0 && w, AUGUR_EINVAL, "cic_base: null pointer");
  hipLaunchKernelGGL(k_cic_base, dim3(blocks_for(n * D, 256)), dim3(256), 0, as_stream(stream), pos, n, D, off, h,
                     i0, w);
  AUGUR_LAUNCHED("cic_base");
  return AUGUR_OK;
}

}  // extern "C"

namespace augur {

// Keys outside [0, nkeys) are clamped into it (documented precondition).
__global__ void k_keys_i64_to_u32(const int64_t* in, uint32_t* out, int64_t n, int64_t nkeys) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = in[i];
    out[i] = (uint32_t)(k < 0 ? 0 : (k >= nkeys ? nkeys - 1 : k));
  }
}
__global__ void k_u32_to_i64(const uint32_t* in, int64_t* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}
__global__ void k_count_keys(const uint32_t* keys, int64_t n, uint32_t* counts) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&counts[keys[i]], 1u);
}

}  // namespace augur

extern "C" {

size_t augur_sort_workspace_bytes(int64_t n, int64_t nkeys) {
  if (n < 0 || nkeys < 0) return 0;
  return 2 * align_up(n * sizeof(uint32_t)) + radix_workspace_bytes(n) + align_up((nkeys + 1) * sizeof(uint32_t)) +
         scan_workspace_bytes(nkeys + 1) + 1024;
}

int augur_sort_by_cell(const int64_t* keys, int64_t n, int64_t nkeys, int64_t* perm, int64_t* starts, void* ws,
                       size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(n >= 0 && nkeys >= 1 && nkeys < ((int64_t)1 << 31), AUGUR_EINVAL,
                "sort_by_cell: need n >= 0 and 1 <= nkeys < 2^31");
  AUGUR_REQUIRE(n < ((int64_t)1 << 32), AUGUR_EINVAL, "sort_by_cell: n must be < 2^32");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  uint32_t* k = c.take<uint32_t>(n);
  uint32_t* v = c.take<uint32_t>(n);
  void* rws = c.take<char>(radix_workspace_bytes(n));
  uint32_t* cnt = c.take<uint32_t>(nkeys + 1);
  void* sws = c.take<char>(scan_workspace_bytes(nkeys + 1));
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "sort_by_cell: workspace too small (%zu < %zu)", ws_bytes, c.used);
  if (n > 0) {
    AUGUR_REQUIRE(keys && perm, AUGUR_EINVAL, "sort_by_cell: null pointer");
    int rc = AUGUR_OK;
    hipLaunchKernelGGL(k_keys_i64_to_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, k, n, nkeys);
    AUGUR_LAUNCHED("keys_to_u32");
    hipLaunchKernelGGL(k_iota_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, v, n);
    AUGUR_LAUNCHED("iota");
    const int bits = ilog2(nkeys);
    uint32_t *ks = k, *vs = v;
    if (bits > 0) {
      rc = radix_sort_pairs(k, v, n, bits, rws, radix_workspace_bytes(n), &ks, &vs, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_u32_to_i64, dim3(blocks_for(n, 256)), dim3(256), 0, s, vs, perm, n);
    AUGUR_LAUNCHED("perm_to_i64");
    if (starts) {
      rc = fill_u32(cnt, nkeys + 1, 0u, s);
      if (rc) return rc;
      hipLaunchKernelGGL(k_count_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, ks, n, cnt);
      AUGUR_LAUNCHED("count_keys");
      rc = exclusive_scan_u32(cnt, cnt, nkeys + 1, sws, s);
      if (rc) return rc;
      hipLaunchKernelGGL(k_u32_to_i64, dim3(blocks_for(nkeys + 1, 256)), dim3(256), 0, s, cnt, starts, nkeys + 1);
      AUGUR_LAUNCHED("starts_to_i64");
    }
  } else if (starts) {
    int rc = fill_u64(starts, nkeys + 1, 0ull, s);
    if (rc) return rc;
  }
  return AUGUR_OK;
}

size_t augur_cic_deposit_workspace_bytes(int64_t B, int D, int64_t G, int C) {
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  return align_up(B * C * GD * sizeof(unsigned long long)) + 256;
}

int augur_cic_deposit(const float* pos, const float* vel, const float* mass, float mass_value, const float* mask,
                      int64_t B, int64_t N, int D, int64_t G, float off, float h, int periodic, float scale, float* out,
                      void* ws, size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "cic_deposit: D must be 2 or 3");
  AUGUR_REQUIRE(B >= 1 && N >= 0 && G >= 2 && h > 0.0f, AUGUR_EINVAL, "cic_deposit: bad sizes");
  AUGUR_REQUIRE(!periodic || is_pow2(G), AUGUR_EUNSUPPORTED, "cic_deposit: periodic needs power-of-two G");
  AUGUR_REQUIRE(G <= 4096, AUGUR_EUNSUPPORTED, "cic_deposit: G too large");
  const int C = vel ? 1 + D : 1;
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  unsigned long long* acc = c.take<unsigned long long>(B * C * GD);
  unsigned* maxbits = c.take<unsigned>(4);
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "cic_deposit: workspace too small (%zu < %zu)", ws_bytes, c.used);
  AUGUR_REQUIRE(out && (N == 0 || pos), AUGUR_EINVAL, "cic_deposit: null pointer");
  int rc = fill_u64(acc, B * C * GD, 0ull, s);
  if (rc) return rc;
  rc = fill_u32(maxbits, 4, 0u, s);
  if (rc) return rc;
  const int64_t tot = B * N;
  if (tot > 0) {
    {
      TimedScope ts_("dep_absmax", s);
      hipLaunchKernelGGL(k_dep_absmax, dim3(blocks_for(tot, 256, 2048)), dim3(256), 0, s, vel, mass, mass_value,
                         mask, tot, D, C, scale, maxbits);
      AUGUR_LAUNCHED("dep_absmax");
    }
    TimedScope ts_("dep_fixed", s);
    if (D == 2)
      hipLaunchKernelGGL(k_dep_fixed<2>, dim3(blocks_for(tot, 256)), dim3(256), 0, s, pos, vel, mass, mass_value,
                         mask, B, N, (const int64_t*)nullptr, tot, (int)G, off, h, periodic, scale, C, maxbits, acc);
    else
      hipLaunchKernelGGL(k_dep_fixed<3>, dim3(blocks_for(tot, 256)), dim3(256), 0, s, pos, vel, mass, mass_value,
                         mask, B, N, (const int64_t*)nullptr, tot, (int)G, off, h, periodic, scale, C, maxbits, acc);
    AUGUR_LAUNCHED("dep_fixed");
  }
  TimedScope ts_("dep_convert", s);
  hipLaunchKernelGGL(k_dep_convert, dim3(blocks_for(B * C * GD, 256)), dim3(256), 0, s, acc, out, B, C, GD,
                     tot > 0 ? tot : 1, maxbits);
  AUGUR_LAUNCHED("dep_convert");
  return AUGUR_OK;
}

int augur_cic_deposit_ragged(const float* pos, const float* vel, const float* mass, float mass_value,
                             const int64_t* starts, int64_t B, int64_t total, int D, int64_t G, float off, float h,
                             int periodic, float scale, float* out, void* ws, size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "cic_deposit_ragged: D must be 2 or 3");
  AUGUR_REQUIRE(B >= 1 && total >= 0 && G >= 2 && h > 0.0f && starts, AUGUR_EINVAL, "cic_deposit_ragged: bad sizes");
  AUGUR_REQUIRE(!periodic || is_pow2(G), AUGUR_EUNSUPPORTED, "cic_deposit_ragged: periodic needs power-of-two G");
  AUGUR_REQUIRE(G <= 4096, AUGUR_EUNSUPPORTED, "cic_deposit_ragged: G too large");
  const int C = vel ? 1 + D : 1;
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  unsigned long long* acc = c.take<unsigned long long>(B * C * GD);
  unsigned* maxbits = c.take<unsigned>(4);
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "cic_deposit_ragged: workspace too small (%zu < %zu)", ws_bytes, c.used);
  AUGUR_REQUIRE(out && (total == 0 || pos), AUGUR_EINVAL, "cic_deposit_ragged: null pointer");
  int rc = fill_u64(acc, B * C * GD, 0ull, s);
  if (rc) return rc;
  rc = fill_u32(maxbits, 4, 0u, s);
  if (rc) return rc;
  if (total > 0) {
    {
      TimedScope ts_("dep_absmax", s);
      hipLaunchKernelGGL(k_dep_absmax, dim3(blocks_for(total, 256, 2048)), dim3(256), 0, s, vel, mass, mass_value,
                         (const float*)nullptr, total, D, C, scale, maxbits);
      AUGUR_LAUNCHED("dep_absmax");
    }
    TimedScope ts_("dep_fixed", s);
    if (D == 2)
      hipLaunchKernelGGL(k_dep_fixed<2>, dim3(blocks_for(total, 256)), dim3(256), 0, s, pos, vel, mass, mass_value,
                         (const float*)nullptr, B, (int64_t)0, starts, total, (int)G, off, h, periodic, scale, C,
                         maxbits, acc);
    else
      hipLaunchKernelGGL(k_dep_fixed<3>, dim3(blocks_for(total, 256)), dim3(256), 0, s, pos, vel, mass, mass_value,
                         (const float*)nullptr, B, (int64_t)0, starts, total, (int)G, off, h, periodic, scale, C,
                         maxbits, acc);
    AUGUR_LAUNCHED("dep_fixed");
  }
  TimedScope ts_("dep_convert", s);
  hipLaunchKernelGGL(k_dep_convert, dim3(blocks_for(B * C * GD, 256)), dim3(256), 0, s, acc, out, B, C, GD,
                     total > 0 ? total : 1, maxbits);
  AUGUR_LAUNCHED("dep_convert");
  return AUGUR_OK;
}

int augur_cic_gather(const float* field, const float* pos, int64_t B, int64_t N, int D, int64_t G, int C, float off,
                     float h, int periodic, float* out, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "cic_gather: D must be 2 or 3");
  AUGUR_REQUIRE(B >= 1 && N >= 0 && G >= 2 && C >= 1 && C <= 8 && h > 0.0f, AUGUR_EINVAL,
                "cic_gather: bad sizes (C must be 1..8)");
  AUGUR_REQUIRE(!periodic || is_pow2(G), AUGUR_EUNSUPPORTED, "cic_gather: periodic needs power-of-two G");
  if (B * N == 0) return AUGUR_OK;
  AUGUR_REQUIRE(field && pos && out, AUGUR_EINVAL, "cic_gather: null pointer");
  hipStream_t s = as_stream(stream);
  if (D == 2)
    hipLaunchKernelGGL(k_gather<2>, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, field, pos, B, N, (int)G, C, off,
                       h, periodic, out);
  else
    hipLaunchKernelGGL(k_gather<3>, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, field, pos, B, N, (int)G, C, off,
                       h, periodic, out);
  AUGUR_LAUNCHED("gather");
  return AUGUR_OK;
}

int augur_tsc_deposit(const float* pos, const float* vel, const float* mass, float mass_value, const float* mask,
                      int64_t B, int64_t N, int D, int64_t G, float off, float h, int periodic, float scale, float* out,
                      void* ws, size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "tsc_deposit: D must be 2 or 3");
  AUGUR_REQUIRE(B >= 1 && N >= 0 && G >= 2 && h > 0.0f, AUGUR_EINVAL, "tsc_deposit: bad sizes");
  AUGUR_REQUIRE(!periodic || is_pow2(G), AUGUR_EUNSUPPORTED, "tsc_deposit: periodic needs power-of-two G");
  AUGUR_REQUIRE(G <= 4096, AUGUR_EUNSUPPORTED, "tsc_deposit: G too large");
  const int C = vel ? 1 + D : 1;
  int64_t GD = 1;
  for (int d = 0; d < D; ++d) GD *= G;
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  unsigned long long* acc = c.take<unsigned long long>(B * C * GD);
  unsigned* maxbits = c.take<unsigned>(4);
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "tsc_deposit: workspace too small (%zu < %zu)", ws_bytes, c.used);
  AUGUR_REQUIRE(out && (N == 0 || pos), AUGUR_EINVAL, "tsc_deposit: null pointer");
  int rc = fill_u64(acc, B * C * GD, 0ull, s);
  if (rc) return rc;
  rc = fill_u32(maxbits, 4, 0u, s);
  if (rc) return rc;
  const int64_t tot = B * N;
  if (tot > 0) {
    hipLaunchKernelGGL(k_dep_absmax, dim3(blocks_for(tot, 256, 2048)), dim3(256), 0, s, vel, mass, mass_value, mask,
                       tot, D, C, scale, maxbits);
    AUGUR_LAUNCHED("dep_absmax");
    TimedScope ts_("dep_tsc", s);
    if (D == 2)
      hipLaunchKernelGGL(k_dep_tsc<2>, dim3(blocks_for(tot, 256)), dim3(256), 0, s, pos, vel, mass, mass_value, mask,
                         B, N, tot, (int)G, off, h, periodic, scale, C, maxbits, acc);
    else
      hipLaunchKernelGGL(k_dep_tsc<3>, dim3(blocks_for(tot, 256)), dim3(256), 0, s, pos, vel, mass, mass_value, mask,
                         B, N, tot, (int)G, off, h, periodic, scale, C, maxbits, acc);
    AUGUR_LAUNCHED("dep_tsc");
  }
  hipLaunchKernelGGL(k_dep_convert, dim3(blocks_for(B * C * GD, 256)), dim3(256), 0, s, acc, out, B, C, GD,
                     tot > 0 ? tot : 1, maxbits);
  AUGUR_LAUNCHED("dep_convert");
  return AUGUR_OK;
}

int augur_tsc_gather(const float* field, const float* pos, int64_t B, int64_t N, int D, int64_t G, int C, float off,
                     float h, int periodic, float* out, void* stream) {
  AUGUR_REQUIRE(D == 2 || D == 3, AUGUR_EINVAL, "tsc_gather: D must be 2 or 3");
  AUGUR_REQUIRE(B >= 1 && N >= 0 && G >= 2 && C >= 1 && C <= 8 && h > 0.0f, AUGUR_EINVAL,
                "tsc_gather: bad sizes (C must be 1..8)");
  AUGUR_REQUIRE(!periodic || is_pow2(G), AUGUR_EUNSUPPORTED, "tsc_gather: periodic needs power-of-two G");
  if (B * N == 0) return AUGUR_OK;
  AUGUR_REQUIRE(field && pos && out, AUGUR_EINVAL, "tsc_gather: null pointer");
  hipStream_t s = as_stream(stream);
  if (D == 2)
    hipLaunchKernelGGL(k_gather_tsc<2>, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, field, pos, B, N, (int)G, C,
                       off, h, periodic, out);
  else
    hipLaunchKernelGGL(k_gather_tsc<3>, dim3(blocks_for(B * N, 256)), dim3(256), 0, s, field, pos, B, N, (int)G, C,
                       off, h, periodic, out);
  AUGUR_LAUNCHED("tsc_gather");
  return AUGUR_OK;
}

size_t augur_reduce_workspace_bytes(int64_t n) {
  (void)n;
  return align_up(RED_BLOCKS * 6 * sizeof(double)) + 256;
}

int augur_reduce_energy_momentum_f64(const float* vel, const float* mass, float mass_value, const float* phi,
                                     int64_t n, double* out, void* ws, size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(n >= 0 && out, AUGUR_EINVAL, "reduce: bad arguments");
  AUGUR_REQUIRE(ws && ws_bytes >= RED_BLOCKS * 6 * sizeof(double), AUGUR_EWORKSPACE, "reduce: workspace too small");
  AUGUR_REQUIRE(n == 0 || vel, AUGUR_EINVAL, "reduce: null velocity");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_red_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, vel, mass, mass_value, phi, n,
                     (double*)ws);
  AUGUR_LAUNCHED("red_partial");
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(RED_THREADS), 0, s, (const double*)ws, out);
  AUGUR_LAUNCHED("red_final");
  return AUGUR_OK;
}

}  // extern "C"
