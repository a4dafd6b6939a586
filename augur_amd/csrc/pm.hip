// Periodic PM stepper on gfx950.
//
// State: particles physically sorted by CIC base cell (key = (z*G + y)*G + x
// of floor((pos/h) - 0.5) wrapped), with cell_start[] (G^3 + 1 offsets) from
// the last sort.  One step:
//   1. deposit   k_dep_tiles: per 32x16x16 tile, the particles whose base
//                cell (as of the last sort) lies in the tile's region grown
//                by the drift bound D, cut into work items planned at sort
//                time; fixed-point CIC unit counts summed exactly in FP64
//                (registers and LDS atomics), bitwise order-independent.
//   2. FFT       x-rows R2C -> y-columns -> z-columns fused with the Green's
//                function and the inverse z transform (two outputs) ->
//                inverse y with three outputs -> x-rows C2R of the three
//                acceleration components written as three FP32 planes.
//   3. gather    k_gather_tiles: CIC gather of the force planes fused with the
//                closing half kick, the next opening half kick and the drift,
//                and the max-speed reduction that bounds the next D.
// D (cells) is the cumulative max drift since the last sort; when it would
// exceed 1 the gated re-sort kernels run (they are always launched and exit
// at once when not needed, so the step has no host sync and is capturable).
#include <algorithm>

#include "fft_core.hpp"

using namespace augur;
using namespace augur_dev;

namespace augur {

// --------------------------------------------------------------- bookkeeping
struct PmDev {
  float cum_cells;      // cumulative max drift in cells since the last sort
  uint32_t vmax_bits;   // max |v| component seen by the last drift (float bits)
  uint32_t need_sort;   // gate of the re-sort kernels
  uint32_t maxbits;     // deposit bound: float bits of max |m * scale|
  uint32_t pad[4];
};

constexpr float DRIFT_MARGIN = 1e-3f;  // cells; covers position rounding
constexpr int DMAX = 1;                // largest drift (cells) between sorts

// The stepper's sort order is tile-major: 32x16x16-cell tiles in x-fastest
// order, and inside a tile its cells in x-fastest order.  A tile's particles
// (and each run of whole z layers of it) are then one contiguous range.
constexpr int TD_X = 32, TD_Y = 16, TD_Z = 16;
constexpr int TD_CELLS = TD_X * TD_Y * TD_Z;

__device__ __forceinline__ uint32_t tile_key(int cx, int cy, int cz, int G) {
  const uint32_t ntx = (uint32_t)G / TD_X, nty = (uint32_t)G / TD_Y;
  const uint32_t t = ((uint32_t)(cz / TD_Z) * nty + (uint32_t)(cy / TD_Y)) * ntx + (uint32_t)(cx / TD_X);
  const uint32_t l = ((uint32_t)(cz % TD_Z) * TD_Y + (uint32_t)(cy % TD_Y)) * TD_X + (uint32_t)(cx % TD_X);
  return t * TD_CELLS + l;
}

__device__ __forceinline__ int drift_cells(const PmDev* st) {
  const float c = st->cum_cells;
  return c <= 0.0f ? 0 : (int)ceilf(c + DRIFT_MARGIN);
}

// cum += dt*vmax/h; vmax = 0; need_sort = (D would exceed DMAX)
__global__ void k_pm_decide(PmDev* st, float dt_over_h, int force_sort) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float v = __uint_as_float(st->vmax_bits);
  st->cum_cells += dt_over_h * v;
  st->vmax_bits = 0u;
  st->need_sort = (force_sort || st->cum_cells + DRIFT_MARGIN > (float)DMAX) ? 1u : 0u;
}

// also, with uniform masses (uniform != 0), the deposit bound |mass_value *
// scale| that k_pm_set_maxbits would write (one launch less per step)
__global__ void k_pm_sorted(PmDev* st, int uniform, float mass_value, float scale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->need_sort) {
    st->cum_cells = 0.0f;
    st->need_sort = 0u;
  }
  if (uniform) st->maxbits = __float_as_uint(fabsf(__fmul_rn(mass_value, scale)));
}

__global__ void k_pm_set_maxbits(PmDev* st, const float* mass, float mass_value, float scale, int64_t n) {
  // uniform mass: the bound is |mass_value * scale| (same FP32 product as the deposit)
  if (mass == nullptr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) st->maxbits = __float_as_uint(fabsf(__fmul_rn(mass_value, scale)));
    return;
  }
  float mx = 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    mx = fmaxf(mx, fabsf(__fmul_rn(mass[i], scale)));
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&st->maxbits, __float_as_uint(mx));
}

// ------------------------------------------------------------- gated sort
__global__ void k_base_keys(const float* __restrict__ pos, int64_t n, int G, float h, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals, uint32_t* __restrict__ ngaps, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ngaps = 0u;  // for k_cell_starts, later in the same sort
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = wrap((int)floorf(cic_u(pos[3 * i + d], 0.0f, h)), G);
    keys[i] = tile_key(c[0], c[1], c[2], G);
    vals[i] = (uint32_t)i;
  }
}

template <class T, int K>
__global__ void k_permute(const T* __restrict__ src, T* __restrict__ dst, const uint32_t* __restrict__ perm, int64_t n,
                          const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = perm[i];
#pragma unroll
    for (int k = 0; k < K; ++k) dst[i * K + k] = src[j * K + k];
  }
}

template <class T>
__global__ void k_copy(const T* __restrict__ src, T* __restrict__ dst, int64_t n, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// The state arrays permuted in one pass into tmp (pos 3n, vel 3n, then ids n
// and mass n when present), and copied back in one pass.
__global__ void k_permute_state(const float* __restrict__ pos, const float* __restrict__ vel,
                                const int32_t* __restrict__ ids, const float* __restrict__ mass,
                                const uint32_t* __restrict__ perm, int64_t n, float* __restrict__ tmp,
                                const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = perm[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      tmp[3 * i + d] = pos[3 * j + d];
      tmp[3 * n + 3 * i + d] = vel[3 * j + d];
    }
    if (ids) reinterpret_cast<int32_t*>(tmp)[6 * n + i] = ids[j];
    if (mass) tmp[7 * n + i] = mass[j];
  }
}
__global__ void k_restore_state(const float* __restrict__ tmp, int64_t n, float* __restrict__ pos,
                                float* __restrict__ vel, int32_t* __restrict__ ids, float* __restrict__ mass,
                                const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pos[3 * i + d] = tmp[3 * i + d];
      vel[3 * i + d] = tmp[3 * n + 3 * i + d];
    }
    if (ids) ids[i] = reinterpret_cast<const int32_t*>(tmp)[6 * n + i];
    if (mass) mass[i] = tmp[7 * n + i];
  }
}

// cell_start[c] = the first sorted slot whose key is >= c, for c in [0, ncells],
// straight from the sorted keys: slot i (0..n) owns the cells (key[i-1], key[i]]
// (key[-1] = -1, key[n] = ncells).  Runs of more than CS_SPAN empty cells go
// to a list that k_cell_fill writes with whole workgroups.
constexpr uint32_t CS_SPAN = 64;
__global__ void k_cell_starts(const uint32_t* __restrict__ ks, int64_t n, uint32_t ncells,
                              uint32_t* __restrict__ cell_start, uint32_t* __restrict__ gaps,
                              uint32_t* __restrict__ ngaps, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t lo = i == 0 ? 0u : ks[i - 1] + 1u;
    const uint32_t hi = i == n ? ncells : ks[i];
    if (hi < lo) continue;  // a repeated key: no cell starts here
    if (hi - lo < CS_SPAN) {
      for (uint32_t c = lo; c <= hi; ++c) cell_start[c] = (uint32_t)i;
    } else {
      const uint32_t k = atomicAdd(ngaps, 1u);
      gaps[3 * k] = lo;
      gaps[3 * k + 1] = hi;
      gaps[3 * k + 2] = (uint32_t)i;
    }
  }
}
__global__ void k_cell_fill(const uint32_t* __restrict__ gaps, const uint32_t* __restrict__ ngaps,
                            uint32_t* __restrict__ cell_start, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  const uint32_t ng = *ngaps;
  for (uint32_t g = blockIdx.x; g < ng; g += gridDim.x) {
    const uint32_t lo = gaps[3 * g], hi = gaps[3 * g + 1], v = gaps[3 * g + 2];
    for (uint32_t c = lo + threadIdx.x; c <= hi; c += blockDim.x) cell_start[c] = v;
  }
}

// ---------------------------------------------------------- tile deposit
// A tile's deposit reads the particles whose base cell, as of the last sort,
// lies in the tile grown by TD_LO = 1 + DMAX cells below and TD_HI = DMAX
// above on each axis (the drift bound never exceeds DMAX between sorts).  In
// the tile-major order that region is at most TD_MAXSEG key runs over the 27
// neighbouring tiles (the tile itself is one run).  At sort time k_dep_plan
// stores each tile's runs and k_tile_items cuts its visit list into work
// items, so a dense clump is spread over many workgroups and a step only
// streams the stored tables.  An item accumulates the fixed-point unit
// counts of its CIC contributions (integers below 2^53, so FP64 sums are
// exact) in a 64 KB LDS tile with FP64 atomics; a tile with one item writes
// its cells directly, the items of a heavier tile add their LDS tiles into a
// zero-kept int64 tile image in HBM with device atomics, and k_dep_finish
// (the next launch) converts and clears the images.  Integer sums make the
// result bitwise independent of order and of the split.
constexpr int TD_LO = 1 + DMAX, TD_HI = DMAX;
constexpr int TD_THREADS = 512;  // 64 KB LDS tile: two workgroups per CU
constexpr int TD_NW = TD_THREADS / 64;
// runs: a piece of the region that is partial in x gives one run per (y, z)
// row, one full in x but partial in y one run per z layer, one full in both
// a single run (x pieces: 2 partial + 1 full; y and z: 2 partial + 1 full)
constexpr int TD_REGY = TD_Y + TD_LO + TD_HI, TD_REGZ = TD_Z + TD_LO + TD_HI;
constexpr int TD_MAXSEG = 2 * TD_REGY * TD_REGZ + 2 * TD_REGZ + 3;
constexpr uint32_t TD_CHUNK = 16384;  // most particle visits per work item
#ifndef AUGUR_TD_K
#define AUGUR_TD_K 8
#endif
constexpr int TD_K = AUGUR_TD_K;      // consecutive visits per lane in a wave batch
constexpr int TD_BLK = TD_K;          // visits per entry of the visit -> segment table: one
                                      // entry per lane batch, so no lane walks other lanes' segments
constexpr int TD_NBLK = TD_CHUNK / TD_BLK;  // table entries per item

struct TileOrigin {
  int x0, y0, z0;
};
__device__ __forceinline__ TileOrigin tile_origin(int tile, int G) {
  const int ntx = G / TD_X, nty = G / TD_Y;
  return TileOrigin{(tile % ntx) * TD_X, ((tile / ntx) % nty) * TD_Y, (tile / (ntx * nty)) * TD_Z};
}

// one axis of the grown region, cut at tile boundaries: tile offset (-1, 0,
// +1) and local cell range [a, b); a single piece when one tile spans the axis
struct Piece {
  int off, a, b;
};
__device__ __forceinline__ int axis_pieces(int ntile, int T, Piece* p) {
  if (ntile == 1) {
    p[0] = Piece{0, 0, T};
    return 1;
  }
  p[0] = Piece{-1, T - TD_LO, T};
  p[1] = Piece{0, 0, T};
  p[2] = Piece{1, 0, TD_HI};
  return 3;
}
__device__ __forceinline__ int box_runs(const Piece& px, const Piece& py, const Piece& pz) {
  const bool fx = px.b - px.a == TD_X, fy = py.b - py.a == TD_Y;
  if (fx && fy) return 1;
  if (fx) return pz.b - pz.a;
  return (py.b - py.a) * (pz.b - pz.a);
}

// Sort-time plan, one wave per tile: the grown region's non-empty key runs as
// sorted-slot starts and an inclusive prefix of their lengths (TD_MAXSEG
// entries each; the unused tail has length 0), and the tile's number of work items of at most
// TD_CHUNK visits; items[ntiles] = 0 so an exclusive scan leaves the total there.
// A batch of grids has scene-major tiles: tile = scene * tps + tile of the
// scene (tps tiles per scene, keys scene * G^3 + tile_key), and a region
// wraps within its own scene's grid.
__device__ __forceinline__ void dep_plan_body(int bid, const uint32_t* __restrict__ cell_start, int G, int ntiles,
                                              int tps, uint32_t* __restrict__ items, uint32_t* __restrict__ seg_lo,
                                              uint32_t* __restrict__ seg_pre) {
  const int tile = bid * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (tile > ntiles) return;
  if (tile == ntiles) {
    if (lane == 0) items[ntiles] = 0u;
    return;
  }
  const int ntx = G / TD_X, nty = G / TD_Y, ntz = G / TD_Z;
  const int scene = tile / tps, ts = tile - scene * tps;
  const int tx = ts % ntx, ty = (ts / ntx) % nty, tz = ts / (ntx * nty);
  Piece PX[3], PY[3], PZ[3];
  const int nx = axis_pieces(ntx, TD_X, PX), ny = axis_pieces(nty, TD_Y, PY), nz = axis_pieces(ntz, TD_Z, PZ);
  int pre[28];  // runs before box b (boxes in z, y, x piece order)
  pre[0] = 0;
  for (int bz = 0, bi = 0; bz < nz; ++bz)
    for (int by = 0; by < ny; ++by)
      for (int bx = 0; bx < nx; ++bx, ++bi) pre[bi + 1] = pre[bi] + box_runs(PX[bx], PY[by], PZ[bz]);
  const int nbox = nx * ny * nz, total_runs = pre[nbox];
  uint32_t* lo_t = seg_lo + (int64_t)tile * TD_MAXSEG;
  uint32_t* pre_t = seg_pre + (int64_t)tile * (TD_MAXSEG + 1);
  if (lane == 0) pre_t[0] = 0u;
  uint32_t carry = 0, cnt = 0;  // visits and non-empty runs stored so far
  for (int r0 = 0; r0 < TD_MAXSEG; r0 += 64) {
    const int r = r0 + lane;
    uint32_t lo = 0, hi = 0;
    if (r < total_runs) {
      int bi = 0;
      while (pre[bi + 1] <= r) ++bi;
      const int bx = bi % nx, by = (bi / nx) % ny, bz = bi / (nx * ny), k = r - pre[bi];
      const Piece &px = PX[bx], &py = PY[by], &pz = PZ[bz];
      const int nt = (((tz + pz.off + ntz) % ntz) * nty + (ty + py.off + nty) % nty) * ntx + (tx + px.off + ntx) % ntx;
      const uint32_t tb = (uint32_t)(scene * tps + nt) * TD_CELLS;
      uint32_t k0, len;
      const bool fx = px.b - px.a == TD_X, fy = py.b - py.a == TD_Y;
      if (fx && fy) {
        k0 = tb + (uint32_t)pz.a * TD_Y * TD_X;
        len = (uint32_t)(pz.b - pz.a) * TD_Y * TD_X;
      } else if (fx) {
        k0 = tb + ((uint32_t)(pz.a + k) * TD_Y + py.a) * TD_X;
        len = (uint32_t)(py.b - py.a) * TD_X;
      } else {
        const int nyr = py.b - py.a;
        k0 = tb + ((uint32_t)(pz.a + k / nyr) * TD_Y + py.a + k % nyr) * TD_X + px.a;
        len = (uint32_t)(px.b - px.a);
      }
      lo = cell_start[k0];
      hi = cell_start[k0 + len];
    }
    uint32_t v = hi - lo;
    // empty runs are left out of the tables (stored runs are compacted), so a
    // lane's walk over its visits never steps through an empty run
    const uint64_t nonempty = __ballot(v != 0u);
    const uint32_t at = cnt + (uint32_t)__popcll(nonempty & ((1ull << lane) - 1ull));
    const bool keep = v != 0u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(v, o, 64);
      if (lane >= o) v += y;
    }
    if (keep) {
      lo_t[at] = lo;
      pre_t[at + 1] = v + carry;
    }
    carry += __shfl(v, 63, 64);
    cnt += (uint32_t)__popcll(nonempty);
  }
  // the unused tail: runs of length 0 after the last visit
  for (uint32_t r = cnt + lane; r < (uint32_t)TD_MAXSEG; r += 64) {
    lo_t[r] = 0u;
    pre_t[r + 1] = carry;
  }
  if (lane == 0) items[tile] = carry == 0u ? 1u : (carry + TD_CHUNK - 1) / TD_CHUNK;
}

// Per work item (one workgroup each; items beyond the total exit): its tile
// and visit range [vlo, vhi) (the tile's visits cut into near-equal parts of
// at most CHUNK), and blk[j] = the segment holding visit vlo + TD_BLK j.
struct DepItem {
  uint32_t tile, vlo, vhi, pad;
};

template <int MAXSEG, uint32_t CHUNK>
__device__ __forceinline__ void tile_items_body(uint32_t item, const uint32_t* __restrict__ item_start, int ntiles,
                                                const uint32_t* __restrict__ tseg_pre, DepItem* __restrict__ info,
                                                uint16_t* __restrict__ blk) {
  constexpr int NBLK = CHUNK / TD_BLK;
  if (item >= item_start[ntiles]) return;
  int a = 0, b = ntiles;  // last tile with item_start[tile] <= item
  while (b - a > 1) {
    const int m = (a + b) >> 1;
    if (item_start[m] <= item) a = m; else b = m;
  }
  const uint32_t* pre = tseg_pre + (int64_t)a * (MAXSEG + 1);
  const uint32_t chunk = item - item_start[a], nch = item_start[a + 1] - item_start[a];
  const uint32_t total = pre[MAXSEG];
  const uint32_t per = ((total + nch - 1) / nch + TD_BLK - 1) / TD_BLK * TD_BLK;  // <= CHUNK
  const uint32_t vlo = min(total, chunk * per), vhi = min(total, vlo + per);
  if (threadIdx.x == 0) info[item] = DepItem{(uint32_t)a, vlo, vhi, 0u};
  for (uint32_t j = threadIdx.x; j < (vhi - vlo + TD_BLK - 1) / TD_BLK; j += blockDim.x) {
    const uint32_t v = vlo + j * TD_BLK;  // last segment with pre[seg] <= v (non-empty: v < total)
    int lo = 0, hi = MAXSEG;
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (pre[m] <= v) lo = m; else hi = m;
    }
    blk[(int64_t)item * NBLK + j] = (uint16_t)lo;
  }
}

// adds a lane's run sums (base cell at tile offsets in key) into the LDS tile.
// Branch-free: a corner outside the tile, and every corner when key < 0 (no
// run), is redirected to the lane's own trash cell past the tile (never
// read), so every lane issues the same 8 atomics.  Byte offsets: an axis
// term outside the tile is TD_OFF, which lifts the sum past every trash
// cell, and min(sum, trash) then picks the trash cell.
constexpr int TD_TRASH = 64;  // one trash cell per lane of a wave
constexpr int TD_OFF = 1 << 20;
__device__ __forceinline__ void td_flush(double* acc, const double* a, int key, int G, int lane) {
  const int ax = key & 1023, ay = (key >> 10) & 1023, az = (key >> 20) & 1023;
  const int x1 = (ax + 1) & (G - 1), y1 = (ay + 1) & (G - 1), z1 = (az + 1) & (G - 1);
  const int cx[2] = {ax < TD_X ? ax * 8 : TD_OFF, x1 < TD_X ? x1 * 8 : TD_OFF};
  const int cy[2] = {ay < TD_Y ? ay * (TD_X * 8) : TD_OFF, y1 < TD_Y ? y1 * (TD_X * 8) : TD_OFF};
  const bool ok = key >= 0;
  const int cz[2] = {ok && az < TD_Z ? az * (TD_X * TD_Y * 8) : TD_OFF,
                     ok && z1 < TD_Z ? z1 * (TD_X * TD_Y * 8) : TD_OFF};
  const int trash = (TD_CELLS + lane) * 8;
  char* base = reinterpret_cast<char*>(acc);
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    atomicAdd(reinterpret_cast<double*>(base + min(cz[oz] + cy[oy] + cx[ox], trash)), a[o]);
  }
}

// perm (may be NULL): particle index of sorted slot i; otherwise slot i is particle i.
// P2: h is a power of two and inv_h = 1/h (multiply instead of divide, same bits).
template <bool P2>
__global__ __launch_bounds__(TD_THREADS) void k_dep_tiles(
    const float* __restrict__ pos, const uint32_t* __restrict__ perm, const float* __restrict__ mass,
    float mass_value, int64_t n, int G, float h, float inv_h, float scale, const PmDev* __restrict__ st,
    const uint32_t* __restrict__ item_start, const DepItem* __restrict__ info, const uint16_t* __restrict__ iblk,
    const uint32_t* __restrict__ tseg_lo, const uint32_t* __restrict__ tseg_pre, int ntiles, int tps,
    unsigned long long* __restrict__ fix, float* __restrict__ rho) {
  // unit counts of the fixed-point contributions: integers below 2^53 in FP64,
  // so the LDS float atomics and register sums are exact in any order
  __shared__ double acc[TD_CELLS + TD_TRASH];
  __shared__ uint32_t seg_lo[TD_MAXSEG];
  __shared__ uint32_t seg_pre[TD_MAXSEG + 1];
  __shared__ uint16_t blk_seg[TD_NBLK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int E = fixed_exponent(st->maxbits, n);
  const float valE0 = ldexpf(__fmul_rn(mass_value, scale), E);
  const uint32_t total_items = item_start[ntiles];
  for (uint32_t item = blockIdx.x; item < total_items; item += gridDim.x) {
    const DepItem it = info[item];
    const int tile = (int)it.tile;
    const uint32_t nch = item_start[tile + 1] - item_start[tile];
    const int scene = tile / tps;  // scene-major tiles of a batch of grids
    const TileOrigin t = tile_origin(tile - scene * tps, G);
    float* const rho_s = rho + (int64_t)scene * G * G * G;
    const uint32_t* lo_t = tseg_lo + (int64_t)tile * TD_MAXSEG;
    const uint32_t* pre_t = tseg_pre + (int64_t)tile * (TD_MAXSEG + 1);
    const uint16_t* blk_t = iblk + (int64_t)item * TD_NBLK;
    const uint32_t plo = it.vlo, phi = it.vhi;
    const uint32_t pblk = (phi - plo + TD_BLK - 1) / TD_BLK;
    for (int i = tid; i < TD_CELLS; i += TD_THREADS) acc[i] = 0.0;
    for (int r = tid; r <= TD_MAXSEG; r += TD_THREADS) {
      if (r < TD_MAXSEG) seg_lo[r] = lo_t[r];
      seg_pre[r] = pre_t[r];
    }
    for (uint32_t j = tid; j < pblk; j += TD_THREADS) blk_seg[j] = blk_t[j];
    __syncthreads();
    {
      // wave batches of 64 * TD_K consecutive visits, TD_K per lane: loads are
      // coalesced, and a lane sums a cell's particles in registers and
      // touches LDS when its base cell changes
      for (uint32_t bb = plo + (uint32_t)wave * 64 * TD_K; bb < phi; bb += TD_NW * 64 * TD_K) {
        const uint32_t l0 = bb + (uint32_t)lane * TD_K;
        if (l0 >= phi) continue;
        const uint32_t l1 = min(phi, l0 + TD_K);
        int seg = blk_seg[(l0 - plo) / TD_BLK];
        while (seg_pre[seg + 1] <= l0) ++seg;
        uint32_t snext = seg_pre[seg + 1];
        int64_t sbase = (int64_t)seg_lo[seg] - seg_pre[seg];
        // slots first (LDS walk), then each kind of load for all visits
        // together, so the position loads are in flight at the same time
        // (a load that depends on another, the perm lookup, is its own pass).
        // A visit past the lane's range reads the lane's first slot (always
        // valid) and adds nothing: no branches around the loads.
        uint32_t sl[TD_K];
        bool vok[TD_K];
#pragma unroll
        for (int u = 0; u < TD_K; ++u) {
          const uint32_t ww = l0 + u;
          vok[u] = ww < l1;
          if (vok[u]) {
            while (ww >= snext) {
              ++seg;
              snext = seg_pre[seg + 1];
              sbase = (int64_t)seg_lo[seg] - seg_pre[seg];
            }
            sl[u] = (uint32_t)(sbase + ww);
          } else {
            sl[u] = sl[0];  // visit 0 is always in range (l0 < phi)
          }
        }
        if (perm) {
#pragma unroll
          for (int u = 0; u < TD_K; ++u) sl[u] = perm[sl[u]];
        }
        float px[TD_K][3], pm[TD_K];
#pragma unroll
        for (int u = 0; u < TD_K; ++u) {
          const float* pp = pos + 3 * (uint64_t)sl[u];
          px[u][0] = pp[0];
          px[u][1] = pp[1];
          px[u][2] = pp[2];
        }
        if (mass) {
#pragma unroll
          for (int u = 0; u < TD_K; ++u) pm[u] = mass[sl[u]];
        }
        int rkey = -1;
        double run[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) run[o] = 0;
        // straight-line per visit: a visit outside the grown tile (or past
        // the lane's range) keeps the current run and adds zero units, so
        // only the flush on a change of base cell is a branch
#pragma unroll
        for (int u = 0; u < TD_K; ++u) {
          int i0[3];
          float wv[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const float uu = P2 ? cic_u_pow2(px[u][d], inv_h) : cic_u(px[u][d], 0.0f, h);
            const float f = floorf(uu);
            i0[d] = (int)f;
            wv[d] = __fsub_rn(uu, f);
          }
          // base cell as offsets from the tile origin (wrapped into [0, G));
          // a corner is inside when its offset is below the tile size
          const int ax = (i0[0] - t.x0) & (G - 1), ay = (i0[1] - t.y0) & (G - 1), az = (i0[2] - t.z0) & (G - 1);
          const bool in = vok[u] & ((ax < TD_X) | (ax == G - 1)) & ((ay < TD_Y) | (ay == G - 1)) &
                          ((az < TD_Z) | (az == G - 1));
          const int key = in ? (az << 20) | (ay << 10) | ax : rkey;
          // units rint(m * scale * w * 2^E), the power-of-two scaling applied to m * scale
          const float valE = in ? (mass ? ldexpf(__fmul_rn(pm[u], scale), E) : valE0) : 0.0f;
          double q[8];
#pragma unroll
          for (int o = 0; o < 8; ++o) {
            const float wx = (o & 1) ? wv[0] : __fsub_rn(1.0f, wv[0]);
            const float wy = ((o >> 1) & 1) ? wv[1] : __fsub_rn(1.0f, wv[1]);
            const float wz = (o >> 2) ? wv[2] : __fsub_rn(1.0f, wv[2]);
            q[o] = (double)rintf(__fmul_rn(valE, __fmul_rn(__fmul_rn(wx, wy), wz)));
          }
          if (key != rkey) {
            if (rkey >= 0) td_flush(acc, run, rkey, G, lane);
#pragma unroll
            for (int o = 0; o < 8; ++o) run[o] = q[o];
            rkey = key;
          } else {
#pragma unroll
            for (int o = 0; o < 8; ++o) run[o] += q[o];
          }
        }
        if (rkey >= 0) td_flush(acc, run, rkey, G, lane);
      }
    }
    __syncthreads();
    if (nch == 1) {
      for (int i = tid; i < TD_CELLS; i += TD_THREADS) {
        const int lx = i % TD_X, ly = (i / TD_X) % TD_Y, lz = i / (TD_X * TD_Y);
        const int64_t cell = ((int64_t)(t.z0 + lz) * G + (t.y0 + ly)) * G + (t.x0 + lx);
        rho_s[cell] = (float)ldexp(acc[i], -E);
      }
    } else {  // heavy tile: add into its image; k_dep_finish converts it
      unsigned long long* img = fix + (int64_t)tile * TD_CELLS;
      for (int i = tid; i < TD_CELLS; i += TD_THREADS) {
        const double v = acc[i];
        if (v != 0.0) atomicAdd(&img[i], (unsigned long long)(long long)v);
      }
    }
    __syncthreads();  // LDS is reused by the next item
  }
}

// heavy tiles (more than one item): image -> density cells, image cleared
// again for the next launch (one workgroup per tile; others exit)
__global__ __launch_bounds__(256) void k_dep_finish(const uint32_t* __restrict__ item_start,
                                                    unsigned long long* __restrict__ fix, const PmDev* __restrict__ st,
                                                    int64_t n, int G, int tps, float* __restrict__ rho) {
  const int tile = blockIdx.x;
  if (item_start[tile + 1] - item_start[tile] <= 1u) return;
  const int E = fixed_exponent(st->maxbits, n);
  const int ntx = G / TD_X, nty = G / TD_Y;
  const int scene = tile / tps, ts = tile - scene * tps;
  rho += (int64_t)scene * G * G * G;
  const int x0 = (ts % ntx) * TD_X, y0 = ((ts / ntx) % nty) * TD_Y, z0 = (ts / (ntx * nty)) * TD_Z;
  // plain loads: the atomics of k_dep_tiles are complete and visible at this
  // launch; each thread reads its 32 cells before writing any
  unsigned long long* img = fix + (int64_t)tile * TD_CELLS;
  constexpr int PER = TD_CELLS / 256;
  unsigned long long v[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) v[k] = img[threadIdx.x + 256 * k];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int lx = i % TD_X, ly = (i / TD_X) % TD_Y, lz = i / (TD_X * TD_Y);
    rho[((int64_t)(z0 + lz) * G + (y0 + ly)) * G + (x0 + lx)] = (float)ldexp((double)(long long)v[k], -E);
    img[i] = 0ull;  // the image stays zero between launches
  }
}

// zero-fills the heavy-tile images (once per workspace set-up)
__global__ void k_zero_u64(unsigned long long* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

// ------------------------------------------------------ fused FFT passes
// Spectral layout after the x R2C pass: (z, y|ky, kx) with nxc = G/2 + 1
// columns per row and a row pitch RP rounded up to 16 complex (128 bytes), so
// a column tile's row segment is one aligned cache line.
__host__ __device__ constexpr int64_t pm_pitch(int64_t G) { return (G / 2 + 1 + 15) / 16 * 16; }
// Column tiles take NL lines (line-fastest thread mapping): the first N / NQ
// tiles hold the Nyquist column kx = N/2 of NQ <= NL adjacent rows (one cache
// line per point, so they start first and overlap the rest; lines NQ .. NL-1
// stay empty), and N * nfull more hold NL adjacent kx < N/2 columns of one
// row each.
template <int N>
struct PmCol {
  static constexpr int T = N >= 512 ? 1024 : FFT_THREADS;
  static constexpr int NL = FftLines<N, T>::value > 16 ? 16 : FftLines<N, T>::value;
  static constexpr int LS = FftStride<N>::value;
  static constexpr int nxc = N / 2 + 1;
  static constexpr int RP = (int)pm_pitch(N);
  static constexpr int nfull = N / 2 / NL;
  // Nyquist rows per tile, z pass and y pass.  Measured at N = 256 (column
  // pass times): 16 rows put all the y pass's one-line-per-point traffic on
  // 16 workgroups and ran 8 us slower than one tile per row; 2 rows (y) and 8
  // (z) ran 2.4 and 3.2 us faster than one tile per row.
  static constexpr int NQZ = NL < 8 ? NL : 8, NQY = 2;
  static constexpr int blocks(int nq) { return N / nq + N * nfull; }  // workgroups per grid
  static_assert((N / 2) % NL == 0 && NQZ <= NL && NQY <= NL && N % NQZ == 0 && N % NQY == 0,
                "column tiles must divide the rows");
};

// Where a column tile's lines sit: line c, point k at base + k * es + c * ls,
// with its kx (and, for the row-indexed pass, its row) per line.
template <int N, bool NYQ, int NQ>
struct PmColTile {
  using C = PmCol<N>;
  int r0, c0;  // first row (ky or z) and first kx
  int64_t ls;  // line stride: one row apart (row_stride) for the Nyquist tiles
  __device__ PmColTile(int b, int64_t row_stride) {
    if (NYQ) {
      r0 = b * NQ;
      c0 = N / 2;
      ls = row_stride;
    } else {
      r0 = (b - N / NQ) / C::nfull;
      c0 = ((b - N / NQ) % C::nfull) * C::NL;
      ls = 1;
    }
  }
  __device__ bool used(int c) const { return !NYQ || NQ == C::NL || c < NQ; }
  __device__ int kx(int c) const { return NYQ ? N / 2 : c0 + c; }
  __device__ int row(int c) const { return NYQ ? r0 + c : r0; }
};

// z columns at fixed ky: forward z FFT, Green's function g = -4 pi G / k^2 / G^3,
// outputs A = IFFTz(g rho) (in place of the input) and B = IFFTz(-i kz g rho)
// (to outB; Nyquist kz zeroed).  One LDS buffer: the Green-scaled spectrum P
// is formed in place, the B transform runs in place too, and each thread
// keeps the P values its first B pass reads, to put P back for the A
// transform (four workgroups per CU instead of two).
template <int N, bool NYQ>
__device__ __forceinline__ void pm_zpass_tile(float2* __restrict__ spec, float2* __restrict__ outB, float kf,
                                              float green, float2* X, const float2* tw) {
  using C = PmCol<N>;
  constexpr int R0 = FftPlan<N>::R0, Q = N / R0;
  const PmColTile<N, NYQ, C::NQZ> tl(blockIdx.x, C::RP);
  const int64_t base = (int64_t)tl.r0 * C::RP + tl.c0, es = (int64_t)N * C::RP;
  auto gsrc = [&](int c, int k) -> float2 {
    return tl.used(c) ? spec[base + k * es + c * tl.ls] : make_float2(0.f, 0.f);
  };
  // forward transform, scaled by the Green's function, into X (in place)
  auto green_dst = [&](int c, int kz, float2 v) {
    const int kx = tl.kx(c), ky = tl.row(c);
    const int ny = ky <= N / 2 ? ky : ky - N;
    const int nz = kz <= N / 2 ? kz : kz - N;
    const float fx = kf * kx, fy = kf * ny, fz = kf * nz;
    const float k2 = fx * fx + fy * fy + fz * fz;
    const float g = (kx == 0 && ky == 0 && kz == 0) ? 0.0f : green / k2;
    X[c * C::LS + fft_pad(kz)] = make_float2(v.x * g, v.y * g);
  };
  fft_lines<N, false, true, false, true>(X, tw, C::NL, gsrc, green_dst);
  // B = -i kz P, transformed in place
  float2 keep[R0];
  int kc = 0;
  auto bsrc = [&](int c, int kz) -> float2 {
    const float2 p = X[c * C::LS + fft_pad(kz)];
    keep[kc++] = p;
    const int nz = kz <= N / 2 ? kz : kz - N;
    const float az = (kz == N / 2) ? 0.0f : kf * nz;
    return make_float2(az * p.y, -az * p.x);
  };
  auto bdst = [&](int c, int k, float2 v) {
    if (tl.used(c)) outB[base + k * es + c * tl.ls] = v;
  };
  fft_lines<N, true, true, true, false>(X, tw, C::NL, bsrc, bdst);
  // P back into X: the first pass of a line-fastest transform gives thread t
  // line t % NL and points t / NL + r Q (r = 0 .. R0 - 1)
  const int t = threadIdx.x;
  if (t < C::NL * Q) {
    const int line = t % C::NL, j = t / C::NL;
#pragma unroll
    for (int r = 0; r < R0; ++r) X[line * C::LS + fft_pad(j + r * Q)] = keep[r];
  }
  __syncthreads();
  auto asrc = [&](int c, int k) { return X[c * C::LS + fft_pad(k)]; };
  auto adst = [&](int c, int k, float2 v) {
    if (tl.used(c)) spec[base + k * es + c * tl.ls] = v;
  };
  fft_lines<N, true, true, true, false>(X, tw, C::NL, asrc, adst);
}

template <int N>
__global__ __launch_bounds__(PmCol<N>::T) void k_pm_zpass(float2* __restrict__ spec, float2* __restrict__ outB,
                                                         float kf, float green) {
  using C = PmCol<N>;
  spec += blockIdx.y * (int64_t)N * N * C::RP;  // grid blockIdx.y of a batch
  outB += blockIdx.y * (int64_t)N * N * C::RP;
  __shared__ float2 X[C::NL * C::LS];
  __shared__ float2 tw[N];
  make_twiddles<N>(tw);
  __syncthreads();
  if (blockIdx.x < (unsigned)(N / C::NQZ))
    pm_zpass_tile<N, true>(spec, outB, kf, green, X, tw);
  else
    pm_zpass_tile<N, false>(spec, outB, kf, green, X, tw);
}

// y columns at fixed z: reads A (s0) and B (s2); writes
//   s1 = IFFTy(-i ky A)      (Nyquist ky zeroed)
//   s0 = IFFTy(A)            (in place; the x pass applies -i kx)
//   s2 = IFFTy(B)            (in place)
template <int N, bool NYQ>
__device__ __forceinline__ void pm_yinv_tile(float2* __restrict__ s0, float2* __restrict__ s1,
                                             float2* __restrict__ s2, float kf, float2* X, const float2* tw) {
  using C = PmCol<N>;
  const PmColTile<N, NYQ, C::NQY> tl(blockIdx.x, (int64_t)N * C::RP);
  const int64_t base = (int64_t)tl.r0 * N * C::RP + tl.c0, es = C::RP;
  auto ld = [&](const float2* p, int c, int k) -> float2 {
    return tl.used(c) ? p[base + k * es + c * tl.ls] : make_float2(0.f, 0.f);
  };
  auto st = [&](float2* p, int c, int k, float2 v) {
    if (tl.used(c)) p[base + k * es + c * tl.ls] = v;
  };
  // s0 is read once: the first pass of the -i ky transform keeps the values
  // it reads, and the first pass of the plain transform (same thread, same
  // points, same order) takes them back
  constexpr int R0 = FftPlan<N>::R0;
  float2 keep[R0];
  int kc = 0;
  auto src_ky = [&](int c, int ky) -> float2 {
    const float2 v = ld(s0, c, ky);
    keep[kc++] = v;
    const float f = (ky == N / 2) ? 0.0f : kf * (float)(ky <= N / 2 ? ky : ky - N);
    return make_float2(f * v.y, -f * v.x);
  };
  fft_lines<N, true, true>(X, tw, C::NL, src_ky, [&](int c, int k, float2 v) { st(s1, c, k, v); });
  int kr = 0;
  fft_lines<N, true, true>(X, tw, C::NL, [&](int, int) { return keep[kr++]; },
                           [&](int c, int k, float2 v) { st(s0, c, k, v); });
  fft_lines<N, true, true>(X, tw, C::NL, [&](int c, int k) { return ld(s2, c, k); },
                           [&](int c, int k, float2 v) { st(s2, c, k, v); });
}

template <int N>
__global__ __launch_bounds__(PmCol<N>::T) void k_pm_yinv(float2* __restrict__ s0, float2* __restrict__ s1,
                                                        float2* __restrict__ s2, float kf) {
  using C = PmCol<N>;
  s0 += blockIdx.y * (int64_t)N * N * C::RP;  // grid blockIdx.y of a batch
  s1 += blockIdx.y * (int64_t)N * N * C::RP;
  s2 += blockIdx.y * (int64_t)N * N * C::RP;
  __shared__ float2 X[C::NL * C::LS];
  __shared__ float2 tw[N];
  make_twiddles<N>(tw);
  __syncthreads();
  if (blockIdx.x < (unsigned)(N / C::NQY))
    pm_yinv_tile<N, true>(s0, s1, s2, kf, X, tw);
  else
    pm_yinv_tile<N, false>(s0, s1, s2, kf, X, tw);
}

// x rows: three half spectra (A~, -i ky A~, B~) -> real rows; the first gets
// -i kx (Nyquist kx zeroed).  Output: three FP32 planes a_x, a_y, a_z of N^3.
template <int N>
struct PmRow {
  static constexpr int NL = FftLines<N, FFT_THREADS>::value;
  static constexpr int LS = FftStride<N>::value;
  static constexpr int nxc = N / 2 + 1;
  static constexpr int RP = (int)pm_pitch(N);
};

template <int N>
__global__ __launch_bounds__(FFT_THREADS) void k_pm_xc2r(const float2* __restrict__ s0,
                                                         const float2* __restrict__ s1,
                                                         const float2* __restrict__ s2, float kf,
                                                         float* __restrict__ force, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  using C = PmRow<N>;
  __shared__ float2 X[C::NL * C::LS];
  __shared__ float2 tw[N];
  constexpr int64_t PLANE = (int64_t)N * N * N;
  s0 += blockIdx.y * (int64_t)N * N * C::RP;  // grid blockIdx.y of a batch
  s1 += blockIdx.y * (int64_t)N * N * C::RP;
  s2 += blockIdx.y * (int64_t)N * N * C::RP;
  force += blockIdx.y * 3 * PLANE;
  make_twiddles<N>(tw);
  __syncthreads();
  const int64_t row0 = (int64_t)blockIdx.x * C::NL;
  auto herm = [&](const float2* p, int l, int k) -> float2 {
    float2 v = p[(row0 + l) * C::RP + (k < C::nxc ? k : N - k)];
    if (k >= C::nxc) v.y = -v.y;
    return v;
  };
  // -i kx v is zero at k = 0 and N/2, where C2R would drop imaginary parts
  auto src0_full = [&](int l, int k) -> float2 {
    const float2 v = herm(s0, l, k);
    const float f = (k == N / 2) ? 0.0f : kf * (float)(k < C::nxc ? k : k - N);
    return make_float2(f * v.y, -f * v.x);
  };
  auto src1 = [&](int l, int k) -> float2 {
    float2 v = herm(s1, l, k);
    if (k == 0 || k == N / 2) v.y = 0.0f;
    return v;
  };
  auto src2 = [&](int l, int k) -> float2 {
    float2 v = herm(s2, l, k);
    if (k == 0 || k == N / 2) v.y = 0.0f;
    return v;
  };
  fft_lines<N, true>(X, tw, C::NL, src0_full, [&](int l, int x, float2 v) { force[(row0 + l) * N + x] = v.x; });
  fft_lines<N, true>(X, tw, C::NL, src1, [&](int l, int x, float2 v) { force[PLANE + (row0 + l) * N + x] = v.x; });
  fft_lines<N, true>(X, tw, C::NL, src2,
                     [&](int l, int x, float2 v) { force[2 * PLANE + (row0 + l) * N + x] = v.x; });
}

// x rows R2C of the density, two real rows per complex transform:
// FFT(a + i b) = Z, A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i.
// Output rows at the padded pitch.
template <int N>
__global__ __launch_bounds__(FFT_THREADS) void k_pm_rows_r2c(const float* __restrict__ rho,
                                                             float2* __restrict__ out) {
  using C = PmRow<N>;
  constexpr int H = C::NL / 2;
  rho += blockIdx.y * (int64_t)N * N * N;  // grid blockIdx.y of a batch
  out += blockIdx.y * (int64_t)N * N * C::RP;
  __shared__ float2 X[H * C::LS];
  __shared__ float2 tw[N];
  make_twiddles<N>(tw);
  __syncthreads();
  const int64_t row0 = (int64_t)blockIdx.x * C::NL;
  fft_lines<N, false, false, false, true>(
      X, tw, H,
      [&](int l, int x) { return make_float2(rho[(row0 + l) * N + x], rho[(row0 + l + H) * N + x]); },
      [&](int l, int k, float2 v) { X[l * C::LS + fft_pad(k)] = v; });
  for (int e = threadIdx.x; e < H * C::nxc; e += FFT_THREADS) {
    const int l = e / C::nxc, k = e - l * C::nxc;
    const float2 z = X[l * C::LS + fft_pad(k)], w = X[l * C::LS + fft_pad((N - k) & (N - 1))];
    out[(row0 + l) * C::RP + k] = make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
    out[(row0 + l + H) * C::RP + k] = make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x));
  }
}

// ------------------------------------------------------------- gather + KDK
// max of v over the block, then one atomicMax per block
__device__ __forceinline__ void block_max_to(float v, uint32_t* target) {
  __shared__ float red[16];
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmaxf(m, red[i]);
    atomicMax(target, __float_as_uint(m));
  }
}

// four FP64 sums over the block in a fixed order (lane tree, then waves in
// order), written by thread 0
__device__ __forceinline__ void block_sum4_to(double (&v)[4], double* out) {
  __shared__ double red4[16][4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) red4[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double m = 0.0;
      for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m += red4[i][c];
      out[c] = m;
    }
}

constexpr int PM_GU = 1;  // particles in flight per thread in the gather

// CIC gather of the force planes fused with the kicks and the drift (grid
// stride over the particles in order, so the cell-sorted state streams
// contiguously); GU particles per thread have their loads in flight at once.
template <bool P2, int GU>
__global__ __launch_bounds__(256) void k_gather_kdk(float* __restrict__ pos, float* __restrict__ vel,
                                                    float* __restrict__ acc, const float* __restrict__ force,
                                                    const uint32_t* __restrict__ perm, int64_t n, int G, float h,
                                                    float inv_h, float hdt, float dt, float L, int do_kick, int drift,
                                                    PmDev* __restrict__ st) {
  const int64_t G3 = (int64_t)G * G * G;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float vmax = 0.0f;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i0 < n; i0 += stride * GU) {
    int64_t pi[GU];
    float p[GU][3], vv[GU][3];
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      const int64_t i = i0 + u * stride;
      pi[u] = -1;
      if (i < n) {
        pi[u] = perm ? (int64_t)perm[i] : i;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          p[u][d] = pos[3 * pi[u] + d];
          vv[u][d] = do_kick ? vel[3 * pi[u] + d] : 0.0f;
        }
      }
    }
    float a[GU][3];
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      a[u][0] = a[u][1] = a[u][2] = 0.0f;
      if (pi[u] < 0) continue;
      int i0c[3];
      float w[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float uu = P2 ? cic_u_pow2(p[u][d], inv_h) : cic_u(p[u][d], 0.0f, h);
        const float f = floorf(uu);
        i0c[d] = (int)f;
        w[d] = __fsub_rn(uu, f);
      }
      const int x0 = wrap(i0c[0], G), x1 = wrap(i0c[0] + 1, G);
      const int y0 = wrap(i0c[1], G), y1 = wrap(i0c[1] + 1, G);
      const int z0 = wrap(i0c[2], G), z1 = wrap(i0c[2] + 1, G);
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
            a[u][0] += force[cell] * wt;
            a[u][1] += force[G3 + cell] * wt;
            a[u][2] += force[2 * G3 + cell] * wt;
          }
    }
#pragma unroll
    for (int u = 0; u < GU; ++u) {
      if (pi[u] < 0) continue;
      const int64_t q = pi[u];
      // between steps of one call the acceleration lives only in the
      // velocities (both half kicks are applied here); it is stored when no
      // drift follows
      if (!drift) {
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[3 * q + d] = a[u][d];
      }
      if (do_kick) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float v = __fmaf_rn(hdt, a[u][d], vv[u][d]);
          if (drift) {
            v = __fmaf_rn(hdt, a[u][d], v);
            float x = __fmaf_rn(dt, v, p[u][d]);
            if (x >= L) x -= L;
            if (x < 0.0f) x += L;
            if (x >= L) x = 0.0f;  // -tiny + L rounds to L
            pos[3 * q + d] = x;
            vmax = fmaxf(vmax, fabsf(v));
          }
          vel[3 * q + d] = v;
        }
      }
    }
  }
  if (st && drift) block_max_to(vmax, &st->vmax_bits);
}

// Tile gather + KDK for the sorted (tile-major) stepper state.  A gather
// piece is TG_SUB whole z layers of a tile (32x16x4 cells), so its particles
// are one contiguous slot range.  A workgroup stages the force planes over
// the piece grown by DMAX cells below and 1 + DMAX above (35x19x7 cells: the
// CIC corners of every particle whose sorted cell is in the piece, for drift
// up to DMAX cells) in LDS, then streams the piece's particles through the
// same arithmetic as k_gather_kdk.  Two workgroups per CU.
constexpr int TG_THREADS = 512;
#ifndef AUGUR_TG_SUB
#define AUGUR_TG_SUB 4
#endif
constexpr int TG_SUB = AUGUR_TG_SUB;  // z layers per piece
constexpr int TG_NSUB = TD_Z / TG_SUB;
constexpr uint32_t TG_CHUNK = 8192;  // most particles per work item
constexpr int TG_U = 4;              // particles per thread in flight
constexpr int TG_RX = TD_X + 1 + 2 * DMAX, TG_RY = TD_Y + 1 + 2 * DMAX, TG_RZ = TG_SUB + 1 + 2 * DMAX;
constexpr int TG_REG = TG_RX * TG_RY * TG_RZ;
// LDS image of the region: rows of TG_P floats holding columns x0 - 4 ..
// x0 + TD_X + 3 (16-byte chunks of the force planes, loaded straight into LDS)
constexpr int TG_XOFF = 4;                    // LDS column of x0
constexpr int TG_P = TD_X + 2 * TG_XOFF;      // row pitch (floats)
constexpr int TG_QR = TG_P / 4;               // 16-byte chunks per row
constexpr int TG_NROW = TG_RY * TG_RZ;        // rows per plane
constexpr int TG_NQ = TG_NROW * TG_QR;        // chunks per plane
constexpr int TG_PLANE = TG_NROW * TG_P;      // floats per plane
static_assert(DMAX < TG_XOFF && TD_X + 1 + DMAX <= TD_X + TG_XOFF, "region columns inside the LDS rows");

// sort-time plan: items per piece (one thread each); items[npieces] = 0
__device__ __forceinline__ void gather_plan_body(int bid, const uint32_t* __restrict__ cell_start, int npieces,
                                                 uint32_t* __restrict__ items) {
  const int pc = bid * (int)blockDim.x + threadIdx.x;
  if (pc > npieces) return;
  if (pc == npieces) {
    items[npieces] = 0u;
    return;
  }
  constexpr uint32_t PC = TG_SUB * TD_Y * TD_X;  // cells per piece; pieces are consecutive in key order
  const uint32_t cnt = cell_start[(uint32_t)(pc + 1) * PC] - cell_start[(uint32_t)pc * PC];
  items[pc] = cnt == 0u ? 1u : (cnt + TG_CHUNK - 1) / TG_CHUNK;
}

// item -> (piece, slot range): the piece's particles cut into near-equal parts
__device__ __forceinline__ void gather_items_body(int bid, const uint32_t* __restrict__ cell_start,
                                                  const uint32_t* __restrict__ item_start, int npieces,
                                                  DepItem* __restrict__ info) {
  const int pc = bid * (int)blockDim.x + threadIdx.x;
  if (pc >= npieces) return;
  constexpr uint32_t PC = TG_SUB * TD_Y * TD_X;
  const uint32_t lo = cell_start[(uint32_t)pc * PC], hi = cell_start[(uint32_t)(pc + 1) * PC];
  const uint32_t i0 = item_start[pc], nch = item_start[pc + 1] - i0;
  const uint32_t per = (hi - lo + nch - 1) / nch;
  for (uint32_t c = 0; c < nch; ++c)
    info[i0 + c] = DepItem{(uint32_t)pc, min(hi, lo + c * per), min(hi, lo + (c + 1) * per), 0u};
}

// The deposit's and the gather's sort-time plans share launches (256
// threads): blocks [0, dep_blocks) plan the tiles (a wave per tile), the
// rest the gather pieces (a thread per piece); then, after the scans of both
// item counts, blocks [0, item_blocks) cut the tiles' work items and the
// rest the pieces'.
__global__ __launch_bounds__(256) void k_plans(const uint32_t* __restrict__ cell_start, int G, int ntiles,
                                               uint32_t* __restrict__ items, uint32_t* __restrict__ seg_lo,
                                               uint32_t* __restrict__ seg_pre, int dep_blocks, int npieces,
                                               uint32_t* __restrict__ gitems, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  if ((int)blockIdx.x < dep_blocks)
    dep_plan_body((int)blockIdx.x, cell_start, G, ntiles, ntiles, items, seg_lo, seg_pre);
  else
    gather_plan_body((int)blockIdx.x - dep_blocks, cell_start, npieces, gitems);
}

template <int MAXSEG, uint32_t CHUNK>
__global__ __launch_bounds__(256) void k_plan_items(const uint32_t* __restrict__ item_start, int ntiles,
                                                    const uint32_t* __restrict__ tseg_pre, DepItem* __restrict__ info,
                                                    uint16_t* __restrict__ blk, int max_items, int item_blocks,
                                                    const uint32_t* __restrict__ cell_start,
                                                    const uint32_t* __restrict__ gitem_start, int npieces,
                                                    DepItem* __restrict__ ginfo, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  if ((int)blockIdx.x < item_blocks) {  // the tiles' items from a capped grid (a short launch when gated off)
    for (int item = blockIdx.x; item < max_items; item += item_blocks)
      tile_items_body<MAXSEG, CHUNK>((uint32_t)item, item_start, ntiles, tseg_pre, info, blk);
  } else {
    gather_items_body((int)blockIdx.x - item_blocks, cell_start, gitem_start, npieces, ginfo);
  }
}

// Raw buffer views of an item's particle arrays: a lane past the item's end
// gets zeros from a load and its store is dropped by the range check, so the
// particle loop has no per-lane branches and the loads of the next batch are
// in flight while the stores of this one drain.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t item_view(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));

// DRIFT: also the next opening half kick and the drift (pos, vel written);
// otherwise the closing half kick only (acc, vel written).  DIAG: per-
// workgroup FP64 sums of m v and m v^2 / 2 at the closing-kick velocity.
template <bool P2, bool DRIFT, bool DIAG>
__global__ __launch_bounds__(TG_THREADS) void k_gather_tiles(
    float* __restrict__ pos, float* __restrict__ vel, float* __restrict__ acc, const float* __restrict__ force,
    int64_t n, int G, float h, float inv_h, float hdt, float dt, float L,
    const uint32_t* __restrict__ item_start, const DepItem* __restrict__ info, int npieces, PmDev* __restrict__ st,
    const float* __restrict__ mass, float mass_value, double* __restrict__ dpart) {
  __shared__ __attribute__((aligned(16))) float fr[3 * TG_PLANE];
  const float* fx = fr;
  const float* fy = fr + TG_PLANE;
  const float* fz = fr + 2 * TG_PLANE;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t G3 = (int64_t)G * G * G;
  float vmax = 0.0f;
  // FP64 sums of m v and m v^2 / 2 at the synchronised (full-step) velocity:
  // one partial per workgroup, summed in a fixed order
  double dsum[4] = {0.0, 0.0, 0.0, 0.0};
  const uint32_t total_items = item_start[npieces];
  // XCD-coherent order: workgroups are dealt to the 8 XCDs round-robin, so
  // workgroup b takes item (b % 8) * per + b / 8 and each XCD works through
  // one contiguous range of pieces, whose shared region halos then hit in
  // that XCD's L2 (half the fetched bytes of the round-robin order)
  const uint32_t per = (total_items + 7) / 8;
  const uint32_t xcd = blockIdx.x % 8, slot0 = blockIdx.x / 8, slots = (gridDim.x + 7 - xcd) / 8;
  const uint32_t lim = min(total_items, (xcd + 1) * per);
  for (uint32_t item = xcd * per + slot0; item < lim; item += slots) {
    const DepItem it = info[item];
    const TileOrigin to = tile_origin((int)it.tile / TG_NSUB, G);
    const int x0 = to.x0, y0 = to.y0, z0 = to.z0 + ((int)it.tile % TG_NSUB) * TG_SUB;
    const uint32_t cnt = it.vhi - it.vlo;  // at most TG_CHUNK
    const __amdgpu_buffer_rsrc_t rpos = item_view(pos + 3 * (int64_t)it.vlo, 12u * cnt);
    const __amdgpu_buffer_rsrc_t rvel = item_view(vel + 3 * (int64_t)it.vlo, 12u * cnt);
    const __amdgpu_buffer_rsrc_t racc = item_view(acc + 3 * (int64_t)it.vlo, 12u * cnt);
    const __amdgpu_buffer_rsrc_t rmass = item_view(mass ? mass + it.vlo : pos, mass ? 4u * cnt : 0u);
    // force planes over the grown piece, origin (x0, y0, z0) - DMAX wrapped,
    // loaded straight into LDS in 16-byte chunks (one wave instruction fills
    // 64 consecutive chunks of the image; chunk (plane, row, j) holds columns
    // x0 - 4 + 4j .. + 3): every load of the region issued at once
    constexpr int NWQ = (3 * TG_NQ + 63) / 64;
    for (int wq = wave; wq < NWQ; wq += TG_THREADS / 64) {
      const int Q = wq * 64 + lane;
      if (Q < 3 * TG_NQ) {
        const int pl = Q / TG_NQ, q = Q - pl * TG_NQ;
        const int row = q / TG_QR, j = q - row * TG_QR;
        const int yy = row % TG_RY, zz = row / TG_RY;
        const int gx = (x0 - TG_XOFF + 4 * j) & (G - 1), gy = (y0 - DMAX + yy) & (G - 1),
                  gz = (z0 - DMAX + zz) & (G - 1);
        const float* src = force + pl * G3 + ((int64_t)gz * G + gy) * G + gx;
        __builtin_amdgcn_global_load_lds(src, fr + wq * 256, 16, 0, 0);
      }
    }
    // particle j of the item (byte offset 12 j) is lane tid + TG_THREADS u of batch j0
    float p[TG_U][3], vel0[TG_U][3], pm[TG_U];
    auto load_particles = [&](uint32_t j0, float (&pp)[TG_U][3], float (&vv)[TG_U][3], float (&mm)[TG_U]) {
#pragma unroll
      for (int u = 0; u < TG_U; ++u) {
        const uint32_t off = 12u * (j0 + tid + u * TG_THREADS);
        const u32x3 a = __builtin_amdgcn_raw_buffer_load_b96(rpos, off, 0, 0);
        const u32x3 b = __builtin_amdgcn_raw_buffer_load_b96(rvel, off, 0, 0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          pp[u][d] = __uint_as_float(a[d]);
          vv[u][d] = __uint_as_float(b[d]);
        }
        mm[u] = DIAG ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rmass, off / 3u, 0, 0)) : 0.0f;
      }
    };
    load_particles(0, p, vel0, pm);
    // the region (LDS) and the first batch (registers) landed: waiting here
    // (vmcnt 0) also leaves no load pending on the loop-carried registers at
    // the loop head, so the compiler puts no wait there (which would stall on
    // the last batch's stores)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    // the piece's particles in batches of TG_THREADS * TG_U, the whole
    // workgroup on every batch (lanes past the end masked by the views)
    for (uint32_t j0 = 0; j0 < cnt; j0 += TG_THREADS * TG_U) {
      // the next batch's loads go out first and land during this batch's
      // arithmetic (and ahead of its stores); with the diagnostics' FP64 sums
      // that does not fit in four waves' registers per SIMD, and they go out
      // after the arithmetic, still ahead of the stores
      float pn[TG_U][3], vn[TG_U][3], pmn[TG_U];
      if (!DIAG) {
        load_particles(j0 + TG_THREADS * TG_U, pn, vn, pmn);
        __builtin_amdgcn_sched_barrier(0);
      }
      float o1[TG_U][3], o2[TG_U][3];  // DRIFT: pos, vel; otherwise acc, vel
#pragma unroll
      for (int u = 0; u < TG_U; ++u) {
        const bool live = j0 + tid + u * TG_THREADS < cnt;
        float w[3];
        int lb[3];
        const int t0[3] = {x0, y0, z0};
        const int tmax[3] = {TD_X + 2 * DMAX - 1, TD_Y + 2 * DMAX - 1, TG_SUB + 2 * DMAX - 1};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float uu = P2 ? cic_u_pow2(p[u][d], inv_h) : cic_u(p[u][d], 0.0f, h);
          const float f = floorf(uu);
          w[d] = __fsub_rn(uu, f);
          // base cell in region coordinates; the drift bound keeps it in range
          // (the clamp only guards the LDS bounds, also for masked lanes)
          lb[d] = min(((int)f - t0[d] + DMAX) & (G - 1), tmax[d]);
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
              const int c = ((lb[2] + oz) * TG_RY + lb[1] + oy) * TG_P + lb[0] + (TG_XOFF - DMAX) + ox;
              a[0] += fx[c] * wt;
              a[1] += fy[c] * wt;
              a[2] += fz[c] * wt;
            }
        float vs[3];  // closing half kick: the synchronised velocity of this step
#pragma unroll
        for (int d = 0; d < 3; ++d) vs[d] = __fmaf_rn(hdt, a[d], vel0[u][d]);
        if (DIAG) {
          const double m = live ? (mass ? (double)pm[u] : (double)mass_value) : 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) dsum[d] += m * (double)vs[d];
          dsum[3] += 0.5 * m * ((double)vs[0] * vs[0] + (double)vs[1] * vs[1] + (double)vs[2] * vs[2]);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (DRIFT) {
            const float vv = __fmaf_rn(hdt, a[d], vs[d]);
            float x = __fmaf_rn(dt, vv, p[u][d]);
            if (x >= L) x -= L;
            if (x < 0.0f) x += L;
            if (x >= L) x = 0.0f;  // -tiny + L rounds to L
            o1[u][d] = x;
            o2[u][d] = vv;
            vmax = fmaxf(vmax, live ? fabsf(vv) : 0.0f);
          } else {
            o1[u][d] = a[d];
            o2[u][d] = vs[d];
          }
        }
      }
      if (DIAG) load_particles(j0 + TG_THREADS * TG_U, pn, vn, pmn);
#pragma unroll
      for (int u = 0; u < TG_U; ++u) {
        const uint32_t off = 12u * (j0 + tid + u * TG_THREADS);
        u32x3 a, b;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          a[d] = __float_as_uint(o1[u][d]);
          b[d] = __float_as_uint(o2[u][d]);
        }
        __builtin_amdgcn_raw_buffer_store_b96(a, DRIFT ? rpos : racc, off, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b96(b, rvel, off, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < TG_U; ++u) {
        pm[u] = pmn[u];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          p[u][d] = pn[u][d];
          vel0[u][d] = vn[u][d];
        }
      }
    }
    __syncthreads();  // LDS is reused by the next item
  }
  if (DRIFT) block_max_to(vmax, &st->vmax_bits);
  if (DIAG) block_sum4_to(dsum, dpart + (int64_t)blockIdx.x * 4);
}

// sums the gather's per-workgroup FP64 partials in a fixed order: out[0..3]
constexpr int DIAG_THREADS = 1024;
__global__ __launch_bounds__(DIAG_THREADS) void k_diag_finish(const double* __restrict__ dpart, int64_t nblk,
                                                              double* __restrict__ out) {
  double acc4[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t b = threadIdx.x; b < nblk; b += DIAG_THREADS)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc4[c] += dpart[b * 4 + c];
  __shared__ double red[DIAG_THREADS][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) red[threadIdx.x][c] = acc4[c];
  __syncthreads();
  for (int w = DIAG_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 4; ++c) red[threadIdx.x][c] += red[threadIdx.x + w][c];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[0][threadIdx.x];
}

// ---------------------------------------- ragged batch on the tiled deposit
// Keys of a ragged batch (scene b holds particles [starts[b], starts[b+1])):
// b G^3 + tile_key of the particle's base cell, so that the batch's sort
// order is scene-major tile-major and its tiles are scene * tps + tile.
__global__ void k_ragged_keys(const float* __restrict__ pos, const int64_t* __restrict__ starts, int B, int64_t n,
                              int G, float h, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                              uint32_t* __restrict__ ngaps, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ngaps = 0u;  // for k_cell_starts, later in the same call
  const uint32_t G3 = (uint32_t)G * G * G;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = B;  // the last scene with starts[scene] <= i
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (starts[m] <= i) lo = m; else hi = m;
    }
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = wrap((int)floorf(cic_u(pos[3 * i + d], 0.0f, h)), G);
    keys[i] = (uint32_t)lo * G3 + tile_key(c[0], c[1], c[2], G);
    vals[i] = (uint32_t)i;
  }
}

__global__ __launch_bounds__(256) void k_dep_plans(const uint32_t* __restrict__ cell_start, int G, int ntiles, int tps,
                                                   uint32_t* __restrict__ items, uint32_t* __restrict__ seg_lo,
                                                   uint32_t* __restrict__ seg_pre, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  dep_plan_body((int)blockIdx.x, cell_start, G, ntiles, tps, items, seg_lo, seg_pre);
}

__global__ __launch_bounds__(256) void k_dep_items(const uint32_t* __restrict__ item_start, int ntiles,
                                                   const uint32_t* __restrict__ tseg_pre, DepItem* __restrict__ info,
                                                   uint16_t* __restrict__ blk, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  tile_items_body<TD_MAXSEG, TD_CHUNK>(blockIdx.x, item_start, ntiles, tseg_pre, info, blk);
}

// the int64 images of the heavy tiles (more than one item) zeroed before
// their items add into them (a transient workspace is not kept zero)
__global__ __launch_bounds__(256) void k_dep_zero_heavy(const uint32_t* __restrict__ item_start,
                                                        unsigned long long* __restrict__ fix, const uint32_t* gate) {
  if (gate && *gate == 0u) return;
  const int tile = blockIdx.x;
  if (item_start[tile + 1] - item_start[tile] <= 1u) return;
  unsigned long long* img = fix + (int64_t)tile * TD_CELLS;
  for (int i = threadIdx.x; i < TD_CELLS; i += 256) img[i] = 0ull;
}

// Every kick and drift is one explicit FMA (v + (dt/2) a, x + dt v) in all the
// stepping kernels, so a step gives the same bits whether its kicks are fused
// into the gather or run here.
__global__ void k_kick_drift(float* __restrict__ pos, float* __restrict__ vel, const float* __restrict__ acc, int64_t n,
                             float hdt, float dt, float L, PmDev* __restrict__ st) {
  float vmax = 0.0f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float v = __fmaf_rn(hdt, acc[3 * i + d], vel[3 * i + d]);
      float x = __fmaf_rn(dt, v, pos[3 * i + d]);
      if (x >= L) x -= L;
      if (x < 0.0f) x += L;
      if (x >= L) x = 0.0f;  // -tiny + L rounds to L
      vel[3 * i + d] = v;
      pos[3 * i + d] = x;
      vmax = fmaxf(vmax, fabsf(v));
    }
  }
  block_max_to(vmax, &st->vmax_bits);
}

__global__ void k_pm_init(PmDev* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->cum_cells = 0.0f;
    st->vmax_bits = 0u;
    st->need_sort = 0u;
    st->maxbits = 0u;
  }
}

// ------------------------------------------------------------- workspace
struct PmBufs {
  PmDev* st;
  uint32_t* cell_start;  // G^3 + 1
  float* rho;            // G^3
  float2* s0;            // G*G*pitch complex (spectrum / A / out0)
  float2* s1;            // out1
  float2* s2;            // B / out2
  float* force;          // 3 G^3: planes a_x, a_y, a_z
  uint32_t* keys;        // n
  uint32_t* vals;        // n
  void* rws;             // radix sort scratch
  void* scan_ws;         // scan scratch (G^3 + 1)
  float* tmp;            // 8n: the permuted state (pos, vel, ids, mass)
  uint32_t* gaps;        // 3 per run of empty cells longer than CS_SPAN
  uint32_t* ngaps;
  uint32_t* perm;        // n (for unsorted force evaluation)
  uint32_t* items;       // ntiles + 1: work items per tile, then their exclusive scan
  unsigned long long* fix;  // G^3: int64 tile images of heavy tiles (kept zero between launches)
  uint32_t* tseg_lo;     // ntiles * TD_MAXSEG: row segment starts per tile
  uint32_t* tseg_pre;    // ntiles * (TD_MAXSEG + 1): inclusive prefix of their lengths
  DepItem* info;         // max_items: tile and visit range of each work item
  uint16_t* iblk;        // max_items * TD_NBLK: visit -> segment table of each item
  // gather plan: work items over the pieces (TG_SUB z layers of a tile)
  uint32_t* gitems;      // npieces + 1
  DepItem* ginfo;        // max_gitems
  double* dpart;         // 4 * max_gitems: per-workgroup FP64 diagnostics of the gather
};

static int64_t pm_tiles(int64_t G) { return (G / TD_X) * (G / TD_Y) * (G / TD_Z); }
static int64_t pm_pieces(int64_t G) { return pm_tiles(G) * TG_NSUB; }
// a cell lies in at most 2 tile regions per axis, so the planned visits are
// at most 8n and every tile has at least one item
static int64_t pm_max_items(int64_t n, int64_t G) { return pm_tiles(G) + (8 * n + TD_CHUNK - 1) / TD_CHUNK; }
static int64_t pm_max_gitems(int64_t n, int64_t G) { return pm_pieces(G) + (n + TG_CHUNK - 1) / TG_CHUNK; }

static bool carve_pm(Carver& c, int64_t n, int64_t G, PmBufs& b) {
  const int64_t G3 = G * G * G, nc = G * G * pm_pitch(G);
  b.st = c.take<PmDev>(1);
  b.cell_start = c.take<uint32_t>(G3 + 1);
  b.rho = c.take<float>(G3);
  b.s0 = c.take<float2>(nc);
  b.s1 = c.take<float2>(nc);
  b.s2 = c.take<float2>(nc);
  b.force = c.take<float>(3 * G3);
  b.keys = c.take<uint32_t>(n);
  b.vals = c.take<uint32_t>(n);
  b.rws = c.take<char>(radix_workspace_bytes(n));
  b.scan_ws = c.take<char>(scan_workspace_bytes(G3 + 1));
  b.tmp = c.take<float>(8 * n);
  b.gaps = c.take<uint32_t>(3 * ((G3 + 1) / CS_SPAN + 2));
  b.ngaps = c.take<uint32_t>(1);
  b.perm = c.take<uint32_t>(n);
  b.items = c.take<uint32_t>(pm_tiles(G) + 1);
  b.fix = c.take<unsigned long long>(G3);
  b.tseg_lo = c.take<uint32_t>(pm_tiles(G) * TD_MAXSEG);
  b.tseg_pre = c.take<uint32_t>(pm_tiles(G) * (TD_MAXSEG + 1));
  b.info = c.take<DepItem>(pm_max_items(n, G));
  b.iblk = c.take<uint16_t>(pm_max_items(n, G) * TD_NBLK);
  b.gitems = c.take<uint32_t>(pm_pieces(G) + 1);
  b.ginfo = c.take<DepItem>(pm_max_gitems(n, G));
  b.dpart = c.take<double>(4 * pm_max_gitems(n, G));
  return c.ok();
}

static int check_pm(int64_t n, int64_t G, float L) {
  AUGUR_REQUIRE(n >= 1 && n < ((int64_t)1 << 31), AUGUR_EINVAL, "pm: need 1 <= n < 2^31");
  AUGUR_REQUIRE(is_pow2(G) && G >= 32 && G <= 512, AUGUR_EUNSUPPORTED, "pm: G must be a power of two in [32, 512]");
  AUGUR_REQUIRE(L > 0.0f, AUGUR_EINVAL, "pm: L must be positive");
  return AUGUR_OK;
}

// fresh state: bookkeeping reset, heavy-tile images zeroed
static int pm_init(PmBufs& b, int64_t G, hipStream_t s) {
  hipLaunchKernelGGL(k_pm_init, dim3(1), dim3(64), 0, s, b.st);
  AUGUR_LAUNCHED("pm_init");
  const int64_t G3 = G * G * G;
  hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(G3, 256, 4096)), dim3(256), 0, s, b.fix, G3);
  AUGUR_LAUNCHED("zero_fix");
  return AUGUR_OK;
}

// Keys of the current positions, stable radix sort, cell_start; with
// physical != 0 the state arrays are permuted into sorted order, else the
// sorted order is left in b.perm.  All kernels are gated by `gate`.
static int pm_sort(float* pos, float* vel, float* acc, float* mass, int32_t* ids, int64_t n, int64_t G, float L,
                   PmBufs& b, const uint32_t* gate, bool physical, hipStream_t s) {
  TimedScope ts_("pm_sort", s);
  const float h = (float)((double)L / (double)G);
  const int64_t G3 = G * G * G;
  const dim3 g(blocks_for(n, 256, 1024)), bl(256);
  hipLaunchKernelGGL(k_base_keys, g, bl, 0, s, pos, n, (int)G, h, b.keys, b.vals, b.ngaps, gate);
  AUGUR_LAUNCHED("base_keys");
  uint32_t *ks, *perm;
  int rc = radix_sort_pairs(b.keys, b.vals, n, 3 * ilog2(G), b.rws, radix_workspace_bytes(n), &ks, &perm, s, gate);
  if (rc) return rc;
  // cell_start from the sorted keys
  hipLaunchKernelGGL(k_cell_starts, dim3(blocks_for(n + 1, 256, 2048)), bl, 0, s, (const uint32_t*)ks, n,
                     (uint32_t)G3, b.cell_start, b.gaps, b.ngaps, gate);
  AUGUR_LAUNCHED("cell_starts");
  hipLaunchKernelGGL(k_cell_fill, dim3(1024), bl, 0, s, (const uint32_t*)b.gaps, (const uint32_t*)b.ngaps,
                     b.cell_start, gate);
  AUGUR_LAUNCHED("cell_fill");
  // work items: the deposit's over the tile regions grown for the largest
  // drift bound, the gather's over the tile cores
  const int64_t ntiles = pm_tiles(G);
  const int64_t npc = pm_pieces(G);
  const int dep_blocks = (int)((ntiles + 1 + 3) / 4);
  hipLaunchKernelGGL(k_plans, dim3((unsigned)(dep_blocks + blocks_for(npc + 1, 256))), dim3(256), 0, s,
                     (const uint32_t*)b.cell_start, (int)G, (int)ntiles, b.items, b.tseg_lo, b.tseg_pre, dep_blocks,
                     (int)npc, b.gitems, gate);
  AUGUR_LAUNCHED("plans");
  rc = exclusive_scan2_u32(b.items, b.items, ntiles + 1, b.gitems, b.gitems, npc + 1, b.scan_ws, s, gate);
  if (rc) return rc;
  const int max_items = (int)pm_max_items(n, G);
  const int item_blocks = max_items < 2048 ? max_items : 2048;
  hipLaunchKernelGGL((k_plan_items<TD_MAXSEG, TD_CHUNK>), dim3((unsigned)(item_blocks + blocks_for(npc, 256))),
                     dim3(256), 0, s, (const uint32_t*)b.items, (int)ntiles, (const uint32_t*)b.tseg_pre, b.info,
                     b.iblk, max_items, item_blocks, (const uint32_t*)b.cell_start, (const uint32_t*)b.gitems,
                     (int)npc, b.ginfo, gate);
  AUGUR_LAUNCHED("plan_items");
  if (!physical) {
    hipLaunchKernelGGL(k_copy<uint32_t>, g, bl, 0, s, perm, b.perm, n, gate);
    AUGUR_LAUNCHED("copy_perm");
    return AUGUR_OK;
  }
  hipLaunchKernelGGL(k_permute_state, g, bl, 0, s, (const float*)pos, (const float*)vel, (const int32_t*)ids,
                     (const float*)mass, perm, n, b.tmp, gate);
  AUGUR_LAUNCHED("permute_state");
  hipLaunchKernelGGL(k_restore_state, g, bl, 0, s, (const float*)b.tmp, n, pos, vel, ids, mass, gate);
  AUGUR_LAUNCHED("restore_state");
  if (acc) {  // only when the caller's acc must follow (prepare)
    hipLaunchKernelGGL((k_permute<float, 3>), g, bl, 0, s, (const float*)acc, b.tmp, perm, n, gate);
    AUGUR_LAUNCHED("permute3");
    hipLaunchKernelGGL(k_copy<float>, dim3(blocks_for(3 * n, 256, 1024)), bl, 0, s, (const float*)b.tmp, acc, 3 * n,
                       gate);
    AUGUR_LAUNCHED("copy3");
  }
  return AUGUR_OK;
}

// launches KERNEL<GG> with GG = G as a compile-time constant usable in grid/block
#define PM_SWITCH_G(G, KERNEL, grid, block, s, ...)                                                   \
  switch (G) {                                                                                          \
    case 32: { constexpr int GG = 32; hipLaunchKernelGGL(KERNEL<GG>, grid, block, 0, s, __VA_ARGS__); break; }   \
    case 64: { constexpr int GG = 64; hipLaunchKernelGGL(KERNEL<GG>, grid, block, 0, s, __VA_ARGS__); break; }   \
    case 128: { constexpr int GG = 128; hipLaunchKernelGGL(KERNEL<GG>, grid, block, 0, s, __VA_ARGS__); break; } \
    case 256: { constexpr int GG = 256; hipLaunchKernelGGL(KERNEL<GG>, grid, block, 0, s, __VA_ARGS__); break; } \
    case 512: { constexpr int GG = 512; hipLaunchKernelGGL(KERNEL<GG>, grid, block, 0, s, __VA_ARGS__); break; } \
    default: set_error("pm: unsupported G %d", (int)(G)); return AUGUR_EUNSUPPORTED;                  \
  }

// the deposit's density scale 1 / h^3
static float pm_dep_scale(int64_t G, float L) {
  const float h = (float)((double)L / (double)G);
  return (float)(1.0 / ((double)h * (double)h * (double)h));
}

// deposit of sorted slots (optional perm) -> b.rho; bound_ready: the deposit
// bound is already in b.st (k_pm_sorted wrote it for uniform masses)
static int pm_deposit(const float* pos, const uint32_t* perm, const float* mass, float mass_value, int64_t n,
                      int64_t G, float L, PmBufs& b, hipStream_t s, bool bound_ready = false) {
  const float h = (float)((double)L / (double)G);
  const float scale = pm_dep_scale(G, L);
  if (!bound_ready) {
    TimedScope ts_("dep_bound", s);
    if (mass) {
      hipLaunchKernelGGL(k_pm_set_maxbits, dim3(1), dim3(64), 0, s, b.st, (const float*)nullptr, 0.0f, scale, n);
      AUGUR_LAUNCHED("dep_bound_reset");
      hipLaunchKernelGGL(k_pm_set_maxbits, dim3(blocks_for(n, 256, 2048)), dim3(256), 0, s, b.st, mass, mass_value,
                         scale, n);
    } else
      hipLaunchKernelGGL(k_pm_set_maxbits, dim3(1), dim3(64), 0, s, b.st, (const float*)nullptr, mass_value, scale, n);
    AUGUR_LAUNCHED("dep_bound");
  }
  const int64_t ntiles = pm_tiles(G);
  {
    TimedScope ts_("dep_tiles", s);
    // a workgroup for every possible item (a cell lies in at most 8 tile
    // regions), so no workgroup works through two items of a dense clump in
    // turn; workgroups past the plan's item count exit at once
    const int64_t grid = pm_max_items(n, G);
    const bool p2 = pow2_float(h);
    const float inv_h = p2 ? 1.0f / h : 0.0f;
    if (p2)
      hipLaunchKernelGGL(k_dep_tiles<true>, dim3((unsigned)grid), dim3(TD_THREADS), 0, s, pos, perm, mass,
                         mass_value, n, (int)G, h, inv_h, scale, (const PmDev*)b.st, (const uint32_t*)b.items,
                         (const DepItem*)b.info, (const uint16_t*)b.iblk, (const uint32_t*)b.tseg_lo,
                         (const uint32_t*)b.tseg_pre,
                         (int)ntiles, (int)ntiles, b.fix, b.rho);
    else
      hipLaunchKernelGGL(k_dep_tiles<false>, dim3((unsigned)grid), dim3(TD_THREADS), 0, s, pos, perm, mass,
                         mass_value, n, (int)G, h, inv_h, scale, (const PmDev*)b.st, (const uint32_t*)b.items,
                         (const DepItem*)b.info, (const uint16_t*)b.iblk, (const uint32_t*)b.tseg_lo,
                         (const uint32_t*)b.tseg_pre,
                         (int)ntiles, (int)ntiles, b.fix, b.rho);
    AUGUR_LAUNCHED("dep_tiles");
    hipLaunchKernelGGL(k_dep_finish, dim3((unsigned)ntiles), dim3(256), 0, s, (const uint32_t*)b.items, b.fix,
                       (const PmDev*)b.st, n, (int)G, (int)ntiles, b.rho);
    AUGUR_LAUNCHED("dep_finish");
  }
  return AUGUR_OK;
}

// FFT Poisson + spectral gradient: b.rho -> force planes b.force
// nb grids in a batch: rho, the spectra and force are nb consecutive blocks
static int pm_solve(int64_t G, float L, float Gconst, PmBufs& b, hipStream_t s, int nb = 1) {
  int rc;
  {
    TimedScope ts_("fft_rows_r2c", s);
    PM_SWITCH_G(G, k_pm_rows_r2c, dim3((unsigned)(G * G / PmRow<GG>::NL), (unsigned)nb), dim3(FFT_THREADS), s,
                (const float*)b.rho, b.s0);
    AUGUR_LAUNCHED("fft_rows_r2c");
  }
  rc = fft_cols_y(b.s0, G, pm_pitch(G), s, nb);  // y columns forward
  if (rc) return rc;
  const double PI = 3.14159265358979323846;
  const float kf = (float)(2.0 * PI / (double)L);
  const float green = (float)(-4.0 * PI * (double)Gconst / ((double)G * G * G));
  {
    TimedScope ts_("pm_zpass", s);
    PM_SWITCH_G(G, k_pm_zpass, dim3((unsigned)PmCol<GG>::blocks(PmCol<GG>::NQZ), (unsigned)nb), dim3(PmCol<GG>::T), s, b.s0, b.s2,
                kf, green);
    AUGUR_LAUNCHED("pm_zpass");
  }
  {
    TimedScope ts_("pm_yinv", s);
    PM_SWITCH_G(G, k_pm_yinv, dim3((unsigned)PmCol<GG>::blocks(PmCol<GG>::NQY), (unsigned)nb), dim3(PmCol<GG>::T), s, b.s0, b.s1,
                b.s2, kf);
    AUGUR_LAUNCHED("pm_yinv");
  }
  {
    TimedScope ts_("pm_xc2r", s);
    PM_SWITCH_G(G, k_pm_xc2r, dim3((unsigned)(G * G / PmRow<GG>::NL), (unsigned)nb), dim3(FFT_THREADS), s,
                (const float2*)b.s0, (const float2*)b.s1, (const float2*)b.s2, kf, b.force, (const uint32_t*)nullptr);
    AUGUR_LAUNCHED("pm_xc2r");
  }
  return AUGUR_OK;
}

// deposit -> FFT Poisson + spectral gradient -> force planes
static int pm_force_grid(const float* pos, const uint32_t* perm, const float* mass, float mass_value, int64_t n,
                         int64_t G, float L, float Gconst, PmBufs& b, hipStream_t s, bool bound_ready = false) {
  const int rc = pm_deposit(pos, perm, mass, mass_value, n, G, L, b, s, bound_ready);
  if (rc) return rc;
  return pm_solve(G, L, Gconst, b, s);
}

// ------------------------------------------------- surrogate residual gate
// Per-step check of a learned force field against plain PM at probe
// particles, decided on the device (no host synchronisation, capturable).
// gate[0]  fallback flag, sticky within a call: the correction conv skips while set
// gate[1]  revert flag of the current step: the force planes are re-made from
//          the PM spectra still in the workspace (augur_pm_force_replay)
// gate[2]  step counter of the call; gate[3] first step that fell back (~0: none)
constexpr int GATE_THREADS = 256;

template <bool P2>
__device__ __forceinline__ void gather_point(const float* __restrict__ force, int G, float h, float inv_h,
                                             const float* p, float a[3]) {
  const int64_t G3 = (int64_t)G * G * G;
  int c[3];
  float w[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float uu = P2 ? cic_u_pow2(p[d], inv_h) : cic_u(p[d], 0.0f, h);
    const float f = floorf(uu);
    c[d] = (int)f;
    w[d] = __fsub_rn(uu, f);
  }
  a[0] = a[1] = a[2] = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
    const float wt = (ox ? w[0] : 1.0f - w[0]) * (oy ? w[1] : 1.0f - w[1]) * (oz ? w[2] : 1.0f - w[2]);
    const int64_t cell = ((int64_t)wrap(c[2] + oz, G) * G + wrap(c[1] + oy, G)) * G + wrap(c[0] + ox, G);
#pragma unroll
    for (int d = 0; d < 3; ++d) a[d] = fmaf(wt, force[d * G3 + cell], a[d]);
  }
}

// phases 1 and 2: one probe per thread over ceil(P / GATE_THREADS) workgroups;
// phase 2 leaves each workgroup's FP64 sums in part[2 blk], part[2 blk + 1]
template <bool P2>
__global__ __launch_bounds__(GATE_THREADS) void k_sur_probe(int phase, const float* __restrict__ force,
                                                            const float* __restrict__ pos,
                                                            const int64_t* __restrict__ probe, int P, int G, float h,
                                                            float inv_h, float* __restrict__ ref,
                                                            double* __restrict__ part, uint32_t* __restrict__ gate) {
  const int tid = threadIdx.x;
  const int i = blockIdx.x * GATE_THREADS + tid;
  double num = 0.0, den = 0.0;
  if (i < P) {
    const int64_t k = probe[i];
    const float p[3] = {pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]};
    float a[3];
    gather_point<P2>(force, G, h, inv_h, p, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (phase == 1) {
        ref[3 * i + d] = a[d];
      } else {
        const double r = (double)ref[3 * i + d], e = (double)a[d] - r;
        num += e * e;
        den += r * r;
      }
    }
  }
  if (phase == 1) {
    if (i == 0) gate[1] = 0u;
    return;
  }
  __shared__ double sn[GATE_THREADS / 64], sd[GATE_THREADS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    num += __shfl_down(num, o);
    den += __shfl_down(den, o);
  }
  if ((tid & 63) == 0) {
    sn[tid >> 6] = num;
    sd[tid >> 6] = den;
  }
  __syncthreads();
  if (tid == 0) {
    double tn = 0.0, td = 0.0;
    for (int w = 0; w < GATE_THREADS / 64; ++w) {
      tn += sn[w];
      td += sd[w];
    }
    part[2 * blockIdx.x] = tn;
    part[2 * blockIdx.x + 1] = td;
  }
}

// phase 0 resets the gate; phase 3 sums the partials in workgroup order and decides
__global__ void k_sur_decide(int phase, int nblk, const double* __restrict__ part, uint32_t* __restrict__ gate,
                             const float* __restrict__ thr, float* __restrict__ log, int nlog) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (phase == 0) {
    gate[0] = 0u;
    gate[1] = 0u;
    gate[2] = 0u;
    gate[3] = 0xffffffffu;
    return;
  }
  double tn = 0.0, td = 0.0;
  for (int b = 0; b < nblk; ++b) {
    tn += part[2 * b];
    td += part[2 * b + 1];
  }
  const uint32_t st = gate[2];
  const double res = td > 0.0 ? sqrt(tn / td) : (tn > 0.0 ? INFINITY : 0.0);
  if ((int)st < nlog) log[st] = (float)res;
  const float t = thr[(int)st < nlog ? st : nlog - 1];
  if (gate[0] == 0u && res > (double)t) {
    gate[0] = 1u;
    gate[1] = 1u;
    gate[3] = st;
  }
  gate[2] = st + 1u;
}

}  // namespace augur

// gather + kicks (+ drift unless last) of one step; with diag, the FP64
// sums [m vx, m vy, m vz, m v^2 / 2] at the step's synchronised velocity go
// to diag[0..3] (mass NULL: mass_value each)
static int pm_step_end(float* pos, float* vel, float* acc, const float* mass, float mass_value, int64_t n, int64_t G,
                       float L, float dt, int last, double* diag, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && vel && acc, AUGUR_EINVAL, "pm_step_end: bad arguments");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_step_end: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  const float h = (float)((double)L / (double)G);
  const float hdt = 0.5f * dt;
  const int64_t npc = pm_pieces(G);
  const int64_t ng = pm_max_gitems(n, G);
  double* dpart = diag ? b.dpart : nullptr;
  {
    TimedScope ts_("gather_kdk", s);
    const dim3 gg((unsigned)ng);
    const bool p2 = pow2_float(h);
    const float inv_h = p2 ? 1.0f / h : 0.0f;
#define AUGUR_GATHER(P2, DRIFT, DIAG)                                                                             \
  hipLaunchKernelGGL((k_gather_tiles<P2, DRIFT, DIAG>), gg, dim3(TG_THREADS), 0, s, pos, vel, acc,                  \
                     (const float*)b.force, n, (int)G, h, inv_h, hdt, dt, L, (const uint32_t*)b.gitems,            \
                     (const DepItem*)b.ginfo, (int)npc, b.st, mass, mass_value, dpart)
    const int variant = (p2 ? 4 : 0) | (last ? 0 : 2) | (dpart ? 1 : 0);
    switch (variant) {
      case 0: AUGUR_GATHER(false, false, false); break;
      case 1: AUGUR_GATHER(false, false, true); break;
      case 2: AUGUR_GATHER(false, true, false); break;
      case 3: AUGUR_GATHER(false, true, true); break;
      case 4: AUGUR_GATHER(true, false, false); break;
      case 5: AUGUR_GATHER(true, false, true); break;
      case 6: AUGUR_GATHER(true, true, false); break;
      default: AUGUR_GATHER(true, true, true); break;
    }
#undef AUGUR_GATHER
    AUGUR_LAUNCHED("gather_kdk");
  }
  if (diag) {
    TimedScope ts_("diag_finish", s);
    hipLaunchKernelGGL(k_diag_finish, dim3(1), dim3(DIAG_THREADS), 0, s, (const double*)dpart, ng, diag);
    AUGUR_LAUNCHED("diag_finish");
  }
  return AUGUR_OK;
}

extern "C" {

size_t augur_pm_workspace_bytes(int64_t n, int64_t G) {
  Carver c(nullptr, 0);
  PmBufs b;
  carve_pm(c, n, G, b);
  return c.used + 4096;
}

int augur_pm_prepare(float* pos, float* vel, float* acc, float* mass, int32_t* ids, int64_t n, int64_t G, float L,
                     void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos, AUGUR_EINVAL, "pm_prepare: null positions");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_prepare: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  rc = pm_init(b, G, s);
  if (rc) return rc;
  return pm_sort(pos, vel, acc, mass, ids, n, G, L, b, nullptr, true, s);
}

int augur_pm_accel(const float* pos, float* acc, const float* mass, float mass_value, int64_t n, int64_t G, float L,
                   float Gconst, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && acc, AUGUR_EINVAL, "pm_accel: null pointer");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_accel: workspace too small (%zu < %zu)", ws_bytes, c.used);
  // any input order: sort an index permutation (D = 0), deposit through it
  rc = pm_init(b, G, s);
  if (rc) return rc;
  rc = pm_sort((float*)pos, nullptr, nullptr, nullptr, nullptr, n, G, L, b, nullptr, false, s);
  if (rc) return rc;
  rc = pm_force_grid(pos, b.perm, mass, mass_value, n, G, L, Gconst, b, s);
  if (rc) return rc;
  const float h = (float)((double)L / (double)G);
  TimedScope ts_("gather", s);
  if (pow2_float(h))
    hipLaunchKernelGGL((k_gather_kdk<true, PM_GU>), dim3(blocks_for(n, 256, 2048)), dim3(256), 0, s, (float*)pos,
                       (float*)nullptr, acc, (const float*)b.force, (const uint32_t*)b.perm, n, (int)G, h, 1.0f / h,
                       0.0f, 0.0f, L, 0, 0, (PmDev*)nullptr);
  else
    hipLaunchKernelGGL((k_gather_kdk<false, PM_GU>), dim3(blocks_for(n, 256, 2048)), dim3(256), 0, s, (float*)pos,
                       (float*)nullptr, acc, (const float*)b.force, (const uint32_t*)b.perm, n, (int)G, h, 0.0f,
                       0.0f, 0.0f, L, 0, 0, (PmDev*)nullptr);
  AUGUR_LAUNCHED("gather");
  return AUGUR_OK;
}

int augur_pm_fields(const float* pos, const float* mass, float mass_value, int64_t n, int64_t G, float L, float Gconst,
                    float* rho, float* force, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && rho && force, AUGUR_EINVAL, "pm_fields: null pointer");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_fields: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  rc = pm_init(b, G, s);
  if (rc) return rc;
  rc = pm_sort((float*)pos, nullptr, nullptr, nullptr, nullptr, n, G, L, b, nullptr, false, s);
  if (rc) return rc;
  b.rho = rho;      // the caller's grids receive the density and the force planes
  b.force = force;
  return pm_force_grid(pos, b.perm, mass, mass_value, n, G, L, Gconst, b, s);
}

int augur_pm_density(const float* pos, const float* mass, float mass_value, int64_t n, int64_t G, float L, float* rho,
                     void* ws, size_t ws_bytes, void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && rho, AUGUR_EINVAL, "pm_density: null pointer");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_density: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  rc = pm_init(b, G, s);
  if (rc) return rc;
  rc = pm_sort((float*)pos, nullptr, nullptr, nullptr, nullptr, n, G, L, b, nullptr, false, s);
  if (rc) return rc;
  b.rho = rho;
  return pm_deposit(pos, b.perm, mass, mass_value, n, G, L, b, s);
}

int augur_pm_step_begin(float* pos, float* vel, float* acc, float* mass, int32_t* ids, float mass_value, int64_t n,
                        int64_t G, float L, float Gconst, float dt, int first, void* ws, size_t ws_bytes,
                        void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && vel && acc, AUGUR_EINVAL, "pm_step_begin: bad arguments");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_step_begin: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  const float h = (float)((double)L / (double)G);
  if (first) {  // opening half kick and drift from the stored acceleration
    TimedScope ts_("kick_drift", s);
    hipLaunchKernelGGL(k_kick_drift, dim3(blocks_for(n, 256, 2048)), dim3(256), 0, s, pos, vel, acc, n, 0.5f * dt, dt,
                       L, b.st);
    AUGUR_LAUNCHED("kick_drift");
  }
  hipLaunchKernelGGL(k_pm_decide, dim3(1), dim3(64), 0, s, b.st, dt / h, 0);
  AUGUR_LAUNCHED("pm_decide");
  // the state arrays (ids included) are permuted in place when the gate
  // opens; acc is not (the last step rewrites all of it in the final order)
  rc = pm_sort(pos, vel, nullptr, mass, ids, n, G, L, b, &b.st->need_sort, true, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pm_sorted, dim3(1), dim3(64), 0, s, b.st, mass ? 0 : 1, mass_value, pm_dep_scale(G, L));
  AUGUR_LAUNCHED("pm_sorted");
  return pm_force_grid(pos, nullptr, mass, mass_value, n, G, L, Gconst, b, s, mass == nullptr);
}

int augur_pm_step_end(float* pos, float* vel, float* acc, int64_t n, int64_t G, float L, float dt, int last, void* ws,
                      size_t ws_bytes, void* stream) {
  return pm_step_end(pos, vel, acc, nullptr, 0.0f, n, G, L, dt, last, nullptr, ws, ws_bytes, stream);
}

int augur_pm_grid_offsets(int64_t n, int64_t G, size_t* rho_offset, size_t* force_offset) {
  AUGUR_REQUIRE(rho_offset && force_offset, AUGUR_EINVAL, "pm_grid_offsets: null pointer");
  // carve against a stand-in base address (never dereferenced) to read offsets
  const uintptr_t base = 4096;
  Carver c(reinterpret_cast<void*>(base), SIZE_MAX);
  PmBufs b;
  carve_pm(c, n, G, b);
  *rho_offset = (size_t)(reinterpret_cast<uintptr_t>(b.rho) - base);
  *force_offset = (size_t)(reinterpret_cast<uintptr_t>(b.force) - base);
  return AUGUR_OK;
}

int augur_pm_kdk_diag(float* pos, float* vel, float* acc, float* mass, int32_t* ids, float mass_value, int64_t n,
                      int64_t G, float L, float Gconst, float dt, int nsteps, double* diag, void* ws, size_t ws_bytes,
                      void* stream) {
  AUGUR_REQUIRE(nsteps >= 0 && diag, AUGUR_EINVAL, "pm_kdk_diag: bad arguments");
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && vel && acc, AUGUR_EINVAL, "pm_kdk_diag: bad arguments");
  for (int st = 0; st < nsteps; ++st) {
    rc = augur_pm_step_begin(pos, vel, acc, mass, ids, mass_value, n, G, L, Gconst, dt, st == 0, ws, ws_bytes, stream);
    if (rc) return rc;
    rc = pm_step_end(pos, vel, acc, mass, mass_value, n, G, L, dt, st == nsteps - 1, diag + 4 * (int64_t)st, ws,
                     ws_bytes, stream);
    if (rc) return rc;
  }
  return AUGUR_OK;
}

int augur_pm_solve_grid(const float* rho, float* force, int64_t G, float L, float Gconst, void* ws, size_t ws_bytes,
                        void* stream) {
  int rc = check_pm(1, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(rho && force, AUGUR_EINVAL, "pm_solve_grid: null pointer");
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, 1, G, b), AUGUR_EWORKSPACE, "pm_solve_grid: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  b.rho = const_cast<float*>(rho);   // read only by the solve
  b.force = force;
  return pm_solve(G, L, Gconst, b, as_stream(stream));
}

size_t augur_pm_solve_grids_workspace_bytes(int64_t B, int64_t G) {
  if (B < 1 || G < 2) return 0;
  return 3 * align_up((size_t)(B * G * G * pm_pitch(G)) * sizeof(float2)) + 256;
}

int augur_pm_solve_grids(const float* rho, float* force, int64_t B, int64_t G, float L, float Gconst, void* ws,
                         size_t ws_bytes, void* stream) {
  int rc = check_pm(1, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(B >= 1 && B <= 65535, AUGUR_EINVAL, "pm_solve_grids: B must be in [1, 65535]");
  AUGUR_REQUIRE(rho && force, AUGUR_EINVAL, "pm_solve_grids: null pointer");
  Carver c(ws, ws_bytes);
  PmBufs b{};
  const int64_t nc = B * G * G * pm_pitch(G);
  b.s0 = c.take<float2>(nc);
  b.s1 = c.take<float2>(nc);
  b.s2 = c.take<float2>(nc);
  AUGUR_REQUIRE(c.ok(), AUGUR_EWORKSPACE, "pm_solve_grids: workspace too small (%zu < %zu)", ws_bytes, c.used);
  // the grids in groups whose three spectra (at most 128 MB, or one grid)
  // stay in the 256 MB Infinity Cache between the five passes, every group
  // on the first group's spectra; each grid's arithmetic is unchanged (at 64
  // x 128^3: 2.66 ms in one group, 1.98 ms in groups of 4, the same bits)
  const int64_t G3 = G * G * G;
  const int64_t per_grid = 3 * G * G * pm_pitch(G) * (int64_t)sizeof(float2);
  const int64_t grp = std::max<int64_t>(1, std::min<int64_t>(B, ((int64_t)128 << 20) / per_grid));
  for (int64_t g0 = 0; g0 < B; g0 += grp) {
    PmBufs bg = b;
    bg.rho = const_cast<float*>(rho) + g0 * G3;  // read only by the solve
    bg.force = force + g0 * 3 * G3;
    rc = pm_solve(G, L, Gconst, bg, as_stream(stream), (int)std::min(grp, B - g0));
    if (rc) return rc;
  }
  return AUGUR_OK;
}

// The workspace of the ragged tiled deposit, carved in this order.
struct RaggedDepBufs {
  PmDev* st;
  uint32_t *keys, *vals, *cell_start, *gaps, *ngaps, *items, *tseg_lo, *tseg_pre;
  char *rws, *scan_ws;
  DepItem* info;
  uint16_t* iblk;
  unsigned long long* fix;
};
static bool carve_ragged_dep(Carver& c, int64_t B, int64_t n, int64_t G, RaggedDepBufs& b) {
  const int64_t ncells = B * G * G * G, ntiles = B * pm_tiles(G);
  const int64_t max_items = ntiles + (8 * n + TD_CHUNK - 1) / TD_CHUNK;
  b.st = c.take<PmDev>(1);
  b.keys = c.take<uint32_t>(n);
  b.vals = c.take<uint32_t>(n);
  b.rws = c.take<char>(radix_workspace_bytes(n));
  b.cell_start = c.take<uint32_t>(ncells + 1);
  b.gaps = c.take<uint32_t>(3 * ((ncells + 1) / CS_SPAN + 2));
  b.ngaps = c.take<uint32_t>(1);
  b.items = c.take<uint32_t>(ntiles + 1);
  b.scan_ws = c.take<char>(scan_workspace_bytes(ntiles + 1));
  b.tseg_lo = c.take<uint32_t>(ntiles * TD_MAXSEG);
  b.tseg_pre = c.take<uint32_t>(ntiles * (TD_MAXSEG + 1));
  b.info = c.take<DepItem>(max_items);
  b.iblk = c.take<uint16_t>(max_items * TD_NBLK);
  b.fix = c.take<unsigned long long>(ncells);
  return c.ok();
}

size_t augur_cic_deposit_ragged_tiled_workspace_bytes(int64_t B, int64_t G, int64_t total) {
  if (B < 1 || G < 32 || total < 0) return 0;
  Carver c(nullptr, 0);
  RaggedDepBufs b;
  carve_ragged_dep(c, B, total > 0 ? total : 1, G, b);
  return c.used + 4096;
}

}  // extern "C"

// The ragged tiled deposit after validation and carving: the sort and plans
// (gated by `gate`, NULL = always), the deposit bound and the tiled deposit
// of every scene's tiles into out (B grids).
static int ragged_deposit(const float* pos, const float* mass, float mass_value, const int64_t* starts, int64_t B,
                          int64_t n, int64_t G, float h, float scale, float* out, RaggedDepBufs& b,
                          const uint32_t* gate, hipStream_t s) {
  const int64_t G3 = G * G * G, tps = pm_tiles(G), ntiles = B * tps;
  const dim3 g(blocks_for(n, 256, 1024)), bl(256);
  {
    TimedScope ts_("ragged_sort", s);
    hipLaunchKernelGGL(k_ragged_keys, g, bl, 0, s, pos, starts, (int)B, n, (int)G, h, b.keys, b.vals, b.ngaps, gate);
    AUGUR_LAUNCHED("ragged_keys");
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < B * G3) ++bits;
    uint32_t *ks, *perm;
    int rc = radix_sort_pairs(b.keys, b.vals, n, bits, b.rws, radix_workspace_bytes(n), &ks, &perm, s, gate);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cell_starts, dim3(blocks_for(n + 1, 256, 2048)), bl, 0, s, (const uint32_t*)ks, n,
                       (uint32_t)(B * G3), b.cell_start, b.gaps, b.ngaps, gate);
    AUGUR_LAUNCHED("cell_starts");
    hipLaunchKernelGGL(k_cell_fill, dim3(1024), bl, 0, s, (const uint32_t*)b.gaps, (const uint32_t*)b.ngaps,
                       b.cell_start, gate);
    AUGUR_LAUNCHED("cell_fill");
    hipLaunchKernelGGL(k_dep_plans, dim3((unsigned)((ntiles + 1 + 3) / 4)), dim3(256), 0, s,
                       (const uint32_t*)b.cell_start, (int)G, (int)ntiles, (int)tps, b.items, b.tseg_lo, b.tseg_pre,
                       gate);
    AUGUR_LAUNCHED("dep_plans");
    rc = exclusive_scan_u32(b.items, b.items, ntiles + 1, b.scan_ws, s, gate);
    if (rc) return rc;
    const int64_t max_items = ntiles + (8 * n + TD_CHUNK - 1) / TD_CHUNK;
    hipLaunchKernelGGL(k_dep_items, dim3((unsigned)max_items), dim3(256), 0, s, (const uint32_t*)b.items, (int)ntiles,
                       (const uint32_t*)b.tseg_pre, b.info, b.iblk, gate);
    AUGUR_LAUNCHED("dep_items");
    hipLaunchKernelGGL(k_dep_zero_heavy, dim3((unsigned)ntiles), dim3(256), 0, s, (const uint32_t*)b.items, b.fix,
                       gate);
    AUGUR_LAUNCHED("dep_zero_heavy");
    b.vals = perm;  // sorted slot -> particle (the same buffer on every call: the pass count is fixed)
  }
  {
    TimedScope ts_("dep_bound", s);
    if (mass) {
      hipLaunchKernelGGL(k_pm_set_maxbits, dim3(1), dim3(64), 0, s, b.st, (const float*)nullptr, 0.0f, scale, n);
      AUGUR_LAUNCHED("dep_bound_reset");
      hipLaunchKernelGGL(k_pm_set_maxbits, dim3(blocks_for(n, 256, 2048)), dim3(256), 0, s, b.st, mass, mass_value,
                         scale, n);
    } else
      hipLaunchKernelGGL(k_pm_set_maxbits, dim3(1), dim3(64), 0, s, b.st, (const float*)nullptr, mass_value, scale, n);
    AUGUR_LAUNCHED("dep_bound");
  }
  {
    TimedScope ts_("dep_tiles", s);
    const int64_t grid = ntiles + (8 * n + TD_CHUNK - 1) / TD_CHUNK;  // a workgroup for every possible item
    const bool p2 = pow2_float(h);
    const float inv_h = p2 ? 1.0f / h : 0.0f;
    if (p2)
      hipLaunchKernelGGL(k_dep_tiles<true>, dim3((unsigned)grid), dim3(TD_THREADS), 0, s, pos,
                         (const uint32_t*)b.vals, mass, mass_value, n, (int)G, h, inv_h, scale, (const PmDev*)b.st,
                         (const uint32_t*)b.items, (const DepItem*)b.info, (const uint16_t*)b.iblk,
                         (const uint32_t*)b.tseg_lo, (const uint32_t*)b.tseg_pre, (int)ntiles, (int)tps, b.fix, out);
    else
      hipLaunchKernelGGL(k_dep_tiles<false>, dim3((unsigned)grid), dim3(TD_THREADS), 0, s, pos,
                         (const uint32_t*)b.vals, mass, mass_value, n, (int)G, h, inv_h, scale, (const PmDev*)b.st,
                         (const uint32_t*)b.items, (const DepItem*)b.info, (const uint16_t*)b.iblk,
                         (const uint32_t*)b.tseg_lo, (const uint32_t*)b.tseg_pre, (int)ntiles, (int)tps, b.fix, out);
    AUGUR_LAUNCHED("dep_tiles");
    hipLaunchKernelGGL(k_dep_finish, dim3((unsigned)ntiles), dim3(256), 0, s, (const uint32_t*)b.items, b.fix,
                       (const PmDev*)b.st, n, (int)G, (int)tps, out);
    AUGUR_LAUNCHED("dep_finish");
  }
  return AUGUR_OK;
}

static int check_ragged_dep(int64_t B, int64_t total, int64_t G, float h, const int64_t* starts, const float* out,
                            const char* who) {
  AUGUR_REQUIRE(is_pow2(G) && G >= 32 && G <= 512, AUGUR_EUNSUPPORTED, "%s: G must be a power of two in [32, 512]",
                who);
  AUGUR_REQUIRE(B >= 1 && B <= 65535 && total >= 0 && total < ((int64_t)1 << 31) && h > 0.0f && starts && out,
                AUGUR_EINVAL, "%s: bad arguments", who);
  AUGUR_REQUIRE(B * G * G * G < ((int64_t)1 << 32), AUGUR_EUNSUPPORTED, "%s: B G^3 must be below 2^32", who);
  return AUGUR_OK;
}

extern "C" {

int augur_cic_deposit_ragged_tiled(const float* pos, const float* mass, float mass_value, const int64_t* starts,
                                   int64_t B, int64_t total, int64_t G, float h, float scale, float* out, void* ws,
                                   size_t ws_bytes, void* stream) {
  int rc = check_ragged_dep(B, total, G, h, starts, out, "cic_deposit_ragged_tiled");
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  if (total == 0) return fill_u32(out, B * G * G * G, 0u, s);
  AUGUR_REQUIRE(pos, AUGUR_EINVAL, "cic_deposit_ragged_tiled: null positions");
  Carver c(ws, ws_bytes);
  RaggedDepBufs b;
  AUGUR_REQUIRE(carve_ragged_dep(c, B, total, G, b), AUGUR_EWORKSPACE,
                "cic_deposit_ragged_tiled: workspace too small (%zu < %zu)", ws_bytes, c.used);
  return ragged_deposit(pos, mass, mass_value, starts, B, total, G, h, scale, out, b, nullptr, s);
}

size_t augur_cic_deposit_ragged_planned_workspace_bytes(int64_t B, int64_t G, int64_t total) {
  return augur_cic_deposit_ragged_tiled_workspace_bytes(B, G, total);
}

int augur_cic_deposit_ragged_planned(const float* pos, const float* mass, float mass_value, const int64_t* starts,
                                     int64_t B, int64_t total, int64_t G, float h, float scale, float drift_dt,
                                     int force_sort, float* out, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_ragged_dep(B, total, G, h, starts, out, "cic_deposit_ragged_planned");
  if (rc) return rc;
  AUGUR_REQUIRE(drift_dt >= 0.0f, AUGUR_EINVAL, "cic_deposit_ragged_planned: drift_dt must be >= 0");
  hipStream_t s = as_stream(stream);
  if (total == 0) return fill_u32(out, B * G * G * G, 0u, s);
  AUGUR_REQUIRE(pos, AUGUR_EINVAL, "cic_deposit_ragged_planned: null positions");
  Carver c(ws, ws_bytes);
  RaggedDepBufs b;
  AUGUR_REQUIRE(carve_ragged_dep(c, B, total, G, b), AUGUR_EWORKSPACE,
                "cic_deposit_ragged_planned: workspace too small (%zu < %zu)", ws_bytes, c.used);
  if (force_sort) {
    hipLaunchKernelGGL(k_pm_init, dim3(1), dim3(64), 0, s, b.st);
    AUGUR_LAUNCHED("pm_init");
  }
  // the drift since the last sort (max |v| of the tracked drifts, dt each) decides the re-sort
  hipLaunchKernelGGL(k_pm_decide, dim3(1), dim3(64), 0, s, b.st, drift_dt / h, force_sort);
  AUGUR_LAUNCHED("pm_decide");
  rc = ragged_deposit(pos, mass, mass_value, starts, B, total, G, h, scale, out, b, &b.st->need_sort, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pm_sorted, dim3(1), dim3(64), 0, s, b.st, 0, 0.0f, 0.0f);
  AUGUR_LAUNCHED("pm_sorted");
  return AUGUR_OK;
}

uint32_t* augur_ragged_plan_drift_word(void* ws, size_t ws_bytes) {
  Carver c(ws, ws_bytes);
  PmDev* st = c.take<PmDev>(1);
  return c.ok() ? &st->vmax_bits : nullptr;
}

int augur_pm_force_replay(int64_t n, int64_t G, float L, const uint32_t* gate, void* ws, size_t ws_bytes,
                          void* stream) {
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(gate, AUGUR_EINVAL, "pm_force_replay: null gate");
  hipStream_t s = as_stream(stream);
  Carver c(ws, ws_bytes);
  PmBufs b;
  AUGUR_REQUIRE(carve_pm(c, n, G, b), AUGUR_EWORKSPACE, "pm_force_replay: workspace too small (%zu < %zu)", ws_bytes,
                c.used);
  const double PI = 3.14159265358979323846;
  const float kf = (float)(2.0 * PI / (double)L);
  TimedScope ts_("pm_force_replay", s);
  PM_SWITCH_G(G, k_pm_xc2r, dim3((unsigned)(G * G / PmRow<GG>::NL)), dim3(FFT_THREADS), s, (const float2*)b.s0,
              (const float2*)b.s1, (const float2*)b.s2, kf, b.force, gate);
  AUGUR_LAUNCHED("pm_force_replay");
  return AUGUR_OK;
}

size_t augur_sur_gate_scratch_bytes(int64_t P) {
  const int64_t nblk = (P + GATE_THREADS - 1) / GATE_THREADS;
  return align_up((size_t)(3 * P) * sizeof(float), 16) + (size_t)(2 * (nblk > 0 ? nblk : 1)) * sizeof(double);
}

int augur_sur_gate(int phase, const float* force, const float* pos, const int64_t* probe, int64_t P, int64_t G,
                   float L, void* scratch, uint32_t* gate, const float* thr, float* log, int nlog, void* stream) {
  AUGUR_REQUIRE(gate && phase >= 0 && phase <= 2, AUGUR_EINVAL, "sur_gate: bad phase or null gate");
  AUGUR_REQUIRE(is_pow2(G) && G >= 2 && G <= 4096 && L > 0.0f, AUGUR_EINVAL, "sur_gate: bad grid");
  hipStream_t s = as_stream(stream);
  TimedScope ts_("sur_gate", s);
  if (phase == 0) {
    hipLaunchKernelGGL(k_sur_decide, dim3(1), dim3(64), 0, s, 0, 0, (const double*)nullptr, gate,
                       (const float*)nullptr, (float*)nullptr, 0);
    AUGUR_LAUNCHED("sur_gate");
    return AUGUR_OK;
  }
  AUGUR_REQUIRE(force && pos && probe && scratch && P >= 1 && P < ((int64_t)1 << 31), AUGUR_EINVAL,
                "sur_gate: bad probe arguments");
  if (phase == 2) AUGUR_REQUIRE(thr && log && nlog >= 1, AUGUR_EINVAL, "sur_gate: phase 2 needs thr, log");
  const float h = (float)((double)L / (double)G);
  const int nblk = (int)((P + GATE_THREADS - 1) / GATE_THREADS);
  float* ref = static_cast<float*>(scratch);
  double* part = reinterpret_cast<double*>(static_cast<char*>(scratch) + align_up((size_t)(3 * P) * sizeof(float), 16));
  if (pow2_float(h))
    hipLaunchKernelGGL(k_sur_probe<true>, dim3(nblk), dim3(GATE_THREADS), 0, s, phase, force, pos, probe, (int)P,
                       (int)G, h, 1.0f / h, ref, part, gate);
  else
    hipLaunchKernelGGL(k_sur_probe<false>, dim3(nblk), dim3(GATE_THREADS), 0, s, phase, force, pos, probe, (int)P,
                       (int)G, h, 0.0f, ref, part, gate);
  AUGUR_LAUNCHED("sur_gate");
  if (phase == 2) {
    hipLaunchKernelGGL(k_sur_decide, dim3(1), dim3(64), 0, s, 3, nblk, (const double*)part, gate, thr, log, nlog);
    AUGUR_LAUNCHED("sur_gate_decide");
  }
  return AUGUR_OK;
}

int augur_pm_kdk(float* pos, float* vel, float* acc, float* mass, int32_t* ids, float mass_value, int64_t n, int64_t G,
                 float L, float Gconst, float dt, int nsteps, void* ws, size_t ws_bytes, void* stream) {
  AUGUR_REQUIRE(nsteps >= 0, AUGUR_EINVAL, "pm_kdk: bad arguments");
  int rc = check_pm(n, G, L);
  if (rc) return rc;
  AUGUR_REQUIRE(pos && vel && acc, AUGUR_EINVAL, "pm_kdk: bad arguments");
  for (int st = 0; st < nsteps; ++st) {
    rc = augur_pm_step_begin(pos, vel, acc, mass, ids, mass_value, n, G, L, Gconst, dt, st == 0, ws, ws_bytes, stream);
    if (rc) return rc;
    rc = augur_pm_step_end(pos, vel, acc, n, G, L, dt, st == nsteps - 1, ws, ws_bytes, stream);
    if (rc) return rc;
  }
  return AUGUR_OK;
}

}  // extern "C"
