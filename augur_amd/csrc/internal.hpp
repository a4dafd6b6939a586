// Host-side helpers shared between the library's translation units.
#pragma once
#include "common.hpp"

namespace augur {

int fill_u64(void* p, int64_t n, unsigned long long v, hipStream_t s);
int fill_u32(void* p, int64_t n, uint32_t v, hipStream_t s);
size_t scan_workspace_bytes(int64_t n);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* ws, hipStream_t s,
                       const uint32_t* gate = nullptr);
// two exclusive scans (one launch when both have at most 16384 elements)
int exclusive_scan2_u32(const uint32_t* in0, uint32_t* out0, int64_t n0, const uint32_t* in1, uint32_t* out1,
                        int64_t n1, void* ws, hipStream_t s, const uint32_t* gate);
size_t radix_workspace_bytes(int64_t n);
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int bits, void* ws, size_t ws_bytes,
                     uint32_t** keys_out, uint32_t** vals_out, hipStream_t s, const uint32_t* gate = nullptr);

int fft_r2c_xy(const float* in, float2* out, int64_t G, int64_t pitch, hipStream_t s);

// compute units of the current device (cached per device; 256 if the query fails)
int device_cu_count();
int fft_cols_y(float2* data, int64_t G, int64_t pitch, hipStream_t s, int64_t nb = 1);

// Per-launch event timing (see timing.hip).  A TimedScope around a launch
// records a start/stop event pair on the stream when timing is enabled and
// does nothing otherwise.
int timing_start(const char* name, hipStream_t s, int* slot);
int timing_stop(int slot, hipStream_t s);

struct TimedScope {
  int slot;
  hipStream_t s;
  TimedScope(const char* name, hipStream_t st) : slot(-1), s(st) { timing_start(name, s, &slot); }
  ~TimedScope() { timing_stop(slot, s); }
};

}  // namespace augur
