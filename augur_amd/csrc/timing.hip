// Optional per-launch timing with HIP events recorded on the launch stream.
// Off by default (no event work at all); when on, every instrumented launch
// records a start/stop pair from a fixed pool, and augur_timing_collect()
// waits on the events and aggregates elapsed time per kernel name.
#include <mutex>
#include <string.h>
#include <string>
#include <vector>

#include "internal.hpp"

namespace augur {

namespace {
constexpr int POOL = 16384;
struct Rec {
  const char* name;
  hipEvent_t a, b;
};
std::mutex mu;
bool enabled = false;
std::vector<hipEvent_t> pool;
std::vector<Rec> recs;
size_t next_event = 0;
std::string only;  // when set, only launches of this name are recorded
}  // namespace

bool timing_on() { return enabled; }

int timing_start(const char* name, hipStream_t s, int* slot) {
  *slot = -1;
  if (!enabled) return AUGUR_OK;
  std::lock_guard<std::mutex> g(mu);
  if (!only.empty() && only != name) return AUGUR_OK;
  if (next_event + 2 > pool.size()) return AUGUR_OK;  // pool exhausted: stop recording
  Rec r{name, pool[next_event], pool[next_event + 1]};
  next_event += 2;
  AUGUR_HIP(hipEventRecord(r.a, s));
  recs.push_back(r);
  *slot = (int)recs.size() - 1;
  return AUGUR_OK;
}

int timing_stop(int slot, hipStream_t s) {
  if (slot < 0) return AUGUR_OK;
  std::lock_guard<std::mutex> g(mu);
  AUGUR_HIP(hipEventRecord(recs[slot].b, s));
  return AUGUR_OK;
}

}  // namespace augur

using namespace augur;

extern "C" {

int augur_timing_enable(int on) {
  std::lock_guard<std::mutex> g(mu);
  if (on && pool.empty()) {
    pool.resize(POOL);
    for (auto& e : pool) AUGUR_HIP(hipEventCreate(&e));
  }
  enabled = on != 0;
  recs.clear();
  next_event = 0;
  return AUGUR_OK;
}

int augur_timing_filter(const char* name) {
  std::lock_guard<std::mutex> g(mu);
  only = name ? name : "";
  return AUGUR_OK;
}

int augur_timing_collect(char* names, size_t names_cap, double* ms, int64_t* counts, int max_entries, int* n_entries) {
  std::lock_guard<std::mutex> g(mu);
  AUGUR_REQUIRE(names && ms && counts && n_entries && max_entries > 0, AUGUR_EINVAL, "timing_collect: bad arguments");
  std::vector<std::string> keys;
  std::vector<double> tot;
  std::vector<int64_t> cnt;
  for (auto& r : recs) {
    AUGUR_HIP(hipEventSynchronize(r.b));
    float t = 0.f;
    AUGUR_HIP(hipEventElapsedTime(&t, r.a, r.b));
    size_t k = 0;
    while (k < keys.size() && keys[k] != r.name) ++k;
    if (k == keys.size()) {
      keys.push_back(r.name);
      tot.push_back(0.0);
      cnt.push_back(0);
    }
    tot[k] += t;
    cnt[k] += 1;
  }
  std::string joined;
  int n = 0;
  for (size_t k = 0; k < keys.size() && n < max_entries; ++k, ++n) {
    joined += keys[k];
    joined += '\n';
    ms[n] = tot[k];
    counts[n] = cnt[k];
  }
  AUGUR_REQUIRE(joined.size() < names_cap, AUGUR_EINVAL, "timing_collect: names buffer too small");
  memcpy(names, joined.c_str(), joined.size() + 1);
  *n_entries = n;
  recs.clear();
  next_event = 0;
  return AUGUR_OK;
}

}  // extern "C"
