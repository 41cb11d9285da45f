// profile.hip -- event-based kernel timing registry (see profile.h).
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "profile.h"

namespace cubql {

namespace {
struct Stat { uint64_t launches = 0; double ms = 0, bytes = 0; uint64_t units = 0; };
struct Pending { hipEvent_t a = nullptr, b = nullptr; const char *name = nullptr; double bytes = 0; uint64_t units = 0; };
std::mutex g_mu;
std::map<std::string, Stat> g_stats;
bool g_on = false;
std::vector<Pending> t_pending;  // guarded by g_mu
}  // namespace

bool prof_on() { return g_on; }

int prof_begin(hipStream_t s) {
  if (!g_on) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  Pending p;
  if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return -1;
  if (hipEventRecord(p.a, s) != hipSuccess) return -1;
  t_pending.push_back(p);
  return (int)t_pending.size() - 1;
}

void prof_end(int slot, hipStream_t s, const char *name) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (slot < 0 || slot >= (int)t_pending.size()) return;
  t_pending[slot].name = name;
  (void)hipEventRecord(t_pending[slot].b, s);
}

void prof_set(int slot, double bytes, uint64_t units) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (slot < 0 || slot >= (int)t_pending.size()) return;
  t_pending[slot].bytes = bytes;
  t_pending[slot].units = units;
}

static void flush_locked() {
  for (auto &p : t_pending) {
    float ms = 0.f;
    if (p.name && p.units > 0 && hipEventSynchronize(p.b) == hipSuccess &&
        hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      Stat &st = g_stats[p.name];
      st.launches++;
      st.ms += ms;
      st.bytes += p.bytes;
      st.units += p.units;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  t_pending.clear();
}

void prof_flush() {
  std::lock_guard<std::mutex> lk(g_mu);
  flush_locked();
}

}  // namespace cubql

extern "C" void cubql_profile_enable(int on) { cubql::g_on = on != 0; }

extern "C" void cubql_profile_reset(void) {
  std::lock_guard<std::mutex> lk(cubql::g_mu);
  cubql::flush_locked();
  cubql::g_stats.clear();
}

extern "C" int cubql_profile_query(const char *kernel, uint64_t *launches, double *totalMs,
                                   double *bytes, uint64_t *units) {
  std::lock_guard<std::mutex> lk(cubql::g_mu);
  cubql::flush_locked();
  auto it = cubql::g_stats.find(kernel ? kernel : "");
  if (it == cubql::g_stats.end()) return -1;
  if (launches) *launches = it->second.launches;
  if (totalMs) *totalMs = it->second.ms;
  if (bytes) *bytes = it->second.bytes;
  if (units) *units = it->second.units;
  return 0;
}
