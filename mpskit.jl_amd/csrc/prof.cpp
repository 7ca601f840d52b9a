// Event-based kernel timing: each bracketed launch records a start/stop event
// pair on its own stream; durations are resolved when queried.
#include "prof.h"

#include "ops.h"

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace mk {
namespace {
struct Rec {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_pending;
std::map<std::string, std::pair<double, int64_t>> g_done;

void resolve_locked() {
  for (Rec& r : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& e = g_done[r.name];
      e.first += ms;
      e.second += 1;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_pending.clear();
}
}  // namespace

ProfScope::ProfScope(const char* n, hipStream_t s) : name(n), stream(s) {
  bool on;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    on = g_on;
  }
  // the allocation-free entry point creates no events
  if (!on || fixed_scratch().base) return;
  // launches being captured into a graph are not timed
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&start) != hipSuccess) {
    start = nullptr;
    return;
  }
  (void)hipEventRecord(start, stream);
}

ProfScope::~ProfScope() {
  if (!start) return;
  hipEvent_t stop;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start);
    return;
  }
  (void)hipEventRecord(stop, stream);
  std::lock_guard<std::mutex> lk(g_mu);
  g_pending.push_back(Rec{name, start, stop});
}

}  // namespace mk

using namespace mk;

extern "C" mk_status mk_prof_enable(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = on != 0;
  return MK_OK;
}

extern "C" mk_status mk_prof_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  resolve_locked();
  g_done.clear();
  return MK_OK;
}

extern "C" mk_status mk_prof_query(const char* kernel, double* total_ms, int64_t* count) {
  MK_REQUIRE(kernel && total_ms && count, MK_BAD_ARG, "prof_query: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  resolve_locked();
  auto it = g_done.find(kernel);
  *total_ms = it == g_done.end() ? 0.0 : it->second.first;
  *count = it == g_done.end() ? 0 : it->second.second;
  return MK_OK;
}
