// Opt-in per-kernel timing with HIP events on the launch stream.
#pragma once
#include "common.h"

namespace mk {
// Bracket one kernel launch: prof_begin() before, prof_end(name) after.  No-ops
// unless profiling is enabled (mk_prof_enable).
struct ProfScope {
  const char* name;
  hipStream_t stream;
  hipEvent_t start = nullptr;
  ProfScope(const char* n, hipStream_t s);
  ~ProfScope();
};
}  // namespace mk
