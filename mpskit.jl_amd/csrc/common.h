// Shared helpers for the MI355X (gfx950) contraction library: status codes,
// thread-local error text, HIP error checks.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mpskit_hip.h"

namespace mk {

void set_error(const char* fmt, ...);
const char* last_error();

#define MK_CHECK_HIP(expr)                                                       \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      ::mk::set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr,              \
                      hipGetErrorString(_e));                                    \
      return MK_HIP_ERROR;                                                       \
    }                                                                            \
  } while (0)

#define MK_REQUIRE(cond, code, ...)                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      ::mk::set_error(__VA_ARGS__);                                              \
      return (code);                                                             \
    }                                                                            \
  } while (0)

#define MK_TRY(expr)                                                             \
  do {                                                                           \
    mk_status _s = (expr);                                                       \
    if (_s != MK_OK) return _s;                                                  \
  } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace mk
