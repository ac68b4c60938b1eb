// Host-side internals shared by the library's translation units: error
// reporting (thread-local message + status code) and small helpers.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gasgraph.h"

namespace gg {

// Record a failure for gg_last_error() and return its code.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Deterministic 64-bit mixing (splitmix64 finaliser) and a counter-based
// stream: value i of stream s under seed k.  Used by every generator so the
// same (parameters, seed) gives the same graph on any host and thread count.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t counter_rng(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}

}  // namespace gg

#define GG_CHECK_ARG(cond, ...)                                   \
  do {                                                            \
    if (!(cond)) return ::gg::set_error(GG_ERR_INVALID, __VA_ARGS__); \
  } while (0)
