// Status strings, ABI version and the per-workspace launch counters of
// libtcuscan_hip.so.
#include <mutex>
#include <unordered_map>

#include "tcs_common.hpp"

namespace tcs {

// Launches on one workspace are stream-ordered by contract, so a counter per
// workspace address gives every launch a tag distinct from the previous
// 2^30 - 1 launches on it.
uint32_t next_launch_tag(const void* ws) {
  static std::mutex mu;
  static std::unordered_map<const void*, uint32_t> epochs;
  std::lock_guard<std::mutex> lock(mu);
  return tag_of_epoch(epochs[ws]++);
}

}  // namespace tcs

extern "C" {

const char* tcs_status_string(int status) {
  switch (status) {
    case TCS_OK: return "ok";
    case TCS_ERR_INVALID_ARG: return "invalid argument";
    case TCS_ERR_UNSUPPORTED_DTYPE: return "unsupported dtype";
    case TCS_ERR_WORKSPACE: return "workspace missing or too small";
    case TCS_ERR_LAUNCH: return "kernel launch failed";
    default: return "unknown status";
  }
}

int tcs_abi_version(void) { return 100; }

}  // extern "C"
