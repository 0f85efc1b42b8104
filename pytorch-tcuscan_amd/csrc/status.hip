// Status strings and ABI version of libtcuscan_hip.so.
#include "tcs_common.hpp"

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

int tcs_abi_version(void) { return 101; }

}  // extern "C"
