// Status strings, ABI version and workspace maintenance of libtcuscan_hip.so.
#include "tcs_common.hpp"

namespace tcs {

// Grid-stride fill of 64-bit words (a kernel rather than hipMemsetAsync, so
// that callers may capture it in a HIP graph).
__global__ __launch_bounds__(256) void fill_words_kernel(uint64_t* __restrict__ p, uint64_t words,
                                                         uint64_t value) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = value;
}

int fill_words(void* p, size_t bytes, uint64_t value, hipStream_t st) {
  if (bytes == 0) return TCS_OK;
  if (!p || (reinterpret_cast<uintptr_t>(p) & 7u) != 0 || (bytes & 7u) != 0) return TCS_ERR_INVALID_ARG;
  const uint64_t words = bytes / 8;
  uint64_t blocks = (words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_words_kernel, dim3((uint32_t)blocks), dim3(256), 0, st,
                     static_cast<uint64_t*>(p), words, value);
  TCS_CHECK_LAUNCH();
  return TCS_OK;
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
    case TCS_ERR_SCRATCH_LIMIT: return "input exceeds the fixed scratch of the no-allocation entry point";
    default: return "unknown status";
  }
}

int tcs_abi_version(void) { return 104; }

int tcs_workspace_reset(void* workspace, size_t workspace_bytes, void* stream) {
  if (!workspace || workspace_bytes < tcs::kHeaderBytes) return TCS_ERR_WORKSPACE;
  return tcs::fill_words(workspace, workspace_bytes & ~(size_t)7, 0, static_cast<hipStream_t>(stream));
}

int tcs_workspace_set_poll_limit(void* workspace, size_t workspace_bytes, int polls, void* stream) {
  if (!workspace || workspace_bytes < tcs::kHeaderBytes) return TCS_ERR_WORKSPACE;
  if (polls > tcs::kFallbackPolls) return TCS_ERR_INVALID_ARG;
  const uint64_t word = polls < 0 ? 0 : (uint64_t)polls + 1;
  return tcs::fill_words(static_cast<char*>(workspace) + tcs::kPollWordOffset, 8, word,
                         static_cast<hipStream_t>(stream));
}

}  // extern "C"
