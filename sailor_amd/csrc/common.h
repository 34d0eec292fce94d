// Shared host/device helpers for the SAILOR MI355X kernels.
//
// Error policy: every extern "C" entry point returns an int status and records
// a thread-local message retrievable with sailor_last_error(); nothing in this
// library calls exit() (the reference's CUDA_CHECK_ERRORS exits the process,
// c_lib/VoxelEncoding/include/cuda_helper.h:31-40).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/sailor_hip.h"

namespace sailor {

// Records a formatted error message for the calling thread and returns `code`.
int set_error(int code, const char* fmt, ...);

// Returns SAILOR_ERR_HIP (with a message) if the last launch failed.
int check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// One-time launch setup (dynamic LDS limits, resident grid sizes) is kept per
// device in std::atomic slots: several host threads may reach it at once,
// and the setup itself is idempotent.
constexpr int kMaxDevices = 64;
// The calling thread's current device as a slot index in [0, kMaxDevices).
int device_slot(int* slot);

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Float -> int conversion with fully defined results for every input:
// NaN -> 0, values at or beyond the int32 range saturate.  In-range values
// truncate toward zero exactly like a C cast.  The CPU oracle uses the same
// rule so that integer outputs derived from degenerate floats (0/0, x/0) agree.
__host__ __device__ __forceinline__ int sat_f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int)(-2147483647 - 1);
  return (int)x;
}

// Largest k with 2^k <= v (v >= 1); equals (int)log2f(v) for the sizes used.
__host__ __device__ __forceinline__ int floor_log2(int v) {
  int k = 0;
  while ((v >> (k + 1)) > 0) ++k;
  return k;
}

// Order-preserving map float -> int32 so that integer atomicMin/atomicMax
// implement float min/max (non-NaN inputs).
__device__ __forceinline__ int fkey(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float fkey_inv(int k) {
  return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff));
}

// Wave-wide (64 lane) reductions.
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_min_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace sailor

#define SAILOR_HIP_TRY(expr)                                                  \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::sailor::set_error(SAILOR_ERR_HIP, "%s failed: %s", #expr,      \
                                 hipGetErrorString(_e));                      \
  } while (0)

#define SAILOR_REQUIRE(cond, ...)                                             \
  do {                                                                        \
    if (!(cond)) return ::sailor::set_error(SAILOR_ERR_ARG, __VA_ARGS__);     \
  } while (0)
