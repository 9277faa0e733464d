// Shared helpers for the petgem_amd library: error reporting, HIP checks,
// complex arithmetic on double2/float2, wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/petgem_amd.h"

namespace pg {

void set_error(const char* fmt, ...);

#define PG_FAIL(code, ...)        \
  do {                            \
    ::pg::set_error(__VA_ARGS__); \
    return (code);                \
  } while (0)

#define PG_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess) {                                                                \
      ::pg::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,                   \
                      hipGetErrorString(_e));                                              \
      return PG_ERR_HIP;                                                                   \
    }                                                                                      \
  } while (0)

#define PG_CHECK_LAUNCH() PG_HIP(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---- complex helpers -------------------------------------------------------
__host__ __device__ inline double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 zadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__host__ __device__ inline double2 zsub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__host__ __device__ inline double2 zdiv(double2 a, double2 b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.y + b.x * r;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int kWave = 64;

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over groups of G consecutive lanes (G power of two <= 64); result in all lanes.
template <int G, typename T>
__device__ inline T group_sum(T v) {
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

}  // namespace pg
