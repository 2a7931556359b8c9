// pq_common.h -- shared host/device helpers for the pequod MI355X kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "pequod_hip.h"

// Record a HIP failure and return the library's status code.
#define PQ_HIP(call)                                                           \
  do {                                                                         \
    hipError_t pq_err_ = (call);                                               \
    if (pq_err_ != hipSuccess) {                                               \
      pq_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,          \
                   hipGetErrorString(pq_err_));                                \
      return PQ_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

#define PQ_CHECK_ARG(cond, ...)                                                \
  do {                                                                         \
    if (!(cond)) {                                                             \
      pq_set_error(__VA_ARGS__);                                               \
      return PQ_ERR_ARG;                                                       \
    }                                                                          \
  } while (0)

void pq_set_error(const char *fmt, ...);

// Supported DST lengths: n + 1 = N with N = 2^a (8 <= N <= 4096) or
// N = 3 * 2^a (12 <= N <= 3072).
bool pq_supported_n(int n);

// Stream argument of the C ABI: NULL selects the handle's own stream,
// PQ_STREAM_NULL the HIP null (default) stream, anything else is a hipStream_t.
inline hipStream_t pq_resolve_stream(void *s, hipStream_t own) {
  if (s == nullptr) return own;
  if (s == PQ_STREAM_NULL) return (hipStream_t)0;
  return (hipStream_t)s;
}

namespace pq {

template <typename T> struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * (-i)
template <typename T> __device__ __forceinline__ cx<T> mul_mi(cx<T> a) {
  return {a.y, -a.x};
}
template <typename T> __device__ __forceinline__ cx<T> scale(cx<T> a, T s) {
  return {a.x * s, a.y * s};
}
// Component-wise conversion (rounds to nearest when narrowing).
template <typename D, typename S> __device__ __forceinline__ cx<D> cx_cast(cx<S> a) {
  return {D(a.x), D(a.y)};
}

// LDS padding: one pad element after every 16, so that a thread-blocked
// access (lane t touching elements 16t..16t+15) spreads over the banks.
__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }

}  // namespace pq
