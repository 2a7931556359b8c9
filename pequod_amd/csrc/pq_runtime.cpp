// pq_runtime.cpp -- error reporting, device memory helpers and host-array
// mirrors for libpequod_hip.so.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "pequod_hip.h"

static thread_local char g_err[1024] = "";

void pq_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

bool pq_supported_n(int n) {
  const int N = n + 1;
  if (N < 8 || N > 4096) return false;
  int m = N;
  if (m % 3 == 0) m /= 3;
  if (m < 4) return false;
  return (m & (m - 1)) == 0;
}

#define RT_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      pq_set_error("%s failed: %s", #call, hipGetErrorString(e_));             \
      return e_ == hipErrorOutOfMemory ? PQ_ERR_NOMEM : PQ_ERR_HIP;            \
    }                                                                          \
  } while (0)

namespace {
struct Mirror {
  size_t bytes;
  void *dev;
};
std::mutex g_mx;
std::map<uintptr_t, Mirror> g_mirrors;  // keyed by host start address

// Mirror containing host address p, or end().
std::map<uintptr_t, Mirror>::iterator find_mirror(uintptr_t p) {
  auto it = g_mirrors.upper_bound(p);
  if (it == g_mirrors.begin()) return g_mirrors.end();
  --it;
  if (p >= it->first && p < it->first + it->second.bytes) return it;
  return g_mirrors.end();
}
}  // namespace

extern "C" {

const char *pq_last_error(void) { return g_err; }

int pq_device_info(char *buf, int len) {
  int dev = 0;
  RT_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  RT_HIP(hipGetDeviceProperties(&p, dev));
  if (buf && len > 0)
    snprintf(buf, (size_t)len, "%s (%s), %d CUs, %.1f GiB, LDS/WG %zu", p.name,
             p.gcnArchName, p.multiProcessorCount,
             (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0),
             (size_t)p.sharedMemPerBlock);
  return PQ_OK;
}

int pq_malloc(void **dev_ptr, size_t bytes) {
  if (!dev_ptr) {
    pq_set_error("pq_malloc: null output pointer");
    return PQ_ERR_ARG;
  }
  RT_HIP(hipMalloc(dev_ptr, bytes));
  return PQ_OK;
}

int pq_free(void *dev_ptr) {
  RT_HIP(hipFree(dev_ptr));
  return PQ_OK;
}

int pq_memcpy_h2d(void *dev, const void *host, size_t bytes) {
  RT_HIP(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
  return PQ_OK;
}

int pq_memcpy_d2h(void *host, const void *dev, size_t bytes) {
  RT_HIP(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
  return PQ_OK;
}

int pq_synchronize(void) {
  RT_HIP(hipDeviceSynchronize());
  return PQ_OK;
}

int pq_mirror_register(void *host, size_t bytes) {
  if (!host || bytes == 0) {
    pq_set_error("pq_mirror_register: empty range");
    return PQ_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mx);
  uintptr_t p = (uintptr_t)host;
  if (find_mirror(p) != g_mirrors.end() || find_mirror(p + bytes - 1) != g_mirrors.end()) {
    pq_set_error("pq_mirror_register: range overlaps an existing mirror");
    return PQ_ERR_ARG;
  }
  void *dev = nullptr;
  RT_HIP(hipMalloc(&dev, bytes));
  hipError_t e = hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    pq_set_error("pq_mirror_register: upload failed: %s", hipGetErrorString(e));
    return PQ_ERR_HIP;
  }
  g_mirrors[p] = Mirror{bytes, dev};
  return PQ_OK;
}

int pq_mirror_unregister(void *host) {
  std::lock_guard<std::mutex> lk(g_mx);
  auto it = g_mirrors.find((uintptr_t)host);
  if (it == g_mirrors.end()) {
    pq_set_error("pq_mirror_unregister: not a registered mirror start");
    return PQ_ERR_ARG;
  }
  (void)hipDeviceSynchronize();
  (void)hipFree(it->second.dev);
  g_mirrors.erase(it);
  return PQ_OK;
}

void *pq_mirror_device_ptr(const void *host) {
  std::lock_guard<std::mutex> lk(g_mx);
  auto it = find_mirror((uintptr_t)host);
  if (it == g_mirrors.end()) return nullptr;
  return (char *)it->second.dev + ((uintptr_t)host - it->first);
}

int pq_mirror_push(const void *host) {
  std::lock_guard<std::mutex> lk(g_mx);
  auto it = find_mirror((uintptr_t)host);
  if (it == g_mirrors.end()) {
    pq_set_error("pq_mirror_push: pointer is not in a registered mirror");
    return PQ_ERR_ARG;
  }
  RT_HIP(hipDeviceSynchronize());
  RT_HIP(hipMemcpy(it->second.dev, (const void *)it->first, it->second.bytes,
                   hipMemcpyHostToDevice));
  return PQ_OK;
}

int pq_mirror_pull(void *host) {
  std::lock_guard<std::mutex> lk(g_mx);
  auto it = find_mirror((uintptr_t)host);
  if (it == g_mirrors.end()) {
    pq_set_error("pq_mirror_pull: pointer is not in a registered mirror");
    return PQ_ERR_ARG;
  }
  RT_HIP(hipDeviceSynchronize());
  RT_HIP(hipMemcpy((void *)it->first, it->second.dev, it->second.bytes,
                   hipMemcpyDeviceToHost));
  return PQ_OK;
}

}  // extern "C"
