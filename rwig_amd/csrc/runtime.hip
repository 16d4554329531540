// Host runtime of librwig_hip.so: error text, device pool, device checks.
#include "common.hpp"

namespace rwig {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

size_t round_size(size_t bytes) {
  // size classes: 256 B granularity up to 1 MiB, then 1/8 of the next power
  // of two, so a class wastes at most 12.5 %
  if (bytes <= (1u << 20)) return (bytes + 255) & ~size_t(255);
  size_t p = 1;
  while (p < bytes) p <<= 1;
  const size_t step = p / 16;
  return (bytes + step - 1) / step * step;
}

DevicePool &DevicePool::instance() {
  static DevicePool *pool = new DevicePool();  // lives until process exit
  return *pool;
}

void *DevicePool::alloc(size_t bytes) {
  const size_t sz = round_size(bytes);
  int dev = 0;
  RWIG_HIP_CHECK(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = free_.find({dev, sz});
    if (it != free_.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(RWIG_ERR_NOMEM, "hipMalloc of " + std::to_string(sz) + " bytes failed");
    }
  }
  return p;
}

void DevicePool::release(void *p, size_t bytes) {
  if (!p) return;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(mu_);
  free_[{dev, round_size(bytes)}].push_back(p);
}

void DevicePool::trim() {
  std::lock_guard<std::mutex> lk(mu_);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : free_) {
    if (kv.first.first != cur) continue;
    for (void *p : kv.second) (void)hipFree(p);
    kv.second.clear();
  }
}

void ensure_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    (void)hipGetLastError();
    throw Error(RWIG_ERR_NODEV, "no HIP device available");
  }
}

}  // namespace rwig

extern "C" {

int rwig_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n > 0 ? 1 : 0;
}

const char *rwig_last_error(void) { return rwig::g_last_error.c_str(); }

}  // extern "C"
