// Shared host and device utilities of librwig_hip.so (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rwig_hip.h"

namespace rwig {

// ---------------------------------------------------------------- errors

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define RWIG_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::rwig::Error(RWIG_ERR_HIP, std::string(#expr) + ": " +           \
                                            hipGetErrorString(e_));           \
  } while (0)

#define RWIG_REQUIRE(cond, msg)                                                \
  do {                                                                         \
    if (!(cond)) throw ::rwig::Error(RWIG_ERR_INVALID, msg);                   \
  } while (0)

// Runs fn, converting exceptions into a status code and a last-error text.
template <class F> int guarded(const char *where, F &&fn) {
  try {
    fn();
    set_last_error("");
    return RWIG_OK;
  } catch (const Error &e) {
    set_last_error(std::string(where) + ": " + e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error(std::string(where) + ": host allocation failed");
    return RWIG_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_last_error(std::string(where) + ": " + e.what());
    return RWIG_ERR_INVALID;
  }
}

// ---------------------------------------------------------------- memory

// Caching device allocator: blocks freed by a call are kept per device and
// handed to later calls of the same size class, so repeated solves do not
// pay hipMalloc/hipFree each time.
class DevicePool {
 public:
  static DevicePool &instance();
  void *alloc(size_t bytes);
  void release(void *p, size_t bytes);
  void trim();

 private:
  std::mutex mu_;
  std::map<std::pair<int, size_t>, std::vector<void *>> free_;
};

size_t round_size(size_t bytes);

template <class T> class DBuf {
 public:
  DBuf() {}
  explicit DBuf(size_t n) { reset(n); }
  ~DBuf() { clear(); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    clear();
    p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0;
    return *this;
  }
  void reset(size_t n) {
    clear();
    n_ = n;
    if (n) p_ = static_cast<T *>(DevicePool::instance().alloc(n * sizeof(T)));
  }
  void clear() {
    if (p_) DevicePool::instance().release(p_, n_ * sizeof(T));
    p_ = nullptr;
    n_ = 0;
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T *p_ = nullptr;
  size_t n_ = 0;
};

// Owned non-blocking stream at default priority.
class Stream {
 public:
  Stream() { RWIG_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~Stream() { if (s_) (void)hipStreamDestroy(s_); }
  Stream(const Stream &) = delete;
  Stream &operator=(const Stream &) = delete;
  hipStream_t get() const { return s_; }
  void sync() const { RWIG_HIP_CHECK(hipStreamSynchronize(s_)); }

 private:
  hipStream_t s_ = nullptr;
};

void ensure_device();

template <class T> void h2d(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) RWIG_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T> void d2h(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) RWIG_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <class T> void d2d(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) RWIG_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

inline void check_launch() { RWIG_HIP_CHECK(hipGetLastError()); }

inline unsigned grid_for(size_t n, unsigned block, unsigned cap = 65535u * 16u) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------- device side

constexpr int kWave = 64;

// Device-resident loop control of one iterative solve. The loop of the
// reference, while (iter < maxiter && err >= tol) { iter++; ... }, is kept on
// the device: every kernel of an iteration returns at once unless active is
// set, and the iteration's finalise kernel updates iter, err and active. The
// host only polls active between batches of enqueued iterations.
struct LoopState {
  int iter;
  int active;
  int maxiter;
  int pad;
  double err;
  double tol;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return v;
}
// Index of the calling wavefront in a one-dimensional workgroup, as a value
// the compiler knows to be uniform across the wavefront (every lane of a
// wavefront has the same threadIdx.x / 64). Loop bounds and column offsets
// derived from it then stay in scalar registers: a column offset of a
// global load becomes one scalar multiply instead of six vector
// instructions per load, two of them quarter-rate multiplies.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Block-wide sum with a fixed reduction tree (deterministic). sh must hold
// blockDim.x / 64 doubles. Every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < nw; ++k) t += sh[k];
  return t;
}
__device__ __forceinline__ double block_min(double v, double *sh) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  v = wave_min(v);
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = sh[0];
  for (int k = 1; k < nw; ++k) t = fmin(t, sh[k]);
  return t;
}

}  // namespace rwig
