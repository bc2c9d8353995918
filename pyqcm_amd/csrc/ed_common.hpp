// ed_common.hpp -- status codes, errors, global parameters, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pyqcm_ed.h"

namespace ed {

// Exception carried inside the library; converted to a status code at the
// C-ABI boundary (no exception crosses it).  Replaces qcm_ED_throw
// (src_python/qcm_lib.cpp:71-74).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

// interrupt requests (ed_request_cancel): check_cancel throws ED_ERR_CANCELLED
// once per request; the solvers call it once per step
void request_cancel();
void check_cancel();

#define ED_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess)                                                                  \
      ::ed::fail(ED_ERR_HIP, std::string("HIP error ") + hipGetErrorString(_e) + " at " + \
                                 __FILE__ + ":" + std::to_string(__LINE__));               \
  } while (0)

// ---------------------------------------------------------------------------
// Global parameters: same names and defaults as the reference registry
// (src_util/global_parameter.cpp:44-90), read once per call.
// ---------------------------------------------------------------------------
struct Params {
  std::map<std::string, bool> b;
  std::map<std::string, long> i;
  std::map<std::string, double> d;
  std::map<std::string, char> c;
  std::mutex mu;
  Params();
  bool getb(const std::string &n);
  long geti(const std::string &n);
  double getd(const std::string &n);
  char getc(const std::string &n);
};
Params &params();

// ---------------------------------------------------------------------------
// Device memory: a size-bucketed caching allocator so that repeated solves do
// not pay hipMalloc/hipFree on every Krylov vector.
// ---------------------------------------------------------------------------
void *dev_alloc(size_t bytes);
void dev_free(void *p, size_t bytes);
void dev_release_cache();
// free device memory plus the blocks this library holds in its cache
size_t dev_available_bytes();
// the calling thread's compute stream (default priority; one per host thread)
hipStream_t stream();
// whether the calling thread has a stream already
bool stream_started();

// host-to-device copy out of pageable memory through a pinned staging ring
// (asynchronous on the calling thread's stream once the data is staged)
void h2d_staged(void *dst, const void *src, size_t bytes);
// device-to-host copy into pageable memory through pinned staging buffers
// (transfers of one chunk overlap the host copy of the previous one);
// synchronizes the calling thread's stream
void d2h_pageable(void *dst, const void *src, size_t bytes);

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  // set for a view: p lies in memory that `owner` keeps alive, and is not freed here
  std::shared_ptr<void> owner;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), owner(std::move(o.owner)) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      owner = std::move(o.owner);
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  // count elements at q inside memory kept alive by keep
  static DevBuf view(T *q, size_t count, std::shared_ptr<void> keep) {
    DevBuf b;
    b.p = q;
    b.n = count;
    b.owner = std::move(keep);
    return b;
  }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) p = static_cast<T *>(dev_alloc(count * sizeof(T)));
  }
  void release() {
    if (p && !owner) dev_free(p, n * sizeof(T));
    owner.reset();
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count) {
    if (count > n) alloc(count);
    if (count) h2d_staged(p, h, count * sizeof(T));
  }
  void upload(const std::vector<T> &h) { upload(h.data(), h.size()); }
  void download(T *h, size_t count) const { d2h_pageable(h, p, count * sizeof(T)); }
  std::vector<T> to_host() const {
    std::vector<T> h(n);
    download(h.data(), n);
    return h;
  }
};

void sync();

// verb_ED progress lines on stderr (the reference prints its solver progress
// under the same flag), each stamped with the wall time in ms since the
// library's first stamp
bool verbose();
void vlog(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
double wall_ms();

}  // namespace ed
