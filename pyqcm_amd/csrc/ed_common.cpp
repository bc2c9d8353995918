// ed_common.cpp -- global parameters, caching device allocator, library stream.
#include "ed_common.hpp"

#include <atomic>
#include <cstring>

#include <chrono>
#include <cstdarg>
#include <unordered_map>

namespace ed {

Params::Params() {
  // src_util/global_parameter.cpp:44-90 (the ED knobs, same names and defaults)
  b["check_lanczos_residual"] = false;
  b["no_degenerate_BL"] = false;
  b["nosym"] = false;
  b["continued_fraction"] = false;
  b["verb_ED"] = false;
  b["verb_Hilbert"] = false;
  b["verb_warning"] = false;
  b["parallel_sectors"] = false;
  b["merge_states"] = false;
  b["combine_mcf"] = false;
  b["block_Lanczos_QR"] = false;
  d["accur_band_lanczos"] = 1e-12;
  d["accur_continued_fraction"] = 0.01;
  d["accur_Davidson"] = 1.0e-5;
  d["accur_deflation"] = 1e-7;
  d["accur_lanczos"] = 1e-12;
  d["accur_Q_matrix"] = 1.0e-5;
  d["band_lanczos_minimum_gap"] = 1e-5;
  d["minimum_weight"] = 0.01;
  d["Qmatrix_wtol"] = 1.0e-5;
  d["Qmatrix_vtol"] = 1.0e-10;
  d["temperature"] = 0.0;
  i["n_states"] = 1;
  i["max_dim_full"] = 256;
  i["max_dim_print"] = 64;
  i["max_iter_BL"] = 600;
  i["max_iter_CF"] = 400;
  i["max_iter_lanczos"] = 1000;
  i["seed"] = 0;
  i["GF_lookup_depth"] = 64;
  // 'G' (added here): GPU factorized format when the model is factorizable,
  // device CSR otherwise.  'F', 'S', 'E', 'N' keep their reference meaning.
  c["Hamiltonian_format"] = 'G';
  c["Ground_state_method"] = 'L';
  c["GF_method"] = 'L';
  // device-side knobs (no reference counterpart)
  i["hv_variant"] = 0;            // 0 auto-tuned, 1 fused row kernel, 2 two-pass, 3 two-pass with LDS slabs
  i["sell_sort_window"] = 256;    // two-pass up table: rows sorted by length within windows of this many states
  i["reduction_blocks_per_cu"] = 1;
  i["vector_block"] = 1024;       // threads per workgroup of the Krylov vector kernels (multi-dot, axpy, combine): 256, 512 or 1024
  i["hv_block"] = 0;              // row-kernel block size: 0 auto, 256 or 1024 threads
  i["hv_row_order"] = 1;          // fused row kernel: 0 label order, 1 breadth-first order of the dn rows
  b["store_lanczos_basis"] = true;  // keep Krylov vectors in HBM instead of replaying the recurrence
  b["poll_reductions"] = true;      // wait for reduction results by polling pinned memory (else stream sync)
  b["host_random_start"] = false;   // Lanczos start vector from the reference's host generator instead of the device one
  i["lanczos_block"] = 6;           // Lanczos steps queued per block without a host wait (multiple of 3; 0: a wait per step)
  i["band_lanczos_block"] = 8;      // band-Lanczos steps queued per block with their scalars computed on the device (< 2: a host wait per step)
  b["lanczos_graph"] = true;        // replay the block as a captured hipGraph
  b["lanczos_reorth"] = false;      // full FP64 reorthogonalisation of the ground-state Lanczos (basis in HBM)
  b["sell16"] = true;               // 16-bit sign-only sliced-ELL entries for single-magnitude hoppings
  d["lanczos_basis_budget_GB"] = -1.0;  // negative: 90 % of the available device memory less 4 GB
}

bool Params::getb(const std::string &n) {
  std::lock_guard<std::mutex> l(mu);
  auto it = b.find(n);
  if (it == b.end()) fail(ED_ERR_INVALID, "unknown boolean global parameter " + n);
  return it->second;
}
long Params::geti(const std::string &n) {
  std::lock_guard<std::mutex> l(mu);
  auto it = i.find(n);
  if (it == i.end()) fail(ED_ERR_INVALID, "unknown integer global parameter " + n);
  return it->second;
}
double Params::getd(const std::string &n) {
  std::lock_guard<std::mutex> l(mu);
  auto it = d.find(n);
  if (it == d.end()) fail(ED_ERR_INVALID, "unknown real global parameter " + n);
  return it->second;
}
char Params::getc(const std::string &n) {
  std::lock_guard<std::mutex> l(mu);
  auto it = c.find(n);
  if (it == c.end()) fail(ED_ERR_INVALID, "unknown char global parameter " + n);
  return it->second;
}

Params &params() {
  static Params P;
  return P;
}

// ---------------------------------------------------------------------------
// caching allocator: exact-size free lists (Krylov vectors of one sector all
// have the same size, so reuse is near perfect); freed on ed_release_cache
// or when an allocation fails.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_alloc_mu;
std::unordered_multimap<size_t, void *> g_free;
size_t g_cached_bytes = 0;

size_t round_up(size_t b) {
  const size_t a = 256;
  return (b + a - 1) / a * a;
}
}  // namespace

void dev_release_cache() {
  std::lock_guard<std::mutex> l(g_alloc_mu);
  for (auto &kv : g_free) (void)hipFree(kv.second);
  g_free.clear();
  g_cached_bytes = 0;
}

namespace {
std::atomic<int> g_cancel{0};
}

void request_cancel() { g_cancel.store(1, std::memory_order_relaxed); }

void check_cancel() {
  if (g_cancel.load(std::memory_order_relaxed) && g_cancel.exchange(0)) fail(ED_ERR_CANCELLED, "interrupted");
}

size_t dev_available_bytes() {
  size_t fr = 0, tot = 0;
  ED_HIP(hipMemGetInfo(&fr, &tot));
  std::lock_guard<std::mutex> l(g_alloc_mu);
  return fr + g_cached_bytes;
}

void *dev_alloc(size_t bytes) {
  size_t b = round_up(bytes);
  {
    std::lock_guard<std::mutex> l(g_alloc_mu);
    auto it = g_free.find(b);
    if (it != g_free.end()) {
      void *p = it->second;
      g_free.erase(it);
      g_cached_bytes -= b;
      return p;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dev_release_cache();
    e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(ED_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
  }
  return p;
}

void dev_free(void *p, size_t bytes) {
  if (!p) return;
  size_t b = round_up(bytes);
  std::lock_guard<std::mutex> l(g_alloc_mu);
  g_free.emplace(b, p);
  g_cached_bytes += b;
}

void h2d_staged(void *dst, const void *src, size_t bytes) {
  // pinned staging ring: slot k is reused once the copy out of it has run
  // (its event); the host copy into pinned memory is the only host wait, so
  // the many small table uploads of a sector no longer each pay a pageable
  // transfer's synchronous staging
  constexpr int kSlots = 16;
  constexpr size_t kSlot = size_t(512) << 10;
  static std::mutex mu;
  static char *pin = nullptr;
  static hipEvent_t ev[kSlots];
  static bool used[kSlots] = {};
  static int next = 0;
  if (!bytes) return;
  std::lock_guard<std::mutex> lock(mu);
  if (!pin) {
    ED_HIP(hipHostMalloc(reinterpret_cast<void **>(&pin), kSlots * kSlot, hipHostMallocDefault));
    for (int i = 0; i < kSlots; ++i) ED_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  }
  for (size_t off = 0; off < bytes; off += kSlot) {
    const size_t m = std::min(kSlot, bytes - off);
    const int k = next;
    next = (next + 1) % kSlots;
    if (used[k]) ED_HIP(hipEventSynchronize(ev[k]));
    char *slot = pin + (size_t)k * kSlot;
    std::memcpy(slot, static_cast<const char *>(src) + off, m);
    ED_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, slot, m, hipMemcpyHostToDevice, stream()));
    ED_HIP(hipEventRecord(ev[k], stream()));
    used[k] = true;
  }
}

void d2h_pageable(void *dst, const void *src, size_t bytes) {
  constexpr size_t kChunk = size_t(8) << 20;
  if (bytes <= kChunk) {
    if (bytes) ED_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    ED_HIP(hipStreamSynchronize(stream()));
    return;
  }
  static std::mutex mu;
  static char *pin[2] = {nullptr, nullptr};
  static hipEvent_t ev[2];
  std::lock_guard<std::mutex> lock(mu);
  if (!pin[0])
    for (int i = 0; i < 2; ++i) {
      ED_HIP(hipHostMalloc(reinterpret_cast<void **>(&pin[i]), kChunk, hipHostMallocDefault));
      ED_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  auto len = [&](size_t c) { return std::min(kChunk, bytes - c * kChunk); };
  auto issue = [&](size_t c) {
    ED_HIP(hipMemcpyAsync(pin[c % 2], static_cast<const char *>(src) + c * kChunk, len(c), hipMemcpyDeviceToHost,
                          stream()));
    ED_HIP(hipEventRecord(ev[c % 2], stream()));
  };
  auto drain = [&](size_t c) {
    ED_HIP(hipEventSynchronize(ev[c % 2]));
    char *d = static_cast<char *>(dst) + c * kChunk;
    const char *h = pin[c % 2];
    const size_t m = len(c), blk = size_t(1) << 20, nb = (m + blk - 1) / blk;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nb; ++i) std::memcpy(d + i * blk, h + i * blk, std::min(blk, m - i * blk));
  };
  issue(0);
  for (size_t c = 1; c < nch; ++c) {
    issue(c);
    drain(c - 1);
  }
  drain(nch - 1);
}

hipStream_t stream() {
  static hipStream_t s = [] {
    hipStream_t st;
    ED_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return st;
  }();
  return s;
}

void sync() { ED_HIP(hipStreamSynchronize(stream())); }

bool verbose() { return params().getb("verb_ED"); }

double wall_ms() {
  using clk = std::chrono::steady_clock;
  static const clk::time_point t0 = clk::now();
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

void vlog(const char *fmt, ...) {
  if (!verbose()) return;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::fprintf(stderr, "[ed %10.3f ms] %s\n", wall_ms(), buf);
  std::fflush(stderr);
}

}  // namespace ed
