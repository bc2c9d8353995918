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
  i["lanczos_block"] = 12;          // Lanczos steps queued per block without a host wait (multiple of 3; 0: a wait per step)
  i["band_lanczos_block"] = 6;      // band-Lanczos steps queued per block with their scalars computed on the device (< 2: a host wait per step)
  b["lanczos_graph"] = true;        // replay the block as a captured hipGraph
  i["gf_streams"] = 2;              // band-Lanczos Green function: hole and electron parts on 2 host threads / streams (1: one after the other)
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
// or when an allocation fails.  Each block belongs to the stream of the thread
// that allocated it and is handed out again only to that stream, whose order
// then covers any work still queued on the block; a block freed from another
// stream waits for that stream first.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_alloc_mu;
struct FreeBlock {
  void *p;
  hipStream_t s;
};
std::unordered_multimap<size_t, FreeBlock> g_free;
std::unordered_map<void *, hipStream_t> g_owner;  // live blocks
size_t g_cached_bytes = 0;

size_t round_up(size_t b) {
  const size_t a = 256;
  return (b + a - 1) / a * a;
}
}  // namespace

void dev_release_cache() {
  std::lock_guard<std::mutex> l(g_alloc_mu);
  for (auto &kv : g_free) (void)hipFree(kv.second.p);
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
  const hipStream_t s = stream();
  {
    std::lock_guard<std::mutex> l(g_alloc_mu);
    auto r = g_free.equal_range(b);
    for (auto it = r.first; it != r.second; ++it)
      if (it->second.s == s) {
        void *p = it->second.p;
        g_free.erase(it);
        g_cached_bytes -= b;
        g_owner[p] = s;
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
  std::lock_guard<std::mutex> l(g_alloc_mu);
  g_owner[p] = s;
  return p;
}

void dev_free(void *p, size_t bytes) {
  if (!p) return;
  size_t b = round_up(bytes);
  hipStream_t owner = nullptr;
  {
    std::lock_guard<std::mutex> l(g_alloc_mu);
    auto it = g_owner.find(p);
    if (it != g_owner.end()) {
      owner = it->second;
      g_owner.erase(it);
    }
  }
  // freed from another thread's stream: its queued work on the block ends
  // before the owner stream may hand the block out again
  if (stream_started() && stream() != owner) (void)hipStreamSynchronize(stream());
  std::lock_guard<std::mutex> l(g_alloc_mu);
  g_free.emplace(b, FreeBlock{p, owner ? owner : stream()});
  g_cached_bytes += b;
}

// Pinned staging buffers belong to one host thread at a time (each thread
// copies through its own, with events recorded on its own stream: an event
// recorded on another thread's stream cannot be waited on while that stream
// is being captured into a graph).  A thread's set comes from a process-wide
// pool on first use and goes back, drained, when the thread ends.
namespace {
struct Staging {
  static constexpr int kUp = 16;                          // upload ring slots
  static constexpr size_t kUpSlot = size_t(512) << 10;
  static constexpr size_t kDown = size_t(8) << 20;        // download double buffer
  char *up = nullptr, *down[2] = {nullptr, nullptr};
  hipEvent_t ev_up[kUp], ev_down[2];
  bool used[kUp] = {};
  int next = 0;
  Staging() {
    ED_HIP(hipHostMalloc(reinterpret_cast<void **>(&up), kUp * kUpSlot, hipHostMallocDefault));
    for (int i = 0; i < kUp; ++i) ED_HIP(hipEventCreateWithFlags(&ev_up[i], hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) ED_HIP(hipEventCreateWithFlags(&ev_down[i], hipEventDisableTiming));
  }
  char *down_buf(int i) {
    if (!down[i]) ED_HIP(hipHostMalloc(reinterpret_cast<void **>(&down[i]), kDown, hipHostMallocDefault));
    return down[i];
  }
  // the uploads still in flight end before another thread may own the set
  void drain() {
    for (int i = 0; i < kUp; ++i)
      if (used[i]) (void)hipEventSynchronize(ev_up[i]);
    for (int i = 0; i < kUp; ++i) used[i] = false;
  }
};
std::mutex g_staging_mu;
std::vector<Staging *> *g_staging_pool = new std::vector<Staging *>;  // never destroyed
struct ThreadStaging {
  Staging *s = nullptr;
  ~ThreadStaging() {
    if (!s) return;
    s->drain();
    std::lock_guard<std::mutex> l(g_staging_mu);
    g_staging_pool->push_back(s);
  }
};
thread_local ThreadStaging t_staging;
Staging &staging() {
  if (!t_staging.s) {
    {
      std::lock_guard<std::mutex> l(g_staging_mu);
      if (!g_staging_pool->empty()) {
        t_staging.s = g_staging_pool->back();
        g_staging_pool->pop_back();
      }
    }
    if (!t_staging.s) t_staging.s = new Staging;
  }
  return *t_staging.s;
}
}  // namespace

void h2d_staged(void *dst, const void *src, size_t bytes) {
  // pinned staging ring: slot k is reused once the copy out of it has run
  // (its event); the host copy into pinned memory is the only host wait, so
  // the many small table uploads of a sector no longer each pay a pageable
  // transfer's synchronous staging
  if (!bytes) return;
  Staging &R = staging();
  for (size_t off = 0; off < bytes; off += Staging::kUpSlot) {
    const size_t m = std::min(Staging::kUpSlot, bytes - off);
    const int k = R.next;
    R.next = (R.next + 1) % Staging::kUp;
    if (R.used[k]) ED_HIP(hipEventSynchronize(R.ev_up[k]));
    char *slot = R.up + (size_t)k * Staging::kUpSlot;
    std::memcpy(slot, static_cast<const char *>(src) + off, m);
    ED_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, slot, m, hipMemcpyHostToDevice, stream()));
    ED_HIP(hipEventRecord(R.ev_up[k], stream()));
    R.used[k] = true;
  }
}

void d2h_pageable(void *dst, const void *src, size_t bytes) {
  constexpr size_t kChunk = Staging::kDown;
  if (bytes <= kChunk) {
    if (bytes) ED_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    ED_HIP(hipStreamSynchronize(stream()));
    return;
  }
  Staging &R = staging();
  char *pin[2] = {R.down_buf(0), R.down_buf(1)};
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  auto len = [&](size_t c) { return std::min(kChunk, bytes - c * kChunk); };
  auto issue = [&](size_t c) {
    ED_HIP(hipMemcpyAsync(pin[c % 2], static_cast<const char *>(src) + c * kChunk, len(c), hipMemcpyDeviceToHost,
                          stream()));
    ED_HIP(hipEventRecord(R.ev_down[c % 2], stream()));
  };
  auto drain = [&](size_t c) {
    ED_HIP(hipEventSynchronize(R.ev_down[c % 2]));
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

// Each host thread queues its work on a stream of its own (non-blocking,
// default priority), taken from a process-wide pool on the thread's first
// use and handed back when the thread ends, so threads that come and go
// reuse streams; work a thread left queued stays ordered before the next
// owner's.
namespace {
std::mutex g_stream_mu;
std::vector<hipStream_t> *g_stream_pool = new std::vector<hipStream_t>;  // never destroyed
struct ThreadStream {
  hipStream_t s = nullptr;
  ~ThreadStream() {
    if (!s) return;
    std::lock_guard<std::mutex> l(g_stream_mu);
    g_stream_pool->push_back(s);
  }
};
thread_local ThreadStream t_stream;
}  // namespace

hipStream_t stream() {
  if (!t_stream.s) {
    std::lock_guard<std::mutex> l(g_stream_mu);
    if (!g_stream_pool->empty()) {
      t_stream.s = g_stream_pool->back();
      g_stream_pool->pop_back();
    } else {
      ED_HIP(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
    }
  }
  return t_stream.s;
}

bool stream_started() { return t_stream.s != nullptr; }

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
