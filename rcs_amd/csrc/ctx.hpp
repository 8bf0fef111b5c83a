// Execution context: one HIP stream, a caching HBM block pool and a staging ring for
// launch parameter arrays (pinned host memory mirrored into HBM in stream order).
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>

#include "rcs_common.hpp"

namespace rcs {

// Best-fit caching allocator over hipMalloc'd blocks. The factorization creates and
// drops thousands of blocks whose sizes repeat (box x box, rank x box, rank x rank);
// cached blocks are reused instead of going back to the driver.
class DevPool {
 public:
  ~DevPool() { release_all(); }
  void* alloc(size_t bytes) {
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_.lower_bound(bytes);
    if (!no_reuse_ && it != free_.end() && it->first <= bytes + bytes / 4) {
      void* p = it->second;
      size_t sz = it->first;
      free_.erase(it);
      used_[p] = sz;
      in_use_ += sz;
      return p;
    }
    // blocks cached here are HBM the HIP runtime cannot use: past the limit (all pools of the
    // process together) this pool's cached blocks go back first
    if (limit_ == 0) {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      limit_ = tot / 100 * 85;
    }
    if (all_pools_total() + bytes > limit_) {
      const size_t others = all_pools_total() - total_;
      const size_t room = limit_ > others + bytes ? limit_ - others - bytes : 0;
      trim_to(room);
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();  // the failure is handled here: do not leave it for the next check
      trim();
      HIPCHK(hipMalloc(&p, bytes));
    }
    used_[p] = bytes;
    total_ += bytes;
    g_total().fetch_add(bytes);
    in_use_ += bytes;
    peak_ = std::max(peak_, in_use_);
    return p;
  }
  void free(void* p) {
    if (!p) return;
    auto it = used_.find(p);
    if (it == used_.end()) throw std::runtime_error("DevPool: free of unknown pointer");
    in_use_ -= it->second;
    free_.insert({it->second, p});
    used_.erase(it);
  }
  void trim() {
    for (auto& kv : free_) { (void)hipFree(kv.second); total_ -= kv.first; g_total().fetch_sub(kv.first); }
    free_.clear();
  }
  // Return cached blocks, largest first, until at most `target` bytes are held.
  void trim_to(size_t target) {
    while (total_ > target && !free_.empty()) {
      auto it = std::prev(free_.end());
      (void)hipFree(it->second);
      total_ -= it->first;
      g_total().fetch_sub(it->first);
      free_.erase(it);
    }
  }
  void release_all() {
    trim();
    for (auto& kv : used_) (void)hipFree(kv.first);
    used_.clear();
    g_total().fetch_sub(total_);
    total_ = in_use_ = 0;
  }
  static size_t all_pools_total() { return g_total().load(); }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }
  void reset_peak() { peak_ = in_use_; }
  size_t total() const { return total_; }

 private:
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> used_;
  size_t total_ = 0, in_use_ = 0, peak_ = 0;
  size_t limit_ = 0;  // 85% of the device memory, set at the first allocation
  static std::atomic<size_t>& g_total() {  // bytes held by every pool of the process
    static std::atomic<size_t> t{0};
    return t;
  }
  bool no_reuse_ = std::getenv("RCS_DEBUG_NO_REUSE") != nullptr;  // debugging: never recycle blocks
};

// Phase timing with HIP events on the library's stream (off unless enabled). Each
// record brackets the launches of one tagged operation; totals are folded in at
// read-out (or when the event pool fills up).
enum ProfTag { PROF_SCHUR = 0, PROF_GETRF = 1, PROF_GETRS = 2, PROF_ENTRIES = 3, PROF_ACA = 4, PROF_RRQR = 5,
               PROF_UPDATE = 6, PROF_NTAGS = 7 };
class Profiler {
 public:
  struct Agg {
    double work = 0, ms = 0;
    long long launches = 0;
  };
  bool on = false;
  ~Profiler() {
    for (auto& e : events_) (void)hipEventDestroy(e);
  }
  int begin(hipStream_t st) {
    if (!on) return -1;
    int i = take();
    HIPCHK(hipEventRecord(events_[i], st));
    return i;
  }
  void end(hipStream_t st, int h, int tag, double work) {
    if (!on || h < 0) return;
    int i = take();
    HIPCHK(hipEventRecord(events_[i], st));
    recs_.push_back({h, i, tag, work});
    if (used_ > 4000) fold(st);
  }
  void fold(hipStream_t st) {
    if (recs_.empty()) { used_ = 0; return; }
    HIPCHK(hipStreamSynchronize(st));
    for (auto& r : recs_) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, events_[r.a], events_[r.b]));
      agg_[r.tag].ms += ms;
      agg_[r.tag].work += r.work;
      agg_[r.tag].launches += 1;
    }
    recs_.clear();
    used_ = 0;
  }
  void reset(hipStream_t st) {
    fold(st);
    for (auto& a : agg_) a = Agg();
  }
  Agg get(hipStream_t st, int tag) {
    fold(st);
    return agg_[tag];
  }

 private:
  struct Rec {
    int a, b, tag;
    double work;
  };
  int take() {
    if (used_ == (int)events_.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      events_.push_back(e);
    }
    return used_++;
  }
  std::vector<hipEvent_t> events_;
  std::vector<Rec> recs_;
  int used_ = 0;
  Agg agg_[PROF_NTAGS];
};

struct Ctx {
  hipStream_t st = nullptr;
  DevPool pool;
  Profiler prof;
  char* h_stage = nullptr;
  char* d_stage = nullptr;
  size_t cap = 0, off = 0;
  double* d_scalars = nullptr;  // scratch for reductions read back by the host
  double* h_scalars = nullptr;
  size_t n_scalars = 0;
  // per-tile arrival counters of split-inner-index GEMMs (zero between launches: the last
  // workgroup of a tile resets its counter)
  static constexpr int n_tile_counters = 256;
  unsigned* d_tile_counters = nullptr;

  explicit Ctx(size_t stage_bytes = size_t(64) << 20, size_t nscal = 1 << 16) {
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    cap = stage_bytes;
    HIPCHK(hipHostMalloc((void**)&h_stage, cap, hipHostMallocDefault));
    HIPCHK(hipMalloc((void**)&d_stage, cap));
    n_scalars = nscal;
    HIPCHK(hipMalloc((void**)&d_scalars, sizeof(double) * nscal));
    HIPCHK(hipHostMalloc((void**)&h_scalars, sizeof(double) * nscal, hipHostMallocDefault));
    HIPCHK(hipMalloc((void**)&d_tile_counters, sizeof(unsigned) * n_tile_counters));
    HIPCHK(hipMemset(d_tile_counters, 0, sizeof(unsigned) * n_tile_counters));
  }
  ~Ctx() {
    if (st) (void)hipStreamSynchronize(st);
    pool.release_all();
    if (h_stage) (void)hipHostFree(h_stage);
    if (d_stage) (void)hipFree(d_stage);
    if (d_scalars) (void)hipFree(d_scalars);
    if (h_scalars) (void)hipHostFree(h_scalars);
    if (d_tile_counters) (void)hipFree(d_tile_counters);
    if (st) (void)hipStreamDestroy(st);
  }
  Ctx(const Ctx&) = delete;
  Ctx& operator=(const Ctx&) = delete;

  // Copy a small parameter array to HBM in stream order; returns the device address.
  void* push(const void* data, size_t bytes) {
    const size_t span = (bytes + 255) & ~size_t(255);  // keeps every block 256-byte aligned
    if (span > cap) throw std::runtime_error("Ctx::push: parameter block larger than the staging ring");
    if (off + span > cap) {
      HIPCHK(hipStreamSynchronize(st));  // every earlier copy out of the ring has completed
      off = 0;
    }
    std::memcpy(h_stage + off, data, bytes);  // exactly the caller's bytes, never past its end
    void* d = d_stage + off;
    HIPCHK(hipMemcpyAsync(d, h_stage + off, bytes, hipMemcpyHostToDevice, st));
    off += span;
    pushed += span;
    return d;
  }
  // Stream-ordered host -> HBM copy of a host buffer of any size into device memory the
  // caller owns: staged through the pinned ring (no pageable-memory DMA), then copied
  // device to device, so the ring space can be reused by later launches.
  void upload(void* dst, const void* src, size_t bytes) {
    const size_t chunk = cap / 4;
    for (size_t o = 0; o < bytes; o += chunk) {
      const size_t n = std::min(chunk, bytes - o);
      void* d = push((const char*)src + o, n);
      HIPCHK(hipMemcpyAsync((char*)dst + o, d, n, hipMemcpyDeviceToDevice, st));
    }
  }
  // Bytes pushed since the counter was last reset: a batch whose parameter blocks exceed
  // the ring would overwrite its own earlier blocks before its launch.
  size_t pushed = 0;
  void sync() { HIPCHK(hipStreamSynchronize(st)); }
  // Fetch n doubles of d_scalars (synchronizes).
  const double* fetch_scalars(size_t n) {
    if (n > n_scalars) throw std::runtime_error("fetch_scalars: too many");
    HIPCHK(hipMemcpyAsync(h_scalars, d_scalars, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    sync();
    return h_scalars;
  }
};

}  // namespace rcs
