// Shared plumbing of libfemhip: HIP error checks, a device buffer owner, the
// stream-ordered scratch allocator and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <condition_variable>
#include <map>
#include <mutex>
#include <vector>

namespace femhip {

// A HIP runtime call failed (the reference's CUDA_CHECK, include/cuda/cuda-check.hpp:17-22):
// message "file:line: expr failed: Name (String)". Non-sticky errors are cleared.
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct OutOfMemory : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t err, const char* expr, const char* file, int line) {
  if (err == hipSuccess) return;
  (void)hipGetLastError();
  std::string msg = std::string(file) + ":" + std::to_string(line) + ": " + expr + " failed: " +
                    hipGetErrorName(err) + " (" + hipGetErrorString(err) + ")";
  if (err == hipErrorOutOfMemory) throw OutOfMemory(msg);
  throw DeviceError(msg);
}

#define HIP_CHECK(expr) ::femhip::hip_check((expr), #expr, __FILE__, __LINE__)
#define HIP_CHECK_LAUNCH() HIP_CHECK(hipGetLastError())

inline void require(bool ok, const std::string& msg) {
  if (!ok) throw std::invalid_argument(msg);
}

// Shared/exclusive lock between hipGraph stream capture (exclusive) and the
// device calls that must not run while another thread records a graph
// (shared): synchronous copies on the legacy stream, allocation and release,
// stream creation and destruction. Writers (captures) are preferred: once one
// waits, new readers wait too. Shared holds are re-entrant per thread, and the
// recording thread passes its own shared requests.
class CaptureLock {
 public:
  void lock_shared();
  void unlock_shared();
  void lock();
  void unlock();

 private:
  std::mutex m_;
  std::condition_variable cv_;
  int readers_ = 0, writers_waiting_ = 0;
  bool writer_ = false;
};
CaptureLock& capture_lock();
struct SharedCapture {
  SharedCapture() { capture_lock().lock_shared(); }
  ~SharedCapture() { capture_lock().unlock_shared(); }
  SharedCapture(const SharedCapture&) = delete;
  SharedCapture& operator=(const SharedCapture&) = delete;
};
struct ExclusiveCapture {
  ExclusiveCapture() { capture_lock().lock(); }
  ~ExclusiveCapture() { capture_lock().unlock(); }
  ExclusiveCapture(const ExclusiveCapture&) = delete;
  ExclusiveCapture& operator=(const ExclusiveCapture&) = delete;
};

// Device blocks kept for reuse by a long-lived owner (the solve server):
// hipFree waits for every stream of the device, so a server that freed each
// job's buffers would stall all other jobs' solves. take() hands out a cached
// block at most 1/8 (+64 KiB) larger than asked (best fit), else hipMalloc's a
// new one; on an out-of-memory failure it frees the cache and tries once more.
// give() caches a block; past max_cached bytes it frees the largest ones.
// A block must be idle when given back (DevBuf synchronizes its stream first).
class BlockCache {
 public:
  explicit BlockCache(size_t max_cached) : max_(max_cached) {}
  ~BlockCache() { trim(0); }
  BlockCache(const BlockCache&) = delete;
  BlockCache& operator=(const BlockCache&) = delete;
  void* take(size_t bytes, size_t* got);
  void give(void* p, size_t bytes);
  void trim(size_t keep);  // frees cached blocks, largest first, down to `keep` bytes
  size_t cached_bytes();

 private:
  std::mutex m_;
  std::multimap<size_t, void*> free_;
  size_t cached_ = 0, max_;
};

// While a CachedAllocScope is alive on a thread, DevBuf allocations on that
// thread come from `cache` and belong to `stream`: all work on them must be
// ordered on that stream, which must outlive them; release synchronizes it and
// gives the block back to the cache.
struct CacheScope {
  BlockCache* cache;
  hipStream_t stream;
};
inline const CacheScope*& cache_scope() {
  static thread_local const CacheScope* scope = nullptr;
  return scope;
}
class CachedAllocScope {
 public:
  CachedAllocScope(BlockCache* cache, hipStream_t s) : scope_{cache, s}, prev_(cache_scope()) {
    cache_scope() = &scope_;
  }
  ~CachedAllocScope() { cache_scope() = prev_; }
  CachedAllocScope(const CachedAllocScope&) = delete;
  CachedAllocScope& operator=(const CachedAllocScope&) = delete;

 private:
  CacheScope scope_;
  const CacheScope* prev_;
};

// Owner of one device allocation (move-only), the counterpart of the reference's
// DeviceArray (include/cuda/device-array.hpp:18-95) without the synchronous copies.
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { allocate(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), cache_(o.cache_), s_(o.s_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.n_ = 0;
    o.cache_ = nullptr;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      cache_ = o.cache_;
      s_ = o.s_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.n_ = 0;
      o.cache_ = nullptr;
    }
    return *this;
  }
  void allocate(size_t n) {
    release();
    if (n == 0) return;
    if (const CacheScope* cs = cache_scope()) {
      p_ = static_cast<T*>(cs->cache->take(n * sizeof(T), &bytes_));
      cache_ = cs->cache;
      s_ = cs->stream;
      n_ = n;
      return;
    }
    SharedCapture lk;
    HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
    n_ = n;
  }
  void upload(const T* host, size_t n) {
    allocate(n);
    if (n == 0) return;
    if (cache_) {  // in the order of the block's stream, complete on return
      HIP_CHECK(hipMemcpyAsync(p_, host, n * sizeof(T), hipMemcpyHostToDevice, s_));
      HIP_CHECK(hipStreamSynchronize(s_));
      return;
    }
    SharedCapture lk;
    HIP_CHECK(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T>& host) { upload(host.data(), host.size()); }
  std::vector<T> download() const {
    std::vector<T> h(n_);
    SharedCapture lk;
    if (n_) HIP_CHECK(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  void release() noexcept {
    if (p_ && cache_) {
      (void)hipStreamSynchronize(s_);  // idle before another job may take it
      cache_->give(p_, bytes_);
    } else if (p_) {
      SharedCapture lk;
      (void)hipFree(p_);
    }
    p_ = nullptr;
    n_ = 0;
    cache_ = nullptr;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
  BlockCache* cache_ = nullptr;  // set when the block came from a BlockCache
  hipStream_t s_ = nullptr;      // the stream a cached block belongs to
  size_t bytes_ = 0;             // the cached block's size
};

// Stream-ordered temporary (hipMallocAsync/hipFreeAsync on the caller's stream),
// for the E/Q-vector scratch of the unfused operator path.
template <class T>
class StreamTmp {
 public:
  StreamTmp(size_t n, hipStream_t s) : s_(s), n_(n) {
    SharedCapture lk;
    if (n) HIP_CHECK(hipMallocAsync(reinterpret_cast<void**>(&p_), n * sizeof(T), s));
  }
  ~StreamTmp() {
    if (p_) {
      SharedCapture lk;
      (void)hipFreeAsync(p_, s_);
    }
  }
  StreamTmp(const StreamTmp&) = delete;
  StreamTmp& operator=(const StreamTmp&) = delete;
  T* data() { return p_; }
  size_t size() const { return n_; }

 private:
  hipStream_t s_;
  T* p_ = nullptr;
  size_t n_ = 0;
};

inline int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

inline unsigned grid_for(size_t items, int threads) {
  const size_t g = (items + threads - 1) / threads;
  return static_cast<unsigned>(g == 0 ? 1 : g);
}

// Fill n elements of p with zero on `s` (a kernel, not hipMemsetAsync, so it is
// safe inside a captured hipGraph).
void zero_fill(void* p, size_t bytes, hipStream_t s);
void stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

// Restriction / basis / qfunction vocabulary shared by the translation units.
enum : int { kNoTranspose = 0, kTranspose = 1 };
enum : int { kEvalNone = 0, kEvalInterp = 1, kEvalGrad = 2, kEvalWeight = 3 };
enum : int { kScatterAtomic = 0, kScatterDeterministic = 1 };

}  // namespace femhip
