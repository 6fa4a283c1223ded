// Device plumbing: the graph-capture-safe fill kernel, device queries and a
// launch probe for the host-forced error tests.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"

namespace femhip {

namespace {

__global__ __launch_bounds__(256) void zero_kernel(uint4* __restrict__ p, size_t n16, unsigned char* tail,
                                                   size_t ntail) {
  const size_t stride = static_cast<size_t>(gridDim.x) * 256;
  for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n16; i += stride)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// dst[i] = src[i] over n16 16-byte words, one word per thread, nontemporal
// both ways (the data is touched once).
__global__ __launch_bounds__(256) void stream_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                          size_t n16) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n16) return;
  const d2 x = __builtin_nontemporal_load(reinterpret_cast<const d2*>(src) + i);
  __builtin_nontemporal_store(x, reinterpret_cast<d2*>(dst) + i);
}

__global__ void probe_kernel(int* sink) {
  if (sink) sink[threadIdx.x] = 0;
}

}  // namespace

namespace {
thread_local int t_shared = 0;       // shared holds of this thread (re-entrant)
thread_local bool t_exclusive = false;  // this thread is recording a graph
}  // namespace

void CaptureLock::lock_shared() {
  if (t_exclusive || t_shared++ > 0) return;
  std::unique_lock<std::mutex> lk(m_);
  cv_.wait(lk, [&] { return !writer_ && writers_waiting_ == 0; });
  ++readers_;
}
void CaptureLock::unlock_shared() {
  if (t_exclusive || --t_shared > 0) return;
  std::lock_guard<std::mutex> lk(m_);
  if (--readers_ == 0) cv_.notify_all();
}
void CaptureLock::lock() {
  if (t_shared > 0) throw std::logic_error("capture lock: recording while holding device setup");
  std::unique_lock<std::mutex> lk(m_);
  ++writers_waiting_;
  cv_.wait(lk, [&] { return !writer_ && readers_ == 0; });
  --writers_waiting_;
  writer_ = true;
  t_exclusive = true;
}
void CaptureLock::unlock() {
  std::lock_guard<std::mutex> lk(m_);
  writer_ = false;
  t_exclusive = false;
  cv_.notify_all();
}
CaptureLock& capture_lock() {
  static CaptureLock l;
  return l;
}

void* BlockCache::take(size_t bytes, size_t* got) {
  const size_t want = (bytes + 255) & ~static_cast<size_t>(255);
  {
    std::lock_guard<std::mutex> lk(m_);
    auto it = free_.lower_bound(want);
    if (it != free_.end() && it->first <= want + want / 8 + (static_cast<size_t>(64) << 10)) {
      void* p = it->second;
      *got = it->first;
      cached_ -= it->first;
      free_.erase(it);
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e;
  {
    SharedCapture lk;
    e = hipMalloc(&p, want);
  }
  if (e == hipErrorOutOfMemory) {  // cached blocks may be what is missing
    (void)hipGetLastError();
    trim(0);
    SharedCapture lk;
    e = hipMalloc(&p, want);
  }
  HIP_CHECK(e);
  *got = want;
  return p;
}

void BlockCache::give(void* p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(m_);
    free_.emplace(bytes, p);
    cached_ += bytes;
    if (cached_ <= max_) return;
  }
  trim(max_);
}

void BlockCache::trim(size_t keep) {
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> lk(m_);
    while (cached_ > keep && !free_.empty()) {
      auto it = std::prev(free_.end());
      cached_ -= it->first;
      drop.push_back(it->second);
      free_.erase(it);
    }
  }
  if (drop.empty()) return;
  SharedCapture lk;
  for (void* p : drop) (void)hipFree(p);
}

size_t BlockCache::cached_bytes() {
  std::lock_guard<std::mutex> lk(m_);
  return cached_;
}

void zero_fill(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  // Allocations are at least 256-byte aligned; callers pass whole arrays.
  const uintptr_t addr = reinterpret_cast<uintptr_t>(p);
  require(addr % 16 == 0, "zero_fill: pointer must be 16-byte aligned");
  const size_t n16 = bytes / 16, ntail = bytes % 16;
  const size_t want = (n16 + 255) / 256;
  const unsigned grid = static_cast<unsigned>(std::min<size_t>(std::max<size_t>(want, 1), 8192));
  zero_kernel<<<grid, 256, 0, s>>>(static_cast<uint4*>(p), n16, static_cast<unsigned char*>(p) + n16 * 16, ntail);
  HIP_CHECK_LAUNCH();
}

void stream_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  require(dst != nullptr && src != nullptr, "stream_copy: pointers must not be null");
  require(reinterpret_cast<uintptr_t>(dst) % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0 && bytes % 16 == 0,
          "stream_copy: pointers and size must be multiples of 16 bytes");
  // One 16-byte word per thread (the fastest form measured); a launch holds at
  // most 2^24 workgroups of 256 (HIP allows fewer than 2^32 work-items per
  // dimension), so larger copies take several launches.
  const size_t n16 = bytes / 16;
  constexpr size_t kChunk = (size_t{1} << 24) * 256;
  for (size_t off = 0; off < n16; off += kChunk) {
    const size_t n = std::min(kChunk, n16 - off), blocks = (n + 255) / 256;
    stream_copy_kernel<<<static_cast<unsigned>(blocks), 256, 0, s>>>(static_cast<const double2*>(src) + off,
                                                                    static_cast<double2*>(dst) + off, n);
    HIP_CHECK_LAUNCH();
  }
}

int probe_launch(int blocks, int threads) {
  hipLaunchKernelGGL(probe_kernel, dim3(blocks), dim3(threads), 0, 0, nullptr);
  const hipError_t e = hipGetLastError();
  return static_cast<int>(e);
}

}  // namespace femhip
