// Caching device allocator and error plumbing.
#include "common.h"

#include <algorithm>
#include <cstring>

namespace dmrg {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& s) { g_last_error = s; }
const char* last_error() { return g_last_error.c_str(); }

namespace {
// Pinned staging arena of upload_bytes: regions are handed out in order;
// when the arena is exhausted every copy issued from it is waited for and
// it starts over.
struct Staging {
  std::mutex mu;
  char* buf = nullptr;
  size_t cap = 0, head = 0;
  std::vector<std::pair<hipEvent_t, hipStream_t>> pending;  // copy done event, its stream
  std::vector<hipEvent_t> spare;
};
constexpr size_t kStagingBytes = size_t(32) << 20;
constexpr size_t kMaxPending = 4096;

Staging& staging() {
  static Staging* s = new Staging();  // leaked, as the device pool
  return *s;
}
}  // namespace

void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return;
  // a copy captured into a graph would re-read the staging arena at every
  // replay, after the arena has been reused
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  DMRG_HIP(hipStreamIsCapturing(stream, &cap));
  if (cap != hipStreamCaptureStatusNone)
    throw Error(kErrInternal, "descriptor upload on a stream that is being captured");
  Staging& st = staging();
  std::lock_guard<std::mutex> g(st.mu);
  if (!st.buf) {
    DMRG_HIP(hipHostMalloc((void**)&st.buf, kStagingBytes, hipHostMallocDefault));
    st.cap = kStagingBytes;
  }
  if (bytes > st.cap / 4) {
    DMRG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    DMRG_HIP(hipStreamSynchronize(stream));
    return;
  }
  const size_t need = (bytes + 255) & ~size_t(255);
  if (st.head + need > st.cap || st.pending.size() >= kMaxPending) {
    bool all = true;
    for (const auto& pe : st.pending) {
      if (hipEventSynchronize(pe.first) != hipSuccess) all = false;
      st.spare.push_back(pe.first);
    }
    if (!all) {
      // an event whose stream was destroyed by its owner cannot be waited
      // on; wait for the whole device instead before the arena is reused
      (void)hipGetLastError();
      DMRG_HIP(hipDeviceSynchronize());
    }
    st.pending.clear();
    st.head = 0;
  }
  char* p = st.buf + st.head;
  st.head += need;
  std::memcpy(p, src, bytes);
  DMRG_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, stream));
  hipEvent_t e;
  if (st.spare.empty()) {
    DMRG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  } else {
    e = st.spare.back();
    st.spare.pop_back();
  }
  st.pending.push_back({e, stream});
  DMRG_HIP(hipEventRecord(e, stream));
}

void upload_forget_stream(hipStream_t stream) {
  Staging& st = staging();
  std::lock_guard<std::mutex> g(st.mu);
  std::vector<std::pair<hipEvent_t, hipStream_t>> keep;
  for (const auto& pe : st.pending) {
    if (pe.second == stream) {
      (void)hipEventSynchronize(pe.first);
      st.spare.push_back(pe.first);
    } else {
      keep.push_back(pe);
    }
  }
  st.pending.swap(keep);
}

DevicePool& DevicePool::get() {
  // Intentionally leaked: buffers may be released by static destructors
  // after the HIP runtime itself has begun tearing down.
  static DevicePool* pool = new DevicePool();
  return *pool;
}

size_t DevicePool::round(size_t b) {
  if (b <= 4096) return (b + 255) & ~size_t(255);
  size_t r = 4096;
  while (r < b) r <<= 1;
  return r;
}

void* DevicePool::alloc(size_t bytes) {
  const size_t r = round(std::max<size_t>(bytes, 1));
  std::lock_guard<std::mutex> g(mu_);
  auto it = free_.find(r);
  if (it != free_.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    cached_ -= r;
    in_use_ += r;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, r);
  if (e != hipSuccess) {
    // release the cache and retry once
    for (auto& kv : free_)
      for (void* q : kv.second) {
        (void)hipFree(q);
        size_of_.erase(q);
      }
    free_.clear();
    cached_ = 0;
    (void)hipGetLastError();
    e = hipMalloc(&p, r);
    if (e != hipSuccess)
      throw Error(kErrHip, std::string("hipMalloc(") + std::to_string(r) + "): " + hipGetErrorString(e));
  }
  size_of_[p] = r;
  in_use_ += r;
  return p;
}

void DevicePool::release(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = size_of_.find(p);
  if (it == size_of_.end()) return;
  free_[it->second].push_back(p);
  in_use_ -= it->second;
  cached_ += it->second;
}

void DevicePool::trim() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& kv : free_)
    for (void* q : kv.second) {
      (void)hipFree(q);
      size_of_.erase(q);
    }
  free_.clear();
  cached_ = 0;
}

}  // namespace dmrg
