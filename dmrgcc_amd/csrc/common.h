// Shared runtime pieces of the MI355X DMRG library: status codes, HIP error
// checks, and a caching device allocator.  Host code is plain C++17; device
// code lives in the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmrg {

// Status codes returned through the C ABI (0 = success).
enum Status : int {
  kOk = 0,
  kErrHip = 1,        // a HIP runtime call failed
  kErrShape = 2,      // inconsistent dimensions / indices (reference: print + exit)
  kErrArgument = 3,   // invalid argument (null pointer, out-of-range enum)
  kErrConvergence = 4,// iterative solver did not converge
  kErrInternal = 5,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DMRG_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::dmrg::Error(::dmrg::kErrHip, std::string(#expr) + ": " +        \
                                               hipGetErrorString(e_));        \
  } while (0)

inline void require(bool ok, const std::string& msg, int code = kErrShape) {
  if (!ok) throw Error(code, msg);
}

// Thread-local last error text for the C ABI.
void set_last_error(const std::string& s);
const char* last_error();

// ---------------------------------------------------------------------------
// Caching device allocator.  hipMalloc/hipFree synchronise the device and
// cost tens of microseconds; a sweep allocates and frees thousands of block
// buffers, so freed buffers are kept in size-class free lists and reused.
// Size classes are powers of two above 4 KiB rounded to 256 B below it.
// ---------------------------------------------------------------------------
class DevicePool {
 public:
  static DevicePool& get();
  void* alloc(size_t bytes);
  void release(void* p);
  size_t bytes_in_use() const { return in_use_; }
  size_t bytes_cached() const { return cached_; }
  void trim();  // return every cached buffer to the driver

 private:
  static size_t round(size_t b);
  std::mutex mu_;
  std::map<size_t, std::vector<void*>> free_;
  std::map<void*, size_t> size_of_;
  size_t in_use_ = 0, cached_ = 0;
};

// RAII device buffer of elements of type T.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  ~DevBuf() { reset(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void resize(size_t n) {
    reset();
    n_ = n;
    if (n) p_ = static_cast<T*>(DevicePool::get().alloc(n * sizeof(T)));
  }
  // Grow without preserving contents.
  void ensure(size_t n) { if (n > n_) resize(n); }
  void reset() {
    if (p_) DevicePool::get().release(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* data() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Asynchronous host-to-device copy of `bytes` from `src` (any host memory,
// reusable as soon as this returns): the bytes are copied into a pinned
// staging arena and a DMA copy is queued on `stream`, ordered with the
// stream's other work; the host does not wait for the stream.  A staging
// region is reused only after the copy that read it has completed (an event
// per copy).  Copies larger than a quarter of the arena go through a
// pageable copy and a stream synchronisation instead.  Not for use inside a
// stream capture.
void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t stream);
// Wait for the staging copies issued on `stream` and stop tracking them; call
// before `stream` is destroyed (an event is not waited on after its stream
// is gone).
void upload_forget_stream(hipStream_t stream);

// Upload a host vector into a (re)sized device buffer on `stream` (see
// upload_bytes: the host vector may be reused once this returns).
template <typename T>
void upload(DevBuf<T>& dst, const std::vector<T>& src, hipStream_t stream) {
  dst.ensure(src.size());
  if (!src.empty()) upload_bytes(dst.data(), src.data(), src.size() * sizeof(T), stream);
}

}  // namespace dmrg
