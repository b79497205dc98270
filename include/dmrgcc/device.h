// Plumbing of the drop-in classes: status checks and device buffers over the
// C ABI of include/dmrgcc_hip.h.  A failing library call prints the library's
// message and exits with -1, as the reference's error() does
// (src/util/general.h:106-111).
#ifndef DMRGCC_DROPIN_DEVICE_H_
#define DMRGCC_DROPIN_DEVICE_H_

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../dmrgcc_hip.h"

namespace dmrgcc_dropin {

inline void check(int rc) {
  if (rc != 0) {
    std::printf("%s\n", dmrgcc_last_error());
    std::exit(-1);
  }
}

// Device memory of n doubles from the library's pool.
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n) { resize(n); }
  ~DeviceVector() { dmrgcc_free(p_); }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  void resize(int64_t n) {
    if (n == n_) return;
    dmrgcc_free(p_);
    p_ = nullptr;
    n_ = n;
    void* q = nullptr;
    check(dmrgcc_alloc(8 * (n > 0 ? n : 1), &q));
    p_ = static_cast<double*>(q);
  }
  void upload(const double* h, int64_t n) {
    resize(n);
    check(dmrgcc_upload(p_, h, 8 * n, nullptr));
  }
  void download(double* h, int64_t n) const { check(dmrgcc_download(h, p_, 8 * n, nullptr)); }
  double* data() const { return p_; }
  int64_t size() const { return n_; }

 private:
  double* p_ = nullptr;
  int64_t n_ = -1;
};

// A block list flattened for the C ABI: per block (channel, l, r, rows,
// cols, element offset) and the concatenated data.
struct FlatBlocks {
  std::vector<int> o, l, r, rows, cols;
  std::vector<int64_t> off;
  std::vector<double> data;
  int n() const { return (int)o.size(); }
};

}  // namespace dmrgcc_dropin

#endif  // DMRGCC_DROPIN_DEVICE_H_
