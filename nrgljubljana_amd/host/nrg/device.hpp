// device.hpp - device context and device-resident matrices of the host layer.
//
// The reference keeps every matrix in host memory and writes eigenvectors and density matrices to
// disk between the two NRG sweeps (c++/eigen.hpp:389-419, c++/operators.hpp:84-111). Here eigenvector
// blocks, operator blocks and density matrices stay in HBM; host copies are made only on request.
#pragma once
#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "../../../include/nrgcore.h"
#include "traits.hpp"

namespace NRG {

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
// Blocking copies of the host layer. Each first waits for the null stream, so that the copy is issued only
// after every fill, copy and library kernel queued before it has completed. Stream order alone should
// give this; under rocprofv3's kernel tracing a strided copy (run on a DMA engine) was seen to overtake
// the fill kernel queued before it, so the order is made explicit here.
inline void dcopy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  hip_check(hipMemcpy(dst, src, bytes, kind), "hipMemcpy");
}
inline void dcopy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                    hipMemcpyKind kind) {
  hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  hip_check(hipMemcpy2D(dst, dpitch, src, spitch, width, height, kind), "hipMemcpy2D");
}
inline void nrg_check(int status, const char *what) {
  if (status != NRG_OK) throw std::runtime_error(std::string(what) + " failed: " + nrg_last_error());
}

// One library context on one device (nrg_ctx_create / nrg_ctx_destroy); calls are issued on `stream`.
class Device {
 public:
  explicit Device(int device = 0) : device_(device) {
    nrg_check(nrg_ctx_create(&ctx_, device), "nrg_ctx_create");
    hip_check(hipSetDevice(device), "hipSetDevice");
    active_ptr() = this;
  }
  ~Device() {
    if (active_ptr() == this) active_ptr() = nullptr;
    if (ctx_) nrg_ctx_destroy(ctx_);
  }
  // the most recently created live context: used by helpers that work on device matrices without one
  // being passed (the diagonal and trace helpers below)
  static Device &active() { return active_ptr() ? *active_ptr() : default_device(); }
  Device(const Device &) = delete;
  Device &operator=(const Device &) = delete;
  nrg_ctx *ctx() const { return ctx_; }
  void *stream() const { return nullptr; }   // the null stream; every library call synchronises it
  int device() const { return device_; }
  // process-wide context used by the single-call drop-ins (diagonalise, product, transform, rotate)
  static Device &default_device() {
    static Device d(0);
    return d;
  }

 private:
  static Device *&active_ptr() {
    static Device *p = nullptr;
    return p;
  }
  int device_ = 0;
  nrg_ctx *ctx_ = nullptr;
};

// Caching device allocator of the host layer. The backward sweeps allocate and free tens of density
// matrices per step, and hipFree waits for the whole device; freed blocks are kept here instead, per
// device and size class (sizes rounded up to 1/8 of their power of two, at most 12.5% slack), and
// handed out again. Up to `limit` bytes are kept; trim() returns them to the runtime (nrg_run_dir does
// at its end). Blocks are never freed from the destructor: the runtime may be gone at process exit.
class DevicePool {
 public:
  static DevicePool &get() {
    static DevicePool *p = new DevicePool();   // never destroyed, see above
    return *p;
  }
  static size_t size_class(size_t bytes) {
    if (bytes <= 4096) return 4096;
    int k = 63 - __builtin_clzll(bytes);
    const size_t step = (size_t)1 << (k - 3);
    return (bytes + step - 1) / step * step;
  }
  void *alloc(size_t cls) {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    {
      std::lock_guard<std::mutex> g(m_);
      auto it = free_.find({dev, cls});
      if (it != free_.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        cached_ -= cls;
        return p;
      }
    }
    void *p = nullptr;
    if (hipMalloc(&p, cls) != hipSuccess) {   // out of memory: give the cache back and try once more
      (void)hipGetLastError();
      trim();
      hip_check(hipMalloc(&p, cls), "hipMalloc");
    }
    return p;
  }
  void release(void *p, size_t cls) {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    {
      std::lock_guard<std::mutex> g(m_);
      if (cached_ + cls <= limit) {
        free_[{dev, cls}].push_back(p);
        cached_ += cls;
        return;
      }
    }
    (void)hipFree(p);
  }
  void trim() {
    std::lock_guard<std::mutex> g(m_);
    for (auto &[k, v] : free_)
      for (void *p : v) (void)hipFree(p);
    free_.clear();
    cached_ = 0;
  }
  size_t cached() const { return cached_; }
  size_t limit = (size_t)64 << 30;

 private:
  std::mutex m_;
  std::map<std::pair<int, size_t>, std::vector<void *>> free_;
  size_t cached_ = 0;
};

// Owning device buffer (move-only), from the DevicePool.
template <typename T>
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(size_t n) { alloc(n); }
  ~DBuf() { release(); }
  DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_), cls_(o.cls_) { o.p_ = nullptr; o.n_ = 0; o.cls_ = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_; cls_ = o.cls_;
      o.p_ = nullptr; o.n_ = 0; o.cls_ = 0;
    }
    return *this;
  }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  void alloc(size_t n) {
    release();
    if (n) {
      cls_ = DevicePool::size_class(n * sizeof(T));
      p_ = (T *)DevicePool::get().alloc(cls_);
    }
    n_ = n;
  }
  void zero() {
    if (n_) hip_check(hipMemset(p_, 0, n_ * sizeof(T)), "hipMemset");
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) DevicePool::get().release(p_, cls_);
    p_ = nullptr;
    n_ = 0;
    cls_ = 0;
  }
  T *p_ = nullptr;
  size_t n_ = 0, cls_ = 0;
};

// Non-owning view of a row-major device matrix: element (i, j) at p[i * ld + j].
struct DView {
  double *p = nullptr;
  size_t rows = 0, cols = 0, ld = 0;
  bool empty() const { return rows == 0 || cols == 0; }
  // rows [r0, r1) x columns [c0, c1) of this view
  DView sub(std::pair<size_t, size_t> r, std::pair<size_t, size_t> c) const {
    NRG_ASSERT(r.first <= r.second && r.second <= rows && c.first <= c.second && c.second <= cols);
    return DView{p + r.first * ld + c.first, r.second - r.first, c.second - c.first, ld};
  }
};

inline Matrix download(const DView &v) {
  Matrix m(v.rows, v.cols);
  if (!v.empty())
    dcopy2d(m.data(), v.cols * sizeof(double), v.p, v.ld * sizeof(double), v.cols * sizeof(double),
                          v.rows, hipMemcpyDeviceToHost);
  return m;
}
inline void upload(const DView &v, const Matrix &m) {
  NRG_ASSERT(m.size1() == v.rows && m.size2() == v.cols);
  if (!v.empty())
    dcopy2d(v.p, v.ld * sizeof(double), m.data(), v.cols * sizeof(double), v.cols * sizeof(double),
                          v.rows, hipMemcpyHostToDevice);
}
// Diagonal entries of many device matrices in one library call (nrg_diagonal_batched): items[k]
// covers entries [first, first + count) of matrix A (leading dimension ld) at values[off ..].
inline void get_diagonals(const std::vector<nrg_diag_item> &items, std::vector<double> &values) {
  nrg_check(nrg_diagonal_batched(Device::active().ctx(), NRG_DIAG_GET, items.data(), (int)items.size(), values.data(),
                                 Device::active().stream()),
            "nrg_diagonal_batched");
}
inline void set_diagonals(const std::vector<nrg_diag_item> &items, std::vector<double> &values) {
  nrg_check(nrg_diagonal_batched(Device::active().ctx(), NRG_DIAG_SET, items.data(), (int)items.size(), values.data(),
                                 Device::active().stream()),
            "nrg_diagonal_batched");
}
// traces of square device matrices, each summed in index order as the reference's trace loop
inline std::vector<double> traces(const std::vector<DView> &ms) {
  std::vector<nrg_diag_item> items;
  int64_t off = 0;
  std::vector<size_t> owner;   // matrix of each item (empty matrices have none)
  for (size_t k = 0; k < ms.size(); k++) {
    const int64_t n = (int64_t)std::min(ms[k].rows, ms[k].cols);
    if (n == 0) continue;
    items.push_back(nrg_diag_item{ms[k].p, (int64_t)ms[k].ld, 0, n, off});
    owner.push_back(k);
    off += n;
  }
  std::vector<double> d((size_t)off + 1), t(ms.size(), 0.0);
  if (off) get_diagonals(items, d);
  for (size_t j = 0; j < items.size(); j++)
    for (int64_t i = 0; i < items[j].count; i++) t[owner[j]] += d[(size_t)(items[j].off + i)];
  return t;
}
// sum of the diagonal of a square device matrix
inline double trace_real(const DView &v) { return traces({v})[0]; }

// Owning row-major device matrix (contiguous, ld = cols), zero-initialised.
class DMatrix {
 public:
  DMatrix() = default;
  DMatrix(size_t r, size_t c) : buf_(r * c), r_(r), c_(c) { buf_.zero(); }
  explicit DMatrix(const Matrix &m) : DMatrix(m.size1(), m.size2()) { upload(view(), m); }
  DView view() const { return DView{buf_.get(), r_, c_, c_}; }
  Matrix host() const { return download(view()); }
  size_t size1() const { return r_; }
  size_t size2() const { return c_; }
  double *data() const { return buf_.get(); }

 private:
  DBuf<double> buf_;
  size_t r_ = 0, c_ = 0;
};
inline size_t size1(const DMatrix &m) { return m.size1(); }
inline size_t size2(const DMatrix &m) { return m.size2(); }
inline double trace_real(const DMatrix &m) { return trace_real(m.view()); }

}  // namespace NRG
