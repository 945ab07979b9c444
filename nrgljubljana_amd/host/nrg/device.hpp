// device.hpp - device context and device-resident matrices of the host layer.
//
// The reference keeps every matrix in host memory and writes eigenvectors and density matrices to
// disk between the two NRG sweeps (c++/eigen.hpp:389-419, c++/operators.hpp:84-111). Here eigenvector
// blocks, operator blocks and density matrices stay in HBM; host copies are made only on request.
#pragma once
#include <hip/hip_runtime_api.h>

#include <memory>
#include <utility>

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
  }
  ~Device() {
    if (ctx_) nrg_ctx_destroy(ctx_);
  }
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
  int device_ = 0;
  nrg_ctx *ctx_ = nullptr;
};

// Owning device buffer (move-only).
template <typename T>
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(size_t n) { alloc(n); }
  ~DBuf() { release(); }
  DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  void alloc(size_t n) {
    release();
    if (n) hip_check(hipMalloc((void **)&p_, n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void zero() {
    if (n_) hip_check(hipMemset(p_, 0, n_ * sizeof(T)), "hipMemset");
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T *p_ = nullptr;
  size_t n_ = 0;
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
// sum of the diagonal of a square device matrix
inline double trace_real(const DView &v) {
  const size_t n = std::min(v.rows, v.cols);
  if (!n) return 0.0;
  std::vector<double> d(n);
  dcopy2d(d.data(), sizeof(double), v.p, (v.ld + 1) * sizeof(double), sizeof(double), n,
                        hipMemcpyDeviceToHost);
  double t = 0;
  for (double x : d) t += x;
  return t;
}

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
