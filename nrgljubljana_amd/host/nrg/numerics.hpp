// numerics.hpp - product / transform / rotate (c++/numerics_ublas.hpp:108-140) on the GPU.
//
// Gemm3Batch collects accumulate-terms over device matrices and evaluates them in one
// nrg_gemm3_grouped call, in the reference's operation order:
//   product   M += f A B^T                     (transA 0, no middle factor, B^T)
//   transform M += f A (O B^T)                 (T = O B^T first, then M += f A T)
//   rotate    M += f (U^T O) U                 (T = U^T O first, then M += f T U)
// Terms whose A or B (U for rotate) is empty are skipped, as finite_size() does in the reference.
// The host-matrix overloads are drop-ins for single calls (copy in, one grouped call, copy out).
#pragma once
#include "device.hpp"

namespace NRG {

class Gemm3Batch {
 public:
  explicit Gemm3Batch(Device &dev) : dev_(dev) {}
  // M += f A B^T; A: m x k, B: n x k, M: m x n
  void product(const DView &M, double f, const DView &A, const DView &B) {
    if (A.empty() || B.empty()) return;
    NRG_ASSERT(M.rows == A.rows && A.cols == B.cols && B.rows == M.cols && std::isfinite(f));
    nrg_gemm3_term t{};
    t.m = (int64_t)A.rows; t.n = (int64_t)B.rows; t.k = (int64_t)A.cols; t.l = (int64_t)A.cols; t.f = f;
    t.A = A.p; t.lda = (int64_t)A.ld; t.transA = 0;
    t.O = nullptr; t.ldo = 0;
    t.B = B.p; t.ldb = (int64_t)B.ld; t.transB = 0;   // opB = B^T
    t.C = M.p; t.ldc = (int64_t)M.ld;
    terms_.push_back(t);
  }
  // M += f A O B^T; A: m x k, O: k x l, B: n x l
  void transform(const DView &M, double f, const DView &A, const DView &O, const DView &B) {
    if (A.empty() || B.empty()) return;
    NRG_ASSERT(M.rows == A.rows && A.cols == O.rows && O.cols == B.cols && B.rows == M.cols && std::isfinite(f));
    nrg_gemm3_term t{};
    t.m = (int64_t)A.rows; t.n = (int64_t)B.rows; t.k = (int64_t)A.cols; t.l = (int64_t)O.cols; t.f = f;
    t.A = A.p; t.lda = (int64_t)A.ld; t.transA = 0;
    t.O = O.p; t.ldo = (int64_t)O.ld;
    t.B = B.p; t.ldb = (int64_t)B.ld; t.transB = 0;   // opB = B^T
    t.C = M.p; t.ldc = (int64_t)M.ld;
    terms_.push_back(t);
  }
  // M += f U^T O U; U: r x c, O: r x r, M: c x c
  void rotate(const DView &M, double f, const DView &U, const DView &O) {
    if (U.empty()) return;
    NRG_ASSERT(M.rows == U.cols && U.rows == O.rows && O.cols == U.rows && U.cols == M.cols && std::isfinite(f));
    nrg_gemm3_term t{};
    t.m = (int64_t)U.cols; t.n = (int64_t)U.cols; t.k = (int64_t)U.rows; t.l = (int64_t)U.rows; t.f = f;
    t.A = U.p; t.lda = (int64_t)U.ld; t.transA = 1;
    t.O = O.p; t.ldo = (int64_t)O.ld;
    t.B = U.p; t.ldb = (int64_t)U.ld; t.transB = 1;   // opB = U
    t.C = M.p; t.ldc = (int64_t)M.ld;
    terms_.push_back(t);
  }
  size_t size() const { return terms_.size(); }
  // Evaluate all queued terms (precision: NRG_GEMM_FP64 or NRG_GEMM_OZAKI) and clear the queue.
  void flush(int precision = NRG_GEMM_FP64) {
    if (!terms_.empty())
      nrg_check(nrg_gemm3_grouped(dev_.ctx(), terms_.data(), (int)terms_.size(), precision, dev_.stream()),
                "nrg_gemm3_grouped");
    terms_.clear();
  }

 private:
  Device &dev_;
  std::vector<nrg_gemm3_term> terms_;
};

// Single-call drop-ins on host matrices.
inline void product(Matrix &M, double factor, const Matrix &A, const Matrix &B) {
  if (!finite_size(A) || !finite_size(B)) return;
  Device &d = Device::default_device();
  DMatrix dM(M), dA(A), dB(B);
  Gemm3Batch g(d);
  g.product(dM.view(), factor, dA.view(), dB.view());
  g.flush();
  M = dM.host();
}
inline void transform(Matrix &M, double factor, const Matrix &A, const Matrix &O, const Matrix &B) {
  if (!finite_size(A) || !finite_size(B)) return;
  Device &d = Device::default_device();
  DMatrix dM(M), dA(A), dO(O), dB(B);
  Gemm3Batch g(d);
  g.transform(dM.view(), factor, dA.view(), dO.view(), dB.view());
  g.flush();
  M = dM.host();
}
inline void rotate(Matrix &M, double factor, const Matrix &U, const Matrix &O) {
  if (!finite_size(U)) return;
  Device &d = Device::default_device();
  DMatrix dM(M), dU(U), dO(O);
  Gemm3Batch g(d);
  g.rotate(dM.view(), factor, dU.view(), dO.view());
  g.flush();
  M = dM.host();
}

}  // namespace NRG
