// traits.hpp - host matrix type, quantum-number labels and numeric helpers of the host layer.
//
// The reference works on row-major ublas matrices (c++/traits.hpp:41) with free functions size1/size2/
// nrvec/dim (c++/numerics.hpp) and labels subspaces by Invar (c++/invar.hpp). This layer keeps those
// names and meanings for real (double) matrices.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace NRG {

// Row-major dense matrix of doubles (the reference's Matrix_traits<double>).
class Matrix {
 public:
  Matrix() = default;
  Matrix(size_t r, size_t c) : r_(r), c_(c), a_(r * c, 0.0) {}
  double &operator()(size_t i, size_t j) { return a_[i * c_ + j]; }
  double operator()(size_t i, size_t j) const { return a_[i * c_ + j]; }
  size_t size1() const { return r_; }
  size_t size2() const { return c_; }
  double *data() { return a_.data(); }
  const double *data() const { return a_.data(); }
  void resize(size_t r, size_t c) {   // non-conserving
    r_ = r;
    c_ = c;
    a_.assign(r * c, 0.0);
  }

 private:
  size_t r_ = 0, c_ = 0;
  std::vector<double> a_;
};

inline size_t size1(const Matrix &m) { return m.size1(); }
inline size_t size2(const Matrix &m) { return m.size2(); }
inline size_t nrvec(const Matrix &m) { return m.size1(); }   // eigenvectors are rows
inline size_t dim(const Matrix &m) { return m.size2(); }
inline bool finite_size(const Matrix &m) { return m.size1() > 0 && m.size2() > 0; }
inline bool is_square(const Matrix &m) { return m.size1() == m.size2(); }
inline Matrix zero_matrix(size_t r, size_t c) { return Matrix(r, c); }
inline Matrix zero_matrix(size_t n) { return Matrix(n, n); }
inline Matrix id_matrix(size_t n) {
  Matrix m(n, n);
  for (size_t i = 0; i < n; i++) m(i, i) = 1.0;
  return m;
}
inline double trace_real(const Matrix &m) {
  double t = 0;
  for (size_t i = 0; i < std::min(m.size1(), m.size2()); i++) t += m(i, i);
  return t;
}

// Invariant-subspace label: a tuple of integer quantum numbers (c++/invar.hpp), ordered lexicographically.
struct Invar {
  std::vector<int> qn;
  Invar() = default;
  Invar(std::initializer_list<int> q) : qn(q) {}
  explicit Invar(std::vector<int> q) : qn(std::move(q)) {}
  bool operator<(const Invar &o) const { return qn < o.qn; }
  bool operator==(const Invar &o) const { return qn == o.qn; }
  bool operator!=(const Invar &o) const { return qn != o.qn; }
  std::string str() const {
    std::string s;
    for (size_t i = 0; i < qn.size(); i++) s += (i ? " " : "") + std::to_string(qn[i]);
    return s;
  }
};
using Twoinvar = std::pair<Invar, Invar>;

// my_fcmp / num_equal (c++/numerics.hpp:107-120): equal if both tiny or within a relative epsilon
inline int my_fcmp(double x, double y, double small_epsilon, double rel_epsilon) {
  if (x == 0.0 && y == 0.0) return 0;
  if (std::abs(x) < small_epsilon && std::abs(y) < small_epsilon) return 0;
  if (std::abs(x - y) < rel_epsilon * (std::abs(x) + std::abs(y))) return 0;
  return x > y ? 1 : (x < y ? -1 : 0);
}
inline bool num_equal(double a, double b, double check_precision = 1e-12) {
  return my_fcmp(a, b, check_precision, check_precision) == 0;
}

// my_assert (c++/portabil.hpp): the reference aborts; the host layer throws so that callers can recover
#define NRG_ASSERT(cond)                                                                                  \
  do {                                                                                                    \
    if (!(cond)) throw std::runtime_error(std::string("assertion failed: ") + #cond + " at " + __FILE__ + \
                                          ":" + std::to_string(__LINE__));                                \
  } while (0)

}  // namespace NRG
