// eigen.hpp - eigenvalue/eigenvector containers of one NRG step (c++/eigen.hpp:30-420).
//
// Same names and meanings as the reference: Values (raw/rel/abs/abs_zero/abs_T/abs_G/corr energies,
// 30-98), Vectors (eigenvectors as rows, M x dim, 100-135), Blocks (the per-ancestor column blocks U(i)
// used by recalc and the density-matrix rotation, 139-177), RawEigen (180-199), Eigen (K/D ranges for
// FDM, truncation, diagonal_exp, 202-300) and DiagInfo (map Invar -> Eigen, 303-420). The difference:
// eigenvectors live in device memory and U(i) blocks are views into them (no copies, no unitary files).
#pragma once
#include <functional>
#include <limits>
#include <map>
#include <memory>

#include "device.hpp"

namespace NRG {

class Values {
 public:
  void resize(size_t size) { v.resize(size); }
  double raw(size_t i) const { return v[i]; }
  double rel(size_t i) const { return v[i]; }
  double abs(size_t i) const { NRG_ASSERT(std::isfinite(scale)); return rel(i) * scale; }
  double rel_zero(size_t i) const { NRG_ASSERT(std::isfinite(shift)); return rel(i) - shift; }
  double abs_zero(size_t i) const { return (rel(i) - shift) * scale; }
  double abs_T(size_t i) const { NRG_ASSERT(std::isfinite(T_shift)); return abs_zero(i) + T_shift; }
  double abs_G(size_t i) const { NRG_ASSERT(std::isfinite(abs_GS_energy)); return abs_T(i) - abs_GS_energy; }
  double corr(size_t i) const { return corrected[i]; }
  size_t size() const { return v.size(); }
  double lowest_rel() const { return v.front(); }
  double highest_corr() const { return corrected.back(); }
  const std::vector<double> &all_rel() const { return v; }
  std::vector<double> all_rel_zero() const { return map([this](double x) { return x - shift; }); }
  std::vector<double> all_abs_zero() const { return map([this](double x) { return (x - shift) * scale; }); }
  std::vector<double> all_abs_T() const { return map([this](double x) { return (x - shift) * scale + T_shift; }); }
  std::vector<double> all_abs_G() const {
    return map([this](double x) { return (x - shift) * scale + T_shift - abs_GS_energy; });
  }
  const std::vector<double> &all_corr() const { return corrected; }
  void set(std::vector<double> in) { v = std::move(in); }
  void set_scale(double s) { scale = s; }
  void set_shift(double s) { shift = s; }
  void set_T_shift(double s) { T_shift = s; }
  void set_abs_GS_energy(double e) { abs_GS_energy = e; }
  void set_corr(std::vector<double> in) { corrected = std::move(in); }
  bool has_abs() const { return std::isfinite(scale); }
  bool has_zero() const { return std::isfinite(shift); }
  bool has_abs_zero() const { return has_abs() && has_zero(); }
  bool has_abs_T() const { return has_abs_zero() && std::isfinite(T_shift); }
  bool has_abs_G() const { return has_abs_T() && std::isfinite(abs_GS_energy); }
  bool has_corr() const { return !corrected.empty(); }

 private:
  template <typename F> std::vector<double> map(F f) const {
    std::vector<double> r(v.size());
    for (size_t i = 0; i < v.size(); i++) r[i] = f(v[i]);
    return r;
  }
  static constexpr double nan = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> v;
  double scale = nan, shift = nan, T_shift = nan, abs_GS_energy = nan;
  std::vector<double> corrected;
};

// Eigenvectors as rows of an M x dim device matrix. `owner` keeps the device storage alive (the
// batch arena of diagonalisations(), or a matrix of its own).
class Vectors {
 public:
  size_t M() const { return v_.rows; }
  size_t dim() const { return v_.cols; }
  void set(const Matrix &m) {
    auto own = std::make_shared<DMatrix>(m);
    v_ = own->view();
    owner_ = own;
  }
  void set_device(DView v, std::shared_ptr<void> owner) {
    NRG_ASSERT(v.rows <= v.cols);
    v_ = v;
    owner_ = std::move(owner);
  }
  Matrix get() const { return download(v_); }   // host copy
  const DView &dev() const { return v_; }
  void standard_basis(size_t size) { set(id_matrix(size)); }
  DView submatrix_const(std::pair<size_t, size_t> r1, std::pair<size_t, size_t> r2) const { return v_.sub(r1, r2); }
  void shrink() { v_.rows = 0; }   // keeps the dimension

 private:
  DView v_;
  std::shared_ptr<void> owner_;
};

// U(i): the columns of the eigenvectors that belong to ancestor subspace i (views, 1-based operator()).
class Blocks {
 public:
  void resize(size_t nr) { blocks.resize(nr); }
  void set(size_t i, DView m) { NRG_ASSERT(i < blocks.size()); blocks[i] = m; }
  const DView &get(size_t i) const { NRG_ASSERT(i < blocks.size()); return blocks[i]; }
  const DView &operator()(size_t i) const { NRG_ASSERT(1 <= i && i <= blocks.size()); return blocks[i - 1]; }
  size_t size() const { return blocks.size(); }
  void truncate(size_t nr) {
    for (auto &b : blocks) {
      NRG_ASSERT(nr <= b.rows);
      b.rows = nr;
    }
  }
  void clear() { blocks.assign(blocks.size(), DView{}); }

 private:
  std::vector<DView> blocks;
};

struct RawEigen {
  std::vector<double> val;
  Matrix vec;
  RawEigen() = default;
  RawEigen(size_t M, size_t dim) : val(M), vec(M, dim) { NRG_ASSERT(M <= dim); }
  size_t getnrcomputed() const { NRG_ASSERT(val.size() == nrvec(vec)); return val.size(); }
  size_t getdim() const { return NRG::dim(vec); }
  void check_diag() const;   // diag.hpp
};

class Eigen {
 public:
  Values values;
  Vectors vectors;
  Blocks U;
  Eigen() = default;
  Eigen(RawEigen &&raw, bool last_step) : last(last_step) {
    values.set(std::move(raw.val));
    vectors.set(raw.vec);
  }
  // eigenvalues from the batched solver, eigenvectors already resident on the device
  Eigen(std::vector<double> val, DView vec, std::shared_ptr<void> owner, bool last_step) : last(last_step) {
    values.set(std::move(val));
    vectors.set_device(vec, std::move(owner));
  }
  size_t getnrcomputed() const { return values.size(); }
  size_t getdim() const { return vectors.dim(); }
  size_t getnrall() const { return getnrcomputed(); }
  size_t getnrkept() const { return getnrpost(); }
  size_t getnrdiscarded() const { return getnrcomputed() - getnrpost(); }
  size_t getnrstored() const { return nrstored == -1 ? values.size() : (size_t)nrstored; }
  // FDM ranges: in the last step every state counts as discarded (c++/eigen.hpp:234,245-246)
  std::pair<size_t, size_t> Drange() const { return {boundary(), getnrall()}; }
  std::pair<size_t, size_t> Krange() const { return {0, boundary()}; }
  bool is_last() const { return last; }
  void truncate_prepare(size_t nrpost_) {
    nrpost = (long)nrpost_;
    NRG_ASSERT(nrpost_ <= getnrcomputed());
  }
  void truncate_perform() {
    nrstored = nrpost;
    U.truncate(getnrpost());
  }
  void diagonal(const std::vector<double> &v) {
    values.set(v);
    vectors.standard_basis(v.size());
  }
  void subtract_Egs(double Egs) { values.set_shift(Egs); }
  void subtract_GS_energy(double GS_energy) { values.set_abs_GS_energy(GS_energy); }
  // diagonal matrix exp(-factor * E_corr) over the stored states (c++/eigen.hpp:270-276)
  Matrix diagonal_exp(double factor) const {
    const size_t d = getnrstored();
    Matrix m(d, d);
    for (size_t i = 0; i < d; i++) m(i, i) = std::exp(-values.corr(i) * factor);
    return m;
  }
  // Tr[fnc(factor E) exp(-factor E)] over rel_zero energies (c++/eigen.hpp:277-280)
  double trace(const std::function<double(double)> &fnc, double factor) const {
    double s = 0.0;
    for (double x : values.all_rel_zero()) s += fnc(factor * x) * std::exp(-factor * x);
    return s;
  }
  void clear_eigenvectors() {
    vectors.shrink();
    U.clear();
  }

 private:
  long nrpost = -1, nrstored = -1;
  bool last = false;
  size_t getnrpost() const { return nrpost == -1 ? getnrcomputed() : (size_t)nrpost; }
  size_t boundary() const { return last ? 0 : getnrkept(); }
};

using MultFnc = std::function<double(const Invar &)>;

class DiagInfo : public std::map<Invar, Eigen> {
 public:
  double find_groundstate() const {
    NRG_ASSERT(!empty());
    double e = std::numeric_limits<double>::infinity();
    for (const auto &[I, eig] : *this)
      if (eig.values.size()) e = std::min(e, eig.values.lowest_rel());
    return e;
  }
  void subtract_Egs(double Egs) { for (auto &kv : *this) kv.second.subtract_Egs(Egs); }
  double Egs_subtraction() {
    const double Egs = find_groundstate();
    subtract_Egs(Egs);
    return Egs;
  }
  void subtract_GS_energy(double GS_energy) { for (auto &kv : *this) kv.second.subtract_GS_energy(GS_energy); }
  std::vector<double> sorted_energies_rel_zero() const {
    std::vector<double> e;
    for (const auto &kv : *this) {
      const auto z = kv.second.values.all_rel_zero();
      e.insert(e.end(), z.begin(), z.end());
    }
    std::sort(e.begin(), e.end());
    return e;
  }
  std::vector<double> sorted_energies_corr() const {
    std::vector<double> e;
    for (const auto &kv : *this) e.insert(e.end(), kv.second.values.all_corr().begin(), kv.second.values.all_corr().end());
    std::sort(e.begin(), e.end());
    return e;
  }
  void truncate_perform() { for (auto &kv : *this) kv.second.truncate_perform(); }
  size_t size_subspace(const Invar &I) const {
    const auto f = find(I);
    return f != end() ? f->second.getnrstored() : 0;
  }
  std::pair<const Eigen &, const Eigen &> subs(const Invar &I1, const Invar &I2) const { return {at(I1), at(I2)}; }
  std::pair<size_t, size_t> dims(const Invar &I1, const Invar &I2) const { return {size_subspace(I1), size_subspace(I2)}; }
  void clear_eigenvectors() { for (auto &kv : *this) kv.second.clear_eigenvectors(); }
  double count_states(const MultFnc &mult) const {
    double s = 0;
    for (const auto &[I, eig] : *this) s += mult(I) * eig.getnrstored();
    return s;
  }
  size_t count_subspaces() const {
    size_t c = 0;
    for (const auto &kv : *this) c += kv.second.getnrstored() > 0;
    return c;
  }
  double trace(const std::function<double(double)> &fnc, double factor, const MultFnc &mult) const {
    double s = 0;
    for (const auto &[I, eig] : *this) s += mult(I) * eig.trace(fnc, factor);
    return s;
  }
};

}  // namespace NRG
