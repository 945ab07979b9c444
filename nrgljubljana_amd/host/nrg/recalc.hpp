// recalc.hpp - recalculation of irreducible matrix elements (c++/recalc.hpp:18-122).
//
// split_in_blocks_Eigen makes U(i) the columns part(i) of the stored eigenvectors; here they are
// views into the device eigenvector matrix instead of copies. recalc_f and recalc_general queue one
// product / transform term per table row into a Gemm3Batch: the caller queues every (I1, Ip) of an
// operator (the batching point of Symmetry::recalc_irreduc / recalc_doublet) and flushes once.
#pragma once
#include <map>

#include "eigen.hpp"
#include "numerics.hpp"

namespace NRG {

// Dimensions of the ancestor subspaces that make up a subspace (c++/subspaces.hpp:18-80).
class SubspaceDimensions {
 public:
  SubspaceDimensions() = default;
  SubspaceDimensions(std::vector<size_t> dims, std::vector<Invar> ancestors)
      : dims_(std::move(dims)), ancestors_(std::move(ancestors)) {
    NRG_ASSERT(dims_.size() == ancestors_.size());
  }
  size_t combs() const { return dims_.size(); }
  size_t rmax(size_t i) const { NRG_ASSERT(i < combs()); return dims_[i]; }
  bool exists(size_t i) const { return rmax(i) > 0; }
  size_t offset(size_t i) const {
    NRG_ASSERT(i < combs());
    size_t o = 0;
    for (size_t k = 0; k < i; k++) o += dims_[k];
    return o;
  }
  std::pair<size_t, size_t> part(size_t i) const { return {offset(i), offset(i) + rmax(i)}; }
  size_t total() const {
    size_t t = 0;
    for (size_t d : dims_) t += d;
    return t;
  }
  Invar ancestor(size_t i) const { return ancestors_[i]; }

 private:
  std::vector<size_t> dims_;
  std::vector<Invar> ancestors_;
};
using SubspaceStructure = std::map<Invar, SubspaceDimensions>;

// c++/recalc.hpp:18-30 (U(i) = vectors[0:nr, part(i)], nr = stored eigenpairs), as views
inline void split_in_blocks_Eigen(Eigen &e, const SubspaceDimensions &sub) {
  const size_t combs = sub.combs(), nr = e.getnrstored();
  NRG_ASSERT(0 < nr && nr <= e.getdim() && sub.total() == e.getdim());
  e.U.resize(combs);
  for (size_t b = 0; b < combs; b++) e.U.set(b, e.vectors.submatrix_const({0, nr}, sub.part(b)));
}
inline void split_in_blocks(DiagInfo &diag, const SubspaceStructure &substruct) {
  for (auto &[I, eig] : diag) split_in_blocks_Eigen(eig, substruct.at(I));
}

// Table rows (c++/recalc.hpp: Recalc_f {i1, ip, factor}, Recalc {i1, ip, IN1, INp, factor}); 1-based i1/ip.
struct Recalc_f {
  size_t i1, ip;
  double factor;
};
struct Recalc {
  size_t i1, ip;
  Invar IN1, INp;
  double factor;
};

using MatrixElements = std::map<Twoinvar, DMatrix>;

// <I1||f||Ip> = sum_rows factor U_I1(i1) U_Ip(ip)^T (c++/recalc.hpp:52-76). Returns the zeroed
// dim1 x dimp device matrix; its terms are queued in `batch`.
inline DMatrix recalc_f(const DiagInfo &diag, const Invar &I1, const Invar &Ip, const std::vector<Recalc_f> &table,
                        Gemm3Batch &batch) {
  const auto [dim1, dimp] = diag.dims(I1, Ip);
  DMatrix f(dim1, dimp);
  if (!dim1 || !dimp) return f;
  const auto &[diagI1, diagIp] = diag.subs(I1, Ip);
  for (const auto &r : table) batch.product(f.view(), r.factor, diagI1.U(r.i1), diagIp.U(r.ip));
  return f;
}

// cn(I1,Ip) = sum_rows factor U_I1(i1) cold(IN1,INp) U_Ip(ip)^T (c++/recalc.hpp:82-108). Rows whose
// (IN1, INp) block is absent from `cold` are skipped, as in the reference.
inline DMatrix recalc_general(const DiagInfo &diag, const MatrixElements &cold, const Invar &I1, const Invar &Ip,
                              const std::vector<Recalc> &table, Gemm3Batch &batch) {
  const auto [dim1, dimp] = diag.dims(I1, Ip);
  DMatrix cn(dim1, dimp);
  if (!dim1 || !dimp) return cn;
  const auto &[diagI1, diagIp] = diag.subs(I1, Ip);
  for (const auto &r : table) {
    NRG_ASSERT(1 <= r.i1 && r.i1 <= diagI1.U.size() && 1 <= r.ip && r.ip <= diagIp.U.size());
    const auto it = cold.find({r.IN1, r.INp});
    if (it == cold.end()) continue;
    batch.transform(cn.view(), r.factor, diagI1.U(r.i1), it->second.view(), diagIp.U(r.ip));
  }
  return cn;
}

// m += value U_I(i1) U_I(ip)^T (c++/recalc.hpp:111-121)
inline void recalc1_global(const DiagInfo &diag, const Invar &I, DMatrix &m, size_t i1, size_t ip, double value,
                           Gemm3Batch &batch) {
  const auto &diagI = diag.at(I);
  batch.product(m.view(), value, diagI.U(i1), diagI.U(ip));
}

}  // namespace NRG
