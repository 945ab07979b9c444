// dmnrg.hpp - density matrices of the backward sweep (c++/dmnrg.hpp:29-217, c++/store.hpp:19-96).
//
// DM-NRG (init_rho, calc_densitymatrix_iterN) and FDM (init_rho_FDM, calc_fulldensitymatrix_iterN)
// with the reference's trace checks. Every cdmI of a step - all subspaces at N-1 times all their
// children, KK and DD sectors - is queued as one rotate term in a Gemm3Batch and the step is evaluated
// in one grouped call; the density matrices stay in device memory between steps.
#pragma once
#include <chrono>
#include <cstdio>
#include <functional>
#include <map>
#include <string>

#include "measurements.hpp"
#include "numerics.hpp"
#include "recalc.hpp"

namespace NRG {

// Density matrices of one step, one per subspace (c++/operators.hpp:77-83), in device memory.
class DensMatElements : public std::map<Invar, DMatrix> {
 public:
  double trace(const MultFnc &mult) const {   // one diagonal read for all subspaces
    std::vector<DView> v;
    for (const auto &[I, m] : *this) v.push_back(m.view());
    const std::vector<double> t = traces(v);
    double s = 0;
    size_t k = 0;
    for (const auto &[I, m] : *this) s += mult(I) * t[k++];
    return s;
  }
};

// c++/dmnrg.hpp:29-33
inline void check_trace_rho(const DensMatElements &m, const MultFnc &mult, double ref_value = 1.0) {
  if (!num_equal(m.trace(mult), ref_value)) throw std::runtime_error("check_trace_rho() failed");
}

// c++/dmnrg.hpp:42-50: rho_N = exp(-scT E_corr) / Z over the stored states, trace checked to be 1
inline DensMatElements init_rho(double scT, const DiagInfo &diag, const MultFnc &mult) {
  DensMatElements rho;
  const double Z = grand_canonical_Z(scT, diag, mult);
  for (const auto &[I, eig] : diag) {
    Matrix m = eig.diagonal_exp(scT);
    for (size_t i = 0; i < m.size1(); i++) m(i, i) /= Z;
    rho[I] = DMatrix(m);
  }
  check_trace_rho(rho, mult);
  return rho;
}

// rho^{N-1}(I) += factor U^T rho^N(I1) U with U = vectors[0:nromega, part(i)] of child I1
// (c++/dmnrg.hpp:54-80), queued in `batch`.
inline void cdmI(size_t i, const Invar &I1, const DMatrix &rhoN, const Eigen &diagI1, DMatrix &rhoNEW, size_t N,
                 double factor, const Store &store, Gemm3Batch &batch) {
  const size_t dim = size2(rhoNEW), nromega = size2(rhoN);
  if (nromega == 0 || dim == 0) return;
  const SubspaceDimensions &sd = store[N].at(I1).rmax;
  NRG_ASSERT(i < sd.combs());
  const size_t rmax = sd.rmax(i);
  if (rmax == 0) return;
  NRG_ASSERT(rmax == dim);
  NRG_ASSERT(nromega <= diagI1.vectors.M());
  const DView U = diagI1.vectors.submatrix_const({0, nromega}, sd.part(i));
  batch.rotate(rhoNEW.view(), factor, U, rhoN.view());
}

// c++/dmnrg.hpp:84-102
inline DensMatElements calc_densitymatrix_iterN(const DiagInfo &diag, const DensMatElements &rho, size_t N,
                                                const Store &store, const Symmetry &sym, Gemm3Batch &batch) {
  DensMatElements rhoPrev;
  for (const auto &[I, dimsub] : store[N - 1]) {
    const size_t dim = dimsub.kept();
    rhoPrev[I] = DMatrix(dim, dim);
    if (dim == 0) continue;
    const auto ns = sym.new_subspaces(I);
    for (size_t i = 0; i < ns.size(); i++) {
      const auto x = rho.find(ns[i]);
      const auto y = diag.find(ns[i]);
      if (x != rho.end() && y != diag.end())
        cdmI(i, ns[i], x->second, y->second, rhoPrev[I], N, sym.mult(ns[i]) / sym.mult(I), store, batch);
    }
  }
  batch.flush();
  return rhoPrev;
}

// c++/dmnrg.hpp:118-133; diags[N] holds the eigenvectors of step N; `keep` (if given) receives
// rho(N-1) for every step, as the reference saves it (rhoPrev.save(N-1, P, filename))
inline void calc_densitymatrix(DensMatElements &rho, const Store &store, const Symmetry &sym,
                               const std::vector<const DiagInfo *> &diags, size_t Ninit, size_t Nmax, Device &dev,
                               const std::function<void(size_t, const DensMatElements &)> &keep = nullptr) {
  check_trace_rho(rho, sym.multfnc());
  Gemm3Batch batch(dev);
  for (size_t N = Nmax - 1; N > Ninit; N--) {
    auto rhoPrev = calc_densitymatrix_iterN(*diags[N], rho, N, store, sym, batch);
    check_trace_rho(rhoPrev, sym.multfnc());
    if (keep) keep(N - 1, rhoPrev);
    rho.swap(rhoPrev);
  }
}

// Shell-N FDM density matrix from the discarded states (c++/dmnrg.hpp:145-162):
// rho(i,i) = exp(-E_abs_zero(i)/T) w_N / Z_N^D on the D states [min, max), trace checked against w_N.
inline DensMatElements init_rho_FDM(size_t N, const Store &store, const Stats &stats, const MultFnc &mult, double T) {
  DensMatElements rhoFDM;
  std::vector<nrg_diag_item> items;
  std::vector<double> d;
  for (const auto &[I, ds] : store[N]) {
    DMatrix m(ds.max(), ds.max());
    if (stats.ZnDNd[N] != 0.0 && ds.max() > ds.min()) {
      items.push_back(nrg_diag_item{m.data(), (int64_t)m.size2(), (int64_t)ds.min(), (int64_t)(ds.max() - ds.min()),
                                    (int64_t)d.size()});
      for (size_t i = ds.min(); i < ds.max(); i++)
        d.push_back(std::exp(-ds.eig.values.abs_zero(i) / T) * stats.wn[N] / stats.ZnDNd[N]);
    }
    rhoFDM[I] = std::move(m);
  }
  if (!items.empty()) set_diagonals(items, d);   // one transfer for all subspaces
  if (stats.wn[N] != 0.0) {
    const double tr = rhoFDM.trace(mult);
    const double diff = (tr - stats.wn[N]) / stats.wn[N];
    if (!num_equal(diff, 0.0, 1e-8) && !(stats.wn[N] < 1e-12)) {
      size_t nonfinite = 0, states = 0;
      for (const auto &[I, ds] : store[N])
        for (size_t i = ds.min(); i < ds.max(); i++, states++) nonfinite += !std::isfinite(ds.eig.values.abs_zero(i));
      char msg[256];
      std::snprintf(msg, sizeof msg,
                    "init_rho_FDM: trace %.17g of the shell-%zu density matrix differs from w_N = %.17g (Z_N^D = %.17g, "
                    "%zu of %zu discarded energies not finite)",
                    tr, N, stats.wn[N], stats.ZnDNd[N], nonfinite, states);
      throw std::runtime_error(msg);
    }
  }
  return rhoFDM;
}

// c++/dmnrg.hpp:164-197: KK contribution from rhoFDM(N) and, except at the last step, the DD
// contribution from init_rho_FDM(N), rotated into every subspace of step N-1.
// wall-clock seconds of the parts of a sweep, accumulated into `sub` (if given) under `key`
class SubTimer {
 public:
  SubTimer(std::map<std::string, double> *sub, const char *key) : sub_(sub), key_(key), t0_(std::chrono::steady_clock::now()) {}
  ~SubTimer() {
    if (sub_) (*sub_)[key_] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
  }

 private:
  std::map<std::string, double> *sub_;
  const char *key_;
  std::chrono::steady_clock::time_point t0_;
};

inline DensMatElements calc_fulldensitymatrix_iterN(bool last_N, const DiagInfo &diag, const DensMatElements &rhoFDM,
                                                    size_t N, const Store &store, const Stats &stats,
                                                    const Symmetry &sym, double T, Gemm3Batch &batch,
                                                    std::map<std::string, double> *subt = nullptr) {
  DensMatElements rhoDD, rhoFDMPrev;
  if (!last_N) {
    SubTimer t(subt, "FDM.init_rho");
    rhoDD = init_rho_FDM(N, store, stats, sym.multfnc(), T);
  }
  {
    SubTimer tq(subt, "FDM.queue");
    for (const auto &[I, ds] : store[N - 1]) {
      const auto subs = sym.new_subspaces(I);
      const size_t dim = ds.kept();
      rhoFDMPrev[I] = DMatrix(dim, dim);
      if (!dim) continue;
      for (size_t i = 0; i < sym.nr_combs(); i++) {
        const Invar &sub = subs[i];
        const double coef = sym.mult(sub) / sym.mult(I);
        const auto y = diag.find(sub);
        if (y == diag.end()) continue;
        if (const auto x1 = rhoFDM.find(sub); x1 != rhoFDM.end())
          cdmI(i, sub, x1->second, y->second, rhoFDMPrev[I], N, coef, store, batch);
        if (!last_N)
          if (const auto x2 = rhoDD.find(sub); x2 != rhoDD.end())
            cdmI(i, sub, x2->second, y->second, rhoFDMPrev[I], N, coef, store, batch);
      }
    }
  }
  SubTimer tg(subt, "FDM.gemm");
  batch.flush();
  return rhoFDMPrev;
}

// c++/dmnrg.hpp:199-217: the backward FDM sweep. rhoFDM enters as rho_FDM(Nmax-1) and leaves as
// rho_FDM(Ninit); `keep` (if given) receives rho_FDM(N-1) for every step. Tr rho_FDM(N-1) must equal
// sum_{n >= N} w_n.
inline void calc_fulldensitymatrix(DensMatElements &rhoFDM, const Store &store, const Stats &stats, const Symmetry &sym,
                                   const std::vector<const DiagInfo *> &diags, size_t Ninit, size_t Nmax, double T,
                                   Device &dev,
                                   const std::function<void(size_t, const DensMatElements &)> &keep = nullptr,
                                   std::map<std::string, double> *subt = nullptr) {
  Gemm3Batch batch(dev);
  for (size_t N = Nmax - 1; N > Ninit; N--) {
    auto rhoFDMPrev =
        calc_fulldensitymatrix_iterN(N == Nmax - 1, *diags[N], rhoFDM, N, store, stats, sym, T, batch, subt);
    double tr;
    {
      SubTimer tt(subt, "FDM.trace");
      tr = rhoFDMPrev.trace(sym.multfnc());
    }
    double expected = 0.0;
    for (size_t n = N; n < Nmax; n++) expected += stats.wn[n];
    const double diff = (tr - expected) / expected;
    if (!num_equal(diff, 0.0)) {
      char msg[160];
      std::snprintf(msg, sizeof msg, "calc_fulldensitymatrix: trace check failed at N=%zu (trace %.17g, expected %.17g)",
                    N, tr, expected);
      throw std::runtime_error(msg);
    }
    if (keep) {
      SubTimer tk(subt, "FDM.keep");
      keep(N - 1, rhoFDMPrev);
    }
    SubTimer tf(subt, "FDM.free");
    rhoFDM.swap(rhoFDMPrev);
  }
}

}  // namespace NRG
