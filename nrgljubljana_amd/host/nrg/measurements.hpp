// measurements.hpp - per-step store, statistical sums and FDM thermodynamics.
//
// Store / Sub (c++/store.hpp:19-96), grand_canonical_Z, calc_ZnD and fdm_thermodynamics
// (c++/measurements.hpp:51-163). The reference accumulates Z_n^D, Z_{ZZG} and the energy moments in
// 400-bit GMP floats and reads them back with mpf_get_d (which truncates toward zero). GMP is not
// used here: the sums are kept in IEEE binary128 (__float128, 113-bit significand). The
// summands are the same doubles (mult * exp(...) formed in double, as the reference forms them before
// mpf_set_d), every addend is positive in the Z sums, and the read-back truncates toward zero as
// mpf_get_d does, so the double results agree with the 400-bit ones unless the exact sum lies within
// ~1e-30 relative of a double boundary.
#pragma once
#include <quadmath.h>

#include <functional>

#include "recalc.hpp"

namespace NRG {

using quad = __float128;

// mpf_get_d: nearest double toward zero
inline double quad_to_double_trunc(quad q) {
  double d = (double)q;
  if (std::isfinite(d) && fabsq((quad)d) > fabsq(q)) d = std::nextafter(d, 0.0);
  return d;
}

// Symmetry information the sweep needs (the relevant subset of c++/symmetry.hpp): the children of a
// subspace of step N-1 in step N (one per combination, ordered as the ancestors of SubspaceDimensions),
// and the multiplicity of a subspace.
class Symmetry {
 public:
  virtual ~Symmetry() = default;
  virtual std::vector<Invar> new_subspaces(const Invar &I) const = 0;
  virtual double mult(const Invar &I) const = 0;
  virtual size_t nr_combs() const = 0;
  MultFnc multfnc() const {
    return [this](const Invar &I) { return mult(I); };
  }
};

struct Sub {   // c++/store.hpp:19-33
  Eigen eig;
  SubspaceDimensions rmax;
  bool is_last = false;
  size_t kept() const { return eig.getnrkept(); }
  size_t total() const { return eig.getdim(); }
  size_t min() const { return is_last ? 0 : kept(); }   // D states of FDM: [min, max)
  size_t max() const { return total(); }
};
using Subs = std::map<Invar, Sub>;

class Store : public std::vector<Subs> {   // c++/store.hpp:49-96
 public:
  const size_t Nbegin, Nend;
  Store(size_t Nbegin_, size_t Nend_) : Nbegin(Nbegin_), Nend(Nend_) { this->resize(Nend ? Nend : 1); }
  void shift_abs_energies(double GS_energy) {
    for (size_t N = Nbegin; N < Nend; N++)
      for (auto &kv : (*this)[N]) kv.second.eig.subtract_GS_energy(GS_energy);
  }
};

struct Stats {   // the FDM fields of c++/stats.hpp
  std::vector<quad> ZnDG, ZnDN;
  std::vector<double> ZnDNd, wn, wnfactor;
  double ZZG = 0, GS_energy = 0;
  double Z_fdm = 0, F_fdm = 0, E_fdm = 0, C_fdm = 0, S_fdm = 0;
  explicit Stats(size_t nlen = 0) : ZnDG(nlen), ZnDN(nlen), ZnDNd(nlen), wn(nlen), wnfactor(nlen) {}
};

// c++/measurements.hpp:51-55: sum over subspaces of mult(I) * sum over kept states of exp(-factor E_corr)
inline double grand_canonical_Z(double factor, const DiagInfo &diag, const MultFnc &mult) {
  double Z = 0.0;
  for (const auto &[I, eig] : diag) {
    double s = 0.0;
    for (size_t i = 0; i < eig.getnrkept(); i++) s += std::exp(-factor * eig.values.corr(i));
    Z += mult(I) * s;
  }
  return Z;
}

// c++/measurements.hpp:59-101
inline void calc_ZnD(const Store &store, Stats &stats, const Symmetry &sym, double T) {
  NRG_ASSERT(stats.ZnDG.size() >= store.Nend);
  for (size_t N = store.Nbegin; N < store.Nend; N++) {
    quad ZnDG = 0, ZnDN = 0;
    for (const auto &[I, ds] : store[N])
      for (size_t i = ds.min(); i < ds.max(); i++) {
        ZnDG += (quad)(sym.mult(I) * std::exp(-ds.eig.values.abs_G(i) / T));
        ZnDN += (quad)(sym.mult(I) * std::exp(-ds.eig.values.abs_zero(i) / T));
      }
    stats.ZnDG[N] = ZnDG;
    stats.ZnDN[N] = ZnDN;
    stats.ZnDNd[N] = quad_to_double_trunc(ZnDN);
  }
  quad ZZG = 0;
  for (size_t N = store.Nbegin; N < store.Nend; N++) {
    quad b = 1;
    for (size_t k = 0; k + N + 1 < store.Nend; k++) b *= (quad)sym.nr_combs();
    ZZG += stats.ZnDG[N] * b;
  }
  stats.ZZG = quad_to_double_trunc(ZZG);
  for (size_t N = store.Nbegin; N < store.Nend; N++) {
    const double w = std::pow((double)sym.nr_combs(), (double)(store.Nend - N - 1)) / stats.ZZG;
    stats.wnfactor[N] = w;
    stats.wn[N] = w * quad_to_double_trunc(stats.ZnDG[N]);
  }
  double sumwn = 0.0;
  for (double x : stats.wn) sumwn += x;
  if (!num_equal(sumwn, 1.0)) throw std::runtime_error("calc_ZnD: sum of w_n differs from 1");
}

// c++/measurements.hpp:117-163
inline void fdm_thermodynamics(const Store &store, Stats &stats, const Symmetry &sym, double T) {
  stats.Z_fdm = stats.ZZG * std::exp(-stats.GS_energy / T);
  stats.F_fdm = -std::log(stats.ZZG) * T + stats.GS_energy;
  quad E = 0, E2 = 0;
  for (size_t N = store.Nbegin; N < store.Nend; N++)
    if (stats.wn[N] > 1e-16)
      for (const auto &[I, ds] : store[N])
        for (size_t i = ds.min(); i < ds.max(); i++) {
          quad weight = (quad)(stats.wn[N] * sym.mult(I) * std::exp(-ds.eig.values.abs_zero(i) / T));
          weight /= stats.ZnDN[N];
          const quad e = (quad)ds.eig.values.abs_T(i);
          E += e * weight;
          E2 += e * e * weight;
        }
  stats.E_fdm = quad_to_double_trunc(E);
  const quad varE = E2 - E * E;
  stats.C_fdm = quad_to_double_trunc(varE) / std::pow(T, 2);
  stats.S_fdm = (stats.E_fdm - stats.F_fdm) / T;
}

}  // namespace NRG
