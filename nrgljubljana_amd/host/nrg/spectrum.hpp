// spectrum.hpp - FDM spectral accumulation of one NRG step on the GPU.
//
// BinGrid follows Bins::setlimits / loggrid_std (c++/bins.hpp:61-100); FDMSpectrum is BaseSpectrum::calc
// (c++/algo.hpp:55-67) with Algo_FDM (c++/algo_FDM.hpp:123-151) and its two ChainBinning accumulators
// (greater and lesser halves, c++/spectrum.hpp:12-31): every subspace pair of a step with both operator
// blocks present is handed to one nrg_fdm_bin call; the step's bins are merged into the running
// spectrum (SpectrumRealFreq::mergeCFS, c++/spectrum.hpp:73-76) on the host; continuous() trims the
// bins (Bins::trim) and broadens them onto the mesh on the GPU (nrg_broaden). DMNRGSpectrum is the
// same loop for Algo_DMNRG (c++/algo_DMNRG.hpp:30-57) through nrg_dmnrg_bin, merged with the N/N+2
// window (SpectrumRealFreq::mergeNN2, c++/spectrum.hpp:77-127). CFSSpectrum is the loop for Algo_CFS
// (c++/algo_CFS.hpp:13-144) through nrg_cfs_bin, each step's bins added as they are (mergeCFS);
// FTSpectrum the loop for Algo_FT (c++/algo_FT.hpp:16-49) through nrg_ft_bin, merged as DM-NRG;
// MatsubaraGF the Matsubara-axis forms FDMmats, DMNRGmats, FTmats (nrg_*_mats); TempSeries the
// temperature-dependent G(T), I1(T), I2(T), chi(T) (nrg_temp_sum).
#pragma once
#include "dmnrg.hpp"
#include <complex>

namespace NRG {

struct BinGrid {
  double log10emin = 0, log10emax = 0;
  double bins = 1000;   // bins per decade (P.bins)
  size_t nbins = 0;
  double discard_immediately = 1e-16;
  // emin / emax as Bins::setlimits computes them from P.emin, P.emax, SCALE(0), last_step_scale()
  static BinGrid standard(double emin_param, double emax_param, double scale0, double last_step_scale, double bins,
                          double discard_immediately = 1e-16) {
    const double emax = emax_param > 0 ? emax_param : scale0 * std::pow(10.0, 2.0);
    const double emin = emin_param > 0 ? emin_param : last_step_scale / std::pow(10.0, 2.0);
    BinGrid g;
    g.log10emin = std::floor(std::log10(emin));
    g.log10emax = std::ceil(std::log10(emax));
    g.bins = bins;
    g.nbins = (size_t)((g.log10emax - g.log10emin) * bins + 1.0);
    g.discard_immediately = discard_immediately;
    return g;
  }
  double energy(size_t i) const { return std::pow(10.0, log10emin + (double)i / bins); }
  nrg_bin_grid c() const { return nrg_bin_grid{log10emin, bins, (int64_t)nbins, discard_immediately}; }
};

// Spectrum of one algorithm half: weights on the grid for positive and negative energies.
struct SpectrumBins {
  std::vector<double> pos, neg;
  explicit SpectrumBins(size_t n = 0) : pos(n, 0.0), neg(n, 0.0) {}
  double total_weight() const {
    double s = 0;
    for (double x : pos) s += x;
    for (double x : neg) s += x;
    return s;
  }
};

// Bins::trim (c++/bins.hpp:164-185): keep bin i if |w_i| >= discard_trim (e_{i+1} - e_i); always
// keep the last bin. Returns the kept (energy, weight) peaks.
inline std::pair<std::vector<double>, std::vector<double>> trim(const BinGrid &g, const std::vector<double> &w,
                                                                double discard_trim) {
  std::vector<double> e, ww;
  if (w.empty()) return {e, ww};
  for (size_t i = 0; i + 1 < w.size(); i++) {
    const double ei = g.energy(i), width = g.energy(i + 1) - ei;
    if (std::abs(w[i]) >= discard_trim * width) {
      e.push_back(ei);
      ww.push_back(w[i]);
    }
  }
  e.push_back(g.energy(w.size() - 1));
  ww.push_back(w.back());
  return {e, ww};
}

// make_mesh (c++/spectrum.hpp:173-178): E = broaden_max, divided by broaden_ratio while E > broaden_min
inline std::vector<double> make_mesh(double broaden_max, double broaden_min, double broaden_ratio) {
  std::vector<double> m;
  for (double E = broaden_max; E > broaden_min; E /= broaden_ratio) m.push_back(E);
  return m;
}

// A broadened spectrum: (energy, value) pairs, negative energies then positive, each ascending, as
// SpectrumRealFreq::continuous writes them.
using Curve = std::vector<std::pair<double, double>>;

// SpectrumRealFreq::continuous (c++/spectrum.hpp:180-209) on the GPU (nrg_broaden).
inline Curve continuous(Device &dev, const BinGrid &g, const SpectrumBins &s, double discard_trim, double alpha,
                        double omega0, const std::vector<double> &mesh) {
  const auto [pe, pw] = trim(g, s.pos, discard_trim);
  const auto [ne, nw] = trim(g, s.neg, discard_trim);
  auto up = [](const std::vector<double> &v) {
    DBuf<double> d(v.size() + 1);
    if (!v.empty()) dcopy(d.get(), v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice);
    return d;
  };
  DBuf<double> dpe = up(pe), dpw = up(pw), dne = up(ne), dnw = up(nw), dm = up(mesh);
  DBuf<double> op(mesh.size() + 1), on(mesh.size() + 1);
  nrg_check(nrg_broaden(dev.ctx(), dpe.get(), dpw.get(), (int64_t)pe.size(), dne.get(), dnw.get(), (int64_t)ne.size(),
                        dm.get(), (int64_t)mesh.size(), alpha, omega0, op.get(), on.get(), dev.stream()),
            "nrg_broaden");
  std::vector<double> vp(mesh.size()), vn(mesh.size());
  if (!mesh.empty()) {
    dcopy(vp.data(), op.get(), sizeof(double) * mesh.size(), hipMemcpyDeviceToHost);
    dcopy(vn.data(), on.get(), sizeof(double) * mesh.size(), hipMemcpyDeviceToHost);
  }
  Curve neg, pos;
  for (size_t k = 0; k < mesh.size(); k++) {
    neg.emplace_back(-mesh[k], vn[k]);
    pos.emplace_back(mesh[k], vp[k]);
  }
  std::sort(neg.begin(), neg.end());
  std::sort(pos.begin(), pos.end());
  neg.insert(neg.end(), pos.begin(), pos.end());
  return neg;
}

using FactorFnc = std::function<double(const Invar &, const Invar &)>;   // ff(Ii, Ij)
using CheckFnc = std::function<bool(const Invar &, const Invar &, int)>;  // cf(Ij, Ii, spin)

namespace detail {
// The pairs of one step for the algorithms whose loops run over kept() and discarded()
// (c++/eigen.hpp:237-242: DM-NRG, CFS, FT): abs_zero energies in device memory, K = the kept states
// (all computed states at the last step when they are all kept; FDM's Krange() is not used here).
struct KeptPairs {   // also the FDM pairs (fdm_pairs)
  DBuf<double> dE;
  std::vector<nrg_fdm_pair> pairs;
  std::vector<std::pair<Invar, Invar>> inv;   // (Ip, I1) of every pair: Ii, Ij of the descriptor
};

inline KeptPairs kept_pairs(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2,
                            const FactorFnc &ff, const CheckFnc &cf, int spin) {
  KeptPairs kp;
  std::map<Invar, size_t> eoff;
  std::vector<double> E;
  for (const auto &[I, eig] : diag) {
    const auto e = eig.values.all_abs_zero();
    eoff[I] = E.size();
    E.insert(E.end(), e.begin(), e.end());
  }
  kp.dE.alloc(E.size() + 1);
  if (!E.empty())
    dcopy(kp.dE.get(), E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice);
  for (const auto &[Ip, diagp] : diag)
    for (const auto &[I1, diag1] : diag) {
      const Twoinvar II{I1, Ip};
      const auto a = op1.find(II), b = op2.find(II);
      if (a == op1.end() || b == op2.end() || (cf && !cf(I1, Ip, spin))) continue;
      const DMatrix &o1 = a->second, &o2 = b->second;
      NRG_ASSERT(size1(o1) == size1(o2) && size2(o1) == size2(o2));
      nrg_fdm_pair p{};
      p.dim_i = (int64_t)size2(o1);
      p.dim_j = (int64_t)size1(o1);
      p.K_i = (int64_t)std::min(diagp.getnrkept(), (size_t)p.dim_i);
      p.K_j = (int64_t)std::min(diag1.getnrkept(), (size_t)p.dim_j);
      p.E_i = kp.dE.get() + eoff[Ip];
      p.E_j = kp.dE.get() + eoff[I1];
      p.op1 = o1.data();
      p.op2 = o2.data();
      p.ld_op = (int64_t)size2(o1);
      p.factor = ff(Ip, I1);
      kp.pairs.push_back(p);
      kp.inv.emplace_back(Ip, I1);
    }
  return kp;
}

// The pairs of one FDM step (Algo_FDM*, c++/algo_FDM.hpp): abs_G energies, K = Krange() (no kept
// states at the last step), rhoFDM where the step has it.
inline KeptPairs fdm_pairs(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2,
                           const DensMatElements &rhoFDM, const FactorFnc &ff, const CheckFnc &cf, int spin) {
  KeptPairs kp;
  std::map<Invar, size_t> eoff;
  std::vector<double> E;
  for (const auto &[I, eig] : diag) {
    const auto e = eig.values.all_abs_G();
    eoff[I] = E.size();
    E.insert(E.end(), e.begin(), e.end());
  }
  kp.dE.alloc(E.size() + 1);
  if (!E.empty())
    dcopy(kp.dE.get(), E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice);
  for (const auto &[Ii, diagi] : diag)
    for (const auto &[Ij, diagj] : diag) {
      const Twoinvar II{Ij, Ii};
      const auto a = op1.find(II), b = op2.find(II);
      if (a == op1.end() || b == op2.end() || (cf && !cf(Ij, Ii, spin))) continue;
      const DMatrix &o1 = a->second, &o2 = b->second;
      NRG_ASSERT(size1(o1) == size1(o2) && size2(o1) == size2(o2));
      nrg_fdm_pair p{};
      p.dim_i = (int64_t)size2(o1);
      p.dim_j = (int64_t)size1(o1);
      NRG_ASSERT((size_t)p.dim_i <= diagi.getnrall() && (size_t)p.dim_j <= diagj.getnrall());
      p.K_i = (int64_t)std::min(diagi.Krange().second, (size_t)p.dim_i);
      p.K_j = (int64_t)std::min(diagj.Krange().second, (size_t)p.dim_j);
      p.E_i = kp.dE.get() + eoff[Ii];
      p.E_j = kp.dE.get() + eoff[Ij];
      p.op1 = o1.data();
      p.op2 = o2.data();
      p.ld_op = (int64_t)size2(o1);
      if (const auto r = rhoFDM.find(Ii); r != rhoFDM.end()) { p.rho_i = r->second.data(); p.n_rho_i = (int64_t)size1(r->second); }
      if (const auto r = rhoFDM.find(Ij); r != rhoFDM.end()) { p.rho_j = r->second.data(); p.n_rho_j = (int64_t)size1(r->second); }
      p.factor = ff(Ii, Ij);
      kp.pairs.push_back(p);
      kp.inv.emplace_back(Ii, Ij);
    }
  return kp;
}

inline void set_rho(nrg_fdm_pair &p, const DensMatElements &rho, const Invar &Ip, const Invar &I1, bool need_i,
                    bool need_j) {
  if (need_i) {   // rho.at(Ip)
    const auto r = rho.find(Ip);
    NRG_ASSERT(r != rho.end());
    p.rho_i = r->second.data(); p.n_rho_i = (int64_t)size1(r->second);
  }
  if (need_j) {   // rho.at(I1)
    const auto r = rho.find(I1);
    NRG_ASSERT(r != rho.end());
    p.rho_j = r->second.data(); p.n_rho_j = (int64_t)size1(r->second);
  }
}

inline std::vector<double> to_host(const DBuf<double> &d, size_t n) {
  std::vector<double> h(n);
  if (n) dcopy(h.data(), d.get(), sizeof(double) * n, hipMemcpyDeviceToHost);
  return h;
}
}  // namespace detail

class FDMSpectrum {
 public:
  FDMSpectrum(Device &dev, const BinGrid &grid, int sign, double T)
      : dev_(dev), grid_(grid), sign_(sign), T_(T), acc_(4 * grid.nbins), gt_(grid.nbins), ls_(grid.nbins) {
    NRG_ASSERT(sign == 1 || sign == -1);
  }
  // One step: energies abs_G of `diag`, operator blocks op(Ij, Ii) = <Ij|op|Ii> (keys {Ij, Ii}),
  // rhoFDM of the same step, wnf = stats.wnfactor[N].
  void calc(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, const DensMatElements &rhoFDM,
            double wnf, const FactorFnc &ff, const CheckFnc &cf = nullptr, int spin = 0) {
    acc_.zero();
    const auto kp = detail::fdm_pairs(diag, op1, op2, rhoFDM, ff, cf, spin);
    const auto &pairs = kp.pairs;
    const nrg_bin_grid g = grid_.c();
    double *base = acc_.get();
    const size_t nb = grid_.nbins;
    nrg_check(nrg_fdm_bin(dev_.ctx(), pairs.data(), (int)pairs.size(), &g, T_, wnf, sign_, base, base + nb, base + 2 * nb,
                          base + 3 * nb, dev_.stream()),
              "nrg_fdm_bin");
    // mergeCFS of the step's chains into the running spectra
    std::vector<double> h(4 * nb);
    dcopy(h.data(), base, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < nb; i++) {
      gt_.pos[i] += h[i];
      gt_.neg[i] += h[nb + i];
      ls_.pos[i] += h[2 * nb + i];
      ls_.neg[i] += h[3 * nb + i];
    }
  }
  const SpectrumBins &gt() const { return gt_; }   // FDMgt
  const SpectrumBins &ls() const { return ls_; }   // FDMls
  SpectrumBins total() const {                     // FDM = gt + ls
    SpectrumBins t(grid_.nbins);
    for (size_t i = 0; i < grid_.nbins; i++) {
      t.pos[i] = gt_.pos[i] + ls_.pos[i];
      t.neg[i] = gt_.neg[i] + ls_.neg[i];
    }
    return t;
  }
  const BinGrid &grid() const { return grid_; }

 private:
  Device &dev_;
  BinGrid grid_;
  int sign_;
  double T_;
  DBuf<double> acc_;
  SpectrumBins gt_, ls_;
};

// Energy window of the N/N+2 patching (c++/params.hpp:650-659, SpectrumRealFreq::mergeNN2half):
// E0 = P.goodE, Efactor = P.getEfactor() (Lambda for one channel, sqrt(Lambda) otherwise or with NN1).
struct NN2Window {
  double E0 = 2.0, Efactor = 2.0, NNtanh = 0.0;
  bool ZBW = false, NN2avg = false;
  double emin(double scale) const { return ZBW ? 0.0 : scale * E0; }
  double ex(double scale) const { return scale * E0 * Efactor; }
  double emax(double scale) const {
    return ZBW ? std::numeric_limits<double>::max() : scale * E0 * Efactor * Efactor;
  }
};

// windowfunction (c++/spectrum.hpp:91-108); first/last: step.first() / step.last()
inline double windowfunction(double E, double Emin, double Ex, double Emax, bool first, bool last, const NN2Window &w) {
  if (E <= Ex && last) return 1.0;
  if (E >= Ex && first) return 1.0;
  if (w.ZBW) return 1.0;
  if (E <= Emin || E >= Emax) return 0.0;
  auto fnc = [&w](double x) {
    if (w.NNtanh <= 0) return x;
    const double f0 = std::tanh(w.NNtanh * (0.0 - 0.5)), fx = std::tanh(w.NNtanh * (x - 0.5)),
                 f1 = std::tanh(w.NNtanh * (1.0 - 0.5));
    return (fx - f0) / (f1 - f0);
  };
  if (E <= Ex) return fnc((E - Emin) / (Ex - Emin));
  return 1.0 - fnc((E - Ex) / (Emax - Ex));
}

// mergeNN2half (c++/spectrum.hpp:115-127): a step's bins into the running half spectrum
inline void mergeNN2half(const BinGrid &g, std::vector<double> &full, const std::vector<double> &cs, double scale,
                         bool first, bool last, const NN2Window &w) {
  NRG_ASSERT(full.size() == cs.size() && cs.size() == g.nbins);
  const double Emin = w.emin(scale), Ex = w.ex(scale), Emax = w.emax(scale);
  for (size_t i = 0; i < g.nbins; i++) {
    const double energy = g.energy(i);
    if (Emin < energy && energy < Emax && cs[i] != 0.0)
      full[i] += (w.NN2avg ? 0.5 : 1.0) * cs[i] * windowfunction(energy, Emin, Ex, Emax, first, last, w);
  }
}


// A step's bins merged with the N/N+2 window (SpectrumRealFreq::mergeNN2, c++/spectrum.hpp:77-81).
class NN2Spectrum {
 public:
  NN2Spectrum(const BinGrid &grid, const NN2Window &window) : grid_(grid), win_(window), spec_(grid.nbins), step_(grid.nbins) {}
  const SpectrumBins &spectrum() const { return spec_; }   // the merged spectrum so far
  const SpectrumBins &last_step() const { return step_; }  // the last step's bins before merging
  const BinGrid &grid() const { return grid_; }

 protected:
  void merge_step(const std::vector<double> &h, double scale, bool merge, bool first, bool last) {
    const size_t nb = grid_.nbins;
    step_.pos.assign(h.begin(), h.begin() + nb);
    step_.neg.assign(h.begin() + nb, h.begin() + 2 * nb);
    if (merge) {
      mergeNN2half(grid_, spec_.pos, step_.pos, scale, first, last, win_);
      mergeNN2half(grid_, spec_.neg, step_.neg, scale, first, last, win_);
    }
  }
  BinGrid grid_;
  NN2Window win_;
  SpectrumBins spec_, step_;
};

// Algo_DMNRG (c++/algo_DMNRG.hpp:16-57) over all pairs of a step, spectrum merged with the N/N+2 window
class DMNRGSpectrum : public NN2Spectrum {
 public:
  DMNRGSpectrum(Device &dev, const BinGrid &grid, int sign, const NN2Window &window)
      : NN2Spectrum(grid, window), dev_(dev), sign_(sign), acc_(2 * grid.nbins) {
    NRG_ASSERT(sign == 1 || sign == -1);
  }
  // One step: abs_zero energies of `diag`, operator blocks op(Ij, Ii) (keys {Ij, Ii}), the DM-NRG
  // density matrices rho of the step. scale = step.scale(); merge = step.N_for_merging();
  // first/last = step.first() / step.last().
  void calc(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, const DensMatElements &rho,
            double scale, bool merge, bool first, bool last, const FactorFnc &ff, const CheckFnc &cf = nullptr,
            int spin = 0) {
    auto kp = detail::kept_pairs(diag, op1, op2, ff, cf, spin);
    std::vector<nrg_fdm_pair> pairs;
    for (size_t k = 0; k < kp.pairs.size(); k++) {
      nrg_fdm_pair p = kp.pairs[k];
      if (!p.K_i || !p.K_j) continue;
      detail::set_rho(p, rho, kp.inv[k].first, kp.inv[k].second, true, true);
      pairs.push_back(p);
    }
    acc_.zero();
    const nrg_bin_grid g = grid_.c();
    const size_t nb = grid_.nbins;
    nrg_check(nrg_dmnrg_bin(dev_.ctx(), pairs.data(), (int)pairs.size(), &g, win_.emin(scale), win_.emax(scale), sign_,
                            acc_.get(), acc_.get() + nb, dev_.stream()),
              "nrg_dmnrg_bin");
    merge_step(detail::to_host(acc_, 2 * nb), scale, merge, first, last);   // Algo_DMNRG::end
  }

 private:
  Device &dev_;
  int sign_;
  DBuf<double> acc_;
};

// Algo_FT (c++/algo_FT.hpp:16-49) over all pairs of a step, merged with the N/N+2 window
class FTSpectrum : public NN2Spectrum {
 public:
  FTSpectrum(Device &dev, const BinGrid &grid, int sign, double T, const NN2Window &window)
      : NN2Spectrum(grid, window), dev_(dev), sign_(sign), T_(T), acc_(2 * grid.nbins) {
    NRG_ASSERT(sign == 1 || sign == -1);
  }
  // One step: Z = stats.Zft; scale, merge, first, last as for DMNRGSpectrum::calc.
  void calc(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, double Z, double scale,
            bool merge, bool first, bool last, const FactorFnc &ff, const CheckFnc &cf = nullptr, int spin = 0) {
    auto kp = detail::kept_pairs(diag, op1, op2, ff, cf, spin);
    std::vector<nrg_fdm_pair> pairs;
    for (const auto &p : kp.pairs)
      if (p.K_i && p.K_j) pairs.push_back(p);
    acc_.zero();
    const nrg_bin_grid g = grid_.c();
    const size_t nb = grid_.nbins;
    nrg_check(nrg_ft_bin(dev_.ctx(), pairs.data(), (int)pairs.size(), &g, T_, Z, sign_, acc_.get(), acc_.get() + nb,
                         dev_.stream()),
              "nrg_ft_bin");
    merge_step(detail::to_host(acc_, 2 * nb), scale, merge, first, last);   // Algo_FT::end
  }

 private:
  Device &dev_;
  int sign_;
  double T_;
  DBuf<double> acc_;
};

// Algo_CFS = Algo_CFSgt + Algo_CFSls (c++/algo_CFS.hpp:13-144) over all pairs of a step. Each step's
// bins are added into the spectra as they are (SpectrumRealFreq::mergeCFS, c++/spectrum.hpp:73-76):
// greater(), lesser() and their sum total() (the reference's CFSgt, CFSls and CFS outputs).
class CFSSpectrum {
 public:
  CFSSpectrum(Device &dev, const BinGrid &grid, int sign)
      : dev_(dev), grid_(grid), sign_(sign), acc_(4 * grid.nbins), gt_(grid.nbins), ls_(grid.nbins),
        tot_(grid.nbins), step_gt_(grid.nbins), step_ls_(grid.nbins) {
    NRG_ASSERT(sign == 1 || sign == -1);
  }
  // One step: abs_zero energies of `diag`, operator blocks op(Ij, Ii) (keys {Ij, Ii}), the DM-NRG
  // density matrices rho of the step (read below the last step only). last = step.last(); T = P.T;
  // Z = stats.Zft.
  void calc(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, const DensMatElements &rho,
            bool last, double T, double Z, const FactorFnc &ff, const CheckFnc &cf = nullptr, int spin = 0) {
    auto kp = detail::kept_pairs(diag, op1, op2, ff, cf, spin);
    std::vector<nrg_fdm_pair> pairs;
    for (size_t k = 0; k < kp.pairs.size(); k++) {
      nrg_fdm_pair p = kp.pairs[k];
      const bool ls = !last && p.K_i > 0 && p.dim_j > p.K_j, gt = !last && p.K_j > 0 && p.dim_i > p.K_i;
      if (last ? !(p.K_i && p.K_j) : !(ls || gt)) continue;
      detail::set_rho(p, rho, kp.inv[k].first, kp.inv[k].second, ls, gt);
      pairs.push_back(p);
    }
    acc_.zero();
    const nrg_bin_grid g = grid_.c();
    const size_t nb = grid_.nbins;
    nrg_check(nrg_cfs_bin(dev_.ctx(), pairs.data(), (int)pairs.size(), &g, T, Z, last ? 1 : 0, sign_, acc_.get(),
                          acc_.get() + nb, acc_.get() + 2 * nb, acc_.get() + 3 * nb, dev_.stream()),
              "nrg_cfs_bin");
    const std::vector<double> h = detail::to_host(acc_, 4 * nb);
    step_gt_.pos.assign(h.begin(), h.begin() + nb);
    step_gt_.neg.assign(h.begin() + nb, h.begin() + 2 * nb);
    step_ls_.pos.assign(h.begin() + 2 * nb, h.begin() + 3 * nb);
    step_ls_.neg.assign(h.begin() + 3 * nb, h.end());
    for (size_t k = 0; k < nb; k++) {   // Algo_CFS::end: spec_tot.mergeCFS(gt) then mergeCFS(ls)
      gt_.pos[k] += step_gt_.pos[k];
      gt_.neg[k] += step_gt_.neg[k];
      ls_.pos[k] += step_ls_.pos[k];
      ls_.neg[k] += step_ls_.neg[k];
      tot_.pos[k] += step_gt_.pos[k];
      tot_.pos[k] += step_ls_.pos[k];
      tot_.neg[k] += step_gt_.neg[k];
      tot_.neg[k] += step_ls_.neg[k];
    }
  }
  const SpectrumBins &greater() const { return gt_; }
  const SpectrumBins &lesser() const { return ls_; }
  const SpectrumBins &total() const { return tot_; }
  const SpectrumBins &last_greater() const { return step_gt_; }   // the last step's bins
  const SpectrumBins &last_lesser() const { return step_ls_; }
  const BinGrid &grid() const { return grid_; }

 private:
  Device &dev_;
  BinGrid grid_;
  int sign_;
  DBuf<double> acc_;
  SpectrumBins gt_, ls_, tot_, step_gt_, step_ls_;
};

// GFMatsubara + ChainMatsubara (c++/spectrum.hpp:34-42, 212-227; Matsubara, c++/matsubara.hpp):
// G(i w_n), n < mats, with each step's values added (GFMatsubara::merge). calc_fdm, calc_dmnrg and
// calc_ft are Algo_FDMmats, Algo_DMNRGmats and Algo_FTmats over all pairs of a step.
class MatsubaraGF {
 public:
  MatsubaraGF(Device &dev, size_t mats, int sign, double T)
      : dev_(dev), mats_(mats), sign_(sign), T_(T), acc_(2 * std::max<size_t>(mats, 1)), G_(mats) {
    NRG_ASSERT(sign == 1 || sign == -1);
  }
  // ww(n) (c++/matsubara.hpp:20-27): T pi (2n + 1) for fermions (sign -1), T pi 2n for bosons
  double frequency(size_t n) const { return T_ * M_PI * (double)(2 * n + (sign_ == 1 ? 0 : 1)); }
  // Algo_FDMmats: abs_G energies, rhoFDM of the step, wnf = stats.wnfactor[N]
  void calc_fdm(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2,
                const DensMatElements &rhoFDM, double wnf, const FactorFnc &ff, const CheckFnc &cf = nullptr,
                int spin = 0) {
    const auto kp = detail::fdm_pairs(diag, op1, op2, rhoFDM, ff, cf, spin);
    run([&] {
      return nrg_fdm_mats(dev_.ctx(), kp.pairs.data(), (int)kp.pairs.size(), (int64_t)mats_, T_, wnf, sign_,
                          acc_.get(), dev_.stream());
    }, "nrg_fdm_mats");
  }
  // Algo_DMNRGmats: abs_zero energies, the DM-NRG rho of the step
  void calc_dmnrg(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, const DensMatElements &rho,
                  const FactorFnc &ff, const CheckFnc &cf = nullptr, int spin = 0) {
    auto kp = detail::kept_pairs(diag, op1, op2, ff, cf, spin);
    std::vector<nrg_fdm_pair> pairs;
    for (size_t k = 0; k < kp.pairs.size(); k++) {
      nrg_fdm_pair p = kp.pairs[k];
      if (!p.K_i || !p.K_j) continue;
      detail::set_rho(p, rho, kp.inv[k].first, kp.inv[k].second, true, true);
      pairs.push_back(p);
    }
    run([&] {
      return nrg_dmnrg_mats(dev_.ctx(), pairs.data(), (int)pairs.size(), (int64_t)mats_, T_, sign_, acc_.get(),
                            dev_.stream());
    }, "nrg_dmnrg_mats");
  }
  // Algo_FTmats: abs_zero energies, Z = stats.Zft
  void calc_ft(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, double Z,
               const FactorFnc &ff, const CheckFnc &cf = nullptr, int spin = 0) {
    auto kp = detail::kept_pairs(diag, op1, op2, ff, cf, spin);
    std::vector<nrg_fdm_pair> pairs;
    for (const auto &p : kp.pairs)
      if (p.K_i && p.K_j) pairs.push_back(p);
    run([&] {
      return nrg_ft_mats(dev_.ctx(), pairs.data(), (int)pairs.size(), (int64_t)mats_, T_, Z, sign_, acc_.get(),
                         dev_.stream());
    }, "nrg_ft_mats");
  }
  const std::vector<std::complex<double>> &values() const { return G_; }   // G(i w_n), n < mats
  const std::vector<std::complex<double>> &last_step() const { return step_; }

 private:
  template <class F> void run(F &&call, const char *what) {
    acc_.zero();
    nrg_check(call(), what);
    const std::vector<double> h = detail::to_host(acc_, 2 * mats_);
    step_.assign(mats_, {0.0, 0.0});
    for (size_t n = 0; n < mats_; n++) {
      step_[n] = {h[2 * n], h[2 * n + 1]};
      G_[n] += step_[n];
    }
  }
  Device &dev_;
  size_t mats_;
  int sign_;
  double T_;
  DBuf<double> acc_;
  std::vector<std::complex<double>> G_, step_;
};

// TempDependence + ChainTempDependence (c++/spectrum.hpp:44-53, 229-245) with Temp::add_value
// (c++/bins.hpp:188-203): (temperature, value) points of G(T), I1(T), I2(T) (Algo_GT<n>) or chi(T)
// (Algo_CHIT), one per step, values at an equal temperature added. kind: NRG_TEMP_*.
class TempSeries {
 public:
  TempSeries(Device &dev, int kind) : dev_(dev), kind_(kind), acc_(1) {
    NRG_ASSERT(kind >= NRG_TEMP_GT && kind <= NRG_TEMP_CHIT);
  }
  // One step: abs_zero energies; temperature = P.gtp * step.scale() (or P.chitp * ...), Z = stats.Zgt
  // (or stats.Zchit). Returns the step's value.
  double calc(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, double temperature, double Z,
              const FactorFnc &ff, const CheckFnc &cf = nullptr, int spin = 0) {
    auto kp = detail::kept_pairs(diag, op1, op2, ff, cf, spin);
    std::vector<nrg_fdm_pair> pairs;
    for (const auto &p : kp.pairs)
      if (p.K_i && p.K_j) pairs.push_back(p);
    acc_.zero();
    nrg_check(nrg_temp_sum(dev_.ctx(), kind_, pairs.data(), (int)pairs.size(), temperature, Z, acc_.get(), dev_.stream()),
              "nrg_temp_sum");
    const double v = detail::to_host(acc_, 1)[0];
    for (auto &[t, w] : pts_)
      if (t == temperature) {
        w += v;
        return v;
      }
    pts_.emplace_back(temperature, v);
    return v;
  }
  const std::vector<std::pair<double, double>> &points() const { return pts_; }

 private:
  Device &dev_;
  int kind_;
  DBuf<double> acc_;
  std::vector<std::pair<double, double>> pts_;
};

}  // namespace NRG
