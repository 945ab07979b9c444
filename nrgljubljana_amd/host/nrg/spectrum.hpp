// spectrum.hpp - FDM spectral accumulation of one NRG step on the GPU.
//
// BinGrid follows Bins::setlimits / loggrid_std (c++/bins.hpp:61-100); FDMSpectrum is BaseSpectrum::calc
// (c++/algo.hpp:55-67) with Algo_FDM (c++/algo_FDM.hpp:123-151) and its two ChainBinning accumulators
// (greater and lesser halves, c++/spectrum.hpp:12-31): every subspace pair of a step with both operator
// blocks present is handed to one nrg_fdm_bin call; the step's bins are merged into the running
// spectrum (SpectrumRealFreq::mergeCFS, c++/spectrum.hpp:73-76) on the host; continuous() trims the
// bins (Bins::trim) and broadens them onto the mesh on the GPU (nrg_broaden). DMNRGSpectrum is the
// same loop for Algo_DMNRG (c++/algo_DMNRG.hpp:30-57) through nrg_dmnrg_bin, merged with the N/N+2
// window (SpectrumRealFreq::mergeNN2, c++/spectrum.hpp:77-127).
#pragma once
#include "dmnrg.hpp"

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
    if (!v.empty()) hip_check(hipMemcpy(d.get(), v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
    return d;
  };
  DBuf<double> dpe = up(pe), dpw = up(pw), dne = up(ne), dnw = up(nw), dm = up(mesh);
  DBuf<double> op(mesh.size() + 1), on(mesh.size() + 1);
  nrg_check(nrg_broaden(dev.ctx(), dpe.get(), dpw.get(), (int64_t)pe.size(), dne.get(), dnw.get(), (int64_t)ne.size(),
                        dm.get(), (int64_t)mesh.size(), alpha, omega0, op.get(), on.get(), dev.stream()),
            "nrg_broaden");
  std::vector<double> vp(mesh.size()), vn(mesh.size());
  if (!mesh.empty()) {
    hip_check(hipMemcpy(vp.data(), op.get(), sizeof(double) * mesh.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(vn.data(), on.get(), sizeof(double) * mesh.size(), hipMemcpyDeviceToHost), "hipMemcpy");
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
    // energies of every subspace, in device memory
    std::map<Invar, std::pair<size_t, size_t>> eoff;
    std::vector<double> E;
    for (const auto &[I, eig] : diag) {
      const auto e = eig.values.all_abs_G();
      eoff[I] = {E.size(), e.size()};
      E.insert(E.end(), e.begin(), e.end());
    }
    DBuf<double> dE(E.size() + 1);
    if (!E.empty())
      hip_check(hipMemcpy(dE.get(), E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice), "hipMemcpy");
    std::vector<nrg_fdm_pair> pairs;
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
        p.E_i = dE.get() + eoff[Ii].first;
        p.E_j = dE.get() + eoff[Ij].first;
        p.op1 = o1.data();
        p.op2 = o2.data();
        p.ld_op = (int64_t)size2(o1);
        if (const auto r = rhoFDM.find(Ii); r != rhoFDM.end()) { p.rho_i = r->second.data(); p.n_rho_i = (int64_t)size1(r->second); }
        if (const auto r = rhoFDM.find(Ij); r != rhoFDM.end()) { p.rho_j = r->second.data(); p.n_rho_j = (int64_t)size1(r->second); }
        p.factor = ff(Ii, Ij);
        pairs.push_back(p);
      }
    const nrg_bin_grid g = grid_.c();
    double *base = acc_.get();
    const size_t nb = grid_.nbins;
    nrg_check(nrg_fdm_bin(dev_.ctx(), pairs.data(), (int)pairs.size(), &g, T_, wnf, sign_, base, base + nb, base + 2 * nb,
                          base + 3 * nb, dev_.stream()),
              "nrg_fdm_bin");
    // mergeCFS of the step's chains into the running spectra
    std::vector<double> h(4 * nb);
    hip_check(hipMemcpy(h.data(), base, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
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

// Algo_DMNRG (c++/algo_DMNRG.hpp:16-57) over all pairs of a step, spectrum merged with the N/N+2 window
class DMNRGSpectrum {
 public:
  DMNRGSpectrum(Device &dev, const BinGrid &grid, int sign, const NN2Window &window)
      : dev_(dev), grid_(grid), sign_(sign), win_(window), acc_(2 * grid.nbins), spec_(grid.nbins), step_(grid.nbins) {
    NRG_ASSERT(sign == 1 || sign == -1);
  }
  // One step: abs_zero energies of `diag`, operator blocks op(Ij, Ii) (keys {Ij, Ii}), the DM-NRG
  // density matrices rho of the step. scale = step.scale(); merge = step.N_for_merging();
  // first/last = step.first() / step.last().
  void calc(const DiagInfo &diag, const MatrixElements &op1, const MatrixElements &op2, const DensMatElements &rho,
            double scale, bool merge, bool first, bool last, const FactorFnc &ff, const CheckFnc &cf = nullptr,
            int spin = 0) {
    std::map<Invar, size_t> eoff;
    std::vector<double> E;
    for (const auto &[I, eig] : diag) {
      const auto e = eig.values.all_abs_zero();
      eoff[I] = E.size();
      E.insert(E.end(), e.begin(), e.end());
    }
    DBuf<double> dE(E.size() + 1);
    if (!E.empty())
      hip_check(hipMemcpy(dE.get(), E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice), "hipMemcpy");
    std::vector<nrg_fdm_pair> pairs;
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
        p.K_i = (int64_t)std::min(diagp.Krange().second, (size_t)p.dim_i);
        p.K_j = (int64_t)std::min(diag1.Krange().second, (size_t)p.dim_j);
        if (!p.K_i || !p.K_j) continue;
        p.E_i = dE.get() + eoff[Ip];
        p.E_j = dE.get() + eoff[I1];
        p.op1 = o1.data();
        p.op2 = o2.data();
        p.ld_op = (int64_t)size2(o1);
        const auto ri = rho.find(Ip), rj = rho.find(I1);
        NRG_ASSERT(ri != rho.end() && rj != rho.end());
        p.rho_i = ri->second.data(); p.n_rho_i = (int64_t)size1(ri->second);
        p.rho_j = rj->second.data(); p.n_rho_j = (int64_t)size1(rj->second);
        p.factor = ff(Ip, I1);
        pairs.push_back(p);
      }
    acc_.zero();
    const nrg_bin_grid g = grid_.c();
    const size_t nb = grid_.nbins;
    nrg_check(nrg_dmnrg_bin(dev_.ctx(), pairs.data(), (int)pairs.size(), &g, win_.emin(scale), win_.emax(scale), sign_,
                            acc_.get(), acc_.get() + nb, dev_.stream()),
              "nrg_dmnrg_bin");
    std::vector<double> h(2 * nb);
    hip_check(hipMemcpy(h.data(), acc_.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    step_.pos.assign(h.begin(), h.begin() + nb);
    step_.neg.assign(h.begin() + nb, h.end());
    if (merge) {   // Algo_DMNRG::end -> SpectrumRealFreq::mergeNN2
      mergeNN2half(grid_, spec_.pos, step_.pos, scale, first, last, win_);
      mergeNN2half(grid_, spec_.neg, step_.neg, scale, first, last, win_);
    }
  }
  const SpectrumBins &spectrum() const { return spec_; }   // the merged spectrum so far
  const SpectrumBins &last_step() const { return step_; }  // the last step's bins before merging
  const BinGrid &grid() const { return grid_; }

 private:
  Device &dev_;
  BinGrid grid_;
  int sign_;
  NN2Window win_;
  DBuf<double> acc_;
  SpectrumBins spec_, step_;
};

}  // namespace NRG
