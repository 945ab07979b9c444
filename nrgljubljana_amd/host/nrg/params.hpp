// params.hpp - the `param` file of an NRG run: the [param] block read into typed settings.
//
// Same keywords, defaults and meanings as the reference's Params (c++/params.hpp:84-662) for the
// settings the per-step path uses; the block is parsed as parse_block / find_block do
// (c++/misc.hpp:91-131: `keyword=value` lines up to the next `[block]`, trailing blanks stripped,
// duplicate keywords an error). Keywords the path does not use are kept in `unused` and reported,
// as "Unused settings" in the reference. The derived quantities (SCALE, nrg_step_scale_factor,
// the FDM/DM flags, channels and combs) follow c++/params.hpp:505-659.
#pragma once
#include <cmath>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace NRG {

enum class RUNTYPE { NRG, DMNRG };

class Params {
 public:
  // discretisation and chain
  std::string symtype, discretization = "Z", tri = "old";
  double Lambda = 2.0, z = 1.0, bandrescale = 1.0;
  bool polarized = false, pol2x2 = false, rungs = false, substeps = false, absolute = false;
  size_t Ninit = 0;
  // iteration
  std::string diag = "default", strategy = "kept";
  double diagratio = 1.0, restartfactor = 2.0, keepenergy = -1.0, safeguard = 1e-5, fixeps = 1e-15;
  bool restart = true;
  size_t keep = 100, keepmin = 0, safeguardmax = 200;
  // temperature and operators
  double T = 0.001, betabar = 1.0, gtp = 0.7, chitp = 1.0, chitp_ratio = -999;
  std::string ops, specs, specd, spect, specq, specot, specgt, speci1t, speci2t, specchit;
  // algorithms
  bool finite = false, dmnrg = false, cfs = false, cfsgt = false, cfsls = false, fdm = false, fdmgt = false,
       fdmls = false, fdmexpv = false, finitemats = false, dmnrgmats = false, fdmmats = false, dm = false;
  size_t fdmexpvn = 0, mats = 100;
  // mesh, binning, broadening
  double broaden_max = 10, broaden_min = -99., broaden_min_ratio = 3.0, broaden_ratio = 1.05, alpha = 0.3,
         omega0 = -1.0, omega0_ratio = 1.0, emin = -1.0, emax = -1.0, accumulation = 0.0, linstep = 0.0,
         discard_trim = 1e-16, discard_immediately = 1e-16, goodE = 2.0, NNtanh = 0.0;
  bool savebins = false, broaden = true, NN1 = false, NN2even = true, NN2avg = false, reim = false;
  size_t bins = 1000;
  // output
  size_t dumpannotated = 0, dumpprecision = 8, dumpdiagonal = 0;
  bool dumpabs = false, dumpscaled = true, dumpgroups = true, dumpsubspaces = false, dumpenergies = false,
       dumpabsenergies = false, checksumrules = false, calc0 = true, lastall = false, lastalloverride = false,
       data_has_rescaled_energies = true, logall = false, done = true;
  double grouptol = 1e-6;
  int width_td = 16, width_custom = 16, prec_td = 10, prec_custom = 10, prec_xy = 10;
  std::string logstr;
  std::map<std::string, std::string> unused;

  // set from the data file (Params::set_channels_and_combs, determine_Nmax_Nlen)
  size_t channels = 0, coeffactor = 1, coefchannels = 0, perchannel = 1, combs = 0, Nmax = 0, Nlen = 0, spin = 2;

  Params() = default;
  explicit Params(const std::string &filename, const std::string &block = "param") {
    std::ifstream F(filename);
    if (!F) throw std::runtime_error("Can't open " + filename + " for reading.");
    std::stringstream ss;
    ss << F.rdbuf();
    parse(ss.str(), block);
  }
  void parse(const std::string &text, const std::string &block = "param") {
    std::istringstream F(text);
    std::string line;
    bool in = false, found = false;
    std::map<std::string, std::string> kv;
    while (std::getline(F, line)) {
      while (!line.empty() && (line.back() == '\r')) line.pop_back();
      if (!in) {
        if (line == "[" + block + "]") in = found = true;
        continue;
      }
      if (line.empty() || line[0] == '#') continue;
      if (line[0] == '[') break;
      const auto eq = line.find('=');
      if (eq == std::string::npos) continue;
      const std::string key = line.substr(0, eq);
      std::string value = line.substr(eq + 1);
      while (!value.empty() && std::isspace((unsigned char)value.back())) value.pop_back();
      if (kv.count(key)) throw std::runtime_error("Duplicate keyword: " + key);
      kv[key] = value;
    }
    if (!found) throw std::runtime_error("Block [" + block + "] not found in the parameter file.");
    for (const auto &[k, v] : kv)
      if (!set(k, v)) unused[k] = v;
    validate();
  }

  // c++/params.hpp:505-531
  void set_channels_and_combs(size_t ch) {
    if (ch < 1) throw std::runtime_error("channels must be >= 1");
    channels = ch;
    coeffactor = pol2x2 ? 4 : (polarized ? 2 : 1);
    coefchannels = coeffactor * channels;
    if (symtype == "U1" || symtype == "SU2" || symtype == "DBLSU2") perchannel = 2;
    else if (symtype == "NONE" || symtype == "P" || symtype == "PP") perchannel = 4;
    else perchannel = 1;
    spin = (symtype == "SL" || symtype == "SL3") ? 1 : 2;
    const size_t states = spin == 1 ? 2 : 4;
    combs = 1;
    if (!substeps)
      for (size_t i = 0; i < channels; i++) combs *= states;
    else
      combs = states;
  }

  bool cfs_flags() const { return cfs || cfsgt || cfsls; }
  bool fdm_flags() const { return fdm || fdmgt || fdmls || fdmmats || fdmexpv; }
  bool dmnrg_flags() const { return dmnrg || dmnrgmats; }
  bool cfs_or_fdm_flags() const { return cfs_flags() || fdm_flags(); }
  bool dm_flags() const { return cfs_flags() || fdm_flags() || dmnrg_flags(); }
  bool keep_all_states_in_last_step() const { return lastall || (cfs_or_fdm_flags() && !lastalloverride); }
  bool need_rho() const { return cfs_flags() || dmnrg_flags(); }
  bool need_rhoFDM() const { return fdm_flags(); }
  bool ZBW() const { return Nmax == Ninit; }
  bool do_recalc_kept(RUNTYPE rt) const {
    return strategy == "kept" && !(cfs_or_fdm_flags() && rt == RUNTYPE::DMNRG) && !ZBW();
  }
  bool do_recalc_all(RUNTYPE rt) const { return !do_recalc_kept(rt) && !ZBW(); }
  bool do_recalc_none() const { return ZBW(); }
  bool logletter(char c) const { return logall || logstr.find(c) != std::string::npos; }

  // c++/params.hpp:619-634: the reduced bandwidth of the length-N chain
  double SCALE(int N) const {
    double scale = 0.0;
    if (discretization == "Y") scale = 0.5 * (1. + 1. / Lambda);
    if (discretization == "C" || discretization == "Z") scale = (1.0 - 1. / Lambda) / std::log(Lambda);
    if (!substeps)
      scale *= std::pow(Lambda, -(N - 1) / 2. + 1 - z);
    else
      scale *= std::pow(Lambda, -N / (2. * channels) + 3 / 2. - z);
    if (scale == 0.0) throw std::runtime_error("SCALE: unknown discretization " + discretization);
    return scale * bandrescale;
  }
  double last_step_scale() const { return SCALE((int)Nmax); }
  double nrg_step_scale_factor() const {
    return absolute ? 1 : (!substeps ? std::sqrt(Lambda) : std::pow(Lambda, 0.5 / channels));
  }
  double get_broaden_min() const { return broaden_min <= 0.0 ? broaden_min_ratio * last_step_scale() : broaden_min; }
  double getEfactor() const { return (channels == 1 && !NN1) ? Lambda : std::sqrt(Lambda); }
  double getE0() const { return goodE; }
  double getEmin() const { return !ZBW() ? getE0() : 0.0; }
  double getEx() const { return getE0() * getEfactor(); }
  double getEmax() const { return !ZBW() ? getE0() * std::pow(getEfactor(), 2) : std::numeric_limits<double>::max(); }

 private:
  static bool to_bool(const std::string &k, const std::string &v) {
    if (v == "true" || v == "1" || v == "yes") return true;
    if (v == "false" || v == "0" || v == "no") return false;
    throw std::runtime_error("bad boolean value for " + k + ": " + v);
  }
  bool set(const std::string &k, const std::string &v) {
    auto D = [&](double &x) { x = std::stod(v); return true; };
    auto Z = [&](size_t &x) { x = (size_t)std::stoull(v); return true; };
    auto I = [&](int &x) { x = std::stoi(v); return true; };
    auto B = [&](bool &x) { x = to_bool(k, v); return true; };
    auto S = [&](std::string &x) { x = v; return true; };
    if (k == "symtype") return S(symtype);
    if (k == "discretization") return S(discretization);
    if (k == "tri") return S(tri);
    if (k == "Lambda") return D(Lambda);
    if (k == "z") return D(z);
    if (k == "bandrescale") return D(bandrescale);
    if (k == "polarized") return B(polarized);
    if (k == "pol2x2") return B(pol2x2);
    if (k == "rungs") return B(rungs);
    if (k == "substeps") return B(substeps);
    if (k == "absolute") return B(absolute);
    if (k == "Ninit") return Z(Ninit);
    if (k == "diag") return S(diag);
    if (k == "strategy") return S(strategy);
    if (k == "diagratio") return D(diagratio);
    if (k == "restartfactor") return D(restartfactor);
    if (k == "keepenergy") return D(keepenergy);
    if (k == "safeguard") return D(safeguard);
    if (k == "fixeps") return D(fixeps);
    if (k == "restart") return B(restart);
    if (k == "keep") return Z(keep);
    if (k == "keepmin") return Z(keepmin);
    if (k == "safeguardmax") return Z(safeguardmax);
    if (k == "T") return D(T);
    if (k == "betabar") return D(betabar);
    if (k == "gtp") return D(gtp);
    if (k == "chitp") return D(chitp);
    if (k == "chitp_ratio") return D(chitp_ratio);
    if (k == "ops") return S(ops);
    if (k == "specs") return S(specs);
    if (k == "specd") return S(specd);
    if (k == "spect") return S(spect);
    if (k == "specq") return S(specq);
    if (k == "specot") return S(specot);
    if (k == "specgt") return S(specgt);
    if (k == "speci1t") return S(speci1t);
    if (k == "speci2t") return S(speci2t);
    if (k == "specchit") return S(specchit);
    if (k == "finite") return B(finite);
    if (k == "dmnrg") return B(dmnrg);
    if (k == "cfs") return B(cfs);
    if (k == "cfsgt") return B(cfsgt);
    if (k == "cfsls") return B(cfsls);
    if (k == "fdm") return B(fdm);
    if (k == "fdmgt") return B(fdmgt);
    if (k == "fdmls") return B(fdmls);
    if (k == "fdmexpv") return B(fdmexpv);
    if (k == "fdmexpvn") return Z(fdmexpvn);
    if (k == "finitemats") return B(finitemats);
    if (k == "dmnrgmats") return B(dmnrgmats);
    if (k == "fdmmats") return B(fdmmats);
    if (k == "mats") return Z(mats);
    if (k == "dm") return B(dm);
    if (k == "broaden_max") return D(broaden_max);
    if (k == "broaden_min") return D(broaden_min);
    if (k == "broaden_min_ratio") return D(broaden_min_ratio);
    if (k == "broaden_ratio") return D(broaden_ratio);
    if (k == "alpha") return D(alpha);
    if (k == "omega0") return D(omega0);
    if (k == "omega0_ratio") return D(omega0_ratio);
    if (k == "emin") return D(emin);
    if (k == "emax") return D(emax);
    if (k == "accumulation") return D(accumulation);
    if (k == "linstep") return D(linstep);
    if (k == "discard_trim") return D(discard_trim);
    if (k == "discard_immediately") return D(discard_immediately);
    if (k == "goodE") return D(goodE);
    if (k == "NNtanh") return D(NNtanh);
    if (k == "savebins") return B(savebins);
    if (k == "broaden") return B(broaden);
    if (k == "NN1") return B(NN1);
    if (k == "NN2even") return B(NN2even);
    if (k == "NN2avg") return B(NN2avg);
    if (k == "reim") return B(reim);
    if (k == "bins") return Z(bins);
    if (k == "dumpannotated") return Z(dumpannotated);
    if (k == "dumpprecision") return Z(dumpprecision);
    if (k == "dumpdiagonal") return Z(dumpdiagonal);
    if (k == "dumpabs") return B(dumpabs);
    if (k == "dumpscaled") return B(dumpscaled);
    if (k == "dumpgroups") return B(dumpgroups);
    if (k == "dumpsubspaces") return B(dumpsubspaces);
    if (k == "dumpenergies") return B(dumpenergies);
    if (k == "dumpabsenergies") return B(dumpabsenergies);
    if (k == "checksumrules") return B(checksumrules);
    if (k == "calc0") return B(calc0);
    if (k == "lastall") return B(lastall);
    if (k == "lastalloverride") return B(lastalloverride);
    if (k == "data_has_rescaled_energies") return B(data_has_rescaled_energies);
    if (k == "logall") return B(logall);
    if (k == "log") return S(logstr);
    if (k == "done") return B(done);
    if (k == "grouptol") return D(grouptol);
    if (k == "width_td") return I(width_td);
    if (k == "width_custom") return I(width_custom);
    if (k == "prec_td") return I(prec_td);
    if (k == "prec_custom") return I(prec_custom);
    if (k == "prec_xy") return I(prec_xy);
    return false;
  }
  // c++/params.hpp:564-577
  void validate() {
    if (!(keep > 1)) throw std::runtime_error("keep must be > 1");
    if (keepenergy > 0.0 && keepmin > keep) throw std::runtime_error("keepmin must not exceed keep");
    if (dm_flags()) dm = true;
    if (!(Lambda > 1.0)) throw std::runtime_error("Lambda must be > 1");
    if (diag == "dsyevr" || diag == "zheevr") {
      if (!(0.0 < diagratio && diagratio <= 1.0)) throw std::runtime_error("diagratio must be in (0,1]");
      if (cfs_flags() && diagratio != 1.0)
        throw std::invalid_argument("CFS/FDM is not compatible with partial diagonalisation.");
    }
    if (dumpabs && dumpscaled) throw std::runtime_error("dumpabs=true and dumpscaled=true is a meaningless combination");
    discretization = discretization.substr(0, 1);
    if (chitp_ratio > 0.0) chitp = chitp_ratio / betabar;
  }
};

// Step (c++/step.hpp:13-87): the iteration index and the energy scales derived from it.
class Step {
 public:
  Step(const Params &P, RUNTYPE rt) : P_(P), runtype_(rt) { init(); }
  void set(int newN) {
    trueN_ = newN;
    ndxN_ = (size_t)std::max(newN, 0);
  }
  void init() { set((int)P_.Ninit); }
  void next() { trueN_++; ndxN_++; }
  size_t N() const { return ndxN_; }
  size_t ndx() const { return ndxN_; }
  double energyscale() const { return P_.SCALE(trueN_ + 1); }
  double scale() const { return P_.absolute ? 1.0 : energyscale(); }
  double unscale() const { return P_.absolute ? energyscale() : 1.0; }
  double Teff() const { return energyscale() / P_.betabar; }
  double TD_factor() const { return P_.betabar / unscale(); }
  double scT() const { return scale() / P_.T; }
  std::pair<size_t, size_t> NM() const {
    const size_t N = ndxN_ / P_.channels;
    return {N, ndxN_ - N * P_.channels};
  }
  bool N_for_merging() const {
    if (P_.NN1 || P_.NN2avg) return true;
    return P_.NN2even ? (ndxN_ % 2 == 0) : (ndxN_ % 2 == 1);
  }
  size_t firstndx() const { return P_.Ninit; }
  size_t lastndx() const { return P_.ZBW() ? P_.Ninit : P_.Nmax - 1; }
  bool first() const { return ndxN_ == firstndx(); }
  bool last(size_t N) const { return N == lastndx() || (P_.ZBW() && N == firstndx()); }
  bool last() const { return last(ndxN_); }
  bool end() const { return ndxN_ >= P_.Nmax; }
  void set_last() {
    set((int)lastndx());
    if (P_.ZBW()) set_ZBW();
  }
  void set_ZBW() {
    trueN_ = (int)P_.Ninit - 1;
    ndxN_ = P_.Ninit;
  }
  bool nrg() const { return runtype_ == RUNTYPE::NRG; }
  bool dmnrg() const { return runtype_ == RUNTYPE::DMNRG; }
  int trueN() const { return trueN_; }

 private:
  const Params &P_;
  RUNTYPE runtype_;
  int trueN_ = 0;
  size_t ndxN_ = 0;
};

}  // namespace NRG
