// driver.hpp - an NRG calculation: the forward sweep, the density matrices and the second sweep.
//
// Follows NRG_calculation (c++/nrg-general.hpp:109-180) and the step loop of c++/core.hpp:
// nrg_loop / iterate (285-293, 234-247), do_diag (122-154), after_diag (194-231), docalc0 (251-262),
// with SubspaceStructure / TaskList (c++/subspaces.hpp, c++/core.hpp:40-70), the round-off
// correction Clusters (c++/splitting.hpp:33-73), truncation (c++/truncation.hpp:14-69), the
// operator recalculation (c++/oprecalc.hpp:70-87, 204-221), the measurements
// (c++/measurements.hpp:17-224) and the output files (c++/output.hpp, c++/outfield.hpp,
// c++/spectrum.hpp:148-227, c++/store.hpp:57-84). Every numerical step runs on the GPU through
// the library: one assembly launch (nrg_hamiltonian) and one batched eigensolve (nrg_syev_batched)
// per step, one grouped product call (nrg_gemm3_grouped) per step for every operator and for the
// <||f||> blocks, one per backward-sweep step for the density matrices, and the spectral kernels.
// Eigenvectors and density matrices stay in device memory between the sweeps (StepStore).
#pragma once
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <set>
#include <unordered_map>

#include "read_input.hpp"
#include "spectrum.hpp"
#include "store.hpp"
#include "symmetry.hpp"

namespace NRG {

// ---------------------------------------------------------------- operators
struct Operators {   // c++/operators.hpp:167-203
  Opch opch;
  std::map<std::string, MatrixElements> ops, opsp, opsg, opd, opt, opq, opot;
};

inline DMatrix clone(const DMatrix &m) {
  DMatrix c(m.size1(), m.size2());
  if (m.size1() && m.size2())
    dcopy(c.data(), m.data(), 8 * m.size1() * m.size2(), hipMemcpyDeviceToDevice);
  return c;
}
inline MatrixElements clone(const MatrixElements &m) {
  MatrixElements c;
  for (const auto &[II, mat] : m) c.emplace(II, clone(mat));
  return c;
}
inline Operators clone(const Operators &a) {
  Operators c;
  for (const auto &ch : a.opch) {
    c.opch.emplace_back();
    for (const auto &o : ch) c.opch.back().push_back(clone(o));
  }
  for (const auto &[n, m] : a.ops) c.ops[n] = clone(m);
  for (const auto &[n, m] : a.opsp) c.opsp[n] = clone(m);
  for (const auto &[n, m] : a.opsg) c.opsg[n] = clone(m);
  for (const auto &[n, m] : a.opd) c.opd[n] = clone(m);
  for (const auto &[n, m] : a.opt) c.opt[n] = clone(m);
  for (const auto &[n, m] : a.opq) c.opq[n] = clone(m);
  for (const auto &[n, m] : a.opot) c.opot[n] = clone(m);
  return c;
}
// trim_matrix (c++/numerics.hpp:267-276) of every block to the stored dimensions (MatrixElements::trim)
inline void trim(MatrixElements &m, const DiagInfo &diag) {
  for (auto &[II, mat] : m) {
    const auto [d1, d2] = diag.dims(II.first, II.second);
    if (mat.size1() == 0 || mat.size2() == 0) continue;
    NRG_ASSERT(d1 <= mat.size1() && d2 <= mat.size2());
    if (d1 == mat.size1() && d2 == mat.size2()) continue;
    DMatrix t(d1, d2);
    if (d1 && d2)
      dcopy2d(t.data(), 8 * d2, mat.data(), 8 * mat.size2(), 8 * d2, d1, hipMemcpyDeviceToDevice);
    mat = std::move(t);
  }
}
inline void trim_matrices(Operators &a, const DiagInfo &diag) {   // Operators::trim_matrices
  for (auto *set : {&a.ops, &a.opsp, &a.opsg, &a.opd, &a.opt, &a.opq, &a.opot})
    for (auto &kv : *set) trim(kv.second, diag);
}
// the diagonal of a device matrix
inline std::vector<double> diagonal(const DMatrix &m, size_t n) {
  std::vector<double> d(n + 1);
  if (n) get_diagonals({nrg_diag_item{m.data(), (int64_t)m.size2(), 0, (int64_t)n, 0}}, d);
  d.resize(n);
  return d;
}

// ---------------------------------------------------------------- output formatting
inline std::string fmt_g(double x, int prec, int width = 0) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%*.*g", width, prec, x);
  return buf;
}

// TD_generic (c++/outfield.hpp:90-118): a header of field names, then one line of values per call
class TDFile {
 public:
  TDFile(std::string filename, std::vector<std::string> fields, int prec, int width)
      : fn_(std::move(filename)), fields_(std::move(fields)), prec_(prec), width_(width) {}
  void set(const std::string &f, double x) {
    if (std::find(fields_.begin(), fields_.end(), f) == fields_.end()) throw std::runtime_error("Invalid field " + f);
    vals_[f] = x;
  }
  void save_values() {
    if (!O_.is_open()) {
      O_.open(fn_);
      O_ << '#';
      for (const auto &f : fields_) O_ << std::setw(width_) << f << " ";
      O_ << '\n';
    }
    O_ << ' ';
    for (const auto &f : fields_) O_ << std::setw(width_) << fmt_g(vals_.at(f), prec_, width_) << " ";
    O_ << '\n';
    O_.flush();
  }

 private:
  std::string fn_;
  std::vector<std::string> fields_;
  int prec_, width_;
  std::map<std::string, double> vals_;
  std::ofstream O_;
};

// ExpvOutput (c++/output.hpp:22-62): 'custom' / 'customfdm'
class ExpvFile {
 public:
  ExpvFile(const std::string &fn, std::vector<std::string> fields, const Params &P) : fields_(std::move(fields)), P_(P) {
    F_.open(fn);
    F_ << '#' << fmt_g(1, P.prec_custom, P.width_custom) << ' ';
    for (size_t c = 1; c <= fields_.size(); c++) F_ << fmt_g((double)(1 + c), P.prec_custom, P.width_custom) << ' ';
    F_ << '\n' << '#' << std::setw(P.width_custom) << "T" << ' ';
    for (const auto &f : fields_) F_ << std::setw(P.width_custom) << f << ' ';
    F_ << '\n';
  }
  void field_values(double label, const std::map<std::string, double> &m) {
    F_ << ' ' << fmt_g(label, P_.prec_custom, P_.width_custom) << ' ';
    for (const auto &f : fields_) F_ << fmt_g(m.at(f), P_.prec_custom, P_.width_custom) << ' ';
    F_ << '\n';
    F_.flush();
  }

 private:
  std::vector<std::string> fields_;
  const Params &P_;
  std::ofstream F_;
};

// ---------------------------------------------------------------- spectral functions
// One requested correlator and its algorithms (BaseSpectrum + Algo_* of c++/algo.hpp, c++/oprecalc.hpp
// prepare_spec / loopover), saved at the end of its sweep.
struct Spectrum {
  std::string name, prefix;   // "A_d-A_d", "spec"
  const MatrixElements *op1 = nullptr, *op2 = nullptr;
  int spin = 0, sign = -1;
  FactorFnc ff;
  CheckFnc cf;
  std::unique_ptr<FDMSpectrum> fdm;
  std::unique_ptr<CFSSpectrum> cfs;
  std::unique_ptr<DMNRGSpectrum> dmnrg;
  std::unique_ptr<FTSpectrum> ft;
  std::unique_ptr<MatsubaraGF> fdmmats, dmnrgmats, ftmats;
  std::unique_ptr<TempSeries> temp;
  int temp_kind = -1;
};

struct Timing {
  std::map<std::string, double> s;   // seconds per section
  std::vector<std::pair<size_t, double>> nrg_steps, dm_steps;   // wall time of every step
  std::vector<std::vector<size_t>> blocks;   // block orders diagonalised at every forward step
};

class NRGCalculation {
 public:
  NRGCalculation(const std::string &dir, const std::string &tables_dir, Device &dev, std::ostream &log = std::cout)
      : dir_(dir), dev_(dev), log_(log), P_(dir + "/param") {
    const auto t0 = clock();
    input_ = read_input_file(dir + "/data", P_.coeffactor);
    if (input_.symtype != P_.symtype) throw std::runtime_error("symtype of param and data files differ");
    P_.set_channels_and_combs(input_.channels);
    Sym_ = std::make_unique<TableSymmetry>(P_, tables_dir);
    // determine_Nmax_Nlen (c++/read-input.hpp:50-60)
    coef_.xi = input_.coef.count("xi") ? input_.coef.at("xi") : std::vector<std::vector<double>>{};
    coef_.zeta = input_.coef.count("zeta") ? input_.coef.at("zeta") : std::vector<std::vector<double>>{};
    if (input_.coef.count("xiR")) coef_.xiR = input_.coef.at("xiR");
    if (input_.coef.count("zetaR")) coef_.zetaR = input_.coef.at("zetaR");
    if (P_.tri == "cpp") throw std::runtime_error("tri=cpp (Wilson-chain tridiagonalisation) is not supported");
    if (coef_.xi.empty() || coef_.xi[0].size() != input_.Nmax + 1)
      throw std::runtime_error("coefficient table length differs from the chain length in the data file");
    P_.Nmax = input_.Nmax;
    P_.Nlen = !P_.ZBW() ? P_.Nmax : P_.Nmax + 1;
    if (P_.ZBW()) throw std::runtime_error("zero-bandwidth calculations are not supported");
    if (P_.accumulation > 0.0) throw std::runtime_error("accumulation > 0 binning is not supported");
    for (const auto &kv : P_.unused) log_ << "Unused setting: " << kv.first << "=" << kv.second << "\n";
    diag0_ = initial_diag();
    operators0_ = initial_operators();
    stats_ = RunStats(P_.Nlen, input_.GS_energy);
    store_ = std::make_unique<Store>(P_.Ninit, P_.Nlen);
    td_ = std::make_unique<TDFile>(dir + "/td", td_fields(), P_.prec_td, P_.width_td);
    timing_.s["setup"] += seconds_since(t0);
  }

  // NRG_calculation::NRG_calculation body
  void run() {
    auto diag = run_nrg(RUNTYPE::NRG);
    if (P_.dm) {
      if (P_.need_rho()) calc_rho(diag);
      if (P_.need_rhoFDM()) calc_rhoFDM();
      run_nrg(RUNTYPE::DMNRG);
    }
    if (P_.done) std::ofstream(dir_ + "/DONE");
  }
  const Timing &timing() const { return timing_; }
  const Params &params() const { return P_; }

 private:
  using clk = std::chrono::steady_clock;
  static clk::time_point clock() { return clk::now(); }
  static double seconds_since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

  struct RunStats {   // c++/stats.hpp:16-72
    double Egs = 0, total_energy = 0, GS_energy = 0, Z = 0, Zft = 0, Zgt = 0, Zchit = 0;
    std::map<std::string, double> expv, fdmexpv;
    Stats fdm;
    RunStats() = default;
    RunStats(size_t nlen, double GS0) : total_energy(GS0), fdm(nlen) {}
  };

  std::vector<std::string> td_fields() const {
    std::vector<std::string> f{"T"};
    for (const auto &s : Sym_->td_fields()) f.push_back(s);
    for (const char *s : {"<E>", "<E^2>", "C", "F", "S"}) f.push_back(s);
    return f;
  }

  // DiagInfo(std::istream&, nsubs, P) (c++/eigen.hpp:307-317): energies, identity vectors, corr = raw
  DiagInfo initial_diag() const {
    DiagInfo d;
    for (const auto &[I, e0] : input_.subspaces) {
      std::vector<double> e = e0;
      if (!(P_.data_has_rescaled_energies || P_.absolute))
        for (auto &x : e) x /= P_.SCALE((int)P_.Ninit);
      auto own = std::make_shared<DMatrix>(id_matrix(e.size()));
      Eigen eig(e, own->view(), own, P_.ZBW());
      eig.values.set_corr(e);
      eig.values.set_shift(0.0);
      eig.values.set_scale(P_.absolute ? 1.0 : P_.SCALE((int)P_.Ninit));
      d[I] = std::move(eig);
    }
    return d;
  }
  static MatrixElements to_device(const IrredMatrices &m) {
    MatrixElements d;
    for (const auto &[II, mat] : m) d.emplace(II, DMatrix(mat));
    return d;
  }
  Operators initial_operators() const {
    Operators a;
    for (const auto &ch : input_.opch) {
      a.opch.emplace_back();
      for (const auto &o : ch) a.opch.back().push_back(to_device(o));
    }
    for (const auto &[type, named] : input_.ops) {
      const auto &[name, m] = named;
      switch (type) {
        case 's': a.ops[name] = to_device(m); break;
        case 'p': a.opsp[name] = to_device(m); break;
        case 'g': a.opsg[name] = to_device(m); break;
        case 'd': a.opd[name] = to_device(m); break;
        case 't': a.opt[name] = to_device(m); break;
        case 'o': a.opot[name] = to_device(m); break;
        case 'q': a.opq[name] = to_device(m); break;
        default: throw std::runtime_error(std::string("unknown operator block ") + type);
      }
    }
    return a;
  }

  // ---------------------------------------------------------- one sweep (run_nrg)
  DiagInfo run_nrg(RUNTYPE rt) {
    runtype_ = rt;
    ops_ = clone(operators0_);
    setup_spectra(rt);
    if (rt == RUNTYPE::NRG) {
      if (P_.dumpenergies) Fenergies_.open(dir_ + "/energies.nrg");
      std::vector<std::string> f;
      for (const auto &kv : ops_.ops) f.push_back(kv.first);
      for (const auto &kv : ops_.opsg) f.push_back(kv.first);
      custom_ = std::make_unique<ExpvFile>(dir_ + "/custom", f, P_);
    } else if (P_.fdmexpv) {
      std::vector<std::string> f;
      for (const auto &kv : ops_.ops) f.push_back(kv.first);
      for (const auto &kv : ops_.opsg) f.push_back(kv.first);
      customfdm_ = std::make_unique<ExpvFile>(dir_ + "/customfdm", f, P_);
    }
    Step step(P_, rt);
    if (step.nrg() && P_.calc0) docalc0(step);
    DiagInfo diag = diag0_;
    for (step.init(); !step.end(); step.next()) {
      const auto t0 = clock();
      diag = iterate(step, diag);
      (rt == RUNTYPE::NRG ? timing_.nrg_steps : timing_.dm_steps).emplace_back(step.ndx(), seconds_since(t0));
    }
    step.set((int)step.lastndx());
    log_ << "Total energy: " << std::setprecision(18) << stats_.total_energy << "\n";
    stats_.GS_energy = stats_.total_energy;
    if (step.nrg()) {
      store_->shift_abs_energies(stats_.GS_energy);
      if (P_.dumpabsenergies) dump_all_absolute_energies();
      if (P_.dumpsubspaces) dump_subspaces();
    }
    save_spectra();
    Fenergies_.close();
    annotated_.close();
    custom_.reset();
    customfdm_.reset();
    return diag;
  }

  // docalc0 (c++/core.hpp:251-262): measurements on the initial data, as step N = Ninit - 1
  void docalc0(Step &step) {
    step.set((int)P_.Ninit - 1);
    perform_basic_measurements(step, diag0_);
    Store empty(0, 0);
    calculate_spectral_and_expv(step, diag0_, empty);
  }

  // SubspaceStructure (c++/core.hpp:40-70)
  SubspaceStructure subspace_structure(const DiagInfo &diagprev) const {
    std::set<Invar> subs;
    for (const auto &[I, eig] : diagprev)
      for (const auto &In : Sym_->new_subspaces(I))
        if (Sym_->Invar_allowed(In)) subs.insert(In);
    SubspaceStructure ss;
    for (const auto &I : subs) {
      const auto anc = Sym_->ancestors(I);
      std::vector<size_t> dims;
      for (size_t i = 0; i < anc.size(); i++)
        dims.push_back(Sym_->triangle_inequality(I, anc[i], Sym_->QN_subspace(i)) ? diagprev.size_subspace(anc[i]) : 0);
      ss[I] = SubspaceDimensions(dims, anc);
    }
    return ss;
  }

  // iterate (c++/core.hpp:234-247)
  DiagInfo iterate(const Step &step, const DiagInfo &diagprev) {
    const SubspaceStructure substruct = subspace_structure(diagprev);
    DiagInfo diag = do_diag(step, diagprev, substruct);
    after_diag(step, diag, substruct);
    trim_matrices(ops_, diag);
    diag.clear_eigenvectors();
    return diag;
  }

  struct NotEnough {};

  // do_diag (c++/core.hpp:122-154)
  DiagInfo do_diag(const Step &step, const DiagInfo &diagprev, const SubspaceStructure &substruct) {
    double diagratio = P_.diagratio;
    for (;;) {
      DiagInfo diag;
      const auto t0 = clock();
      if (step.nrg()) {
        // TaskList (c++/subspaces.hpp:103-123): non-empty subspaces, largest first
        std::vector<std::pair<size_t, Invar>> tasks;
        for (const auto &[I, rm] : substruct)
          if (rm.total()) tasks.emplace_back(rm.total(), I);
        std::sort(tasks.begin(), tasks.end(), std::greater<>());
        StepHamiltonians H;
        for (const auto &[n, I] : tasks) {
          const size_t blk = H.add_block(I, substruct.at(I));
          Sym_->hamiltonian(H, blk, I, substruct.at(I), diagprev, ops_.opch, coef_, step.N(), step.scale());
        }
        DiagParams DP(P_.diag == "default" ? "default" : P_.diag, diagratio);   // check_diag on the device
        std::vector<size_t> sizes;
        double n3 = 0;
        for (const auto &t : tasks) {
          sizes.push_back(t.first);
          n3 += std::pow((double)t.first, 3);
        }
        timing_.blocks.push_back(sizes);
        log_ << "N=" << step.ndx() << " blocks " << tasks.size() << " max " << (tasks.empty() ? 0 : tasks.front().first)
             << " sum n^3 " << n3 << "\n";
        diag = diagonalisations(H, DP, dev_, step.last());
      } else {
        diag = steps_.load(step.ndx());
        diag.subtract_GS_energy(stats_.GS_energy);
      }
      timing_.s["diag"] += seconds_since(t0);
      stats_.Egs = diag.Egs_subtraction();
      clusters(diag, P_.fixeps);
      try {
        truncate_prepare(step, diag);
        return diag;
      } catch (NotEnough &) {
        log_ << "Insufficient number of states computed.\n";
        if (!(step.nrg() && P_.restart)) return diag;
        diagratio = std::min(diagratio * P_.restartfactor, 1.0);
        log_ << "Restarting this iteration step. diagratio=" << diagratio << "\n";
      }
    }
  }

  // Clusters (c++/splitting.hpp:33-73): eigenvalues within epsilon of a cluster's lowest value are
  // replaced by it in the corrected energies
  static void clusters(DiagInfo &diag, double epsilon) {
    const auto energies = diag.sorted_energies_rel_zero();
    NRG_ASSERT(!energies.empty());
    std::unordered_map<double, double> mapping;
    double e0 = energies.front();
    size_t i0 = 0, size = 1;
    for (size_t i = 0; i < energies.size(); i++) {
      if ((energies[i] - e0) < epsilon) {
        size++;
      } else {
        if (size > 1) {
          bool split = false;
          for (size_t a = i0; a < i && !split; a++)
            for (size_t b = a + 1; b < i; b++)
              if (energies[a] != energies[b]) { split = true; break; }
          if (split)
            for (size_t j = i0 + 1; j < i; j++)
              if (energies[j] != energies[i0]) mapping.insert({energies[j], energies[i0]});
        }
        e0 = energies[i];
        i0 = i;
        size = 1;
      }
    }
    for (auto &[I, eig] : diag) {
      auto v = eig.values.all_rel_zero();
      for (auto &r : v)
        if (const auto m = mapping.find(r); m != mapping.end()) r = m->second;
      eig.values.set_corr(std::move(v));
    }
  }

  // highest_retained_energy + truncate_prepare (c++/truncation.hpp:14-69)
  void truncate_prepare(const Step &step, DiagInfo &diag) {
    const auto energies = diag.sorted_energies_corr();
    const size_t total = energies.size();
    NRG_ASSERT(total != 0 && energies.front() == 0.0);
    size_t nrkeep = P_.keep;
    if (P_.keepenergy > 0.0) {
      const double ke = P_.keepenergy * step.unscale();
      size_t cnt = 0;
      for (double e : energies) cnt += e <= ke;
      nrkeep = std::clamp<size_t>(1 + cnt, P_.keepmin, P_.keep);
    }
    if (P_.safeguard > 0.0) {
      size_t extra = 0;
      while (nrkeep < total && (energies[nrkeep] - energies[nrkeep - 1]) <= P_.safeguard && extra < P_.safeguardmax) {
        nrkeep++;
        extra++;
      }
    }
    nrkeep = std::clamp<size_t>(nrkeep, 1, total);
    const double Emax = energies[nrkeep - 1];
    for (auto &[I, eig] : diag) {
      size_t n = 0;
      if (step.last() && P_.keep_all_states_in_last_step())
        n = eig.getnrcomputed();
      else
        for (double e : eig.values.all_corr()) n += e <= Emax;
      eig.truncate_prepare(n);
    }
    size_t nrall = 0, nrkept = 0;
    for (const auto &[I, eig] : diag) {
      nrall += eig.getdim();
      nrkept += eig.getnrkept();
    }
    log_ << "N=" << step.ndx() << " Emax=" << Emax / step.unscale() << " kept " << nrkept << " of " << nrall << "\n";
    for (const auto &[I, eig] : diag)
      if (eig.getnrkept() == eig.getnrcomputed() && eig.values.highest_corr() != Emax && eig.getnrcomputed() < eig.getdim())
        throw NotEnough{};
  }

  // after_diag (c++/core.hpp:194-231)
  void after_diag(const Step &step, DiagInfo &diag, const SubspaceStructure &substruct) {
    // stats.update
    stats_.total_energy += stats_.Egs * step.scale();
    if (step.nrg()) {
      for (auto &[I, eig] : diag) {   // calc_abs_energies
        eig.values.set_scale(step.scale());
        eig.values.set_T_shift(stats_.total_energy);
      }
      if (P_.dm) steps_.save(step.ndx(), diag);
      perform_basic_measurements(step, diag);
    }
    split_in_blocks(diag, substruct);
    if (P_.do_recalc_all(runtype_)) {
      recalculate_operators(step, diag);
      calculate_spectral_and_expv(step, diag, *store_);
    }
    diag.truncate_perform();
    Subs subs;
    for (const auto &[I, eig] : diag) {
      const auto it = substruct.find(I);
      subs[I] = Sub{eig, it == substruct.end() ? SubspaceDimensions() : it->second, step.last()};
    }
    (*store_)[step.ndx()] = std::move(subs);
    if (!step.last() && step.nrg()) {   // recalc_irreducible (c++/oprecalc.hpp:204-221)
      const auto t0 = clock();
      Gemm3Batch batch(dev_);
      Opch opch = Sym_->recalc_irreduc(diag, batch);
      batch.flush();
      ops_.opch = std::move(opch);
      timing_.s["recalc f"] += seconds_since(t0);
    }
    if (P_.do_recalc_kept(runtype_)) {
      recalculate_operators(step, diag);
      calculate_spectral_and_expv(step, diag, *store_);
    }
    if (P_.do_recalc_none()) calculate_spectral_and_expv(step, diag, *store_);
  }

  // Oprecalc::Ops::do_s / do_g and the spectra's operator list
  bool do_s(const std::string &name, const Step &step) const {
    if (step.nrg()) return true;
    if (step.dmnrg() && P_.fdmexpv && step.N() <= P_.fdmexpvn) return true;
    return needed_.count({"s", name}) > 0;
  }
  bool do_g(const std::string &name, const Step &step) const {
    if (step.nrg()) return true;
    if (step.dmnrg() && P_.fdmexpv && step.N() <= P_.fdmexpvn) return true;
    return needed_.count({"g", name}) > 0;
  }

  // Oprecalc::recalculate_operators (c++/oprecalc.hpp:70-87): every selected operator of the step in
  // one grouped product call
  void recalculate_operators(const Step &step, const DiagInfo &diag) {
    const auto t0 = clock();
    Gemm3Batch batch(dev_);
    std::map<std::string, MatrixElements> s, g, d, t;
    for (const auto &[name, m] : ops_.ops)
      if (do_s(name, step)) s[name] = Sym_->recalc_singlet(diag, m, batch);
    for (const auto &[name, m] : ops_.opd)
      if (needed_.count({"d", name})) d[name] = Sym_->recalc_doublet(diag, m, batch);
    for (const auto &[name, m] : ops_.opt)
      if (needed_.count({"t", name})) t[name] = Sym_->recalc_triplet(diag, m, batch);
    for (const auto &[name, m] : ops_.opsp)
      if (needed_.count({"p", name})) throw std::runtime_error("odd-parity singlets do not exist for " + P_.symtype);
    for (const auto &[name, m] : ops_.opsg)
      if (do_g(name, step)) {
        g[name] = Sym_->recalc_singlet(diag, m, batch);
        Sym_->recalc_global(diag, name, g[name], batch);
      }
    for (const auto &[name, m] : ops_.opq)
      if (needed_.count({"q", name})) throw std::runtime_error("quadruplet operators are not supported");
    for (const auto &[name, m] : ops_.opot)
      if (needed_.count({"ot", name})) throw std::runtime_error("orbital triplet operators are not supported");
    batch.flush();
    for (auto &[name, m] : ops_.ops) m = s.count(name) ? std::move(s[name]) : MatrixElements();
    for (auto &[name, m] : ops_.opd) m = d.count(name) ? std::move(d[name]) : MatrixElements();
    for (auto &[name, m] : ops_.opt) m = t.count(name) ? std::move(t[name]) : MatrixElements();
    for (auto &[name, m] : ops_.opsg) m = g.count(name) ? std::move(g[name]) : MatrixElements();
    for (auto &[name, m] : ops_.opsp) m = MatrixElements();
    for (auto &[name, m] : ops_.opq) m = MatrixElements();
    for (auto &[name, m] : ops_.opot) m = MatrixElements();
    timing_.s["recalc"] += seconds_since(t0);
  }

  // perform_basic_measurements (c++/measurements.hpp:229-234)
  void perform_basic_measurements(const Step &step, const DiagInfo &diag) {
    dump_energies(step.ndx(), diag);
    annotated_dump(step, diag);
    calculate_TD(step, diag);
  }

  // calculate_TD (c++/measurements.hpp:168-186)
  void calculate_TD(const Step &step, const DiagInfo &diag) {
    const double rf = step.TD_factor();
    double Z = 0, E = 0, E2 = 0;
    std::map<Invar, double> sumZ;
    for (const auto &[I, eig] : diag) {
      const auto v = eig.values.all_rel_zero();
      double z = 0, e = 0, e2 = 0;
      for (double x : v) {
        const double w = std::exp(-rf * x);
        z += w;
        e += rf * x * w;
        e2 += std::pow(rf * x, 2) * w;
      }
      const double m = Sym_->mult(I);
      Z += m * z;
      E += m * e;
      E2 += m * e2;
      // calculate_Z over the measured states, for the symmetry's fields
      double zm = 0;
      const auto &c = eig.values.all_corr();
      for (size_t i = 0; i < eig.getnrstored() && i < c.size(); i++) zm += std::exp(-rf * c[i]);
      sumZ[I] = m * zm;
    }
    stats_.Z = Z;
    td_->set("T", step.Teff());
    td_->set("<E>", E / Z);
    td_->set("<E^2>", E2 / Z);
    td_->set("C", E2 / Z - std::pow(E / Z, 2));
    td_->set("F", -std::log(Z));
    td_->set("S", E / Z + std::log(Z));
    for (const auto &[f, x] : Sym_->td_values(sumZ, Z)) td_->set(f, x);
    td_->save_values();
  }

  // calculate_spectral_and_expv (c++/measurements.hpp:197-224)
  void calculate_spectral_and_expv(const Step &step, const DiagInfo &diag, const Store &store) {
    const auto t0 = clock();
    DensMatElements rho, rhoFDM;
    if (step.dmnrg()) {
      if (P_.need_rho()) {
        rho = steps_.load_rho(step.ndx(), "rho");
        check_trace_rho(rho, Sym_->multfnc());
      }
      if (P_.need_rhoFDM()) rhoFDM = steps_.load_rho(step.ndx(), "rhofdm");
    }
    const auto mult = Sym_->multfnc();
    stats_.Zft = grand_canonical_Z(step.scT(), diag, mult);
    if (!P_.specgt.empty() || !P_.speci1t.empty() || !P_.speci2t.empty())
      stats_.Zgt = grand_canonical_Z(1.0 / P_.gtp, diag, mult);
    if (!P_.specchit.empty()) stats_.Zchit = grand_canonical_Z(1.0 / P_.chitp, diag, mult);
    for (auto &sp : spectra_) calc_spectrum(sp, step, diag, rho, rhoFDM);
    timing_.s["spec"] += seconds_since(t0);
    if (step.nrg()) {   // measure_singlet
      const double f = step.TD_factor();
      double Z = 0;
      for (const auto &[I, eig] : diag) {
        double s = 0;
        for (size_t i = 0; i < eig.getnrstored(); i++) s += std::exp(-f * eig.values.corr(i));
        Z += mult(I) * s;
      }
      for (const auto *set : {&ops_.ops, &ops_.opsg})
      for (const auto &[name, m] : *set) {
        double tr = 0;
        for (const auto &[I, eig] : diag) {
          const size_t n = eig.getnrstored();
          const auto it = m.find({I, I});
          if (it == m.end()) throw std::runtime_error("measure_singlet: missing block of " + name);
          NRG_ASSERT(it->second.size1() == n && it->second.size2() == n);
          const auto d = diagonal(it->second, n);
          double s = 0;
          for (size_t i = 0; i < n; i++) s += std::exp(-f * eig.values.corr(i)) * d[i];
          tr += mult(I) * s;
        }
        stats_.expv[name] = tr / Z;
      }
      custom_->field_values(step.Teff(), stats_.expv);
    }
    if (step.dmnrg() && P_.fdmexpv && step.N() == P_.fdmexpvn) {   // measure_singlet_fdm
      for (const auto *set : {&ops_.ops, &ops_.opsg})
      for (const auto &[name, m] : *set) {
        double tr = 0;
        for (const auto &[I, r] : rhoFDM) {
          const size_t k = store[step.ndx()].at(I).kept();
          if (!k) continue;
          const Matrix R = download(r.view().sub({0, k}, {0, k}));
          const Matrix O = download(m.at({I, I}).view().sub({0, k}, {0, k}));
          double s = 0;
          for (size_t i = 0; i < k; i++)
            for (size_t j = 0; j < k; j++) s += R(i, j) * O(j, i);
          tr += mult(I) * s;
        }
        stats_.fdmexpv[name] = tr;
      }
      customfdm_->field_values(P_.T, stats_.fdmexpv);
    }
  }

  // Oprecalc constructor (c++/oprecalc.hpp:139-193): the requested spectra of this sweep
  void setup_spectra(RUNTYPE rt) {
    spectra_.clear();
    needed_.clear();
    const auto tokens = [](const std::string &s) {
      std::set<std::string> t;
      std::istringstream ss(s);
      std::string w;
      while (ss >> w) t.insert(w);
      return t;
    };
    const TableSymmetry *S = Sym_.get();
    const FactorFnc corr_ff = [S](const Invar &, const Invar &I1) { return S->mult(I1); };
    const FactorFnc spin_ff = [S](const Invar &Ip, const Invar &I1) { return S->dynamicsusceptibility_factor(Ip, I1); };
    const FactorFnc spec_ff = [S](const Invar &Ip, const Invar &I1) { return S->specdens_factor(Ip, I1); };
    const CheckFnc trivial = nullptr;
    const CheckFnc spincheck = [S](const Invar &I1, const Invar &Ip, int spin) { return S->check_SPIN(I1, Ip, spin); };
    using Set = std::map<std::string, MatrixElements>;
    auto loop = [&](Set &set1, Set &set2, const std::set<std::string> &want, const std::string &prefix,
                    const std::string &type1, const std::string &type2, const FactorFnc &ff, const CheckFnc &cf,
                    int sign, int spin) {
      for (auto &[n1, op1] : set1)
        for (auto &[n2, op2] : set2) {
          const std::string name = n1 + "-" + n2 + (spin == 0 ? "" : (spin == 1 ? "-u" : "-d"));
          if (!want.count(name)) continue;
          if (prepare(prefix, name, &op1, &op2, ff, cf, sign, spin, rt)) {
            needed_.insert({type1, n1});
            needed_.insert({type2, n2});
          }
        }
    };
    const auto specs = tokens(P_.specs), specchit = tokens(P_.specchit);
    loop(ops_.ops, ops_.ops, specs, "corr", "s", "s", corr_ff, trivial, +1, 0);
    loop(ops_.opsp, ops_.opsp, specs, "corr", "p", "p", corr_ff, trivial, +1, 0);
    loop(ops_.opsg, ops_.opsg, specs, "corr", "g", "g", corr_ff, trivial, +1, 0);
    loop(ops_.ops, ops_.opsg, specs, "corr", "s", "g", corr_ff, trivial, +1, 0);
    loop(ops_.opsg, ops_.ops, specs, "corr", "g", "s", corr_ff, trivial, +1, 0);
    loop(ops_.ops, ops_.ops, specchit, "chit", "s", "s", corr_ff, trivial, +1, 0);
    loop(ops_.ops, ops_.opsg, specchit, "chit", "s", "g", corr_ff, trivial, +1, 0);
    loop(ops_.opsg, ops_.ops, specchit, "chit", "g", "s", corr_ff, trivial, +1, 0);
    loop(ops_.opsg, ops_.opsg, specchit, "chit", "g", "g", corr_ff, trivial, +1, 0);
    loop(ops_.opt, ops_.opt, tokens(P_.spect), "spin", "t", "t", spin_ff, trivial, +1, 0);
    const int vmin = Sym_->isfield() ? -1 : 0, vmax = Sym_->isfield() ? 1 : 0;
    for (int spin = vmin; spin <= vmax; spin += 2) {
      loop(ops_.opd, ops_.opd, tokens(P_.specd), "spec", "d", "d", spec_ff, spincheck, -1, spin);
      loop(ops_.opd, ops_.opd, tokens(P_.specgt), "gt", "d", "d", spec_ff, spincheck, -1, spin);
      loop(ops_.opd, ops_.opd, tokens(P_.speci1t), "i1t", "d", "d", spec_ff, spincheck, -1, spin);
      loop(ops_.opd, ops_.opd, tokens(P_.speci2t), "i2t", "d", "d", spec_ff, spincheck, -1, spin);
    }
    if (!P_.specq.empty() || !P_.specot.empty()) throw std::runtime_error("specq/specot are not supported");
  }

  // prepare_spec (c++/oprecalc.hpp:100-133)
  bool prepare(const std::string &prefix, const std::string &name, const MatrixElements *op1, const MatrixElements *op2,
               const FactorFnc &ff, const CheckFnc &cf, int sign, int spin, RUNTYPE rt) {
    Spectrum sp;
    sp.name = name;
    sp.prefix = prefix;
    sp.op1 = op1;
    sp.op2 = op2;
    sp.ff = ff;
    sp.cf = cf;
    sp.sign = sign;
    sp.spin = spin;
    if (prefix == "gt" || prefix == "i1t" || prefix == "i2t" || prefix == "chit") {
      if (rt != RUNTYPE::NRG) return false;
      sp.temp_kind = prefix == "gt" ? NRG_TEMP_GT : prefix == "i1t" ? NRG_TEMP_I1T : prefix == "i2t" ? NRG_TEMP_I2T : NRG_TEMP_CHIT;
      sp.temp = std::make_unique<TempSeries>(dev_, sp.temp_kind);
      spectra_.push_back(std::move(sp));
      return true;
    }
    const BinGrid grid = BinGrid::standard(P_.emin, P_.emax, P_.SCALE(0), P_.last_step_scale(), (double)P_.bins,
                                           P_.discard_immediately);
    NN2Window w;
    w.E0 = P_.getE0();
    w.Efactor = P_.getEfactor();
    w.NNtanh = P_.NNtanh;
    w.ZBW = P_.ZBW();
    w.NN2avg = P_.NN2avg;
    bool any = false;
    if (rt == RUNTYPE::NRG) {
      if (P_.finite) { sp.ft = std::make_unique<FTSpectrum>(dev_, grid, sign, P_.T, w); any = true; }
      if (P_.finitemats) { sp.ftmats = std::make_unique<MatsubaraGF>(dev_, P_.mats, sign, P_.T); any = true; }
    } else {
      if (P_.dmnrg) { sp.dmnrg = std::make_unique<DMNRGSpectrum>(dev_, grid, sign, w); any = true; }
      if (P_.dmnrgmats) { sp.dmnrgmats = std::make_unique<MatsubaraGF>(dev_, P_.mats, sign, P_.T); any = true; }
      if (P_.cfs_flags()) { sp.cfs = std::make_unique<CFSSpectrum>(dev_, grid, sign); any = true; }
      if (P_.fdm || P_.fdmgt || P_.fdmls) { sp.fdm = std::make_unique<FDMSpectrum>(dev_, grid, sign, P_.T); any = true; }
      if (P_.fdmmats) { sp.fdmmats = std::make_unique<MatsubaraGF>(dev_, P_.mats, sign, P_.T); any = true; }
    }
    if (any) spectra_.push_back(std::move(sp));
    return any;
  }

  // BaseSpectrum::calc for every algorithm of one correlator
  void calc_spectrum(Spectrum &sp, const Step &step, const DiagInfo &diag, const DensMatElements &rho,
                     const DensMatElements &rhoFDM) {
    const MatrixElements &o1 = *sp.op1, &o2 = *sp.op2;
    if (sp.temp) {
      const double temperature = (sp.temp_kind == NRG_TEMP_CHIT ? P_.chitp : P_.gtp) * step.scale();
      sp.temp->calc(diag, o1, o2, temperature, sp.temp_kind == NRG_TEMP_CHIT ? stats_.Zchit : stats_.Zgt, sp.ff, sp.cf,
                    sp.spin);
      return;
    }
    if (sp.ft) sp.ft->calc(diag, o1, o2, stats_.Zft, step.scale(), step.N_for_merging(), step.first(), step.last(), sp.ff,
                           sp.cf, sp.spin);
    if (sp.ftmats) sp.ftmats->calc_ft(diag, o1, o2, stats_.Zft, sp.ff, sp.cf, sp.spin);
    if (sp.dmnrg)
      sp.dmnrg->calc(diag, o1, o2, rho, step.scale(), step.N_for_merging(), step.first(), step.last(), sp.ff, sp.cf,
                     sp.spin);
    if (sp.dmnrgmats) sp.dmnrgmats->calc_dmnrg(diag, o1, o2, rho, sp.ff, sp.cf, sp.spin);
    if (sp.cfs) sp.cfs->calc(diag, o1, o2, rho, step.last(), P_.T, stats_.Zft, sp.ff, sp.cf, sp.spin);
    if (sp.fdm) sp.fdm->calc(diag, o1, o2, rhoFDM, stats_.fdm.wnfactor.at(step.ndx()), sp.ff, sp.cf, sp.spin);
    if (sp.fdmmats) sp.fdmmats->calc_fdm(diag, o1, o2, rhoFDM, stats_.fdm.wnfactor.at(step.ndx()), sp.ff, sp.cf, sp.spin);
  }

  // SpectrumRealFreq::save (c++/spectrum.hpp:82-88, 148-209), GFMatsubara::save, TempDependence::save
  void save_real(const std::string &algo, const Spectrum &sp, const BinGrid &g, const SpectrumBins &b) {
    const std::string fn = dir_ + "/" + sp.prefix + "_" + algo + "_dens_" + sp.name;
    if (P_.savebins) {
      std::ofstream F(fn + ".bin", std::ios::binary);
      const auto [pe, pw] = trim(g, b.pos, P_.discard_trim);
      const auto [ne, nw] = trim(g, b.neg, P_.discard_trim);
      auto put = [&](double e, double w) {
        F.write((const char *)&e, sizeof e);
        F.write((const char *)&w, sizeof w);
        if (P_.reim) {
          const double z = 0.0;
          F.write((const char *)&z, sizeof z);
        }
      };
      for (size_t i = 0; i < pe.size(); i++) put(pe[i], pw[i]);
      for (size_t i = 0; i < ne.size(); i++) put(-ne[i], nw[i]);
    }
    if (P_.broaden) {
      const double omega0 = P_.omega0 < 0.0 ? P_.omega0_ratio * P_.T : P_.omega0;
      const auto mesh = make_mesh(P_.broaden_max, P_.get_broaden_min(), P_.broaden_ratio);
      const Curve c = continuous(dev_, g, b, P_.discard_trim, P_.alpha, omega0, mesh);
      std::ofstream F(fn + ".dat");
      F << std::setprecision(P_.prec_xy);
      for (const auto &[e, w] : c) {
        F << e << " " << w;
        if (P_.reim) F << " " << 0;
        F << "\n";
      }
    }
  }
  void save_mats(const std::string &algo, const Spectrum &sp, const MatsubaraGF &gf) {
    std::ofstream F(dir_ + "/" + sp.prefix + "_" + algo + "_dens_" + sp.name + ".dat");
    F << std::setprecision(P_.prec_xy);
    for (size_t n = 0; n < gf.values().size(); n++) {
      const auto z = gf.values()[n];
      F << gf.frequency(n) << " " << z.real() << " " << (std::abs(z.imag()) > std::abs(z.real()) * 1e-10 ? z.imag() : 0)
        << "\n";
    }
  }
  void save_spectra() {
    for (const auto &sp : spectra_) {
      if (sp.temp) {
        auto pts = sp.temp->points();
        std::sort(pts.begin(), pts.end());
        const char *algo = sp.temp_kind == NRG_TEMP_CHIT ? "CHIT" : sp.temp_kind == NRG_TEMP_GT ? "GT" :
                           sp.temp_kind == NRG_TEMP_I1T ? "I1T" : "I2T";
        std::ofstream F(dir_ + "/" + sp.prefix + "_" + algo + "_dens_" + sp.name + ".dat");
        F << std::setprecision(P_.prec_xy);
        for (const auto &[t, v] : pts) {
          F << t << " " << v;
          if (P_.reim) F << " " << 0;
          F << "\n";
        }
        continue;
      }
      if (sp.ft) save_real("FT", sp, sp.ft->grid(), sp.ft->spectrum());
      if (sp.ftmats) save_mats("FTmats", sp, *sp.ftmats);
      if (sp.dmnrg) save_real("DMNRG", sp, sp.dmnrg->grid(), sp.dmnrg->spectrum());
      if (sp.dmnrgmats) save_mats("DMNRGmats", sp, *sp.dmnrgmats);
      if (sp.cfs) {
        if (P_.cfs) save_real("CFS", sp, sp.cfs->grid(), sp.cfs->total());
        if (P_.cfsgt) save_real("CFSgt", sp, sp.cfs->grid(), sp.cfs->greater());
        if (P_.cfsls) save_real("CFSls", sp, sp.cfs->grid(), sp.cfs->lesser());
      }
      if (sp.fdm) {
        if (P_.fdm) save_real("FDM", sp, sp.fdm->grid(), sp.fdm->total());
        if (P_.fdmgt) save_real("FDMgt", sp, sp.fdm->grid(), sp.fdm->gt());
        if (P_.fdmls) save_real("FDMls", sp, sp.fdm->grid(), sp.fdm->ls());
      }
      if (sp.fdmmats) save_mats("FDMmats", sp, *sp.fdmmats);
    }
    spectra_.clear();
  }

  // ---------------------------------------------------------- density matrices between the sweeps
  std::vector<DiagInfo> loaded_steps() {
    std::vector<DiagInfo> d(P_.Nmax);
    for (size_t N = P_.Ninit + 1; N < P_.Nmax; N++) d[N] = steps_.load(N);
    return d;
  }
  static std::vector<const DiagInfo *> pointers(const std::vector<DiagInfo> &d) {
    std::vector<const DiagInfo *> p;
    for (const auto &x : d) p.push_back(&x);
    return p;
  }
  // NRG_calculation::calc_rho
  void calc_rho(const DiagInfo &diag) {
    const auto t0 = clock();
    Step step(P_, RUNTYPE::NRG);
    step.set_last();
    DensMatElements rho = init_rho(step.scT(), diag, Sym_->multfnc());
    DensMatElements keep;
    for (const auto &[I, m] : rho) keep.emplace(I, clone(m));
    steps_.save_rho(step.lastndx(), "rho", std::move(keep));
    const auto d = loaded_steps();
    calc_densitymatrix(rho, *store_, *Sym_, pointers(d), P_.Ninit, P_.Nmax, dev_,
                       [this](size_t N, const DensMatElements &r) {
                         DensMatElements c;
                         for (const auto &[I, m] : r) c.emplace(I, clone(m));
                         steps_.save_rho(N, "rho", std::move(c));
                       });
    timing_.s["DM"] += seconds_since(t0);
  }
  // NRG_calculation::calc_rhoFDM
  void calc_rhoFDM() {
    const auto t0 = clock();
    Step step(P_, RUNTYPE::NRG);
    step.set_last();
    stats_.fdm.GS_energy = stats_.GS_energy;
    {
      SubTimer ts(&timing_.s, "FDM.stats");
      calc_ZnD(*store_, stats_.fdm, *Sym_, P_.T);
      fdm_thermodynamics(*store_, stats_.fdm, *Sym_, P_.T);
    }
    TDFile tdfdm(dir_ + "/tdfdm", {"T", "E_fdm", "C_fdm", "F_fdm", "S_fdm"}, P_.prec_td, P_.width_td);
    tdfdm.set("T", P_.T);
    tdfdm.set("F_fdm", stats_.fdm.F_fdm);
    tdfdm.set("E_fdm", stats_.fdm.E_fdm);
    tdfdm.set("C_fdm", stats_.fdm.C_fdm);
    tdfdm.set("S_fdm", stats_.fdm.S_fdm);
    tdfdm.save_values();
    DensMatElements rhoFDM = init_rho_FDM(step.lastndx(), *store_, stats_.fdm, Sym_->multfnc(), P_.T);
    DensMatElements keep;
    for (const auto &[I, m] : rhoFDM) keep.emplace(I, clone(m));
    steps_.save_rho(step.lastndx(), "rhofdm", std::move(keep));
    std::vector<DiagInfo> d;
    {
      SubTimer tl(&timing_.s, "FDM.load");
      d = loaded_steps();
    }
    calc_fulldensitymatrix(rhoFDM, *store_, stats_.fdm, *Sym_, pointers(d), P_.Ninit, P_.Nmax, P_.T, dev_,
                           [this](size_t N, const DensMatElements &r) {
                             DensMatElements c;
                             for (const auto &[I, m] : r) c.emplace(I, clone(m));
                             steps_.save_rho(N, "rhofdm", std::move(c));
                           },
                           &timing_.s);
    timing_.s["FDM"] += seconds_since(t0);
  }

  // ---------------------------------------------------------- plain-text dumps
  void dump_energies(size_t N, const DiagInfo &diag) {   // Output::dump_energies, DiagInfo::dump_energies
    if (!Fenergies_.is_open()) return;
    Fenergies_ << "\n===== Iteration number: " << N << "\n";
    for (const auto &[I, eig] : diag) {
      Fenergies_ << "Subspace: " << I.str() << "\n";
      for (double x : eig.values.all_rel()) Fenergies_ << x << " ";
      Fenergies_ << "\n";
    }
  }
  // Annotated::dump (c++/output.hpp:65-118)
  void annotated_dump(const Step &step, const DiagInfo &diag) {
    if (!P_.dumpannotated) return;
    if (!annotated_.is_open()) {
      annotated_.open(dir_ + "/annotated.dat");
      annotated_ << std::setprecision((int)P_.dumpprecision);
    }
    std::vector<std::pair<double, Invar>> seznam;
    for (const auto &[I, eig] : diag)
      for (double e : eig.values.all_rel_zero()) seznam.emplace_back(e, I);
    std::sort(seznam.begin(), seznam.end());
    size_t len = std::min<size_t>(seznam.size(), P_.dumpannotated);
    while (len < seznam.size() - 1 && my_fcmp(seznam[len].first, seznam[len - 1].first, P_.grouptol, P_.grouptol) == 0)
      len++;
    const auto scale = [&](double x) {
      return x * (P_.dumpscaled ? 1.0 : step.scale()) + (P_.dumpabs ? stats_.total_energy : 0.0);
    };
    if (P_.dumpgroups) {
      for (size_t i = 0; i < len;) {
        const auto [e0, I0] = seznam[i];
        annotated_ << scale(e0);
        std::vector<std::string> qn;
        size_t deg = 0;
        while (i < len && my_fcmp(seznam[i].first, e0, P_.grouptol, P_.grouptol) == 0) {
          qn.push_back(seznam[i].second.str());
          deg += (size_t)Sym_->mult(seznam[i].second);
          i++;
        }
        std::sort(qn.begin(), qn.end());
        for (const auto &s : qn) annotated_ << " (" << s << ")";
        annotated_ << " [" << deg << "]\n";
      }
    } else {
      seznam.resize(len);
      for (const auto &[e, I] : seznam) annotated_ << scale(e) << " " << I.str() << "\n";
    }
    annotated_ << "\n";
    annotated_.flush();
  }
  void dump_subspaces() {   // Store::dump_subspaces
    std::ofstream O(dir_ + "/subspaces.dat");
    for (size_t N = store_->Nbegin; N < store_->Nend; N++) {
      O << "Iteration " << N << "\n" << "len_dm=" << (*store_)[N].size() << "\n";
      for (const auto &[I, sub] : (*store_)[N]) O << "I=" << I.str() << " kept=" << sub.kept() << " total=" << sub.total() << "\n";
      O << "\n";
    }
  }
  void dump_all_absolute_energies() {   // Store::dump_abs_G
    std::ofstream F(dir_ + "/absolute_energies.dat");
    for (size_t N = store_->Nbegin; N < store_->Nend; N++) {
      F << "\n===== Iteration number: " << N + 1 << "\n";
      for (const auto &[I, ds] : (*store_)[N]) {
        F << "Subspace: " << I.str() << "\n";
        for (double x : ds.eig.values.all_abs_G()) F << x << " ";
        F << "\n";
      }
    }
  }

  std::string dir_;
  Device &dev_;
  std::ostream &log_;
  Params P_;
  InputData input_;
  std::unique_ptr<TableSymmetry> Sym_;
  Coef coef_;
  DiagInfo diag0_;
  Operators operators0_, ops_;
  RunStats stats_;
  std::unique_ptr<Store> store_;
  StepStore steps_;
  RUNTYPE runtype_ = RUNTYPE::NRG;
  std::list<Spectrum> spectra_;
  std::set<std::pair<std::string, std::string>> needed_;
  std::unique_ptr<TDFile> td_;
  std::unique_ptr<ExpvFile> custom_, customfdm_;
  std::ofstream Fenergies_, annotated_;
  Timing timing_;
};

}  // namespace NRG
