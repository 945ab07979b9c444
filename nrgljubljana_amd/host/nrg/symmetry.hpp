// symmetry.hpp - the QS and QSZ symmetry types, driven by their coefficient tables.
//
// Follows Symmetry<S> (c++/symmetry.hpp:57-242) with SymmetryQS (c++/symmetry/sym-QS-impl.hpp) and
// SymmetryQSZ (c++/symmetry/sym-QSZ-impl.hpp): the ancestor differences In and shell quantum numbers
// QN (load()), multiplicities, the triangle inequality, allowed subspaces, the thermodynamic fields,
// the spectral-function factors, make_matrix (the OFFDIAG / DIAG rows of the Hamiltonian, as terms of
// a StepHamiltonians batch) and the recalculation drivers recalc_irreduc, recalc_doublet,
// recalc_triplet and recalc_singlet (c++/symmetry/nrg-recalc-QS.hpp:31-666,
// c++/symmetry/nrg-recalc-QSZ.hpp), whose table rows are queued into one Gemm3Batch per operator.
//
// The tables are the reference's data, converted to plain text (tables/<SYM>-<n>ch.tab, written by
// tables/make_tables.py): each factor or quantum number is an expression evaluated here (expr.hpp).
#pragma once
#include <fstream>
#include <set>
#include <sstream>

#include "expr.hpp"
#include "hamiltonian.hpp"
#include "measurements.hpp"
#include "params.hpp"

namespace NRG {

// The coefficient tables of one symmetry type and channel count: section name -> rows of fields.
class CoefTables {
 public:
  CoefTables() = default;
  explicit CoefTables(const std::string &path) {
    std::ifstream F(path);
    if (!F) throw std::runtime_error("Can't open coefficient tables " + path);
    std::string line, cur;
    while (std::getline(F, line)) {
      if (line.empty() || line[0] == '#') continue;
      if (line[0] == '@') {
        cur = line.substr(1);
        sections_[cur];
        continue;
      }
      if (cur.empty()) throw std::runtime_error(path + ": row outside a section");
      std::vector<std::string> fields;
      std::string f;
      std::istringstream ss(line);
      while (std::getline(ss, f, '|')) {
        const auto b = f.find_first_not_of(" \t"), e = f.find_last_not_of(" \t");
        fields.push_back(b == std::string::npos ? "" : f.substr(b, e - b + 1));
      }
      sections_[cur].push_back(std::move(fields));
    }
  }
  bool has(const std::string &name) const { return sections_.count(name) > 0; }
  const std::vector<std::vector<std::string>> &at(const std::string &name) const {
    const auto it = sections_.find(name);
    if (it == sections_.end()) throw std::runtime_error("coefficient table '" + name + "' not found");
    return it->second;
  }

 private:
  std::map<std::string, std::vector<std::vector<std::string>>> sections_;
};

// Recalc_f {i1, ip, factor} and Recalc {i1, ip, IN1, INp, factor} rows as expressions
struct ExprRowF {
  size_t i1, ip;
  Expr factor;
};
struct ExprRow {
  size_t i1, ip;
  std::vector<Expr> IN1, INp;
  Expr factor;
};
struct OffdiagRow {
  size_t i, j, ch;
  Expr factor;
};
struct DiagRow {
  size_t i, ch;
  Expr number;
};

inline bool su2_triangle_inequality(int ss1, int ss2, int ss3) {   // c++/symmetry.hpp:31-33
  return (std::abs(ss1 - ss2) <= ss3 - 1) && (std::abs(ss2 - ss3) <= ss1 - 1) && (std::abs(ss3 - ss1) <= ss2 - 1);
}

using OpchChannel = std::vector<MatrixElements>;
using Opch = std::vector<OpchChannel>;   // [channel][index < perchannel]

// Wilson-chain coefficient tables (c++/coef.hpp): xi(N, ch), zeta(N, ch), one table per coefficient
// channel, N = 0..Nmax.
struct Coef {
  std::vector<std::vector<double>> xi, zeta, xiR, zetaR;
  static double at(const std::vector<std::vector<double>> &t, size_t N, size_t ch, const char *what) {
    if (ch >= t.size() || N >= t[ch].size())
      throw std::runtime_error(std::string("coefficient ") + what + "(" + std::to_string(N) + "," + std::to_string(ch) +
                               ") out of range");
    return t[ch][N];
  }
  double get_xi(size_t N, size_t ch) const { return at(xi, N, ch, "xi"); }
  double get_zeta(size_t N, size_t ch) const { return at(zeta, N, ch, "zeta"); }
  double get_xiR(size_t N, size_t ch) const { return at(xiR, N, ch, "xiR"); }
  double get_zetaR(size_t N, size_t ch) const { return at(zetaR, N, ch, "zetaR"); }
};

class TableSymmetry : public Symmetry {
 public:
  TableSymmetry(const Params &P, const std::string &tables_dir) : P_(P) {
    if (P.symtype == "QS") qsz_ = false;
    else if (P.symtype == "QSZ") qsz_ = true;
    else throw std::runtime_error("symmetry type " + P.symtype + " is not supported (QS and QSZ are)");
    if (P.channels < 1 || P.channels > 2) throw std::runtime_error("QS/QSZ tables are built for 1 and 2 channels");
    if (P.substeps) throw std::runtime_error("substeps=true is not supported");
    if (qsz_ && P.polarized) throw std::runtime_error("polarized=true is not supported");
    const size_t ch = P.channels;
    tab_ = CoefTables(tables_dir + "/" + P.symtype + "-" + std::to_string(ch) + "ch.tab");
    // load(): In[i], QN[i] (1-based in the tables), then erase_first()
    In_.resize(P.combs);
    QN_.resize(P.combs);
    for (const auto &r : tab_.at("In")) In_.at(std::stoul(r.at(0)) - 1) = Invar{std::stoi(r.at(1)), std::stoi(r.at(2))};
    for (const auto &r : tab_.at("QN")) QN_.at(std::stoul(r.at(0)) - 1) = Invar{std::stoi(r.at(1)), std::stoi(r.at(2))};
    for (const auto &r : tab_.at("offdiag")) offdiag_.push_back(offdiag_row(r));
    for (const auto &r : tab_.at("diag")) diag_.push_back({std::stoul(r.at(0)), std::stoul(r.at(1)), Expr(r.at(2))});
    if (P.rungs) {
      for (const auto &r : tab_.at("offdiag-mix")) mix_.push_back(offdiag_row(r));
      for (const auto &r : tab_.at("runghop")) {   // RUNGHOP(i, j, factor)
        OffdiagRow o{std::stoul(r.at(0)), std::stoul(r.at(1)), 0, Expr(r.at(2))};
        rung_.push_back(std::move(o));
      }
    }
    for (const char *name : {"doubletp", "doubletm", "tripletp", "triplets", "tripletm"})
      for (const auto &r : tab_.at(name)) {
        ExprRow e{std::stoul(r.at(0)), std::stoul(r.at(1)), {Expr(r.at(2)), Expr(r.at(3))}, {Expr(r.at(4)), Expr(r.at(5))},
                  Expr(r.at(6))};
        general_[name].push_back(std::move(e));
      }
    for (const char *name : {"SZtot", "Q1u", "Q1d", "Qdiff", "Q1", "Q2", "Qtot"}) {
      const std::string sec = std::string("global:") + name;
      if (!tab_.has(sec)) continue;
      for (const auto &r : tab_.at(sec)) global_[name].push_back({std::stoul(r.at(0)), std::stoul(r.at(1)), Expr(r.at(3))});
    }
    const char *fnames[] = {"spinupa", "spindowna", "spinupb", "spindownb"};
    for (const char *name : fnames) {
      if (!tab_.has(name)) continue;
      for (const auto &r : tab_.at(name)) f_[name].push_back({std::stoul(r.at(0)), std::stoul(r.at(1)), Expr(r.at(2))});
    }
  }

  // --- quantum numbers (c++/symmetry.hpp:90-140)
  const std::string &symtype() const { return P_.symtype; }
  bool isfield() const { return qsz_; }
  size_t nr_combs() const override { return P_.combs; }
  Invar ancestor(const Invar &I, size_t i) const { return combine(I, In_.at(i)); }
  std::vector<Invar> ancestors(const Invar &I) const {
    std::vector<Invar> a;
    for (size_t i = 0; i < P_.combs; i++) a.push_back(ancestor(I, i));
    return a;
  }
  std::vector<Invar> new_subspaces(const Invar &I) const override {
    std::vector<Invar> a;
    for (size_t i = 0; i < P_.combs; i++) a.push_back(Invar{I.qn[0] - In_[i].qn[0], I.qn[1] - In_[i].qn[1]});
    return a;
  }
  const Invar &QN_subspace(size_t i) const { return QN_.at(i); }
  double mult(const Invar &I) const override { return qsz_ ? 1.0 : (double)I.qn[1]; }
  bool triangle_inequality(const Invar &I1, const Invar &I2, const Invar &I3) const {
    if (qsz_) return I1.qn[0] == I2.qn[0] + I3.qn[0] && I1.qn[1] == I2.qn[1] + I3.qn[1];
    return I1.qn[0] == I2.qn[0] + I3.qn[0] && su2_triangle_inequality(I1.qn[1], I2.qn[1], I3.qn[1]);
  }
  bool Invar_allowed(const Invar &I) const { return qsz_ ? true : I.qn[1] > 0; }
  Invar invar_singlet() const { return qsz_ ? Invar{0, 0} : Invar{0, 1}; }

  // --- thermodynamics (calculate_TD of the symmetry classes)
  std::vector<std::string> td_fields() const {
    if (qsz_) return {"<Sz^2>", "<Sz>", "<Q>", "<Q^2>"};
    return {"<Sz^2>", "<Q>", "<Q^2>"};
  }
  // Tr-weighted sums over the subspaces; sumZ(I) = mult(I) sum exp(-factor E_corr) over measured states
  std::map<std::string, double> td_values(const std::map<Invar, double> &sumZ, double Z) const {
    double trSZ = 0, trSZ2 = 0, trQ = 0, trQ2 = 0;
    for (const auto &[I, s] : sumZ) {
      const int q = I.qn[0];
      trQ += s * q;
      trQ2 += s * q * q;
      if (qsz_) {
        const double sz = (I.qn[1] - 1.0) / 2.0;
        trSZ += s * sz;
        trSZ2 += s * std::pow(sz, 2);
      } else {
        const int ss = I.qn[1];
        trSZ += s * (ss * ss - 1) / 12.;
      }
    }
    if (qsz_) return {{"<Sz^2>", trSZ2 / Z}, {"<Sz>", trSZ / Z}, {"<Q>", trQ / Z}, {"<Q^2>", trQ2 / Z}};
    return {{"<Sz^2>", trSZ / Z}, {"<Q>", trQ / Z}, {"<Q^2>", trQ2 / Z}};
  }

  // --- spectral-function factors (c++/symmetry/sym-QS-impl.hpp:41-65; defaults of Symmetry for QSZ)
  double specdens_factor(const Invar &Ip, const Invar &I1) const {
    if (qsz_) return 1.0;
    if (I1.qn[0] - Ip.qn[0] != 1) throw std::runtime_error("specdens_factor: charge difference must be 1");
    const int ssp = Ip.qn[1], ss1 = I1.qn[1];
    if (std::abs(ss1 - ssp) != 1) throw std::runtime_error("specdens_factor: spin difference must be 1");
    const double Sp = (ssp - 1.0) / 2.0;
    return ss1 == ssp + 1 ? Sp + 1.0 : Sp;
  }
  double dynamicsusceptibility_factor(const Invar &Ip, const Invar &I1) const {
    if (qsz_) return 1.0;
    if (I1.qn[0] != Ip.qn[0]) throw std::runtime_error("dynamicsusceptibility_factor: charges must agree");
    const int ssp = Ip.qn[1], ss1 = I1.qn[1];
    if (ss1 == ssp + 2) return 1. + (ssp - 1) / 3.;
    if (ss1 == ssp) return ssp / 3.;
    if (ss1 == ssp - 2) return (-2. + ssp) / 3.;
    throw std::runtime_error("dynamicsusceptibility_factor: no match");
  }
  bool check_SPIN(const Invar &I1, const Invar &Ip, int SPIN) const {
    if (!qsz_) return true;
    return I1.qn[1] - Ip.qn[1] == SPIN;
  }

  // --- Hamiltonian (hamiltonian(), c++/core.hpp:72-88, with make_matrix of the symmetry type):
  // diagonal sqrt(Lambda) E_corr of the ancestors, then the OFFDIAG and DIAG rows (and with
  // rungs=true the OFFDIAG_MIX and RUNGHOP rows), in table order.
  void hamiltonian(StepHamiltonians &H, size_t blk, const Invar &I, const SubspaceDimensions &qq,
                   const DiagInfo &diagprev, const Opch &opch, const Coef &coef, size_t N, double step_scale) const {
    const auto anc = ancestors(I);
    for (size_t i = 0; i < P_.combs; i++)
      if (qq.rmax(i)) H.energies(blk, i, diagprev.at(anc[i]).values.all_corr(), P_.nrg_step_scale_factor());
    ExprVars v;
    v["ss"] = I.qn[1];
    v["q"] = I.qn[0];
    auto offdiag = [&](const OffdiagRow &r, double c) {
      if (!qq.exists(r.i - 1) || !qq.exists(r.j - 1)) return;   // offdiag_contributes: factor not evaluated
      const auto &m = opch.at(r.ch).at(0);
      const auto it = m.find({anc.at(r.i - 1), anc.at(r.j - 1)});
      if (it == m.end())
        throw std::runtime_error("offdiag_function(): matrix not found " + std::to_string(r.i) + " " + std::to_string(r.j) +
                                 " " + std::to_string(r.ch));
      H.offdiag(blk, r.i, r.j, it->second, r.factor(v) * c, step_scale);
    };
    auto diagrow = [&](const DiagRow &r) {
      H.diag(blk, r.i, r.number(v), coef.get_zeta(N + 1, r.ch), step_scale, (double)P_.spin / 2);
    };
    if (P_.channels == 2 && !qsz_) {   // the two-channel QS code includes the diagonal rows first
      for (const auto &r : diag_) diagrow(r);
      for (const auto &r : offdiag_) offdiag(r, coef.get_xi(N, r.ch));
    } else {
      for (const auto &r : offdiag_) offdiag(r, coef.get_xi(N, r.ch));
      for (const auto &r : diag_) diagrow(r);
    }
    if (P_.rungs) {
      for (const auto &r : mix_) offdiag(r, coef.get_xiR(N, r.ch));
      for (const auto &r : rung_) H.diag_offdiag(blk, r.i, r.j, r.factor(v) * coef.get_zetaR(N + 1, 0), step_scale);
    }
  }

  // recalc_global (c++/symmetry/nrg-recalc-QS.hpp:680-741, nrg-recalc-QSZ.hpp:497-547): a global
  // singlet operator gets, on top of its recalc_singlet value, value * U_I(i1) U_I(ip)^T per table row
  // (recalc1_global, c++/recalc.hpp:111-122) in every subspace
  bool has_global(const std::string &name) const { return global_.count(name) > 0; }
  void recalc_global(const DiagInfo &diag, const std::string &name, MatrixElements &cnew, Gemm3Batch &batch) const {
    const auto it = global_.find(name);
    if (it == global_.end()) throw std::runtime_error("no recalc_global table for operator " + name);
    for (const auto &[I1, eig] : diag) {
      auto &cn = cnew[{I1, I1}];
      if (!eig.getnrstored()) continue;
      for (const auto &r : it->second) recalc1_global(diag, I1, cn, r.i1, r.ip, r.factor(ExprVars{}), batch);
    }
  }

  // --- recalculation drivers. Each returns the new matrix elements; the products are queued in
  // `batch` and evaluated by the caller's flush().
  // recalc_irreduc: <I1||f^dag||Ip> of every channel, I1 = Ip + (1, +-1) (nrg-recalc-QS.hpp:191-453)
  Opch recalc_irreduc(const DiagInfo &diag, Gemm3Batch &batch) const {
    Opch opch(P_.channels);
    for (auto &c : opch) c.resize(P_.perchannel);
    const char *up[] = {"spinupa", "spinupb"}, *down[] = {"spindowna", "spindownb"};
    for (const auto &[Ip, eig] : diag) {
      ExprVars v;
      v["qp"] = Ip.qn[0];
      v[qsz_ ? "sszp" : "ssp"] = Ip.qn[1];
      for (int s : {+1, -1}) {
        const Invar I1{Ip.qn[0] + 1, Ip.qn[1] + s};
        if (!diag.count(I1) || !diag.at(I1).getnrstored() || !eig.getnrstored()) continue;
        for (size_t ch = 0; ch < P_.channels; ch++) {
          std::vector<Recalc_f> t;
          for (const auto &r : f_.at(s > 0 ? up[ch] : down[ch])) t.push_back({r.i1, r.ip, r.factor(v)});
          opch[ch][0][{I1, Ip}] = recalc_f(diag, I1, Ip, t, batch);
        }
      }
    }
    return opch;
  }
  // recalc_doublet (nrg-recalc-QS.hpp:31-182): Ip = I1 - (1, -+1) with the doubletp / doubletm tables
  MatrixElements recalc_doublet(const DiagInfo &diag, const MatrixElements &cold, Gemm3Batch &batch) const {
    MatrixElements cnew;
    for (const auto &[I1, eig] : diag) {
      recalc_pair(diag, cold, I1, Invar{I1.qn[0] - 1, I1.qn[1] + 1}, "doubletp", cnew, batch);
      recalc_pair(diag, cold, I1, Invar{I1.qn[0] - 1, I1.qn[1] - 1}, "doubletm", cnew, batch);
    }
    return cnew;
  }
  // recalc_triplet (nrg-recalc-QS.hpp:496-666): Ip = I1, I1 + (0, 2), I1 - (0, 2)
  MatrixElements recalc_triplet(const DiagInfo &diag, const MatrixElements &cold, Gemm3Batch &batch) const {
    MatrixElements cnew;
    for (const auto &[I1, eig] : diag) {
      recalc_pair(diag, cold, I1, I1, "triplets", cnew, batch);
      recalc_pair(diag, cold, I1, Invar{I1.qn[0], I1.qn[1] + 2}, "tripletp", cnew, batch);
      recalc_pair(diag, cold, I1, Invar{I1.qn[0], I1.qn[1] - 2}, "tripletm", cnew, batch);
    }
    return cnew;
  }
  // recalc_singlet (c++/symmetry.hpp:185-200): the diagonal table {i, i, anc_i, anc_i, 1} for every
  // subspace (even parity only: QS and QSZ have no parity quantum number)
  MatrixElements recalc_singlet(const DiagInfo &diag, const MatrixElements &nold, Gemm3Batch &batch) const {
    MatrixElements nnew;
    for (const auto &[I, eig] : diag) {
      std::vector<Recalc> t;
      for (size_t i = 0; i < P_.combs; i++) {
        const Invar anc = ancestor(I, i);
        if (Invar_allowed(anc)) t.push_back({i + 1, i + 1, anc, anc, 1.0});
      }
      nnew.emplace(Twoinvar{I, I}, recalc_general(diag, nold, I, I, t, batch));
    }
    return nnew;
  }

 private:
  static Invar combine(const Invar &a, const Invar &b) { return Invar{a.qn[0] + b.qn[0], a.qn[1] + b.qn[1]}; }
  static OffdiagRow offdiag_row(const std::vector<std::string> &r) {
    return OffdiagRow{std::stoul(r.at(0)), std::stoul(r.at(1)), std::stoul(r.at(2)), Expr(r.at(3))};
  }
  // One RECALC block of the generated code: if both subspaces exist with stored states, evaluate the
  // table rows for I1 and queue recalc_general (rows with a disallowed ancestor are skipped).
  void recalc_pair(const DiagInfo &diag, const MatrixElements &cold, const Invar &I1, const Invar &Ip,
                   const std::string &table, MatrixElements &cnew, Gemm3Batch &batch) const {
    if (!diag.count(I1) || !diag.count(Ip)) return;
    if (!diag.at(I1).getnrstored() || !diag.at(Ip).getnrstored()) return;
    ExprVars v;
    v["q1"] = I1.qn[0];
    v[qsz_ ? "ssz1" : "ss1"] = I1.qn[1];
    std::vector<Recalc> t;
    for (const auto &r : general_.at(table)) {
      const Invar IN1{(int)std::lround(r.IN1[0](v)), (int)std::lround(r.IN1[1](v))};
      const Invar INp{(int)std::lround(r.INp[0](v)), (int)std::lround(r.INp[1](v))};
      if (!Invar_allowed(IN1) || !Invar_allowed(INp)) continue;
      if (IN1 != ancestor(I1, r.i1 - 1) || INp != ancestor(Ip, r.ip - 1))
        throw std::runtime_error("recalc table " + table + ": row does not match the ancestors");
      t.push_back({r.i1, r.ip, IN1, INp, r.factor(v)});
    }
    cnew.emplace(Twoinvar{I1, Ip}, recalc_general(diag, cold, I1, Ip, t, batch));
  }

  const Params &P_;
  bool qsz_ = false;
  CoefTables tab_;
  std::vector<Invar> In_, QN_;
  std::vector<OffdiagRow> offdiag_, mix_, rung_;
  std::vector<DiagRow> diag_;
  std::map<std::string, std::vector<ExprRow>> general_;
  std::map<std::string, std::vector<ExprRowF>> f_, global_;
};

}  // namespace NRG
