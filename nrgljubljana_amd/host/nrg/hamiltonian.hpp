// hamiltonian.hpp - the Hamiltonian blocks of an NRG step assembled on the GPU (nrg_hamiltonian).
//
// StepHamiltonians collects, for every block I of the step, what hamiltonian() (c++/core.hpp:72-88)
// does: the rescaled energies of the ancestors on the diagonal, then the calls the symmetry's
// make_matrix table makes, with the reference's names and argument meanings (indices i, j 1-based as
// in the tables): offdiag_function (c++/matrix.hpp:32-58), diag_function / diag_function_half
// (c++/matrix.hpp:67-92) and diag_offdiag_function (c++/matrix.hpp:96-111). assemble() builds every
// block in one device arena, which diagonalisations() hands to the eigensolver without a host copy.
#pragma once
#include <memory>

#include "diag.hpp"
#include "recalc.hpp"

namespace NRG {

class StepHamiltonians {
 public:
  // A new block with its sub-block structure (SubspaceDimensions of I); returns its index.
  size_t add_block(const Invar &I, const SubspaceDimensions &qq) {
    blocks_.push_back({I, qq});
    return blocks_.size() - 1;
  }
  // h(r,r) = scale_factor * corr(n) over sub-block i (0-based): P.nrg_step_scale_factor() times the
  // ancestor's corrected energies (diagprev.at(anc[i]).values.corr), c++/core.hpp:78-82
  void energies(size_t blk, size_t i, const std::vector<double> &corr, double scale_factor) {
    const auto &qq = blocks_.at(blk).qq;
    const size_t d = qq.rmax(i);
    NRG_ASSERT(corr.size() >= d);
    const size_t at = staged_.size();
    staged_.insert(staged_.end(), corr.begin(), corr.begin() + d);
    push(NRG_HAM_DIAG, blk, qq.offset(i), qq.offset(i), d, d, nullptr, 0, scale_factor, 0.0, 0, (int64_t)at);
  }
  // Symmetry::offdiag_function_impl: the block (i, j) of <||f||> times factor / step.scale(), or its
  // transpose into the upper block (j, i) when i > j. Skipped when a sub-block is empty
  // (offdiag_contributes); a non-finite factor throws as in the reference.
  void offdiag(size_t blk, size_t i, size_t j, const DMatrix &mat, double factor, double step_scale) {
    const auto &qq = blocks_.at(blk).qq;
    NRG_ASSERT(1 <= i && i <= qq.combs() && 1 <= j && j <= qq.combs());
    if (!qq.exists(i - 1) || !qq.exists(j - 1)) return;
    if (!std::isfinite(factor)) throw std::runtime_error("offdiag_function(): factor not finite");
    NRG_ASSERT(mat.size1() == qq.rmax(i - 1) && mat.size2() == qq.rmax(j - 1));
    const double fs = factor / step_scale;
    if (i > j)
      push(NRG_HAM_BLOCK, blk, qq.offset(j - 1), qq.offset(i - 1), qq.rmax(j - 1), qq.rmax(i - 1), mat.data(),
           (int64_t)mat.size2(), fs, 0.0, 1, -1);
    else
      push(NRG_HAM_BLOCK, blk, qq.offset(i - 1), qq.offset(j - 1), qq.rmax(i - 1), qq.rmax(j - 1), mat.data(),
           (int64_t)mat.size2(), fs, 0.0, 0, -1);
  }
  // diag_function_impl: h(j,j) += sc_zeta * (number - f * avgoccup) / step.scale() over sub-block i
  // (avgoccup = P.spin / 2; f = 1 for diag_function, 0.5 for diag_function_half)
  void diag(size_t blk, size_t i, double number, double sc_zeta, double step_scale, double avgoccup, double f = 1.0) {
    const auto &qq = blocks_.at(blk).qq;
    NRG_ASSERT(1 <= i && i <= qq.combs());
    const double shift = sc_zeta * (number - f * avgoccup) / step_scale;
    push(NRG_HAM_DIAG, blk, qq.offset(i - 1), qq.offset(i - 1), qq.rmax(i - 1), qq.rmax(i - 1), nullptr, 0, 0.0, shift,
         0, -1);
  }
  // diag_offdiag_function: h(begin1 + l, begin2 + l) += factor / step.scale() for i <= j
  void diag_offdiag(size_t blk, size_t i, size_t j, double factor, double step_scale) {
    const auto &qq = blocks_.at(blk).qq;
    NRG_ASSERT(1 <= i && i <= qq.combs() && 1 <= j && j <= qq.combs());
    if (i > j) return;
    const size_t s1 = qq.rmax(i - 1), s2 = qq.rmax(j - 1);
    if (!(s1 > 0 && s2 > 0)) return;
    NRG_ASSERT(s1 == s2);
    push(NRG_HAM_IDENT, blk, qq.offset(i - 1), qq.offset(j - 1), s1, s1, nullptr, 0, factor / step_scale, 0.0, 0, -1);
  }

  size_t size() const { return blocks_.size(); }
  const Invar &invar(size_t blk) const { return blocks_.at(blk).I; }
  size_t order(size_t blk) const { return blocks_.at(blk).qq.total(); }

  // Every block into one device arena (block b row-major, upper triangle filled as diagonalise expects)
  struct Arena {
    std::shared_ptr<DBuf<double>> buf;
    std::vector<int64_t> n, off;
  };
  Arena assemble(Device &dev) const {
    Arena a;
    int64_t tot = 0;
    for (const auto &b : blocks_) {
      a.n.push_back((int64_t)b.qq.total());
      a.off.push_back(tot);
      tot += a.n.back() * a.n.back();
    }
    a.buf = std::make_shared<DBuf<double>>((size_t)tot + 1);
    DBuf<double> st(staged_.size() + 1);
    if (!staged_.empty())
      dcopy(st.get(), staged_.data(), sizeof(double) * staged_.size(), hipMemcpyHostToDevice);
    std::vector<nrg_ham_term> t = terms_;
    for (size_t k = 0; k < t.size(); k++)
      if (stage_[k] >= 0) t[k].src = st.get() + stage_[k];
    nrg_check(nrg_hamiltonian(dev.ctx(), (int)a.n.size(), a.n.data(), a.off.data(), a.buf->get(), t.data(), (int)t.size(),
                              dev.stream()),
              "nrg_hamiltonian");
    return a;
  }

 private:
  struct Block {
    Invar I;
    SubspaceDimensions qq;
  };
  void push(int kind, size_t blk, size_t row, size_t col, size_t rows, size_t cols, const double *src, int64_t ld,
            double factor, double shift, int trans, int64_t stage) {
    nrg_ham_term t{};
    t.kind = kind;
    t.trans = trans;
    t.block = (int64_t)blk;
    t.row = (int64_t)row;
    t.col = (int64_t)col;
    t.rows = (int64_t)rows;
    t.cols = (int64_t)cols;
    t.src = src;
    t.ld = ld;
    t.factor = factor;
    t.shift = shift;
    terms_.push_back(t);
    stage_.push_back(stage);
  }
  std::vector<Block> blocks_;
  std::vector<nrg_ham_term> terms_;
  std::vector<int64_t> stage_;    // per term: offset of its energies in staged_, or -1
  std::vector<double> staged_;
};

// diagonalisations (c++/core.hpp:90-120) of a step whose Hamiltonians were assembled on the GPU: no
// host matrices, no copy; results keyed by Invar as diagnew[I] = Eigen(std::move(e), step).
inline DiagInfo diagonalisations(const StepHamiltonians &h, const DiagParams &DP, Device &dev, bool last_step = false) {
  auto a = h.assemble(dev);
  std::vector<int64_t> rows;
  for (int64_t n : a.n) rows.push_back((int64_t)DP.nr_wanted((size_t)n));
  auto res = detail::syev_arena(dev, a.buf, a.n, a.off, DP.mode(), DP.check ? &rows : nullptr);
  DiagInfo diag;
  for (size_t b = 0; b < h.size(); b++) {
    const size_t n = (size_t)res.n[b], M = DP.nr_wanted(n);
    std::vector<double> val(res.val[b].begin(), res.val[b].begin() + M);
    DView vec{res.arena->get() + res.off[b], M, n, n};
    diag[h.invar(b)] = Eigen(std::move(val), vec, res.arena, last_step);
  }
  return diag;
}

}  // namespace NRG
