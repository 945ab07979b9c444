// diag.hpp - block diagonalisation through the batched GPU eigensolver.
//
// diagonalise() keeps the signature and contract of the reference (c++/diag.hpp:261-292): the matrix
// has its upper triangle filled (is_matrix_upper), is consumed, and the eigenpairs come back ascending
// with eigenvectors as rows, checked by check_diag (c++/numerics.hpp:340-404). diagonalisations()
// is the batching point one level up (c++/core.hpp:90-120): every block of a step in one
// nrg_syev_batched call, the eigenvectors left in device memory for recalc and the density matrices.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "eigen.hpp"

namespace NRG {

// DiagParams (c++/params.hpp:664-685). diag selects the solver form: "default", "dsyev", "dsyevd" and
// "dsyevr" run the FP64 form ("dsyevr" keeps the lowest ceil(diagratio * dim) pairs, as the reference's
// RANGE='I' call does); "mixed" runs the FP32 + FP64-refinement form.
class DiagParams {
 public:
  std::string diag = "default";
  double diagratio = 1.0;
  bool logall = false;
  std::string logstr;
  bool check = true;   // run check_diag on the results (the reference always does)
  DiagParams() = default;
  DiagParams(std::string d, double ratio) : diag(std::move(d)), diagratio(ratio) {}
  bool logletter(char c) const { return logall || logstr.find(c) != std::string::npos; }
  int mode() const {
    if (diag == "default" || diag == "dsyev" || diag == "dsyevd" || diag == "dsyevr") return NRG_SYEV_FP64;
    if (diag == "mixed") return NRG_SYEV_MIXED;
    throw std::runtime_error("DiagParams: unknown diag method '" + diag + "'");
  }
  size_t nr_wanted(size_t dim) const {   // c++/diag.hpp:123-128
    if (diag != "dsyevr" || diagratio == 1.0) return dim;
    return std::clamp<size_t>((size_t)std::ceil(diagratio * (double)dim), 1, dim);
  }
};

inline bool is_matrix_upper(const Matrix &m) {   // lower triangle must be zero (c++/numerics.hpp)
  if (!is_square(m)) return false;
  for (size_t i = 1; i < m.size1(); i++)
    for (size_t j = 0; j < i; j++)
      if (m(i, j) != 0.0) return false;
  return true;
}

// c++/numerics.hpp:342-367
inline bool is_unitary(const Matrix &vec, double NORMALIZATION_EPSILON = 1e-12, double ORTHOGONALITY_EPSILON = 1e-10) {
  const size_t M = nrvec(vec), d = dim(vec);
  for (size_t r = 0; r < M; r++) {
    double s = 0;
    for (size_t j = 0; j < d; j++) s += vec(r, j) * vec(r, j);
    if (!num_equal(std::abs(s), 1.0, NORMALIZATION_EPSILON)) return false;
  }
  for (size_t r1 = 0; r1 < M; r1++)
    for (size_t r2 = r1 + 1; r2 < M; r2++) {
      double s = 0;
      for (size_t j = 0; j < d; j++) s += vec(r1, j) * vec(r2, j);
      if (!num_equal(std::abs(s), 0.0, ORTHOGONALITY_EPSILON)) return false;
    }
  return true;
}

inline void check_diag(const std::vector<double> &val, const Matrix &vec) {   // c++/numerics.hpp:398-403
  NRG_ASSERT(val.size() == nrvec(vec));
  for (double v : val) NRG_ASSERT(std::isfinite(v));
  NRG_ASSERT(is_unitary(vec));
}
inline void RawEigen::check_diag() const { NRG::check_diag(val, vec); }

namespace detail {
// One nrg_syev_batched call over `mats` (each square, upper triangle filled). Returns the device arena
// holding the eigenvectors (row b of block b's region = eigenvector, ascending) and the eigenvalues.
struct BatchResult {
  std::shared_ptr<DBuf<double>> arena;
  std::vector<int64_t> n, off;
  std::vector<std::vector<double>> val;
};
// check_diag (c++/numerics.hpp:340-404) of the first rows[b] eigenpairs of every block, on the device
// (nrg_check_diag_batched); thresholds as is_unitary's num_equal: |<v,v> - 1| < 1e-12 (|<v,v>| + 1) and
// |<v_r,v_s>| < 1e-10
inline void check_diag_device(Device &dev, const std::vector<int64_t> &n, const std::vector<int64_t> &rows,
                              const std::vector<int64_t> &off, const double *V, const std::vector<int64_t> &offw,
                              const double *w, size_t *first_bad = nullptr) {
  const size_t nb = n.size();
  std::vector<double> dn(nb), dof(nb);
  std::vector<int> info(nb, 0);
  const int r = nrg_check_diag_batched(dev.ctx(), (int)nb, n.data(), rows.data(), off.data(), V, offw.data(), w, 2e-12,
                                       1e-10, dn.data(), dof.data(), info.data(), dev.stream());
  if (r == NRG_OK) return;
  if (r != NRG_E_NOCONV) nrg_check(r, "nrg_check_diag_batched");
  for (size_t b = 0; b < nb; b++)
    if (info[b]) {
      if (first_bad) *first_bad = b;
      char msg[200];
      std::snprintf(msg, sizeof msg, "check_diag failed for block %zu (dim %lld): max |<v,v> - 1| = %.3e, max |<v_r,v_s>| = %.3e%s",
                    b, (long long)n[b], dn[b], dof[b], info[b] == 3 ? ", non-finite eigenvalue" : "");
      throw std::runtime_error(msg);
    }
}

// nrg_syev_batched on blocks already in a device arena (block b: n[b] x n[b] at off[b]); with
// check_rows, check_diag_device on the first check_rows[b] eigenpairs of each block
inline BatchResult syev_arena(Device &dev, std::shared_ptr<DBuf<double>> arena, std::vector<int64_t> n,
                              std::vector<int64_t> off, int mode, const std::vector<int64_t> *check_rows = nullptr) {
  BatchResult r;
  const size_t nb = n.size();
  r.arena = std::move(arena);
  r.n = std::move(n);
  r.off = std::move(off);
  std::vector<int64_t> offw(nb);
  int64_t totw = 0;
  for (size_t b = 0; b < nb; b++) {
    offw[b] = totw;
    totw += r.n[b];
  }
  DBuf<double> w((size_t)totw + 1);
  std::vector<int> info(nb, 0);
  // diagnostic: with NRG_DUMP_DIR set, the input blocks are kept on the host, and a block that fails
  // check_diag is written to NRG_DUMP_DIR/check_diag_block_<b>.bin (int64 order, then the row-major
  // matrix as handed to the eigensolver)
  const char *dump_dir = check_rows ? std::getenv("NRG_DUMP_DIR") : nullptr;
  std::vector<double> input;
  if (dump_dir) {
    const int64_t tot = nb ? r.off[nb - 1] + r.n[nb - 1] * r.n[nb - 1] : 0;
    input.resize((size_t)tot);
    if (tot) dcopy(input.data(), r.arena->get(), sizeof(double) * (size_t)tot, hipMemcpyDeviceToHost);
  }
  nrg_check(nrg_syev_batched(dev.ctx(), (int)nb, r.n.data(), r.off.data(), r.arena->get(), offw.data(), w.get(), mode,
                             info.data(), dev.stream()),
            "nrg_syev_batched");
  for (size_t b = 0; b < nb; b++)
    if (info[b] != 0)   // the reference throws on INFO != 0 (c++/diag.hpp:61-82)
      throw std::runtime_error("diagonalisation of block " + std::to_string(b) + " failed, info=" + std::to_string(info[b]));
  if (check_rows) {
    size_t bad = nb;
    try {
      check_diag_device(dev, r.n, *check_rows, r.off, r.arena->get(), offw, w.get(), &bad);
    } catch (const std::runtime_error &) {
      if (dump_dir && bad < nb) {
        const std::string fn = std::string(dump_dir) + "/check_diag_block_" + std::to_string(bad) + ".bin";
        if (FILE *f = std::fopen(fn.c_str(), "wb")) {
          std::fwrite(&r.n[bad], sizeof(int64_t), 1, f);
          std::fwrite(input.data() + r.off[bad], sizeof(double), (size_t)(r.n[bad] * r.n[bad]), f);
          std::fclose(f);
        }
      }
      throw;
    }
  }
  std::vector<double> wh((size_t)totw);
  dcopy(wh.data(), w.get(), sizeof(double) * (size_t)totw, hipMemcpyDeviceToHost);
  r.val.resize(nb);
  for (size_t b = 0; b < nb; b++) r.val[b].assign(wh.begin() + offw[b], wh.begin() + offw[b] + r.n[b]);
  return r;
}

inline BatchResult syev_batch(Device &dev, const std::vector<const Matrix *> &mats, int mode,
                              const std::vector<int64_t> *check_rows = nullptr) {
  const size_t nb = mats.size();
  std::vector<int64_t> n(nb), off(nb);
  int64_t tot = 0;
  for (size_t b = 0; b < nb; b++) {
    NRG_ASSERT(is_square(*mats[b]));
    n[b] = (int64_t)mats[b]->size1();
    off[b] = tot;
    tot += n[b] * n[b];
  }
  std::vector<double> host((size_t)tot);
  for (size_t b = 0; b < nb; b++)
    if (n[b]) std::memcpy(host.data() + off[b], mats[b]->data(), sizeof(double) * (size_t)(n[b] * n[b]));
  auto arena = std::make_shared<DBuf<double>>((size_t)tot + 1);
  dcopy(arena->get(), host.data(), sizeof(double) * (size_t)tot, hipMemcpyHostToDevice);
  return syev_arena(dev, std::move(arena), std::move(n), std::move(off), mode, check_rows);
}
}  // namespace detail

// Drop-in for diagonalise(M &m, const DiagParams &DP, int myrank) (c++/diag.hpp:261-292).
inline RawEigen diagonalise(Matrix &m, const DiagParams &DP, int myrank = -1) {
  (void)myrank;
  NRG_ASSERT(is_matrix_upper(m));
  auto res = detail::syev_batch(Device::default_device(), {&m}, DP.mode());
  const size_t n = m.size1(), M = DP.nr_wanted(n);
  RawEigen d(M, n);
  Matrix all = download(DView{res.arena->get(), n, n, n});
  for (size_t i = 0; i < M; i++) {
    d.val[i] = res.val[0][i];
    for (size_t j = 0; j < n; j++) d.vec(i, j) = all(i, j);
  }
  NRG_ASSERT(d.getnrcomputed() > 0);
  if (DP.check) d.check_diag();
  m = Matrix();   // consumed, as LAPACK destroys it
  return d;
}

// Batched diagonalisation of all blocks of a step (c++/core.hpp:90-120): one GPU call; results keyed by
// Invar as in diagnew[I] = Eigen(std::move(e), step). `tasks` are consumed.
inline DiagInfo diagonalisations(std::vector<std::pair<Invar, Matrix>> &tasks, const DiagParams &DP, Device &dev,
                                 bool last_step = false) {
  std::vector<const Matrix *> mats;
  for (auto &t : tasks) {
    NRG_ASSERT(is_matrix_upper(t.second));
    mats.push_back(&t.second);
  }
  std::vector<int64_t> rows;
  for (const Matrix *m : mats) rows.push_back((int64_t)DP.nr_wanted(m->size1()));
  auto res = detail::syev_batch(dev, mats, DP.mode(), DP.check ? &rows : nullptr);
  DiagInfo diag;
  for (size_t b = 0; b < tasks.size(); b++) {
    const size_t n = (size_t)res.n[b], M = DP.nr_wanted(n);
    std::vector<double> val(res.val[b].begin(), res.val[b].begin() + M);
    DView vec{res.arena->get() + res.off[b], M, n, n};
    diag[tasks[b].first] = Eigen(std::move(val), vec, res.arena, last_step);
  }
  tasks.clear();
  return diag;
}

}  // namespace NRG
