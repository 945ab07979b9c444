// Host-layer tests on the GPU: the reference-shaped C++ interfaces (diagonalise, diagonalisations,
// product/transform/rotate, recalc_general, the FDM density-matrix sweep, the FDM spectral step)
// checked against the oracle (LAPACK dsyev / BLAS dgemm restatement in oracle/liboracle.so) and
// against the reference's own invariants (check_diag, Tr rho = sum w_n, sum w_n = 1).
// Usage: test_host_gpu --blas=<path of the BLAS/LAPACK library the oracle loads> [--filter=...]
#include <random>

#include "minitest.hpp"
#include "nrg/nrg.hpp"

extern "C" {
int oracle_init(const char *blas_path);
void oracle_blas_threads(int n);
int oracle_syev_batched(int nblk, const int64_t *n, const int64_t *off_H, double *H, const int64_t *off_w, double *w,
                        int threads, int *info);
int oracle_gemm3(const nrg_gemm3_term *terms, int nterms);
int oracle_fdm_bin(const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double wnf, int sign,
                   double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg);
int oracle_dmnrg_bin(const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double emin, double emax, int sign,
                     double *out_pos, double *out_neg);
int oracle_cfs_bin(const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z, int last, int sign,
                   double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg);
int oracle_ft_bin(const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z, int sign,
                  double *out_pos, double *out_neg);
int oracle_mats(int mode, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double wnf, double Z, int sign,
                double *G);
double oracle_temp_sum(int kind, const nrg_fdm_pair *pairs, int npairs, double temperature, double Z);
void oracle_merge_nn2(const nrg_bin_grid *g, double *full, const double *cs, double Emin, double Ex, double Emax, int first,
                      int last, int zbw, double NNtanh, int nn2avg);
int oracle_broaden(const double *pos_e, const double *pos_w, int64_t npos, const double *neg_e, const double *neg_w,
                   int64_t nneg, const double *mesh, int64_t nmesh, double alpha, double omega0, double *out_pos,
                   double *out_neg);
}

using namespace NRG;

static std::mt19937_64 rng(20261203);

static Matrix random_upper(size_t n, double scale = 1.0) {   // symmetric, upper triangle stored
  std::normal_distribution<double> nd(0.0, scale);
  Matrix m(n, n);
  for (size_t i = 0; i < n; i++)
    for (size_t j = i; j < n; j++) m(i, j) = nd(rng);
  return m;
}
static Matrix random_matrix(size_t r, size_t c) {
  std::normal_distribution<double> nd(0.0, 1.0);
  Matrix m(r, c);
  for (size_t i = 0; i < r; i++)
    for (size_t j = 0; j < c; j++) m(i, j) = nd(rng);
  return m;
}
static double max_abs_diff(const Matrix &a, const Matrix &b) {
  NRG_ASSERT(a.size1() == b.size1() && a.size2() == b.size2());
  double e = 0;
  for (size_t i = 0; i < a.size1(); i++)
    for (size_t j = 0; j < a.size2(); j++) e = std::max(e, std::abs(a(i, j) - b(i, j)));
  return e;
}
static double max_abs(const Matrix &a) {
  double e = 0;
  for (size_t i = 0; i < a.size1(); i++)
    for (size_t j = 0; j < a.size2(); j++) e = std::max(e, std::abs(a(i, j)));
  return e;
}
// oracle rotate/transform/product on host matrices: one term, host pointers
static void oracle_term(Matrix &C, const nrg_gemm3_term &shape, const Matrix &A, const Matrix *O, const Matrix &B) {
  nrg_gemm3_term t = shape;
  t.A = A.data(); t.lda = (int64_t)A.size2();
  t.O = O ? O->data() : nullptr; t.ldo = O ? (int64_t)O->size2() : 0;
  t.B = B.data(); t.ldb = (int64_t)B.size2();
  t.C = C.data(); t.ldc = (int64_t)C.size2();
  NRG_ASSERT(oracle_gemm3(&t, 1) == 0);
}

constexpr double sq2 = 0.70710678118654752440;

TEST(Diag, dsyev_pauli) {   // the known answers of test/unit/diag_test/diag_test.cpp:74-107
  for (const char *form : {"dsyev", "mixed"}) {
    DiagParams DP(form, 1.0);
    {
      Matrix m(2, 2);   // sigma_z
      m(0, 0) = -1.0; m(1, 1) = 1.0;
      const auto res = diagonalise(m, DP, -1);
      EXPECT_EQ(res.getnrcomputed(), 2u);
      EXPECT_EQ(res.getdim(), 2u);
      EXPECT_DOUBLE_EQ(res.val[0], -1.0);
      EXPECT_DOUBLE_EQ(res.val[1], 1.0);
      EXPECT_DOUBLE_EQ(std::abs(res.vec(0, 0)), 1.0);
      EXPECT_NEAR(res.vec(0, 1), 0.0, 1e-16);
      EXPECT_NEAR(res.vec(1, 0), 0.0, 1e-16);
      EXPECT_DOUBLE_EQ(std::abs(res.vec(1, 1)), 1.0);
    }
    {
      Matrix m(2, 2);   // sigma_x
      m(0, 1) = 1.0;
      const auto res = diagonalise(m, DP, -1);
      EXPECT_DOUBLE_EQ(res.val[0], -1.0);
      EXPECT_DOUBLE_EQ(res.val[1], 1.0);
      EXPECT_DOUBLE_EQ(std::abs(res.vec(0, 0)), sq2);
      EXPECT_DOUBLE_EQ(std::abs(res.vec(0, 1)), sq2);
      EXPECT_DOUBLE_EQ(res.vec(0, 0) * res.vec(0, 1), -0.5);
      EXPECT_DOUBLE_EQ(res.vec(1, 0) * res.vec(1, 1), 0.5);
    }
  }
}

TEST(Diag, dsyevr_keeps_lowest_fraction) {
  Matrix m = random_upper(40);
  Matrix ref = m;
  const auto all = diagonalise(ref, DiagParams("dsyev", 1.0), -1);
  const auto part = diagonalise(m, DiagParams("dsyevr", 0.3), -1);
  EXPECT_EQ(part.getnrcomputed(), 12u);
  for (size_t i = 0; i < 12; i++) EXPECT_NEAR(part.val[i], all.val[i], 1e-13);
  EXPECT_EQ(m.size1(), 0u);   // consumed
}

TEST(Diag, rejects_lower_triangle_and_reports_failure) {
  Matrix m(3, 3);
  m(2, 0) = 1.0;
  EXPECT_THROW(diagonalise(m, DiagParams(), -1));
  Matrix bad = random_upper(5);
  bad(1, 3) = std::nan("");
  EXPECT_THROW(diagonalise(bad, DiagParams(), -1));
}

TEST(Diag, diagonalisations_match_oracle) {
  Device &dev = Device::default_device();
  std::vector<std::pair<Invar, Matrix>> tasks;
  std::vector<Matrix> copies;
  std::uniform_int_distribution<int> nd(1, 300);
  for (int b = 0; b < 40; b++) {
    const int n = b < 3 ? b + 1 : nd(rng);
    tasks.push_back({Invar{b, b % 3}, random_upper((size_t)n)});
    copies.push_back(tasks.back().second);
  }
  for (const char *form : {"dsyevd", "mixed"}) {
    auto t = tasks;
    DiagInfo diag = diagonalisations(t, DiagParams(form, 1.0), dev);
    EXPECT_EQ(diag.size(), tasks.size());
    // oracle: dsyev per block
    std::vector<int64_t> n, offH, offw;
    int64_t tot = 0, totw = 0;
    for (auto &c : copies) {
      n.push_back((int64_t)c.size1());
      offH.push_back(tot);
      offw.push_back(totw);
      tot += n.back() * n.back();
      totw += n.back();
    }
    std::vector<double> H((size_t)tot), w((size_t)totw);
    for (size_t b = 0; b < copies.size(); b++)
      std::copy(copies[b].data(), copies[b].data() + n[b] * n[b], H.begin() + offH[b]);
    std::vector<int> info(copies.size());
    EXPECT_EQ(oracle_syev_batched((int)copies.size(), n.data(), offH.data(), H.data(), offw.data(), w.data(), 4, info.data()), 0);
    double worst = 0;
    for (size_t b = 0; b < copies.size(); b++) {
      const Eigen &e = diag.at(tasks[b].first);
      EXPECT_EQ(e.getnrcomputed(), (size_t)n[b]);
      double hn = 0;
      for (int64_t i = 0; i < n[b]; i++) hn = std::max(hn, std::abs(w[offw[b] + i]));
      for (int64_t i = 0; i < n[b]; i++)
        worst = std::max(worst, std::abs(e.values.rel(i) - w[offw[b] + i]) / std::max(hn, 1e-300));
    }
    EXPECT_TRUE(worst <= 1e-12);
    std::printf("    %s: max |lambda - lambda_oracle| / ||H|| = %.2e\n", form, worst);
  }
}

TEST(Hamiltonian, assembled_blocks_have_the_closed_form_spectra) {
  Device &dev = Device::default_device();
  const double step_scale = 0.37, sf = std::sqrt(2.0);
  StepHamiltonians hs;
  // block 0: two sub-blocks of 150 states, energies a and b, coupled by t times the identity
  // (diag_offdiag_function): eigenvalues (a+b)/2 -+ sqrt(((a-b)/2)^2 + t^2), 150 times each
  const size_t m = 150;
  const double c1 = 0.2, c2 = 0.9, t = 0.35;
  const size_t b0 = hs.add_block(Invar{0}, SubspaceDimensions({m, m}, {Invar{0}, Invar{1}}));
  hs.energies(b0, 0, std::vector<double>(m, c1), sf);
  hs.energies(b0, 1, std::vector<double>(m, c2), sf);
  hs.diag_offdiag(b0, 1, 2, t * step_scale, step_scale);
  hs.diag_offdiag(b0, 2, 1, 5.0, step_scale);   // i > j: ignored, as in the reference
  // block 1: sub-blocks of 90 and 130 states coupled by a term given for (2, 1): a 130 x 90 matrix
  // F added transposed into the upper block, times f / step.scale().
  // Eigenvalues: -+ the singular values of (f / scale) F, and 40 zeros
  const size_t r1 = 90, r2 = 130;
  const double f = -0.8;
  Matrix F = random_matrix(r2, r1);
  DMatrix Fd(F);
  const size_t b1 = hs.add_block(Invar{1}, SubspaceDimensions({r1, r2}, {Invar{0}, Invar{1}}));
  hs.offdiag(b1, 2, 1, Fd, f, step_scale);
  EXPECT_EQ(hs.order(b1), r1 + r2);
  const DiagParams DP;
  DiagInfo diag = diagonalisations(hs, DP, dev);
  {
    const auto &v = diag.at(Invar{0}).values;
    const double a = sf * c1, b = sf * c2, r = std::sqrt((a - b) / 2 * (a - b) / 2 + t * t);
    double err = 0;
    for (size_t k = 0; k < 2 * m; k++) err = std::max(err, std::abs(v.rel(k) - ((a + b) / 2 + (k < m ? -r : r))));
    EXPECT_TRUE(err <= 1e-14);
  }
  {
    // singular values of F from the oracle's dsyev of F^T F (90 x 90)
    Matrix G(r1, r1);
    for (size_t i = 0; i < r1; i++)
      for (size_t j = 0; j < r1; j++) {
        double s = 0;
        for (size_t k = 0; k < r2; k++) s += F(k, i) * F(k, j);
        G(i, j) = j >= i ? s : 0.0;
      }
    std::vector<double> w(r1);
    std::vector<int64_t> n{(int64_t)r1}, off{0};
    std::vector<int> info(1);
    EXPECT_EQ(oracle_syev_batched(1, n.data(), off.data(), G.data(), off.data(), w.data(), 1, info.data()), 0);
    std::vector<double> want;
    const double g = std::abs(f / step_scale);
    for (double x : w) {
      const double sv = g * std::sqrt(std::max(x, 0.0));
      want.push_back(sv);
      want.push_back(-sv);
    }
    for (size_t k = 0; k < r2 - r1; k++) want.push_back(0.0);
    std::sort(want.begin(), want.end());
    const auto &v = diag.at(Invar{1}).values;
    double err = 0, big = 0;
    for (size_t k = 0; k < want.size(); k++) {
      err = std::max(err, std::abs(v.rel(k) - want[k]));
      big = std::max(big, std::abs(want[k]));
    }
    EXPECT_TRUE(err <= 1e-13 * big);
    std::printf("    identity-coupled block and transposed-coupling block: max deviation %.2e of %.2e\n", err, big);
  }
}

TEST(Numerics, product_transform_rotate_match_oracle) {
  const size_t m = 70, n = 90, k = 130, l = 65;
  {   // product M += f A B^T
    Matrix A = random_matrix(m, k), B = random_matrix(n, k), M = random_matrix(m, n), R = M;
    product(M, 0.75, A, B);
    nrg_gemm3_term s{}; s.m = m; s.n = n; s.k = k; s.l = k; s.f = 0.75; s.transA = 0; s.transB = 0;
    oracle_term(R, s, A, nullptr, B);
    EXPECT_TRUE(max_abs_diff(M, R) <= 1e-13 * max_abs(R));
  }
  {   // transform M += f A O B^T
    Matrix A = random_matrix(m, k), O = random_matrix(k, l), B = random_matrix(n, l), M = random_matrix(m, n), R = M;
    transform(M, -1.25, A, O, B);
    nrg_gemm3_term s{}; s.m = m; s.n = n; s.k = k; s.l = l; s.f = -1.25; s.transA = 0; s.transB = 0;
    oracle_term(R, s, A, &O, B);
    EXPECT_TRUE(max_abs_diff(M, R) <= 1e-13 * max_abs(R));
  }
  {   // rotate M += f U^T O U
    Matrix U = random_matrix(k, m), O = random_matrix(k, k), M = random_matrix(m, m), R = M;
    rotate(M, 0.5, U, O);
    nrg_gemm3_term s{}; s.m = m; s.n = m; s.k = k; s.l = k; s.f = 0.5; s.transA = 1; s.transB = 1;
    oracle_term(R, s, U, &O, U);
    EXPECT_TRUE(max_abs_diff(M, R) <= 1e-13 * max_abs(R));
  }
  {   // empty operands contribute nothing (finite_size)
    Matrix M = random_matrix(3, 3), R = M;
    product(M, 1.0, Matrix(3, 0), Matrix(3, 0));
    EXPECT_EQ(max_abs_diff(M, R), 0.0);
  }
}

// ---- a toy NRG chain: subspaces {Q}; step N's {Q} has ancestors {Q} (empty site) and {Q-1}
// (occupied site) of step N-1; multiplicity 1 + Q % 2.
class ToySym : public Symmetry {
 public:
  std::vector<Invar> new_subspaces(const Invar &I) const override { return {Invar{I.qn[0]}, Invar{I.qn[0] + 1}}; }
  double mult(const Invar &I) const override { return 1.0 + (I.qn[0] % 2); }
  size_t nr_combs() const override { return 2; }
};

struct ToyChain {
  size_t Nmax = 6, keep = 24;
  double T = 0.05;
  ToySym sym;
  std::vector<DiagInfo> diags;
  std::vector<SubspaceStructure> sstruct;
  Store store{0, 6};
  Stats stats{6};
  ToyChain() {
    Device &dev = Device::default_device();
    diags.resize(Nmax);
    sstruct.resize(Nmax);
    double tshift = 0, gs = 1e300;
    for (size_t N = 0; N < Nmax; N++) {
      std::vector<std::pair<Invar, Matrix>> tasks;
      const double scale = std::pow(2.0, -(double)N / 2);
      for (int Q = 0; Q <= (int)N + 1; Q++) {
        const Invar I{Q};
        size_t d0, d1;
        if (N == 0) {
          d0 = 3;
          d1 = 0;
        } else {
          d0 = diags[N - 1].count(Invar{Q}) ? diags[N - 1].at(Invar{Q}).getnrkept() : 0;
          d1 = diags[N - 1].count(Invar{Q - 1}) ? diags[N - 1].at(Invar{Q - 1}).getnrkept() : 0;
        }
        if (d0 + d1 == 0) continue;
        sstruct[N][I] = SubspaceDimensions({d0, d1}, {Invar{Q}, Invar{Q - 1}});
        Matrix H = random_upper(d0 + d1, 0.3);
        for (size_t i = 0; i < d0 + d1; i++) H(i, i) += 0.5 * Q + 0.05 * (double)i;
        tasks.push_back({I, std::move(H)});
      }
      diags[N] = diagonalisations(tasks, DiagParams("dsyev", 1.0), dev, N == Nmax - 1);
      const double Egs = diags[N].Egs_subtraction();
      for (auto &[I, eig] : diags[N]) {
        eig.values.set_scale(scale);
        eig.values.set_T_shift(tshift);
        eig.values.set_corr(eig.values.all_rel_zero());
        eig.truncate_prepare(std::min(keep, eig.getnrcomputed()));
        for (double x : eig.values.all_abs_T()) gs = std::min(gs, x);
      }
      tshift += Egs * scale;
      for (auto &[I, eig] : diags[N]) store[N][I] = Sub{eig, sstruct[N].at(I), N == Nmax - 1};
    }
    for (size_t N = 0; N < Nmax; N++) {
      diags[N].subtract_GS_energy(gs);
      store.shift_abs_energies(gs);
    }
    stats.GS_energy = gs;
    calc_ZnD(store, stats, sym, T);
  }
};

static ToyChain &chain() {
  static ToyChain c;
  return c;
}

TEST(Measurements, calc_ZnD_sum_rule) {
  auto &c = chain();
  double s = 0;
  for (double w : c.stats.wn) s += w;
  EXPECT_NEAR(s, 1.0, 1e-12);
  for (size_t N = 0; N < c.Nmax; N++) EXPECT_TRUE(c.stats.wn[N] >= 0.0);
  fdm_thermodynamics(c.store, c.stats, c.sym, c.T);
  EXPECT_TRUE(std::isfinite(c.stats.E_fdm) && c.stats.C_fdm >= 0.0);
}

TEST(DMNRG, fulldensitymatrix_sweep_traces_and_oracle_parity) {
  auto &c = chain();
  Device &dev = Device::default_device();
  const auto mult = c.sym.multfnc();
  DensMatElements rho = init_rho_FDM(c.Nmax - 1, c.store, c.stats, mult, c.T);
  EXPECT_NEAR(rho.trace(mult), c.stats.wn[c.Nmax - 1], 1e-12);
  // host replica of every step with the oracle's dgemm for the rotations
  std::map<Invar, Matrix> host;
  for (auto &[I, m] : rho) host[I] = m.host();
  std::vector<const DiagInfo *> dptr;
  for (auto &d : c.diags) dptr.push_back(&d);
  double worst = 0;
  size_t steps = 0;
  calc_fulldensitymatrix(rho, c.store, c.stats, c.sym, dptr, 0, c.Nmax, c.T, dev,
                         [&](size_t Nm1, const DensMatElements &got) {
    const size_t N = Nm1 + 1;
    const bool last = N == c.Nmax - 1;
    std::map<Invar, Matrix> dd;
    if (!last)
      for (auto &[I, m] : init_rho_FDM(N, c.store, c.stats, mult, c.T)) dd[I] = m.host();
    std::map<Invar, Matrix> prev;
    for (const auto &[I, ds] : c.store[N - 1]) {
      Matrix acc(ds.kept(), ds.kept());
      const auto subs = c.sym.new_subspaces(I);
      for (size_t i = 0; i < 2; i++) {
        const auto y = c.diags[N].find(subs[i]);
        if (y == c.diags[N].end() || !ds.kept()) continue;
        const auto &sd = c.store[N].at(subs[i]).rmax;
        const Matrix V = y->second.vectors.get();
        for (const auto *src : {&host, last ? nullptr : &dd}) {
          if (!src || !src->count(subs[i])) continue;
          const Matrix &r = src->at(subs[i]);
          const size_t nom = r.size1();
          if (!nom || !sd.rmax(i)) continue;
          Matrix U(nom, sd.rmax(i));
          for (size_t a = 0; a < nom; a++)
            for (size_t b = 0; b < sd.rmax(i); b++) U(a, b) = V(a, sd.offset(i) + b);
          nrg_gemm3_term s{}; s.m = U.size2(); s.n = U.size2(); s.k = nom; s.l = nom;
          s.f = c.sym.mult(subs[i]) / c.sym.mult(I); s.transA = 1; s.transB = 1;
          oracle_term(acc, s, U, &r, U);
        }
      }
      const Matrix g = got.at(I).host();
      worst = std::max(worst, max_abs_diff(g, acc) / std::max(max_abs(acc), 1e-300));
      // symmetric and positive semidefinite diagonal
      for (size_t a = 0; a < g.size1(); a++) {
        EXPECT_TRUE(g(a, a) >= -1e-15);
        for (size_t b = 0; b < a; b++) EXPECT_NEAR(g(a, b), g(b, a), 1e-14 * std::max(max_abs(g), 1e-300));
      }
      prev[I] = acc;
    }
    host.swap(prev);
    steps++;
  });
  EXPECT_EQ(steps, c.Nmax - 1);
  EXPECT_TRUE(worst <= 1e-12);
  std::printf("    max relative deviation from the oracle over the sweep: %.2e\n", worst);
  // rho_FDM(Ninit) carries the whole weight
  EXPECT_NEAR(rho.trace(mult), 1.0, 1e-12);
}

TEST(DMNRG, densitymatrix_sweep_conserves_trace) {
  auto &c = chain();
  Device &dev = Device::default_device();
  const auto mult = c.sym.multfnc();
  const size_t last = c.Nmax - 1;
  // rho_N over the kept states of the last step (init_rho runs after truncation), c++/dmnrg.hpp:42-50
  DiagInfo dl = c.diags[last];
  dl.truncate_perform();
  DensMatElements rho = init_rho(1.0 / c.T * std::pow(2.0, -(double)last / 2), dl, mult);
  std::vector<const DiagInfo *> dptr;
  for (auto &d : c.diags) dptr.push_back(&d);
  calc_densitymatrix(rho, c.store, c.sym, dptr, 0, c.Nmax, dev);   // checks Tr = 1 at every step
  EXPECT_NEAR(rho.trace(mult), 1.0, 1e-12);
  for (const auto &[I, m] : rho) EXPECT_EQ(m.size1(), c.store[0].at(I).kept());
}

TEST(Recalc, recalc_general_matches_oracle) {
  auto &c = chain();
  Device &dev = Device::default_device();
  const size_t N = 3;
  DiagInfo &diag = c.diags[N];
  split_in_blocks(diag, c.sstruct[N]);
  // <Q+1||d^dag||Q> at N-1, random; table rows (1,1) and (2,2) with a fermionic sign
  MatrixElements cold;
  std::map<Twoinvar, Matrix> cold_h;
  for (const auto &[I, e] : c.diags[N - 1]) {
    const Invar Ip{I.qn[0] + 1};
    if (!c.diags[N - 1].count(Ip)) continue;
    Matrix m = random_matrix(c.diags[N - 1].at(Ip).getnrkept(), e.getnrkept());
    cold_h[{Ip, I}] = m;
    cold[{Ip, I}] = DMatrix(m);
  }
  Gemm3Batch batch(dev);
  std::map<Twoinvar, DMatrix> out;
  for (const auto &[I, e] : diag) {
    const Invar I1{I.qn[0] + 1};
    if (!diag.count(I1)) continue;
    const std::vector<Recalc> table = {{1, 1, Invar{I1.qn[0]}, Invar{I.qn[0]}, 1.0},
                                       {2, 2, Invar{I1.qn[0] - 1}, Invar{I.qn[0] - 1}, -1.0}};
    out[{I1, I}] = recalc_general(diag, cold, I1, I, table, batch);
  }
  batch.flush();
  EXPECT_TRUE(!out.empty());
  double worst = 0;
  for (auto &[II, m] : out) {
    const auto &[I1, Ip] = II;
    const Eigen &e1 = diag.at(I1), &ep = diag.at(Ip);
    Matrix ref(e1.getnrstored(), ep.getnrstored());
    const Matrix V1 = e1.vectors.get(), Vp = ep.vectors.get();
    const auto &s1 = c.sstruct[N].at(I1), &sp = c.sstruct[N].at(Ip);
    for (size_t row = 0; row < 2; row++) {
      const Twoinvar key{s1.ancestor(row), sp.ancestor(row)};
      if (!cold_h.count(key) || !s1.rmax(row) || !sp.rmax(row)) continue;
      Matrix U1(V1.size1(), s1.rmax(row)), Up(Vp.size1(), sp.rmax(row));
      for (size_t a = 0; a < V1.size1(); a++)
        for (size_t b = 0; b < s1.rmax(row); b++) U1(a, b) = V1(a, s1.offset(row) + b);
      for (size_t a = 0; a < Vp.size1(); a++)
        for (size_t b = 0; b < sp.rmax(row); b++) Up(a, b) = Vp(a, sp.offset(row) + b);
      nrg_gemm3_term s{}; s.m = U1.size1(); s.n = Up.size1(); s.k = U1.size2(); s.l = Up.size2();
      s.f = row ? -1.0 : 1.0; s.transA = 0; s.transB = 0;
      oracle_term(ref, s, U1, &cold_h.at(key), Up);
    }
    worst = std::max(worst, max_abs_diff(m.host(), ref) / std::max(max_abs(ref), 1e-300));
  }
  EXPECT_TRUE(worst <= 1e-13);
  std::printf("    recalc_general: max relative deviation %.2e over %zu blocks\n", worst, out.size());
}

TEST(Spectrum, fdm_step_matches_oracle) {
  auto &c = chain();
  Device &dev = Device::default_device();
  const auto mult = c.sym.multfnc();
  const size_t N = c.Nmax - 1;   // the last step: every state discarded, rho_FDM = init_rho_FDM
  for (size_t step : {N, N - 1}) {   // step N-1 is the first with discarded states below the last
    DiagInfo &diag = c.diags[step];
    DensMatElements rho;
    if (step == N) {
      rho = init_rho_FDM(N, c.store, c.stats, mult, c.T);
    } else {   // rho_FDM(step) restricted to kept states, from the sweep
      DensMatElements r = init_rho_FDM(N, c.store, c.stats, mult, c.T);
      std::vector<const DiagInfo *> dptr;
      for (auto &d : c.diags) dptr.push_back(&d);
      calc_fulldensitymatrix(r, c.store, c.stats, c.sym, dptr, step, c.Nmax, c.T, dev);
      rho = std::move(r);
    }
    MatrixElements op;
    std::map<Twoinvar, Matrix> op_h;
    for (const auto &[I, e] : diag) {
      const Invar Ij{I.qn[0] + 1};
      if (!diag.count(Ij)) continue;
      Matrix m = random_matrix(diag.at(Ij).getnrall(), e.getnrall());
      op_h[{Ij, I}] = m;
      op[{Ij, I}] = DMatrix(m);
    }
    const BinGrid grid = BinGrid::standard(-1, -1, 1.0, std::pow(2.0, -(double)N / 2), 100, 1e-16);
    FDMSpectrum spec(dev, grid, -1, c.T);
    const auto ff = [](const Invar &Ii, const Invar &) { return 1.0 + 0.25 * Ii.qn[0]; };
    spec.calc(diag, op, op, rho, c.stats.wnfactor[step], ff);
    // oracle on host copies of the same pairs
    std::vector<nrg_fdm_pair> pairs;
    std::map<Invar, std::vector<double>> E;
    std::map<Invar, Matrix> rho_h;
    for (const auto &[I, e] : diag) E[I] = e.values.all_abs_G();
    for (auto &[I, m] : rho) rho_h[I] = m.host();
    for (const auto &[II, m] : op_h) {
      const auto &[Ij, Ii] = II;
      nrg_fdm_pair p{};
      p.dim_i = (int64_t)m.size2(); p.dim_j = (int64_t)m.size1();
      p.K_i = (int64_t)diag.at(Ii).Krange().second; p.K_j = (int64_t)diag.at(Ij).Krange().second;
      p.E_i = E[Ii].data(); p.E_j = E[Ij].data();
      p.op1 = m.data(); p.op2 = m.data(); p.ld_op = (int64_t)m.size2();
      if (rho_h.count(Ii)) { p.rho_i = rho_h[Ii].data(); p.n_rho_i = (int64_t)rho_h[Ii].size1(); }
      if (rho_h.count(Ij)) { p.rho_j = rho_h[Ij].data(); p.n_rho_j = (int64_t)rho_h[Ij].size1(); }
      p.factor = ff(Ii, Ij);
      pairs.push_back(p);
    }
    const nrg_bin_grid g = grid.c();
    std::vector<double> gp(grid.nbins), gn(grid.nbins), lp(grid.nbins), ln(grid.nbins);
    EXPECT_EQ(oracle_fdm_bin(pairs.data(), (int)pairs.size(), &g, c.T, c.stats.wnfactor[step], -1, gp.data(), gn.data(),
                             lp.data(), ln.data()),
              0);
    double scale = 0, err = 0;
    for (size_t i = 0; i < grid.nbins; i++) {
      scale += std::abs(gp[i]) + std::abs(gn[i]) + std::abs(lp[i]) + std::abs(ln[i]);
      err = std::max({err, std::abs(spec.gt().pos[i] - gp[i]), std::abs(spec.gt().neg[i] - gn[i]),
                      std::abs(spec.ls().pos[i] - lp[i]), std::abs(spec.ls().neg[i] - ln[i])});
    }
    EXPECT_TRUE(scale > 0);
    EXPECT_TRUE(err <= 1e-13 * scale);
    std::printf("    step %zu: %zu pairs, max bin deviation %.2e of total |weight| %.3e\n", step, pairs.size(), err, scale);
  }
}

TEST(Spectrum, dmnrg_step_matches_oracle) {
  auto &c = chain();
  Device &dev = Device::default_device();
  const auto mult = c.sym.multfnc();
  const size_t last = c.Nmax - 1;
  DiagInfo dl = c.diags[last];
  dl.truncate_perform();
  DensMatElements rho = init_rho(1.0 / c.T * std::pow(2.0, -(double)last / 2), dl, mult);
  std::vector<const DiagInfo *> dptr;
  for (auto &d : c.diags) dptr.push_back(&d);
  const size_t step = last - 2;
  calc_densitymatrix(rho, c.store, c.sym, dptr, step, c.Nmax, dev);   // DM-NRG rho of `step`
  DiagInfo &diag = c.diags[step];
  MatrixElements op;
  std::map<Twoinvar, Matrix> op_h;
  for (const auto &[I, e] : diag) {
    const Invar Ij{I.qn[0] + 1};
    if (!diag.count(Ij)) continue;
    Matrix m = random_matrix(diag.at(Ij).getnrall(), e.getnrall());
    op_h[{Ij, I}] = m;
    op[{Ij, I}] = DMatrix(m);
  }
  const BinGrid grid = BinGrid::standard(-1, -1, 1.0, std::pow(2.0, -(double)last / 2), 100, 1e-16);
  NN2Window win;
  win.E0 = 1.5;
  win.Efactor = 2.0;
  win.NN2avg = true;
  const double scale = std::pow(2.0, -(double)step / 2);
  const auto ff = [](const Invar &Ii, const Invar &) { return 1.0 + 0.25 * Ii.qn[0]; };
  DMNRGSpectrum spec(dev, grid, -1, win);
  spec.calc(diag, op, op, rho, scale, true, false, false, ff);
  // oracle on host copies of the same pairs, then the N/N+2 merge of both halves
  std::vector<nrg_fdm_pair> pairs;
  std::map<Invar, std::vector<double>> E;
  std::map<Invar, Matrix> rho_h;
  for (const auto &[I, e] : diag) E[I] = e.values.all_abs_zero();
  for (auto &[I, m] : rho) rho_h[I] = m.host();
  for (const auto &[II, m] : op_h) {
    const auto &[Ij, Ii] = II;
    nrg_fdm_pair p{};
    p.dim_i = (int64_t)m.size2(); p.dim_j = (int64_t)m.size1();
    p.K_i = (int64_t)diag.at(Ii).getnrkept(); p.K_j = (int64_t)diag.at(Ij).getnrkept();   // kept()
    p.E_i = E[Ii].data(); p.E_j = E[Ij].data();
    p.op1 = m.data(); p.op2 = m.data(); p.ld_op = (int64_t)m.size2();
    p.rho_i = rho_h[Ii].data(); p.n_rho_i = (int64_t)rho_h[Ii].size1();
    p.rho_j = rho_h[Ij].data(); p.n_rho_j = (int64_t)rho_h[Ij].size1();
    p.factor = ff(Ii, Ij);
    pairs.push_back(p);
  }
  const nrg_bin_grid g = grid.c();
  std::vector<double> bp(grid.nbins), bn(grid.nbins), fp(grid.nbins), fn(grid.nbins);
  EXPECT_EQ(oracle_dmnrg_bin(pairs.data(), (int)pairs.size(), &g, win.emin(scale), win.emax(scale), -1, bp.data(), bn.data()),
            0);
  oracle_merge_nn2(&g, fp.data(), bp.data(), win.emin(scale), win.ex(scale), win.emax(scale), 0, 0, 0, 0.0, 1);
  oracle_merge_nn2(&g, fn.data(), bn.data(), win.emin(scale), win.ex(scale), win.emax(scale), 0, 0, 0, 0.0, 1);
  double scale_b = 0, err_b = 0, scale_f = 0, err_f = 0;
  for (size_t i = 0; i < grid.nbins; i++) {
    scale_b += std::abs(bp[i]) + std::abs(bn[i]);
    scale_f += std::abs(fp[i]) + std::abs(fn[i]);
    err_b = std::max({err_b, std::abs(spec.last_step().pos[i] - bp[i]), std::abs(spec.last_step().neg[i] - bn[i])});
    err_f = std::max({err_f, std::abs(spec.spectrum().pos[i] - fp[i]), std::abs(spec.spectrum().neg[i] - fn[i])});
  }
  EXPECT_TRUE(scale_b > 0 && scale_f > 0);
  EXPECT_TRUE(err_b <= 1e-13 * scale_b);
  EXPECT_TRUE(err_f <= 1e-13 * scale_f);
  std::printf("    step %zu: %zu pairs, max bin deviation %.2e of %.3e (merged %.2e of %.3e)\n", step, pairs.size(), err_b,
              scale_b, err_f, scale_f);
}

TEST(Spectrum, cfs_steps_match_oracle) {
  // a step below the last one (DM-NRG rho from the sweep) and the last step, accumulated as mergeCFS does
  auto &c = chain();
  Device &dev = Device::default_device();
  const auto mult = c.sym.multfnc();
  const size_t last = c.Nmax - 1;
  DiagInfo dl = c.diags[last];
  dl.truncate_perform();
  DensMatElements rho = init_rho(1.0 / c.T * std::pow(2.0, -(double)last / 2), dl, mult);
  std::vector<const DiagInfo *> dptr;
  for (auto &d : c.diags) dptr.push_back(&d);
  const size_t step = last - 1;   // a step with discarded states (24 kept per subspace)
  calc_densitymatrix(rho, c.store, c.sym, dptr, step, c.Nmax, dev);
  const BinGrid grid = BinGrid::standard(-1, -1, 1.0, std::pow(2.0, -(double)last / 2), 100, 1e-16);
  const auto ff = [](const Invar &Ii, const Invar &) { return 1.0 + 0.25 * Ii.qn[0]; };
  const double Z = 2.5;
  CFSSpectrum spec(dev, grid, -1);
  const nrg_bin_grid g = grid.c();
  std::vector<double> tot(2 * grid.nbins, 0.0);
  double err = 0, scale = 0;
  for (const bool is_last : {false, true}) {
    DiagInfo &diag = c.diags[is_last ? last : step];
    MatrixElements op;
    std::map<Twoinvar, Matrix> op_h;
    for (const auto &[I, e] : diag) {
      const Invar Ij{I.qn[0] + 1};
      if (!diag.count(Ij)) continue;
      Matrix m = random_matrix(diag.at(Ij).getnrall(), e.getnrall());
      op_h[{Ij, I}] = m;
      op[{Ij, I}] = DMatrix(m);
    }
    spec.calc(diag, op, op, rho, is_last, c.T, Z, ff);
    std::vector<nrg_fdm_pair> pairs;
    std::map<Invar, std::vector<double>> E;
    std::map<Invar, Matrix> rho_h;
    for (const auto &[I, e] : diag) E[I] = e.values.all_abs_zero();
    for (auto &[I, m] : rho) rho_h[I] = m.host();
    for (const auto &[II, m] : op_h) {
      const auto &[Ij, Ii] = II;
      nrg_fdm_pair p{};
      p.dim_i = (int64_t)m.size2(); p.dim_j = (int64_t)m.size1();
      p.K_i = (int64_t)diag.at(Ii).getnrkept(); p.K_j = (int64_t)diag.at(Ij).getnrkept();   // kept()
      p.E_i = E[Ii].data(); p.E_j = E[Ij].data();
      p.op1 = m.data(); p.op2 = m.data(); p.ld_op = (int64_t)m.size2();
      if (!is_last) {
        p.rho_i = rho_h[Ii].data(); p.n_rho_i = (int64_t)rho_h[Ii].size1();
        p.rho_j = rho_h[Ij].data(); p.n_rho_j = (int64_t)rho_h[Ij].size1();
      }
      p.factor = ff(Ii, Ij);
      pairs.push_back(p);
    }
    std::vector<double> b(4 * grid.nbins, 0.0);
    const size_t nb = grid.nbins;
    EXPECT_EQ(oracle_cfs_bin(pairs.data(), (int)pairs.size(), &g, c.T, Z, is_last ? 1 : 0, -1, b.data(), b.data() + nb,
                             b.data() + 2 * nb, b.data() + 3 * nb),
              0);
    for (size_t i = 0; i < nb; i++) {
      scale += std::abs(b[i]) + std::abs(b[nb + i]) + std::abs(b[2 * nb + i]) + std::abs(b[3 * nb + i]);
      err = std::max({err, std::abs(spec.last_greater().pos[i] - b[i]), std::abs(spec.last_greater().neg[i] - b[nb + i]),
                      std::abs(spec.last_lesser().pos[i] - b[2 * nb + i]),
                      std::abs(spec.last_lesser().neg[i] - b[3 * nb + i])});
      tot[i] += b[i] + b[2 * nb + i];
      tot[nb + i] += b[nb + i] + b[3 * nb + i];
    }
    double w = 0;
    for (double x : b) w += std::abs(x);
    EXPECT_TRUE(w > 0);
    std::printf("    %s step: %zu pairs, total |weight| %.3e\n", is_last ? "last" : "inner", pairs.size(), w);
  }
  double err_t = 0;
  for (size_t i = 0; i < grid.nbins; i++)
    err_t = std::max({err_t, std::abs(spec.total().pos[i] - tot[i]), std::abs(spec.total().neg[i] - tot[grid.nbins + i])});
  EXPECT_TRUE(scale > 0);
  EXPECT_TRUE(err <= 1e-13 * scale);
  EXPECT_TRUE(err_t <= 2e-13 * scale);
  std::printf("    max bin deviation %.2e (total %.2e) of %.3e\n", err, err_t, scale);
}

TEST(Spectrum, ft_steps_match_oracle) {
  // an inner step and the last step (all 24 kept states of each subspace), bins and N/N+2 merge
  auto &c = chain();
  Device &dev = Device::default_device();
  const size_t last = c.Nmax - 1;
  const BinGrid grid = BinGrid::standard(-1, -1, 1.0, std::pow(2.0, -(double)last / 2), 100, 1e-16);
  NN2Window win;
  win.E0 = 1.5;
  win.Efactor = 2.0;
  const auto ff = [](const Invar &Ii, const Invar &) { return 1.0 - 0.125 * Ii.qn[0]; };
  const double Z = 3.5;
  FTSpectrum spec(dev, grid, -1, c.T, win);
  const nrg_bin_grid g = grid.c();
  const size_t nb = grid.nbins;
  std::vector<double> fp(nb, 0.0), fn(nb, 0.0);
  double err = 0, scale = 0;
  for (const size_t step : {last - 1, last}) {
    DiagInfo &diag = c.diags[step];
    MatrixElements op;
    std::map<Twoinvar, Matrix> op_h;
    for (const auto &[I, e] : diag) {
      const Invar Ij{I.qn[0] + 1};
      if (!diag.count(Ij)) continue;
      Matrix m = random_matrix(diag.at(Ij).getnrall(), e.getnrall());
      op_h[{Ij, I}] = m;
      op[{Ij, I}] = DMatrix(m);
    }
    const double sc = std::pow(2.0, -(double)step / 2);
    const bool is_last = step == last;
    spec.calc(diag, op, op, Z, sc, true, false, is_last, ff);
    std::vector<nrg_fdm_pair> pairs;
    std::map<Invar, std::vector<double>> E;
    for (const auto &[I, e] : diag) E[I] = e.values.all_abs_zero();
    for (const auto &[II, m] : op_h) {
      const auto &[Ij, Ii] = II;
      nrg_fdm_pair p{};
      p.dim_i = (int64_t)m.size2(); p.dim_j = (int64_t)m.size1();
      p.K_i = (int64_t)diag.at(Ii).getnrkept(); p.K_j = (int64_t)diag.at(Ij).getnrkept();
      p.E_i = E[Ii].data(); p.E_j = E[Ij].data();
      p.op1 = m.data(); p.op2 = m.data(); p.ld_op = (int64_t)m.size2();
      p.factor = ff(Ii, Ij);
      pairs.push_back(p);
    }
    std::vector<double> bp(nb, 0.0), bn(nb, 0.0);
    EXPECT_EQ(oracle_ft_bin(pairs.data(), (int)pairs.size(), &g, c.T, Z, -1, bp.data(), bn.data()), 0);
    oracle_merge_nn2(&g, fp.data(), bp.data(), win.emin(sc), win.ex(sc), win.emax(sc), 0, is_last, 0, 0.0, 0);
    oracle_merge_nn2(&g, fn.data(), bn.data(), win.emin(sc), win.ex(sc), win.emax(sc), 0, is_last, 0, 0.0, 0);
    double w = 0;
    for (size_t i = 0; i < nb; i++) {
      w += std::abs(bp[i]) + std::abs(bn[i]);
      err = std::max({err, std::abs(spec.last_step().pos[i] - bp[i]), std::abs(spec.last_step().neg[i] - bn[i])});
    }
    EXPECT_TRUE(w > 0);
    scale += w;
    std::printf("    step %zu: %zu pairs, total |weight| %.3e\n", step, pairs.size(), w);
  }
  double err_m = 0;
  for (size_t i = 0; i < nb; i++)
    err_m = std::max({err_m, std::abs(spec.spectrum().pos[i] - fp[i]), std::abs(spec.spectrum().neg[i] - fn[i])});
  EXPECT_TRUE(err <= 1e-13 * scale);
  EXPECT_TRUE(err_m <= 1e-13 * scale);
  std::printf("    max bin deviation %.2e (merged %.2e) of %.3e\n", err, err_m, scale);
}

TEST(Spectrum, matsubara_steps_match_oracle) {
  // FDMmats at the step below the last (rho_FDM from the sweep), FTmats at the last step, both
  // accumulated into one MatsubaraGF, as GFMatsubara::merge does across steps
  auto &c = chain();
  Device &dev = Device::default_device();
  const auto mult = c.sym.multfnc();
  const size_t N = c.Nmax - 1, step = N - 1;
  DensMatElements rho = init_rho_FDM(N, c.store, c.stats, mult, c.T);
  std::vector<const DiagInfo *> dptr;
  for (auto &d : c.diags) dptr.push_back(&d);
  calc_fulldensitymatrix(rho, c.store, c.stats, c.sym, dptr, step, c.Nmax, c.T, dev);
  const size_t mats = 64;
  const double Z = 2.0;
  const auto ff = [](const Invar &Ii, const Invar &) { return 1.0 + 0.25 * Ii.qn[0]; };
  MatsubaraGF gf(dev, mats, -1, c.T);
  std::vector<double> want(2 * mats, 0.0);
  double err = 0, scale = 0;
  for (const bool ft : {false, true}) {
    DiagInfo &diag = c.diags[ft ? N : step];
    MatrixElements op;
    std::map<Twoinvar, Matrix> op_h;
    for (const auto &[I, e] : diag) {
      const Invar Ij{I.qn[0] + 1};
      if (!diag.count(Ij)) continue;
      Matrix m = random_matrix(diag.at(Ij).getnrall(), e.getnrall());
      op_h[{Ij, I}] = m;
      op[{Ij, I}] = DMatrix(m);
    }
    if (ft) gf.calc_ft(diag, op, op, Z, ff);
    else gf.calc_fdm(diag, op, op, rho, c.stats.wnfactor[step], ff);
    std::vector<nrg_fdm_pair> pairs;
    std::map<Invar, std::vector<double>> E;
    std::map<Invar, Matrix> rho_h;
    for (const auto &[I, e] : diag) E[I] = ft ? e.values.all_abs_zero() : e.values.all_abs_G();
    for (auto &[I, m] : rho) rho_h[I] = m.host();
    for (const auto &[II, m] : op_h) {
      const auto &[Ij, Ii] = II;
      nrg_fdm_pair p{};
      p.dim_i = (int64_t)m.size2(); p.dim_j = (int64_t)m.size1();
      p.K_i = (int64_t)(ft ? diag.at(Ii).getnrkept() : diag.at(Ii).Krange().second);
      p.K_j = (int64_t)(ft ? diag.at(Ij).getnrkept() : diag.at(Ij).Krange().second);
      p.E_i = E[Ii].data(); p.E_j = E[Ij].data();
      p.op1 = m.data(); p.op2 = m.data(); p.ld_op = (int64_t)m.size2();
      if (!ft && rho_h.count(Ii)) { p.rho_i = rho_h[Ii].data(); p.n_rho_i = (int64_t)rho_h[Ii].size1(); }
      if (!ft && rho_h.count(Ij)) { p.rho_j = rho_h[Ij].data(); p.n_rho_j = (int64_t)rho_h[Ij].size1(); }
      p.factor = ff(Ii, Ij);
      pairs.push_back(p);
    }
    std::vector<double> G(2 * mats, 0.0);
    EXPECT_EQ(oracle_mats(ft ? 2 : 0, pairs.data(), (int)pairs.size(), (int64_t)mats, c.T, c.stats.wnfactor[step], Z, -1,
                          G.data()),
              0);
    for (size_t n = 0; n < mats; n++) {
      const std::complex<double> o(G[2 * n], G[2 * n + 1]);
      err = std::max(err, std::abs(gf.last_step()[n] - o));
      scale = std::max(scale, std::abs(o));
      want[2 * n] += G[2 * n];
      want[2 * n + 1] += G[2 * n + 1];
    }
    std::printf("    %s: %zu pairs, |G(i w_0)| = %.3e\n", ft ? "FTmats (last step)" : "FDMmats", pairs.size(),
                std::hypot(G[0], G[1]));
  }
  double err_t = 0;
  for (size_t n = 0; n < mats; n++)
    err_t = std::max(err_t, std::abs(gf.values()[n] - std::complex<double>(want[2 * n], want[2 * n + 1])));
  EXPECT_TRUE(scale > 0);
  EXPECT_TRUE(err <= 1e-12 * scale);
  EXPECT_TRUE(err_t <= 2e-12 * scale);
  std::printf("    max deviation %.2e (accumulated %.2e) of max |G| %.3e\n", err, err_t, scale);
}

TEST(Spectrum, temperature_series_match_oracle) {
  // G(T) and chi(T) of two steps, at the temperatures gtp * scale(N)
  auto &c = chain();
  Device &dev = Device::default_device();
  const size_t N = c.Nmax - 1;
  const auto ff = [](const Invar &Ii, const Invar &) { return 1.0 + 0.5 * Ii.qn[0]; };
  for (const int kind : {NRG_TEMP_GT, NRG_TEMP_CHIT}) {
    TempSeries ts(dev, kind);
    for (const size_t step : {N - 1, N}) {
      DiagInfo &diag = c.diags[step];
      MatrixElements op;
      std::map<Twoinvar, Matrix> op_h;
      for (const auto &[I, e] : diag) {
        const Invar Ij{I.qn[0] + 1};
        if (!diag.count(Ij)) continue;
        Matrix m = random_matrix(diag.at(Ij).getnrall(), e.getnrall());
        op_h[{Ij, I}] = m;
        op[{Ij, I}] = DMatrix(m);
      }
      const double temperature = 0.4 * std::pow(2.0, -(double)step / 2), Z = 1.7;
      const double got = ts.calc(diag, op, op, temperature, Z, ff);
      std::vector<nrg_fdm_pair> pairs;
      std::map<Invar, std::vector<double>> E;
      for (const auto &[I, e] : diag) E[I] = e.values.all_abs_zero();
      double scale = 0;
      for (const auto &[II, m] : op_h) {
        const auto &[Ij, Ii] = II;
        nrg_fdm_pair p{};
        p.dim_i = (int64_t)m.size2(); p.dim_j = (int64_t)m.size1();
        p.K_i = (int64_t)diag.at(Ii).getnrkept(); p.K_j = (int64_t)diag.at(Ij).getnrkept();
        p.E_i = E[Ii].data(); p.E_j = E[Ij].data();
        p.op1 = m.data(); p.op2 = m.data(); p.ld_op = (int64_t)m.size2();
        p.factor = ff(Ii, Ij);
        pairs.push_back(p);
        for (size_t r = 0; r < (size_t)p.K_j; r++)
          for (size_t q = 0; q < (size_t)p.K_i; q++) scale = std::max(scale, std::abs(m(r, q) * m(r, q) * p.factor));
      }
      const double want = oracle_temp_sum(kind, pairs.data(), (int)pairs.size(), temperature, Z);
      EXPECT_TRUE(std::abs(got - want) <= 1e-12 * std::max(std::abs(want), 1e-300) + 1e-15 * scale);
      std::printf("    kind %d step %zu: %.6e (oracle %.6e)\n", kind, step, got, want);
    }
    EXPECT_EQ(ts.points().size(), (size_t)2);
  }
}

TEST(Store, steps_round_trip_through_the_budget) {
  // every step's eigenvectors (with their U(i) blocks) and a density matrix per step, under a
  // device budget of about two steps: the earliest move to pinned host memory and come back intact
  auto &c = chain();
  std::vector<DiagInfo> orig(c.Nmax);
  size_t step_bytes = 0;
  for (size_t N = 1; N < c.Nmax; N++) {
    orig[N] = c.diags[N];
    split_in_blocks(orig[N], c.sstruct[N]);
    for (const auto &[I, e] : orig[N]) step_bytes = std::max(step_bytes, 8 * e.vectors.dim() * e.vectors.M());
  }
  StepStore store(3 * step_bytes);
  std::map<size_t, std::map<Invar, Matrix>> rho_h;
  for (size_t N = 1; N < c.Nmax; N++) {
    store.save(N, orig[N]);
    DensMatElements rho;   // any matrices of the kept sizes
    for (const auto &[I, e] : c.diags[N]) {
      rho_h[N][I] = random_matrix(e.getnrkept(), e.getnrkept());
      rho.emplace(I, DMatrix(rho_h[N][I]));
    }
    store.save_rho(N, "rho", std::move(rho));
    EXPECT_TRUE(store.device_bytes() <= store.budget());
  }
  EXPECT_TRUE(store.host_bytes() > 0);
  std::printf("    device %zu B (budget %zu B), host %zu B\n", store.device_bytes(), store.budget(), store.host_bytes());
  for (size_t N = c.Nmax - 1; N >= 1; N--) {   // the backward sweep's order
    const DiagInfo d = store.load(N);
    EXPECT_EQ(d.size(), orig[N].size());
    for (const auto &[I, e] : orig[N]) {
      const Eigen &g = d.at(I);
      EXPECT_TRUE(g.values.all_rel() == e.values.all_rel());
      EXPECT_TRUE(max_abs_diff(g.vectors.get(), e.vectors.get()) == 0.0);
      EXPECT_EQ(g.U.size(), e.U.size());
      for (size_t b = 0; b < e.U.size(); b++) {
        const DView u = e.U.get(b), v = g.U.get(b);
        EXPECT_TRUE(u.rows == v.rows && u.cols == v.cols);
        if (!u.empty()) EXPECT_TRUE(max_abs_diff(download(u), download(v)) == 0.0);
      }
    }
    const DensMatElements r = store.load_rho(N, "rho", true);
    EXPECT_FALSE(store.has_rho(N, "rho"));
    for (const auto &[I, m] : rho_h[N]) EXPECT_TRUE(max_abs_diff(r.at(I).host(), m) == 0.0);
  }
  (void)store.load(2, true);
  EXPECT_FALSE(store.has(2));
  bool threw = false;
  try {
    (void)store.load(2);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  EXPECT_TRUE(threw);
}

TEST(Spectrum, continuous_matches_oracle) {
  // the FDM spectrum of the last step of the toy chain, trimmed and broadened as the reference saves it
  auto &c = chain();
  Device &dev = Device::default_device();
  const size_t N = c.Nmax - 1;
  DiagInfo &diag = c.diags[N];
  MatrixElements op;
  for (const auto &[I, e] : diag) {
    const Invar Ij{I.qn[0] + 1};
    if (diag.count(Ij)) op[{Ij, I}] = DMatrix(random_matrix(diag.at(Ij).getnrall(), e.getnrall()));
  }
  const BinGrid grid = BinGrid::standard(-1, -1, 1.0, std::pow(2.0, -(double)N / 2), 200);
  FDMSpectrum spec(dev, grid, -1, c.T);
  DensMatElements rho = init_rho_FDM(N, c.store, c.stats, c.sym.multfnc(), c.T);
  spec.calc(diag, op, op, rho, c.stats.wnfactor[N], [](const Invar &, const Invar &) { return 1.0; });
  const auto mesh = make_mesh(10.0, 3.0 * std::pow(2.0, -(double)N / 2) * 1e-2, 1.05);
  const SpectrumBins tot = spec.total();
  const Curve cv = continuous(dev, grid, tot, 1e-16, 0.3, c.T, mesh);
  EXPECT_EQ(cv.size(), 2 * mesh.size());
  const auto [pe, pw] = trim(grid, tot.pos, 1e-16);
  const auto [ne, nw] = trim(grid, tot.neg, 1e-16);
  std::vector<double> rp(mesh.size()), rn(mesh.size()), sp(mesh.size()), sn(mesh.size());
  oracle_broaden(pe.data(), pw.data(), (int64_t)pe.size(), ne.data(), nw.data(), (int64_t)ne.size(), mesh.data(),
                 (int64_t)mesh.size(), 0.3, c.T, rp.data(), rn.data());
  std::vector<double> apw(pw), anw(nw);
  for (auto &x : apw) x = std::abs(x);
  for (auto &x : anw) x = std::abs(x);
  oracle_broaden(pe.data(), apw.data(), (int64_t)pe.size(), ne.data(), anw.data(), (int64_t)ne.size(), mesh.data(),
                 (int64_t)mesh.size(), 0.3, c.T, sp.data(), sn.data());
  double worst = 0;
  const size_t M = mesh.size();
  for (size_t k = 0; k < M; k++) {
    // cv: negative energies ascending (-mesh[k] at k, the mesh descends) then positive ascending
    const double gneg = cv[k].second, gpos = cv[M + (M - 1 - k)].second;
    EXPECT_NEAR(cv[k].first, -mesh[k], 0.0);
    EXPECT_NEAR(cv[M + (M - 1 - k)].first, mesh[k], 0.0);
    worst = std::max(worst, std::abs(gpos - rp[k]) / (sp[k] + 1e-300));
    worst = std::max(worst, std::abs(gneg - rn[k]) / (sn[k] + 1e-300));
  }
  EXPECT_TRUE(worst <= 1e-13);
  std::printf("    continuous: %zu + %zu peaks after trim, %zu mesh points, max relative deviation %.2e\n", pe.size(),
              ne.size(), M, worst);
}

int main(int argc, char **argv) {
  const char *blas = nullptr;
  for (int i = 1; i < argc; i++)
    if (!std::strncmp(argv[i], "--blas=", 7)) blas = argv[i] + 7;
  if (!blas || oracle_init(blas) != 0) {
    std::printf("usage: %s --blas=<BLAS/LAPACK library for the oracle>\n", argv[0]);
    return 2;
  }
  oracle_blas_threads(1);
  return minitest_main(argc, argv);
}
