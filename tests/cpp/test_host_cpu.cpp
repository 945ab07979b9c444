// Host-layer tests that need no GPU: containers (c++/eigen.hpp), helpers (c++/numerics.hpp),
// subspace dimensions (c++/subspaces.hpp), store ranges (c++/store.hpp), bin grid (c++/bins.hpp).
// With "--thermo FILE" it instead reads a store description and prints calc_ZnD /
// fdm_thermodynamics results as hex floats for the exact-arithmetic check in Python.
#include <cstdio>
#include <fstream>
#include <sstream>

#include "minitest.hpp"
#include "nrg/nrg.hpp"

using namespace NRG;

static Eigen values_only(std::vector<double> v, bool last = false) {
  Eigen e(std::move(v), DView{nullptr, 0, 0, 0}, nullptr, last);
  e.vectors.set_device(DView{nullptr, 0, e.getnrcomputed(), e.getnrcomputed()}, nullptr);
  return e;
}

TEST(Values, derived_energies) {
  Values v;
  v.set({1.0, 2.0, 4.0});
  EXPECT_FALSE(v.has_abs());
  v.set_scale(0.5);
  v.set_shift(1.0);
  v.set_T_shift(10.0);
  v.set_abs_GS_energy(3.0);
  EXPECT_TRUE(v.has_abs_G());
  EXPECT_DOUBLE_EQ(v.abs(2), 2.0);
  EXPECT_DOUBLE_EQ(v.rel_zero(2), 3.0);
  EXPECT_DOUBLE_EQ(v.abs_zero(2), 1.5);
  EXPECT_DOUBLE_EQ(v.abs_T(2), 11.5);
  EXPECT_DOUBLE_EQ(v.abs_G(2), 8.5);
  EXPECT_EQ(v.all_abs_G().size(), 3u);
  EXPECT_DOUBLE_EQ(v.all_abs_G()[0], 7.0);
}

TEST(Eigen, kept_discarded_ranges) {
  auto e = values_only({0.0, 0.1, 0.2, 0.3, 0.4});
  EXPECT_EQ(e.getnrkept(), 5u);
  e.truncate_prepare(2);
  EXPECT_EQ(e.getnrkept(), 2u);
  EXPECT_EQ(e.getnrdiscarded(), 3u);
  EXPECT_EQ(e.Krange().second, 2u);
  EXPECT_EQ(e.Drange().first, 2u);
  EXPECT_EQ(e.Drange().second, 5u);
  auto l = values_only({0.0, 0.1, 0.2}, true);   // last step: every state is "discarded"
  l.truncate_prepare(1);
  EXPECT_EQ(l.Krange().second, 0u);
  EXPECT_EQ(l.Drange().first, 0u);
  EXPECT_EQ(l.Drange().second, 3u);
}

TEST(Eigen, diagonal_exp_and_trace) {
  auto e = values_only({0.0, 1.0, 2.0});
  e.values.set_corr({0.0, 1.0, 2.0});
  e.values.set_shift(0.0);
  const auto m = e.diagonal_exp(0.5);
  EXPECT_DOUBLE_EQ(m(1, 1), std::exp(-0.5));
  EXPECT_DOUBLE_EQ(m(0, 1), 0.0);
  const double z = e.trace([](double) { return 1.0; }, 0.5);
  EXPECT_DOUBLE_EQ(z, 1.0 + std::exp(-0.5) + std::exp(-1.0));
}

TEST(DiagInfo, groundstate_and_counts) {
  DiagInfo d;
  d[Invar{0}] = values_only({0.5, 1.0});
  d[Invar{1}] = values_only({-0.25, 2.0, 3.0});
  EXPECT_DOUBLE_EQ(d.find_groundstate(), -0.25);
  EXPECT_DOUBLE_EQ(d.Egs_subtraction(), -0.25);
  EXPECT_DOUBLE_EQ(d.at(Invar{0}).values.rel_zero(0), 0.75);
  auto mult = [](const Invar &I) { return I.qn[0] == 1 ? 2.0 : 1.0; };
  EXPECT_DOUBLE_EQ(d.count_states(mult), 2.0 + 6.0);
  EXPECT_EQ(d.count_subspaces(), 2u);
  EXPECT_EQ(d.dims(Invar{0}, Invar{1}).second, 3u);
  EXPECT_EQ(d.size_subspace(Invar{7}), 0u);
  const auto s = d.sorted_energies_rel_zero();
  EXPECT_EQ(s.size(), 5u);
  EXPECT_DOUBLE_EQ(s.front(), 0.0);
}

TEST(Numerics, num_equal_semantics) {   // c++/numerics.hpp:107-120
  EXPECT_TRUE(num_equal(1.0, 1.0 + 1e-13));
  EXPECT_FALSE(num_equal(1.0, 1.0 + 1e-11));
  EXPECT_TRUE(num_equal(1e-13, -1e-13));   // both tiny
  EXPECT_TRUE(num_equal(0.0, 0.0));
}

TEST(Diag, check_is_matrix_upper) {   // as test/unit/diag_test/diag_test.cpp
  Matrix m(2, 2);
  m(0, 0) = m(0, 1) = m(1, 1) = 1.0;
  EXPECT_TRUE(is_square(m));
  EXPECT_TRUE(is_matrix_upper(m));
  m(1, 0) = 1.0;
  EXPECT_FALSE(is_matrix_upper(m));
}

TEST(Diag, is_unitary) {
  Matrix r(2, 2);
  const double c = std::cos(0.3), s = std::sin(0.3);
  r(0, 0) = c; r(0, 1) = s; r(1, 0) = -s; r(1, 1) = c;
  EXPECT_TRUE(is_unitary(r));
  r(1, 1) = 1.1;
  EXPECT_FALSE(is_unitary(r));
  EXPECT_THROW(check_diag({1.0, 2.0}, r));
}

TEST(Diag, params_select_forms) {
  EXPECT_EQ(DiagParams("default", 1.0).mode(), NRG_SYEV_FP64);
  EXPECT_EQ(DiagParams("dsyevd", 1.0).mode(), NRG_SYEV_FP64);
  EXPECT_EQ(DiagParams("mixed", 1.0).mode(), NRG_SYEV_MIXED);
  EXPECT_THROW(DiagParams("zheev", 1.0).mode());
  EXPECT_EQ(DiagParams("dsyevr", 0.5).nr_wanted(9), 5u);   // ceil(4.5), c++/diag.hpp:123-128
  EXPECT_EQ(DiagParams("dsyevr", 1e-9).nr_wanted(9), 1u);
  EXPECT_EQ(DiagParams("dsyev", 0.5).nr_wanted(9), 9u);
}

TEST(Subspaces, dimensions_and_parts) {
  SubspaceDimensions sd({3, 0, 2}, {Invar{0}, Invar{1}, Invar{2}});
  EXPECT_EQ(sd.combs(), 3u);
  EXPECT_EQ(sd.offset(2), 3u);
  EXPECT_EQ(sd.part(2).second, 5u);
  EXPECT_FALSE(sd.exists(1));
  EXPECT_EQ(sd.total(), 5u);
}

// The data-file reader's error paths (c++/read-input.hpp: my_assert on the version, "Corrupted input
// file." for an unknown subspace, std::invalid_argument for an unknown block)
static InputData parse(const std::string &text) {
  std::istringstream in(text);
  return read_input(in);
}
TEST(ReadInput, header_blocks_and_errors) {
  const std::string head = "# symtype QS\n#!9\n1 2 2\n0 1\n1\n0.\n1 2\n2\n0.5 1.5\nf 0 0\n1\n1 2 0 1\n0.25\n-0.5\n";
  const InputData d = parse(head + "e\n-0.125\nsn_d\n1\n0 1 0 1\n2.\nz\n2\n1\n0.5\n0.25\n2\n0\n0\n0\n");
  EXPECT_EQ(d.symtype, std::string("QS"));
  EXPECT_EQ(d.nsubs, (size_t)2);
  EXPECT_EQ(d.subspaces[1].second.size(), (size_t)2);
  EXPECT_EQ(d.opch.size(), (size_t)1);
  const Matrix &f = d.opch[0][0].at({Invar{1, 2}, Invar{0, 1}});
  EXPECT_TRUE(f.size1() == 2 && f.size2() == 1 && f(0, 0) == 0.25 && f(1, 0) == -0.5);
  EXPECT_TRUE(d.has_GS_energy && d.GS_energy == -0.125);
  EXPECT_TRUE(d.ops.size() == 1 && d.ops[0].first == 's' && d.ops[0].second.first == "n_d");
  EXPECT_TRUE(d.coef.at("xi")[0].size() == 3 && d.coef.at("zeta")[0][2] == 0.0);
  EXPECT_THROW(parse("# symtype QS\n#!8\n1 2 2\n"));                       // version
  EXPECT_THROW(parse("# symtype XYZ\n#!9\n1 2 2\n"));                      // symmetry type
  EXPECT_THROW(parse(head + "w\n"));                                         // unknown block
  EXPECT_THROW(parse("# symtype QS\n#!9\n1 2 1\n0 1\n1\n0.\nf 0 0\n1\n1 2 0 1\n0.25\n"));   // unknown subspace
  EXPECT_THROW(parse("# symtype QS\n#!9\n1 2 1\n0 1\n2\n0.\n"));       // truncated energies
}

TEST(Store, fdm_ranges) {
  Sub s{values_only({0, 1, 2, 3}), SubspaceDimensions(), false};
  s.eig.truncate_prepare(1);
  EXPECT_EQ(s.min(), 1u);
  EXPECT_EQ(s.max(), 4u);
  s.is_last = true;
  EXPECT_EQ(s.min(), 0u);
}

TEST(Bins, standard_grid) {   // c++/bins.hpp:61-100
  const auto g = BinGrid::standard(-1.0, -1.0, 1.0, 3e-9, 10);
  EXPECT_DOUBLE_EQ(g.log10emin, -11.0);   // floor(log10(3e-11))
  EXPECT_DOUBLE_EQ(g.log10emax, 2.0);
  EXPECT_EQ(g.nbins, 131u);
  EXPECT_DOUBLE_EQ(g.energy(10), 1e-10);
}

TEST(Quad, truncating_readback) {
  const quad third = (quad)1 / 3;
  const double d = quad_to_double_trunc(third);
  EXPECT_TRUE((quad)d <= third);
  EXPECT_TRUE((quad)std::nextafter(d, 1.0) > third);
}

// --thermo FILE: T combs Nbegin Nend GS_energy, then per subspace:
//   N mult n kept is_last scale shift T_shift abs_GS  e_0 .. e_{n-1}
class ListSym : public Symmetry {
 public:
  size_t combs = 1;
  std::vector<Invar> new_subspaces(const Invar &) const override { return {}; }
  double mult(const Invar &I) const override { return (double)I.qn[1]; }
  size_t nr_combs() const override { return combs; }
};

static int thermo(const char *path) {
  std::ifstream in(path);
  double T, GS;
  size_t combs, Nb, Ne, nsub;
  in >> T >> combs >> Nb >> Ne >> GS >> nsub;
  Store store(Nb, Ne);
  ListSym sym;
  sym.combs = combs;
  for (size_t s = 0; s < nsub; s++) {
    size_t N, n, kept;
    int mult, last;
    double scale, shift, tshift, absgs;
    in >> N >> mult >> n >> kept >> last >> scale >> shift >> tshift >> absgs;
    std::vector<double> e(n);
    for (auto &x : e) in >> x;
    Sub sub{values_only(e, last != 0), SubspaceDimensions(), last != 0};
    sub.eig.truncate_prepare(kept);
    sub.eig.values.set_scale(scale);
    sub.eig.values.set_shift(shift);
    sub.eig.values.set_T_shift(tshift);
    sub.eig.values.set_abs_GS_energy(absgs);
    store[N][Invar{(int)s, mult}] = std::move(sub);
  }
  Stats stats(Ne);
  stats.GS_energy = GS;
  calc_ZnD(store, stats, sym, T);
  fdm_thermodynamics(store, stats, sym, T);
  for (size_t N = Nb; N < Ne; N++) std::printf("%zu %a %a %a\n", N, stats.ZnDNd[N], stats.wn[N], stats.wnfactor[N]);
  std::printf("ZZG %a\nZ_fdm %a\nF_fdm %a\nE_fdm %a\nC_fdm %a\nS_fdm %a\n", stats.ZZG, stats.Z_fdm, stats.F_fdm,
              stats.E_fdm, stats.C_fdm, stats.S_fdm);
  return 0;
}

// --data IN OUT: read_input(IN), write_input to OUT, read OUT back and require the same data
static bool same(const IrredMatrices &a, const IrredMatrices &b) {
  if (a.size() != b.size()) return false;
  for (const auto &[k, m] : a) {
    const auto it = b.find(k);
    if (it == b.end() || it->second.size1() != m.size1() || it->second.size2() != m.size2()) return false;
    for (size_t i = 0; i < m.size1(); i++)
      for (size_t j = 0; j < m.size2(); j++)
        if (m(i, j) != it->second(i, j)) return false;
  }
  return true;
}

static int data_round_trip(const char *in, const char *out) {
  const InputData d = read_input_file(in);
  {
    std::ofstream o(out);
    write_input(o, d);
  }
  const InputData e = read_input_file(out);
  bool ok = d.symtype == e.symtype && d.channels == e.channels && d.Nmax == e.Nmax && d.nsubs == e.nsubs &&
            d.subspaces.size() == d.nsubs && d.subspaces == e.subspaces && d.GS_energy == e.GS_energy &&
            d.has_GS_energy == e.has_GS_energy && d.coef == e.coef && d.coef_blocks == e.coef_blocks &&
            d.opch.size() == e.opch.size() && d.ops.size() == e.ops.size();
  for (size_t c = 0; ok && c < d.opch.size(); c++) {
    ok = d.opch[c].size() == e.opch[c].size();
    for (size_t j = 0; ok && j < d.opch[c].size(); j++) ok = same(d.opch[c][j], e.opch[c][j]);
  }
  for (size_t k = 0; ok && k < d.ops.size(); k++)
    ok = d.ops[k].first == e.ops[k].first && d.ops[k].second.first == e.ops[k].second.first &&
         same(d.ops[k].second.second, e.ops[k].second.second);
  size_t nstates = 0;
  for (const auto &[I, en] : d.subspaces) nstates += en.size();
  std::printf("%s %zu %zu %zu %zu %zu %zu %zu %s\n", d.symtype.c_str(), d.channels, d.Nmax, d.nsubs, nstates,
              d.opch.empty() ? 0 : d.opch[0].size(), d.ops.size(), d.coef.size(), ok ? "same" : "DIFFERENT");
  return ok ? 0 : 1;
}

int main(int argc, char **argv) {
  if (argc == 4 && !std::strcmp(argv[1], "--data")) {
    try {
      return data_round_trip(argv[2], argv[3]);
    } catch (const std::exception &e) {
      std::printf("error: %s\n", e.what());
      return 2;
    }
  }
  if (argc == 3 && !std::strcmp(argv[1], "--thermo")) {
    try {
      return thermo(argv[2]);
    } catch (const std::exception &e) {
      std::printf("error: %s\n", e.what());
      return 2;
    }
  }
  return minitest_main(argc, argv);
}
