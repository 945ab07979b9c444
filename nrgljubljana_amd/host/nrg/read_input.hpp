// read_input.hpp - the initial-data file of NRG Ljubljana (`data`, format version 9).
//
// Follows parse_datafile_header, get_next_block and InputData (c++/read-input.hpp:23-135) with the
// readers they call: DiagInfo(std::istream&, nsubs, P) (c++/eigen.hpp:307-317), Opch(std::istream&)
// (c++/operators.hpp:142-155), MatrixElements(std::istream&, diag) (c++/operators.hpp:30-41),
// Coef::read / read_values (c++/coef.hpp:21-23, 49-52), read_std_vector and skip_comments
// (c++/basicio.hpp:140-154, c++/misc.hpp:145-160). The number of quantum numbers of an invariant
// subspace (Invar::invdim) and of f-operators per channel (P.perchannel) follow from the symmetry
// type as in the symmetry classes' initInvar lists and Params::set_channels_and_combs
// (c++/params.hpp:505-531); the number of coefficient tables is coeffactor * channels, coeffactor
// being 1, 2 (polarized) or 4 (pol2x2), a parameter-file setting.
//
// The energies are returned as written. The reference divides them by SCALE(Ninit) unless
// P.data_has_rescaled_energies or P.absolute: that is the caller's step. write_input writes the same
// format back (the numbers at full precision), so that a file can be checked by a round trip.
#pragma once
#include <cctype>
#include <fstream>
#include <iomanip>
#include <istream>
#include <map>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "recalc.hpp"

namespace NRG {

struct InputLayout {
  size_t invdim = 0;       // quantum numbers per invariant subspace
  size_t perchannel = 0;   // f-operator matrices per channel
};

inline InputLayout input_layout(const std::string &symtype) {
  static const std::map<std::string, size_t> invdim = {
      {"DBLISOSZ", 3}, {"DBLSU2", 2}, {"ISO", 2},    {"ISO2", 2},    {"ISO2LR", 3}, {"ISOLR", 3},  {"ISOSZ", 2},
      {"ISOSZLR", 3},  {"NONE", 1},   {"P", 1},      {"PP", 2},      {"QJ", 2},     {"QS", 2},     {"QSC3", 3},
      {"QSLR", 3},     {"QST", 3},    {"QSTZ", 3},   {"QSZ", 2},     {"QSZLR", 3},  {"QSZTZ", 3},  {"SL", 1},
      {"SL3", 3},      {"SPSU2", 1},  {"SPSU2C3", 2}, {"SPSU2LR", 2}, {"SPSU2T", 2}, {"SPU1", 1},   {"SPU1LR", 2},
      {"SU2", 1},      {"U1", 1}};
  const auto it = invdim.find(symtype);
  if (it == invdim.end()) throw std::invalid_argument("Unknown symmetry type " + symtype);
  InputLayout l;
  l.invdim = it->second;
  if (symtype == "U1" || symtype == "SU2" || symtype == "DBLSU2") l.perchannel = 2;   // spin up and down
  else if (symtype == "NONE" || symtype == "P" || symtype == "PP") l.perchannel = 4;  // CR/AN x UP/DO
  else l.perchannel = 1;
  return l;
}

using IrredMatrices = std::map<Twoinvar, Matrix>;   // <I1||O||I2>, rows = states of I1, cols = of I2

struct InputData {
  std::string symtype;
  size_t channels = 0, Nmax = 0, nsubs = 0;
  std::vector<std::pair<Invar, std::vector<double>>> subspaces;   // in file order, with their energies
  std::vector<std::vector<IrredMatrices>> opch;                   // [channel][index < perchannel]
  std::vector<std::pair<char, std::pair<std::string, IrredMatrices>>> ops;   // 's','p','g','d','t','o','q' blocks
  double GS_energy = 0.0;
  bool has_GS_energy = false;
  // coefficient tables by name (xi, zeta, delta, kappa, xiR, zetaR, ep, em, u0p, u0m), one per
  // coefficient channel
  std::map<std::string, std::vector<std::vector<double>>> coef;
  std::vector<char> coef_blocks;   // the order of the 'z', 'Z', 'X', 'T' blocks in the file
  size_t states(const Invar &I) const {
    for (const auto &[J, e] : subspaces)
      if (J == I) return e.size();
    throw std::runtime_error("Corrupted input file.");
  }
};

namespace detail {
template <class T> T read_value(std::istream &f) {
  T v;
  f >> v;
  if (f.fail()) throw std::runtime_error("Input file is corrupted.");
  return v;
}
inline void skip_comments(std::istream &f) {
  while (f) {
    const int ch = f.peek();
    if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r') {
      f.ignore();
      continue;
    }
    if (ch != '#') break;
    std::string line;
    std::getline(f, line);
  }
}
inline Invar read_invar(std::istream &f, size_t invdim) {
  std::vector<int> q(invdim);
  for (auto &x : q) x = read_value<int>(f);
  return Invar(std::move(q));
}
inline std::vector<double> read_vector(std::istream &f, bool nr_is_max_index) {
  const auto nr = read_value<size_t>(f);
  std::vector<double> v(nr_is_max_index ? nr + 1 : nr);
  for (auto &x : v) x = read_value<double>(f);
  return v;
}
inline IrredMatrices read_irred(std::istream &f, const InputData &d, size_t invdim) {
  IrredMatrices m;
  const auto nf = read_value<size_t>(f);
  for (size_t k = 0; k < nf; k++) {
    const Invar I1 = read_invar(f, invdim), I2 = read_invar(f, invdim);
    Matrix a(d.states(I1), d.states(I2));
    for (size_t i = 0; i < a.size1(); i++)
      for (size_t j = 0; j < a.size2(); j++) a(i, j) = read_value<double>(f);
    m[{I1, I2}] = std::move(a);
  }
  if (m.size() != nf) throw std::runtime_error("Corrupted input file.");
  return m;
}
}  // namespace detail

// parse_datafile_header (c++/read-input.hpp:23-47): the '#' lines before the data, with '#!' the
// version and '# symtype X' the symmetry type
inline std::string parse_datafile_header(std::istream &f, int expected_version = 9) {
  std::string sym;
  int version = -1;
  while (f.peek() == '#') {
    f.ignore();
    if (f.peek() == '!') {
      f.ignore();
      f >> version;
    } else {
      std::string line;
      std::getline(f, line);
      if (line.find("symtype", 1) != std::string::npos) {
        const auto p = line.find_last_of(" \t");
        if (p != std::string::npos && p < line.size() - 1) sym = line.substr(p + 1);
      }
    }
    if (f.peek() == '\n') f.ignore();
  }
  if (version != expected_version)
    throw std::runtime_error("Data file version " + std::to_string(version) + ", expected " +
                             std::to_string(expected_version));
  return sym;
}

// InputData (c++/read-input.hpp:78-132). coeffactor: 1, or 2 / 4 for polarized / pol2x2 bands.
inline InputData read_input(std::istream &f, size_t coeffactor = 1) {
  using namespace detail;
  InputData d;
  d.symtype = parse_datafile_header(f);
  const InputLayout L = input_layout(d.symtype);
  d.channels = read_value<size_t>(f);
  d.Nmax = read_value<size_t>(f);
  d.nsubs = read_value<size_t>(f);
  skip_comments(f);
  for (size_t s = 0; s < d.nsubs; s++) {
    Invar I = read_invar(f, L.invdim);
    d.subspaces.emplace_back(std::move(I), read_vector(f, false));
  }
  skip_comments(f);
  d.opch.assign(d.channels, std::vector<IrredMatrices>(L.perchannel));
  for (auto &oc : d.opch)
    for (auto &o : oc) {
      if (read_value<char>(f) != 'f') throw std::runtime_error("Corrupted input file: expected an f block.");
      (void)read_value<size_t>(f);
      (void)read_value<size_t>(f);
      o = read_irred(f, d, L.invdim);
    }
  const size_t coefch = coeffactor * d.channels;
  auto tables = [&](std::initializer_list<const char *> names) {
    for (const char *n : names) {
      auto &t = d.coef[n];
      t.clear();
      for (size_t c = 0; c < coefch; c++) t.push_back(read_vector(f, true));
    }
  };
  while (true) {   // get_next_block
    while (!f.eof() && std::isspace(f.peek())) f.get();
    const int ch = f.get();
    if (f.eof()) break;
    std::string name;
    std::getline(f, name);
    switch (ch) {
      case '#': break;
      case 'e':
        d.GS_energy = read_value<double>(f);
        d.has_GS_energy = true;
        break;
      case 's': case 'p': case 'g': case 'd': case 't': case 'o': case 'q':
        d.ops.push_back({(char)ch, {name, read_irred(f, d, L.invdim)}});
        break;
      case 'z': tables({"xi", "zeta"}); d.coef_blocks.push_back('z'); break;
      case 'Z': tables({"delta", "kappa"}); d.coef_blocks.push_back('Z'); break;
      case 'X': tables({"xiR", "zetaR"}); d.coef_blocks.push_back('X'); break;
      case 'T': tables({"ep", "em", "u0p", "u0m"}); d.coef_blocks.push_back('T'); break;
      default: throw std::invalid_argument(std::string("Unknown block ") + (char)ch + " in data file.");
    }
  }
  return d;
}

inline InputData read_input_file(const std::string &path, size_t coeffactor = 1) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Can't load initial data.");
  return read_input(f, coeffactor);
}

// The same format, numbers at full precision (17 significant digits), no comments besides the header.
inline void write_input(std::ostream &o, const InputData &d) {
  o << "# symtype " << d.symtype << "\n#!9\n" << d.channels << ' ' << d.Nmax << ' ' << d.nsubs << '\n';
  o << std::setprecision(17);
  auto inv = [&o](const Invar &I) {
    for (size_t k = 0; k < I.qn.size(); k++) o << (k ? " " : "") << I.qn[k];
  };
  auto irred = [&](const IrredMatrices &m) {
    o << m.size() << '\n';
    for (const auto &[II, a] : m) {
      inv(II.first);
      o << ' ';
      inv(II.second);
      o << '\n';
      for (size_t i = 0; i < a.size1(); i++) {
        for (size_t j = 0; j < a.size2(); j++) o << (j ? " " : "") << a(i, j);
        o << '\n';
      }
    }
  };
  for (const auto &[I, e] : d.subspaces) {
    inv(I);
    o << '\n' << e.size() << '\n';
    for (size_t k = 0; k < e.size(); k++) o << (k ? " " : "") << e[k];
    o << '\n';
  }
  for (size_t c = 0; c < d.opch.size(); c++)
    for (size_t j = 0; j < d.opch[c].size(); j++) {
      o << "f " << c << ' ' << j << '\n';
      irred(d.opch[c][j]);
    }
  if (d.has_GS_energy) o << "e\n" << d.GS_energy << '\n';
  for (const auto &[ch, op] : d.ops) {
    o << ch << op.first << '\n';
    irred(op.second);
  }
  static const std::map<char, std::vector<const char *>> names = {
      {'z', {"xi", "zeta"}}, {'Z', {"delta", "kappa"}}, {'X', {"xiR", "zetaR"}}, {'T', {"ep", "em", "u0p", "u0m"}}};
  for (char b : d.coef_blocks) {
    o << b << '\n';
    for (const char *n : names.at(b))
      for (const auto &t : d.coef.at(n)) {
        o << t.size() - 1 << '\n';
        for (double x : t) o << x << '\n';
      }
  }
}

// The matrices of a block in device memory, as the spectral and recalc paths take them.
inline MatrixElements to_device(const IrredMatrices &m) {
  MatrixElements out;
  for (const auto &[k, a] : m) out.emplace(k, DMatrix(a));
  return out;
}

}  // namespace NRG
