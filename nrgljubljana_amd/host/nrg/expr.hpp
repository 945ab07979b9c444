// expr.hpp - arithmetic expressions of the symmetry coefficient tables.
//
// The QS/QSZ tables (c++/symmetry/coefnew/qs*/ of the reference, #include'd there as C++ code) give
// their factors and quantum numbers as expressions in the subspace variables: numbers, + - * /,
// parentheses, unary signs, the functions S(x) = (x-1)/2 (c++/invar.hpp: S, ISO, SZ), sqrt(x) and
// Power(x, y), and variable names (q1, ss1, ssz1, qp, ssp, sszp, ss). Expr parses such a string once
// into a small tree and evaluates it for given variable values.
#pragma once
#include <cctype>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace NRG {

using ExprVars = std::map<std::string, double>;

class Expr {
 public:
  Expr() = default;
  explicit Expr(const std::string &text) : text_(text) {
    size_t pos = 0;
    root_ = parse_sum(pos);
    skip(pos);
    if (pos != text_.size()) fail("unexpected '" + text_.substr(pos) + "'");
  }
  double operator()(const ExprVars &vars) const {
    if (!root_) throw std::runtime_error("Expr: empty expression");
    return eval(*root_, vars);
  }
  const std::string &text() const { return text_; }
  // true if the expression has no variables (a number)
  bool is_constant() const { return root_ && constant(*root_); }

 private:
  struct Node {
    char op;   // 'n' number, 'v' variable, 'f' function, '+', '-', '*', '/', 'u' (unary minus)
    double value = 0;
    std::string name;
    std::vector<std::unique_ptr<Node>> args;
  };
  using P = std::unique_ptr<Node>;

  [[noreturn]] void fail(const std::string &what) const {
    throw std::runtime_error("table expression '" + text_ + "': " + what);
  }
  void skip(size_t &pos) const {
    while (pos < text_.size() && std::isspace((unsigned char)text_[pos])) pos++;
  }
  static P node(char op) {
    auto n = std::make_unique<Node>();
    n->op = op;
    return n;
  }
  P parse_sum(size_t &pos) const {
    P left = parse_product(pos);
    for (;;) {
      skip(pos);
      if (pos >= text_.size() || (text_[pos] != '+' && text_[pos] != '-')) return left;
      const char op = text_[pos++];
      P n = node(op);
      n->args.push_back(std::move(left));
      n->args.push_back(parse_product(pos));
      left = std::move(n);
    }
  }
  P parse_product(size_t &pos) const {
    P left = parse_unary(pos);
    for (;;) {
      skip(pos);
      if (pos >= text_.size() || (text_[pos] != '*' && text_[pos] != '/')) return left;
      const char op = text_[pos++];
      P n = node(op);
      n->args.push_back(std::move(left));
      n->args.push_back(parse_unary(pos));
      left = std::move(n);
    }
  }
  P parse_unary(size_t &pos) const {
    skip(pos);
    if (pos < text_.size() && (text_[pos] == '-' || text_[pos] == '+')) {
      const char op = text_[pos++];
      P inner = parse_unary(pos);
      if (op == '+') return inner;
      P n = node('u');
      n->args.push_back(std::move(inner));
      return n;
    }
    return parse_primary(pos);
  }
  P parse_primary(size_t &pos) const {
    skip(pos);
    if (pos >= text_.size()) fail("unexpected end");
    const char c = text_[pos];
    if (c == '(') {
      pos++;
      P inner = parse_sum(pos);
      skip(pos);
      if (pos >= text_.size() || text_[pos] != ')') fail("missing ')'");
      pos++;
      return inner;
    }
    if (std::isdigit((unsigned char)c) || c == '.') {
      size_t end = pos;
      while (end < text_.size() && (std::isdigit((unsigned char)text_[end]) || text_[end] == '.')) end++;
      if (end < text_.size() && (text_[end] == 'e' || text_[end] == 'E')) {
        size_t e = end + 1;
        if (e < text_.size() && (text_[e] == '+' || text_[e] == '-')) e++;
        if (e < text_.size() && std::isdigit((unsigned char)text_[e])) {
          end = e;
          while (end < text_.size() && std::isdigit((unsigned char)text_[end])) end++;
        }
      }
      P n = node('n');
      n->value = std::stod(text_.substr(pos, end - pos));
      pos = end;
      return n;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t end = pos;
      while (end < text_.size() && (std::isalnum((unsigned char)text_[end]) || text_[end] == '_')) end++;
      const std::string name = text_.substr(pos, end - pos);
      pos = end;
      skip(pos);
      if (pos < text_.size() && text_[pos] == '(') {   // function call
        pos++;
        P n = node('f');
        n->name = name;
        for (;;) {
          n->args.push_back(parse_sum(pos));
          skip(pos);
          if (pos < text_.size() && text_[pos] == ',') { pos++; continue; }
          if (pos < text_.size() && text_[pos] == ')') { pos++; break; }
          fail("bad argument list of " + name);
        }
        const size_t want = name == "Power" ? 2 : 1;
        if ((name != "S" && name != "SZ" && name != "ISO" && name != "sqrt" && name != "Power") || n->args.size() != want)
          fail("unknown function " + name + " with " + std::to_string(n->args.size()) + " arguments");
        return n;
      }
      P n = node('v');
      n->name = name;
      return n;
    }
    fail(std::string("unexpected '") + c + "'");
  }
  double eval(const Node &n, const ExprVars &vars) const {
    switch (n.op) {
      case 'n': return n.value;
      case 'v': {
        const auto it = vars.find(n.name);
        if (it == vars.end()) fail("variable " + n.name + " not set");
        return it->second;
      }
      case 'u': return -eval(*n.args[0], vars);
      case '+': return eval(*n.args[0], vars) + eval(*n.args[1], vars);
      case '-': return eval(*n.args[0], vars) - eval(*n.args[1], vars);
      case '*': return eval(*n.args[0], vars) * eval(*n.args[1], vars);
      case '/': return eval(*n.args[0], vars) / eval(*n.args[1], vars);
      case 'f': {
        const double a = eval(*n.args[0], vars);
        if (n.name == "sqrt") return std::sqrt(a);
        if (n.name == "Power") return std::pow(a, eval(*n.args[1], vars));
        return (a - 1.0) / 2.0;   // S, SZ, ISO: multiplicity -> quantum number
      }
    }
    fail("corrupt expression tree");
  }
  static bool constant(const Node &n) {
    if (n.op == 'v') return false;
    for (const auto &a : n.args)
      if (!constant(*a)) return false;
    return true;
  }
  std::string text_;
  std::shared_ptr<Node> root_;
};

}  // namespace NRG
