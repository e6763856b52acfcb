// OpenQASM 2.0 front end for the ket:: API (reference interface
// include/ket/qasm.hpp: from_qasm / to_qasm; behaviour of
// bindings/qasm/qasm.cpp:308-544 for parsing and :551-782 for writing).
//
// from_qasm: registers, the qelib1 gates ket has, U/u/u1/u2/u3/p/phase,
// measure (single and register-wide), barrier, user `gate` definitions
// (appended as composite blocks, usable before they are defined), `pi`
// expressions with + - * / ^, unary signs, parentheses and
// sin/cos/tan/exp/ln/sqrt. Anything else (opaque, reset, if, classical
// control) throws std::runtime_error. Every input ket's parser accepts gives
// the same circuit here.
//
// to_qasm: the text ket writes, byte for byte: one `gate` definition per
// distinct composite (dependencies first), pi-fraction angles, probes left
// out.
//
// Written as a tokenizer plus recursive-descent parser; the circuit is built
// after a first pass has collected every register and a second every gate
// definition, so several qregs become one qubit range in declaration order.
#include <ket/qasm.hpp>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <ket/circuit.hpp>
#include <ket/dag.hpp>
#include <ket/gate.hpp>

namespace ket {
namespace {

[[noreturn]] void bad(const std::string& what, int line) {
  throw std::runtime_error("qasm: line " + std::to_string(line) + ": " + what);
}

struct Tok {
  enum Kind { Id, Num, Str, Sym, End } kind;
  std::string text;
  double num = 0.0;
  int line = 1;
};

std::vector<Tok> lex(const std::string& src) {
  std::vector<Tok> out;
  int line = 1;
  for (std::size_t i = 0; i < src.size();) {
    const char c = src[i];
    if (c == '\n') {
      ++line;
      ++i;
    } else if (std::isspace(static_cast<unsigned char>(c))) {
      ++i;
    } else if (c == '/' && i + 1 < src.size() && src[i + 1] == '/') {
      while (i < src.size() && src[i] != '\n') ++i;
    } else if (c == '/' && i + 1 < src.size() && src[i + 1] == '*') {
      i += 2;
      while (i + 1 < src.size() && !(src[i] == '*' && src[i + 1] == '/')) {
        if (src[i] == '\n') ++line;
        ++i;
      }
      i += 2;
    } else if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
      std::size_t j = i;
      while (j < src.size() && (std::isalnum(static_cast<unsigned char>(src[j])) || src[j] == '_'))
        ++j;
      out.push_back({Tok::Id, src.substr(i, j - i), 0.0, line});
      i = j;
    } else if (std::isdigit(static_cast<unsigned char>(c)) ||
               (c == '.' && i + 1 < src.size() && std::isdigit(static_cast<unsigned char>(src[i + 1])))) {
      std::size_t used = 0;
      const double v = std::stod(src.substr(i), &used);
      out.push_back({Tok::Num, src.substr(i, used), v, line});
      i += used;
    } else if (c == '"') {
      const std::size_t j = src.find('"', i + 1);
      if (j == std::string::npos) bad("unterminated string", line);
      out.push_back({Tok::Str, src.substr(i + 1, j - i - 1), 0.0, line});
      i = j + 1;
    } else if (c == '-' && i + 1 < src.size() && src[i + 1] == '>') {
      out.push_back({Tok::Sym, "->", 0.0, line});
      i += 2;
    } else if (c == '=' && i + 1 < src.size() && src[i + 1] == '=') {
      out.push_back({Tok::Sym, "==", 0.0, line});
      i += 2;
    } else {
      out.push_back({Tok::Sym, std::string(1, c), 0.0, line});
      ++i;
    }
  }
  out.push_back({Tok::End, "", 0.0, line});
  return out;
}

// Parameter expressions: an AST evaluated against gate-parameter bindings.
struct Expr {
  enum Op { Num, Var, Neg, Add, Sub, Mul, Div, Pow, Fn } op;
  double value = 0.0;
  std::string name;
  std::shared_ptr<Expr> a, b;
  double eval(const std::map<std::string, double>& env, int line) const {
    switch (op) {
      case Num: return value;
      case Var: {
        auto it = env.find(name);
        if (it == env.end()) bad("unknown parameter '" + name + "'", line);
        return it->second;
      }
      case Neg: return -a->eval(env, line);
      case Add: return a->eval(env, line) + b->eval(env, line);
      case Sub: return a->eval(env, line) - b->eval(env, line);
      case Mul: return a->eval(env, line) * b->eval(env, line);
      case Div: return a->eval(env, line) / b->eval(env, line);
      case Pow: return std::pow(a->eval(env, line), b->eval(env, line));
      case Fn: {
        const double x = a->eval(env, line);
        if (name == "sin") return std::sin(x);
        if (name == "cos") return std::cos(x);
        if (name == "tan") return std::tan(x);
        if (name == "exp") return std::exp(x);
        if (name == "ln") return std::log(x);
        if (name == "sqrt") return std::sqrt(x);
        bad("unknown function '" + name + "'", line);
      }
    }
    return 0.0;
  }
};
using ExprP = std::shared_ptr<Expr>;

ExprP node(Expr::Op op) {
  auto e = std::make_shared<Expr>();
  e->op = op;
  return e;
}

struct GateDef {
  std::vector<std::string> params, args;
  struct Stmt {
    std::string name;
    std::vector<ExprP> params;
    std::vector<std::string> args;
    int line;
  };
  std::vector<Stmt> body;
};

struct Reg {
  std::size_t offset, size;
};

class Parser {
 public:
  explicit Parser(std::vector<Tok> toks) : t_(std::move(toks)) {}

  Circuit parse() {
    // Pass 1: registers, so the circuit width is known before any gate.
    std::size_t nq = 0, nc = 0;
    for (std::size_t i = 0; i + 4 < t_.size(); ++i) {
      if (t_[i].kind == Tok::Id && (t_[i].text == "qreg" || t_[i].text == "creg") &&
          t_[i + 1].kind == Tok::Id && t_[i + 2].text == "[" && t_[i + 3].kind == Tok::Num) {
        const std::size_t n = static_cast<std::size_t>(t_[i + 3].num);
        if (t_[i].text == "qreg") {
          qregs_[t_[i + 1].text] = {nq, n};
          nq += n;
        } else {
          cregs_[t_[i + 1].text] = {nc, n};
          nc += n;
        }
      }
    }
    // Pass 2: gate definitions, so a gate may be used before (or inside a
    // definition written before) its own definition, as ket allows. The
    // first definition of a name is kept (bindings/qasm/qasm.cpp:488-500).
    while (peek().kind != Tok::End) {
      if (peek().kind == Tok::Id && peek().text == "gate") {
        next();
        gate_def(true);
      } else {
        while (peek().kind != Tok::End && !accept(";")) next();
      }
    }
    p_ = 0;
    Circuit c(nq);
    while (peek().kind != Tok::End) statement(c);
    return c;
  }

 private:
  const Tok& peek() const { return t_[p_]; }
  const Tok& next() { return t_[p_++]; }
  bool accept(const std::string& s) {
    if (peek().text == s && peek().kind != Tok::Str) {
      ++p_;
      return true;
    }
    return false;
  }
  void expect(const std::string& s) {
    if (!accept(s)) bad("expected '" + s + "', found '" + peek().text + "'", peek().line);
  }
  std::string ident() {
    if (peek().kind != Tok::Id) bad("expected a name, found '" + peek().text + "'", peek().line);
    return next().text;
  }

  ExprP primary() {
    const Tok& k = peek();
    if (accept("-")) {
      auto e = node(Expr::Neg);
      e->a = primary();
      return e;
    }
    if (accept("+")) return primary();
    if (accept("(")) {
      ExprP e = expr();
      expect(")");
      return e;
    }
    if (k.kind == Tok::Num) {
      ++p_;
      auto e = node(Expr::Num);
      e->value = k.num;
      return e;
    }
    if (k.kind == Tok::Id) {
      const std::string name = next().text;
      if (name == "pi") {
        auto e = node(Expr::Num);
        e->value = 3.14159265358979323846;
        return e;
      }
      if (accept("(")) {
        auto e = node(Expr::Fn);
        e->name = name;
        e->a = expr();
        expect(")");
        return e;
      }
      auto e = node(Expr::Var);
      e->name = name;
      return e;
    }
    bad("bad expression at '" + k.text + "'", k.line);
  }
  ExprP power() {
    ExprP a = primary();
    if (accept("^")) {
      auto e = node(Expr::Pow);
      e->a = a;
      e->b = power();
      return e;
    }
    return a;
  }
  ExprP term() {
    ExprP a = power();
    for (;;) {
      Expr::Op op;
      if (accept("*")) op = Expr::Mul;
      else if (accept("/")) op = Expr::Div;
      else return a;
      auto e = node(op);
      e->a = a;
      e->b = power();
      a = e;
    }
  }
  ExprP expr() {
    ExprP a = term();
    for (;;) {
      Expr::Op op;
      if (accept("+")) op = Expr::Add;
      else if (accept("-")) op = Expr::Sub;
      else return a;
      auto e = node(op);
      e->a = a;
      e->b = term();
      a = e;
    }
  }

  // A register operand: name or name[i]; returns the absolute indices.
  std::vector<std::size_t> operand(const std::map<std::string, Reg>& regs, int line) {
    const std::string name = ident();
    auto it = regs.find(name);
    if (it == regs.end()) bad("unknown register '" + name + "'", line);
    if (accept("[")) {
      if (peek().kind != Tok::Num) bad("expected an index", line);
      const std::size_t i = static_cast<std::size_t>(next().num);
      expect("]");
      if (i >= it->second.size) bad("index out of range for '" + name + "'", line);
      return {it->second.offset + i};
    }
    std::vector<std::size_t> all;
    for (std::size_t i = 0; i < it->second.size; ++i) all.push_back(it->second.offset + i);
    return all;
  }

  void statement(Circuit& c) {
    const Tok& k = peek();
    const int line = k.line;
    if (k.kind != Tok::Id) bad("unexpected '" + k.text + "'", line);
    const std::string w = next().text;
    if (w == "OPENQASM") {
      next();
      expect(";");
    } else if (w == "include") {
      if (peek().kind != Tok::Str) bad("include expects a file name", line);
      next();
      expect(";");
    } else if (w == "qreg" || w == "creg") {
      ident();
      expect("[");
      next();
      expect("]");
      expect(";");
    } else if (w == "gate") {
      gate_def(false);
    } else if (w == "measure") {
      const auto q = operand(qregs_, line);
      expect("->");
      const auto cl = operand(cregs_, line);
      expect(";");
      if (q.size() != cl.size()) bad("measure register sizes differ", line);
      for (std::size_t i = 0; i < q.size(); ++i) c.measure(q[i], cl[i]);
    } else if (w == "barrier") {
      std::vector<std::size_t> qs;
      do {
        for (std::size_t q : operand(qregs_, line)) qs.push_back(q);
      } while (accept(","));
      expect(";");
      c.barrier(qs);
    } else if (w == "opaque" || w == "reset" || w == "if") {
      bad("'" + w + "' is not supported", line);
    } else {
      apply(c, w, line);
    }
  }

  void gate_def(bool keep) {
    const std::string name = ident();
    GateDef d;
    if (accept("(")) {
      if (!accept(")")) {
        do d.params.push_back(ident());
        while (accept(","));
        expect(")");
      }
    }
    do d.args.push_back(ident());
    while (accept(","));
    expect("{");
    while (!accept("}")) {
      GateDef::Stmt s;
      s.line = peek().line;
      s.name = ident();
      if (accept("(")) {
        if (!accept(")")) {
          do s.params.push_back(expr());
          while (accept(","));
          expect(")");
        }
      }
      do s.args.push_back(ident());
      while (accept(","));
      expect(";");
      d.body.push_back(std::move(s));
    }
    if (keep) defs_.emplace(name, std::move(d));
  }

  void apply(Circuit& c, const std::string& name, int line) {
    std::vector<double> params;
    if (accept("(")) {
      if (!accept(")")) {
        do params.push_back(expr()->eval({}, line));
        while (accept(","));
        expect(")");
      }
    }
    std::vector<std::vector<std::size_t>> args;
    do args.push_back(operand(qregs_, line));
    while (accept(","));
    expect(";");
    // Register broadcast: every operand a whole register of one size, or a
    // single qubit.
    std::size_t width = 1;
    for (const auto& a : args)
      if (a.size() > 1) {
        if (width > 1 && a.size() != width) bad("register sizes differ", line);
        width = a.size();
      }
    for (std::size_t i = 0; i < width; ++i) {
      std::vector<std::size_t> qs;
      for (const auto& a : args) qs.push_back(a.size() == 1 ? a[0] : a[i]);
      emit(c, name, params, qs, line);
    }
  }

  void emit(Circuit& c, const std::string& name, const std::vector<double>& p,
            const std::vector<std::size_t>& q, int line) {
    // Gates without parameters ignore a parameter list, as ket's parser does.
    constexpr std::size_t kAny = static_cast<std::size_t>(-1);
    auto need = [&](std::size_t nq, std::size_t np) {
      if (q.size() != nq || (np != kAny && p.size() != np))
        bad("'" + name + "' takes " + std::to_string(nq) + " qubits and " +
                std::to_string(np) + " parameters",
            line);
    };
    const double pi = 3.14159265358979323846;
    if (name == "id") { need(1, kAny); return; }
    if (name == "h") { need(1, kAny); c.h(q[0]); return; }
    if (name == "x") { need(1, kAny); c.x(q[0]); return; }
    if (name == "y") { need(1, kAny); c.y(q[0]); return; }
    if (name == "z") { need(1, kAny); c.z(q[0]); return; }
    if (name == "s") { need(1, kAny); c.s(q[0]); return; }
    if (name == "sdg") { need(1, kAny); c.sdg(q[0]); return; }
    if (name == "t") { need(1, kAny); c.t(q[0]); return; }
    if (name == "tdg") { need(1, kAny); c.tdg(q[0]); return; }
    if (name == "rx") { need(1, 1); c.rx(q[0], p[0]); return; }
    if (name == "ry") { need(1, 1); c.ry(q[0], p[0]); return; }
    if (name == "rz") { need(1, 1); c.rz(q[0], p[0]); return; }
    if (name == "u" || name == "U" || name == "u3") { need(1, 3); c.u(q[0], p[0], p[1], p[2]); return; }
    if (name == "u2") { need(1, 2); c.u(q[0], pi / 2, p[0], p[1]); return; }
    if (name == "u1" || name == "p" || name == "phase") { need(1, 1); c.u(q[0], 0.0, 0.0, p[0]); return; }
    if (name == "cx" || name == "CX") { need(2, kAny); c.cx(q[0], q[1]); return; }
    if (name == "cy") { need(2, kAny); c.cy(q[0], q[1]); return; }
    if (name == "cz") { need(2, kAny); c.cz(q[0], q[1]); return; }
    if (name == "ch") { need(2, kAny); c.ch(q[0], q[1]); return; }
    if (name == "crx") { need(2, 1); c.crx(q[0], q[1], p[0]); return; }
    if (name == "cry") { need(2, 1); c.cry(q[0], q[1], p[0]); return; }
    if (name == "crz") { need(2, 1); c.crz(q[0], q[1], p[0]); return; }
    if (name == "cu1" || name == "cp") { need(2, 1); c.cp(q[0], q[1], p[0]); return; }
    if (name == "cu3" || name == "cu") { need(2, 3); c.cu(q[0], q[1], p[0], p[1], p[2]); return; }
    if (name == "swap") { need(2, kAny); c.swap(q[0], q[1]); return; }
    if (name == "ccx") { need(3, kAny); c.ccx(q[0], q[1], q[2]); return; }
    if (name == "cswap") { need(3, kAny); c.cswap(q[0], q[1], q[2]); return; }
    auto it = defs_.find(name);
    if (it == defs_.end()) bad("unknown gate '" + name + "'", line);
    const GateDef& d = it->second;
    need(d.args.size(), d.params.size());
    c.append(instantiate(name, d, p, line), q, name);
  }

  // The definition's body as a sub-circuit over its own arguments.
  Circuit instantiate(const std::string& name, const GateDef& d, const std::vector<double>& p,
                      int line) {
    std::map<std::string, double> env;
    for (std::size_t i = 0; i < d.params.size(); ++i) env[d.params[i]] = p[i];
    std::map<std::string, std::size_t> arg;
    for (std::size_t i = 0; i < d.args.size(); ++i) arg[d.args[i]] = i;
    Circuit sub(d.args.size(), name);
    for (const GateDef::Stmt& s : d.body) {
      std::vector<double> vals;
      for (const ExprP& e : s.params) vals.push_back(e->eval(env, s.line));
      std::vector<std::size_t> qs;
      for (const std::string& a : s.args) {
        auto f = arg.find(a);
        if (f == arg.end()) bad("unknown argument '" + a + "' in gate '" + name + "'", s.line);
        qs.push_back(f->second);
      }
      if (s.name == "barrier")
        sub.barrier(qs);
      else
        emit(sub, s.name, vals, qs, s.line);
    }
    (void)line;
    return sub;
  }

  std::vector<Tok> t_;
  std::size_t p_ = 0;
  std::map<std::string, Reg> qregs_, cregs_;
  std::map<std::string, GateDef> defs_;
};

// Angle text of the serializer. Multiples n*pi/d with d <= 12 (to 1e-9 of
// the ratio) print as pi fractions, |angle| < 1e-12 as 0, anything else with
// 15 significant digits. The rule fixes the text ket writes
// (bindings/qasm/qasm.cpp:554-575), so it is kept as ket has it; the text
// re-parses to the same text (pi fractions evaluate back within the 1e-9
// window, 15-digit decimals print unchanged).
std::string angle_text(double a) {
  if (std::fabs(a) < 1e-12) return "0";
  const double pi = std::acos(-1.0);
  const double ratio = a / pi;
  for (int d = 1; d <= 12; ++d) {
    const double scaled = ratio * d;
    const long n = std::lround(scaled);
    if (n == 0 || !(std::fabs(scaled - static_cast<double>(n)) < 1e-9)) continue;
    const long m = n < 0 ? -n : n;
    std::string t = m == 1 ? std::string("pi") : std::to_string(m) + "*pi";
    if (d > 1) t += "/" + std::to_string(d);
    return n < 0 ? "-" + t : t;
  }
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.15g", a);
  return buf;
}

// OpenQASM spelling of each GateType, and whether its parameters print.
// ket writes U for the general one-qubit gate, cu3 for CU and cp for CP
// (bindings/qasm/qasm.cpp:629-700).
const char* gate_word(GateType t) {
  switch (t) {
    case GateType::H: return "h";
    case GateType::X: return "x";
    case GateType::Y: return "y";
    case GateType::Z: return "z";
    case GateType::S: return "s";
    case GateType::Sdg: return "sdg";
    case GateType::T: return "t";
    case GateType::Tdg: return "tdg";
    case GateType::Rx: return "rx";
    case GateType::Ry: return "ry";
    case GateType::Rz: return "rz";
    case GateType::U: return "U";
    case GateType::CH: return "ch";
    case GateType::CX: return "cx";
    case GateType::CY: return "cy";
    case GateType::CZ: return "cz";
    case GateType::CRx: return "crx";
    case GateType::CRy: return "cry";
    case GateType::CRz: return "crz";
    case GateType::CU: return "cu3";
    case GateType::CP: return "cp";
    case GateType::Swap: return "swap";
    case GateType::CCX: return "ccx";
    case GateType::CSwap: return "cswap";
    default: return nullptr;
  }
}

// Writes a circuit as OpenQASM. Every distinct composite becomes one `gate`
// definition over formal qubits q0, q1, ... and is invoked by name; a
// definition is written before any definition or statement that uses it
// (rendering a body registers the composites inside it first). Two
// composites share a definition when label, arity and rendered body agree;
// a label reused with another body gets the first free suffix _2, _3, ...
// Unlabelled composites are called `block`. Probes have no OpenQASM form.
class Writer {
 public:
  std::string circuit(const Circuit& c) {
    std::string body;
    for (const DagNode& node : c.dag().nodes())
      statement(body, node.gate, [](std::size_t q) { return "q[" + std::to_string(q) + "]"; });
    std::string out = "OPENQASM 2.0;\ninclude \"qelib1.inc\";\n";
    out += defs_;
    out += "qreg q[" + std::to_string(c.n_qubits()) + "];\n";
    if (c.n_clbits() > 0) out += "creg c[" + std::to_string(c.n_clbits()) + "];\n";
    return out + body;
  }

 private:
  template <class Ref>
  void operands(std::string& os, const Gate& g, Ref ref) {
    for (std::size_t k = 0; k < g.qubits.size(); ++k) {
      os += k == 0 ? " " : ",";
      os += ref(g.qubits[k].index);
    }
  }

  template <class Ref>
  void statement(std::string& os, const Gate& g, Ref ref) {
    switch (g.type) {
      case GateType::Probe:
        return;
      case GateType::Measure:
        os += "measure " + ref(g.qubits[0].index) + " -> c[" + std::to_string(g.clbit) + "];\n";
        return;
      case GateType::Barrier:
        os += "barrier";
        if (g.qubits.empty()) os += " ";
        operands(os, g, ref);
        os += ";\n";
        return;
      case GateType::Composite:
        if (!g.definition) return;
        os += definition(*g.definition, g.label);
        operands(os, g, ref);
        os += ";\n";
        return;
      default:
        break;
    }
    os += gate_word(g.type);
    if (!g.params.empty()) {
      os += "(";
      for (std::size_t i = 0; i < g.params.size(); ++i) {
        if (i) os += ",";
        os += angle_text(g.params[i]);
      }
      os += ")";
    }
    operands(os, g, ref);
    os += ";\n";
  }

  // The name a composite is invoked by; writes its definition on first use.
  std::string definition(const Circuit& def, const std::string& label) {
    std::string body;
    for (const DagNode& node : def.dag().nodes())
      statement(body, node.gate, [](std::size_t q) { return "q" + std::to_string(q); });
    std::string key = label;
    key += '#';
    key += std::to_string(def.n_qubits());
    key += '\x1f';
    key += body;
    if (auto it = named_.find(key); it != named_.end()) return it->second;
    std::string name = label.empty() ? "block" : label;
    if (taken_.count(name)) {
      int k = 2;
      while (taken_.count(name + "_" + std::to_string(k))) ++k;
      name += "_" + std::to_string(k);
    }
    taken_.insert(name);
    named_.emplace(std::move(key), name);
    defs_ += "gate " + name;
    for (std::size_t i = 0; i < def.n_qubits(); ++i) defs_ += (i ? ",q" : " q") + std::to_string(i);
    defs_ += " {\n";
    std::size_t a = 0;
    while (a < body.size()) {
      std::size_t b = body.find('\n', a);
      if (b == std::string::npos) b = body.size();
      if (b > a) defs_ += "  " + body.substr(a, b - a) + "\n";
      a = b + 1;
    }
    defs_ += "}\n";
    return name;
  }

  std::string defs_;
  std::map<std::string, std::string> named_;
  std::set<std::string> taken_;
};

}  // namespace

Circuit from_qasm(const std::string& source) { return Parser(lex(source)).parse(); }

std::string to_qasm(const Circuit& circuit) { return Writer().circuit(circuit); }

}  // namespace ket
