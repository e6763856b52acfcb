// OpenQASM 2.0 front end for the ket:: API (reference interface
// include/ket/qasm.hpp: from_qasm / to_qasm; behaviour of
// bindings/qasm/qasm.cpp:308-544 and :761): registers, the qelib1 gates ket
// has, u/u1/u2/u3/p, measure (single and register-wide), barrier, user
// `gate` definitions (appended as composite blocks), `pi` expressions with
// + - * / ^, unary signs, parentheses and sin/cos/tan/exp/ln/sqrt. Anything
// else (opaque, reset, if, classical control) throws std::runtime_error.
//
// Written as a tokenizer plus recursive-descent parser; the circuit is built
// after a first pass has collected every register, so several qregs become
// one qubit range in declaration order.
#include <ket/qasm.hpp>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
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
      gate_def();
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

  void gate_def() {
    const int line = peek().line;
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
      if (s.name == "barrier") {
        while (!accept(";")) next();
        continue;
      }
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
    if (defs_.count(name)) bad("gate '" + name + "' defined twice", line);
    defs_[name] = std::move(d);
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
    auto need = [&](std::size_t nq, std::size_t np) {
      if (q.size() != nq || p.size() != np)
        bad("'" + name + "' takes " + std::to_string(nq) + " qubits and " +
                std::to_string(np) + " parameters",
            line);
    };
    const double pi = 3.14159265358979323846;
    if (name == "id") { need(1, 0); return; }
    if (name == "h") { need(1, 0); c.h(q[0]); return; }
    if (name == "x") { need(1, 0); c.x(q[0]); return; }
    if (name == "y") { need(1, 0); c.y(q[0]); return; }
    if (name == "z") { need(1, 0); c.z(q[0]); return; }
    if (name == "s") { need(1, 0); c.s(q[0]); return; }
    if (name == "sdg") { need(1, 0); c.sdg(q[0]); return; }
    if (name == "t") { need(1, 0); c.t(q[0]); return; }
    if (name == "tdg") { need(1, 0); c.tdg(q[0]); return; }
    if (name == "rx") { need(1, 1); c.rx(q[0], p[0]); return; }
    if (name == "ry") { need(1, 1); c.ry(q[0], p[0]); return; }
    if (name == "rz") { need(1, 1); c.rz(q[0], p[0]); return; }
    if (name == "u" || name == "U" || name == "u3") { need(1, 3); c.u(q[0], p[0], p[1], p[2]); return; }
    if (name == "u2") { need(1, 2); c.u(q[0], pi / 2, p[0], p[1]); return; }
    if (name == "u1" || name == "p") { need(1, 1); c.u(q[0], 0.0, 0.0, p[0]); return; }
    if (name == "cx" || name == "CX") { need(2, 0); c.cx(q[0], q[1]); return; }
    if (name == "cy") { need(2, 0); c.cy(q[0], q[1]); return; }
    if (name == "cz") { need(2, 0); c.cz(q[0], q[1]); return; }
    if (name == "ch") { need(2, 0); c.ch(q[0], q[1]); return; }
    if (name == "crx") { need(2, 1); c.crx(q[0], q[1], p[0]); return; }
    if (name == "cry") { need(2, 1); c.cry(q[0], q[1], p[0]); return; }
    if (name == "crz") { need(2, 1); c.crz(q[0], q[1], p[0]); return; }
    if (name == "cu1" || name == "cp") { need(2, 1); c.cp(q[0], q[1], p[0]); return; }
    if (name == "cu3" || name == "cu") { need(2, 3); c.cu(q[0], q[1], p[0], p[1], p[2]); return; }
    if (name == "swap") { need(2, 0); c.swap(q[0], q[1]); return; }
    if (name == "ccx") { need(3, 0); c.ccx(q[0], q[1], q[2]); return; }
    if (name == "cswap") { need(3, 0); c.cswap(q[0], q[1], q[2]); return; }
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

void write_gates(std::ostringstream& os, const Circuit& c, const std::vector<std::size_t>& wire) {
  static const char* names[] = {"h", "x", "y", "z", "s", "sdg", "t", "tdg", "rx", "ry", "rz",
                                "u3", "ch", "cx", "cy", "cz", "crx", "cry", "crz", "cu3", "cu1",
                                "swap", "ccx", "cswap"};
  char buf[64];
  for (const DagNode& node : c.dag().nodes()) {
    const Gate& g = node.gate;
    switch (g.type) {
      case GateType::Measure:
        os << "measure q[" << wire[g.qubits[0].index] << "] -> c[" << g.clbit << "];\n";
        break;
      case GateType::Barrier: {
        os << "barrier ";
        for (std::size_t i = 0; i < g.qubits.size(); ++i)
          os << (i ? "," : "") << "q[" << wire[g.qubits[i].index] << "]";
        os << ";\n";
        break;
      }
      case GateType::Probe:
        break;
      case GateType::Composite: {
        std::vector<std::size_t> sub(g.definition->n_qubits());
        for (std::size_t j = 0; j < g.qubits.size(); ++j) sub[j] = wire[g.qubits[j].index];
        write_gates(os, *g.definition, sub);
        break;
      }
      default: {
        os << names[static_cast<int>(g.type)];
        if (!g.params.empty()) {
          os << "(";
          for (std::size_t i = 0; i < g.params.size(); ++i) {
            std::snprintf(buf, sizeof buf, "%.17g", g.params[i]);
            os << (i ? "," : "") << buf;
          }
          os << ")";
        }
        for (std::size_t i = 0; i < g.qubits.size(); ++i)
          os << (i ? "," : " ") << "q[" << wire[g.qubits[i].index] << "]";
        os << ";\n";
      }
    }
  }
}

}  // namespace

Circuit from_qasm(const std::string& source) { return Parser(lex(source)).parse(); }

std::string to_qasm(const Circuit& circuit) {
  std::ostringstream os;
  os << "OPENQASM 2.0;\ninclude \"qelib1.inc\";\n";
  os << "qreg q[" << circuit.n_qubits() << "];\n";
  if (circuit.n_clbits()) os << "creg c[" << circuit.n_clbits() << "];\n";
  std::vector<std::size_t> wire(circuit.n_qubits());
  for (std::size_t i = 0; i < wire.size(); ++i) wire[i] = i;
  write_gates(os, circuit, wire);
  return os.str();
}

}  // namespace ket
