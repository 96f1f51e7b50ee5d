// Symbolic integer algebra (see include/loop_tool/symbolic.h).
//
// Behaviour follows src/core/symbolic.cpp of the reference:
//   simplify          symbolic.cpp:354-495   (here: polynomial normal form)
//   evaluate          symbolic.cpp:304-352   (here: int64, C division)
//   isolate/can_isolate symbolic.cpp:497-590 (here: linear isolation)
//   unify             symbolic.cpp:597-859   (here: fixed-point size solver)
//   evaluate(constraints) symbolic.cpp:861-905
//   differentiate     symbolic.cpp:907-962
//   intercept         symbolic.cpp:964-970
#include "loop_tool/symbolic.h"

#include <algorithm>
#include <atomic>
#include <limits>
#include <map>
#include <sstream>
#include <unordered_set>

namespace loop_tool {
namespace symbolic {

struct ExprNode {
  Expr::Type type = Expr::Type::value;
  Op op = Op::constant;
  int64_t val = 0;
  Symbol sym{"", -1};
  std::vector<std::shared_ptr<const ExprNode>> args;
  uint64_t hash_sym = 0;   // symbol sensitive
  uint64_t hash_shape = 0;  // symbol insensitive
  bool simplified = false;
};

namespace {

using NodeP = std::shared_ptr<const ExprNode>;

NodeP finish(std::shared_ptr<ExprNode> n) {
  uint64_t hs = hash((uint64_t)n->type * 31 + (uint64_t)n->op);
  uint64_t hi = hs;
  if (n->type == Expr::Type::value) {
    hs = hash_combine(hs, (uint64_t)n->val);
    hi = hs;
  } else if (n->type == Expr::Type::symbol) {
    hi = hash_combine(hi, 0x5eed);
    hs = hash_combine(hi, n->sym.hash());
  }
  for (const auto &a : n->args) {
    hs = hash_combine(hs, a->hash_sym);
    hi = hash_combine(hi, a->hash_shape);
  }
  n->hash_sym = hs;
  n->hash_shape = hi;
  return n;
}

NodeP make_value(int64_t v) {
  auto n = std::make_shared<ExprNode>();
  n->type = Expr::Type::value;
  n->val = v;
  n->simplified = true;
  return finish(n);
}

NodeP make_symbol(const Symbol &s) {
  auto n = std::make_shared<ExprNode>();
  n->type = Expr::Type::symbol;
  n->sym = s;
  n->simplified = true;
  return finish(n);
}

NodeP make_fn(Op op, std::vector<NodeP> args, bool simplified) {
  auto n = std::make_shared<ExprNode>();
  n->type = Expr::Type::function;
  n->op = op;
  n->args = std::move(args);
  n->simplified = simplified;
  return finish(n);
}

// Structural total order (deterministic within a process).
int cmp(const ExprNode *a, const ExprNode *b) {
  if (a == b) return 0;
  if (a->type != b->type) return a->type < b->type ? -1 : 1;
  switch (a->type) {
    case Expr::Type::value:
      return a->val == b->val ? 0 : (a->val < b->val ? -1 : 1);
    case Expr::Type::symbol:
      return a->sym.id() == b->sym.id() ? 0 : (a->sym.id() < b->sym.id() ? -1 : 1);
    default:
      break;
  }
  if (a->op != b->op) return a->op < b->op ? -1 : 1;
  if (a->args.size() != b->args.size())
    return a->args.size() < b->args.size() ? -1 : 1;
  for (size_t i = 0; i < a->args.size(); ++i) {
    int c = cmp(a->args[i].get(), b->args[i].get());
    if (c) return c;
  }
  return 0;
}

bool same(const ExprNode *a, const ExprNode *b) {
  if (a == b) return true;
  if (a->hash_sym != b->hash_sym) return false;
  return cmp(a, b) == 0;
}

// ---- polynomial normal form ------------------------------------------------
// A monomial is a sorted list of atoms (repeated atoms are powers); a
// polynomial maps monomials to non-zero integer coefficients.
using Mono = std::vector<NodeP>;

struct MonoLess {
  bool operator()(const Mono &a, const Mono &b) const {
    // higher degree first, then structural order of the atoms
    if (a.size() != b.size()) return a.size() > b.size();
    for (size_t i = 0; i < a.size(); ++i) {
      int c = cmp(a[i].get(), b[i].get());
      if (c) return c < 0;
    }
    return false;
  }
};

using Poly = std::map<Mono, int64_t, MonoLess>;

void poly_add_term(Poly &p, const Mono &m, int64_t c) {
  if (!c) return;
  auto it = p.find(m);
  if (it == p.end()) {
    p.emplace(m, c);
  } else {
    it->second += c;
    if (!it->second) p.erase(it);
  }
}

Poly poly_mul(const Poly &a, const Poly &b) {
  Poly out;
  for (const auto &ta : a) {
    for (const auto &tb : b) {
      Mono m = ta.first;
      m.insert(m.end(), tb.first.begin(), tb.first.end());
      std::sort(m.begin(), m.end(), [](const NodeP &x, const NodeP &y) {
        return cmp(x.get(), y.get()) < 0;
      });
      poly_add_term(out, m, ta.second * tb.second);
    }
  }
  return out;
}

Poly to_poly(const NodeP &n);

Poly atom_poly(const NodeP &n) {
  Poly p;
  p.emplace(Mono{n}, 1);
  return p;
}

Poly to_poly(const NodeP &n) {
  Poly p;
  switch (n->type) {
    case Expr::Type::value:
      if (n->val) p.emplace(Mono{}, n->val);
      return p;
    case Expr::Type::symbol:
      return atom_poly(n);
    default:
      break;
  }
  switch (n->op) {
    case Op::add: {
      p = to_poly(n->args[0]);
      for (const auto &t : to_poly(n->args[1])) poly_add_term(p, t.first, t.second);
      return p;
    }
    case Op::multiply:
      return poly_mul(to_poly(n->args[0]), to_poly(n->args[1]));
    case Op::negate: {
      for (const auto &t : to_poly(n->args[0])) p.emplace(t.first, -t.second);
      return p;
    }
    default:
      return atom_poly(n);
  }
}

NodeP from_poly(const Poly &p) {
  if (p.empty()) return make_value(0);
  NodeP sum;
  NodeP constant;
  for (const auto &t : p) {
    if (t.first.empty()) {
      constant = make_value(t.second);
      continue;
    }
    NodeP prod = t.first[0];
    for (size_t i = 1; i < t.first.size(); ++i)
      prod = make_fn(Op::multiply, {prod, t.first[i]}, true);
    if (t.second != 1) prod = make_fn(Op::multiply, {prod, make_value(t.second)}, true);
    sum = sum ? make_fn(Op::add, {sum, prod}, true) : prod;
  }
  if (constant) sum = sum ? make_fn(Op::add, {sum, constant}, true) : constant;
  return sum;
}

bool poly_divisible(const Poly &p, int64_t d) {
  for (const auto &t : p)
    if (t.second % d) return false;
  return true;
}

NodeP simplify_node(const NodeP &n);

NodeP simplify_node(const NodeP &n) {
  if (n->simplified) return n;
  if (n->type != Expr::Type::function) return n;
  std::vector<NodeP> a;
  for (const auto &x : n->args) a.emplace_back(simplify_node(x));
  switch (n->op) {
    case Op::add:
    case Op::multiply:
    case Op::negate: {
      auto tmp = make_fn(n->op, a, true);
      return from_poly(to_poly(tmp));
    }
    case Op::divide: {
      const auto &l = a[0];
      const auto &r = a[1];
      if (r->type == Expr::Type::value) {
        int64_t d = r->val;
        if (d == 1) return l;
        if (d == -1) return from_poly(to_poly(make_fn(Op::negate, {l}, true)));
        if (d != 0) {
          if (l->type == Expr::Type::value && l->val % d == 0)
            return make_value(l->val / d);
          auto p = to_poly(l);
          if (l->type != Expr::Type::value && poly_divisible(p, d)) {
            Poly q;
            for (const auto &t : p) q.emplace(t.first, t.second / d);
            return from_poly(q);
          }
        }
      }
      if (l->type == Expr::Type::value && l->val == 0) return l;
      return make_fn(Op::divide, a, true);
    }
    case Op::modulo: {
      const auto &l = a[0];
      const auto &r = a[1];
      if (r->type == Expr::Type::value) {
        int64_t d = r->val;
        if (d == 1 || d == -1) return make_value(0);
        if (d != 0) {
          if (l->type == Expr::Type::value) return make_value(l->val % d);
          if (poly_divisible(to_poly(l), d)) return make_value(0);
        }
      }
      return make_fn(Op::modulo, a, true);
    }
    case Op::max: {
      if (a[0]->type == Expr::Type::value && a[1]->type == Expr::Type::value)
        return make_value(std::max(a[0]->val, a[1]->val));
      if (a[0]->type == Expr::Type::value &&
          a[0]->val == std::numeric_limits<int64_t>::min())
        return a[1];
      if (same(a[0].get(), a[1].get())) return a[0];
      if (cmp(a[0].get(), a[1].get()) > 0) std::swap(a[0], a[1]);
      return make_fn(Op::max, a, true);
    }
    case Op::size: {
      if (a[0]->type == Expr::Type::value) return a[0];
      return make_fn(Op::size, a, true);
    }
    case Op::reciprocal: {
      if (a[0]->type == Expr::Type::value && (a[0]->val == 1 || a[0]->val == -1))
        return a[0];
      return make_fn(Op::reciprocal, a, true);
    }
    default:
      return make_fn(n->op, a, true);
  }
}

}  // namespace

// ---------------------------------------------------------------- Symbol ----
int32_t Symbol::getNewId() {
  static std::atomic<int32_t> counter{0};
  return counter++;
}
Symbol::operator Expr() const { return Expr(*this); }
Expr Symbol::operator+(const Symbol &rhs) const { return Expr(*this) + Expr(rhs); }
Expr Symbol::operator*(const Symbol &rhs) const { return Expr(*this) * Expr(rhs); }
Expr Symbol::operator+(const Expr &rhs) const { return Expr(*this) + rhs; }
Expr Symbol::operator*(const Expr &rhs) const { return Expr(*this) * rhs; }

// ------------------------------------------------------------------ Expr ----
Expr::Expr() : impl_(make_value(-1)) {}
Expr::Expr(int64_t v) : impl_(make_value(v)) {}
Expr::Expr(const Symbol &s) : impl_(make_symbol(s)) {}
Expr::Expr(Op op, const Expr &a) : impl_(make_fn(op, {a.impl_}, false)) {}
Expr::Expr(Op op, const Expr &a, const Expr &b)
    : impl_(make_fn(op, {a.impl_, b.impl_}, false)) {}

Expr::Type Expr::type() const { return impl_->type; }
Op Expr::op() const { return impl_->op; }
bool Expr::simplified() const { return impl_->simplified; }

int64_t Expr::value() const {
  ASSERT(type() == Type::value)
      << "attempted to get a value from a symbolic expression: " << dump();
  return impl_->val;
}

const Symbol &Expr::symbol() const {
  ASSERT(type() == Type::symbol)
      << "attempted to get a symbol from a non-symbol expression: " << dump();
  return impl_->sym;
}

std::vector<Expr> Expr::args() const {
  std::vector<Expr> out;
  for (const auto &a : impl_->args) out.emplace_back(a);
  return out;
}

Expr Expr::arg(int idx) const { return Expr(impl_->args.at(idx)); }

Expr Expr::size(const Expr &arg) {
  ASSERT(arg.type() == Type::symbol) << "size() takes a symbol, got " << arg.dump();
  return Expr(Op::size, arg);
}

Expr Expr::reciprocal() const {
  if (type() == Type::value) {
    ASSERT(value() != 0) << "cannot calculate 1/0";
  }
  return Expr(Op::reciprocal, *this);
}

uint64_t Expr::hash(bool symbol_sensitive) const {
  return symbol_sensitive ? impl_->hash_sym : impl_->hash_shape;
}

bool Expr::operator==(const Expr &rhs) const {
  return same(impl_.get(), rhs.impl_.get());
}

Expr Expr::walk(const std::function<Expr(const Expr &)> &f) const {
  if (type() == Type::function) {
    std::vector<NodeP> new_args;
    bool changed = false;
    for (const auto &a : impl_->args) {
      auto w = Expr(a).walk(f);
      changed |= (w.impl_.get() != a.get());
      new_args.emplace_back(w.impl_);
    }
    if (changed) return f(Expr(make_fn(op(), new_args, false)));
  }
  return f(*this);
}

void Expr::visit(const std::function<void(const Expr &)> &f) const {
  for (const auto &a : impl_->args) Expr(a).visit(f);
  f(*this);
}

Expr Expr::replace(Symbol A, Symbol B) const {
  return replace(A, Expr(B));
}

Expr Expr::replace(Symbol A, Expr e) const {
  return walk([&](const Expr &x) {
    if (x.type() == Type::symbol && x.symbol() == A) return e;
    return x;
  });
}

Expr Expr::replace(const Expr &target, Symbol B) const {
  return walk([&](const Expr &x) {
    if (x == target) return Expr(B);
    return x;
  });
}

Expr Expr::replace(const Expr &target, int64_t c) const {
  return walk([&](const Expr &x) {
    if (x == target) return Expr(c);
    return x;
  });
}

Expr Expr::replace(Symbol A, int64_t c) const { return replace(A, Expr(c)); }

bool Expr::contains(const Symbol &s) const {
  const auto *n = impl_.get();
  if (n->type == Type::symbol) return n->sym == s;
  for (const auto &a : n->args)
    if (Expr(a).contains(s)) return true;
  return false;
}

std::vector<Symbol> Expr::symbols(bool include_sized) const {
  std::vector<Symbol> out;
  std::unordered_set<int32_t> seen;
  std::function<void(const ExprNode *)> rec = [&](const ExprNode *n) {
    if (!include_sized && n->type == Type::function && n->op == Op::size) return;
    for (const auto &a : n->args) rec(a.get());
    if (n->type == Type::symbol && !seen.count(n->sym.id())) {
      seen.insert(n->sym.id());
      out.emplace_back(n->sym);
    }
  };
  rec(impl_.get());
  return out;
}

std::string Expr::dump(
    bool short_form,
    const std::unordered_map<Symbol, std::string, Hash<Symbol>> &repl) const {
  std::stringstream ss;
  const auto *n = impl_.get();
  auto sub = [&](const NodeP &a) {
    Expr e(a);
    bool atomic = e.type() != Type::function || e.op() == Op::size ||
                  e.op() == Op::max || e.op() == Op::reciprocal;
    if (atomic) return e.dump(short_form, repl);
    return "(" + e.dump(short_form, repl) + ")";
  };
  if (n->type == Type::value) {
    ss << n->val;
  } else if (n->type == Type::symbol) {
    auto it = repl.find(n->sym);
    if (it != repl.end()) {
      ss << it->second;
    } else {
      ss << n->sym.name();
      if (!short_form) ss << "[id:" << n->sym.id() << "]";
    }
  } else {
    switch (n->op) {
      case Op::size:
        ss << "|" << Expr(n->args[0]).dump(short_form, repl) << "|";
        break;
      case Op::max:
        ss << "max(" << Expr(n->args[0]).dump(short_form, repl) << ", "
           << Expr(n->args[1]).dump(short_form, repl) << ")";
        break;
      case Op::negate:
        ss << "-" << sub(n->args[0]);
        break;
      case Op::reciprocal:
        ss << "(1 / " << Expr(n->args[0]).dump(short_form, repl) << ")";
        break;
      case Op::add:
        ss << sub(n->args[0]) << "+" << sub(n->args[1]);
        break;
      case Op::multiply:
        ss << sub(n->args[0]) << "*" << sub(n->args[1]);
        break;
      case Op::divide:
        ss << sub(n->args[0]) << "/" << sub(n->args[1]);
        break;
      case Op::modulo:
        ss << sub(n->args[0]) << "%" << sub(n->args[1]);
        break;
      default:
        ASSERT(0) << "cannot print op " << (int)n->op;
    }
  }
  return ss.str();
}

Expr Expr::simplify() const { return Expr(simplify_node(impl_)); }

bool Expr::can_evaluate() const {
  if (type() == Type::function && op() == Op::max)
    return arg(0).can_evaluate() || arg(1).can_evaluate();
  bool can = true;
  visit([&](const Expr &e) {
    if (e.type() == Type::symbol) can = false;
  });
  return can;
}

int64_t Expr::evaluate() const {
  ASSERT(can_evaluate()) << "cannot evaluate " << dump();
  const auto *n = impl_.get();
  if (n->type == Type::value) return n->val;
  auto ev = [&](int i) { return Expr(n->args[i]).evaluate(); };
  switch (n->op) {
    case Op::add:
      return ev(0) + ev(1);
    case Op::multiply:
      return ev(0) * ev(1);
    case Op::divide: {
      auto d = ev(1);
      ASSERT(d != 0) << "division by zero evaluating " << dump();
      return ev(0) / d;
    }
    case Op::modulo: {
      auto d = ev(1);
      ASSERT(d != 0) << "modulo by zero evaluating " << dump();
      return ev(0) % d;
    }
    case Op::max: {
      bool l = arg(0).can_evaluate(), r = arg(1).can_evaluate();
      if (l && !r) return ev(0);
      if (r && !l) return ev(1);
      return std::max(ev(0), ev(1));
    }
    case Op::negate:
      return -ev(0);
    case Op::reciprocal: {
      auto d = ev(0);
      ASSERT(d != 0) << "reciprocal of zero";
      return 1 / d;
    }
    case Op::size:
      return ev(0);
    default:
      ASSERT(0) << "cannot evaluate " << dump();
  }
  return 0;
}

// ------------------------------------------------------------ isolation ----
namespace {

// Linear decomposition of `e` (simplified) in `sym`: e = coef * sym + rest,
// with neither coef nor rest containing sym.  Fails when sym appears inside
// a non-polynomial atom or with degree > 1.
bool linear_in(const Expr &e, const Symbol &sym, Expr &coef, Expr &rest) {
  auto p = to_poly(e.simplify().impl_);
  Poly c, r;
  for (const auto &t : p) {
    int count = 0;
    Mono other;
    for (const auto &atom : t.first) {
      if (atom->type == Expr::Type::symbol && atom->sym == sym) {
        count++;
        continue;
      }
      if (Expr(atom).contains(sym)) return false;
      other.emplace_back(atom);
    }
    if (count > 1) return false;
    if (count == 1)
      poly_add_term(c, other, t.second);
    else
      poly_add_term(r, t.first, t.second);
  }
  if (c.empty()) return false;
  coef = Expr(from_poly(c));
  rest = Expr(from_poly(r));
  return true;
}

}  // namespace

bool can_isolate(const Constraint &c, const Symbol &sym) {
  bool l = c.first.contains(sym), r = c.second.contains(sym);
  if (l == r) return false;
  Expr coef, rest;
  return linear_in(l ? c.first : c.second, sym, coef, rest);
}

Constraint isolate(const Constraint &c, const Symbol &sym) {
  bool l = c.first.contains(sym), r = c.second.contains(sym);
  ASSERT(l != r) << "cannot isolate " << sym.name()
                 << " appearing on both or neither side of " << c.first.dump()
                 << " = " << c.second.dump();
  const auto &with = l ? c.first : c.second;
  const auto &without = l ? c.second : c.first;
  if (with == Expr(sym)) return {Expr(sym), without};
  Expr coef, rest;
  ASSERT(linear_in(with, sym, coef, rest))
      << "cannot isolate " << sym.name() << " in " << with.dump();
  // coef*sym + rest = without  ->  sym = (without - rest) / coef
  Expr rhs = (without - rest).simplify();
  if (coef.type() == Expr::Type::value && coef.value() == 1) return {Expr(sym), rhs};
  if (coef.type() == Expr::Type::value && coef.value() == -1)
    return {Expr(sym), (-rhs).simplify()};
  return {Expr(sym), (rhs / coef).simplify()};
}

// ---------------------------------------------------------------- unify ----
// Fixed-point solver over sizes.  Every symbol gets |s| from
//   1. size equations (constraints without plain symbols), which win, else
//   2. index equations s = f(y..): |s| = f(|y|-1 ..) + 1 (largest index + 1),
// and every isolatable symbol of every index equation is reported as
// `s = rhs`.  Conflicting numeric sizes are an error, as upstream.
std::vector<Constraint> unify(std::vector<Constraint> constraints_in) {
  // dedupe
  std::vector<Constraint> constraints;
  for (const auto &c : constraints_in) {
    Constraint s{c.first.simplify(), c.second.simplify()};
    bool dup = false;
    for (const auto &o : constraints)
      if (o.first == s.first && o.second == s.second) dup = true;
    if (!dup) constraints.emplace_back(s);
  }

  std::vector<Symbol> all_syms;
  std::unordered_set<int32_t> seen;
  for (const auto &c : constraints) {
    for (const auto &e : {c.first, c.second})
      for (const auto &s : e.symbols())
        if (!seen.count(s.id())) {
          seen.insert(s.id());
          all_syms.emplace_back(s);
        }
  }

  // Size symbols stand for |s| during solving.
  std::unordered_map<int32_t, Symbol> size_sym;
  std::unordered_map<int32_t, Symbol> size_owner;  // size sym id -> sym
  for (const auto &s : all_syms) {
    Symbol ss(s.name() + "_size");
    size_sym.emplace(s.id(), ss);
    size_owner.emplace(ss.id(), s);
  }
  auto to_size_syms = [&](const Expr &e) {
    return e
        .walk([&](const Expr &x) {
          if (x.type() == Expr::Type::function && x.op() == Op::size &&
              x.arg(0).type() == Expr::Type::symbol &&
              size_sym.count(x.arg(0).symbol().id()))
            return Expr(size_sym.at(x.arg(0).symbol().id()));
          return x;
        })
        .simplify();
  };
  auto from_size_syms = [&](const Expr &e) {
    return e
        .walk([&](const Expr &x) {
          if (x.type() == Expr::Type::symbol && size_owner.count(x.symbol().id()))
            return Expr::size(Expr(size_owner.at(x.symbol().id())));
          return x;
        })
        .simplify();
  };

  std::vector<Constraint> mapped;
  for (const auto &c : constraints)
    mapped.emplace_back(to_size_syms(c.first), to_size_syms(c.second));

  auto has_plain = [&](const Constraint &c) {
    for (const auto &s : all_syms)
      if (c.first.contains(s) || c.second.contains(s)) return true;
    return false;
  };

  // candidate size expressions (in size symbols) per symbol
  std::map<int32_t, std::vector<Expr>> size_eq;
  for (const auto &c : mapped) {
    if (has_plain(c)) continue;
    for (const auto &s : all_syms) {
      const auto &ss = size_sym.at(s.id());
      if (!can_isolate(c, ss)) continue;
      auto e = isolate(c, ss).second;
      auto &v = size_eq[s.id()];
      bool dup = false;
      for (const auto &o : v) dup |= (o == e);
      if (!dup) v.emplace_back(e);
    }
  }

  // isolated index equations
  std::map<int32_t, std::vector<Expr>> index_eq;
  for (const auto &c : mapped) {
    if (!has_plain(c)) continue;
    for (const auto &s : all_syms) {
      if (!can_isolate(c, s)) continue;
      auto e = isolate(c, s).second;
      auto &v = index_eq[s.id()];
      bool dup = false;
      for (const auto &o : v) dup |= (o == e);
      if (!dup) v.emplace_back(e);
    }
  }

  std::unordered_map<int32_t, int64_t> known;
  auto substitute = [&](Expr e) {
    for (const auto &p : known) e = e.replace(size_sym.at(p.first), p.second);
    return e.simplify();
  };
  auto is_value = [](const Expr &e) { return e.type() == Expr::Type::value; };

  for (int iter = 0; iter < 1000; ++iter) {
    bool changed = false;
    for (const auto &s : all_syms) {
      if (!size_eq.count(s.id())) continue;
      for (const auto &e : size_eq.at(s.id())) {
        auto v = substitute(e);
        if (!is_value(v)) continue;
        if (known.count(s.id())) {
          ASSERT(known.at(s.id()) == v.value())
              << "size of " << s.name()
              << " set multiple times to different values (new value: "
              << v.value() << " old: " << known.at(s.id()) << ")";
          continue;
        }
        known[s.id()] = v.value();
        changed = true;
      }
    }
    for (const auto &s : all_syms) {
      if (size_eq.count(s.id()) || known.count(s.id())) continue;
      if (!index_eq.count(s.id())) continue;
      bool any = false;
      int64_t best = 0;
      for (const auto &e : index_eq.at(s.id())) {
        // largest index: each symbol y at |y|-1
        auto m = e.walk([&](const Expr &x) {
          if (x.type() == Expr::Type::symbol && size_sym.count(x.symbol().id()))
            return Expr(size_sym.at(x.symbol().id())) - Expr(1);
          return x;
        });
        auto v = substitute(m + Expr(1));
        if (!is_value(v)) continue;
        best = any ? std::max(best, v.value()) : v.value();
        any = true;
      }
      if (any) {
        known[s.id()] = best;
        changed = true;
      }
    }
    if (!changed) break;
  }

  std::vector<Constraint> out;
  for (const auto &s : all_syms) {
    if (known.count(s.id())) {
      out.emplace_back(Expr::size(Expr(s)), Expr(known.at(s.id())));
      continue;
    }
    std::vector<Expr> cands;
    if (size_eq.count(s.id())) {
      for (const auto &e : size_eq.at(s.id())) cands.emplace_back(substitute(e));
    } else if (index_eq.count(s.id())) {
      const auto &e = index_eq.at(s.id()).front();
      auto m = e.walk([&](const Expr &x) {
        if (x.type() == Expr::Type::symbol && size_sym.count(x.symbol().id()))
          return Expr(size_sym.at(x.symbol().id())) - Expr(1);
        return x;
      });
      cands.emplace_back(substitute(m + Expr(1)));
    }
    if (cands.empty()) continue;
    Expr e = cands[0];
    for (size_t i = 1; i < cands.size(); ++i) e = Expr::max(e, cands[i]);
    out.emplace_back(Expr::size(Expr(s)), from_size_syms(e));
  }
  for (const auto &s : all_syms) {
    if (!index_eq.count(s.id())) continue;
    for (const auto &e : index_eq.at(s.id()))
      out.emplace_back(Expr(s), from_size_syms(substitute(e)));
  }
  return out;
}

std::vector<Constraint> evaluate(const std::vector<Constraint> &old) {
  std::vector<Constraint> constraints = old;
  std::unordered_map<int32_t, int64_t> sizes;
  for (int iter = 0; iter < 1000; ++iter) {
    bool changed = false;
    for (const auto &c : constraints) {
      if (c.first.type() == Expr::Type::function && c.first.op() == Op::size &&
          c.second.can_evaluate()) {
        const auto &sym = c.first.arg(0).symbol();
        if (sizes.count(sym.id())) continue;
        sizes[sym.id()] = c.second.evaluate();
        changed = true;
      }
    }
    if (!changed) break;
    for (auto &c : constraints) {
      auto e = c.second;
      for (const auto &s : e.symbols()) {
        if (sizes.count(s.id()))
          e = e.replace(Expr::size(Expr(s)), sizes.at(s.id()));
      }
      c.second = e.simplify();
    }
  }
  return constraints;
}

Expr differentiate(Expr e, Symbol sym) {
  if (!e.contains(sym)) return Expr(0);
  if (e.type() == Expr::Type::symbol) return Expr(1);
  if (e.op() == Op::size) return Expr(0);
  const auto a = e.arg(0);
  switch (e.op()) {
    case Op::add:
      return (differentiate(a, sym) + differentiate(e.arg(1), sym)).simplify();
    case Op::multiply: {
      const auto b = e.arg(1);
      return (differentiate(a, sym) * b + a * differentiate(b, sym)).simplify();
    }
    case Op::divide: {
      const auto b = e.arg(1);
      if (!b.contains(sym)) return (differentiate(a, sym) / b).simplify();
      return ((differentiate(a, sym) * b - a * differentiate(b, sym)) / (b * b))
          .simplify();
    }
    case Op::negate:
      return (-differentiate(a, sym)).simplify();
    case Op::reciprocal:
      return (-(differentiate(a, sym) / (a * a))).simplify();
    default:
      ASSERT(0) << "Cannot differentiate " << e.dump() << " with respect to "
                << sym.name();
  }
  return Expr(0);
}

Expr intercept(Expr e) {
  for (const auto &s : e.symbols()) e = e.replace(s, Expr(0)).simplify();
  return e.simplify();
}

}  // namespace symbolic
}  // namespace loop_tool
