/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Symbolic integer expressions over named symbols: the index and size algebra
// that the lazy frontend, the IR views and the Compiler analysis share.
//
// Public surface follows include/loop_tool/symbolic.h:19-249 of the reference
// (Symbol, Expr, Op, Constraint, unify, isolate, evaluate, differentiate,
// intercept).  The implementation is this project's own: expressions are kept
// in a polynomial normal form by simplify() (sum of integer-coefficient
// monomials over atoms), which makes equality, differentiation and intercepts
// exact for the quasi-affine index equations that views produce.
//
// Deliberate differences from the reference, kept small and documented:
//  * Expr::evaluate() returns int64_t (the reference returns float and loses
//    integers above 2^24, symbolic.cpp:304); division and modulo evaluate with
//    C semantics, as the generated code does at run time.
//  * differentiate() treats |x| (Op::size) as a constant instead of failing.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "error.h"

namespace loop_tool {
namespace symbolic {

// 64-bit finalizer (SplitMix64's published mixing steps) used for every hash
// in the frontend; values are only ever compared within one process.
inline uint64_t hash(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

inline uint64_t hash_combine(uint64_t a, uint64_t b) {
  return hash(a * 0x100000001b3ULL ^ (b + 0x7f4a7c159e3779b9ULL + (a << 6)));
}

template <typename T>
struct Hash {
  std::size_t operator()(const T &k) const { return k.hash(); }
};

enum struct Op {
  constant = 0,
  negate,
  reciprocal,
  size,
  max,
  add,
  multiply,
  divide,
  modulo
};

struct Expr;
struct ExprNode;

struct Symbol {
  std::string name_;
  int32_t id_ = -1;
  Symbol() : name_("X"), id_(getNewId()) {}
  explicit Symbol(std::string name) : name_(std::move(name)), id_(getNewId()) {}
  // Rebuilds a symbol with a known id (deserialization); takes no new id.
  Symbol(std::string name, int32_t id) : name_(std::move(name)), id_(id) {}
  Symbol(const Symbol &) = default;
  Symbol &operator=(const Symbol &) = default;

  static int32_t getNewId();
  int32_t id() const { return id_; }
  size_t hash() const { return symbolic::hash((uint64_t)(uint32_t)id_); }
  bool operator==(const Symbol &s) const { return s.id_ == id_; }
  bool operator!=(const Symbol &s) const { return s.id_ != id_; }
  std::string name() const { return name_; }

  operator Expr() const;
  Expr operator+(const Symbol &rhs) const;
  Expr operator*(const Symbol &rhs) const;
  Expr operator+(const Expr &rhs) const;
  Expr operator*(const Expr &rhs) const;
};

struct Expr {
  enum class Type { value, symbol, function };
  std::shared_ptr<const ExprNode> impl_;

  // The default expression is the value -1 (an "unset" marker, as upstream).
  Expr();
  explicit Expr(std::shared_ptr<const ExprNode> impl) : impl_(std::move(impl)) {}
  explicit Expr(int64_t v);
  explicit Expr(int v) : Expr((int64_t)v) {}
  explicit Expr(long long v) : Expr((int64_t)v) {}
  explicit Expr(unsigned long v) : Expr((int64_t)v) {}
  explicit Expr(unsigned v) : Expr((int64_t)v) {}
  explicit Expr(const Symbol &s);
  Expr(Op op, const Expr &a);
  Expr(Op op, const Expr &a, const Expr &b);

  Type type() const;
  Op op() const;
  int64_t value() const;
  const Symbol &symbol() const;
  std::vector<Expr> args() const;
  Expr arg(int idx) const;
  bool simplified() const;

  Expr walk(const std::function<Expr(const Expr &)> &f) const;
  void visit(const std::function<void(const Expr &)> &f) const;
  Expr replace(Symbol A, Symbol B) const;
  Expr replace(Symbol A, Expr e) const;
  Expr replace(const Expr &e, Symbol B) const;
  Expr replace(const Expr &e, int64_t c) const;
  Expr replace(Symbol A, int64_t c) const;

  std::string dump(bool short_form = false,
                   const std::unordered_map<Symbol, std::string, Hash<Symbol>>
                       &replacements = {}) const;

  // symbol_sensitive=false hashes every symbol alike (structure only).
  uint64_t hash(bool symbol_sensitive = false) const;
  bool contains(const Symbol &s) const;
  std::vector<Symbol> symbols(bool include_sized = true) const;

  Expr operator+(const Expr &rhs) const { return Expr(Op::add, *this, rhs); }
  Expr operator*(const Expr &rhs) const {
    return Expr(Op::multiply, *this, rhs);
  }
  Expr operator-() const { return Expr(Op::negate, *this); }
  Expr operator-(const Expr &rhs) const { return *this + (-rhs); }
  Expr operator/(const Expr &rhs) const { return Expr(Op::divide, *this, rhs); }
  Expr operator%(const Expr &rhs) const { return Expr(Op::modulo, *this, rhs); }
  static Expr size(const Expr &arg);
  static Expr max(const Expr &lhs, const Expr &rhs) {
    return Expr(Op::max, lhs, rhs);
  }
  Expr reciprocal() const;

  bool operator==(const Expr &rhs) const;
  bool operator!=(const Expr &rhs) const { return !(*this == rhs); }

  Expr simplify() const;
  bool can_evaluate() const;
  int64_t evaluate() const;
};

// A constraint is `lhs = rhs`; the lhs is either Expr(Symbol) (an index
// equation) or Expr::size(Symbol) (a size equation).
using Constraint = std::pair<Expr, Expr>;

// Derive |s| and the isolated index equations for every symbol mentioned.
std::vector<Constraint> unify(std::vector<Constraint> constraints);
bool can_isolate(const Constraint &c, const Symbol &sym);
Constraint isolate(const Constraint &c, const Symbol &sym);

// Substitute every size that evaluates to a number into the other rhs.
std::vector<Constraint> evaluate(const std::vector<Constraint> &constraints);

Expr differentiate(Expr, Symbol);
// zero out every symbol
Expr intercept(Expr);

}  // namespace symbolic
}  // namespace loop_tool
