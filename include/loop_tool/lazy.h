/*
Copyright (c) Facebook, Inc. and its affiliates.

This source code is licensed under the MIT license found in the
LICENSE file in the root directory of this source tree.

The declarations below restate the reference loop_tool's public interface
(include/loop_tool/) so that code written against it compiles unchanged.
*/
// Lazy symbolic tensor frontend of loop_tool (`lz::Tensor`, `Symbol`,
// `.as/.to/.sum`).  A Tensor records an op graph over named symbols; data()
// unifies sizes, lowers the graph to IR with the default schedule, builds the
// LoopTree, compiles it with the default backend (cached by a shape-structural
// hash) and runs it.
//
// Interface kept from include/loop_tool/lazy.h:19-688 of the reference so that
// programs written against it drive this build unchanged.  Most method bodies
// live in lazy.cpp; the compile/lower caches are additionally guarded by a
// mutex so that several serving threads may use the frontend.
#pragma once

#include <memory>
#include <mutex>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "backend.h"
#include "compile.h"
#include "hardware.h"
#include "ir.h"
#include "symbolic.h"

namespace detail {

template <typename... Conds>
struct and_ : std::true_type {};
template <typename Cond, typename... Conds>
struct and_<Cond, Conds...>
    : std::conditional<Cond::value, and_<Conds...>, std::false_type>::type {};
template <typename Target, typename... Ts>
using areT = and_<std::is_same<Ts, Target>...>;

}  // namespace detail

namespace loop_tool {
namespace lazy {

using Symbol = loop_tool::symbolic::Symbol;
using Expr = loop_tool::symbolic::Expr;
using Constraint = loop_tool::symbolic::Constraint;

template <typename... Args>
std::vector<Expr> Index(const Args &...args) {
  return std::vector<Expr>{Expr(args)...};
}

struct CachedLowered {
  IR ir;
  LoopTree loop_tree;
  int64_t output_size;
  std::vector<int64_t> sizes;
};

std::unordered_map<size_t, std::shared_ptr<Compiled>> &getCompilationCache();
std::unordered_map<size_t, CachedLowered> &getLoweredCache();
// Guards both caches above for multi-threaded callers.
std::recursive_mutex &getCacheMutex();

struct TensorImpl {
  uint64_t hash_ = 0;
  bool unified_ = false;
  Operation op_ = Operation::constant;
  mutable Memory memory_;
  mutable bool owning_ = true;
  mutable bool force_recompute_ = false;
  std::vector<Symbol> shape_;
  mutable std::vector<int64_t> sizes_;
  std::vector<Constraint> constraints_;
  std::vector<std::shared_ptr<TensorImpl>> deps_;

  TensorImpl(void *data, std::vector<int64_t> sizes);
  explicit TensorImpl(std::vector<Symbol> shape);
  TensorImpl(Operation op, std::vector<Symbol> shape,
             std::vector<std::shared_ptr<TensorImpl>> deps,
             std::vector<Constraint> constraints = {});
  TensorImpl(const TensorImpl &other);
  ~TensorImpl();

  std::unordered_map<int, Expr> size_constraints() const;
  std::vector<Constraint> constraints() const { return constraints_; }
  void updateHash(bool force_unify = false);
  uint64_t hash() const { return hash_; }
  bool owning() const { return owning_; }
  void bind(void *data, std::vector<int64_t> sizes);

  const std::vector<Symbol> &shape() const { return shape_; }
  int64_t size(int dim) const;
  std::vector<int64_t> sizes() const;

  void collectSymbolMap(std::unordered_map<int, Symbol> &symbol_map,
                        std::unordered_set<TensorImpl *> &seen);
  void propagateSymbolMap(const std::unordered_map<int, Symbol> &symbol_map,
                          std::unordered_set<TensorImpl *> &seen);
  void unifySymbols();
  void collectConstraints(std::vector<Constraint> &constraints,
                          std::unordered_set<TensorImpl *> &seen,
                          std::unordered_set<int64_t> &seen_constraint_hashes);
  void propagateConstraints(const std::vector<Constraint> &constraints,
                            std::unordered_set<TensorImpl *> &seen);
  void unifyConstraints();
  void unify(bool force = false);
  void populateCompilationCache();
  void populateLoweredCache();

  std::vector<void *> getInputBuffers(
      std::unordered_set<const TensorImpl *> &seen) const;

  void force_recompute() const { force_recompute_ = true; }
  void clear_cache() const;

  // Compiles (on a cache miss) and runs the graph; returns the output buffer.
  void *data_ptr() const;
  template <typename T>
  T *data() const {
    return static_cast<T *>(data_ptr());
  }

  std::pair<IR, std::unordered_map<int, std::pair<IR::VarRef, int64_t>>> lower()
      const;
  IR ir() const;
  LoopTree loop_tree() const;
  void compile();
  std::string code() const { return compiled()->dump(); }
  void set(const IR &ir);
  void set(const LoopTree &loop_tree);
  std::shared_ptr<Compiled> compiled() const;

  LoopTree schedule(
      IR &ir,
      const std::unordered_map<int, std::pair<IR::VarRef, int64_t>> &var_map) const;
  IR::NodeRef resolve(
      IR &ir, std::unordered_map<int, std::pair<IR::VarRef, int64_t>> &var_map,
      std::unordered_map<const TensorImpl *, IR::NodeRef> &impl_map) const;
};

struct Tensor {
  std::shared_ptr<TensorImpl> impl_;
  Tensor(std::shared_ptr<TensorImpl> impl) : impl_(std::move(impl)) {}
  Tensor(std::vector<Symbol> shape) : impl_(std::make_shared<TensorImpl>(shape)) {}

  template <typename... Symbols,
            std::enable_if_t<::detail::areT<Symbol, Symbols...>::value, int> = 0>
  Tensor(Symbols... symbols)
      : impl_(std::make_shared<TensorImpl>(std::vector<Symbol>{symbols...})) {}

  template <typename... Sizes,
            std::enable_if_t<::detail::areT<int, Sizes...>::value, int> = 0>
  Tensor(Sizes... sizes)
      : impl_(std::make_shared<TensorImpl>(
            nullptr, std::vector<int64_t>{static_cast<int64_t>(sizes)...})) {}

  template <typename... Sizes,
            std::enable_if_t<::detail::areT<int64_t, Sizes...>::value, int> = 0>
  Tensor(Sizes... sizes)
      : impl_(std::make_shared<TensorImpl>(nullptr, std::vector<int64_t>{sizes...})) {}

  template <typename... Sizes,
            std::enable_if_t<::detail::areT<size_t, Sizes...>::value, int> = 0>
  Tensor(Sizes... sizes)
      : impl_(std::make_shared<TensorImpl>(
            nullptr, std::vector<int64_t>{static_cast<int64_t>(sizes)...})) {}

  Tensor(void *data, std::vector<int64_t> sizes)
      : impl_(std::make_shared<TensorImpl>(data, sizes)) {}
  Tensor(std::vector<int64_t> sizes)
      : impl_(std::make_shared<TensorImpl>(nullptr, sizes)) {}
  void bind(void *data, std::vector<int64_t> sizes) { impl()->bind(data, sizes); }

  std::vector<Symbol> broadcast_shape(const Tensor &rhs) const;

  Tensor operator*(const Tensor &rhs) const { return binary(Operation::multiply, rhs); }
  Tensor operator/(const Tensor &rhs) const { return binary(Operation::divide, rhs); }
  Tensor operator+(const Tensor &rhs) const { return binary(Operation::add, rhs); }
  Tensor operator-(const Tensor &rhs) const { return binary(Operation::subtract, rhs); }
  Tensor max(const Tensor &rhs) const { return binary(Operation::max, rhs); }
  Tensor min(const Tensor &rhs) const { return binary(Operation::min, rhs); }
  // concatenation along every differing dimension
  Tensor operator|(const Tensor &rhs) const;

  Tensor pad(Symbol padded_dim, int64_t pre, int64_t post) const;
  Tensor pad(Symbol padded_dim, int64_t amount) const {
    return pad(padded_dim, amount, amount);
  }
  Tensor transpose(std::vector<int> order) const;
  Tensor transpose(std::vector<Symbol> new_shape) const;
  Tensor flatten(std::vector<Symbol> symbols, Symbol new_symbol) const;

  Tensor reduce(Operation op, const std::vector<Symbol> &reduction_vars) const;
  Tensor sum(const std::vector<Symbol> &v) const { return reduce(Operation::add, v); }
  Tensor prod(const std::vector<Symbol> &v) const {
    return reduce(Operation::multiply, v);
  }
  Tensor max(const std::vector<Symbol> &v) const { return reduce(Operation::max, v); }
  Tensor min(const std::vector<Symbol> &v) const { return reduce(Operation::min, v); }
  template <typename... Args>
  Tensor sum(const Args &...args) const {
    return sum(std::vector<Symbol>{args...});
  }

  Tensor exp() const { return unary(Operation::exp); }
  Tensor log() const { return unary(Operation::log); }
  Tensor sqrt() const { return unary(Operation::sqrt); }
  Tensor abs() const { return unary(Operation::abs); }
  Tensor operator-() const { return unary(Operation::negate); }
  Tensor reciprocal() const { return unary(Operation::reciprocal); }

  Tensor as(std::vector<Symbol> shape) const {
    return Tensor(std::make_shared<TensorImpl>(
        Operation::name, shape, std::vector<std::shared_ptr<TensorImpl>>{impl_}));
  }
  template <typename... Args>
  Tensor as(const Args &...args) const {
    return as(std::vector<Symbol>{args...});
  }

  Tensor to(std::vector<Symbol> shape, std::vector<Constraint> constraints) const {
    return Tensor(std::make_shared<TensorImpl>(
        Operation::view, shape, std::vector<std::shared_ptr<TensorImpl>>{impl_},
        constraints));
  }
  template <typename... Constraints>
  Tensor to(std::vector<Symbol> shape, const Constraints &...args) const {
    return to(shape, std::vector<Constraint>{args...});
  }

  std::shared_ptr<TensorImpl> impl() const { return impl_; }
  bool owning() const { return impl_->owning(); }
  std::vector<Symbol> shape() const { return impl_->shape(); }
  size_t size(int dim) const { return impl_->size(dim); }
  std::vector<int64_t> sizes() const { return impl_->sizes(); }
  size_t numel() const {
    size_t total = 1;
    for (auto s : sizes()) total *= s;
    return total;
  }
  LoopTree loop_tree() const { return impl_->loop_tree(); }
  IR ir() const { return impl_->ir(); }
  std::string code() const { return impl_->code(); }
  void set(const IR &ir) { impl_->set(ir); }
  void set(const LoopTree &loop_tree) { impl_->set(loop_tree); }
  void force_recompute() const { impl()->force_recompute(); }
  void clear_cache() const { impl()->clear_cache(); }
  size_t hash() const { return impl()->hash(); }

  template <typename T>
  T *data() const {
    return impl()->template data<T>();
  }

  void unify(bool force = false) const { impl_->unify(force); }
  void compile() const { impl_->compile(); }
  std::shared_ptr<Compiled> compiled() const { return impl_->compiled(); }
  bool has_deps() const { return !impl_->deps_.empty(); }

 private:
  Tensor binary(Operation op, const Tensor &rhs) const {
    return Tensor(std::make_shared<TensorImpl>(
        op, broadcast_shape(rhs),
        std::vector<std::shared_ptr<TensorImpl>>{impl_, rhs.impl()}));
  }
  Tensor unary(Operation op) const {
    return Tensor(std::make_shared<TensorImpl>(
        op, impl_->shape(), std::vector<std::shared_ptr<TensorImpl>>{impl_}));
  }
};

}  // namespace lazy
}  // namespace loop_tool
