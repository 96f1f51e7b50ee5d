// Lazy frontend (see include/loop_tool/lazy.h).
// Reference behaviour: include/loop_tool/lazy.h:82-327 and
// src/frontends/lazy.cpp (bind :23-70, input buffers :72-88, default schedule
// :90-126, resolve :154-273, unification :275-412, caches :414-438).
#include "loop_tool/lazy.h"

#include <algorithm>
#include <sstream>

namespace loop_tool {
namespace lazy {

using symbolic::Op;

std::unordered_map<size_t, std::shared_ptr<Compiled>> &getCompilationCache() {
  static std::unordered_map<size_t, std::shared_ptr<Compiled>> cache;
  return cache;
}

std::unordered_map<size_t, CachedLowered> &getLoweredCache() {
  static std::unordered_map<size_t, CachedLowered> cache;
  return cache;
}

std::recursive_mutex &getCacheMutex() {
  static std::recursive_mutex m;
  return m;
}

namespace {
uint64_t constraint_hash(const Constraint &c) {
  return symbolic::hash_combine(c.first.hash(true), c.second.hash(true));
}
bool is_symbol_size(const Expr &e) {
  return e.type() == Expr::Type::function && e.op() == Op::size &&
         e.arg(0).type() == Expr::Type::symbol;
}
}  // namespace

// ------------------------------------------------------------ TensorImpl ----
TensorImpl::TensorImpl(void *data, std::vector<int64_t> sizes)
    : op_(Operation::constant) {
  bind(data, sizes);
  updateHash();
}

TensorImpl::TensorImpl(std::vector<Symbol> shape)
    : op_(Operation::constant), shape_(std::move(shape)) {
  updateHash();
}

TensorImpl::TensorImpl(Operation op, std::vector<Symbol> shape,
                       std::vector<std::shared_ptr<TensorImpl>> deps,
                       std::vector<Constraint> constraints)
    : op_(op),
      shape_(std::move(shape)),
      constraints_(constraints),
      deps_(std::move(deps)) {
  updateHash();
  if (op != Operation::view) return;
  // A view may introduce output symbols that only appear on the right of its
  // constraints; record every such symbol solved for, so sizes can flow both
  // ways during unification.
  std::unordered_set<int> known;
  for (const auto &c : constraints)
    for (const auto &s : c.first.symbols()) known.insert(s.id());
  for (const auto &d : deps_)
    for (const auto &s : d->shape()) known.insert(s.id());
  for (const auto &sym : shape_) {
    if (known.count(sym.id())) continue;
    for (const auto &c : constraints) {
      for (const auto &s : c.second.symbols()) {
        if (known.count(s.id())) continue;
        if (symbolic::can_isolate(c, s)) constraints_.emplace_back(symbolic::isolate(c, s));
      }
    }
  }
}

TensorImpl::TensorImpl(const TensorImpl &o)
    : hash_(o.hash_),
      unified_(o.unified_),
      op_(o.op_),
      memory_(o.memory_),
      owning_(o.owning_),
      force_recompute_(o.force_recompute_),
      shape_(o.shape_),
      sizes_(o.sizes_),
      constraints_(o.constraints_),
      deps_(o.deps_) {
  // A copy never takes over ownership of the original's buffer.
  if (o.owning_ && o.memory_.address) {
    if (o.deps_.empty()) {
      owning_ = false;  // constant data is shared
    } else {
      memory_ = Memory{};  // computed data is recomputed on demand
    }
  }
}

// An owned buffer goes back to the hardware that allocated it: the highest
// id in its compatibility mask (CPU memory is 0x1; a device's host-visible
// memory carries 0x1 | 1 << id).  The reference frees through whatever
// hardware is the default at destruction (lazy.h:120-123), which hands such
// memory to the wrong allocator once the default has changed.
TensorImpl::~TensorImpl() {
  if (!owning_ || !memory_.address) return;
  int id = 0;
  for (int b = 0; b < 32; ++b)
    if (memory_.compatible & (1 << b)) id = b;
  for (const auto &hw : getHardware())
    if (hw->id() == id) {
      hw->free(memory_);
      return;
    }
  getDefaultHardware()->free(memory_);
}

std::unordered_map<int, Expr> TensorImpl::size_constraints() const {
  std::unordered_map<int, Expr> out;
  for (const auto &c : constraints_)
    if (is_symbol_size(c.first)) out.emplace(c.first.arg(0).symbol().id(), c.second);
  return out;
}

void TensorImpl::updateHash(bool force_unify) {
  if (force_unify) unify(true);
  uint64_t h = symbolic::hash((uint64_t)op_);
  h = symbolic::hash_combine(h, symbolic::hash(shape_.size()));
  for (const auto &c : constraints_) {
    h = symbolic::hash_combine(h, c.first.hash());
    h = symbolic::hash_combine(h, c.second.hash());
  }
  // symbols are hashed by first-appearance position, not identity
  std::unordered_map<int, uint64_t> numbering;
  auto hash_sym = [&](const Symbol &s) {
    if (op_ == Operation::name) return symbolic::hash(0);
    auto it = numbering.find(s.id());
    if (it == numbering.end()) it = numbering.emplace(s.id(), numbering.size()).first;
    return symbolic::hash(it->second);
  };
  for (const auto &d : deps_) {
    h = symbolic::hash_combine(h, d->hash());
    for (const auto &s : d->shape()) h = symbolic::hash_combine(h, hash_sym(s));
  }
  for (const auto &s : shape_) h = symbolic::hash_combine(h, hash_sym(s));
  hash_ = h;
}

void TensorImpl::bind(void *data, std::vector<int64_t> sizes) {
  memory_.address = data;
  if (data) {
    owning_ = false;
    if (!deps_.empty()) force_recompute_ = true;  // caller-provided output
  } else {
    owning_ = true;
  }
  if (shape_.empty()) shape_.resize(sizes.size());
  ASSERT(sizes.size() == shape_.size())
      << "Invalid binding, expected " << shape_.size() << " dims got "
      << sizes.size() << " dims";
  std::unordered_set<size_t> already;
  for (size_t i = 0; i < sizes.size(); ++i) {
    for (const auto &c : constraints_) {
      if (c.first == Expr::size(Expr(shape_[i])) &&
          c.second.type() == Expr::Type::value) {
        already.insert(i);
        ASSERT(c.second.value() == sizes[i])
            << "Already bound " << c.first.dump() << " to " << c.second.dump()
            << ", can't change that to " << sizes[i];
      }
    }
  }
  bool changed = false;
  for (size_t i = 0; i < sizes.size(); ++i) {
    if (already.count(i)) continue;
    constraints_.emplace_back(Expr::size(Expr(shape_[i])), Expr(sizes[i]));
    changed = true;
  }
  updateHash(changed);
}

std::vector<void *> TensorImpl::getInputBuffers(
    std::unordered_set<const TensorImpl *> &seen) const {
  if (seen.count(this)) return {};
  seen.insert(this);
  if (op_ == Operation::constant) return {data<void>()};
  std::vector<void *> out;
  for (const auto &d : deps_)
    for (auto *b : d->getInputBuffers(seen)) out.emplace_back(b);
  return out;
}

void TensorImpl::clear_cache() const {
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  getCompilationCache().erase(hash());
  getLoweredCache().erase(hash());
}

void *TensorImpl::data_ptr() const {
  if (memory_.address && !force_recompute_) return memory_.address;
  force_recompute_ = false;
  auto alloc = [](size_t bytes) { return getDefaultHardware()->alloc(bytes); };

  if (owning_ && !memory_.address && deps_.empty()) {
    size_t n = 1;
    auto sc = size_constraints();
    for (const auto &s : shape_) {
      ASSERT(sc.count(s.id())) << "cannot allocate owned tensor, size for "
                               << s.name() << " not provided";
      n *= sc.at(s.id()).value();
    }
    memory_ = alloc(sizeof(float) * n);
    return memory_.address;
  }

  std::shared_ptr<Compiled> cc;
  int64_t out_size = 0;
  {
    std::lock_guard<std::recursive_mutex> g(getCacheMutex());
    if (!getCompilationCache().count(hash()))
      const_cast<TensorImpl *>(this)->populateCompilationCache();
    cc = getCompilationCache().at(hash());
    out_size = getLoweredCache().at(hash()).output_size;
  }
  if (owning_) {
    if (!memory_.address) memory_ = alloc(sizeof(float) * out_size);
  } else {
    ASSERT(memory_.address) << "non-owning output without a buffer";
  }
  std::unordered_set<const TensorImpl *> seen;
  auto buffers = getInputBuffers(seen);
  buffers.emplace_back(memory_.address);
  cc->run(buffers, true);
  return memory_.address;
}

std::pair<IR, std::unordered_map<int, std::pair<IR::VarRef, int64_t>>>
TensorImpl::lower() const {
  const_cast<TensorImpl *>(this)->unify();
  IR ir;
  std::unordered_map<int, std::pair<IR::VarRef, int64_t>> var_map;
  std::unordered_map<const TensorImpl *, IR::NodeRef> impl_map;
  auto node_ref = resolve(ir, var_map, impl_map);
  std::vector<IR::VarRef> vars;
  for (const auto &s : shape_) vars.emplace_back(var_map.at(s.id()).first);
  auto out = ir.create_node(Operation::write, {node_ref}, vars);
  ir.set_outputs({out});
  return {ir, var_map};
}

IR TensorImpl::ir() const {
  {
    std::lock_guard<std::recursive_mutex> g(getCacheMutex());
    auto &lc = getLoweredCache();
    if (lc.count(hash())) return lc.at(hash()).ir;
  }
  auto ll = lower();
  return schedule(ll.first, ll.second).ir;
}

LoopTree TensorImpl::loop_tree() const {
  {
    std::lock_guard<std::recursive_mutex> g(getCacheMutex());
    auto &lc = getLoweredCache();
    if (lc.count(hash())) return lc.at(hash()).loop_tree;
  }
  auto ll = lower();
  return schedule(ll.first, ll.second);
}

void TensorImpl::compile() {
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  if (!getCompilationCache().count(hash())) populateCompilationCache();
}

std::shared_ptr<Compiled> TensorImpl::compiled() const {
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  if (!getCompilationCache().count(hash())) const_cast<TensorImpl *>(this)->compile();
  return getCompilationCache().at(hash());
}

void TensorImpl::set(const IR &ir) { set(LoopTree(ir)); }

void TensorImpl::set(const LoopTree &loop_tree) {
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  auto h = hash();
  clear_cache();
  if (!getLoweredCache().count(h)) populateLoweredCache();
  auto &lowered = getLoweredCache().at(h);
  lowered = CachedLowered{loop_tree.ir, loop_tree, lowered.output_size, lowered.sizes};
}

// Default schedule: reads and views stay unscheduled; every other node loops
// over its loop vars (VarRef order) at full size with no tail.
LoopTree TensorImpl::schedule(
    IR &ir,
    const std::unordered_map<int, std::pair<IR::VarRef, int64_t>> &var_map) const {
  std::unordered_map<IR::VarRef, int64_t> var_sizes;
  for (const auto &p : var_map) var_sizes[p.second.first] = p.second.second;
  for (auto n : ir.nodes()) {
    auto op = ir.node(n).op();
    if (op == Operation::read || op == Operation::view) {
      ir.set_order(n, {});
      continue;
    }
    std::vector<std::pair<IR::VarRef, IR::LoopSize>> order;
    for (auto v : ir.loop_vars(n))
      order.emplace_back(v, IR::LoopSize{var_sizes.at(v), 0});
    ir.set_order(n, order);
  }
  return LoopTree(ir);
}

std::vector<int64_t> TensorImpl::sizes() const {
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  if (!getLoweredCache().count(hash()))
    const_cast<TensorImpl *>(this)->populateLoweredCache();
  return getLoweredCache().at(hash()).sizes;
}

int64_t TensorImpl::size(int dim) const {
  const_cast<TensorImpl *>(this)->unify();
  ASSERT(dim >= 0 && dim < (int)shape_.size()) << "dim out of range";
  auto id = shape_[dim].id();
  auto sc = size_constraints();
  ASSERT(sc.count(id)) << "couldn't find size of " << Expr(shape_[dim]).dump();
  auto e = sc.at(id);
  if (!e.can_evaluate()) {
    const_cast<TensorImpl *>(this)->unify(true);
    e = size_constraints().at(id);
  }
  ASSERT(e.can_evaluate()) << "cannot resolve symbol " << Expr(shape_[dim]).dump()
                           << " got expr " << e.dump();
  return e.evaluate();
}

IR::NodeRef TensorImpl::resolve(
    IR &ir, std::unordered_map<int, std::pair<IR::VarRef, int64_t>> &var_map,
    std::unordered_map<const TensorImpl *, IR::NodeRef> &impl_map) const {
  auto found = impl_map.find(this);
  if (found != impl_map.end()) return found->second;

  std::vector<IR::NodeRef> node_deps;
  for (const auto &d : deps_) node_deps.emplace_back(d->resolve(ir, var_map, impl_map));

  std::vector<IR::VarRef> vars;
  for (const auto &s : shape_) {
    if (!var_map.count(s.id())) {
      auto sc = size_constraints();
      ASSERT(sc.count(s.id())) << "unbound variable in compute " << s.name()
                               << " (id: " << s.id() << ")";
      if (!sc.at(s.id()).can_evaluate()) {
        const_cast<TensorImpl *>(this)->unify(true);
        sc = size_constraints();
      }
      const auto &e = sc.at(s.id());
      ASSERT(e.can_evaluate()) << "can't resolve size for sym " << s.name()
                               << " expr " << e.dump();
      std::stringstream name;
      name << s.name() << "_" << s.id();
      var_map[s.id()] = {ir.create_var(name.str()), e.evaluate()};
    }
    vars.emplace_back(var_map.at(s.id()).first);
  }
  std::unordered_map<int, IR::VarRef> sym_var_map;
  for (const auto &p : var_map) sym_var_map[p.first] = p.second.first;

  std::vector<Constraint> node_constraints;
  std::unordered_set<uint64_t> seen;
  auto mapped = [&](const Expr &e) {
    for (const auto &s : e.symbols())
      if (!sym_var_map.count(s.id())) return false;
    return true;
  };
  for (const auto &c : constraints_) {
    if (!seen.insert(constraint_hash(c)).second) continue;
    if (mapped(c.first) && mapped(c.second)) node_constraints.emplace_back(c);
  }

  IR::NodeRef node_ref = -1;
  switch (op_) {
    case Operation::name:
      ASSERT(node_deps.size() == 1) << "invalid rename (only 1 input allowed)";
      node_ref = node_deps[0];
      break;
    case Operation::view:
      node_ref = ir.create_node(Operation::view, node_deps, vars, node_constraints,
                                sym_var_map);
      break;
    case Operation::constant:
      node_ref = ir.create_node(Operation::read, {}, vars);
      ir.add_input(node_ref);
      break;
    case Operation::add:
    case Operation::subtract:
    case Operation::multiply:
    case Operation::divide:
    case Operation::min:
    case Operation::max:
    case Operation::exp:
    case Operation::log:
    case Operation::sqrt:
    case Operation::abs:
    case Operation::reciprocal:
    case Operation::negate:
      node_ref = ir.create_node(op_, node_deps, vars);
      break;
    default:
      break;
  }
  ASSERT(node_ref > -1) << "couldn't resolve node op: " << loop_tool::dump(op_);
  impl_map.emplace(this, node_ref);
  return node_ref;
}

void TensorImpl::collectConstraints(std::vector<Constraint> &constraints,
                                    std::unordered_set<TensorImpl *> &seen,
                                    std::unordered_set<int64_t> &seen_hashes) {
  for (const auto &c : constraints_)
    if (seen_hashes.insert((int64_t)constraint_hash(c)).second)
      constraints.emplace_back(c);
  seen.insert(this);
  for (const auto &d : deps_)
    if (!seen.count(d.get())) d->collectConstraints(constraints, seen, seen_hashes);
}

void TensorImpl::propagateConstraints(const std::vector<Constraint> &constraints,
                                      std::unordered_set<TensorImpl *> &seen) {
  std::vector<Symbol> mine;
  for (const auto &c : constraints_) {
    for (const auto &s : c.first.symbols()) mine.emplace_back(s);
    for (const auto &s : c.second.symbols()) mine.emplace_back(s);
  }
  for (const auto &s : shape_) mine.emplace_back(s);
  constraints_.clear();
  for (const auto &c : constraints) {
    bool relevant = false;
    for (const auto &s : mine)
      if (c.first.contains(s) || c.second.contains(s)) {
        relevant = true;
        break;
      }
    if (relevant) constraints_.emplace_back(c);
  }
  seen.insert(this);
  for (const auto &d : deps_)
    if (!seen.count(d.get())) d->propagateConstraints(constraints, seen);
}

void TensorImpl::unifyConstraints() {
  std::vector<Constraint> constraints;
  std::unordered_set<TensorImpl *> seen;
  std::unordered_set<int64_t> seen_hashes;
  collectConstraints(constraints, seen, seen_hashes);
  auto solved = symbolic::evaluate(symbolic::unify(constraints));
  seen.clear();
  propagateConstraints(solved, seen);
}

// `.as()` renames: every symbol of the renamed tensor maps to the new name.
void TensorImpl::collectSymbolMap(std::unordered_map<int, Symbol> &symbol_map,
                                  std::unordered_set<TensorImpl *> &seen) {
  if (op_ == Operation::name) {
    ASSERT(deps_.size() == 1) << "invalid rename";
    const auto &dep_shape = deps_[0]->shape();
    ASSERT(dep_shape.size() == shape_.size())
        << "found shape size of " << dep_shape.size() << " dims expected "
        << shape_.size();
    for (size_t i = 0; i < shape_.size(); ++i) {
      auto s = shape_[i];
      std::unordered_set<int> chain{s.id()};
      while (symbol_map.count(s.id())) {
        s = symbol_map.at(s.id());
        ASSERT(!chain.count(s.id())) << "unexpected cycle found in symbol map";
        chain.insert(s.id());
      }
      if (!chain.count(dep_shape[i].id())) symbol_map[dep_shape[i].id()] = shape_[i];
    }
  }
  seen.insert(this);
  for (const auto &d : deps_)
    if (!seen.count(d.get())) d->collectSymbolMap(symbol_map, seen);
}

void TensorImpl::propagateSymbolMap(const std::unordered_map<int, Symbol> &symbol_map,
                                    std::unordered_set<TensorImpl *> &seen) {
  for (auto &s : shape_) {
    while (symbol_map.count(s.id())) {
      auto old = s;
      s = symbol_map.at(old.id());
      for (auto &c : constraints_) {
        c.first = c.first.replace(old, s);
        c.second = c.second.replace(old, s);
      }
    }
  }
  seen.insert(this);
  for (const auto &d : deps_)
    if (!seen.count(d.get())) d->propagateSymbolMap(symbol_map, seen);
}

void TensorImpl::unifySymbols() {
  std::unordered_map<int, Symbol> symbol_map;
  std::unordered_set<TensorImpl *> seen;
  collectSymbolMap(symbol_map, seen);
  seen.clear();
  propagateSymbolMap(symbol_map, seen);
}

void TensorImpl::unify(bool force) {
  if (!force && unified_) return;
  for (const auto &d : deps_) d->unify(force);
  unified_ = true;
  unifySymbols();
  unifyConstraints();
}

void TensorImpl::populateLoweredCache() {
  auto h = hash();
  auto ll = lower();
  auto loop_tree = schedule(ll.first, ll.second);
  int64_t total = 1;
  std::vector<int64_t> sizes;
  for (const auto &s : shape_) {
    auto n = ll.second.at(s.id()).second;
    total *= n;
    sizes.emplace_back(n);
  }
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  getLoweredCache().emplace(h, CachedLowered{ll.first, loop_tree, total, sizes});
}

void TensorImpl::populateCompilationCache() {
  std::lock_guard<std::recursive_mutex> g(getCacheMutex());
  auto h = hash();
  if (!getLoweredCache().count(h)) populateLoweredCache();
  auto loop_tree = getLoweredCache().at(h).loop_tree;
  std::shared_ptr<Compiled> cc = getDefaultBackend()->compile(loop_tree);
  getCompilationCache().emplace(h, cc);
}

// ---------------------------------------------------------------- Tensor ----
std::vector<Symbol> Tensor::broadcast_shape(const Tensor &rhs) const {
  std::unordered_set<int> ids;
  std::vector<Symbol> out;
  for (const auto &s : impl_->shape()) {
    ids.insert(s.id());
    out.emplace_back(s);
  }
  for (const auto &s : rhs.shape())
    if (!ids.count(s.id())) out.emplace_back(s);
  return out;
}

Tensor Tensor::operator|(const Tensor &rhs) const {
  ASSERT(impl_->shape().size() == rhs.impl()->shape().size())
      << "concatenation needs tensors of equal rank";
  std::vector<Symbol> out_shape;
  std::vector<Constraint> constraints;
  for (size_t i = 0; i < impl_->shape().size(); ++i) {
    auto a = impl_->shape()[i];
    auto b = rhs.impl()->shape()[i];
    if (a == b) {
      out_shape.emplace_back(a);
      continue;
    }
    Symbol joined(a.name() + b.name());
    out_shape.emplace_back(joined);
    constraints.emplace_back(Expr(joined), Expr(a));
    constraints.emplace_back(Expr(joined), Expr(b) + Expr::size(Expr(a)));
    constraints.emplace_back(Expr::size(Expr(joined)),
                             Expr::size(Expr(a)) + Expr::size(Expr(b)));
  }
  return to(out_shape, constraints) + rhs.to(out_shape, constraints);
}

Tensor Tensor::pad(Symbol padded_dim, int64_t pre, int64_t post) const {
  ASSERT(pre >= 0) << "cannot pad by a negative number";
  ASSERT(post >= 0) << "cannot pad by a negative number";
  ASSERT(pre || post) << "unecessary pad operation (padded by zero)";
  Symbol padded(padded_dim.name() + "_p_" + std::to_string(pre) + "_" +
                std::to_string(post));
  std::vector<Symbol> out_shape;
  for (const auto &s : shape()) out_shape.emplace_back(s == padded_dim ? padded : s);
  return to(out_shape, Constraint(Expr(padded), Expr(padded_dim) + Expr(pre)),
            Constraint(Expr::size(Expr(padded)),
                       Expr::size(Expr(padded_dim)) + Expr(post + pre)));
}

Tensor Tensor::transpose(std::vector<int> order) const {
  ASSERT(order.size() == shape().size()) << "invalid transpose";
  std::vector<Symbol> new_shape;
  for (auto idx : order) {
    ASSERT(idx >= 0 && idx < (int)order.size()) << "invalid transpose";
    new_shape.emplace_back(shape().at(idx));
  }
  return transpose(new_shape);
}

Tensor Tensor::transpose(std::vector<Symbol> new_shape) const {
  ASSERT(new_shape.size() == shape().size()) << "invalid transpose";
  auto impl = std::make_shared<TensorImpl>(*impl_);
  impl->shape_ = new_shape;
  return Tensor(impl);
}

Tensor Tensor::flatten(std::vector<Symbol> symbols, Symbol new_symbol) const {
  ASSERT(!symbols.empty()) << "flatten needs at least one symbol";
  std::unordered_set<int> flat;
  for (const auto &s : symbols) flat.insert(s.id());
  std::vector<Symbol> out_shape;
  bool in_run = false, ended = false;
  for (const auto &s : shape()) {
    if (flat.count(s.id())) {
      ASSERT(!ended) << "flatten only works on contiguous dimensions. "
                        "Otherwise, transpose first.";
      in_run = true;
      continue;
    }
    if (in_run) ended = true;
    out_shape.emplace_back(s);
  }
  out_shape.emplace_back(new_symbol);
  // new = s0*|s1|*..*|sn| + ... + sn  (row-major over the flattened symbols)
  Expr index(0);
  Expr stride(1);
  for (auto it = symbols.rbegin(); it != symbols.rend(); ++it) {
    index = Expr(*it) * stride + index;
    stride = Expr::size(Expr(*it)) * stride;
  }
  return to(out_shape, Constraint(Expr(new_symbol), index));
}

Tensor Tensor::reduce(Operation op, const std::vector<Symbol> &reduction_vars) const {
  ASSERT(!reduction_vars.empty()) << "reduction variables required (got none)";
  std::unordered_set<int> red;
  for (const auto &s : reduction_vars) red.insert(s.id());
  std::vector<Symbol> out_shape;
  for (const auto &s : shape())
    if (!red.count(s.id())) out_shape.emplace_back(s);
  ASSERT(out_shape.size() != shape().size()) << "reduction variables not over any input";
  return Tensor(std::make_shared<TensorImpl>(
      op, out_shape, std::vector<std::shared_ptr<TensorImpl>>{impl_}));
}

}  // namespace lazy
}  // namespace loop_tool
