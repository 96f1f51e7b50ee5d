// C ABI of the loop_tool frontend (see include/loop_tool_c.h).
//
// Written against the public loop_tool C++ API only, so the same source also
// builds against another implementation of that API (the CPU oracle builds it
// against the reference headers to drive the reference library with
// identical calls).
#include <algorithm>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>

#include "loop_tool/backend.h"
#include "loop_tool/hardware.h"
#include "loop_tool/ir.h"
#include "loop_tool/lazy.h"
#include "loop_tool/mutate.h"
#include "loop_tool/nn.h"
#include "loop_tool/serialization.h"
#include "loop_tool_c.h"

namespace lz = loop_tool::lazy;
using loop_tool::LoopTree;

struct ltc_symbol {
  lz::Symbol s;
};
struct ltc_expr {
  lz::Expr e;
};
struct ltc_tensor {
  lz::Tensor t;
};
struct ltc_looptree {
  LoopTree lt;
};
struct ltc_compiled {
  std::shared_ptr<loop_tool::Compiled> cc;
};

namespace {

thread_local std::string g_error;

int copy_out(const std::string &s, char *buf, int cap) {
  if (buf && cap > 0) {
    size_t n = std::min(s.size(), (size_t)cap - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return (int)s.size();
}

template <typename F>
auto guard_ptr(F f) -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_error = e.what();
  } catch (...) {
    g_error = "unknown error";
  }
  return nullptr;
}

template <typename F>
int guard_int(F f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_error = e.what();
  } catch (...) {
    g_error = "unknown error";
  }
  return -1;
}

std::vector<lz::Symbol> syms_of(ltc_symbol *const *syms, int n) {
  std::vector<lz::Symbol> out;
  for (int i = 0; i < n; ++i) out.push_back(syms[i]->s);
  return out;
}

ltc_tensor *wrap(lz::Tensor t) { return new ltc_tensor{t}; }
ltc_looptree *wrap(const LoopTree &lt) { return new ltc_looptree{lt}; }

}  // namespace

extern "C" {

const char *ltc_last_error(void) { return g_error.c_str(); }
int ltc_abi_version(void) { return 1; }

int ltc_backends(char *buf, int cap) {
  return guard_int([&] {
    std::vector<std::string> names;
    for (const auto &p : loop_tool::getBackends()) names.push_back(p.first);
    std::sort(names.begin(), names.end());
    std::string s;
    for (size_t i = 0; i < names.size(); ++i) s += (i ? "," : "") + names[i];
    return copy_out(s, buf, cap);
  });
}

int ltc_set_default_backend(const char *name) {
  return guard_int([&] {
    loop_tool::setDefaultBackend(name);
    return 0;
  });
}

int ltc_default_backend(char *buf, int cap) {
  return guard_int([&] { return copy_out(loop_tool::getDefaultBackend()->name(), buf, cap); });
}

int ltc_load_library(const char *path) {
  return guard_int([&] {
    loop_tool::loadLibrary(path);
    return 0;
  });
}

int ltc_set_default_hardware(const char *name) {
  return guard_int([&] {
    for (const auto &hw : loop_tool::getHardware()) {
      if (hw->name() == name) {
        loop_tool::setDefaultHardwareId(hw->id());
        return 0;
      }
    }
    throw std::runtime_error(std::string("no hardware named ") + name);
  });
}

// ---------------------------------------------------------- symbols/exprs ----
ltc_symbol *ltc_symbol_new(const char *name) {
  return guard_ptr([&] { return new ltc_symbol{lz::Symbol(std::string(name))}; });
}
void ltc_symbol_free(ltc_symbol *s) { delete s; }
int ltc_symbol_id(const ltc_symbol *s) { return s->s.id(); }
int ltc_symbol_name(const ltc_symbol *s, char *buf, int cap) {
  return copy_out(s->s.name(), buf, cap);
}

ltc_expr *ltc_expr_value(int64_t v) {
  return guard_ptr([&] { return new ltc_expr{lz::Expr((int64_t)v)}; });
}
ltc_expr *ltc_expr_symbol(const ltc_symbol *s) {
  return guard_ptr([&] { return new ltc_expr{lz::Expr(s->s)}; });
}
ltc_expr *ltc_expr_size(const ltc_symbol *s) {
  return guard_ptr([&] { return new ltc_expr{lz::Expr::size(lz::Expr(s->s))}; });
}
ltc_expr *ltc_expr_op(int op, const ltc_expr *a, const ltc_expr *b) {
  return guard_ptr([&]() -> ltc_expr * {
    switch (op) {
      case LTC_E_ADD: return new ltc_expr{a->e + b->e};
      case LTC_E_SUB: return new ltc_expr{a->e - b->e};
      case LTC_E_MUL: return new ltc_expr{a->e * b->e};
      case LTC_E_DIV: return new ltc_expr{a->e / b->e};
      case LTC_E_MOD: return new ltc_expr{a->e % b->e};
      case LTC_E_MAX: return new ltc_expr{lz::Expr::max(a->e, b->e)};
      case LTC_E_NEG: return new ltc_expr{-a->e};
      default: throw std::runtime_error("unknown expression op");
    }
  });
}
void ltc_expr_free(ltc_expr *e) { delete e; }
int ltc_expr_dump(const ltc_expr *e, char *buf, int cap) {
  return guard_int([&] { return copy_out(e->e.dump(true), buf, cap); });
}

// ---------------------------------------------------------------- tensors ----
ltc_tensor *ltc_tensor_from_sizes(const int64_t *sizes, int n) {
  return guard_ptr([&] { return wrap(lz::Tensor(std::vector<int64_t>(sizes, sizes + n))); });
}
ltc_tensor *ltc_tensor_from_symbols(ltc_symbol *const *syms, int n) {
  return guard_ptr([&] { return wrap(lz::Tensor(syms_of(syms, n))); });
}
ltc_tensor *ltc_tensor_from_data(void *data, const int64_t *sizes, int n) {
  return guard_ptr(
      [&] { return wrap(lz::Tensor(data, std::vector<int64_t>(sizes, sizes + n))); });
}
int ltc_tensor_bind(ltc_tensor *t, void *data, const int64_t *sizes, int n) {
  return guard_int([&] {
    t->t.bind(data, std::vector<int64_t>(sizes, sizes + n));
    return 0;
  });
}
void ltc_tensor_free(ltc_tensor *t) { delete t; }

ltc_tensor *ltc_tensor_binary(int op, const ltc_tensor *a, const ltc_tensor *b) {
  return guard_ptr([&]() -> ltc_tensor * {
    switch (op) {
      case LTC_T_ADD: return wrap(a->t + b->t);
      case LTC_T_SUB: return wrap(a->t - b->t);
      case LTC_T_MUL: return wrap(a->t * b->t);
      case LTC_T_DIV: return wrap(a->t / b->t);
      case LTC_T_MAX: return wrap(a->t.max(b->t));
      case LTC_T_MIN: return wrap(a->t.min(b->t));
      case LTC_T_CONCAT: return wrap(a->t | b->t);
      default: throw std::runtime_error("unknown binary op");
    }
  });
}

ltc_tensor *ltc_tensor_unary(int op, const ltc_tensor *t) {
  return guard_ptr([&]() -> ltc_tensor * {
    switch (op) {
      case LTC_U_EXP: return wrap(t->t.exp());
      case LTC_U_LOG: return wrap(t->t.log());
      case LTC_U_SQRT: return wrap(t->t.sqrt());
      case LTC_U_ABS: return wrap(t->t.abs());
      case LTC_U_NEG: return wrap(-t->t);
      case LTC_U_RECIP: return wrap(t->t.reciprocal());
      default: throw std::runtime_error("unknown unary op");
    }
  });
}

ltc_tensor *ltc_tensor_reduce(int op, const ltc_tensor *t, ltc_symbol *const *syms, int n) {
  return guard_ptr([&]() -> ltc_tensor * {
    auto s = syms_of(syms, n);
    switch (op) {
      case LTC_R_SUM: return wrap(t->t.sum(s));
      case LTC_R_PROD: return wrap(t->t.prod(s));
      case LTC_R_MAX: return wrap(t->t.max(s));
      case LTC_R_MIN: return wrap(t->t.min(s));
      default: throw std::runtime_error("unknown reduction op");
    }
  });
}

ltc_tensor *ltc_tensor_as(const ltc_tensor *t, ltc_symbol *const *syms, int n) {
  return guard_ptr([&] { return wrap(t->t.as(syms_of(syms, n))); });
}

ltc_tensor *ltc_tensor_to(const ltc_tensor *t, ltc_symbol *const *syms, int n,
                          ltc_expr *const *lhs, ltc_expr *const *rhs, int nc) {
  return guard_ptr([&] {
    std::vector<lz::Constraint> cs;
    for (int i = 0; i < nc; ++i) cs.emplace_back(lhs[i]->e, rhs[i]->e);
    return wrap(t->t.to(syms_of(syms, n), cs));
  });
}

ltc_tensor *ltc_tensor_pad(const ltc_tensor *t, const ltc_symbol *dim, int64_t pre,
                           int64_t post) {
  return guard_ptr([&] { return wrap(t->t.pad(dim->s, pre, post)); });
}

ltc_tensor *ltc_tensor_transpose(const ltc_tensor *t, ltc_symbol *const *syms, int n) {
  return guard_ptr([&] { return wrap(t->t.transpose(syms_of(syms, n))); });
}

ltc_tensor *ltc_tensor_flatten(const ltc_tensor *t, ltc_symbol *const *syms, int n,
                               const ltc_symbol *new_sym) {
  return guard_ptr([&] { return wrap(t->t.flatten(syms_of(syms, n), new_sym->s)); });
}

ltc_tensor *ltc_nn_convolve(const ltc_tensor *X, const ltc_tensor *W,
                            ltc_symbol *const *spatial, ltc_symbol *const *window, int n,
                            int stride) {
  return guard_ptr([&] {
    return wrap(loop_tool::nn::convolve(X->t, W->t, syms_of(spatial, n),
                                        syms_of(window, n), stride));
  });
}

ltc_tensor *ltc_nn_maxpool(const ltc_tensor *X, ltc_symbol *const *spatial, int n, int k,
                           int stride) {
  return guard_ptr(
      [&] { return wrap(loop_tool::nn::maxpool(X->t, syms_of(spatial, n), k, stride)); });
}

int ltc_tensor_ndim(const ltc_tensor *t) { return (int)t->t.shape().size(); }

int ltc_tensor_shape(const ltc_tensor *t, ltc_symbol **out, int cap) {
  return guard_int([&] {
    auto shape = t->t.shape();
    for (int i = 0; i < (int)shape.size() && i < cap; ++i) out[i] = new ltc_symbol{shape[i]};
    return (int)shape.size();
  });
}

int ltc_tensor_sizes(const ltc_tensor *t, int64_t *out, int cap) {
  return guard_int([&] {
    auto s = t->t.sizes();
    for (int i = 0; i < (int)s.size() && i < cap; ++i) out[i] = s[i];
    return (int)s.size();
  });
}

int64_t ltc_tensor_size(const ltc_tensor *t, int dim) {
  try {
    return (int64_t)t->t.size(dim);
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

float *ltc_tensor_data(const ltc_tensor *t) {
  return guard_ptr([&] { return t->t.data<float>(); });
}

uint64_t ltc_tensor_hash(const ltc_tensor *t) { return t->t.hash(); }

int ltc_tensor_code(const ltc_tensor *t, char *buf, int cap) {
  return guard_int([&] { return copy_out(t->t.code(), buf, cap); });
}

int ltc_tensor_compile(const ltc_tensor *t) {
  return guard_int([&] {
    t->t.compile();
    return 0;
  });
}

ltc_looptree *ltc_tensor_loop_tree(const ltc_tensor *t) {
  return guard_ptr([&] { return wrap(t->t.loop_tree()); });
}

int ltc_tensor_set_loop_tree(const ltc_tensor *t, const ltc_looptree *lt) {
  return guard_int([&] {
    const_cast<lz::Tensor &>(t->t).set(lt->lt);
    return 0;
  });
}

ltc_compiled *ltc_tensor_compiled(const ltc_tensor *t) {
  return guard_ptr([&] { return new ltc_compiled{t->t.compiled()}; });
}

int ltc_tensor_clear_cache(const ltc_tensor *t) {
  return guard_int([&] {
    t->t.clear_cache();
    return 0;
  });
}

int ltc_tensor_force_recompute(const ltc_tensor *t) {
  return guard_int([&] {
    t->t.force_recompute();
    return 0;
  });
}

// --------------------------------------------------------------- LoopTree ----
void ltc_looptree_free(ltc_looptree *lt) { delete lt; }

int ltc_looptree_dump(const ltc_looptree *lt, char *buf, int cap) {
  return guard_int([&] { return copy_out(lt->lt.dump(), buf, cap); });
}

int ltc_looptree_size(const ltc_looptree *lt) { return (int)lt->lt.nodes.size(); }

int ltc_looptree_roots(const ltc_looptree *lt, int *out, int cap) {
  const auto &r = lt->lt.roots;
  for (int i = 0; i < (int)r.size() && i < cap; ++i) out[i] = r[i];
  return (int)r.size();
}

int ltc_looptree_children(const ltc_looptree *lt, int ref, int *out, int cap) {
  return guard_int([&] {
    const auto &c = lt->lt.children(ref);
    for (int i = 0; i < (int)c.size() && i < cap; ++i) out[i] = c[i];
    return (int)c.size();
  });
}

int ltc_looptree_parent(const ltc_looptree *lt, int ref) {
  // -1 is a root's parent (LoopTree::parent), so errors return -2
  try {
    return (int)lt->lt.parent(ref);
  } catch (const std::exception &e) {
    g_error = e.what();
  } catch (...) {
    g_error = "unknown error";
  }
  return -2;
}

int ltc_looptree_is_loop(const ltc_looptree *lt, int ref) {
  return guard_int([&] { return lt->lt.kind(ref) == LoopTree::LOOP ? 1 : 0; });
}

int ltc_looptree_loop(const ltc_looptree *lt, int ref, int *var, int *var_depth,
                      int64_t *size, int64_t *tail) {
  return guard_int([&] {
    auto l = lt->lt.loop(ref);
    *var = l.var;
    *var_depth = l.var_depth;
    *size = l.size;
    *tail = l.tail;
    return 0;
  });
}

int ltc_looptree_var_name(const ltc_looptree *lt, int var, char *buf, int cap) {
  return guard_int([&] { return copy_out(lt->lt.ir.var(var).name(), buf, cap); });
}

int ltc_looptree_node(const ltc_looptree *lt, int ref) {
  return guard_int([&] { return (int)lt->lt.node(ref); });
}

int ltc_looptree_node_vars(const ltc_looptree *lt, int ref, int *out, int cap) {
  return guard_int([&] {
    const auto vars = lt->lt.ir.node(lt->lt.node(ref)).vars();
    for (int i = 0; i < (int)vars.size() && i < cap; ++i) out[i] = (int)vars[i];
    return (int)vars.size();
  });
}

int ltc_looptree_node_op(const ltc_looptree *lt, int ref, char *buf, int cap) {
  return guard_int([&] {
    return copy_out(loop_tool::dump(lt->lt.ir.node(lt->lt.node(ref)).op()), buf, cap);
  });
}

int ltc_looptree_annotation(const ltc_looptree *lt, int ref, char *buf, int cap) {
  return guard_int([&] { return copy_out(lt->lt.annotation(ref), buf, cap); });
}

int ltc_looptree_num_inputs(const ltc_looptree *lt) {
  return (int)lt->lt.ir.inputs().size();
}
int ltc_looptree_num_outputs(const ltc_looptree *lt) {
  return (int)lt->lt.ir.outputs().size();
}

int ltc_looptree_io_numel(const ltc_looptree *lt, int64_t *out, int cap) {
  return guard_int([&] {
    loop_tool::Compiler c(lt->lt);
    std::vector<int64_t> n;
    for (auto nr : lt->lt.ir.inputs()) {
      int64_t e = 1;
      for (auto v : lt->lt.ir.node(nr).vars()) e *= (int64_t)c.var_sizes.at(v);
      n.push_back(e);
    }
    for (auto nr : lt->lt.ir.outputs()) {
      int64_t e = 1;
      for (auto v : lt->lt.ir.node(nr).vars()) e *= (int64_t)c.var_sizes.at(v);
      n.push_back(e);
    }
    for (int i = 0; i < (int)n.size() && i < cap; ++i) out[i] = n[i];
    return (int)n.size();
  });
}

int64_t ltc_looptree_flops(const ltc_looptree *lt) {
  try {
    return loop_tool::flops(lt->lt);
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

int ltc_looptree_trivially_parallel(const ltc_looptree *lt, int ref) {
  return guard_int([&] { return loop_tool::is_trivially_parallel(lt->lt, ref) ? 1 : 0; });
}

ltc_looptree *ltc_looptree_split(const ltc_looptree *lt, int ref, int64_t size) {
  return guard_ptr([&] { return wrap(loop_tool::split(lt->lt, ref, size)); });
}
ltc_looptree *ltc_looptree_merge(const ltc_looptree *lt, int ref) {
  return guard_ptr([&] { return wrap(loop_tool::merge(lt->lt, ref)); });
}
ltc_looptree *ltc_looptree_swap_loops(const ltc_looptree *lt, int a, int b) {
  return guard_ptr([&] { return wrap(loop_tool::swap_loops(lt->lt, a, b)); });
}
ltc_looptree *ltc_looptree_swap_nodes(const ltc_looptree *lt, int a, int b) {
  return guard_ptr([&] { return wrap(loop_tool::swap_nodes(lt->lt, a, b)); });
}
ltc_looptree *ltc_looptree_try_swap(const ltc_looptree *lt, int a, int b) {
  return guard_ptr([&] { return wrap(loop_tool::try_swap(lt->lt, a, b)); });
}
ltc_looptree *ltc_looptree_annotate(const ltc_looptree *lt, int ref, const char *annot) {
  return guard_ptr([&] { return wrap(loop_tool::annotate(lt->lt, ref, std::string(annot))); });
}
ltc_looptree *ltc_looptree_unroll_inner(const ltc_looptree *lt, int amount) {
  return guard_ptr([&] { return wrap(loop_tool::unroll_inner_loops(lt->lt, amount)); });
}
ltc_looptree *ltc_looptree_maximize_reuse(const ltc_looptree *lt) {
  return guard_ptr([&] { return wrap(loop_tool::maximize_reuse(lt->lt)); });
}
ltc_looptree *ltc_looptree_disable_reuse(const ltc_looptree *lt, int loop_ref, int node) {
  return guard_ptr([&] { return wrap(loop_tool::disable_reuse(lt->lt, loop_ref, node)); });
}
ltc_looptree *ltc_looptree_enable_reuse(const ltc_looptree *lt, int loop_ref, int node) {
  return guard_ptr([&] { return wrap(loop_tool::enable_reuse(lt->lt, loop_ref, node)); });
}
int ltc_looptree_map_ref(const ltc_looptree *new_lt, int old_ref, const ltc_looptree *old_lt) {
  return guard_int([&] { return (int)loop_tool::map_ref(new_lt->lt, old_ref, old_lt->lt); });
}

int ltc_looptree_serialize(const ltc_looptree *lt, char *buf, int cap) {
  return guard_int([&] { return copy_out(loop_tool::serialize(lt->lt.ir), buf, cap); });
}

ltc_looptree *ltc_looptree_deserialize(const char *text) {
  return guard_ptr([&] { return wrap(LoopTree(loop_tool::deserialize(text))); });
}

// --------------------------------------------------------------- compiled ----
ltc_compiled *ltc_compile(const ltc_looptree *lt, const char *backend) {
  return guard_ptr([&] {
    const auto &b = loop_tool::getBackends();
    auto it = b.find(backend);
    if (it == b.end()) throw std::runtime_error(std::string("no backend named ") + backend);
    std::shared_ptr<loop_tool::Compiled> cc = it->second->compile(lt->lt);
    return new ltc_compiled{cc};
  });
}

int ltc_compiled_run(const ltc_compiled *cc, void *const *mem, int n, int sync) {
  return guard_int([&] {
    cc->cc->run(std::vector<void *>(mem, mem + n), sync != 0);
    return 0;
  });
}

int ltc_compiled_dump(const ltc_compiled *cc, char *buf, int cap) {
  return guard_int([&] { return copy_out(cc->cc->dump(), buf, cap); });
}

int ltc_compiled_int_property(const ltc_compiled *cc, const char *key, int64_t *out) {
  return guard_int([&] {
    auto it = cc->cc->int_properties.find(key);
    if (it == cc->cc->int_properties.end()) return 0;
    *out = it->second;
    return 1;
  });
}

int ltc_compiled_string_property(const ltc_compiled *cc, const char *key, char *buf,
                                 int cap) {
  return guard_int([&] {
    auto it = cc->cc->string_properties.find(key);
    if (it == cc->cc->string_properties.end()) return copy_out("", buf, cap);
    return copy_out(it->second, buf, cap);
  });
}

int ltc_compiled_name(const ltc_compiled *cc, char *buf, int cap) {
  return copy_out(cc->cc->name, buf, cap);
}

void ltc_compiled_free(ltc_compiled *cc) { delete cc; }

}  // extern "C"
