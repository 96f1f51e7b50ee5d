/*
 * loop_tool_c.h -- C ABI of the loop_tool frontend for foreign-function
 * callers (Python ctypes, Go cgo, Java JNI, ...).
 *
 * It exposes the surface that the reference binds for Python with pybind11
 * (src/frontends/python.cpp:86-698): backend selection, symbols and index
 * expressions, lazy tensors and their ops (python.cpp:339-575), LoopTree
 * inspection and schedule transforms (python.cpp:228-337), and compiling /
 * running a LoopTree with a named backend (python.cpp:181-218).  Handles are
 * opaque; every handle returned must be released with its *_free function.
 *
 * Errors: functions returning a pointer return NULL, functions returning int
 * return a negative value; the message is ltc_last_error() (per thread).  No
 * C++ exception crosses the boundary (the reference surfaces its ASSERT
 * failures as Python RuntimeError, python.cpp via pybind11).
 *
 * String out-parameters follow snprintf: the return value is the full
 * length; at most cap-1 bytes plus a NUL are written.
 */
#ifndef LOOP_TOOL_C_H_
#define LOOP_TOOL_C_H_

#include <stddef.h>
#include <stdint.h>

#ifndef LTC_API
#define LTC_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#pragma GCC visibility push(default)

typedef struct ltc_symbol ltc_symbol;
typedef struct ltc_expr ltc_expr;
typedef struct ltc_tensor ltc_tensor;
typedef struct ltc_looptree ltc_looptree;
typedef struct ltc_compiled ltc_compiled;

const char *ltc_last_error(void);
int ltc_abi_version(void);

/* backends & hardware: python.cpp:87-110 (load_lib, backends,
 * set_default_hardware, set_default_backend, get_default_backend) */
int ltc_backends(char *buf, int cap);           /* comma separated */
int ltc_set_default_backend(const char *name);
int ltc_default_backend(char *buf, int cap);
int ltc_load_library(const char *path);
int ltc_set_default_hardware(const char *name);

/* symbols & expressions: python.cpp:339-366 (Symbol, Expr, Size) */
ltc_symbol *ltc_symbol_new(const char *name);
void ltc_symbol_free(ltc_symbol *s);
int ltc_symbol_id(const ltc_symbol *s);
int ltc_symbol_name(const ltc_symbol *s, char *buf, int cap);

/* expression ops */
enum { LTC_E_ADD = 0, LTC_E_SUB, LTC_E_MUL, LTC_E_DIV, LTC_E_MOD, LTC_E_MAX, LTC_E_NEG };
ltc_expr *ltc_expr_value(int64_t v);
ltc_expr *ltc_expr_symbol(const ltc_symbol *s);
ltc_expr *ltc_expr_size(const ltc_symbol *s);
ltc_expr *ltc_expr_op(int op, const ltc_expr *a, const ltc_expr *b); /* b unused for NEG */
void ltc_expr_free(ltc_expr *e);
int ltc_expr_dump(const ltc_expr *e, char *buf, int cap);

/* lazy tensors: python.cpp:376-575 */
ltc_tensor *ltc_tensor_from_sizes(const int64_t *sizes, int n);   /* lz::Tensor(sizes) */
ltc_tensor *ltc_tensor_from_symbols(ltc_symbol *const *syms, int n);
ltc_tensor *ltc_tensor_from_data(void *data, const int64_t *sizes, int n); /* not copied */
int ltc_tensor_bind(ltc_tensor *t, void *data, const int64_t *sizes, int n);
void ltc_tensor_free(ltc_tensor *t);

enum { LTC_T_ADD = 0, LTC_T_SUB, LTC_T_MUL, LTC_T_DIV, LTC_T_MAX, LTC_T_MIN, LTC_T_CONCAT };
ltc_tensor *ltc_tensor_binary(int op, const ltc_tensor *a, const ltc_tensor *b);
enum { LTC_U_EXP = 0, LTC_U_LOG, LTC_U_SQRT, LTC_U_ABS, LTC_U_NEG, LTC_U_RECIP };
ltc_tensor *ltc_tensor_unary(int op, const ltc_tensor *t);
enum { LTC_R_SUM = 0, LTC_R_PROD, LTC_R_MAX, LTC_R_MIN };
ltc_tensor *ltc_tensor_reduce(int op, const ltc_tensor *t, ltc_symbol *const *syms, int n);
ltc_tensor *ltc_tensor_as(const ltc_tensor *t, ltc_symbol *const *syms, int n);
/* t.to(shape, {lhs[i] = rhs[i]}) */
ltc_tensor *ltc_tensor_to(const ltc_tensor *t, ltc_symbol *const *syms, int n,
                          ltc_expr *const *lhs, ltc_expr *const *rhs, int nc);
ltc_tensor *ltc_tensor_pad(const ltc_tensor *t, const ltc_symbol *dim, int64_t pre,
                           int64_t post);
ltc_tensor *ltc_tensor_transpose(const ltc_tensor *t, ltc_symbol *const *syms, int n);
ltc_tensor *ltc_tensor_flatten(const ltc_tensor *t, ltc_symbol *const *syms, int n,
                               const ltc_symbol *new_sym);
/* nn helpers: src/frontends/nn.cpp:15-76 */
ltc_tensor *ltc_nn_convolve(const ltc_tensor *X, const ltc_tensor *W,
                            ltc_symbol *const *spatial, ltc_symbol *const *window, int n,
                            int stride);
ltc_tensor *ltc_nn_maxpool(const ltc_tensor *X, ltc_symbol *const *spatial, int n, int k,
                           int stride);

int ltc_tensor_ndim(const ltc_tensor *t);
int ltc_tensor_shape(const ltc_tensor *t, ltc_symbol **out, int cap); /* new handles */
int ltc_tensor_sizes(const ltc_tensor *t, int64_t *out, int cap);
int64_t ltc_tensor_size(const ltc_tensor *t, int dim);
float *ltc_tensor_data(const ltc_tensor *t); /* compiles and runs when needed */
uint64_t ltc_tensor_hash(const ltc_tensor *t);
int ltc_tensor_code(const ltc_tensor *t, char *buf, int cap);
int ltc_tensor_compile(const ltc_tensor *t);
ltc_looptree *ltc_tensor_loop_tree(const ltc_tensor *t);
int ltc_tensor_set_loop_tree(const ltc_tensor *t, const ltc_looptree *lt);
ltc_compiled *ltc_tensor_compiled(const ltc_tensor *t);
/* drop the program's compiled and lowered cache entries (python.cpp:566) */
int ltc_tensor_clear_cache(const ltc_tensor *t);
/* the next data access recomputes the tensor (python.cpp:565) */
int ltc_tensor_force_recompute(const ltc_tensor *t);

/* LoopTree: python.cpp:228-337 */
void ltc_looptree_free(ltc_looptree *lt);
int ltc_looptree_dump(const ltc_looptree *lt, char *buf, int cap);
int ltc_looptree_size(const ltc_looptree *lt);          /* number of tree refs */
int ltc_looptree_roots(const ltc_looptree *lt, int *out, int cap);
int ltc_looptree_children(const ltc_looptree *lt, int ref, int *out, int cap);
int ltc_looptree_parent(const ltc_looptree *lt, int ref); /* -1: a root; -2: error */
int ltc_looptree_is_loop(const ltc_looptree *lt, int ref);
int ltc_looptree_loop(const ltc_looptree *lt, int ref, int *var, int *var_depth,
                      int64_t *size, int64_t *tail);
int ltc_looptree_var_name(const ltc_looptree *lt, int var, char *buf, int cap);
int ltc_looptree_node(const ltc_looptree *lt, int ref);   /* IR node of a leaf */
/* vars (ids, as in ltc_looptree_loop) of a leaf's IR node, in its order */
int ltc_looptree_node_vars(const ltc_looptree *lt, int ref, int *out, int cap);
int ltc_looptree_node_op(const ltc_looptree *lt, int ref, char *buf, int cap);
int ltc_looptree_annotation(const ltc_looptree *lt, int ref, char *buf, int cap);
int ltc_looptree_num_inputs(const ltc_looptree *lt);
int ltc_looptree_num_outputs(const ltc_looptree *lt);
/* element counts of each input then each output buffer */
int ltc_looptree_io_numel(const ltc_looptree *lt, int64_t *out, int cap);
int64_t ltc_looptree_flops(const ltc_looptree *lt);
int ltc_looptree_trivially_parallel(const ltc_looptree *lt, int ref);

/* schedule transforms (src/frontends/mutate.cpp); each returns a new tree */
ltc_looptree *ltc_looptree_split(const ltc_looptree *lt, int ref, int64_t size);
ltc_looptree *ltc_looptree_merge(const ltc_looptree *lt, int ref);
ltc_looptree *ltc_looptree_swap_loops(const ltc_looptree *lt, int a, int b);
ltc_looptree *ltc_looptree_swap_nodes(const ltc_looptree *lt, int a, int b);
ltc_looptree *ltc_looptree_try_swap(const ltc_looptree *lt, int a, int b);
ltc_looptree *ltc_looptree_annotate(const ltc_looptree *lt, int ref, const char *annot);
ltc_looptree *ltc_looptree_unroll_inner(const ltc_looptree *lt, int amount);
ltc_looptree *ltc_looptree_maximize_reuse(const ltc_looptree *lt);
ltc_looptree *ltc_looptree_disable_reuse(const ltc_looptree *lt, int loop_ref, int node);
ltc_looptree *ltc_looptree_enable_reuse(const ltc_looptree *lt, int loop_ref, int node);
int ltc_looptree_map_ref(const ltc_looptree *new_lt, int old_ref, const ltc_looptree *old_lt);
int ltc_looptree_serialize(const ltc_looptree *lt, char *buf, int cap); /* IR text */
ltc_looptree *ltc_looptree_deserialize(const char *text);

/* compile & run: Backend::compile / Compiled::run (backend.h:18-68) */
ltc_compiled *ltc_compile(const ltc_looptree *lt, const char *backend);
/* mem: one pointer per input then per output (host or device memory) */
int ltc_compiled_run(const ltc_compiled *cc, void *const *mem, int n, int sync);
int ltc_compiled_dump(const ltc_compiled *cc, char *buf, int cap);
int ltc_compiled_int_property(const ltc_compiled *cc, const char *key, int64_t *out);
int ltc_compiled_string_property(const ltc_compiled *cc, const char *key, char *buf, int cap);
int ltc_compiled_name(const ltc_compiled *cc, char *buf, int cap);
void ltc_compiled_free(ltc_compiled *cc);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif

#endif /* LOOP_TOOL_C_H_ */
