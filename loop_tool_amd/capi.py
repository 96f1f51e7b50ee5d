"""ctypes binding of the loop_tool C ABI (include/loop_tool_c.h).

Mirrors the Python API of the reference's pybind11 module `loop_tool_py`
(src/frontends/python.cpp:86-698): Symbol, Expr, Size, Tensor (lazy), LoopTree
(inspection + schedule transforms), backends / set_default_backend, and
compiling a LoopTree with a named backend.  Any library exporting the C ABI can
be wrapped (the product library, or another build of the same API).
"""
import contextlib
import ctypes as C
import fcntl
import os

import numpy as np

_P = C.c_void_p


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


class LoopToolError(RuntimeError):
    pass


class Library:
    """One loaded implementation of the loop_tool C ABI."""

    def __init__(self, path, mode=C.RTLD_LOCAL, compile_lock=None):
        """compile_lock: path of a lock file held across every call that may
        compile.  The reference's C backend writes its code to fixed names
        (/tmp/fn_<n>.c and .so, src/backends/cpu/cpp.cpp:462-472), so two
        processes compiling through it at once can load each other's half
        written library; the lock serialises them."""
        self.path = path
        self.compile_lock_path = compile_lock
        self.lib = C.CDLL(path, mode=mode)
        L = self.lib
        s = lambda *a: _sig(L, *a)
        s("ltc_last_error", C.c_char_p)
        s("ltc_abi_version", C.c_int)
        s("ltc_backends", C.c_int, C.c_char_p, C.c_int)
        s("ltc_set_default_backend", C.c_int, C.c_char_p)
        s("ltc_default_backend", C.c_int, C.c_char_p, C.c_int)
        s("ltc_load_library", C.c_int, C.c_char_p)
        s("ltc_set_default_hardware", C.c_int, C.c_char_p)
        s("ltc_symbol_new", _P, C.c_char_p)
        s("ltc_symbol_free", None, _P)
        s("ltc_symbol_id", C.c_int, _P)
        s("ltc_symbol_name", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_expr_value", _P, C.c_int64)
        s("ltc_expr_symbol", _P, _P)
        s("ltc_expr_size", _P, _P)
        s("ltc_expr_op", _P, C.c_int, _P, _P)
        s("ltc_expr_free", None, _P)
        s("ltc_expr_dump", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_tensor_from_sizes", _P, C.POINTER(C.c_int64), C.c_int)
        s("ltc_tensor_from_symbols", _P, C.POINTER(_P), C.c_int)
        s("ltc_tensor_from_data", _P, _P, C.POINTER(C.c_int64), C.c_int)
        s("ltc_tensor_bind", C.c_int, _P, _P, C.POINTER(C.c_int64), C.c_int)
        s("ltc_tensor_free", None, _P)
        s("ltc_tensor_binary", _P, C.c_int, _P, _P)
        s("ltc_tensor_unary", _P, C.c_int, _P)
        s("ltc_tensor_reduce", _P, C.c_int, _P, C.POINTER(_P), C.c_int)
        s("ltc_tensor_as", _P, _P, C.POINTER(_P), C.c_int)
        s("ltc_tensor_to", _P, _P, C.POINTER(_P), C.c_int, C.POINTER(_P), C.POINTER(_P), C.c_int)
        s("ltc_tensor_pad", _P, _P, _P, C.c_int64, C.c_int64)
        s("ltc_tensor_transpose", _P, _P, C.POINTER(_P), C.c_int)
        s("ltc_tensor_flatten", _P, _P, C.POINTER(_P), C.c_int, _P)
        s("ltc_nn_convolve", _P, _P, _P, C.POINTER(_P), C.POINTER(_P), C.c_int, C.c_int)
        s("ltc_nn_maxpool", _P, _P, C.POINTER(_P), C.c_int, C.c_int, C.c_int)
        s("ltc_tensor_ndim", C.c_int, _P)
        s("ltc_tensor_shape", C.c_int, _P, C.POINTER(_P), C.c_int)
        s("ltc_tensor_sizes", C.c_int, _P, C.POINTER(C.c_int64), C.c_int)
        s("ltc_tensor_size", C.c_int64, _P, C.c_int)
        s("ltc_tensor_data", C.POINTER(C.c_float), _P)
        s("ltc_tensor_hash", C.c_uint64, _P)
        s("ltc_tensor_code", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_tensor_compile", C.c_int, _P)
        s("ltc_tensor_loop_tree", _P, _P)
        s("ltc_tensor_set_loop_tree", C.c_int, _P, _P)
        s("ltc_tensor_compiled", _P, _P)
        s("ltc_tensor_clear_cache", C.c_int, _P)
        s("ltc_tensor_force_recompute", C.c_int, _P)
        s("ltc_looptree_free", None, _P)
        s("ltc_looptree_dump", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_looptree_size", C.c_int, _P)
        s("ltc_looptree_roots", C.c_int, _P, C.POINTER(C.c_int), C.c_int)
        s("ltc_looptree_children", C.c_int, _P, C.c_int, C.POINTER(C.c_int), C.c_int)
        s("ltc_looptree_parent", C.c_int, _P, C.c_int)
        s("ltc_looptree_is_loop", C.c_int, _P, C.c_int)
        s("ltc_looptree_loop", C.c_int, _P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
          C.POINTER(C.c_int64), C.POINTER(C.c_int64))
        s("ltc_looptree_var_name", C.c_int, _P, C.c_int, C.c_char_p, C.c_int)
        s("ltc_looptree_node", C.c_int, _P, C.c_int)
        s("ltc_looptree_node_op", C.c_int, _P, C.c_int, C.c_char_p, C.c_int)
        s("ltc_looptree_node_vars", C.c_int, _P, C.c_int, C.POINTER(C.c_int), C.c_int)
        s("ltc_looptree_annotation", C.c_int, _P, C.c_int, C.c_char_p, C.c_int)
        s("ltc_looptree_num_inputs", C.c_int, _P)
        s("ltc_looptree_num_outputs", C.c_int, _P)
        s("ltc_looptree_io_numel", C.c_int, _P, C.POINTER(C.c_int64), C.c_int)
        s("ltc_looptree_flops", C.c_int64, _P)
        s("ltc_looptree_trivially_parallel", C.c_int, _P, C.c_int)
        s("ltc_looptree_split", _P, _P, C.c_int, C.c_int64)
        s("ltc_looptree_merge", _P, _P, C.c_int)
        s("ltc_looptree_swap_loops", _P, _P, C.c_int, C.c_int)
        s("ltc_looptree_swap_nodes", _P, _P, C.c_int, C.c_int)
        s("ltc_looptree_try_swap", _P, _P, C.c_int, C.c_int)
        s("ltc_looptree_annotate", _P, _P, C.c_int, C.c_char_p)
        s("ltc_looptree_unroll_inner", _P, _P, C.c_int)
        s("ltc_looptree_maximize_reuse", _P, _P)
        s("ltc_looptree_disable_reuse", _P, _P, C.c_int, C.c_int)
        s("ltc_looptree_enable_reuse", _P, _P, C.c_int, C.c_int)
        s("ltc_looptree_map_ref", C.c_int, _P, C.c_int, _P)
        s("ltc_looptree_serialize", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_looptree_deserialize", _P, C.c_char_p)
        s("ltc_compile", _P, _P, C.c_char_p)
        s("ltc_compiled_run", C.c_int, _P, C.POINTER(_P), C.c_int, C.c_int)
        s("ltc_compiled_dump", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_compiled_int_property", C.c_int, _P, C.c_char_p, C.POINTER(C.c_int64))
        s("ltc_compiled_string_property", C.c_int, _P, C.c_char_p, C.c_char_p, C.c_int)
        s("ltc_compiled_name", C.c_int, _P, C.c_char_p, C.c_int)
        s("ltc_compiled_free", None, _P)

    # -- helpers ---------------------------------------------------------
    @contextlib.contextmanager
    def compiling(self):
        if self.compile_lock_path is None:
            yield
            return
        with open(self.compile_lock_path, "a") as f:
            fcntl.flock(f, fcntl.LOCK_EX)
            try:
                yield
            finally:
                fcntl.flock(f, fcntl.LOCK_UN)

    def error(self):
        return self.lib.ltc_last_error().decode()

    def check_ptr(self, p):
        if not p:
            raise LoopToolError(self.error())
        return p

    def check_int(self, r):
        if r < 0:
            raise LoopToolError(self.error())
        return r

    def string(self, fn, *args):
        n = self.check_int(fn(*args, None, 0))
        buf = C.create_string_buffer(n + 1)
        self.check_int(fn(*args, buf, n + 1))
        return buf.value.decode()

    # -- module level API (python.cpp:87-110) ----------------------------
    def backends(self):
        return [b for b in self.string(self.lib.ltc_backends).split(",") if b]

    def set_default_backend(self, name):
        self.check_int(self.lib.ltc_set_default_backend(name.encode()))

    def get_default_backend(self):
        return self.string(self.lib.ltc_default_backend)

    def load_lib(self, path):
        self.check_int(self.lib.ltc_load_library(os.fspath(path).encode()))

    def set_default_hardware(self, name):
        self.check_int(self.lib.ltc_set_default_hardware(name.encode()))

    def Symbol(self, name="X"):
        return Symbol(self, name)

    def Size(self, sym):
        return Expr(self, self.check_ptr(self.lib.ltc_expr_size(sym.h)))

    def Expr(self, v):
        return Expr.of(self, v)

    def Tensor(self, *args, data=None):
        return Tensor.create(self, *args, data=data)

    def deserialize(self, text):
        return LoopTree(self, self.check_ptr(self.lib.ltc_looptree_deserialize(text.encode())))

    def compile(self, loop_tree, backend):
        with self.compiling():
            return Compiled(self, self.check_ptr(
                self.lib.ltc_compile(loop_tree.h, backend.encode())))

    def convolve(self, X, W, spatial, window, stride=1):
        return _derived(self, self.lib.ltc_nn_convolve(
            X.h, W.h, _handles(spatial), _handles(window), len(spatial), stride), X, W)

    def maxpool(self, X, spatial, k, stride=1):
        return _derived(self, self.lib.ltc_nn_maxpool(
            X.h, _handles(spatial), len(spatial), k, stride), X)


def _handles(objs):
    arr = (_P * max(len(objs), 1))()
    for i, o in enumerate(objs):
        arr[i] = o.h
    return arr


class _Handle:
    _free = None

    def __init__(self, lib, h):
        self.L = lib
        self.h = h

    def __del__(self):
        try:
            if self.h and self._free:
                getattr(self.L.lib, self._free)(self.h)
                self.h = None
        except Exception:
            pass


class Symbol(_Handle):
    _free = "ltc_symbol_free"

    def __init__(self, lib, name="X", handle=None):
        super().__init__(lib, handle or lib.check_ptr(lib.lib.ltc_symbol_new(name.encode())))

    @property
    def name(self):
        return self.L.string(self.L.lib.ltc_symbol_name, self.h)

    @property
    def id(self):
        return self.L.lib.ltc_symbol_id(self.h)

    def __eq__(self, o):
        return isinstance(o, Symbol) and o.id == self.id

    def __hash__(self):
        return hash(self.id)

    def expr(self):
        return Expr(self.L, self.L.check_ptr(self.L.lib.ltc_expr_symbol(self.h)))

    def __add__(self, o):
        return self.expr() + o

    def __radd__(self, o):
        return Expr.of(self.L, o) + self.expr()

    def __mul__(self, o):
        return self.expr() * o

    def __rmul__(self, o):
        return Expr.of(self.L, o) * self.expr()

    def __sub__(self, o):
        return self.expr() - o

    def __repr__(self):
        return "Symbol(%s#%d)" % (self.name, self.id)


class Expr(_Handle):
    _free = "ltc_expr_free"

    @staticmethod
    def of(lib, v):
        if isinstance(v, Expr):
            return v
        if isinstance(v, Symbol):
            return v.expr()
        return Expr(lib, lib.check_ptr(lib.lib.ltc_expr_value(int(v))))

    def _op(self, op, o):
        o = Expr.of(self.L, o)
        return Expr(self.L, self.L.check_ptr(self.L.lib.ltc_expr_op(op, self.h, o.h)))

    def __add__(self, o):
        return self._op(0, o)

    def __radd__(self, o):
        return Expr.of(self.L, o)._op(0, self)

    def __sub__(self, o):
        return self._op(1, o)

    def __mul__(self, o):
        return self._op(2, o)

    def __rmul__(self, o):
        return Expr.of(self.L, o)._op(2, self)

    def __floordiv__(self, o):
        return self._op(3, o)

    def __mod__(self, o):
        return self._op(4, o)

    def __neg__(self):
        return Expr(self.L, self.L.check_ptr(self.L.lib.ltc_expr_op(6, self.h, self.h)))

    def __repr__(self):
        return self.L.string(self.L.lib.ltc_expr_dump, self.h)


def _derived(L, h, *srcs):
    """A tensor computed from `srcs`: it keeps their bound host arrays alive,
    since the native graph holds only their addresses."""
    t = Tensor(L, L.check_ptr(h))
    keep = []
    for s in srcs:
        keep += getattr(s, "_keepalive", [])
    t._keepalive = keep
    return t


class Tensor(_Handle):
    _free = "ltc_tensor_free"

    @staticmethod
    def create(lib, *args, data=None):
        if data is not None:
            arr = data
            sizes = (C.c_int64 * len(arr.shape))(*arr.shape)
            t = Tensor(lib, lib.check_ptr(lib.lib.ltc_tensor_from_data(
                arr.ctypes.data, sizes, len(arr.shape))))
            t._keepalive = [arr]
            return t
        if args and all(isinstance(a, Symbol) for a in args):
            return Tensor(lib, lib.check_ptr(lib.lib.ltc_tensor_from_symbols(
                _handles(args), len(args))))
        sizes = [int(a) for a in args]
        arr = (C.c_int64 * max(len(sizes), 1))(*sizes)
        return Tensor(lib, lib.check_ptr(lib.lib.ltc_tensor_from_sizes(arr, len(sizes))))

    def _bin(self, op, o):
        return _derived(self.L, self.L.lib.ltc_tensor_binary(op, self.h, o.h), self, o)

    def _un(self, op):
        return _derived(self.L, self.L.lib.ltc_tensor_unary(op, self.h), self)

    def __add__(self, o):
        return self._bin(0, o)

    def __sub__(self, o):
        return self._bin(1, o)

    def __mul__(self, o):
        return self._bin(2, o)

    def __truediv__(self, o):
        return self._bin(3, o)

    def max(self, o):
        return self._bin(4, o)

    def min(self, o):
        return self._bin(5, o)

    def __or__(self, o):
        return self._bin(6, o)

    def exp(self):
        return self._un(0)

    def log(self):
        return self._un(1)

    def sqrt(self):
        return self._un(2)

    def abs(self):
        return self._un(3)

    def __neg__(self):
        return self._un(4)

    def reciprocal(self):
        return self._un(5)

    def _red(self, op, syms):
        return _derived(self.L, self.L.lib.ltc_tensor_reduce(
            op, self.h, _handles(syms), len(syms)), self)

    def sum(self, *syms):
        return self._red(0, syms)

    def prod(self, *syms):
        return self._red(1, syms)

    def max_reduce(self, *syms):
        return self._red(2, syms)

    def min_reduce(self, *syms):
        return self._red(3, syms)

    def as_(self, *syms):
        return _derived(self.L, self.L.lib.ltc_tensor_as(self.h, _handles(syms), len(syms)),
                        self)

    def to(self, syms, constraints=()):
        lhs = [Expr.of(self.L, c[0]) for c in constraints]
        rhs = [Expr.of(self.L, c[1]) for c in constraints]
        return _derived(self.L, self.L.lib.ltc_tensor_to(
            self.h, _handles(syms), len(syms), _handles(lhs), _handles(rhs), len(lhs)), self)

    def pad(self, sym, pre, post=None):
        post = pre if post is None else post
        return _derived(self.L, self.L.lib.ltc_tensor_pad(self.h, sym.h, pre, post), self)

    def transpose(self, *syms):
        return _derived(self.L, self.L.lib.ltc_tensor_transpose(
            self.h, _handles(syms), len(syms)), self)

    def flatten(self, syms, new_sym):
        return _derived(self.L, self.L.lib.ltc_tensor_flatten(
            self.h, _handles(syms), len(syms), new_sym.h), self)

    def bind(self, data, sizes=None):
        if data is not None:
            self._keepalive = getattr(self, "_keepalive", []) + [data]
            sizes = data.shape if sizes is None else sizes
        arr = (C.c_int64 * len(sizes))(*sizes)
        ptr = data.ctypes.data if data is not None else None
        self.L.check_int(self.L.lib.ltc_tensor_bind(self.h, ptr, arr, len(sizes)))
        return self

    def set_size(self, *sizes):
        return self.bind(None, sizes)

    @property
    def shape(self):
        n = self.L.lib.ltc_tensor_ndim(self.h)
        out = (_P * max(n, 1))()
        self.L.check_int(self.L.lib.ltc_tensor_shape(self.h, out, n))
        return [Symbol(self.L, handle=out[i]) for i in range(n)]

    @property
    def sizes(self):
        out = (C.c_int64 * 16)()
        n = self.L.check_int(self.L.lib.ltc_tensor_sizes(self.h, out, 16))
        return [out[i] for i in range(n)]

    def size(self, dim):
        r = self.L.lib.ltc_tensor_size(self.h, dim)
        if r < 0:
            raise LoopToolError(self.L.error())
        return r

    def numpy(self):
        sizes = self.sizes
        with self.L.compiling():
            p = self.L.check_ptr(self.L.lib.ltc_tensor_data(self.h))
        n = int(np.prod(sizes)) if sizes else 1
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(sizes).copy()

    def numpy_into(self, out):
        """Evaluate into a caller array, as the reference's `numpy()` does
        (python.cpp:548-561): the array is bound as the tensor's output
        buffer, then the program runs on the default backend."""
        assert out.dtype == np.float32 and out.flags["C_CONTIGUOUS"]
        sizes = self.sizes
        assert out.size == (int(np.prod(sizes)) if sizes else 1)
        self.bind(out, sizes)
        with self.L.compiling():
            self.L.check_ptr(self.L.lib.ltc_tensor_data(self.h))
        return out.reshape(sizes)

    def resolve(self):
        with self.L.compiling():
            self.L.check_ptr(self.L.lib.ltc_tensor_data(self.h))

    def clear_cache(self):
        """Drop this program's compiled and lowered cache entries
        (python.cpp:566)."""
        self.L.check_int(self.L.lib.ltc_tensor_clear_cache(self.h))

    def force_recompute(self):
        self.L.check_int(self.L.lib.ltc_tensor_force_recompute(self.h))

    @property
    def code(self):
        return self.L.string(self.L.lib.ltc_tensor_code, self.h)

    def compile(self):
        with self.L.compiling():
            self.L.check_int(self.L.lib.ltc_tensor_compile(self.h))

    @property
    def loop_tree(self):
        return LoopTree(self.L, self.L.check_ptr(self.L.lib.ltc_tensor_loop_tree(self.h)))

    def set(self, lt):
        self.L.check_int(self.L.lib.ltc_tensor_set_loop_tree(self.h, lt.h))

    @property
    def compiled(self):
        with self.L.compiling():
            return Compiled(self.L, self.L.check_ptr(self.L.lib.ltc_tensor_compiled(self.h)))

    def __hash__(self):
        return self.L.lib.ltc_tensor_hash(self.h)


class LoopTree(_Handle):
    _free = "ltc_looptree_free"

    def _new(self, h):
        return LoopTree(self.L, self.L.check_ptr(h))

    def dump(self):
        return self.L.string(self.L.lib.ltc_looptree_dump, self.h)

    __str__ = dump

    def _ints(self, fn, *args):
        n = self.L.check_int(fn(self.h, *args, None, 0))
        out = (C.c_int * max(n, 1))()
        fn(self.h, *args, out, n)
        return [out[i] for i in range(n)]

    @property
    def roots(self):
        return self._ints(self.L.lib.ltc_looptree_roots)

    def children(self, ref):
        return self._ints(self.L.lib.ltc_looptree_children, ref)

    def parent(self, ref):
        """The parent's ref, -1 for a root (LoopTree::parent, ir.h)."""
        r = self.L.lib.ltc_looptree_parent(self.h, ref)
        if r < -1:
            raise LoopToolError(self.L.error())
        return r

    def is_loop(self, ref):
        return bool(self.L.check_int(self.L.lib.ltc_looptree_is_loop(self.h, ref)))

    def loop(self, ref):
        v, d, s, t = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        self.L.check_int(self.L.lib.ltc_looptree_loop(self.h, ref, v, d, s, t))
        return dict(var=v.value, depth=d.value, size=s.value, tail=t.value)

    def var_name(self, v):
        return self.L.string(self.L.lib.ltc_looptree_var_name, self.h, v)

    def node_op(self, ref):
        return self.L.string(self.L.lib.ltc_looptree_node_op, self.h, ref)

    def node(self, ref):
        return self.L.check_int(self.L.lib.ltc_looptree_node(self.h, ref))

    def node_vars(self, ref):
        """Var ids of a leaf's IR node, in the node's order (the last is
        the contiguous dim of its value)."""
        return self._ints(self.L.lib.ltc_looptree_node_vars, ref)

    def annotation(self, ref):
        return self.L.string(self.L.lib.ltc_looptree_annotation, self.h, ref)

    @property
    def num_refs(self):
        return self.L.lib.ltc_looptree_size(self.h)

    @property
    def loops(self):
        return [r for r in self.walk() if self.is_loop(r)]

    def walk(self):
        out = []

        def rec(r):
            out.append(r)
            for c in self.children(r):
                rec(c)
        for r in self.roots:
            rec(r)
        return out

    @property
    def num_inputs(self):
        return self.L.lib.ltc_looptree_num_inputs(self.h)

    @property
    def num_outputs(self):
        return self.L.lib.ltc_looptree_num_outputs(self.h)

    def io_numel(self):
        out = (C.c_int64 * 64)()
        n = self.L.check_int(self.L.lib.ltc_looptree_io_numel(self.h, out, 64))
        return [out[i] for i in range(n)]

    def flops(self):
        return self.L.lib.ltc_looptree_flops(self.h)

    def trivially_parallel(self, ref):
        return bool(self.L.check_int(self.L.lib.ltc_looptree_trivially_parallel(self.h, ref)))

    def split(self, ref, size):
        return self._new(self.L.lib.ltc_looptree_split(self.h, ref, size))

    def merge(self, ref):
        return self._new(self.L.lib.ltc_looptree_merge(self.h, ref))

    def swap_loops(self, a, b):
        return self._new(self.L.lib.ltc_looptree_swap_loops(self.h, a, b))

    def swap_nodes(self, a, b):
        return self._new(self.L.lib.ltc_looptree_swap_nodes(self.h, a, b))

    def try_swap(self, a, b):
        return self._new(self.L.lib.ltc_looptree_try_swap(self.h, a, b))

    def annotate(self, ref, annot):
        return self._new(self.L.lib.ltc_looptree_annotate(self.h, ref, annot.encode()))

    def unroll_inner_loops(self, amount):
        return self._new(self.L.lib.ltc_looptree_unroll_inner(self.h, amount))

    def maximize_reuse(self):
        return self._new(self.L.lib.ltc_looptree_maximize_reuse(self.h))

    def disable_reuse(self, loop_ref, node):
        return self._new(self.L.lib.ltc_looptree_disable_reuse(self.h, loop_ref, node))

    def enable_reuse(self, loop_ref, node):
        return self._new(self.L.lib.ltc_looptree_enable_reuse(self.h, loop_ref, node))

    def map_ref(self, old_ref, old_lt):
        return self.L.check_int(self.L.lib.ltc_looptree_map_ref(self.h, old_ref, old_lt.h))

    def serialize(self):
        return self.L.string(self.L.lib.ltc_looptree_serialize, self.h)


class Compiled(_Handle):
    _free = "ltc_compiled_free"

    def run(self, buffers, sync=True):
        """buffers: numpy float32 arrays (host) or raw device pointers (int)."""
        ptrs = (_P * len(buffers))()
        for i, b in enumerate(buffers):
            if isinstance(b, int):
                ptrs[i] = b
            else:
                assert b.dtype == np.float32 and b.flags["C_CONTIGUOUS"]
                ptrs[i] = b.ctypes.data
        self.L.check_int(self.L.lib.ltc_compiled_run(self.h, ptrs, len(buffers), int(sync)))

    __call__ = run

    def dump(self):
        return self.L.string(self.L.lib.ltc_compiled_dump, self.h)

    @property
    def name(self):
        return self.L.string(self.L.lib.ltc_compiled_name, self.h)

    def int_property(self, key):
        v = C.c_int64()
        found = self.L.check_int(self.L.lib.ltc_compiled_int_property(self.h, key.encode(), v))
        return v.value if found else None

    def string_property(self, key):
        return self.L.string(self.L.lib.ltc_compiled_string_property, self.h, key.encode())
