"""The lazy frontend evaluating through the hip backend.

With `hip` as the default backend, `Tensor.numpy()` / `data()` compile the
program's LoopTree through Backend::compile, keep the Compiled in the
hash-keyed compilation cache and run it with Compiled::run on the tensors'
buffers -- the path of the reference's TensorImpl::data<T>()
(include/loop_tool/lazy.h:219-260) and populateCompilationCache
(src/frontends/lazy.cpp:430-438).  Three ways to hold the output:

  owned-cpu  output allocated by the default hardware, the CPU (malloc);
  pinned     default hardware "hip": pinned host memory, compatible = 1 | 2;
  bound      the caller's array bound as the output, as the reference's
             Python `numpy()` does (src/frontends/python.cpp:548-561).

Each result is compared with the same LoopTree run on the C oracle: bit-exact
for index-only programs (pad, concat, transpose, max-pool), <= 1e-5
normwise for FP32 arithmetic, and softmax within 1e-5 of float64 and within
the oracle's own float32 error of the oracle.
"""
import contextlib

import numpy as np
import pytest

from loop_tool_amd import programs as P

pytestmark = pytest.mark.gpu


def normwise(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


CASES = [
    ("matmul", dict(M=128, N=128, K=128), lambda s: [P.uniform(1, s[0]), P.uniform(2, s[1])]),
    ("softmax", dict(R=2048, Cn=8192), lambda s: [P.normal(3, s[0], 3.0)]),
    ("conv2d", dict(N=4, Ci=64, Co=64, H=28, W=28),
     lambda s: [P.normal(4, s[0], 576 ** -0.5), P.normal(5, s[1], 576 ** -0.5)]),
    ("maxpool2d", dict(N=8, C=64, H=112, W=112), lambda s: [P.normal(6, s[0])]),
    ("pad2d", dict(N=32, C=128, H=56, W=56), lambda s: [P.normal(7, s[0])]),
    ("concat2d", dict(B=4096, M0=3000, M1=5192), lambda s: [P.normal(8, s[0]), P.normal(9, s[1])]),
    ("transpose2d", dict(M=4096, N=8192), lambda s: [P.normal(10, s[0])]),
]
INDEX_ONLY = {"maxpool2d", "pad2d", "concat2d", "transpose2d"}


@contextlib.contextmanager
def defaults(L, backend, hardware):
    L.set_default_backend(backend)
    L.set_default_hardware(hardware)
    try:
        yield
    finally:
        L.set_default_backend("cpp")
        L.set_default_hardware("cpu")


def lazy_eval(L, name, kw, inputs, mode, backend="hip"):
    """Build the program over the input arrays and let the lazy frontend
    evaluate it on `backend`; returns (output, Compiled name, LoopTree)."""
    with defaults(L, backend, "hip" if mode == "pinned" else "cpu"):
        out, _ = getattr(P, name)(L, data=inputs, **kw)
        out.clear_cache()  # another test may have compiled this hash on cpp
        if mode == "bound":
            y = np.full(int(np.prod(out.sizes)), np.nan, np.float32)
            out.numpy_into(y)
        else:
            y = out.numpy()
        compiled = out.compiled.name
        tree = out.loop_tree
        out.clear_cache()
        del out
    return y.ravel(), compiled, tree


def oracle(L, tree, inputs):
    y = np.zeros(tree.io_numel()[len(inputs)], np.float32)
    L.compile(tree, "cpp").run([np.ascontiguousarray(a, np.float32) for a in inputs] + [y])
    return y


@pytest.mark.parametrize("mode", ["owned-cpu", "pinned", "bound"])
@pytest.mark.parametrize("name,kw,gen", CASES, ids=[c[0] for c in CASES])
def test_lazy_numpy_on_hip(lt, hip_dev, name, kw, gen, mode):
    shapes = getattr(P, name)(lt, **kw)[1]
    inputs = gen(shapes)
    y, compiled, tree = lazy_eval(lt, name, kw, inputs, mode)
    assert compiled == "hip"
    y_ref = oracle(lt, tree, inputs)
    f64 = P.reference_output(name, inputs).ravel()
    print(name, kw, mode, "hip-vs-oracle %.3g hip-vs-f64 %.3g oracle-vs-f64 %.3g" % (
        normwise(y, y_ref), normwise(y, f64), normwise(y_ref, f64)))
    if name in INDEX_ONLY:
        assert np.array_equal(y.view(np.uint32), y_ref.view(np.uint32))
        assert np.array_equal(y, f64.astype(np.float32))
    elif name == "softmax":
        assert normwise(y, f64) <= 1e-5
        assert normwise(y, y_ref) <= normwise(y_ref, f64) + 1e-5
    else:
        assert normwise(y, y_ref) <= 1e-5


def test_lazy_cache_reuses_hip_compiled(lt, hip_dev):
    """A second program with the same structure and sizes hits the
    hash-keyed compilation cache (lazy.cpp:430-438): the same Compiled runs
    again on the new program's buffers, and recomputes after new data."""
    a, b = P.uniform(1, (128, 128)), P.uniform(2, (128, 128))
    with defaults(lt, "hip", "cpu"):
        c1, _ = P.matmul(lt, 128, 128, 128, data=[a, b])
        c1.clear_cache()
        y1 = c1.numpy()
        k1 = c1.compiled.int_property("code_cached")
        a2 = P.uniform(3, (128, 128))
        c2, _ = P.matmul(lt, 128, 128, 128, data=[a2, b])
        assert hash(c2) == hash(c1)
        y2 = c2.numpy()
        assert c2.compiled.name == "hip"
        c1.clear_cache()
    assert k1 is not None
    assert normwise(y1, a.astype(np.float64) @ b) <= 1e-5
    assert normwise(y2, a2.astype(np.float64) @ b) <= 1e-5


def test_lazy_bf16_backend_conv(lt, hip_dev):
    """Default backend hip_bf16: the lazy conv runs with bf16 operands on
    MFMA; <= 1e-5 of float64 over the bf16-rounded operands, <= 1e-2 of
    float64 over the float operands."""
    kw = dict(N=2, Ci=64, Co=128, H=28, W=28)
    shapes = P.conv2d(lt, **kw)[1]
    inputs = [P.normal(4, shapes[0], 576 ** -0.5), P.normal(5, shapes[1], 576 ** -0.5)]
    y, compiled, _ = lazy_eval(lt, "conv2d", kw, inputs, "owned-cpu", backend="hip_bf16")
    assert compiled == "hip_bf16"

    def bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32)
    f64_bf = P.reference_output("conv2d", [bf16(x) for x in inputs]).ravel()
    f64 = P.reference_output("conv2d", inputs).ravel()
    assert normwise(y, f64_bf) <= 1e-5
    assert normwise(y, f64) <= 1e-2
