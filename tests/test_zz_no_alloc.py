"""The allocation-free launch (lth_run_static, include/loop_tool_hip.h): the
plan's scratch comes from the library's static device array instead of
device memory the Compiled allocates.

On the GPU it must give the same bits as the allocating entry point
(ltc_compiled_run on device buffers) for the same plan: the same kernels in
the same order, only the scratch addresses differ.  Two calls in a row, the
second with a smaller input, so the second sees the first's leftovers in the
shared array; the first is preceded by a call on NaN inputs, so stale
scratch contents would show.  On the CPU: plans needing more scratch than
the array holds, and Compileds whose code is not loaded yet, are refused
before anything touches the device.
"""
import ctypes as C

import numpy as np
import pytest

from loop_tool_amd import programs as P

STATIC_BYTES = 64 << 20


def _lib():
    from loop_tool_amd import lib
    L = lib.lib
    L.lth_run_static.restype = C.c_int
    L.lth_run_static.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_void_p]
    L.lth_last_error.restype = C.c_char_p
    return L


def run_static(cc, bufs, stream):
    L = _lib()
    arr = (C.c_void_p * len(bufs))(*[b.ptr if hasattr(b, "ptr") else b for b in bufs])
    rc = L.lth_run_static(cc.h, arr, len(bufs), stream)
    if rc != 0:
        raise RuntimeError(L.lth_last_error().decode())


def scratch_bytes(sizes):
    return sum((n * 4 + 255) // 256 * 256 for n in sizes)


def test_refuses_plans_beyond_the_static_scratch(lt):
    """bf16 matmul 4096 x 4096 x 8192 stages 2 x 64 MiB of operands: more than the
    64 MiB array (refused before any device call)."""
    from loop_tool_amd import hip
    out, _ = P.matmul(lt, 4096, 4096, 8192)
    cc = lt.compile(out.loop_tree, "hip_bf16")
    assert scratch_bytes(hip.scratch_sizes(cc)) > STATIC_BYTES
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        run_static(cc, [0, 0, 0], None)


def test_refuses_unloaded_code_and_short_buffer_lists(lt):
    out, _ = P.softmax(lt, 64, 512)
    cc = lt.compile(out.loop_tree, "hip")
    with pytest.raises(RuntimeError, match="expected 2 device buffers"):
        run_static(cc, [0], None)
    with pytest.raises(RuntimeError, match="not loaded"):
        run_static(cc, [0, 0], None)


CASES = [  # (program, bigger sizes, smaller sizes, backend, knobs)
    ("conv2d", dict(N=8, Ci=128, Co=128, H=28, W=28), dict(N=2, Ci=64, Co=64, H=14, W=14),
     "hip_bf16", ""),
    ("matmul", dict(M=1024, N=768, K=1536), dict(M=256, N=128, K=320), "hip_bf16", ""),
    ("softmax", dict(R=512, Cn=4096), dict(R=64, Cn=1000), "hip", "hip:plain"),
    ("conv2d", dict(N=4, Ci=64, Co=64, H=28, W=28), dict(N=1, Ci=32, Co=16, H=9, W=9),
     "hip", "contract=staged"),
]


def _inputs(name, shapes, poison=False):
    if name == "matmul":
        xs = [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    elif name == "softmax":
        xs = [P.normal(3, shapes[0], 3.0)]
    else:
        fan = shapes[1][1] * 9
        xs = [P.normal(4, shapes[0], fan ** -0.5), P.normal(5, shapes[1], fan ** -0.5)]
    return [np.full_like(x, np.nan) for x in xs] if poison else xs


@pytest.mark.gpu
@pytest.mark.parametrize("name,big,small,backend,knobs", CASES,
                         ids=["%s_%s" % (c[0], c[3]) for c in CASES])
def test_static_scratch_matches_allocating_run(lt, hip_dev, name, big, small, backend, knobs):
    from loop_tool_amd import hip
    stream = hip.Stream()
    for step, kw in enumerate((big, small)):
        out, shapes = getattr(P, name)(lt, **kw)
        tree = out.loop_tree
        if knobs:
            tree = tree.annotate(tree.roots[0], knobs)
        n_out = tree.io_numel()[len(shapes)]
        d_in = [hip.DeviceArray.from_numpy(x) for x in _inputs(name, shapes)]
        # the allocating entry point: Compiled-owned scratch
        ref = lt.compile(tree, backend)
        d_ref = hip.DeviceArray(n_out)
        ref.run([b.ptr for b in d_in + [d_ref]])
        want = d_ref.download()
        # a second Compiled of the same plan, loaded but owning no scratch
        cc = lt.compile(tree, backend)
        hip.compiled_stream(cc)
        sizes = hip.scratch_sizes(cc)
        print(name, kw, backend, knobs, "scratch bytes", scratch_bytes(sizes), flush=True)
        assert 0 < scratch_bytes(sizes) <= STATIC_BYTES
        d_out = hip.DeviceArray(n_out)
        if step == 0:  # leave NaN-derived contents in the shared scratch
            d_nan = [hip.DeviceArray.from_numpy(x) for x in _inputs(name, shapes, True)]
            run_static(cc, d_nan + [d_out], stream.h)
            stream.synchronize()
            assert np.isnan(d_out.download()).any()
        d_out.upload(np.full(n_out, 7.0, np.float32))
        run_static(cc, d_in + [d_out], stream.h)
        stream.synchronize()
        got = d_out.download()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, kw)
