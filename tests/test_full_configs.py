"""GPU parity of the benchmark configurations at their full sizes, with the
plans the benchmark times.

  * conv2d BF16 N=32, C 128 -> 128, 56 x 56, 3 x 3 pad 1, default knobs: the
    halo tile of 128 x 256 x 32 with one halo buffer and 8 waves.  A sample
    of 122k outputs -- every border pixel of 256 output planes plus 65,536
    uniformly drawn outputs -- against float64 of the bf16-rounded operands
    (<= 1e-5 normwise) and of the float operands (<= 1e-2).
  * row softmax 8192 x 8192: every row against the C oracle and float64,
    and each row's sum against 1.
  * the 512 uneven GEMMs (M, N, K log-uniform in [32, 2048], seed 6): the
    grouped graph replay bitwise against one launch per GEMM, and sampled
    rows of every GEMM against float64 of the rounded operands.
  * FP16 operands (`mfma_f16`) through the backend, matmul and conv.
"""
import numpy as np
import pytest

from loop_tool_amd import programs as P

pytestmark = pytest.mark.gpu


def normwise(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def bf16(a):
    u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def f16(a):
    return a.astype(np.float16).astype(np.float32)


conv_sample_indices = P.conv_sample_indices
conv_f64_at = P.conv_f64_at


def test_conv2d_bf16_benchmark_plan(lt, hip_dev):
    N, C, H = 32, 128, 56
    out, shapes = P.conv2d(lt, N, C, C, H, H)
    std = 1.0 / np.sqrt(C * 9)
    X, Wt = P.normal(4, shapes[0], std), P.normal(5, shapes[1], std)
    cc = lt.compile(out.loop_tree, "hip_bf16")
    desc = cc.dump().splitlines()[0]
    print(desc)
    assert cc.string_property("kinds") == "stage,contract_bf16"
    assert "tile 128x256x32 pipelined, 8 waves" in desc, desc
    assert "Q halo of" in desc and "(one buffer)" in desc, desc
    dx = hip_dev.DeviceArray.from_numpy(X)
    dw = hip_dev.DeviceArray.from_numpy(Wt)
    dy = hip_dev.DeviceArray(N * C * H * H)
    dy.upload(np.full(N * C * H * H, np.nan, np.float32))
    cc.run([dx.ptr, dw.ptr, dy.ptr])
    y = dy.download((N, C, H, H))
    assert np.isfinite(y).all()
    idx = conv_sample_indices(N, C, H, H)
    assert len(idx) >= 65536 + 2 * C * (4 * H - 4)
    got = y[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
    f64_bf = conv_f64_at(bf16(X), bf16(Wt), idx)
    f64 = conv_f64_at(X, Wt, idx)
    e_bf, e = normwise(got, f64_bf), normwise(got, f64)
    print("conv bf16 %d sampled outputs: vs f64(bf16 operands) %.3g, vs f64 %.3g" % (
        len(idx), e_bf, e))
    assert e_bf <= 1e-5
    assert e <= 1e-2
    # the same plan again on the same buffers: the same bits
    y1 = y.copy()
    cc.run([dx.ptr, dw.ptr, dy.ptr])
    assert np.array_equal(dy.download((N, C, H, H)).view(np.uint32), y1.view(np.uint32))


def test_conv2d_bf16_direct_halo_bit_identical(lt, hip_dev):
    """The benchmark conv with the halo read from the float source
    (halo=direct: only the weights are staged) gives the bits of the staged
    halo plan."""
    N, C, H = 32, 128, 56
    out, shapes = P.conv2d(lt, N, C, C, H, H)
    std = 1.0 / np.sqrt(C * 9)
    X, Wt = P.normal(4, shapes[0], std), P.normal(5, shapes[1], std)
    dx = hip_dev.DeviceArray.from_numpy(X)
    dw = hip_dev.DeviceArray.from_numpy(Wt)
    ys = []
    for knobs in ("", "halo=direct"):
        t = out.loop_tree
        if knobs:
            t = t.annotate(t.roots[0], knobs)
        cc = lt.compile(t, "hip_bf16")
        assert ("from the float source" in cc.dump().splitlines()[0]) == bool(knobs)
        dy = hip_dev.DeviceArray(N * C * H * H)
        cc.run([dx.ptr, dw.ptr, dy.ptr])
        ys.append(dy.download())
    assert np.array_equal(ys[0].view(np.uint32), ys[1].view(np.uint32))


def test_softmax_8192_every_row(lt, hip_dev):
    R = Cn = 8192
    out, shapes = P.softmax(lt, R, Cn)
    x = P.normal(3, shapes[0], 3.0)
    tree = out.loop_tree
    y_or = np.zeros(R * Cn, np.float32)
    lt.compile(tree, "cpp").run([x, y_or])
    cc = lt.compile(tree, "hip")
    assert cc.string_property("kinds") == "rowreduce"
    y = np.zeros(R * Cn, np.float32)
    cc.run([x, y])
    xd = x.astype(np.float64)
    e = np.exp(xd - xd.max(axis=1, keepdims=True))
    f64 = (e / e.sum(axis=1, keepdims=True)).ravel()
    del e, xd
    e_hip, e_or = normwise(y, f64), normwise(y_or, f64)
    print("softmax 8192^2: hip-vs-f64 %.3g oracle-vs-f64 %.3g hip-vs-oracle %.3g" % (
        e_hip, e_or, normwise(y, y_or)))
    assert e_hip <= 1e-5
    assert normwise(y, y_or) <= e_or + 1e-5
    rows = y.reshape(R, Cn).astype(np.float64).sum(axis=1)
    assert np.abs(rows - 1).max() < 1e-4


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_batched_512_uneven_gemms(hip_dev, dtype):
    from loop_tool_amd import batched as bat
    dims = bat.shapes(512, seed=6)
    b = bat.Batch(dims, dtype=dtype, seed=6)
    b.run_plain(bat.DEFAULT_CFG)
    b.synchronize()
    want = [b.result(i) for i in range(len(dims))]
    for c in b.C:
        c.fill_zero()
    b.capture()
    for rep in range(2):
        b.replay()
        b.synchronize()
        for i in range(len(dims)):
            assert np.array_equal(b.result(i).view(np.uint32), want[i].view(np.uint32)), (
                rep, i, dims[i])
    r = (lambda a: bf16(a)) if dtype == "bf16" else f16
    rng = np.random.RandomState(13)
    worst = worst16 = 0.0
    for i, (m, n, k) in enumerate(dims):
        a, bb = b.host[i]
        rows = np.sort(rng.choice(m, min(m, 48), replace=False))
        got = want[i][rows].astype(np.float64)
        ref = a[rows].astype(np.float64) @ bb.astype(np.float64)
        ref16 = r(a[rows]).astype(np.float64) @ r(bb).astype(np.float64)
        e, e16 = normwise(got, ref), normwise(got, ref16)
        worst, worst16 = max(worst, e), max(worst16, e16)
        assert e16 <= 1e-5, (i, dims[i], e16)
        assert e <= 1e-2, (i, dims[i], e)
    print("512 GEMMs %s: worst vs f64 %.3g, worst vs f64 of rounded operands %.3g" % (
        dtype, worst, worst16))


@pytest.mark.parametrize("name,kw", [("matmul", dict(M=512, N=384, K=640)),
                                     ("matmul", dict(M=37, N=53, K=29)),
                                     ("conv2d", dict(N=4, Ci=64, Co=128, H=28, W=28)),
                                     ("conv2d", dict(N=2, Ci=16, Co=72, H=15, W=15, stride=2))])
@pytest.mark.parametrize("knobs", ["mfma_f16", "mfma_f16 igemm=pipe", "mfma_f16 contract=gather"])
def test_f16_operands_through_backend(lt, hip_dev, name, kw, knobs):
    """FP16 MFMA operands with float32 accumulation, selected by the
    `mfma_f16` annotation on the plain `hip` backend: <= 1e-5 of float64 over
    the fp16-rounded operands."""
    out, shapes = getattr(P, name)(lt, **kw)
    if name == "matmul":
        inputs = [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    else:
        fan = shapes[1][1] * 9
        inputs = [P.normal(4, shapes[0], fan ** -0.5), P.normal(5, shapes[1], fan ** -0.5)]
    t = out.loop_tree
    t = t.annotate(t.roots[0], knobs)
    cc = lt.compile(t, "hip")
    assert cc.string_property("kinds").split(",")[-1] == "contract_f16"
    y = np.zeros(t.io_numel()[len(inputs)], np.float32)
    cc.run(inputs + [y])
    rk = {k: kw[k] for k in ("stride",) if k in kw}
    f64_16 = P.reference_output(name, [f16(a) for a in inputs], **rk).ravel()
    f64 = P.reference_output(name, inputs, **rk).ravel()
    print(name, kw, knobs, "vs f64(f16 operands) %.3g vs f64 %.3g" % (
        normwise(y, f64_16), normwise(y, f64)))
    assert normwise(y, f64_16) <= 1e-5
    assert normwise(y, f64) <= 1e-2
