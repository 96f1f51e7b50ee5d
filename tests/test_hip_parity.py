"""GPU parity of the hip backend against the C oracle ("cpp", a restatement of
the reference C backend) on the same seeded inputs, through the C ABI.

Bars: bit-exact for the plain (pointwise) lowering of + - * / max min and
index-only nests; <= 1e-5 normwise relative for FP32 arithmetic nests lowered
with a different accumulation order; every result is also compared with a
float64 evaluation of the same program and its error is printed.
"""
import numpy as np
import pytest

from loop_tool_amd import programs as P

pytestmark = pytest.mark.gpu


def run_backend(L, out, shapes, inputs, backend, annotate=None):
    lt = out.loop_tree
    if annotate:
        lt = lt.annotate(lt.roots[0], annotate)
    cc = L.compile(lt, backend)
    n_out = lt.io_numel()[len(inputs)]
    y = np.zeros(n_out, np.float32)
    cc.run([np.ascontiguousarray(a, np.float32) for a in inputs] + [y])
    return y, cc


def normwise(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


CASES = [
    ("matmul", dict(M=128, N=128, K=128), lambda s: [P.uniform(1, s[0]), P.uniform(2, s[1])]),
    ("matmul", dict(M=37, N=53, K=29), lambda s: [P.uniform(1, s[0]), P.uniform(2, s[1])]),
    ("softmax", dict(R=64, Cn=1000), lambda s: [P.normal(3, s[0], 3.0)]),
    ("conv2d", dict(N=2, Ci=8, Co=16, H=12, W=12), lambda s: [P.normal(4, s[0]), P.normal(5, s[1])]),
    ("maxpool2d", dict(N=2, C=3, H=16, W=16), lambda s: [P.normal(6, s[0])]),
]


@pytest.mark.parametrize("name,kw,gen", CASES, ids=[c[0] + str(i) for i, c in enumerate(CASES)])
def test_plain_hip_bit_exact_vs_oracle(lt, hip_dev, name, kw, gen):
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = gen(shapes)
    y_ref, _ = run_backend(lt, out, shapes, inputs, "cpp")
    y_hip, cc = run_backend(lt, out, shapes, inputs, "hip", annotate="hip:plain")
    f64 = P.reference_output(name, inputs).ravel()
    print(name, kw, "oracle err vs f64 %.3g, hip err vs f64 %.3g" % (
        normwise(y_ref, f64), normwise(y_hip, f64)))
    if name == "softmax":
        # exp is double libm on the host and ocml on the device: equal after
        # rounding to float except in double-rounding corner cases
        assert normwise(y_hip, y_ref) <= 1e-6
    else:
        assert np.array_equal(y_hip.view(np.uint32), y_ref.view(np.uint32))


@pytest.mark.parametrize("R,Cn", [(512, 8192), (100, 1000), (33, 130), (8, 65536)])
def test_rowreduce_softmax_vs_oracle(lt, hip_dev, R, Cn):
    """Default hip lowering of softmax (row-cooperative kernel): tree-ordered
    sums, so the bar is 1e-5 normwise relative to the C oracle."""
    out, shapes = P.softmax(lt, R, Cn)
    x = P.normal(3, shapes[0], 3.0)
    y_ref, _ = run_backend(lt, out, shapes, [x], "cpp")
    y_hip, cc = run_backend(lt, out, shapes, [x], "hip")
    f64 = P.reference_output("softmax", [x]).ravel()
    e_or, e_hip = normwise(y_ref, f64), normwise(y_hip, f64)
    print("softmax %dx%d kinds=%s oracle-vs-f64 %.3g hip-vs-f64 %.3g hip-vs-oracle %.3g" % (
        R, Cn, cc.string_property("kinds"), e_or, e_hip, normwise(y_hip, y_ref)))
    # the oracle sums each row sequentially in float32 and is itself off the
    # float64 result by e_or (2e-5 at 8192 columns); the hip kernel sums in a
    # tree and is closer to float64.  Bar: within 1e-5 of float64, and within
    # the oracle's own error (+1e-5) of the oracle.
    assert e_hip <= 1e-5
    assert normwise(y_hip, y_ref) <= e_or + 1e-5
    rows = y_hip.reshape(R, Cn).astype(np.float64).sum(axis=1)
    assert np.abs(rows - 1).max() < 1e-4  # rows of a softmax sum to 1


CONTRACT = [
    ("matmul", dict(M=128, N=128, K=128)),
    ("matmul", dict(M=37, N=53, K=29)),
    ("matmul", dict(M=300, N=200, K=700)),
    ("conv2d", dict(N=2, Ci=8, Co=16, H=12, W=12)),
    ("conv2d", dict(N=3, Ci=32, Co=40, H=14, W=14)),
    ("matmul", dict(M=300, N=200, K=704)),
    ("conv2d", dict(N=2, Ci=16, Co=72, H=15, W=15, stride=2)),
    ("conv2d", dict(N=4, Ci=64, Co=128, H=28, W=28)),
    ("conv2d", dict(N=1, Ci=32, Co=8, H=9, W=12, pad=2, k=5)),
    ("conv2d", dict(N=1, Ci=40, Co=24, H=5, W=300)),
    ("bmm", dict(Bn=3, M=100, N=72, K=96)),
]


def ref_kw(name, kw):
    return {k: kw[k] for k in ("pad", "stride") if name == "conv2d" and k in kw}


def contract_inputs(name, shapes):
    if name in ("matmul", "bmm"):
        return [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    fan = shapes[1][1] * shapes[1][2] * shapes[1][3]
    return [P.normal(4, shapes[0], fan ** -0.5), P.normal(5, shapes[1], fan ** -0.5)]


@pytest.mark.parametrize("name,kw", CONTRACT, ids=["%s%d" % (c[0], i) for i, c in enumerate(CONTRACT)])
def test_contraction_f32_vs_oracle(lt, hip_dev, name, kw):
    """FP32 contraction on MFMA (v_mfma_f32_32x32x2_f32): <= 1e-5 normwise vs
    the C oracle; the error of both against float64 is printed."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    y_ref, _ = run_backend(lt, out, shapes, inputs, "cpp")
    y_hip, cc = run_backend(lt, out, shapes, inputs, "hip")
    assert cc.string_property("kinds") == "contract_f32"
    f64 = P.reference_output(name, inputs, **ref_kw(name, kw)).ravel()
    print(name, kw, "oracle-vs-f64 %.3g hip-vs-f64 %.3g hip-vs-oracle %.3g" % (
        normwise(y_ref, f64), normwise(y_hip, f64), normwise(y_hip, y_ref)))
    assert normwise(y_hip, y_ref) <= 1e-5


@pytest.mark.parametrize("name,kw", CONTRACT, ids=["%s%d" % (c[0], i) for i, c in enumerate(CONTRACT)])
@pytest.mark.parametrize("knobs", ["contract=staged", "contract=staged tile=128x128",
                                   "contract=staged tile=64x128 waves=8",
                                   "contract=staged stage=flat"])
def test_contraction_f32_staged_vs_oracle(lt, hip_dev, name, kw, knobs):
    """FP32 contraction on staged float operands (one staging launch, then
    the pipelined float tile on v_mfma_f32_32x32x2_f32): <= 1e-5 normwise vs
    the C oracle and vs float64."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    y_ref, _ = run_backend(lt, out, shapes, inputs, "cpp")
    y_hip, cc = run_backend(lt, out, shapes, inputs, "hip", annotate=knobs)
    assert cc.string_property("kinds") == "stage,contract_f32"
    f64 = P.reference_output(name, inputs, **ref_kw(name, kw)).ravel()
    print(name, kw, knobs, "hip-vs-oracle %.3g hip-vs-f64 %.3g" % (
        normwise(y_hip, y_ref), normwise(y_hip, f64)))
    assert normwise(y_hip, y_ref) <= 1e-5
    assert normwise(y_hip, f64) <= 1e-5


@pytest.mark.parametrize("name,kw", CONTRACT, ids=["%s%d" % (c[0], i) for i, c in enumerate(CONTRACT)])
def test_contraction_bf16_vs_f64(lt, hip_dev, name, kw):
    """BF16 operands on MFMA with FP32 accumulation: <= 1e-2 normwise vs a
    float64 evaluation, and exactly the float64 result of the bf16-rounded
    operands up to float32 accumulation error (<= 1e-5)."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    y_hip, cc = run_backend(lt, out, shapes, inputs, "hip_bf16")
    assert cc.string_property("kinds").split(",")[-1] == "contract_bf16"
    f64 = P.reference_output(name, inputs, **ref_kw(name, kw)).ravel()
    # round operands to bf16 (round to nearest even) and evaluate in float64
    def bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32)
    f64_bf = P.reference_output(name, [bf16(a) for a in inputs], **ref_kw(name, kw)).ravel()
    e, e_bf = normwise(y_hip, f64), normwise(y_hip, f64_bf)
    print(name, kw, "bf16 hip-vs-f64 %.3g hip-vs-f64(bf16 operands) %.3g" % (e, e_bf))
    assert e <= 1e-2
    assert e_bf <= 1e-5


STAGED = CONTRACT


@pytest.mark.parametrize("name,kw", STAGED, ids=["%s%d" % (c[0], i) for i, c in enumerate(STAGED)])
@pytest.mark.parametrize("tile", ["", "tile=64x64", "tile=128x64", "tile=64x128", "tile=128x256",
                                  "tile=128x256 bk=64", "tile=64x64 bk=64", "igemm=plain",
                                  "tile=128x256 bk=64 igemm=plain", "contract=gather",
                                  "stage=flat", "waves=8", "waves=2", "lds=swz", "lds=pad",
                                  "tile=256x256 bk=64 waves=8 lds=swz", "kb=outer",
                                  "tile=256x256 bk=32 lds=dl", "tile=64x64 bk=64 lds=dl"])
def test_contraction_bf16_staged_tiles(lt, hip_dev, name, kw, tile):
    """Staged-operand lowering (both operands permuted -- k-blocked or flat --,
    zero-haloed and rounded to bf16 by one staging launch, then the
    implicit-GEMM tile, pipelined or single-stage, with 2, 4 or 8 waves and
    padded or XOR-permuted LDS rows filled through registers or by
    direct-to-LDS loads) at every tile shape, and the gathered
    lowering, against float64 of the bf16-rounded operands."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    t = out.loop_tree
    if tile:
        t = t.annotate(t.roots[0], tile)
    cc = lt.compile(t, "hip_bf16")
    kinds = cc.string_property("kinds")
    assert kinds == ("contract_bf16" if tile == "contract=gather" else
                     "stage,contract_bf16"), kinds
    y = np.zeros(t.io_numel()[len(inputs)], np.float32)
    cc.run(inputs + [y])

    def bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32)
    f64_bf = P.reference_output(name, [bf16(a) for a in inputs], **ref_kw(name, kw)).ravel()
    e_bf = normwise(y, f64_bf)
    print(name, kw, tile, "hip-vs-f64(bf16 operands) %.3g" % e_bf)
    assert e_bf <= 1e-5


HALO = [
    ("conv2d", dict(N=4, Ci=64, Co=128, H=28, W=28)),
    ("conv2d", dict(N=3, Ci=128, Co=40, H=14, W=14)),
    ("conv2d", dict(N=2, Ci=64, Co=72, H=15, W=15, stride=2)),
    ("conv2d", dict(N=1, Ci=64, Co=8, H=9, W=12, pad=2, k=5)),
    ("conv2d", dict(N=1, Ci=40, Co=24, H=5, W=300)),
    ("conv2d", dict(N=2, Ci=96, Co=64, H=7, W=7, pad=0)),
]


@pytest.mark.parametrize("name,kw", HALO, ids=["%s%d" % (c[0], i) for i, c in enumerate(HALO)])
@pytest.mark.parametrize("tile", ["", "tile=128x128", "tile=64x128", "tile=128x256 waves=4",
                                  "tile=128x512"])
def test_conv_bf16_halo_tile(lt, hip_dev, name, kw, tile):
    """The halo form of the staged tile (the Q rows of all window taps of a
    channel block loaded once into LDS, in two buffers or one, from a staged
    16-bit copy or straight from the float source) computes the same K order
    as the
    pipelined tile over the same k-blocked staging (blocks outer, taps
    inner), with LDS filled through registers or by direct-to-LDS loads: the
    results must be bit-identical, and within float32
    accumulation error of float64 over the bf16-rounded operands.  Covers
    partial tiles, stride 2, a 5x5 window, rows longer than a tile, pad 0
    and channel counts that are not a multiple of the K block."""
    if tile == "tile=128x512" and kw.get("W") == 300:
        pytest.skip("two halo buffers of 1122 rows exceed the LDS of a CU")
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    ys = []
    for knobs in ("igemm=pipe stage=blocked kb=outer", "igemm=halo", "igemm=halo1",
                  "igemm=halo halo=direct", "igemm=halo1 halo=direct", "igemm=halo1 lds=dl",
                  "igemm=pipe stage=blocked kb=outer lds=swz"):
        if (knobs == "igemm=halo halo=direct" and tile == "tile=128x256 waves=4"
                and kw.get("W") == 300):
            # two halo buffers of 864 rows plus the float source's landing
            # slots exceed the LDS of a CU (the one-buffer form runs)
            continue
        t = out.loop_tree
        t = t.annotate(t.roots[0], (knobs + " " + tile).strip())
        cc = lt.compile(t, "hip_bf16")
        assert cc.string_property("kinds") == "stage,contract_bf16"
        assert ("Q halo" in cc.dump().splitlines()[0]) == knobs.startswith("igemm=halo")
        assert ("from the float source" in cc.dump().splitlines()[0]) == ("direct" in knobs)
        y = np.zeros(t.io_numel()[len(inputs)], np.float32)
        cc.run(inputs + [y])
        ys.append(y)
    for y in ys[1:]:
        assert np.array_equal(y.view(np.uint32), ys[0].view(np.uint32))

    def bf16(a):
        u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32)
    f64_bf = P.reference_output(name, [bf16(a) for a in inputs], **ref_kw(name, kw)).ravel()
    assert normwise(ys[0], f64_bf) <= 1e-5


@pytest.mark.parametrize("name,kw", HALO, ids=["%s%d" % (c[0], i) for i, c in enumerate(HALO)])
def test_plane_staging_rows_bit_identical(lt, hip_dev, name, kw):
    """srows=N sets how many rows of a source plane one workgroup of the
    plane-staging pass handles; the staged operands, and so the results, are
    the same bits for every N."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    ys = []
    for knobs in ("", "srows=1", "srows=3", "srows=64"):
        t = out.loop_tree
        if knobs:
            t = t.annotate(t.roots[0], knobs)
        y = np.zeros(t.io_numel()[len(inputs)], np.float32)
        lt.compile(t, "hip_bf16").run(inputs + [y])
        ys.append(y)
    for y in ys[1:]:
        assert np.array_equal(y.view(np.uint32), ys[0].view(np.uint32))


@pytest.mark.parametrize("name,kw", CONTRACT, ids=["%s%d" % (c[0], i) for i, c in enumerate(CONTRACT)])
@pytest.mark.parametrize("backend", ["hip", "hip_bf16"])
def test_ksplit_tile_vs_whole_tile(lt, hip_dev, name, kw, backend):
    """The gathered tile with each K tile split over four waves (32 x 32
    tiles, partial sums added in wave order) against the same nest on 64 x
    64 tiles: within float32 rounding of each other (<= 1e-5 normwise), the
    float path within 1e-5 of the C oracle, 16-bit operands within 1e-5 of
    float64 of the rounded operands; two runs give the same bits."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    extra = " contract=gather" if backend == "hip_bf16" else ""
    y1, cc1 = run_backend(lt, out, shapes, inputs, backend, annotate="ksplit=1" + extra)
    assert "split over 4 waves" in cc1.dump().splitlines()[0]
    y0, cc0 = run_backend(lt, out, shapes, inputs, backend, annotate="ksplit=0" + extra)
    assert "split over" not in cc0.dump().splitlines()[0]
    assert normwise(y1, y0) <= 1e-5
    if backend == "hip":
        y_or, _ = run_backend(lt, out, shapes, inputs, "cpp")
        assert normwise(y1, y_or) <= 1e-5
    else:
        f64 = P.reference_output(name, [_bf16_round(a) for a in inputs], **ref_kw(name, kw)).ravel()
        assert normwise(y1, f64) <= 1e-5
    y2, _ = run_backend(lt, out, shapes, inputs, backend, annotate="ksplit=1" + extra)
    assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))


def test_pointwise_past_2_to_the_30_elements(lt, hip_dev):
    """Buffers of more than 2^30 elements take the 64-bit index form of the
    pointwise lowering (32-bit offsets are used only below that size): a
    broadcast add over 32769 x 32768 floats must equal numpy's float32 add bit
    for bit everywhere, including past element 2^31 / 4 of the output."""
    M, N = 32769, 32768
    assert M * N > 1 << 30
    out, shapes = P.add_broadcast(lt, M, N)
    a = ((np.arange(M * N, dtype=np.int64) % 4093) - 2046).astype(np.float32) * 0.25
    b = (np.arange(N, dtype=np.int64) % 17).astype(np.float32) * 0.5
    t = out.loop_tree
    cc = lt.compile(t, "hip")
    assert cc.string_property("kinds") == "pointwise"
    y = np.zeros(M * N, np.float32)
    cc.run([a, b, y])
    expect = (a.reshape(M, N) + b).ravel()
    assert np.array_equal(y, expect)


def run_plan(lt, out, inputs, knobs, backend="hip"):
    t = out.loop_tree
    if knobs:
        t = t.annotate(t.roots[0], knobs)
    cc = lt.compile(t, backend)
    y = np.zeros(t.io_numel()[len(inputs)], np.float32)
    cc.run(inputs + [y])
    return y, cc


def test_streamk_float_matmul_vs_f64(lt, hip_dev):
    """576 tiles on 512 workgroup slots: the float tile runs stream-K (every
    workgroup an equal share of all tiles' K tiles, tiles split between two
    workgroups accumulated with float atomics into the zeroed output).
    <= 1e-5 normwise vs float64, and two runs give the same bits."""
    out, shapes = P.matmul(lt, 3072, 3072, 256)
    inputs = [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    y, cc = run_plan(lt, out, inputs, "")
    assert "stream-K" in cc.dump().splitlines()[0]
    f64 = P.reference_output("matmul", inputs).ravel()
    assert normwise(y, f64) <= 1e-5
    y2, _ = run_plan(lt, out, inputs, "")
    assert np.array_equal(y.view(np.uint32), y2.view(np.uint32))
    y0, cc0 = run_plan(lt, out, inputs, "streamk=0")
    assert "stream-K" not in cc0.dump().splitlines()[0]
    assert normwise(y, y0) <= 1e-5


def test_streamk_float_conv(lt, hip_dev):
    """The float conv of the benchmark shape (784 tiles of 128 x 128): the
    stream-K plan against the one-tile-per-workgroup plan (<= 1e-5
    normwise, the float32 bar; they differ only in how split tiles' two
    partial sums are rounded), repeatable to the bit, and against float64 on
    a sample of output channels."""
    out, shapes = P.conv2d(lt, 32, 128, 128, 56, 56)
    std = 1.0 / np.sqrt(128 * 9)
    inputs = [P.normal(4, shapes[0], std), P.normal(5, shapes[1], std)]
    y, cc = run_plan(lt, out, inputs, "")
    assert "stream-K" in cc.dump().splitlines()[0]
    y2, _ = run_plan(lt, out, inputs, "")
    assert np.array_equal(y.view(np.uint32), y2.view(np.uint32))
    y0, _ = run_plan(lt, out, inputs, "streamk=0")
    assert normwise(y, y0) <= 1e-5
    # float64 for 8 output channels spread over the range
    co = np.arange(0, 128, 16)
    f64 = P.reference_output("conv2d", [inputs[0], inputs[1][co]])
    got = y.reshape(32, 128, 56, 56)[:, co]
    assert normwise(got.ravel(), f64.ravel()) <= 1e-5


def test_streamk_whole_tiles_only_bit_identical(lt, hip_dev):
    """1024 tiles of 64 x 64 over 512 workgroups: every range boundary falls
    between tiles, so forced stream-K shares no tile (no zeroing, no atomic
    adds) and equals the one-tile-per-workgroup plan to the bit."""
    out, shapes = P.matmul(lt, 2048, 2048, 96)
    inputs = [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    y, cc = run_plan(lt, out, inputs, "contract=staged tile=64x64 streamk=1")
    assert "stream-K" in cc.dump().splitlines()[0] and "split_[" not in cc.dump()
    y0, _ = run_plan(lt, out, inputs, "contract=staged tile=64x64 streamk=0")
    assert np.array_equal(y.view(np.uint32), y0.view(np.uint32))
    assert normwise(y, P.reference_output("matmul", inputs).ravel()) <= 1e-5


def _bf16_round(a):
    u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


KNOB_OPTIONS = [["", "tile=64x64", "tile=128x128", "tile=64x128", "tile=128x256",
                 "tile=256x256 bk=64"],
                ["", "waves=2", "waves=8"], ["", "lds=swz", "lds=pad", "lds=dl"], ["", "stage=flat"],
                ["", "kb=inner", "kb=outer"],
                ["", "igemm=plain", "igemm=pipe", "igemm=halo", "igemm=halo1"],
                ["", "contract=staged", "contract=gather"], ["", "streamk=1", "streamk=0"],
                ["", "ksplit=1", "ksplit=0"]]


def random_knob_cases(n=48, seed=21):
    import random
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        name, kw = rng.choice(CONTRACT)
        knobs = " ".join(x for x in (rng.choice(o) for o in KNOB_OPTIONS) if x)
        out.append((name, kw, rng.choice(["hip", "hip_bf16"]), knobs))
    return out


@pytest.mark.parametrize("name,kw,backend,knobs", random_knob_cases(),
                         ids=lambda v: v.replace(" ", "_") if isinstance(v, str) else "")
def test_random_knob_sets(lt, hip_dev, name, kw, backend, knobs):
    """Seeded random combinations of every contraction knob (tile, waves,
    LDS layout, staging layout and K-block order, tile form, staged or
    gathered, stream-K): float operands within 1e-5 of float64, 16-bit
    operands within 1e-5 of float64 of the rounded operands."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = contract_inputs(name, shapes)
    y, cc = run_backend(lt, out, shapes, inputs, backend, annotate=knobs)
    ops = inputs if backend == "hip" else [_bf16_round(a) for a in inputs]
    f64 = P.reference_output(name, ops, **ref_kw(name, kw)).ravel()
    assert normwise(y, f64) <= 1e-5, (cc.string_property("kinds"), cc.dump().splitlines()[0])


DEGENERATE = [
    ("matmul", dict(M=1, N=1, K=1)), ("matmul", dict(M=1, N=300, K=1)),
    ("softmax", dict(R=1, Cn=1)), ("softmax", dict(R=3, Cn=1)),
    ("conv2d", dict(N=1, Ci=1, Co=1, H=1, W=1)),
    ("conv2d", dict(N=1, Ci=3, Co=2, H=2, W=2, pad=0, k=2)),
    ("maxpool2d", dict(N=1, C=1, H=2, W=2)), ("concat2d", dict(B=1, M0=1, M1=1)),
    ("transpose2d", dict(M=1, N=1)), ("bmm", dict(Bn=1, M=1, N=1, K=1)),
]


@pytest.mark.parametrize("name,kw", DEGENERATE, ids=["%s%d" % (c[0], i) for i, c in enumerate(DEGENERATE)])
@pytest.mark.parametrize("backend", ["hip", "hip_bf16"])
def test_size_one_dims(lt, hip_dev, name, kw, backend):
    """Size-1 dims everywhere (1x1 convs and images, single rows and
    columns): the float path equals the oracle within 1e-5 (bit for bit
    where lowered pointwise); 16-bit contractions match float64 of the
    rounded operands."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = [P.normal(40 + i, s) for i, s in enumerate(shapes)]
    y_or, _ = run_backend(lt, out, shapes, inputs, "cpp")
    y, cc = run_backend(lt, out, shapes, inputs, backend)
    kinds = cc.string_property("kinds")
    if kinds.endswith("contract_bf16"):
        f64 = P.reference_output(name, [_bf16_round(a) for a in inputs], **ref_kw(name, kw)).ravel()
        assert normwise(y, f64) <= 1e-5, kinds
    elif all(k == "pointwise" for k in kinds.split(",")):
        assert np.array_equal(y.view(np.uint32), y_or.view(np.uint32)), kinds
    else:
        assert normwise(y, y_or) <= 1e-5, kinds


def test_run_without_sync_on_device_buffers(lt, hip_dev):
    """Compiled::run(memory, sync=false) returns after enqueueing
    (backend.h:18-22): several runs queued back to back on device buffers
    give the oracle's result once the device is synchronised; with host
    buffers the call stages through device memory and returns complete."""
    out, shapes = P.softmax(lt, 512, 4096)
    x = P.normal(3, shapes[0], 3.0)
    y_or = np.zeros(512 * 4096, np.float32)
    lt.compile(out.loop_tree, "cpp").run([x, y_or])
    cc = lt.compile(out.loop_tree, "hip")
    dx, dy = hip_dev.DeviceArray.from_numpy(x), hip_dev.DeviceArray(512 * 4096)
    for _ in range(4):
        cc.run([dx.ptr, dy.ptr], sync=False)
    hip_dev.synchronize()
    y = dy.download()
    assert normwise(y, y_or) <= 1e-5
    yh = np.zeros_like(y_or)
    cc.run([x, yh], sync=False)
    assert np.array_equal(yh.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("prog", [("conv2d", (4, 64, 64, 28, 28), 64 * 9),
                                  ("matmul", (368, 1600, 1984), 1984)],
                         ids=["conv-halo", "matmul-direct-lds"])
def test_runs_from_threads_run_apart(lt, hip_dev, prog):
    """Compiled::run called from several host threads at once: each thread
    launches on its own stream with its own scratch and device copies of
    host buffers, so one Compiled serves them concurrently (the reference's
    cpp backend is not re-entrant, cpp.cpp:454-455).  Every thread's results
    equal the single-threaded run bit for bit, for host buffers (sync) and
    for device buffers queued with sync=false and then synchronised.  The
    programs: a conv on the halo tile and a matmul on the direct-to-LDS tile,
    whose kernels then share CUs."""
    import threading
    name, args, fan = prog
    out, shapes = getattr(P, name)(lt, *args)
    sets = [[P.normal(60 + 2 * i, shapes[0], fan ** -0.5),
             P.normal(61 + 2 * i, shapes[1], fan ** -0.5)] for i in range(4)]
    t = out.loop_tree
    cc = lt.compile(t, "hip_bf16")
    n_out = t.io_numel()[2]
    want = []
    for x, w in sets:
        y = np.zeros(n_out, np.float32)
        cc.run([x, w, y])
        want.append(y)
    d_in = [[hip_dev.DeviceArray.from_numpy(a) for a in s] for s in sets]
    d_out = [hip_dev.DeviceArray(n_out) for _ in sets]
    got, streams, errors = {}, {}, []
    start = threading.Barrier(len(sets))

    def worker(i):
        try:
            streams[i] = hip_dev.compiled_stream(cc)
            start.wait()
            for r in range(3):
                y = np.zeros(n_out, np.float32)
                cc.run(sets[i] + [y])
                got[(i, r)] = y
            for _ in range(3):
                cc.run([d_in[i][0].ptr, d_in[i][1].ptr, d_out[i].ptr], sync=False)
            hip_dev.stream_synchronize(streams[i])
            got[(i, "dev")] = d_out[i].download()
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(sets))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    assert len(set(streams.values())) == len(sets)  # one stream per thread
    assert hip_dev.compiled_stream(cc) not in streams.values()
    for i in range(len(sets)):
        for r in list(range(3)) + ["dev"]:
            assert np.array_equal(got[(i, r)].view(np.uint32), want[i].view(np.uint32)), (i, r)


def test_overlapped_runs_keep_each_run_apart(lt, hip_dev):
    """With overlap=1 a staged bf16 plan overlaps a run's operand staging
    with the previous run's MFMA kernel (two scratch sets, a second stream).
    Runs queued back to back without synchronisation, each on its own inputs
    and output, and mixed with host-pointer runs, all give the bits of the
    serial plan (the default)."""
    out, shapes = P.conv2d(lt, 4, 64, 64, 28, 28)
    fan = 64 * 9
    sets = [[P.normal(40 + 2 * i, shapes[0], fan ** -0.5),
             P.normal(41 + 2 * i, shapes[1], fan ** -0.5)] for i in range(5)]
    t = out.loop_tree
    cc = lt.compile(t.annotate(t.roots[0], "overlap=1"), "hip_bf16")
    assert cc.int_property("overlap") == 1
    serial = lt.compile(t, "hip_bf16")
    assert serial.int_property("overlap") == 0
    n_out = t.io_numel()[2]
    want = []
    for x, w in sets:
        y = np.zeros(n_out, np.float32)
        serial.run([x, w, y])
        want.append(y)
    d_in = [[hip_dev.DeviceArray.from_numpy(a) for a in s] for s in sets]
    d_out = [hip_dev.DeviceArray(n_out) for _ in sets]
    for rep in range(3):
        for i in range(len(sets)):
            cc.run([d_in[i][0].ptr, d_in[i][1].ptr, d_out[i].ptr], sync=False)
        if rep == 1:  # a host-pointer run in the middle of the queue
            y = np.zeros(n_out, np.float32)
            cc.run(sets[2] + [y])
            assert np.array_equal(y.view(np.uint32), want[2].view(np.uint32))
        hip_dev.synchronize()
        for i in range(len(sets)):
            got = d_out[i].download()
            assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (rep, i)
            d_out[i].fill_zero()
