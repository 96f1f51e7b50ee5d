"""Packs of compiled nests in one hipGraph (include/loop_tool_hip.h,
lth_pack_*).

Nests of every plan kind -- pointwise, index-only, row reductions, gathered
and staged contractions in float32 and bf16, convolutions -- are captured into
one graph, spread over one or several concurrent streams.  Each replay must
give every nest's output bit for bit as one Compiled::run of that nest does,
over repeated replays with eager work (the outputs cleared) between them.
"""
import numpy as np
import pytest

from loop_tool_amd import programs as P

pytestmark = pytest.mark.gpu

MIX = [
    ("softmax", dict(R=64, Cn=1000), "hip"),
    ("softmax", dict(R=300, Cn=4096), "hip"),
    ("matmul", dict(M=128, N=128, K=128), "hip"),
    ("matmul", dict(M=300, N=200, K=704), "hip_bf16"),
    ("matmul", dict(M=37, N=53, K=29), "hip"),
    ("conv2d", dict(N=2, Ci=32, Co=40, H=14, W=14), "hip_bf16"),
    ("conv2d", dict(N=2, Ci=8, Co=16, H=12, W=12), "hip"),
    ("maxpool2d", dict(N=2, C=3, H=16, W=16), "hip"),
    ("pad2d", dict(N=2, C=5, H=9, W=11), "hip"),
    ("concat2d", dict(B=17, M0=33, M1=20), "hip"),
    ("transpose2d", dict(M=300, N=77), "hip"),
    ("add_broadcast", dict(M=513, N=129), "hip"),
]


def build(lt, hip_dev):
    """(Compiled, device buffers) per nest of the mix, each with its own
    seeded inputs and an output buffer."""
    nests = []
    for i, (name, kw, backend) in enumerate(MIX):
        out, shapes = getattr(P, name)(lt, **kw)
        tree = out.loop_tree
        inputs = [P.normal(100 + 7 * i + j, s) for j, s in enumerate(shapes)]
        cc = lt.compile(tree, backend)
        d_in = [hip_dev.DeviceArray.from_numpy(a) for a in inputs]
        d_out = hip_dev.DeviceArray(tree.io_numel()[len(inputs)])
        nests.append((cc, d_in + [d_out]))
    return nests


@pytest.mark.parametrize("lanes", [1, 4, 8])
def test_pack_replays_match_each_run(lt, hip_dev, lanes):
    nests = build(lt, hip_dev)
    want = []
    for cc, io in nests:
        cc.run([b.ptr for b in io])
        want.append(io[-1].download())
    for _, io in nests:
        io[-1].fill_zero()
    hip_dev.synchronize()
    pack = hip_dev.Pack(nests, lanes=lanes)
    try:
        for rep in range(3):
            pack.replay()
            pack.synchronize()
            for i, (_, io) in enumerate(nests):
                got = io[-1].download()
                assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (
                    rep, MIX[i])
            for _, io in nests:  # eager work between replays
                io[-1].fill_zero()
            hip_dev.synchronize()
    finally:
        pack.free()


def test_pack_refuses_bad_arguments(lt, hip_dev):
    nests = build(lt, hip_dev)[:2]
    with pytest.raises(Exception, match="lanes"):
        hip_dev.Pack(nests, lanes=0)
    cc, io = nests[0]
    with pytest.raises(Exception, match="device buffers"):
        hip_dev.Pack([(cc, io[:-1])], lanes=1)
    # flags other than LTH_PACK_GROUP_MATMULS are refused
    import ctypes as C
    L = hip_dev._L
    s = hip_dev.Stream()
    h = L.lth_pack_create((C.c_void_p * 1)(cc.h), hip_dev._ptrs(io), (C.c_int * 1)(len(io)), 1, 1,
                          2, s.h)
    assert not h and b"unknown pack flags" in L.lth_last_error()


def test_pack_runs_one_compiled_on_several_buffer_sets(lt, hip_dev):
    """One Compiled packed three times over different inputs and outputs
    (each entry gets scratch of its own): every output is its own run's."""
    out, shapes = P.conv2d(lt, N=2, Ci=32, Co=40, H=14, W=14)
    tree = out.loop_tree
    cc = lt.compile(tree, "hip_bf16")
    nests, want = [], []
    for r in range(3):
        d_in = [hip_dev.DeviceArray.from_numpy(P.normal(500 + 10 * r + j, s))
                for j, s in enumerate(shapes)]
        d_out = hip_dev.DeviceArray(tree.io_numel()[len(shapes)])
        cc.run([b.ptr for b in d_in + [d_out]])
        want.append(d_out.download())
        d_out.fill_zero()
        nests.append((cc, d_in + [d_out]))
    hip_dev.synchronize()
    pack = hip_dev.Pack(nests, lanes=3)
    try:
        for rep in range(2):
            pack.replay()
            pack.synchronize()
            for i, (_, io) in enumerate(nests):
                assert np.array_equal(io[-1].download().view(np.uint32),
                                      want[i].view(np.uint32)), (rep, i)
                io[-1].fill_zero()
            hip_dev.synchronize()
    finally:
        pack.free()


GROUP_MIX = [
    ("matmul", dict(M=300, N=200, K=704), "hip_bf16", ""),
    ("matmul", dict(M=1024, N=512, K=256), "hip_bf16", ""),
    ("matmul", dict(M=64, N=2000, K=48), "hip_bf16", ""),
    ("matmul", dict(M=37, N=53, K=29), "hip_bf16", ""),  # K not a multiple of 8: not grouped
    ("matmul", dict(M=256, N=96, K=512), "hip", "mfma_f16"),
    ("matmul", dict(M=128, N=128, K=128), "hip", ""),  # float32: not grouped
    ("softmax", dict(R=300, Cn=4096), "hip", ""),
    ("conv2d", dict(N=2, Ci=32, Co=40, H=14, W=14), "hip_bf16", ""),
]


@pytest.mark.parametrize("lanes", [1, 3])
def test_pack_groups_matmul_nests(lt, hip_dev, lanes):
    """group_matmuls: the plain 16-bit matmul nests (bf16 and fp16) run as
    one grouped launch per operand type, within float32 rounding of their
    own runs (<= 1e-5 normwise; the same rounded operands, another
    accumulation order) and of float64 over the rounded operands; the other
    nests stay bit-identical to their own runs, over repeated replays."""
    nests, want, grouped = [], [], []
    for i, (name, kw, backend, ann) in enumerate(GROUP_MIX):
        out, shapes = getattr(P, name)(lt, **kw)
        tree = out.loop_tree
        if ann:
            tree = tree.annotate(tree.roots[0], ann)
        cc = lt.compile(tree, backend)
        ins = [P.normal(700 + 5 * i + j, s) for j, s in enumerate(shapes)]
        d_in = [hip_dev.DeviceArray.from_numpy(a) for a in ins]
        d_out = hip_dev.DeviceArray(tree.io_numel()[len(shapes)])
        cc.run([b.ptr for b in d_in + [d_out]])
        want.append(d_out.download())
        kinds = cc.string_property("kinds").split(",")
        grouped.append(cc.int_property("plain_matmul") == 1 and kw.get("K", 1) % 8 == 0 and
                       kinds[-1] in ("contract_bf16", "contract_f16"))
        if grouped[-1]:
            r = P.round_bf16 if kinds[-1] == "contract_bf16" else (
                lambda a: a.astype(np.float16).astype(np.float32))
            ref = r(ins[0]).astype(np.float64) @ r(ins[1]).astype(np.float64)
            grouped[-1] = ref
        d_out.fill_zero()
        nests.append((cc, d_in + [d_out]))
    assert sum(g is not False for g in grouped) == 4
    hip_dev.synchronize()
    pack = hip_dev.Pack(nests, lanes=lanes, group_matmuls=True)
    try:
        for rep in range(3):
            pack.replay()
            pack.synchronize()
            for i, (_, io) in enumerate(nests):
                got = io[-1].download()
                if grouped[i] is False:
                    assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (
                        rep, GROUP_MIX[i])
                else:
                    ref = grouped[i].ravel()
                    own = want[i].astype(np.float64)
                    g64 = got.astype(np.float64)
                    assert np.linalg.norm(g64 - own) <= 1e-5 * np.linalg.norm(own), GROUP_MIX[i]
                    assert np.linalg.norm(g64 - ref) <= 1e-5 * np.linalg.norm(ref), GROUP_MIX[i]
                io[-1].fill_zero()
            hip_dev.synchronize()
    finally:
        pack.free()


def test_pack_back_to_back_replays_of_the_serving_mix(lt, hip_dev):
    """Regression: 64 nests of the serving mix (every plan kind, many small
    direct-to-LDS contraction tiles) on 8 lanes, replayed back to back with
    no synchronisation between replays, so kernels of different nests share
    CUs.  Every output must stay bit-identical to a one-stream launch
    sequence.  (A barrier of the direct-to-LDS tiles that let LDS reads stay
    in flight across it gave wrong output blocks here under contention.)"""
    from loop_tool_amd import packed
    specs = packed.serving_mix(n=64, seed=7)
    compiled = packed.compile_nests(lt, specs, workers=8)
    rng = np.random.default_rng(7)
    nests, scratch = [], []
    for cc, tree, shapes in compiled:
        d_in = [hip_dev.DeviceArray.from_numpy(rng.standard_normal(int(np.prod(sh)), np.float32))
                for sh in shapes]
        d_out = hip_dev.DeviceArray(tree.io_numel()[len(shapes)])
        nests.append((cc, d_in + [d_out]))
        scratch.append([hip_dev.DeviceArray(max(w, 1)) for w in hip_dev.scratch_sizes(cc)])
    stream = hip_dev.Stream()
    for (cc, io), sc in zip(nests, scratch):
        hip_dev.run_on(cc, [b.ptr for b in io], [b.ptr for b in sc], stream.h)
    stream.synchronize()
    want = [io[-1].download() for _, io in nests]
    pack = hip_dev.Pack(nests, lanes=8, stream=stream.h)
    try:
        for rnd in range(4):
            for _ in range(25):
                pack.replay()
            stream.synchronize()
            bad = [i for i, (_, io) in enumerate(nests)
                   if not np.array_equal(io[-1].download().view(np.uint32),
                                         want[i].view(np.uint32))]
            assert not bad, (rnd, [specs[i] for i in bad])
    finally:
        pack.free()
