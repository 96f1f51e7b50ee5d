"""Serving runtime: structural program keys (CPU) and served results (GPU).

Served inputs are materialised on the device from (seed, scale) with the same
SplitMix64 uniform as programs.uniform, so every result is checked against a
host evaluation of bit-identical inputs: float64 closed forms (and the C
oracle for the exact max-pool).  Contractions run on 16-bit operands with
float32 accumulation: <= 1e-2 normwise of float64 and <= 1e-5 of float64 on
the bf16-rounded operands.  Softmax: <= 1e-5 normwise of float64.
"""
import numpy as np
import pytest

from loop_tool_amd import programs as P


def bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def normwise(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def host_inputs(shapes, seeds, scales):
    return [P.uniform(s, sh) * np.float32(c) for sh, s, c in zip(shapes, seeds, scales)]


# ------------------------------------------------------------------- CPU --
def test_structural_key_masks_sizes(lt):
    from loop_tool_amd import serve as S
    h = S.structural_hash
    assert h(P.matmul(lt, 3, 5, 7)[0]) == h(P.matmul(lt, 300, 200, 704)[0])
    assert h(P.matmul(lt, 3, 5, 7)[0]) == S.template_hash("matmul")
    assert h(P.softmax(lt, 100, 1000)[0]) == S.template_hash("softmax")
    assert h(P.conv2d(lt, 2, 16, 72, 15, 15)[0]) == S.template_hash("conv2d")
    assert h(P.maxpool2d(lt, 2, 3, 16, 16)[0]) == S.template_hash("maxpool2d")
    keys = {S.template_hash(f) for f in S.FAMILIES[:4]}
    assert len(keys) == 4


def test_structural_key_sees_structure(lt):
    from loop_tool_amd import serve as S
    h = S.structural_hash
    conv = S.template_hash("conv2d")
    assert h(P.conv2d(lt, 2, 16, 72, 15, 15, pad=2)[0]) != conv
    assert h(P.conv2d(lt, 2, 16, 72, 15, 15, stride=2)[0]) != conv
    assert h(P.maxpool2d(lt, 2, 3, 16, 16, k=3, stride=1)[0]) != S.template_hash("maxpool2d")
    assert h(P.add_broadcast(lt, 4, 5)[0]) not in {S.template_hash(f) for f in S.FAMILIES[:4]}
    # the window size lives in the bound sizes only: a 5x5 convolution has the
    # 3x3 structure, and the server's family plan rejects it by its sizes
    assert h(P.conv2d(lt, 2, 16, 72, 15, 15, k=5)[0]) == conv


# ------------------------------------------------------------------- GPU --
CASES = [
    ("matmul", dict(M=64, N=48, K=80), (11, 12), (1.0, 1.0)),
    ("matmul", dict(M=300, N=200, K=704), (13, 14), (1.0, 1.0)),
    ("matmul", dict(M=33, N=17, K=29), (15, 16), (1.0, 1.0)),
    ("softmax", dict(R=37, Cn=300), (21,), (3.0,)),
    ("softmax", dict(R=256, Cn=8192), (22,), (3.0,)),
    ("softmax", dict(R=5, Cn=20000), (23,), (3.0,)),
    ("conv2d", dict(N=2, Ci=40, Co=40, H=14, W=14), (31, 32), (1.0, 1 / np.sqrt(360))),
    ("conv2d", dict(N=3, Ci=64, Co=48, H=9, W=11), (33, 34), (1.0, 1 / np.sqrt(576))),
    ("maxpool2d", dict(N=2, C=3, H=16, W=18), (41,), (1.0,)),
    ("maxpool2d", dict(N=1, C=5, H=15, W=9), (42,), (1.0,)),
]


@pytest.fixture(scope="module")
def server(hip_dev):
    from loop_tool_amd import serve as S
    s = S.Server(budget_gb=16, slots=4, keep_outputs=True, checksums=True)
    yield s
    s.close()


def check_result(name, kw, y, inputs):
    ref = P.reference_output(name, inputs).ravel()
    if name in ("matmul", "conv2d"):
        ref16 = P.reference_output(name, [bf16(a) for a in inputs]).ravel()
        e, e16 = normwise(y, ref), normwise(y, ref16)
        print(name, kw, "vs f64 %.3g, vs f64 of bf16 operands %.3g" % (e, e16))
        assert e <= 1e-2 and e16 <= 1e-5
    elif name == "softmax":
        e = normwise(y, ref)
        print(name, kw, "vs f64 %.3g" % e)
        assert e <= 1e-5
    else:
        assert np.array_equal(y, ref.astype(np.float32)), name


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw,seeds,scales", CASES,
                         ids=["%s%d" % (c[0], i) for i, c in enumerate(CASES)])
def test_served_family_result(lt, server, name, kw, seeds, scales):
    out, shapes = getattr(P, name)(lt, **kw)
    assert server.family(out) == name
    rid = server.submit(out, seeds, scales)
    server.drain()
    info = server.info(rid)
    assert info["state"] == 2 and info["family"] == name, info
    y = server.output(rid)
    server.release(rid)
    check_result(name, kw, y, host_inputs(shapes, seeds, scales))


@pytest.mark.gpu
def test_served_maxpool_bit_exact_vs_oracle(lt, server):
    out, shapes = P.maxpool2d(lt, 2, 8, 12, 10)
    rid = server.submit(out, [5], [1.0])
    server.drain()
    y = server.output(rid)
    server.release(rid)
    x = host_inputs(shapes, [5], [1.0])
    ref = np.zeros(out.loop_tree.io_numel()[1], np.float32)
    lt.compile(out.loop_tree, "cpp").run(x + [ref])
    assert np.array_equal(y, ref)


@pytest.mark.gpu
def test_unknown_structure_served_by_hip_backend(lt, server):
    """A program of no family (broadcast add, a pad nest) is lowered and
    compiled by the hip backend once, then reused for the same lazy hash."""
    for name, kw, n_in in (("add_broadcast", dict(M=40, N=70), 2),
                           ("pad1d", dict(N=1000, pre=3, post=5), 1)):
        out, shapes = getattr(P, name)(lt, **kw)
        assert server.family(out) == "generic"
        seeds = list(range(50, 50 + n_in))
        before = server.stats()
        ids = [server.submit(out, seeds, [1.0] * n_in) for _ in range(2)]
        server.drain()
        after = server.stats()
        assert after["compiles"] == before["compiles"] + 1
        assert after["exact_hits"] == before["exact_hits"] + 1
        x = host_inputs(shapes, seeds, [1.0] * n_in)
        ref = np.zeros(out.loop_tree.io_numel()[n_in], np.float32)
        lt.compile(out.loop_tree, "cpp").run(x + [ref])
        for rid in ids:
            assert server.info(rid)["family"] == "generic"
            assert np.array_equal(server.output(rid), ref), name
            server.release(rid)


@pytest.mark.gpu
def test_compiled_cache_is_bounded(lt, hip_dev):
    """Programs of no family stay compiled up to `compiled_cap`, least
    recently used dropped first; a dropped program compiles again on its
    next request and still gives the oracle's result."""
    from loop_tool_amd import serve as S
    s = S.Server(budget_gb=1, slots=2, keep_outputs=True, compiled_cap=2)
    progs = [P.pad1d(lt, 100 + 10 * i, 1, 2) for i in range(3)]

    def run(k):
        out, shapes = progs[k]
        rid = s.submit(out, [7], [1.0])
        s.drain()
        x = host_inputs(shapes, [7], [1.0])
        ref = np.zeros(out.loop_tree.io_numel()[1], np.float32)
        lt.compile(out.loop_tree, "cpp").run(x + [ref])
        assert np.array_equal(s.output(rid), ref)
        s.release(rid)
    for k in (0, 1, 0, 2):  # 1 is the least recently used when 2 arrives
        run(k)
    st = s.stats()
    assert st["compiles"] == 3 and st["exact_hits"] == 1 and st["compiled_evictions"] == 1
    run(0)  # still cached
    run(1)  # evicted: compiled again
    st = s.stats()
    assert st["compiles"] == 4 and st["exact_hits"] == 2 and st["compiled_evictions"] == 2
    s.close()


def mixed_programs(lt, n=36, seed=5):
    rng = np.random.RandomState(seed)
    progs = []
    for i in range(n):
        f = i % 4
        if f == 0:
            M, N, K = (int(x) for x in rng.randint(2, 40, 3) * 16)
            progs.append((P.matmul(lt, M, N, K)[0], [100 + 2 * i, 101 + 2 * i], [1.0, 1.0]))
        elif f == 1:
            progs.append((P.softmax(lt, int(rng.randint(1, 300)), int(rng.randint(1, 5000)))[0],
                          [100 + 2 * i], [3.0]))
        elif f == 2:
            N, C, H = int(rng.randint(1, 4)), int(rng.randint(8, 80)), int(rng.randint(4, 30))
            progs.append((P.conv2d(lt, N, C, C, H, H)[0], [100 + 2 * i, 101 + 2 * i],
                          [1.0, 1 / np.sqrt(9 * C)]))
        else:
            N, C, H = int(rng.randint(1, 4)), int(rng.randint(1, 40)), int(rng.randint(2, 60))
            progs.append((P.maxpool2d(lt, N, C, H, H)[0], [100 + 2 * i], [1.0]))
    return progs


@pytest.mark.gpu
def test_same_result_alone_and_in_the_mix(lt, hip_dev):
    """Every request's output is bitwise the same when it runs alone (one slot,
    drained before the next) and when all of them are in flight together on
    eight slots: device checksums of every output, and the outputs of a few."""
    from loop_tool_amd import serve as S
    progs = mixed_programs(lt)
    mix = S.Server(budget_gb=16, slots=8, keep_outputs=True, checksums=True)
    ids = [mix.submit(p, s, c) for p, s, c in progs]
    mix.drain()
    st = mix.stats()
    assert st["completed"] == len(progs) and st["failed"] == 0
    alone = S.Server(budget_gb=16, slots=1, keep_outputs=True, checksums=True)
    for k, (p, s, c) in enumerate(progs):
        rid = alone.submit(p, s, c)
        alone.drain()
        a, b = mix.info(ids[k]), alone.info(rid)
        assert a["checksum"] == b["checksum"] != 0, (k, a["plan"])
        if k < 8:
            assert np.array_equal(mix.output(ids[k]), alone.output(rid))
        alone.release(rid)
    mix.close()
    alone.close()


@pytest.mark.gpu
def test_admission_by_memory_budget(lt, hip_dev):
    """With a budget below the sum of the requests, admission waits for
    completions; a request larger than the whole budget fails cleanly."""
    from loop_tool_amd import serve as S
    s = S.Server(budget_gb=0.2, slots=8)
    out, _ = P.softmax(lt, 2048, 4096)  # 2 x 32 MiB per request
    ids = [s.submit(out, [i], [3.0]) for i in range(12)]
    big, _ = P.softmax(lt, 8192, 8192)  # 512 MiB
    bad = s.submit(big, [1], [3.0])
    s.drain()
    st = s.stats()
    assert all(s.info(i)["state"] == 2 for i in ids)
    assert s.info(bad)["state"] == -1
    assert st["admission_waits"] > 0 and st["peak_bytes"] <= 0.2e9
    s.close()


@pytest.mark.gpu
def test_open_loop_workload_short(lt, hip_dev):
    """Poisson arrivals through the client thread: every request completes and
    the latency percentiles are ordered."""
    from loop_tool_amd import serve as S
    s = S.Server(budget_gb=32, slots=8)
    st = s.run_workload(rate=400, duration_s=1.0, seed=3)
    print(st)
    assert st["submitted"] > 200 and st["completed"] == st["submitted"]
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["struct_misses"] <= 4 and st["compiles"] == 0
    s.close()
