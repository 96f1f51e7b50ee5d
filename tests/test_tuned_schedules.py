"""Schedules drive the hip lowering, and tuned schedules come back from the
store.

The lowering reads a LoopTree's schedule (loop_tool_amd/autotune.py states
the mapping): split sizes set contraction tiles and K tiles, and the values
per lane and lanes per workgroup of row and pointwise kernels; loops marked
"unroll" (mutate.cpp:836-863) are unrolled in pointwise kernels.  The tuned
store keys a program by its IR with var names replaced by positions; the hip
backend's compile lowers a stored schedule of the same program in place of
the tree it is handed.

CPU tests compile only (hipRTC needs no GPU); GPU tests run the schedules
against the C oracle and float64.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from loop_tool_amd import autotune as A, programs as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cfg(cc):
    return {k: cc.int_property(k) for k in ("tile_m", "tile_n", "bk", "vec", "block")
            if cc.int_property(k) is not None}


def split(tree, op, var_name, *sizes):
    var = next(tree.loop(r)["var"] for r in tree.loops
               if tree.var_name(tree.loop(r)["var"]).split("_")[0] == var_name)
    for s in sizes:
        tree = A.split_var(tree, op, var, s)
        assert tree is not None
    return tree


def test_split_sets_contraction_tiles(lt):
    out, _ = P.matmul(lt, 512, 384, 640)
    t = out.loop_tree
    base = cfg(lt.compile(t, "hip_bf16"))
    t2 = split(split(split(t, "add", "M", 128), "add", "N", 256), "add", "K", 64)
    cc = lt.compile(t2, "hip_bf16")
    assert cfg(cc) == dict(tile_m=128, tile_n=256, bk=64) != base
    assert "tile 128x256x64" in cc.dump().splitlines()[0]
    # an annotation takes precedence over the schedule
    cc = lt.compile(t2.annotate(t2.roots[0], "tile=64x64"), "hip_bf16")
    assert (cfg(cc)["tile_m"], cfg(cc)["tile_n"]) == (64, 64)
    # the float gathered tile reads the same splits
    cc = lt.compile(split(split(t, "add", "M", 64), "add", "N", 128), "hip")
    assert cc.string_property("kinds") == "contract_f32"
    assert (cfg(cc)["tile_m"], cfg(cc)["tile_n"]) == (64, 128)


def test_split_sets_row_lanes_and_vector_width(lt):
    out, _ = P.softmax(lt, 256, 4096)
    t = split(out.loop_tree, "write", "C", 1024, 2)
    cc = lt.compile(t, "hip")
    assert cc.string_property("kinds") == "rowreduce"
    assert cfg(cc) == dict(vec=2, block=512)
    assert "512 lanes x 2 wide" in cc.dump().splitlines()[0]


def test_split_sets_pointwise_lanes_and_vector_width(lt):
    out, _ = P.add_broadcast(lt, 1024, 4096)
    t = split(out.loop_tree, "write", "N", 1024, 2)
    cc = lt.compile(t, "hip")
    assert cfg(cc) == dict(vec=2, block=512)
    assert "__launch_bounds__(512)" in cc.dump() and "lt_f2" in cc.dump()


def test_unroll_marks_reduction_loops(lt):
    out, _ = P.conv2d(lt, 1, 8, 4, 6, 6)
    t = out.loop_tree
    plain = lt.compile(t.annotate(t.roots[0], "hip:plain"), "hip").dump()
    t = t.unroll_inner_loops(16)
    unrolled = lt.compile(t.annotate(t.roots[0], "hip:plain"), "hip").dump()
    assert "#pragma unroll" not in plain and "#pragma unroll" in unrolled


# ------------------------------------------------------------------ store --
@pytest.fixture
def store(lt, tmp_path):
    """A store in a fresh directory; the backend's lookups return to the
    session's directory afterwards."""
    s = A.TunedStore(lt, str(tmp_path / "tuned"))
    yield s
    A.TunedStore.reset(lt)


KEY_CHILD = r"""
import sys; sys.path.insert(0, %r)
import loop_tool_amd as lt
from loop_tool_amd import autotune as A, programs as P
out, _ = P.softmax(lt.lib, 300, 2000)
print(A.TunedStore(lt.lib, %r).key(out.loop_tree))
"""


def test_key_is_stable_across_processes(lt, store):
    keys = set()
    for _ in range(2):
        r = subprocess.run([sys.executable, "-c", KEY_CHILD % (ROOT, store.path)],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        keys.add(r.stdout.split()[-1])
    # this process built other programs first: different var numbering
    P.matmul(lt, 8, 8, 8)
    out, _ = P.softmax(lt, 300, 2000)
    keys.add(store.key(out.loop_tree))
    assert len(keys) == 1
    out2, _ = P.softmax(lt, 300, 2001)
    assert store.key(out2.loop_tree) not in keys


def test_store_round_trip_and_compile_reads_it(lt, store):
    out, _ = P.matmul(lt, 512, 384, 640)
    prog = out.loop_tree
    assert store.get(prog) is None
    tuned = split(split(prog, "add", "M", 256), "add", "N", 128)
    key = store.put(prog, tuned)
    assert store.keys() == [key]
    assert store.get(prog).dump() == tuned.dump()
    cc = lt.compile(prog, "hip_bf16")
    assert cc.string_property("tuned_schedule") == key
    assert (cfg(cc)["tile_m"], cfg(cc)["tile_n"]) == (256, 128)
    assert "tuned schedule " + key in cc.dump()
    # the lazy frontend's compile goes through the same backend
    out_again, _ = P.matmul(lt, 512, 384, 640)
    lt.set_default_backend("hip_bf16")
    try:
        out_again.clear_cache()
        assert out_again.compiled.string_property("tuned_schedule") == key
        out_again.clear_cache()
    finally:
        lt.set_default_backend("cpp")
    # lookups off
    os.environ["LT_HIP_TUNED"] = "0"
    try:
        assert lt.compile(prog, "hip_bf16").string_property("tuned_schedule") == ""
    finally:
        del os.environ["LT_HIP_TUNED"]


def test_store_refuses_a_schedule_of_another_program(lt, store):
    a, _ = P.matmul(lt, 512, 384, 640)
    b, _ = P.matmul(lt, 512, 384, 512)
    with pytest.raises(RuntimeError, match="not a schedule of the program"):
        store.put(a.loop_tree, b.loop_tree)


def test_candidates_keep_their_plan_kind(lt):
    """Every schedule candidate the tuner makes compiles for gfx950 and
    keeps the plan kind of the program (the mapping never falls back to a
    slower kind)."""
    for name, kw, backend in [("matmul", dict(M=512, N=384, K=640), "hip_bf16"),
                              ("softmax", dict(R=64, Cn=4096), "hip"),
                              ("add_broadcast", dict(M=64, N=4096), "hip")]:
        out, _ = getattr(P, name)(lt, **kw)
        t = out.loop_tree
        kind = lt.compile(t, backend).string_property("kinds")
        cands = A.schedule_candidates(t, kind.split(",")[-1])
        assert cands, name
        for desc, c in cands:
            assert lt.compile(c, backend).string_property("kinds") == kind, (name, desc)


# -------------------------------------------------------------------- GPU --
def _normwise(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def _run(lt, tree, inputs, backend):
    y = np.zeros(tree.io_numel()[len(inputs)], np.float32)
    cc = lt.compile(tree, backend)
    cc.run(inputs + [y])
    return y, cc


@pytest.mark.gpu
def test_split_schedules_run_and_match(lt, hip_dev):
    """A split schedule changes the plan and the result still matches:
    bf16 matmul against float64 of the rounded operands, softmax and a
    broadcast add against the C oracle running the same tree."""
    out, sh = P.matmul(lt, 512, 384, 640)
    ins = [P.uniform(1, sh[0]), P.uniform(2, sh[1])]
    t = split(split(split(out.loop_tree, "add", "M", 128), "add", "N", 256), "add", "K", 64)
    y, cc = _run(lt, t, ins, "hip_bf16")
    assert cfg(cc) == dict(tile_m=128, tile_n=256, bk=64)
    f64 = P.reference_output("matmul", [P.round_bf16(a) for a in ins]).ravel()
    assert _normwise(y, f64) <= 1e-5

    out, sh = P.softmax(lt, 256, 4096)
    x = [P.normal(3, sh[0], 3.0)]
    for sizes in ((1024, 2), (256, 4), (128, 1)):
        t = split(out.loop_tree, "write", "C", *sizes)
        y, cc = _run(lt, t, x, "hip")
        assert cfg(cc)["vec"] == sizes[1] and cfg(cc)["block"] == sizes[0] // sizes[1]
        y_or, _ = _run(lt, t, x, "cpp")
        assert _normwise(y, y_or) <= 2e-5

    out, sh = P.add_broadcast(lt, 1024, 4096)
    ins = [P.normal(4, sh[0]), P.normal(5, sh[1])]
    t = split(out.loop_tree, "write", "N", 1024, 2)
    y, cc = _run(lt, t, ins, "hip")
    y_or, _ = _run(lt, t, ins, "cpp")
    assert np.array_equal(y, y_or)


@pytest.mark.gpu
def test_swapped_and_unrolled_reductions_follow_the_tree(lt, hip_dev):
    """The pointwise lowering accumulates in the order of the tree's
    reduction loops: swapped and unrolled schedules of a convolution equal
    the C oracle on the same tree bit for bit."""
    out, sh = P.conv2d(lt, 2, 8, 4, 10, 10)
    ins = [P.normal(4, sh[0]), P.normal(5, sh[1])]
    t = out.loop_tree.annotate(out.loop_tree.roots[0], "hip:plain")
    trees = [t.unroll_inner_loops(16)]
    for desc, c in A.schedule_candidates(t, "pointwise"):
        if desc.startswith("swap"):
            trees.append(c)
    assert len(trees) > 3
    for c in trees:
        y, _ = _run(lt, c, ins, "hip")
        y_or, _ = _run(lt, c, ins, "cpp")
        assert np.array_equal(y.view(np.uint32), y_or.view(np.uint32))


FRESH_CHILD = r"""
import json, sys
sys.path.insert(0, %(root)r)
import numpy as np
import loop_tool_amd as lt
from loop_tool_amd import hip, programs as P
lt.load_lib(%(oracle)r)
hip.init()
L = lt.lib
x = P.normal(3, (1024, 8192), 3.0)
L.set_default_backend("hip")
out, _ = P.softmax(L, 1024, 8192, data=[x])
y = out.numpy().ravel()
cc = out.compiled
rec = dict(key=cc.string_property("tuned_schedule"), dump=cc.dump()[:4000])
tree = out.loop_tree
y_or = np.zeros(y.size, np.float32)
L.compile(tree, "cpp").run([x, y_or])
rec["err_vs_oracle"] = float(np.abs(y - y_or).max() / np.abs(y_or).max())
print("RESULT " + json.dumps(rec))
"""


@pytest.mark.gpu
def test_tuned_schedule_reused_by_a_fresh_process(lt, hip_dev, store):
    """Tune the softmax in this process (schedules searched, checked against
    the oracle's result, timed); a fresh process evaluating the same program
    through the lazy frontend on `hip` compiles the stored schedule (its key
    in the Compiled's properties, its tree in dump()) and matches the
    oracle."""
    import json
    out, sh = P.softmax(lt, 1024, 8192)
    x = P.normal(3, sh[0], 3.0)
    prog = out.loop_tree
    y_or, _ = _run(lt, prog, [x], "cpp")
    best, ms, rep = A.tune(lt, prog, [x], baseline=y_or, tol=2e-5, iters=5, store=store,
                           candidates=[])
    assert all(c.get("ok", True) is not None for c in rep["candidates"])
    if rep["key"] is None:  # the tree as given won: store the best schedule
        cands = [c for c in rep["candidates"][1:] if c.get("ok") and
                 not c["schedule"].startswith("knobs")]
        win = min(cands, key=lambda c: c["ms"])
        best = dict(A.schedule_candidates(prog, "rowreduce"))[win["schedule"]]
        store.put(prog, best)
    key = store.key(prog)
    env = dict(os.environ, LT_HIP_TUNED_DIR=store.path)
    code = FRESH_CHILD % dict(root=ROOT, oracle=os.path.join(ROOT, "oracle",
                                                              "liblt_oracle.so"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                       text=True, timeout=300)
    lines = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")]
    assert lines, r.stdout[-2000:] + r.stderr[-2000:]
    rec = json.loads(lines[-1][7:])
    assert rec["key"] == key
    assert "tuned schedule " + key in rec["dump"]
    assert best.dump().splitlines()[0] in rec["dump"]
    assert rec["err_vs_oracle"] <= 2e-5
