"""Random schedules: the correctness harness of the reference's test/test.py.

The reference draws random splits per loop var (`rand_sched`,
test/test.py:85-96), runs each schedule on every backend and requires
max|y - ref| < 1e-3 * mean|ref| against numpy (`check_exec`, :99-123), over
pointwise add, sum-reduce and matmul at powers of two, odd sizes and primes
with seed 1337 (:284-313).  Here the same programs come from the lazy
frontend, and a schedule is drawn with the LoopTree transforms the tuner
uses: 1-3 splits per loop with the inner size uniform in [2, sqrt(size)]
(tails allowed), then up to two swaps of a loop with its parent loop.

The reference's C backend mis-indexes one family of these schedules: a tail
block pins its iterator to the loop's own size (cpp.cpp:338-346), also where
an enclosing tail block has shrunk that loop, so the block indexes past the
buffers and the last elements are never computed.  `c_backend_misindexes`
finds those trees from the loop structure alone.

* CPU: the C oracle equals the reference library's own C backend bit for
  bit on every schedule, out-of-range writes included (buffers sit inside
  guard regions); where the reference indexes correctly both meet the
  reference's bar against float64, and where it does not, the interpreter
  (the reference's other CPU backend, restated as the secondary oracle) does.
* GPU: the hip backend equals the C oracle bit for bit where it lowers the
  tree pointwise, and within 1e-5 normwise where it regroups a reduction; on
  the mis-indexed trees it is held to the interpreter instead, since it
  starts each tail block where the enclosing block's iterations end.

Beyond the reference's programs, a softmax chain and a padded conv are drawn
the same way.
"""
import math
import random
import re

import numpy as np
import pytest

from loop_tool_amd import programs as P
from loop_tool_amd.capi import LoopToolError

# test/test.py:287-292: powers of two, odd numbers near them, primes
SIZES = ([2 ** i for i in range(1, 10)] + [2 ** i + 1 for i in range(1, 8)] +
         [x for x in range(8, 30) if all(x % y for y in range(2, x))])
MM = [(m, n, k) for m in range(1, 17, 3) for n in range(5, 7) for k in range(8, 11)]


def pw_add(L, n):
    a = L.Symbol("a")
    return L.Tensor(n).as_(a) + L.Tensor(n).as_(a), [(n,), (n,)]


def reduce_add(L, n):
    s = L.Symbol("N")
    return L.Tensor(n).as_(s).sum(s), [(n,)]


def build(L, prog, size):
    if prog == "pw":
        return pw_add(L, size)
    if prog == "reduce":
        return reduce_add(L, size)
    if prog == "softmax":
        return P.softmax(L, *size)
    if prog == "conv":
        return P.conv2d(L, *size)
    return P.matmul(L, *size)


def rand_schedule(t, rng):
    """Random splits (1-3 per loop of the tree as built), then up to two
    parent/child loop swaps; illegal swaps are skipped."""
    for _ in range(rng.randint(1, 3) * max(1, len(t.loops))):
        ref = rng.choice(t.loops)
        size = t.loop(ref)["size"]
        if size >= 4:
            t = t.split(ref, rng.randint(2, math.isqrt(size)))
    for _ in range(rng.randint(0, 2)):
        pairs = [(t.parent(r), r) for r in t.loops
                 if t.parent(r) >= 0 and t.is_loop(t.parent(r))]
        if not pairs:
            break
        a, b = rng.choice(pairs)
        try:
            t = t.swap_loops(a, b)
        except LoopToolError:
            pass
    return t


def inner_extent(t, ref):
    """Extent of the var covered by one iteration of loop `ref` (the C
    backend's inner_sizes): the first loop of the same var below it."""
    v = t.loop(ref)["var"]

    def find(r):
        for c in t.children(r):
            if t.is_loop(c):
                if t.loop(c)["var"] == v:
                    return c
                d = find(c)
                if d is not None:
                    return d
        return None
    d = find(ref)
    if d is None:
        return 1
    L = t.loop(d)
    return L["size"] * inner_extent(t, d) + L["tail"]


def c_backend_misindexes(t):
    """True when the C backend's tail override (cpp.cpp:295-351) shrinks a
    loop that then has a tail block of its own: that block's iterator is set
    to the loop's full size, not to the iterations run before it."""
    def walk(ref, overrides):
        if not t.is_loop(ref):
            return False
        L = t.loop(ref)
        v, size, tail = L["var"], L["size"], L["tail"]
        ov = dict(overrides)
        if v in ov:
            inner = inner_extent(t, ref)
            size, tail = ov[v] // inner, ov[v] % inner
            del ov[v]
            if tail > 0 and size != L["size"]:
                return True
        if any(walk(c, ov) for c in t.children(ref)):
            return True
        if tail > 0:
            o2 = dict(ov)
            o2[v] = tail
            return any(walk(c, o2) for c in t.children(ref))
        return False
    return any(walk(r, {}) for r in t.roots)


GUARD = 1 << 12
SENTINEL = -7.0  # fills the output's guard regions


def guarded(a, fill=0.0):
    """`a` inside a buffer whose surroundings hold `fill`: reads and writes
    past `a` by a C backend that mis-indexes stay in the guard, and are
    compared too."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    buf = np.full(a.size + 2 * GUARD, fill, np.float32)
    buf[GUARD:GUARD + a.size] = a
    return buf


def run_guarded(L, tree, backend, xs):
    """The whole guarded output buffer: inputs read 0 past their ends, the
    output's guards start as SENTINEL."""
    n = tree.io_numel()[len(xs)]
    bufs = [guarded(a) for a in xs] + [guarded(np.zeros(n, np.float32), SENTINEL)]
    L.compile(tree, backend).run([b[GUARD:b.size - GUARD] for b in bufs])
    return bufs[-1]


def guards_untouched(b, n):
    return bool((b[:GUARD] == SENTINEL).all() and (b[GUARD + n:] == SENTINEL).all())


def cases(seed=1337, reps_1d=2, reps_mm=1):
    """(program, size, schedule seed) in the reference's order and counts
    (fewer repetitions: every case also compiles C with the host compiler)."""
    rng = random.Random(seed)
    out = []
    for prog in ("pw", "reduce"):
        for s in SIZES:
            out += [(prog, s, rng.randrange(1 << 30)) for _ in range(reps_1d)]
    for mnk in MM:
        out += [("mm", mnk, rng.randrange(1 << 30)) for _ in range(reps_mm)]
    # beyond the reference's list: a softmax chain and a padded conv
    for rc in ((7, 33), (16, 64), (3, 129)):
        out += [("softmax", rc, rng.randrange(1 << 30)) for _ in range(2)]
    for conv in ((1, 3, 4, 6, 6), (2, 4, 2, 5, 7)):
        out += [("conv", conv, rng.randrange(1 << 30)) for _ in range(3)]
    return out


CASES = cases()


def inputs_for(shapes, seed):
    # test/test.py uses randn inputs
    return [P.normal(seed * 7 + i, s) for i, s in enumerate(shapes)]


def f64_reference(prog, xs):
    x = [np.asarray(a, np.float64) for a in xs]
    if prog == "pw":
        return x[0] + x[1]
    if prog == "reduce":
        return np.array([x[0].sum()])
    if prog == "softmax":
        return P.reference_output("softmax", xs)
    if prog == "conv":
        return P.reference_output("conv2d", xs)
    return x[0] @ x[1]


def reference_bar(y, ref):
    """test/test.py:107-113: max_diff < 1e-3 * mean|ref|."""
    return float(np.abs(y - ref.ravel()).max()) < 1e-3 * float(np.abs(ref).mean())


def test_schedules_are_random_and_legal(lt):
    """The draw changes the tree (splits with tails, swaps) and keeps the
    loop extents: every schedule's I/O sizes equal the unscheduled tree's.
    Some draws are trees the C backend mis-indexes."""
    changed = misindexed = 0
    for prog, size, seed in CASES:
        out, _ = build(lt, prog, size)
        t0 = out.loop_tree
        t = rand_schedule(t0, random.Random(seed))
        assert t.io_numel() == t0.io_numel()
        changed += t.dump() != t0.dump()
        misindexed += c_backend_misindexes(t)
    assert changed >= len(CASES) // 2
    assert 0 < misindexed < len(CASES) // 2


def test_misindexed_tail_detector(lt):
    """33 elements split 4 x (2 x 4) + 1: the tail's element is computed
    at index 40 by the C oracle (past the buffer, into the guard) and never
    at 32; the detector flags the tree and not its two-level form."""
    out, shapes = pw_add(lt, 33)
    t2 = out.loop_tree.split(out.loop_tree.loops[0], 4)     # 8 r 1 x 4
    t3 = t2.split(t2.loops[0], 2)                            # 4 r 1 x 2 x 4
    assert not c_backend_misindexes(t2) and c_backend_misindexes(t3)
    xs = [np.arange(33, dtype=np.float32), np.ones(33, np.float32)]
    y2 = run_guarded(lt, t2, "cpp", xs)[GUARD:GUARD + 33]
    np.testing.assert_array_equal(y2, xs[0] + 1)
    b = run_guarded(lt, t3, "cpp", xs)
    assert b[GUARD + 32] == 0 and b[GUARD + 40] == 0  # 0 + 0 read from the guards
    assert not guards_untouched(b, 33)
    y = run_guarded(lt, t3, "cpu_interpreted", xs)[GUARD:GUARD + 33]
    np.testing.assert_array_equal(y, xs[0] + 1)


@pytest.mark.parametrize("prog,size,seed", CASES[::3],
                         ids=["%s-%s-%d" % (c[0], c[1], i) for i, c in enumerate(CASES[::3])])
def test_random_schedule_oracle_vs_reference(lt, ref, prog, size, seed):
    """The oracle and the reference's own C backend agree bit for bit on a
    random schedule, guard regions included; the result meets the
    reference's bar against float64 unless the C backend mis-indexes the
    tree, where the interpreter does."""
    out, shapes = build(lt, prog, size)
    t = rand_schedule(out.loop_tree, random.Random(seed))
    # the same draw replayed on the reference's own tree gives the same
    # schedule: tree structure and transforms are restated exactly
    # (symbol ids count per library: compare the dumps without them)
    r_out, _ = build(ref, prog, size)
    t_ref = rand_schedule(r_out.loop_tree, random.Random(seed))
    assert re.sub(r"_\d+", "", t_ref.dump()) == re.sub(r"_\d+", "", t.dump())
    xs = inputs_for(shapes, seed)
    b_or = run_guarded(lt, t, "cpp", xs)
    b_ref = run_guarded(ref, t_ref, "cpp", xs)
    np.testing.assert_array_equal(b_or.view(np.uint32), b_ref.view(np.uint32))
    want = f64_reference(prog, xs)
    n = want.size
    if c_backend_misindexes(t):
        if prog == "conv":  # the interpreter's padded views differ as well
            return
        y = run_guarded(lt, t, "cpu_interpreted", xs)[GUARD:GUARD + n]
    else:
        y = b_or[GUARD:GUARD + n]
        assert guards_untouched(b_or, n)
    assert reference_bar(y, want), t.dump()


@pytest.mark.gpu
@pytest.mark.parametrize("prog,size,seed", CASES,
                         ids=["%s-%s-%d" % (c[0], c[1], i) for i, c in enumerate(CASES)])
def test_random_schedule_hip_vs_oracle(lt, hip_dev, prog, size, seed):
    out, shapes = build(lt, prog, size)
    t = rand_schedule(out.loop_tree, random.Random(seed))
    xs = inputs_for(shapes, seed)
    n = t.io_numel()[len(xs)]
    oracle = "cpu_interpreted" if c_backend_misindexes(t) else "cpp"
    if oracle == "cpp" or prog != "conv":
        y_or = run_guarded(lt, t, oracle, xs)[GUARD:GUARD + n]
    else:  # the interpreter's padded views differ: float64 instead
        y_or = f64_reference(prog, xs).astype(np.float32).ravel()
        oracle = "float64"
    cc = lt.compile(t, "hip")
    y = np.zeros(n, np.float32)
    cc.run(list(xs) + [y])
    kinds = cc.string_property("kinds").split(",")
    if all(k == "pointwise" for k in kinds) and (oracle == "cpp" or prog == "pw"):
        np.testing.assert_array_equal(y, y_or, err_msg=t.dump())
    else:
        scale = max(float(np.abs(y_or).max()), 1e-30)
        assert float(np.abs(y - y_or).max()) / scale <= 1e-5, (kinds, oracle, t.dump())
    assert reference_bar(y, f64_reference(prog, xs)), t.dump()
