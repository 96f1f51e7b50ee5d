"""Pins the secondary CPU oracle (oracle/cpu_interpreted.cpp, registered as
"cpu_interpreted") bit-for-bit against the interpreter of the reference
library built from /root/reference (oracle/_ref), on every oracle program and
on split schedules with tails; against the C oracle where both are exact; and
checks the "[backend]" sub-tree offload (src/core/compile.cpp:273-386)."""
import numpy as np
import pytest

from loop_tool_amd import programs as P
from test_oracle import PROGRAMS, run


GUARD = 1 << 16


def guarded(a, fill):
    """`a` as a view into a larger buffer whose surroundings hold `fill`.
    The reference interpreter's bound predicates take their offset from the
    access intercept (src/core/compile.cpp:912-914), so on padded views it
    reads past the input buffer; a fixed guard makes those reads -- and the
    outputs -- deterministic, and the same for both interpreters."""
    a = np.ascontiguousarray(a, np.float32).ravel()
    buf = np.full(a.size + 2 * GUARD, fill, np.float32)
    buf[GUARD:GUARD + a.size] = a
    return buf[GUARD:GUARD + a.size]


@pytest.mark.parametrize("name,kw,gen", PROGRAMS, ids=["%s%d" % (p[0], i) for i, p in enumerate(PROGRAMS)])
@pytest.mark.parametrize("fill", [0.0, 1.0])
def test_interpreter_matches_reference_interpreter(lt, ref, name, kw, gen, fill):
    out_a, shapes = getattr(P, name)(lt, **kw)
    out_b, _ = getattr(P, name)(ref, **kw)
    inputs = [guarded(a, fill) for a in gen(shapes)]
    ya, _ = run(lt, out_a, inputs, "cpu_interpreted")
    yb, _ = run(ref, out_b, inputs, "cpu_interpreted")
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), name


@pytest.mark.parametrize("split", [3, 5, 7])
def test_interpreter_split_schedules_match_reference(lt, ref, split):
    def tree(L):
        out, shapes = P.matmul(L, 13, 11, 17)
        t = out.loop_tree
        t = t.split(t.roots[0], split)  # M loop: split with a tail
        return t, shapes
    ta, shapes = tree(lt)
    tb, _ = tree(ref)
    inputs = [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    ys = []
    for L, t in ((lt, ta), (ref, tb)):
        y = np.zeros(t.io_numel()[2], np.float32)
        L.compile(t, "cpu_interpreted").run(inputs + [y])
        ys.append(y)
    assert np.array_equal(ys[0].view(np.uint32), ys[1].view(np.uint32))
    np.testing.assert_allclose(ys[0], (inputs[0].astype(np.float64) @ inputs[1]).ravel(),
                               rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name,kw,gen", [p for p in PROGRAMS if p[0] in
                                         ("matmul", "maxpool2d", "transpose2d", "add_broadcast")],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_interpreter_equals_c_oracle_without_views(lt, name, kw, gen):
    """No padded view and no transcendental: the interpreter's float
    arithmetic in loop order equals the C oracle bit for bit."""
    out, shapes = getattr(P, name)(lt, **kw)
    inputs = gen(shapes)
    ya, _ = run(lt, out, inputs, "cpu_interpreted")
    yb, _ = run(lt, out, inputs, "cpp")
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))


@pytest.mark.parametrize("prog,kw", [("matmul", dict(M=6, N=7, K=5)),
                                     ("add_broadcast", dict(M=6, N=9)),
                                     ("transpose2d", dict(M=7, N=5))])
def test_backend_annotation_behaves_as_reference(lt, ref, prog, kw):
    """"[backend]" annotations hand a sub-tree to another backend
    (compile.cpp:273-386).  The reference's stride lookup covers only the loops
    strictly above the annotated one (compile.cpp:444-460, 330-337), so on
    these trees -- annotated at any loop -- the reference interpreter refuses
    to compile; the restatement refuses exactly where the reference does."""
    for L_ in (lt, ref):
        out, _ = getattr(P, prog)(L_, **kw)
        t = out.loop_tree
        for r in t.loops:
            with pytest.raises(Exception):
                L_.compile(t.annotate(r, "[cpp]"), "cpu_interpreted")
