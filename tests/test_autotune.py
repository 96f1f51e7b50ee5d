"""The schedule autotuner (loop_tool_amd/autotune.py).

CPU: every knob candidate of every plan kind lowers and compiles for gfx950
(hipRTC needs no GPU) and keeps the plan kind it is searched for.
GPU: `tune` searches schedules made with the mutate moves, knob candidates
and random schedules (the reference's test/test.py:85-123 harness: a
schedule must not change the result) and returns the fastest candidate whose
result stays within the tolerance of the oracle's.  (The tuned-schedule
store: tests/test_tuned_schedules.py.)
"""
import random

import numpy as np
import pytest

from loop_tool_amd import autotune, programs as P
from test_schedules import rand_schedule

KIND_PROGRAMS = [
    ("rowreduce", "softmax", dict(R=64, Cn=1000), "hip"),
    ("contract_f32", "matmul", dict(M=256, N=192, K=128), "hip"),
    ("contract_bf16", "conv2d", dict(N=2, Ci=64, Co=64, H=14, W=14), "hip_bf16"),
    ("pointwise", "transpose2d", dict(M=300, N=200), "hip"),
]


@pytest.mark.parametrize("kind,name,kw,backend", KIND_PROGRAMS, ids=[k[0] for k in KIND_PROGRAMS])
def test_every_candidate_compiles(lt, kind, name, kw, backend):
    out, _ = getattr(P, name)(lt, **kw)
    t = out.loop_tree
    assert lt.compile(t, backend).string_property("kinds").split(",")[-1] == kind
    cands = autotune.knob_candidates(kind)
    assert len(cands) >= 2 and len(set(cands)) == len(cands)
    for k in cands:
        cc = lt.compile(autotune.annotate_all(t, k), backend)
        got = cc.string_property("kinds").split(",")[-1]
        if "contract=gather" in k:
            assert got == "contract_f32", k
        else:
            assert got == kind, k


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw,tol", [("softmax", dict(R=256, Cn=1000), 1e-5),
                                         ("matmul", dict(M=192, N=160, K=96), 1e-5)])
def test_tune_returns_a_correct_winner(lt, hip_dev, name, kw, tol):
    out, shapes = getattr(P, name)(lt, **kw)
    if name == "softmax":
        inputs = [P.normal(3, shapes[0], 3.0)]
    else:
        inputs = [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
    tree = out.loop_tree
    y_or = np.zeros(tree.io_numel()[len(inputs)], np.float32)
    lt.compile(tree, "cpp").run(inputs + [y_or])
    rng = random.Random(11)
    extra = [rand_schedule(tree, rng) for _ in range(3)]
    best, ms, report = autotune.tune(lt, tree, inputs, baseline=y_or, tol=tol, iters=5,
                                     extra_trees=extra)
    cands = report["candidates"]
    kind = cands[0]["kind"]
    assert len(cands) == (1 + len(autotune.schedule_candidates(tree, kind)) +
                          len(autotune.knob_candidates(kind)) + len(extra))
    assert any(c["schedule"].startswith("split") for c in cands) or kind == "contract_f32"
    timed = [c for c in cands if "ms" in c]
    assert all(c["ms"] > 0 for c in timed)
    # the winner is the fastest candidate that kept the result
    ok = [c for c in timed[1:] if c["ok"]] + [timed[0]]
    assert ms == min(c["ms"] for c in ok)
    y = np.zeros_like(y_or)
    lt.compile(best, "hip").run(inputs + [y])
    scale = float(np.abs(y_or).max())
    assert float(np.abs(y - y_or).max()) / scale <= tol
