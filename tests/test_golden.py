"""The C-oracle restatement against committed golden fixtures written by the
reference loop_tool library built from its own sources
(tests/golden/make_golden.py): the same LoopTree structure, flop count,
trivially-parallel flags and -- bit for bit, or within 1e-6 for nests with
exp -- the same outputs on the same seeded inputs.  These run without the
reference build (oracle/_ref) present."""
import hashlib
import json
import os

import numpy as np
import pytest

from loop_tool_amd import programs as P
from test_frontend import BUILDS, STEPS
from test_oracle import PROGRAMS, norm_ids, run

GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                     "oracle_programs.json")))


def check_output(y, e):
    if e["exact"]:
        assert hashlib.sha256(np.ascontiguousarray(y, np.float32).tobytes()).hexdigest() == \
            e["sha256"], e["id"]
    if "output" in e:
        ref = np.array(e["output"], np.float32)
        if e["exact"]:
            assert np.array_equal(y.view(np.uint32), ref.view(np.uint32)), e["id"]
        else:  # exp: libm versions may differ in the last bit
            assert np.abs(y - ref).max() <= 1e-6 * np.abs(ref).max(), e["id"]


@pytest.mark.parametrize("i", range(len(PROGRAMS)), ids=[e["id"] for e in GOLDEN["programs"]])
def test_oracle_matches_golden_program(lt, i):
    e = GOLDEN["programs"][i]
    name, kw, gen = PROGRAMS[i]
    assert (name, kw) == (e["program"], e["kw"]), "fixture out of date: run make_golden.py"
    out, shapes = getattr(P, name)(lt, **kw)
    y, tree = run(lt, out, gen(shapes))
    assert norm_ids(tree.dump()) == e["dump"]
    assert tree.flops() == e["flops"]
    assert y.size == e["numel"]
    check_output(y, e)


@pytest.mark.parametrize("e", GOLDEN["transforms"], ids=[e["id"] for e in GOLDEN["transforms"]])
def test_transform_matches_golden(lt, e):
    build, gen = BUILDS[e["program"]]
    out, shapes = build(lt)
    t = out.loop_tree
    if e.get("rejected"):
        with pytest.raises(Exception):
            for f in STEPS[e["step"]]:
                t = f(t)
        return
    for f in STEPS[e["step"]]:
        t = f(t)
    assert norm_ids(t.dump()) == e["dump"]
    assert t.flops() == e["flops"]
    assert [bool(t.trivially_parallel(r)) for r in t.walk()] == e["parallel"]
    inputs = gen(shapes)
    y = np.zeros(t.io_numel()[len(inputs)], np.float32)
    lt.compile(t, "cpp").run([np.ascontiguousarray(a) for a in inputs] + [y])
    e = dict(e, numel=y.size)
    check_output(y, e)
