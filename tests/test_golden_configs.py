"""The benchmark configurations at their full sizes against golden fixtures
written by the reference loop_tool library built from its own sources
(tests/golden/make_golden.py --configs -> tests/golden/config_programs.json):
the LoopTree dump and flop count of each configuration's lazy graph, and the
SHA-256 of the reference C backend's float32 output on the seeded inputs
(sampled outputs, to 1e-6, for the softmax, whose exp comes from libm).

CPU tests pin the C-oracle restatement to the fixtures.  GPU tests hold the
hip backend to the same hashes wherever its lowering computes in the oracle's
order: the index-only programs (pad, concat, transpose, max-pool) in their
default lowering, and matmul and the FP32 convolution in the exact pointwise
lowering (`hip:plain`).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from loop_tool_amd import programs as P
from test_oracle import norm_ids

_STD = 1152 ** -0.5
# (id, program, sizes, inputs, exact)
CONFIGS = [
    ("matmul_128", "matmul", dict(M=128, N=128, K=128),
     lambda s: [P.uniform(1, s[0]), P.uniform(2, s[1])], True),
    ("softmax_8192", "softmax", dict(R=8192, Cn=8192),
     lambda s: [P.normal(3, s[0], 3.0)], False),
    ("conv2d_f32_32x128x56", "conv2d", dict(N=32, Ci=128, Co=128, H=56, W=56),
     lambda s: [P.normal(4, s[0], _STD), P.normal(5, s[1], _STD)], True),
    ("maxpool_32x128x112", "maxpool2d", dict(N=32, C=128, H=112, W=112),
     lambda s: [P.normal(6, s[0])], True),
    ("pad2d_32x128x56", "pad2d", dict(N=32, C=128, H=56, W=56),
     lambda s: [P.normal(4, s[0], _STD)], True),
    ("concat_8192x4096", "concat2d", dict(B=8192, M0=4096, M1=4096),
     lambda s: [P.normal(8, s[0]), P.normal(9, s[1])], True),
    ("transpose_8192", "transpose2d", dict(M=8192, N=8192),
     lambda s: [P.normal(10, s[0])], True),
]
SAMPLE_STRIDE = 65537  # sampled outputs of the non-exact programs

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                     "config_programs.json")
GOLDEN = {e["id"]: e for e in json.load(open(_PATH))["programs"]} if os.path.exists(_PATH) else {}


def sha256(y):
    return hashlib.sha256(np.ascontiguousarray(y, np.float32).tobytes()).hexdigest()


def sample(y):
    return np.ascontiguousarray(y, np.float32).ravel()[::SAMPLE_STRIDE]


def inputs_of(L, cid):
    _, name, kw, gen, _ = next(c for c in CONFIGS if c[0] == cid)
    out, shapes = getattr(P, name)(L, **kw)
    return out, gen(shapes)


def run_tree(L, tree, inputs, backend):
    y = np.zeros(tree.io_numel()[len(inputs)], np.float32)
    L.compile(tree, backend).run([np.ascontiguousarray(a, np.float32) for a in inputs] + [y])
    return y


def check(y, e):
    if e["exact"]:
        assert sha256(y) == e["sha256"], e["id"]
    else:
        ref = np.array(e["sample"], np.float32)
        assert np.abs(sample(y) - ref).max() <= 1e-6 * np.abs(ref).max(), e["id"]


def golden(cid):
    if cid not in GOLDEN:
        pytest.fail("tests/golden/config_programs.json lacks %s: run "
                    "tests/golden/make_golden.py --configs" % cid)
    return GOLDEN[cid]


@pytest.mark.parametrize("cid", [c[0] for c in CONFIGS])
def test_oracle_matches_reference_at_full_size(lt, cid):
    e = golden(cid)
    out, inputs = inputs_of(lt, cid)
    tree = out.loop_tree
    assert norm_ids(tree.dump()) == e["dump"]
    assert tree.flops() == e["flops"]
    y = run_tree(lt, tree, inputs, "cpp")
    assert y.size == e["numel"]
    check(y, e)


GPU_EXACT = [("matmul_128", "hip:plain"), ("conv2d_f32_32x128x56", "hip:plain"),
             ("maxpool_32x128x112", ""), ("pad2d_32x128x56", ""),
             ("concat_8192x4096", ""), ("transpose_8192", "")]


@pytest.mark.gpu
@pytest.mark.parametrize("cid,knobs", GPU_EXACT, ids=[c[0] for c in GPU_EXACT])
def test_hip_matches_reference_hash_at_full_size(lt, hip_dev, cid, knobs):
    """The hip backend's output bytes hash to the reference C backend's."""
    e = golden(cid)
    out, inputs = inputs_of(lt, cid)
    tree = out.loop_tree
    if knobs:
        tree = tree.annotate(tree.roots[0], knobs)
    cc = lt.compile(tree, "hip")
    if not knobs:
        assert set(cc.string_property("kinds").split(",")) == {"pointwise"}
    y = np.zeros(tree.io_numel()[len(inputs)], np.float32)
    cc.run(inputs + [y])
    assert sha256(y) == e["sha256"], cid
