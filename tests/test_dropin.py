"""Drop-in check: the reference loop_tool library (compiled from
/root/reference, default visibility, as its own build produces it) loads this
repository's hip backend as an out-of-tree plugin through its own
loop_tool::loadLibrary (include/loop_tool/backend.h:92), and the reference's
own lazy frontend, IR and LoopTree then drive it unchanged.

Runs in a child process holding only the reference stack.  The child ends with
os._exit because the reference unloads plugin libraries before its backend
registry during static destruction (src/core/backend.cpp:16,63-65).
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PUB = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_pub.so")
PLUGIN = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_hip.so")

CHILD = r"""
import json, os, sys
import numpy as np
sys.path.insert(0, %(root)r)
from loop_tool_amd import capi, programs as P
L = capi.Library(%(pub)r)
L.load_lib(%(plugin)r)
out = {"backends": L.backends(), "default": L.get_default_backend(), "cases": []}
run = %(run)r
cases = [("matmul", dict(M=64, N=48, K=80)), ("softmax", dict(R=64, Cn=1024)),
         ("conv2d", dict(N=2, Ci=8, Co=16, H=12, W=12)), ("maxpool2d", dict(N=2, C=3, H=16, W=16))]
for name, kw in cases:
    prog, shapes = getattr(P, name)(L, **kw)
    tree = prog.loop_tree
    cc = L.compile(tree, "hip")
    rec = {"name": name, "compiled": cc.name, "kinds": cc.string_property("kinds")}
    if run:
        gens = {"matmul": lambda s: [P.uniform(1, s[0]), P.uniform(2, s[1])],
                "softmax": lambda s: [P.normal(3, s[0], 3.0)],
                "conv2d": lambda s: [P.normal(4, s[0]), P.normal(5, s[1])],
                "maxpool2d": lambda s: [P.normal(6, s[0])]}
        ins = gens[name](shapes)
        n = tree.io_numel()[len(ins)]
        y_hip = np.zeros(n, np.float32); y_ref = np.zeros(n, np.float32)
        cc.run(ins + [y_hip])
        L.compile(tree, "cpp").run(ins + [y_ref])
        rec["err"] = float(np.abs(y_hip - y_ref).max() / np.abs(y_ref).max())
        # the reference's own lazy frontend evaluating the program with hip as
        # its default backend: data<float>() compiles through the plugin and
        # runs it (include/loop_tool/lazy.h:219-260), with the output owned
        # by the CPU hardware, by the plugin's pinned-memory hardware, or
        # bound to a caller array (the reference Python numpy())
        L.set_default_backend("hip")
        lazy = {}
        for hw in ("cpu", "hip", "bound"):
            L.set_default_hardware("cpu" if hw == "bound" else hw)
            prog2, _ = getattr(P, name)(L, data=ins, **kw)
            prog2.clear_cache()
            if hw == "bound":
                y_lazy = np.full(n, np.nan, np.float32)
                prog2.numpy_into(y_lazy)
            else:
                y_lazy = prog2.numpy().ravel()
            rec.setdefault("lazy_compiled", []).append(prog2.compiled.name)
            prog2.clear_cache()
            # the reference frees owned outputs through the default hardware
            # (lazy.h:120-123): release it before the default changes back
            del prog2
            lazy[hw] = y_lazy
        L.set_default_hardware("cpu")
        L.set_default_backend("cpp")
        rec["lazy_err"] = {hw: float(np.abs(v - y_ref).max() / np.abs(y_ref).max())
                           for hw, v in lazy.items()}
        rec["lazy_bitwise_vs_compile"] = {
            hw: bool(np.array_equal(v.view(np.uint32), y_hip.view(np.uint32)))
            for hw, v in lazy.items()}
    out["cases"].append(rec)
print("RESULT " + json.dumps(out))
sys.stdout.flush()
os._exit(0)
"""


def run_child(run):
    if not (os.path.exists(PUB) and os.path.exists(PLUGIN)):
        pytest.skip("reference plugin stack not built (oracle/build.py)")
    code = CHILD % dict(root=ROOT, pub=PUB, plugin=PLUGIN, run=run)
    r = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    lines = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")]
    assert lines, r.stdout[-3000:]
    return json.loads(lines[-1][7:])


def test_reference_library_registers_and_lowers_hip_plugin():
    res = run_child(False)
    assert "hip" in res["backends"] and "cpp" in res["backends"]
    assert res["default"] == "cpp"  # the reference keeps its own default
    kinds = {c["name"]: c["kinds"] for c in res["cases"]}
    assert kinds == {"matmul": "contract_f32", "softmax": "rowreduce",
                     "conv2d": "contract_f32", "maxpool2d": "pointwise"}


@pytest.mark.gpu
def test_reference_library_runs_hip_plugin_on_gpu(hip_dev):
    res = run_child(True)
    for c in res["cases"]:
        print(c)
        assert c["compiled"] == "hip"
        assert c["err"] <= 1e-5, c
        # the reference's lazy frontend drove the plugin, three ways
        assert c["lazy_compiled"] == ["hip"] * 3, c
        for hw, e in c["lazy_err"].items():
            assert e <= 1e-5, (c["name"], hw, e)
        # same plugin, same LoopTree, same inputs: the same bits as the
        # direct Backend::compile run
        assert all(c["lazy_bitwise_vs_compile"].values()), c
