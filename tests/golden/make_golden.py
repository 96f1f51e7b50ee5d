"""Writes tests/golden/oracle_programs.json from the reference loop_tool
library built from its own sources (oracle/_ref/libloop_tool_ref.so, see
oracle/build.py), so the C-oracle restatement stays pinned where that build
is absent.

For every program of tests/test_oracle.py (PROGRAMS) and every schedule
transform of tests/test_frontend.py the fixture records, as the reference
produced them on the repository's seeded inputs:
  * the LoopTree dump with symbol ids removed, and the flop count;
  * the SHA-256 of the reference C backend's float32 output bytes, and the
    output itself when it has at most 4096 elements;
  * for each transform, the per-ref trivially-parallel flags, or that the
    reference rejected it.
Programs with exp are pinned by tolerance only (libm versions differ).

With --configs it writes tests/golden/config_programs.json instead: the
benchmark configurations of tests/test_golden_configs.py (CONFIGS) at their
full sizes -- LoopTree dump, flop count, and the SHA-256 of the reference C
backend's output (strided output samples for the softmax, which uses exp).

    python tests/golden/make_golden.py [--configs]
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from loop_tool_amd import capi, programs as P  # noqa: E402
from test_frontend import BUILDS, STEPS  # noqa: E402
from test_oracle import PROGRAMS, norm_ids, run  # noqa: E402

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_ref.so")
OUT = os.path.join(HERE, "oracle_programs.json")
OUT_CONFIGS = os.path.join(HERE, "config_programs.json")


def configs(ref):
    import test_golden_configs as G
    entries = []
    for cid, name, kw, gen, exact in G.CONFIGS:
        out, shapes = getattr(P, name)(ref, **kw)
        inputs = gen(shapes)
        tree = out.loop_tree
        y = G.run_tree(ref, tree, inputs, "cpp")
        e = dict(id=cid, program=name, kw=kw, dump=norm_ids(tree.dump()),
                 flops=int(tree.flops()), numel=int(y.size), exact=exact)
        if exact:
            e["sha256"] = G.sha256(y)
        else:
            e["sample_stride"] = G.SAMPLE_STRIDE
            e["sample"] = [float(v) for v in G.sample(y)]
        entries.append(e)
        print(cid, "done", flush=True)
    with open(OUT_CONFIGS, "w") as f:
        json.dump(dict(source="reference loop_tool built from its sources (oracle/_ref)",
                       programs=entries), f, indent=1)
    print("wrote %d configurations to %s" % (len(entries), OUT_CONFIGS))


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit("build the reference library first: python -m oracle.build")
    ref = capi.Library(REF_LIB)
    if "--configs" in sys.argv[1:]:
        return configs(ref)
    entries = []
    for i, (name, kw, gen) in enumerate(PROGRAMS):
        out, shapes = getattr(P, name)(ref, **kw)
        inputs = gen(shapes)
        y, lt = run(ref, out, inputs)
        e = dict(id="%s%d" % (name, i), program=name, kw=kw, dump=norm_ids(lt.dump()),
                 flops=int(lt.flops()), exact=name != "softmax",
                 sha256=hashlib.sha256(np.ascontiguousarray(y, np.float32).tobytes()).hexdigest(),
                 numel=int(y.size))
        if y.size <= 4096:
            e["output"] = [float(v) for v in y]
        entries.append(e)
    transforms = []
    for prog in sorted(BUILDS):
        build, gen = BUILDS[prog]
        for step in sorted(STEPS):
            e = dict(id="%s/%s" % (prog, step), program=prog, step=step)
            try:
                out, shapes = build(ref)
                t = out.loop_tree
                for f in STEPS[step]:
                    t = f(t)
            except Exception:  # the reference rejects this transform here
                e["rejected"] = True
                transforms.append(e)
                continue
            inputs = gen(shapes)
            y = np.zeros(t.io_numel()[len(inputs)], np.float32)
            ref.compile(t, "cpp").run([np.ascontiguousarray(a) for a in inputs] + [y])
            e.update(dump=norm_ids(t.dump()), flops=int(t.flops()),
                     parallel=[bool(t.trivially_parallel(r)) for r in t.walk()],
                     exact=prog != "softmax",
                     sha256=hashlib.sha256(y.tobytes()).hexdigest(),
                     output=[float(v) for v in y])
            transforms.append(e)
    with open(OUT, "w") as f:
        json.dump(dict(source="reference loop_tool built from its sources (oracle/_ref)",
                       programs=entries, transforms=transforms), f, indent=1)
    print("wrote %d programs and %d transforms to %s" % (len(entries), len(transforms), OUT))


if __name__ == "__main__":
    main()
