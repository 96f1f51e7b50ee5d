"""Schedule autotuner for the hip backend and the on-disk tuned-schedule store.

The search space is the LoopTree itself: candidate schedules are produced with
the reference's schedule transforms (split, swap, merge, annotate;
src/frontends/mutate.cpp) plus the hip backend's annotations (block=N, vec=N,
tile=MxN, bk=N, waves=N, stage=..., igemm=..., hip:plain, mfma_bf16).  Every
candidate is compiled, checked against
a baseline result on the same seeded inputs (the harness of the reference's
test/test.py:85-123: a schedule must not change the result beyond the stated
tolerance), and timed with HIP events on device-resident buffers.  The winner
is stored as IR text (loop_tool::serialize) keyed by the program's
shape-structural hash, so a later process reuses it without searching.
"""
import itertools
import json
import os

import numpy as np


def _default_store_path():
    return os.environ.get("LT_TUNED_STORE") or os.path.join(
        os.path.expanduser("~"), ".cache", "loop_tool_amd", "tuned.jsonl")


class TunedStore:
    """Append-only JSON-lines file: {"key", "arch", "ir", "ms", "knobs"}."""

    def __init__(self, path=None):
        self.path = path or _default_store_path()
        self.entries = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                for line in f:
                    line = line.strip()
                    if line:
                        e = json.loads(line)
                        self.entries[(e["key"], e["arch"])] = e

    def get(self, key, arch="gfx950"):
        return self.entries.get((str(key), arch))

    def put(self, key, tree, ms, knobs, arch="gfx950"):
        e = dict(key=str(key), arch=arch, ir=tree.serialize(), ms=ms, knobs=knobs)
        self.entries[(str(key), arch)] = e
        os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
        with open(self.path, "a") as f:
            f.write(json.dumps(e) + "\n")
        return e

    def load_tree(self, L, key, arch="gfx950"):
        e = self.get(key, arch)
        return None if e is None else L.deserialize(e["ir"])


def annotate_all(tree, annot):
    """Annotate the outermost loop (the knobs are read from any annotation)."""
    return tree.annotate(tree.roots[0], annot) if tree.roots else tree


def knob_candidates(kind):
    """Knob sets worth trying for a plan kind."""
    if kind == "rowreduce":
        return ["block=%d vec=%d" % (b, v) for b, v in
                itertools.product((64, 128, 256, 512, 1024), (4, 2, 1))]
    if kind == "contract_f32":  # staged float tile, or the gathered tile
        return ["tile=%dx%d" % t for t in ((64, 64), (128, 64), (64, 128), (128, 128))] + [
            "contract=staged", "contract=staged tile=128x64", "contract=staged waves=8",
            "contract=staged streamk=0", "contract=staged streamk=1", "contract=gather"]
    if kind.startswith("contract"):  # staged 16-bit operands
        tiles = ["tile=128x256", "tile=256x128", "tile=128x128 bk=64", "tile=256x256 bk=64",
                 "tile=64x128", "tile=128x64"]
        return tiles + ["stage=flat", "waves=8", "waves=2", "igemm=halo",
                        "igemm=halo tile=128x128", "igemm=plain"]
    if kind == "pointwise":
        return ["block=%d vec=%d" % (b, v) for b, v in
                itertools.product((128, 256, 512), (4, 2, 1))]
    return [""]


def tune(L, tree, inputs, baseline=None, candidates=None, tol=1e-5, iters=20,
         backend="hip", extra_trees=()):
    """Search annotated variants of `tree` (plus `extra_trees`, e.g. split or
    swapped schedules) and return (best_tree, best_ms, report).

    inputs: numpy float32 arrays; baseline: expected output (default: the
    untuned hip result).  Candidates that fail to compile, or whose result
    differs from the baseline by more than `tol` (normwise), are rejected."""
    from . import hip
    d_in = [hip.DeviceArray.from_numpy(a) for a in inputs]
    n_out = tree.io_numel()[len(inputs)]
    d_out = hip.DeviceArray(n_out)
    io = d_in + [d_out]

    def measure(t):
        cc = L.compile(t, backend)
        cc.run([b.ptr for b in io])
        hip.synchronize()
        y = d_out.download()
        ms = hip.time_ms(cc, io, warmup=2, iters=iters)
        return cc, y, ms

    cc0, y0, ms0 = measure(tree)
    if baseline is None:
        baseline = y0
    scale = max(float(np.abs(baseline).max()), 1e-30)
    kind = cc0.string_property("kinds").split(",")[-1]
    cands = list(candidates) if candidates is not None else knob_candidates(kind)
    report = [dict(knobs="", ms=ms0, kind=kind,
                   err=float(np.abs(y0 - baseline).max() / scale))]
    best = (ms0, tree, "")
    trees = [(annotate_all(tree, k), k) for k in cands if k]
    trees += [(t, "schedule%d" % i) for i, t in enumerate(extra_trees)]
    for t, k in trees:
        try:
            cc, y, ms = measure(t)
        except Exception as e:  # not lowerable with these knobs
            report.append(dict(knobs=k, error=str(e).splitlines()[0][:200]))
            continue
        err = float(np.abs(y - baseline).max() / scale)
        ok = err <= tol
        report.append(dict(knobs=k, ms=ms, err=err, ok=ok,
                           kind=cc.string_property("kinds").split(",")[-1]))
        if ok and ms < best[0]:
            best = (ms, t, k)
    return best[1], best[0], dict(best_knobs=best[2], candidates=report)
