"""Schedule autotuner for the hip backend and its tuned-schedule store.

The search space is the LoopTree itself.  Candidates come from the
reference's schedule transforms (src/frontends/mutate.cpp: split :226-343,
swap :345-481, unroll_inner_loops :836-863) applied to the program's default
tree, and the hip lowering reads each schedule as follows (explicit knob
annotations take precedence):

  split        a var split into nested loops gives its innermost loop's
               extent: for contractions the tile's rows / columns (the row
               and column var, 64..256) and the K tile (the innermost reduced
               var, 32 or 64); for row reductions and pointwise kernels the
               values per lane of the contiguous var (1, 2, 4) and, one loop
               level up, the lanes per workgroup (64..1024)
  swap         the pointwise kernels accumulate reductions in the order of
               the tree's reduction loops
  unroll       reduction loops annotated "unroll" are fully unrolled in the
               pointwise kernels

Every candidate is compiled, checked against a baseline result on the same
seeded inputs (the harness of the reference's test/test.py:85-123: a
schedule must not change the result beyond the stated tolerance), and timed
with HIP events on device-resident buffers.  The winner goes to the
tuned-schedule store, which the hip backend's compile consults for every
program (include/loop_tool_hip.h, lth_tuned_*), so a later process -- the
lazy frontend included -- runs the tuned schedule without searching.
"""
import ctypes as C
import itertools
import os

import numpy as np

_P = C.c_void_p


def _lib(L):
    lib = L.lib
    if not getattr(lib, "_lt_tuned_sigs", False):
        for name, res, args in (("lth_tuned_set_dir", C.c_int, [C.c_char_p]),
                                ("lth_schedule_key", C.c_int, [_P, C.c_char_p, C.c_int]),
                                ("lth_tuned_put", C.c_int, [_P, _P]),
                                ("lth_tuned_get", _P, [_P]),
                                ("lth_last_error", C.c_char_p, [])):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        lib._lt_tuned_sigs = True
    return lib


class TunedStore:
    """The hip backend's tuned-schedule store: one IR text file per program
    under <path>/gfx950/<key>.ir.  Opening a store directs the backend's own
    lookups (Backend::compile) in this process to it; `reset` returns them
    to $LT_HIP_TUNED_DIR (else $HOME/.cache/loop_tool_amd/tuned)."""

    def __init__(self, L, path=None):
        self.L = L
        self.path = path or os.environ.get("LT_HIP_TUNED_DIR") or os.path.join(
            os.path.expanduser("~"), ".cache", "loop_tool_amd", "tuned")
        lib = _lib(L)
        if lib.lth_tuned_set_dir(self.path.encode()) != 0:
            raise RuntimeError(lib.lth_last_error().decode())

    @staticmethod
    def reset(L):
        _lib(L).lth_tuned_set_dir(b"")

    def key(self, program):
        lib = _lib(self.L)
        buf = C.create_string_buffer(64)
        if lib.lth_schedule_key(program.h, buf, 64) < 0:
            raise RuntimeError(lib.lth_last_error().decode())
        return buf.value.decode()

    def put(self, program, schedule):
        lib = _lib(self.L)
        if lib.lth_tuned_put(program.h, schedule.h) != 0:
            raise RuntimeError(lib.lth_last_error().decode())
        return self.key(program)

    def get(self, program):
        from .capi import LoopTree
        h = _lib(self.L).lth_tuned_get(program.h)
        return LoopTree(self.L, h) if h else None

    def keys(self):
        d = os.path.join(self.path, "gfx950")
        if not os.path.isdir(d):
            return []
        return sorted(f[:-3] for f in os.listdir(d) if f.endswith(".ir"))


# ------------------------------------------------------- schedule moves ----
def leaves(tree, op=None):
    return [r for r in tree.walk() if not tree.is_loop(r) and (op is None or tree.node_op(r) == op)]


def loops_above(tree, ref):
    out = []
    p = tree.parent(ref)
    while p != -1:
        out.append(p)
        p = tree.parent(p)
    return out[::-1]  # outer to inner


def split_var(tree, leaf_op, var, size, nth=0):
    """Split the innermost loop over `var` above the nth leaf computing
    `leaf_op` by `size` (mutate.cpp:226-343); None when there is no such
    loop or it is not longer than `size`."""
    ls = leaves(tree, leaf_op)
    if len(ls) <= nth:
        return None
    cand = [r for r in loops_above(tree, ls[nth]) if tree.loop(r)["var"] == var]
    if not cand or tree.loop(cand[-1])["size"] <= size:
        return None
    return tree.split(cand[-1], size)


def _reduced_vars(tree, leaf):
    node_vars = set(tree.node_vars(leaf))
    return [tree.loop(r)["var"] for r in loops_above(tree, leaf)
            if tree.loop(r)["var"] not in node_vars]


def schedule_candidates(tree, kind):
    """(description, LoopTree) pairs: schedules of `tree` built with the
    mutate moves, chosen for what the lowering of plan kind `kind` reads."""
    out = []

    def add(desc, t):
        if t is not None:
            out.append((desc, t))

    if kind.startswith("contract"):
        adds = leaves(tree, "add")
        if not adds:
            return out
        red = adds[0]
        nonred = tree.node_vars(red)
        kvars = _reduced_vars(tree, red)
        extent = {}
        for r in loops_above(tree, red):
            lp = tree.loop(r)
            extent[lp["var"]] = extent.get(lp["var"], 1) * max(lp["size"], 1)
        names = {v: tree.var_name(v) for v in list(nonred) + kvars}
        tiles = (64, 128, 256)
        combos = [[(v, s)] for v in nonred for s in tiles]
        if len(nonred) == 2:
            combos += [[(nonred[0], a), (nonred[1], b)] for a in tiles for b in tiles]
        kv = max(kvars, key=lambda v: extent.get(v, 0)) if kvars else None
        for combo in combos:
            t, desc = tree, []
            for v, sz in combo:
                t = split_var(t, "add", v, sz) if t is not None else None
                desc.append("split %s %d" % (names[v], sz))
            if t is None:
                continue
            add(", ".join(desc), t)
            if kv is not None and len(combo) == len(nonred[:2]):
                for bk in (32, 64):
                    add(", ".join(desc + ["split %s %d" % (names[kv], bk)]),
                        split_var(t, "add", kv, bk))
        return out

    writes = leaves(tree, "write")
    if not writes:
        return out
    w = writes[-1]
    c = tree.node_vars(w)[-1]
    name = tree.var_name(c)
    if kind == "rowreduce":
        for vec, lanes in itertools.product((1, 2, 4), (64, 128, 256, 512, 1024)):
            t = split_var(tree, "write", c, vec * lanes)
            if t is not None:
                t = split_var(t, "write", c, vec)
            add("split %s %d x %d" % (name, lanes, vec), t)
        return out
    # pointwise: values per lane and lanes per workgroup of the contiguous
    # var; unrolled reduction loops; swapped loops
    for vec, block in itertools.product((1, 2, 4), (128, 256, 512)):
        t = split_var(tree, "write", c, vec * block)
        if t is not None:
            t = split_var(t, "write", c, vec)
        add("split %s %d x %d" % (name, block, vec), t)
    for n in (4, 16):
        try:
            add("unroll %d" % n, tree.unroll_inner_loops(n))
        except Exception:
            pass
    for r in tree.loops:
        for ch in tree.children(r):
            if tree.is_loop(ch):
                try:
                    add("swap L%d L%d" % (r, ch), tree.try_swap(r, ch))
                except Exception:
                    pass
    return out


def knob_candidates(kind):
    """Annotation knob sets worth trying for a plan kind (the backend's
    launch knobs beyond what a schedule expresses)."""
    if kind == "rowreduce":
        return ["block=%d vec=%d" % (b, v) for b, v in
                itertools.product((64, 256, 1024), (4, 2))]
    if kind == "contract_f32":  # staged float tile, or the gathered tile
        return ["contract=staged", "contract=staged waves=8", "contract=staged streamk=0",
                "contract=staged streamk=1", "contract=gather"]
    if kind.startswith("contract"):  # staged 16-bit operands
        return ["stage=flat", "waves=8", "waves=2", "igemm=halo", "igemm=plain",
                "tile=256x256 bk=64", "tile=128x256", "srows=2"]
    if kind == "pointwise":
        return ["block=%d vec=%d" % (b, v) for b, v in itertools.product((256,), (4, 1))]
    return [""]


def annotate_all(tree, annot):
    """Annotate the outermost loop (the knobs are read from any annotation)."""
    return tree.annotate(tree.roots[0], annot) if tree.roots else tree


def tune(L, tree, inputs, baseline=None, candidates=None, tol=1e-5, iters=20,
         backend="hip", extra_trees=(), schedules=True, store=None):
    """Search schedules of `tree` (mutate moves, plus annotation knob sets
    and `extra_trees`) and return (best_tree, best_ms, report).

    inputs: numpy float32 arrays; baseline: expected output (default: the
    untuned hip result).  Candidates that fail to compile, or whose result
    differs from the baseline by more than `tol` (normwise), are rejected.
    With a `store` (TunedStore), the winner is stored as the program's tuned
    schedule when it beats the tree as given."""
    from . import hip
    d_in = [hip.DeviceArray.from_numpy(a) for a in inputs]
    n_out = tree.io_numel()[len(inputs)]
    d_out = hip.DeviceArray(n_out)
    io = d_in + [d_out]
    saved = os.environ.get("LT_HIP_TUNED")
    os.environ["LT_HIP_TUNED"] = "0"  # candidates compile as given

    def measure(t):
        cc = L.compile(t, backend)
        cc.run([b.ptr for b in io])
        hip.synchronize()
        y = d_out.download()
        ms = hip.time_ms(cc, io, warmup=2, iters=iters)
        return cc, y, ms

    try:
        cc0, y0, ms0 = measure(tree)
        if baseline is None:
            baseline = y0
        scale = max(float(np.abs(baseline).max()), 1e-30)
        kind = cc0.string_property("kinds").split(",")[-1]
        trees = []
        if schedules:
            trees += schedule_candidates(tree, kind)
        cands = list(candidates) if candidates is not None else knob_candidates(kind)
        trees += [("knobs " + k, annotate_all(tree, k)) for k in cands if k]
        trees += [("schedule%d" % i, t) for i, t in enumerate(extra_trees)]
        report = [dict(schedule="as given", ms=ms0, kind=kind,
                       err=float(np.abs(y0 - baseline).max() / scale))]
        best = (ms0, tree, "as given")
        for desc, t in trees:
            try:
                cc, y, ms = measure(t)
            except Exception as e:  # not lowerable with these knobs
                report.append(dict(schedule=desc, error=str(e).splitlines()[0][:200]))
                continue
            err = float(np.abs(y - baseline).max() / scale)
            ok = err <= tol
            report.append(dict(schedule=desc, ms=ms, err=err, ok=ok,
                               kind=cc.string_property("kinds").split(",")[-1],
                               plan=cc.dump().splitlines()[0][:160]))
            if ok and ms < best[0]:
                best = (ms, t, desc)
    finally:
        if saved is None:
            os.environ.pop("LT_HIP_TUNED", None)
        else:
            os.environ["LT_HIP_TUNED"] = saved
    key = None
    if store is not None and best[2] != "as given":
        key = store.put(tree, best[1])
    return best[1], best[0], dict(best=best[2], candidates=report, key=key)
