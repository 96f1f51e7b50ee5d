"""Mixed nests for packs (hip.Pack, include/loop_tool_hip.h lth_pack_*).

`serving_mix` draws nests the way the serving workload draws requests: 40 %
matmul (M, N, K log-uniform in [32, 2048], multiples of 16, bf16 operands),
30 % row softmax (rows and columns uniform in [256, 8192]), 20 % 3x3 pad-1
convolutions (N 1-16, C_in and C_out 32-256 in steps of 32, H = W 14-112,
bf16 operands), 10 % 2x2 max-pools (N 1-16, C 32-256, H = W even in
14-112).  `compile_nests` compiles them with the hip backend, first in worker
processes (hipRTC compiles of one process run one at a time), which leave
the code objects in the backend's on-disk cache.
"""
import math

import numpy as np

from . import programs as P


def serving_mix(n=64, seed=7):
    rng = np.random.RandomState(seed)

    def logu(lo, hi, mult):
        v = math.exp(math.log(lo) + rng.rand() * (math.log(hi) - math.log(lo)))
        return int(min(hi, max(lo, round(v / mult) * mult)))

    out = []
    for _ in range(n):
        u = rng.rand()
        if u < 0.4:
            out.append(("matmul", dict(M=logu(32, 2048, 16), N=logu(32, 2048, 16),
                                       K=logu(32, 2048, 16)), "hip_bf16"))
        elif u < 0.7:
            out.append(("softmax", dict(R=int(rng.randint(256, 8193)),
                                        Cn=int(rng.randint(256, 8193))), "hip"))
        elif u < 0.9:
            h = int(rng.randint(14, 113))
            out.append(("conv2d", dict(N=int(rng.randint(1, 17)), Ci=32 * int(rng.randint(1, 9)),
                                       Co=32 * int(rng.randint(1, 9)), H=h, W=h), "hip_bf16"))
        else:
            h = 2 * int(rng.randint(7, 57))
            out.append(("maxpool2d", dict(N=int(rng.randint(1, 17)),
                                          C=32 * int(rng.randint(1, 9)), H=h, W=h), "hip"))
    return out


def _compile_worker(specs):
    import loop_tool_amd as lt_mod
    L = lt_mod.lib
    for name, kw, backend in specs:
        out, _ = getattr(P, name)(L, **kw)
        L.compile(out.loop_tree, backend)
    return len(specs)


def compile_nests(L, specs, workers=16):
    """[(Compiled, LoopTree, input shapes)] of the specs, compiled after a
    parallel warm-up of the code cache.  The warm-up is only a speed-up:
    when its worker processes cannot start (a calling script without an
    `if __name__ == "__main__":` guard, which spawned workers re-import), the
    nests are compiled in this process instead."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from concurrent.futures.process import BrokenProcessPool
    uniq = sorted(set((n, tuple(sorted(kw.items())), b) for n, kw, b in specs))
    uniq = [(n, dict(kw), b) for n, kw, b in uniq]
    workers = max(1, min(workers, len(uniq)))
    if workers > 1:
        try:
            with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
                list(ex.map(_compile_worker, [uniq[i::workers] for i in range(workers)]))
        except (BrokenProcessPool, RuntimeError):
            pass
    out = []
    for name, kw, backend in specs:
        prog, shapes = getattr(P, name)(L, **kw)
        out.append((L.compile(prog.loop_tree, backend), prog.loop_tree, shapes))
    return out
