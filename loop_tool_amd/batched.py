"""Batches of independent matmul nests of uneven, run-time sizes.

Workload: n GEMMs C_i = A_i B_i with M, N, K drawn log-uniform in [lo, hi]
and rounded to multiples of 16 (SplitMix64 seed), 16-bit (bf16 or fp16)
operands resident in HBM, float32 accumulation and output.

Each GEMM is a loop_tool matmul nest; the autotuner picks a schedule per
shape bucket -- the bucket's representative matmul nest split by the tile
(BM, BN, BK) -- by timing the whole batch's grouped launch under candidate
tile configurations, and stores that schedule in the hip backend's tuned
store.  A batch takes its tile configurations from the stored schedules:
each is lowered by the hip backend, and the tile it chooses is the GEMMs'
configuration.  The tuned batch runs as one persistent launch over every
output tile of every GEMM, captured once into a hipGraph.  The plain forms
are one hip-backend Compiled::run per matmul nest (PlainNests) and one
launch of the library's tile per GEMM.
"""
import ctypes as C
import math

import numpy as np

from . import hip, programs as P
from .hip import _L, _P, DeviceArray, HipError

_L.lth_batch_last_error.restype = C.c_char_p
_L.lth_batch_create.restype = _P
_L.lth_batch_create.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                C.POINTER(C.c_int), C.c_int]
for _n, _a in (("lth_batch_set_operands", [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
               ("lth_batch_set_configs", [_P, C.POINTER(C.c_int)]),
               ("lth_batch_num_items", [_P]),
               ("lth_batch_run_plain", [_P, C.c_int, _P]),
               ("lth_batch_run_grouped", [_P, C.c_int, _P]),
               ("lth_batch_capture", [_P, C.c_int, _P]),
               ("lth_batch_replay", [_P, _P]),
               ("lth_convert", [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P])):
    getattr(_L, _n).restype = C.c_int
    getattr(_L, _n).argtypes = _a
_L.lth_batch_free.restype = None
_L.lth_batch_free.argtypes = [_P]

TILES = [(64, 64), (128, 64), (64, 128), (128, 128), (256, 128), (128, 256), (128, 128)]
TILE_K = [32, 32, 32, 64, 32, 32, 32]
DEFAULT_CFG = 3


def _chk(rc):
    if rc < 0:
        raise HipError(_L.lth_batch_last_error().decode())
    return rc


def pitch(k):
    """Row pitch (elements) of the batch's 16-bit operands: K rounded up to
    a multiple of 64 (include/loop_tool_hip.h, LTH_BATCH_PITCH)."""
    return (k + 63) // 64 * 64


def shapes(n=512, seed=6, lo=32, hi=2048, mult=16):
    u = P.uniform(seed, (n, 3)).astype(np.float64) * 0.5 + 0.5  # [0, 1)
    v = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    v = np.clip(np.round(v / mult) * mult, lo, hi).astype(np.int64)
    return [tuple(int(x) for x in row) for row in v]  # (M, N, K)


def bucket(m, n, k):
    return (max(5, math.ceil(math.log2(m))), max(5, math.ceil(math.log2(n))))


class Batch:
    """Device operands and the native batch plan for a list of (M, N, K)."""

    def __init__(self, dims, dtype="bf16", seed=6, stream=None):
        self.dims = list(dims)
        self.dtype = 0 if dtype == "bf16" else 1
        if stream is None:  # graph capture needs a stream of its own
            self._own_stream = hip.Stream()
            stream = self._own_stream.h
        self.stream = stream
        n = len(self.dims)
        arr = lambda i: (C.c_int * n)(*[d[i] for d in self.dims])
        self.h = _L.lth_batch_create(n, arr(0), arr(1), arr(2), self.dtype)
        if not self.h:
            raise HipError(_L.lth_batch_last_error().decode())
        self.A, self.Bt, self.C, self.host = [], [], [], []
        for i, (m, nn, k) in enumerate(self.dims):
            a = P.uniform(seed * 1000003 + 2 * i, (m, k))
            b = P.uniform(seed * 1000003 + 2 * i + 1, (k, nn))
            self.host.append((a, b))
            a_d, b_d = DeviceArray.from_numpy(a), DeviceArray.from_numpy(b)
            kp = pitch(k)  # rows padded to whole 128-byte lines
            A16, Bt16 = DeviceArray(m * kp // 2 + 1), DeviceArray(nn * kp // 2 + 1)
            _chk(_L.lth_convert(a_d.ptr, A16.ptr, m, k, 0, self.dtype, self.stream))
            _chk(_L.lth_convert(b_d.ptr, Bt16.ptr, k, nn, 1, self.dtype, self.stream))
            _L.lt_hip_stream_synchronize(self.stream)  # before a_d, b_d are freed
            self.A.append(A16)
            self.Bt.append(Bt16)
            self.C.append(DeviceArray(m * nn))
        hip.synchronize()
        ptrs = lambda xs: (_P * n)(*[x.ptr for x in xs])
        _chk(_L.lth_batch_set_operands(self.h, ptrs(self.A), ptrs(self.Bt), ptrs(self.C)))
        self.cfg = [DEFAULT_CFG] * n

    def flops(self):
        return sum(2 * m * n * k for m, n, k in self.dims)

    def set_configs(self, cfg):
        self.cfg = list(cfg)
        _chk(_L.lth_batch_set_configs(self.h, (C.c_int * len(cfg))(*cfg)))

    def run_plain(self, cfg=DEFAULT_CFG):
        _chk(_L.lth_batch_run_plain(self.h, cfg, self.stream))

    def run_grouped(self, grid=0):
        _chk(_L.lth_batch_run_grouped(self.h, grid, self.stream))

    def capture(self, grid=0):
        _chk(_L.lth_batch_capture(self.h, grid, self.stream))

    def replay(self):
        _chk(_L.lth_batch_replay(self.h, self.stream))

    def num_items(self):
        return _chk(_L.lth_batch_num_items(self.h))

    def synchronize(self):
        _L.lt_hip_stream_synchronize(self.stream)

    def result(self, i):
        self.synchronize()
        m, n, _ = self.dims[i]
        return self.C[i].download((m, n))

    def check(self, idx, rounding=True):
        """(error vs float64 of the float32 operands, error vs float64 of the
        16-bit-rounded operands) for GEMM idx, normwise."""
        a, b = self.host[idx]
        y = self.result(idx).astype(np.float64)
        ref = a.astype(np.float64) @ b.astype(np.float64)
        e = np.abs(y - ref).max() / np.abs(ref).max()
        if self.dtype == 0:
            r = lambda x: _bf16(x)
        else:
            r = lambda x: x.astype(np.float16).astype(np.float32)
        ref16 = r(a).astype(np.float64) @ r(b).astype(np.float64)
        e16 = np.abs(y - ref16).max() / np.abs(ref16).max()
        return float(e), float(e16)

    def free(self):
        if getattr(self, "h", None):
            _L.lth_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _bf16(a):
    u = a.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def time_ms(fn, iters=10, stream=None):
    e0, e1 = hip.Event(), hip.Event()
    fn()
    hip.synchronize()
    e0.record(stream)
    for _ in range(iters):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / iters


def tune_buckets(batch, grid=0, iters=3, store=None, L=None, rounds=1):
    """Choose a tile configuration per shape bucket *in the context of the
    whole batch*: start from the best uniform configuration, then, bucket by
    bucket (most flops first), keep any configuration that makes the whole
    grouped launch faster by more than 1%.  (Timing a bucket alone favours
    small tiles, which only pay off when the bucket has the GPU to itself.)
    Returns the per-GEMM configurations and a report.  With a `store` (an
    autotune.TunedStore) and `L`, each bucket's choice is stored as a
    schedule of the bucket's representative matmul nest (bucket_program):
    the nest split by the tile and the K tile."""
    groups = {}
    for i, (m, n, k) in enumerate(batch.dims):
        groups.setdefault(bucket(m, n, k), []).append(i)
    n = len(batch.dims)
    stream = batch.stream

    def measure(cfg):
        batch.set_configs(cfg)
        return time_ms(lambda: batch.run_grouped(grid), iters, stream)

    uniform = {c: measure([c] * n) for c in range(len(TILES))}
    best_c = min(uniform, key=uniform.get)
    cfg = [best_c] * n
    best = uniform[best_c]
    order = sorted(groups, key=lambda key: -sum(2 * batch.dims[i][0] * batch.dims[i][1]
                                                 * batch.dims[i][2] for i in groups[key]))
    for _ in range(rounds):
        for key in order:
            idx = groups[key]
            for c in range(len(TILES)):
                if c == cfg[idx[0]]:
                    continue
                trial = list(cfg)
                for i in idx:
                    trial[i] = c
                t = measure(trial)
                if t < best * 0.99:
                    best, cfg = t, trial
    batch.set_configs(cfg)
    report = dict(uniform_ms={"%dx%dx%d" % (TILES[c] + (TILE_K[c],)): round(t, 4)
                              for c, t in uniform.items()},
                  tuned_ms=round(best, 4), buckets={})
    for key in order:
        c = cfg[groups[key][0]]
        report["buckets"]["%dx%d" % (2 ** key[0], 2 ** key[1])] = dict(
            gemms=len(groups[key]), tile="%dx%dx%d" % (TILES[c] + (TILE_K[c],)))
        if store is not None and L is not None:
            dt = "bf16" if batch.dtype == 0 else "f16"
            prog = bucket_program(L, key, dt)
            store.put(prog, tuned_tree(L, prog, c))
    return cfg, report


def bucket_program(L, key, dtype="bf16"):
    """The representative matmul nest of a shape bucket (one per bucket, so
    that each bucket keeps a schedule of its own): M and N are 16 times the
    bucket's upper bounds 2^key[0] and 2^key[1] (every tile of the library
    is then a split of their loops), K = 512; fp16 batches annotate it
    `mfma_f16`, so the two operand types keep schedules of their own."""
    out, _ = P.matmul(L, 16 * 2 ** key[0], 16 * 2 ** key[1], 512)
    t = out.loop_tree
    return t if dtype == "bf16" else t.annotate(t.roots[0], "mfma_f16")


def tuned_tree(L, program, cfg):
    """A schedule of a matmul nest for tile configuration `cfg`: its M, N
    and K loops split by the tile's rows, columns and K tile (what the hip
    lowering reads as the tile of a contraction)."""
    from .autotune import split_var, leaves
    t = program
    red = leaves(t, "add")[0]
    m, n = t.node_vars(red)
    k = next(t.loop(r)["var"] for r in t.loops if t.loop(r)["var"] not in (m, n))
    bm, bn = TILES[cfg]
    for var, size in ((m, bm), (n, bn), (k, TILE_K[cfg])):
        nt = split_var(t, "add", var, size)
        if nt is None:
            raise ValueError("matmul nest too small for the tile %dx%dx%d" % (
                bm, bn, TILE_K[cfg]))
        t = nt
    return t


def config_of(L, schedule, dtype="bf16"):
    """The library's tile configuration that the hip backend lowers
    `schedule` to: it is compiled for 16-bit operands and its tile read back
    from the Compiled."""
    backend = "hip_bf16" if dtype == "bf16" else "hip"
    t = schedule
    if dtype != "bf16" and "mfma_f16" not in t.dump():
        t = t.annotate(t.roots[0], "mfma_f16")
    cc = L.compile(t, backend)
    got = (cc.int_property("tile_m"), cc.int_property("tile_n"), cc.int_property("bk"))
    for c, (tile, bk) in enumerate(zip(TILES, TILE_K)):
        if tile + (bk,) == got:
            return c
    raise ValueError("no tile configuration of the batch library is %dx%dx%d" % got)


def configs_from_store(L, store, dims, dtype="bf16"):
    """Per-GEMM tile configurations from the schedules stored for the
    GEMMs' shape buckets (None where a bucket has no stored schedule)."""
    by_bucket = {}
    out = []
    for m, n, k in dims:
        key = bucket(m, n, k)
        if key not in by_bucket:
            sched = store.get(bucket_program(L, key, dtype))
            by_bucket[key] = None if sched is None else config_of(L, sched, dtype)
        out.append(by_bucket[key])
    return out


def _matmul_nest(L, dims, dtype):
    out, _ = P.matmul(L, *dims)
    t = out.loop_tree
    return t if dtype == "bf16" else t.annotate(t.roots[0], "mfma_f16")


def _compile_nests(args):
    """Worker: compile matmul nests with the hip backend, which leaves their
    code objects in the on-disk code cache.  No GPU is touched."""
    dims, dtype = args
    import loop_tool_amd as lt_mod
    L = lt_mod.lib
    backend = "hip_bf16" if dtype == "bf16" else "hip"
    for d in dims:
        L.compile(_matmul_nest(L, d, dtype), backend)
    return len(dims)


def prewarm_code_cache(dims, dtype="bf16", workers=16):
    """Compile the matmul nests of `dims` in worker processes into the hip
    backend's on-disk code cache (hipRTC compiles of one process run one at
    a time; processes run in parallel), so that compiling them here only
    loads cached code objects.  Returns the number compiled there: 0 when
    the workers cannot start (a calling script without an
    `if __name__ == "__main__":` guard), and then they compile on first use."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from concurrent.futures.process import BrokenProcessPool
    dims = sorted(set(tuple(d) for d in dims))
    workers = max(1, min(workers, len(dims)))
    try:
        with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
            return sum(ex.map(_compile_nests, [(dims[i::workers], dtype)
                                               for i in range(workers)]))
    except (BrokenProcessPool, RuntimeError):
        return 0


class PlainNests:
    """The plain form of the batch: one matmul nest per GEMM compiled by the
    hip backend (default schedule, 16-bit MFMA operands staged from float32
    inputs in HBM by the plan itself), launched one after another on one
    stream with lth_run_on and shared scratch.  `workers` processes compile
    the nests into the on-disk code cache first (none: compile here)."""

    def __init__(self, L, batch, dtype="bf16", workers=16):
        self.batch = batch
        backend = "hip_bf16" if dtype == "bf16" else "hip"
        if workers:
            prewarm_code_cache(batch.dims, dtype, workers)
        self.cc = [L.compile(_matmul_nest(L, d, dtype), backend) for d in batch.dims]
        sizes = [hip.scratch_sizes(c) for c in self.cc]
        nsc = max(len(s) for s in sizes)
        self.scratch = [DeviceArray(max(s[j] for s in sizes if len(s) > j)) for j in range(nsc)]
        self.io = []
        for i, (m, n, k) in enumerate(batch.dims):
            a, b = batch.host[i]
            self.io.append([DeviceArray.from_numpy(a), DeviceArray.from_numpy(b),
                            DeviceArray(m * n)])
        hip.synchronize()

    def run(self):
        for cc, io in zip(self.cc, self.io):
            hip.run_on(cc, io, self.scratch[:len(hip.scratch_sizes(cc))], self.batch.stream)

    def result(self, i):
        self.batch.synchronize()
        m, n, _ = self.batch.dims[i]
        return self.io[i][2].download((m, n))
