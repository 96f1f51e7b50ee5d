"""Batches of independent matmul nests of uneven, run-time sizes.

Workload: n GEMMs C_i = A_i B_i with M, N, K drawn log-uniform in [lo, hi]
and rounded to multiples of 16 (SplitMix64 seed), 16-bit (bf16 or fp16)
operands resident in HBM, float32 accumulation and output.

Each GEMM is a loop_tool matmul nest; the autotuner picks a schedule per
shape bucket -- the nest split by the tile (BM, BN, BK) and annotated
`tile=BMxBN bk=BK` -- by timing the whole batch's grouped launch under
candidate tile configurations, and stores the tuned LoopTree (IR text) per
bucket.  The tuned batch then
runs as one persistent launch over every output tile of every GEMM, captured
once into a hipGraph; the plain form is one launch per GEMM with the default
128x128 tile.
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
            A16, Bt16 = DeviceArray(m * k // 2 + 1), DeviceArray(nn * k // 2 + 1)
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
    Returns the per-GEMM configurations and a report; stores one tuned
    LoopTree per bucket."""
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
            m, nn, k = batch.dims[groups[key][0]]
            store.put("matmul-bucket-%dx%d-%s" % (2 ** key[0], 2 ** key[1],
                                                  "bf16" if batch.dtype == 0 else "f16"),
                      tuned_tree(L, m, nn, k, c), best, "tile=%dx%d bk=%d" % (
                          TILES[c] + (TILE_K[c],)))
    return cfg, report


def tuned_tree(L, M, N, K, cfg):
    """The matmul nest scheduled for tile configuration `cfg`: M and N split
    by the tile, K split by 32, annotated with the tile for the hip backend."""
    out, _ = P.matmul(L, M, N, K)
    t = out.loop_tree
    bm, bn = TILES[cfg]
    if M > bm:
        t = t.split(t.roots[0], bm)
    return t.annotate(t.roots[0], "tile=%dx%d bk=%d mfma_bf16" % (bm, bn, TILE_K[cfg]))
