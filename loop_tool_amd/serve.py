"""Python binding of the serving runtime (include/loop_tool_serve.h).

    from loop_tool_amd import serve, programs as P
    s = serve.Server(budget_gb=200, slots=8)
    out, shapes = P.softmax(lt, 1024, 4096)
    rid = s.submit(out, seeds=[3], scales=[3.0])
    s.drain(); print(s.info(rid))

Programs are lazy tensors of the C++ frontend (loop_tool_amd.capi); their
inputs are materialised on the device from (seed, scale) pairs with the same
SplitMix64 uniform as programs.uniform, so host references can be rebuilt
bit-for-bit.
"""
import ctypes as C

import numpy as np

from . import lib as _lib_obj
from .capi import _P

_L = _lib_obj.lib

FAMILIES = ["matmul", "softmax", "conv2d", "maxpool2d", "generic"]


class Config(C.Structure):
    _fields_ = [("budget_bytes", C.c_double), ("slots", C.c_int), ("keep_outputs", C.c_int),
                ("checksums", C.c_int), ("dtype", C.c_int), ("compiled_cap", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_int), ("family", C.c_int), ("state", C.c_int),
                ("arrival_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double),
                ("out_numel", C.c_int64), ("flops", C.c_int64), ("bytes", C.c_int64),
                ("checksum", C.c_uint64), ("plan", C.c_char * 96)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["plan"] = self.plan.decode()
        d["family"] = FAMILIES[self.family]
        d["latency_ms"] = (self.done_s - self.arrival_s) * 1e3 if self.state == 2 else None
        return d


class Stats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("wall_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("flops", C.c_double),
                ("bytes", C.c_double), ("family_count", C.c_int64 * 5),
                ("struct_hits", C.c_int64), ("struct_misses", C.c_int64),
                ("exact_hits", C.c_int64), ("compiles", C.c_int64),
                ("peak_bytes", C.c_double), ("admission_waits", C.c_int64),
                ("client_lag_ms", C.c_double), ("host_plan_ms", C.c_double),
                ("host_launch_ms", C.c_double), ("host_retire_ms", C.c_double),
                ("device_allocs", C.c_int64), ("cached_bytes", C.c_double),
                ("compiled_evictions", C.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["family_count"] = dict(zip(FAMILIES, list(self.family_count)))
        return d


class Workload(C.Structure):
    _fields_ = [("rate", C.c_double), ("duration_s", C.c_double), ("seed", C.c_uint64),
                ("mix", C.c_double * 4), ("max_requests", C.c_int)]


def _sig(name, res, *args):
    f = getattr(_L, name)
    f.restype = res
    f.argtypes = list(args)


_sig("lts_last_error", C.c_char_p)
_sig("lts_create", _P, C.POINTER(Config))
_sig("lts_free", None, _P)
_sig("lts_structural_hash", C.c_uint64, _P)
_sig("lts_template_hash", C.c_uint64, C.c_int)
_sig("lts_family", C.c_int, _P, _P)
_sig("lts_submit", C.c_int, _P, _P, C.POINTER(C.c_uint64), C.POINTER(C.c_float), C.c_int)
_sig("lts_poll", C.c_int, _P)
_sig("lts_drain", C.c_int, _P)
_sig("lts_info", C.c_int, _P, C.c_int, C.POINTER(RequestInfo))
_sig("lts_output", C.c_int, _P, C.c_int, C.POINTER(C.c_float), C.c_int64)
_sig("lts_release", C.c_int, _P, C.c_int)
_sig("lts_stats_get", C.c_int, _P, C.POINTER(Stats))
_sig("lts_stats_reset", C.c_int, _P)
_sig("lts_run_workload", C.c_int, _P, C.POINTER(Workload), C.POINTER(Stats))


class ServeError(RuntimeError):
    pass


def _chk(rc):
    if rc < 0:
        raise ServeError(_L.lts_last_error().decode())
    return rc


def structural_hash(program):
    """Key of a lazy program's structure with its sizes masked."""
    return _L.lts_structural_hash(program.h)


def template_hash(family):
    return _L.lts_template_hash(FAMILIES.index(family))


class Server:
    def __init__(self, budget_gb=200.0, slots=8, keep_outputs=False, checksums=False,
                 dtype="bf16", compiled_cap=0):
        cfg = Config(budget_gb * 1e9, slots, int(keep_outputs), int(checksums),
                     0 if dtype == "bf16" else 1, compiled_cap)
        self.h = _L.lts_create(C.byref(cfg))
        if not self.h:
            raise ServeError(_L.lts_last_error().decode())

    def family(self, program):
        return FAMILIES[_chk(_L.lts_family(self.h, program.h))]

    def submit(self, program, seeds, scales):
        n = len(seeds)
        s = (C.c_uint64 * max(n, 1))(*seeds)
        f = (C.c_float * max(n, 1))(*scales)
        return _chk(_L.lts_submit(self.h, program.h, s, f, n))

    def poll(self):
        return _chk(_L.lts_poll(self.h))

    def drain(self):
        return _chk(_L.lts_drain(self.h))

    def info(self, rid):
        i = RequestInfo()
        _chk(_L.lts_info(self.h, rid, C.byref(i)))
        return i.as_dict()

    def output(self, rid):
        n = self.info(rid)["out_numel"]
        y = np.empty(n, np.float32)
        _chk(_L.lts_output(self.h, rid, y.ctypes.data_as(C.POINTER(C.c_float)), n))
        return y

    def release(self, rid):
        _chk(_L.lts_release(self.h, rid))

    def stats(self):
        s = Stats()
        _chk(_L.lts_stats_get(self.h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _chk(_L.lts_stats_reset(self.h))

    def run_workload(self, rate=2000.0, duration_s=60.0, seed=7, mix=(0.4, 0.3, 0.2, 0.1),
                     max_requests=0):
        w = Workload(rate, duration_s, seed, (C.c_double * 4)(*mix), max_requests)
        s = Stats()
        _chk(_L.lts_run_workload(self.h, C.byref(w), C.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None):
            _L.lts_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
