"""Device memory, streams and events through the C ABI of include/lt_hip.h.

Used by tests and the benchmark to keep inputs resident in HBM and to time
kernels with HIP events on the stream the kernels are launched on.
"""
import ctypes as C

import numpy as np

from . import lib as _lib

_L = _lib.lib
_P = C.c_void_p


def _sig(name, res, *args):
    f = getattr(_L, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("lt_hip_last_error", C.c_char_p)
_sig("lt_hip_init", C.c_int)
_sig("lt_hip_device_count", C.c_int, C.POINTER(C.c_int))
_sig("lt_hip_malloc", C.c_int, C.POINTER(_P), C.c_size_t)
_sig("lt_hip_free", C.c_int, _P)
_sig("lt_hip_memcpy", C.c_int, _P, _P, C.c_size_t, C.c_int)
_sig("lt_hip_memset", C.c_int, _P, C.c_int, C.c_size_t)
_sig("lt_hip_is_device_pointer", C.c_int, _P)
_sig("lt_hip_mem_info", C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t))
_sig("lt_hip_stream_create", C.c_int, C.POINTER(_P))
_sig("lt_hip_stream_destroy", C.c_int, _P)
_sig("lt_hip_stream_synchronize", C.c_int, _P)
_sig("lt_hip_device_synchronize", C.c_int)
_sig("lt_hip_event_create", C.c_int, C.POINTER(_P))
_sig("lt_hip_event_destroy", C.c_int, _P)
_sig("lt_hip_event_record", C.c_int, _P, _P)
_sig("lt_hip_event_synchronize", C.c_int, _P)
_sig("lt_hip_event_elapsed_ms", C.c_int, _P, _P, C.POINTER(C.c_float))


class HipError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise HipError(_L.lt_hip_last_error().decode())


class Props(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64),
                ("compute_units", C.c_int), ("max_threads_per_block", C.c_int),
                ("warp_size", C.c_int), ("clock_khz", C.c_int),
                ("memory_clock_khz", C.c_int), ("memory_bus_width", C.c_int),
                ("l2_cache_bytes", C.c_int), ("lds_bytes_per_block", C.c_int),
                ("total_global_mem", C.c_size_t)]


_sig("lt_hip_device_props", C.c_int, C.c_int, C.POINTER(Props))


def available():
    n = C.c_int(0)
    if _L.lt_hip_device_count(C.byref(n)) != 0:
        return False
    return n.value > 0


def init():
    _check(_L.lt_hip_init())


def props(device=0):
    p = Props()
    _check(_L.lt_hip_device_props(device, C.byref(p)))
    return {k: (getattr(p, k).decode() if isinstance(getattr(p, k), bytes) else getattr(p, k))
            for k, _ in Props._fields_}


def synchronize():
    _check(_L.lt_hip_device_synchronize())


def stream_synchronize(stream):
    """Wait for the work queued on a raw stream handle."""
    _check(_L.lt_hip_stream_synchronize(stream))


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    _check(_L.lt_hip_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


class DeviceArray:
    """A float32 buffer in HBM."""

    def __init__(self, numel, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.numel = int(numel)
        self.nbytes = self.numel * self.dtype.itemsize
        p = _P()
        _check(_L.lt_hip_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @staticmethod
    def from_numpy(a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        _check(_L.lt_hip_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0))

    def download(self, shape=None):
        out = np.empty(self.numel, self.dtype)
        _check(_L.lt_hip_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out.reshape(shape) if shape is not None else out

    def fill_zero(self):
        _check(_L.lt_hip_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            _L.lt_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _P()
        _check(_L.lt_hip_event_create(C.byref(p)))
        self.h = p.value

    def record(self, stream=None):
        _check(_L.lt_hip_event_record(self.h, stream))

    def synchronize(self):
        _check(_L.lt_hip_event_synchronize(self.h))

    def elapsed_ms(self, end):
        ms = C.c_float()
        _check(_L.lt_hip_event_elapsed_ms(self.h, end.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            _L.lt_hip_event_destroy(self.h)
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = _P()
        _check(_L.lt_hip_stream_create(C.byref(p)))
        self.h = p.value

    def synchronize(self):
        _check(_L.lt_hip_stream_synchronize(self.h))

    def __del__(self):
        try:
            _L.lt_hip_stream_destroy(self.h)
        except Exception:
            pass


# ---- hip backend extensions (include/loop_tool_hip.h) ----------------------
_sig("lth_last_error", C.c_char_p)
_sig("lth_is_hip", C.c_int, _P)
_sig("lth_compiled_stream", C.c_int, _P, C.POINTER(_P))
_sig("lth_num_stages", C.c_int, _P)
_sig("lth_stage_info", C.c_int, _P, C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int,
     C.POINTER(C.c_int64), C.POINTER(C.c_int64))
_sig("lth_scratch_sizes", C.c_int, _P, C.POINTER(C.c_int64), C.c_int)
_sig("lth_run_on", C.c_int, _P, C.POINTER(_P), C.c_int, C.POINTER(_P), C.c_int, _P)
_sig("lth_time", C.c_int, _P, C.POINTER(_P), C.c_int, C.c_int, C.c_int,
     C.POINTER(C.c_float))
_sig("lth_time_stage", C.c_int, _P, C.POINTER(_P), C.c_int, C.c_int, C.c_int, C.c_int,
     C.POINTER(C.c_float))


def _lth(rc):
    if rc < 0:
        raise HipError(_L.lth_last_error().decode())
    return rc


def _ptrs(bufs):
    arr = (_P * max(len(bufs), 1))()
    for i, b in enumerate(bufs):
        arr[i] = b.ptr if isinstance(b, DeviceArray) else b
    return arr


def stages(compiled):
    out = []
    for i in range(_lth(_L.lth_num_stages(compiled.h))):
        name, kind = C.create_string_buffer(128), C.create_string_buffer(64)
        by, fl = C.c_int64(), C.c_int64()
        _lth(_L.lth_stage_info(compiled.h, i, name, 128, kind, 64, by, fl))
        out.append(dict(kernel=name.value.decode(), kind=kind.value.decode(),
                        bytes=by.value, flops=fl.value))
    return out


def compiled_stream(compiled):
    """The calling thread's stream of a hip Compiled (its runs launch there)."""
    s = _P()
    _lth(_L.lth_compiled_stream(compiled.h, C.byref(s)))
    return s.value


def scratch_sizes(compiled):
    out = (C.c_int64 * 64)()
    n = _lth(_L.lth_scratch_sizes(compiled.h, out, 64))
    return [out[i] for i in range(n)]


def run_on(compiled, io, scratch=(), stream=None):
    _lth(_L.lth_run_on(compiled.h, _ptrs(io), len(io), _ptrs(scratch), len(scratch), stream))


def time_ms(compiled, io, warmup=3, iters=20):
    ms = C.c_float()
    _lth(_L.lth_time(compiled.h, _ptrs(io), len(io), warmup, iters, C.byref(ms)))
    return ms.value


def time_stage_ms(compiled, io, stage, warmup=3, iters=20):
    """Mean ms of one launch of plan stage `stage` alone (HIP events on the
    calling thread's stream of the Compiled)."""
    ms = C.c_float()
    _lth(_L.lth_time_stage(compiled.h, _ptrs(io), len(io), stage, warmup, iters,
                           C.byref(ms)))
    return ms.value


_sig("lth_pack_create", _P, C.POINTER(_P), C.POINTER(_P), C.POINTER(C.c_int), C.c_int, C.c_int,
     C.c_int, _P)
_sig("lth_pack_replay", C.c_int, _P, _P)
_L.lth_pack_free.restype = None
_L.lth_pack_free.argtypes = [_P]
PACK_GROUP_MATMULS = 1  # include/loop_tool_hip.h LTH_PACK_GROUP_MATMULS


class Pack:
    """Many compiled nests captured into one hipGraph (include/loop_tool_hip.h,
    lth_pack_*): `nests` is a list of (Compiled, [device buffers, inputs then
    outputs]); the nests are spread over `lanes` streams that run
    concurrently; `replay()` runs all of them with one graph launch on
    `stream` (a created Stream's handle, by default the pack's own).  With
    `group_matmuls`, the plain 16-bit matmul nests run as one persistent
    grouped GEMM launch per operand type."""

    def __init__(self, nests, lanes=4, stream=None, group_matmuls=False):
        if stream is None:
            self._own_stream = Stream()
            stream = self._own_stream.h
        self.stream = stream
        self._keep = [c for c, _ in nests]  # Compiled handles outlive the graph
        cc = (_P * len(nests))(*[c.h for c, _ in nests])
        flat = [b for _, io in nests for b in io]
        n_io = (C.c_int * len(nests))(*[len(io) for _, io in nests])
        flags = PACK_GROUP_MATMULS if group_matmuls else 0
        self.h = _L.lth_pack_create(cc, _ptrs(flat), n_io, len(nests), lanes, flags, stream)
        if not self.h:
            raise HipError(_L.lth_last_error().decode())

    def replay(self):
        _lth(_L.lth_pack_replay(self.h, self.stream))

    def synchronize(self):
        _check(_L.lt_hip_stream_synchronize(self.stream))

    def free(self):
        if getattr(self, "h", None):
            _L.lth_pack_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
