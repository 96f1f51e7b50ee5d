"""The C-ABI libraries load and export every entry point their headers
declare (no compute call: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "loop_tool_amd", "liblt_amd.so")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:ltc|lth|lts|lt_hip)_[a-z0-9_]+)\s*\(", text)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header", ["loop_tool_c.h", "lt_hip.h", "loop_tool_hip.h",
                                    "loop_tool_serve.h"])
def test_product_exports_every_declared_symbol(header):
    names = declared(header)
    assert len(names) > 5
    missing = [n for n in names if n not in exported(LIB)]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for n in names:
        getattr(lib, n)


def test_reference_driver_exports_c_abi(ref):
    names = declared("loop_tool_c.h")
    missing = [n for n in names if n not in exported(ref.path)]
    assert not missing, missing
    # nothing but the C ABI is visible from the reference build
    assert all(n.startswith("ltc_") for n in exported(ref.path))


def test_errors_cross_the_boundary_as_codes(lt):
    import loop_tool_amd as lt_mod
    with pytest.raises(lt_mod.LoopToolError, match="couldn't find backend"):
        lt.set_default_backend("nonexistent")
    s = lt.Symbol("X")
    with pytest.raises(lt_mod.LoopToolError):
        lt.Tensor(s).numpy()  # unbound size


def test_batch_refuses_operands_of_2_gib():
    """lth_batch_create refuses a GEMM whose 16-bit operand reaches 2^31
    bytes (the tile's buffer ranges and row offsets are 32-bit)."""
    lib = ctypes.CDLL(LIB)
    lib.lth_batch_create.restype = ctypes.c_void_p
    lib.lth_batch_last_error.restype = ctypes.c_char_p
    arr = lambda v: (ctypes.c_int * 1)(v)
    h = lib.lth_batch_create(1, arr(65536), arr(64), arr(16384), 0)
    assert not h and b"2^31 bytes" in lib.lth_batch_last_error()
    h = lib.lth_batch_create(1, arr(65535), arr(64), arr(16384), 0)
    assert h
    lib.lth_batch_free.argtypes = [ctypes.c_void_p]
    lib.lth_batch_free(h)
