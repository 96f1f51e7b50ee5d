"""loop_tool_amd: loop_tool's frontend and an MI355X (gfx950) "hip" backend.

The native library (liblt_amd.so, built by `python loop_tool_amd/build.py`)
holds the C++ loop_tool core, lazy frontend, schedule transforms and the hip
backend; this package binds its C ABI (include/loop_tool_c.h) with ctypes and
exposes the API of the reference's `loop_tool_py` module
(src/frontends/python.cpp:86-698): Symbol, Expr, Size, Tensor, LoopTree,
backends(), set_default_backend(), ...
"""
import os

from . import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblt_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "loop_tool_amd native library missing at %s; build it with "
        "`python loop_tool_amd/build.py`" % LIB_PATH)

lib = capi.Library(LIB_PATH)

Symbol = lib.Symbol
Size = lib.Size
Expr = lib.Expr
Tensor = lib.Tensor
backends = lib.backends
set_default_backend = lib.set_default_backend
get_default_backend = lib.get_default_backend
set_default_hardware = lib.set_default_hardware
load_lib = lib.load_lib
deserialize = lib.deserialize
compile = lib.compile
convolve = lib.convolve
maxpool = lib.maxpool
LoopToolError = capi.LoopToolError


class Backend:
    """`with lt.Backend("hip"): ...` (python/__init__.py of the reference)."""

    def __init__(self, name):
        self.name = name

    def __enter__(self):
        self.old = get_default_backend()
        set_default_backend(self.name)

    def __exit__(self, *args):
        set_default_backend(self.old)
