"""TEST INFRASTRUCTURE ONLY -- builds the CPU oracles used by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.

  oracle/liblt_oracle.so
      The C-backend restatement (oracle/cpp_backend.cpp) registered as "cpp",
      loaded into a process with loop_tool::loadLibrary next to the product.

  oracle/_ref/libloop_tool_ref.so   (only where /root/reference exists)
      The reference loop_tool compiled from its own sources under
      /root/reference (core, cpp backend, lazy/mutate/nn frontends) plus this
      repository's C ABI driver (loop_tool_amd/csrc/capi.cpp) compiled against
      the reference headers.  Everything except the ltc_* entry points is
      hidden and the library is linked -Bsymbolic, so it can be loaded in the
      same process as the product without symbol interposition.  No reference
      source is copied: the files are compiled where they lie.
"""
import concurrent.futures as cf
import os
import sys
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
REF = os.environ.get("LT_REFERENCE_DIR", "/root/reference")
REF_OUT = os.path.join(ORACLE, "_ref")
REF_LIB = os.path.join(REF_OUT, "libloop_tool_ref.so")
ORACLE_LIB = os.path.join(ORACLE, "liblt_oracle.so")

REF_SOURCES = [
    "src/core/backend.cpp", "src/core/compile.cpp", "src/core/hardware.cpp",
    "src/core/ir.cpp", "src/core/serialization.cpp", "src/core/symbolic.cpp",
    "src/core/tensor.cpp", "src/backends/cpu/cpp.cpp", "src/frontends/lazy.cpp",
    "src/frontends/mutate.cpp", "src/frontends/nn.cpp",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)


def _stale(out, deps):
    return not os.path.exists(out) or os.path.getmtime(out) < max(
        os.path.getmtime(d) for d in deps)


def build_oracle(force=False):
    srcs = [os.path.join(ORACLE, f) for f in sorted(os.listdir(ORACLE)) if f.endswith(".cpp")]
    lib = os.path.join(ROOT, "loop_tool_amd", "liblt_amd.so")
    if force or _stale(ORACLE_LIB, srcs + [lib]):
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared",
              "-I" + os.path.join(ROOT, "include"), "-o", ORACLE_LIB] + srcs +
             ["-L" + os.path.dirname(lib), "-llt_amd",
              "-Wl,-rpath,$ORIGIN/../loop_tool_amd"])
    return ORACLE_LIB


def build_reference(force=False, jobs=8):
    if not os.path.isdir(os.path.join(REF, "src")):
        return None  # no reference sources here: a prebuilt copy is used if present
    os.makedirs(os.path.join(REF_OUT, "obj"), exist_ok=True)
    flags = ["-std=c++17", "-O2", "-fPIC", "-fvisibility=hidden",
             "-fvisibility-inlines-hidden", "-I" + os.path.join(REF, "include"),
             "-include", "chrono", "-w"]
    srcs = [os.path.join(REF, s) for s in REF_SOURCES]
    srcs.append(os.path.join(ROOT, "loop_tool_amd", "csrc", "capi.cpp"))
    objs, todo = [], []
    for s in srcs:
        obj = os.path.join(REF_OUT, "obj", os.path.basename(s)[:-4] + ".o")
        objs.append(obj)
        extra = [os.path.join(ROOT, "include", "loop_tool_c.h")]
        if force or _stale(obj, [s] + extra):
            inc = ["-I" + os.path.join(ROOT, "include")] if s.endswith("capi.cpp") else []
            # the driver must see the reference headers first
            todo.append(["g++"] + flags[:6] + ["-I" + os.path.join(REF, "include")] +
                        inc + flags[6:] + ["-c", s, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    if todo or force or not os.path.exists(REF_LIB):
        vs = os.path.join(REF_OUT, "exports.map")
        with open(vs, "w") as f:
            f.write("{ global: ltc_*; local: *; };\n")
        _run(["g++", "-shared", "-Wl,-Bsymbolic", "-Wl,--version-script=" + vs,
              "-o", REF_LIB] + objs + ["-ldl"])
    return REF_LIB


def build_all(force=False):
    out = [build_oracle(force)]
    for fn in (build_reference, build_reference_plugin):
        r = fn(force)
        if r:
            out.append(r)
    return out



PUB_LIB = os.path.join(REF_OUT, "libloop_tool_pub.so")
PLUGIN_LIB = os.path.join(REF_OUT, "libloop_tool_hip.so")


def build_reference_plugin(force=False, jobs=8):
    """Drop-in check: the reference library with default symbol visibility
    (as its own CMake build produces it) plus this repository's hip backend
    sources compiled against the *reference* headers into an out-of-tree
    plugin, loaded by the reference's loop_tool::loadLibrary."""
    if not os.path.isdir(os.path.join(REF, "src")):
        return None
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    objdir = os.path.join(REF_OUT, "obj_pub")
    os.makedirs(objdir, exist_ok=True)
    sys.path.insert(0, ROOT)
    from loop_tool_amd import build as product_build
    product_build._embed_sources()  # device headers the hip sources embed
    base = ["-std=c++17", "-O2", "-fPIC", "-I" + os.path.join(REF, "include"),
            "-I" + os.path.join(ROOT, "include"), "-include", "chrono", "-w",
            "-I" + os.path.join(product_build.BUILD, "gen")]
    core = [os.path.join(REF, s) for s in REF_SOURCES] + [
        os.path.join(ROOT, "loop_tool_amd", "csrc", "capi.cpp")]
    hipdir = os.path.join(ROOT, "loop_tool_amd", "csrc", "hip")
    plugin = [os.path.join(hipdir, f) for f in sorted(os.listdir(hipdir)) if f.endswith(".cpp")]
    hdrs = [os.path.join(hipdir, f) for f in os.listdir(hipdir) if f.endswith(".h")]
    todo, core_objs, plugin_objs = [], [], []
    for s in core + plugin:
        obj = os.path.join(objdir, ("hip_" if s in plugin else "") + os.path.basename(s)[:-4] + ".o")
        (plugin_objs if s in plugin else core_objs).append(obj)
        extra = ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(rocm, "include")] if s in plugin else []
        if force or _stale(obj, [s] + hdrs):
            todo.append(["g++"] + base + extra + ["-c", s, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    if todo or force or not os.path.exists(PUB_LIB):
        _run(["g++", "-shared", "-o", PUB_LIB] + core_objs + ["-ldl"])
    if todo or force or not os.path.exists(PLUGIN_LIB):
        _run(["g++", "-shared", "-o", PLUGIN_LIB] + plugin_objs +
             ["-L" + REF_OUT, "-lloop_tool_pub", "-Wl,-rpath,$ORIGIN",
              "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-lhiprtc",
              "-Wl,-rpath," + os.path.join(rocm, "lib")])
    return PLUGIN_LIB


if __name__ == "__main__":
    for p in build_all():
        print(p)
