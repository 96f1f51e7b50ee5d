"""Build the native libraries of loop_tool_amd in-tree.

    python loop_tool_amd/build.py            # product library only
    python loop_tool_amd/build.py --all      # + the test oracle libraries

Outputs (git-ignored, shipped with the tree):
    loop_tool_amd/liblt_amd.so     C++ loop_tool core, lazy frontend, schedule
                                   transforms, the "hip" backend and the C ABIs
                                   (include/loop_tool_c.h, include/lt_hip.h)
    loop_tool_amd/kernels/*.hsaco  hand-written gfx950 kernels built with hipcc

Host C++ is compiled with g++ against the HIP runtime headers; device code is
generated per LoopTree and compiled at run time by hipRTC for gfx950, except
the hand-written kernel library, which hipcc builds here.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "loop_tool_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LIB = os.path.join(PKG, "liblt_amd.so")

CXXFLAGS = [
    "-std=c++17", "-O2", "-g0", "-fPIC", "-Wall", "-Wno-sign-compare",
    "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
    "-I" + os.path.join(ROCM, "include"), "-I" + os.path.join(BUILD, "gen"),
]
# device headers whose text is embedded in the library for hipRTC modules
EMBEDDED = ["igemm_tile.h"]
# every HIP compile: gfx950 only, packed FP32 instructions disabled
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    """Host C++ (.cpp, built with g++) and HIP (.hip, built with hipcc: the
    static device scratch of the allocation-free launch) of the library."""
    out = []
    for sub in ("core", "frontends", "hip", "serve"):
        d = os.path.join(CSRC, sub)
        out += [os.path.join(d, f) for f in sorted(os.listdir(d))
                if f.endswith(".cpp") or f.endswith(".hip")]
    out.append(os.path.join(CSRC, "capi.cpp"))
    out.append(os.path.join(CSRC, "capi_hip.cpp"))
    return out


def _headers():
    hs = []
    for base in (os.path.join(ROOT, "include"), CSRC):
        for dp, _, fs in os.walk(base):
            hs += [os.path.join(dp, f) for f in fs if f.endswith(".h")]
    return hs


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _embed_sources():
    """csrc/kernels/<name>.h -> _build/gen/<name-without-.h>.inc (a raw string)."""
    gen = os.path.join(BUILD, "gen")
    os.makedirs(gen, exist_ok=True)
    for h in EMBEDDED:
        src = os.path.join(CSRC, "kernels", h)
        text = open(src).read()
        assert ")LTSRC\"" not in text
        body = 'R"LTSRC(' + text + ')LTSRC"\n'
        out = os.path.join(gen, h[:-2] + ".inc")
        if not os.path.exists(out) or open(out).read() != body:
            with open(out, "w") as f:
                f.write(body)


def build_product(force=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    _embed_sources()
    newest_header = max(os.path.getmtime(h) for h in _headers())
    objs, todo = [], []
    for src in sources():
        rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
        obj = os.path.join(BUILD, rel.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        stale = (force or not os.path.exists(obj)
                 or os.path.getmtime(obj) < max(os.path.getmtime(src), newest_header))
        if stale and src.endswith(".hip"):
            flags = [f for f in CXXFLAGS if f != "-D__HIP_PLATFORM_AMD__"]
            todo.append([os.path.join(ROCM, "bin", "hipcc")] + HIPCC_FLAGS + flags +
                        ["-c", src, "-o", obj])
        elif stale:
            todo.append(["g++"] + CXXFLAGS + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    if todo or force or not os.path.exists(LIB):
        _run(["g++", "-shared", "-o", LIB] + objs +
             ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lhiprtc", "-ldl",
              "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return LIB


def build_kernels(force=False):
    """hipcc-compile the hand-written kernel library to gfx950 code objects."""
    kdir = os.path.join(CSRC, "kernels")
    out_dir = os.path.join(PKG, "kernels")
    os.makedirs(out_dir, exist_ok=True)
    built = []
    if not os.path.isdir(kdir):
        return built
    for f in sorted(os.listdir(kdir)):
        if not f.endswith(".hip"):
            continue
        src = os.path.join(kdir, f)
        out = os.path.join(out_dir, f[:-4] + ".hsaco")
        deps = [src] + [os.path.join(kdir, h) for h in os.listdir(kdir) if h.endswith(".h")]
        if force or not os.path.exists(out) or os.path.getmtime(out) < max(
                os.path.getmtime(d) for d in deps):
            bundle = out + ".bundle"
            _run([os.path.join(ROCM, "bin", "hipcc"), "--genco"] + HIPCC_FLAGS +
                 ["-I" + kdir, "-o", bundle, src])
            # keep the bare gfx950 code object (what hipModuleLoadData maps)
            _run([os.path.join(ROCM, "llvm", "bin", "clang-offload-bundler"), "--unbundle",
                  "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                  "--input=" + bundle, "--output=" + out])
            os.remove(bundle)
        built.append(out)
    return built


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true", help="also build the test oracles")
    ap.add_argument("--force", action="store_true")
    args = ap.parse_args(argv)
    print(build_product(force=args.force))
    for k in build_kernels(force=args.force):
        print(k)
    if args.all:
        sys.path.insert(0, ROOT)
        from oracle import build as oracle_build  # test infrastructure
        for p in oracle_build.build_all(force=args.force):
            print(p)


if __name__ == "__main__":
    main()
