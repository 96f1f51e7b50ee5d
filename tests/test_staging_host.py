"""Staging passes of the 16-bit contraction lowering, executed on the CPU.

The hip backend emits the staging of each contraction operand as a device
function `lt_stage_op<o>(in, S, workgroup)`, called by one fused staging
kernel.  Here those functions are compiled for the host with ROCm's clang and
run workgroup by workgroup: all 256 threads run up to the barrier, then all
run again through the write phase (the LDS tile is a static array, and the
pre-barrier phase is idempotent).  The staged buffers must hold exactly the
layout the MFMA tile reads: the source permuted so the shared reduced dim is
innermost, a zero halo wherever a view indexes out of range, zeros up to the
padded K extent, values rounded to bf16 (RNE).  The host build runs under
AddressSanitizer, so a read outside an input fails; guard words after each
staged buffer catch writes past its end, and a nonzero fill catches elements
never written.
Expected layouts come from numpy relayouts of the inputs, not from the
emitter.
"""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from loop_tool_amd import programs as P

CLANG = "/opt/rocm/llvm/bin/clang++"

PRELUDE = r"""
#include <cstdio>
#include <cstdlib>
#include <cstring>
struct { unsigned x; } threadIdx;
static int g_phase;
#define __device__
#define __forceinline__ inline
#define __shared__ static
#define __syncthreads() do { if (g_phase == 0) return; } while (0)
"""

DRIVER = r"""
typedef void (*stage_fn)(const float*, LT_ELT*, unsigned);
static void* slurp(const char* path, size_t n) {
  void* p = malloc(n);
  FILE* fp = fopen(path, "rb");
  if (!fp || fread(p, 1, n, fp) != n) { fprintf(stderr, "cannot read %s\n", path); exit(2); }
  fclose(fp);
  return p;
}
int main(int argc, char** argv) {
  // argv: dir numel0 staged0 numel1 staged1
  stage_fn fn[2] = {lt_stage_op0, lt_stage_op1};
  for (int o = 0; o < 2; ++o) {
    size_t n = strtoull(argv[2 + 2 * o], 0, 10), s = strtoull(argv[3 + 2 * o], 0, 10);
    char path[4096];
    snprintf(path, sizeof path, "%s/in%d.bin", argv[1], o);
    float* in = (float*)slurp(path, n * 4);
    unsigned char* S = (unsigned char*)malloc((s + 64) * sizeof(LT_ELT));
    memset(S, 0xA5, (s + 64) * sizeof(LT_ELT));
    for (unsigned b = 0; b < lt_stage_blocks[o]; ++b)
      for (g_phase = 0; g_phase < 2; ++g_phase)
        for (threadIdx.x = 0; threadIdx.x < 256; ++threadIdx.x) fn[o](in, (LT_ELT*)S, b);
    snprintf(path, sizeof path, "%s/S%d.bin", argv[1], o);
    FILE* fp = fopen(path, "wb");
    fwrite(S, sizeof(LT_ELT), s + 64, fp);
    fclose(fp);
    free(in);
    free(S);
  }
  return 0;
}
"""


def bf16_bits(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def extract(src, name):
    """The text of one emitted device function, by brace matching."""
    i = src.index("__device__ __forceinline__ void " + name + "(")
    depth, j = 0, src.index("{", i)
    while True:
        depth += {"{": 1, "}": -1}.get(src[j], 0)
        j += 1
        if depth == 0:
            return src[i:j]


def f32_bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def staged_conv_input(X, pad, k, stride, kip, bits=None):
    """X[N][C][H][W] -> X'[N][H+lo+hi][W+lo+hi][kip] (bf16 bits, or `bits`)."""
    N, C, H, W = X.shape
    Ho, Wo = (H + 2 * pad - k) // stride + 1, (W + 2 * pad - k) // stride + 1
    hi_h = max(0, (Ho - 1) * stride + k - 1 - pad - (H - 1))
    hi_w = max(0, (Wo - 1) * stride + k - 1 - pad - (W - 1))
    bits = bits or bf16_bits
    v = bits(X.transpose(0, 2, 3, 1))
    out = np.zeros((N, H + pad + hi_h, W + pad + hi_w, kip), v.dtype)
    out[:, pad:pad + H, pad:pad + W, :C] = v
    return out


def staged_weights(Wt, kip, bits=None):
    """W[Co][C][k][k] -> W'[Co][k][k][kip] (bf16 bits, or `bits`)."""
    Co, C, k, _ = Wt.shape
    bits = bits or bf16_bits
    v = bits(Wt.transpose(0, 2, 3, 1))
    out = np.zeros((Co, k, k, kip), v.dtype)
    out[..., :C] = v
    return out


def run_staging(lt, out, inputs, tmp, annotate="", backend="hip_bf16"):
    t = out.loop_tree
    if annotate:
        t = t.annotate(t.roots[0], annotate)
    cc = lt.compile(t, backend)
    kind = "contract_bf16" if backend == "hip_bf16" else "contract_f32"
    assert cc.string_property("kinds") == "stage," + kind
    src = cc.string_property("code")
    blocks = re.search(r"static constexpr unsigned lt_stage_blocks\[2\] = \{[^}]*\};", src).group(0)
    elt = "__bf16" if backend == "hip_bf16" else "float"
    body = (PRELUDE + "#define LT_ELT " + elt + "\n" + blocks + "\n" +
            extract(src, "lt_stage_op0") + "\n" + extract(src, "lt_stage_op1") + "\n" + DRIVER)
    with open(os.path.join(tmp, "s.cpp"), "w") as f:
        f.write(body)
    # AddressSanitizer (host build): a source read outside the input is an error
    subprocess.run([CLANG, "-std=c++17", "-O1", "-w", "-fsanitize=address",
                    "-o", os.path.join(tmp, "s"), os.path.join(tmp, "s.cpp")], check=True)
    return cc


def check_staged(tmp, exp, numels):
    """exp: expected staged operands as uint16 (16-bit staging) or uint32
    (float staging) bit patterns."""
    args = [os.path.join(tmp, "s"), tmp]
    for o in range(2):
        args += [str(numels[o]), str(exp[o].size)]
    subprocess.run(args, check=True, timeout=300)
    for o in range(2):
        dt = exp[o].dtype
        S = np.fromfile(os.path.join(tmp, "S%d.bin" % o), dt)
        guard = dt.type(0xA5A5A5A5 if dt == np.uint32 else 0xA5A5)
        assert np.all(S[exp[o].size:] == guard), "operand %d: write past the staged buffer" % o
        bad = np.flatnonzero(S[:exp[o].size] != exp[o].ravel())
        assert bad.size == 0, "operand %d: %d wrong elements, first at %s of %s" % (
            o, bad.size, np.unravel_index(bad[0], exp[o].shape), exp[o].shape)


CONV = [
    dict(N=2, Ci=128, Co=16, H=56, W=56),            # 16-byte loads, 64 channels per workgroup
    dict(N=1, Ci=40, Co=8, H=15, W=15),              # scalar loads (odd rows), K 40 padded to 64
    dict(N=2, Ci=16, Co=72, H=15, W=15, stride=2),   # stride-2 halo, 32 channels per workgroup
    dict(N=1, Ci=32, Co=8, H=9, W=12, pad=2, k=5),   # 5x5 window, two-wide halo
    dict(N=1, Ci=64, Co=8, H=5, W=300),              # rows longer than 256: one row per workgroup
]


@pytest.fixture
def tmpdir_():
    d = tempfile.mkdtemp(prefix="lt_stage_host_")
    yield d
    shutil.rmtree(d)


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm clang for the host build")
@pytest.mark.parametrize("kw", CONV, ids=[str(i) for i in range(len(CONV))])
def test_conv_staging_on_host(lt, kw, tmpdir_):
    pad, k, stride = kw.get("pad", 1), kw.get("k", 3), kw.get("stride", 1)
    out, shapes = P.conv2d(lt, **kw)
    X, Wt = P.uniform(7, shapes[0]), P.uniform(8, shapes[1])
    cc = run_staging(lt, out, [X, Wt], tmpdir_, "stage=flat")
    d = cc.dump()
    kin = kw["Ci"]
    m = re.search(r"K=\d+x%d(?: \(padded to (\d+)\))?" % kin, d)
    kip = int(m.group(1)) if m.group(1) else kin
    exp = [staged_conv_input(X, pad, k, stride, kip), staged_weights(Wt, kip)]
    X.tofile(os.path.join(tmpdir_, "in0.bin"))
    Wt.tofile(os.path.join(tmpdir_, "in1.bin"))
    check_staged(tmpdir_, exp, [X.size, Wt.size])


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm clang for the host build")
@pytest.mark.parametrize("M,N,K", [(64, 96, 128), (37, 53, 29), (300, 200, 704)])
def test_matmul_staging_on_host(lt, M, N, K, tmpdir_):
    """A[M][K] rounded in place (K padded), B[K][N] transposed to Bt[N][K]."""
    out, shapes = P.matmul(lt, M, N, K)
    A, B = P.uniform(9, shapes[0]), P.uniform(10, shapes[1])
    cc = run_staging(lt, out, [A, B], tmpdir_, "stage=flat")
    m = re.search(r"K=1x%d(?: \(padded to (\d+)\))?" % K, cc.dump())
    kip = int(m.group(1)) if m.group(1) else K
    ea = np.zeros((M, kip), np.uint16)
    ea[:, :K] = bf16_bits(A)
    eb = np.zeros((N, kip), np.uint16)
    eb[:, :K] = bf16_bits(B.T)
    A.tofile(os.path.join(tmpdir_, "in0.bin"))
    B.tofile(os.path.join(tmpdir_, "in1.bin"))
    check_staged(tmpdir_, [ea, eb], [A.size, B.size])


def blocked(flat, bk, row_axes):
    """flat staged layout [...][kip] -> k-blocked [kip/bk][row_axes order][bk]."""
    kb = flat.shape[-1] // bk
    b = flat.reshape(flat.shape[:-1] + (kb, bk))
    nd = flat.ndim - 1
    return b.transpose((nd,) + tuple(row_axes) + (nd + 1,))


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm clang for the host build")
@pytest.mark.parametrize("kw", [CONV[0], CONV[1], CONV[2], CONV[3]], ids=["0", "1", "2", "3"])
@pytest.mark.parametrize("bk", [32, 64])
def test_conv_blocked_staging_on_host(lt, kw, bk, tmpdir_):
    """stage=blocked (the default where k_in spans several K tiles):
    X'[C/bk][N][H+2][W+2][bk] and W'[C/bk][k][k][Co][bk] (the window dims,
    indexed by reduced vars only, before the output channels)."""
    pad, k, stride = kw.get("pad", 1), kw.get("k", 3), kw.get("stride", 1)
    out, shapes = P.conv2d(lt, **kw)
    X, Wt = P.uniform(7, shapes[0]), P.uniform(8, shapes[1])
    cc = run_staging(lt, out, [X, Wt], tmpdir_, "stage=blocked bk=%d" % bk)
    kin = kw["Ci"]
    kip = -(-kin // bk) * bk
    d = cc.dump().splitlines()[0]  # the plan description
    if kip == bk:  # a single K block: the flat layout
        assert "k-blocked" not in d
        exp = [staged_conv_input(X, pad, k, stride, kip), staged_weights(Wt, kip)]
    else:
        assert "k-blocked" in d
        exp = [blocked(staged_conv_input(X, pad, k, stride, kip), bk, (0, 1, 2)),
               blocked(staged_weights(Wt, kip), bk, (1, 2, 0))]
    X.tofile(os.path.join(tmpdir_, "in0.bin"))
    Wt.tofile(os.path.join(tmpdir_, "in1.bin"))
    check_staged(tmpdir_, exp, [X.size, Wt.size])


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm clang for the host build")
@pytest.mark.parametrize("M,N,K", [(64, 96, 128), (300, 200, 704)])
def test_matmul_blocked_staging_on_host(lt, M, N, K, tmpdir_):
    """stage=blocked: A'[K/32][M][32] and Bt'[K/32][N][32]."""
    out, shapes = P.matmul(lt, M, N, K)
    A, B = P.uniform(9, shapes[0]), P.uniform(10, shapes[1])
    cc = run_staging(lt, out, [A, B], tmpdir_, "stage=blocked bk=32")
    kip = -(-K // 32) * 32
    ea = np.zeros((M, kip), np.uint16)
    ea[:, :K] = bf16_bits(A)
    eb = np.zeros((N, kip), np.uint16)
    eb[:, :K] = bf16_bits(B.T)
    A.tofile(os.path.join(tmpdir_, "in0.bin"))
    B.tofile(os.path.join(tmpdir_, "in1.bin"))
    check_staged(tmpdir_, [blocked(ea, 32, (0,)), blocked(eb, 32, (0,))], [A.size, B.size])


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm clang for the host build")
@pytest.mark.parametrize("kw", CONV, ids=[str(i) for i in range(len(CONV))])
def test_conv_float_staging_on_host(lt, kw, tmpdir_):
    """Float operands ("hip" backend, contract=staged): the same k-blocked
    layouts as 16-bit staging, with the values copied exactly."""
    pad, k, stride = kw.get("pad", 1), kw.get("k", 3), kw.get("stride", 1)
    out, shapes = P.conv2d(lt, **kw)
    X, Wt = P.uniform(7, shapes[0]), P.uniform(8, shapes[1])
    cc = run_staging(lt, out, [X, Wt], tmpdir_, "contract=staged", backend="hip")
    kip = -(-kw["Ci"] // 32) * 32
    ex = staged_conv_input(X, pad, k, stride, kip, f32_bits)
    ew = staged_weights(Wt, kip, f32_bits)
    if kip > 32:
        ex, ew = blocked(ex, 32, (0, 1, 2)), blocked(ew, 32, (1, 2, 0))
    X.tofile(os.path.join(tmpdir_, "in0.bin"))
    Wt.tofile(os.path.join(tmpdir_, "in1.bin"))
    check_staged(tmpdir_, [ex, ew], [X.size, Wt.size])


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm clang for the host build")
@pytest.mark.parametrize("M,N,K", [(64, 96, 128), (37, 53, 29), (300, 200, 704)])
def test_matmul_float_staging_on_host(lt, M, N, K, tmpdir_):
    out, shapes = P.matmul(lt, M, N, K)
    A, B = P.uniform(9, shapes[0]), P.uniform(10, shapes[1])
    run_staging(lt, out, [A, B], tmpdir_, "contract=staged", backend="hip")
    kip = -(-K // 32) * 32
    ea = np.zeros((M, kip), np.uint32)
    ea[:, :K] = f32_bits(A)
    eb = np.zeros((N, kip), np.uint32)
    eb[:, :K] = f32_bits(B.T)
    if kip > 32:
        ea, eb = blocked(ea, 32, (0,)), blocked(eb, 32, (0,))
    A.tofile(os.path.join(tmpdir_, "in0.bin"))
    B.tofile(os.path.join(tmpdir_, "in1.bin"))
    check_staged(tmpdir_, [ea, eb], [A.size, B.size])
