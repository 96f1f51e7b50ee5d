"""Lowering decisions and gfx950 code generation of the hip backend, checked
on the CPU: every plan is generated and compiled by hipRTC (which needs no
GPU); launching is covered by the gpu tests."""
import pytest

from loop_tool_amd import programs as P

CASES = [
    ("softmax", dict(R=8192, Cn=8192), "hip", "rowreduce"),
    ("softmax", dict(R=33, Cn=130), "hip", "rowreduce"),
    ("matmul", dict(M=128, N=128, K=128), "hip", "contract_f32"),
    ("matmul", dict(M=37, N=53, K=29), "hip_bf16", "contract_bf16"),
    ("conv2d", dict(N=32, Ci=128, Co=128, H=56, W=56), "hip", "contract_f32"),
    ("conv2d", dict(N=32, Ci=128, Co=128, H=56, W=56), "hip_bf16", "contract_bf16"),
    ("conv2d", dict(N=2, Ci=16, Co=72, H=15, W=15, stride=2), "hip_bf16", "contract_bf16"),
    ("matmul", dict(M=300, N=200, K=704), "hip_bf16", "contract_bf16"),
    ("maxpool2d", dict(N=2, C=3, H=16, W=16), "hip", "pointwise"),
    ("concat2d", dict(B=3, M0=5, M1=4), "hip", "pointwise"),
    ("pad1d", dict(N=11, pre=2, post=3), "hip", "pointwise"),
    ("transpose2d", dict(M=7, N=5), "hip", "pointwise"),
]


@pytest.mark.parametrize("name,kw,backend,kind", CASES,
                         ids=["%s-%s-%d" % (c[0], c[2], i) for i, c in enumerate(CASES)])
def test_plan_kind_and_compile(lt, name, kw, backend, kind):
    out, _ = getattr(P, name)(lt, **kw)
    cc = lt.compile(out.loop_tree, backend)
    assert cc.string_property("kinds").split(",")[-1] == kind
    assert "gfx950" in cc.dump()


def test_plain_annotation_forces_exact_lowering(lt):
    out, _ = P.softmax(lt, 64, 1000)
    t = out.loop_tree
    t = t.annotate(t.roots[0], "hip:plain")
    cc = lt.compile(t, "hip")
    assert cc.string_property("kinds") == "pointwise,pointwise,pointwise"
    # the annotation is ignored by the C oracle, which still compiles it
    lt.compile(t, "cpp")


def test_split_schedule_lowers_with_tails(lt):
    out, _ = P.matmul(lt, 16, 16, 16)
    t = out.loop_tree
    t = t.split(t.children(t.roots[0])[0], 10)  # K loop: 1 x 10 + 6 tail
    t = t.annotate(t.roots[0], "hip:plain")
    src = lt.compile(t, "hip").dump()
    assert "for (int " in src or "for (long " in src


@pytest.mark.parametrize("tile", ["tile=64x64", "tile=128x64", "tile=64x128", "tile=128x128",
                                  "tile=128x256", "tile=256x128", "tile=64x256",
                                  "tile=64x64 bk=64", "tile=128x256 bk=64"])
def test_staged_contraction_plans(lt, tile):
    """16-bit conv: one staging launch (zero-haloed NHWC input, K-innermost
    weights) then the implicit-GEMM tile, at every tile shape."""
    out, _ = P.conv2d(lt, 2, 16, 72, 15, 15, stride=2)
    t = out.loop_tree
    t = t.annotate(t.roots[0], tile)
    cc = lt.compile(t, "hip_bf16")
    assert cc.string_property("kinds") == "stage,contract_bf16"
    bk = "64" if "bk=64" in tile else "32"
    assert "staged P" in cc.dump() and tile.split()[0][5:] + "x" + bk in cc.dump()


def test_staged_layout_of_conv_operands(lt):
    out, _ = P.conv2d(lt, 32, 128, 128, 56, 56)
    t = out.loop_tree
    d = lt.compile(t, "hip_bf16").dump()
    # k-blocked by the K tile of 32: X'[4][32][58][58][32] and W'[4][3][3][128][32],
    # K = (4 blocks x 9 taps) x 32
    assert "K=36x32" in d and "k-blocked staging" in d
    assert "staged P 147456 and Q 13778944 elements" in d
    # flat: X'[32][58][58][128] and W'[128][3][3][128], K = 9 x 128
    d = lt.compile(t.annotate(t.roots[0], "stage=flat"), "hip_bf16").dump()
    assert "K=9x128" in d and "staged P 147456 and Q 13778944 elements" in d
    assert "k-blocked" not in d.splitlines()[0]


def test_streamk_plans(lt):
    """The float conv of the benchmark shape splits 784 tiles over 512
    workgroups and zeroes only the shared tiles' output blocks at staging;
    forced on a shape whose range boundaries all fall between tiles, the
    plan has no zeroing branch (and still compiles)."""
    out, _ = P.conv2d(lt, 32, 128, 128, 56, 56)
    d = lt.compile(out.loop_tree, "hip").dump()
    assert "stream-K over 512 workgroups" in d.splitlines()[0]
    n_split = int(d.split("split_[", 1)[1].split("]", 1)[0])
    assert 0 < n_split < 512
    t = P.matmul(lt, 2048, 2048, 96)[0].loop_tree
    d = lt.compile(t.annotate(t.roots[0], "contract=staged tile=64x64 streamk=1"), "hip").dump()
    assert "stream-K" in d.splitlines()[0] and "split_[" not in d


@pytest.mark.parametrize("knob,msg", [("waves=3", "waves=N must be 2, 4 or 8"),
                                      ("tile=100x64", "multiples of 32"),
                                      ("tile=abc", "expected tile=MxN"),
                                      ("bk=7", "bk=N must be 32 or 64"),
                                      ("block=3", "block=N must be a multiple of 64"),
                                      ("vec=3", "vec=N must be 1, 2 or 4"),
                                      ("igemm=foo", "unknown value"),
                                      ("streamk=2", "unknown value"),
                                      ("srows=0", "srows=N must be in \\[1, 64\\]"),
                                      ("waves=-1", "expected a number")])
def test_bad_knob_values_are_errors(lt, knob, msg):
    """A token whose key is a knob of the hip backend must carry a valid
    value; annotations meant for other consumers pass through."""
    import loop_tool_amd as lt_mod
    out, _ = P.matmul(lt, 64, 64, 64)
    t = out.loop_tree
    with pytest.raises(lt_mod.LoopToolError, match=msg):
        lt.compile(t.annotate(t.roots[0], knob), "hip")
    lt.compile(t.annotate(t.roots[0], "parallel unroll=4 tile=64x64"), "hip")


@pytest.mark.parametrize("name,kw", [("matmul", dict(M=0, N=4, K=4)),
                                     ("softmax", dict(R=0, Cn=8))])
def test_empty_extent_is_an_error(lt, name, kw):
    """A zero-sized loop var is refused when the tree is built for a
    backend, with the variable named (no kernel with an empty grid)."""
    import loop_tool_amd as lt_mod
    out, _ = getattr(P, name)(lt, **kw)
    with pytest.raises(lt_mod.LoopToolError, match="empty extent"):
        lt.compile(out.loop_tree, "hip")


def test_code_object_disk_cache(tmp_path):
    """A second process loads the code object that the first compiled; an
    entry whose stored source differs from the module's is not used."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import loop_tool_amd as lt\n"
            "from loop_tool_amd import programs as P\n"
            "out, _ = P.softmax(lt.lib, 64, 1000)\n"
            "print(lt.lib.compile(out.loop_tree, 'hip').int_property('code_cached'))\n" % root)
    env = dict(os.environ, LT_HIP_CACHE_DIR=str(tmp_path))

    def cached():
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        return int(r.stdout.strip().splitlines()[-1])
    assert cached() == 0 and cached() == 1
    (entry,) = list(tmp_path.iterdir())
    b = entry.read_bytes()
    entry.write_bytes(b[:8] + bytes([b[8] ^ 1]) + b[9:])  # one source byte changed
    assert cached() == 0
    # the write is published by rename: no temporary file stays behind
    assert [p.name for p in tmp_path.iterdir()] == [entry.name]


@pytest.mark.gpu
def test_damaged_disk_cache_entry_is_recompiled(tmp_path, hip_dev):
    """A disk entry with the right source but a code object the runtime
    rejects (not an ELF image) is dropped and the module compiled again:
    the run succeeds and matches a run with a clean cache."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import numpy as np\n"
            "import loop_tool_amd as lt\n"
            "from loop_tool_amd import hip, programs as P\n"
            "hip.init()\n"
            "out, sh = P.softmax(lt.lib, 64, 1000)\n"
            "cc = lt.lib.compile(out.loop_tree, 'hip')\n"
            "y = np.zeros(64000, np.float32)\n"
            "cc.run([P.normal(3, sh[0], 3.0), y])\n"
            "print(cc.int_property('code_cached'), float(y.sum()))\n" % root)
    env = dict(os.environ, LT_HIP_CACHE_DIR=str(tmp_path))

    def run():
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        c, total = r.stdout.strip().splitlines()[-1].split()
        return int(c), float(total)
    c0, t0 = run()
    assert c0 == 0
    (entry,) = list(tmp_path.iterdir())
    b = entry.read_bytes()
    n = int.from_bytes(b[:8], "little")
    entry.write_bytes(b[:8 + n] + b"not a code object")
    c1, t1 = run()
    assert c1 == 1 and t1 == t0
    assert entry.read_bytes()[8 + n:8 + n + 4] == b"\x7fELF"  # rewritten


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N above 1 exits with a message."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)


def test_build_script_runs_without_the_package():
    """The build script named in the package's error messages runs as a
    plain script, so it works in a tree whose native library is not built
    yet (importing the package needs that library)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "loop_tool_amd", "build.py"),
                        "--help"], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, PYTHONPATH=""))
    assert r.returncode == 0 and "--all" in r.stdout
    src = open(os.path.join(root, "loop_tool_amd", "build.py")).read()
    assert "import loop_tool_amd" not in src and "from loop_tool_amd" not in src


def test_plain_matmul_descriptor(lt):
    """Plans record when their nest is a plain row-major matmul (what a pack
    may group into one batched GEMM launch); convolutions and batched
    products are not."""
    cases = [("matmul", dict(M=300, N=200, K=704), "hip_bf16", 1),
             ("matmul", dict(M=128, N=128, K=128), "hip", 1),
             ("conv2d", dict(N=2, Ci=32, Co=40, H=14, W=14), "hip_bf16", None),
             ("bmm", dict(Bn=2, M=64, N=64, K=64), "hip_bf16", None)]
    for name, kw, backend, want in cases:
        out, _ = getattr(P, name)(lt, **kw)
        cc = lt.compile(out.loop_tree, backend)
        assert cc.int_property("plain_matmul") == want, (name, kw)
