"""Benchmark of the loop_tool hip backend on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--nest softmax|matmul|matmul_f32_<n>|matmul_bf16_<n>|conv2d|conv2d_bf16|
                            maxpool|copy|batched|serving]

Default workload: FP32 row softmax over an 8192 x 8192 tensor (max-reduce,
subtract, exp, sum-reduce, divide over the columns), written with the lazy
frontend and compiled by the "hip" backend.  A step is one run of the
compiled nest over the whole tensor with the input already resident in HBM.
`value` is algorithmic HBM bytes (one read + one write of the tensor) per
second over the timed steps.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E datasheet
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "f16": 2500.0, "f32": 157.3}
# HBM bytes per launch of the dominant kernel, from FETCH_SIZE / WRITE_SIZE
# counter runs scaled by a copy kernel of known size
# (profiles/lt_row0_softmax_8192x8192_hbm_traffic.md)
MEASURED_TRAFFIC = {"softmax": 537084416}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nest", default="softmax")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--duration", type=float, default=60.0,
                    help="serving: seconds of Poisson arrivals")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: arrivals per second")
    ap.add_argument("--slots", type=int, default=8, help="serving: requests in flight")
    ap.add_argument("--annotate", default="",
                    help="hip backend knobs for the nest, e.g. 'tile=128x256'")
    ap.add_argument("--quick", action="store_true",
                    help="batched: no plain forms and no tuning (every GEMM on the "
                         "128x128x32 tile), for profiling the grouped kernel")
    ap.add_argument("--no-conv", action="store_true",
                    help="softmax run: skip the conv2d bf16 kernel and plan figures")
    ap.add_argument("--no-batched", action="store_true",
                    help="softmax run: skip the batched-GEMM graph replay figure")
    return ap.parse_args()


SETTLE_S = 0.2  # untimed GPU work before the warmup steps


def settle(step, sync, seconds=SETTLE_S, max_steps=5000):
    """Run `step` untimed for about `seconds` (at most `max_steps` times):
    the GPU's clocks ramp up over the first milliseconds of work after idle,
    and a short timed region measured before that reads 4-20 % slow.
    Returns the number of steps run."""
    n, t0 = 0, time.perf_counter()
    while n < max_steps and time.perf_counter() - t0 < seconds:
        for _ in range(10):
            step()
        n += 10
        sync()
    return n


def workload(lt, name):
    from loop_tool_amd import programs as P
    if name == "softmax":
        R, C = 8192, 8192
        out, shapes = P.softmax(lt, R, C)
        gen = lambda: [P.normal(3, shapes[0], 3.0)]
        cfg = dict(workload="row softmax fp32 (max, subtract, exp, sum, divide over C)",
                   rows=R, cols=C)
        return out, shapes, gen, cfg, "hbm", "f32", "hip"
    if name == "matmul":
        out, shapes = P.matmul(lt, 128, 128, 128)
        gen = lambda: [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
        return out, shapes, gen, dict(workload="fp32 matmul M=N=K=128"), "mfma", "f32", "hip"
    if name.startswith("matmul_f32"):
        n = int(name.split("_")[-1]) if name.count("_") == 2 else 4096
        out, shapes = P.matmul(lt, n, n, n)
        gen = lambda: [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
        return out, shapes, gen, dict(workload="matmul fp32, M=N=K=%d" % n), "mfma", "f32", "hip"
    if name.startswith("matmul_bf16"):
        n = int(name.split("_")[-1]) if name.count("_") == 2 else 8192
        out, shapes = P.matmul(lt, n, n, n)
        gen = lambda: [P.uniform(1, shapes[0]), P.uniform(2, shapes[1])]
        return out, shapes, gen, dict(workload="matmul bf16 operands, fp32 accumulate, "
                                               "M=N=K=%d" % n), "mfma", "bf16", "hip_bf16"
    if name in ("conv2d", "conv2d_bf16"):
        N, Ci, Co, H = 32, 128, 128, 56
        out, shapes = P.conv2d(lt, N, Ci, Co, H, H)
        std = 1.0 / np.sqrt(Ci * 9)
        gen = lambda: [P.normal(4, shapes[0], std), P.normal(5, shapes[1], std)]
        bf = name == "conv2d_bf16"
        cfg = dict(workload="conv2d 3x3 pad 1 stride 1 (pad nest + window sum)",
                   N=N, C_in=Ci, C_out=Co, H=H, W=H, operands="bf16" if bf else "fp32")
        return out, shapes, gen, cfg, "mfma", "bf16" if bf else "f32", (
            "hip_bf16" if bf else "hip")
    if name == "copy":
        R, C = 8192, 8192
        r, c = lt.Symbol("R"), lt.Symbol("C")
        out = lt.Tensor(R, C).as_(r, c)
        gen = lambda: [P.normal(3, (R, C))]
        return out, [(R, C)], gen, dict(workload="fp32 copy (calibration of HBM counters)",
                                         rows=R, cols=C), "hbm", "f32", "hip"
    if name == "maxpool":
        out, shapes = P.maxpool2d(lt, 32, 128, 112, 112)
        gen = lambda: [P.normal(6, shapes[0])]
        return out, shapes, gen, dict(workload="2x2 max-pool stride 2", N=32, C=128,
                                      H=112, W=112), "hbm", "f32", "hip"
    raise SystemExit("unknown nest " + name)


# error bars per nest: error vs float64 of the float operands ("f64"), and of
# the 16-bit-rounded operands for 16-bit contractions ("f64_rounded")
TOLERANCE = {"softmax": {"f64": 1e-5}, "matmul": {"f64": 1e-5}, "matmul_f32": {"f64": 1e-5},
             "matmul_bf16": {"f64": 1e-2, "f64_rounded": 1e-5},
             "conv2d": {"f64": 1e-5}, "conv2d_bf16": {"f64": 1e-2, "f64_rounded": 1e-5},
             "maxpool": {"f64": 0.0}, "copy": {"f64": 0.0}}


def nest_errors(nest, inputs_np, y):
    """Normwise error of the device result y against float64 closed forms:
    every row of the softmax; every element of matmul 128, max-pool and copy;
    128 rows of large matmuls; for convolutions, every border pixel of two
    batch planes of every output channel plus 65,536 random outputs."""
    from loop_tool_amd import programs as P

    def nw(a, b):
        return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))
    kind = "matmul_bf16" if nest.startswith("matmul_bf16") else (
        "matmul_f32" if nest.startswith("matmul_f32") else nest)
    out = {}
    if nest == "softmax":
        x = inputs_np[0]
        R, C = x.shape
        yy = y.reshape(R, C)
        worst, scale = 0.0, 0.0
        for r0 in range(0, R, 1024):
            ref = P.reference_output("softmax", [x[r0:r0 + 1024]])
            worst = max(worst, float(np.abs(yy[r0:r0 + 1024] - ref).max()))
            scale = max(scale, float(np.abs(ref).max()))
        out["f64"] = worst / scale
        out["rows_checked"] = R
    elif nest == "copy":
        out["f64"] = float(np.abs(y - inputs_np[0].ravel()).max())
    elif nest == "maxpool":
        ref = P.reference_output("maxpool2d", inputs_np).ravel()
        out["f64"] = float(np.abs(y - ref).max())
    elif nest == "matmul":
        out["f64"] = nw(y, P.reference_output("matmul", inputs_np).ravel())
    elif kind in ("matmul_f32", "matmul_bf16"):
        a, b = inputs_np
        M = a.shape[0]
        rows = np.unique(np.concatenate([np.arange(64), np.random.RandomState(1).randint(0, M, 64)]))
        yy = y.reshape(M, b.shape[1])[rows]
        out["f64"] = nw(yy, a[rows].astype(np.float64) @ b.astype(np.float64))
        if kind == "matmul_bf16":
            out["f64_rounded"] = nw(yy, P.round_bf16(a[rows]).astype(np.float64)
                                    @ P.round_bf16(b).astype(np.float64))
        out["rows_checked"] = int(rows.size)
    elif nest in ("conv2d", "conv2d_bf16"):
        X, W = inputs_np
        N, Co, H = X.shape[0], W.shape[0], X.shape[2]
        idx = P.conv_sample_indices(N, Co, H, H)
        got = y.reshape(N, Co, H, H)[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]
        out["f64"] = nw(got, P.conv_f64_at(X, W, idx))
        if nest == "conv2d_bf16":
            out["f64_rounded"] = nw(got, P.conv_f64_at(P.round_bf16(X), P.round_bf16(W), idx))
        out["outputs_checked"] = int(len(idx))
        out["sample"] = "every border pixel of batch planes 0 and N-1 of every output " \
                        "channel, plus 65,536 uniformly drawn outputs"
    tol = TOLERANCE.get(kind, {})
    out["ok"] = all(out[k] <= t for k, t in tol.items() if k in out)
    out["tolerance"] = tol
    return out


def cpu_baseline(name, inputs_np, shapes, seconds, bytes_per_step, flops_per_step):
    """Time the reference C backend (oracle/_ref, else its restatement) on
    host cores over a bounded sample of the same workload."""
    from loop_tool_amd import capi, programs as P
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_ref.so")
    kind = "reference"
    if os.path.exists(ref_path):
        L = capi.Library(ref_path)
    else:
        import loop_tool_amd as lt_mod
        lt_mod.load_lib(os.path.join(ROOT, "oracle", "liblt_oracle.so"))
        L, kind = lt_mod.lib, "port"
    # sample: a slice of the rows for softmax, the full nest otherwise
    if name == "softmax":
        rows = 512
        out, sh = P.softmax(L, rows, shapes[0][1])
        ins = [inputs_np[0][:rows]]
        scale = rows / shapes[0][0]
        sample = "%d of %d rows" % (rows, shapes[0][0])
    elif name.startswith("matmul_bf16") or name.startswith("matmul_f32"):
        rows = 16
        M, K = shapes[0]
        out, sh = P.matmul(L, rows, shapes[1][1], K)
        ins = [inputs_np[0][:rows], inputs_np[1]]
        scale = rows / M
        sample = "%d of %d rows of A (fp32)" % (rows, M)
    else:
        out, sh = workload(L, name)[:2]
        ins = inputs_np
        scale = 1.0
        sample = "full workload"
    lt = out.loop_tree
    cc = L.compile(lt, "cpp")
    y = np.zeros(lt.io_numel()[len(ins)], np.float32)
    ins = [np.ascontiguousarray(a, np.float32) for a in ins]
    cc.run(ins + [y])  # warm-up (page-in)
    reps, t0 = 0, time.perf_counter()
    while True:
        cc.run(ins + [y])
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds or reps >= 1000:
            break
    per_rep = el / reps
    if flops_per_step and not bytes_per_step:
        value, unit = flops_per_step * scale / per_rep / 1e12, "TFLOP/s"
    else:
        value, unit = bytes_per_step * scale / per_rep / 1e9, "GB/s"
    return dict(value=round(value, 4), unit=unit, cores=1, kind=kind,
                sample="%s, %d reps, %.1f s, single thread (reference C backend, "
                       "cc -O3 -std=c99)" % (sample, reps, el))


def batched_cpu_baseline(dims, seconds):
    """The reference C backend on a sample of the batch's GEMMs (fp32 matmul
    nests; compile time excluded), as TFLOP/s."""
    from loop_tool_amd import capi, programs as P
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_ref.so")
    kind = "reference"
    if os.path.exists(ref_path):
        L = capi.Library(ref_path)
    else:
        import loop_tool_amd as lt_mod
        lt_mod.load_lib(os.path.join(ROOT, "oracle", "liblt_oracle.so"))
        L, kind = lt_mod.lib, "port"
    flops, el, done = 0, 0.0, 0
    for i in range(len(dims)):
        M, N, K = dims[i]
        out, sh = P.matmul(L, M, N, K)
        cc = L.compile(out.loop_tree, "cpp")
        a, b = P.uniform(1, sh[0]), P.uniform(2, sh[1])
        y = np.zeros(M * N, np.float32)
        t0 = time.perf_counter()
        cc.run([a, b, y])
        el += time.perf_counter() - t0
        flops += 2 * M * N * K
        done += 1
        if el >= seconds:
            break
    return dict(value=round(flops / el / 1e12, 6), unit="TFLOP/s", cores=1, kind=kind,
                sample="the first %d of the %d GEMMs, fp32, %.1f s, single thread "
                       "(reference C backend, cc -O3 -std=c99)" % (done, len(dims), el))


def bench_batched(args):
    """512 independent matmul nests of uneven sizes (M, N, K log-uniform in
    [32, 2048], multiples of 16), bf16 (and fp16) operands, fp32 accumulation.
    The tuned form: per shape bucket a schedule is tuned in the context of
    the whole batch and stored in the hip backend's tuned store; the batch
    takes its tile configurations from the stored schedules and runs as one
    persistent grouped launch replayed from a hipGraph.  Plain forms: one
    hip-backend Compiled::run per matmul nest (default schedule, float32
    operands in HBM staged to 16 bits by each plan), and one launch of the
    library's default tile per GEMM.  A step is one pass over the batch."""
    import tempfile
    import loop_tool_amd as lt_mod
    from loop_tool_amd import autotune as A, batched as B, hip
    hip.init()
    L = lt_mod.lib
    dims = B.shapes(n=512, seed=6)
    stream = hip.Stream()
    flops = sum(2 * m * n * k for m, n, k in dims)
    # algorithmic HBM bytes: 16-bit operands read once, fp32 output written once
    nbytes = sum(2 * (m * k + k * n) + 4 * m * n for m, n, k in dims)
    store = A.TunedStore(L, tempfile.mkdtemp(prefix="lt_bench_tuned_"))
    out = {}
    dtypes = ("bf16",) if args.quick else ("bf16", "f16")

    def progress(msg):
        print("bench.py batched: " + msg, file=sys.stderr, flush=True)
    for dt in dtypes:
        progress("%s: operands" % dt)
        b = B.Batch(dims, dtype=dt, seed=6, stream=stream.h)
        if args.quick:
            b.set_configs([3] * len(dims))  # the tile the tuner picks for every bucket
            b.capture()
            for _ in range(args.warmup):
                b.replay()
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                b.replay()
            stream.synchronize()
            el = time.perf_counter() - t0
            e, e16 = batch_errors(b)
            out[dt] = dict(graph_ms=round(el / args.steps * 1e3, 5), configs="128x128x64 tile",
                           max_err_vs_f64=e, max_err_vs_f64_of_rounded_operands=e16)
            settle_steps = 0
            b.free()
            continue
        b.run_plain(B.DEFAULT_CFG)
        stream.synchronize()
        plain_ms = B.time_ms(lambda: b.run_plain(B.DEFAULT_CFG), iters=10, stream=stream.h)
        progress("%s: compiling 512 plain nests" % dt)
        t0 = time.perf_counter()
        pn = B.PlainNests(L, b, dtype=dt)
        compile_s = time.perf_counter() - t0
        progress("%s: compiled in %.1f s; timing" % (dt, compile_s))
        pn.run()
        stream.synchronize()
        plain_hip_ms = B.time_ms(pn.run, iters=5, stream=stream.h)
        del pn
        progress("%s: tuning the buckets" % dt)
        cfg, report = B.tune_buckets(b, iters=3, store=store, L=L)
        progress("%s: tuned; timing the graph replay" % dt)
        stored = B.configs_from_store(L, store, dims, dt)
        assert stored == cfg or [(B.TILES[c], B.TILE_K[c]) for c in stored] == [
            (B.TILES[c], B.TILE_K[c]) for c in cfg]
        b.set_configs(stored)
        grouped_ms = B.time_ms(lambda: b.run_grouped(), iters=10, stream=stream.h)
        b.capture()
        if dt == "bf16":
            settle_steps = settle(b.replay, stream.synchronize)
            for _ in range(args.warmup):
                b.replay()
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                b.replay()
            stream.synchronize()
            el = time.perf_counter() - t0
        # the dominant (only) kernel of a replay: HIP events on its stream
        graph_ms = B.time_ms(b.replay, iters=max(args.steps, 20), stream=stream.h)
        e, e16 = batch_errors(b)
        out[dt] = dict(plain_hip_backend_ms=round(plain_hip_ms, 5),
                       plain_hip_backend_compile_s=round(compile_s, 2),
                       plain_library_ms=round(plain_ms, 5), grouped_ms=round(grouped_ms, 5),
                       graph_ms=round(graph_ms, 5),
                       speedup_vs_plain_hip_backend=round(plain_hip_ms / graph_ms, 3),
                       speedup_vs_plain_library=round(plain_ms / graph_ms, 3),
                       work_items=b.num_items(), tuning=report,
                       configs_from="schedules stored per shape bucket (%d) in the hip "
                                    "backend's tuned store" % len(report["buckets"]),
                       max_err_vs_f64=e, max_err_vs_f64_of_rounded_operands=e16,
                       error_sample="48 random rows (all rows when fewer) of every GEMM, "
                                    "after the graph replays")
        b.free()
    A.TunedStore.reset(L)
    kern_ms = out["bf16"]["graph_ms"]
    achieved = nbytes / (kern_ms * 1e-3) / 1e9
    line = {
        "metric": "TFLOP/s (algorithmic flops / time), batched uneven GEMMs, hip backend",
        "value": round(flops * args.steps / el / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (SplitMix64-seeded sizes and operands)",
        "config": dict(workload="512 independent matmul nests, M,N,K log-uniform in "
                                "[32, 2048] (multiples of 16), bf16 operands, fp32 accumulate",
                       gemms=len(dims), algorithmic_flops=flops, algorithmic_bytes=nbytes,
                       settle_steps=settle_steps, settle_s=SETTLE_S,
                       bf16=out["bf16"], f16=out.get("f16"), quick=bool(args.quick),
                       note="the batch's arithmetic intensity (flops/bytes) is below the "
                            "MI355X ridge point (2.5 PFLOP/s / 8 TB/s), so HBM bounds it"),
        "roofline": dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                         unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = batched_cpu_baseline(dims, args.cpu_seconds)
    print(json.dumps(line))
    bad = [dt for dt in out if out[dt]["max_err_vs_f64"] > 1e-2
           or out[dt]["max_err_vs_f64_of_rounded_operands"] > 1e-5]
    if bad:
        print("bench.py: batched result outside its error bar: %s" % bad, file=sys.stderr)
        sys.exit(3)


def batch_errors(b, rows_per_gemm=48, seed=13):
    """(worst normwise error vs float64 of the float operands, worst vs
    float64 of the 16-bit-rounded operands) over sampled rows of every GEMM
    of the batch."""
    from loop_tool_amd import programs as P
    rnd = P.round_bf16 if b.dtype == 0 else (lambda a: a.astype(np.float16).astype(np.float32))
    rng = np.random.RandomState(seed)
    worst = worst16 = 0.0
    for i, (m, n, k) in enumerate(b.dims):
        a, bb = b.host[i]
        rows = np.sort(rng.choice(m, min(m, rows_per_gemm), replace=False))
        got = b.result(i)[rows].astype(np.float64)
        ref = a[rows].astype(np.float64) @ bb.astype(np.float64)
        ref16 = rnd(a[rows]).astype(np.float64) @ rnd(bb).astype(np.float64)
        worst = max(worst, float(np.abs(got - ref).max() / np.abs(ref).max()))
        worst16 = max(worst16, float(np.abs(got - ref16).max() / np.abs(ref16).max()))
    return worst, worst16


def serving_cpu_baseline(seconds, seed=7):
    """The reference C backend serving the same request mix one at a time on
    one host core: every new shape is compiled (cc -O3) and run, as the
    reference's lazy compile cache does for unseen shapes."""
    from loop_tool_amd import capi, programs as P
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libloop_tool_ref.so")
    kind = "reference"
    if os.path.exists(ref_path):
        L = capi.Library(ref_path)
    else:
        import loop_tool_amd as lt_mod
        lt_mod.load_lib(os.path.join(ROOT, "oracle", "liblt_oracle.so"))
        L, kind = lt_mod.lib, "port"
    rng = np.random.RandomState(seed)
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        u = rng.rand()
        if u < 0.4:
            M, N, K = (int(np.clip(np.round(np.exp(rng.uniform(np.log(32), np.log(2048))) / 16)
                                   * 16, 32, 2048)) for _ in range(3))
            out, sh = P.matmul(L, M, N, K)
        elif u < 0.7:
            out, sh = P.softmax(L, int(rng.randint(256, 8193)), int(rng.randint(256, 8193)))
        elif u < 0.9:
            N, C, H = int(rng.randint(1, 17)), int(rng.randint(32, 257)), int(rng.randint(14, 113))
            out, sh = P.conv2d(L, N, C, C, H, H)
        else:
            N, C, H = int(rng.randint(1, 17)), int(rng.randint(32, 257)), int(rng.randint(14, 113))
            out, sh = P.maxpool2d(L, N, C, H, H)
        t = out.loop_tree
        cc = L.compile(t, "cpp")
        ins = [P.uniform(i + 1, s_) for i, s_ in enumerate(sh)]
        y = np.zeros(t.io_numel()[len(ins)], np.float32)
        cc.run(ins + [y])
        done += 1
    el = time.perf_counter() - t0
    return dict(value=round(done / el, 4), unit="req/s", cores=1, kind=kind,
                sample="%d requests of the same mix served one at a time in %.1f s "
                       "(compile per new shape + run, reference C backend, single thread)"
                       % (done, el))


def bench_serving(args):
    """Open-loop serving of lazy programs (40% matmul, 30% softmax, 20% conv2d,
    10% max-pool, every shape drawn afresh) arriving as a Poisson process."""
    from loop_tool_amd import hip, serve as S
    hip.init()
    srv = S.Server(budget_gb=200.0, slots=args.slots)
    srv.run_workload(rate=args.rate, duration_s=1.0, seed=99)  # warm-up: code, pools
    srv.reset_stats()
    st = srv.run_workload(rate=args.rate, duration_s=args.duration, seed=7)
    srv.close()
    wall = st["wall_s"]
    tflops = st["flops"] / wall / 1e12
    gbs = st["bytes"] / wall / 1e9
    line = {
        "metric": "served requests/s, Poisson arrivals of mixed-shape lazy programs "
                  "(p50/p99 latency in config)",
        "value": round(st["completed"] / wall, 2),
        "unit": "req/s",
        "n_gpus": 1,
        "steps": int(st["completed"]),
        "warmup": 1,
        "ms_per_step": round(st["mean_ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bf16 operands for contractions)",
        "data": "synthetic (SplitMix64-seeded shapes and inputs, materialised in HBM)",
        "config": dict(workload="serving: Poisson %.0f req/s for %.0f s; 40%% matmul "
                                "(M,N,K log-uniform 32-2048), 30%% softmax (rows, cols "
                                "256-8192), 20%% conv2d 3x3 (N 1-16, C 32-256, H=W 14-112), "
                                "10%% 2x2 max-pool; budget 200 GB" % (args.rate, args.duration),
                       slots=args.slots, p50_ms=round(st["p50_ms"], 4),
                       p99_ms=round(st["p99_ms"], 4), max_ms=round(st["max_ms"], 4),
                       submitted=st["submitted"], completed=st["completed"],
                       failed=st["failed"], families=st["family_count"],
                       aggregate_tflops=round(tflops, 3), aggregate_gbs=round(gbs, 2),
                       structural_cache=dict(hits=st["struct_hits"], misses=st["struct_misses"]),
                       hip_backend_compiles=st["compiles"],
                       peak_device_gb=round(st["peak_bytes"] / 1e9, 3),
                       admission_waits=st["admission_waits"],
                       client_lag_ms=round(st["client_lag_ms"], 4),
                       host_ms=dict(plan=round(st["host_plan_ms"], 1),
                                    launch=round(st["host_launch_ms"], 1),
                                    retire=round(st["host_retire_ms"], 1)),
                       device_allocs=st["device_allocs"]),
        "roofline": dict(bound="hbm", achieved=round(gbs, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = serving_cpu_baseline(args.cpu_seconds)
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.nest == "batched":
        return bench_batched(args)
    if args.nest == "serving":
        return bench_serving(args)
    if args.nest == "pack":
        return bench_pack(args)
    import loop_tool_amd as lt_mod
    from loop_tool_amd import hip

    L = lt_mod.lib
    hip.init()
    out, shapes, gen, cfg, bound, dtype, backend = workload(L, args.nest)
    lt = out.loop_tree
    if args.annotate:
        lt = lt.annotate(lt.roots[0], args.annotate)
        cfg = dict(cfg, knobs=args.annotate)
    cc = L.compile(lt, backend)
    inputs_np = gen()
    n_out = lt.io_numel()[len(inputs_np)]
    d_in = [hip.DeviceArray.from_numpy(a) for a in inputs_np]
    d_out = hip.DeviceArray(n_out)
    io = d_in + [d_out]
    st = hip.stages(cc)
    bytes_step = sum(s["bytes"] for s in st)
    flops_step = sum(s["flops"] for s in st)

    # correctness against float64 on the device result
    cc.run([b.ptr for b in io])
    hip.synchronize()
    errors = nest_errors(args.nest, inputs_np, d_out.download())

    ptrs = [b.ptr for b in io]
    settle_steps = settle(lambda: cc.run(ptrs, sync=False), hip.synchronize)
    for _ in range(args.warmup):
        cc.run([b.ptr for b in io], sync=False)
    hip.synchronize()
    # the timed steps, bracketed by HIP events on the plan's stream as well
    stream = hip.compiled_stream(cc)
    e0, e1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        cc.run([b.ptr for b in io], sync=False)
    e1.record(stream)
    hip.synchronize()
    el = time.perf_counter() - t0
    ms = el / args.steps * 1e3
    gpu_step_ms = e0.elapsed_ms(e1) / args.steps

    # dominant kernel: a one-kernel plan is timed by the events around the
    # timed steps; otherwise its own launches are timed with HIP events on
    # the plan's stream (median of three runs)
    key = "bytes" if bound == "hbm" else "flops"
    dom = max(range(len(st)), key=lambda i: st[i][key])
    if len(st) == 1:
        kern_ms = gpu_step_ms
    else:
        kern_ms = sorted(hip.time_stage_ms(cc, io, dom, warmup=3, iters=max(args.steps, 20))
                         for _ in range(3))[1]
    plan_ms = hip.time_ms(cc, io, warmup=3, iters=max(args.steps, 20))
    if bound == "hbm":
        value = bytes_step * args.steps / el / 1e9
        unit = "GB/s"
        achieved = st[dom]["bytes"] / (kern_ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=MEASURED_TRAFFIC.get(args.nest))
        metric = "HBM GB/s (algorithmic bytes / time), hip backend"
        plan_frac = bytes_step / (plan_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    else:
        value = flops_step * args.steps / el / 1e12
        unit = "TFLOP/s"
        achieved = st[dom]["flops"] / (kern_ms * 1e-3) / 1e12
        peak = MFMA_PEAK_TFLOPS[dtype]
        roof = dict(bound="mfma", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                    frac=round(achieved / peak, 4), traffic=None)
        metric = "TFLOP/s (algorithmic flops / time), hip backend"
        plan_frac = flops_step / (plan_ms * 1e-3) / 1e12 / peak

    line = {
        "metric": metric,
        "value": round(value, 3),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (SplitMix64-seeded inputs)",
        "config": dict(cfg, kernels=[s["kernel"] + ":" + s["kind"] for s in st],
                       algorithmic_bytes=bytes_step, algorithmic_flops=flops_step,
                       dominant_kernel=st[dom]["kernel"], kernel_ms=round(kern_ms, 5),
                       gpu_ms_per_step=round(gpu_step_ms, 5),
                       plan_ms=round(plan_ms, 5), plan_frac=round(plan_frac, 4),
                       max_err_vs_f64=errors.get("f64"), errors=errors,
                       settle_steps=settle_steps, settle_s=SETTLE_S),
        "roofline": roof,
    }
    if args.nest == "softmax" and not args.no_conv:
        line["config"]["conv2d_bf16"] = conv_bf16_figures(L, hip)
    if args.nest == "softmax" and not args.no_batched:
        line["config"]["batched_bf16"] = batched_figures(hip)
    if args.nest == "matmul":
        line["config"]["launch_floor"] = launch_floor(L, hip)
        line["config"]["workgroups"] = st_grid(cc)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.nest, inputs_np, shapes, args.cpu_seconds,
                                            bytes_step if bound == "hbm" else 0,
                                            flops_step if bound != "hbm" else 0)
    print(json.dumps(line))
    bad = [k for k, v in [("", errors)] + [
        (k, line["config"].get(k, {}).get("errors", {"ok": True}))
        for k in ("conv2d_bf16", "batched_bf16")]
        if not v["ok"]]
    if bad:
        print("bench.py: result outside its error bar: %s" % (bad,), file=sys.stderr)
        sys.exit(3)


def st_grid(cc):
    """Workgroups of each stage of a Compiled (from its plan text)."""
    import re
    return [int(x) for x in re.findall(r"grid (\d+)", cc.string_property("grid") or "")] or None


def launch_floor(L, hip, iters=200):
    """The smallest hip-backend kernel: a one-element pointwise nest, launched
    back to back on its plan's stream; mean milliseconds per launch between
    HIP events (the floor under any one-kernel plan's step time)."""
    from loop_tool_amd import programs as P
    out, _ = P.add_broadcast(L, 1, 1)
    cc = L.compile(out.loop_tree, "hip")
    io = [hip.DeviceArray(1), hip.DeviceArray(1), hip.DeviceArray(1)]
    ms = sorted(hip.time_ms(cc, io, warmup=20, iters=iters) for _ in range(3))[1]
    return dict(kernel="one-element pointwise nest", ms_per_launch=round(ms, 5),
                note="HIP events around %d back-to-back launches, median of 3" % iters)


def conv_bf16_figures(L, hip):
    """The BF16 convolution N=32, C 128 -> 128, 56 x 56, 3 x 3 pad 1 measured
    in the same run as the softmax: the MFMA kernel's and the whole plan's
    (staging + kernel) fraction of the 2.5 PFLOP/s bf16 peak, each the median
    of three HIP-event timings of 20 launches, and the border-inclusive
    sampled error against float64."""
    out, shapes, gen, cfg, _, _, backend = workload(L, "conv2d_bf16")
    cc = L.compile(out.loop_tree, backend)
    inputs_np = gen()
    d_in = [hip.DeviceArray.from_numpy(a) for a in inputs_np]
    d_out = hip.DeviceArray(out.loop_tree.io_numel()[2])
    io = d_in + [d_out]
    st = hip.stages(cc)
    flops = sum(s["flops"] for s in st)
    dom = max(range(len(st)), key=lambda i: st[i]["flops"])
    cc.run([b.ptr for b in io])
    hip.synchronize()
    errors = nest_errors("conv2d_bf16", inputs_np, d_out.download())
    ptrs = [b.ptr for b in io]
    settle(lambda: cc.run(ptrs, sync=False), hip.synchronize)
    kern_ms = sorted(hip.time_stage_ms(cc, io, dom, warmup=3, iters=20) for _ in range(3))[1]
    plan_ms = sorted(hip.time_ms(cc, io, warmup=3, iters=20) for _ in range(3))[1]
    peak = MFMA_PEAK_TFLOPS["bf16"]
    kt = st[dom]["flops"] / (kern_ms * 1e-3) / 1e12
    pt = flops / (plan_ms * 1e-3) / 1e12
    return dict(workload=cfg["workload"] + ", bf16 operands, fp32 accumulate",
                kernels=[s["kernel"] + ":" + s["kind"] for s in st],
                plan=cc.dump().splitlines()[0], algorithmic_flops=flops,
                kernel_ms=round(kern_ms, 5), kernel_tflops=round(kt, 2),
                kernel_frac=round(kt / peak, 4), plan_ms=round(plan_ms, 5),
                plan_tflops=round(pt, 2), plan_frac=round(pt / peak, 4), peak_tflops=peak,
                errors=errors)


def bench_pack(args):
    """64 nests of the serving mix (loop_tool_amd/packed.py: matmul, row
    softmax, 3x3 conv, 2x2 max-pool of random sizes, seed 7), each compiled by
    the hip backend with its own device inputs, output and scratch.  A step is
    one pass over all of them: as one graph replay of a pack over 8 streams
    (lth_pack_*).  Beside it, the same pass as one launch sequence on one
    stream (lth_run_on per nest) and as packs over 1 and 4 streams; every
    form's outputs are compared with the launch sequence's, bit for bit."""
    import loop_tool_amd as lt_mod
    from loop_tool_amd import hip, packed
    hip.init()
    L = lt_mod.lib
    specs = packed.serving_mix(n=64, seed=7)
    t0 = time.perf_counter()
    compiled = packed.compile_nests(L, specs)
    compile_s = time.perf_counter() - t0
    rng = np.random.default_rng(7)
    nests, scratch = [], []
    for cc, tree, shapes in compiled:
        d_in = [hip.DeviceArray.from_numpy(rng.standard_normal(int(np.prod(sh)), np.float32))
                for sh in shapes]
        d_out = hip.DeviceArray(tree.io_numel()[len(shapes)])
        nests.append((cc, d_in + [d_out]))
        scratch.append([hip.DeviceArray(max(w, 1)) for w in hip.scratch_sizes(cc)])
    flops = sum(sum(st["flops"] for st in hip.stages(cc)) for cc, _ in nests)
    nbytes = sum(sum(st["bytes"] for st in hip.stages(cc)) for cc, _ in nests)
    stream = hip.Stream()

    def sequence():
        for (cc, io), sc in zip(nests, scratch):
            hip.run_on(cc, io, sc, stream.h)

    def timed(fn, iters=10):
        fn()
        stream.synchronize()
        settle(fn, stream.synchronize)
        e0, e1 = hip.Event(), hip.Event()
        best = []
        for _ in range(3):
            e0.record(stream.h)
            for _ in range(iters):
                fn()
            e1.record(stream.h)
            e1.synchronize()
            best.append(e0.elapsed_ms(e1) / iters)
        return sorted(best)[1]

    seq_ms = timed(sequence)
    want = [io[-1].download() for _, io in nests]
    forms, same = {}, {}
    for lanes in (1, 4, 8):
        for _, io in nests:
            io[-1].fill_zero()
        hip.synchronize()
        pk = hip.Pack(nests, lanes=lanes, stream=stream.h)
        forms[lanes] = timed(pk.replay)
        same[lanes] = all(np.array_equal(io[-1].download().view(np.uint32), w.view(np.uint32))
                          for (_, io), w in zip(nests, want))
        if lanes == 8:
            for _ in range(args.warmup):
                pk.replay()
            stream.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.steps):
                pk.replay()
            stream.synchronize()
            el = time.perf_counter() - t1
        pk.free()
    # the same packs with the plain 16-bit matmul nests grouped into one
    # persistent GEMM launch: those within 1e-5 (normwise) of the sequence,
    # the rest bit for bit
    grouped = [cc.int_property("plain_matmul") == 1 and
               hip.stages(cc)[-1]["kind"] in ("contract_bf16", "contract_f16")
               for cc, _ in nests]
    gforms, gerr, gsame = {}, 0.0, True
    for lanes in (1, 8):
        for _, io in nests:
            io[-1].fill_zero()
        hip.synchronize()
        pk = hip.Pack(nests, lanes=lanes, stream=stream.h, group_matmuls=True)
        gforms[lanes] = timed(pk.replay)
        for (_, io), w, g in zip(nests, want, grouped):
            got = io[-1].download()
            if g:
                w64 = w.astype(np.float64)
                gerr = max(gerr, float(np.linalg.norm(got.astype(np.float64) - w64) /
                                       max(np.linalg.norm(w64), 1e-30)))
            else:
                gsame = gsame and np.array_equal(got.view(np.uint32), w.view(np.uint32))
        pk.free()
    counts = {}
    for name, _, _ in specs:
        counts[name] = counts.get(name, 0) + 1
    line = {
        "metric": "mixed nests per second (one graph replay of a pack), hip backend",
        "value": round(len(nests) * args.steps / el, 1),
        "unit": "nests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+f32",
        "data": "synthetic (numpy normal inputs, seeded sizes)",
        "config": dict(workload="64 nests of the serving mix in one pass", nests=counts,
                       algorithmic_flops=flops, algorithmic_bytes=nbytes,
                       compile_s=round(compile_s, 1),
                       one_stream_sequence_ms=round(seq_ms, 4),
                       pack_ms={str(k): round(v, 4) for k, v in forms.items()},
                       speedup_vs_sequence={str(k): round(seq_ms / v, 3) for k, v in forms.items()},
                       bit_identical_to_sequence={str(k): v for k, v in same.items()},
                       grouped_matmuls=dict(
                           nests=int(sum(grouped)),
                           pack_ms={str(k): round(v, 4) for k, v in gforms.items()},
                           max_normwise_err_vs_sequence=gerr, tolerance=1e-5,
                           other_nests_bit_identical=gsame)),
    }
    print(json.dumps(line))
    if not all(same.values()) or not gsame or gerr > 1e-5:
        print("bench.py: a pack's outputs differ from the launch sequence's", file=sys.stderr)
        sys.exit(3)


def batched_figures(hip):
    """The 512 uneven bf16 GEMMs of `--nest batched` (M, N, K log-uniform in
    [32, 2048], multiples of 16, seed 6) measured in the same run as the
    softmax: one graph-replayed launch with the 128x128x64 tile on every GEMM
    (the tile the tuner picks for every shape bucket of this batch; `--nest
    batched` runs the tuning itself), the median of three HIP-event timings
    of 20 replays, its fraction of the HBM roofline, and the error on 48
    sampled rows of every GEMM."""
    from loop_tool_amd import batched as B
    dims = B.shapes(n=512, seed=6)
    flops = sum(2 * m * n * k for m, n, k in dims)
    nbytes = sum(2 * (m * k + k * n) + 4 * m * n for m, n, k in dims)
    stream = hip.Stream()
    b = B.Batch(dims, dtype="bf16", seed=6, stream=stream.h)
    b.set_configs([3] * len(dims))
    b.capture()
    settle(b.replay, stream.synchronize)
    ms = sorted(B.time_ms(b.replay, iters=20, stream=stream.h) for _ in range(3))[1]
    e, e16 = batch_errors(b)
    items = b.num_items()
    b.free()
    gbs = nbytes / (ms * 1e-3) / 1e9
    tol = {"f64": 1e-2, "f64_rounded": 1e-5}
    return dict(workload="512 independent matmul nests, M,N,K log-uniform in [32, 2048] "
                         "(multiples of 16), bf16 operands, fp32 accumulate, one "
                         "graph-replayed launch, 128x128x64 tile on every GEMM",
                gemms=len(dims), work_items=items, algorithmic_flops=flops,
                algorithmic_bytes=nbytes, graph_ms=round(ms, 5),
                tflops=round(flops / (ms * 1e-3) / 1e12, 2), hbm_gbs=round(gbs, 1),
                hbm_frac=round(gbs / HBM_PEAK_GBS, 4),
                errors=dict(f64=e, f64_rounded=e16, rows="48 random rows of every GEMM",
                            ok=e <= tol["f64"] and e16 <= tol["f64_rounded"], tolerance=tol))


if __name__ == "__main__":
    main()
