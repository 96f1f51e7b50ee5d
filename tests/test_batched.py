"""Batches of uneven, run-time-sized matmul nests on the GPU.

Every GEMM of a batch is checked against float64 of its 16-bit-rounded
operands (the only rounding left is the float32 accumulation order, bound
1e-5 normwise) and against float64 of the float32 operands (bf16 operand
rounding, bound 1e-2).  The grouped persistent launch, its hipGraph replay and
the tuned per-bucket tile choice must agree bit-for-bit with one launch per
GEMM of the same tile shape: every output element is reduced by one lane in
the same K order whichever way the tiles are scheduled.
"""
import numpy as np
import pytest

from loop_tool_amd import programs as P

pytestmark = pytest.mark.gpu

EDGE_DIMS = [(32, 32, 32), (48, 80, 16), (2048, 32, 64), (33 * 16, 2000, 16 * 7),
             (16, 16, 8), (100, 36, 24), (256, 128, 2048)]


@pytest.fixture(scope="module")
def bat(hip_dev):
    from loop_tool_amd import batched
    return batched


@pytest.mark.parametrize("dtype", ["bf16", "f16"])
def test_plain_every_tile_vs_f64(bat, hip_dev, dtype):
    b = bat.Batch(EDGE_DIMS, dtype=dtype, seed=11)
    for cfg in range(len(bat.TILES)):
        b.run_plain(cfg)
        hip_dev.synchronize()
        for i, d in enumerate(EDGE_DIMS):
            e, e16 = b.check(i)
            print(dtype, bat.TILES[cfg], d, e, e16, flush=True)
            assert e16 <= 1e-5, (cfg, d, e16)
            assert e <= 1e-2, (cfg, d, e)


@pytest.mark.parametrize("cfgs", [None, [0, 1, 2], [0, 1, 2, 6]], ids=["all", "small", "mid"])
def test_grouped_and_graph_match_plain(bat, hip_dev, cfgs):
    """Each register class of the grouped kernel (chosen by the batch's
    largest configuration) matches the plain launches to the bit."""
    dims = bat.shapes(n=64, seed=7, lo=32, hi=1024)
    b = bat.Batch(dims, dtype="bf16", seed=3)
    rng = np.random.RandomState(0)
    choice = cfgs if cfgs is not None else list(range(len(bat.TILES)))
    cfg = [int(choice[c]) for c in rng.randint(0, len(choice), len(dims))]
    b.set_configs(cfg)
    b.run_plain(-1)
    hip_dev.synchronize()
    want = [b.result(i) for i in range(len(dims))]
    for c in b.C:
        c.fill_zero()
    # the work queues reset themselves after each launch: launches in a row,
    # with fewer workgroups than XCD segments (each runs whole segments
    # alone), fewer than the GPU holds, as many, and more than there are
    # items (the surplus finds the queues empty)
    for grid in (97, 3, 0, 100000, 0):
        b.run_grouped(grid=grid)
        hip_dev.synchronize()
        for i in range(len(dims)):
            assert np.array_equal(b.result(i), want[i]), ("grouped", grid, i, dims[i], cfg[i])
        for c in b.C:
            c.fill_zero()
    b.capture(grid=256 if cfgs is None else 0)  # 0: the occupancy-sized grid
    for rep in range(3):
        b.replay()
        hip_dev.synchronize()
        for i in range(len(dims)):
            assert np.array_equal(b.result(i), want[i]), ("replay", rep, i)
        for c in b.C:  # eager work between replays
            c.fill_zero()
    # reconfiguring invalidates the captured graph
    b.set_configs([0] * len(dims))
    with pytest.raises(Exception):
        b.replay()


@pytest.mark.parametrize("cfg", range(7))
def test_one_configuration_lists_match_plain(bat, hip_dev, cfg):
    """A batch whose GEMMs all use one tile configuration runs that
    configuration's kernel over per-workgroup lists built for the launch's
    grid, each item's first K tile loaded during the previous item's last.
    It matches the plain launches to the bit at every grid size (fewer
    workgroups than XCD segments, more than there are items, the occupancy-
    sized grid) and under graph replay, with eager work between replays."""
    dims = bat.shapes(n=48, seed=17, lo=16, hi=700) + EDGE_DIMS
    b = bat.Batch(dims, dtype="bf16", seed=4)
    b.set_configs([cfg] * len(dims))
    b.run_plain(-1)
    hip_dev.synchronize()
    want = [b.result(i) for i in range(len(dims))]
    for c in b.C:
        c.fill_zero()
    for grid in (97, 3, 0, 100000, 5, 0):
        b.run_grouped(grid=grid)
        hip_dev.synchronize()
        for i in range(len(dims)):
            assert np.array_equal(b.result(i), want[i]), ("grouped", cfg, grid, i, dims[i])
        for c in b.C:
            c.fill_zero()
    b.capture()
    for rep in range(3):
        if rep == 2:  # a launch with another grid builds lists of its own
            b.run_grouped(grid=97)
            hip_dev.synchronize()
            for c in b.C:
                c.fill_zero()
        b.replay()
        hip_dev.synchronize()
        for i in range(len(dims)):
            assert np.array_equal(b.result(i), want[i]), ("replay", cfg, rep, i)
        for c in b.C:
            c.fill_zero()


def test_tuned_batch_vs_f64(bat, hip_dev, lt, tmp_path):
    """Tune the buckets, store one schedule per bucket, and build a fresh
    batch from the stored schedules: its configurations are the tuned ones,
    its graph replay equals one launch per GEMM of those configurations bit
    for bit, and every GEMM is within the bars of float64."""
    from loop_tool_amd.autotune import TunedStore
    dims = bat.shapes(n=48, seed=5, lo=32, hi=512)
    b = bat.Batch(dims, dtype="bf16", seed=9)
    store = TunedStore(lt, str(tmp_path / "tuned"))
    try:
        cfg, report = bat.tune_buckets(b, iters=2, store=store, L=lt)
        assert set(cfg) <= set(range(len(bat.TILES)))
        assert len(store.keys()) == len(report["buckets"])
        assert report["tuned_ms"] <= min(report["uniform_ms"].values()) * 1.0001
        stored = bat.configs_from_store(lt, store, dims)
        want = [(bat.TILES[c], bat.TILE_K[c]) for c in cfg]
        assert [(bat.TILES[c], bat.TILE_K[c]) for c in stored] == want
        b2 = bat.Batch(dims, dtype="bf16", seed=9)
        b2.set_configs(stored)
        b2.run_plain(-1)
        b2.synchronize()
        plain = [b2.result(i) for i in range(len(dims))]
        for c in b2.C:
            c.fill_zero()
        b2.capture()
        b2.replay()
        b2.synchronize()
        for i in range(len(dims)):
            assert np.array_equal(b2.result(i), plain[i]), (i, dims[i])
            e, e16 = b2.check(i)
            assert e16 <= 1e-5 and e <= 1e-2, (i, dims[i], e, e16)
    finally:
        TunedStore.reset(lt)


def test_plain_nests_match_the_batch(bat, hip_dev, lt):
    """The plain form of the batch -- one hip-backend Compiled per matmul
    nest, launched one after another -- against float64 of the rounded
    operands, GEMM by GEMM."""
    dims = bat.shapes(n=24, seed=5, lo=32, hi=512)
    b = bat.Batch(dims, dtype="bf16", seed=9)
    pn = bat.PlainNests(lt, b)
    pn.run()
    for i in range(len(dims)):
        a, bb = b.host[i]
        ref16 = P.round_bf16(a).astype(np.float64) @ P.round_bf16(bb).astype(np.float64)
        y = pn.result(i)
        assert np.abs(y - ref16).max() / np.abs(ref16).max() <= 1e-5, (i, dims[i])


def test_batch_rejects_bad_sizes(bat, hip_dev):
    with pytest.raises(Exception, match="multiple of 8"):
        bat.Batch([(32, 32, 12)])
