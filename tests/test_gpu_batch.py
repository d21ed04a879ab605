"""GPU tests of the multi-scene persistent batch, Morton ordering, hipGraph
capture/replay and the autotuner: every item's hits must equal the CPU
oracle's for that scene, bit for bit, in every form."""
import gc
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-batch] {msg}", flush=True)
    sys.stdout.flush()


def bits(h):
    return np.ascontiguousarray(h, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def batch_set(cw, orc):
    """Uneven items over six scenes (two with 128-byte nodes), oracle hits."""
    from cwbvh_amd import workloads as W

    rng = np.random.default_rng(21)
    scenes, items, refs = [], [], []
    sizes = [0, 1, 63, 65, 4096, 30000, 7, 100000, 2500, 50000, 333, 12345]
    for i, n in enumerate(sizes):
        if i < 6:
            m = W.heightfield_of(int(rng.integers(1000, 60000)), seed=100 + i) if i % 2 == 0 \
                else W.icospheres_of(int(rng.integers(1000, 60000)), seed=100 + i)
            params = cw.default_build_params(cwbvh_node_stride=8 if i in (1, 4) else 5)
            _, lay = cw.build_layouts(m, params, bvh2=False)
            scenes.append((m, lay, cw.Scene(lay)))
        m, lay, sc = scenes[i % 6]
        diag = m.diagonal()
        rays = cw.rays_ao_surface(m, (n + 7) // 8 + 1, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=200 + i)[:n]
        items.append((sc, lay, rays))
        refs.append({a: orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, rays, any_hit=a) for a in (False, True)})
    say(f"batch set: {len(items)} items, {sum(len(r) for _, _, r in items)} rays")
    return scenes, items, refs


def device_items(cw, items):
    bufs, out = [], []
    for sc, lay, rays in items:
        dr = cw.DeviceBuffer.from_numpy(rays if len(rays) else np.zeros((1, 8), np.float32))
        dh = cw.DeviceBuffer(max(16, 16 * len(rays)))
        bufs.append((dr, dh))
        out.append((sc, dr.ptr, dh.ptr, len(rays)))
    return bufs, out


def check_items(cw, items, bufs, refs, any_hit, what):
    for k, ((sc, lay, rays), (dr, dh), ref) in enumerate(zip(items, bufs, refs)):
        if len(rays) == 0:
            continue
        g = dh.download((len(rays), 4))
        diff = (bits(g) != bits(ref[any_hit])).any(axis=1)
        assert not diff.any(), f"{what}: item {k} has {int(diff.sum())} of {len(rays)} rays differing"


@pytest.mark.parametrize("morton", [False, True])
@pytest.mark.parametrize("any_hit", [False, True])
def test_batch_matches_oracle(cw, batch_set, morton, any_hit):
    _, items, refs = batch_set
    bufs, ditems = device_items(cw, items)
    say(f"batch launch morton={morton} any={any_hit}")
    b = cw.Batch(ditems, cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST, morton=morton)
    info = b.info()
    assert info["num_jobs"] == sum(1 for _, _, r in items if len(r)) and info["total_rays"] == sum(len(r) for _, _, r in items)
    b.launch()
    cw.synchronize()
    check_items(cw, items, bufs, refs, any_hit, f"batch morton={morton}")
    assert not cw.stack_overflowed()


@pytest.mark.parametrize("kw", [dict(block_size=64, lds_depth=4, fetch_chunk=1),
                                dict(block_size=256, lds_depth=32, fetch_chunk=256, prefetch=1),
                                dict(postpone_pct=0, refill_idle=0, refill_budget=1, grid_blocks=5),
                                dict(loop_mode=1, block_size=64, lds_depth=2, fetch_chunk=3),
                                dict(loop_mode=2, block_size=128, lds_depth=8, fetch_chunk=64, prefetch=1),
                                # the autotuner's usual picks on many-scene batches
                                dict(loop_mode=2, block_size=256, lds_depth=8),
                                dict(loop_mode=2, block_size=64, lds_depth=8),
                                dict(loop_mode=1, block_size=64, lds_depth=8),
                                # every tree's stack fits the LDS depth: the LDS-only batch kernels
                                dict(loop_mode=1, block_size=128, lds_depth=64),
                                dict(loop_mode=1, block_size=256, lds_depth=40, prefetch=1)])
def test_batch_configs(cw, batch_set, kw):
    _, items, refs = batch_set
    bufs, ditems = device_items(cw, items)
    say(f"batch config {kw}")
    b = cw.Batch(ditems, cw.TRACE_CLOSEST, cw.default_trace_config(cw.KIND_CWBVH, **kw))
    b.launch()
    cw.synchronize()
    check_items(cw, items, bufs, refs, False, f"batch {kw}")


def test_graph_replays_identical(cw, batch_set):
    _, items, refs = batch_set
    bufs, ditems = device_items(cw, items)
    st = cw.Stream()
    for morton in (False, True):
        say(f"graph capture/replay morton={morton}")
        b = cw.Batch(ditems, cw.TRACE_CLOSEST, morton=morton)
        b.capture(st)
        for rep in range(4):
            for _, dh in bufs:
                dh.zero()
            b.replay(st)
            st.synchronize()
            check_items(cw, items, bufs, refs, False, f"replay {rep} morton={morton}")
            # eager work on the same stream between replays
            b.launch(st)
            st.synchronize()
            check_items(cw, items, bufs, refs, False, f"eager after replay {rep}")


def test_prim_id_batch(cw, batch_set):
    _, items, refs = batch_set
    bufs, ditems = device_items(cw, items)
    b = cw.Batch(ditems, cw.TRACE_PRIM_ID)
    b.launch()
    cw.synchronize()
    for (sc, lay, rays), (dr, dh), ref in zip(items, bufs, refs):
        if len(rays) == 0:
            continue
        ids = cw.hit_ids(ref[False])
        want = np.where(ids >= 0, lay.prim_idx[np.clip(ids, 0, None)], -1)
        assert np.array_equal(cw.hit_ids(dh.download((len(rays), 4))), want)


def test_morton_order_single_scene(cw, orc, batch_set):
    scenes, items, refs = batch_set
    sc, lay, rays = items[7]
    n = len(rays)
    lo, hi = sc.bounds()
    dr = cw.DeviceBuffer.from_numpy(rays)
    do = cw.DeviceBuffer(4 * n)
    cw.morton_order(dr.ptr, n, lo, hi, do.ptr)
    order = do.download((n,), np.int32)
    assert np.array_equal(np.sort(order), np.arange(n))
    # sorted by the ray key (direction octant, then origin Morton code)
    assert np.all(np.diff(_ray_keys(rays[order], lo, hi)) >= 0)
    g = sc.trace_numpy(rays, 0, order=order)
    assert np.array_equal(bits(g), bits(refs[7][False]))


def _ray_keys(rays, lo, hi):
    """Keys of the ray-order pre-pass, in float32: scale = 128 / (hi - lo),
    q = trunc(clamp((o - lo) * scale, 0, 127)) per axis; key = direction
    octant (x sign highest) << 21 | Morton code of q (x bit highest)."""
    lo, hi = np.float32(lo), np.float32(hi)
    scale = np.float32(128.0) / (hi - lo)
    f = np.minimum(np.maximum((rays[:, :3] - lo) * scale, np.float32(0)), np.float32(127))
    q = f.astype(np.uint32).astype(np.int64)
    code = np.zeros(len(rays), np.int64)
    for b in range(7):
        for axis, shift in ((0, 2), (1, 1), (2, 0)):
            code |= ((q[:, axis] >> b) & 1) << (3 * b + shift)
    d = rays[:, 4:7]
    octant = ((d[:, 0] < 0).astype(np.int64) << 2) | ((d[:, 1] < 0).astype(np.int64) << 1) | (d[:, 2] < 0)
    return (octant << 21) | code


def test_morton_order_reuses_its_workspace(cw):
    """Calls of shrinking and growing size on one stream reuse and grow the
    kept workspace; each order is the stable sort of the float32 keys (ties
    keep ray order; directions in every octant), origins outside [lo, hi]
    clamp; destroying the stream
    releases its workspace and later calls on the null stream still work."""
    rng = np.random.default_rng(21)
    lo = np.array([-1.0, -2.0, -3.0], np.float32)
    hi = np.array([2.0, 1.0, 0.5], np.float32)

    def check(n, stream):
        say(f"morton workspace n={n} stream={'own' if stream else 'null'}")
        rays = np.zeros((n, 8), np.float32)
        rays[:, :3] = rng.uniform(lo - 0.1, hi + 0.1, size=(n, 3)).astype(np.float32)
        rays[: n // 4, :3] = np.round(rays[: n // 4, :3], 1)   # repeated keys
        rays[:, 4:7] = rng.normal(size=(n, 3)).astype(np.float32)
        rays[: n // 8, 4:7] = np.abs(rays[: n // 8, 4:7])      # one octant
        rays[:, 7] = np.inf
        dr = cw.DeviceBuffer.from_numpy(rays)
        do = cw.DeviceBuffer(4 * n)
        cw.morton_order(dr.ptr, n, lo, hi, do.ptr, stream)
        if stream is not None:
            stream.synchronize()
        order = do.download((n,), np.int32)
        assert np.array_equal(order, np.argsort(_ray_keys(rays, lo, hi), kind="stable"))

    st = cw.Stream()
    # sizes around the sort's 4096-key tiles and 64-key waves, and one of
    # 4,100 tiles (65 chunks of the column scan)
    for n in (50_000, 1_000, 300_000, 7, 1, 64, 65, 4096, 4097, 8191, (1 << 20) + 5, (1 << 24) + 12_345):
        check(n, st)
    del st
    gc.collect()
    check(20_000, None)
    check(60_000, None)


def test_morton_order_equal_keys_keep_ray_order(cw):
    """Every origin the same: the stable sort returns the identity order
    (ties keep ray order across tiles, waves and lanes)."""
    for n in (5000, 3 * 4096 + 17):
        rays = np.zeros((n, 8), np.float32)
        rays[:, :3] = 0.25
        rays[:, 4], rays[:, 7] = 1.0, np.inf
        dr, do = cw.DeviceBuffer.from_numpy(rays), cw.DeviceBuffer(4 * n)
        cw.morton_order(dr.ptr, n, [0, 0, 0], [1, 1, 1], do.ptr)
        cw.synchronize()
        assert np.array_equal(do.download((n,), np.int32), np.arange(n, dtype=np.int32))


def test_autotune_small(cw, batch_set):
    _, items, refs = batch_set
    bufs, ditems = device_items(cw, items)
    say("autotune")
    best, recs = cw.autotune_batch(ditems, block_sizes=(64, 128), lds_depths=(8, 16), morton=(0, 1), prefetch=(0, 1),
                                   loop_modes=(0, 2), reps=2)
    assert len(recs) == 32 and best.ms == min(r.ms for r in recs) and best.ms > 0
    assert {r.cfg.loop_mode for r in recs} == {0, 2}
    b = cw.Batch(ditems, cw.TRACE_CLOSEST, best.cfg, morton=bool(best.morton))
    b.launch()
    cw.synchronize()
    check_items(cw, items, bufs, refs, False, "autotuned")


def test_autotune_lds_depth_of_the_stack_need(cw, batch_set):
    """An LDS depth of 0 in the tuning space stands for the deepest stack need
    of the items' trees, the smallest depth at which the batch keeps the whole
    stack in LDS; that configuration's hits equal the oracle's."""
    _, items, refs = batch_set
    bufs, ditems = device_items(cw, items)
    need = max(lay.stats()["stack_need"] for _, lay, rays in items if len(rays))
    say(f"autotune at the stack need ({need})")
    best, recs = cw.autotune_batch(ditems, block_sizes=(128,), lds_depths=(0,), morton=(0,), prefetch=(0,),
                                   loop_modes=(1,), reps=1)
    assert len(recs) == 1 and recs[0].cfg.lds_depth == need and best.cfg.lds_depth == need
    b = cw.Batch(ditems, cw.TRACE_CLOSEST, best.cfg)
    b.launch()
    cw.synchronize()
    check_items(cw, items, bufs, refs, False, "LDS depth = stack need")
