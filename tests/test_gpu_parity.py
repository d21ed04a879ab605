"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Contract: with triangle postponing off (the default), every kernel variant
returns hits bit-identical to the scalar oracle -- t, u, v and triAddr.  With
postponing on, closest hits may differ only on grazing ties (classified by
oracle.compare_hits) and any-hit occlusion must agree on >= 99.999% of rays.
"""
import json
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def bits(h):
    return np.ascontiguousarray(h, np.float32).view(np.uint32)


def assert_bitwise(gpu, ref, what):
    diff = (bits(gpu) != bits(ref)).any(axis=1)
    if diff.any():
        i = np.nonzero(diff)[0][:5]
        raise AssertionError(f"{what}: {int(diff.sum())} of {len(ref)} rays differ; first {i.tolist()}: "
                             f"gpu={gpu[i].tolist()} ref={ref[i].tolist()}")


def oracle_trace(orc, lay, rays, any_hit=False, counters=False):
    if lay.kind == 1:
        return orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, rays, any_hit=any_hit, counters=counters)
    return orc.trace_bvh2(lay.nodes, lay.woop, rays, any_hit=any_hit, counters=counters)


@pytest.fixture(scope="module")
def c0(cw):
    from cwbvh_amd import workloads as W

    say("building C0")
    wl = W.c0()
    bvh2, cwb = cw.build_layouts(wl.mesh)
    return wl, {"bvh2": bvh2, "cwbvh": cwb}


@pytest.fixture(scope="module")
def c0_ref(orc, c0):
    wl, lays = c0
    out = {}
    for k, lay in lays.items():
        for any_hit in (False, True):
            out[(k, any_hit)] = oracle_trace(orc, lay, wl.rays, any_hit)
    return out


@pytest.mark.parametrize("kind", ["cwbvh", "bvh2"])
@pytest.mark.parametrize("any_hit", [False, True])
def test_c0_bit_identical(cw, c0, c0_ref, kind, any_hit):
    wl, lays = c0
    say(f"C0 {kind} any={any_hit} n={len(wl.rays)}")
    sc = cw.Scene(lays[kind])
    g = sc.trace_numpy(wl.rays, cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST)
    assert_bitwise(g, c0_ref[(kind, any_hit)], f"C0 {kind} any={any_hit}")
    assert not cw.stack_overflowed()


CWBVH_VARIANTS = [
    dict(block_size=64, lds_depth=8),
    dict(block_size=256, lds_depth=4, fetch_chunk=256),
    dict(block_size=128, lds_depth=1, fetch_chunk=1),
    dict(block_size=128, lds_depth=32, refill_idle=0, refill_budget=1),
    dict(block_size=64, lds_depth=64, refill_idle=64, refill_budget=1 << 20),
    dict(block_size=128, lds_depth=16, grid_blocks=7),
    dict(work_slices=1),
    dict(work_slices=3, block_size=64, prefetch=1),
    dict(prefetch=1, lds_depth=2, fetch_chunk=7),
    dict(persistent=0),
    dict(persistent=0, block_size=256, lds_depth=8, postpone_pct=0),
    dict(fetch_chunk=0, work_slices=8),
    dict(fetch_chunk=1024, work_slices=5),
    dict(loop_mode=0),
    dict(loop_mode=0, persistent=0, block_size=64, lds_depth=4),
    dict(loop_mode=0, prefetch=1, refill_idle=8, refill_budget=32, fetch_chunk=33),
    dict(loop_mode=1),
    dict(loop_mode=1, occupancy=7),
    dict(loop_mode=1, occupancy=7, block_size=256, lds_depth=16, fetch_chunk=64),
    dict(loop_mode=1, occupancy=1),
    dict(loop_mode=1, persistent=0, block_size=64),
    dict(loop_mode=1, prefetch=1, lds_depth=3, fetch_chunk=5),
    dict(loop_mode=2),
    dict(loop_mode=2, prefetch=1, lds_depth=1, fetch_chunk=1),
    dict(loop_mode=2, block_size=256, refill_idle=0, refill_budget=1, work_slices=8),
    dict(loop_mode=2, persistent=0, block_size=128, lds_depth=64),
    dict(loop_mode=2, postpone_pct=50),
    dict(loop_mode=2, postpone_pct=90, lds_depth=4, block_size=256),
    dict(loop_mode=2, postpone_pct=100, block_size=64, fetch_chunk=3, prefetch=1),
    dict(loop_mode=1, lds_depth=2, persistent=0),
    dict(loop_mode=2, postpone_pct=40, block_size=256, refill_idle=0, refill_budget=1),
    dict(loop_mode=3),
    dict(loop_mode=3, postpone_pct=2),
    dict(loop_mode=3, postpone_pct=100, block_size=256, lds_depth=4),
    dict(loop_mode=3, persistent=0, block_size=64),
    dict(loop_mode=3, prefetch=1, fetch_chunk=7, lds_depth=2),
    dict(loop_mode=3, refill_idle=0, refill_budget=1, work_slices=3),
    dict(loop_mode=3, block_size=64, lds_depth=64, refill_idle=64, refill_budget=1 << 20),
]


def assert_any_hit_valid(orc, lay, rays, g, ref, what):
    """Any-hit returns the first hit its traversal order meets.  A schedule
    that reorders triangle tests may report another triangle, but occlusion
    (hit or miss) is order-independent and every reported hit is real."""
    ids, rids = orc.hit_ids(g), orc.hit_ids(ref)
    assert np.array_equal(ids >= 0, rids >= 0), f"{what}: occlusion differs on {int(((ids >= 0) != (rids >= 0)).sum())} rays"
    same = ids == rids
    assert_bitwise(g[same], ref[same], what)
    for i in np.nonzero(~same)[0][:3000]:
        ok, t, u, v = orc.woop_test(lay.woop[ids[i]:ids[i] + 3], rays[i])
        assert ok and np.float32(t) == g[i, 0] and np.float32(u) == g[i, 2] and np.float32(v) == g[i, 3], (what, i)


QUEUE_OFF = ("loop_mode 3 was taken out: a GPU run that included its first launches faulted the device "
             "together with another untested path, and no cause has been shown")


@pytest.mark.parametrize("variant", [pytest.param(i, marks=pytest.mark.skip(reason=QUEUE_OFF))
                                     if CWBVH_VARIANTS[i].get("loop_mode") == 3 else i
                                     for i in range(len(CWBVH_VARIANTS))])
def test_cwbvh_config_variants_bit_identical(cw, orc, c0, c0_ref, variant):
    """Every schedule is bit-identical to the oracle for closest hit; for
    any-hit too, except speculation (loop_mode 2), whose contract is
    assert_any_hit_valid."""
    wl, lays = c0
    kw = CWBVH_VARIANTS[variant]
    say(f"CWBVH variant {kw}")
    cfg = cw.default_trace_config(cw.KIND_CWBVH, **kw)
    sc = cw.Scene(lays["cwbvh"])
    for any_hit in (False, True):
        g = sc.trace_numpy(wl.rays, int(any_hit), cfg)
        what = f"variant {kw} any={any_hit}"
        if any_hit and kw.get("loop_mode", 0) == 2:
            assert_any_hit_valid(orc, lays["cwbvh"], wl.rays, g, c0_ref[("cwbvh", True)], what)
        else:
            assert_bitwise(g, c0_ref[("cwbvh", any_hit)], what)
        assert not cw.stack_overflowed(), what


@pytest.mark.parametrize("kw", [
    dict(block_size=64, lds_depth=1), dict(block_size=256, lds_depth=8), dict(block_size=128, lds_depth=64),
    # the persistent grid: each ray's visits and tests are those of one ray per thread
    dict(persistent=1, refill_idle=4, refill_budget=16, work_slices=8),
    dict(persistent=2, refill_idle=4, refill_budget=16, work_slices=8, block_size=128, lds_depth=4),
    dict(persistent=1, refill_idle=0, refill_budget=1, work_slices=1, fetch_chunk=64, block_size=64, lds_depth=2),
    dict(persistent=1, refill_idle=32, refill_budget=1000, work_slices=3, fetch_chunk=1000, grid_blocks=7),
    dict(persistent=1, refill_idle=64, refill_budget=1, work_slices=8, fetch_chunk=1, grid_blocks=4096)])
def test_bvh2_config_variants_bit_identical(cw, c0, c0_ref, kw):
    wl, lays = c0
    say(f"BVH2 variant {kw}")
    cfg = cw.default_trace_config(cw.KIND_BVH2, **kw)
    sc = cw.Scene(lays["bvh2"])
    for any_hit in (False, True):
        g = sc.trace_numpy(wl.rays, int(any_hit), cfg)
        assert_bitwise(g, c0_ref[("bvh2", any_hit)], f"bvh2 {kw} any={any_hit}")


@pytest.mark.parametrize("kw", [dict(loop_mode=0, postpone_pct=20), dict(loop_mode=0, postpone_pct=40, prefetch=1, refill_idle=4,
                                                           refill_budget=8)])
def test_cwbvh_postponing_within_contract(cw, orc, c0, c0_ref, kw, tmp_path):
    wl, lays = c0
    say(f"CWBVH triangle postponing {kw}")
    cfg = cw.default_trace_config(cw.KIND_CWBVH, **kw)
    lay = lays["cwbvh"]
    sc = cw.Scene(lay)
    g = sc.trace_numpy(wl.rays, 0, cfg)
    r = c0_ref[("cwbvh", False)]
    cmp = orc.compare_hits(g, r, lay.woop, wl.rays, lay.prim_idx)
    ga = sc.trace_numpy(wl.rays, 1, cfg)
    ca = orc.compare_hits(ga, c0_ref[("cwbvh", True)], lay.woop, wl.rays, lay.prim_idx, any_hit=True)
    # every disagreement is logged with its ray
    log = cmp["log"] + ca["log"]
    (tmp_path / "disagreements.json").write_text(json.dumps(log))
    say(f"{len(log)} disagreements logged ({cmp['ties']} grazing ties, {len(cmp['mismatches'])} closest "
        f"mismatches, {len(ca['mismatches'])} occlusion)")
    for e in log[:20]:
        say(f"  {e}")
    assert len(log) == cmp["ties"] + len(cmp["mismatches"]) + len(ca["mismatches"])
    assert not cmp["mismatches"], cmp["mismatches"][:10]
    assert cmp["ties"] <= 1e-5 * len(r)
    assert ca["exact"] >= (1 - 1e-5) * len(r)
    # every reported any-hit triangle is a real hit of its ray
    ids = cw.hit_ids(ga)
    for i in np.nonzero(ids >= 0)[0][:2000]:
        ok, t, u, v = orc.woop_test(lay.woop[ids[i]:ids[i] + 3], wl.rays[i])
        assert ok and np.float32(t) == ga[i, 0]


def test_stride8_relayout_bit_identical(cw, c0, c0_ref):
    wl, lays = c0
    say("CWBVH 128-byte nodes")
    sc = cw.Scene(lays["cwbvh"].restride(8))
    for any_hit in (False, True):
        assert_bitwise(sc.trace_numpy(wl.rays, int(any_hit)), c0_ref[("cwbvh", any_hit)], f"stride8 any={any_hit}")


def test_stats_counters_match_oracle(cw, orc, c0):
    """CWBVH visits exactly the oracle's nodes and triangles (same order).  The
    BVH2 while-while loop lets a lane that holds a postponed leaf keep
    traversing until every lane of its wave found one, so it visits at least
    the oracle's work (its hits are still identical)."""
    wl, lays = c0
    rays = wl.rays[:200000]
    for k, lay in lays.items():
        say(f"counters {k}")
        _, rc = oracle_trace(orc, lay, rays, counters=True)
        _, gc = cw.Scene(lay).trace_numpy(rays, 0, with_stats=True)
        if k == "cwbvh":
            assert np.array_equal(rc, gc), (k, rc, gc)
        else:
            assert np.all(gc >= rc) and np.all(gc <= 1.2 * rc), (k, rc, gc)


@pytest.mark.parametrize("loop_mode,postpone", [(0, 0), (1, 0), (2, 0), (2, 60)])
def test_stats_counters_loop_modes(cw, orc, c0, loop_mode, postpone):
    """The grouped and interleaved schedules keep each ray's order of node
    visits and triangle tests, so their counters equal the oracle's.  Speculation tests a
    node's triangles later than the oracle does, so the closest hit shrinks
    tmax later and a few more nodes pass the slab test; the work stays close."""
    wl, lays = c0
    rays = wl.rays[:200000]
    lay = lays["cwbvh"]
    say(f"counters loop_mode={loop_mode}")
    _, rc = oracle_trace(orc, lay, rays, counters=True)
    cfg = cw.default_trace_config(cw.KIND_CWBVH, loop_mode=loop_mode, postpone_pct=postpone)
    _, gc = cw.Scene(lay).trace_numpy(rays, 0, cfg, with_stats=True)
    if loop_mode in (0, 1):
        assert np.array_equal(rc, gc), (rc, gc)
    else:
        assert np.all(gc >= 0.9 * rc) and np.all(gc <= 1.5 * rc), (rc, gc)


@pytest.mark.parametrize("kw", [dict(loop_mode=0), dict(loop_mode=1), dict(loop_mode=2, postpone_pct=60)])
def test_loop_profile_counters(cw, c0, kw):
    """The occupancy counters of a STATS launch agree with its work counters:
    one node step per visited node, one triangle use per test, and no part of
    the loop uses more lanes than it ran with."""
    wl, lays = c0
    say(f"loop profile {kw}")
    cw.loop_profile()  # clear
    cfg = cw.default_trace_config(cw.KIND_CWBVH, **kw)
    _, gc = cw.Scene(lays["cwbvh"]).trace_numpy(wl.rays[:300000], 0, cfg, with_stats=True)
    p = cw.loop_profile()
    say(f"  {p}")
    assert p["node_steps"] == gc[0] and p["tri_tests"] == gc[1]
    for part in ("iteration", "node", "tri"):
        assert 0 < p[f"{part}_util"] <= 1.0
    assert p["node_slots"] <= p["iteration_slots"] and p["active_lanes"] <= p["iteration_slots"]
    assert cw.loop_profile()["iteration_slots"] == 0  # cleared on read


@pytest.mark.parametrize("count", [1, 63, 100_000, 600_000])
def test_automatic_persistence_by_ray_count(cw, c0, c0_ref, count):
    """The default (persistent = 2) traces small batches one ray per thread
    and large ones with the persistent grid; both match the oracle."""
    wl, lays = c0
    say(f"default config on {count} rays")
    sc = cw.Scene(lays["cwbvh"])
    assert sc.default_config().persistent == 2
    for any_hit in (False, True):
        g = sc.trace_numpy(wl.rays[:count], int(any_hit))
        assert_bitwise(g, c0_ref[("cwbvh", any_hit)][:count], f"auto persistence n={count} any={any_hit}")


def test_wrapped_scene_and_scene_default_config(cw, c0, c0_ref):
    """A scene wrapped around caller buffers traces like the owning upload,
    and the per-scene default picks the schedule by scene size."""
    wl, lays = c0
    for k in ("cwbvh", "bvh2"):
        lay = lays[k]
        say(f"wrapped scene {k}")
        bufs = [cw.DeviceBuffer.from_numpy(a) for a in (lay.nodes, lay.woop, lay.prim_idx)]
        sc = cw.Scene.wrap(lay.kind, lay.node_stride, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                           lay.stats()["stack_need"], keep=bufs)
        for any_hit in (False, True):
            assert_bitwise(sc.trace_numpy(wl.rays, int(any_hit)), c0_ref[(k, any_hit)], f"wrapped {k} any={any_hit}")
        # bounds read back from the wrapped root equal the upload's
        lo_w, hi_w = sc.bounds()
        lo_o, hi_o = cw.Scene(lay).bounds()
        assert np.array_equal(lo_w, lo_o) and np.array_equal(hi_w, hi_o), (lo_w, lo_o, hi_w, hi_o)
    small = cw.Scene(lays["cwbvh"]).default_config()
    assert small.loop_mode == cw.default_trace_config(cw.KIND_CWBVH).loop_mode == 1
    assert small.ray_order == 0
    # a wrapped CWBVH tree larger than the 256 MB Infinity Cache: no bounds to
    # sort rays in, so rays in the given order with the grouped schedule
    big_nodes = cw.DeviceBuffer(300 << 20)
    big = cw.Scene.wrap(cw.KIND_CWBVH, 5, big_nodes.ptr, big_nodes.ptr, None, 8, keep=(big_nodes,))
    assert big.default_config().loop_mode == 0 and big.default_config().ray_order == 0


def test_scene_node_stride_modes(cw, c0, c0_ref):
    """Uploads with 128 B nodes trace bit-identically to the 80 B layout; the
    automatic mode keeps cache-resident scenes at 80 B."""
    wl, lays = c0
    for mode, want in ((cw.SCENE_STRIDE_AUTO, 5), (cw.SCENE_STRIDE_KEEP, 5), (cw.SCENE_STRIDE_128, 8)):
        say(f"stride mode {mode}")
        sc = cw.Scene(lays["cwbvh"], stride_mode=mode)
        assert sc.node_stride == want
        assert sc.device_bytes() > 16 * want * lays["cwbvh"].stats()["num_nodes"]
        for any_hit in (False, True):
            assert_bitwise(sc.trace_numpy(wl.rays, int(any_hit)), c0_ref[("cwbvh", any_hit)],
                           f"stride mode {mode} any={any_hit}")
    assert cw.Scene(lays["bvh2"], stride_mode=cw.SCENE_STRIDE_128).node_stride == 4
    with pytest.raises(cw.CwbvhError):
        cw.Scene(lays["cwbvh"], stride_mode=3)


def test_large_scene_uploads_128b_nodes(cw):
    """A CWBVH scene larger than the Infinity Cache (here 5 M triangles,
    about 320 MB of nodes and Woop rows) gets 128 B nodes by default, and its
    default configuration sorts launches of 2^20 rays or more by origin
    Morton code first; hits equal those of the 80 B layout traced in the
    given order.  The BVH2 scene of the same mesh sorts by default too and
    traces on the persistent grid; its hits equal one ray per thread's."""
    say("building 5 M-triangle soup")
    mesh = cw.Mesh.cluster_soup(5_000_000, 1024, 0.005, 0.05, float(np.log(1e-3)), 0.5, seed=41)
    bvh2, cwb = cw.build_layouts(mesh)
    st = cwb.stats()
    assert 16 * (st["nodes_f4"] + st["woop_f4"]) > (256 << 20)
    auto = cw.Scene(cwb)
    keep = cw.Scene(cwb, stride_mode=cw.SCENE_STRIDE_KEEP)
    assert auto.node_stride == 8 and keep.node_stride == 5
    d = auto.default_config()
    assert d.loop_mode == 1 and d.prefetch == 1 and d.ray_order == 1
    as_given = keep.default_config()
    as_given.ray_order = 0
    diag = mesh.diagonal()
    rays = cw.rays_sphere_from_surface(mesh, (1 << 20) + 4321, 1e-4 * diag, 0.0, np.inf, seed=42)
    for any_hit in (False, True):
        say(f"large scene any={any_hit}")
        a = auto.trace_numpy(rays, int(any_hit))
        b = keep.trace_numpy(rays, int(any_hit), as_given)
        assert_bitwise(a, b, f"128 B nodes, sorted vs 80 B nodes, given order, any={any_hit}")
    # the reference entry points on the same arrays (80 B nodes): the wrapped
    # scene reads its bounds from the root, so rtTraceCWBVH sorts too
    bufs = [cw.DeviceBuffer.from_numpy(a) for a in (cwb.nodes, cwb.woop, cwb.prim_idx)]
    wrapped = cw.Scene.wrap(cw.KIND_CWBVH, 5, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, st["stack_need"], keep=bufs)
    d = wrapped.default_config()
    assert d.ray_order == 1 and d.loop_mode == 1 and d.prefetch == 1
    n = len(rays)
    d_rays = cw.DeviceBuffer.from_numpy(rays)
    d_hits = cw.DeviceBuffer(16 * n)
    say("large scene through rtBindCWBVHData / rtTraceCWBVH")
    cw.rtBindCWBVHData(*(b.ptr for b in bufs))
    cw.rtTraceCWBVH(d_rays.ptr, d_hits.ptr, n)
    assert_bitwise(d_hits.download((n, 4)), keep.trace_numpy(rays, 0, as_given), "rtTraceCWBVH sorted vs given order")
    # a launch above 2^26 rays is sorted and traced in pieces (here 2^26 and
    # 4,321 rays); the rays repeat the set above, so do the hits
    n_big = (1 << 26) + 4321
    say(f"large scene, {n_big} rays in two sorted pieces")
    reps = -(-n_big // len(rays))
    big = auto.trace_numpy(np.tile(rays, (reps, 1))[:n_big], 0)
    small = keep.trace_numpy(rays, 0, as_given)
    for k in range(reps):
        part = big[k * len(rays):(k + 1) * len(rays)]
        assert_bitwise(part, small[:len(part)], f"piecewise sorted launch, repeat {k}")
    del big
    # fused surfel trace with the scene's default (which has node prefetch,
    # not used by surfel traces): equal to the same schedule without prefetch
    rs = np.random.default_rng(43)
    pos = np.zeros((2048, 4), np.float32)
    pos[:, :3] = mesh.bounds()[0] + rs.random((2048, 3)) * (mesh.bounds()[1] - mesh.bounds()[0])
    nrm = np.zeros((2048, 4), np.float32)
    nv = rs.normal(size=(2048, 3))
    nrm[:, :3] = nv / np.linalg.norm(nv, axis=1, keepdims=True)
    bufs_s = [cw.DeviceBuffer.from_numpy(a) for a in (pos, nrm, np.full(2048, 1e-3, np.float32))]
    hs = [cw.DeviceBuffer(16 * 2048 * 8) for _ in range(2)]
    say("large scene fused surfel trace, default configuration")
    auto.trace_surfels(*(b.ptr for b in bufs_s), 2048, 8, 0.0, np.inf, 7, d_hits=hs[0].ptr)
    auto.trace_surfels(*(b.ptr for b in bufs_s), 2048, 8, 0.0, np.inf, 7, d_hits=hs[1].ptr,
                       cfg=auto.default_config(prefetch=0))
    a, b = (h.download((2048 * 8, 4)) for h in hs)
    assert_bitwise(a, b, "surfel trace, scene default vs explicit")
    assert (cw.hit_ids(a) >= 0).mean() > 0.05
    s2 = cw.Scene(bvh2)
    d2 = s2.default_config()
    assert d2.ray_order == 1 and d2.persistent == 2
    # against one ray per thread in the given order: the reference's launch
    b2_given = s2.default_config(persistent=0)
    b2_given.ray_order = 0
    for any_hit in (False, True):
        say(f"large BVH2 scene any={any_hit}")
        a = s2.trace_numpy(rays, int(any_hit))
        b = s2.trace_numpy(rays, int(any_hit), b2_given)
        assert_bitwise(a, b, f"BVH2 sorted vs given order any={any_hit}")
        if not any_hit:
            # the two kinds agree on every hit distance
            assert np.array_equal(a[:, 0].view(np.uint32), auto.trace_numpy(rays, 0)[:, 0].view(np.uint32))


def test_invalid_loop_configs_rejected(cw, c0):
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    for kw in (dict(loop_mode=3), dict(loop_mode=-1), dict(loop_mode=1, postpone_pct=20), dict(occupancy=6),
               dict(occupancy=-1), dict(ray_order=2), dict(ray_order=-1)):
        cfg = cw.default_trace_config(cw.KIND_CWBVH, **kw)
        with pytest.raises(cw.CwbvhError) as e:
            sc.trace_numpy(wl.rays[:64], 0, cfg)
        assert e.value.status == 1, kw
    # BVH2: the CWBVH-only fields stay 0; a persistent launch needs a refill budget
    s2 = cw.Scene(lays["bvh2"])
    for kw in (dict(loop_mode=1), dict(prefetch=1), dict(postpone_pct=10), dict(occupancy=7), dict(persistent=3),
               dict(persistent=1, refill_budget=0), dict(persistent=2, refill_budget=4, work_slices=9),
               dict(persistent=1, refill_budget=4, refill_idle=65)):
        cfg = cw.default_trace_config(cw.KIND_BVH2, **kw)
        with pytest.raises(cw.CwbvhError) as e:
            s2.trace_numpy(wl.rays[:64], 0, cfg)
        assert e.value.status == 1, kw


def test_misaligned_buffers_rejected(cw, c0):
    """Rays and hits are read and written as 16-byte vectors: buffers that
    are not 16-byte aligned are refused before any launch."""
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    dr = cw.DeviceBuffer.from_numpy(wl.rays[:65])
    dh = cw.DeviceBuffer(16 * 65)
    for rays, hits in ((dr.ptr + 4, dh.ptr), (dr.ptr, dh.ptr + 8)):
        with pytest.raises(cw.CwbvhError) as e:
            sc.trace(rays, hits, 64, 0)
        assert e.value.status == 1
    with pytest.raises(cw.CwbvhError):
        cw.Batch([(sc, dr.ptr + 16 + 4, dh.ptr, 64)])


@pytest.mark.parametrize("k,kw", [("cwbvh", None), ("bvh2", None), ("cwbvh", dict(persistent=0)),
                                  ("cwbvh", dict(loop_mode=2, prefetch=1)), ("bvh2", dict(persistent=0)),
                                  ("bvh2", dict(persistent=1, refill_idle=4, refill_budget=16, work_slices=8))])
def test_order_entries_out_of_range_stay_in_bounds(cw, c0, c0_ref, k, kw):
    """An ordered trace whose order holds entries outside [0, n) traces ray i
    into slot i for those entries: with an otherwise identity order the hits
    equal the plain trace's, and no access leaves the buffers."""
    wl, lays = c0
    n = len(wl.rays)
    order = np.arange(n, dtype=np.int32)
    order[[3, 100, 777, n - 1]] = [-5, n, 1 << 30, -(1 << 31)]
    say(f"out-of-range order entries {k} {kw}")
    sc = cw.Scene(lays[k])
    kind = cw.KIND_CWBVH if k == "cwbvh" else cw.KIND_BVH2
    cfg = cw.default_trace_config(kind, **kw) if kw is not None else None
    g = sc.trace_numpy(wl.rays, 0, cfg, order=order)
    assert_bitwise(g, c0_ref[(k, False)], f"out-of-range order entries {k} {kw}")


def test_sorted_launches_on_two_threads_and_streams(cw, c0, c0_ref):
    """ray_order = 1 on any scene sorts launches of 2^20 rays or more in the
    stream's own workspace: two host threads tracing on two streams at once
    (different sizes, several rounds) each get the hits of the given order."""
    import concurrent.futures as cf
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    base = len(wl.rays)
    ref = c0_ref[("cwbvh", False)]
    cfg = sc.default_config(ray_order=1)
    sizes = ((1 << 20) + 77, (1 << 20) + 3 * base // 2)
    jobs = []
    for n in sizes:
        reps = -(-n // base)
        rays = np.tile(wl.rays, (reps, 1))[:n]
        jobs.append((n, cw.DeviceBuffer.from_numpy(rays), cw.DeviceBuffer(16 * n), cw.Stream(),
                     np.tile(ref, (reps, 1))[:n]))

    def run(job):
        n, dr, dh, st, want = job
        for _ in range(3):
            sc.trace(dr.ptr, dh.ptr, n, 0, cfg, None, st)
        st.synchronize()
        return dh.download((n, 4)), want

    say("sorted launches on two threads")
    with cf.ThreadPoolExecutor(2) as ex:
        for got, want in ex.map(run, jobs):
            assert_bitwise(got, want, "sorted launch on its own thread and stream")


def test_sorted_launch_prim_ids_and_counts(cw, c0):
    """With the Morton pre-pass, PRIM_ID output equals the unsorted launch's
    and the STATS totals (sums over rays of visited nodes and tested
    triangles) are the same numbers: sorting only reorders the rays."""
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    n = (1 << 20) + 5
    rays = np.tile(wl.rays, (-(-n // len(wl.rays)), 1))[:n]
    dr = cw.DeviceBuffer.from_numpy(rays)
    out = {}
    for order in (0, 1):
        say(f"PRIM_ID | STATS with ray_order={order}")
        dh, dc = cw.DeviceBuffer(16 * n), cw.DeviceBuffer(16)
        dc.zero()
        sc.trace(dr.ptr, dh.ptr, n, cw.TRACE_PRIM_ID | cw.TRACE_STATS, sc.default_config(ray_order=order), dc.ptr)
        cw.synchronize()
        out[order] = (dh.download((n, 4)), dc.download((2,), np.uint64))
    assert_bitwise(out[1][0], out[0][0], "PRIM_ID, sorted vs given order")
    assert np.array_equal(out[1][1], out[0][1]) and out[0][1][0] > n


def test_prim_id_output(cw, c0, c0_ref):
    wl, lays = c0
    for k, lay in lays.items():
        say(f"PRIM_ID {k}")
        g = cw.Scene(lay).trace_numpy(wl.rays, cw.TRACE_PRIM_ID)
        ids = cw.hit_ids(c0_ref[(k, False)])
        want = np.where(ids >= 0, lay.prim_idx[np.clip(ids, 0, None)], -1)
        assert np.array_equal(cw.hit_ids(g), want)
        assert np.array_equal(bits(g)[:, [0, 2, 3]], bits(c0_ref[(k, False)])[:, [0, 2, 3]])


def test_ordered_trace(cw, c0, c0_ref):
    wl, lays = c0
    perm = np.random.default_rng(3).permutation(len(wl.rays)).astype(np.int32)
    for k, lay in lays.items():
        say(f"ordered {k}")
        g = cw.Scene(lay).trace_numpy(wl.rays, 0, order=perm)
        assert_bitwise(g, c0_ref[(k, False)], f"ordered {k}")


def test_known_answer_planes(cw):
    n_tri, n_ray = 5000, 100000
    mesh = cw.Mesh.validation_planes(n_tri)
    rays = cw.rays_validation(n_ray, n_tri)
    k = (np.arange(n_ray, dtype=np.uint64) * np.uint64(12345678) % np.uint64(n_tri)).astype(np.int64)
    bvh2, cwb = cw.build_layouts(mesh)
    for lay in (cwb, bvh2):
        say(f"KAT kind={lay.kind}")
        g = cw.Scene(lay).trace_numpy(rays, cw.TRACE_PRIM_ID)
        assert np.array_equal(cw.hit_ids(g), k)


def test_edge_cases(cw, orc):
    rng = np.random.default_rng(11)
    mesh = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=2)
    bvh2, cwb = cw.build_layouts(mesh)
    rays = np.zeros((1000, 8), np.float32)
    rays[:, :3] = rng.uniform(-0.2, 1.2, size=(1000, 3))
    d = rng.normal(size=(1000, 3))
    d[:100] = [0, 0, -1]            # axis-aligned
    d[100:200] = [-0.0, 0.0, 1.0]   # signed zeros
    d[200:300] = [1, 0, 0]          # grazing the heightfield
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf             # tmax = inf
    rays[300:400, 7] = 1e-6         # tiny tmax
    rays[400:500, 0] += 100         # misses everything
    for lay in (cwb, bvh2):
        sc = cw.Scene(lay)
        for n in (1, 63, 65, 1000):
            say(f"edge kind={lay.kind} n={n}")
            for any_hit in (False, True):
                g = sc.trace_numpy(rays[:n], int(any_hit))
                assert_bitwise(g, oracle_trace(orc, lay, rays[:n], any_hit), f"edge n={n}")
        sc.trace(0, 0, 0)  # n == 0 is a no-op
        cw.synchronize()


def _pathological_rays():
    """Rays from above a heightfield, 64 per kind: NaN origin, NaN direction,
    zero direction, tmin > tmax, negative tmin, tmin = -0, huge origins,
    NaN tmax, NaN tmin, a direction of length 1e-30; then direction z
    components of -1e38, exactly -2^125 and the next float below it (the ray
    setup's reciprocal takes the IEEE division beyond 2^125), and -infinity."""
    rng = np.random.default_rng(12)
    r = np.zeros((768, 8), np.float32)
    r[:, :3] = rng.uniform(-0.2, 1.2, size=(768, 3))
    r[:, 2] = 1.0
    d = rng.normal(size=(768, 3))
    d[:, 2] = -np.abs(d[:, 2]) - 0.5
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 7] = np.inf
    nan = np.float32(np.nan)
    r[0:64, 0] = nan
    r[64:128, 5] = nan
    r[128:192, 4:7] = 0.0
    r[192:256, 3], r[192:256, 7] = 5.0, 1.0
    r[256:320, 3] = -1.0
    r[320:384, 3] = -0.0
    r[384:448, :3] *= 1e30
    r[448:512, 7] = nan
    r[512:576, 3] = nan
    r[576:640, 4:7] = [0.0, 0.0, -1e-30]
    two125 = np.float32(2.0 ** 125)
    r[640:672, 6] = -1e38
    r[672:688, 6] = -two125
    r[688:704, 6] = -np.nextafter(two125, np.float32(np.inf))
    r[704:768, 6] = -np.inf
    return r


@pytest.mark.parametrize("kind,kw", [("both", None), ("cwbvh", dict(loop_mode=0)), ("cwbvh", dict(loop_mode=2)),
                                     ("cwbvh", dict(persistent=0)), ("cwbvh", dict(persistent=1, refill_budget=1)),
                                     ("cwbvh", dict(persistent=1, occupancy=7)),
                                     ("bvh2", dict(persistent=1, refill_idle=2, refill_budget=8, work_slices=8))])
def test_pathological_rays(cw, orc, kind, kw):
    """NaN, zero-length, inverted-interval and huge-direction rays terminate
    and return what the oracle returns, bit for bit (NaN rays miss)."""
    mesh = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=2)
    bvh2, cwb = cw.build_layouts(mesh)
    rays = _pathological_rays()
    for lay in (cwb, bvh2):
        if kind != "both" and lay.kind != (cw.KIND_CWBVH if kind == "cwbvh" else cw.KIND_BVH2):
            continue
        sc = cw.Scene(lay)
        cfg = cw.default_trace_config(lay.kind, **kw) if kw is not None else None
        for any_hit in (False, True):
            say(f"pathological rays kind={lay.kind} {kw} any={any_hit}")
            g = sc.trace_numpy(rays, int(any_hit), cfg)
            ref = oracle_trace(orc, lay, rays, any_hit)
            assert_bitwise(g, ref, f"pathological kind={lay.kind} any={any_hit}")
            assert np.all(orc.hit_ids(ref)[:128] == -1)


def test_refined_reciprocal_equals_division(cw):
    """The ray setup takes 1/d as v_rcp_f32 plus one Newton step when every
    clamped component lies in [2^-125, 2^125]; on this device that equals the
    IEEE quotient for every such float (all 2^32 patterns checked)."""
    say("reciprocal check over all float patterns")
    c = cw.rcp_check()
    assert c["in_range"] == 4_194_304_002
    assert c["in_range_mismatches"] == 0
    assert c["outside_mismatches"] > 0  # outside the range the two do differ, and the check sees it


def test_single_triangle_and_empty_scene(cw, orc):
    one = cw.Mesh.from_arrays([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])
    empty = cw.Mesh.from_arrays(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    rays = np.array([[0.2, 0.2, 1, 0, 0, 0, -1, 10], [2, 2, 1, 0, 0, 0, -1, 10]], np.float32)
    for mesh in (one, empty):
        for lay in cw.build_layouts(mesh):
            say(f"tiny scene kind={lay.kind}")
            g = cw.Scene(lay).trace_numpy(rays)
            assert_bitwise(g, oracle_trace(orc, lay, rays), "tiny scene")


@pytest.fixture(scope="module")
def c1_small(cw, orc):
    from cwbvh_amd import workloads as W

    say("building C1 (980k triangles)")
    mesh = W.c1_mesh()
    bvh2, cwb = cw.build_layouts(mesh)

    def primary(m, rays):
        h = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, rays)
        ids = orc.hit_ids(h)
        return h, np.where(ids >= 0, cwb.prim_idx[np.clip(ids, 0, None)], -1).astype(np.int32)

    wl = W.c1(primary, width=320, height=180, mesh=mesh)
    return wl, {"bvh2": bvh2, "cwbvh": cwb}


@pytest.mark.parametrize("kind", ["cwbvh", "bvh2"])
def test_c1_sample_bit_identical(cw, orc, c1_small, kind):
    wl, lays = c1_small
    lay = lays[kind]
    sc = cw.Scene(lay)
    for name, rays in (("primary", wl.extra["primary"]), ("ao", wl.rays)):
        for any_hit in (False, True):
            say(f"C1 {kind} {name} any={any_hit} n={len(rays)}")
            g = sc.trace_numpy(rays, int(any_hit))
            assert_bitwise(g, oracle_trace(orc, lay, rays, any_hit), f"C1 {kind} {name}")


def test_compat_entry_points(cw, orc, c0, c0_ref):
    wl, lays = c0
    rays = wl.rays
    n = len(rays)
    d_rays = cw.DeviceBuffer.from_numpy(rays)
    d_hits = cw.DeviceBuffer(16 * n)
    bufs = {}
    for k, lay in lays.items():
        bufs[k] = [cw.DeviceBuffer.from_numpy(lay.nodes), cw.DeviceBuffer.from_numpy(lay.woop),
                   cw.DeviceBuffer.from_numpy(lay.prim_idx)]
    say("rtBindCWBVHData / rtTraceCWBVH")
    cw.rtBindCWBVHData(*(b.ptr for b in bufs["cwbvh"]))
    ms = cw.rtTraceCWBVH(d_rays.ptr, d_hits.ptr, n)
    assert ms > 0
    assert_bitwise(d_hits.download((n, 4)), c0_ref[("cwbvh", False)], "rtTraceCWBVH")
    say("rtBindBVH2Data / rtTraceBVH2")
    cw.rtBindBVH2Data(*(b.ptr for b in bufs["bvh2"]))
    cw.rtTraceBVH2(d_rays.ptr, d_hits.ptr, n)
    assert_bitwise(d_hits.download((n, 4)), c0_ref[("bvh2", False)], "rtTraceBVH2")
    # rebinding replaces the previous binding
    cw.rtBindCWBVHData(*(b.ptr for b in bufs["cwbvh"]))
    cw.rtTraceCWBVH(d_rays.ptr, d_hits.ptr, 1000)
    assert_bitwise(d_hits.download((1000, 4)), c0_ref[("cwbvh", False)][:1000], "rebind")


def deep_chain_cwbvh(cw, depth):
    """CWBVH whose every node has two inner children covering the same box, the
    chain child visited first: the sibling is pushed at every level, so a ray
    needs `depth` stack entries."""
    def node(p, e, imask, base, tbase, metas, q_lo, q_hi):
        n = np.zeros((5, 4), np.float32)
        n[0, :3] = p
        b = n[0].view(np.uint8)
        b[12:15] = np.array(e, np.int8).view(np.uint8)
        b[15] = imask
        n[1].view(np.int32)[0] = base
        n[1].view(np.int32)[1] = tbase
        n[1].view(np.uint8)[8:16] = metas
        q = n[2:5].view(np.uint8).reshape(-1)
        for s in range(8):
            if metas[s]:
                q[0 * 8 + s] = q[1 * 8 + s] = q[2 * 8 + s] = q_lo
                q[24 + 0 * 8 + s] = q[24 + 8 + s] = q[24 + 16 + s] = q_hi
        return n

    nodes = {}
    e = [-7, -7, -7]  # 255 * 2^-7 covers [0, 1.99]
    inner0 = (1 << 5) | 24
    inner1 = (1 << 5) | 25
    leaf1 = (1 << 5) | 0
    for i in range(depth):
        idx = 0 if i == 0 else 2 * i - 1
        if i < depth - 1:
            nodes[idx] = node((0, 0, 0), e, 0b11, 2 * i + 1, 0, [inner0, inner1, 0, 0, 0, 0, 0, 0], 0, 255)
            nodes[2 * i + 2] = node((0, 0, 0), e, 0, 0, i + 1, [leaf1, 0, 0, 0, 0, 0, 0, 0], 0, 255)
        else:
            nodes[idx] = node((0, 0, 0), e, 0, 0, 0, [leaf1, 0, 0, 0, 0, 0, 0, 0], 0, 255)
    arr = np.concatenate([nodes[i] for i in range(max(nodes) + 1)])
    far = cw.woopify([1000, 0, 0], [1001, 0, 0], [1000, 1, 0])
    woop = np.concatenate([far] * (depth + 1)).astype(np.float32)
    prim = np.tile(np.array([0, -1, -1], np.int32), depth + 1)
    return arr, woop, prim


@pytest.fixture(scope="module")
def split_scene(cw):
    """Long thin triangles crossing the unit cube among small ones, built
    with spatial splits (triangles referenced from several leaves)."""
    rng = np.random.default_rng(31)
    a = rng.uniform(0, 1, size=(2000, 3))
    b = rng.uniform(0, 1, size=(2000, 3))
    c = a + rng.normal(scale=0.01, size=(2000, 3))
    small = rng.uniform(0, 1, size=(40000, 1, 3)) + rng.normal(scale=0.01, size=(40000, 3, 3))
    v = np.concatenate([np.stack([a, b, c], axis=1), small]).astype(np.float32).reshape(-1, 3)
    mesh = cw.Mesh.from_arrays(v, np.arange(len(v), dtype=np.int32).reshape(-1, 3))
    say("building split scene")
    bvh2, cwb = cw.build_layouts(mesh, cw.default_build_params(split_alpha=1e-5, split_budget=0.5))
    assert cwb.stats()["num_tri_refs"] > mesh.num_triangles
    rays = np.zeros((200_000, 8), np.float32)
    rays[:, :3] = rng.uniform(-0.2, 1.2, size=(len(rays), 3))
    d = rng.normal(size=(len(rays), 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf
    return {"bvh2": bvh2, "cwbvh": cwb}, rays


@pytest.mark.parametrize("kind,kw", [("bvh2", None), ("cwbvh", dict(loop_mode=0)), ("cwbvh", dict(loop_mode=1)),
                                     ("cwbvh", dict(loop_mode=2)), ("cwbvh", dict(loop_mode=2, postpone_pct=40)),
                                     ("cwbvh", dict(persistent=0, block_size=64, lds_depth=2))])
def test_spatial_split_layouts(cw, orc, split_scene, kind, kw):
    """A split triangle sits at several addresses: which one a closest hit
    reports follows the traversal order, so the contract is equal t, u, v
    bits and the same triangle; occlusion agrees exactly."""
    lays, rays = split_scene
    lay = lays[kind]
    sc = cw.Scene(lay)
    cfg = cw.default_trace_config(cw.KIND_CWBVH, **kw) if kw is not None else None
    for any_hit in (False, True):
        say(f"split scene {kind} {kw} any={any_hit} n={len(rays)}")
        ref = oracle_trace(orc, lay, rays, any_hit)
        g = sc.trace_numpy(rays, cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST, cfg)
        ig, ir = orc.hit_ids(g), orc.hit_ids(ref)
        assert np.array_equal(ig >= 0, ir >= 0)
        if any_hit:
            continue
        assert np.array_equal(bits(g)[:, [0, 2, 3]], bits(ref)[:, [0, 2, 3]])
        hit = ig >= 0
        assert np.array_equal(lay.prim_idx[ig[hit]], lay.prim_idx[ir[hit]])
        assert 0.05 < hit.mean() < 0.95
    assert not cw.stack_overflowed()


def test_failed_allocation_does_not_poison_later_calls(cw, c0, c0_ref):
    """A device allocation that cannot succeed (1 PiB) fails with
    CWBVH_ERR_OUT_OF_MEMORY; the error is not reported again by the next
    trace, which returns the oracle's hits."""
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    with pytest.raises(cw.CwbvhError) as e:
        cw.DeviceBuffer(1 << 50)
    assert e.value.status == 3
    assert_bitwise(sc.trace_numpy(wl.rays, 0), c0_ref[("cwbvh", False)], "trace after a failed allocation")


def test_sort_workspace_limit_falls_back_to_given_order(cw, c0, c0_ref):
    """A ray_order = 1 launch whose sort workspace exceeds the workspace limit
    traces its rays in the given order (same hits), leaves no error for the
    next call, and a direct cwbvh_morton_order call fails with
    CWBVH_ERR_OUT_OF_MEMORY under the limit."""
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    n = (1 << 20) + 11
    reps = -(-n // len(wl.rays))
    rays = np.tile(wl.rays, (reps, 1))[:n]
    want = np.tile(c0_ref[("cwbvh", False)], (reps, 1))[:n]
    dr, dh, do = cw.DeviceBuffer.from_numpy(rays), cw.DeviceBuffer(16 * n), cw.DeviceBuffer(4 * n)
    st = cw.Stream()
    cfg = sc.default_config(ray_order=1)
    try:
        cw.set_workspace_limit(1 << 20)   # far below the ~16 B per ray the sort needs
        say("sorted launch under a 1 MiB workspace limit")
        sc.trace(dr.ptr, dh.ptr, n, 0, cfg, None, st)
        st.synchronize()
        assert cw.lib.cwbvh_last_error().decode() == ""
        assert_bitwise(dh.download((n, 4)), want, "sorted launch without room for the sort")
        lo, hi = sc.bounds()
        with pytest.raises(cw.CwbvhError) as e:
            cw.morton_order(dr.ptr, n, lo, hi, do.ptr, st)
        assert e.value.status == 3
    finally:
        cw.set_workspace_limit(0)
    # the next call is not affected and sorts again
    dh.zero()
    sc.trace(dr.ptr, dh.ptr, n, 0, cfg, None, st)
    st.synchronize()
    assert_bitwise(dh.download((n, 4)), want, "sorted launch after the limit is lifted")


def test_per_thread_stream_handle_refused(cw, c0):
    """The per-thread stream handle names a different stream on every host
    thread; stream-keyed counters and workspaces cannot follow it, so traces
    on it are refused instead of sharing state between threads."""
    wl, lays = c0
    sc = cw.Scene(lays["cwbvh"])
    dr, dh = cw.DeviceBuffer.from_numpy(wl.rays[:4096]), cw.DeviceBuffer(16 * 4096)
    per_thread = 2  # hipStreamPerThread
    rc = cw.lib.cwbvh_trace(sc.handle, dr.ptr, dh.ptr, 4096, 0, None, None, per_thread)
    assert rc == 1 and "hipStreamPerThread" in cw.lib.cwbvh_last_error().decode()


def test_non_finite_vertices(cw, orc):
    """A mesh with NaN, +inf and -inf corners: the triangles they touch stay
    out of the tree, and both kernels return the oracle's hits bit for bit,
    closest and any hit."""
    rng = np.random.default_rng(31)
    v = rng.random((300, 3)).astype(np.float32)
    idx = rng.integers(0, 300, (400, 3)).astype(np.int32)
    v[5, 1], v[7, 0], v[9, 2] = np.nan, np.inf, -np.inf
    bvh2, cwb = cw.build_layouts(cw.Mesh.from_arrays(v, idx))
    rays = np.zeros((4096, 8), np.float32)
    rays[:, :3] = rng.uniform(-0.2, 1.2, size=(4096, 3))
    d = rng.normal(size=(4096, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf
    for lay in (cwb, bvh2):
        sc = cw.Scene(lay)
        for any_hit in (False, True):
            say(f"non-finite vertices kind={lay.kind} any={any_hit}")
            g = sc.trace_numpy(rays, int(any_hit))
            ref = oracle_trace(orc, lay, rays, any_hit)
            assert 0.1 < (orc.hit_ids(ref) >= 0).mean() < 0.9
            assert_bitwise(g, ref, f"non-finite vertices kind={lay.kind} any={any_hit}")


def test_stack_overflow_is_flagged_not_faulted(cw, orc):
    ray = np.array([[-1, -1, -1, 0, *(np.ones(3) / np.sqrt(3)), 1e30]], np.float32)
    # within capacity: traversal equals the oracle
    nodes, woop, prim = deep_chain_cwbvh(cw, 40)
    lay = cw.Layout.from_arrays(cw.KIND_CWBVH, 5, nodes, woop, prim)
    assert lay.stats()["stack_need"] == 41
    say("deep chain within capacity")
    g, cnt = cw.Scene(lay).trace_numpy(np.repeat(ray, 64, axis=0), 0, with_stats=True)
    r, rc = orc.trace_cwbvh(nodes, 5, woop, np.repeat(ray, 64, axis=0), counters=True)
    assert_bitwise(g, r, "deep chain")
    assert np.array_equal(cnt, rc) and rc[0] == 64 * (2 * 40 - 1)
    assert not cw.stack_overflowed()
    # deeper than the kernels' 64 entries: scene creation refuses it ...
    nodes, woop, prim = deep_chain_cwbvh(cw, 100)
    lay = cw.Layout.from_arrays(cw.KIND_CWBVH, 5, nodes, woop, prim)
    with pytest.raises(cw.CwbvhError) as e:
        cw.Scene(lay)
    assert e.value.status == 4
    # ... and the unchecked bind path reports the overflow instead of faulting
    say("deep chain beyond capacity through rtBindCWBVHData")
    dn, dw, dp = (cw.DeviceBuffer.from_numpy(a) for a in (nodes, woop, prim))
    dr = cw.DeviceBuffer.from_numpy(np.repeat(ray, 64, axis=0))
    dh = cw.DeviceBuffer(16 * 64)
    cw.rtBindCWBVHData(dn.ptr, dw.ptr, dp.ptr)
    with pytest.raises(cw.CwbvhError) as e:
        cw.rtTraceCWBVH(dr.ptr, dh.ptr, 64)
    assert e.value.status == 4
    assert not cw.stack_overflowed()  # the flag was consumed by the failing call
