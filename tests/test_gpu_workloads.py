"""GPU parity on the benchmark workloads themselves, at their full sizes,
against the CPU oracle:

  C1  all 32,986,368 AO rays in one launch with the default configuration
      (the 7-wave build of the interleaved kernel, which the automatic
      occupancy picks from 2^22 rays up), every ray, closest and any hit;
  C2  the whole 256-scene plan (66 M triangles, 154 M rays) through one
      persistent batch launch, its Morton pre-pass and its graph replay;
  C3  the 20 M-triangle clustered soup with all 64 M rays traced in one launch
      with the scene default (128 B nodes, ray-order pre-pass in the launch,
      interleaved schedule with node prefetch);
  C4  the 32-scene serving set (100 M triangles) through the continuous-
      batching server, every served hit of overlapping requests.

Closest hits must be bit-identical to the oracle's, and so must any hits of
schedules that keep each ray's order of tests.  The speculative schedule's
any hits must agree on occlusion (it may report another valid triangle,
which the oracle's Woop test re-verifies)."""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-workloads] {msg}", flush=True)
    sys.stdout.flush()


def bits(h):
    return np.ascontiguousarray(h, np.float32).view(np.uint32)


def threads():
    import oracle

    return oracle.default_threads()


def assert_any_hit_valid(orc, got, ref, rays, woop, what):
    occ_g, occ_r = orc.hit_ids(got) >= 0, orc.hit_ids(ref) >= 0
    assert np.array_equal(occ_g, occ_r), f"{what}: occlusion differs on {int((occ_g != occ_r).sum())} rays"
    # every reported triangle is a valid hit of its ray (oracle Woop test)
    ids = orc.hit_ids(got)
    for i in np.nonzero(occ_g)[0][:: max(1, int(occ_g.sum()) // 2000)]:
        ok, t, u, v = orc.woop_test(woop[ids[i]:ids[i] + 3], rays[i])
        assert ok and np.float32(t) == got[i, 0], f"{what}: ray {i} reports an invalid hit"


# ------------------------------------------------------------------- C1


def test_c1_full_ao_batch_every_ray_matches_oracle(cw, orc):
    from cwbvh_amd import workloads as W

    say("building C1")
    wl = W.c1(W.gpu_primary_tracer())
    n = len(wl.rays)
    assert n > (1 << 22)
    _, cwb = cw.build_layouts(wl.mesh, bvh2=False)
    sc = cw.Scene(cwb)
    cfg = sc.default_config()
    assert cfg.occupancy == 0 and cfg.loop_mode == 1  # automatic: the 7-wave build at this size
    dr, dh = cw.DeviceBuffer.from_numpy(wl.rays), cw.DeviceBuffer(16 * n)
    for flags in (cw.TRACE_CLOSEST, cw.TRACE_ANY):
        say(f"C1 {n} rays flags={flags}")
        dh.zero()
        sc.trace(dr.ptr, dh.ptr, n, flags, cfg)
        cw.synchronize()
        got = dh.download((n, 4))
        ref = orc.trace_cwbvh(cwb.nodes, cwb.node_stride, cwb.woop, wl.rays, any_hit=bool(flags), threads=threads())
        diff = (bits(got) != bits(ref)).any(axis=1)
        assert not diff.any(), f"C1 flags={flags}: {int(diff.sum())} of {n} rays differ from the oracle"
    assert not cw.stack_overflowed()


# ------------------------------------------------------------------- C2


@pytest.fixture(scope="module")
def c2_full(cw):
    from cwbvh_amd import workloads as W

    plan = W.c2_plan()
    rng = np.random.default_rng(77)
    say(f"building the C2 plan: {len(plan)} scenes")

    def make(entry):
        mesh, rays, _ = W.c2_scene(entry)
        _, lay = cw.build_layouts(mesh, cw.default_build_params(num_threads=2), bvh2=False)
        return mesh.num_triangles, lay, rays

    with ThreadPoolExecutor(max(1, threads() // 2)) as ex:
        built = list(ex.map(make, plan))
    items, keep = [], []
    for nt, lay, rays in built:
        sc = cw.Scene(lay)
        dr = cw.DeviceBuffer.from_numpy(rays)
        dh = cw.DeviceBuffer(16 * len(rays))
        idx = np.sort(rng.choice(len(rays), size=min(512, len(rays)), replace=False))
        items.append((sc, dr.ptr, dh.ptr, len(rays)))
        keep.append((lay, sc, dr, dh, idx, rays[idx]))
    say(f"C2: {sum(b[0] for b in built)} triangles, {sum(it[3] for it in items)} rays")
    del built
    return items, keep


@pytest.mark.parametrize("any_hit", [False, True])
def test_c2_full_plan_batch_graph_replay_matches_oracle(cw, orc, c2_full, any_hit):
    items, keep = c2_full
    flags = cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST
    st = cw.Stream()
    # an autotuner pick on this batch: Morton pre-pass, speculative schedule
    cfg = cw.default_trace_config(cw.KIND_CWBVH, block_size=128, lds_depth=16, loop_mode=2)
    b = cw.Batch(items, flags, cfg, morton=True)
    b.capture(st)
    for k in range(2):
        for _, _, _, dh, _, _ in keep:
            dh.zero()
        say(f"C2 graph replay {k} any={any_hit}")
        b.replay(st)
        st.synchronize()
        for j, (lay, sc, dr, dh, idx, srays) in enumerate(keep):
            got = dh.download((items[j][3], 4))[idx]
            ref = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, srays, any_hit=any_hit, threads=threads())
            if any_hit:
                assert_any_hit_valid(orc, got, ref, srays, lay.woop, f"C2 scene {j}")
            else:
                diff = (bits(got) != bits(ref)).any(axis=1)
                assert not diff.any(), f"C2 scene {j}: {int(diff.sum())} of {len(idx)} sampled rays differ"
    assert not cw.stack_overflowed(st)


def test_c2_full_plan_plain_launches_match_oracle(cw, orc, c2_full):
    """The plain per-scene launches (each scene's default configuration) that
    the batch is measured against."""
    items, keep = c2_full
    say("C2 plain launches")
    for dh in (k[3] for k in keep):
        dh.zero()
    for sc, r, h, n in items:
        sc.trace(r, h, n, 0)
    cw.synchronize()
    for j, (lay, sc, dr, dh, idx, srays) in enumerate(keep):
        got = dh.download((items[j][3], 4))[idx]
        ref = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, srays, threads=threads())
        assert np.array_equal(bits(got), bits(ref)), f"C2 plain launch, scene {j}"


# ------------------------------------------------------------------- C3


@pytest.fixture(scope="module")
def c3_full(cw):
    from cwbvh_amd import workloads as W

    say("building C3: 20 M triangles, 64 M rays")
    mesh = W.c3_mesh()
    diag = mesh.diagonal()
    rays = cw.rays_sphere_from_surface(mesh, 64 * 1024 * 1024, 1e-4 * diag, 0.0, np.inf, seed=W.SEEDS["C3"] + 1)
    rays = cw.shuffle_rays(rays, seed=W.SEEDS["C3"] + 2)
    bvh2, cwb = cw.build_layouts(mesh)
    del mesh
    idx = np.sort(np.random.default_rng(78).choice(len(rays), size=1 << 20, replace=False))
    srays = rays[idx]
    dr = cw.DeviceBuffer.from_numpy(rays)
    n = len(rays)
    del rays
    return bvh2, cwb, dr, n, idx, srays


@pytest.mark.parametrize("kind", ["cwbvh", "bvh2"])
def test_c3_full_launch_scene_default_matches_oracle(cw, orc, c3_full, kind):
    bvh2, cwb, dr, n, idx, srays = c3_full
    lay = cwb if kind == "cwbvh" else bvh2
    sc = cw.Scene(lay)
    cfg = sc.default_config()
    if kind == "cwbvh":
        assert sc.node_stride == 8 and cfg.loop_mode == 1 and cfg.prefetch == 1
    assert cfg.ray_order == 1
    dh = cw.DeviceBuffer(16 * n)
    for flags in (cw.TRACE_CLOSEST, cw.TRACE_ANY):
        say(f"C3 {kind} flags={flags}: {n} rays in one launch")
        dh.zero()
        sc.trace(dr.ptr, dh.ptr, n, flags, cfg)
        cw.synchronize()
        got = dh.download((n, 4))[idx]
        if kind == "cwbvh":
            ref = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, srays, any_hit=bool(flags), threads=threads())
        else:
            ref = orc.trace_bvh2(lay.nodes, lay.woop, srays, any_hit=bool(flags), threads=threads())
        assert 0.2 < (orc.hit_ids(ref) >= 0).mean() < 1.0
        if flags == cw.TRACE_ANY and kind == "cwbvh" and cfg.loop_mode == 2:
            assert_any_hit_valid(orc, got, ref, srays, lay.woop, "C3 any")
        else:
            diff = (bits(got) != bits(ref)).any(axis=1)
            assert not diff.any(), f"C3 {kind} flags={flags}: {int(diff.sum())} of {len(idx)} sampled rays differ"
    assert not cw.stack_overflowed()


# ------------------------------------------------------------------- C4


@pytest.fixture(scope="module")
def c4_full(cw):
    from cwbvh_amd import workloads as W

    specs = W.c4_plan()
    say(f"building the C4 serving set: {len(specs)} scenes")

    def make(entry):
        mesh = W.c4_mesh(entry)
        _, lay = cw.build_layouts(mesh, cw.default_build_params(num_threads=4), bvh2=False)
        diag = mesh.diagonal()
        rays = cw.rays_ao_surface(mesh, 131072, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=entry["seed"] + 1)
        return mesh.num_triangles, lay, rays

    with ThreadPoolExecutor(max(1, threads() // 4)) as ex:
        built = list(ex.map(make, specs))
    scenes, lays, pools = [], [], []
    for nt, lay, rays in built:
        scenes.append(cw.Scene(lay))
        lays.append(lay)
        pool = cw.PinnedArray(rays.shape, np.float32)
        pool.array[:] = rays
        pools.append(pool)
    say(f"C4: {sum(b[0] for b in built)} triangles resident")
    return scenes, lays, pools


@pytest.mark.parametrize("any_hit", [False, True])
def test_c4_served_requests_match_oracle(cw, orc, c4_full, any_hit):
    scenes, lays, pools = c4_full
    rng = np.random.default_rng(79 + int(any_hit))
    flags = cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST
    srv = cw.Server(scenes, cw.default_server_config(memory_budget=2 << 30, max_inflight=3, max_batch_requests=16,
                                                     flags=flags))
    reqs = []
    for _ in range(48):
        s = int(rng.integers(0, len(scenes)))
        n = int(np.exp(rng.uniform(np.log(1000), np.log(1_000_000))))
        off = int(rng.integers(0, len(pools[s].array) - n))
        h = cw.PinnedArray((n, 4), np.float32)
        reqs.append((srv.submit(s, pools[s].ptr + 32 * off, n, h.ptr), s, off, n, h))
    say(f"C4 any={any_hit}: {len(reqs)} requests, {sum(r[3] for r in reqs)} rays submitted")
    for rid, s, off, n, h in reqs:
        srv.wait(rid, 120.0)
        rays = pools[s].array[off:off + n]
        ref = orc.trace_cwbvh(lays[s].nodes, lays[s].node_stride, lays[s].woop, rays, any_hit=any_hit,
                              threads=threads())
        diff = (bits(h.array) != bits(ref)).any(axis=1)
        assert not diff.any(), f"C4 request {rid} (scene {s}, {n} rays): {int(diff.sum())} hits differ"
    st = srv.stats()
    say(f"C4 stats {st}")
    assert st["completed"] == len(reqs) and st["max_batch_requests_seen"] > 1
    srv.close()
