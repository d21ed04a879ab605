"""GPU tests of the continuous-batching server: every request's hits equal the
CPU oracle's and the same rays traced alone (bit for bit) while requests
overlap on the device, HBM-budget admission, serial mode and the Poisson load
generator."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-serve] {msg}", flush=True)
    sys.stdout.flush()


def bits(h):
    return np.ascontiguousarray(h, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def serving_set(cw):
    from cwbvh_amd import workloads as W

    scenes, pools, lays = [], [], []
    for i in range(4):
        m = W.heightfield_of(20000 + 15000 * i, seed=300 + i) if i % 2 == 0 else W.icospheres_of(30000, seed=300 + i)
        _, lay = cw.build_layouts(m, bvh2=False)
        sc = cw.Scene(lay)
        diag = m.diagonal()
        rays = cw.rays_ao_surface(m, 40000, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=400 + i)
        pool = cw.PinnedArray(rays.shape, np.float32)
        pool.array[:] = rays
        scenes.append(sc)
        pools.append(pool)
        lays.append(lay)
    return scenes, pools, lays


def alone(cw, sc, rays):
    return sc.trace_numpy(rays, cw.TRACE_CLOSEST)


def oracle_hits(lay, rays):
    import oracle

    return oracle.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, rays)


# lds_depth 0 (the default): each launch's LDS depth is its trees' deepest
# stack need (the LDS-only kernels); 3: the LDS + scratch stack
@pytest.mark.parametrize("serial,lds_depth", [(0, 0), (1, 0), (0, 3), (1, 3)])
def test_requests_same_alone_and_in_mix(cw, serving_set, serial, lds_depth):
    scenes, pools, lays = serving_set
    rng = np.random.default_rng(5)
    cfg = cw.default_server_config(serial=serial, max_inflight=3, max_batch_requests=8,
                                   memory_budget=64 << 20)
    assert cfg.trace.lds_depth == 0
    cfg.trace.lds_depth = lds_depth
    srv = cw.Server(scenes, cfg)
    reqs = []
    for k in range(40):
        s = int(rng.integers(0, len(scenes)))
        n = int(np.exp(rng.uniform(np.log(100), np.log(60000))))
        off = int(rng.integers(0, len(pools[s].array) - n))
        hits = cw.PinnedArray((n, 4), np.float32)
        rid = srv.submit(s, pools[s].ptr + 32 * off, n, hits.ptr)
        reqs.append((rid, s, off, n, hits))
    say(f"submitted {len(reqs)} requests serial={serial}")
    for rid, s, off, n, hits in reqs:
        lat = srv.wait(rid, 60.0)
        assert lat >= 0
        rays = pools[s].array[off:off + n]
        ref = alone(cw, scenes[s], rays)
        assert np.array_equal(bits(hits.array), bits(ref)), f"request {rid} differs from its solo trace"
        assert np.array_equal(bits(hits.array), bits(oracle_hits(lays[s], rays))), f"request {rid} vs oracle"
    st = srv.stats()
    say(f"stats {st}")
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= cfg.memory_budget
    if serial:
        assert st["max_batch_requests_seen"] == 1 and st["max_inflight_seen"] == 1
    else:
        assert st["max_batch_requests_seen"] > 1  # requests were batched together
    srv.close()


def test_budget_admission(cw, serving_set):
    scenes, pools, lays = serving_set
    # budget for ~2.5 requests of 50k rays: admission must wait for completions
    budget = int(2.5 * 50000 * 48)
    srv = cw.Server(scenes, cw.default_server_config(memory_budget=budget, max_inflight=4))
    big = cw.PinnedArray((200000, 4), np.float32)
    with pytest.raises(cw.CwbvhError) as e:
        srv.submit(0, pools[0].ptr, 200000, big.ptr)  # 9.6 MB > 6 MB budget
    assert e.value.status == 7
    outs = []
    for k in range(12):
        h = cw.PinnedArray((50000, 4), np.float32)
        outs.append((srv.submit(k % 4, pools[k % 4].ptr, 50000, h.ptr), k % 4, h))
    srv.drain()
    for rid, s, h in outs:
        srv.wait(rid)
        assert np.array_equal(bits(h.array), bits(alone(cw, scenes[s], pools[s].array[:50000])))
        assert np.array_equal(bits(h.array), bits(oracle_hits(lays[s], pools[s].array[:50000])))
    st = srv.stats()
    assert st["peak_bytes"] <= budget and st["completed"] == 12
    srv.close()


def test_zero_ray_request(cw, serving_set):
    scenes, pools, lays = serving_set
    srv = cw.Server(scenes)
    rid = srv.submit(1, 0, 0, 0)
    assert srv.wait(rid, 5.0) == 0.0
    srv.close()


def test_poisson_harness(cw, serving_set):
    scenes, pools, lays = serving_set
    srv = cw.Server(scenes, cw.default_server_config(memory_budget=256 << 20))
    hit_pool = cw.PinnedArray((400000, 4), np.float32)
    say("poisson 3000 req/s x 60")
    res, lat = srv.run_poisson(3000.0, 60, 1000, 100000, pools, hit_pool, seed=9)
    say(f"poisson {res}")
    assert res["rays"] > 0 and 0 < res["p50_ms"] <= res["p99_ms"] <= res["max_ms"]
    assert np.all(lat > 0) and len(lat) == 60
    assert srv.stats()["completed"] == 60
    srv.close()


@pytest.fixture(scope="module")
def surfel_pools(cw, serving_set):
    """Per scene: 20,000 surfels on the surface (pinned positions, normals,
    radii), the window the surfel requests below take from."""
    scenes, pools, lays = serving_set
    rng = np.random.default_rng(71)
    out = []
    for k, pool in enumerate(pools):
        pts = pool.array[::8][:20000]
        pos = cw.PinnedArray((len(pts), 4), np.float32)
        pos.array[:] = 0
        pos.array[:, :3] = pts[:, :3]
        nrm = cw.PinnedArray((len(pts), 4), np.float32)
        nrm.array[:] = 0
        nrm.array[:, :3] = pts[:, 4:7]
        rad = cw.PinnedArray((len(pts),), np.float32)
        rad.array[:] = rng.uniform(1e-4, 1e-3, size=len(pts)).astype(np.float32)
        out.append((pos, nrm, rad))
    return out


def test_device_and_host_requests_in_one_mix_match_oracle(cw, serving_set):
    """Device-resident requests (no copies) and host requests share batches;
    every request's hits equal the oracle's."""
    scenes, pools, lays = serving_set
    rng = np.random.default_rng(6)
    dpools = [cw.DeviceBuffer.from_numpy(p.array) for p in pools]
    srv = cw.Server(scenes, cw.default_server_config(max_inflight=3, max_batch_requests=8, memory_budget=64 << 20))
    reqs = []
    for k in range(40):
        s = int(rng.integers(0, len(scenes)))
        n = int(np.exp(rng.uniform(np.log(100), np.log(60000))))
        off = int(rng.integers(0, len(pools[s].array) - n))
        if k % 2:
            dh = cw.DeviceBuffer(16 * n)
            rid = srv.submit_device(s, dpools[s].ptr + 32 * off, n, dh.ptr)
            reqs.append((rid, s, off, n, dh, True))
        else:
            h = cw.PinnedArray((n, 4), np.float32)
            reqs.append((srv.submit(s, pools[s].ptr + 32 * off, n, h.ptr), s, off, n, h, False))
    say(f"submitted {len(reqs)} host and device requests")
    for rid, s, off, n, buf, dev in reqs:
        srv.wait(rid, 60.0)
        got = buf.download((n, 4)) if dev else buf.array
        ref = oracle_hits(lays[s], pools[s].array[off:off + n])
        assert np.array_equal(bits(got), bits(ref)), f"request {rid} (device={dev}) differs from the oracle"
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["max_batch_requests_seen"] > 1
    srv.close()


@pytest.mark.parametrize("any_hit", [False, True])
def test_surfel_requests_match_oracle(cw, serving_set, surfel_pools, any_hit):
    """Surfel requests return, per surfel, the number of its generated AO
    rays that hit: the oracle's count on the host generator's rays."""
    import oracle

    scenes, pools, lays = serving_set
    rng = np.random.default_rng(8)
    flags = cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST
    srv = cw.Server(scenes, cw.default_server_config(max_inflight=3, memory_budget=64 << 20, flags=flags))
    reqs = []
    for k in range(12):
        s = int(rng.integers(0, len(scenes)))
        pos, nrm, rad = surfel_pools[s]
        t = int(rng.integers(1, 4000))
        off = int(rng.integers(0, len(pos.array) - t))
        rpt = int(rng.integers(1, 20))
        diag = 1.0
        counts = cw.PinnedArray((t,), np.uint32)
        rid = srv.submit_surfels(s, pos.ptr + 16 * off, nrm.ptr + 16 * off, rad.ptr + 4 * off, t, rpt, 0.0,
                                 0.05 * diag, 100 + k, counts.ptr)
        reqs.append((rid, s, off, t, rpt, 100 + k, counts))
    say(f"submitted {len(reqs)} surfel requests any={any_hit}")
    for rid, s, off, t, rpt, seed, counts in reqs:
        srv.wait(rid, 60.0)
        pos, nrm, rad = surfel_pools[s]
        rays = cw.rays_ao_from_samples(pos.array[off:off + t], nrm.array[off:off + t], rad.array[off:off + t], rpt,
                                       0.0, np.float32(0.05), seed)
        ref = oracle.trace_cwbvh(lays[s].nodes, lays[s].node_stride, lays[s].woop, rays, any_hit=any_hit)
        want = (oracle.hit_ids(ref) >= 0).reshape(t, rpt).sum(axis=1).astype(np.uint32)
        assert np.array_equal(counts.array, want), f"surfel request {rid}: counts differ from the oracle"
    assert srv.stats()["completed"] == len(reqs)
    srv.close()


def test_poisson_device_and_surfel_harnesses(cw, serving_set, surfel_pools):
    scenes, pools, lays = serving_set
    srv = cw.Server(scenes, cw.default_server_config(memory_budget=256 << 20))
    dpools = [cw.DeviceBuffer.from_numpy(p.array) for p in pools]
    dhits = cw.DeviceBuffer(16 * 400000)
    say("poisson device 3000 req/s x 60")
    res, lat = srv.run_poisson_device(3000.0, 60, 1000, 100000, [d.ptr for d in dpools],
                                      [len(p.array) for p in pools], dhits.ptr, 400000, seed=9)
    assert res["rays"] > 0 and 0 < res["p50_ms"] <= res["p99_ms"] and len(lat) == 60
    counts = cw.PinnedArray((100000,), np.uint32)
    say("poisson surfels 3000 req/s x 60")
    res, lat = srv.run_poisson_surfels(3000.0, 60, 1000, 100000, surfel_pools, 8, 0.0, 0.05, counts, seed=10)
    assert res["rays"] > 0 and 0 < res["p50_ms"] <= res["p99_ms"] and np.all(lat > 0)
    assert srv.stats()["completed"] == 120
    srv.close()


def test_scene_admitted_while_serving(cw, serving_set):
    """A scene built and admitted while requests on the resident scenes are in
    flight serves its own requests with the oracle's hits; the earlier
    requests are unaffected."""
    from cwbvh_amd import workloads as W

    scenes, pools, lays = serving_set
    srv = cw.Server(scenes[:2], cw.default_server_config(max_inflight=3, memory_budget=64 << 20))
    early = []
    for k in range(10):
        n = 20000
        h = cw.PinnedArray((n, 4), np.float32)
        early.append((srv.submit(k % 2, pools[k % 2].ptr, n, h.ptr), k % 2, n, h))
    mesh = W.icospheres_of(25000, seed=909)
    _, lay = cw.build_layouts(mesh, bvh2=False)
    new = cw.Scene(lay)
    idx = srv.add_scene(new)
    assert idx == 2
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 5000, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=910)
    pool = cw.PinnedArray(rays.shape, np.float32)
    pool.array[:] = rays
    h = cw.PinnedArray((len(rays), 4), np.float32)
    rid = srv.submit(idx, pool.ptr, len(rays), h.ptr)
    srv.wait(rid, 60.0)
    assert np.array_equal(bits(h.array), bits(oracle_hits(lay, rays)))
    for rid, s, n, hh in early:
        srv.wait(rid, 60.0)
        assert np.array_equal(bits(hh.array), bits(oracle_hits(lays[s], pools[s].array[:n])))
    with pytest.raises(cw.CwbvhError):
        srv.submit(3, pool.ptr, 10, h.ptr)   # no scene 3
    srv.close()
