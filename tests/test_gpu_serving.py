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


@pytest.mark.parametrize("serial", [0, 1])
def test_requests_same_alone_and_in_mix(cw, serving_set, serial):
    scenes, pools, lays = serving_set
    rng = np.random.default_rng(5)
    cfg = cw.default_server_config(serial=serial, max_inflight=3, max_batch_requests=8,
                                   memory_budget=64 << 20)
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
