"""Pins for the CPU oracle (oracle/oracle.c).

The reference ships no tests, fixtures or golden vectors for this path, so the
oracle is pinned by: the reference's own known-answer scene
(src/ValidationKernels.cu:6-39), closed-form single-triangle answers, brute
force over every triangle (independent of any tree), agreement of the two
independent traversal restatements (BVH2 and CWBVH), and invariants of the
any-hit / closest-hit relation.
"""
import numpy as np
import pytest


def _prims(hits, prim_idx, orc):
    ids = orc.hit_ids(hits)
    return np.where(ids >= 0, prim_idx[np.clip(ids, 0, None)], -1)


def _ray(o, d, tmin=0.0, tmax=1e30):
    return np.array([[*o, tmin, *d, tmax]], np.float32)


# ---------------------------------------------------------- known answers


def test_validation_planes_known_answer(cw, orc):
    """Reference KAT: triangle k in plane z=k; ray r starts at z=k-0.5 with
    k = (r*12345678) mod n (64-bit), so it hits triangle k at t = 0.5/d_z."""
    n_tri, n_ray = 1000, 4000
    mesh = cw.Mesh.validation_planes(n_tri)
    rays = cw.rays_validation(n_ray, n_tri)
    k = (np.arange(n_ray, dtype=np.uint64) * np.uint64(12345678) % np.uint64(n_tri)).astype(np.int64)
    bvh2, cwb = cw.build_layouts(mesh)
    d = np.array([-0.01, 0.5, 1.0])
    t_exact = 0.5 / (d[2] / np.linalg.norm(d))
    for hits, prim_idx in ((orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, rays), cwb.prim_idx),
                           (orc.trace_bvh2(bvh2.nodes, bvh2.woop, rays), bvh2.prim_idx)):
        assert np.array_equal(_prims(hits, prim_idx, orc), k)
        # Woop rows of a plane at z=k lose ~k ulps in Oz; error bound grows with k
        assert np.all(np.abs(hits[:, 0] - t_exact) <= 2e-7 * (1 + k) + 1e-6)


def test_validation_rays_use_64bit_modulo(cw):
    rays = cw.rays_validation(400, 7)
    k = np.array([(r * 12345678) % 7 for r in range(400)], np.float32)
    assert np.array_equal(rays[:, 2], k - 0.5)  # no int32 overflow past r=173


def test_single_triangle_closed_form(cw, orc):
    """v0=(1,0,0), v1=(0,1,0), v2=(0,0,0): hit point (x,y,0) has u=x (weight of
    v0), v=y (weight of v1); a ray from z=2 straight down hits at t=2."""
    w = cw.woopify([1, 0, 0], [0, 1, 0], [0, 0, 0])
    ok, t, u, v = orc.woop_test(w, _ray((0.25, 0.5, 2.0), (0, 0, -1)))
    assert ok and t == 2.0 and u == 0.25 and v == 0.5
    ok, t, u, v = orc.woop_test(w, _ray((0.125, 0.25, -3.0), (0, 0, 1)))  # from below: both faces
    assert ok and t == 3.0 and u == 0.125 and v == 0.25
    assert not orc.woop_test(w, _ray((0.75, 0.5, 2.0), (0, 0, -1)))[0]  # u+v > 1
    assert not orc.woop_test(w, _ray((-0.1, 0.5, 2.0), (0, 0, -1)))[0]  # u < 0


def test_woop_edges_inclusive_t_strict(cw, orc):
    w = cw.woopify([1, 0, 0], [0, 1, 0], [0, 0, 0])
    assert orc.woop_test(w, _ray((0.0, 0.5, 1.0), (0, 0, -1)))[0]       # u == 0 edge
    assert orc.woop_test(w, _ray((0.5, 0.5, 1.0), (0, 0, -1)))[0]       # u+v == 1 edge
    assert orc.woop_test(w, _ray((0.0, 0.0, 1.0), (0, 0, -1)))[0]       # vertex v2
    assert not orc.woop_test(w, _ray((0.2, 0.2, 1.0), (0, 0, -1), tmax=1.0))[0]   # t == tmax
    assert orc.woop_test(w, _ray((0.2, 0.2, 1.0), (0, 0, -1), tmax=np.nextafter(np.float32(1), np.float32(2))))[0]
    assert not orc.woop_test(w, _ray((0.2, 0.2, 1.0), (0, 0, -1), tmin=1.0))[0]   # t == tmin


def test_woopify_matches_reference_cofactor_route(cw, orc):
    """Product Woopify (adjugate in double) vs the reference's float32 4x4
    cofactor inverse restated in the oracle: equal up to float rounding."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        a = cw.woopify(*v)
        b = orc.woopify_ref(*v)
        scale = np.abs(b).max()
        assert np.allclose(a, b, rtol=2e-4, atol=2e-5 * scale)
        # local coordinates of the vertices: v0 -> (u,v)=(1,0), v1 -> (0,1), v2 -> (0,0); w = 0
        for vert, uv in ((v[0], (1, 0)), (v[1], (0, 1)), (v[2], (0, 0))):
            x = np.append(vert.astype(np.float64), 1.0)
            assert abs(a[1].astype(np.float64) @ x - uv[0]) < 1e-4
            assert abs(a[2].astype(np.float64) @ x - uv[1]) < 1e-4
            assert abs(a[0, :3].astype(np.float64) @ x[:3] - a[0, 3]) < 1e-4 * max(1, abs(a[0, 3]))


def test_woopify_negative_zero_and_degenerate(cw, orc):
    # axis-aligned triangle in z=0 whose row0.x would be -0.0
    w = cw.woopify([0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert w[0].view(np.uint32)[0] != 0x80000000
    # degenerate (collinear) triangle never reports a hit
    d = cw.woopify([0, 0, 0], [1, 1, 1], [2, 2, 2])
    rng = np.random.default_rng(1)
    for _ in range(100):
        o = rng.normal(size=3)
        dirv = rng.normal(size=3)
        dirv /= np.linalg.norm(dirv)
        assert not orc.woop_test(d, _ray(o, dirv, 0.0, np.inf))[0]


def test_tie_rule_lowest_address_wins(cw, orc):
    """Two coincident triangles: closest hit reports the lower Woop address
    whatever the test order."""
    w = cw.woopify([1, 0, 0], [0, 1, 0], [0, 0, 0])
    woop = np.concatenate([w, w]).astype(np.float32)
    ray = _ray((0.2, 0.3, 1.0), (0, 0, -1))
    for order in ([0, 3], [3, 0]):
        h = orc.trace_brute(woop, np.array(order, np.int32), ray)
        assert orc.hit_ids(h)[0] == 0


# ------------------------------------------------------- brute-force pins


def _soup(rng, n, spread=1.0, size=0.2):
    c = rng.uniform(0, spread, size=(n, 1, 3))
    v = (c + rng.normal(scale=size, size=(n, 3, 3))).astype(np.float32).reshape(-1, 3)
    return v, np.arange(3 * n, dtype=np.int32).reshape(-1, 3)


def _random_rays(rng, n, lo=-0.5, hi=1.5, tmax=np.float32(1e30)):
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 4:7] = d
    r[:, 3] = 0.0
    r[:, 7] = tmax
    return r


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_traversals_equal_brute_force(cw, orc, seed):
    rng = np.random.default_rng(seed)
    v, idx = _soup(rng, 700)
    mesh = cw.Mesh.from_arrays(v, idx)
    rays = _random_rays(rng, 3000)
    # axis-aligned and signed-zero directions exercise the 2^-80 clamp
    rays[:200, 4:7] = [0.0, 0.0, -1.0]
    rays[200:400, 4:7] = [-0.0, 1.0, 0.0]
    bvh2, cwb = cw.build_layouts(mesh, cw.default_build_params(bvh2_max_leaf=4))
    for lay in (bvh2, cwb):
        addrs = np.nonzero(lay.prim_idx >= 0)[0].astype(np.int32)
        for any_hit in (False, True):
            brute = orc.trace_brute(lay.woop, addrs, rays, any_hit=any_hit)
            if lay.kind == cw.KIND_CWBVH:
                tr = orc.trace_cwbvh(lay.nodes, 5, lay.woop, rays, any_hit=any_hit)
            else:
                tr = orc.trace_bvh2(lay.nodes, lay.woop, rays, any_hit=any_hit)
            pb, pt = _prims(brute, lay.prim_idx, orc), _prims(tr, lay.prim_idx, orc)
            if any_hit:
                assert np.array_equal(pb >= 0, pt >= 0)
            else:
                assert np.array_equal(pb, pt)
                assert np.array_equal(brute[:, [0, 2, 3]].view(np.uint32), tr[:, [0, 2, 3]].view(np.uint32))


def test_bvh2_and_cwbvh_agree_on_c0_sample(cw, orc):
    from cwbvh_amd import workloads as W

    wl = W.c0(num_points=4096)
    bvh2, cwb = cw.build_layouts(wl.mesh)
    h8 = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays)
    h2 = orc.trace_bvh2(bvh2.nodes, bvh2.woop, wl.rays)
    assert np.array_equal(_prims(h8, cwb.prim_idx, orc), _prims(h2, bvh2.prim_idx, orc))
    # same Woop rows per primitive -> same t/u/v bits
    assert np.array_equal(h8[:, [0, 2, 3]].view(np.uint32), h2[:, [0, 2, 3]].view(np.uint32))
    assert 0.05 < (orc.hit_ids(h8) >= 0).mean() < 0.9


def test_any_hit_occlusion_equals_closest_found(cw, orc):
    from cwbvh_amd import workloads as W

    wl = W.c0(num_points=2048)
    bvh2, cwb = cw.build_layouts(wl.mesh)
    c = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays)
    a = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays, any_hit=True)
    assert np.array_equal(orc.hit_ids(c) >= 0, orc.hit_ids(a) >= 0)
    # any-hit t is never closer than the closest hit; misses report tmax
    hit = orc.hit_ids(a) >= 0
    assert np.all(a[hit, 0] >= c[hit, 0])
    assert np.array_equal(a[~hit, 0], wl.rays[~hit, 7])
    assert np.all(a[~hit, 2:] == 0)


def test_oracle_stack_limit_reported(cw, orc):
    from cwbvh_amd import workloads as W

    wl = W.c0(num_points=256)
    _, cwb = cw.build_layouts(wl.mesh)
    need = cwb.stats()["stack_need"]
    orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays, stack_capacity=need)
    with pytest.raises(RuntimeError):
        orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays, stack_capacity=1)


def test_oracle_counters_and_threads_deterministic(cw, orc):
    from cwbvh_amd import workloads as W

    wl = W.c0(num_points=2048)
    _, cwb = cw.build_layouts(wl.mesh)
    h1, c1 = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays, threads=1, counters=True)
    h8, c8 = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays, threads=8, counters=True)
    assert np.array_equal(h1.view(np.uint32), h8.view(np.uint32))
    assert np.array_equal(c1, c8) and c1[0] >= len(wl.rays)


def test_compare_hits_logs_every_disagreement(cw, orc):
    """Planted disagreements (on the CPU, in a copy of the oracle's hits) are
    each logged with their ray and both hits."""
    mesh = cw.Mesh.heightfield(16, 8, 4, 0.1, seed=5)
    _, cwb = cw.build_layouts(mesh, bvh2=False)
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 300, 4, 1e-4 * diag, 0.0, 0.2 * diag, seed=6)
    ref = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, rays)
    ids = orc.hit_ids(ref)
    hit, miss = np.nonzero(ids >= 0)[0], np.nonzero(ids < 0)[0]
    assert len(hit) > 3 and len(miss) > 3
    got = ref.copy()
    got[miss[0], 1] = ref[hit[0], 1]           # a miss reported as a hit
    got[hit[1], 1] = np.array([-1], np.int32).view(np.float32)[0]   # a hit reported as a miss
    clean = orc.compare_hits(ref, ref, cwb.woop, rays, cwb.prim_idx)
    assert clean["exact"] == len(rays) and clean["log"] == []
    c = orc.compare_hits(got, ref, cwb.woop, rays, cwb.prim_idx)
    assert sorted(c["mismatches"]) == sorted([int(miss[0]), int(hit[1])])
    logged = {e["ray_index"]: e for e in c["log"]}
    assert set(logged) == {int(miss[0]), int(hit[1])}
    e = logged[int(miss[0])]
    assert e["kind"] == "mismatch" and np.array_equal(np.float32(e["ray"]), rays[miss[0]])
    assert e["oracle_prim"] == -1 and e["gpu_prim"] >= 0
    a = orc.compare_hits(got, ref, cwb.woop, rays, cwb.prim_idx, any_hit=True)
    assert {x["ray_index"] for x in a["log"]} == {int(miss[0]), int(hit[1])}
    assert all(x["kind"] == "occlusion differs" for x in a["log"])
