"""Converter invariants for the reference BVH2 / CWBVH layouts
(reference src/GPUBVHConverter.cpp:10-378) and the 128-byte re-lay."""
import numpy as np
import pytest


def walk_cwbvh(lay):
    """Yield (node_index, slot, kind, payload) for every non-empty child slot."""
    nodes = lay.nodes
    S = lay.node_stride
    n_nodes = len(nodes) // S
    out = []
    for ni in range(n_nodes):
        nd = nodes[ni * S:ni * S + 5]
        n0b = nd[0].view(np.uint8)
        imask = int(n0b[15])
        e = n0b[12:15].view(np.int8).astype(int)
        base = int(nd[1].view(np.int32)[0])
        tbase = int(nd[1].view(np.int32)[1])
        meta = nd[1].view(np.uint8)[8:16]
        for s in range(8):
            m = int(meta[s])
            if m == 0:
                assert not (imask >> s) & 1
                continue
            if (m & 0b11000) == 0b11000 and (m >> 5) == 1:
                assert (imask >> s) & 1, "inner meta without imask bit"
                assert m & 31 == 24 + s
                child = base + bin(imask & ((1 << s) - 1)).count("1")
                out.append((ni, s, "inner", child, e))
            else:
                assert not (imask >> s) & 1
                unary = m >> 5
                assert unary in (1, 3, 7)
                cnt = bin(unary).count("1")
                off = m & 31
                assert off + cnt <= 24
                out.append((ni, s, "leaf", [3 * (tbase + off + k) for k in range(cnt)], e))
    return out


def check_coverage(mesh, leaf_prims, leaf_lo, leaf_hi, samples=6, seed=0):
    """Every point of every triangle lies in the box of some leaf that holds
    the triangle, every box on that leaf's root path included (leaf_lo/hi
    are those intersections).  Points: the vertices and random interior
    points.  With spatial splits a triangle's leaves hold only parts of it,
    so this is the conservativeness that matters for traversal."""
    v = mesh.vertices.astype(np.float64)[mesh.indices]                 # (nt, 3, 3)
    rng = np.random.default_rng(seed)
    w = rng.dirichlet([1.0, 1.0, 1.0], size=samples)
    pts = np.concatenate([v, np.einsum("kj,tjc->tkc", w, v)], axis=1)  # (nt, 3 + samples, 3)
    covered = np.zeros(pts.shape[:2], bool)
    prim = np.asarray(leaf_prims, np.int64)
    lo = np.asarray(leaf_lo, np.float64)[:, None, :]
    hi = np.asarray(leaf_hi, np.float64)[:, None, :]
    q = pts[prim]
    tol = 1e-12 * np.abs(q)  # the rounding of the float64 interior points
    inside = np.all((q + tol >= lo) & (q - tol <= hi), axis=2)
    np.logical_or.at(covered, prim, inside)
    bad = np.argwhere(~covered)
    assert len(bad) == 0, f"uncovered triangle points {bad[:5]}"


def check_cwbvh(mesh, lay, orc):
    nodes = lay.nodes
    S = lay.node_stride
    slots = walk_cwbvh(lay)
    refs = [a for (_, _, kind, p, _) in slots if kind == "leaf" for a in p]
    assert len(refs) == len(set(refs)) == lay.stats()["num_tri_refs"]
    prims = lay.prim_idx[np.array(refs, np.int64)]
    assert np.all(prims >= 0)
    # every triangle is referenced, at most once per leaf (spatial splits may
    # reference it from several leaves)
    assert set(prims.tolist()) == set(range(mesh.num_triangles))
    for (_, _, kind, p, _) in slots:
        if kind == "leaf":
            ids = lay.prim_idx[np.array(p)].tolist()
            assert len(ids) == len(set(ids))
    # every node except the root is the child of exactly one inner slot
    inner_children = [p for (_, _, kind, p, _) in slots if kind == "inner"]
    assert sorted(inner_children) == list(range(1, len(nodes) // S))
    for (_, _, _, _, e) in slots:
        assert np.all((e >= -126) & (e <= 127))
    # conservativeness: the decoded boxes on the path to a leaf contain the
    # parts of its triangles that the leaf stands for
    by_node = {}
    for ni, s, kind, p, e in slots:
        by_node.setdefault(ni, []).append((s, kind, p))
    eff = {0: (np.full(3, -np.inf), np.full(3, np.inf))}
    order = [0]
    leaf_prims, leaf_lo, leaf_hi = [], [], []
    while order:
        ni = order.pop()
        plo, phi = eff[ni]
        lo, hi = orc.cwbvh_decode(nodes[ni * S:ni * S + 5])
        for s, kind, p in by_node.get(ni, []):
            blo, bhi = np.maximum(plo, lo[s]), np.minimum(phi, hi[s])
            if kind == "inner":
                eff[p] = (blo, bhi)
                order.append(p)
            else:
                for pr in lay.prim_idx[np.array(p)]:
                    leaf_prims.append(pr)
                    leaf_lo.append(blo)
                    leaf_hi.append(bhi)
    check_coverage(mesh, leaf_prims, leaf_lo, leaf_hi)


def check_bvh2(mesh, lay):
    nodes = lay.nodes
    woop = lay.woop
    seen = []
    stack = [(0, np.full(3, -np.inf), np.full(3, np.inf))]
    visited = 0
    leaf_prims, leaf_lo, leaf_hi = [], [], []
    while stack:
        a, plo, phi = stack.pop()
        assert a % 4 == 0 and a + 3 < len(nodes)
        visited += 1
        n = nodes[a:a + 4]
        boxes = [(np.array([n[0, 0], n[0, 2], n[2, 0]]), np.array([n[0, 1], n[0, 3], n[2, 1]])),
                 (np.array([n[1, 0], n[1, 2], n[2, 2]]), np.array([n[1, 1], n[1, 3], n[2, 3]]))]
        ch = n[3].view(np.int32)[:2]
        for c, (blo, bhi) in zip(ch, boxes):
            elo, ehi = np.maximum(plo, blo), np.minimum(phi, bhi)
            if c >= 0:
                stack.append((int(c), elo, ehi))
                continue
            w = ~int(c)
            prims = []
            while woop[w].view(np.uint32)[0] != 0x80000000:
                prims.append(int(lay.prim_idx[w]))
                assert np.all(lay.prim_idx[w + 1:w + 3] == -1)
                w += 3
            assert lay.prim_idx[w] == -1
            assert len(prims) == len(set(prims))
            for pr in prims:
                leaf_prims.append(pr)
                leaf_lo.append(elo)
                leaf_hi.append(ehi)
            seen += prims
    assert visited == lay.stats()["num_nodes"]
    assert len(seen) == lay.stats()["num_tri_refs"]
    assert set(seen) == set(range(mesh.num_triangles))
    check_coverage(mesh, leaf_prims, leaf_lo, leaf_hi)


@pytest.mark.parametrize("which", ["heightfield", "soup", "planes", "spheres"])
def test_layout_invariants(cw, orc, which):
    rng = np.random.default_rng(4)
    if which == "heightfield":
        mesh = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=3)
    elif which == "soup":
        c = rng.uniform(0, 1, size=(800, 1, 3))
        v = (c + rng.normal(scale=0.05, size=(800, 3, 3))).astype(np.float32).reshape(-1, 3)
        mesh = cw.Mesh.from_arrays(v, np.arange(len(v), dtype=np.int32).reshape(-1, 3))
    elif which == "planes":
        mesh = cw.Mesh.validation_planes(300)  # flat nodes: zero z-extent
    else:
        mesh = cw.Mesh.icosphere_grid(2, 2, 2, 3, 4.0, 0.5, 1.5, ground_margin=0.5, seed=2)
    bvh2, cwb = cw.build_layouts(mesh)
    check_cwbvh(mesh, cwb, orc)
    check_cwbvh(mesh, cwb.restride(8), orc)
    check_bvh2(mesh, bvh2)


def test_restride_same_traversal(cw, orc):
    mesh = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=3)
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 500, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=1)
    _, cwb = cw.build_layouts(mesh)
    c8 = cwb.restride(8)
    assert c8.node_stride == 8 and len(c8.nodes) == 8 * (len(cwb.nodes) // 5)
    assert np.all(c8.nodes.view(np.uint32).reshape(-1, 8, 4)[:, 5:] == 0)
    a = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, rays)
    b = orc.trace_cwbvh(c8.nodes, 8, c8.woop, rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # stride-8 build straight from the builder is the same tree
    _, d = cw.build_layouts(mesh, cw.default_build_params(cwbvh_node_stride=8))
    assert np.array_equal(d.nodes.view(np.uint32), c8.nodes.view(np.uint32))


def test_single_and_empty_meshes(cw, orc):
    one = cw.Mesh.from_arrays([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])
    bvh2, cwb = cw.build_layouts(one)
    ray = np.array([[0.2, 0.2, 1, 0, 0, 0, -1, 10]], np.float32)
    for h, pi in ((orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, ray), cwb.prim_idx),
                  (orc.trace_bvh2(bvh2.nodes, bvh2.woop, ray), bvh2.prim_idx)):
        assert pi[orc.hit_ids(h)[0]] == 0 and h[0, 0] == 1.0
    empty = cw.Mesh.from_arrays(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    bvh2, cwb = cw.build_layouts(empty)
    assert orc.hit_ids(orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, ray))[0] == -1
    assert orc.hit_ids(orc.trace_bvh2(bvh2.nodes, bvh2.woop, ray))[0] == -1


def test_layout_from_arrays_roundtrip(cw):
    mesh = cw.Mesh.heightfield(16, 8, 8, 0.1, seed=3)
    bvh2, cwb = cw.build_layouts(mesh)
    for lay in (bvh2, cwb):
        r = cw.Layout.from_arrays(lay.kind, lay.node_stride, lay.nodes, lay.woop, lay.prim_idx)
        a, b = lay.stats(), r.stats()
        for k in ("kind", "node_stride", "nodes_f4", "woop_f4", "max_depth", "stack_need", "num_nodes",
                  "num_tri_refs"):
            assert a[k] == b[k], k
    with pytest.raises(cw.CwbvhError):
        cw.Layout.from_arrays(cw.KIND_CWBVH, 6, cwb.nodes, cwb.woop, cwb.prim_idx)


def test_depth_limit_respected(cw):
    # all centroids equal: only median splits are possible
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (5000, 1))
    mesh = cw.Mesh.from_arrays(v, np.arange(len(v), dtype=np.int32).reshape(-1, 3))
    bvh2, cwb = cw.build_layouts(mesh, cw.default_build_params(max_depth=20))
    assert bvh2.stats()["stack_need"] <= 21
    assert cwb.stats()["stack_need"] <= 21


@pytest.mark.parametrize("split_alpha", [0.0, 1e-5])
def test_build_same_for_any_thread_count(cw, split_alpha):
    """The parallel builder (work queue over subtrees, parallel binning of
    large nodes, Woop rows in parallel, both layouts side by side) returns
    the same bytes as the single-threaded one, spatial splits included."""
    mesh = cw.Mesh.cluster_soup(300_000, 16, 0.002, 0.2, -5.0, 0.5, seed=8)
    ref = cw.build_layouts(mesh, cw.default_build_params(num_threads=1, split_alpha=split_alpha))
    if split_alpha > 0:
        assert ref[1].stats()["num_tri_refs"] > mesh.num_triangles
    for t in (3, 8):
        got = cw.build_layouts(mesh, cw.default_build_params(num_threads=t, split_alpha=split_alpha))
        for a, b in zip(ref, got):
            assert np.array_equal(a.nodes.view(np.uint32), b.nodes.view(np.uint32))
            assert np.array_equal(a.woop.view(np.uint32), b.woop.view(np.uint32))
            assert np.array_equal(a.prim_idx, b.prim_idx)
            assert a.stats() == b.stats()


def test_sah_quality_cwbvh_fewer_bytes(cw, orc):
    """The 8-wide collapse must fetch fewer node bytes per ray than BVH2 on C0."""
    from cwbvh_amd import workloads as W

    wl = W.c0(num_points=2048)
    bvh2, cwb = cw.build_layouts(wl.mesh)
    _, c8 = orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, wl.rays, counters=True)
    _, c2 = orc.trace_bvh2(bvh2.nodes, bvh2.woop, wl.rays, counters=True)
    assert 80 * c8[0] + 48 * c8[1] < 0.8 * (64 * c2[0] + 48 * c2[1])


def _strip_mesh(cw, xs, y_extent, z_extent, base):
    """Triangles side by side along x: triangle i spans [xs[i], xs[i+1]] in x,
    the full y_extent and z_extent, offset by base."""
    v, idx = [], []
    for i in range(len(xs) - 1):
        x0, x1 = xs[i], xs[i + 1]
        k = len(v)
        v += [(x0, 0.0, 0.0), (x1, y_extent, 0.0), (x0, y_extent, z_extent)]
        idx.append((k, k + 1, k + 2))
    v = np.asarray(v, np.float64) + np.asarray(base, np.float64)
    return cw.Mesh.from_arrays(v.astype(np.float32), np.asarray(idx, np.int32))


@pytest.mark.parametrize("extent,y_extent,z_extent,base", [
    (255 * 2.0 ** -3, 1.0, 0.0, (1.0, 1.0, 1.0)),                          # exactly 255 * 2^e
    (255 * 2.0 ** -3 * (1 + 2.0 ** -20), 0.3, 0.25, (0.3, -0.7, 0.1)),     # just above a power of two
    (255 * 2.0 ** 4 * (1 - 2.0 ** -22), 255 * 2.0 ** 4, 1e-3, (-3.0, 2.0, 5.0)),  # just below
    (1e-30, 1e-30, 0.0, (0.0, 0.0, 0.0)),                                  # tiny extents
    (1e30, 3e29, 2e29, (-5e29, -1e29, 0.0)),                               # huge extents
    (1e-3, 1e-3, 1e-3, (-1e6, 3e5, -2e4)),                                 # small box far from origin
    (0.0, 0.0, 0.0, (1.0, 2.0, 3.0)),                                      # every triangle one point
])
def test_quantization_edge_cases(cw, orc, extent, y_extent, z_extent, base):
    """Child boxes decoded from the quantized planes contain their triangles
    at the exponent boundaries the reference converter gets wrong
    (src/GPUBVHConverter.cpp:284-316: log2f rounding, zero extent)."""
    xs = np.linspace(0.0, extent, 9)
    mesh = _strip_mesh(cw, xs, y_extent, z_extent, base)
    bvh2, cwb = cw.build_layouts(mesh)
    check_cwbvh(mesh, cwb, orc)
    check_bvh2(mesh, bvh2)
    # traversal through the quantized boxes finds what brute force finds
    lo, hi = mesh.bounds()
    c = (lo.astype(np.float64) + hi) / 2
    rng = np.random.default_rng(9)
    o = c + (hi - lo + 1e-30) * rng.uniform(-0.6, 0.6, size=(256, 3))
    o[:, 2] = lo[2] - max(extent, 1e-20)
    d = np.tile([0.0, 0.0, 1.0], (256, 1)) + rng.normal(scale=1e-3, size=(256, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((256, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, np.inf
    ref = orc.trace_brute(cwb.woop, np.arange(0, len(cwb.woop), 3, dtype=np.int32)[cwb.prim_idx[::3] >= 0], rays)
    got = orc.trace_cwbvh(cwb.nodes, cwb.node_stride, cwb.woop, rays)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_from_arrays_refuses_malformed_layouts(cw):
    """Caller arrays are walked on the host before any kernel sees them:
    indices a kernel would derive out of range are refused, not traced."""
    mesh = cw.Mesh.heightfield(8, 8, 4, 0.1, seed=5)
    bvh2, cwb = cw.build_layouts(mesh)

    def refuses(kind, stride, nodes, woop, prim, what):
        with pytest.raises(cw.CwbvhError) as e:
            cw.Layout.from_arrays(kind, stride, nodes, woop, prim)
        assert e.value.status == 1 and what in str(e.value), str(e.value)

    # CWBVH: leaf triangles beyond the Woop array
    refuses(cw.KIND_CWBVH, 5, cwb.nodes, cwb.woop[:9], cwb.prim_idx[:9], "out of range")
    # CWBVH: inner meta naming another slot than its imask bit
    n = cwb.nodes.copy()
    meta = n[1].view(np.uint8)[8:16]
    imask = n[0].view(np.uint8)[15]
    s = next(k for k in range(8) if imask >> k & 1)
    meta[s] = 0x38 | ((s + 1) % 8)
    refuses(cw.KIND_CWBVH, 5, n, cwb.woop, cwb.prim_idx, "meta")
    # CWBVH: a child pointing back at the root (a cycle)
    n = cwb.nodes.copy()
    n[1].view(np.int32)[0] = 0
    refuses(cw.KIND_CWBVH, 5, n, cwb.woop, cwb.prim_idx, "twice")
    # BVH2: a child address that is not a node boundary
    n = bvh2.nodes.copy()
    c = n[3].view(np.int32)
    j = 0 if c[0] >= 0 else 1
    c[j] = c[j] + 1 if c[j] >= 0 else 1
    refuses(cw.KIND_BVH2, 4, n, bvh2.woop, bvh2.prim_idx, "node address")
    # BVH2: a leaf whose terminator is cut off
    refuses(cw.KIND_BVH2, 4, bvh2.nodes, bvh2.woop[:5], bvh2.prim_idx[:5], "Woop array")
    # prim index shorter than the Woop array
    with pytest.raises(cw.CwbvhError):
        cw.Layout.from_arrays(cw.KIND_CWBVH, 5, cwb.nodes, cwb.woop, cwb.prim_idx[:-3])
    # the untouched arrays are accepted
    assert cw.Layout.from_arrays(cw.KIND_CWBVH, 5, cwb.nodes, cwb.woop, cwb.prim_idx).stats()["num_nodes"] == \
        cwb.stats()["num_nodes"]


def assert_same_hits(a, b, prim_idx, orc, prim_idx_b=None):
    """Equal t, u, v bits and the same triangle behind the two hit addresses."""
    prim_idx_b = prim_idx if prim_idx_b is None else prim_idx_b
    assert np.array_equal(a[:, [0, 2, 3]].view(np.uint32), b[:, [0, 2, 3]].view(np.uint32))
    ia, ib = orc.hit_ids(a), orc.hit_ids(b)
    assert np.array_equal(ia >= 0, ib >= 0)
    hit = ia >= 0
    assert np.array_equal(prim_idx[ia[hit]], prim_idx_b[ib[hit]])


def _slivers(cw, n_long, n_small, seed):
    """Long thin triangles crossing the unit cube among small ones: the case
    spatial splits are for (object splits leave heavily overlapping boxes)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 1, size=(n_long, 3))
    b = rng.uniform(0, 1, size=(n_long, 3))
    c = a + rng.normal(scale=0.01, size=(n_long, 3))
    cen = rng.uniform(0, 1, size=(n_small, 1, 3))
    small = cen + rng.normal(scale=0.01, size=(n_small, 3, 3))
    v = np.concatenate([np.stack([a, b, c], axis=1), small]).astype(np.float32).reshape(-1, 3)
    return cw.Mesh.from_arrays(v, np.arange(len(v), dtype=np.int32).reshape(-1, 3))


@pytest.mark.parametrize("stride", [5, 8])
def test_spatial_splits(cw, orc, stride):
    """Spatial splits reference a triangle from several leaves, each with the
    box of its part: layouts stay conservative (every triangle point is in
    a leaf box that holds it), traversal equals brute force bit for bit, and
    the per-ray work drops against the object-split build."""
    mesh = _slivers(cw, 400, 4000, seed=12)
    nt = mesh.num_triangles
    off = cw.build_layouts(mesh, cw.default_build_params(split_alpha=0.0, cwbvh_node_stride=stride))
    on = cw.build_layouts(mesh, cw.default_build_params(split_alpha=1e-5, split_budget=0.5,
                                                        cwbvh_node_stride=stride))
    for lay in off:
        assert lay.stats()["num_tri_refs"] == nt
    bvh2, cwb = on
    assert nt < bvh2.stats()["num_tri_refs"] <= nt + int(0.5 * nt)
    assert nt < cwb.stats()["num_tri_refs"] <= nt + int(0.5 * nt)
    check_cwbvh(mesh, cwb, orc)
    check_bvh2(mesh, bvh2)
    rng = np.random.default_rng(13)
    rays = np.zeros((3000, 8), np.float32)
    rays[:, :3] = rng.uniform(-0.2, 1.2, size=(3000, 3))
    d = rng.normal(size=(3000, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf
    rows = np.arange(0, len(cwb.woop), 3, dtype=np.int32)[cwb.prim_idx[::3] >= 0]
    ref = orc.trace_brute(cwb.woop, rows, rays)
    got, k_on = orc.trace_cwbvh(cwb.nodes, stride, cwb.woop, rays, counters=True)
    # a split triangle has several addresses: brute force reports the lowest,
    # traversal the first copy it tests; t, u, v and the triangle agree
    assert_same_hits(got, ref, cwb.prim_idx, orc)
    for any_hit in (False, True):
        h2 = orc.trace_bvh2(bvh2.nodes, bvh2.woop, rays, any_hit=any_hit)
        h8 = orc.trace_cwbvh(cwb.nodes, stride, cwb.woop, rays, any_hit=any_hit)
        assert np.array_equal(orc.hit_ids(h2) >= 0, orc.hit_ids(h8) >= 0)
    # the same nearest triangle and distance as the object-split layout
    base = orc.trace_cwbvh(off[1].nodes, stride, off[1].woop, rays)
    assert_same_hits(got, base, cwb.prim_idx, orc, off[1].prim_idx)
    _, k_off = orc.trace_cwbvh(off[1].nodes, stride, off[1].woop, rays, counters=True)
    assert k_on[0] + k_on[1] < k_off[0] + k_off[1]


def _mesh_with_non_finite_vertices(cw):
    """400 random triangles over 300 vertices, three of them NaN, +inf and
    -inf; returns the mesh and the mask of triangles whose corners are finite."""
    rng = np.random.default_rng(31)
    v = rng.random((300, 3)).astype(np.float32)
    idx = rng.integers(0, 300, (400, 3)).astype(np.int32)
    v[5, 1], v[7, 0], v[9, 2] = np.nan, np.inf, -np.inf
    finite = ~np.isin(idx, [5, 7, 9]).any(axis=1)
    return cw.Mesh.from_arrays(v, idx), v, idx, finite


@pytest.mark.parametrize("split_alpha", [0.0, 1e-5])
def test_non_finite_vertices_stay_out_of_the_tree(cw, orc, split_alpha):
    """A triangle with a NaN or infinite corner can never be hit (its Woop
    rows are the never-hit rows) and is left out of the tree, so its box does
    not make the boxes around its neighbours infinite: both layouts hold
    exactly the finite triangles, and traversal equals brute force over the
    Woop rows of every triangle of the mesh, closest and any hit."""
    mesh, v, idx, finite = _mesh_with_non_finite_vertices(cw)
    assert 0 < (~finite).sum() < len(idx)
    bvh2, cwb = cw.build_layouts(mesh, cw.default_build_params(split_alpha=split_alpha))
    for lay in (bvh2, cwb):
        present = set(int(p) for p in lay.prim_idx if p >= 0)
        assert present == set(np.flatnonzero(finite).tolist())
    rows = np.concatenate([cw.woopify(v[a], v[b], v[c]) for a, b, c in idx]).astype(np.float32)
    assert np.array_equal(rows[3 * np.flatnonzero(~finite)], np.tile([[0, 0, 0, 1]], ((~finite).sum(), 1)))
    rng = np.random.default_rng(32)
    rays = np.zeros((3000, 8), np.float32)
    rays[:, :3] = rng.uniform(-0.2, 1.2, size=(3000, 3))
    d = rng.normal(size=(3000, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = np.inf
    addrs = np.arange(0, len(rows), 3, dtype=np.int32)
    for any_hit in (False, True):
        ref = orc.trace_brute(rows, addrs, rays, any_hit=any_hit)
        hit_ref = orc.hit_ids(ref) >= 0
        assert 0.1 < hit_ref.mean() < 0.9
        for lay, got in ((cwb, orc.trace_cwbvh(cwb.nodes, cwb.node_stride, cwb.woop, rays, any_hit=any_hit)),
                         (bvh2, orc.trace_bvh2(bvh2.nodes, bvh2.woop, rays, any_hit=any_hit))):
            ids = orc.hit_ids(got)
            assert np.array_equal(ids >= 0, hit_ref), (lay.kind, any_hit)
            if not any_hit:
                # the same triangle at the same distance (the layouts store
                # the triangles at their own addresses)
                assert np.array_equal(got[hit_ref, 0].view(np.uint32), ref[hit_ref, 0].view(np.uint32))
                assert np.array_equal(lay.prim_idx[ids[hit_ref]], orc.hit_ids(ref)[hit_ref] // 3)


def test_spatial_split_budget_and_params(cw):
    mesh = _slivers(cw, 300, 2000, seed=14)
    nt = mesh.num_triangles
    assert cw.default_build_params().split_alpha == 0.0
    for budget in (0.0, 0.02, 0.2):
        bvh2, cwb = cw.build_layouts(mesh, cw.default_build_params(split_alpha=1e-5, split_budget=budget))
        assert nt <= cwb.stats()["num_tri_refs"] <= nt + int(budget * nt)
        assert nt <= bvh2.stats()["num_tri_refs"] <= nt + int(budget * nt)
    for bad in (dict(split_alpha=-1.0), dict(split_alpha=2.0), dict(split_budget=-0.1), dict(split_budget=5.0),
                dict(split_alpha=float("nan"))):
        with pytest.raises(cw.CwbvhError):
            cw.build_layouts(mesh, cw.default_build_params(**bad))
