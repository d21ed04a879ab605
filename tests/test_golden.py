"""Committed vectors (tests/golden/c0_small.npz, written by make_golden.py):
generators, Woop rows and oracle traversal must reproduce them bit for bit."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden" / "c0_small.npz"
SEED = 0x60DE


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_generators_reproduce_fixture(cw, gold):
    mesh = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=SEED)
    assert np.array_equal(mesh.vertices.view(np.uint32), gold["vertices"].view(np.uint32))
    assert np.array_equal(mesh.indices, gold["indices"])
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 512, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=SEED + 1)
    assert np.array_equal(rays.view(np.uint32), gold["rays"].view(np.uint32))


def test_woop_rows_reproduce_fixture(cw, gold):
    v, idx = gold["vertices"], gold["indices"]
    w = np.concatenate([cw.woopify(v[a], v[b], v[c]) for a, b, c in idx])
    assert np.array_equal(w.view(np.uint32), gold["woop"].view(np.uint32))


def _check(hits, prim_idx, gold, orc, any_hit=False):
    ids = orc.hit_ids(hits)
    prims = np.where(ids >= 0, prim_idx[np.clip(ids, 0, None)], -1)
    if any_hit:
        assert np.array_equal(prims >= 0, gold["any_occluded"])
    else:
        assert np.array_equal(prims, gold["closest_prim"])
        assert np.array_equal(hits[:, [0, 2, 3]].view(np.uint32), gold["closest_tuv"].view(np.uint32))


def test_oracle_traversals_reproduce_fixture(cw, orc, gold):
    mesh = cw.Mesh.from_arrays(gold["vertices"], gold["indices"])
    bvh2, cwb = cw.build_layouts(mesh)
    rays = gold["rays"]
    for any_hit in (False, True):
        _check(orc.trace_cwbvh(cwb.nodes, 5, cwb.woop, rays, any_hit=any_hit), cwb.prim_idx, gold, orc, any_hit)
        _check(orc.trace_bvh2(bvh2.nodes, bvh2.woop, rays, any_hit=any_hit), bvh2.prim_idx, gold, orc, any_hit)
    assert 0.05 < (gold["closest_prim"] >= 0).mean() < 0.95


@pytest.mark.gpu
def test_gpu_reproduces_fixture(cw, orc, gold):
    mesh = cw.Mesh.from_arrays(gold["vertices"], gold["indices"])
    bvh2, cwb = cw.build_layouts(mesh)
    for lay in (cwb, bvh2):
        sc = cw.Scene(lay)
        for any_hit in (False, True):
            h = sc.trace_numpy(gold["rays"], cw.TRACE_ANY if any_hit else cw.TRACE_CLOSEST)
            _check(h, lay.prim_idx, gold, orc, any_hit)
