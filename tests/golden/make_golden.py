"""Write tests/golden/c0_small.npz: a small C0-style scene, its AO rays and the
brute-force closest/any hits (every triangle tested, no tree involved).

The reference holds no fixtures for this path; these vectors pin the product's
generators (mesh + rays bit for bit) and the oracle's results across rounds.
Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import cwbvh_amd as cw  # noqa: E402
import oracle  # noqa: E402

SEED = 0x60DE


def make():
    mesh = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=SEED)
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 512, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=SEED + 1)
    # Woop rows per primitive, in primitive order
    v, idx = mesh.vertices, mesh.indices
    woop = np.concatenate([cw.woopify(v[a], v[b], v[c]) for a, b, c in idx]).astype(np.float32)
    addrs = np.arange(0, len(woop), 3, dtype=np.int32)
    closest = oracle.trace_brute(woop, addrs, rays)
    anyh = oracle.trace_brute(woop, addrs, rays, any_hit=True)
    ids = oracle.hit_ids(closest)
    return dict(vertices=v, indices=idx, rays=rays, woop=woop,
                closest_prim=np.where(ids >= 0, ids // 3, -1).astype(np.int32),
                closest_tuv=closest[:, [0, 2, 3]].copy(),
                any_occluded=(oracle.hit_ids(anyh) >= 0))


if __name__ == "__main__":
    out = Path(__file__).resolve().parent / "c0_small.npz"
    np.savez_compressed(out, **make())
    print("wrote", out, out.stat().st_size, "bytes")
