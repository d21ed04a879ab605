"""Host-side checks: C ABI exports, argument validation, scene and ray generators."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol(cw):
    declared = declared_functions()
    assert {"rtBindCWBVHData", "rtTraceCWBVH", "rtBindBVH2Data", "rtTraceBVH2", "cwbvh_trace"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", cw.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    missing = sorted(declared - exported)
    assert not missing, missing
    # the Python binding covers the same surface
    assert declared <= set(cw.EXPORTED_SYMBOLS) | {"cwbvh_abi_version"}


def test_abi_version_and_status_strings(cw):
    assert cw.lib.cwbvh_abi_version() == 2
    assert cw.lib.cwbvh_status_string(0) == b"ok"
    assert cw.lib.cwbvh_status_string(4).startswith(b"tree deeper")


def test_argument_validation(cw):
    with pytest.raises(cw.CwbvhError) as e:
        cw.Mesh.from_arrays([[0, 0, 0]], [[0, 1, 2]])
    assert e.value.status == 1 and "out of range" in str(e.value)
    with pytest.raises(cw.CwbvhError):
        cw.Mesh.heightfield(0, 4)
    with pytest.raises(cw.CwbvhError):
        cw.build_layouts(cw.Mesh.heightfield(4, 4), cw.default_build_params(cwbvh_node_stride=6))
    with pytest.raises(cw.CwbvhError):
        cw.build_layouts(cw.Mesh.heightfield(4, 4), cw.default_build_params(max_depth=200))


def test_trace_config_defaults(cw):
    c = cw.default_trace_config(cw.KIND_CWBVH)
    assert c.block_size in (64, 128, 256) and 1 <= c.lds_depth <= cw.STACK_CAPACITY
    assert c.fetch_chunk >= 0 and c.refill_budget >= 1 and c.persistent in (0, 1, 2)
    assert c.loop_mode in (0, 1, 2)
    assert c.occupancy == 0  # automatic
    assert c.ray_order == 0  # rays in the given order
    b = cw.default_trace_config(cw.KIND_BVH2)
    assert b.block_size in (64, 128, 256)
    with pytest.raises(TypeError):
        cw.default_trace_config(cw.KIND_CWBVH, loop_mod=1)  # misspelt field


# ------------------------------------------------------------- generators


def test_heightfield_shape(cw):
    m = cw.Mesh.heightfield(256, 128, 8, 0.1, seed=1)
    assert m.num_triangles == 65536 and m.num_vertices == 257 * 129
    lo, hi = m.bounds()
    assert lo[0] == 0 and lo[1] == 0 and hi[0] == 1 and hi[1] == 1
    assert -0.2 <= lo[2] < 0 < hi[2] <= 0.2  # |h| <= sum 0.1*2^-k < 0.2
    # normals of the generated triangles point up
    v, idx = m.vertices, m.indices
    n = np.cross(v[idx[:, 1]] - v[idx[:, 0]], v[idx[:, 2]] - v[idx[:, 0]])
    assert np.all(n[:, 2] > 0)


def test_icosphere_grid_counts_and_closed_surfaces(cw):
    m = cw.Mesh.icosphere_grid(2, 1, 1, 7, 4.0, 0.5, 1.5, ground_margin=-1.0, seed=3)
    assert m.num_triangles == 2 * 980
    idx = m.indices
    # every edge of each sphere is shared by exactly two triangles (closed, deduplicated)
    e = np.sort(np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)
    g = cw.Mesh.icosphere_grid(10, 10, 10, 7, 4.0, 0.5, 1.5, ground_margin=1.0, seed=3)
    assert g.num_triangles == 980_002


def test_ao_rays_distribution(cw):
    m = cw.Mesh.heightfield(64, 32, 8, 0.1, seed=2)
    diag = m.diagonal()
    r = cw.rays_ao_surface(m, 20000, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=4)
    d = r[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(r[:, 3] == 0) and np.allclose(r[:, 7], np.float32(0.1 * diag))
    # 8 rays share an origin
    o = r[:, :3].reshape(-1, 8, 3)
    assert np.all(o == o[:, :1])
    # heightfield normals point up on average, so do the cosine-weighted directions
    assert d[:, 2].mean() > 0.3


def test_cosine_moments_flat_plane(cw):
    """On a flat plane (normal +z) the direction law is exactly cosine-weighted."""
    m = cw.Mesh.from_arrays([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], [[0, 1, 2], [0, 2, 3]])
    r = cw.rays_ao_surface(m, 50000, 4, 1e-3, 0.0, 1.0, seed=9)
    cz = r[:, 6].astype(np.float64)
    assert np.all(cz >= 0)
    assert abs(cz.mean() - 2 / 3) < 0.005
    assert abs((cz ** 2).mean() - 1 / 2) < 0.005
    assert abs(r[:, 4].mean()) < 0.01 and abs(r[:, 5].mean()) < 0.01
    assert np.allclose(r[:, 2], 1e-3)  # origin offset along the normal
    # area-uniform points
    assert abs(r[:, 0].mean() - 0.5) < 0.01 and abs(r[:, 1].mean() - 0.5) < 0.01


def test_generators_deterministic(cw):
    m = cw.Mesh.heightfield(32, 16, 8, 0.1, seed=5)
    a = cw.rays_ao_surface(m, 1000, 8, 1e-4, 0.0, 0.1, seed=1)
    b = cw.rays_ao_surface(m, 1000, 8, 1e-4, 0.0, 0.1, seed=1)
    c = cw.rays_ao_surface(m, 1000, 8, 1e-4, 0.0, 0.1, seed=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    s1 = cw.Mesh.cluster_soup(5000, 16, seed=3)
    s2 = cw.Mesh.cluster_soup(5000, 16, seed=3)
    assert np.array_equal(s1.vertices, s2.vertices)
    lo, hi = s1.bounds()
    assert np.all(lo > -0.1) and np.all(hi < 1.1)


def test_sphere_rays_and_shuffle(cw):
    m = cw.Mesh.cluster_soup(2000, 8, seed=1)
    r = cw.rays_sphere_from_surface(m, 20000, 1e-6, 0.0, np.inf, seed=2)
    assert np.all(np.isinf(r[:, 7]))
    d = r[:, 4:7].astype(np.float64)
    assert abs(d.mean(axis=0)).max() < 0.02  # uniform sphere
    s = cw.shuffle_rays(r, seed=3)
    assert not np.array_equal(s, r)
    assert np.array_equal(np.sort(s.view(np.uint64).reshape(len(s), -1), axis=0),
                          np.sort(r.view(np.uint64).reshape(len(r), -1), axis=0))


def test_pinhole_camera(cw):
    r = cw.rays_pinhole(64, 36, (0, 0, 0), (0, 1, 0), (0, 0, 1), 60.0)
    d = r[:, 4:7].reshape(36, 64, 3)
    assert np.allclose(np.linalg.norm(r[:, 4:7], axis=1), 1, atol=1e-6)
    assert np.all(d[:, :, 1] > 0)                   # looking along +y
    assert d[0, 0, 2] > 0 > d[-1, 0, 2]             # row 0 is the top
    assert d[0, 0, 0] < 0 < d[0, -1, 0]             # column 0 is the left
