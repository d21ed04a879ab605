"""On-disk formats: the reference's mesh and surfel files (src/main.cpp:31-64),
AO rays from surfels (src/ValidationKernels.cu:63-101), and the layout cache.

The reference's own files (mesh.dat, samplecache.dat) are not in its checkout,
so each reader is checked against files written here byte by byte with numpy
in the layout the reference's fread calls expect."""
import numpy as np
import pytest


def _write_reference_mesh(path, v, idx, mat):
    with open(path, "wb") as f:
        np.array([len(v), len(idx)], "<i4").tofile(f)
        np.asarray(v, "<f4").tofile(f)
        np.asarray(idx, "<i4").tofile(f)
        np.asarray(mat, "<i4").tofile(f)


def _write_reference_samples(path, sx, sy, pos, nrm, rad):
    with open(path, "wb") as f:
        np.array([sx, sy], "<i4").tofile(f)
        np.asarray(pos, "<f4").tofile(f)
        np.asarray(nrm, "<f4").tofile(f)
        np.asarray(rad, "<f4").tofile(f)


def test_mesh_dat_matches_reference_layout(cw, tmp_path):
    rng = np.random.default_rng(1)
    v = rng.normal(size=(50, 3)).astype(np.float32)
    idx = rng.integers(0, 50, size=(80, 3)).astype(np.int32)
    mat = rng.integers(0, 7, size=80).astype(np.int32)
    p = tmp_path / "mesh.dat"
    _write_reference_mesh(p, v, idx, mat)
    m = cw.Mesh.read_dat(p)
    assert np.array_equal(m.vertices, v) and np.array_equal(m.indices, idx) and np.array_equal(m.materials, mat)
    # our writer produces the same bytes
    q = tmp_path / "again.dat"
    m.write_dat(q)
    assert q.read_bytes() == p.read_bytes()
    # a generated mesh without materials writes zeros
    h = cw.Mesh.heightfield(4, 3, 2, 0.1, seed=1)
    h.write_dat(q)
    back = cw.Mesh.read_dat(q)
    assert np.array_equal(back.vertices, h.vertices) and np.array_equal(back.materials, np.zeros(24, np.int32))


def test_mesh_dat_refuses_bad_files(cw, tmp_path):
    v = np.zeros((3, 3), np.float32)
    p = tmp_path / "m.dat"
    _write_reference_mesh(p, v, [[0, 1, 2]], [0])
    data = p.read_bytes()
    (tmp_path / "short.dat").write_bytes(data[:-1])
    with pytest.raises(cw.CwbvhError) as e:
        cw.Mesh.read_dat(tmp_path / "short.dat")
    assert e.value.status == 10
    (tmp_path / "long.dat").write_bytes(data + b"\0")
    with pytest.raises(cw.CwbvhError) as e:
        cw.Mesh.read_dat(tmp_path / "long.dat")
    assert e.value.status == 10
    _write_reference_mesh(tmp_path / "idx.dat", v, [[0, 1, 3]], [0])
    with pytest.raises(cw.CwbvhError) as e:
        cw.Mesh.read_dat(tmp_path / "idx.dat")
    assert e.value.status == 1
    with pytest.raises(cw.CwbvhError) as e:
        cw.Mesh.read_dat(tmp_path / "missing.dat")
    assert e.value.status == 10


def test_samples_match_reference_layout(cw, tmp_path):
    rng = np.random.default_rng(2)
    sx, sy = 7, 5
    pos = rng.normal(size=(sx * sy, 4)).astype(np.float32)
    nrm = rng.normal(size=(sx * sy, 4)).astype(np.float32)
    rad = rng.uniform(0.01, 0.1, size=sx * sy).astype(np.float32)
    p = tmp_path / "samplecache.dat"
    _write_reference_samples(p, sx, sy, pos, nrm, rad)
    P, N, R, size = cw.read_samples(p)
    assert size == (sx, sy) and np.array_equal(P, pos) and np.array_equal(N, nrm) and np.array_equal(R, rad)
    q = tmp_path / "again.dat"
    cw.write_samples(q, P, N, R, sx, sy)
    assert q.read_bytes() == p.read_bytes()
    (tmp_path / "short.dat").write_bytes(p.read_bytes()[:-4])
    with pytest.raises(cw.CwbvhError) as e:
        cw.read_samples(tmp_path / "short.dat")
    assert e.value.status == 10


def test_ao_rays_from_samples(cw):
    """Reference semantics (src/ValidationKernels.cu:63-101): origin p + n*r/2,
    tmin 0.01, tmax 1e20, unit cosine-weighted directions around n."""
    rng = np.random.default_rng(3)
    n_tex, k = 400, 32
    pos = np.zeros((n_tex, 4), np.float32)
    pos[:, :3] = rng.uniform(-5, 5, size=(n_tex, 3))
    nrm = np.zeros((n_tex, 4), np.float32)
    d = rng.normal(size=(n_tex, 3))
    d[:50] = [0, 0, 1]       # normal along z: the frame falls back to n x y
    d[50:100] = [0, 0, -1]
    nrm[:, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rad = rng.uniform(0.01, 0.2, size=n_tex).astype(np.float32)
    rays = cw.rays_ao_from_samples(pos, nrm, rad, k, seed=11)
    assert rays.shape == (n_tex * k, 8)
    o = rays[:, :3].reshape(n_tex, k, 3)
    want = pos[:, :3] + nrm[:, :3] * (rad * np.float32(0.5))[:, None]
    assert np.array_equal(o, np.broadcast_to(want[:, None, :], o.shape))
    assert np.all(rays[:, 3] == np.float32(0.01)) and np.all(rays[:, 7] == np.float32(1e20))
    dirs = rays[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-6)
    cos = np.einsum("ij,ij->i", dirs, np.repeat(nrm[:, :3], k, axis=0))
    assert cos.min() > -1e-6
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2
    assert abs(cos.mean() - 2 / 3) < 0.01 and abs((cos ** 2).mean() - 0.5) < 0.01
    assert np.array_equal(rays, cw.rays_ao_from_samples(pos, nrm, rad, k, seed=11))
    assert not np.array_equal(rays, cw.rays_ao_from_samples(pos, nrm, rad, k, seed=12))


@pytest.mark.parametrize("which", ["bvh2", "cwbvh", "cwbvh128"])
def test_layout_cache_roundtrip(cw, orc, tmp_path, which):
    mesh = cw.Mesh.icosphere_grid(2, 2, 2, 3, 4.0, 0.5, 1.5, ground_margin=0.5, seed=4)
    bvh2, cwb = cw.build_layouts(mesh)
    lay = {"bvh2": bvh2, "cwbvh": cwb, "cwbvh128": cwb.restride(8) if cwb else None}[which]
    p = tmp_path / "scene.lay"
    lay.save(p)
    back = cw.Layout.load(p)
    for a in ("nodes", "woop", "prim_idx"):
        assert np.array_equal(getattr(back, a).view(np.uint32) if a != "prim_idx" else back.prim_idx,
                              getattr(lay, a).view(np.uint32) if a != "prim_idx" else lay.prim_idx)
    assert back.stats() == lay.stats()
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 2000, 4, 1e-4 * diag, 0.0, 0.2 * diag, seed=5)
    if lay.kind == cw.KIND_CWBVH:
        a = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, rays)
        b = orc.trace_cwbvh(back.nodes, back.node_stride, back.woop, rays)
    else:
        a = orc.trace_bvh2(lay.nodes, lay.woop, rays)
        b = orc.trace_bvh2(back.nodes, back.woop, rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_layout_cache_refuses_damaged_files(cw, tmp_path):
    _, cwb = cw.build_layouts(cw.Mesh.heightfield(8, 8, 4, 0.1, seed=6))
    p = tmp_path / "s.lay"
    cwb.save(p)
    data = bytearray(p.read_bytes())

    def load_bytes(b):
        q = tmp_path / "bad.lay"
        q.write_bytes(bytes(b))
        with pytest.raises(cw.CwbvhError) as e:
            cw.Layout.load(q)
        return e.value

    flipped = bytearray(data)
    flipped[100] ^= 0x40          # inside the node array
    assert "checksum" in str(load_bytes(flipped))
    assert load_bytes(data[:-1]).status == 10
    wrong_magic = bytearray(data)
    wrong_magic[0] ^= 1
    assert "not a layout" in str(load_bytes(wrong_magic))
    wrong_version = bytearray(data)
    wrong_version[8] = 99
    assert "version" in str(load_bytes(wrong_version))
