"""GPU tests of the file paths: a scene reloaded from the layout cache traces
like the built one, and the reference-file workload (mesh file + surfel file,
written here in the reference's formats) traces bit-identically to the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-files] {msg}", flush=True)
    sys.stdout.flush()


def bits(h):
    return np.ascontiguousarray(h, np.float32).view(np.uint32)


def test_layout_cache_scene_traces_identically(cw, orc, tmp_path):
    """A scene reloaded from the layout cache returns the oracle's hits (on
    the reloaded arrays, which equal the built ones) and the built scene's."""
    mesh = cw.Mesh.icosphere_grid(3, 3, 3, 4, 4.0, 0.5, 1.5, ground_margin=0.5, seed=8)
    diag = mesh.diagonal()
    rays = cw.rays_ao_surface(mesh, 20000, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=9)
    for lay in cw.build_layouts(mesh):
        say(f"layout cache kind={lay.kind}")
        p = tmp_path / f"k{lay.kind}.lay"
        lay.save(p)
        re = cw.Layout.load(p)
        assert np.array_equal(bits(re.nodes), bits(lay.nodes)) and np.array_equal(bits(re.woop), bits(lay.woop))
        a = cw.Scene(lay)
        b = cw.Scene(re)
        for fl in (cw.TRACE_CLOSEST, cw.TRACE_ANY):
            if re.kind == cw.KIND_CWBVH:
                ref = orc.trace_cwbvh(re.nodes, re.node_stride, re.woop, rays, any_hit=bool(fl))
            else:
                ref = orc.trace_bvh2(re.nodes, re.woop, rays, any_hit=bool(fl))
            got = b.trace_numpy(rays, fl)
            assert np.array_equal(bits(got), bits(ref)), "reloaded scene differs from the oracle"
            assert np.array_equal(bits(a.trace_numpy(rays, fl)), bits(got))


def test_reference_file_workload(cw, orc, tmp_path):
    from cwbvh_amd import workloads as W

    mesh = cw.Mesh.heightfield(64, 32, 8, 0.1, seed=10)
    mesh.write_dat(tmp_path / "mesh.dat", materials=np.arange(mesh.num_triangles) % 5)
    # surfels: area samples of the surface with their normals and a texel radius
    diag = mesh.diagonal()
    pts = cw.rays_ao_surface(mesh, 4096, 1, 0.0, 0.0, 1.0, seed=11)
    pos = np.zeros((4096, 4), np.float32)
    pos[:, :3] = pts[:, :3]
    nrm = np.zeros((4096, 4), np.float32)
    nrm[:, 2] = 1.0
    rad = np.full(4096, 2e-3 * diag, np.float32)
    cw.write_samples(tmp_path / "samplecache.dat", pos, nrm, rad, 64, 64)
    wl = W.reference_files(tmp_path / "mesh.dat", tmp_path / "samplecache.dat", rays_per_texel=32)
    assert len(wl.rays) == 4096 * 32 and np.array_equal(wl.mesh.materials, np.arange(mesh.num_triangles) % 5)
    bvh2, cwb = cw.build_layouts(wl.mesh)
    for lay in (cwb, bvh2):
        say(f"reference-file workload kind={lay.kind}")
        g = cw.Scene(lay).trace_numpy(wl.rays, cw.TRACE_CLOSEST)
        if lay.kind == cw.KIND_CWBVH:
            r = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, wl.rays)
        else:
            r = orc.trace_bvh2(lay.nodes, lay.woop, wl.rays)
        assert np.array_equal(bits(g), bits(r))
