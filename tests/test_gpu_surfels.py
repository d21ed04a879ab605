"""GPU tests of AO from surfels: the device generator returns the host
generator's rays bit for bit, and the fused generate-and-trace kernel returns
what the CPU oracle and the plain kernel return for those rays (hits, or
per-surfel hit counts)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-surfels] {msg}", flush=True)
    sys.stdout.flush()


def bits(h):
    return np.ascontiguousarray(h, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def surfels(cw):
    mesh = cw.Mesh.icosphere_grid(4, 4, 2, 5, 4.0, 0.5, 1.5, ground_margin=0.5, seed=21)
    diag = mesh.diagonal()
    # surfels on the surface: positions and normals from area samples
    pts = cw.rays_ao_surface(mesh, 6000, 1, 0.0, 0.0, 1.0, seed=22)
    rng = np.random.default_rng(23)
    pos = np.zeros((6000, 4), np.float32)
    pos[:, :3] = pts[:, :3]
    nrm = np.zeros((6000, 4), np.float32)
    n = rng.normal(size=(6000, 3))
    n[:300] = [0, 0, 1]          # frame fallback (n x z too short)
    n[300:600] = [0, 0, -1]
    nrm[:, :3] = n / np.linalg.norm(n, axis=1, keepdims=True)
    rad = rng.uniform(1e-4, 1e-3, size=6000).astype(np.float32) * np.float32(diag)
    _, cwb = cw.build_layouts(mesh, bvh2=False)
    return mesh, cwb, pos, nrm, rad, diag


def test_device_generator_matches_host(cw, surfels):
    mesh, cwb, pos, nrm, rad, diag = surfels
    k = 24
    host = cw.rays_ao_from_samples(pos, nrm, rad, k, 0.0, np.float32(0.05 * diag), seed=5)
    bufs = [cw.DeviceBuffer.from_numpy(a) for a in (pos, nrm, rad)]
    out = cw.DeviceBuffer(32 * len(host))
    say(f"device generator {len(host)} rays")
    cw.rays_ao_from_samples_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, len(pos), k, 0.0,
                                   np.float32(0.05 * diag), 5, out.ptr)
    cw.synchronize()
    dev = out.download((len(host), 8), np.float32)
    assert np.array_equal(bits(dev), bits(host))


@pytest.mark.parametrize("kw", [dict(), dict(loop_mode=0), dict(loop_mode=2, postpone_pct=30),
                                dict(persistent=0), dict(persistent=1, fetch_chunk=7, block_size=64),
                                dict(occupancy=7), dict(occupancy=7, persistent=0, block_size=256)])
def test_fused_trace_equals_trace_of_generated_rays(cw, orc, surfels, kw):
    mesh, cwb, pos, nrm, rad, diag = surfels
    k = 16
    tmax = np.float32(0.05 * diag)
    rays = cw.rays_ao_from_samples(pos, nrm, rad, k, 0.0, tmax, seed=9)
    sc = cw.Scene(cwb)
    cfg = cw.default_trace_config(cw.KIND_CWBVH, **kw)
    bufs = [cw.DeviceBuffer.from_numpy(a) for a in (pos, nrm, rad)]
    n = len(rays)
    dh = cw.DeviceBuffer(16 * n)
    dc = cw.DeviceBuffer(4 * len(pos))
    for flags in (cw.TRACE_CLOSEST, cw.TRACE_ANY):
        say(f"fused {kw} flags={flags}")
        ref = sc.trace_numpy(rays, flags, cfg)
        oref = orc.trace_cwbvh(cwb.nodes, cwb.node_stride, cwb.woop, rays, any_hit=bool(flags))
        sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, len(pos), k, 0.0, tmax, 9, flags, d_hits=dh.ptr,
                         cfg=cfg)
        cw.synchronize()
        got = dh.download((n, 4), np.float32)
        if flags == cw.TRACE_ANY and kw.get("loop_mode") == 2:
            assert np.array_equal(cw.hit_ids(got) >= 0, cw.hit_ids(ref) >= 0)
            assert np.array_equal(cw.hit_ids(got) >= 0, cw.hit_ids(oref) >= 0)
        else:
            assert np.array_equal(bits(got), bits(ref))
            assert np.array_equal(bits(got), bits(oref)), "fused hits differ from the oracle"
        # per-surfel counts of rays that hit
        dc.zero()
        sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, len(pos), k, 0.0, tmax, 9, flags,
                         d_texel_hits=dc.ptr, cfg=cfg)
        cw.synchronize()
        counts = dc.download((len(pos),), np.uint32)
        want = (cw.hit_ids(oref) >= 0).reshape(len(pos), k).sum(axis=1)
        assert np.array_equal(counts, want.astype(np.uint32))
        assert 0 < counts.sum() < n


def test_fused_trace_rejects_bad_calls(cw, surfels):
    mesh, cwb, pos, nrm, rad, diag = surfels
    bvh2, _ = cw.build_layouts(mesh, cwbvh=False)
    bufs = [cw.DeviceBuffer.from_numpy(a) for a in (pos, nrm, rad)]
    dh = cw.DeviceBuffer(16 * len(pos))
    dc = cw.DeviceBuffer(4 * len(pos))
    args = (bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, len(pos), 1, 0.0, 1.0, 1)
    with pytest.raises(cw.CwbvhError):
        cw.Scene(bvh2).trace_surfels(*args, d_hits=dh.ptr)
    sc = cw.Scene(cwb)
    with pytest.raises(cw.CwbvhError):
        sc.trace_surfels(*args, d_hits=dh.ptr, d_texel_hits=dc.ptr)
    with pytest.raises(cw.CwbvhError):
        sc.trace_surfels(*args)
    with pytest.raises(cw.CwbvhError):
        sc.trace_surfels(*args, d_hits=dh.ptr, cfg=cw.default_trace_config(cw.KIND_CWBVH, prefetch=1))
