"""Seeded workload definitions (scenes + rays) for the benchmark configurations.

C0  heightfield 256x128 cells (65,536 triangles) over [0,1]^2, 8 sine octaves
    with amplitudes 0.1*2^-k; 131,072 area-uniform points x 8 cosine AO rays,
    origin offset 1e-4*diag, tmin 0, tmax 0.1*diag.
C1  10x10x10 randomly rotated geodesic spheres (frequency 7 = 980 triangles,
    radius U[0.5,1.5], pitch 4) + a 2-triangle ground square (980,002
    triangles); 1920x1080 pinhole primaries, 16 cosine AO rays per primary hit
    (offset 1e-4*diag, tmax 0.05*diag).
C2  multi-scene batch, C3 clustered soup, C4 serving scene set (see the
    respective functions).

"diag" is the diagonal of the mesh bounds.  Generation is deterministic for
a seed (host C++ generators); the AO rays of C1 depend on the primary hits,
which every traversal path returns identically.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from . import (KIND_CWBVH, Layout, Mesh, Scene, TRACE_CLOSEST, build_layouts, default_build_params, hit_ids,
               rays_ao_from_hits, rays_ao_from_samples, rays_ao_surface, rays_pinhole, rays_sphere_from_surface,
               read_samples)

SEEDS = {"C0": 0xC0C0_0001, "C1": 0xC1C1_0001, "C2": 0xC2C2_0001, "C3": 0xC3C3_0001, "C4": 0xC4C4_0001}

C1_CAMERA = dict(eye=(18.0, -20.0, 60.0), target=(18.0, 20.0, -3.5), up=(0.0, 0.0, 1.0), fov=40.0)


@dataclass
class Workload:
    name: str
    mesh: Mesh
    rays: np.ndarray                     # (n, 8) float32
    diag: float
    seed: int
    description: str
    extra: dict = field(default_factory=dict)


def reference_files(mesh_path, samples_path, rays_per_texel: int = 32, seed: int = 0) -> Workload:
    """The reference's own workload (src/main.cpp:31-73): the mesh file and
    rays_per_texel AO rays per surfel of the sample file (tmin 0.01, tmax
    1e20), on the library's random stream."""
    m = Mesh.read_dat(mesh_path)
    pos, nrm, rad, (sx, sy) = read_samples(samples_path)
    rays = rays_ao_from_samples(pos, nrm, rad, rays_per_texel, 0.01, 1e20, seed)
    return Workload("files", m, rays, m.diagonal(), seed,
                    f"reference files: {m.num_triangles} tris, {sx}x{sy} surfels x {rays_per_texel} AO rays",
                    extra={"surfels": sx * sy})


def c0_mesh(seed: int = SEEDS["C0"]) -> Mesh:
    return Mesh.heightfield(256, 128, 8, 0.1, seed=seed)


def c0(num_points: int = 131072, rays_per_point: int = 8, seed: int = SEEDS["C0"]) -> Workload:
    m = c0_mesh(seed)
    diag = m.diagonal()
    rays = rays_ao_surface(m, num_points, rays_per_point, 1e-4 * diag, 0.0, 0.1 * diag, seed=seed + 1)
    return Workload("C0", m, rays, diag, seed,
                    f"heightfield 256x128 ({m.num_triangles} tris), {len(rays)} AO rays "
                    f"({num_points} points x {rays_per_point}), tmax=0.1*diag")


def c1_mesh(seed: int = SEEDS["C1"]) -> Mesh:
    return Mesh.icosphere_grid(10, 10, 10, 7, 4.0, 0.5, 1.5, ground_margin=1.0, seed=seed)


def c1_primary(width: int = 1920, height: int = 1080) -> np.ndarray:
    c = C1_CAMERA
    return rays_pinhole(width, height, c["eye"], c["target"], c["up"], c["fov"], 0.0, 1e30)


def c1(trace_primary: Callable[[Mesh, np.ndarray], tuple[np.ndarray, np.ndarray]], width: int = 1920,
       height: int = 1080, rays_per_hit: int = 16, seed: int = SEEDS["C1"], mesh: Mesh | None = None) -> Workload:
    """trace_primary(mesh, rays) -> (hits (n,4), prim ids (n,)) closest hits of the primaries."""
    m = mesh if mesh is not None else c1_mesh(seed)
    diag = m.diagonal()
    prim = c1_primary(width, height)
    hits, pids = trace_primary(m, prim)
    ao = rays_ao_from_hits(m, prim, hits, pids, rays_per_hit, 1e-4 * diag, 0.05 * diag, seed=seed + 1)
    return Workload("C1", m, ao, diag, seed,
                    f"10^3 icospheres f7 + ground ({m.num_triangles} tris), {width}x{height} primaries, "
                    f"{len(ao)} AO rays ({rays_per_hit}/hit), tmax=0.05*diag",
                    extra={"primary": prim, "primary_hit_rate": float((pids >= 0).mean())})


def gpu_primary_tracer(params=None):
    """trace_primary for c1() using the GPU CWBVH path (closest hit, prim ids)."""

    def run(mesh: Mesh, rays: np.ndarray):
        _, lay = build_layouts(mesh, params, bvh2=False, cwbvh=True)
        scene = Scene(lay)
        hits = scene.trace_numpy(rays, TRACE_CLOSEST)
        ids = hit_ids(hits)
        pidx = lay.prim_idx
        pids = np.where(ids >= 0, pidx[np.clip(ids, 0, None)], -1).astype(np.int32)
        return hits, pids

    return run


# ------------------------------------------------------------------- C2


def heightfield_of(ntri: int, seed: int) -> Mesh:
    cells = max(1, ntri // 2)
    ny = max(1, int(math.sqrt(cells / 2)))
    nx = max(1, cells // ny)
    return Mesh.heightfield(nx, ny, 8, 0.1, seed=seed)


def icospheres_of(ntri: int, seed: int) -> Mesh:
    per = 980  # frequency 7
    g = max(1, round((ntri / per) ** (1.0 / 3.0)))
    return Mesh.icosphere_grid(g, g, g, 7, 4.0, 0.5, 1.5, ground_margin=0.25, seed=seed)


def c2_plan(num_scenes: int = 256, tri_range=(1_000, 2_000_000), ray_range=(4_096, 4_194_304),
            seed: int = SEEDS["C2"]) -> list[dict]:
    """Scene/ray-batch plan: log-uniform triangle and ray counts, alternating generators."""
    rng = np.random.default_rng(seed)
    plan = []
    for i in range(num_scenes):
        ntri = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        nray = int(math.exp(rng.uniform(math.log(ray_range[0]), math.log(ray_range[1]))))
        kind = "heightfield" if i % 2 == 0 else "icospheres"
        plan.append({"index": i, "kind": kind, "target_tris": ntri, "rays": nray, "seed": seed + 7919 * (i + 1)})
    return plan


def c2_scene(entry: dict) -> tuple[Mesh, np.ndarray, float]:
    m = heightfield_of(entry["target_tris"], entry["seed"]) if entry["kind"] == "heightfield" \
        else icospheres_of(entry["target_tris"], entry["seed"])
    diag = m.diagonal()
    n = entry["rays"]
    rays = rays_ao_surface(m, (n + 7) // 8, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=entry["seed"] + 1)[:n]
    return m, rays, diag


# ------------------------------------------------------------------- C3


def c3_mesh(num_triangles: int = 20_000_000, seed: int = SEEDS["C3"]) -> Mesh:
    return Mesh.cluster_soup(num_triangles, 4096, 0.005, 0.05, math.log(1e-3), 0.5, seed=seed)


def c3(num_triangles: int = 20_000_000, num_rays: int = 64 * 1024 * 1024, seed: int = SEEDS["C3"]) -> Workload:
    m = c3_mesh(num_triangles, seed)
    diag = m.diagonal()
    rays = rays_sphere_from_surface(m, num_rays, 1e-4 * diag, 0.0, np.inf, seed=seed + 1)
    return Workload("C3", m, rays, diag, seed,
                    f"cluster soup {m.num_triangles} tris / 4096 clusters, {len(rays)} diffuse rays, tmax=inf")


# ------------------------------------------------------------------- C4


def c4_plan(num_scenes: int = 32, total_tris: int = 100_000_000, seed: int = SEEDS["C4"]) -> list[dict]:
    """Resident scene set: heightfields, icosphere lattices and clustered soups
    in rotation, sizes log-uniform around total_tris / num_scenes."""
    rng = np.random.default_rng(seed)
    raw = np.exp(rng.uniform(np.log(0.25), np.log(4.0), size=num_scenes))
    sizes = (raw / raw.sum() * total_tris).astype(np.int64)
    kinds = ("heightfield", "icospheres", "soup")
    return [{"index": i, "kind": kinds[i % 3], "target_tris": int(max(1000, sizes[i])), "seed": seed + 104729 * (i + 1)}
            for i in range(num_scenes)]


def c4_mesh(entry: dict) -> Mesh:
    if entry["kind"] == "heightfield":
        return heightfield_of(entry["target_tris"], entry["seed"])
    if entry["kind"] == "icospheres":
        return icospheres_of(entry["target_tris"], entry["seed"])
    return Mesh.cluster_soup(entry["target_tris"], 512, 0.005, 0.05, math.log(1e-3), 0.5, seed=entry["seed"])
