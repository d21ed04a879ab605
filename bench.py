#!/usr/bin/env python3
"""Benchmark: AO ray traversal Mrays/s on one MI355X (CWBVH vs BVH2).

Default workload is C1: 980,002-triangle icosphere lattice, ~33M AO rays
(16 per primary hit of a 1920x1080 camera).  One "step" traverses the whole
AO batch twice with the CWBVH kernel: closest hit, then any hit.  Rays, BVH
and hit buffers are resident in HBM before the timed region; the timed region
holds only kernel launches on one stream, bracketed by HIP events.

Prints ONE JSON line.  `value` = CWBVH AO rays traversed per second over the
step (closest + any), in Mrays/s.  Extra keys report each mode separately, the
BVH2 kernel on the same rays, the roofline of the dominant kernel (CWBVH
closest hit) and the CPU baseline (scalar CPU restatement on host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s AO any-hit + closest-hit, CWBVH vs BVH2, 1 GPU; serving p99 ms"
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = ROOT / "profiles" / "cwbvh_closest_c1_traffic.json"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(kind: int, n: int, nodes: int, tris: int, leaves: int = 0) -> int:
    """Reference-layout bytes: ray 32 + hit 16 per ray, 80 B per CWBVH node or
    64 B per BVH2 node, 48 B per triangle test, 16 B per BVH2 leaf terminator."""
    if kind == 1:
        return 48 * n + 80 * nodes + 48 * tris
    return 48 * n + 64 * nodes + 48 * tris + 16 * leaves


def time_launches(fns, steps: int, warmup: int, cw):
    """Run fns (list of callables) per step; per-launch times from HIP events."""
    for _ in range(warmup):
        for f in fns:
            f()
    cw.synchronize()
    evs = [[cw.Event() for _ in range(len(fns) + 1)] for _ in range(steps)]
    for s in range(steps):
        evs[s][0].record()
        for i, f in enumerate(fns):
            f()
            evs[s][i + 1].record()
    evs[-1][-1].synchronize()
    per = np.array([[evs[s][i].elapsed_ms(evs[s][i + 1]) for i in range(len(fns))] for s in range(steps)])
    total_ms = evs[0][0].elapsed_ms(evs[-1][-1])
    return total_ms, per


def cpu_baseline(layout, rays, target_s: float, threads: int) -> dict:
    import oracle

    nodes, woop = layout.nodes, layout.woop
    n0 = min(len(rays), 262144)
    t0 = time.perf_counter()
    oracle.trace_cwbvh(nodes, layout.node_stride, woop, rays[:n0], threads=threads)
    dt = time.perf_counter() - t0
    n = int(min(len(rays), max(n0, n0 * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    oracle.trace_cwbvh(nodes, layout.node_stride, woop, rays[:n], threads=threads)
    dt = time.perf_counter() - t0
    # one host thread on a small sample, for the per-core figure
    n1 = min(len(rays), 65536)
    t1 = time.perf_counter()
    oracle.trace_cwbvh(nodes, layout.node_stride, woop, rays[:n1], threads=1)
    dt1 = time.perf_counter() - t1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {n} AO rays of the workload, CWBVH closest hit, scalar C restatement "
                      f"(oracle/oracle.c) on {threads} host threads, {dt:.1f} s wall",
            "single_thread_mrays_s": round(n1 / dt1 / 1e6, 3), "host_cpu": model,
            "host_threads": os.cpu_count()}


def bench_c2(args, cw, W, dev) -> int:
    """C2: many scenes, uneven AO batches.  Plain = one cwbvh_trace launch per
    scene (default configuration) back to back on one stream; batched = one
    autotuned persistent multi-scene launch replayed from a hipGraph (the
    Morton pre-pass, when the tuner picks it, is inside the timed region)."""
    from concurrent.futures import ThreadPoolExecutor

    plan = W.c2_plan(num_scenes=args.c2_scenes)
    t0 = time.perf_counter()

    def make(entry):
        mesh, rays, _ = W.c2_scene(entry)
        _, lay = cw.build_layouts(mesh, cw.default_build_params(num_threads=2), bvh2=False)
        return entry, mesh.num_triangles, lay, rays

    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1) // 2)) as ex:
        built = list(ex.map(make, plan))
    build_s = time.perf_counter() - t0
    tris = sum(b[1] for b in built)
    log(f"C2: {len(built)} scenes, {tris} triangles, {sum(len(b[3]) for b in built)} rays, build {build_s:.1f} s")
    scenes, bufs, items = [], [], []
    for entry, nt, lay, rays in built:
        sc = cw.Scene(lay)
        dr = cw.DeviceBuffer.from_numpy(rays)
        dh = cw.DeviceBuffer(16 * len(rays))
        scenes.append(sc)
        bufs.append((dr, dh))
        items.append((sc, dr.ptr, dh.ptr, len(rays)))
    del built
    total = sum(it[3] for it in items)
    st = cw.Stream()
    cfg0 = cw.default_trace_config(cw.KIND_CWBVH)

    def plain():
        for sc, r, h, n in items:
            sc.trace(r, h, n, 0, None, None, st)

    def plain_sync():
        for sc, r, h, n in items:
            sc.trace(r, h, n, 0, None, None, st)
            st.synchronize()

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        st.synchronize()
        e0, e1 = cw.Event(), cw.Event()
        e0.record(st)
        for _ in range(steps):
            fn()
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_ms(e1) / steps

    ms_plain = timed(plain, args.steps, args.warmup)
    t = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        plain_sync()
    ms_plain_sync = (time.perf_counter() - t) * 1e3 / max(1, args.steps // 2)
    log(f"plain {ms_plain:.2f} ms, plain+sync {ms_plain_sync:.2f} ms")
    t = time.perf_counter()
    best, recs = cw.autotune_batch(items, block_sizes=(64, 128, 256), lds_depths=(8, 16, 32), morton=(0, 1),
                                   prefetch=(0, 1), loop_modes=(0, 1, 2), reps=args.tune_reps, stream=st)
    tune_s = time.perf_counter() - t
    log(f"autotune {tune_s:.1f} s best {best.as_dict()}")
    batch = cw.Batch(items, cw.TRACE_CLOSEST, best.cfg, morton=bool(best.morton))
    batch.capture(st)
    ms_batch = timed(lambda: batch.replay(st), args.steps, args.warmup)
    untuned = cw.Batch(items, cw.TRACE_CLOSEST, cfg0, morton=False)
    ms_untuned = timed(lambda: untuned.launch(st), args.steps, args.warmup)
    value = total / (ms_batch / 1e3) / 1e6
    # algorithmic traffic of the batch (counting launches per scene, outside
    # the timed region): every scene's rays, hits, node visits and triangle
    # tests in reference-layout bytes
    d_cnt = cw.DeviceBuffer(16)
    nodes_c = tris_c = 0
    for sc, r, h, n in items:
        d_cnt.zero()
        sc.trace(r, h, n, cw.TRACE_STATS, cfg0, d_cnt.ptr, st)
        st.synchronize()
        c = d_cnt.download((2,), np.uint64)
        nodes_c += int(c[0])
        tris_c += int(c[1])
    bytes_c = algorithmic_bytes(1, total, nodes_c, tris_c)
    achieved = bytes_c / (ms_batch / 1e3) / 1e9
    out = {
        "metric": METRIC + " [C2 batched]",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batch, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C2: {len(items)} seeded scenes (heightfields / icosphere lattices), "
                               f"{tris} triangles, {total} AO rays (closest hit)",
                   "triangles": tris, "rays": total, "scenes": len(items), "seed": W.SEEDS["C2"]},
        "plain_ms": round(ms_plain, 4),
        "plain_mrays_s": round(total / ms_plain / 1e3, 2),
        "plain_sync_ms": round(ms_plain_sync, 4),
        "batched_untuned_ms": round(ms_untuned, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "cwbvh_trace<closest, MULTI> (autotuned persistent batch, graph replay)",
                     "algorithmic_bytes_per_launch": bytes_c, "avg_launch_ms": round(ms_batch, 4),
                     "per_ray": {"nodes": nodes_c / total, "triangles": tris_c / total, "bytes": bytes_c / total},
                     "plain_achieved": round(bytes_c / (ms_plain / 1e3) / 1e9, 1)},
        "speedup_vs_plain": round(ms_plain / ms_batch, 3),
        "speedup_vs_plain_sync": round(ms_plain_sync / ms_batch, 3),
        "autotune": {"best": best.as_dict(), "seconds": round(tune_s, 1),
                     "records": [r.as_dict() for r in recs]},
        "build_s": round(build_s, 1),
        "device": dev["arch"],
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_c3(args, cw, W, dev) -> int:
    """C3: large incoherent scene, closest hit; rays shuffled vs Morton-ordered
    (Morton order applied through the ordered-trace indirection; the sort is
    timed separately)."""
    t0 = time.perf_counter()
    mesh = W.c3_mesh(args.c3_tris)
    diag = mesh.diagonal()
    rays = cw.rays_sphere_from_surface(mesh, args.c3_rays, 1e-4 * diag, 0.0, np.inf, seed=W.SEEDS["C3"] + 1)
    rays = cw.shuffle_rays(rays, seed=W.SEEDS["C3"] + 2)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    bvh2, cwb = cw.build_layouts(mesh)
    build_s = time.perf_counter() - t0
    n = len(rays)
    log(f"C3: {mesh.num_triangles} tris, {n} rays, gen {gen_s:.1f} s, build {build_s:.1f} s, "
        f"cwbvh {cwb.stats()}, bvh2 {bvh2.stats()}")
    # the library's default upload (a scene this large gets 128 B nodes: one
    # cache line per node), plus the 80 B reference layout for the table below
    s_cw, s_b2 = cw.Scene(cwb), cw.Scene(bvh2)
    s_cw80 = s_cw if s_cw.node_stride == 5 else cw.Scene(cwb, stride_mode=cw.SCENE_STRIDE_KEEP)
    s_cw8 = s_cw if s_cw.node_stride == 8 else cw.Scene(cwb, stride_mode=cw.SCENE_STRIDE_128)
    log(f"default upload: {s_cw.node_stride * 16} B nodes")
    lo, hi = s_cw.bounds()
    del bvh2
    dr = cw.DeviceBuffer.from_numpy(rays)
    del rays
    dh = cw.DeviceBuffer(16 * n)
    do = cw.DeviceBuffer(4 * n)
    e0, e1 = cw.Event(), cw.Event()
    # the first call sizes the library's per-stream sort workspace; the
    # timed one runs the key kernel and the radix sort alone
    cw.morton_order(dr.ptr, n, lo, hi, do.ptr)
    e0.record()
    cw.morton_order(dr.ptr, n, lo, hi, do.ptr)
    e1.record()
    e1.synchronize()
    sort_ms = e0.elapsed_ms(e1)
    # the library's configuration for each scene: at this size it sorts a
    # launch's rays by origin Morton code before traversing them (the sort
    # runs inside the timed launch)
    cfg_cw, cfg_b2 = s_cw.default_config(), s_b2.default_config()
    given_cw, given_b2 = s_cw.default_config(ray_order=0), s_b2.default_config(ray_order=0)

    def t(sc, cfg, order):
        fn = lambda: sc.trace(dr.ptr, dh.ptr, n, 0, cfg, None, None, order)
        total, per = time_launches([fn], args.steps, args.warmup, cw)
        return float(per[:, 0].mean())

    res = {}
    for name, sc, cfg, given in (("cwbvh", s_cw, cfg_cw, given_cw), ("bvh2", s_b2, cfg_b2, given_b2)):
        # default: shuffled rays, library default (sorted inside the launch);
        # shuffled: traced in the given order; morton: pre-ordered, sort not counted
        for oname, c, order in (("default", cfg, None), ("shuffled", given, None), ("morton", given, do.ptr)):
            ms = t(sc, c, order)
            res[f"{name}_{oname}"] = {"ms": round(ms, 3), "mrays_per_s": round(n / ms / 1e3, 1)}
            log(f"{name} {oname}: {ms:.2f} ms {n / ms / 1e3:.0f} Mrays/s")
    log(f"morton sort {sort_ms:.2f} ms")
    # the reference's 80 B nodes with their scene default, as rtTraceCWBVH
    # traces caller arrays bound in that layout
    ms80 = t(s_cw80, s_cw80.default_config(), None)
    res["cwbvh_node80_default"] = {"ms": round(ms80, 3), "mrays_per_s": round(n / ms80 / 1e3, 1)}
    log(f"cwbvh 80 B nodes, scene default: {ms80:.2f} ms {n / ms80 / 1e3:.0f} Mrays/s")
    # node layout x inner-loop schedule: this scene is far larger than L2 and
    # the Infinity Cache, so line-aligned nodes and access order matter
    layouts = []
    for lname, sc in (("node80", s_cw80), ("node128", s_cw8)):
        for mode in (0, 1, 2):
            cfg = cw.default_trace_config(cw.KIND_CWBVH, loop_mode=mode)
            for oname, order in (("shuffled", None), ("morton", do.ptr)):
                ms = t(sc, cfg, order)
                layouts.append({"nodes": lname, "loop_mode": mode, "order": oname, "ms": round(ms, 3),
                                "mrays_per_s": round(n / ms / 1e3, 1)})
                log(f"cwbvh {lname} loop_mode {mode} {oname}: {ms:.2f} ms {n / ms / 1e3:.0f} Mrays/s")
    d_cnt = cw.DeviceBuffer(16)
    d_cnt.zero()
    s_cw.trace(dr.ptr, dh.ptr, n, cw.TRACE_STATS, cfg_cw, d_cnt.ptr)
    cw.synchronize()
    c = d_cnt.download((2,), np.uint64)
    bytes_c = algorithmic_bytes(1, n, int(c[0]), int(c[1]))
    ms_c = res["cwbvh_default"]["ms"]
    achieved = bytes_c / (ms_c / 1e3) / 1e9
    value = n / ms_c / 1e3
    out = {
        "metric": METRIC + " [C3 large incoherent]",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_c,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C3: clustered soup {mesh.num_triangles} triangles (4096 clusters), {n} diffuse "
                               f"rays from surface points, tmax=inf, closest hit, shuffled input, library default "
                               f"configuration (Morton pre-pass inside the launch)",
                   "triangles": mesh.num_triangles, "rays": n, "seed": W.SEEDS["C3"],
                   "node_bytes": 16 * s_cw.node_stride},
        "results": res,
        "cwbvh_layout_schedule": layouts,
        "best_cwbvh_shuffled": min((r for r in layouts if r["order"] == "shuffled"), key=lambda r: r["ms"]),
        "morton_sort_ms": round(sort_ms, 3),
        "cwbvh_vs_bvh2": {"default": round(res["bvh2_default"]["ms"] / res["cwbvh_default"]["ms"], 3),
                          "shuffled": round(res["bvh2_shuffled"]["ms"] / res["cwbvh_shuffled"]["ms"], 3),
                          "morton": round(res["bvh2_morton"]["ms"] / res["cwbvh_morton"]["ms"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "per_ray": {"nodes": int(c[0]) / n, "triangles": int(c[1]) / n, "bytes": bytes_c / n}},
        "build_s": round(build_s, 1),
        "device": dev["arch"],
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_ao(args, cw, W, dev) -> int:
    """AO from surfels on the C1 scene: the primary hits of the 1920x1080
    camera become surfels (hit point, geometric normal facing the camera,
    texel radius 2e-4 * diag), 16 AO rays each (tmin 0, tmax 0.05 * diag).
    Three ways to run the same rays, closest and any hit:
      unfused      ray generation kernel writes rays, then cwbvh_trace;
      fused_hits   cwbvh_trace_surfels generates rays in the traversal kernel;
      fused_counts as fused_hits, with per-surfel hit counts instead of hits."""
    mesh = W.c1_mesh()
    diag = mesh.diagonal()
    prim = W.c1_primary()
    hits, pids = W.gpu_primary_tracer()(mesh, prim)
    sel = np.nonzero(pids >= 0)[0]
    v, idx = mesh.vertices, mesh.indices
    tri = v[idx[pids[sel]]].astype(np.float64)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    dirs = prim[sel, 4:7].astype(np.float64)
    nrm[np.einsum("ij,ij->i", nrm, dirs) > 0] *= -1
    pos = np.zeros((len(sel), 4), np.float32)
    pos[:, :3] = prim[sel, :3] + hits[sel, :1] * prim[sel, 4:7]
    nrm4 = np.zeros((len(sel), 4), np.float32)
    nrm4[:, :3] = nrm
    rad = np.full(len(sel), 2e-4 * diag, np.float32)
    k, tmax, seed = 16, np.float32(0.05 * diag), W.SEEDS["C1"] + 7
    ntex, n = len(sel), len(sel) * 16
    _, cwb = cw.build_layouts(mesh, bvh2=False)
    sc = cw.Scene(cwb)
    bufs = [cw.DeviceBuffer.from_numpy(a) for a in (pos, nrm4, rad)]
    d_rays = cw.DeviceBuffer(32 * n)
    d_hits = cw.DeviceBuffer(16 * n)
    d_counts = cw.DeviceBuffer(4 * ntex)
    cfg = sc.default_config()
    log(f"AO surfels: {ntex} surfels x {k} = {n} rays")

    def unfused(fl):
        def f():
            cw.rays_ao_from_samples_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, d_rays.ptr)
            sc.trace(d_rays.ptr, d_hits.ptr, n, fl, cfg)
        return f

    def fused_hits(fl):
        return lambda: sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, fl,
                                        d_hits=d_hits.ptr, cfg=cfg)

    def fused_counts(fl):
        def f():
            d_counts.zero()
            sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, fl,
                             d_texel_hits=d_counts.ptr, cfg=cfg)
        return f

    table = {}
    for name, mk in (("unfused", unfused), ("fused_hits", fused_hits), ("fused_counts", fused_counts)):
        for fname, fl in (("closest", cw.TRACE_CLOSEST), ("any", cw.TRACE_ANY)):
            _, per = time_launches([mk(fl)], args.steps, args.warmup, cw)
            ms = float(per[:, 0].mean())
            table[f"{name}_{fname}"] = {"ms": round(ms, 4), "mrays_per_s": round(n / ms / 1e3, 1)}
            log(f"{name} {fname}: {ms:.3f} ms {n / ms / 1e3:.0f} Mrays/s")
    # the fused kernel returns what tracing the generated rays returns
    sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, cw.TRACE_CLOSEST,
                     d_hits=d_hits.ptr, cfg=cfg)
    cw.synchronize()
    fused = d_hits.download((n, 4), np.float32)
    cw.rays_ao_from_samples_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, d_rays.ptr)
    sc.trace(d_rays.ptr, d_hits.ptr, n, cw.TRACE_CLOSEST, cfg)
    cw.synchronize()
    identical = bool(np.array_equal(fused.view(np.uint32), d_hits.download((n, 4), np.float32).view(np.uint32)))
    best = max(table.items(), key=lambda kv: kv[1]["mrays_per_s"])
    # algorithmic traffic of the reported kernel (fused, per-surfel counts,
    # any hit): 36 B of surfel read and 4 B of count per surfel, plus the node
    # visits and triangle tests of the same rays traced as a buffer
    d_cnt = cw.DeviceBuffer(16)
    d_cnt.zero()
    sc.trace(d_rays.ptr, d_hits.ptr, n, cw.TRACE_ANY | cw.TRACE_STATS, cfg, d_cnt.ptr)
    cw.synchronize()
    c = d_cnt.download((2,), np.uint64)
    bytes_c = 40 * ntex + 80 * int(c[0]) + 48 * int(c[1])
    ms_v = table["fused_counts_any"]["ms"]
    achieved = bytes_c / (ms_v / 1e3) / 1e9
    out = {
        "metric": METRIC + " [AO from surfels, fused ray generation]",
        "value": table["fused_counts_any"]["mrays_per_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": table["fused_counts_any"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C1 scene, {ntex} surfels (primary hits) x {k} AO rays, tmax=0.05*diag",
                   "triangles": mesh.num_triangles, "surfels": ntex, "rays": n, "seed": seed},
        "results": table,
        "fused_hits_identical_to_unfused": identical,
        "speedup_fused_counts_vs_unfused_any": round(table["unfused_any"]["ms"] / table["fused_counts_any"]["ms"], 3),
        "speedup_fused_hits_vs_unfused_closest": round(table["unfused_closest"]["ms"] / table["fused_hits_closest"]["ms"], 3),
        "bytes_per_ray": {"unfused": 32 + 32 + 16, "fused_hits": 16, "fused_counts": "4 per surfel"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "cwbvh_trace<any, SRC=surfels> with per-surfel counts",
                     "algorithmic_bytes_per_launch": bytes_c, "avg_launch_ms": ms_v,
                     "per_ray": {"nodes": int(c[0]) / n, "triangles": int(c[1]) / n, "bytes": bytes_c / n}},
        "best": best[0],
        "device": dev["arch"],
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_c4(args, cw, W, dev) -> int:
    """C4: resident scenes, Poisson AO query requests (log-uniform 1k..1M rays,
    uniform scene), host-resident rays copied in and hits copied out per
    request.  Continuous batching vs serial per-request plain launches."""
    from concurrent.futures import ThreadPoolExecutor

    specs = W.c4_plan(num_scenes=args.c4_scenes, total_tris=args.c4_tris)
    t0 = time.perf_counter()

    def make(entry):
        mesh = W.c4_mesh(entry)
        _, lay = cw.build_layouts(mesh, cw.default_build_params(num_threads=4), bvh2=False)
        diag = mesh.diagonal()
        rays = cw.rays_ao_surface(mesh, 131072, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=entry["seed"] + 1)
        return entry, mesh.num_triangles, lay, rays

    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1) // 4)) as ex:
        built = list(ex.map(make, specs))
    build_s = time.perf_counter() - t0
    tris = sum(b[1] for b in built)
    scenes, pools, hbm = [], [], 0
    for entry, nt, lay, rays in built:
        sc = cw.Scene(lay)
        hbm += sc.device_bytes()
        pool = cw.PinnedArray(rays.shape, np.float32)
        pool.array[:] = rays
        scenes.append(sc)
        pools.append(pool)
    del built
    log(f"C4: {len(scenes)} resident scenes, {tris} triangles, {hbm / 2**30:.2f} GiB of BVH in HBM, "
        f"build {build_s:.1f} s")
    hit_pool = cw.PinnedArray((16 << 20, 4), np.float32)
    rates = [float(r) for r in args.c4_rates.split(",")]
    table = []
    for serial in (0, 1):
        cfg = cw.default_server_config(serial=serial, memory_budget=8 << 30, max_inflight=3,
                                       max_batch_requests=64, max_batch_rays=16 << 20)
        srv = cw.Server(scenes, cfg)
        srv.run_poisson(2000.0, 20, 1000, 1_000_000, pools, hit_pool, seed=1)  # warm-up
        for rate in rates:
            res, _ = srv.run_poisson(rate, args.c4_requests, 1000, 1_000_000, pools, hit_pool, seed=int(rate))
            st = srv.stats()
            row = {"mode": "serial" if serial else "continuous", "rate_req_s": rate,
                   "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
                   "mean_ms": round(res["mean_ms"], 3), "mrays_per_s": round(res["mrays_per_s"], 1),
                   "wall_s": round(res["wall_s"], 3), "max_inflight_seen": st["max_inflight_seen"],
                   "max_batch_requests_seen": st["max_batch_requests_seen"]}
            log(str(row))
            table.append(row)
        srv.close()
    cont = [r for r in table if r["mode"] == "continuous"]
    ser = [r for r in table if r["mode"] == "serial"]
    out = {
        "metric": METRIC + " [C4 serving]",
        "value": cont[0]["p99_ms"],
        "unit": "ms (p99 latency at the lowest rate)",
        "n_gpus": 1,
        "steps": args.c4_requests,
        "warmup": 20,
        "ms_per_step": cont[0]["mean_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C4: {len(scenes)} resident scenes ({tris} triangles), Poisson requests of "
                               f"log-uniform 1k..1M AO rays, host rays in / hits out (PCIe included)",
                   "triangles": tris, "scenes": len(scenes), "requests_per_rate": args.c4_requests,
                   "rates": rates, "seed": W.SEEDS["C4"]},
        "table": table,
        "max_mrays_per_s": {"continuous": max(r["mrays_per_s"] for r in cont),
                            "serial": max(r["mrays_per_s"] for r in ser)},
        "bvh_hbm_gib": round(hbm / 2**30, 2),
        "build_s": round(build_s, 1),
        "device": dev["arch"],
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c2", "c3", "c4", "files", "ao"])
    ap.add_argument("--mesh-dat", help="--config files: the reference's mesh file (src/main.cpp:31-45)")
    ap.add_argument("--samples-dat", help="--config files: the reference's surfel file (src/main.cpp:52-64)")
    ap.add_argument("--rays-per-texel", type=int, default=32)
    ap.add_argument("--c3-tris", type=int, default=20_000_000)
    ap.add_argument("--c3-rays", type=int, default=64 * 1024 * 1024)
    ap.add_argument("--c4-scenes", type=int, default=32)
    ap.add_argument("--c4-tris", type=int, default=100_000_000, help="total triangles of the C4 scene set")
    ap.add_argument("--c4-requests", type=int, default=2000)
    ap.add_argument("--c4-rates", default="1000,2000,5000,10000,20000,50000")
    ap.add_argument("--c2-scenes", type=int, default=256)
    ap.add_argument("--tune-reps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--node-stride", type=int, default=5, choices=[5, 8])
    ap.add_argument("--skip-bvh2", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): rays are independent, so N GPUs "
              "would each run this same single-GPU workload on 1/N of the rays with a replicated BVH; "
              "multi-GPU runs are not part of this build.", file=sys.stderr)
        return 2

    import cwbvh_amd as cw
    from cwbvh_amd import workloads as W

    if cw.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 3
    dev = cw.device_info(0)
    log(f"device {dev['arch']} CUs={dev['compute_units']}")

    if args.config == "c2":
        return bench_c2(args, cw, W, dev)
    if args.config == "c4":
        return bench_c4(args, cw, W, dev)
    if args.config == "c3":
        return bench_c3(args, cw, W, dev)
    if args.config == "ao":
        return bench_ao(args, cw, W, dev)

    params = cw.default_build_params(cwbvh_node_stride=args.node_stride)
    t0 = time.perf_counter()
    if args.config == "c0":
        wl = W.c0()
    elif args.config == "files":
        if not (args.mesh_dat and args.samples_dat):
            print("--config files needs --mesh-dat and --samples-dat", file=sys.stderr)
            return 2
        wl = W.reference_files(args.mesh_dat, args.samples_dat, args.rays_per_texel)
    else:
        wl = W.c1(W.gpu_primary_tracer(params))
    gen_s = time.perf_counter() - t0
    rays = wl.rays
    n = len(rays)
    log(f"workload {wl.name}: {wl.description} ({gen_s:.1f} s)")

    t0 = time.perf_counter()
    bvh2, cwb = cw.build_layouts(wl.mesh, params)
    build_s = time.perf_counter() - t0
    log(f"build {build_s:.2f} s: cwbvh {cwb.stats()} bvh2 {bvh2.stats()}")

    s_cw = cw.Scene(cwb)
    s_b2 = cw.Scene(bvh2)
    d_rays = cw.DeviceBuffer.from_numpy(rays)
    d_hits = cw.DeviceBuffer(16 * n)
    d_cnt = cw.DeviceBuffer(16)
    cfg_cw = s_cw.default_config()
    cfg_b2 = cw.default_trace_config(cw.KIND_BVH2)

    # algorithmic traffic (counting pass, outside the timed region)
    stats = {}
    for name, sc, cfg, fl in (("cwbvh_closest", s_cw, cfg_cw, 0), ("cwbvh_any", s_cw, cfg_cw, 1),
                              ("bvh2_closest", s_b2, cfg_b2, 0), ("bvh2_any", s_b2, cfg_b2, 1)):
        d_cnt.zero()
        sc.trace(d_rays.ptr, d_hits.ptr, n, fl | cw.TRACE_STATS, cfg, d_cnt.ptr)
        cw.synchronize()
        c = d_cnt.download((2,), np.uint64)
        stats[name] = (int(c[0]), int(c[1]))

    def run(sc, fl, cfg):
        return lambda: sc.trace(d_rays.ptr, d_hits.ptr, n, fl, cfg)

    log("timing CWBVH closest + any")
    total_ms, per = time_launches([run(s_cw, 0, cfg_cw), run(s_cw, 1, cfg_cw)], args.steps, args.warmup, cw)
    ms_step = total_ms / args.steps
    ms_closest = float(per[:, 0].mean())
    ms_any = float(per[:, 1].mean())
    modes = {"cwbvh_closest": n / ms_closest / 1e3, "cwbvh_any": n / ms_any / 1e3}
    if not args.skip_bvh2:
        log("timing BVH2 closest + any")
        _, per2 = time_launches([run(s_b2, 0, cfg_b2), run(s_b2, 1, cfg_b2)], args.steps, args.warmup, cw)
        modes["bvh2_closest"] = n / float(per2[:, 0].mean()) / 1e3
        modes["bvh2_any"] = n / float(per2[:, 1].mean()) / 1e3
    overflow = cw.stack_overflowed()

    nodes_c, tris_c = stats["cwbvh_closest"]
    bytes_c = algorithmic_bytes(1, n, nodes_c, tris_c)
    achieved = bytes_c / (ms_closest / 1e3) / 1e9
    # HBM-side bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    # WRITE_SIZE calibrated on a 1 GiB copy) for this workload and config
    traffic = None
    if PROFILE_TRAFFIC.exists():
        try:
            t = json.loads(PROFILE_TRAFFIC.read_text())
            if t.get("workload") == wl.name and t.get("rays") == n and t.get("config") == cfg_cw.as_dict():
                traffic = t.get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    compulsory = 48 * n + s_cw.device_bytes()
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "cwbvh_trace<closest> (persistent CWBVH, closest hit)",
            "algorithmic_bytes_per_launch": bytes_c, "avg_launch_ms": round(ms_closest, 4),
            "per_ray": {"nodes": nodes_c / n, "triangles": tris_c / n, "bytes": bytes_c / n},
            # frac > 1 is possible: algorithmic node/triangle bytes are mostly
            # served by L2; the measured HBM-side rate and the compulsory floor
            # (rays + hits + scene once) bracket the real DRAM load
            "hbm_side_gbs": round(traffic / (ms_closest / 1e3) / 1e9, 1) if traffic else None,
            "compulsory_bytes_per_launch": compulsory,
            "compulsory_gbs": round(compulsory / (ms_closest / 1e3) / 1e9, 1),
            "limiter": "VALU issue (SQ_INSTS_VALU x 4 cycles / 1024 SIMDs ~ launch time)"}

    cpu = None
    if not args.no_cpu_baseline:
        threads = max(1, min(16, os.cpu_count() or 1))
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(cwb, rays, args.cpu_seconds, threads)

    value = 2 * n / (ms_step / 1e3) / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "reference files" if args.config == "files" else "synthetic",
        "config": {"workload": f"{wl.name}: {wl.description}", "triangles": wl.mesh.num_triangles,
                   "rays_per_step": 2 * n, "ao_rays": n, "seed": wl.seed,
                   "step": "CWBVH closest-hit pass + CWBVH any-hit pass over the AO batch",
                   "node_stride_f4": args.node_stride},
        "roofline": roof,
        "cpu_baseline": cpu,
        "modes_mrays_s": {k: round(v, 2) for k, v in modes.items()},
        "cwbvh_vs_bvh2": {k: round(modes[f"cwbvh_{k}"] / modes[f"bvh2_{k}"], 3) for k in ("closest", "any")
                          if f"bvh2_{k}" in modes},
        "traversal_counts": {k: {"nodes_per_ray": v[0] / n, "tris_per_ray": v[1] / n} for k, v in stats.items()},
        "trace_config": {"cwbvh": cfg_cw.as_dict(), "bvh2": cfg_b2.as_dict()},
        "build_s": round(build_s, 2),
        "stack_overflow": overflow,
        "device": dev["arch"],
    }
    if wl.extra.get("primary_hit_rate") is not None:
        out["config"]["primary_hit_rate"] = round(wl.extra["primary_hit_rate"], 4)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
