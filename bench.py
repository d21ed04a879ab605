#!/usr/bin/env python3
"""Benchmark: AO ray traversal Mrays/s on one MI355X (CWBVH vs BVH2).

The default run (`python bench.py`) measures every configuration, each in a
process of its own, and prints ONE JSON line:

  C1  (the line's `value`) 980,002-triangle icosphere lattice, 32,986,368 AO
      rays (16 per primary hit of a 1920x1080 camera).  One "step" traverses
      the AO batch twice with the CWBVH kernel: closest hit, then any hit.
      Rays, BVH and hits are resident in HBM before the timed region; the
      timed region holds only kernel launches on one stream, bracketed by HIP
      events.  The BVH2 kernel runs on the same rays; the 2,073,600 coherent
      primaries are timed alone and together with the AO batch.  After the
      timed steps one closest-hit and one any-hit launch with the timed
      configuration are checked against the oracle on every ray ("parity");
      a mismatch makes the run exit non-zero.
  C0  heightfield, 1,048,576 AO rays, both kernels, both modes.
  C2  256 scenes in one persistent batch (graph replay) vs 256 plain launches,
      autotuned vs untuned.
  C3  20 M-triangle clustered soup, 64 M diffuse rays: scene default (Morton
      pre-pass in the launch), given order, pre-ordered.
  C4  serving: 32 resident scenes, Poisson requests, p50/p99 latency.
  AO  C1 surfels: ray generation fused into the traversal kernel.
Each of C0/C2/C3/C4/AO appears under "configs" with its own seed, parity
sample and roofline.  `--config cN` runs one of them alone and prints its
line.  `--config profile` runs the short launch sequence the rocprofv3
summaries under profiles/ are made from.

Rooflines: `achieved` counts the bytes the path must move through HBM per
launch (32 B of ray read and 16 B of hit written per ray, plus the scene's
bytes once), so `frac` <= 1 by construction; `traffic` is the HBM-side bytes
measured with rocprofv3 (FETCH_SIZE / WRITE_SIZE) where a profile of that
workload is committed.  The node and triangle bytes the kernel reads per ray
(mostly served by L2) are reported separately as `cache_served`, and the
VALU issue rate against a calibrated ceiling as `valu_issue`.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s AO any-hit + closest-hit, CWBVH vs BVH2, 1 GPU; serving p99 ms"
HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"
PARITY_FAIL_RC = 5


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ helpers


def host_threads() -> int:
    """CPUs this process may run on (the box's share), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_info() -> dict:
    info = {"affinity_cpus": host_threads(), "os_cpu_count": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cache_bytes(kind: int, n: int, nodes: int, tris: int, leaves: int = 0) -> int:
    """Reference-layout bytes a launch reads and writes at the kernel's side:
    ray 32 + hit 16 per ray, 80 B per CWBVH node visit or 64 B per BVH2 node
    visit, 48 B per triangle test, 16 B per BVH2 leaf terminator.  On
    cache-resident scenes L2 serves most of the node and triangle bytes."""
    if kind == 1:
        return 48 * n + 80 * nodes + 48 * tris
    return 48 * n + 64 * nodes + 48 * tris + 16 * leaves


def time_launches(fns, steps: int, warmup: int, cw, stream=None):
    """Run fns (list of callables) per step; per-launch times from HIP events
    recorded on the launch stream."""
    for _ in range(warmup):
        for f in fns:
            f()
    cw.synchronize()
    evs = [[cw.Event() for _ in range(len(fns) + 1)] for _ in range(steps)]
    for s in range(steps):
        evs[s][0].record(stream)
        for i, f in enumerate(fns):
            f()
            evs[s][i + 1].record(stream)
    evs[-1][-1].synchronize()
    per = np.array([[evs[s][i].elapsed_ms(evs[s][i + 1]) for i in range(len(fns))] for s in range(steps)])
    total_ms = evs[0][0].elapsed_ms(evs[-1][-1])
    return total_ms, per


def profile_record(name: str, **match) -> dict | None:
    """A committed profile summary (profiles/<name>.json) if it was measured on
    this workload and configuration."""
    p = PROFILES / f"{name}.json"
    if not p.exists():
        return None
    try:
        rec = json.loads(p.read_text())
    except (ValueError, OSError):
        return None
    return rec if all(rec.get(k) == v for k, v in match.items()) else None


def roofline(kernel: str, n: int, t_ms: float, scene_bytes: int, cache_b: int | None = None,
             per_ray: dict | None = None, traffic: int | None = None, valu: dict | None = None,
             note: str | None = None) -> dict:
    """HBM roofline of one launch.  achieved = compulsory bytes (48 B per ray +
    the scene once) / launch time; traffic = measured HBM-side bytes."""
    t = t_ms / 1e3
    compulsory = 48 * n + int(scene_bytes)
    achieved = compulsory / t / 1e9
    r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int(traffic) if traffic else None,
         "kernel": kernel, "algorithmic_bytes_per_launch": compulsory,
         "algorithmic_bytes": "48 B per ray (32 B ray read, 16 B hit write) + scene bytes once",
         "avg_launch_ms": round(t_ms, 4)}
    if traffic:
        r["hbm_measured"] = {"gbs": round(traffic / t / 1e9, 1), "frac": round(traffic / t / 1e9 / HBM_PEAK_GBS, 4),
                             "traffic_over_algorithmic": round(traffic / compulsory, 3)}
    if cache_b is not None:
        r["cache_served"] = {"bytes_per_launch": int(cache_b), "gbs": round(cache_b / t / 1e9, 1),
                             "what": "reference-layout node/triangle bytes per visit + ray/hit bytes, mostly L2 hits",
                             **({"per_ray": per_ray} if per_ray else {})}
    if valu:
        r["valu_issue"] = valu
    if note:
        r["note"] = note
    return r


def valu_issue(workload: str, kernel: str, n: int, t_ms: float) -> dict | None:
    """VALU issue rate of a launch (instructions counted by rocprofv3 for this
    workload) against the issue ceiling the calibration probes measured."""
    rec = profile_record("valu_issue", workload=workload, rays=n)
    if not rec or kernel not in rec.get("kernels", {}):
        return None
    insts = rec["kernels"][kernel]["valu_insts_per_launch"]
    ceil = rec["ceiling_insts_per_s"]
    rate = insts / (t_ms / 1e3)
    return {"valu_insts_per_launch": insts, "insts_per_s": round(rate, 1), "ceiling_insts_per_s": ceil,
            "ceiling_probe": rec.get("ceiling_probe"), "frac": round(rate / ceil, 4),
            "source": "profiles/valu_issue.md"}


def parity_block(orc, gpu, ref, rays, woop, any_hit: bool, what: str) -> dict:
    """Bitwise comparison with the oracle; non-identical rays are classified
    (closest: grazing ties vs mismatches; any: occlusion)."""
    gb = np.ascontiguousarray(gpu, np.float32).view(np.uint32)
    rb = np.ascontiguousarray(ref, np.float32).view(np.uint32)
    diff = np.nonzero((gb != rb).any(axis=1))[0]
    occl_g, occl_r = orc.hit_ids(gpu) >= 0, orc.hit_ids(ref) >= 0
    occl_diff = int((occl_g != occl_r).sum())
    out = {"what": what, "rays": int(len(ref)), "bitwise_mismatches": int(len(diff)),
           "occlusion_mismatches": occl_diff, "hit_rate": round(float(occl_r.mean()), 4) if len(ref) else 0.0}
    if len(diff):
        sub = diff[:2000]
        c = orc.compare_hits(gpu[sub], ref[sub], woop, rays[sub], any_hit=any_hit)
        out["classified_sample"] = {"rays": int(len(sub)), "ties": c["ties"], "mismatches": len(c["mismatches"])}
        out["log"] = c["log"][:5]
    # the result contract: closest hits bit-identical except grazing ties;
    # any-hit occlusion agreeing on >= 99.999% of rays
    if any_hit:
        out["ok"] = occl_diff <= 1e-5 * len(ref)
    else:
        out["ok"] = len(diff) == 0 or (out["classified_sample"]["mismatches"] == 0 and len(diff) <= 2000)
    return out


def cpu_baseline(layout, rays, threads: int, target_s: float) -> dict:
    """The CPU restatement (oracle.c built -O3 -march=native on this host)
    timed on the host cores on a bounded sample of the workload's rays."""
    import oracle

    native = True
    try:
        oracle.native_lib()
    except (RuntimeError, OSError) as e:
        log(f"native oracle build failed ({e}); timing the checker build")
        native = False
    nodes, woop = layout.nodes, layout.woop
    n0 = min(len(rays), 262144)
    t0 = time.perf_counter()
    oracle.trace_cwbvh(nodes, layout.node_stride, woop, rays[:n0], threads=threads, native=native)
    dt = time.perf_counter() - t0
    n = int(min(len(rays), max(n0, n0 * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    oracle.trace_cwbvh(nodes, layout.node_stride, woop, rays[:n], threads=threads, native=native)
    dt = time.perf_counter() - t0
    n1 = min(len(rays), 65536)
    t1 = time.perf_counter()
    oracle.trace_cwbvh(nodes, layout.node_stride, woop, rays[:n1], threads=1, native=native)
    dt1 = time.perf_counter() - t1
    cpu = cpu_info()
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first {n} AO rays of the workload, CWBVH closest hit, scalar C restatement (oracle/oracle.c, "
                      f"{'-O3 -march=native' if native else '-O2 checker build'}) on {threads} host threads, "
                      f"{dt:.1f} s wall ({dt * threads:.0f} thread-seconds)",
            "single_thread_mrays_s": round(n1 / dt1 / 1e6, 3), "host_cpu": cpu.get("model", ""), "lscpu": cpu}


def base_line(metric: str, value: float, unit: str, args, ms: float, config: dict, higher=True) -> dict:
    return {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config}


# ------------------------------------------------------- C0 / C1 / files


def bench_single(args, cw, W, dev, name: str) -> dict:
    import oracle

    params = cw.default_build_params(cwbvh_node_stride=args.node_stride)
    t0 = time.perf_counter()
    if name == "c0":
        wl = W.c0()
    elif name == "files":
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

    s_cw, s_b2 = cw.Scene(cwb), cw.Scene(bvh2)
    d_rays = cw.DeviceBuffer.from_numpy(rays)
    d_hits = cw.DeviceBuffer(16 * n)
    d_cnt = cw.DeviceBuffer(16)
    cfg_cw, cfg_b2 = s_cw.default_config(), s_b2.default_config()

    # traversal counts (counting launches, outside the timed region)
    stats = {}
    for key, sc, cfg, fl in (("cwbvh_closest", s_cw, cfg_cw, 0), ("cwbvh_any", s_cw, cfg_cw, 1),
                             ("bvh2_closest", s_b2, cfg_b2, 0), ("bvh2_any", s_b2, cfg_b2, 1)):
        d_cnt.zero()
        sc.trace(d_rays.ptr, d_hits.ptr, n, fl | cw.TRACE_STATS, cfg, d_cnt.ptr)
        cw.synchronize()
        c = d_cnt.download((2,), np.uint64)
        stats[key] = (int(c[0]), int(c[1]))

    def run(sc, fl, cfg, dr=d_rays, dh=d_hits, cnt=n):
        return lambda: sc.trace(dr.ptr, dh.ptr, cnt, fl, cfg)

    # the timed launches are a profiler region (rocprofv3 --selected-regions --marker-trace
    # collects these and nothing of the set-up, counting or parity launches)
    log("timing CWBVH closest + any")
    cw.profiler_start()
    total_ms, per = time_launches([run(s_cw, 0, cfg_cw), run(s_cw, 1, cfg_cw)], args.steps, args.warmup, cw)
    ms_step = total_ms / args.steps
    ms = {"cwbvh_closest": float(per[:, 0].mean()), "cwbvh_any": float(per[:, 1].mean())}
    log("timing BVH2 closest + any")
    _, per2 = time_launches([run(s_b2, 0, cfg_b2), run(s_b2, 1, cfg_b2)], args.steps, args.warmup, cw)
    cw.profiler_stop()
    ms["bvh2_closest"] = float(per2[:, 0].mean())
    ms["bvh2_any"] = float(per2[:, 1].mean())
    modes = {k: n / v / 1e3 for k, v in ms.items()}
    overflow = cw.stack_overflowed()

    primary = None
    prim = wl.extra.get("primary")
    if prim is not None:
        # the coherent camera rays alone, and one frame = primaries + AO closest + AO any
        npr = len(prim)
        d_prim = cw.DeviceBuffer.from_numpy(prim)
        d_phit = cw.DeviceBuffer(16 * npr)
        pc, pb = s_cw.default_config(), s_b2.default_config()
        log("timing primaries (CWBVH, BVH2) and the whole frame")
        _, pp = time_launches([run(s_cw, 0, pc, d_prim, d_phit, npr), run(s_cw, 1, pc, d_prim, d_phit, npr),
                               run(s_b2, 0, pb, d_prim, d_phit, npr), run(s_b2, 1, pb, d_prim, d_phit, npr)],
                              args.steps, args.warmup, cw)
        frame_ms, _ = time_launches([run(s_cw, 0, pc, d_prim, d_phit, npr), run(s_cw, 0, cfg_cw),
                                     run(s_cw, 1, cfg_cw)], args.steps, args.warmup, cw)
        pm = {k: float(pp[:, i].mean()) for i, k in enumerate(("cwbvh_closest", "cwbvh_any", "bvh2_closest",
                                                               "bvh2_any"))}
        primary = {"rays": npr, "ms": {k: round(v, 4) for k, v in pm.items()},
                   "mrays_s": {k: round(npr / v / 1e3, 1) for k, v in pm.items()},
                   "cwbvh_vs_bvh2": {m: round(pm[f"bvh2_{m}"] / pm[f"cwbvh_{m}"], 3) for m in ("closest", "any")},
                   "frame_ms": round(frame_ms / args.steps, 4),
                   "frame": "CWBVH primaries closest + AO closest + AO any, one stream",
                   "frame_mrays_s": round((npr + 2 * n) / (frame_ms / args.steps) / 1e3, 1)}
        # primaries against the oracle too
        cw.synchronize()
        s_cw.trace(d_prim.ptr, d_phit.ptr, npr, 0, pc)
        cw.synchronize()
        primary["parity"] = parity_block(oracle, d_phit.download((npr, 4)),
                                         oracle.trace_cwbvh(cwb.nodes, cwb.node_stride, cwb.woop, prim),
                                         prim, cwb.woop, False, "CWBVH primaries closest, every ray")
        del d_prim, d_phit

    # parity: the timed configuration, every ray, against the oracle
    threads = host_threads()
    parity = []
    for key, sc, lay, cfg, fl in (("cwbvh_closest", s_cw, cwb, cfg_cw, 0), ("cwbvh_any", s_cw, cwb, cfg_cw, 1),
                                  ("bvh2_closest", s_b2, bvh2, cfg_b2, 0), ("bvh2_any", s_b2, bvh2, cfg_b2, 1)):
        d_hits.zero()
        sc.trace(d_rays.ptr, d_hits.ptr, n, fl, cfg)
        cw.synchronize()
        got = d_hits.download((n, 4))
        t0 = time.perf_counter()
        if lay.kind == cw.KIND_CWBVH:
            ref = oracle.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, rays, any_hit=bool(fl), threads=threads)
        else:
            ref = oracle.trace_bvh2(lay.nodes, lay.woop, rays, any_hit=bool(fl), threads=threads)
        p = parity_block(oracle, got, ref, rays, lay.woop, bool(fl), f"{key}, timed configuration, every ray")
        p["oracle_s"] = round(time.perf_counter() - t0, 2)
        log(f"parity {key}: {p['bitwise_mismatches']} of {n} rays differ from the oracle ({p['oracle_s']} s)")
        parity.append(p)
        del got, ref

    nodes_c, tris_c = stats["cwbvh_closest"]
    kname = "cwbvh_trace<closest>"
    rec = profile_record("cwbvh_closest_c1_traffic", workload=wl.name, rays=n, config=cfg_cw.as_dict())
    roof = roofline(f"{kname} (persistent CWBVH, closest hit, {cfg_cw.as_dict()})", n, ms["cwbvh_closest"],
                    s_cw.device_bytes(), cache_bytes(1, n, nodes_c, tris_c),
                    {"nodes": nodes_c / n, "triangles": tris_c / n,
                     "bytes": cache_bytes(1, n, nodes_c, tris_c) / n},
                    traffic=rec.get("hbm_bytes_per_launch") if rec else None,
                    valu=valu_issue(wl.name, "cwbvh_closest", n, ms["cwbvh_closest"]))

    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline on {threads} threads")
        cpu = cpu_baseline(cwb, rays, threads, args.cpu_seconds)

    value = 2 * n / (ms_step / 1e3) / 1e6
    metric = METRIC if name == "c1" else f"{METRIC} [{wl.name}]"
    out = base_line(metric, value, "Mrays/s", args, ms_step,
                    {"workload": f"{wl.name}: {wl.description}", "triangles": wl.mesh.num_triangles,
                     "rays_per_step": 2 * n, "ao_rays": n, "seed": wl.seed,
                     "step": "CWBVH closest-hit pass + CWBVH any-hit pass over the AO batch",
                     "node_stride_f4": args.node_stride})
    if name == "files":
        out["data"] = "reference files"
    out.update({
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": {"rays": n, "mismatches": sum(0 if p["ok"] else max(1, p["bitwise_mismatches"]) for p in parity),
                   "ok": all(p["ok"] for p in parity), "checks": parity},
        "modes_ms": {k: round(v, 4) for k, v in ms.items()},
        "modes_mrays_s": {k: round(v, 2) for k, v in modes.items()},
        "cwbvh_vs_bvh2": {k: round(modes[f"cwbvh_{k}"] / modes[f"bvh2_{k}"], 3) for k in ("closest", "any")},
        "traversal_counts": {k: {"nodes_per_ray": v[0] / n, "tris_per_ray": v[1] / n} for k, v in stats.items()},
        "bvh2_roofline": roofline("bvh2_trace<closest>", n, ms["bvh2_closest"], s_b2.device_bytes(),
                                  cache_bytes(0, n, *stats["bvh2_closest"]),
                                  valu=valu_issue(wl.name, "bvh2_closest", n, ms["bvh2_closest"])),
        "trace_config": {"cwbvh": cfg_cw.as_dict(), "bvh2": cfg_b2.as_dict()},
        "build_s": round(build_s, 2),
        "stack_overflow": overflow,
        "device": dev["arch"],
    })
    if primary:
        out["primary"] = primary
        out["parity"]["checks"].append(primary["parity"])
        out["parity"]["ok"] = out["parity"]["ok"] and primary["parity"]["ok"]
    if wl.extra.get("primary_hit_rate") is not None:
        out["config"]["primary_hit_rate"] = round(wl.extra["primary_hit_rate"], 4)
    return out


# ------------------------------------------------------------------- C2


def build_c2(args, cw, W) -> list:
    """C2's scenes and AO rays: [(plan entry, triangles, CWBVH layout, rays)]."""
    from concurrent.futures import ThreadPoolExecutor

    def make(entry):
        mesh, rays, _ = W.c2_scene(entry)
        _, lay = cw.build_layouts(mesh, cw.default_build_params(num_threads=2), bvh2=False)
        return entry, mesh.num_triangles, lay, rays

    with ThreadPoolExecutor(max_workers=max(1, host_threads() // 2)) as ex:
        return list(ex.map(make, W.c2_plan(num_scenes=args.c2_scenes)))


def bench_c2(args, cw, W, dev) -> dict:
    """C2: many scenes, uneven AO batches.  Plain = one cwbvh_trace launch per
    scene (default configuration) back to back on one stream; batched = one
    autotuned persistent multi-scene launch replayed from a hipGraph (the
    Morton pre-pass, when the tuner picks it, is inside the timed region)."""
    import oracle

    t0 = time.perf_counter()
    built = build_c2(args, cw, W)
    build_s = time.perf_counter() - t0
    tris = sum(b[1] for b in built)
    log(f"C2: {len(built)} scenes, {tris} triangles, {sum(len(b[3]) for b in built)} rays, build {build_s:.1f} s")
    scenes, bufs, items, lays, samples = [], [], [], [], []
    rng = np.random.default_rng(W.SEEDS["C2"] + 99)
    for entry, nt, lay, rays in built:
        sc = cw.Scene(lay)
        dr = cw.DeviceBuffer.from_numpy(rays)
        dh = cw.DeviceBuffer(16 * len(rays))
        scenes.append(sc)
        bufs.append((dr, dh))
        items.append((sc, dr.ptr, dh.ptr, len(rays)))
        lays.append(lay)
        # parity sample: up to 1024 rays of every scene
        idx = np.sort(rng.choice(len(rays), size=min(1024, len(rays)), replace=False))
        samples.append((idx, rays[idx]))
    del built
    total = sum(it[3] for it in items)
    scene_bytes = sum(sc.device_bytes() for sc in scenes)
    st = cw.Stream()
    cfg0 = cw.default_trace_config(cw.KIND_CWBVH)

    def plain():
        for sc, r, h, n in items:
            sc.trace(r, h, n, 0, None, None, st)

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
    log(f"plain {ms_plain:.2f} ms")
    t = time.perf_counter()
    best, recs = cw.autotune_batch(items, block_sizes=(64, 128, 256), lds_depths=(0, 8, 16, 32), morton=(0, 1),
                                   prefetch=(0, 1), loop_modes=(0, 1, 2), reps=args.tune_reps, stream=st)
    tune_s = time.perf_counter() - t
    log(f"autotune {tune_s:.1f} s best {best.as_dict()}")
    batch = cw.Batch(items, cw.TRACE_CLOSEST, best.cfg, morton=bool(best.morton))
    batch.capture(st)
    ms_batch = timed(lambda: batch.replay(st), args.steps, args.warmup)
    # the library's batch default (no configuration given)
    untuned = cw.Batch(items, cw.TRACE_CLOSEST, None, morton=False)
    ms_untuned = timed(lambda: untuned.launch(st), args.steps, args.warmup)
    value = total / (ms_batch / 1e3) / 1e6

    # parity: the tuned graph replay's hits against the oracle on every
    # scene's sample
    for _, dh in bufs:
        dh.zero()
    batch.replay(st)
    st.synchronize()
    got_all, ref_all, ray_all = [], [], []
    mism = 0
    for (idx, srays), (_, dh), lay, it in zip(samples, bufs, lays, items):
        got = dh.download((it[3], 4))[idx]
        ref = oracle.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, srays, threads=host_threads())
        mism += int((got.view(np.uint32) != ref.view(np.uint32)).any(axis=1).sum())
        got_all.append(got)
        ref_all.append(ref)
    parity = {"rays": int(sum(len(s[0]) for s in samples)), "mismatches": mism, "ok": mism == 0,
              "what": "autotuned graph replay, up to 1024 random rays of every scene, closest hit, bitwise",
              "hit_rate": round(float(np.mean(np.concatenate([oracle.hit_ids(r) >= 0 for r in ref_all]))), 4)}
    log(f"parity C2: {mism} of {parity['rays']} sampled rays differ from the oracle")

    # any hit (AO occlusion): the tuned configuration as an any-hit graph
    # replay against 256 plain any-hit launches; its sampled hits against the
    # oracle's any hit
    def plain_any():
        for sc, r, h, n in items:
            sc.trace(r, h, n, cw.TRACE_ANY, None, None, st)

    ms_plain_any = timed(plain_any, args.steps, args.warmup)
    batch_any = cw.Batch(items, cw.TRACE_ANY, best.cfg, morton=bool(best.morton))
    batch_any.capture(st)
    ms_batch_any = timed(lambda: batch_any.replay(st), args.steps, args.warmup)
    for _, dh in bufs:
        dh.zero()
    batch_any.replay(st)
    st.synchronize()
    mism_any = occl_any = unverified = 0
    for (idx, srays), (_, dh), lay, it in zip(samples, bufs, lays, items):
        got = dh.download((it[3], 4))[idx]
        ref = oracle.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, srays, any_hit=True, threads=host_threads())
        differ = (got.view(np.uint32) != ref.view(np.uint32)).any(axis=1)
        mism_any += int(differ.sum())
        gid, rid = oracle.hit_ids(got), oracle.hit_ids(ref)
        occl_any += int(((gid >= 0) != (rid >= 0)).sum())
        # a speculative schedule may report another triangle on the ray: it
        # must be one the oracle's Woop test accepts with the same t, u, v
        for i in np.nonzero(differ & (gid >= 0) & (rid >= 0))[0]:
            ok, t, u, v = oracle.woop_test(lay.woop[gid[i]:gid[i] + 3], srays[i])
            if not (ok and np.float32(t) == got[i, 0] and np.float32(u) == got[i, 2] and np.float32(v) == got[i, 3]):
                unverified += 1
    any_ok = occl_any == 0 and unverified == 0 and (mism_any == 0 or best.cfg.loop_mode == 2)
    any_hit = {"batched_ms": round(ms_batch_any, 4), "batched_mrays_s": round(total / ms_batch_any / 1e3, 2),
               "plain_ms": round(ms_plain_any, 4), "plain_mrays_s": round(total / ms_plain_any / 1e3, 2),
               "speedup_vs_plain": round(ms_plain_any / ms_batch_any, 3),
               "what": "the tuned configuration as an any-hit graph replay; plain = 256 any-hit launches",
               "parity": {"rays": parity["rays"], "bitwise_mismatches": mism_any, "occlusion_mismatches": occl_any,
                          "unverified_hits": unverified, "ok": any_ok,
                          "what": "any-hit graph replay, the same samples, against the oracle's any hit: bitwise "
                                  "(the speculative schedule: occlusion identical and every other reported hit "
                                  "accepted by the oracle's Woop test)"}}
    parity["ok"] = parity["ok"] and any_ok
    log(f"C2 any hit: batched {ms_batch_any:.2f} ms, plain {ms_plain_any:.2f} ms, "
        f"{mism_any} sampled rays differ from the oracle")

    # cache-side traffic of the batch (counting launches per scene, outside
    # the timed region)
    d_cnt = cw.DeviceBuffer(16)
    nodes_c = tris_c = 0
    for sc, r, h, n in items:
        d_cnt.zero()
        sc.trace(r, h, n, cw.TRACE_STATS, cfg0, d_cnt.ptr, st)
        st.synchronize()
        c = d_cnt.download((2,), np.uint64)
        nodes_c += int(c[0])
        tris_c += int(c[1])
    bytes_c = cache_bytes(1, total, nodes_c, tris_c)
    # measured L2 <-> fabric bytes of the library-default batch launch over the same rays
    rec = profile_record("cwbvh_batch_c2_traffic", rays=total, scenes=len(items), batch="library default")
    out = base_line(METRIC + " [C2 batched]", value, "Mrays/s", args, ms_batch,
                    {"workload": f"C2: {len(items)} seeded scenes (heightfields / icosphere lattices), "
                                 f"{tris} triangles, {total} AO rays (closest hit)",
                     "triangles": tris, "rays": total, "scenes": len(items), "seed": W.SEEDS["C2"]})
    out.update({
        "plain_ms": round(ms_plain, 4),
        "plain_mrays_s": round(total / ms_plain / 1e3, 2),
        "batched_untuned_ms": round(ms_untuned, 4),
        "batched_untuned_roofline": roofline(
            "cwbvh_trace<closest, MULTI> (persistent batch, library default configuration)", total, ms_untuned,
            scene_bytes, bytes_c, traffic=rec.get("hbm_bytes_per_launch") if rec else None),
        "roofline": roofline("cwbvh_trace<closest, MULTI> (autotuned persistent batch, graph replay)", total,
                             ms_batch, scene_bytes, bytes_c,
                             {"nodes": nodes_c / total, "triangles": tris_c / total, "bytes": bytes_c / total}),
        "plain_roofline": roofline("cwbvh_trace<closest> x 256 plain launches", total, ms_plain, scene_bytes),
        "speedup_vs_plain": round(ms_plain / ms_batch, 3),
        "speedup_tuned_vs_untuned": round(ms_untuned / ms_batch, 4),
        "any_hit": any_hit,
        "autotune": {"best": best.as_dict(), "seconds": round(tune_s, 1), "points": len(recs),
                     "fastest_5": [r.as_dict() for r in sorted(recs, key=lambda r: r.ms)[:5]],
                     "slowest_ms": round(max(r.ms for r in recs), 4)},
        "parity": parity,
        "build_s": round(build_s, 1),
        "device": dev["arch"],
    })
    return out


# ------------------------------------------------------------------- C3


def bench_c3(args, cw, W, dev) -> dict:
    """C3: large incoherent scene, closest hit; rays shuffled vs Morton-ordered
    (Morton order applied through the ordered-trace indirection; the sort is
    timed separately)."""
    import oracle

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
    # cache line per node), plus the 80 B reference layout
    s_cw, s_b2 = cw.Scene(cwb), cw.Scene(bvh2)
    s_cw80 = s_cw if s_cw.node_stride == 5 else cw.Scene(cwb, stride_mode=cw.SCENE_STRIDE_KEEP)
    lo, hi = s_cw.bounds()
    rng = np.random.default_rng(W.SEEDS["C3"] + 3)
    sample = np.sort(rng.choice(n, size=min(n, 1 << 20), replace=False))
    srays = rays[sample]
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
    ms80 = t(s_cw80, s_cw80.default_config(), None)
    res["cwbvh_node80_default"] = {"ms": round(ms80, 3), "mrays_per_s": round(n / ms80 / 1e3, 1)}
    log(f"cwbvh 80 B nodes, scene default: {ms80:.2f} ms {n / ms80 / 1e3:.0f} Mrays/s")

    # parity: the default launch (sort inside), 2^20 random rays
    parity = []
    for name, sc, lay, cfg in (("cwbvh", s_cw, cwb, cfg_cw), ("bvh2", s_b2, bvh2, cfg_b2)):
        dh.zero()
        sc.trace(dr.ptr, dh.ptr, n, 0, cfg)
        cw.synchronize()
        got = dh.download((n, 4))[sample]
        t0 = time.perf_counter()
        if lay.kind == cw.KIND_CWBVH:
            ref = oracle.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, srays, threads=host_threads())
        else:
            ref = oracle.trace_bvh2(lay.nodes, lay.woop, srays, threads=host_threads())
        p = parity_block(oracle, got, ref, srays, lay.woop, False,
                         f"{name} scene default (sorted in the launch), {len(sample)} random rays of {n}")
        p["oracle_s"] = round(time.perf_counter() - t0, 2)
        log(f"parity C3 {name}: {p['bitwise_mismatches']} of {len(sample)} differ ({p['oracle_s']} s)")
        parity.append(p)

    d_cnt = cw.DeviceBuffer(16)
    d_cnt.zero()
    s_cw.trace(dr.ptr, dh.ptr, n, cw.TRACE_STATS, cfg_cw, d_cnt.ptr)
    cw.synchronize()
    c = d_cnt.download((2,), np.uint64)
    bytes_c = cache_bytes(1, n, int(c[0]), int(c[1]))
    ms_c = res["cwbvh_default"]["ms"]
    rec = profile_record("cwbvh_closest_c3_traffic", rays=n, config=cfg_cw.as_dict())
    out = base_line(METRIC + " [C3 large incoherent]", n / ms_c / 1e3, "Mrays/s", args, ms_c,
                    {"workload": f"C3: clustered soup {mesh.num_triangles} triangles (4096 clusters), {n} diffuse "
                                 f"rays from surface points, tmax=inf, closest hit, shuffled input, library default "
                                 f"configuration (Morton pre-pass inside the launch)",
                     "triangles": mesh.num_triangles, "rays": n, "seed": W.SEEDS["C3"],
                     "node_bytes": 16 * s_cw.node_stride})
    out.update({
        "results": res,
        "morton_sort_ms": round(sort_ms, 3),
        "cwbvh_vs_bvh2": {"default": round(res["bvh2_default"]["ms"] / res["cwbvh_default"]["ms"], 3),
                          "shuffled": round(res["bvh2_shuffled"]["ms"] / res["cwbvh_shuffled"]["ms"], 3),
                          "morton": round(res["bvh2_morton"]["ms"] / res["cwbvh_morton"]["ms"], 3)},
        "roofline": roofline("cwbvh_trace<closest> (ray-order pre-pass + interleaved schedule with prefetch, 128 B nodes)", n, ms_c,
                             s_cw.device_bytes(), bytes_c,
                             {"nodes": int(c[0]) / n, "triangles": int(c[1]) / n, "bytes": bytes_c / n},
                             traffic=rec.get("hbm_bytes_per_launch") if rec else None),
        "parity": {"rays": len(sample), "mismatches": sum(p["bitwise_mismatches"] for p in parity),
                   "ok": all(p["ok"] for p in parity), "checks": parity},
        "build_s": round(build_s, 1),
        "device": dev["arch"],
    })
    return out


# ------------------------------------------------------------------- AO


def bench_ao(args, cw, W, dev) -> dict:
    """AO from surfels on the C1 scene: the primary hits of the 1920x1080
    camera become surfels (hit point, geometric normal facing the camera,
    texel radius 2e-4 * diag), 16 AO rays each (tmin 0, tmax 0.05 * diag).
    Three ways to run the same rays, closest and any hit:
      unfused      ray generation kernel writes rays, then cwbvh_trace;
      fused_hits   cwbvh_trace_surfels generates rays in the traversal kernel;
      fused_counts as fused_hits, with per-surfel hit counts instead of hits."""
    import oracle

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
    # parity: fused hits and fused per-surfel counts against the oracle
    # tracing the host generator's rays (every ray)
    rays = cw.rays_ao_from_samples(pos, nrm4, rad, k, 0.0, tmax, seed)
    parity = []
    for fname, fl in (("closest", cw.TRACE_CLOSEST), ("any", cw.TRACE_ANY)):
        ref = oracle.trace_cwbvh(cwb.nodes, cwb.node_stride, cwb.woop, rays, any_hit=bool(fl), threads=host_threads())
        d_hits.zero()
        sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, fl, d_hits=d_hits.ptr,
                         cfg=cfg)
        d_counts.zero()
        sc.trace_surfels(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, fl,
                         d_texel_hits=d_counts.ptr, cfg=cfg)
        cw.synchronize()
        p = parity_block(oracle, d_hits.download((n, 4)), ref, rays, cwb.woop, bool(fl),
                         f"fused generate-and-trace hits, {fname}, every ray")
        want = (oracle.hit_ids(ref) >= 0).reshape(ntex, k).sum(axis=1).astype(np.uint32)
        p["count_mismatches"] = int((d_counts.download((ntex,), np.uint32) != want).sum())
        p["ok"] = p["ok"] and p["count_mismatches"] == 0
        parity.append(p)
        log(f"parity AO {fname}: {p['bitwise_mismatches']} hits, {p['count_mismatches']} counts differ")
    d_cnt = cw.DeviceBuffer(16)
    d_cnt.zero()
    cw.rays_ao_from_samples_device(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, ntex, k, 0.0, tmax, seed, d_rays.ptr)
    sc.trace(d_rays.ptr, d_hits.ptr, n, cw.TRACE_ANY | cw.TRACE_STATS, cfg, d_cnt.ptr)
    cw.synchronize()
    c = d_cnt.download((2,), np.uint64)
    ms_v = table["fused_counts_any"]["ms"]
    best = max(table.items(), key=lambda kv: kv[1]["mrays_per_s"])
    # the fused kernel reads 36 B of surfel and adds to a 4 B count per surfel
    # instead of moving 48 B of ray and hit per ray
    out = base_line(METRIC + " [AO from surfels, fused ray generation]", table["fused_counts_any"]["mrays_per_s"],
                    "Mrays/s", args, ms_v,
                    {"workload": f"C1 scene, {ntex} surfels (primary hits) x {k} AO rays, tmax=0.05*diag",
                     "triangles": mesh.num_triangles, "surfels": ntex, "rays": n, "seed": seed})
    roof = roofline("cwbvh_trace<any, SRC=surfels> with per-surfel counts", 0, ms_v, sc.device_bytes() + 40 * ntex,
                    80 * int(c[0]) + 48 * int(c[1]) + 40 * ntex,
                    note="algorithmic bytes: 36 B surfel read + 4 B count per surfel + the scene once (no ray or "
                         "hit buffers)")
    out.update({
        "results": table,
        "speedup_fused_counts_vs_unfused_any": round(table["unfused_any"]["ms"] / table["fused_counts_any"]["ms"], 3),
        "speedup_fused_hits_vs_unfused_closest": round(table["unfused_closest"]["ms"] /
                                                       table["fused_hits_closest"]["ms"], 3),
        "bytes_per_ray": {"unfused": 32 + 32 + 16, "fused_hits": 16, "fused_counts": "4 per surfel"},
        "roofline": roof,
        "parity": {"rays": n, "mismatches": sum(p["bitwise_mismatches"] + p["count_mismatches"] for p in parity),
                   "ok": all(p["ok"] for p in parity), "checks": parity},
        "best": best[0],
        "device": dev["arch"],
    })
    return out


# ------------------------------------------------------------------- C4


def bench_c4(args, cw, W, dev) -> dict:
    """C4: resident scenes, Poisson AO query requests (log-uniform 1k..1M rays,
    uniform scene), host-resident rays copied in and hits copied out per
    request.  Continuous batching vs serial per-request plain launches."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle

    specs = W.c4_plan(num_scenes=args.c4_scenes, total_tris=args.c4_tris)
    t0 = time.perf_counter()

    def make(entry):
        mesh = W.c4_mesh(entry)
        _, lay = cw.build_layouts(mesh, cw.default_build_params(num_threads=4), bvh2=False)
        diag = mesh.diagonal()
        rays = cw.rays_ao_surface(mesh, 131072, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=entry["seed"] + 1)
        log(f"C4 scene {entry['index']}: {mesh.num_triangles} triangles")
        return entry, mesh.num_triangles, lay, rays

    with ThreadPoolExecutor(max_workers=max(1, host_threads() // 4)) as ex:
        built = list(ex.map(make, specs))
    build_s = time.perf_counter() - t0
    tris = sum(b[1] for b in built)
    scenes, pools, lays, hbm = [], [], [], 0
    for entry, nt, lay, rays in built:
        sc = cw.Scene(lay)
        hbm += sc.device_bytes()
        pool = cw.PinnedArray(rays.shape, np.float32)
        pool.array[:] = rays
        scenes.append(sc)
        pools.append(pool)
        lays.append(lay)
    del built
    log(f"C4: {len(scenes)} resident scenes, {tris} triangles, {hbm / 2**30:.2f} GiB of BVH in HBM, "
        f"build {build_s:.1f} s")
    hit_pool = cw.PinnedArray((16 << 20, 4), np.float32)
    # device-resident copies of the ray pools and a device hit ring, for the
    # requests that keep their buffers on the GPU
    d_pools = [cw.DeviceBuffer.from_numpy(p.array) for p in pools]
    d_hits = cw.DeviceBuffer(16 * (16 << 20))
    # surfel pools: the pools' ray origins and directions as surfel positions
    # and normals (8 rays per surfel), radius 2e-4 of the unit scene scale
    rpt = 8
    surf = []
    for p in pools:
        pts = p.array[::rpt]
        sp = cw.PinnedArray((len(pts), 4), np.float32)
        sp.array[:] = 0
        sp.array[:, :3] = pts[:, :3]
        sn = cw.PinnedArray((len(pts), 4), np.float32)
        sn.array[:] = 0
        sn.array[:, :3] = pts[:, 4:7]
        sr = cw.PinnedArray((len(pts),), np.float32)
        sr.array[:] = np.float32(2e-4)
        surf.append((sp, sn, sr))
    counts = cw.PinnedArray((1 << 21,), np.uint32)
    rates = [float(r) for r in args.c4_rates.split(",")]
    table = []
    modes = (("host", 0), ("host", 1), ("device", 0), ("device", 1), ("surfels", 0))
    for kind, serial in modes:
        cfg = cw.default_server_config(serial=serial, memory_budget=8 << 30, max_inflight=3,
                                       max_batch_requests=64, max_batch_rays=16 << 20)
        srv = cw.Server(scenes, cfg)

        def run(rate, num, seed):
            if kind == "device":
                return srv.run_poisson_device(rate, num, 1000, 1_000_000, [d.ptr for d in d_pools],
                                              [len(p.array) for p in pools], d_hits.ptr, 16 << 20, seed=seed)
            if kind == "surfels":
                return srv.run_poisson_surfels(rate, num, 1000, 1_000_000, surf, rpt, 0.0, 0.05, counts, seed=seed)
            return srv.run_poisson(rate, num, 1000, 1_000_000, pools, hit_pool, seed=seed)

        run(2000.0, 20, 1)  # warm-up
        for rate in rates:
            res, _ = run(rate, args.c4_requests, int(rate))
            st = srv.stats()
            row = {"requests": kind, "mode": "serial" if serial else "continuous", "rate_req_s": rate,
                   "p50_ms": round(res["p50_ms"], 3), "p99_ms": round(res["p99_ms"], 3),
                   "mean_ms": round(res["mean_ms"], 3), "mrays_per_s": round(res["mrays_per_s"], 1),
                   "wall_s": round(res["wall_s"], 3), "max_inflight_seen": st["max_inflight_seen"],
                   "max_batch_requests_seen": st["max_batch_requests_seen"]}
            log(str(row))
            table.append(row)
        srv.close()
    # scene admission: a new scene of the set's mean size is generated,
    # built, uploaded and added while host-ray requests arrive at 5k req/s on
    # the resident scenes; its first request is timed after the add
    from concurrent.futures import ThreadPoolExecutor as _TP

    admit = {}
    srv = cw.Server(scenes, cw.default_server_config(memory_budget=8 << 30, max_inflight=3))
    with _TP(1) as ex:
        load = ex.submit(srv.run_poisson, 5000.0, 2000, 1000, 1_000_000, pools, hit_pool, 77)
        entry = {"index": len(specs), "kind": "icospheres", "target_tris": int(tris / len(specs)),
                 "seed": W.SEEDS["C4"] + 4242}
        t0 = time.perf_counter()
        mesh = W.c4_mesh(entry)
        t1 = time.perf_counter()
        _, lay_new = cw.build_layouts(mesh, bvh2=False)
        t2 = time.perf_counter()
        sc_new = cw.Scene(lay_new)
        t3 = time.perf_counter()
        idx = srv.add_scene(sc_new)
        t4 = time.perf_counter()
        diag = mesh.diagonal()
        nrays = cw.rays_ao_surface(mesh, 16384, 8, 1e-4 * diag, 0.0, 0.1 * diag, seed=entry["seed"] + 1)
        npool = cw.PinnedArray(nrays.shape, np.float32)
        npool.array[:] = nrays
        nh = cw.PinnedArray((len(nrays), 4), np.float32)
        t5 = time.perf_counter()
        first_ms = srv.wait(srv.submit(idx, npool.ptr, len(nrays), nh.ptr), 60.0)
        res_load, _ = load.result()
    ref_new = oracle.trace_cwbvh(lay_new.nodes, lay_new.node_stride, lay_new.woop, nrays, threads=host_threads())
    admit = {"triangles": mesh.num_triangles, "generate_s": round(t1 - t0, 3), "build_s": round(t2 - t1, 3),
             "upload_s": round(t3 - t2, 3), "add_ms": round((t4 - t3) * 1e3, 3),
             "first_request_ms": round(first_ms, 3), "first_request_rays": len(nrays),
             "first_request_matches_oracle": bool(np.array_equal(nh.array.view(np.uint32), ref_new.view(np.uint32))),
             "admission_s": round(t4 - t0, 3),
             "concurrent_load": {"rate_req_s": 5000.0, "p50_ms": round(res_load["p50_ms"], 3),
                                 "p99_ms": round(res_load["p99_ms"], 3)}}
    srv.close()
    log(f"scene admission while serving: {admit}")

    # parity: 24 overlapping requests with their own hit buffers, every served
    # hit against the oracle
    rng = np.random.default_rng(W.SEEDS["C4"] + 5)
    srv = cw.Server(scenes, cw.default_server_config(memory_budget=8 << 30, max_inflight=3))
    reqs = []
    for _ in range(24):
        s = int(rng.integers(0, len(scenes)))
        nr = int(np.exp(rng.uniform(np.log(1000), np.log(200000))))
        off = int(rng.integers(0, len(pools[s].array) - nr))
        h = cw.PinnedArray((nr, 4), np.float32)
        reqs.append((srv.submit(s, pools[s].ptr + 32 * off, nr, h.ptr), s, off, nr, h))
    mism = checked = 0
    for rid, s, off, nr, h in reqs:
        srv.wait(rid, 60.0)
        rr = pools[s].array[off:off + nr]
        ref = oracle.trace_cwbvh(lays[s].nodes, lays[s].node_stride, lays[s].woop, rr, threads=host_threads())
        mism += int((h.array.view(np.uint32) != ref.view(np.uint32)).any(axis=1).sum())
        checked += nr
    srv.close()
    log(f"parity C4: {mism} of {checked} served rays differ from the oracle")
    cont = [r for r in table if r["mode"] == "continuous" and r["requests"] == "host"]
    ser = [r for r in table if r["mode"] == "serial" and r["requests"] == "host"]

    def peak(kind, mode):
        rows = [r["mrays_per_s"] for r in table if r["requests"] == kind and r["mode"] == mode]
        return max(rows) if rows else None
    args.steps = args.c4_requests
    args.warmup = 20
    out = base_line(METRIC + " [C4 serving]", cont[0]["p99_ms"], "ms (p99 latency at the lowest rate)", args,
                    cont[0]["mean_ms"],
                    {"workload": f"C4: {len(scenes)} resident scenes ({tris} triangles), Poisson requests of "
                                 f"log-uniform 1k..1M AO rays; value: host rays in / hits out (PCIe included)",
                     "triangles": tris, "scenes": len(scenes), "requests_per_rate": args.c4_requests,
                     "rates": rates, "seed": W.SEEDS["C4"]}, higher=False)
    out.update({
        "table": table,
        "max_mrays_per_s": {f"{k}_{m}": peak(k, m) for k, m in (("host", "continuous"), ("host", "serial"),
                                                                 ("device", "continuous"), ("device", "serial"),
                                                                 ("surfels", "continuous"))},
        "p99_ms_by_rate": {f"{k}_{m}": {int(r["rate_req_s"]): r["p99_ms"] for r in table
                                         if r["requests"] == k and r["mode"] == m}
                           for k, m in (("host", "continuous"), ("host", "serial"), ("device", "continuous"),
                                        ("device", "serial"), ("surfels", "continuous"))},
        "request_kinds": {"host": "rays copied in, hits copied out (48 B per ray over PCIe)",
                          "device": "rays and hits in device memory, used in place (no copies)",
                          "surfels": f"surfels copied in (36 B), {rpt} AO rays generated per surfel in the traversal "
                                     f"kernel, one 4 B hit count per surfel copied out"},
        "parity": {"rays": checked, "requests": len(reqs), "mismatches": mism,
                   "ok": mism == 0 and admit["first_request_matches_oracle"],
                   "what": "24 overlapping served requests, every hit, closest, bitwise; the admitted scene's "
                           "first request"},
        "scene_admission": admit,
        "bvh_hbm_gib": round(hbm / 2**30, 2),
        "build_s": round(build_s, 1),
        "device": dev["arch"],
    })
    return out


# ------------------------------------------------------- profile launches


def profile_launches(args, cw, W) -> dict:
    """Launch sequence for rocprofv3 summaries: a 1 GiB copy (calibrates the
    byte counters), the three VALU issue probes (calibrate the issue
    ceiling), then --launches CWBVH and --launches BVH2 launches over the
    workload's rays with each scene's default configuration."""
    if args.workload == "c2":
        return profile_launches_c2(args, cw, W)
    if args.workload == "c0":
        wl = W.c0()
    elif args.workload == "c3":
        wl = W.c3(args.c3_tris, args.c3_rays)
        wl.rays = cw.shuffle_rays(wl.rays, seed=W.SEEDS["C3"] + 2)
    else:
        wl = W.c1(W.gpu_primary_tracer())
    bvh2, cwb = cw.build_layouts(wl.mesh)
    n = len(wl.rays)
    s_cw, s_b2 = cw.Scene(cwb), cw.Scene(bvh2)
    cfg, cfg_b2 = s_cw.default_config(), s_b2.default_config()
    dr = cw.DeviceBuffer.from_numpy(wl.rays)
    dh = cw.DeviceBuffer(16 * n)
    copy_bytes = args.copy_mib << 20
    src, dst = cw.DeviceBuffer(copy_bytes), cw.DeviceBuffer(copy_bytes)
    src.zero()
    cw.synchronize()
    # everything below is one profiler region (rocprofv3 --selected-regions --marker-trace)
    cw.profiler_start()
    cw.check(cw.lib.cwbvh_copy_u4(dst.ptr, src.ptr, copy_bytes // 16, None))
    probe_blocks = 8192
    probe = cw.DeviceBuffer(4 * 256 * probe_blocks)
    for mode in (0, 1, 2):
        cw.check(cw.lib.cwbvh_valu_probe(mode, args.probe_iters, probe_blocks, probe.ptr, None))
    flags = cw.TRACE_ANY if args.any else cw.TRACE_CLOSEST
    for _ in range(args.launches):
        s_cw.trace(dr.ptr, dh.ptr, n, flags, cfg)
    for _ in range(args.launches):
        s_b2.trace(dr.ptr, dh.ptr, n, flags, cfg_b2)
    cw.synchronize()
    cw.profiler_stop()
    return {"workload": wl.name, "rays": n, "copy_bytes": copy_bytes, "launches": args.launches, "flags": flags,
            "probe": {"blocks": probe_blocks, "threads_per_block": 256, "iters": args.probe_iters},
            "cwbvh_config": cfg.as_dict(), "bvh2_config": cfg_b2.as_dict(), "node_bytes": 16 * s_cw.node_stride,
            "triangles": wl.mesh.num_triangles,
            "scene_bytes": {"cwbvh": s_cw.device_bytes(), "bvh2": s_b2.device_bytes()}}


def profile_launches_c2(args, cw, W) -> dict:
    """C2 launch sequence for rocprofv3: the 1 GiB copy, the three VALU
    probes, then --launches launches of one batch over C2's 256 scenes with
    the library's default batch configuration."""
    built = build_c2(args, cw, W)
    scenes = [cw.Scene(lay) for _, _, lay, _ in built]
    pools = [cw.DeviceBuffer.from_numpy(rays) for _, _, _, rays in built]
    sizes = [len(rays) for _, _, _, rays in built]
    hits = [cw.DeviceBuffer(16 * n) for n in sizes]
    items = [(sc, p.ptr, h.ptr, n) for sc, p, h, n in zip(scenes, pools, hits, sizes)]
    batch = cw.Batch(items, cw.TRACE_CLOSEST, None)
    copy_bytes = args.copy_mib << 20
    src, dst = cw.DeviceBuffer(copy_bytes), cw.DeviceBuffer(copy_bytes)
    src.zero()
    cw.synchronize()
    cw.profiler_start()
    cw.check(cw.lib.cwbvh_copy_u4(dst.ptr, src.ptr, copy_bytes // 16, None))
    probe_blocks = 8192
    probe = cw.DeviceBuffer(4 * 256 * probe_blocks)
    for mode in (0, 1, 2):
        cw.check(cw.lib.cwbvh_valu_probe(mode, args.probe_iters, probe_blocks, probe.ptr, None))
    for _ in range(args.launches):
        batch.launch()
    cw.synchronize()
    cw.profiler_stop()
    return {"workload": "C2", "rays": int(sum(sizes)), "scenes": len(items), "batch": "library default",
            "copy_bytes": copy_bytes, "launches": args.launches, "flags": cw.TRACE_CLOSEST,
            "probe": {"blocks": probe_blocks, "threads_per_block": 256, "iters": args.probe_iters},
            "triangles": int(sum(b[1] for b in built)),
            "scene_bytes": int(sum(sc.device_bytes() for sc in scenes))}


# ------------------------------------------------------------ orchestrator

CHILD_ORDER = ("c1", "c0", "c2", "c3", "c4", "ao")
# rough lower bounds of a child's run time on the box (seconds): a child is
# skipped when less than this is left before the deadline
CHILD_MIN_S = {"c1": 60, "c0": 20, "c2": 60, "c3": 60, "c4": 60, "ao": 30}


def orchestrate(args) -> int:
    """Run every configuration in a child process of its own (this process
    never touches the GPU) and print one merged line: C1's keys plus the
    others under "configs"."""
    t_start = time.perf_counter()
    lines, status = {}, {}
    stop = False
    for name in CHILD_ORDER:
        if name not in args.configs.split(","):
            continue
        left = args.deadline_s - (time.perf_counter() - t_start)
        if stop or left < CHILD_MIN_S[name]:
            status[name] = "skipped: " + ("an earlier child failed" if stop else f"{left:.0f} s left")
            log(f"{name}: {status[name]}")
            continue
        cmd = [sys.executable, str(ROOT / "bench.py"), "--config", name, "--steps", str(args.steps),
               "--warmup", str(args.warmup)]
        if args.no_cpu_baseline:
            cmd.append("--no-cpu-baseline")
        log(f"--- {name}: {' '.join(cmd[2:])} (timeout {left - 5:.0f} s)")
        t0 = time.perf_counter()
        try:
            proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=max(10.0, left - 5))
        except subprocess.TimeoutExpired:
            status[name] = "timed out"
            log(f"{name}: timed out; no further configuration is run")
            stop = True
            continue
        wall = time.perf_counter() - t0
        line = next((l for l in reversed(proc.stdout.splitlines()) if l.startswith("{")), None)
        if line:
            rec = json.loads(line)
            rec["child_wall_s"] = round(wall, 1)
            lines[name] = rec
        if proc.returncode == 0:
            status[name] = "ok"
        elif proc.returncode == PARITY_FAIL_RC and line:
            status[name] = "parity mismatch"
        else:
            status[name] = f"exit {proc.returncode}"
            # a crashed or killed child may have left the device in a bad
            # state: run nothing more on it
            stop = True
        log(f"{name}: {status[name]} ({wall:.0f} s)")
    if "c1" not in lines:
        print(json.dumps({"metric": METRIC, "value": None, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "error": status.get("c1", "not run"), "status": status}),
              flush=True)
        return 1
    out = lines.pop("c1")
    out["configs"] = {k.upper(): v for k, v in lines.items()}
    out["config_status"] = status
    out["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    print(json.dumps(out), flush=True)
    bad = [k for k, v in status.items() if v == "parity mismatch"]
    return PARITY_FAIL_RC if bad else 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the first ~20 C1 launches after idle run slower while the GPU clocks ramp
    # up (2.97 ms falling to 2.14 ms over about 45 ms of launches); 30 warm-up
    # steps (about 125 ms of C1 launches) keep the ramp out of the timed steps
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--config", default="all", choices=["all", "c0", "c1", "c2", "c3", "c4", "files", "ao", "profile"])
    ap.add_argument("--configs", default=",".join(CHILD_ORDER), help="--config all: which configurations to run")
    ap.add_argument("--deadline-s", type=float, default=540.0, help="--config all: wall-clock budget of the run")
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
    ap.add_argument("--cpu-seconds", type=float, default=2.0, help="target wall seconds of the CPU baseline sample")
    ap.add_argument("--node-stride", type=int, default=5, choices=[5, 8])
    # --config profile
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c3"])
    ap.add_argument("--launches", type=int, default=2)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--probe-iters", type=int, default=4096)
    ap.add_argument("--any", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): rays are independent, so N GPUs "
              "would each run this same single-GPU workload on 1/N of the rays with a replicated BVH; "
              "multi-GPU runs are not part of this build.", file=sys.stderr)
        return 2
    if args.config == "all":
        return orchestrate(args)

    import cwbvh_amd as cw
    from cwbvh_amd import workloads as W

    if cw.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        return 3
    dev = cw.device_info(0)
    log(f"device {dev['arch']} CUs={dev['compute_units']}")
    if args.config == "profile":
        print(json.dumps(profile_launches(args, cw, W)), flush=True)
        return 0
    if args.config == "files" and not (args.mesh_dat and args.samples_dat):
        print("--config files needs --mesh-dat and --samples-dat", file=sys.stderr)
        return 2
    fn = {"c2": bench_c2, "c3": bench_c3, "c4": bench_c4, "ao": bench_ao}.get(args.config)
    out = fn(args, cw, W, dev) if fn else bench_single(args, cw, W, dev, args.config)
    print(json.dumps(out), flush=True)
    par = out.get("parity")
    return PARITY_FAIL_RC if par is not None and not par.get("ok", True) else 0


if __name__ == "__main__":
    sys.exit(main())
