"""Minimal launch sequence for hardware-counter runs under rocprofv3.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m cwbvh_amd.profile_run --workload c1

Runs, in order: one HBM copy of --copy-mib MiB through `copy_u4_kernel`
(calibrates the byte counters against a known size), then --launches
CWBVH closest-hit launches and --launches BVH2 closest-hit launches over the
workload's rays, each with its scene's default trace configuration
(--loop-mode and --ray-order override it).  Prints one JSON line with the
launch sizes so a counter summary can be matched to the kernels.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c3"])
    ap.add_argument("--c3-tris", type=int, default=5_000_000)
    ap.add_argument("--c3-rays", type=int, default=8 * 1024 * 1024)
    ap.add_argument("--loop-mode", type=int, default=None, help="CWBVH schedule (default: library default)")
    ap.add_argument("--ray-order", type=int, default=None,
                    help="trace config ray_order for both kinds (default: each scene's default)")
    ap.add_argument("--launches", type=int, default=2)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--any", action="store_true")
    args = ap.parse_args()

    import cwbvh_amd as cw
    from cwbvh_amd import workloads as W

    if args.workload == "c0":
        wl = W.c0()
    elif args.workload == "c3":
        wl = W.c3(args.c3_tris, args.c3_rays)
        # the benchmark's shuffled order
        wl.rays = cw.shuffle_rays(wl.rays, seed=W.SEEDS["C3"] + 2)
    else:
        wl = W.c1(W.gpu_primary_tracer())
    bvh2, cwb = cw.build_layouts(wl.mesh)
    n = len(wl.rays)
    s_cw, s_b2 = cw.Scene(cwb), cw.Scene(bvh2)
    cfg, cfg_b2 = s_cw.default_config(), s_b2.default_config()
    if args.loop_mode is not None:
        cfg.loop_mode = args.loop_mode
    if args.ray_order is not None:
        cfg.ray_order = cfg_b2.ray_order = args.ray_order
    dr = cw.DeviceBuffer.from_numpy(wl.rays)
    dh = cw.DeviceBuffer(16 * n)
    copy_bytes = args.copy_mib << 20
    src = cw.DeviceBuffer(copy_bytes)
    dst = cw.DeviceBuffer(copy_bytes)
    src.zero()
    cw.synchronize()
    cw.check(cw.lib.cwbvh_copy_u4(dst.ptr, src.ptr, copy_bytes // 16, None))
    flags = cw.TRACE_ANY if args.any else cw.TRACE_CLOSEST
    for _ in range(args.launches):
        s_cw.trace(dr.ptr, dh.ptr, n, flags, cfg)
    for _ in range(args.launches):
        s_b2.trace(dr.ptr, dh.ptr, n, flags, cfg_b2)
    cw.synchronize()
    print(json.dumps({"workload": wl.name, "rays": n, "copy_bytes": copy_bytes, "launches": args.launches,
                      "flags": flags, "cwbvh_nodes": cwb.stats()["num_nodes"], "loop_mode": cfg.loop_mode,
                      "ray_order": {"cwbvh": cfg.ray_order, "bvh2": cfg_b2.ray_order},
                      "node_bytes": 16 * s_cw.node_stride, "triangles": wl.mesh.num_triangles,
                      "scene_bytes": {"cwbvh": s_cw.device_bytes(), "bvh2": s_b2.device_bytes()}}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
