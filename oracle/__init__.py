"""CPU oracle for the traversal paths -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, always as the checker; the product (cwbvh_amd) never does.

The C restatement (oracle.c) follows the reference kernels
src/TraversalKernelCWBVH.cu:56-371 and src/TraversalKernelBVH2.cu:14-177 with
the arithmetic contract written in oracle.h.  ``build()`` compiles it into
oracle/liboracle.so with gcc.

The reference itself cannot be compiled here: every source includes
cuda_runtime.h (src/helper_math.h:28, src/WoopTriangleHelper.h:4) and the
builder needs Embree3 + TBB headers, none of which exist in this image, so
there is no oracle/_ref build.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "liboracle.so"

CFLAGS = ["-O2", "-std=c11", "-fPIC", "-shared", "-pthread", "-mfma", "-ffp-contract=off", "-fno-fast-math",
          "-Wall"]


def build() -> Path:
    """Compile oracle.c -> liboracle.so (always rebuilds)."""
    return _compile(LIB_PATH, CFLAGS)


def _compile(out: Path, flags) -> Path:
    out.parent.mkdir(exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    cmd = ["gcc", *flags, str(_DIR / "oracle.c"), "-o", str(tmp), "-lm"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    os.replace(tmp, out)
    return out


# The same restatement compiled for speed on the machine that runs it (the
# CPU baseline of bench.py): -O3 and the host's own instruction set.  FMA
# contraction stays off, so it returns the checker's bits.
NATIVE_PATH = _DIR / "_native" / "liboracle_native.so"
NATIVE_CFLAGS = ["-O3", "-march=native", "-std=c11", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
                 "-fno-fast-math", "-Wall"]


def _bind(L: C.CDLL) -> C.CDLL:
    P, I64, I32 = C.c_void_p, C.c_int64, C.c_int
    L.oracle_woop_test.argtypes = [P, P, C.c_float, C.c_float, P, P, P]
    L.oracle_woop_test.restype = I32
    L.oracle_trace_cwbvh.argtypes = [P, I32, P, P, P, I64, I32, P, I32, I32]
    L.oracle_trace_cwbvh.restype = I32
    L.oracle_trace_bvh2.argtypes = [P, P, P, P, I64, I32, P, I32, I32]
    L.oracle_trace_bvh2.restype = I32
    L.oracle_trace_brute.argtypes = [P, P, I64, P, P, I64, I32, I32]
    L.oracle_trace_brute.restype = I32
    L.oracle_woopify_ref.argtypes = [P, P, P, P]
    L.oracle_woopify_ref.restype = None
    L.oracle_cwbvh_decode.argtypes = [P, P, P]
    L.oracle_cwbvh_decode.restype = None
    return L


_lib = None
_native = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


def native_lib() -> C.CDLL:
    """The -O3 -march=native build, compiled on first use on this host."""
    global _native
    if _native is None:
        _compile(NATIVE_PATH, NATIVE_CFLAGS)
        _native = _bind(C.CDLL(str(NATIVE_PATH)))
    return _native


def _f32(a, cols):
    return np.ascontiguousarray(a, np.float32).reshape(-1, cols)


def default_threads() -> int:
    """Host threads for the oracle: the CPUs this process may run on, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def trace_cwbvh(nodes, node_stride, woop, rays, any_hit=False, threads=None, stack_capacity=64, counters=False,
                native=False):
    """Scalar CWBVH traversal; returns hits (n,4) float32 [, counters (2,) uint64].
    native=True runs the -O3 -march=native build (same bits)."""
    nodes, woop, rays = _f32(nodes, 4), _f32(woop, 4), _f32(rays, 8)
    hits = np.zeros((len(rays), 4), np.float32)
    cnt = np.zeros(2, np.uint64)
    rc = (native_lib() if native else lib()).oracle_trace_cwbvh(nodes.ctypes.data, node_stride, woop.ctypes.data, rays.ctypes.data,
                                  hits.ctypes.data, len(rays), 1 if any_hit else 0, cnt.ctypes.data,
                                  threads or default_threads(), stack_capacity)
    if rc != 0:
        raise RuntimeError("oracle CWBVH traversal stack overflow")
    return (hits, cnt) if counters else hits


def trace_bvh2(nodes, woop, rays, any_hit=False, threads=None, stack_capacity=64, counters=False, native=False):
    nodes, woop, rays = _f32(nodes, 4), _f32(woop, 4), _f32(rays, 8)
    hits = np.zeros((len(rays), 4), np.float32)
    cnt = np.zeros(2, np.uint64)
    rc = (native_lib() if native else lib()).oracle_trace_bvh2(nodes.ctypes.data, woop.ctypes.data, rays.ctypes.data, hits.ctypes.data,
                                 len(rays), 1 if any_hit else 0, cnt.ctypes.data, threads or default_threads(),
                                 stack_capacity)
    if rc != 0:
        raise RuntimeError("oracle BVH2 traversal stack overflow")
    return (hits, cnt) if counters else hits


def trace_brute(woop, tri_addrs, rays, any_hit=False, threads=None):
    woop, rays = _f32(woop, 4), _f32(rays, 8)
    addrs = np.ascontiguousarray(tri_addrs, np.int32)
    hits = np.zeros((len(rays), 4), np.float32)
    lib().oracle_trace_brute(woop.ctypes.data, addrs.ctypes.data, len(addrs), rays.ctypes.data, hits.ctypes.data,
                             len(rays), 1 if any_hit else 0, threads or default_threads())
    return hits


def woop_test(woop3, ray, tmin=None, tmax=None):
    """(hit, t, u, v) of one ray against one Woop triangle (3 float4 rows)."""
    w = _f32(woop3, 4)
    r = _f32(ray, 8)
    tmin = float(r[0, 3]) if tmin is None else tmin
    tmax = float(r[0, 7]) if tmax is None else tmax
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    ok = lib().oracle_woop_test(w.ctypes.data, r.ctypes.data, tmin, tmax, C.byref(t), C.byref(u), C.byref(v))
    return bool(ok), t.value, u.value, v.value


def woopify_ref(v0, v1, v2):
    a, b, c = (np.ascontiguousarray(x, np.float32).reshape(3) for x in (v0, v1, v2))
    out = np.zeros((3, 4), np.float32)
    lib().oracle_woopify_ref(a.ctypes.data, b.ctypes.data, c.ctypes.data, out.ctypes.data)
    return out


def cwbvh_decode(node5):
    n = _f32(node5, 4)
    lo = np.zeros((8, 3), np.float64)
    hi = np.zeros((8, 3), np.float64)
    lib().oracle_cwbvh_decode(n.ctypes.data, lo.ctypes.data, hi.ctypes.data)
    return lo, hi


def hit_ids(hits):
    return np.ascontiguousarray(hits, np.float32).view(np.int32).reshape(-1, 4)[:, 1].copy()


def compare_hits(gpu, ref, woop, rays, prim_idx=None, t_rel=1e-5, any_hit=False):
    """Classify GPU hits against oracle hits.

    Returns dict(total, exact, ties, mismatches=[indices], log=[...]).  A
    closest-hit mismatch counts as a grazing tie when the GPU's triangle is
    itself a valid hit of the ray (oracle Woop test) whose t is within t_rel of
    the oracle's.  For any-hit only the occluded/not-occluded outcome is
    compared, plus the validity of the reported triangle.  `log` holds one
    entry per ray that is not exact (ties included): its index, the ray, both
    hits and the kind of disagreement.
    """
    gpu = _f32(gpu, 4)
    ref = _f32(ref, 4)
    rays = _f32(rays, 8)
    gid, rid = hit_ids(gpu), hit_ids(ref)
    if prim_idx is not None:
        pi = np.asarray(prim_idx)
        gp = np.where(gid >= 0, pi[np.clip(gid, 0, None)], -1)
        rp = np.where(rid >= 0, pi[np.clip(rid, 0, None)], -1)
    else:
        gp, rp = gid, rid
    out = {"total": len(ref), "exact": 0, "ties": 0, "mismatches": [], "log": []}

    def log(i, kind):
        out["log"].append({"ray_index": int(i), "ray": rays[i].tolist(), "gpu_hit": gpu[i].tolist(),
                           "gpu_prim": int(gp[i]), "oracle_hit": ref[i].tolist(), "oracle_prim": int(rp[i]),
                           "kind": kind})

    if any_hit:
        same = (gp >= 0) == (rp >= 0)
        out["exact"] = int(same.sum())
        out["mismatches"] = [int(i) for i in np.nonzero(~same)[0]]
        for i in out["mismatches"]:
            log(i, "occlusion differs")
        return out
    same = (gp == rp) & ((rp < 0) | (np.abs(gpu[:, 0] - ref[:, 0]) <= t_rel * np.abs(ref[:, 0])))
    out["exact"] = int(same.sum())
    for i in np.nonzero(~same)[0]:
        if gid[i] >= 0 and rid[i] >= 0:
            ok, t, _, _ = woop_test(woop[gid[i]:gid[i] + 3], rays[i])
            if ok and abs(t - ref[i, 0]) <= t_rel * abs(ref[i, 0]):
                out["ties"] += 1
                log(i, "grazing tie")
                continue
        out["mismatches"].append(int(i))
        log(i, "mismatch")
    return out
