"""cwbvh_amd: MI355X (gfx950) ray traversal -- CWBVH and BVH2, closest and any hit.

Python host side over the C ABI of ``libcwbvh_amd.so`` (include/cwbvh_amd.h).
The library is built in-tree by ``python cwbvh_amd/build.py``; importing this
package without it raises ImportError -- there is no fallback path.

Mirrors the reference's host surface:
  * ``rtBindCWBVHData`` / ``rtTraceCWBVH`` and ``rtBindBVH2Data`` / ``rtTraceBVH2``
    (reference src/TraversalKernelCWBVH.h:3-11, src/TraversalKernelBVH2.h:3-11);
  * scene + ray generators standing in for src/main.cpp:29-86 and
    src/ValidationKernels.cu;
  * ``build_layouts`` standing in for the Embree build + GPUBVHConverter
    (src/main.cpp:88-200).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().with_name("libcwbvh_amd.so")

KIND_BVH2 = 0
KIND_CWBVH = 1
TRACE_CLOSEST = 0
TRACE_ANY = 1
TRACE_PRIM_ID = 2
TRACE_STATS = 4
STACK_CAPACITY = 64
BVH2_STACK_CAPACITY = 64
SCENE_STRIDE_AUTO = 0
SCENE_STRIDE_KEEP = 1
SCENE_STRIDE_128 = 2

RAY_DTYPE = np.float32  # rays: (n, 8) float32 = origin.xyz, tmin, dir.xyz, tmax
HIT_DTYPE = np.float32  # hits: (n, 4) float32 = t, triAddr bits, u, v


class CwbvhError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"cwbvh status {status}: {message}")
        self.status = status


class BuildParams(C.Structure):
    _fields_ = [
        ("num_threads", C.c_int32),
        ("sah_bins", C.c_int32),
        ("bvh2_node_cost", C.c_float),
        ("bvh2_tri_cost", C.c_float),
        ("bvh2_max_leaf", C.c_int32),
        ("cwbvh_node_cost", C.c_float),
        ("cwbvh_tri_cost", C.c_float),
        ("cwbvh_node_stride", C.c_int32),
        ("max_depth", C.c_int32),
        ("split_alpha", C.c_float),
        ("split_budget", C.c_float),
    ]


class LayoutView(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("node_stride", C.c_int32),
        ("nodes", C.c_void_p),
        ("nodes_f4", C.c_int64),
        ("woop", C.c_void_p),
        ("woop_f4", C.c_int64),
        ("prim_idx", C.c_void_p),
        ("prim_idx_len", C.c_int64),
        ("max_depth", C.c_int32),
        ("stack_need", C.c_int32),
        ("num_nodes", C.c_int64),
        ("num_leaves", C.c_int64),
        ("num_tri_refs", C.c_int64),
        ("sah_cost", C.c_double),
    ]


class TraceConfig(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32),
        ("lds_depth", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("fetch_chunk", C.c_int32),
        ("refill_idle", C.c_int32),
        ("refill_budget", C.c_int32),
        ("postpone_pct", C.c_int32),
        ("prefetch", C.c_int32),
        ("work_slices", C.c_int32),
        ("persistent", C.c_int32),
        ("loop_mode", C.c_int32),
        ("occupancy", C.c_int32),
        ("ray_order", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("d_rays", C.c_void_p), ("d_hits", C.c_void_p), ("n", C.c_int64)]


class TuneSpace(C.Structure):
    _fields_ = [
        ("block_sizes", C.c_int32 * 4), ("n_block", C.c_int32),
        ("lds_depths", C.c_int32 * 4), ("n_lds", C.c_int32),
        ("morton", C.c_int32 * 2), ("n_morton", C.c_int32),
        ("prefetch", C.c_int32 * 2), ("n_prefetch", C.c_int32),
        ("loop_modes", C.c_int32 * 3), ("n_loop_modes", C.c_int32),
    ]


class TuneRecord(C.Structure):
    _fields_ = [("cfg", TraceConfig), ("morton", C.c_int32), ("ms", C.c_float)]

    def as_dict(self) -> dict:
        d = self.cfg.as_dict()
        d.update(morton=self.morton, ms=self.ms)
        return d


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_int64), ("max_inflight", C.c_int32), ("max_batch_requests", C.c_int32),
                ("max_batch_rays", C.c_int64), ("serial", C.c_int32), ("flags", C.c_uint32), ("trace", TraceConfig)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("rays", C.c_int64), ("batches", C.c_int64),
                ("max_inflight_seen", C.c_int32), ("max_batch_requests_seen", C.c_int32),
                ("peak_bytes", C.c_int64), ("busy_s", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class PoissonSpec(C.Structure):
    _fields_ = [("rate", C.c_double), ("num_requests", C.c_int32), ("min_rays", C.c_int64),
                ("max_rays", C.c_int64), ("seed", C.c_uint64)]


class SurfelRequest(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("radius", C.c_void_p), ("num_texels", C.c_int64),
                ("rays_per_texel", C.c_int32), ("tmin", C.c_float), ("tmax", C.c_float), ("seed", C.c_uint64),
                ("texel_hits", C.c_void_p)]


class SurfelPool(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("radius", C.c_void_p), ("num_texels", C.c_int64)]


REQUEST_HOST_RAYS, REQUEST_DEVICE_RAYS, REQUEST_SURFELS = 0, 1, 2


class PoissonResult(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("max_ms", C.c_double), ("rays", C.c_int64), ("mrays_per_s", C.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class DeviceInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("arch", C.c_char * 32),
        ("compute_units", C.c_int32),
        ("wave_size", C.c_int32),
        ("lds_per_block", C.c_int64),
        ("l2_bytes", C.c_int64),
        ("global_mem_bytes", C.c_int64),
        ("clock_khz", C.c_int32),
        ("mem_clock_khz", C.c_int32),
        ("mem_bus_width", C.c_int32),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float
_PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
_SIGS = {
    "cwbvh_abi_version": (_I32, []),
    "cwbvh_status_string": (C.c_char_p, [_I32]),
    "cwbvh_last_error": (C.c_char_p, []),
    "cwbvh_mesh_from_arrays": (_I32, [_P, _I64, _P, _I64, _PP]),
    "cwbvh_mesh_read_dat": (_I32, [C.c_char_p, _PP]),
    "cwbvh_mesh_write_dat": (_I32, [_P, _P, C.c_char_p]),
    "cwbvh_mesh_materials": (_I32, [_P, _PP, C.POINTER(_I64)]),
    "cwbvh_samples_read": (_I32, [C.c_char_p, _PP]),
    "cwbvh_samples_write": (_I32, [C.c_char_p, _I32, _I32, _P, _P, _P]),
    "cwbvh_samples_view": (_I32, [_P, C.POINTER(_I32), C.POINTER(_I32), _PP, _PP, _PP]),
    "cwbvh_samples_free": (None, [_P]),
    "cwbvh_rays_ao_from_samples": (_I32, [_P, _P, _P, _I64, _I32, _F, _F, _U64, _P]),
    "cwbvh_layout_save": (_I32, [_P, C.c_char_p]),
    "cwbvh_rays_ao_from_samples_device": (_I32, [_P, _P, _P, _I64, _I32, _F, _F, _U64, _P, _P]),
    "cwbvh_trace_surfels": (_I32, [_P, _P, _P, _P, _I64, _I32, _F, _F, _U64, _U32, _P, _P, _P, _P, _P]),
    "cwbvh_layout_load": (_I32, [C.c_char_p, _PP]),
    "cwbvh_mesh_view": (_I32, [_P, _PP, C.POINTER(_I64), _PP, C.POINTER(_I64)]),
    "cwbvh_mesh_bounds": (_I32, [_P, _P, _P]),
    "cwbvh_mesh_free": (None, [_P]),
    "cwbvh_mesh_heightfield": (_I32, [_I32, _I32, _I32, _F, _U64, _PP]),
    "cwbvh_mesh_icosphere_grid": (_I32, [_I32, _I32, _I32, _I32, _F, _F, _F, _F, _U64, _PP]),
    "cwbvh_mesh_cluster_soup": (_I32, [_I64, _I32, _F, _F, _F, _F, _U64, _PP]),
    "cwbvh_mesh_validation_planes": (_I32, [_I32, _PP]),
    "cwbvh_rays_ao_surface": (_I32, [_P, _I64, _I32, _F, _F, _F, _U64, _P]),
    "cwbvh_rays_pinhole": (_I32, [_I32, _I32, _P, _P, _P, _F, _F, _F, _P]),
    "cwbvh_rays_ao_from_hits": (_I32, [_P, _P, _P, _P, _I64, _I32, _F, _F, _U64, _P, C.POINTER(_I64)]),
    "cwbvh_rays_sphere_from_surface": (_I32, [_P, _I64, _F, _F, _F, _U64, _P]),
    "cwbvh_rays_validation": (_I32, [_I64, _I32, _P]),
    "cwbvh_rays_shuffle": (_I32, [_P, _I64, _U64]),
    "cwbvh_build_params_default": (_I32, [_P]),
    "cwbvh_build_layouts": (_I32, [_P, _P, _PP, _PP]),
    "cwbvh_layout_from_arrays": (_I32, [_I32, _I32, _P, _I64, _P, _I64, _P, _I64, _PP]),
    "cwbvh_layout_get_view": (_I32, [_P, _P]),
    "cwbvh_layout_restride": (_I32, [_P, _I32, _PP]),
    "cwbvh_layout_free": (None, [_P]),
    "cwbvh_woopify": (_I32, [_P, _P, _P, _P]),
    "cwbvh_device_count": (_I32, [C.POINTER(_I32)]),
    "cwbvh_get_device_info": (_I32, [_I32, _P]),
    "cwbvh_malloc": (_I32, [_PP, _I64]),
    "cwbvh_free": (_I32, [_P]),
    "cwbvh_memcpy_htod": (_I32, [_P, _P, _I64]),
    "cwbvh_memcpy_dtoh": (_I32, [_P, _P, _I64]),
    "cwbvh_memcpy_dtod": (_I32, [_P, _P, _I64]),
    "cwbvh_memset": (_I32, [_P, _I32, _I64]),
    "cwbvh_stream_create": (_I32, [_PP]),
    "cwbvh_stream_destroy": (_I32, [_P]),
    "cwbvh_stream_synchronize": (_I32, [_P]),
    "cwbvh_device_synchronize": (_I32, []),
    "cwbvh_event_create": (_I32, [_PP]),
    "cwbvh_event_destroy": (_I32, [_P]),
    "cwbvh_event_record": (_I32, [_P, _P]),
    "cwbvh_event_synchronize": (_I32, [_P]),
    "cwbvh_event_elapsed_ms": (_I32, [_P, _P, C.POINTER(_F)]),
    "cwbvh_copy_u4": (_I32, [_P, _P, _I64, _P]),
    "cwbvh_valu_probe": (_I32, [_I32, _I64, _I32, _P, _P]),
    "cwbvh_scene_create": (_I32, [_P, _P, _PP]),
    "cwbvh_scene_create_ex": (_I32, [_P, _P, C.c_int32, _PP]),
    "cwbvh_scene_node_stride": (_I32, [_P, C.POINTER(C.c_int32)]),
    "cwbvh_scene_wrap_device": (_I32, [_I32, _I32, _P, _P, _P, _I32, _PP]),
    "cwbvh_scene_device_bytes": (_I32, [_P, C.POINTER(_I64)]),
    "cwbvh_scene_kind": (_I32, [_P, C.POINTER(_I32)]),
    "cwbvh_scene_destroy": (None, [_P]),
    "cwbvh_trace_config_default": (_I32, [_I32, _P]),
    "cwbvh_trace": (_I32, [_P, _P, _P, _I64, _U32, _P, _P, _P]),
    "cwbvh_trace_ordered": (_I32, [_P, _P, _P, _P, _I64, _U32, _P, _P, _P]),
    "cwbvh_stack_overflowed": (_I32, [_P, C.POINTER(_I32)]),
    "cwbvh_loop_profile": (_I32, [_P, _P]),
    "cwbvh_rcp_check": (_I32, [_P, _P]),
    "cwbvh_trace_config_for_scene": (_I32, [_P, _P]),
    "cwbvh_scene_bounds": (_I32, [_P, _P, _P]),
    "cwbvh_morton_order": (_I32, [_P, _I64, _P, _P, _P, _P]),
    "cwbvh_set_workspace_limit": (_I32, [_I64]),
    "cwbvh_batch_create": (_I32, [_P, _I32, _U32, _P, _I32, _PP]),
    "cwbvh_batch_launch": (_I32, [_P, _P]),
    "cwbvh_batch_capture": (_I32, [_P, _P]),
    "cwbvh_batch_replay": (_I32, [_P, _P]),
    "cwbvh_batch_info": (_I32, [_P, C.POINTER(_I64), C.POINTER(_I32), C.POINTER(_I64)]),
    "cwbvh_batch_destroy": (None, [_P]),
    "cwbvh_autotune_batch": (_I32, [_P, _I32, _U32, _P, _I32, _P, _I32, C.POINTER(_I32), _P, _P]),
    "cwbvh_host_alloc": (_I32, [_PP, _I64]),
    "cwbvh_host_free": (_I32, [_P]),
    "cwbvh_server_config_default": (_I32, [_P]),
    "cwbvh_server_create": (_I32, [_P, _I32, _P, _PP]),
    "cwbvh_server_submit": (_I32, [_P, _I32, _P, _I64, _P, C.POINTER(_U64)]),
    "cwbvh_server_wait": (_I32, [_P, _U64, C.c_double, C.POINTER(C.c_double)]),
    "cwbvh_server_drain": (_I32, [_P]),
    "cwbvh_server_get_stats": (_I32, [_P, _P]),
    "cwbvh_server_destroy": (None, [_P]),
    "cwbvh_server_run_poisson": (_I32, [_P, _P, _P, _P, _P, _I64, _P, _P]),
    "cwbvh_server_submit_device": (_I32, [_P, _I32, _P, _I64, _P, C.POINTER(_U64)]),
    "cwbvh_server_add_scene": (_I32, [_P, _P, C.POINTER(_I32)]),
    "cwbvh_server_submit_surfels": (_I32, [_P, _I32, _P, C.POINTER(_U64)]),
    "cwbvh_server_run_poisson_device": (_I32, [_P, _P, _P, _P, _P, _I64, _P, _P]),
    "cwbvh_server_run_poisson_surfels": (_I32, [_P, _P, _P, _I32, _F, _F, _P, _I64, _P, _P]),
    "rtBindCWBVHData": (None, [_P, _P, _P]),
    "rtTraceCWBVH": (None, [_P, _P, C.c_int]),
    "rtBindBVH2Data": (None, [_P, _P, _P]),
    "rtTraceBVH2": (None, [_P, _P, C.c_int]),
    "cwbvh_last_status": (_I32, []),
    "cwbvh_profiler_start": (_I32, []),
    "cwbvh_profiler_stop": (_I32, []),
    "cwbvh_last_trace_ms": (_F, []),
    "cwbvh_set_log": (None, [_I32]),
}


def _load() -> C.CDLL:
    if not _LIB_PATH.exists():
        raise ImportError(f"{_LIB_PATH} is missing; build it with `python cwbvh_amd/build.py`")
    lib = C.CDLL(str(_LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
LIB_PATH = str(_LIB_PATH)
EXPORTED_SYMBOLS = tuple(_SIGS)


def check(status: int) -> None:
    if status != 0:
        raise CwbvhError(status, f"{lib.cwbvh_status_string(status).decode()}: {lib.cwbvh_last_error().decode()}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------------------ meshes


class Mesh:
    """Host triangle mesh (vertices float32 (nv,3), indices int32 (nt,3))."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.cwbvh_mesh_free(h)
            self._h = C.c_void_p(0)

    @staticmethod
    def _make(fn, *args) -> "Mesh":
        out = C.c_void_p()
        check(fn(*args, C.byref(out)))
        return Mesh(out.value)

    @classmethod
    def from_arrays(cls, vertices, indices) -> "Mesh":
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
        return cls._make(lib.cwbvh_mesh_from_arrays, _ptr(v), len(v), _ptr(i), len(i))

    @classmethod
    def read_dat(cls, path) -> "Mesh":
        """Read the reference's mesh file format (materials kept, see .materials)."""
        return cls._make(lib.cwbvh_mesh_read_dat, os.fsencode(path))

    def write_dat(self, path, materials=None) -> None:
        m = None if materials is None else np.ascontiguousarray(materials, np.int32).reshape(-1)
        if m is not None and len(m) != self.num_triangles:
            raise ValueError("one material per triangle")
        check(lib.cwbvh_mesh_write_dat(self._h, _ptr(m) if m is not None else None, os.fsencode(path)))

    @property
    def materials(self) -> np.ndarray:
        p, n = C.c_void_p(), C.c_int64()
        check(lib.cwbvh_mesh_materials(self._h, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int32)), shape=(n.value,)).copy()

    @classmethod
    def heightfield(cls, nx=256, ny=128, octaves=8, amp0=0.1, seed=0) -> "Mesh":
        return cls._make(lib.cwbvh_mesh_heightfield, nx, ny, octaves, amp0, seed)

    @classmethod
    def icosphere_grid(cls, gx=10, gy=10, gz=10, frequency=7, spacing=4.0, scale_min=0.5, scale_max=1.5,
                       ground_margin=0.5, seed=0) -> "Mesh":
        """ground_margin < 0 leaves out the ground square."""
        return cls._make(lib.cwbvh_mesh_icosphere_grid, gx, gy, gz, frequency, spacing, scale_min, scale_max,
                         float(ground_margin), seed)

    @classmethod
    def cluster_soup(cls, num_triangles, num_clusters=4096, sigma_min=0.005, sigma_max=0.05,
                     edge_mu=float(np.log(1e-3)), edge_sigma=0.5, seed=0) -> "Mesh":
        return cls._make(lib.cwbvh_mesh_cluster_soup, num_triangles, num_clusters, sigma_min, sigma_max, edge_mu,
                         edge_sigma, seed)

    @classmethod
    def validation_planes(cls, count) -> "Mesh":
        return cls._make(lib.cwbvh_mesh_validation_planes, count)

    def _view(self):
        vp, ip = C.c_void_p(), C.c_void_p()
        nv, nt = C.c_int64(), C.c_int64()
        check(lib.cwbvh_mesh_view(self._h, C.byref(vp), C.byref(nv), C.byref(ip), C.byref(nt)))
        return vp.value, nv.value, ip.value, nt.value

    @property
    def num_triangles(self) -> int:
        return self._view()[3]

    @property
    def num_vertices(self) -> int:
        return self._view()[1]

    @property
    def vertices(self) -> np.ndarray:
        vp, nv, _, _ = self._view()
        if nv == 0:
            return np.zeros((0, 3), np.float32)
        return np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_float)), shape=(nv, 3)).copy()

    @property
    def indices(self) -> np.ndarray:
        _, _, ip, nt = self._view()
        if nt == 0:
            return np.zeros((0, 3), np.int32)
        return np.ctypeslib.as_array(C.cast(ip, C.POINTER(C.c_int32)), shape=(nt, 3)).copy()

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        check(lib.cwbvh_mesh_bounds(self._h, _ptr(lo), _ptr(hi)))
        return lo, hi

    def diagonal(self) -> float:
        lo, hi = self.bounds()
        return float(np.linalg.norm(hi.astype(np.float64) - lo.astype(np.float64)))


# -------------------------------------------------------------------- rays


def rays_ao_surface(mesh: Mesh, num_points: int, rays_per_point: int, offset: float, tmin: float, tmax: float,
                    seed: int = 0) -> np.ndarray:
    out = np.empty((num_points * rays_per_point, 8), np.float32)
    check(lib.cwbvh_rays_ao_surface(mesh.handle, num_points, rays_per_point, offset, tmin, tmax, seed, _ptr(out)))
    return out


def read_samples(path):
    """Read the reference's surfel file: (positions (n,4), normals (n,4), radius (n,), (size_x, size_y))."""
    h = C.c_void_p()
    check(lib.cwbvh_samples_read(os.fsencode(path), C.byref(h)))
    try:
        sx, sy = C.c_int32(), C.c_int32()
        pp, pn, pr = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib.cwbvh_samples_view(h, C.byref(sx), C.byref(sy), C.byref(pp), C.byref(pn), C.byref(pr)))
        n = sx.value * sy.value
        if n == 0:
            return np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32), np.zeros(0, np.float32), (sx.value, sy.value)
        f4 = lambda p: np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n, 4)).copy()
        rad = np.ctypeslib.as_array(C.cast(pr, C.POINTER(C.c_float)), shape=(n,)).copy()
        return f4(pp), f4(pn), rad, (sx.value, sy.value)
    finally:
        lib.cwbvh_samples_free(h)


def rays_ao_from_samples_device(d_pos: int, d_nrm: int, d_rad: int, num_texels: int, rays_per_texel: int,
                                tmin: float, tmax: float, seed: int, d_out: int, stream: Stream | None = None) -> None:
    check(lib.cwbvh_rays_ao_from_samples_device(d_pos, d_nrm, d_rad, num_texels, rays_per_texel, tmin, tmax, seed,
                                                d_out, stream.handle if stream else None))


def write_samples(path, positions, normals, radius, size_x: int, size_y: int) -> None:
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 4)
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 4)
    r = np.ascontiguousarray(radius, np.float32).reshape(-1)
    if not (len(p) == len(n) == len(r) == size_x * size_y):
        raise ValueError("size_x * size_y samples of each array")
    check(lib.cwbvh_samples_write(os.fsencode(path), size_x, size_y, _ptr(p), _ptr(n), _ptr(r)))


def rays_ao_from_samples(positions, normals, radius, rays_per_texel: int = 32, tmin: float = 0.01,
                         tmax: float = 1e20, seed: int = 0) -> np.ndarray:
    """AO rays per surfel with the reference's frame, origin offset and ray
    interval (src/ValidationKernels.cu:63-101), on the library's random stream."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 4)
    n = np.ascontiguousarray(normals, np.float32).reshape(-1, 4)
    r = np.ascontiguousarray(radius, np.float32).reshape(-1)
    out = np.empty((len(p) * rays_per_texel, 8), np.float32)
    check(lib.cwbvh_rays_ao_from_samples(_ptr(p), _ptr(n), _ptr(r), len(p), rays_per_texel, tmin, tmax, seed,
                                         _ptr(out)))
    return out


def rays_pinhole(width: int, height: int, eye, target, up, fov_deg: float, tmin=0.0, tmax=1e30) -> np.ndarray:
    out = np.empty((width * height, 8), np.float32)
    e = _f32(eye)
    t = _f32(target)
    u = _f32(up)
    check(lib.cwbvh_rays_pinhole(width, height, _ptr(e), _ptr(t), _ptr(u), fov_deg, tmin, tmax, _ptr(out)))
    return out


def rays_ao_from_hits(mesh: Mesh, primary: np.ndarray, hits: np.ndarray, prim_ids: np.ndarray, rays_per_hit: int,
                      offset: float, tmax: float, seed: int = 0) -> np.ndarray:
    primary = np.ascontiguousarray(primary, np.float32)
    hits = np.ascontiguousarray(hits, np.float32)
    prim_ids = np.ascontiguousarray(prim_ids, np.int32)
    n = len(primary)
    out = np.empty((n * rays_per_hit, 8), np.float32)
    cnt = C.c_int64()
    check(lib.cwbvh_rays_ao_from_hits(mesh.handle, _ptr(primary), _ptr(hits), _ptr(prim_ids), n, rays_per_hit, offset,
                                      tmax, seed, _ptr(out), C.byref(cnt)))
    return out[: cnt.value]


def rays_sphere_from_surface(mesh: Mesh, num_rays: int, offset: float, tmin: float = 0.0, tmax: float = np.inf,
                             seed: int = 0) -> np.ndarray:
    out = np.empty((num_rays, 8), np.float32)
    check(lib.cwbvh_rays_sphere_from_surface(mesh.handle, num_rays, offset, tmin, tmax, seed, _ptr(out)))
    return out


def rays_validation(count: int, num_triangles: int) -> np.ndarray:
    out = np.empty((count, 8), np.float32)
    check(lib.cwbvh_rays_validation(count, num_triangles, _ptr(out)))
    return out


def shuffle_rays(rays: np.ndarray, seed: int = 0) -> np.ndarray:
    out = np.ascontiguousarray(rays, np.float32).copy()
    check(lib.cwbvh_rays_shuffle(_ptr(out), len(out), seed))
    return out


# ------------------------------------------------------------------ layouts


def _set_fields(struct, overrides: dict):
    """Apply keyword overrides to a ctypes struct, rejecting unknown names."""
    names = {f for f, _ in type(struct)._fields_}
    for k, v in overrides.items():
        if k not in names:
            raise TypeError(f"{type(struct).__name__} has no field {k!r}")
        setattr(struct, k, v)
    return struct


def default_build_params(**overrides) -> BuildParams:
    p = BuildParams()
    check(lib.cwbvh_build_params_default(C.byref(p)))
    return _set_fields(p, overrides)


class Layout:
    """Host GPU layout (reference BVH2 or CWBVH arrays)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.cwbvh_layout_free(h)
            self._h = C.c_void_p(0)

    def view(self) -> LayoutView:
        v = LayoutView()
        check(lib.cwbvh_layout_get_view(self._h, C.byref(v)))
        return v

    @property
    def kind(self) -> int:
        return self.view().kind

    @property
    def node_stride(self) -> int:
        return self.view().node_stride

    @property
    def nodes(self) -> np.ndarray:
        v = self.view()
        return np.ctypeslib.as_array(C.cast(v.nodes, C.POINTER(C.c_float)), shape=(v.nodes_f4, 4)).copy()

    @property
    def woop(self) -> np.ndarray:
        v = self.view()
        if v.woop_f4 == 0:
            return np.zeros((0, 4), np.float32)
        return np.ctypeslib.as_array(C.cast(v.woop, C.POINTER(C.c_float)), shape=(v.woop_f4, 4)).copy()

    @property
    def prim_idx(self) -> np.ndarray:
        v = self.view()
        if v.prim_idx_len == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(C.cast(v.prim_idx, C.POINTER(C.c_int32)), shape=(v.prim_idx_len,)).copy()

    def stats(self) -> dict:
        v = self.view()
        return {"kind": v.kind, "node_stride": v.node_stride, "nodes_f4": v.nodes_f4, "woop_f4": v.woop_f4,
                "max_depth": v.max_depth, "stack_need": v.stack_need, "num_nodes": v.num_nodes,
                "num_leaves": v.num_leaves, "num_tri_refs": v.num_tri_refs, "sah_cost": v.sah_cost}

    def save(self, path) -> None:
        """Write the layout cache file (cwbvh_layout_save)."""
        check(lib.cwbvh_layout_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path) -> "Layout":
        out = C.c_void_p()
        check(lib.cwbvh_layout_load(os.fsencode(path), C.byref(out)))
        return cls(out.value)

    def restride(self, node_stride: int) -> "Layout":
        out = C.c_void_p()
        check(lib.cwbvh_layout_restride(self._h, node_stride, C.byref(out)))
        return Layout(out.value)

    @classmethod
    def from_arrays(cls, kind: int, node_stride: int, nodes, woop, prim_idx) -> "Layout":
        n = np.ascontiguousarray(nodes, np.float32).reshape(-1, 4)
        w = np.ascontiguousarray(woop, np.float32).reshape(-1, 4)
        p = np.ascontiguousarray(prim_idx, np.int32).reshape(-1)
        out = C.c_void_p()
        check(lib.cwbvh_layout_from_arrays(kind, node_stride, _ptr(n), len(n), _ptr(w) if len(w) else None, len(w),
                                           _ptr(p) if len(p) else None, len(p), C.byref(out)))
        return cls(out.value)


def build_layouts(mesh: Mesh, params: BuildParams | None = None, bvh2: bool = True, cwbvh: bool = True):
    """Build (bvh2_layout, cwbvh_layout) from one SAH tree; either may be None."""
    a, b = C.c_void_p(), C.c_void_p()
    check(lib.cwbvh_build_layouts(mesh.handle, C.byref(params) if params is not None else None,
                                  C.byref(a) if bvh2 else None, C.byref(b) if cwbvh else None))
    return (Layout(a.value) if bvh2 else None), (Layout(b.value) if cwbvh else None)


def woopify(v0, v1, v2) -> np.ndarray:
    a, b, c = (_f32(x).reshape(3) for x in (v0, v1, v2))
    out = np.empty((3, 4), np.float32)
    check(lib.cwbvh_woopify(_ptr(a), _ptr(b), _ptr(c), _ptr(out)))
    return out


# ----------------------------------------------------------------- device


def device_count() -> int:
    n = C.c_int32()
    check(lib.cwbvh_device_count(C.byref(n)))
    return n.value


def device_info(device: int = 0) -> dict:
    d = DeviceInfo()
    check(lib.cwbvh_get_device_info(device, C.byref(d)))
    return {"name": d.name.decode(), "arch": d.arch.decode(), "compute_units": d.compute_units,
            "wave_size": d.wave_size, "lds_per_block": d.lds_per_block, "l2_bytes": d.l2_bytes,
            "global_mem_bytes": d.global_mem_bytes, "clock_khz": d.clock_khz, "mem_clock_khz": d.mem_clock_khz,
            "mem_bus_width": d.mem_bus_width}


class DeviceBuffer:
    """Owning HIP device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib.cwbvh_malloc(C.byref(p), self.nbytes))
        self._p = p

    @property
    def ptr(self) -> int:
        return self._p.value or 0

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 1))
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib.cwbvh_memcpy_htod(self._p, _ptr(a), a.nbytes))

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(lib.cwbvh_memcpy_dtoh(_ptr(out), self._p, out.nbytes))
        return out

    def zero(self) -> None:
        check(lib.cwbvh_memset(self._p, 0, self.nbytes))

    def free(self) -> None:
        if self._p is not None and self._p.value:
            lib.cwbvh_free(self._p)
            self._p = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib.cwbvh_stream_create(C.byref(s)))
        self._s = s

    @property
    def handle(self) -> C.c_void_p:
        return self._s

    def synchronize(self) -> None:
        check(lib.cwbvh_stream_synchronize(self._s))

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib.cwbvh_stream_destroy(s)
            self._s = C.c_void_p(0)


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib.cwbvh_event_create(C.byref(e)))
        self._e = e

    def record(self, stream: Stream | None = None) -> None:
        check(lib.cwbvh_event_record(self._e, stream.handle if stream else None))

    def synchronize(self) -> None:
        check(lib.cwbvh_event_synchronize(self._e))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib.cwbvh_event_elapsed_ms(self._e, end._e, C.byref(ms)))
        return ms.value

    def __del__(self):
        e = getattr(self, "_e", None)
        if e is not None and e.value:
            lib.cwbvh_event_destroy(e)
            self._e = C.c_void_p(0)


def synchronize() -> None:
    check(lib.cwbvh_device_synchronize())


def default_trace_config(kind: int, **overrides) -> TraceConfig:
    c = TraceConfig()
    check(lib.cwbvh_trace_config_default(kind, C.byref(c)))
    return _set_fields(c, overrides)


class Scene:
    """A BVH resident in device memory (owning upload of a host Layout)."""

    def __init__(self, layout: Layout, stream: Stream | None = None, stride_mode: int = 0):
        """stride_mode: SCENE_STRIDE_AUTO (128 B nodes for CWBVH scenes above
        256 MiB), SCENE_STRIDE_KEEP or SCENE_STRIDE_128."""
        s = C.c_void_p()
        check(lib.cwbvh_scene_create_ex(layout.handle, stream.handle if stream else None, stride_mode, C.byref(s)))
        self._s = s
        self.kind = layout.kind
        self._keep = ()

    @classmethod
    def wrap(cls, kind: int, node_stride: int, d_nodes: int, d_woop: int, d_prim_idx: int | None = None,
             stack_need: int = 0, keep=()) -> "Scene":
        """Non-owning scene over caller device buffers (cwbvh_scene_wrap_device);
        `keep` holds the buffers alive as long as the scene."""
        s = C.c_void_p()
        check(lib.cwbvh_scene_wrap_device(kind, node_stride, d_nodes, d_woop, d_prim_idx, stack_need, C.byref(s)))
        obj = cls.__new__(cls)
        obj._s = s
        obj.kind = kind
        obj._keep = tuple(keep)
        return obj

    @property
    def handle(self) -> C.c_void_p:
        return self._s

    @property
    def node_stride(self) -> int:
        """float4 per node of the device node array."""
        v = C.c_int32()
        check(lib.cwbvh_scene_node_stride(self._s, C.byref(v)))
        return v.value

    def device_bytes(self) -> int:
        b = C.c_int64()
        check(lib.cwbvh_scene_device_bytes(self._s, C.byref(b)))
        return b.value

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo = np.zeros(3, np.float32)
        hi = np.zeros(3, np.float32)
        check(lib.cwbvh_scene_bounds(self._s, _ptr(lo), _ptr(hi)))
        return lo, hi

    def default_config(self, **overrides) -> TraceConfig:
        """The configuration a trace without cfg uses for this scene."""
        c = TraceConfig()
        check(lib.cwbvh_trace_config_for_scene(self._s, C.byref(c)))
        return _set_fields(c, overrides)

    def trace(self, d_rays: int, d_hits: int, n: int, flags: int = TRACE_CLOSEST, cfg: TraceConfig | None = None,
              d_counters: int | None = None, stream: Stream | None = None, d_order: int | None = None) -> None:
        cfgp = C.byref(cfg) if cfg is not None else None
        st = stream.handle if stream else None
        if d_order is None:
            check(lib.cwbvh_trace(self._s, d_rays, d_hits, n, flags, cfgp, d_counters, st))
        else:
            check(lib.cwbvh_trace_ordered(self._s, d_rays, d_order, d_hits, n, flags, cfgp, d_counters, st))

    def trace_surfels(self, d_pos: int, d_nrm: int, d_rad: int, num_texels: int, rays_per_texel: int, tmin: float,
                      tmax: float, seed: int, flags: int = TRACE_CLOSEST, d_hits: int | None = None,
                      d_texel_hits: int | None = None, cfg: TraceConfig | None = None, d_counters: int | None = None,
                      stream: Stream | None = None) -> None:
        """Fused AO: rays generated from surfels inside the traversal kernel."""
        check(lib.cwbvh_trace_surfels(self._s, d_pos, d_nrm, d_rad, num_texels, rays_per_texel, tmin, tmax, seed,
                                      flags, d_hits, d_texel_hits, C.byref(cfg) if cfg is not None else None,
                                      d_counters, stream.handle if stream else None))

    def trace_numpy(self, rays: np.ndarray, flags: int = TRACE_CLOSEST, cfg: TraceConfig | None = None,
                    order: np.ndarray | None = None, with_stats: bool = False):
        """Upload rays, trace, download hits (n,4) float32; optionally node/tri counts."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        dr = DeviceBuffer.from_numpy(rays)
        dh = DeviceBuffer(max(16 * n, 16))
        do = DeviceBuffer.from_numpy(np.ascontiguousarray(order, np.int32)) if order is not None else None
        dc = None
        if with_stats:
            dc = DeviceBuffer(16)
            dc.zero()
            flags |= TRACE_STATS
        self.trace(dr.ptr, dh.ptr, n, flags, cfg, dc.ptr if dc else None, None, do.ptr if do else None)
        synchronize()
        hits = dh.download((n, 4), np.float32)
        if with_stats:
            return hits, dc.download((2,), np.uint64)
        return hits

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib.cwbvh_scene_destroy(s)
            self._s = C.c_void_p(0)


def morton_order(d_rays: int, n: int, lo, hi, d_order: int, stream: Stream | None = None) -> None:
    lo = _f32(lo)
    hi = _f32(hi)
    check(lib.cwbvh_morton_order(d_rays, n, _ptr(lo), _ptr(hi), d_order, stream.handle if stream else None))


def profiler_start() -> None:
    """Resume profiler collection (rocprofv3 --selected-regions --marker-trace)."""
    check(lib.cwbvh_profiler_start())


def profiler_stop() -> None:
    """Pause profiler collection."""
    check(lib.cwbvh_profiler_stop())


def set_workspace_limit(nbytes: int) -> None:
    """Largest sort workspace one stream may hold (0 = no limit)."""
    check(lib.cwbvh_set_workspace_limit(int(nbytes)))


def _items(items) -> C.Array:
    arr = (BatchItem * max(1, len(items)))()
    for i, (scene, d_rays, d_hits, n) in enumerate(items):
        arr[i] = BatchItem(scene.handle.value, d_rays, d_hits, n)
    return arr


class Batch:
    """One persistent multi-scene launch over items [(scene, d_rays, d_hits, n), ...]."""

    def __init__(self, items, flags: int = TRACE_CLOSEST, cfg: TraceConfig | None = None, morton: bool = False):
        self._keep = [it[0] for it in items]
        arr = _items(items)
        b = C.c_void_p()
        check(lib.cwbvh_batch_create(arr, len(items), flags, C.byref(cfg) if cfg is not None else None,
                                     1 if morton else 0, C.byref(b)))
        self._b = b

    def launch(self, stream: Stream | None = None) -> None:
        check(lib.cwbvh_batch_launch(self._b, stream.handle if stream else None))

    def capture(self, stream: Stream) -> None:
        check(lib.cwbvh_batch_capture(self._b, stream.handle))

    def replay(self, stream: Stream | None = None) -> None:
        check(lib.cwbvh_batch_replay(self._b, stream.handle if stream else None))

    def info(self) -> dict:
        r, j, b = C.c_int64(), C.c_int32(), C.c_int64()
        check(lib.cwbvh_batch_info(self._b, C.byref(r), C.byref(j), C.byref(b)))
        return {"total_rays": r.value, "num_jobs": j.value, "device_bytes": b.value}

    def __del__(self):
        b = getattr(self, "_b", None)
        if b is not None and b.value:
            lib.cwbvh_batch_destroy(b)
            self._b = C.c_void_p(0)


def autotune_batch(items, flags: int = TRACE_CLOSEST, block_sizes=(64, 128, 256), lds_depths=(0, 8, 16, 32),
                   morton=(0, 1), prefetch=(0,), loop_modes=(), reps: int = 3, stream: Stream | None = None):
    """Time every configuration of the space on the batch; returns (best, records).
    An empty loop_modes keeps the default inner-loop schedule.  An LDS depth
    of 0 stands for the deepest stack need of the items' trees."""
    sp = TuneSpace()
    for name, vals, lo, cap in (("block_sizes", block_sizes, 1, 4), ("lds_depths", lds_depths, 1, 4),
                                ("morton", morton, 1, 2), ("prefetch", prefetch, 1, 2), ("loop_modes", loop_modes, 0, 3)):
        assert lo <= len(vals) <= cap
        arr = getattr(sp, name)
        for i, v in enumerate(vals):
            arr[i] = v
    sp.n_block, sp.n_lds, sp.n_morton, sp.n_prefetch = len(block_sizes), len(lds_depths), len(morton), len(prefetch)
    sp.n_loop_modes = len(loop_modes)
    cap = 4 * 4 * 2 * 2 * 3
    recs = (TuneRecord * cap)()
    nrec = C.c_int32()
    best = TuneRecord()
    arr = _items(items)
    check(lib.cwbvh_autotune_batch(arr, len(items), flags, C.byref(sp), reps, recs, cap, C.byref(nrec),
                                   C.byref(best), stream.handle if stream else None))
    return best, [recs[i] for i in range(nrec.value)]


class PinnedArray:
    """Pinned host allocation exposed as a numpy array (freed with the object)."""

    def __init__(self, shape, dtype=np.float32):
        dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        check(lib.cwbvh_host_alloc(C.byref(p), max(nbytes, 1)))
        self._p = p
        buf = (C.c_char * max(nbytes, 1)).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    @property
    def ptr(self) -> int:
        return self._p.value

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            self.array = None
            lib.cwbvh_host_free(p)
            self._p = C.c_void_p(0)


def default_server_config(**overrides) -> ServerConfig:
    c = ServerConfig()
    check(lib.cwbvh_server_config_default(C.byref(c)))
    return _set_fields(c, overrides)


class Server:
    """Continuous-batching AO query server over resident CWBVH scenes."""

    def __init__(self, scenes, cfg: ServerConfig | None = None):
        self._scenes = list(scenes)
        arr = (C.c_void_p * len(self._scenes))(*[sc.handle.value for sc in self._scenes])
        s = C.c_void_p()
        check(lib.cwbvh_server_create(arr, len(self._scenes), C.byref(cfg) if cfg is not None else None, C.byref(s)))
        self._s = s

    def submit(self, scene: int, host_rays_ptr: int, n: int, host_hits_ptr: int) -> int:
        rid = C.c_uint64()
        check(lib.cwbvh_server_submit(self._s, scene, host_rays_ptr, n, host_hits_ptr, C.byref(rid)))
        return rid.value

    def add_scene(self, scene) -> int:
        """Admit a resident scene while serving; returns its index."""
        idx = C.c_int32()
        check(lib.cwbvh_server_add_scene(self._s, scene.handle, C.byref(idx)))
        self._scenes.append(scene)
        return idx.value

    def submit_device(self, scene: int, d_rays: int, n: int, d_hits: int) -> int:
        rid = C.c_uint64()
        check(lib.cwbvh_server_submit_device(self._s, scene, d_rays, n, d_hits, C.byref(rid)))
        return rid.value

    def submit_surfels(self, scene: int, pos_ptr: int, nrm_ptr: int, rad_ptr: int, num_texels: int,
                       rays_per_texel: int, tmin: float, tmax: float, seed: int, counts_ptr: int) -> int:
        """Host (pinned) surfel arrays in, per-surfel hit counts (uint32) out."""
        q = SurfelRequest(pos_ptr, nrm_ptr, rad_ptr, num_texels, rays_per_texel, tmin, tmax, seed, counts_ptr)
        rid = C.c_uint64()
        check(lib.cwbvh_server_submit_surfels(self._s, scene, C.byref(q), C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_s: float = -1.0) -> float:
        lat = C.c_double()
        check(lib.cwbvh_server_wait(self._s, rid, timeout_s, C.byref(lat)))
        return lat.value

    def drain(self) -> None:
        check(lib.cwbvh_server_drain(self._s))

    def stats(self) -> dict:
        st = ServerStats()
        check(lib.cwbvh_server_get_stats(self._s, C.byref(st)))
        return st.as_dict()

    def run_poisson(self, rate: float, num_requests: int, min_rays: int, max_rays: int, ray_pools, hit_pool,
                    seed: int = 0):
        """ray_pools: list of PinnedArray (n,8) per scene; hit_pool: PinnedArray (m,4)."""
        spec = PoissonSpec(rate, num_requests, min_rays, max_rays, seed)
        pools = (C.c_void_p * len(ray_pools))(*[p.ptr for p in ray_pools])
        sizes = (C.c_int64 * len(ray_pools))(*[len(p.array) for p in ray_pools])
        lat = np.zeros(num_requests, np.float64)
        res = PoissonResult()
        check(lib.cwbvh_server_run_poisson(self._s, C.byref(spec), pools, sizes, hit_pool.ptr, len(hit_pool.array),
                                           lat.ctypes.data, C.byref(res)))
        return res.as_dict(), lat

    def run_poisson_device(self, rate: float, num_requests: int, min_rays: int, max_rays: int, d_ray_pools,
                           pool_sizes, d_hit_pool: int, hit_pool_size: int, seed: int = 0):
        """d_ray_pools: device pointers (one ray pool per scene), pool_sizes: rays per pool."""
        spec = PoissonSpec(rate, num_requests, min_rays, max_rays, seed)
        pools = (C.c_void_p * len(d_ray_pools))(*d_ray_pools)
        sizes = (C.c_int64 * len(pool_sizes))(*pool_sizes)
        lat = np.zeros(num_requests, np.float64)
        res = PoissonResult()
        check(lib.cwbvh_server_run_poisson_device(self._s, C.byref(spec), pools, sizes, d_hit_pool, hit_pool_size,
                                                  lat.ctypes.data, C.byref(res)))
        return res.as_dict(), lat

    def run_poisson_surfels(self, rate: float, num_requests: int, min_rays: int, max_rays: int, pools,
                            rays_per_texel: int, tmin: float, tmax: float, count_pool, seed: int = 0):
        """pools: per scene (positions, normals, radius) PinnedArrays; count_pool: PinnedArray uint32."""
        spec = PoissonSpec(rate, num_requests, min_rays, max_rays, seed)
        arr = (SurfelPool * len(pools))(*[SurfelPool(p.ptr, nrm.ptr, r.ptr, len(p.array)) for p, nrm, r in pools])
        lat = np.zeros(num_requests, np.float64)
        res = PoissonResult()
        check(lib.cwbvh_server_run_poisson_surfels(self._s, C.byref(spec), arr, rays_per_texel, tmin, tmax,
                                                   count_pool.ptr, len(count_pool.array), lat.ctypes.data,
                                                   C.byref(res)))
        return res.as_dict(), lat

    def close(self) -> None:
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib.cwbvh_server_destroy(s)
            self._s = C.c_void_p(0)

    def __del__(self):
        self.close()


def stack_overflowed(stream: Stream | None = None) -> bool:
    f = C.c_int32()
    check(lib.cwbvh_stack_overflowed(stream.handle if stream else None, C.byref(f)))
    return bool(f.value)


def rcp_check(stream: Stream | None = None) -> dict:
    """Runs cwbvh_rcp_check over all 2^32 float patterns and returns its
    counts: in-range mismatches of the refined reciprocal (0 when the ray
    setup's identity holds on this device), in-range patterns, and
    mismatches outside the range."""
    d = DeviceBuffer(24)
    check(lib.cwbvh_rcp_check(d.ptr, stream.handle if stream else None))
    if stream is not None:
        stream.synchronize()
    else:
        synchronize()
    c = d.download((3,), np.uint64)
    return {"in_range_mismatches": int(c[0]), "in_range": int(c[1]), "outside_mismatches": int(c[2])}


def loop_profile(stream: Stream | None = None) -> dict:
    """Loop-occupancy counters of the CWBVH STATS launches on `stream` since
    the last call (cleared on read), with the lane utilization of each part."""
    out = np.zeros(6, np.uint64)
    check(lib.cwbvh_loop_profile(stream.handle if stream else None, out.ctypes.data))
    names = ("iteration_slots", "active_lanes", "node_slots", "node_steps", "tri_slots", "tri_tests")
    d = {k: int(v) for k, v in zip(names, out)}
    for part, (u, s) in {"iteration": ("active_lanes", "iteration_slots"), "node": ("node_steps", "node_slots"),
                         "tri": ("tri_tests", "tri_slots")}.items():
        d[f"{part}_util"] = d[u] / d[s] if d[s] else 0.0
    return d


# ------------------------------------------------ reference entry points


def rtBindCWBVHData(nodes_ptr: int, woop_ptr: int, map_ptr: int | None) -> None:
    lib.rtBindCWBVHData(nodes_ptr, woop_ptr, map_ptr)
    check(lib.cwbvh_last_status())


def rtTraceCWBVH(rays_ptr: int, hits_ptr: int, count: int) -> float:
    lib.rtTraceCWBVH(rays_ptr, hits_ptr, count)
    check(lib.cwbvh_last_status())
    return lib.cwbvh_last_trace_ms()


def rtBindBVH2Data(nodes_ptr: int, woop_ptr: int, map_ptr: int | None) -> None:
    lib.rtBindBVH2Data(nodes_ptr, woop_ptr, map_ptr)
    check(lib.cwbvh_last_status())


def rtTraceBVH2(rays_ptr: int, hits_ptr: int, count: int) -> float:
    lib.rtTraceBVH2(rays_ptr, hits_ptr, count)
    check(lib.cwbvh_last_status())
    return lib.cwbvh_last_trace_ms()


def set_log(enabled: bool) -> None:
    lib.cwbvh_set_log(1 if enabled else 0)


def hit_ids(hits: np.ndarray) -> np.ndarray:
    """triAddr (or prim id) column of a hit array as int32."""
    return np.ascontiguousarray(hits, np.float32).view(np.int32).reshape(-1, 4)[:, 1].copy()
