"""ctypes binding of the C-ABI in include/pdx_amd.h (libpdx_amd.so).

The library is the only compute path: if it is missing or no GPU is visible, every call raises.
torch is imported first when available so that the library binds to the same HIP runtime
(libamdhip64.so.7) as torch, which lets device tensors be handed over by pointer.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

try:  # one HIP runtime per process: let torch load it first
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from . import pdxfile as pf

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdx_amd.so")

PDXG_OK = 0
_STATUS = {1: "invalid argument", 2: "HIP error", 3: "unsupported on the GPU path", 4: "out of device memory"}


class PdxgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class ClusterDesc(ctypes.Structure):
    _fields_ = [("num_embeddings", ctypes.c_uint32), ("used_capacity", ctypes.c_uint32),
                ("max_capacity", ctypes.c_uint32), ("data", ctypes.c_void_p),
                ("ids", ctypes.POINTER(ctypes.c_uint32)), ("tombstones", ctypes.POINTER(ctypes.c_uint32)),
                ("num_tombstones", ctypes.c_uint32)]


class LevelDesc(ctypes.Structure):
    _fields_ = [("num_clusters", ctypes.c_uint32), ("clusters", ctypes.POINTER(ClusterDesc)),
                ("centroids", ctypes.POINTER(ctypes.c_float))]


class IndexDesc(ctypes.Structure):
    _fields_ = [("index_type", ctypes.c_uint32), ("num_dimensions", ctypes.c_uint32),
                ("num_vertical_dimensions", ctypes.c_uint32), ("num_horizontal_dimensions", ctypes.c_uint32),
                ("is_normalized", ctypes.c_uint32), ("quantization_base", ctypes.c_float),
                ("quantization_scale", ctypes.c_float), ("rotation", ctypes.POINTER(ctypes.c_float)),
                ("l1", LevelDesc), ("l0", LevelDesc)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_uint32), ("max_k", ctypes.c_uint32), ("max_nprobe", ctypes.c_uint32),
                ("max_slots", ctypes.c_uint32), ("scratch_budget_bytes", ctypes.c_size_t),
                ("scan_block", ctypes.c_uint32 * 2), ("pipeline_min", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("slots", ctypes.c_uint32), ("scratch_bytes", ctypes.c_size_t), ("gpu_ms", ctypes.c_double),
                ("host_ms", ctypes.c_double), ("turnaround_ms", ctypes.c_double),
                ("summary_capacity_bytes", ctypes.c_size_t), ("admitted_summary_bytes", ctypes.c_double),
                ("footprint_limited_batches", ctypes.c_uint64)]


class SearchArgs(ctypes.Structure):
    _fields_ = [("num_queries", ctypes.c_uint32), ("queries", ctypes.c_void_p),
                ("queries_transformed", ctypes.c_uint32), ("k", ctypes.c_void_p), ("nprobe", ctypes.c_void_p),
                ("k_all", ctypes.c_uint32), ("nprobe_all", ctypes.c_uint32), ("k_max", ctypes.c_uint32),
                ("nprobe_max", ctypes.c_uint32), ("out_stride", ctypes.c_uint32), ("out_ids", ctypes.c_void_p),
                ("out_distances", ctypes.c_void_p), ("out_int_distances", ctypes.c_void_p),
                ("out_counts", ctypes.c_void_p), ("out_bytes", ctypes.c_void_p)]


class DenseSearchArgs(ctypes.Structure):
    _fields_ = [("num_queries", ctypes.c_uint32), ("queries", ctypes.c_void_p),
                ("queries_transformed", ctypes.c_uint32), ("k", ctypes.c_uint32), ("oversample", ctypes.c_uint32),
                ("selection", ctypes.c_void_p), ("out_stride", ctypes.c_uint32), ("out_ids", ctypes.c_void_p),
                ("out_distances", ctypes.c_void_p), ("out_counts", ctypes.c_void_p)]


PDXG_DENSE_BF16, PDXG_DENSE_FP8, PDXG_DENSE_NO_RESIDUAL = 0, 1, 0x100

# every symbol include/pdx_amd.h declares
EXPORTS = ["pdxg_index_create", "pdxg_index_destroy", "pdxg_index_update_cluster", "pdxg_index_device_bytes",
           "pdxg_search", "pdxg_workspace_create", "pdxg_workspace_destroy", "pdxg_workspace_bytes",
           "pdxg_search_device", "pdxg_search_device_noalloc", "pdxg_preprocess", "pdxg_filtered_search", "pdxg_last_error",
           "pdxg_device_info", "pdxg_search_meso", "pdxg_workspace_enable_timing", "pdxg_workspace_stage_times",
           "pdxg_workspace_stage_bytes", "pdxg_workspace_meso_times", "pdxg_workspace_meso_pairs",
           "pdxg_workspace_set_plan", "pdxg_workspace_get_plan", "pdxg_workspace_autotune",
           "pdxg_workspace_set_meso_summary", "pdxg_workspace_meso_summary", "pdxg_workspace_set_order",
           "pdxg_workspace_order", "pdxg_server_create", "pdxg_server_destroy", "pdxg_server_submit", "pdxg_server_poll",
           "pdxg_server_get_stats", "pdxg_server_replay", "pdxg_server_autotune", "pdxg_server_get_plans",
           "pdxg_dense_create", "pdxg_dense_destroy",
           "pdxg_dense_search", "pdxg_dense_append", "pdxg_dense_delete", "pdxg_dense_info",
           "pdxg_dense_enable_timing", "pdxg_dense_stage_times", "pdxg_graph_create", "pdxg_graph_launch",
           "pdxg_graph_nodes", "pdxg_graph_destroy", "pdxg_build_rotation", "pdxg_build_transform",
           "pdxg_build_layout"]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (python -m pdx_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, f32p, u8p = ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint8)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.pdxg_last_error.restype = ctypes.c_char_p
    L.pdxg_index_create.argtypes = [ctypes.c_int, ctypes.POINTER(IndexDesc), ctypes.POINTER(vp)]
    L.pdxg_index_destroy.argtypes = [vp]
    L.pdxg_index_update_cluster.argtypes = [vp, u32, u32, ctypes.POINTER(ClusterDesc)]
    L.pdxg_index_device_bytes.restype = ctypes.c_size_t
    L.pdxg_index_device_bytes.argtypes = [vp]
    L.pdxg_search.argtypes = [vp, ctypes.POINTER(SearchArgs)]
    L.pdxg_workspace_create.argtypes = [vp, u32, u32, u32, ctypes.POINTER(vp)]
    L.pdxg_workspace_destroy.argtypes = [vp]
    L.pdxg_workspace_bytes.restype = ctypes.c_size_t
    L.pdxg_workspace_bytes.argtypes = [vp, u32, u32, u32]
    L.pdxg_search_device.argtypes = [vp, vp, ctypes.POINTER(SearchArgs), vp]
    L.pdxg_search_device_noalloc.argtypes = [vp, ctypes.POINTER(SearchArgs), vp]
    L.pdxg_preprocess.argtypes = [vp, f32p, u32, u32, f32p, u8p]
    L.pdxg_filtered_search.argtypes = [vp, f32p, u32, u32, u32, u8p, u32p, f32p, u32p, u32p]
    L.pdxg_device_info.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
    L.pdxg_search_meso.argtypes = [vp, f32p, u32, u32, u32, u32p, f32p, u32p]
    L.pdxg_workspace_enable_timing.argtypes = [vp, ctypes.c_int]
    L.pdxg_workspace_stage_times.argtypes = [vp, f32p]
    L.pdxg_workspace_stage_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.pdxg_workspace_meso_times.argtypes = [vp, f32p]
    L.pdxg_workspace_meso_times.restype = ctypes.c_int
    L.pdxg_workspace_meso_pairs.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.pdxg_workspace_meso_pairs.restype = ctypes.c_int
    L.pdxg_workspace_set_plan.argtypes = [vp, u32, u32]
    L.pdxg_workspace_get_plan.argtypes = [vp, u32p]
    L.pdxg_workspace_autotune.argtypes = [vp, vp, ctypes.POINTER(SearchArgs), vp]
    L.pdxg_workspace_set_meso_summary.argtypes = [vp, ctypes.c_int]
    L.pdxg_workspace_set_meso_summary.restype = ctypes.c_int
    L.pdxg_workspace_meso_summary.argtypes = [vp]
    L.pdxg_workspace_meso_summary.restype = ctypes.c_int
    L.pdxg_workspace_set_order.argtypes = [vp, ctypes.c_int]
    L.pdxg_workspace_set_order.restype = ctypes.c_int
    L.pdxg_workspace_order.argtypes = [vp]
    L.pdxg_workspace_order.restype = ctypes.c_int
    u64p, dblp = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)
    L.pdxg_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    L.pdxg_server_destroy.argtypes = [vp]
    L.pdxg_server_submit.argtypes = [vp, f32p, u32, u32, ctypes.c_uint64]
    L.pdxg_server_poll.argtypes = [vp, u32, ctypes.c_int, u64p, u32, u32p, f32p, u32p, f32p, u32p]
    L.pdxg_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.pdxg_server_replay.argtypes = [vp, f32p, u32p, u32p, dblp, u32, f32p, dblp]
    L.pdxg_server_autotune.argtypes = [vp, f32p, u32p, u32p, u32]
    L.pdxg_server_get_plans.argtypes = [vp, u32p, u32, u32p]
    L.pdxg_dense_create.argtypes = [vp, u32, ctypes.c_uint64, u32, ctypes.POINTER(vp)]
    L.pdxg_dense_destroy.argtypes = [vp]
    L.pdxg_dense_search.argtypes = [vp, ctypes.POINTER(DenseSearchArgs), vp]
    L.pdxg_dense_append.argtypes = [vp, f32p, u32p, u32, u32]
    L.pdxg_dense_delete.argtypes = [vp, u32p, u32]
    L.pdxg_dense_info.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.pdxg_dense_enable_timing.argtypes = [vp, ctypes.c_int]
    L.pdxg_dense_stage_times.argtypes = [vp, f32p]
    L.pdxg_graph_create.argtypes = [vp, vp, ctypes.POINTER(SearchArgs), ctypes.POINTER(vp)]
    L.pdxg_graph_create.restype = ctypes.c_int
    L.pdxg_graph_launch.argtypes = [vp, vp]
    L.pdxg_graph_launch.restype = ctypes.c_int
    L.pdxg_graph_nodes.argtypes = [vp]
    L.pdxg_graph_nodes.restype = ctypes.c_uint32
    L.pdxg_graph_destroy.argtypes = [vp]
    L.pdxg_graph_destroy.restype = ctypes.c_int
    L.pdxg_build_rotation.argtypes = [vp, u32, vp, vp, vp]
    L.pdxg_build_transform.argtypes = [vp, ctypes.c_uint64, u32, vp, ctypes.c_int, vp, vp, vp]
    L.pdxg_build_layout.argtypes = [vp, u32, u32, ctypes.c_int, ctypes.c_float, ctypes.c_float, vp, ctypes.c_uint64,
                                    u32, u64p, u32p, u64p, vp, ctypes.c_uint64, vp]
    for name in ("pdxg_build_rotation", "pdxg_build_transform", "pdxg_build_layout"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("pdxg_dense_create", "pdxg_dense_destroy", "pdxg_dense_search", "pdxg_dense_append",
                 "pdxg_dense_delete", "pdxg_dense_info", "pdxg_dense_enable_timing", "pdxg_dense_stage_times"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("pdxg_index_create", "pdxg_index_destroy", "pdxg_index_update_cluster", "pdxg_search",
                 "pdxg_workspace_create", "pdxg_workspace_destroy", "pdxg_search_device",
                 "pdxg_search_device_noalloc", "pdxg_preprocess",
                 "pdxg_filtered_search", "pdxg_device_info", "pdxg_search_meso", "pdxg_workspace_enable_timing",
                 "pdxg_workspace_stage_times", "pdxg_workspace_stage_bytes", "pdxg_workspace_set_plan",
                 "pdxg_workspace_get_plan", "pdxg_workspace_autotune", "pdxg_server_create", "pdxg_server_destroy", "pdxg_server_submit",
                 "pdxg_server_poll", "pdxg_server_get_stats", "pdxg_server_replay", "pdxg_server_autotune",
                 "pdxg_server_get_plans"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(status: int) -> None:
    if status != PDXG_OK:
        raise PdxgError(status, lib().pdxg_last_error().decode())


def build_rotation(g_ptr: int, d: int, q32_ptr: int, q64_ptr: int = 0, stream: int = 0) -> None:
    """pdxg_build_rotation on device buffers (g: d x d float64; q32 / q64: d x d outputs)."""
    check(lib().pdxg_build_rotation(g_ptr, d, q32_ptr, q64_ptr or None, stream or None))


def build_transform(x_ptr: int, n: int, d: int, m_ptr: int, normalize: bool, y_ptr: int, minmax_ptr: int,
                    stream: int = 0) -> None:
    """pdxg_build_transform on device buffers (x, y: n x d float32; m: d x d; minmax: 2 floats)."""
    check(lib().pdxg_build_transform(x_ptr, n, d, m_ptr, 1 if normalize else 0, y_ptr, minmax_ptr, stream or None))


def build_layout(y_ptr: int, d: int, v: int, u8: bool, base: float, scale: float, order_ptr: int, n: int,
                 first: np.ndarray, cap: np.ndarray, offset: np.ndarray, out_ptr: int, out_elems: int,
                 stream: int = 0) -> None:
    """pdxg_build_layout: device rows / order / out, host per-cluster first (nc + 1), cap, offset."""
    first = np.ascontiguousarray(first, np.uint64)
    cap = np.ascontiguousarray(cap, np.uint32)
    offset = np.ascontiguousarray(offset, np.uint64)
    nc = len(cap)
    if len(first) != nc + 1 or len(offset) != nc:
        raise ValueError("first must have nc + 1 entries and offset nc")
    u64p = ctypes.POINTER(ctypes.c_uint64)
    check(lib().pdxg_build_layout(y_ptr, d, v, 1 if u8 else 0, base, scale, order_ptr, n, nc, _ptr(first, u64p),
                                  _ptr(cap, ctypes.POINTER(ctypes.c_uint32)), _ptr(offset, u64p), out_ptr, out_elems,
                                  stream or None))


def device_info(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    check(lib().pdxg_device_info(device, buf, 256))
    return buf.value.decode()


def _ptr(a: np.ndarray, t=ctypes.c_void_p):
    return a.ctypes.data_as(t) if a is not None else None


def _cluster_desc(c: pf.Cluster, keep: list) -> ClusterDesc:
    tomb = np.array(sorted(c.tombstones), np.uint32)
    data = np.ascontiguousarray(c.data)
    ids = np.ascontiguousarray(c.ids, np.uint32)
    keep += [tomb, data, ids]
    return ClusterDesc(c.n_live, c.n_used, c.cap, data.ctypes.data_as(ctypes.c_void_p),
                       ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                       tomb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(tomb))


def _level_desc(L: pf.Level | None, keep: list) -> LevelDesc:
    if L is None:
        return LevelDesc(0, None, None)
    arr = (ClusterDesc * L.nc)(*[_cluster_desc(c, keep) for c in L.clusters])
    keep.append(arr)
    cent = None
    if L.centroids is not None:
        ca = np.ascontiguousarray(L.centroids, np.float32)
        keep.append(ca)
        cent = ca.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return LevelDesc(L.nc, arr, cent)


class GpuIndex:
    """A device-resident PDX index (pdxg_index) built from an in-memory IndexData."""

    def __init__(self, index: pf.IndexData, device: int = 0):
        self.data = index
        self.device = device
        keep = []
        m = np.ascontiguousarray(index.matrix, np.float32)
        keep.append(m)
        L = index.l1
        desc = IndexDesc(index.type, L.d, L.v, L.h, int(L.normalized), L.base, L.scale,
                         m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _level_desc(L, keep),
                         _level_desc(index.l0, keep))
        h = ctypes.c_void_p()
        check(lib().pdxg_index_create(device, ctypes.byref(desc), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().pdxg_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def device_bytes(self) -> int:
        return int(lib().pdxg_index_device_bytes(self._h))

    def update_cluster(self, level: int, cluster_id: int) -> None:
        L = self.data.l1 if level == 1 else self.data.l0
        keep = []
        cd = _cluster_desc(L.clusters[cluster_id], keep)
        check(lib().pdxg_index_update_cluster(self._h, level, cluster_id, ctypes.byref(cd)))

    def search(self, queries: np.ndarray, k, nprobe, transformed: bool = False, want_bytes: bool = False,
               want_int: bool = True):
        """Batched search; returns (ids [nq,kmax], dists, int_dists, counts, bytes)."""
        d = self.data.d
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, d)
        nq = q.shape[0]
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (nq,)))
        npb = np.ascontiguousarray(np.broadcast_to(np.asarray(nprobe, np.uint32), (nq,)))
        kmax = max(1, int(kk.max()) if nq else 1)
        npmax = 0 if (nq == 0 or (npb == 0).any()) else int(npb.max())
        ids = np.zeros((nq, kmax), np.uint32)
        dists = np.zeros((nq, kmax), np.float32)
        idists = np.zeros((nq, kmax), np.uint32) if want_int else None
        counts = np.zeros(nq, np.uint32)
        nbytes = np.zeros(nq, np.uint64) if want_bytes else None
        a = SearchArgs(nq, _ptr(q), int(transformed), _ptr(kk), _ptr(npb), 0, 0, kmax, npmax, kmax,
                       _ptr(ids), _ptr(dists), _ptr(idists), _ptr(counts), _ptr(nbytes))
        check(lib().pdxg_search(self._h, ctypes.byref(a)))
        return ids, dists, idists, counts, nbytes

    def preprocess(self, queries: np.ndarray, transformed: bool = False):
        d = self.data.d
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, d)
        rot = np.zeros_like(q)
        qu = np.zeros(q.shape, np.uint8) if self.data.is_u8 else None
        check(lib().pdxg_preprocess(self._h, _ptr(q, ctypes.POINTER(ctypes.c_float)), q.shape[0], int(transformed),
                                    _ptr(rot, ctypes.POINTER(ctypes.c_float)),
                                    _ptr(qu, ctypes.POINTER(ctypes.c_uint8)) if qu is not None else None))
        return rot, qu

    def search_meso(self, queries_transformed: np.ndarray, k: int, nprobe: int):
        """Meso-level search of a two-level index (pdxg_search_meso): (ids [nq,k], dists, counts)."""
        q = np.ascontiguousarray(queries_transformed, np.float32).reshape(-1, self.data.d)
        nq = q.shape[0]
        ids = np.zeros((nq, k), np.uint32)
        dists = np.zeros((nq, k), np.float32)
        counts = np.zeros(nq, np.uint32)
        check(lib().pdxg_search_meso(self._h, _ptr(q, ctypes.POINTER(ctypes.c_float)), nq, k, nprobe,
                                     _ptr(ids, ctypes.POINTER(ctypes.c_uint32)),
                                     _ptr(dists, ctypes.POINTER(ctypes.c_float)),
                                     _ptr(counts, ctypes.POINTER(ctypes.c_uint32))))
        return ids, dists, counts

    def filtered_search(self, query: np.ndarray, k: int, nprobe: int, selection: np.ndarray,
                        transformed: bool = False):
        q = np.ascontiguousarray(query, np.float32).reshape(self.data.d)
        sel = np.ascontiguousarray(selection, np.uint8)
        ids = np.zeros(k, np.uint32)
        dists = np.zeros(k, np.float32)
        idists = np.zeros(k, np.uint32)
        cnt = np.zeros(1, np.uint32)
        check(lib().pdxg_filtered_search(self._h, _ptr(q, ctypes.POINTER(ctypes.c_float)), int(transformed), k,
                                         nprobe, _ptr(sel, ctypes.POINTER(ctypes.c_uint8)),
                                         _ptr(ids, ctypes.POINTER(ctypes.c_uint32)),
                                         _ptr(dists, ctypes.POINTER(ctypes.c_float)),
                                         _ptr(idists, ctypes.POINTER(ctypes.c_uint32)),
                                         _ptr(cnt, ctypes.POINTER(ctypes.c_uint32))))
        n = int(cnt[0])
        return ids[:n], dists[:n], idists[:n]


def search_device_noalloc(index: GpuIndex, queries_ptr: int, nq: int, k_ptr: int, nprobe_ptr: int, k_max: int,
                          nprobe_max: int, out_stride: int, out_ids: int, out_dists: int, out_idists: int,
                          out_counts: int, out_bytes: int = 0, transformed: bool = False, stream: int = 0,
                          k_all: int = 0, nprobe_all: int = 0) -> None:
    """pdxg_search_device_noalloc: the search of Workspace.search with no workspace and no device
    allocation (device pointers as ints, 0 = NULL; asynchronous on `stream`)."""
    a = SearchArgs(nq, queries_ptr or None, int(transformed), k_ptr or None, nprobe_ptr or None, k_all,
                   nprobe_all, k_max, nprobe_max, out_stride, out_ids or None, out_dists or None,
                   out_idists or None, out_counts or None, out_bytes or None)
    check(lib().pdxg_search_device_noalloc(index.handle, ctypes.byref(a), ctypes.c_void_p(stream or 0)))


class Workspace:
    """pdxg_workspace for device-buffer searches (pdxg_search_device)."""

    def __init__(self, index: GpuIndex, max_queries: int, max_k: int, max_nprobe: int):
        self.index = index
        h = ctypes.c_void_p()
        check(lib().pdxg_workspace_create(index.handle, max_queries, max_k, max_nprobe, ctypes.byref(h)))
        self._h = h
        self.bytes = int(lib().pdxg_workspace_bytes(index.handle, max_queries, max_k, max_nprobe))

    def close(self):
        if getattr(self, "_h", None):
            lib().pdxg_workspace_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def enable_timing(self, on: bool = True) -> None:
        check(lib().pdxg_workspace_enable_timing(self._h, int(on)))

    def stage_times(self):
        """Milliseconds of [preprocess, probe, meso scan, data scan] of the last search."""
        out = np.zeros(4, np.float32)
        check(lib().pdxg_workspace_stage_times(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_plan(self, block_meso: int = 0, block_data: int = 0) -> None:
        check(lib().pdxg_workspace_set_plan(self._h, block_meso, block_data))

    def set_meso_summary(self, enable: bool) -> None:
        check(lib().pdxg_workspace_set_meso_summary(self._h, int(enable)))

    def set_order(self, mode: int) -> None:
        """0 query order, 1 longest first, 2 longest first through a persistent work queue."""
        check(lib().pdxg_workspace_set_order(self._h, int(mode)))

    @property
    def order(self) -> int:
        return int(lib().pdxg_workspace_order(self._h))

    @property
    def meso_summary(self) -> bool:
        return bool(lib().pdxg_workspace_meso_summary(self._h))

    def plan(self):
        out = np.zeros(2, np.uint32)
        check(lib().pdxg_workspace_get_plan(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return int(out[0]), int(out[1])

    def autotune(self, queries_ptr: int, nq: int, k_ptr: int, nprobe_ptr: int, k_max: int, nprobe_max: int,
                 out_stride: int, out_ids: int, out_dists: int, out_idists: int, out_counts: int, stream: int = 0,
                 k_all: int = 0, nprobe_all: int = 0):
        """Pick the fastest scan block size per level on this search; returns (meso, data)."""
        a = SearchArgs(nq, queries_ptr or None, 0, k_ptr or None, nprobe_ptr or None, k_all, nprobe_all, k_max,
                       nprobe_max, out_stride, out_ids or None, out_dists or None, out_idists or None,
                       out_counts or None, None)
        check(lib().pdxg_workspace_autotune(self.index.handle, self._h, ctypes.byref(a), ctypes.c_void_p(stream or 0)))
        return self.plan()

    def meso_times(self):
        """Milliseconds of [two-phase first phase, meso scan / replay] of the last timed search."""
        out = np.zeros(2, np.float32)
        check(lib().pdxg_workspace_meso_times(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def meso_pairs(self) -> int:
        """(query, meso member) pairs summarised by the two-phase scan in the last timed search."""
        out = np.zeros(1, np.uint64)
        check(lib().pdxg_workspace_meso_pairs(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return int(out[0])

    def stage_bytes(self):
        """Algorithmic bytes of [meso scan, data scan] of the last search."""
        out = np.zeros(2, np.uint64)
        check(lib().pdxg_workspace_stage_bytes(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return out

    def search(self, queries_ptr: int, nq: int, k_ptr: int, nprobe_ptr: int, k_max: int, nprobe_max: int,
               out_stride: int, out_ids: int, out_dists: int, out_idists: int, out_counts: int,
               out_bytes: int = 0, transformed: bool = False, stream: int = 0, k_all: int = 0,
               nprobe_all: int = 0) -> None:
        """Asynchronous search on device pointers (ints; 0 = NULL) on `stream` (hipStream_t)."""
        a = SearchArgs(nq, queries_ptr or None, int(transformed), k_ptr or None, nprobe_ptr or None, k_all,
                       nprobe_all, k_max, nprobe_max, out_stride, out_ids or None, out_dists or None,
                       out_idists or None, out_counts or None, out_bytes or None)
        check(lib().pdxg_search_device(self.index.handle, self._h, ctypes.byref(a), ctypes.c_void_p(stream or 0)))

    def capture(self, queries_ptr: int, nq: int, k_ptr: int, nprobe_ptr: int, k_max: int, nprobe_max: int,
                out_stride: int, out_ids: int, out_dists: int, out_idists: int, out_counts: int,
                out_bytes: int = 0, transformed: bool = False, k_all: int = 0, nprobe_all: int = 0) -> "Graph":
        """The search `search` would run with these arguments, captured into a HIP graph."""
        a = SearchArgs(nq, queries_ptr or None, int(transformed), k_ptr or None, nprobe_ptr or None, k_all,
                       nprobe_all, k_max, nprobe_max, out_stride, out_ids or None, out_dists or None,
                       out_idists or None, out_counts or None, out_bytes or None)
        h = ctypes.c_void_p()
        check(lib().pdxg_graph_create(self.index.handle, self._h, ctypes.byref(a), ctypes.byref(h)))
        return Graph(h, self)


class Graph:
    """A captured pdxg_search_device call (pdxg_graph_*); replays read the captured buffers."""

    def __init__(self, handle, workspace):
        self._h = handle
        self.workspace = workspace  # kept alive while the graph exists

    @property
    def nodes(self) -> int:
        return int(lib().pdxg_graph_nodes(self._h))

    def launch(self, stream: int = 0) -> None:
        check(lib().pdxg_graph_launch(self._h, ctypes.c_void_p(stream or 0)))

    def close(self):
        if getattr(self, "_h", None):
            lib().pdxg_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dense:
    """Dense low-precision first pass + exact FP32 rerank over a copy of an F32 index (pdxg_dense_*)."""

    def __init__(self, index: GpuIndex, precision: int = PDXG_DENSE_BF16, extra_rows: int = 0,
                 max_queries: int = 1024):
        self.index = index
        self.precision = precision
        h = ctypes.c_void_p()
        check(lib().pdxg_dense_create(index.handle, precision, extra_rows, max_queries, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().pdxg_dense_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        out = np.zeros(6, np.uint64)
        check(lib().pdxg_dense_info(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return {"rows": int(out[0]), "capacity": int(out[1]), "id_capacity": int(out[2]),
                "device_bytes": int(out[3]), "gemm_rows": int(out[4]), "precision": int(out[5])}

    def enable_timing(self, on: bool = True) -> None:
        check(lib().pdxg_dense_enable_timing(self._h, int(on)))

    def stage_times(self):
        """Milliseconds of [queries + filter, GEMM, top-R select, FP32 rerank] of the last search."""
        out = np.zeros(4, np.float32)
        check(lib().pdxg_dense_stage_times(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def search(self, queries_ptr: int, nq: int, k: int, oversample: int, out_stride: int, out_ids: int,
               out_dists: int, out_counts: int, selection_ptr: int = 0, transformed: bool = False,
               stream: int = 0) -> None:
        """Search on device pointers (ints; 0 = NULL) on `stream` (hipStream_t)."""
        a = DenseSearchArgs(nq, queries_ptr or None, int(transformed), k, oversample, selection_ptr or None,
                            out_stride, out_ids or None, out_dists or None, out_counts or None)
        check(lib().pdxg_dense_search(self._h, ctypes.byref(a), ctypes.c_void_p(stream or 0)))

    def append(self, rows: np.ndarray, ids: np.ndarray, transformed: bool = False) -> None:
        r = np.ascontiguousarray(rows, np.float32).reshape(-1, self.index.data.d)
        i = np.ascontiguousarray(ids, np.uint32).reshape(-1)
        assert r.shape[0] == i.size
        check(lib().pdxg_dense_append(self._h, _ptr(r, ctypes.POINTER(ctypes.c_float)),
                                      _ptr(i, ctypes.POINTER(ctypes.c_uint32)), i.size, int(transformed)))

    def delete(self, ids: np.ndarray) -> None:
        i = np.ascontiguousarray(ids, np.uint32).reshape(-1)
        check(lib().pdxg_dense_delete(self._h, _ptr(i, ctypes.POINTER(ctypes.c_uint32)), i.size))


class Server:
    """Continuous-batching server over a resident index (pdxg_server_*)."""

    def __init__(self, index: GpuIndex, max_batch: int = 1024, max_k: int = 100, max_nprobe: int = 0,
                 max_slots: int = 4, scratch_budget_bytes: int = 0, scan_block=(0, 0), pipeline_min: int = 0):
        self.index = index
        self.max_k = max_k
        cfg = ServerConfig(max_batch, max_k, max_nprobe, max_slots, scratch_budget_bytes,
                           (ctypes.c_uint32 * 2)(*scan_block), pipeline_min)
        h = ctypes.c_void_p()
        check(lib().pdxg_server_create(index.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            check(lib().pdxg_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, query: np.ndarray, k: int, nprobe: int, tag: int) -> None:
        q = np.ascontiguousarray(query, np.float32)
        check(lib().pdxg_server_submit(self._h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), k, nprobe, tag))

    def poll(self, max_results: int = 1024, timeout_ms: int = 100):
        """Returns a list of (tag, ids, dists, latency_ms)."""
        K = self.max_k
        tags = np.zeros(max_results, np.uint64)
        ids = np.zeros((max_results, K), np.uint32)
        d = np.zeros((max_results, K), np.float32)
        cnt = np.zeros(max_results, np.uint32)
        lat = np.zeros(max_results, np.float32)
        n = ctypes.c_uint32()
        p = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
        check(lib().pdxg_server_poll(self._h, max_results, timeout_ms, p(tags, ctypes.c_uint64), K,
                                     p(ids, ctypes.c_uint32), p(d, ctypes.c_float), p(cnt, ctypes.c_uint32),
                                     p(lat, ctypes.c_float), ctypes.byref(n)))
        return [(int(tags[i]), ids[i, :cnt[i]].copy(), d[i, :cnt[i]].copy(), float(lat[i])) for i in range(n.value)]

    def stats(self) -> dict:
        s = ServerStats()
        check(lib().pdxg_server_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def autotune(self, queries: np.ndarray, k, nprobe):
        """Tune the scan plan per batch size on a query sample (idle server); returns the plans as
        rows of (largest batch, meso block, data block, two-phase meso scan, scan order mode)."""
        q = np.ascontiguousarray(queries, np.float32)
        n = q.shape[0]
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (n,)))
        npb = np.ascontiguousarray(np.broadcast_to(np.asarray(nprobe, np.uint32), (n,)))
        check(lib().pdxg_server_autotune(self._h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         kk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         npb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n))
        return self.plans()

    def plans(self):
        out = np.zeros((32, 5), np.uint32)
        n = ctypes.c_uint32()
        check(lib().pdxg_server_get_plans(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 32,
                                          ctypes.byref(n)))
        return [tuple(int(v) for v in out[r]) for r in range(min(n.value, 32))]

    def replay(self, queries: np.ndarray, k, nprobe, arrival_s: np.ndarray):
        """Timed load: query i is submitted at arrival_s[i]; returns (latency_ms per query, elapsed_s)."""
        q = np.ascontiguousarray(queries, np.float32)
        n = q.shape[0]
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.uint32), (n,)))
        npb = np.ascontiguousarray(np.broadcast_to(np.asarray(nprobe, np.uint32), (n,)))
        arr = np.ascontiguousarray(arrival_s, np.float64)
        lat = np.zeros(n, np.float32)
        el = ctypes.c_double()
        check(lib().pdxg_server_replay(self._h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       kk.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       npb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                       lat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(el)))
        return lat, el.value
