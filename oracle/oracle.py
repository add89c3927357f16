"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates the reference PDX search (include/pdx/searcher.hpp and the kernels it
calls) on the CPU; see pdx_oracle.cpp for the function-by-function citations. Only the test
suite, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this module, and only
as the checker / baseline, never as part of the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return "avx512f" in line.split()
    except OSError:
        pass
    return False


def build() -> None:
    """Compile the oracle variants (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def isa() -> str:
    return "x86-64-v4 (AVX-512)" if _has_avx512() else "x86-64-v3 (AVX2+FMA)"


def lib():
    global _lib
    if _lib is not None:
        return _lib
    name = "liboracle_v4.so" if _has_avx512() else "liboracle_v3.so"
    # PDX_ORACLE_LIB names another build of pdx_oracle.cpp (used to check that the pinning
    # tests fail on a deliberately altered copy of the oracle)
    path = os.environ.get("PDX_ORACLE_LIB") or os.path.join(_BUILD, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
    L = ctypes.CDLL(path)
    L.orc_last_error.restype = ctypes.c_char_p
    L.orc_load.restype = ctypes.c_void_p
    L.orc_load.argtypes = [ctypes.c_char_p]
    L.orc_free.argtypes = [ctypes.c_void_p]
    L.orc_info.argtypes = [ctypes.c_void_p, _u32p]
    L.orc_quant_params.argtypes = [ctypes.c_void_p, _f32p]
    L.orc_split.argtypes = [ctypes.c_uint32, _u32p, _u32p]
    L.orc_ratios.argtypes = [ctypes.c_uint32, _f32p]
    L.orc_l2sq16.restype = ctypes.c_float
    L.orc_l2sq16.argtypes = [_f32p, _f32p, ctypes.c_uint32]
    L.orc_l2sq_u8.restype = ctypes.c_uint32
    L.orc_start_distances.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f32p]
    L.orc_l2sq_u8.argtypes = [_u8p, _u8p, ctypes.c_uint32]
    L.orc_normalize.argtypes = [_f32p, _f32p, ctypes.c_uint32]
    L.orc_quantize.argtypes = [_f32p, ctypes.c_float, ctypes.c_float, _u8p, ctypes.c_uint32]
    L.orc_preprocess.restype = ctypes.c_int
    L.orc_preprocess.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, _f32p]
    L.orc_search.restype = ctypes.c_int
    L.orc_search.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_int, _u32p, _u32p,
                             ctypes.c_uint32, _u32p, _f32p, _u32p, _u32p, _u64p, ctypes.c_int]
    L.orc_filtered_search.restype = ctypes.c_int
    L.orc_filtered_search.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.c_uint32, _u8p, _u32p, _f32p, _u32p, _u32p]
    L.orc_search_meso.restype = ctypes.c_int
    L.orc_search_meso.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  _u32p, _f32p, _u32p]
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                             _f32p, ctypes.c_uint32, ctypes.c_void_p, _f32p, ctypes.c_uint32, ctypes.c_void_p, _f32p]
    L.orc_cluster_offsets.argtypes = [ctypes.c_void_p, _u64p]
    L.orc_cluster_ids.restype = ctypes.c_uint32
    L.orc_cluster_ids.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _u32p]
    L.orc_delete_slot.restype = ctypes.c_int
    L.orc_delete_slot.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def split(d: int):
    h = ctypes.c_uint32()
    v = ctypes.c_uint32()
    lib().orc_split(d, ctypes.byref(h), ctypes.byref(v))
    return h.value, v.value


def ratios(d: int) -> np.ndarray:
    out = np.empty(d, np.float32)
    lib().orc_ratios(d, _p(out, _f32p))
    return out


def l2sq16(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return float(lib().orc_l2sq16(_p(a, _f32p), _p(b, _f32p), a.size))


def start_distances(q: np.ndarray, x: np.ndarray, v: int) -> np.ndarray:
    """F32 distances of the rows of x (row-major, rotated space) to q in the order Start computes
    them for a vector it scans whole (searcher.hpp:218-281): vertical chain over [0, v), then the
    64-dim horizontal blocks."""
    q = np.ascontiguousarray(q, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, q.size)
    out = np.empty(x.shape[0], np.float32)
    lib().orc_start_distances(_p(q, _f32p), _p(x, _f32p), x.shape[0], q.size, v, _p(out, _f32p))
    return out


def l2sq_u8(a: np.ndarray, b: np.ndarray) -> int:
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    return int(lib().orc_l2sq_u8(_p(a, _u8p), _p(b, _u8p), a.size))


def normalize(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_normalize(_p(x, _f32p), _p(out, _f32p), x.size)
    return out


def quantize(x: np.ndarray, base: float, scale: float) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.size, np.uint8)
    lib().orc_quantize(_p(x, _f32p), base, scale, _p(out, _u8p), x.size)
    return out


class _OrcCluster(ctypes.Structure):
    _fields_ = [("n_live", ctypes.c_uint32), ("n_used", ctypes.c_uint32), ("cap", ctypes.c_uint32),
                ("data", ctypes.c_void_p), ("ids", ctypes.c_void_p), ("tombstones", ctypes.c_void_p),
                ("n_tombstones", ctypes.c_uint32)]


def _clusters(level, keep):
    arr = (_OrcCluster * level.nc)()
    for i, c in enumerate(level.clusters):
        tomb = np.array(sorted(c.tombstones), np.uint32)
        data = np.ascontiguousarray(c.data)
        ids = np.ascontiguousarray(c.ids, np.uint32)
        keep += [tomb, data, ids]
        arr[i] = _OrcCluster(c.n_live, c.n_used, c.cap, data.ctypes.data, ids.ctypes.data, tomb.ctypes.data, len(tomb))
    keep.append(arr)
    return arr


class OracleIndex:
    """A PDX index loaded by the oracle from a file (or built from an in-memory IndexData)."""

    def __init__(self, path: str = None, data=None):
        if data is not None:
            keep = []
            L, M = data.l1, data.l0
            rot = np.ascontiguousarray(data.matrix, np.float32)
            c1 = np.ascontiguousarray(L.centroids, np.float32) if L.centroids is not None else None
            c0 = np.ascontiguousarray(M.centroids, np.float32) if M is not None else None
            keep += [rot, c1, c0]
            h = lib().orc_create(data.type, L.d, int(L.normalized), L.base, L.scale, _p(rot, _f32p), L.nc,
                                 ctypes.cast(_clusters(L, keep), ctypes.c_void_p),
                                 _p(c1, _f32p) if c1 is not None else None,
                                 M.nc if M is not None else 0,
                                 ctypes.cast(_clusters(M, keep), ctypes.c_void_p) if M is not None else None,
                                 _p(c0, _f32p) if c0 is not None else None)
        else:
            h = lib().orc_load(path.encode())
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = ctypes.c_void_p(h)
        info = np.zeros(9, np.uint32)
        lib().orc_info(self._h, _p(info, _u32p))
        (self.type, self.d, self.v_d, self.h_d, self.num_clusters, self.num_meso_clusters,
         self.normalized, self.max_cap, self.max_cap_l0) = [int(x) for x in info]
        qp = np.zeros(4, np.float32)
        lib().orc_quant_params(self._h, _p(qp, _f32p))
        self.quant_base, self.quant_scale, self.quant_scale_sq, self.inv_scale_sq = [float(x) for x in qp]

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    @property
    def is_u8(self) -> bool:
        return self.type in (1, 3)

    def preprocess(self, raw: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, self.d)
        out = np.empty_like(raw)
        if lib().orc_preprocess(self._h, _p(raw, _f32p), raw.shape[0], _p(out, _f32p)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return out

    def search(self, queries: np.ndarray, k, nprobe, transformed: bool = False, threads: int = 1):
        """Returns (ids [nq,kmax] u32, dists f32, idists u32, counts u32, bytes u64)."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, self.d)
        nq = q.shape[0]
        kk = np.broadcast_to(np.asarray(k, np.uint32), (nq,)).copy()
        npb = np.broadcast_to(np.asarray(nprobe, np.uint32), (nq,)).copy()
        kmax = max(1, int(kk.max()) if nq else 1)
        ids = np.zeros((nq, kmax), np.uint32)
        dists = np.zeros((nq, kmax), np.float32)
        idists = np.zeros((nq, kmax), np.uint32)
        counts = np.zeros(nq, np.uint32)
        nbytes = np.zeros(nq, np.uint64)
        rc = lib().orc_search(self._h, _p(q, _f32p), nq, int(transformed), _p(kk, _u32p), _p(npb, _u32p),
                              kmax, _p(ids, _u32p), _p(dists, _f32p), _p(idists, _u32p),
                              _p(counts, _u32p), _p(nbytes, _u64p), threads)
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return ids, dists, idists, counts, nbytes

    def search_meso(self, queries_transformed: np.ndarray, k: int, nprobe: int):
        """Meso-level search of a two-level index (the cluster choice of PDXTreeIndex::Append,
        index.hpp:596-607). Returns (ids [nq,k] L1 cluster ids, dists, counts)."""
        q = np.ascontiguousarray(queries_transformed, np.float32).reshape(-1, self.d)
        nq = q.shape[0]
        ids = np.zeros((nq, k), np.uint32)
        dists = np.zeros((nq, k), np.float32)
        counts = np.zeros(nq, np.uint32)
        if lib().orc_search_meso(self._h, _p(q, _f32p), nq, k, nprobe, _p(ids, _u32p), _p(dists, _f32p),
                                 _p(counts, _u32p)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return ids, dists, counts

    def cluster_offsets(self) -> np.ndarray:
        out = np.zeros(self.num_clusters + 1, np.uint64)
        lib().orc_cluster_offsets(self._h, _p(out, _u64p))
        return out

    def cluster_ids(self, c: int) -> np.ndarray:
        out = np.zeros(self.max_cap, np.uint32)
        n = lib().orc_cluster_ids(self._h, c, _p(out, _u32p))
        return out[:n].copy()

    def delete_slot(self, c: int, slot: int) -> None:
        if lib().orc_delete_slot(self._h, c, slot) != 0:
            raise RuntimeError("invalid slot")

    def filtered_search(self, query: np.ndarray, k: int, nprobe: int, slot_sel: np.ndarray,
                        transformed: bool = False):
        q = np.ascontiguousarray(query, np.float32).reshape(self.d)
        sel = np.ascontiguousarray(slot_sel, np.uint8)
        ids = np.zeros(max(k, 1), np.uint32)
        dists = np.zeros(max(k, 1), np.float32)
        idists = np.zeros(max(k, 1), np.uint32)
        cnt = np.zeros(1, np.uint32)
        rc = lib().orc_filtered_search(self._h, _p(q, _f32p), int(transformed), k, nprobe, _p(sel, _u8p),
                                       _p(ids, _u32p), _p(dists, _f32p), _p(idists, _u32p), _p(cnt, _u32p))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        n = int(cnt[0])
        return ids[:n], dists[:n], idists[:n]
