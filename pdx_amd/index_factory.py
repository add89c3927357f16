"""Python index classes with the reference's interface (python/pdxearch/index_factory.py:1-250):
IndexPDXIVF, IndexPDXIVFSQ8, IndexPDXIVFTree, IndexPDXIVFTreeSQ8 and load_index, backed by the
pdx_amd.compiled module (host C++ index over the HIP kernels). `search_batch` is added for the
batched GPU path.
"""
from __future__ import annotations

import numpy as np

try:  # the HIP runtime must be the one torch loads (same SONAME); import torch first
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass

from pdx_amd.compiled import PDXIndex as _PDXIndex, Server, load_index as _load_index

METRIC_MAP = {"l2sq": 0, "cosine": 1, "ip": 2}
DEFAULT_NPROBE = 32


class _IndexBase:
    _TYPE = ""
    _TREE = False

    def __init__(self, *, num_dimensions: int, distance_metric: str = "l2sq", normalize: bool = True,
                 seed: int = 42, num_clusters: int = 0, num_meso_clusters: int = 0, sampling_fraction: float = 0.0,
                 kmeans_iters: int = 10, hierarchical_indexing: bool = True, n_threads: int = 0) -> None:
        self._index = _PDXIndex(
            self._TYPE, num_dimensions, METRIC_MAP[distance_metric], seed, num_clusters,
            num_meso_clusters if self._TREE else 0, normalize, sampling_fraction, kmeans_iters,
            hierarchical_indexing, n_threads,
        )

    def build(self, data: np.ndarray) -> None:
        self._index.build_index(np.ascontiguousarray(data, dtype=np.float32))

    def search(self, query: np.ndarray, knn: int, nprobe: int = DEFAULT_NPROBE):
        self._index.set_nprobe(nprobe)
        return self._index.search(np.ascontiguousarray(query, dtype=np.float32), knn)

    def search_batch(self, queries: np.ndarray, knn, nprobe=DEFAULT_NPROBE):
        """All queries in one GPU call; knn / nprobe: scalars or one value per query.
        Returns (ids [nq, kmax] u32, dists [nq, kmax] f32, counts [nq] u32)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        kk = int(knn) if np.isscalar(knn) else np.ascontiguousarray(knn, np.uint32).tolist()
        npb = int(nprobe) if np.isscalar(nprobe) else np.ascontiguousarray(nprobe, np.uint32).tolist()
        return self._index.search_batch(q, kk, npb)

    def filtered_search(self, query: np.ndarray, knn: int, row_ids: np.ndarray, nprobe: int = DEFAULT_NPROBE):
        self._index.set_nprobe(nprobe)
        return self._index.filtered_search(
            np.ascontiguousarray(query, dtype=np.float32), knn, np.ascontiguousarray(row_ids, dtype=np.uint64),
        )

    def append(self, row_id: int, embedding: np.ndarray) -> None:
        self._index.append(row_id, np.ascontiguousarray(embedding, dtype=np.float32))

    def delete(self, row_id: int) -> None:
        self._index.delete(row_id)

    def save(self, path: str) -> None:
        self._index.save(path)

    def serve(self, *, max_batch: int = 1024, max_k: int = 100, max_nprobe: int = 0, max_slots: int = 2,
              scratch_budget_bytes: int = 0, pipeline_min: int = 0) -> Server:
        """A continuous-batching server over this index's device copy: ``submit(query, knn, nprobe,
        tag)`` queues one query, ``poll(max_results, timeout_ms)`` returns completed ones as
        ``(tag, ids, distances, latency_ms)``, ``stats()`` the counters, ``autotune(queries, knn,
        nprobe)`` per-batch-size scan plans, ``close()`` stops it. Slots follow the scratch budget
        (0: ``max_slots``). Append, delete, save and build raise while a server of the index runs.
        Not in the reference, whose API serves one query per call."""
        return Server(self._index, max_batch, max_k, max_nprobe, max_slots, scratch_budget_bytes, pipeline_min)

    @property
    def num_dimensions(self) -> int:
        return self._index.get_num_dimensions()

    @property
    def num_clusters(self) -> int:
        return self._index.get_num_clusters()

    @property
    def in_memory_size_bytes(self) -> int:
        return self._index.get_in_memory_size_in_bytes()


class IndexPDXIVF(_IndexBase):
    """Single-level IVF index (F32)."""
    _TYPE = "pdx_f32"

    def __init__(self, *, num_dimensions: int, distance_metric: str = "l2sq", normalize: bool = True, seed: int = 42,
                 num_clusters: int = 0, sampling_fraction: float = 0.0, kmeans_iters: int = 10,
                 hierarchical_indexing: bool = True, n_threads: int = 0) -> None:
        super().__init__(num_dimensions=num_dimensions, distance_metric=distance_metric, normalize=normalize,
                         seed=seed, num_clusters=num_clusters, sampling_fraction=sampling_fraction,
                         kmeans_iters=kmeans_iters, hierarchical_indexing=hierarchical_indexing, n_threads=n_threads)


class IndexPDXIVFSQ8(IndexPDXIVF):
    """Single-level IVF index (U8 scalar quantization)."""
    _TYPE = "pdx_u8"


class IndexPDXIVFTree(_IndexBase):
    """Two-level IVF index (F32)."""
    _TYPE = "pdx_tree_f32"
    _TREE = True


class IndexPDXIVFTreeSQ8(_IndexBase):
    """Two-level IVF index (U8 scalar quantization)."""
    _TYPE = "pdx_tree_u8"
    _TREE = True


def load_index(path: str):
    """Load a PDX index from a single file (auto-detects type)."""
    return _load_index(path)
