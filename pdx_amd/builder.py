"""Index construction (the BuildIndex pipeline of include/pdx/index.hpp:329-394 and 640-757).

normalise (if required) -> random orthogonal rotation (FP64 QR of a seeded Gaussian matrix,
pruners/adsampling.hpp:44-57) -> global SQ8 parameters (quantizers/scalar.hpp:60-74) -> k-means
-> PDX clusters (index.hpp:101-163) -> for two-level indexes, k-means over the centroids with
floor(sqrt(nlist)) meso clusters whose members are the L1 centroids (index.hpp:713-757).

The reference clusters with SuperKMeans (absent from the reference tree); this module runs a
seeded Lloyd k-means of its own (sampled training set, spherical when the data is normalised).
Search parity does not depend on the clustering: any valid index in the reference layout is
searched identically by the oracle and the GPU path.

With a GPU the build runs on the device with the rows resident in HBM: the rotation is the FP64
Householder QR of the seeded Gaussian (pdxg_build_rotation), the rows are normalised and rotated
by the search's own kernels with the global range reduced alongside (pdxg_build_transform),
k-means runs on torch over the device-resident rows, and the quantised PDX layout of every
cluster is written by pdxg_build_layout. Without one the same steps run on the host (numpy).
"""
from __future__ import annotations

import math

import numpy as np

from . import pdxfile as pf

MIN_EMBEDDINGS_TO_SAMPLE = 30720  # clustering.hpp:22
METRICS = {"l2sq": 0, "cosine": 1, "ip": 2}


def compute_number_of_clusters(n: int) -> int:
    """common.hpp:191-199."""
    if n < 500000:
        return int(math.ceil(2 * math.sqrt(n)))
    if n < 2500000:
        return int(math.ceil(4 * math.sqrt(n)))
    return int(math.ceil(8 * math.sqrt(n)))


def rotation_gaussian(d: int, seed: int) -> np.ndarray:
    """The seeded d x d standard-normal matrix whose QR gives the rotation."""
    return np.random.default_rng(seed).standard_normal((d, d))


def random_rotation(d: int, seed: int) -> np.ndarray:
    """Random orthogonal matrix, QR in float64 with the sign of diag(R) folded into Q."""
    g = rotation_gaussian(d, seed)
    q, r = np.linalg.qr(g)
    q *= np.sign(np.diag(r))[None, :]
    return np.ascontiguousarray(q.astype(np.float32))


def normalize_rows(x: np.ndarray) -> np.ndarray:
    n = np.sqrt(np.einsum("ij,ij->i", x, x, dtype=np.float32)).astype(np.float32)
    n[n == 0] = 1.0
    return (x / n[:, None]).astype(np.float32)


def quantization_params(y: np.ndarray):
    """Global (base, scale): ScalarQuantizer::ComputeQuantizationParams (scalar.hpp:60-74)."""
    lo = np.float32(y.min())
    hi = np.float32(y.max())
    rng = np.float32(hi - lo)
    scale = np.float32(255.0) / rng if rng > 0 else np.float32(1.0)
    return float(lo), float(scale)


def quantize(y: np.ndarray, base: float, scale: float) -> np.ndarray:
    """clamp(round_half_away((y - base) * scale), 0, 255) in float32 (scalar.hpp:76-94)."""
    t = (y.astype(np.float32) - np.float32(base)) * np.float32(scale)
    r = np.sign(t) * np.floor(np.abs(t) + np.float32(0.5))
    return np.clip(r, 0, 255).astype(np.uint8)


class _Backend:
    def __init__(self, device=None):
        self.torch = None
        self.dev = None
        try:
            import torch

            if device is not False and torch.cuda.is_available():
                self.torch = torch
                self.dev = torch.device("cuda", 0 if device in (None, True) else int(device))
        except ImportError:
            pass

    def matmul(self, a: np.ndarray, b: np.ndarray, chunk: int = 1 << 18) -> np.ndarray:
        if self.torch is None:
            return a @ b
        t = self.torch
        bt = t.from_numpy(np.ascontiguousarray(b)).to(self.dev)
        out = np.empty((a.shape[0], b.shape[1]), np.float32)
        for s in range(0, a.shape[0], chunk):
            at = t.from_numpy(np.ascontiguousarray(a[s:s + chunk])).to(self.dev)
            out[s:s + chunk] = (at @ bt).cpu().numpy()
        return out

    def preprocess(self, x: np.ndarray, m: np.ndarray, normalize: bool, chunk: int = 1 << 17):
        """Rows normalised (optionally) and rotated, y = x . M; returns (y, min, max)."""
        if self.torch is None:
            y = (normalize_rows(x) if normalize else x) @ m
            return y.astype(np.float32), float(y.min()), float(y.max())
        t = self.torch
        mt = t.from_numpy(np.ascontiguousarray(m)).to(self.dev)
        y = np.empty(x.shape, np.float32)
        lo, hi = np.inf, -np.inf
        for s in range(0, x.shape[0], chunk):
            xt = t.from_numpy(np.ascontiguousarray(x[s:s + chunk])).to(self.dev)
            if normalize:
                nrm = t.linalg.vector_norm(xt, dim=1, keepdim=True)
                xt = xt / t.where(nrm == 0, t.ones_like(nrm), nrm)
            yt = xt @ mt
            lo, hi = min(lo, float(yt.min())), max(hi, float(yt.max()))
            y[s:s + chunk] = yt.cpu().numpy()
        return y, lo, hi

    def quantize(self, y: np.ndarray, base: float, scale: float, chunk: int = 1 << 18) -> np.ndarray:
        if self.torch is None:
            return quantize(y, base, scale)
        t = self.torch
        out = np.empty(y.shape, np.uint8)
        for s in range(0, y.shape[0], chunk):
            yt = t.from_numpy(np.ascontiguousarray(y[s:s + chunk])).to(self.dev)
            v = (yt - np.float32(base)) * np.float32(scale)
            r = t.sign(v) * t.floor(t.abs(v) + 0.5)
            out[s:s + chunk] = r.clamp(0, 255).to(t.uint8).cpu().numpy()
        return out

    def kmeans_resident(self, train, k: int, iters: int, rng, angular: bool) -> np.ndarray:
        """Lloyd iterations with the training sample (host array or device tensor) resident on the
        GPU; returns centroids."""
        t = self.torch
        tr = train if isinstance(train, t.Tensor) else t.from_numpy(np.ascontiguousarray(train)).to(self.dev)
        n, d = tr.shape
        cent = self.seed_plusplus(tr, k, rng)
        chunk = 1 << 16
        for _ in range(iters):
            cn = (cent * cent).sum(1)
            a = t.empty(n, dtype=t.int64, device=self.dev)
            for s in range(0, n, chunk):
                a[s:s + chunk] = t.argmin(cn[None, :] - 2.0 * (tr[s:s + chunk] @ cent.T), dim=1)
            counts = t.bincount(a, minlength=k)
            sums = t.zeros((k, d), dtype=t.float64, device=self.dev)
            for s in range(0, n, chunk):
                sums.index_add_(0, a[s:s + chunk], tr[s:s + chunk].double())
            empty = counts == 0
            cent = (sums / counts.clamp(min=1)[:, None].double()).float()
            ne = int(empty.sum())
            if ne:
                cent[empty] = tr[t.from_numpy(rng.choice(n, ne, replace=False)).to(self.dev)]
            if angular:
                nrm = t.linalg.vector_norm(cent, dim=1, keepdim=True)
                cent = cent / t.where(nrm == 0, t.ones_like(nrm), nrm)
        out = cent.cpu().numpy()
        del tr
        self.torch.cuda.empty_cache()
        return out

    def seed_plusplus(self, tr, k: int, rng, sub: int = 40):
        """k-means++ seeding (D^2 sampling) over a random subsample of at most sub*k rows (and at
        most ~128 MB of them)."""
        t = self.torch
        n, d = tr.shape
        m = min(n, sub * k, max(k, (1 << 25) // d))
        pool = tr[t.from_numpy(np.sort(rng.choice(n, m, replace=False))).to(self.dev)]
        u = t.from_numpy(rng.random(k)).to(self.dev, t.float64)
        first = int(rng.integers(0, m))
        chosen = t.empty(k, dtype=t.int64, device=self.dev)
        chosen[0] = first
        d2 = ((pool - pool[first]) ** 2).sum(1).double()
        for i in range(1, k):
            cdf = t.cumsum(d2, 0)
            j = t.searchsorted(cdf, (u[i] * cdf[-1]).reshape(1)).clamp(max=m - 1)
            chosen[i] = j[0]
            d2 = t.minimum(d2, ((pool - pool[j]) ** 2).sum(1).double())
        return pool[chosen].clone()

    def cluster_sums(self, x: np.ndarray, a: np.ndarray, k: int) -> np.ndarray:
        """Per-cluster float64 sums of the rows of x grouped by assignment a."""
        if self.torch is None:
            order = np.argsort(a, kind="stable")
            xs = x[order].astype(np.float64)
            starts = np.searchsorted(a[order], np.arange(k))
            sums = np.zeros((k, x.shape[1]), np.float64)
            nz = np.nonzero(np.bincount(a, minlength=k))[0]
            if len(nz):
                red = np.add.reduceat(xs, starts[nz], axis=0)
                sums[nz] = red
            return sums
        t = self.torch
        out = t.zeros((k, x.shape[1]), dtype=t.float64, device=self.dev)
        chunk = 1 << 16
        for s in range(0, x.shape[0], chunk):
            xt = t.from_numpy(np.ascontiguousarray(x[s:s + chunk])).to(self.dev).double()
            at = t.from_numpy(a[s:s + chunk]).to(self.dev)
            out.index_add_(0, at, xt)
        return out.cpu().numpy()

    def assign_resident(self, y, c: np.ndarray, chunk: int = 1 << 16) -> np.ndarray:
        """Nearest centroid (L2, ties: lowest index) of every row of a device tensor."""
        t = self.torch
        ct = t.from_numpy(np.ascontiguousarray(c)).to(self.dev)
        cnt = (ct * ct).sum(1)
        out = t.empty(y.shape[0], dtype=t.int64, device=self.dev)
        for s in range(0, y.shape[0], chunk):
            out[s:s + chunk] = t.argmin(cnt[None, :] - 2.0 * (y[s:s + chunk] @ ct.T), dim=1)
        return out.cpu().numpy()

    def assign(self, x: np.ndarray, c: np.ndarray, chunk: int = 1 << 16) -> np.ndarray:
        """Nearest centroid by L2 (ties: lowest index)."""
        cn = np.einsum("ij,ij->i", c, c).astype(np.float32)
        out = np.empty(x.shape[0], np.int64)
        if self.torch is None:
            for s in range(0, x.shape[0], chunk):
                xs = x[s:s + chunk]
                dd = cn[None, :] - 2.0 * (xs @ c.T)
                out[s:s + chunk] = np.argmin(dd, axis=1)
            return out
        t = self.torch
        ct = t.from_numpy(np.ascontiguousarray(c)).to(self.dev)
        cnt = t.from_numpy(cn).to(self.dev)
        for s in range(0, x.shape[0], chunk):
            xt = t.from_numpy(np.ascontiguousarray(x[s:s + chunk])).to(self.dev)
            dd = cnt[None, :] - 2.0 * (xt @ ct.T)
            out[s:s + chunk] = t.argmin(dd, dim=1).cpu().numpy()
        return out


def kmeans(x: np.ndarray, k: int, iters: int, seed: int, sampling_fraction: float, angular: bool,
           backend: _Backend):
    """Seeded Lloyd k-means; returns (centroids [k,d] float32, assignment of every row)."""
    n, d = x.shape
    if k == 1:
        return x[:1].copy(), np.zeros(n, np.int64)
    rng, rows = _sample_rows(n, k, seed, sampling_fraction)
    train = x if rows is None else x[rows]
    if backend.torch is not None:
        cent = backend.kmeans_resident(train, k, iters, rng, angular)
        return cent, backend.assign(x, cent)
    cent = train[rng.choice(train.shape[0], k, replace=False)].astype(np.float32).copy()
    for _ in range(iters):
        a = backend.assign(train, cent)
        counts = np.bincount(a, minlength=k).astype(np.float32)
        sums = backend.cluster_sums(train, a, k)
        empty = counts == 0
        cent = (sums / np.maximum(counts, 1)[:, None]).astype(np.float32)
        if empty.any():
            cent[empty] = train[rng.choice(train.shape[0], int(empty.sum()), replace=False)]
        if angular:
            cent = normalize_rows(cent)
    return cent, backend.assign(x, cent)


def _clusters_from_assignment(vecs, assign, k, level, row_ids):
    order = np.argsort(assign, kind="stable")
    bounds = np.searchsorted(assign[order], np.arange(k + 1))
    clusters = []
    for c in range(k):
        members = order[bounds[c]:bounds[c + 1]]
        clusters.append(pf.make_cluster(vecs[members], row_ids[members].astype(np.uint32), level))
    return clusters


def _sample_rows(n: int, k: int, seed: int, sampling_fraction: float):
    """(rng after drawing, sorted training rows or None for all rows) -- the draw kmeans() makes."""
    rng = np.random.default_rng(seed)
    frac = sampling_fraction if sampling_fraction > 0 else (1.0 if n < MIN_EMBEDDINGS_TO_SAMPLE else 0.3)
    ns = max(k, min(n, int(n * frac)))
    return rng, (np.sort(rng.choice(n, ns, replace=False)) if ns < n else None)


def _build_level_on_device(x: np.ndarray, m: np.ndarray | None, d: int, seed: int, nlist: int, norm: bool, u8: bool,
                           kmeans_iters: int, sampling_fraction: float, be: _Backend):
    """The device build of the data level: returns (rotation, Level)."""
    from . import _capi

    t, dev = be.torch, be.dev
    stream = t.cuda.current_stream(dev).cuda_stream
    n = x.shape[0]
    if m is None:
        g = t.from_numpy(rotation_gaussian(d, seed)).to(dev)
        mt = t.empty((d, d), dtype=t.float32, device=dev)
        _capi.build_rotation(g.data_ptr(), d, mt.data_ptr(), 0, stream)
        del g
        m = mt.cpu().numpy()
    else:
        mt = t.from_numpy(np.ascontiguousarray(m, np.float32)).to(dev)
    xt = t.from_numpy(x).to(dev)
    y = t.empty((n, d), dtype=t.float32, device=dev)
    mm = t.empty(2, dtype=t.float32, device=dev)
    _capi.build_transform(xt.data_ptr(), n, d, mt.data_ptr(), norm, y.data_ptr(), mm.data_ptr(), stream)
    del xt
    lo, hi = (float(v) for v in mm.cpu().numpy())
    h, v = pf.dimension_split(d)
    if u8:  # ScalarQuantizer::ComputeQuantizationParams (scalar.hpp:60-74)
        rng_ = np.float32(np.float32(hi) - np.float32(lo))
        base, scale = float(np.float32(lo)), float(np.float32(255.0) / rng_ if rng_ > 0 else np.float32(1.0))
    else:
        base, scale = 0.0, 1.0
    if nlist == 1:
        cent, assign = y[:1].cpu().numpy(), np.zeros(n, np.int64)
    else:
        rng, rows = _sample_rows(n, nlist, seed, sampling_fraction)
        train = y if rows is None else y[t.from_numpy(rows).to(dev)]
        cent = be.kmeans_resident(train, nlist, kmeans_iters, rng, norm)
        del train
        assign = be.assign_resident(y, cent)
    order = np.argsort(assign, kind="stable")
    first = np.searchsorted(assign[order], np.arange(nlist + 1)).astype(np.uint64)
    counts = np.diff(first).astype(np.int64)
    caps = np.array([pf.default_capacity(int(c)) for c in counts], np.uint64)
    offs = np.concatenate([[0], np.cumsum(caps * np.uint64(d))]).astype(np.uint64)
    out = t.empty(int(offs[-1]), dtype=t.uint8 if u8 else t.float32, device=dev)
    ot = t.from_numpy(order.astype(np.int32)).to(dev)
    _capi.build_layout(y.data_ptr(), d, v, u8, base, scale, ot.data_ptr(), n, first, caps, offs[:-1],
                       out.data_ptr(), int(offs[-1]), stream)
    del y, ot
    buf = out.cpu().numpy()
    del out
    L = pf.Level(d, v, h, u8, norm, [], cent, base, scale)
    for c in range(nlist):
        nc, cap, o = int(counts[c]), int(caps[c]), int(offs[c])
        ids = np.zeros(cap, np.uint32)
        ids[:nc] = order[first[c]:first[c + 1]]
        L.clusters.append(pf.Cluster(n_live=nc, n_used=nc, cap=cap, data=buf[o:o + cap * d], ids=ids))
    t.cuda.empty_cache()
    return m, L


def build_index(data: np.ndarray, index_type: str = "pdx_tree_u8", *, distance_metric: str = "l2sq",
                normalize: bool = False, seed: int = 42, num_clusters: int = 0, num_meso_clusters: int = 0,
                sampling_fraction: float = 0.0, kmeans_iters: int = 10, device=None,
                rotation: np.ndarray | None = None) -> pf.IndexData:
    """Build an index of `index_type` over `data` (n x d float32) with row ids 0..n-1."""
    t = pf.TYPE_NAMES[index_type]
    x = np.ascontiguousarray(data, np.float32)
    n, d = x.shape
    if d == 0 or d > 16384:
        raise ValueError("num_dimensions must be between 1 and 16384")
    if num_clusters and num_clusters > n:
        raise ValueError(f"num_clusters ({num_clusters}) exceeds num_embeddings ({n})")
    nlist = num_clusters or compute_number_of_clusters(n)
    norm = bool(normalize) or distance_metric in ("cosine", "ip")
    be = _Backend(device)
    u8 = t in (pf.TYPE_U8, pf.TYPE_TREE_U8)
    if be.torch is not None:
        m, L = _build_level_on_device(x, rotation, d, seed, nlist, norm, u8, kmeans_iters, sampling_fraction, be)
        cent = L.centroids
    else:
        m = random_rotation(d, seed) if rotation is None else np.ascontiguousarray(rotation, np.float32)
        y, lo, hi = be.preprocess(x, m, norm)
        h, v = pf.dimension_split(d)
        if u8:  # ScalarQuantizer::ComputeQuantizationParams (scalar.hpp:60-74)
            rng_ = np.float32(np.float32(hi) - np.float32(lo))
            base, scale = float(np.float32(lo)), float(np.float32(255.0) / rng_ if rng_ > 0 else np.float32(1.0))
        else:
            base, scale = 0.0, 1.0
        cent, assign = kmeans(y, nlist, kmeans_iters, seed, sampling_fraction, norm, be)
        L = pf.Level(d, v, h, u8, norm, [], cent, base, scale)
        stored = be.quantize(y, base, scale) if u8 else y
        L.clusters = _clusters_from_assignment(stored, assign, nlist, L, np.arange(n))
    idx = pf.IndexData(type=t, matrix=m, l1=L)
    if t >= 2:
        n0 = num_meso_clusters or int(math.sqrt(nlist))
        if n0 >= nlist:
            raise ValueError("num_meso_clusters must be smaller than num_clusters")
        c0, a0 = kmeans(cent, n0, 10, seed, 1.0, norm, be)
        M = pf.Level(d, L.v, L.h, False, norm, [], c0)
        M.clusters = _clusters_from_assignment(cent, a0, n0, M, np.arange(nlist))
        idx.l0 = M
    return idx


def float_twin(index: pf.IndexData, data: np.ndarray, device=None) -> pf.IndexData:
    """The F32 counterpart of an SQ8 index with the same clustering: every cluster keeps its members
    (in slot order), row ids and capacity, centroids and the meso level are shared, and the rows are
    stored as the normalised + rotated float32 vectors instead of their SQ8 codes. `data` holds the
    raw rows by row id."""
    twin = {pf.TYPE_U8: pf.TYPE_F32, pf.TYPE_TREE_U8: pf.TYPE_TREE_F32}
    if index.type not in twin:
        raise ValueError("float_twin needs an SQ8 index")
    L = index.l1
    be = _Backend(device)
    x = np.ascontiguousarray(data, np.float32)
    if be.torch is not None:  # the search's own normalise + rotate kernels
        from . import _capi

        t = be.torch
        stream = t.cuda.current_stream(be.dev).cuda_stream
        xt = t.from_numpy(x).to(be.dev)
        mt = t.from_numpy(np.ascontiguousarray(index.matrix, np.float32)).to(be.dev)
        yt = t.empty_like(xt)
        mm = t.empty(2, dtype=t.float32, device=be.dev)
        _capi.build_transform(xt.data_ptr(), x.shape[0], x.shape[1], mt.data_ptr(), L.normalized, yt.data_ptr(),
                              mm.data_ptr(), stream)
        y = yt.cpu().numpy()
        del xt, yt
    else:
        y, _, _ = be.preprocess(x, index.matrix, L.normalized)
    F = pf.Level(L.d, L.v, L.h, False, L.normalized, [], L.centroids)
    for c in L.clusters:
        if c.tombstones:
            raise ValueError("float_twin needs clusters without tombstones")
        ids = c.ids[:c.n_used]
        F.clusters.append(pf.make_cluster(y[ids.astype(np.int64)], ids, F, cap=c.cap))
    return pf.IndexData(type=twin[index.type], matrix=index.matrix, l1=F, l0=index.l0)
