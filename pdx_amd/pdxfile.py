"""In-memory PDX index model and the reference's single-file on-disk format.

Layout and format follow the reference exactly so that one file can be loaded by either side:

* PDX layout of one cluster (``include/pdx/layout.hpp:21-87``; ``cluster.hpp:341-384``), stride
  S = max_capacity vectors:
    F32 vertical dim j < v:          data[j*S + i]
    U8 vertical 4-dim group at j:    data[j*S + 4*i + (j' - j)]   (tail of r = v % 4 dims: data[j*S + r*i + t])
    horizontal 64-dim block at hd:   data[(v + hd)*S + 64*i + t]
* File (``index.hpp:213-239`` / ``483-514``; ``ivf_wrapper.hpp:133-165`` / ``314-380``;
  ``cluster.hpp:185-221``): type byte, rotation (rows, cols, floats), then the IVF or IVFTree blob
  with compact (stride = n) cluster data.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

H_DIM = 64          # common.hpp:45
U8_GROUP = 4        # common.hpp:46
MIN_MAX_CAPACITY = 256  # cluster.hpp:24

TYPE_F32, TYPE_U8, TYPE_TREE_F32, TYPE_TREE_U8 = 0, 1, 2, 3  # common.hpp:78
TYPE_NAMES = {"pdx_f32": TYPE_F32, "pdx_u8": TYPE_U8, "pdx_tree_f32": TYPE_TREE_F32, "pdx_tree_u8": TYPE_TREE_U8}


def dimension_split(d: int):
    """(horizontal, vertical) dims; common.hpp:133-159."""
    prop = np.float32(0.25) if d <= 128 else np.float32(0.75)
    h = int(np.float32(d) * prop)
    v = d - h
    if h % H_DIM > 0:
        h = ((h + H_DIM // 2) // H_DIM) * H_DIM
        v = d - h
    if not v:
        h = H_DIM
        v = d - h
    if d <= H_DIM:
        h, v = 0, d
    return h, v


def default_capacity(n: int) -> int:
    """max(n * 1.3f, 256) with float arithmetic and truncation; cluster.hpp:26-30."""
    return max(int(np.float32(n) * np.float32(1.3)), MIN_MAX_CAPACITY)


@dataclass
class Cluster:
    n_live: int
    n_used: int
    cap: int
    data: np.ndarray   # flat strided buffer, cap * d elements (float32 or uint8)
    ids: np.ndarray    # uint32[cap]
    tombstones: set = field(default_factory=set)


@dataclass
class Level:
    d: int
    v: int
    h: int
    u8: bool
    normalized: bool
    clusters: list
    centroids: np.ndarray | None
    base: float = 0.0
    scale: float = 1.0

    @property
    def nc(self) -> int:
        return len(self.clusters)

    def cluster_offsets(self) -> np.ndarray:
        caps = np.array([c.cap for c in self.clusters], np.uint64)
        return np.concatenate([[0], np.cumsum(caps)]).astype(np.uint64)


@dataclass
class IndexData:
    type: int
    matrix: np.ndarray   # d x d float32, y = x . M
    l1: Level
    l0: Level | None = None

    @property
    def is_tree(self) -> bool:
        return self.type >= 2

    @property
    def is_u8(self) -> bool:
        return self.type in (TYPE_U8, TYPE_TREE_U8)

    @property
    def d(self) -> int:
        return self.l1.d


# ------------------------------------------------------------------------------------------
# PDX transposition
# ------------------------------------------------------------------------------------------
def _regions(d: int, v: int, h: int, u8: bool):
    """(start_dim, width) of each region of the PDX layout, in storage order."""
    regs = []
    if not u8:
        regs += [(j, 1) for j in range(v)]
    else:
        j = 0
        while j + U8_GROUP <= v:
            regs.append((j, U8_GROUP))
            j += U8_GROUP
        if j < v:
            regs.append((j, v - j))
    regs += [(v + hd, H_DIM) for hd in range(0, h, H_DIM)]
    return regs


def _blocks(buf: np.ndarray, cap: int, v: int, h: int, u8: bool):
    """Views of a strided PDX buffer: (vertical blocks [groups, cap, width], vertical tail
    [cap, r] or None, horizontal blocks [h/64, cap, 64]). F32 vertical dims are groups of width 1,
    U8 ones groups of 4 with a tail of v % 4 dims."""
    w = U8_GROUP if u8 else 1
    g = v // w
    vert = buf[:g * w * cap].reshape(g, cap, w)
    r = v - g * w
    tail = buf[g * w * cap: v * cap].reshape(cap, r) if r else None
    hor = buf[v * cap: (v + h) * cap].reshape(h // H_DIM, cap, H_DIM)
    return vert, tail, hor


def to_pdx(vecs: np.ndarray, cap: int, v: int, h: int, u8: bool) -> np.ndarray:
    """Row-major vectors (n x d) -> strided PDX buffer of cap * d elements."""
    n, d = vecs.shape
    assert n <= cap
    buf = np.zeros(cap * d, np.uint8 if u8 else np.float32)
    vert, tail, hor = _blocks(buf, cap, v, h, u8)
    w = vert.shape[2]
    g = vert.shape[0]
    vert[:, :n, :] = vecs[:, :g * w].reshape(n, g, w).transpose(1, 0, 2)
    if tail is not None:
        tail[:n] = vecs[:, g * w:v]
    hor[:, :n, :] = vecs[:, v:v + h].reshape(n, h // H_DIM, H_DIM).transpose(1, 0, 2)
    return buf


def from_pdx(buf: np.ndarray, n: int, cap: int, d: int, v: int, h: int, u8: bool) -> np.ndarray:
    """Strided PDX buffer -> row-major vectors of the first n slots."""
    out = np.zeros((n, d), buf.dtype)
    vert, tail, hor = _blocks(buf, cap, v, h, u8)
    w = vert.shape[2]
    g = vert.shape[0]
    out[:, :g * w] = vert[:, :n, :].transpose(1, 0, 2).reshape(n, g * w)
    if tail is not None:
        out[:, g * w:v] = tail[:n]
    out[:, v:v + h] = hor[:, :n, :].transpose(1, 0, 2).reshape(n, h)
    return out


def write_slot(buf: np.ndarray, slot: int, vec: np.ndarray, cap: int, v: int, h: int, u8: bool) -> None:
    """Scatter one vector into slot `slot` (Cluster::InsertEmbedding, cluster.hpp:341-384)."""
    for start, width in _regions(vec.size, v, h, u8):
        buf[start * cap + slot * width: start * cap + (slot + 1) * width] = vec[start:start + width]


def make_cluster(vecs: np.ndarray, ids: np.ndarray, level: Level, cap: int | None = None) -> Cluster:
    n = vecs.shape[0]
    cap = default_capacity(n) if cap is None else cap
    idb = np.zeros(cap, np.uint32)
    idb[:n] = ids
    return Cluster(n_live=n, n_used=n, cap=cap, data=to_pdx(vecs, cap, level.v, level.h, level.u8), ids=idb)


# ------------------------------------------------------------------------------------------
# File I/O
# ------------------------------------------------------------------------------------------
def _compact(c: Cluster, L: Level) -> bytes:
    if c.tombstones:
        raise ValueError("compact the cluster before saving (tombstones present)")
    parts = []
    es = 1 if L.u8 else 4
    for start, width in _regions(L.d, L.v, L.h, L.u8):
        parts.append(c.data[start * c.cap: start * c.cap + width * c.n_live].tobytes())
    assert sum(len(p) for p in parts) == c.n_live * L.d * es
    return b"".join(parts)


def _write_level_clusters(out: list, L: Level) -> None:
    for c in L.clusters:
        out.append(struct.pack("<II", c.n_live, c.cap))
    for c in L.clusters:
        out.append(_compact(c, L))
    for c in L.clusters:
        out.append(np.ascontiguousarray(c.ids[:c.n_live], np.uint32).tobytes())


def save(index: IndexData, path: str) -> None:
    out = [struct.pack("<B", index.type)]
    m = np.ascontiguousarray(index.matrix, np.float32)
    out.append(struct.pack("<II", m.shape[0], m.shape[1]))
    out.append(m.tobytes())
    L = index.l1
    if not index.is_tree:
        out.append(struct.pack("<IIII", L.d, L.v, L.h, L.nc))
        _write_level_clusters(out, L)
        out.append(struct.pack("<B", int(L.normalized)))
        out.append(np.ascontiguousarray(L.centroids, np.float32).tobytes())
    else:
        M = index.l0
        out.append(struct.pack("<IIIII", L.d, L.v, L.h, L.nc, M.nc))
        _write_level_clusters(out, M)
        _write_level_clusters(out, L)
        out.append(struct.pack("<B", int(L.normalized)))
        out.append(np.ascontiguousarray(M.centroids, np.float32).tobytes())
    if index.is_u8:
        out.append(struct.pack("<ff", L.base, L.scale))
    with open(path, "wb") as f:
        f.write(b"".join(out))


class _Reader:
    def __init__(self, buf: bytes):
        self.b = memoryview(buf)
        self.p = 0

    def take(self, n: int) -> memoryview:
        if self.p + n > len(self.b):
            raise ValueError("index file truncated")
        v = self.b[self.p:self.p + n]
        self.p += n
        return v

    def u32(self, n=1):
        v = struct.unpack("<%dI" % n, self.take(4 * n))
        return v if n > 1 else v[0]

    def arr(self, dtype, count):
        es = np.dtype(dtype).itemsize
        return np.frombuffer(self.take(es * count), dtype=dtype).copy()


def _read_level_clusters(r: _Reader, L: Level, nc: int) -> None:
    hdr = np.frombuffer(r.take(8 * nc), np.uint32).reshape(nc, 2)
    dt = np.uint8 if L.u8 else np.float32
    for i in range(nc):
        n, cap = int(hdr[i, 0]), int(hdr[i, 1])
        if n > cap:
            raise ValueError("cluster size exceeds capacity")
        buf = np.zeros(cap * L.d, dt)
        for start, width in _regions(L.d, L.v, L.h, L.u8):
            buf[start * cap: start * cap + width * n] = r.arr(dt, width * n)
        L.clusters.append(Cluster(n_live=n, n_used=n, cap=cap, data=buf, ids=np.zeros(cap, np.uint32)))
    for c in L.clusters:
        c.ids[:c.n_live] = r.arr(np.uint32, c.n_live)


def load(path: str) -> IndexData:
    with open(path, "rb") as f:
        r = _Reader(f.read())
    t = struct.unpack("<B", r.take(1))[0]
    if t > 3:
        raise ValueError(f"Unknown PDX index type: {t}")
    rows, cols = r.u32(2)
    matrix = r.arr(np.float32, rows * cols).reshape(rows, cols)
    u8 = t in (TYPE_U8, TYPE_TREE_U8)
    if t < 2:
        d, v, h, nc = r.u32(4)
        L = Level(d, v, h, u8, False, [], None)
        _read_level_clusters(r, L, nc)
        L.normalized = bool(struct.unpack("<B", r.take(1))[0])
        L.centroids = r.arr(np.float32, nc * d).reshape(nc, d)
        l0 = None
    else:
        d, v, h, n1, n0 = r.u32(5)
        l0 = Level(d, v, h, False, False, [], None)
        _read_level_clusters(r, l0, n0)
        L = Level(d, v, h, u8, False, [], None)
        _read_level_clusters(r, L, n1)
        norm = bool(struct.unpack("<B", r.take(1))[0])
        L.normalized = l0.normalized = norm
        l0.centroids = r.arr(np.float32, n0 * d).reshape(n0, d)
    if u8:
        L.base, L.scale = struct.unpack("<ff", r.take(8))
    return IndexData(type=t, matrix=matrix, l1=L, l0=l0)
