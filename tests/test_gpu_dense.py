"""Dense low-precision first pass (BF16 / FP8 GEMM) + exact FP32 rerank vs the oracle.

Every returned distance must equal, bit for bit, the oracle's Start-order distance of the returned
row (oracle.start_distances: vertical chain, then 64-dim horizontal blocks). Ids are checked
exactly wherever every live, passing row is a rerank candidate (k * oversample >= rows), and
through recall against the exact top-k otherwise. Distances are also compared with FP64.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleIndex
from pdx_amd import builder, pdxfile as pf
from tests.pdx_testdata import index_file, mixture

pytestmark = pytest.mark.gpu

# residual coding is on by default where the level has centroids; "_direct" codes the rows themselves
PRECISIONS = [("bf16", 0), ("fp8", 1), ("fp8_direct", 0x101)]


@pytest.fixture(scope="module")
def gpu_lib():
    from pdx_amd import _capi

    print("device:", _capi.device_info(0), flush=True)
    return _capi


def _rows(idx):
    """Row-major stored (rotated) rows and ids of every used slot, cluster by cluster."""
    L = idx.l1
    xs, ids = [], []
    for c in L.clusters:
        xs.append(pf.from_pdx(c.data, c.n_used, c.cap, L.d, L.v, L.h, False))
        ids.append(c.ids[:c.n_used])
    return np.concatenate(xs), np.concatenate(ids).astype(np.uint32)


def _exact_topk(qrot, rows, ids, v, k):
    """Exact (Start-order FP32 distance, id) top-k of each query over the given rows."""
    out_i, out_d = [], []
    for q in qrot:
        dist = orc.start_distances(q, rows, v)
        order = np.lexsort((ids, dist))[:k]
        out_i.append(ids[order])
        out_d.append(dist[order])
    return np.array(out_i), np.array(out_d)


class _Run:
    def __init__(self, gpu_lib, g, prec, nq, extra_rows=0):
        import torch

        self.torch = torch
        self.dev = torch.device("cuda", 0)
        self.dn = gpu_lib.Dense(g, prec, extra_rows=extra_rows, max_queries=nq)

    def search(self, q, k, oversample, selection=None, transformed=False):
        torch = self.torch
        nq = q.shape[0]
        qd = torch.from_numpy(np.ascontiguousarray(q, np.float32)).to(self.dev)
        ids = torch.full((nq, k), -1, dtype=torch.int32, device=self.dev)
        dd = torch.zeros((nq, k), dtype=torch.float32, device=self.dev)
        cnt = torch.zeros(nq, dtype=torch.int32, device=self.dev)
        sel = None if selection is None else torch.from_numpy(selection.astype(np.uint8)).to(self.dev)
        self.dn.search(qd.data_ptr(), nq, k, oversample, k, ids.data_ptr(), dd.data_ptr(), cnt.data_ptr(),
                       selection_ptr=0 if sel is None else sel.data_ptr(), transformed=transformed,
                       stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return ids.cpu().numpy().view(np.uint32), dd.cpu().numpy(), cnt.cpu().numpy()


def _check_exact_rerank(qrot, gi, gd, gc, row_of, rows, v):
    """Returned distances equal the Start-order distance of the returned rows; (dist, id) ascending."""
    for q in range(len(qrot)):
        n = int(gc[q])
        r = np.array([row_of[int(i)] for i in gi[q, :n]], np.int64)
        ref = orc.start_distances(qrot[q], rows[r], v)
        np.testing.assert_array_equal(gd[q, :n].view(np.uint32), ref.view(np.uint32), err_msg=f"q={q}")
        keys = list(zip(gd[q, :n].tolist(), gi[q, :n].tolist()))
        assert keys == sorted(keys), f"q={q} not ascending"
        # FP64 distance of the same rows
        d64 = ((rows[r].astype(np.float64) - qrot[q].astype(np.float64)) ** 2).sum(1)
        np.testing.assert_allclose(gd[q, :n], d64, rtol=1e-5, atol=1e-6)


def _recall(gi, ei, k):
    return np.mean([len(set(gi[q, :k].tolist()) & set(ei[q, :k].tolist())) / k for q in range(len(gi))])


@pytest.mark.parametrize("name,prec", PRECISIONS)
def test_all_rows_candidates_gives_exact_topk(gpu_lib, name, prec):
    """k * oversample >= rows: the result is the exact FP32 top-k whatever the first pass does."""
    path, x, q = index_file("pdx_f32", 240, 128, 4)
    idx = pf.load(path)
    g = gpu_lib.GpuIndex(idx)
    rows, ids = _rows(idx)
    qrot, _ = g.preprocess(q)
    run = _Run(gpu_lib, g, prec, len(q))
    for k in (1, 10):
        print(f"{name} k={k}", flush=True)
        gi, gd, gc = run.search(q, k, (250 + k - 1) // k)
        ei, ed = _exact_topk(qrot, rows, ids, idx.l1.v, k)
        assert (gc == k).all()
        np.testing.assert_array_equal(gi, ei)
        np.testing.assert_array_equal(gd.view(np.uint32), ed.view(np.uint32))


@pytest.mark.parametrize("name,prec", PRECISIONS)
@pytest.mark.parametrize("kind,n,d,nlist,norm", [("pdx_f32", 6000, 256, 24, True),
                                                 ("pdx_tree_f32", 6000, 384, 30, False)])
def test_first_pass_recall_and_exact_rerank(gpu_lib, name, prec, kind, n, d, nlist, norm):
    path, x, q0 = index_file(kind, n, d, nlist, normalize=norm)
    idx = pf.load(path)
    g = gpu_lib.GpuIndex(idx)
    rows, ids = _rows(idx)
    row_of = {int(i): r for r, i in enumerate(ids)}
    rng = np.random.default_rng(5)
    q = np.concatenate([q0, x[rng.integers(0, n, 176)] + np.float32(0.05) * rng.standard_normal((176, d), dtype=np.float32)])
    q = np.ascontiguousarray(q, np.float32)
    qrot, _ = g.preprocess(q)
    run = _Run(gpu_lib, g, prec, len(q))
    k = 10
    ei, ed = _exact_topk(qrot, rows, ids, idx.l1.v, k)
    for os_ in (1, 4):
        gi, gd, gc = run.search(q, k, os_)
        assert (gc == k).all()
        _check_exact_rerank(qrot, gi, gd, gc, row_of, rows, idx.l1.v)
        rec = _recall(gi, ei, k)
        print(f"{kind} d={d} {name} oversample={os_} recall@{k}={rec:.4f}", flush=True)
        if os_ == 4:
            assert rec >= (0.97 if name == "bf16" else 0.93)
    # pre-rotated queries give the same result
    gi2, gd2, _ = run.search(qrot, k, 4, transformed=True)
    gi1, gd1, _ = run.search(q, k, 4)
    np.testing.assert_array_equal(gi2, gi1)
    np.testing.assert_array_equal(gd2.view(np.uint32), gd1.view(np.uint32))


@pytest.mark.parametrize("name,prec", PRECISIONS)
def test_filter_delete_append(gpu_lib, name, prec):
    """10% row filter, deletes and appends (with an id replaced); every passing row is a candidate,
    so results are the exact top-k of the live passing rows."""
    path, x, q = index_file("pdx_f32", 2000, 192, 8, normalize=True)
    idx = pf.load(path)
    g = gpu_lib.GpuIndex(idx)
    rows, ids = _rows(idx)
    v = idx.l1.v
    qrot, _ = g.preprocess(q)
    run = _Run(gpu_lib, g, prec, len(q), extra_rows=64)
    info = run.dn.info()
    assert info["rows"] == len(ids) and info["id_capacity"] == int(ids.max()) + 1 + 64
    # 0x100 reports residual coding: on unless turned off (this index has centroids)
    assert info["precision"] == (prec & 0xFF) | (0 if prec & 0x100 else 0x100)
    rng = np.random.default_rng(3)
    sel = np.zeros(info["id_capacity"], np.uint8)
    sel[rng.choice(ids, 200, replace=False)] = 1
    k = 10

    def expect(live_rows, live_ids):
        m = sel[live_ids] == 1
        return _exact_topk(qrot, live_rows[m], live_ids[m], v, k)

    print(f"{name} filtered", flush=True)
    gi, gd, gc = run.search(q, k, 25, selection=sel)
    ei, ed = expect(rows, ids)
    np.testing.assert_array_equal(gi, ei)
    np.testing.assert_array_equal(gd.view(np.uint32), ed.view(np.uint32))
    assert run.dn.info()["gemm_rows"] == 200

    print(f"{name} delete", flush=True)
    gone = np.unique(gi[:, :3].reshape(-1))  # the current best passing rows
    run.dn.delete(gone)
    keep = ~np.isin(ids, gone)
    rows2, ids2 = rows[keep], ids[keep]
    gi, gd, gc = run.search(q, k, 25, selection=sel)
    ei, ed = expect(rows2, ids2)
    np.testing.assert_array_equal(gi, ei)
    np.testing.assert_array_equal(gd.view(np.uint32), ed.view(np.uint32))
    assert not np.isin(gi, gone).any()

    print(f"{name} append", flush=True)
    base = int(ids.max()) + 1
    new_raw = np.ascontiguousarray(q[:8] + np.float32(0.01) * rng.standard_normal((8, q.shape[1]), dtype=np.float32))
    new_ids = np.arange(base, base + 8, dtype=np.uint32)
    new_ids[0] = ids2[0]  # replaces a live row (an upsert)
    run.dn.append(new_raw, new_ids)
    new_rot, _ = g.preprocess(new_raw)
    sel[new_ids] = 1
    keep = ~np.isin(ids2, new_ids)
    rows3 = np.concatenate([rows2[keep], new_rot])
    ids3 = np.concatenate([ids2[keep], new_ids])
    gi, gd, gc = run.search(q, k, 25, selection=sel)
    ei, ed = expect(rows3, ids3)
    np.testing.assert_array_equal(gi, ei)
    np.testing.assert_array_equal(gd.view(np.uint32), ed.view(np.uint32))
    # each appended row is the nearest row to the query it was made from
    np.testing.assert_array_equal(gi[1:8, 0], new_ids[1:8])
    # unfiltered search after the updates: exact rerank, deleted rows absent
    gi, gd, gc = run.search(q, k, 4)
    row_of = {int(i): r for r, i in enumerate(ids3)}
    _check_exact_rerank(qrot, gi, gd, gc, row_of, rows3, v)
    assert not np.isin(gi, np.setdiff1d(gone, new_ids)).any()


def _synthetic_flat(x, nlist, seed):
    """F32 flat index with rows assigned to clusters at random (the dense pass ignores clusters)."""
    n, d = x.shape
    h, v = pf.dimension_split(d)
    M = builder.random_rotation(d, seed)
    y = np.ascontiguousarray(x @ M, np.float32)
    assign = np.random.default_rng(seed).integers(0, nlist, n)
    L = pf.Level(d, v, h, False, False, [], None)
    cents = []
    for c in range(nlist):
        s = np.nonzero(assign == c)[0]
        L.clusters.append(pf.make_cluster(y[s], s.astype(np.uint32), L))
        cents.append(y[s].mean(0) if len(s) else np.zeros(d, np.float32))
    L.centroids = np.ascontiguousarray(np.stack(cents), np.float32)
    return pf.IndexData(pf.TYPE_F32, M, L), y


@pytest.mark.parametrize("name,prec", PRECISIONS)
def test_several_gemm_chunks_and_single_query(gpu_lib, name, prec):
    """600k unit-norm rows x 2048 queries: the GEMM runs in three row chunks and the select in
    several segments per chunk; one query alone gives the same answer as in the batch (a query's
    low-precision form does not depend on its batch)."""
    n, d, nq, k = 600_000, 64, 2048, 10
    x, q = mixture(n, d, 256, 0.05, 21, unit=True, n_queries=nq)
    idx, y = _synthetic_flat(x, 48, 4)
    rows, ids = _rows(idx)
    row_of = np.empty(n, np.int64)
    row_of[ids] = np.arange(n)
    g = gpu_lib.GpuIndex(idx)
    run = _Run(gpu_lib, g, prec, nq)
    print(f"{name} batch {nq}", flush=True)
    gi, gd, gc = run.search(q, k, 4)
    assert (gc == k).all()
    qrot, _ = g.preprocess(q)
    sub = np.arange(0, nq, 16)
    _check_exact_rerank(qrot[sub], gi[sub], gd[sub], gc[sub], row_of, rows, idx.l1.v)
    # recall against FP64 brute force on a subset
    q64 = qrot[sub].astype(np.float64)
    r64 = rows.astype(np.float64)
    best = []
    for qq in q64:
        dd = ((r64 - qq) ** 2).sum(1)
        best.append(ids[np.argsort(dd, kind="stable")[:k]])
    rec = _recall(gi[sub], np.array(best), k)
    print(f"{name} recall@{k} vs fp64 = {rec:.4f}", flush=True)
    assert rec >= (0.97 if name == "bf16" else 0.8)
    print(f"{name} single query", flush=True)
    for t in (0, 777):
        si, sd, sc = run.search(q[t:t + 1], k, 4)
        np.testing.assert_array_equal(si[0], gi[t])
        np.testing.assert_array_equal(sd[0].view(np.uint32), gd[t].view(np.uint32))


def test_residual_coding_raises_fp8_recall(gpu_lib):
    """Tight clusters of unit-norm rows (noise norm 0.1): neighbours' distances differ by less than
    FP8's rounding of whole rows, but not of residuals to the cluster centroid. Residual FP8 must
    reach high recall@10 at 4x oversampling and beat FP8 of the rows themselves."""
    n, d, nq, k = 40_000, 512, 128, 10
    rng = np.random.default_rng(17)
    c = rng.standard_normal((64, d)).astype(np.float32)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    x = c[rng.integers(0, 64, n)] + np.float32(0.1 / np.sqrt(d)) * rng.standard_normal((n, d), dtype=np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    q = c[rng.integers(0, 64, nq)] + np.float32(0.1 / np.sqrt(d)) * rng.standard_normal((nq, d), dtype=np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    idx = builder.build_index(np.ascontiguousarray(x), "pdx_f32", num_clusters=128, seed=3, kmeans_iters=4,
                              device=0)
    g = gpu_lib.GpuIndex(idx)
    rows, ids = _rows(idx)
    qrot, _ = g.preprocess(np.ascontiguousarray(q))
    q64, r64 = qrot.astype(np.float64), rows.astype(np.float64)
    exact = np.array([ids[np.argsort(((r64 - qq) ** 2).sum(1), kind="stable")[:k]] for qq in q64])
    rec = {}
    for name, prec in (("fp8", 1), ("fp8_direct", 0x101), ("bf16", 0)):
        run = _Run(gpu_lib, g, prec, nq)
        gi, gd, gc = run.search(q, k, 4)
        rec[name] = _recall(gi, exact, k)
        run.dn.close()
    print("recall@10 at 4x oversampling:", rec, flush=True)
    assert rec["fp8"] >= 0.9 and rec["bf16"] >= 0.97
    assert rec["fp8"] > rec["fp8_direct"]
