"""The HIP path against the oracle on indexes whose search data has been changed in place: live
tombstones (searcher.hpp:169-179; cluster.hpp:148-154), filtered search above and below 20 %
per-cluster selectivity on flat and two-level indexes (searcher.hpp:284-372,376-409,726-856;
index.hpp:568-578), and the cluster an append lands in (index.hpp:596-607). The indexes are
handed to both sides in memory, so tombstones are not compacted away as a saved file would."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import builder
from tests.pdx_testdata import mixture

pytestmark = pytest.mark.gpu

_BUILT = {}


def built(kind, n=3000, d=192, nlist=24, seed=5, normalize=False):
    key = (kind, n, d, nlist, seed, normalize)
    if key not in _BUILT:
        x, q = mixture(n, d, 16, 0.3, seed, n_queries=16)
        idx = builder.build_index(x, kind, num_clusters=nlist, normalize=normalize, seed=seed, kmeans_iters=4,
                                  device=False)
        _BUILT[key] = (idx, x, q)
    return _BUILT[key]


def tombstone(level, rng, frac, heavy=()):
    """Tombstone `frac` of every cluster's used slots, and all but one slot of the clusters in
    `heavy` (those reach the warmup's 80 % exit on tombstones alone)."""
    for c, cl in enumerate(level.clusters):
        n = cl.n_used
        if c in heavy:
            dead = set(range(1, n))
        else:
            dead = set(rng.choice(n, int(frac * n), replace=False).tolist())
        cl.tombstones = dead
        cl.n_live = n - len(dead)


def compare(o_res, g_res, u8, what):
    oi, od, oid, oc, ob = o_res
    gi, gd, gid, gc, gb = g_res
    np.testing.assert_array_equal(gc, oc, err_msg=f"{what}: counts")
    for r in range(len(oc)):
        n = int(oc[r])
        np.testing.assert_array_equal(gi[r, :n], oi[r, :n], err_msg=f"{what}: ids q={r}")
        np.testing.assert_array_equal(gd[r, :n].view(np.uint32), od[r, :n].view(np.uint32),
                                      err_msg=f"{what}: dists q={r}")
        if u8:
            np.testing.assert_array_equal(gid[r, :n], oid[r, :n], err_msg=f"{what}: int dists q={r}")
    np.testing.assert_array_equal(gb, ob, err_msg=f"{what}: bytes")


@pytest.mark.parametrize("kind", ["pdx_tree_u8", "pdx_tree_f32", "pdx_u8", "pdx_f32"])
def test_live_tombstones_match_oracle(kind):
    """10 % of every cluster tombstoned, two clusters tombstoned down to one live slot (warmup
    exit on masked slots), k above the live count of a single probed cluster (max/2 rows)."""
    import copy

    from pdx_amd import _capi

    idx0, x, q = built(kind)
    idx = copy.deepcopy(idx0)
    rng = np.random.default_rng(1)
    tombstone(idx.l1, rng, 0.10, heavy=(0, 5))
    o = OracleIndex(data=idx)
    g = _capi.GpuIndex(idx)
    qr = o.preprocess(q)
    # queries aimed at the heavily tombstoned clusters too
    for k in (1, 10, 50, 300):
        for nprobe in (1, 4, 0):
            print(f"{kind} k={k} nprobe={nprobe}", flush=True)
            compare(o.search(qr, k, nprobe, transformed=True),
                    g.search(qr, k, nprobe, transformed=True, want_bytes=True), o.is_u8, f"k={k} nprobe={nprobe}")
    # a query that visits a cluster with one live slot first: Start pushes its masked slots when
    # k exceeds the live count
    L = idx.l1
    if L.centroids is not None:
        qc = np.ascontiguousarray(L.centroids[[0, 5]], np.float32)
        compare(o.search(qc, 20, 1, transformed=True), g.search(qc, 20, 1, transformed=True, want_bytes=True),
                o.is_u8, "centroid queries")


@pytest.mark.parametrize("kind", ["pdx_tree_u8", "pdx_tree_f32"])
def test_meso_level_tombstones_match_oracle(kind):
    """Tombstoned entries at the meso level (what merging an L1 cluster leaves there) make the GPU
    fall back from its two-phase meso scan to the pruned one; results still equal the oracle's."""
    import copy

    import torch

    from pdx_amd import _capi

    idx0, x, q = built(kind)
    idx = copy.deepcopy(idx0)
    M = idx.l0
    for c, cl in enumerate(M.clusters):
        if cl.n_used > 1:
            cl.tombstones = {c % cl.n_used}
            cl.n_live = cl.n_used - 1
    o = OracleIndex(data=idx)
    g = _capi.GpuIndex(idx)
    qr = o.preprocess(q)
    for k in (1, 20):
        for nprobe in (2, 12, 0):
            compare(o.search(qr, k, nprobe, transformed=True),
                    g.search(qr, k, nprobe, transformed=True, want_bytes=True), o.is_u8, f"k={k} nprobe={nprobe}")
    # the device-buffer path with the two-phase scan requested reports the fallback
    nq, K, NP = len(qr), 20, 12
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(np.ascontiguousarray(qr)).to(dev)
    ids = torch.zeros((nq, K), dtype=torch.int32, device=dev)
    dd = torch.zeros((nq, K), dtype=torch.float32, device=dev)
    cnt = torch.zeros(nq, dtype=torch.int32, device=dev)
    ws = _capi.Workspace(g, nq, K, NP)
    ws.set_meso_summary(True)
    assert ws.meso_summary is False
    ws.search(qd.data_ptr(), nq, 0, 0, K, NP, K, ids.data_ptr(), dd.data_ptr(), 0, cnt.data_ptr(), transformed=True,
              k_all=K, nprobe_all=NP, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    oi, od, _, oc, _ = o.search(qr, K, NP, transformed=True)
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), oc)
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oi)
    ws.close()


def _selection(o, passing_rows):
    off = o.cluster_offsets()
    sel = np.zeros(int(off[-1]), np.uint8)
    for c in range(o.num_clusters):
        cid = o.cluster_ids(c)
        sel[int(off[c]):int(off[c]) + len(cid)] = passing_rows[cid]
    return sel


@pytest.mark.parametrize("kind", ["pdx_u8", "pdx_f32", "pdx_tree_u8", "pdx_tree_f32"])
@pytest.mark.parametrize("frac", [0.1, 0.3, 0.6])
def test_filtered_search_matches_oracle(kind, frac):
    """Row filters of 10 %, 30 % and 60 % (FilteredStart's gather and full-scan branches, Warmup<true>
    with and without its early return), plus clusters passing entirely or not at all. Two-level
    indexes run the data level's own IVF probe over the L1 centroids (index.hpp:568-578)."""
    from pdx_amd import _capi

    idx, x, q = built(kind)
    o = OracleIndex(data=idx)
    g = _capi.GpuIndex(idx)
    rng = np.random.default_rng(int(frac * 100))
    passing = (rng.random(len(x)) < frac).astype(np.uint8)
    # one cluster passes entirely, one not at all
    L = idx.l1
    passing[L.clusters[1].ids[:L.clusters[1].n_used]] = 1
    passing[L.clusters[2].ids[:L.clusters[2].n_used]] = 0
    sel = _selection(o, passing)
    qr = o.preprocess(q)
    for t in range(len(qr)):
        for k, nprobe in ((1, 6), (10, 6), (100, 0), (10, 1)):
            gi, gd, gid = g.filtered_search(qr[t], k, nprobe, sel, transformed=True)
            oi, od, oid = o.filtered_search(qr[t], k, nprobe, sel, transformed=True)
            np.testing.assert_array_equal(gi, oi, err_msg=f"q={t} k={k} nprobe={nprobe}")
            np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
            if o.is_u8:
                np.testing.assert_array_equal(gid, oid)
            assert all(passing[i] for i in gi)


@pytest.mark.parametrize("kind", ["pdx_tree_u8", "pdx_tree_f32"])
def test_append_cluster_choice_matches_oracle(kind):
    """The meso-level search an append runs (max(1, n_meso / 8) probes, k = 1, transformed vector;
    index.hpp:596-607), and the same search at other k / nprobe."""
    from pdx_amd import _capi

    idx, x, q = built(kind, n=4000, nlist=64)
    o = OracleIndex(data=idx)
    g = _capi.GpuIndex(idx)
    rng = np.random.default_rng(9)
    rows = np.ascontiguousarray(x[rng.integers(0, len(x), 64)] + np.float32(0.05) * rng.standard_normal((64, x.shape[1]),
                                                                                                        dtype=np.float32))
    rr = o.preprocess(rows)
    n0 = idx.l0.nc
    for k, nprobe in ((1, max(1, n0 // 8)), (1, 0), (5, 2), (16, n0 // 2)):
        gi, gd, gc = g.search_meso(rr, k, nprobe)
        oi, od, oc = o.search_meso(rr, k, nprobe)
        np.testing.assert_array_equal(gc, oc)
        np.testing.assert_array_equal(gi, oi, err_msg=f"k={k} nprobe={nprobe}")
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("kind", ["IndexPDXIVFTree", "IndexPDXIVFTreeSQ8"])
def test_api_append_lands_in_the_oracle_cluster(kind, tmp_path):
    """Through the kept Python API: each appended row lands in the L1 cluster the oracle's meso
    search picks for it on the same (saved) index."""
    import pdx_amd

    x, q = mixture(3000, 128, 16, 0.3, 21, n_queries=24)
    idx = getattr(pdx_amd, kind)(num_dimensions=128, num_clusters=40)
    idx.build(x)
    path = str(tmp_path / "t.pdx")
    idx.save(path)
    o = OracleIndex(path)
    rows = np.ascontiguousarray(q + np.float32(0.01))
    want, _, _ = o.search_meso(o.preprocess(rows), 1, max(1, o.num_meso_clusters // 8))
    raw = idx._index
    for i, r in enumerate(rows):
        idx.append(3000 + i, r)
    for i in range(len(rows)):
        where = [c for c in range(raw.get_num_clusters()) if 3000 + i in set(raw.get_cluster_row_ids(c).tolist())]
        assert where == [int(want[i, 0])], f"row {3000 + i}: cluster {where} vs oracle {int(want[i, 0])}"
