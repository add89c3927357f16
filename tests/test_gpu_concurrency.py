"""Updates that overlap searches in flight: a cluster re-upload (pdxg_index_update_cluster) while a
continuous-batching server has batches on its own streams, and a dense-copy append / delete while a
dense search runs on a non-blocking stream. The update calls wait for the device, so every search
sees either the old or the new data in full, and searches after the update equal the oracle on
the updated index."""
import copy

import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import builder, pdxfile as pf
from tests.pdx_testdata import mixture

pytestmark = pytest.mark.gpu


def _collect(sv, n):
    out = {}
    while len(out) < n:
        for tag, ids, d, lat in sv.poll(256, 2000):
            out[tag] = (ids, d)
    return out


def test_cluster_updates_while_a_server_runs():
    from pdx_amd import _capi

    x, q = mixture(3000, 256, 16, 0.3, 31, n_queries=32)
    idx = builder.build_index(x, "pdx_tree_u8", num_clusters=24, seed=3, kmeans_iters=4, device=False)
    idx = copy.deepcopy(idx)
    g = _capi.GpuIndex(idx)
    rng = np.random.default_rng(2)
    qs = np.ascontiguousarray(np.repeat(q, 8, axis=0) + np.float32(0.01) * rng.standard_normal((256, 256), dtype=np.float32))
    sv = _capi.Server(g, max_batch=16, max_k=20, max_nprobe=0, max_slots=3)
    for i in range(len(qs)):
        sv.submit(qs[i], 20, 0 if i % 3 == 0 else 8, i)
    # while those batches run: tombstone a slot and append a vector to several clusters (a grown
    # used_capacity is what a scan planned before the update must never see)
    L = idx.l1
    for c in range(0, L.nc, 3):
        cl = L.clusters[c]
        cl.tombstones.add(0)
        cl.n_live -= 1
        if cl.n_used < cl.cap:
            src = pf.from_pdx(cl.data, cl.n_used, cl.cap, L.d, L.v, L.h, True)[cl.n_used - 1]
            pf.write_slot(cl.data, cl.n_used, src, cl.cap, L.v, L.h, True)
            cl.ids[cl.n_used] = 100000 + c
            cl.n_used += 1
            cl.n_live += 1
        g.update_cluster(1, c)
    first = _collect(sv, len(qs))
    assert all(len(v[0]) == 20 for v in first.values())
    # searches after the updates see them
    for i in range(len(qs)):
        sv.submit(qs[i], 20, 8, 1000 + i)
    after = _collect(sv, len(qs))
    sv.close()
    o = OracleIndex(data=idx)
    oi, od, _, oc, _ = o.search(qs, 20, 8)
    for i in range(len(qs)):
        ids, d = after[1000 + i]
        np.testing.assert_array_equal(ids, oi[i, :oc[i]])
        np.testing.assert_array_equal(d.view(np.uint32), od[i, :oc[i]].view(np.uint32))


def test_dense_append_and_delete_wait_for_a_search_in_flight():
    import torch

    from pdx_amd import _capi

    x, q = mixture(4000, 256, 16, 0.3, 32, n_queries=64)
    idx = builder.build_index(x, "pdx_f32", num_clusters=16, seed=3, kmeans_iters=4, device=False)
    g = _capi.GpuIndex(idx)
    dn = _capi.Dense(g, _capi.PDXG_DENSE_BF16, extra_rows=256, max_queries=64)
    dev = torch.device("cuda", 0)
    nq, K = len(q), 10
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(dev)

    def outputs():
        return (torch.full((nq, K), -1, dtype=torch.int32, device=dev), torch.zeros((nq, K), device=dev),
                torch.zeros(nq, dtype=torch.int32, device=dev))

    def host(o):
        torch.cuda.synchronize()
        return [t.cpu().numpy().copy() for t in o]

    ref = outputs()
    dn.search(qd.data_ptr(), nq, K, 4, K, ref[0].data_ptr(), ref[1].data_ptr(), ref[2].data_ptr(),
              stream=torch.cuda.current_stream().cuda_stream)
    ref = host(ref)
    # the same search on a non-blocking side stream, immediately followed by an append of rows that
    # would enter its results (copies of the queries) and by a delete of its best rows
    side = torch.cuda.Stream()
    out = outputs()
    torch.cuda.synchronize()
    dn.search(qd.data_ptr(), nq, K, 4, K, out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
              stream=side.cuda_stream)
    new_ids = np.arange(4000, 4000 + nq, dtype=np.uint32)
    dn.append(q, new_ids)
    dn.delete(ref[0][:, 0].view(np.uint32))
    got = host(out)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    # a search issued after the updates finds the appended rows and not the deleted ones
    after = outputs()
    dn.search(qd.data_ptr(), nq, K, 4, K, after[0].data_ptr(), after[1].data_ptr(), after[2].data_ptr(),
              stream=torch.cuda.current_stream().cuda_stream)
    ai = host(after)[0].view(np.uint32)
    assert (ai[:, 0] == new_ids).all()
    assert not np.isin(ref[0][:, 0].view(np.uint32), ai).any()
