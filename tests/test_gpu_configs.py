"""The benchmark configurations' own regimes at reduced N, each against the oracle:

* C2: flat IVF F32 at d = 768, k = 10, nprobe = 32, L2 and inner product (normalised), searched
  through the device-buffer path in batches of 1, 64 and 1024 queries;
* C4: two-level SQ8 at d = 1536 with per-query k in {10, 20, 100} and nprobe in {16, 64, 256},
  served by the continuous-batching server under a scratch budget;
* C5: the dense BF16 / FP8 first pass at d = 3072 with a batch of 2048 queries, a 10 % row filter
  and 1 % of the rows appended and deleted: every returned distance equals the oracle's
  Start-order distance of the returned row, filtered and deleted rows never come back, and recall
  against the exact top-k of the live passing rows stays high.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleIndex
from pdx_amd import builder, pdxfile as pf
from tests.pdx_testdata import mixture

pytestmark = pytest.mark.gpu


def _unit_mixture(n, d, centers, noise, seed, nq):
    """Unit-norm rows / queries: unit centres plus noise of norm ~`noise`, renormalised."""
    x, q = mixture(n, d, centers, noise / np.sqrt(d), seed, unit=True, n_queries=nq)
    return np.ascontiguousarray(x), np.ascontiguousarray(q)


def _same(o_res, gi, gd, gidist, gc, u8, what):
    oi, od, oid, oc, _ = o_res
    np.testing.assert_array_equal(gc, oc, err_msg=f"{what}: counts")
    for r in range(len(oc)):
        n = int(oc[r])
        np.testing.assert_array_equal(gi[r, :n], oi[r, :n], err_msg=f"{what}: ids q={r}")
        np.testing.assert_array_equal(gd[r, :n].view(np.uint32), od[r, :n].view(np.uint32), err_msg=f"{what} q={r}")
        if u8:
            np.testing.assert_array_equal(gidist[r, :n], oid[r, :n], err_msg=f"{what}: int dists q={r}")


@pytest.mark.parametrize("metric", ["l2sq", "ip"])
def test_c2_regime_flat_f32_d768_batches(metric):
    import torch

    from pdx_amd import _capi

    x, q = _unit_mixture(40_000, 768, 64, 0.15, 61, 1024)
    idx = builder.build_index(x, "pdx_f32", num_clusters=100, distance_metric=metric, seed=42, kmeans_iters=5)
    g = _capi.GpuIndex(idx)
    o = OracleIndex(data=idx)
    K, NP = 10, 32
    dev = torch.device("cuda", 0)
    ws = _capi.Workspace(g, 1024, K, NP)
    qd = torch.from_numpy(q).to(dev)
    ids = torch.zeros((1024, K), dtype=torch.int32, device=dev)
    dd = torch.zeros((1024, K), dtype=torch.float32, device=dev)
    cnt = torch.zeros(1024, dtype=torch.int32, device=dev)
    nb = torch.zeros(1024, dtype=torch.int64, device=dev)
    ref = o.search(q, K, NP, threads=16)
    for b in (1, 64, 1024):
        print(f"C2 {metric} batch {b}", flush=True)
        m = 16 if b == 1 else 1024          # 16 single-query calls, 16 batches of 64, one of 1024
        for s in range(0, m, b):
            ws.search(qd[s:s + b].data_ptr(), b, 0, 0, K, NP, K, ids[s:s + b].data_ptr(), dd[s:s + b].data_ptr(), 0,
                      cnt[s:s + b].data_ptr(), nb[s:s + b].data_ptr(), k_all=K, nprobe_all=NP,
                      stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        gi = ids[:m].cpu().numpy().view(np.uint32)
        gd = dd[:m].cpu().numpy()
        gc = cnt[:m].cpu().numpy().view(np.uint32)
        _same(tuple(a[:m] for a in ref), gi, gd, None, gc, False, f"batch {b}")
        np.testing.assert_array_equal(nb[:m].cpu().numpy().astype(np.uint64), ref[4][:m])


def test_c4_regime_served_tree_sq8_d1536():
    from pdx_amd import _capi

    x, q = mixture(30_000, 1536, 64, 0.2, 62, n_queries=192)
    idx = builder.build_index(x, "pdx_tree_u8", num_clusters=300, normalize=True, seed=42, kmeans_iters=4)
    g = _capi.GpuIndex(idx)
    rng = np.random.default_rng(62)
    k = rng.choice(np.array([10, 20, 100], np.uint32), len(q))
    nprobe = rng.choice(np.array([16, 64, 256], np.uint32), len(q))
    probe = _capi.Server(g, max_batch=64, max_k=100, max_nprobe=256, max_slots=1)
    st = probe.stats()
    probe.close()
    per_query = st["summary_capacity_bytes"] // 64   # the largest footprint of one request
    fixed = st["scratch_bytes"] - st["summary_capacity_bytes"]
    # three slots whose meso-summary buffers hold ten requests of the largest footprint each:
    # admission by footprint closes the batches of this burst before max_batch
    sv = _capi.Server(g, max_batch=64, max_k=100, max_nprobe=256, max_slots=3,
                      scratch_budget_bytes=3 * (fixed + 10 * per_query))
    assert sv.stats()["slots"] == 3 and sv.stats()["summary_capacity_bytes"] == 10 * per_query
    for i in range(len(q)):
        sv.submit(q[i], int(k[i]), int(nprobe[i]), i)
    got = {}
    while len(got) < len(q):
        for tag, ids, d, lat in sv.poll(256, 2000):
            got[tag] = (ids, d)
    assert sv.stats()["footprint_limited_batches"] > 0
    sv.close()
    o = OracleIndex(data=idx)
    oi, od, _, oc, _ = o.search(q, k, nprobe, threads=16)
    for i in range(len(q)):
        np.testing.assert_array_equal(got[i][0], oi[i, :oc[i]], err_msg=f"q={i} k={k[i]} nprobe={nprobe[i]}")
        np.testing.assert_array_equal(got[i][1].view(np.uint32), od[i, :oc[i]].view(np.uint32))


@pytest.mark.parametrize("name,prec", [("bf16", 0), ("fp8", 1)])
def test_c5_regime_dense_d3072_filter_updates(name, prec):
    import torch

    from pdx_amd import _capi

    n, d, nq, K, OS = 20_000, 3072, 2048, 10, 4
    n_up = n // 100
    x_all, q = _unit_mixture(n + n_up, d, 64, 0.1, 63, nq)
    x, x_new = x_all[:n], x_all[n:]
    idx = builder.build_index(x, "pdx_f32", num_clusters=64, seed=42, kmeans_iters=4)
    g = _capi.GpuIndex(idx)
    L = idx.l1
    rows = np.concatenate([pf.from_pdx(c.data, c.n_used, c.cap, d, L.v, L.h, False) for c in L.clusters])
    ids = np.concatenate([c.ids[:c.n_used] for c in L.clusters]).astype(np.uint32)
    dn = _capi.Dense(g, prec, extra_rows=n_up, max_queries=nq)
    rng = np.random.default_rng(64)
    sel = (rng.random(n + n_up) < 0.10).astype(np.uint8)
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(q).to(dev)
    seld = torch.from_numpy(sel).to(dev)
    qrot, _ = g.preprocess(q)
    sub = np.arange(0, nq, 8)

    def search():
        oi = torch.full((nq, K), -1, dtype=torch.int32, device=dev)
        od = torch.zeros((nq, K), device=dev)
        oc = torch.zeros(nq, dtype=torch.int32, device=dev)
        dn.search(qd.data_ptr(), nq, K, OS, K, oi.data_ptr(), od.data_ptr(), oc.data_ptr(),
                  selection_ptr=seld.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return oi.cpu().numpy().view(np.uint32), od.cpu().numpy(), oc.cpu().numpy()

    def check(gi, gd, gc, live_rows, live_ids, tag):
        m = sel[live_ids] == 1
        pr, pi = live_rows[m], live_ids[m]
        row_of = {int(i): r for r, i in enumerate(pi)}
        hits = 0
        for t in sub:
            c = int(gc[t])
            assert c == K
            r = np.array([row_of[int(i)] for i in gi[t, :c]])         # only live passing rows come back
            ref = orc.start_distances(qrot[t], pr[r], L.v)
            np.testing.assert_array_equal(gd[t, :c].view(np.uint32), ref.view(np.uint32), err_msg=f"{tag} q={t}")
            exact = orc.start_distances(qrot[t], pr, L.v)
            best = pi[np.lexsort((pi, exact))[:K]]
            hits += len(set(gi[t, :K].tolist()) & set(best.tolist()))
        rec = hits / (K * len(sub))
        print(f"C5 {name} {tag}: recall@{K} vs exact FP32 top-k of the live passing rows {rec:.4f}", flush=True)
        assert rec >= 0.95

    gi, gd, gc = search()
    check(gi, gd, gc, rows, ids, "before updates")
    # 1 % appended and 1 % deleted (among them rows the queries returned)
    new_ids = np.arange(n, n + n_up, dtype=np.uint32)
    dn.append(x_new, new_ids)
    gone = np.unique(np.concatenate([gi[:n_up // 2, 0], rng.choice(ids, n_up, replace=False)]))[:n_up]
    dn.delete(gone)
    new_rot, _ = g.preprocess(x_new)
    keep = ~np.isin(ids, gone)
    rows2 = np.concatenate([rows[keep], new_rot])
    ids2 = np.concatenate([ids[keep], new_ids])
    gi2, gd2, gc2 = search()
    assert not np.isin(gi2, gone).any()
    check(gi2, gd2, gc2, rows2, ids2, "after 1% append/delete")
