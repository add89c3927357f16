"""Captured searches (pdxg_graph_*): every replay of the graph equals the eager search of what the
captured buffers hold at replay time, and the oracle's results."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf
from tests.pdx_testdata import index_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,order", [("pdx_tree_u8", 1), ("pdx_tree_u8", 2), ("pdx_f32", 2)])
def test_graph_replays_equal_eager_and_oracle(kind, order):
    """order 2: the scans run through the persistent work queue, whose counters the graph clears
    with a fill kernel on every replay."""
    import torch

    from pdx_amd import _capi

    path, x, q0 = index_file(kind, 4000, 256, 40, normalize=kind == "pdx_tree_u8")
    rng = np.random.default_rng(12)
    nq, K = 48, 20
    qa = np.ascontiguousarray(x[rng.integers(0, len(x), nq)] + np.float32(0.05) * rng.standard_normal((nq, 256), dtype=np.float32))
    qb = np.ascontiguousarray(x[rng.integers(0, len(x), nq)] + np.float32(0.05) * rng.standard_normal((nq, 256), dtype=np.float32))
    kq = rng.choice([1, 10, 20], nq).astype(np.uint32)
    npb = rng.integers(1, 13, nq).astype(np.uint32)
    o = OracleIndex(path)
    g = _capi.GpuIndex(pf.load(path))
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(qa).to(dev)
    kd = torch.from_numpy(kq.view(np.int32)).to(dev)
    nd = torch.from_numpy(npb.view(np.int32)).to(dev)
    ids = torch.zeros((nq, K), dtype=torch.int32, device=dev)
    dd = torch.zeros((nq, K), dtype=torch.float32, device=dev)
    cnt = torch.zeros(nq, dtype=torch.int32, device=dev)
    ws = _capi.Workspace(g, nq, K, 12)
    ws.set_order(order)
    sp = torch.cuda.current_stream().cuda_stream
    args = (qd.data_ptr(), nq, kd.data_ptr(), nd.data_ptr(), K, 12, K, ids.data_ptr(), dd.data_ptr(), 0,
            cnt.data_ptr())

    def read():
        torch.cuda.synchronize()
        return ids.cpu().numpy().view(np.uint32).copy(), dd.cpu().numpy().copy(), cnt.cpu().numpy().view(np.uint32).copy()

    def expect(qs):
        oi, od, _, oc, _ = o.search(qs, kq, npb)
        return oi, od, oc

    def same(got, exp):
        gi, gd, gc = got
        oi, od, oc = exp
        np.testing.assert_array_equal(gc, oc)
        for r in range(nq):
            np.testing.assert_array_equal(gi[r, :oc[r]], oi[r, :oc[r]])
            np.testing.assert_array_equal(gd[r, :oc[r]].view(np.uint32), od[r, :oc[r]].view(np.uint32))

    ws.search(*args, stream=sp)
    eager_a = read()
    same(eager_a, expect(qa))
    graph = ws.capture(*args)
    print(f"{kind} order {order}: {graph.nodes} nodes", flush=True)
    assert graph.nodes >= 4
    exp_b = expect(qb)
    for rep in range(4):
        # eager work and synchronisations between the replays; outputs cleared each time
        ids.zero_()
        dd.zero_()
        cnt.zero_()
        src = qa if rep % 2 == 0 else qb
        qd.copy_(torch.from_numpy(src).to(dev))
        graph.launch(sp)
        got = read()
        if rep % 2 == 0:
            for a, b in zip(got, eager_a):
                np.testing.assert_array_equal(a, b)
        else:
            same(got, exp_b)
        ws2 = _capi.Workspace(g, 8, K, 12)  # unrelated eager search in between
        ws2.search(qd.data_ptr(), 8, kd.data_ptr(), nd.data_ptr(), K, 12, K, ids.data_ptr(), dd.data_ptr(), 0,
                   cnt.data_ptr(), stream=sp)
        torch.cuda.synchronize()
        ws2.close()
    graph.close()
    ws.close()
