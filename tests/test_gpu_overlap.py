"""Several batches in flight at once, each on its own stream with its own workspace and outputs
over one shared index (how the batched benchmark and the server's slots run): every batch equals
the oracle's results for its queries, whatever runs beside it."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf
from tests.pdx_testdata import index_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,summ", [("pdx_tree_u8", True), ("pdx_tree_u8", False), ("pdx_f32", False)])
def test_batches_in_flight_equal_oracle(kind, summ):
    import torch

    from pdx_amd import _capi

    path, x, _ = index_file(kind, 4000, 256, 40, normalize=kind == "pdx_tree_u8")
    rng = np.random.default_rng(21)
    nslots, nbatch, nq, K, NPM = 3, 6, 64, 20, 12
    qs = [np.ascontiguousarray(x[rng.integers(0, len(x), nq)]
                               + np.float32(0.05) * rng.standard_normal((nq, 256), dtype=np.float32))
          for _ in range(nbatch)]
    ks = [rng.choice([1, 10, 20], nq).astype(np.uint32) for _ in range(nbatch)]
    nps = [rng.integers(1, NPM + 1, nq).astype(np.uint32) for _ in range(nbatch)]
    o = OracleIndex(path)
    exp = [o.search(qs[b], ks[b], nps[b]) for b in range(nbatch)]
    g = _capi.GpuIndex(pf.load(path))
    dev = torch.device("cuda", 0)
    qd = [torch.from_numpy(a).to(dev) for a in qs]
    kd = [torch.from_numpy(a.view(np.int32)).to(dev) for a in ks]
    nd = [torch.from_numpy(a.view(np.int32)).to(dev) for a in nps]
    streams = [torch.cuda.Stream(device=dev) for _ in range(nslots)]
    wss = []
    for _ in range(nslots):
        w = _capi.Workspace(g, nq, K, NPM)
        if kind == "pdx_tree_u8":
            w.set_meso_summary(summ)
        wss.append(w)
    # every batch gets its own outputs so that all of them can be read after one synchronisation
    outs = [(torch.zeros((nq, K), dtype=torch.int32, device=dev), torch.zeros((nq, K), dtype=torch.float32, device=dev),
             torch.zeros((nq, K), dtype=torch.int32, device=dev), torch.zeros(nq, dtype=torch.int32, device=dev))
            for _ in range(nbatch)]
    for rep in range(2):
        for b in range(nbatch):
            j = (b + rep) % nslots
            ids, dd, ii, cnt = outs[b]
            wss[j].search(qd[b].data_ptr(), nq, kd[b].data_ptr(), nd[b].data_ptr(), K, NPM, K, ids.data_ptr(),
                          dd.data_ptr(), ii.data_ptr(), cnt.data_ptr(), 0, stream=streams[j].cuda_stream)
        torch.cuda.synchronize()
        for b in range(nbatch):
            oi, od, oid, oc, _ = exp[b]
            gi = outs[b][0].cpu().numpy().view(np.uint32)
            gd = outs[b][1].cpu().numpy()
            gid = outs[b][2].cpu().numpy().view(np.uint32)
            gc = outs[b][3].cpu().numpy().view(np.uint32)
            np.testing.assert_array_equal(gc, oc)
            for r in range(nq):
                n = oc[r]
                np.testing.assert_array_equal(gi[r, :n], oi[r, :n], err_msg=f"batch {b} query {r}")
                np.testing.assert_array_equal(gd[r, :n].view(np.uint32), od[r, :n].view(np.uint32))
                if kind == "pdx_tree_u8":
                    np.testing.assert_array_equal(gid[r, :n], oid[r, :n])
            for t in outs[b]:
                t.zero_()
