"""Packed tails of the two-phase meso scan on a hand-built two-level index whose meso clusters have
chosen sizes: last 64-member chunks of 1-16 members (4 query groups share a wave), 17-32 (2 groups)
and 33-64 (ordinary items), each cluster visited by several 8-query groups, so packed items hold
full, partial and empty groups. Summaries + replay against the oracle, ids, distances and bytes."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf
from tests import boundary_cases as bc
from tests.test_gpu_parity import _compare, gpu_lib  # noqa: F401  (gpu_lib is a fixture)

pytestmark = pytest.mark.gpu

# members per meso cluster (L1 clusters grouped under it)
MESO_SIZES = [1, 15, 16, 17, 31, 32, 33, 63, 64, 65, 80, 81, 96, 97, 128, 129, 144, 160, 161]


def tails_index(u8: bool, d: int, seed: int) -> pf.IndexData:
    rng = np.random.default_rng(seed)
    nl1 = sum(MESO_SIZES)
    c1 = rng.standard_normal((nl1, d)).astype(np.float32)
    clusters, nid = [], 0
    for c in range(nl1):  # data level: 1-3 vectors around each L1 centroid
        m = int(rng.integers(1, 4))
        rows = c1[c] + np.float32(0.1) * rng.standard_normal((m, d)).astype(np.float32)
        if u8:
            rows = np.clip(np.round(rows * 20 + 128), 0, 255).astype(np.uint8)
        clusters.append((rows, list(range(nid, nid + m)), []))
        nid += m
    L1 = bc.level(d, u8, clusters, None, scale=0.05 if u8 else 1.0)
    perm = rng.permutation(nl1)
    meso, cent0, o = [], [], 0
    for s in MESO_SIZES:  # meso level: the L1 centroids in groups of the chosen sizes
        members = perm[o:o + s]
        o += s
        meso.append((c1[members], members.tolist(), []))
        cent0.append(c1[members].mean(0))
    L0 = bc.level(d, False, meso, np.asarray(cent0, np.float32))
    return pf.IndexData(type=pf.TYPE_TREE_U8 if u8 else pf.TYPE_TREE_F32, matrix=np.eye(d, dtype=np.float32),
                        l1=L1, l0=L0)


@pytest.mark.parametrize("u8,d", [(True, 192), (False, 128)], ids=["u8-d192", "f32-d128"])
def test_packed_tails_match_oracle(gpu_lib, u8, d):
    import torch

    idx = tails_index(u8, d, 5)
    nl1 = sum(MESO_SIZES)
    o = OracleIndex(data=idx)
    g = gpu_lib.GpuIndex(idx)
    rng = np.random.default_rng(7)
    nq, K = 44, 10
    q = np.ascontiguousarray(rng.standard_normal((nq, d)), np.float32)
    kq = rng.choice([1, 5, 10], nq).astype(np.uint32)
    npb = rng.integers(1, nl1 + 1, nq).astype(np.uint32)  # half of the meso clusters, or all
    oi, od, oid, oc, ob = o.search(q, kq, npb)
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(q).to(dev)
    kd = torch.from_numpy(kq.view(np.int32)).to(dev)
    nd = torch.from_numpy(npb.view(np.int32)).to(dev)
    ws = gpu_lib.Workspace(g, nq, K, nl1)
    ws.set_meso_summary(True)
    assert ws.meso_summary
    sp = torch.cuda.current_stream().cuda_stream
    for b in (64, 256):
        print(f"packed tails u8={u8} d={d} block {b}", flush=True)
        ws.set_plan(b, 0)
        ids = torch.zeros((nq, K), dtype=torch.int32, device=dev)
        dd = torch.zeros((nq, K), dtype=torch.float32, device=dev)
        idd = torch.zeros((nq, K), dtype=torch.int32, device=dev)
        cnt = torch.zeros(nq, dtype=torch.int32, device=dev)
        nb = torch.zeros(nq, dtype=torch.int64, device=dev)
        ws.search(qd.data_ptr(), nq, kd.data_ptr(), nd.data_ptr(), K, nl1, K, ids.data_ptr(), dd.data_ptr(),
                  idd.data_ptr(), cnt.data_ptr(), nb.data_ptr(), stream=sp)
        torch.cuda.synchronize()
        res = (ids.cpu().numpy().view(np.uint32), dd.cpu().numpy(), idd.cpu().numpy().view(np.uint32),
               cnt.cpu().numpy().view(np.uint32), nb.cpu().numpy().view(np.uint64))
        _compare((oi, od, oid, oc, ob), res, o.is_u8)
