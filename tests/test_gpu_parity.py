"""HIP path vs the CPU oracle on the same index files and queries (bit-exact).

F32 distances are compared bit-for-bit: both sides use the reference's AVX-512 reduction order
for horizontal blocks / centroid distances and one fma chain in dimension order for vertical
blocks. SQ8 integer distances and ids are exact by construction.
"""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf
from tests.pdx_testdata import index_file

pytestmark = pytest.mark.gpu

CASES = [
    ("pdx_f32", 3000, 128, 24, False),
    ("pdx_u8", 3000, 128, 24, False),
    ("pdx_tree_f32", 3000, 128, 24, False),
    ("pdx_tree_u8", 3000, 128, 24, False),
    ("pdx_tree_u8", 3000, 128, 24, True),     # normalised (cosine / ip)
    ("pdx_f32", 2000, 100, 16, False),        # d <= 128, h = 0: no horizontal block
    ("pdx_u8", 2000, 130, 16, False),         # v = 2: U8 vertical tail, unaligned horizontal query
    ("pdx_u8", 1500, 33, 12, False),          # d <= 64: all vertical, tail of 1 dim
    ("pdx_f32", 1500, 64, 12, False),
    ("pdx_f32", 2500, 384, 20, True),
    ("pdx_u8", 2500, 384, 20, False),
    ("pdx_tree_u8", 4000, 768, 40, False),
    ("pdx_tree_f32", 4000, 768, 40, True),
]


def _ids(c):
    return "%s-n%d-d%d-c%d%s" % (c[0], c[1], c[2], c[3], "-norm" if c[4] else "")


@pytest.fixture(scope="module")
def gpu_lib():
    from pdx_amd import _capi

    print("device:", _capi.device_info(0), flush=True)
    return _capi


def _compare(o_res, g_res, u8):
    oi, od, oid, oc, ob = o_res
    gi, gd, gid, gc, gb = g_res
    np.testing.assert_array_equal(gc, oc)
    for q in range(len(oc)):
        n = int(oc[q])
        np.testing.assert_array_equal(gi[q, :n], oi[q, :n], err_msg=f"ids q={q}")
        np.testing.assert_array_equal(gd[q, :n].view(np.uint32), od[q, :n].view(np.uint32), err_msg=f"dists q={q}")
        if u8:
            np.testing.assert_array_equal(gid[q, :n], oid[q, :n], err_msg=f"int dists q={q}")
    if gb is not None:
        np.testing.assert_array_equal(gb, ob, err_msg="algorithmic bytes")


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_search_matches_oracle(gpu_lib, case):
    kind, n, d, nlist, norm = case
    path, x, q = index_file(kind, n, d, nlist, normalize=norm)
    o = OracleIndex(path)
    g = gpu_lib.GpuIndex(pf.load(path))
    for k in (1, 10, 50):
        for nprobe in (1, 4, 0):
            print(f"{_ids(case)} k={k} nprobe={nprobe}", flush=True)
            o_res = o.search(q, k, nprobe)
            g_res = g.search(q, k, nprobe, want_bytes=True)
            _compare(o_res, g_res, o.is_u8)


def sq8_quantize(rot, base, scale):
    """The SQ8 form of a rotated query (searcher.hpp:638-643 applies QuantizeEmbedding,
    scalar.hpp:76-94): std::round((x - base) * scale) with the float operations of the reference,
    rounding half away from zero, clamped to [0, 255]."""
    v = ((rot - np.float32(base)) * np.float32(scale)).astype(np.float64)
    r = np.sign(v) * np.floor(np.abs(v) + 0.5)
    return np.clip(r, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("case", CASES[:5] + [CASES[6], CASES[7]], ids=_ids)
def test_preprocess_matches_oracle(gpu_lib, case):
    kind, n, d, nlist, norm = case
    path, x, q = index_file(kind, n, d, nlist, normalize=norm)
    o = OracleIndex(path)
    idx = pf.load(path)
    g = gpu_lib.GpuIndex(idx)
    rot, qu = g.preprocess(q)
    orot = o.preprocess(q)
    np.testing.assert_array_equal(rot.view(np.uint32), orot.view(np.uint32))
    if qu is not None:  # the quantised query the U8 scan uses, from the oracle's rotated query
        np.testing.assert_array_equal(qu, sq8_quantize(orot, idx.l1.base, idx.l1.scale))
    # a batch large enough for the tiled rotation (64 queries x 64 outputs per tile, used from 1024
    # queries when there are two tiles per CU of the MI355X's 256), with a ragged last tile; its
    # first rows also go through the 16-query kernel (1089 queries make too few tiles at this d)
    # and the wide one below
    rng = np.random.default_rng(3)
    nb = 64 * -(-512 // ((d + 63) // 64)) + 65
    qb = np.ascontiguousarray(x[rng.integers(0, n, nb)] + np.float32(0.01) * rng.standard_normal((nb, d), dtype=np.float32))
    rot_b, qu_b = g.preprocess(qb)
    orot_b = o.preprocess(qb)
    np.testing.assert_array_equal(rot_b.view(np.uint32), orot_b.view(np.uint32))
    if qu_b is not None:  # the tiled rotation's fused quantisation
        np.testing.assert_array_equal(qu_b, sq8_quantize(orot_b, idx.l1.base, idx.l1.scale))
    np.testing.assert_array_equal(rot_b[:1089].view(np.uint32), g.preprocess(qb[:1089])[0].view(np.uint32))
    np.testing.assert_array_equal(rot_b[:len(q)].view(np.uint32), g.preprocess(qb[:len(q)])[0].view(np.uint32))
    # one and a few queries (the wide small-batch rotation)
    for m in (1, 5, 8):
        np.testing.assert_array_equal(g.preprocess(qb[:m])[0].view(np.uint32), rot_b[:m].view(np.uint32))


def test_mixed_k_and_nprobe_per_query(gpu_lib):
    path, x, q = index_file("pdx_tree_u8", 4000, 768, 40)
    o = OracleIndex(path)
    g = gpu_lib.GpuIndex(pf.load(path))
    rng = np.random.default_rng(3)
    k = rng.choice([1, 10, 20, 100], len(q)).astype(np.uint32)
    nprobe = rng.integers(1, 41, len(q)).astype(np.uint32)
    _compare(o.search(q, k, nprobe), g.search(q, k, nprobe, want_bytes=True), True)


def test_pre_rotated_queries(gpu_lib):
    path, x, q = index_file("pdx_u8", 2500, 384, 20)
    o = OracleIndex(path)
    g = gpu_lib.GpuIndex(pf.load(path))
    qr = o.preprocess(q)
    _compare(o.search(qr, 10, 5, transformed=True), g.search(qr, 10, 5, transformed=True, want_bytes=True), True)


@pytest.mark.parametrize("kind", ["pdx_tree_u8", "pdx_f32"])
def test_every_scan_block_size_gives_the_oracle_results(gpu_lib, kind):
    import torch

    path, x, q = index_file(kind, 4000, 768, 40) if kind == "pdx_tree_u8" else index_file(kind, 2500, 384, 20, normalize=True)
    o = OracleIndex(path)
    g = gpu_lib.GpuIndex(pf.load(path))
    nq, K, npb = len(q), 20, 6
    oi, od, oid, oc, _ = o.search(q, K, npb)
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(q).to(dev)
    ids = torch.zeros((nq, K), dtype=torch.int32, device=dev)
    dd = torch.zeros((nq, K), dtype=torch.float32, device=dev)
    cnt = torch.zeros(nq, dtype=torch.int32, device=dev)
    ws = gpu_lib.Workspace(g, nq, K, npb)
    sp = torch.cuda.current_stream().cuda_stream
    for b in (64, 128, 256, 512, 1024):
        for mode in (0, 1, 2):
            print(f"{kind} block {b} order mode {mode}", flush=True)
            ws.set_plan(b, b)
            ws.set_order(mode)
            assert ws.order == mode
            ws.search(qd.data_ptr(), nq, 0, 0, K, npb, K, ids.data_ptr(), dd.data_ptr(), 0, cnt.data_ptr(), 0,
                      stream=sp, k_all=K, nprobe_all=npb)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), oc)
            np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oi)
            np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint32), od.view(np.uint32))
    plan = ws.autotune(qd.data_ptr(), nq, 0, 0, K, npb, K, ids.data_ptr(), dd.data_ptr(), 0, cnt.data_ptr(),
                       stream=sp, k_all=K, nprobe_all=npb)
    assert plan[1] in (64, 128, 256, 512, 1024)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32), oi)


@pytest.mark.parametrize("kind,n,d,nlist,n_meso,nq", [("pdx_tree_u8", 4000, 768, 40, 0, 24),
                                                      ("pdx_tree_f32", 3000, 128, 24, 0, 24),
                                                      ("pdx_tree_u8", 2000, 130, 16, 0, 24),
                                                      ("pdx_tree_f32", 20000, 256, 400, 4, 150),
                                                      ("pdx_tree_u8", 12000, 1024, 300, 3, 70),
                                                      ("pdx_tree_u8", 24000, 128, 1200, 300, 64)])
def test_two_phase_meso_scan_matches_pruned_and_oracle(gpu_lib, kind, n, d, nlist, n_meso, nq):
    """The meso level through summaries + replay and through the pruned scan: both equal the oracle.

    The fourth and fifth cases have meso clusters of 75-100 members (several 64-lane chunks)
    visited by many queries (several 8-query work items per cluster). The last one visits 300 meso
    clusters per query when nprobe >= 600 (more than one 256-visit pass of the summary offsets)
    and inverts about 15,000 (query, visit) entries (several workgroups of the visit-list
    kernels)."""
    import torch

    path, x, q0 = index_file(kind, n, d, nlist, n_meso=n_meso)
    rng = np.random.default_rng(11)
    q = np.concatenate([q0, x[rng.integers(0, n, max(0, nq - len(q0)))] +
                        np.float32(0.05) * rng.standard_normal((max(0, nq - len(q0)), d), dtype=np.float32)])[:nq]
    q = np.ascontiguousarray(q, np.float32)
    o = OracleIndex(path)
    g = gpu_lib.GpuIndex(pf.load(path))
    K = 20
    npb = rng.integers(1, nlist + 1, nq).astype(np.uint32)   # halved and full meso visit counts
    kq = rng.choice([1, 5, 20], nq).astype(np.uint32)
    oi, od, oid, oc, ob = o.search(q, kq, npb)
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(q).to(dev)
    kd = torch.from_numpy(kq.view(np.int32)).to(dev)
    nd = torch.from_numpy(npb.view(np.int32)).to(dev)
    ws = gpu_lib.Workspace(g, nq, K, nlist)
    sp = torch.cuda.current_stream().cuda_stream
    for summ in (True, False):
        ws.set_meso_summary(summ)
        assert ws.meso_summary == summ
        for b in (64, 256, 1024):
            mode = {64: 2, 256: 0, 1024: 1}[b]  # per-query visit counts differ: every scan order
            print(f"{kind} d={d} summary={summ} block {b} order mode {mode}", flush=True)
            ws.set_plan(b, 0)
            ws.set_order(mode)
            ids = torch.zeros((nq, K), dtype=torch.int32, device=dev)
            dd = torch.zeros((nq, K), dtype=torch.float32, device=dev)
            idd = torch.zeros((nq, K), dtype=torch.int32, device=dev)
            cnt = torch.zeros(nq, dtype=torch.int32, device=dev)
            nb = torch.zeros(nq, dtype=torch.int64, device=dev)
            ws.search(qd.data_ptr(), nq, kd.data_ptr(), nd.data_ptr(), K, nlist, K, ids.data_ptr(), dd.data_ptr(),
                      idd.data_ptr(), cnt.data_ptr(), nb.data_ptr(), stream=sp)
            torch.cuda.synchronize()
            res = (ids.cpu().numpy().view(np.uint32), dd.cpu().numpy(), idd.cpu().numpy().view(np.uint32),
                   cnt.cpu().numpy().view(np.uint32), nb.cpu().numpy().view(np.uint64))
            _compare((oi, od, oid, oc, ob), res, o.is_u8)


def test_c1_configuration_one_query_at_a_time(gpu_lib):
    """The reference's CPU case (IVF F32, 65,536 x 128, 256 lists, k = 10, nprobe = 16), each query
    searched alone as the reference does, bit-exact with the oracle."""
    from tests.pdx_testdata import c1_index

    path, x, q = c1_index()
    o = OracleIndex(path)
    g = gpu_lib.GpuIndex(pf.load(path))
    oi, od, oid, oc, ob = o.search(q, 10, 16)
    for t in range(len(q)):
        gi, gd, gid, gc, gb = g.search(q[t:t + 1], 10, 16, want_bytes=True)
        np.testing.assert_array_equal(gc[0], oc[t])
        np.testing.assert_array_equal(gi[0, :oc[t]], oi[t, :oc[t]])
        np.testing.assert_array_equal(gd[0, :oc[t]].view(np.uint32), od[t, :oc[t]].view(np.uint32))
        np.testing.assert_array_equal(gb[0], ob[t])
