"""The device build (pdxg_build_rotation / _transform / _layout and builder.build_index on a GPU)
against the mathematics and the oracle:

* the rotation is orthogonal to FP64 tolerance and equals numpy's (LAPACK) QR of the same seeded
  Gaussian with the sign of diag(R) folded into Q -- the unique QR with a positive diagonal --
  and its float32 form is the double one rounded;
* every transformed row is bit-identical to the oracle's normalise + rotate (orc_preprocess, what
  a search computes for the same vector), through both rotation kernels, and the reported range
  is the exact min / max;
* every cluster buffer of the layout is byte-identical to the oracle's SQ8 quantisation
  (orc_quantize, quantizers/scalar.hpp:76-94) -- or the float rows -- put in the PDX layout by the
  host transposition (pdxfile.to_pdx), empty clusters and spare capacity included;
* build_index on the device stores exactly that for its own clustering, with the SQ8 parameters of
  scalar.hpp:60-74 over the oracle's rotated rows, and the oracle's search of the device-built
  index equals the GPU's.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleIndex
from pdx_amd import builder, pdxfile as pf
from tests.pdx_testdata import mixture

pytestmark = pytest.mark.gpu


def _dev():
    import torch

    return torch, torch.device("cuda", 0), torch.cuda.current_stream().cuda_stream


def _oracle_with(m: np.ndarray, d: int, normalize: bool) -> OracleIndex:
    """An oracle handle that only carries a rotation and the normalise flag (one tiny cluster)."""
    h, v = pf.dimension_split(d)
    L = pf.Level(d, v, h, False, normalize, [], np.zeros((1, d), np.float32))
    L.clusters.append(pf.make_cluster(np.zeros((1, d), np.float32), np.zeros(1, np.uint32), L))
    return OracleIndex(data=pf.IndexData(type=pf.TYPE_F32, matrix=m, l1=L))


@pytest.mark.parametrize("d", [64, 100, 384, 1024])
def test_rotation_is_the_positive_diagonal_qr(d):
    from pdx_amd import _capi

    torch, dev, st = _dev()
    g = builder.rotation_gaussian(d, 7 + d)
    gt = torch.from_numpy(g).to(dev)
    q32 = torch.empty((d, d), dtype=torch.float32, device=dev)
    q64 = torch.empty((d, d), dtype=torch.float64, device=dev)
    _capi.build_rotation(gt.data_ptr(), d, q32.data_ptr(), q64.data_ptr(), st)
    q = q64.cpu().numpy()
    np.testing.assert_allclose(q.T @ q, np.eye(d), rtol=0, atol=1e-12)
    qn, rn = np.linalg.qr(g)
    qn = qn * np.sign(np.diag(rn))[None, :]
    assert np.abs(q - qn).max() < 1e-11
    r = q.T @ g  # R = Q^T G: upper triangular with a positive diagonal
    assert np.abs(np.tril(r, -1)).max() < 1e-10 * np.abs(r).max()
    assert (np.diag(r) > 0).all()
    np.testing.assert_array_equal(q32.cpu().numpy(), q.astype(np.float32))


@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("n,d", [(500, 100), (3000, 768)])
def test_transform_matches_oracle_preprocess(n, d, normalize):
    from pdx_amd import _capi

    torch, dev, st = _dev()
    x, _ = mixture(n, d, 8, 0.5, 21 + n)
    x[3] = 0.0  # a zero row: normalisation leaves it as it is
    m = builder.random_rotation(d, 5)
    xt, mt = torch.from_numpy(x).to(dev), torch.from_numpy(m).to(dev)
    y = torch.empty_like(xt)
    mm = torch.empty(2, dtype=torch.float32, device=dev)
    _capi.build_transform(xt.data_ptr(), n, d, mt.data_ptr(), normalize, y.data_ptr(), mm.data_ptr(), st)
    ref = _oracle_with(m, d, normalize).preprocess(x)
    got = y.cpu().numpy()
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(mm.cpu().numpy(), np.array([ref.min(), ref.max()], np.float32))


@pytest.mark.parametrize("u8", [True, False])
@pytest.mark.parametrize("d", [100, 130, 384])
def test_layout_matches_host_transposition(d, u8):
    from pdx_amd import _capi

    torch, dev, st = _dev()
    n, nc = 2500, 7
    x, _ = mixture(n, d, 8, 0.5, 31 + d)
    y = _oracle_with(builder.random_rotation(d, 3), d, False).preprocess(x)
    rng = np.random.default_rng(d)
    assign = rng.integers(0, nc, n)
    assign[assign == 4] = 5  # cluster 4 is empty
    order = np.argsort(assign, kind="stable")
    first = np.searchsorted(assign[order], np.arange(nc + 1)).astype(np.uint64)
    counts = np.diff(first).astype(np.int64)
    caps = np.array([pf.default_capacity(int(counts[c])) + (17 if c == 2 else 0) for c in range(nc)], np.uint64)
    offs = np.concatenate([[0], np.cumsum(caps * np.uint64(d))]).astype(np.uint64)
    base, scale = float(np.float32(y.min())), float(np.float32(255.0) / np.float32(np.float32(y.max()) - np.float32(y.min())))
    h, v = pf.dimension_split(d)
    out = torch.empty(int(offs[-1]), dtype=torch.uint8 if u8 else torch.float32, device=dev)
    yt = torch.from_numpy(y).to(dev)
    ot = torch.from_numpy(order.astype(np.int32)).to(dev)
    _capi.build_layout(yt.data_ptr(), d, v, u8, base, scale, ot.data_ptr(), n, first, caps, offs[:-1], out.data_ptr(),
                       int(offs[-1]), st)
    buf = out.cpu().numpy()
    for c in range(nc):
        rows = order[first[c]:first[c + 1]]
        vals = orc.quantize(y[rows], base, scale).reshape(len(rows), d) if u8 else y[rows]
        want = pf.to_pdx(vals.reshape(len(rows), d), int(caps[c]), v, h, u8)
        got = buf[int(offs[c]):int(offs[c]) + int(caps[c]) * d]
        np.testing.assert_array_equal(got.view(np.uint8), want.view(np.uint8), err_msg=f"cluster {c} ({counts[c]} rows)")


@pytest.mark.parametrize("kind", ["pdx_tree_u8", "pdx_f32"])
def test_device_build_index_is_the_host_layout_of_oracle_rows(kind):
    import torch

    x, q = mixture(6000, 192, 24, 0.3, 41, n_queries=24)
    idx = builder.build_index(x, kind, num_clusters=40, seed=9, kmeans_iters=4, device=0)
    m = builder.random_rotation(192, 9)
    assert np.abs(idx.matrix - m).max() <= 2.0 ** -24  # the same Q up to the last float bit
    L = idx.l1
    y = _oracle_with(idx.matrix, 192, L.normalized).preprocess(x)
    if L.u8:  # ScalarQuantizer::ComputeQuantizationParams (scalar.hpp:60-74)
        lo, hi = np.float32(y.min()), np.float32(y.max())
        assert np.float32(L.base) == lo and np.float32(L.scale) == np.float32(255.0) / np.float32(hi - lo)
    seen = np.zeros(len(x), bool)
    for c in L.clusters:
        ids = c.ids[:c.n_used].astype(np.int64)
        seen[ids] = True
        vals = orc.quantize(y[ids], L.base, L.scale).reshape(len(ids), 192) if L.u8 else y[ids]
        np.testing.assert_array_equal(c.data.view(np.uint8), pf.to_pdx(vals, c.cap, L.v, L.h, L.u8).view(np.uint8))
        assert c.cap == pf.default_capacity(c.n_used)
    assert seen.all()
    # and the search of the device-built index is the oracle's
    from pdx_amd import _capi

    g = _capi.GpuIndex(idx)
    oi, od, _, oc, _ = OracleIndex(data=idx).search(q, 10, 8)
    gi, gd, _, gc, _ = g.search(q, 10, 8)
    np.testing.assert_array_equal(gc, oc)
    for t in range(len(q)):
        np.testing.assert_array_equal(gi[t, :oc[t]], oi[t, :oc[t]])
        np.testing.assert_array_equal(gd[t, :oc[t]].view(np.uint32), od[t, :oc[t]].view(np.uint32))
    torch.cuda.synchronize()
