"""Index construction on the host (no GPU): the vectorised PDX transposition equals the per-region
layout formulas (layout.hpp:21-87, cluster.hpp:341-384), and the F32 twin of an SQ8 index keeps
its clustering with rows whose SQ8 codes are the index's stored codes."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import builder, pdxfile as pf
from tests.pdx_testdata import mixture


@pytest.mark.parametrize("d", [4, 33, 64, 65, 100, 128, 130, 256, 384, 1024])
@pytest.mark.parametrize("u8", [True, False])
def test_pdx_layout_matches_address_formulas(d, u8):
    rng = np.random.default_rng(d)
    h, v = pf.dimension_split(d)
    n, cap = 37, 50
    x = rng.integers(0, 255, (n, d)).astype(np.uint8) if u8 else rng.standard_normal((n, d)).astype(np.float32)
    buf = pf.to_pdx(x, cap, v, h, u8)
    g4 = (v // 4) * 4
    for i in (0, 17, n - 1):
        for j in range(d):
            if j < v and not u8:
                addr = j * cap + i                                   # F32 vertical
            elif j < g4:
                addr = (j - j % 4) * cap + 4 * i + j % 4             # U8 vertical 4-dim group
            elif j < v:
                r = v - g4
                addr = g4 * cap + i * r + (j - g4)                   # U8 vertical tail
            else:
                hd = j - v
                addr = v * cap + (hd // 64) * 64 * cap + i * 64 + hd % 64   # horizontal 64-dim block
            assert buf[addr] == x[i, j], (i, j)
    np.testing.assert_array_equal(pf.from_pdx(buf, n, cap, d, v, h, u8), x)


@pytest.mark.parametrize("kind", ["pdx_tree_u8", "pdx_u8"])
def test_float_twin_keeps_the_clustering(kind):
    x, q = mixture(3000, 192, 16, 0.3, 8, n_queries=8)
    u = builder.build_index(x, kind, num_clusters=20, normalize=True, device=False, kmeans_iters=3)
    f = builder.float_twin(u, x, device=False)
    assert f.type == {"pdx_tree_u8": pf.TYPE_TREE_F32, "pdx_u8": pf.TYPE_F32}[kind]
    assert f.l0 is u.l0
    L, F = u.l1, f.l1
    for cu, cf in zip(L.clusters, F.clusters):
        assert (cu.n_used, cu.n_live, cu.cap) == (cf.n_used, cf.n_live, cf.cap)
        np.testing.assert_array_equal(cu.ids[:cu.n_used], cf.ids[:cf.n_used])
        rows = pf.from_pdx(cf.data, cf.n_used, cf.cap, F.d, F.v, F.h, False)
        codes = pf.from_pdx(cu.data, cu.n_used, cu.cap, L.d, L.v, L.h, True)
        np.testing.assert_array_equal(builder.quantize(rows, L.base, L.scale), codes)
    # the twin is a valid index: exhaustive search returns each query's exact top-k
    o = OracleIndex(data=f)
    qr = o.preprocess(q)
    ids, dists, _, cnt, _ = o.search(qr, 5, 0, transformed=True)
    allrows = np.concatenate([pf.from_pdx(c.data, c.n_used, c.cap, F.d, F.v, F.h, False) for c in F.clusters])
    allids = np.concatenate([c.ids[:c.n_used] for c in F.clusters])
    for t in range(len(qr)):
        dd = ((allrows.astype(np.float64) - qr[t]) ** 2).sum(1)
        assert set(ids[t, :5].tolist()) == set(allids[np.argsort(dd)[:5]].tolist())
