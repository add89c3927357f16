"""The reference's Python/C++ index API on the GPU path: build, save/load, search, batch search,
filtered search, append/delete -- checked against the oracle on the saved file and against the
reference's own behavioural tests (test_maintenance.cpp, test_filtered_search.cpp,
test_index_properties.cpp)."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from tests.pdx_testdata import mixture

pytestmark = pytest.mark.gpu

KINDS = ["IndexPDXIVF", "IndexPDXIVFSQ8", "IndexPDXIVFTree", "IndexPDXIVFTreeSQ8"]


def _make(kind, x, **kw):
    import pdx_amd

    idx = getattr(pdx_amd, kind)(num_dimensions=x.shape[1], **kw)
    idx.build(x)
    return idx


@pytest.mark.parametrize("kind", KINDS)
def test_api_search_matches_oracle_on_saved_file(kind, tmp_path):
    x, q = mixture(4000, 192, 24, 0.3, 11, n_queries=16)
    idx = _make(kind, x, num_clusters=30)
    path = str(tmp_path / "i.pdx")
    idx.save(path)
    o = OracleIndex(path)
    oi, od, _, oc, _ = o.search(q, 10, 8)
    for t in range(len(q)):
        ids, dists = idx.search(q[t], 10, nprobe=8)
        np.testing.assert_array_equal(ids, oi[t, :oc[t]])
        np.testing.assert_array_equal(dists.view(np.uint32), od[t, :oc[t]].view(np.uint32))
    bi, bd, bc = idx.search_batch(q, 10, 8)
    np.testing.assert_array_equal(bc, oc)
    np.testing.assert_array_equal(bi, oi)
    # reload through load_index
    import pdx_amd

    l = pdx_amd.load_index(path)
    l.set_nprobe(8)
    ids, dists = l.search(q[0], 10)
    np.testing.assert_array_equal(ids, oi[0, :oc[0]])


@pytest.mark.parametrize("kind", ["IndexPDXIVF", "IndexPDXIVFSQ8"])
def test_filtered_search_semantics(kind, tmp_path):
    x, q = mixture(3000, 128, 24, 0.3, 12, n_queries=8)
    idx = _make(kind, x, num_clusters=20)
    rng = np.random.default_rng(0)
    rows = np.sort(rng.choice(3000, 300, replace=False)).astype(np.uint64)
    for t in range(len(q)):
        ids, _ = idx.filtered_search(q[t], 10, rows, nprobe=20)
        assert set(ids.tolist()) <= set(rows.tolist())            # filtered subset of passing rows
    allrows = np.arange(3000, dtype=np.uint64)
    for t in range(len(q)):                                       # all-pass equals unfiltered
        a, ad = idx.filtered_search(q[t], 10, allrows, nprobe=8)
        b, bd = idx.search(q[t], 10, nprobe=8)
        np.testing.assert_array_equal(np.sort(a), np.sort(b))
    e, _ = idx.filtered_search(q[0], 10, np.array([], np.uint64), nprobe=20)
    assert len(e) == 0                                            # empty filter
    one = np.array([1234], np.uint64)
    s, _ = idx.filtered_search(q[0], 10, one, nprobe=20)
    assert s.tolist() == [1234]                                   # single id
    # GPU filtered search vs the oracle's on the same selection vector
    path = str(tmp_path / "f.pdx")
    idx.save(path)
    o = OracleIndex(path)
    off = o.cluster_offsets()
    sel = np.zeros(int(off[-1]), np.uint8)
    wanted = set(rows.tolist())
    for c in range(o.num_clusters):
        cid = o.cluster_ids(c)
        for slot, r in enumerate(cid):
            if int(r) in wanted:
                sel[int(off[c]) + slot] = 1
    from pdx_amd import _capi, pdxfile as pf

    g = _capi.GpuIndex(pf.load(path))
    for t in range(len(q)):
        gi, gd, _ = g.filtered_search(q[t], 10, 6, sel)
        oi, od, _ = o.filtered_search(q[t], 10, 6, sel)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("kind", ["IndexPDXIVFTree", "IndexPDXIVFTreeSQ8"])
def test_append_then_found_delete_then_absent(kind):
    x, q = mixture(3000, 128, 24, 0.3, 13, n_queries=4)
    idx = _make(kind, x, num_clusters=20)
    n = 3000
    new = q[:2] + 0.001
    for i, v in enumerate(new):
        idx.append(n + i, v)
    for i, v in enumerate(new):
        ids, _ = idx.search(v, 5, nprobe=20)
        assert n + i in ids.tolist()
    idx.delete(n)
    ids, _ = idx.search(new[0], 5, nprobe=20)
    assert n not in ids.tolist()
    with pytest.raises(ValueError):
        idx.delete(n)                      # already deleted
    with pytest.raises(ValueError):
        idx.append(n + 10, new[0])         # not sequential


def test_flat_indexes_reject_maintenance():
    x, _ = mixture(500, 32, 4, 0.3, 14)
    idx = _make("IndexPDXIVF", x, num_clusters=4)
    with pytest.raises(RuntimeError):
        idx.append(500, x[0])
    with pytest.raises(RuntimeError):
        idx.delete(0)


def test_knn_larger_than_n_and_tiny_indexes():
    x, q = mixture(40, 16, 2, 0.3, 15, n_queries=2)
    idx = _make("IndexPDXIVFSQ8", x, num_clusters=2)
    ids, d = idx.search(q[0], 100, nprobe=0)
    assert len(ids) == 40 and len(set(ids.tolist())) == 40
    one = _make("IndexPDXIVF", x[:1], num_clusters=1)
    ids, d = one.search(q[0], 5, nprobe=1)
    assert ids.tolist() == [0]
