"""Pins for the CPU oracle (no GPU): closed forms, exact arithmetic and search invariants.

The reference's own known answers for this path are its static_asserts on the dimension split
(common.hpp:161-189), the ADSampling ratio formula (adsampling.hpp:147-157), the integer SQ8
distance semantics (scalar_computers.hpp:77-134) and its search-level invariants
(test_search.cpp, test_index_properties.cpp). Its recall table needs data that is not shipped.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import OracleIndex
from pdx_amd import builder, pdxfile as pf
from tests.pdx_testdata import index_file, mixture


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("d,h,v", [(4, 0, 4), (33, 0, 33), (64, 0, 64), (65, 0, 65), (100, 0, 100),
                                   (127, 0, 127), (128, 64, 64), (256, 192, 64), (1024, 768, 256),
                                   (1028, 768, 260)])
def test_split_matches_reference_static_asserts(d, h, v):
    assert orc.split(d) == (h, v)
    assert pf.dimension_split(d) == (h, v)


def test_split_for_bench_shapes():
    # horizontal = 0.75 d rounded to a multiple of 64 (common.hpp:140-146)
    assert orc.split(768) == (576, 192)
    assert orc.split(1536) == (1152, 384)
    assert orc.split(3072) == (2304, 768)


def test_ratio_table_closed_form():
    for D in (100, 128, 768, 1024):
        r = orc.ratios(D)
        assert r[0] == 1.0
        i = np.arange(1, D)
        f = (np.float32(i) / np.float32(D)).astype(np.float64)
        a = 1.0 + np.float64(np.float32(1.5)) / np.sqrt(i.astype(np.float64))
        np.testing.assert_array_equal(r[1:], (f * a * a).astype(np.float32))
    # hand-computed anchor: D=768, i=16 -> (16/768) * (1 + 1.5/4)^2 = 0.0208333 * 1.890625
    assert abs(float(orc.ratios(768)[16]) - 0.0393880208) < 4e-9  # one float32 ulp


def test_l2sq16_uses_the_avx512_reduction_order():
    # lane l accumulates dims l, l+16, ...; lanes combine (i,i+8), (i,i+4), (i,i^1), (i,i^2).
    a = np.zeros(32, np.float32)
    a[0] = 4096.0          # 2^24 after squaring
    a[1] = 1.0             # lanes 1 and 9 meet first (1 + 1 = 2) and then join 2^24 exactly
    a[9] = 1.0
    b = np.zeros(32, np.float32)
    assert orc.l2sq16(a, b) == 16777218.0
    # a sequential sum would absorb each 1 separately (2^24 + 1 rounds to 2^24)
    seq = np.float32(0)
    for x in a:
        seq = np.float32(seq + np.float32(x * x))
    assert seq == 16777216.0
    # lanes 1 and 5 pair at the second level
    a2 = np.zeros(16, np.float32)
    a2[[0, 1, 5]] = [4096.0, 1.0, 1.0]
    assert orc.l2sq16(a2, np.zeros(16, np.float32)) == 16777218.0
    # lanes 1 and 9 pair at the first level: the 1 is absorbed before lane 0's 1 can join it
    a3 = np.zeros(16, np.float32)
    a3[[1, 9, 0]] = [4096.0, 1.0, 1.0]
    assert orc.l2sq16(a3, np.zeros(16, np.float32)) == 16777216.0


@pytest.mark.parametrize("d", [1, 7, 15, 16, 31, 64, 100, 384, 768])
def test_l2sq16_close_to_fp64(d):
    rng = np.random.default_rng(d)
    for _ in range(20):
        a = rng.uniform(-10, 10, d).astype(np.float32)
        b = rng.uniform(-10, 10, d).astype(np.float32)
        exact = float(np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2))
        assert abs(orc.l2sq16(a, b) - exact) <= 1e-5 * max(exact, 1e-6)


def test_l2sq16_exact_on_small_integers():
    rng = np.random.default_rng(1)
    a = rng.integers(-50, 50, 300).astype(np.float32)
    b = rng.integers(-50, 50, 300).astype(np.float32)
    assert orc.l2sq16(a, b) == float(np.sum((a - b).astype(np.int64) ** 2))
    assert orc.l2sq16(a, a) == 0.0


def test_l2sq_u8_exact():
    rng = np.random.default_rng(2)
    for d in (1, 4, 63, 64, 384, 1024):
        a = rng.integers(0, 256, d, dtype=np.uint8)
        b = rng.integers(0, 256, d, dtype=np.uint8)
        assert orc.l2sq_u8(a, b) == int(np.sum((a.astype(np.int64) - b.astype(np.int64)) ** 2))
    full = np.full(16384, 255, np.uint8)
    assert orc.l2sq_u8(full, np.zeros(16384, np.uint8)) == 16384 * 255 * 255


def test_quantize_rounds_half_away_from_zero_and_clamps():
    x = np.array([0.5, 1.5, 2.5, -0.4, -3.0, 254.5, 300.0, 0.49999997], np.float32)
    np.testing.assert_array_equal(orc.quantize(x, 0.0, 1.0), [1, 2, 3, 0, 0, 255, 255, 0])
    # (x - base) * scale in float32, then round
    x = np.array([-1.0, 0.0, 1.0], np.float32)
    np.testing.assert_array_equal(orc.quantize(x, -1.0, 127.5), [0, 128, 255])
    np.testing.assert_array_equal(builder.quantize(x, -1.0, 127.5), [0, 128, 255])


def test_normalize():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(300).astype(np.float32)
    y = orc.normalize(x)
    np.testing.assert_allclose(y, x / np.linalg.norm(x.astype(np.float64)), rtol=2e-6, atol=1e-7)
    z = np.zeros(17, np.float32)
    np.testing.assert_array_equal(orc.normalize(z), z)


# ---------------------------------------------------------------- file format
def test_pdx_layout_round_trip():
    rng = np.random.default_rng(4)
    for d, u8 in ((130, True), (130, False), (33, True), (768, True), (100, False)):
        h, v = pf.dimension_split(d)
        vecs = rng.integers(0, 256, (37, d)).astype(np.uint8 if u8 else np.float32)
        buf = pf.to_pdx(vecs, 256, v, h, u8)
        np.testing.assert_array_equal(pf.from_pdx(buf, 37, 256, d, v, h, u8), vecs)
        # reference address formulas (layout.hpp:37-47, 68-86)
        i, j = 5, 0
        if not u8:
            assert buf[j * 256 + i] == vecs[i, j]
        elif v >= 4:
            assert buf[j * 256 + 4 * i] == vecs[i, j]
        else:  # tail group of v < 4 dims stored as v bytes per vector (layout.hpp:75-79)
            assert buf[j * 256 + v * i] == vecs[i, j]
        if h:
            assert buf[v * 256 + 64 * i + 3] == vecs[i, v + 3]


@pytest.mark.parametrize("kind", ["pdx_f32", "pdx_u8", "pdx_tree_f32", "pdx_tree_u8"])
def test_file_round_trip_is_byte_identical(kind, tmp_path):
    path, _, _ = index_file(kind, 1500, 130, 12)
    idx = pf.load(path)
    p2 = str(tmp_path / "again.pdx")
    pf.save(idx, p2)
    assert open(path, "rb").read() == open(p2, "rb").read()
    o = OracleIndex(p2)
    assert o.num_clusters == 12 and o.d == 130


# ---------------------------------------------------------------- search invariants
def _decoded(path):
    """Stored (rotated, quantised for SQ8) vectors of every live slot and their row ids."""
    idx = pf.load(path)
    L = idx.l1
    vecs, ids = [], []
    for c in L.clusters:
        vecs.append(pf.from_pdx(c.data, c.n_used, c.cap, L.d, L.v, L.h, L.u8))
        ids.append(c.ids[:c.n_used])
    return idx, np.concatenate(vecs), np.concatenate(ids)


@pytest.mark.parametrize("kind", ["pdx_f32", "pdx_u8", "pdx_tree_f32", "pdx_tree_u8"])
def test_k_above_n_returns_every_vector_sorted(kind):
    # the heap never fills, so every cluster goes through Start: the result is the full,
    # unpruned ranking by (distance, id) (searcher.hpp:218-281, 568-585)
    path, x, q = index_file(kind, 300, 128, 4)
    o = OracleIndex(path)
    idx, vecs, ids = _decoded(path)
    qr = o.preprocess(q[:4])
    k = 400
    rid, rd, rint, cnt, _ = o.search(qr, k, 0, transformed=True)
    for t in range(4):
        assert cnt[t] == 300
        if idx.is_u8:
            qq = orc.quantize(qr[t], o.quant_base, o.quant_scale).astype(np.int64)
            exact = ((vecs.astype(np.int64) - qq[None, :]) ** 2).sum(1)
            order = np.lexsort((ids, exact))
            np.testing.assert_array_equal(rid[t, :300], ids[order])
            np.testing.assert_array_equal(rint[t, :300], exact[order])
        else:
            exact = ((vecs.astype(np.float64) - qr[t][None, :].astype(np.float64)) ** 2).sum(1)
            got = dict(zip(rid[t, :300].tolist(), rd[t, :300].tolist()))
            for i, e in zip(ids, exact):
                assert abs(got[int(i)] - e) <= 1e-5 * max(e, 1e-6)
            assert np.all(np.diff(rd[t, :300]) >= 0)


@pytest.mark.parametrize("kind", ["pdx_f32", "pdx_u8", "pdx_tree_f32", "pdx_tree_u8"])
def test_returned_distances_are_full_distances(kind):
    path, x, q = index_file(kind, 3000, 384, 20)
    o = OracleIndex(path)
    idx, vecs, ids = _decoded(path)
    pos = {int(i): n for n, i in enumerate(ids)}
    qr = o.preprocess(q)
    rid, rd, rint, cnt, _ = o.search(qr, 20, 6, transformed=True)
    for t in range(len(q)):
        n = int(cnt[t])
        assert n == 20
        assert len(set(rid[t, :n].tolist())) == n
        assert np.all(np.diff(rd[t, :n]) >= 0)
        for r in range(n):
            v = vecs[pos[int(rid[t, r])]]
            if idx.is_u8:
                qq = orc.quantize(qr[t], o.quant_base, o.quant_scale).astype(np.int64)
                e = int(((v.astype(np.int64) - qq) ** 2).sum())
                assert int(rint[t, r]) == e
                assert rd[t, r] == np.float32(np.float32(e) * np.float32(o.inv_scale_sq))
            else:
                e = float(((v.astype(np.float64) - qr[t].astype(np.float64)) ** 2).sum())
                assert abs(rd[t, r] - e) <= 1e-5 * max(e, 1e-6)


@pytest.mark.parametrize("d", [48, 128, 130, 384])
def test_start_distances_equal_the_search_of_a_single_cluster(d):
    """With one cluster the search scans every vector through Start, so its distances are the
    Start-order distances: start_distances must reproduce them bit for bit (ids in the same
    (distance, id) order), and stay within FP32 rounding of the FP64 distance."""
    path, x, q = index_file("pdx_f32", 300, d, 1)
    o = OracleIndex(path)
    idx, vecs, ids = _decoded(path)
    qr = o.preprocess(q)
    rid, rd, _, cnt, _ = o.search(qr, 300, 1, transformed=True)
    for t in range(len(q)):
        sd = orc.start_distances(qr[t], vecs, idx.l1.v)
        order = np.lexsort((ids, sd))
        assert int(cnt[t]) == 300
        np.testing.assert_array_equal(rid[t], ids[order])
        np.testing.assert_array_equal(rd[t].view(np.uint32), sd[order].view(np.uint32))
        e = ((vecs.astype(np.float64) - qr[t].astype(np.float64)) ** 2).sum(1)
        np.testing.assert_allclose(sd, e, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("kind", ["pdx_f32", "pdx_u8", "pdx_tree_f32", "pdx_tree_u8"])
def test_exhaustive_recall_and_nprobe_equivalences(kind):
    path, x, q = index_file(kind, 3000, 128, 24)
    o = OracleIndex(path)
    idx, vecs, ids = _decoded(path)
    qr = o.preprocess(q)
    res = [o.search(qr, 10, npb, transformed=True) for npb in (24, 25, 0)]
    for r in res[1:]:
        np.testing.assert_array_equal(r[0], res[0][0])
        np.testing.assert_array_equal(r[1], res[0][1])
    # exhaustive recall against the exact ranking of the stored vectors
    rec = []
    for t in range(len(q)):
        if idx.is_u8:
            qq = orc.quantize(qr[t], o.quant_base, o.quant_scale).astype(np.int64)
            ex = ((vecs.astype(np.int64) - qq[None, :]) ** 2).sum(1)
        else:
            ex = ((vecs.astype(np.float64) - qr[t][None, :]) ** 2).sum(1)
        gt = set(ids[np.argsort(ex, kind="stable")[:10]].tolist())
        rec.append(len(gt & set(res[0][0][t].tolist())) / 10)
    assert np.mean(rec) >= 0.98


def test_recall_grows_with_nprobe():
    path, x, q = index_file("pdx_u8", 3000, 128, 24)
    o = OracleIndex(path)
    xr = o.preprocess(x)
    qr = o.preprocess(q)
    d2 = ((qr[:, None, :].astype(np.float64) - xr[None, :, :]) ** 2).sum(-1)
    gt = np.argsort(d2, axis=1)[:, :10]
    prev = 0.0
    for npb in (1, 2, 4, 8, 24):
        ids = o.search(qr, 10, npb, transformed=True)[0]
        r = np.mean([len(set(ids[t]) & set(gt[t])) / 10 for t in range(len(q))])
        assert r >= prev - 0.02
        prev = r
    assert prev >= 0.9


@pytest.mark.parametrize("n", [1, 2])
def test_tiny_indexes(n, tmp_path):
    x, q = mixture(n, 16, 1, 0.1, 5, n_queries=3)
    for kind in ("pdx_f32", "pdx_u8"):
        idx = builder.build_index(x, kind, num_clusters=1, device=False)
        p = str(tmp_path / f"{kind}_{n}.pdx")
        pf.save(idx, p)
        o = OracleIndex(p)
        rid, rd, _, cnt, _ = o.search(q, 5, 1)
        assert list(cnt) == [n] * 3
        for t in range(3):
            assert sorted(rid[t, :n].tolist()) == list(range(n))


def test_deleted_slots_are_never_returned_when_enough_live_vectors():
    path, x, q = index_file("pdx_u8", 3000, 128, 24)
    o = OracleIndex(path)
    qr = o.preprocess(q)
    before = o.search(qr, 10, 4, transformed=True)[0]
    victim = int(before[0, 0])
    # find and tombstone the victim's slot
    for c in range(o.num_clusters):
        ids = o.cluster_ids(c)
        hit = np.nonzero(ids == victim)[0]
        if len(hit):
            o.delete_slot(c, int(hit[0]))
            break
    after = o.search(qr, 10, 4, transformed=True)[0]
    assert victim not in after[0].tolist()


def test_rotation_is_the_matrix_product_and_preserves_norms():
    """PreprocessQuery: y = x . M with M the stored row-major matrix (adsampling.hpp:225-245). The
    oracle's fma chain agrees with an FP64 matrix product to FP32 rounding, and the orthogonal
    matrix keeps L2 norms (the reference's cosine = L2 test relies on it, test_search.cpp:289-333)."""
    path, x, q = index_file("pdx_f32", 2000, 96, 8)
    idx = pf.load(path)
    o = OracleIndex(path)
    y = o.preprocess(q)
    ref = q.astype(np.float64) @ idx.matrix.astype(np.float64)
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    np.testing.assert_allclose(np.linalg.norm(y, axis=1), np.linalg.norm(q.astype(np.float64), axis=1), rtol=1e-5)
    # a transposed matrix is a different map: the test would catch y = M . x
    wrong = q.astype(np.float64) @ idx.matrix.astype(np.float64).T
    assert np.abs(wrong - ref).max() > 1e-2


def test_c1_configuration_recall():
    """The reference's CPU case (IVF F32, 65,536 x 128, 256 lists, k = 10, nprobe = 16, one query at
    a time): the oracle's answers against FP64 brute force."""
    from tests.pdx_testdata import c1_index

    path, x, q = c1_index()
    o = OracleIndex(path)
    ids, dist, _, cnt, _ = o.search(q, 10, 16)
    x64 = x.astype(np.float64)
    gt = [np.argsort(((x64 - qq.astype(np.float64)) ** 2).sum(1), kind="stable")[:10] for qq in q]
    rec = np.mean([len(set(ids[i, :cnt[i]].tolist()) & set(gt[i].tolist())) / 10 for i in range(len(q))])
    assert rec >= 0.99
