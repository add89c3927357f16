"""Continuous-batching server: every result is the same alone, in a mixed batch and under timed
load, and equals the oracle's; admission follows the scratch budget."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf
from tests.pdx_testdata import index_file

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    from pdx_amd import _capi

    path, x, q = index_file("pdx_tree_u8", 4000, 768, 40)
    rng = np.random.default_rng(5)
    qs = np.concatenate([q, q + 0.01 * rng.standard_normal(q.shape).astype(np.float32)])
    k = rng.choice([10, 20, 100], len(qs)).astype(np.uint32)
    nprobe = rng.choice([4, 16, 0], len(qs)).astype(np.uint32)
    g = _capi.GpuIndex(pf.load(path))
    return _capi, path, g, qs, k, nprobe


def _collect(sv, n):
    out = {}
    while len(out) < n:
        for tag, ids, d, lat in sv.poll(256, 2000):
            out[tag] = (ids, d, lat)
    return out


def test_alone_equals_mixed_equals_oracle(setup):
    _capi, path, g, qs, k, nprobe = setup
    sv = _capi.Server(g, max_batch=16, max_k=100, max_nprobe=0, max_slots=3)
    alone = {}
    for i in range(len(qs)):
        sv.submit(qs[i], int(k[i]), int(nprobe[i]), i)
        alone.update(_collect(sv, 1))
    for i in range(len(qs)):
        sv.submit(qs[i], int(k[i]), int(nprobe[i]), 1000 + i)
    mixed = _collect(sv, len(qs))
    st = sv.stats()
    assert st["completed"] == 2 * len(qs) and st["slots"] == 3
    o = OracleIndex(path)
    oi, od, _, oc, _ = o.search(qs, k, nprobe)
    for i in range(len(qs)):
        a_ids, a_d, _ = alone[i]
        m_ids, m_d, _ = mixed[1000 + i]
        np.testing.assert_array_equal(a_ids, m_ids)
        np.testing.assert_array_equal(a_d.view(np.uint32), m_d.view(np.uint32))
        np.testing.assert_array_equal(a_ids, oi[i, :oc[i]])
        np.testing.assert_array_equal(a_d.view(np.uint32), od[i, :oc[i]].view(np.uint32))
    sv.close()


def test_replay_returns_every_result_with_latency(setup):
    _capi, path, g, qs, k, nprobe = setup
    sv = _capi.Server(g, max_batch=8, max_k=100, max_nprobe=0, max_slots=4)
    rng = np.random.default_rng(6)
    arrivals = np.cumsum(rng.exponential(1e-4, len(qs)))
    lat, elapsed = sv.replay(qs, k, nprobe, arrivals)
    assert np.all(lat > 0) and elapsed > 0
    assert sv.stats()["completed"] == len(qs)
    sv.close()


def test_scratch_budget_bounds_the_slots(setup):
    _capi, path, g, qs, k, nprobe = setup
    one = _capi.Server(g, max_batch=32, max_k=20, max_nprobe=16, max_slots=8)
    st = one.stats()
    one.close()
    per_slot = st["scratch_bytes"] // st["slots"]
    per_query = st["summary_capacity_bytes"] // 32      # the largest footprint of one request
    fixed = per_slot - st["summary_capacity_bytes"]
    # a budget for two slots of one request each shares out as two slots
    two = _capi.Server(g, max_batch=32, max_k=20, max_nprobe=16, max_slots=8,
                       scratch_budget_bytes=2 * (fixed + per_query) + 1)
    assert two.stats()["slots"] == 2 and two.stats()["summary_capacity_bytes"] == per_query
    two.close()
    # a generous budget: max_slots slots, each with whole summary buffers
    many = _capi.Server(g, max_batch=32, max_k=20, max_nprobe=16, max_slots=3, scratch_budget_bytes=4 * per_slot)
    assert many.stats()["slots"] == 3 and many.stats()["summary_capacity_bytes"] == st["summary_capacity_bytes"]
    many.close()
    with pytest.raises(_capi.PdxgError):
        _capi.Server(g, max_batch=32, max_k=20, max_nprobe=16, max_slots=8, scratch_budget_bytes=fixed // 2)
    s = _capi.Server(g, max_batch=4, max_k=20, max_nprobe=16, max_slots=1)
    with pytest.raises(_capi.PdxgError):
        s.submit(qs[0], 50, 4, 0)       # k above max_k
    with pytest.raises(_capi.PdxgError):
        s.submit(qs[0], 10, 0, 0)       # nprobe "all" above max_nprobe
    s.close()


def test_admission_by_footprint(setup):
    """A budget below one whole slot leaves one slot whose meso-summary buffer holds three requests
    of the largest footprint (nprobe "all": every meso cluster): batches close by footprint before
    max_batch, smaller footprints (nprobe 4: half the meso clusters) fit more per batch, and every
    result still equals the oracle's."""
    _capi, path, g, qs, k, nprobe = setup
    free = _capi.Server(g, max_batch=16, max_k=100, max_nprobe=0, max_slots=1)
    st = free.stats()
    free.close()
    per_query = st["summary_capacity_bytes"] // 16
    fixed = st["scratch_bytes"] - st["summary_capacity_bytes"]
    assert per_query > 0 and fixed > 0
    sv = _capi.Server(g, max_batch=16, max_k=100, max_nprobe=0, max_slots=1, scratch_budget_bytes=fixed + 3 * per_query)
    st = sv.stats()
    assert st["slots"] == 1 and st["summary_capacity_bytes"] == 3 * per_query
    o = OracleIndex(path)
    n = len(qs)
    for np_all in (0, 4):
        npb = np.full(n, np_all, np.uint32)
        before = sv.stats()
        for i in range(n):
            sv.submit(qs[i], int(k[i]), np_all, i)
        got = _collect(sv, n)
        after = sv.stats()
        batches = after["batches"] - before["batches"]
        admitted = after["admitted_summary_bytes"] - before["admitted_summary_bytes"]
        assert admitted <= batches * st["summary_capacity_bytes"]
        if np_all == 0:
            assert batches >= -(-n // 3)
            assert after["footprint_limited_batches"] > before["footprint_limited_batches"]
        oi, od, _, oc, _ = o.search(qs, k, npb)
        for i in range(n):
            ids, d, _ = got[i]
            np.testing.assert_array_equal(ids, oi[i, :oc[i]])
            np.testing.assert_array_equal(d.view(np.uint32), od[i, :oc[i]].view(np.uint32))
    sv.close()


def test_autotuned_plans_keep_results(setup):
    _capi, path, g, qs, k, nprobe = setup
    sv = _capi.Server(g, max_batch=20, max_k=100, max_nprobe=0, max_slots=2)
    assert sv.plans() == []
    plans = sv.autotune(qs[:12], k[:12], nprobe[:12])
    # batches of 1, 4, 16 and max_batch; blocks are workgroup sizes; the two-phase flag is 0/1
    assert [p[0] for p in plans] == [1, 4, 16, 20]
    assert all(p[1] in (64, 128, 256, 512, 1024) and p[2] in (64, 128, 256, 512, 1024) and p[3] in (0, 1)
               and p[4] in (0, 1, 2) for p in plans)
    for i in range(len(qs)):
        sv.submit(qs[i], int(k[i]), int(nprobe[i]), i)
    got = _collect(sv, len(qs))
    o = OracleIndex(path)
    oi, od, _, oc, _ = o.search(qs, k, nprobe)
    for i in range(len(qs)):
        ids, d, _ = got[i]
        np.testing.assert_array_equal(ids, oi[i, :oc[i]])
        np.testing.assert_array_equal(d.view(np.uint32), od[i, :oc[i]].view(np.uint32))
    sv.close()


@pytest.mark.parametrize("kind", ["IndexPDXIVFTreeSQ8", "IndexPDXIVF"])
def test_python_api_server(kind, tmp_path):
    """The server through the kept Python API (IndexPDXIVF*.serve): results equal the one-query
    search and the oracle, counters add up, the index cannot be changed while it runs."""
    import pdx_amd
    from tests.pdx_testdata import mixture

    x, q = mixture(4000, 128, 16, 0.3, 41, n_queries=40)
    idx = getattr(pdx_amd, kind)(num_dimensions=128, num_clusters=40)
    idx.build(x)
    path = str(tmp_path / "s.pdx")
    idx.save(path)
    rng = np.random.default_rng(4)
    k = rng.choice([1, 10, 20], len(q)).astype(np.uint32)
    nprobe = rng.choice([4, 16, 0], len(q)).astype(np.uint32)
    sv = idx.serve(max_batch=16, max_k=20, max_nprobe=0, max_slots=2)
    assert sv.autotune(q[:8], k[:8].tolist(), nprobe[:8].tolist())
    for i in range(len(q)):
        sv.submit(q[i], int(k[i]), int(nprobe[i]), 100 + i)
    got = {}
    while len(got) < len(q):
        for tag, ids, dists, lat in sv.poll(64, 2000):
            assert lat > 0
            got[tag] = (ids, dists)
    st = sv.stats()
    assert st["submitted"] == st["completed"] == len(q) and st["slots"] == 2
    with pytest.raises(RuntimeError):
        idx.save(path)                   # the server holds the device copy
    o = OracleIndex(path)
    oi, od, _, oc, _ = o.search(q, k, nprobe)
    for i in range(len(q)):
        ids, dists = got[100 + i]
        np.testing.assert_array_equal(ids, oi[i, :oc[i]])
        np.testing.assert_array_equal(dists.view(np.uint32), od[i, :oc[i]].view(np.uint32))
        a_ids, a_d = idx.search(q[i], int(k[i]), nprobe=int(nprobe[i]))
        np.testing.assert_array_equal(a_ids, ids)
    sv.close()
    with pytest.raises(RuntimeError):
        sv.submit(q[0], 10, 4, 0)
    idx.save(path)                       # allowed again once the server is closed
