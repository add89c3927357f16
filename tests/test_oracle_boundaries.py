"""The oracle against hand-derived search fixtures (tests/boundary_cases.py): the warmup exit rule
(count of dist >= tau against ceil(0.8 n)), the fetch schedule checkpoints, the two-level meso
halving at nprobe = nlist / 2, the U8 threshold truncation at cur == d, the filtered warmup skip
below 20 % selectivity, and tombstone masking in Warmup and Start (searcher.hpp:169-179,376-540,
726-856; index.hpp:545-566). Each expected id, distance and byte count is written out by hand in
the case's docstring; a change of any of those rules in the oracle fails at least one case."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from tests import boundary_cases as bc

CASES = bc.all_cases()


def run_oracle(case: bc.Case):
    o = OracleIndex(data=case.index)
    if case.selection is not None:
        ids, dists, idists = o.filtered_search(case.queries[0], case.k, case.nprobe, case.selection, transformed=True)
        return ids, dists, idists, None
    ids, dists, idists, counts, nbytes = o.search(case.queries, case.k, case.nprobe, transformed=True)
    n = int(counts[0])
    return ids[0, :n], dists[0, :n], idists[0, :n], int(nbytes[0])


def check(case: bc.Case, ids, dists, idists, nbytes):
    assert ids.tolist() == case.ids, f"{case.name}: ids"
    np.testing.assert_array_equal(np.asarray(dists, np.float32), np.asarray(case.dists, np.float32),
                                  err_msg=f"{case.name}: distances")
    if case.int_dists is not None:
        assert [int(x) for x in idists] == case.int_dists, f"{case.name}: integer distances"
    if case.nbytes is not None and nbytes is not None:
        assert nbytes == case.nbytes, f"{case.name}: algorithmic bytes {nbytes} != {case.nbytes}"


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_oracle_matches_hand_derived_case(case):
    check(case, *run_oracle(case))


def test_every_rule_is_pinned_by_a_case():
    rules = {m for c in CASES for m in c.mutations}
    assert {"warmup count dist >= tau (not >)", "selectivity 0.8", "tuples needed = ceil(0.8 n)",
            "DIMENSIONS_FETCHING_SIZES", "meso halving nprobe < nlist / 2", "U8 tau truncation at cur == d",
            "filtered warmup skip below 20 %", "tombstones masked in Warmup",
            "tombstones masked in Start, k > live"} <= rules


def test_oracle_meso_search_exhaustive_is_sorted_start_distances():
    """The meso-level search alone (the cluster choice of an append, index.hpp:596-607) with k = all
    L1 clusters runs Start on every meso cluster: it returns every L1 centroid, ascending by the
    distance Start computes for a vector it scans whole, ties by id; k = 1 returns the first."""
    from oracle import oracle as orc
    from pdx_amd import builder
    from tests.pdx_testdata import mixture

    x, q = mixture(2000, 192, 16, 0.3, 4, n_queries=8)
    idx = builder.build_index(x, "pdx_tree_u8", num_clusters=30, seed=4, kmeans_iters=3, device=False)
    o = OracleIndex(data=idx)
    qr = o.preprocess(q)
    n1, v = idx.l1.nc, idx.l1.v
    ids, dists, counts = o.search_meso(qr, n1, 0)
    assert (counts == n1).all()
    for t in range(len(qr)):
        sd = orc.start_distances(qr[t], idx.l1.centroids, v)
        order = np.lexsort((np.arange(n1), sd))
        np.testing.assert_array_equal(ids[t], order)
        np.testing.assert_array_equal(dists[t].view(np.uint32), sd[order].view(np.uint32))
    i1, d1, c1 = o.search_meso(qr, 1, 0)
    np.testing.assert_array_equal(i1[:, 0], ids[:, 0])
