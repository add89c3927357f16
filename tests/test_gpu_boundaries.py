"""The HIP path on the hand-derived boundary fixtures of tests/boundary_cases.py: ids, float and
integer distances and algorithmic bytes equal the values worked out by hand from the reference's
Warmup / Prune / Merge rules (and therefore the oracle's, which tests/test_oracle_boundaries.py
checks against the same values)."""
import numpy as np
import pytest

from tests import boundary_cases as bc
from tests.test_oracle_boundaries import check, run_oracle

pytestmark = pytest.mark.gpu

CASES = bc.all_cases()


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_gpu_matches_hand_derived_case(case):
    from pdx_amd import _capi

    print(case.name, flush=True)
    g = _capi.GpuIndex(case.index)
    if case.selection is not None:
        ids, dists, idists = g.filtered_search(case.queries[0], case.k, case.nprobe, case.selection, transformed=True)
        got = (ids, dists, idists, None)
    else:
        ids, dists, idists, counts, nbytes = g.search(case.queries, case.k, case.nprobe, transformed=True,
                                                      want_bytes=True)
        n = int(counts[0])
        got = (ids[0, :n], dists[0, :n], idists[0, :n], int(nbytes[0]))
    check(case, *got)
    # and the same as the oracle, bit for bit
    oi, od, oid, ob = run_oracle(case)
    np.testing.assert_array_equal(got[0], oi)
    np.testing.assert_array_equal(np.asarray(got[1], np.float32).view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("u8", [True, False], ids=["u8", "f32"])
def test_gpu_checkpoints_batched(u8):
    """Each checkpoint case searched as a batch of four queries: every query's byte count stops at
    its checkpoint (16, 32, ..., 256 dims) inside one batched launch."""
    from pdx_amd import _capi

    cases = [bc.case_fetch_checkpoint(u8, t) for t in range(len(bc.CHECKPOINTS_1024))]
    for c in cases:
        g = _capi.GpuIndex(c.index)
        q = np.repeat(c.queries, 4, axis=0)
        ids, dists, idists, counts, nbytes = g.search(q, c.k, c.nprobe, transformed=True, want_bytes=True)
        assert (nbytes == c.nbytes).all(), (c.name, nbytes)
        assert (ids[:, 0] == c.ids[0]).all()
