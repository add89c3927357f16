"""Maintenance through the kept Python API on the GPU path: appends concentrated on one region fill
clusters until they split (index.hpp:1314-1611), deletes empty clusters until they merge
(index.hpp:1257-1311), and after each phase the GPU search equals the oracle's on the saved index
(the device copy is re-created after every split or merge), each saved meso level holds every L1
cluster exactly once, and every live row is found."""
import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf
from tests.pdx_testdata import mixture

pytestmark = pytest.mark.gpu


def _check_saved(idx, path, q, live):
    idx.save(path)
    data = pf.load(path)
    n1 = data.l1.nc
    members = np.concatenate([c.ids[:c.n_used] for c in data.l0.clusters])
    assert sorted(members.tolist()) == list(range(n1)), "meso level must hold every L1 cluster once"
    rows = np.concatenate([c.ids[:c.n_used] for c in data.l1.clusters])
    assert sorted(rows.tolist()) == sorted(np.nonzero(live)[0].tolist())
    o = OracleIndex(path)
    for nprobe in (0, 6):
        oi, od, _, oc, _ = o.search(q, 10, nprobe)
        gi, gd, gc = idx.search_batch(q, 10, nprobe)
        np.testing.assert_array_equal(gc, oc)
        for t in range(len(q)):   # a query may return fewer than k rows; the rest of a row is unspecified
            np.testing.assert_array_equal(gi[t, :oc[t]], oi[t, :oc[t]], err_msg=f"nprobe={nprobe} q={t}")
            np.testing.assert_array_equal(gd[t, :oc[t]].view(np.uint32), od[t, :oc[t]].view(np.uint32))
    return n1, data.l0.nc


@pytest.mark.parametrize("kind", ["IndexPDXIVFTree", "IndexPDXIVFTreeSQ8"])
def test_splits_and_merges_keep_oracle_parity(kind, tmp_path):
    import pdx_amd

    d, n = 64, 3000
    x, q = mixture(n, d, 8, 0.3, 71, n_queries=32)
    idx = getattr(pdx_amd, kind)(num_dimensions=d, num_clusters=40, num_meso_clusters=6, normalize=False)
    idx.build(x)
    live = np.ones(n, bool)
    nc0, m0 = _check_saved(idx, str(tmp_path / "a.pdx"), q, live)
    rng = np.random.default_rng(71)
    new = (x[0] + np.float32(0.05) * rng.standard_normal((2500, d), dtype=np.float32)).astype(np.float32)
    for i, v in enumerate(new):
        idx.append(n + i, v)
    live = np.concatenate([live, np.ones(len(new), bool)])
    nc1, m1 = _check_saved(idx, str(tmp_path / "b.pdx"), np.concatenate([q, new[:16]]), live)
    print(f"{kind}: {nc0} -> {nc1} clusters after appends, meso {m0} -> {m1}", flush=True)
    assert nc1 > nc0
    # every appended row is found by its own vector
    ids, dists, cnt = idx.search_batch(new[:64], 1, 0)
    if kind == "IndexPDXIVFTree":
        np.testing.assert_array_equal(ids[:, 0], np.arange(n, n + 64))
    gone = rng.permutation(len(live))[: int(0.85 * len(live))]
    for r in gone:
        idx.delete(int(r))
    live[gone] = False
    nc2, m2 = _check_saved(idx, str(tmp_path / "c.pdx"), q, live)
    print(f"{kind}: {nc2} clusters after deletes, meso {m2}", flush=True)
    assert nc2 < nc1
