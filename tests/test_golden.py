"""Golden fixtures (tests/golden/, written by make_golden.py): two tiny indexes in the reference's
file format, seeded queries and the oracle's results at fixed k / nprobe. The oracle must still
produce them, and the GPU path must produce them bit for bit (ids, float and integer distances,
counts, algorithmic bytes)."""
import os

import numpy as np
import pytest

from oracle.oracle import OracleIndex
from pdx_amd import pdxfile as pf

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def _expected(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    tags = sorted({k.rsplit("_", 1)[0] for k in z.files if k != "queries"})
    return z, tags


def _kn(tag):
    k, npb = tag.split("_")
    return int(k[1:]), int(npb[2:])


def _check(z, tag, got, u8):
    ids, dist, idist, cnt, nbytes = got
    np.testing.assert_array_equal(cnt, z[tag + "_count"])
    np.testing.assert_array_equal(ids, z[tag + "_ids"])
    np.testing.assert_array_equal(dist.view(np.uint32), z[tag + "_dist"].view(np.uint32))
    if u8:
        np.testing.assert_array_equal(idist, z[tag + "_idist"])
    np.testing.assert_array_equal(nbytes, z[tag + "_bytes"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    z, tags = _expected(name)
    o = OracleIndex(os.path.join(GOLDEN, name + ".pdx"))
    assert len(tags) == 9
    for tag in tags:
        k, npb = _kn(tag)
        _check(z, tag, o.search(z["queries"], k, npb), o.is_u8)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_reproduces_golden(name):
    from pdx_amd import _capi

    z, tags = _expected(name)
    idx = pf.load(os.path.join(GOLDEN, name + ".pdx"))
    g = _capi.GpuIndex(idx)
    for tag in tags:
        k, npb = _kn(tag)
        print(f"{name} {tag}", flush=True)
        gi, gd, gid, gc, gb = g.search(z["queries"], k, npb, want_bytes=True)
        _check(z, tag, (gi, gd, gid, gc, gb), idx.is_u8)
