"""Writes the golden fixtures of this directory: two tiny indexes in the reference's single-file
format (include/pdx/index.hpp:213-239, 483-514) with seeded queries, and the oracle's results for
them (ids, distances, integer distances, counts, algorithmic bytes) at fixed k / nprobe.

The fixtures freeze the oracle's answers so that a change to the oracle or to the GPU path that
alters any result is caught by a test that reads only these files. Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import OracleIndex  # noqa: E402
from pdx_amd import builder, pdxfile as pf  # noqa: E402

CASES = {
    # name: (index type, n, d, nlist, normalize, k list, nprobe list)
    "tree_u8_n800_d96": ("pdx_tree_u8", 800, 96, 12, True, (1, 5, 20), (1, 3, 0)),
    "ivf_f32_n400_d136": ("pdx_f32", 400, 136, 8, False, (1, 10, 50), (2, 4, 0)),
}


def make(name, kind, n, d, nlist, normalize, ks, nprobes):
    rng = np.random.default_rng(20261204 + n + d)
    c = rng.standard_normal((6, d)).astype(np.float32)
    x = (c[rng.integers(0, 6, n)] + 0.4 * rng.standard_normal((n, d))).astype(np.float32)
    q = (c[rng.integers(0, 6, 12)] + 0.4 * rng.standard_normal((12, d))).astype(np.float32)
    idx = builder.build_index(x, kind, num_clusters=nlist, normalize=normalize, seed=5, kmeans_iters=3, device=False)
    path = os.path.join(HERE, name + ".pdx")
    pf.save(idx, path)
    o = OracleIndex(path)
    out = {"queries": q}
    for k in ks:
        for npb in nprobes:
            ids, dist, idist, cnt, nbytes = o.search(q, k, npb)
            tag = f"k{k}_np{npb}"
            out[tag + "_ids"] = ids
            out[tag + "_dist"] = dist
            out[tag + "_idist"] = idist
            out[tag + "_count"] = cnt
            out[tag + "_bytes"] = nbytes
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


if __name__ == "__main__":
    for name, args in CASES.items():
        make(name, *args)
        print("wrote", name)
