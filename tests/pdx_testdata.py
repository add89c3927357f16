"""Seeded synthetic data and small indexes shared by the tests."""
from __future__ import annotations

import os
import tempfile

import numpy as np

from pdx_amd import builder, pdxfile as pf

_CACHE = {}
_DIR = tempfile.mkdtemp(prefix="pdx_tests_")


def mixture(n, d, n_centers, sigma, seed, unit=False, n_queries=0, qnoise=0.05):
    """Gaussian mixture: centres ~ N(0, I), rows = centre + sigma * N(0, I)."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n_centers, d)).astype(np.float32)
    if unit:
        c /= np.linalg.norm(c, axis=1, keepdims=True)
    lab = rng.integers(0, n_centers, n)
    x = (c[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)
    ql = rng.integers(0, n_centers, n_queries)
    q = (c[ql] + sigma * rng.standard_normal((n_queries, d)) + qnoise * rng.standard_normal((n_queries, d))).astype(np.float32)
    if unit:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        if n_queries:
            q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, q


def index_file(kind: str, n: int, d: int, nlist: int, seed: int = 7, normalize: bool = False, n_meso: int = 0,
               n_centers: int = 32, sigma: float = 0.3):
    """Build (once per session) and save an index; returns (path, data, queries)."""
    key = (kind, n, d, nlist, seed, normalize, n_meso, n_centers, sigma)
    if key not in _CACHE:
        x, q = mixture(n, d, n_centers, sigma, seed, n_queries=24)
        idx = builder.build_index(x, kind, num_clusters=nlist, num_meso_clusters=n_meso, normalize=normalize,
                                  seed=seed, kmeans_iters=4, device=False)
        path = os.path.join(_DIR, "%s_%d_%d_%d_%d_%d_%d.pdx" % (kind, n, d, nlist, seed, int(normalize), n_meso))
        pf.save(idx, path)
        _CACHE[key] = (path, x, q)
    return _CACHE[key]


def c1_index():
    """The reference's CPU configuration: 65,536 x 128 from a 64-component mixture (centres N(0, I),
    sigma 0.25), 256 IVF lists, 100 queries from the mixture plus N(0, 0.05^2) noise."""
    key = "c1"
    if key not in _CACHE:
        rng = np.random.default_rng(65536)
        c = rng.standard_normal((64, 128)).astype(np.float32)
        x = (c[rng.integers(0, 64, 65536)] + 0.25 * rng.standard_normal((65536, 128))).astype(np.float32)
        q = (c[rng.integers(0, 64, 100)] + 0.25 * rng.standard_normal((100, 128))
             + 0.05 * rng.standard_normal((100, 128))).astype(np.float32)
        idx = builder.build_index(x, "pdx_f32", num_clusters=256, seed=42, kmeans_iters=10, device=False)
        path = os.path.join(_DIR, "c1.pdx")
        pf.save(idx, path)
        _CACHE[key] = (path, x, q)
    return _CACHE[key]
