"""Hand-derived search fixtures that pin the schedule boundaries of the PDX search.

Each case is a tiny index whose vectors are zero except in a few chosen dimensions, searched with
a pre-rotated all-zero query, so every partial distance, threshold, survivor set and algorithmic
byte count can be worked out by hand from the reference's search (include/pdx/searcher.hpp):

* Warmup (searcher.hpp:376-435): tuples_needed = ceil(0.8f * n) in float; the loop fetches the
  cumulative DIMENSIONS_FETCHING_SIZES checkpoints 16, 32, 64, 96, 128, 160, 224, ... of the
  vertical dims (common.hpp:47-49) and exits once count(dist >= tau) >= tuples_needed
  (searcher.hpp:95-103) or the vertical dims are exhausted.
* thresholds (adsampling.hpp:91-98, searcher.hpp:75-91): tau = top * ratio[cur] (ratio 1 at
  cur == d); for U8, tau = (uint32)(tau_f * scale^2), truncating.
* Prune (searcher.hpp:438-540): survivors are dist < tau (strict), per 64-dim horizontal block,
  then the vertical remainder; Merge (searcher.hpp:542-566) inserts on strict less.
* two-level probe (index.hpp:545-566): the meso search visits half of the meso clusters only when
  nprobe < num_clusters / 2 (integer halving).
* filtered search (searcher.hpp:376-409,726-856): Warmup<true> returns before the first fetch when
  passing / n < 1 - 0.8f in float.
* tombstones (searcher.hpp:169-179): masked to max/2, counted as pruned by Warmup, pushed by Start
  when k exceeds the live vectors.

The ratios used below follow adsampling.hpp:147-157 at D = 128:
ratio[16] = (16/128) * (1 + 1.5/4)^2 = 0.236328125 and ratio[64] = 0.5 * (1 + 1.5/8)^2 =
0.705078125 (both exact in float); ratio[32] = 0.25 * (1 + 1.5/sqrt(32))^2 = 0.40016...,
ratio[80] = 0.625 * (1 + 1.5/sqrt(80))^2 = 0.85220...; at D = 1024 ratio[16] = 0.029541015625.

Algorithmic bytes (the per-query count both the oracle and the GPU report) are
4 * nc * d for a flat probe, s * n_used * d + 4 * m for a Start that pushes m candidates,
s * n_used * w for a Warmup that stops after w vertical dims, s * 64 * |P| per horizontal block and
s * |P| * width per vertical remainder step of Prune, and 4 * |P| for a Merge; s = 1 (U8) or 4 (F32).

Every expected value below is written out by hand from those rules, not computed by the oracle.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from pdx_amd import pdxfile as pf

U8_MAX_HALF = 0x7FFFFFFF                     # std::numeric_limits<uint32_t>::max() / 2
F32_MAX_HALF = np.float32(3.4028234663852886e38) / np.float32(2)


@dataclass
class Case:
    name: str
    index: pf.IndexData
    queries: np.ndarray           # pre-rotated (searched with transformed = True)
    k: int
    nprobe: int
    ids: list                     # expected result ids, ascending distance
    dists: list                   # expected float distances (as float32 values)
    int_dists: list | None        # expected exact integer distances (U8 data levels)
    nbytes: int | None            # expected algorithmic bytes of the query
    selection: np.ndarray | None = None   # per-slot selection (filtered cases), cluster-offset order
    note: str = ""
    mutations: list = field(default_factory=list)  # reference rules this case pins


def vec(d: int, vals: dict, u8: bool) -> np.ndarray:
    v = np.zeros(d, np.uint8 if u8 else np.float32)
    for j, x in vals.items():
        v[j] = x
    return v


def level(d: int, u8: bool, clusters, centroids, scale: float = 1.0) -> pf.Level:
    """clusters: list of (rows, ids, tombstoned slots)."""
    h, v = pf.dimension_split(d)
    cent = None if centroids is None else np.ascontiguousarray(centroids, np.float32)
    L = pf.Level(d=d, v=v, h=h, u8=u8, normalized=False, clusters=[], centroids=cent, base=0.0, scale=scale)
    for rows, ids, tomb in clusters:
        rows = np.asarray(rows, np.uint8 if u8 else np.float32).reshape(len(ids), d)
        c = pf.make_cluster(rows, np.asarray(ids, np.uint32), L)
        c.tombstones = set(int(t) for t in tomb)
        c.n_live = len(ids) - len(c.tombstones)
        L.clusters.append(c)
    return L


def flat(d: int, u8: bool, clusters, scale: float = 1.0) -> pf.IndexData:
    """Flat IVF; centroid c = c * e_0, so an all-zero query probes the clusters in order."""
    cent = np.zeros((len(clusters), d), np.float32)
    cent[:, 0] = np.arange(len(clusters), dtype=np.float32)
    L = level(d, u8, clusters, cent, scale)
    return pf.IndexData(type=pf.TYPE_U8 if u8 else pf.TYPE_F32, matrix=np.eye(d, dtype=np.float32), l1=L)


def zeros(nq: int, d: int) -> np.ndarray:
    return np.zeros((nq, d), np.float32)


def _s(u8):
    return 1 if u8 else 4


# ---------------------------------------------------------------------------------------------
# A. Warmup exit: count(dist >= tau) against ceil(0.8 n), with distances equal to tau
# ---------------------------------------------------------------------------------------------
def case_warmup_tie(u8: bool) -> Case:
    """d = 128 (v = 64, h = 64), k = 1. Cluster 0: one vector at distance 256 (the heap top).
    Cluster 1: five slots; slots 0-3 reach exactly tau(16) = 256 * ratio[16] in dims [0,16) and
    nothing else, slot 4 is 0 there and 1600 in the horizontal block.

    U8: tau(16) = (uint32)(256 * 0.236328125 = 60.5) = 60 and slots 0-3 sum to 49+9+1+1 = 60.
    F32: tau(16) = 60.5 and slots 0-3 sum to 49 + 9 + 2.25 + 0.25 = 60.5.
    Warmup: need = ceil(0.8f * 5) = 4; after the first fetch count(dist >= tau) = 4 >= 4, exit at
    cur = 16. Prune: P = {4}; its block makes 1600 >= tau(80) = 256 * 0.8522 ~ 218: pruned. The
    four slots at exactly tau are pruned although their full distance (60 / 60.5) beats the top:
    with `>` in the count, or with need = ceil(0.81 n) = 5, or tau rounded to 61, the warmup would
    go on to cur = 64, they would survive, and the result would be slot 0 (id 200).
    bytes = probe 4*2*128 + Start s*128 + 4 + Warmup s*5*16 + Prune s*64*1."""
    d = 128
    a = vec(d, {0: 16}, u8)
    hit = vec(d, {0: 7, 1: 3, 2: 1, 3: 1} if u8 else {0: 7, 1: 3, 2: 1.5, 3: 0.5}, u8)
    far = vec(d, {64: 40}, u8)
    idx = flat(d, u8, [([a], [100], []), ([hit, hit, hit, hit, far], [200, 201, 202, 203, 204], [])])
    s = _s(u8)
    nbytes = 4 * 2 * 128 + s * 128 + 4 + s * 5 * 16 + s * 64
    return Case("warmup_tie_" + ("u8" if u8 else "f32"), idx, zeros(1, d), 1, 2, [100], [256.0],
                [256] if u8 else None, nbytes,
                mutations=["warmup count dist >= tau (not >)", "selectivity 0.8", "U8 tau truncation"])


def case_warmup_ceil(u8: bool) -> Case:
    """As case A with seven slots in cluster 1: slots 0-4 at exactly tau(16), slots 5-6 zero in the
    vertical dims and 1600 in the horizontal block. need = ceil(0.8f * 7 = 5.6) = 6 > 5, so the
    warmup goes on: tau(32) = 256 * 0.40016 ~ 102 (U8 102, F32 102.44), count 0; tau(64) =
    256 * 0.705078125 = 180.5 (U8 180), count 0; cur = v = 64 ends it (s * 7 * 64 bytes).
    Prune at tau(64): every slot survives; the horizontal block (s * 64 * 7 bytes) brings slots
    5-6 to 1600; at cur = d tau = 256, slots 0-4 survive; Merge (4 * 5 bytes) keeps the first of
    the equal distances in slot order, id 200. A floor instead of ceil (need 5) would exit at
    cur = 16 and return id 100."""
    d = 128
    a = vec(d, {0: 16}, u8)
    hit = vec(d, {0: 7, 1: 3, 2: 1, 3: 1} if u8 else {0: 7, 1: 3, 2: 1.5, 3: 0.5}, u8)
    far = vec(d, {64: 40}, u8)
    idx = flat(d, u8, [([a], [100], []), ([hit] * 5 + [far] * 2, list(range(200, 207)), [])])
    s = _s(u8)
    nbytes = 4 * 2 * 128 + s * 128 + 4 + s * 7 * 64 + s * 64 * 7 + 4 * 5
    return Case("warmup_ceil_" + ("u8" if u8 else "f32"), idx, zeros(1, d), 1, 2, [200],
                [60.0 if u8 else 60.5], [60] if u8 else None, nbytes,
                mutations=["tuples needed = ceil(0.8 n)"])


# ---------------------------------------------------------------------------------------------
# B. Fetch schedule: one index per checkpoint, the warmup must stop exactly there
# ---------------------------------------------------------------------------------------------
CHECKPOINTS_1024 = [16, 32, 64, 96, 128, 160, 224, 256]   # cumulative 16,16,32,32,32,32,64,64 capped at v


def case_fetch_checkpoint(u8: bool, t: int) -> Case:
    """d = 1024 (v = 256, h = 768), k = 1. Cluster 0: one vector at distance 256. Cluster 1: five
    slots whose only nonzero dim is the last dim of checkpoint t (cum[t] - 1), 255 (65025).
    Before that dim is fetched every partial distance is 0 < tau (tau(16) = 256 * 0.0295 = 7.56,
    U8 7, and larger later); the fetch that reaches cum[t] makes 65025 >= tau for all five, so the
    warmup stops with cur = cum[t] and Prune has no survivor.
    bytes = probe 4*2*1024 + Start s*1024 + 4 + Warmup s*5*cum[t]."""
    d = 1024
    cum = CHECKPOINTS_1024[t]
    a = vec(d, {0: 16}, u8)
    x = vec(d, {cum - 1: 255}, u8)
    idx = flat(d, u8, [([a], [100], []), ([x] * 5, [200, 201, 202, 203, 204], [])])
    s = _s(u8)
    nbytes = 4 * 2 * 1024 + s * 1024 + 4 + s * 5 * cum
    return Case(f"fetch_checkpoint_{cum}_" + ("u8" if u8 else "f32"), idx, zeros(1, d), 1, 2, [100], [256.0],
                [256] if u8 else None, nbytes, mutations=["DIMENSIONS_FETCHING_SIZES"])


# ---------------------------------------------------------------------------------------------
# C. Two-level probe: meso halving at nprobe = num_clusters / 2 and num_clusters / 2 - 1
# ---------------------------------------------------------------------------------------------
def tree_index(u8: bool) -> pf.IndexData:
    """d = 128. Four data clusters of one vector each at distances 16, 9, 1, 4 (ids 10..13).
    Meso level (F32): meso 0 holds the L1 centroids of clusters 0 and 1 at distances 4 and 9,
    meso 1 those of clusters 2 and 3 at distances 1 and 25; meso centroids at distances 0 and 1."""
    d = 128
    data = [([vec(d, {0: r}, u8)], [10 + c], []) for c, r in enumerate([4, 3, 1, 2])]
    L1 = level(d, u8, data, None)
    meso = [([vec(d, {0: 2}, False), vec(d, {0: 3}, False)], [0, 1], []),
            ([vec(d, {0: 1}, False), vec(d, {0: 5}, False)], [2, 3], [])]
    cent0 = np.zeros((2, d), np.float32)
    cent0[1, 0] = 1.0
    L0 = level(d, False, meso, cent0)
    return pf.IndexData(type=pf.TYPE_TREE_U8 if u8 else pf.TYPE_TREE_F32, matrix=np.eye(d, dtype=np.float32),
                        l1=L1, l0=L0)


def case_meso_halving(u8: bool, nprobe: int) -> Case:
    """nprobe = 2 = 4 / 2: no halving (2 < 2 is false), both meso clusters are searched with k = 2.
      meso probe 4*2*128 = 1024; meso 0 Start (n = 2, push 2): 4*2*128 + 8 = 1032, heap {4, 9};
      meso 1 Warmup (need = ceil(1.6) = 2): dists 1 and 25 never give two pruned, so it fetches
      [0,16) [16,32) [32,64): 4*2*64 = 512; Prune at tau(64) = 9 * 0.705 = 6.35 keeps the slot at 1,
      one horizontal block 4*64 = 256, at cur = d tau = 9, kept; Merge 4: order = [2, 0].
      Meso bytes 2828. Data level, k = 1: cluster 2 Start 128*s + 4 (distance 1); cluster 0 Warmup
      (need 1): first fetch 16*s, 16 >= tau(16) (U8 (uint32)0.236 = 0, F32 0.236): exit, no
      survivor. Result id 12 at 1.
    nprobe = 1 < 2: halved, meso 0 only, k = 1: probe 1024 + Start 1024 + 4; data cluster 0
      Start 128*s + 4: id 10 at 16.
    nprobe = 3: no halving, k = 3: meso 0 Start pushes 2 (1032), meso 1 Start pushes 1 (1028):
      order [2, 0, 1]; data: cluster 2 Start, clusters 0 and 1 stop after 16 dims each.
    With `<=` in the halving rule nprobe = 2 would search meso 0 only and return id 10."""
    s = _s(u8)
    if nprobe == 2:
        nbytes = 2828 + s * 128 + 4 + s * 16
        ids, dists, ints = [12], [1.0], [1]
    elif nprobe == 1:
        nbytes = 1024 + 1024 + 4 + s * 128 + 4
        ids, dists, ints = [10], [16.0], [16]
    elif nprobe == 3:
        nbytes = 1024 + 1032 + 1028 + s * 128 + 4 + s * 16 + s * 16
        ids, dists, ints = [12], [1.0], [1]
    else:
        raise ValueError(nprobe)
    return Case(f"meso_halving_np{nprobe}_" + ("u8" if u8 else "f32"), tree_index(u8), zeros(1, 128), 1, nprobe,
                ids, dists, ints if u8 else None, nbytes, mutations=["meso halving nprobe < nlist / 2"])


# ---------------------------------------------------------------------------------------------
# D. U8 threshold at cur == d: (uint32)(float(I) / scale^2 * scale^2) = I - 1
# ---------------------------------------------------------------------------------------------
def case_final_truncation() -> Case:
    """U8, d = 128, scale 5 (scale^2 = 25, inverse 0.04f). Cluster 0: one vector at integer
    distance 20 (16 + 4), heap top 20 * 0.04f. Cluster 1: one vector at 19 (9 + 9 + 1), all in the
    horizontal block. Warmup (need 1) stays at 0 < tau for 16, 32, 64 dims (tau 4, 8, 14), stops
    at v (64 bytes). Prune: survives tau(64) with 0, the block makes 19, and at cur = d
    tau = (uint32)(0.8000000119f * 25) = (uint32)19.9999... = 19: 19 < 19 fails, so the closer
    vector is pruned and the result stays id 100 (int 20). An exact threshold (20) would return it.
    bytes = probe 1024 + Start 128 + 4 + Warmup 64 + Prune 64."""
    d = 128
    a = vec(d, {0: 4, 1: 2}, True)
    b = vec(d, {64: 3, 65: 3, 66: 1}, True)
    idx = flat(d, True, [([a], [100], []), ([b], [200], [])], scale=5.0)
    top = float(np.float32(20) * (np.float32(1) / np.float32(25)))
    return Case("final_threshold_truncation_u8", idx, zeros(1, d), 1, 2, [100], [top], [20],
                1024 + 128 + 4 + 64 + 64, mutations=["U8 tau truncation at cur == d"])


# ---------------------------------------------------------------------------------------------
# E. Filtered search at exactly 20 % and at 10 % per-cluster selectivity
# ---------------------------------------------------------------------------------------------
def case_filtered_selectivity(u8: bool, n1: int) -> Case:
    """d = 128, k = 1, one passing slot in cluster 1 out of n1 (n1 = 5: 20 %, n1 = 10: 10 %).
    Cluster 0: one passing vector at 256 (FilteredStart). Cluster 1 slot 0 sums to exactly tau(16)
    (60 U8 / 60.5 F32) in dims [0,16); the other slots do not pass (masked to max/2).
    20 %: 0.2f < 1 - 0.8f = 0.19999999f is false, so Warmup<true> runs: after the first fetch the
    four masked slots and slot 0 (>= tau) give 5 >= 4, exit at cur = 16, Prune has no survivor:
    result id 100. 10 %: 0.1f < 0.19999999f, Warmup<true> returns before any fetch; Prune starts
    at tau(0) = 256, slot 0 survives the horizontal block (0 < tau(64)) and the vertical remainder
    (60 < 256 at cur = d): result id 200."""
    d = 128
    a = vec(d, {0: 16}, u8)
    hit = vec(d, {0: 7, 1: 3, 2: 1, 3: 1} if u8 else {0: 7, 1: 3, 2: 1.5, 3: 0.5}, u8)
    z = vec(d, {}, u8)
    idx = flat(d, u8, [([a], [100], []), ([hit] + [z] * (n1 - 1), list(range(200, 200 + n1)), [])])
    offs = idx.l1.cluster_offsets()
    sel = np.zeros(int(offs[-1]), np.uint8)
    sel[int(offs[0])] = 1
    sel[int(offs[1])] = 1
    if n1 == 5:
        ids, dists, ints = [100], [256.0], [256]
    else:
        ids, dists, ints = [200], [60.0 if u8 else 60.5], [60]
    return Case(f"filtered_{100 // n1}pct_" + ("u8" if u8 else "f32"), idx, zeros(1, d), 1, 2, ids, dists,
                ints if u8 else None, None, selection=sel, mutations=["filtered warmup skip below 20 %"])


# ---------------------------------------------------------------------------------------------
# F/G. Tombstones: counted as pruned in Warmup; pushed by Start when k exceeds the live vectors
# ---------------------------------------------------------------------------------------------
def case_tombstones(u8: bool, k: int) -> Case:
    """d = 128. Cluster 0: one vector at 256. Cluster 1: five used slots, slots 1-4 tombstoned
    (their data is zero), slot 0 live at exactly tau(16) in dims [0,16).
    k = 1: Warmup masks slots 1-4 to max/2; after the first fetch (s*5*16 bytes: tombstoned slots
    are read) count = 4 masked + slot 0 (>= tau) = 5 >= 4, exit; Prune keeps nothing: id 100.
    Without masking in the warmup the count would be 1, the warmup would go on and slot 0 (60)
    would win.
    k = 3: cluster 1 is a Start (heap 1 < 3): m = min(3 - 1, 5) = 2 smallest of
    [60, max/2 x 4]: slot 0 and, among the equal masked slots, the lowest id (201) under this
    build's (distance, id) order -- the reference's partial_sort leaves that tie open. Result:
    (60, 200), (256, 100), (max/2, 201) with max/2 = float(2^31 - 1) * inv_scale^2 (U8) or
    FLT_MAX / 2 (F32). bytes = probe 1024 + Start s*128 + 4 + Start s*5*128 + 8."""
    d = 128
    a = vec(d, {0: 16}, u8)
    hit = vec(d, {0: 7, 1: 3, 2: 1, 3: 1} if u8 else {0: 7, 1: 3, 2: 1.5, 3: 0.5}, u8)
    z = vec(d, {}, u8)
    idx = flat(d, u8, [([a], [100], []), ([hit, z, z, z, z], [200, 201, 202, 203, 204], [1, 2, 3, 4])])
    s = _s(u8)
    hv = 60.0 if u8 else 60.5
    mx = float(np.float32(U8_MAX_HALF)) if u8 else float(F32_MAX_HALF)
    if k == 1:
        return Case("tombstones_warmup_" + ("u8" if u8 else "f32"), idx, zeros(1, d), 1, 2, [100], [256.0],
                    [256] if u8 else None, 1024 + s * 128 + 4 + s * 5 * 16,
                    mutations=["tombstones masked in Warmup"])
    return Case("tombstones_k_over_live_" + ("u8" if u8 else "f32"), idx, zeros(1, d), 3, 2, [200, 100, 201],
                [hv, 256.0, mx], [60, 256, U8_MAX_HALF] if u8 else None, 1024 + s * 128 + 4 + s * 5 * 128 + 8,
                mutations=["tombstones masked in Start, k > live"])


def all_cases() -> list:
    out = []
    for u8 in (True, False):
        out.append(case_warmup_tie(u8))
        out.append(case_warmup_ceil(u8))
        for t in range(len(CHECKPOINTS_1024)):
            out.append(case_fetch_checkpoint(u8, t))
        for npb in (1, 2, 3):
            out.append(case_meso_halving(u8, npb))
        out.append(case_filtered_selectivity(u8, 5))
        out.append(case_filtered_selectivity(u8, 10))
        out.append(case_tombstones(u8, 1))
        out.append(case_tombstones(u8, 3))
    out.append(case_final_truncation())
    return out
