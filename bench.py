#!/usr/bin/env python3
"""Benchmark of the PDX search path on one MI355X.

Default workload (--config c3): two-level IVF SQ8 (IndexPDXIVFTreeSQ8). N = 5,000,000, d = 1,024,
4,096-component mixture (centres ~ N(0, I), per-coordinate sigma 0.2), normalised as the Python
API does by default, 17,889 lists / 133 meso clusters, k = 20, batches of 4,096 queries with
per-query nprobe drawn uniformly from [8, 128]; results are checked bit-exact against the oracle
in the run.

--config c5: dense low-precision first pass (BF16 and FP8 e4m3 GEMM, FP32 accumulation) with a 4x
oversampled exact FP32 rerank over a copy of an F32 IVF index. N = 2,000,000 unit-norm vectors of
d = 3,072 (2,048-component mixture, noise norm 0.1), batches of 2,048 queries, k = 10, a seeded 10%
row filter and 1% of the rows appended and deleted between the timed batches; the error of the
returned distances against FP64 is reported.

--config c2: IVF float32 with ADSampling. N = 1,000,000 unit-norm vectors of d = 768 from a
1,024-component Gaussian mixture (centres ~ N(0, I) normalised, noise norm 0.15, rows
renormalised), 1,000 IVF lists, 10,000 queries from the mixture, k = 10, nprobe = 32, L2.

A step is one batched search of --batch queries through the device-buffer C-ABI
(pdxg_search_device) with queries, index and outputs resident in HBM. The JSON line also carries
the plain per-query GPU path (one query per call), batch-64 latency percentiles, recall@k against
an FP64 brute-force ground truth, the continuous-batching server (saturated QPS, p50/p99 under
Poisson load), the roofline of the dominant scan kernel and the CPU baseline (the oracle, i.e.
the reference search semantics restated in C++, on the host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 (vector, with packed math)

CONFIGS = {
    "c2": dict(n=1_000_000, d=768, centers=1024, sigma=0.15, unit=True, kind="pdx_f32", nlist=1000, n_meso=0,
               normalize=False, nq=10_000, batch=1024, k=10, nprobe=32, nprobe_range=None, gt=1000),
    "c3": dict(n=5_000_000, d=1024, centers=4096, sigma=0.2, unit=False, kind="pdx_tree_u8", nlist=0, n_meso=0,
               normalize=True, nq=8192, batch=4096, k=20, nprobe=0, nprobe_range=(8, 128), gt=512, f32_twin=True,
               scratch_gib=3),
    # serving over a resident two-level SQ8 index; each query draws k and nprobe from the sets
    "c4": dict(n=10_000_000, d=1536, centers=8192, sigma=0.2, unit=False, kind="pdx_tree_u8", nlist=0, n_meso=0,
               normalize=True, nq=8192, batch=4096, k=100, k_choices=(10, 20, 100), nprobe=0, nprobe_range=None,
               nprobe_choices=(16, 64, 256), gt=256, loads=(0.1, 0.25, 0.5, 0.75, 0.9, 1.0), scratch_gib=24),
    # dense low-precision first pass + FP32 rerank with a row filter and interleaved updates
    "c5": dict(n=2_000_000, d=3072, centers=2048, sigma=0.1, unit=True, kind="pdx_f32", nlist=0, n_meso=0,
               normalize=False, nq=8192, batch=2048, k=10, nprobe=0, nprobe_range=None, gt=512, dense=True,
               oversample=4, filter=0.10, update_frac=0.01, precisions=("bf16", "fp8", "fp8_direct")),
}
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp8": 5000.0, "fp8_direct": 5000.0}  # MI355X dense (no sparsity)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def mixture(n, d, n_centers, sigma, seed, n_queries, unit=True, per_coord=False):
    """Gaussian mixture rows and queries.

    unit=True: centres normalised, rows renormalised, and sigma is the noise scale relative to the
    unit centres (per-coordinate std sigma / sqrt(d), noise norm ~sigma) unless per_coord.
    unit=False: centres ~ N(0, I) and per-coordinate std sigma."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n_centers, d), dtype=np.float32)
    if unit:
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        if not per_coord:
            sigma = sigma / np.sqrt(d)
    x = np.empty((n, d), np.float32)
    chunk = 1 << 17
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        blk = c[rng.integers(0, n_centers, m)] + np.float32(sigma) * rng.standard_normal((m, d), dtype=np.float32)
        if unit:
            blk /= np.linalg.norm(blk, axis=1, keepdims=True)
        x[s:s + m] = blk
        if (s // chunk) % 16 == 15:
            log(f"data: {s + m} of {n} rows")
    q = c[rng.integers(0, n_centers, n_queries)] + np.float32(sigma) * rng.standard_normal((n_queries, d), dtype=np.float32)
    if unit:
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, q.astype(np.float32)


def ground_truth(x, q, k, normalize, torch, dev, chunk=None, mask=None):
    """Exact top-k ids by L2 in float64 on the GPU (rows and queries normalised if the index is);
    rows with mask False are excluded."""
    chunk = chunk or min(1 << 20, (6 << 30) // (8 * x.shape[1]))  # about 6 GiB of float64 rows per chunk
    qt = torch.from_numpy(q).to(dev, torch.float64)
    if normalize:
        qt = qt / torch.linalg.vector_norm(qt, dim=1, keepdim=True)
    qn = (qt * qt).sum(1)
    best_d = torch.full((q.shape[0], k), float("inf"), dtype=torch.float64, device=dev)
    best_i = torch.zeros((q.shape[0], k), dtype=torch.int64, device=dev)
    for s in range(0, x.shape[0], chunk):
        xt = torch.from_numpy(x[s:s + chunk]).to(dev, torch.float64)
        if normalize:
            xt = xt / torch.linalg.vector_norm(xt, dim=1, keepdim=True)
        dd = (xt * xt).sum(1)[None, :] - 2.0 * (qt @ xt.T) + qn[:, None]
        if mask is not None:
            mt = torch.from_numpy(~mask[s:s + chunk]).to(dev)
            dd[:, mt] = float("inf")
        vd, vi = torch.topk(dd, min(k, dd.shape[1]), dim=1, largest=False)
        cd = torch.cat([best_d, vd], 1)
        ci = torch.cat([best_i, vi + s], 1)
        o = torch.argsort(cd, dim=1)[:, :k]
        best_d, best_i = torch.gather(cd, 1, o), torch.gather(ci, 1, o)
    out = best_i.cpu().numpy()
    torch.cuda.empty_cache()
    return out


def recall(ids, counts, gt, k):
    return float(np.mean([len(set(ids[i, :counts[i]].tolist()) & set(gt[i, :k].tolist())) / k for i in range(len(gt))]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    for key in ("n", "d", "centers", "nlist", "nq", "batch", "k", "nprobe"):
        ap.add_argument("--" + key, type=int, default=None)
    ap.add_argument("--sigma", type=float, default=None)
    ap.add_argument("--metric", default="l2sq", choices=["l2sq", "ip"])
    ap.add_argument("--plain-queries", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU work per baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-slots", default="1,2:256,2,4",
                    help="batches in flight (server slots); a list runs the serving leg once per value")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--precisions", default="", help="dense configurations: comma-separated subset of the precisions")
    ap.add_argument("--sigma-per-coord", action="store_true", help="unit-norm mixtures: sigma per coordinate")
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--parity-queries", type=int, default=-1,
                    help="queries checked against the oracle (-1: one whole timed batch)")
    ap.add_argument("--no-f32-twin", action="store_true",
                    help="skip the F32 two-level leg over the same clustering (the recall-gated number)")
    ap.add_argument("--index-cache", default="",
                    help="index file: loaded when present, else written after the build (e.g. to profile without "
                         "the build's kernels)")
    ap.add_argument("--slots", type=int, default=0,
                    help="batches in flight in the timed leg, each on its own stream and workspace (0: the best of "
                         "1-4 on a short trial)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="also run a plain device copy of this size (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()
    cfg = dict(CONFIGS[args.config])
    for key in ("n", "d", "centers", "nlist", "nq", "batch", "k", "nprobe", "sigma"):
        if getattr(args, key) is not None:
            cfg[key] = getattr(args, key)
    if args.metric == "ip":
        cfg["normalize"] = True

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (each GPU would hold a "
              "replica of the index and serve its own queries)", file=sys.stderr)
        sys.exit(2)

    import torch

    from pdx_amd import _capi, builder

    dev = torch.device("cuda", 0)
    log("device:", _capi.device_info(0))
    if cfg.get("dense"):
        return bench_dense(args, cfg, torch, _capi, builder, dev)
    t0 = time.time()
    x, q = mixture(cfg["n"], cfg["d"], cfg["centers"], cfg["sigma"], args.seed, cfg["nq"], cfg["unit"],
                   args.sigma_per_coord)
    log(f"data {x.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    from pdx_amd import pdxfile
    if args.index_cache and os.path.exists(args.index_cache):
        idx = pdxfile.load(args.index_cache)
        log(f"index loaded from {args.index_cache}")
    else:
        idx = builder.build_index(x, cfg["kind"], num_clusters=cfg["nlist"], num_meso_clusters=cfg["n_meso"],
                                  normalize=cfg["normalize"], distance_metric=args.metric, seed=42, device=0)
        if args.index_cache:
            pdxfile.save(idx, args.index_cache)
    build_s = time.time() - t0
    sizes = np.array([c.n_live for c in idx.l1.clusters])
    nlist = idx.l1.nc
    log(f"index {cfg['kind']} built in {build_s:.1f}s; {nlist} lists"
        + (f", {idx.l0.nc} meso" if idx.l0 is not None else "")
        + f"; sizes min/mean/max {sizes.min()}/{sizes.mean():.0f}/{sizes.max()}")
    g = _capi.GpuIndex(idx)
    log(f"index resident: {g.device_bytes / 2**30:.2f} GiB")

    gtq = min(cfg["gt"], cfg["nq"])
    t0 = time.time()
    gt = ground_truth(x, q[:gtq], cfg["k"], cfg["normalize"], torch, dev)
    log(f"fp64 ground truth for {gtq} queries in {time.time() - t0:.1f}s")

    if args.copy_calib_mib:
        a = torch.empty(args.copy_calib_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        bcopy = torch.empty_like(a)
        bcopy.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calib_mib} MiB device-to-device copy")
        del a, bcopy

    copy_gbs = device_copy_gbs(torch, dev)
    log(f"1 GiB device-to-device copy: {copy_gbs:.0f} GB/s (read + write)")

    B, K, NQ = cfg["batch"], cfg["k"], cfg["nq"]
    rng = np.random.default_rng(args.seed + 1)
    if cfg.get("nprobe_choices"):
        nprobe_q = rng.choice(np.array(cfg["nprobe_choices"], np.uint32), NQ)
    elif cfg["nprobe_range"]:
        lo, hi = cfg["nprobe_range"]
        nprobe_q = rng.integers(lo, hi + 1, NQ).astype(np.uint32)
    else:
        nprobe_q = np.full(NQ, cfg["nprobe"], np.uint32)
    # per-query k (k is the largest when the configuration draws k per query)
    k_q = rng.choice(np.array(cfg["k_choices"], np.uint32), NQ) if cfg.get("k_choices") else np.full(NQ, K, np.uint32)
    per_k = bool(cfg.get("k_choices"))
    kd = torch.from_numpy(k_q.astype(np.int32)).to(dev)
    qd = torch.from_numpy(q).to(dev)
    out_ids = torch.zeros((B, K), dtype=torch.int32, device=dev)
    out_d = torch.zeros((B, K), dtype=torch.float32, device=dev)
    out_i = torch.zeros((B, K), dtype=torch.int32, device=dev)
    out_c = torch.zeros(B, dtype=torch.int32, device=dev)
    sp = torch.cuda.current_stream().cuda_stream
    state = {}
    tree = idx.l0 is not None

    def workspace(nmax, np_max):
        # two-level searches size the meso-level result by the largest nprobe; flat ones take any
        return _capi.Workspace(g, nmax, K, np_max if tree else 0)

    def run(s, m, w=None):
        w = w or state["ws"]
        npd = state["np_dev"]
        w.search(qd[s:s + m].data_ptr(), m, kd[s:s + m].data_ptr() if per_k else 0, npd[s:s + m].data_ptr(), K,
                 state["np_max"], K, out_ids.data_ptr(), out_d.data_ptr(), out_i.data_ptr(), out_c.data_ptr(), 0,
                 stream=sp, k_all=K)

    def set_nprobe(arr):
        state["np_dev"] = torch.from_numpy(arr.astype(np.int32)).to(dev)
        state["np_max"] = int(arr.max())
        state["ws"] = workspace(B, state["np_max"])
        state["ws"].enable_timing(True)

    # operating point: with a fixed nprobe, double it until recall@k on the ground-truth subset
    # reaches the target (the metric is QPS at recall@10 >= 0.95)
    def recall_now():
        ids = np.zeros((gtq, K), np.uint32)
        cnt = np.zeros(gtq, np.uint32)
        idist = np.zeros((gtq, K), np.uint32)
        for s in range(0, gtq, B):
            m = min(B, gtq - s)
            run(s, m)
            torch.cuda.synchronize()
            ids[s:s + m] = out_ids[:m].cpu().numpy().view(np.uint32)
            cnt[s:s + m] = out_c[:m].cpu().numpy()
            idist[s:s + m] = out_i[:m].cpu().numpy().view(np.uint32)
        state["gt_idist"] = idist
        k10 = min(10, K)
        state["recall10"] = recall(ids[:, :k10], np.minimum(cnt, k10), gt, k10)
        state["gt_ids"], state["gt_cnt"] = ids, cnt
        if per_k:  # recall@k with each query's own k
            return float(np.mean([len(set(ids[i, :cnt[i]].tolist()) & set(gt[i, :k_q[i]].tolist())) / k_q[i]
                                  for i in range(gtq)]))
        return recall(ids, cnt, gt, K)

    sweep = {}
    while True:
        set_nprobe(nprobe_q)
        if cfg.get("nprobe_choices"):
            key = "{" + ",".join(str(x) for x in cfg["nprobe_choices"]) + "}"
        elif cfg["nprobe_range"]:
            key = "U[%d,%d]" % cfg["nprobe_range"]
        else:
            key = str(int(nprobe_q[0]))
        sweep[key] = recall_now()
        log(f"recall@{'k' if per_k else K} = {sweep[key]:.4f}, recall@10 = {state['recall10']:.4f} at nprobe {key}")
        if cfg["nprobe_range"] or cfg.get("nprobe_choices") or sweep[key] >= args.target_recall or nprobe_q[0] >= nlist:
            break
        nprobe_q[:] = min(nlist, int(nprobe_q[0]) * 2)
    rec = sweep[key]
    ceiling = None
    if cfg["nprobe_range"] or cfg.get("nprobe_choices"):
        # recall with every query probing 4x the configuration's largest nprobe: when it equals the
        # recall above, the probe is not what limits recall (SQ8 rounding of close neighbours is)
        saved = nprobe_q.copy()
        big = min(nlist, 4 * int(saved.max()))
        set_nprobe(np.full(NQ, big, np.uint32))
        r_big = recall_now()
        ceiling = {"nprobe": big, ("recall_at_k_per_query" if per_k else f"recall_at_{K}"): round(r_big, 4),
                   "recall_at_10": round(state["recall10"], 4)}
        log(f"recall at nprobe {big} for every query: {r_big:.4f} (recall@10 {state['recall10']:.4f})")
        nprobe_q = saved
        set_nprobe(nprobe_q)
        recall_now()  # restores recall10 of the configuration's own nprobe

    # autotune the scan workgroup size per level on one batch ("waves per cluster")
    def tune(w, m):
        npd = state["np_dev"]
        return w.autotune(qd[:m].data_ptr(), m, kd[:m].data_ptr() if per_k else 0, npd[:m].data_ptr(), K,
                          state["np_max"], K, out_ids.data_ptr(), out_d.data_ptr(), out_i.data_ptr(),
                          out_c.data_ptr(), stream=sp, k_all=K)

    plan = tune(state["ws"], B)
    log(f"autotuned scan block sizes (meso, data): {plan}"
        + (f"; two-phase meso scan: {state['ws'].meso_summary}" if tree else ""))

    parity = None
    n_par = B if args.parity_queries < 0 else min(args.parity_queries, B)
    if n_par:
        parity = check_parity(idx, q, nprobe_q, k_q, n_par, run, out_ids, out_d, out_i, out_c, torch)

    # one batch in flight at a time: stage and kernel times (HIP events on the search stream)
    nb = max(1, NQ // B)
    for i in range(args.warmup):
        run((i % nb) * B, B)
    torch.cuda.synchronize()
    stage = np.zeros(4)
    sbytes = np.zeros(2)
    meso = np.zeros(2)
    pairs = 0
    t0 = time.perf_counter()
    ws = state["ws"]
    for i in range(args.steps):
        run((i % nb) * B, B)
        stage += ws.stage_times()
        sbytes += ws.stage_bytes()
        if tree:
            meso += ws.meso_times()
            pairs += ws.meso_pairs()
    torch.cuda.synchronize()
    solo_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stage /= args.steps
    sbytes /= args.steps
    meso /= args.steps
    pairs /= args.steps
    log(f"one batch in flight: {solo_ms:.3f} ms/batch, {B / solo_ms * 1e3:.0f} QPS; stage ms "
        f"preprocess/probe/meso/scan = {stage.round(4)}; scan bytes meso/data = {sbytes[0]:.3e}/{sbytes[1]:.3e}"
        + (f"; meso phase1/replay ms = {meso.round(4)}, pairs {pairs:.4g}" if tree else ""))

    # timed leg: `steps` batches with S in flight, each slot a stream and a workspace of its own
    # (the tuned plan), so that one batch's HBM-bound data scan runs beside another's VALU-bound
    # meso phase; inputs resident in HBM, the device synchronised on both sides
    def slot_search(w, s0, o, stream):
        w.search(qd[s0:s0 + B].data_ptr(), B, kd[s0:s0 + B].data_ptr() if per_k else 0,
                 state["np_dev"][s0:s0 + B].data_ptr(), K, state["np_max"], K, o[0].data_ptr(), o[1].data_ptr(),
                 o[2].data_ptr(), o[3].data_ptr(), 0, stream=stream, k_all=K)

    ov = Overlap(torch, dev, ws, (out_ids, out_d, out_i, out_c), lambda: workspace(B, state["np_max"]),
                 slot_search, nb, B, tree)
    n_slots, trial, elapsed, overlap_equal = ov.timed(args.slots, args.steps)
    ms_per_step = elapsed * 1e3 / args.steps
    qps = B * args.steps / elapsed
    log(f"batched, {n_slots} in flight: {ms_per_step:.3f} ms/batch, {qps:.0f} QPS (trial by slots: {trial}); "
        f"overlapped results equal to one batch at a time: {overlap_equal}")
    # Kernels timed with HIP events on the search stream: the data-level scan (HBM-bound:
    # algorithmic bytes = what the reference's schedule reads), and for two-level indexes the
    # meso stage, split into the two-phase scan's summary phase (FP32 vector-ALU bound:
    # 3 FLOP per (query, member, dim) plus a 15-add tree per 64-dim block) and the replay.
    u8 = cfg["kind"].endswith("u8")
    scan_ms = float(stage[3])
    scan_gbs = sbytes[1] / (scan_ms * 1e-3) / 1e9
    scan_roof = {"bound": "hbm", "kernel": ("k_scan<U8>" if u8 else "k_scan<F32>") + " (data-level pruned PDX scan)",
                 "achieved": round(scan_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(scan_gbs / HBM_PEAK_GBS, 4), "ms": round(scan_ms, 4),
                 "algorithmic_bytes_per_launch": int(sbytes[1])}
    roofs = {"data_scan": scan_roof}
    if tree and pairs > 0:
        flops = pairs * (3 * cfg["d"] + 15 * (idx.l1.h // 64))
        tf = flops / (meso[0] * 1e-3) / 1e12
        roofs["meso_phase1"] = {"bound": "valu",
                                "kernel": "k_meso_summary + k_meso_tails + their visit-list kernels (two-phase meso scan, phase 1)",
                                "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": round(tf / FP32_PEAK_TFLOPS, 4), "ms": round(float(meso[0]), 4),
                                "pairs_per_launch": int(pairs), "flops_per_launch": int(flops)}
        roofs["meso_replay"] = {"kernel": "k_scan<F32, replay> (two-phase meso scan, phase 2)",
                                "ms": round(float(meso[1]), 4),
                                "reference_algorithmic_bytes": int(sbytes[0])}
    elif tree:
        mgbs = sbytes[0] / (float(stage[2]) * 1e-3) / 1e9
        roofs["meso_scan"] = {"bound": "hbm", "kernel": "k_scan<F32> (meso-level pruned PDX scan)",
                              "achieved": round(mgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(mgbs / HBM_PEAK_GBS, 4), "ms": round(float(stage[2]), 4),
                              "algorithmic_bytes_per_launch": int(sbytes[0])}
    dom_key = max((k for k in roofs if "bound" in roofs[k]), key=lambda k: roofs[k]["ms"])
    kdesc = ("{" + ",".join(str(x) for x in cfg["k_choices"]) + "}") if per_k else str(K)
    workload = (f"{cfg['kind']}, N={cfg['n']}, d={cfg['d']}, {nlist} lists, {args.metric}, k={kdesc}, nprobe={key}, "
                f"batch={B}")
    # HBM bytes per launch of each kernel from the committed FETCH_SIZE / WRITE_SIZE profiles
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f).get(workload, {})
    except (OSError, ValueError):
        tr = {}
    for kname, r in roofs.items():
        r["traffic"] = (tr.get(kname) or {}).get("traffic_bytes_per_launch")

    # batch-64 latency and the plain per-query path (one query per call)
    ws64 = workspace(64, state["np_max"])
    plan64 = tune(ws64, 64)
    lat64 = []
    for i in range(min(64, NQ // 64)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(i * 64, 64, w=ws64)
        torch.cuda.synchronize()
        lat64.append(time.perf_counter() - t)
    ws1 = workspace(1, state["np_max"])
    plan1 = tune(ws1, 1)
    log(f"autotuned plans: batch 64 {plan64}, one query {plan1}")
    lat1 = []
    for i in range(args.plain_queries + 5):
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(i, 1, w=ws1)
        torch.cuda.synchronize()
        if i >= 5:
            lat1.append(time.perf_counter() - t)
    lat1 = np.array(lat1) * 1e3
    lat64 = np.array(lat64[2:]) * 1e3
    plain_qps = 1e3 / lat1.mean()

    # the same searches captured in HIP graphs (pdxg_graph_*): fixed device buffers that each
    # query (batch) is copied into, then one graph launch
    def graph_latency(w, m, reps):
        qb = torch.empty((m, cfg["d"]), dtype=torch.float32, device=dev)
        kb = kd[:m].clone()
        nb_ = state["np_dev"][:m].clone()
        gr = w.capture(qb.data_ptr(), m, kb.data_ptr() if per_k else 0, nb_.data_ptr(), K, state["np_max"], K,
                       out_ids.data_ptr(), out_d.data_ptr(), out_i.data_ptr(), out_c.data_ptr(), k_all=K)
        lat = []
        for i in range(reps + 5):
            s0 = (i * m) % (NQ - m + 1)
            torch.cuda.synchronize()
            t = time.perf_counter()
            qb.copy_(qd[s0:s0 + m])
            kb.copy_(kd[s0:s0 + m])
            nb_.copy_(state["np_dev"][s0:s0 + m])
            gr.launch(sp)
            torch.cuda.synchronize()
            if i >= 5:
                lat.append(time.perf_counter() - t)
        nodes = gr.nodes
        gr.close()
        return np.array(lat) * 1e3, nodes

    # service time of one batch by size (what a serving loop's batch costs)
    svc = {}
    for m in (1, 16, 64, 256, 1024, B):
        w = ws1 if m == 1 else (ws64 if m <= 64 else state["ws"])
        tt = []
        for i in range(8):
            torch.cuda.synchronize()
            t = time.perf_counter()
            run((i * m) % (NQ - m + 1), m, w=w)
            torch.cuda.synchronize()
            tt.append(time.perf_counter() - t)
        svc[str(m)] = round(float(np.median(tt[2:])) * 1e3, 4)
    log(f"batch service time ms by size: {svc}")

    lat1g, nodes1 = graph_latency(ws1, 1, args.plain_queries)
    lat64g, _ = graph_latency(ws64, 64, 60)
    # a whole batch through the persistent work queue (order 2), captured in one graph
    wsb = state["ws"]
    order_before = wsb.order
    wsb.set_order(2)
    latbg, nodesb = graph_latency(wsb, B, 10)
    wsb.set_order(order_before)
    log(f"graph-captured batch of {B} with the persistent work queue: {np.median(latbg):.3f} ms "
        f"({B / np.median(latbg) * 1e3:.0f} QPS, {nodesb} nodes)")
    log(f"graph-captured: one query p50 {np.percentile(lat1g, 50):.3f} ms p99 {np.percentile(lat1g, 99):.3f} ms "
        f"({nodes1} nodes); batch-64 p50 {np.percentile(lat64g, 50):.3f} ms")
    log(f"plain per-query: {plain_qps:.0f} QPS p50 {np.percentile(lat1, 50):.3f} ms p99 {np.percentile(lat1, 99):.3f} ms; "
        f"batch-64 p50 {np.percentile(lat64, 50):.3f} ms")

    serving = None
    if not args.no_serve:
        # one serving leg per slot count; the headline is the one with the lowest p99 at 50% load
        specs = str(args.serve_slots).split(",")
        legs = [serve_leg(_capi, g, q, nprobe_q, k_q, cfg, args, plan, sp_) for sp_ in specs]
        serving = min(legs, key=lambda l: l["load"]["50%"]["p99_ms"])
        serving["by_slots"] = {sp_: {"saturated_qps": l["saturated_qps"], "p99_ms_at_50%": l["load"]["50%"]["p99_ms"]}
                               for sp_, l in zip(specs, legs)}
    cpu = None if args.no_cpu else cpu_baseline(idx, q, nprobe_q, k_q, args)
    decomposition = None
    if u8 and not per_k:
        decomposition = sq8_exhaustive(idx, g, q[:gtq], gt, state["gt_idist"], state["gt_cnt"], K, torch, dev)
    twin = None
    if cfg.get("f32_twin") and not args.no_f32_twin:
        twin = f32_twin_leg(args, cfg, idx, x, q, gt, nprobe_q, torch, dev, _capi, builder, sp)
    gate_met = bool(state["recall10"] >= args.target_recall)

    line = {
        "metric": ("QPS at recall@10>=%.2f (batched GPU search; plain per-query GPU path and p50/p99 reported beside "
                   "it)" % args.target_recall) if gate_met else
                  ("QPS of the batched GPU search at the configuration's own nprobe; its recall@10 is %.4f, below the "
                   "%.2f gate (SQ8 rounding caps it on this data, see recall_decomposition); the gated figure is "
                   "f32_twin (the F32 two-level path over the same clustering)" % (state["recall10"], args.target_recall)),
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 (int32 accumulation)" if u8 else "f32",
        "data": "synthetic (seeded Gaussian mixture; index built in-process)",
        "config": {"workload": workload, "global_batch": B, "parallelism": "single GPU",
                   "batches_in_flight": n_slots,
                   "scan_block": {"batch": list(plan), "batch64": list(plan64), "one_query": list(plan1)},
                   "meso_two_phase": {"batch": ws.meso_summary, "batch64": ws64.meso_summary,
                                      "one_query": ws1.meso_summary} if tree else None,
                   "scan_order": {"batch": ws.order, "batch64": ws64.order, "one_query": ws1.order}},
        ("recall_at_k_per_query" if per_k else f"recall_at_{K}"): round(rec, 4),
        "recall_at_10": round(state["recall10"], 4),
        "recall_gate_met": gate_met,
        "recall_sweep": sweep,
        "recall_decomposition": decomposition,
        "f32_twin": twin,
        "recall_at_4x_max_nprobe": ceiling,
        "parity_vs_oracle": parity,
        "batches_in_flight": n_slots,
        "batches_in_flight_trial_qps": trial,
        "one_batch_in_flight_qps": round(B / solo_ms * 1e3, 1),
        "overlapped_equal_to_one_at_a_time": overlap_equal,
        "plain_per_query_qps": round(plain_qps, 1),
        "batched_over_plain": round(qps / plain_qps, 2),
        "latency_ms": {"plain_p50": round(float(np.percentile(lat1, 50)), 4),
                       "plain_p99": round(float(np.percentile(lat1, 99)), 4),
                       "batch64_p50": round(float(np.percentile(lat64, 50)), 4),
                       "batch64_p99": round(float(np.percentile(lat64, 99)), 4),
                       "graph_one_query_p50": round(float(np.percentile(lat1g, 50)), 4),
                       "graph_one_query_p99": round(float(np.percentile(lat1g, 99)), 4),
                       "graph_batch64_p50": round(float(np.percentile(lat64g, 50)), 4),
                       f"graph_batch{B}_persistent_queue_median": round(float(np.median(latbg)), 4),
                       f"batch{B}_mean_one_in_flight": round(solo_ms, 4)},
        "batch_service_ms": svc,
        "stage_ms": {"preprocess": round(float(stage[0]), 4), "probe": round(float(stage[1]), 4),
                     "meso_scan": round(float(stage[2]), 4), "scan": round(float(stage[3]), 4)},
        "roofline": dict(roofs[dom_key], stage=dom_key),
        "kernels": roofs,
        "serving": serving,
        "cpu_baseline": cpu,
        "device_copy_gbs": {"value": round(copy_gbs, 1), "what": "1 GiB device-to-device copy (torch), read + write "
                                                              "bytes per second on this GPU: the rate a plain "
                                                              "streaming kernel reaches, for scale against the "
                                                              "scans' HBM traffic rates"},
        "build_s": round(build_s, 1),
    }
    print(json.dumps(line), flush=True)


def device_copy_gbs(torch, dev, mib=1024, reps=5):
    """Read + write GB/s of a plain device-to-device copy (HIP events around `reps` copies)."""
    a = torch.empty(mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * (mib << 20) / (ms * 1e-3) / 1e9


def bench_dense(args, cfg, torch, _capi, builder, dev):
    """Dense first pass (BF16 / FP8 e4m3 GEMM, FP32 accumulation) + k*oversample FP32 rerank over
    a copy of an F32 index, with a seeded row filter and append/delete of update_frac of the rows
    interleaved with the timed batches. One JSON line; value = queries / (search + update time) of
    the fastest precision at recall@10 >= the target."""
    from pdx_amd import pdxfile

    n, d, B, K, NQ = cfg["n"], cfg["d"], cfg["batch"], cfg["k"], cfg["nq"]
    OS = cfg["oversample"]
    n_up = int(round(n * cfg["update_frac"]))
    t0 = time.time()
    x_all, q = mixture(n + n_up, d, cfg["centers"], cfg["sigma"], args.seed, NQ, cfg["unit"], args.sigma_per_coord)
    x, x_new = x_all[:n], x_all[n:]
    log(f"data {x.shape} (+{n_up} rows to append) in {time.time() - t0:.1f}s")
    t0 = time.time()
    if args.index_cache and os.path.exists(args.index_cache):
        idx = pdxfile.load(args.index_cache)
        log(f"index loaded from {args.index_cache}")
    else:
        idx = builder.build_index(x, cfg["kind"], num_clusters=cfg["nlist"], normalize=cfg["normalize"],
                                  distance_metric=args.metric, seed=42, device=0)
        if args.index_cache:
            pdxfile.save(idx, args.index_cache)
    build_s = time.time() - t0
    log(f"index {cfg['kind']} with {idx.l1.nc} lists in {build_s:.1f}s")
    g = _capi.GpuIndex(idx)
    rng = np.random.default_rng(args.seed + 2)
    sel = (rng.random(n + n_up) < cfg["filter"]).astype(np.uint8)  # row filter by id, appended ids included
    sel_d = torch.from_numpy(sel).to(dev)
    gtq = min(cfg["gt"], NQ)
    t0 = time.time()
    gt = ground_truth(x, q[:gtq], K, cfg["normalize"], torch, dev, mask=sel[:n] == 1)
    log(f"fp64 filtered ground truth for {gtq} queries in {time.time() - t0:.1f}s")
    qd = torch.from_numpy(q).to(dev)
    out_ids = torch.zeros((B, K), dtype=torch.int32, device=dev)
    out_d = torch.zeros((B, K), dtype=torch.float32, device=dev)
    out_c = torch.zeros(B, dtype=torch.int32, device=dev)
    sp = torch.cuda.current_stream().cuda_stream
    xq = q.astype(np.float64)
    if cfg["normalize"]:
        xq /= np.linalg.norm(xq, axis=1, keepdims=True)

    def search(dn, s, m, filtered=True):
        dn.search(qd[s:s + m].data_ptr(), m, K, OS, K, out_ids.data_ptr(), out_d.data_ptr(), out_c.data_ptr(),
                  selection_ptr=sel_d.data_ptr() if filtered else 0, stream=sp)

    def measure(dn, rows, live_mask, label):
        """recall@k vs the FP64 ground truth of the live passing rows, and the error of the returned
        FP32 distances against FP64 distances of the same (unrotated) rows."""
        g_t = gt if label == "before" else ground_truth(rows, q[:gtq], K, cfg["normalize"], torch, dev,
                                                        mask=live_mask)
        ids = np.zeros((gtq, K), np.uint32)
        dst = np.zeros((gtq, K), np.float32)
        cnt = np.zeros(gtq, np.uint32)
        for s0 in range(0, gtq, B):
            m = min(B, gtq - s0)
            search(dn, s0, m)
            torch.cuda.synchronize()
            ids[s0:s0 + m] = out_ids[:m].cpu().numpy().view(np.uint32)
            dst[s0:s0 + m] = out_d[:m].cpu().numpy()
            cnt[s0:s0 + m] = out_c[:m].cpu().numpy()
        rec = recall(ids, cnt, g_t, K)
        errs = []
        for t in range(gtq):
            r = rows[ids[t, :cnt[t]]].astype(np.float64)
            if cfg["normalize"]:
                r /= np.linalg.norm(r, axis=1, keepdims=True)
            e = ((r - xq[t]) ** 2).sum(1)
            errs.append(np.abs(dst[t, :cnt[t]].astype(np.float64) - e) / np.maximum(e, 1e-30))
        errs = np.concatenate(errs)
        return rec, {"max_rel": float(errs.max()), "mean_rel": float(errs.mean())}

    results = {}
    for pname in (args.precisions.split(",") if args.precisions else cfg["precisions"]):
        # "fp8_direct": FP8 of the rows themselves (no residual coding), for comparison
        prec = {"bf16": _capi.PDXG_DENSE_BF16, "fp8": _capi.PDXG_DENSE_FP8,
                "fp8_direct": _capi.PDXG_DENSE_FP8 | _capi.PDXG_DENSE_NO_RESIDUAL}[pname]
        dn = _capi.Dense(g, prec, extra_rows=n_up, max_queries=B)
        log(f"{pname}: dense copy {dn.info()['device_bytes'] / 2**30:.1f} GiB")
        rec0, err0 = measure(dn, x, sel[:n] == 1, "before")
        log(f"{pname}: recall@{K} {rec0:.4f}; distance error vs fp64 max {err0['max_rel']:.2e} "
            f"mean {err0['mean_rel']:.2e}")
        # unfiltered batch (the full-size GEMM) for reference
        for _ in range(2):
            search(dn, 0, B, filtered=False)
        torch.cuda.synchronize()
        dn.enable_timing(True)
        t = time.perf_counter()
        reps = 3
        st_unf = np.zeros(4)
        for i in range(reps):
            search(dn, 0, B, filtered=False)
            st_unf += dn.stage_times()
        torch.cuda.synchronize()
        unf_ms = (time.perf_counter() - t) * 1e3 / reps
        st_unf /= reps
        unf_rows = dn.info()["gemm_rows"]
        # timed: filtered batches with append/delete interleaved (update_frac of the rows in total)
        nb = max(1, NQ // B)
        for i in range(args.warmup):
            search(dn, (i % nb) * B, B)
        torch.cuda.synchronize()
        live = np.zeros(n + n_up, bool)
        live[:n] = True
        per = [n_up // args.steps + (1 if i < n_up % args.steps else 0) for i in range(args.steps)]
        urng = np.random.default_rng(args.seed + 3)
        # the update schedule (rows appended and ids deleted at each step) is synthetic input: drawn
        # before the timed loop, so the timed updates are the library's append and delete calls
        sched, nxt = [], 0
        for i in range(args.steps):
            a = per[i]
            if not a:
                sched.append(None)
                continue
            new_ids = np.arange(n + nxt, n + nxt + a, dtype=np.uint32)
            live[new_ids] = True
            cand = np.nonzero(live[:n])[0]
            gone = urng.choice(cand, a, replace=False).astype(np.uint32)
            live[gone] = False
            sched.append((np.ascontiguousarray(x_new[nxt:nxt + a]), new_ids, gone))
            nxt += a
        t_search = t_upd = 0.0
        stage = np.zeros(4)
        gemm_rows = 0
        for i in range(args.steps):
            t = time.perf_counter()
            search(dn, (i % nb) * B, B)
            torch.cuda.synchronize()
            t_search += time.perf_counter() - t
            stage += dn.stage_times()
            gemm_rows += dn.info()["gemm_rows"]
            t = time.perf_counter()
            if sched[i] is not None:
                rows_i, new_ids, gone = sched[i]
                dn.append(rows_i, new_ids)
                dn.delete(gone)
            t_upd += time.perf_counter() - t
        stage /= args.steps
        gemm_rows /= args.steps
        qps_search = B * args.steps / t_search
        qps_total = B * args.steps / (t_search + t_upd)
        dn.enable_timing(False)
        rec1, err1 = measure(dn, x_all, live & (sel == 1), "after")
        flops = 2.0 * B * gemm_rows * d
        tf = flops / (stage[1] * 1e-3) / 1e12
        unf_tf = 2.0 * B * unf_rows * d / (st_unf[1] * 1e-3) / 1e12
        log(f"{pname}: {qps_total:.0f} QPS with updates ({qps_search:.0f} search only); stage ms "
            f"filter/first-pass/merge/rerank = {stage.round(4)}; first pass {tf:.0f} TFLOP/s over "
            f"{gemm_rows:.0f} rows; unfiltered batch {unf_ms:.2f} ms, first pass {unf_tf:.0f} TFLOP/s; "
            f"recall after updates {rec1:.4f}")
        results[pname] = {
            "qps": round(qps_total, 1), "qps_search_only": round(qps_search, 1),
            "ms_per_step": round((t_search + t_upd) * 1e3 / args.steps, 4),
            "update_ms_per_step": round(t_upd * 1e3 / args.steps, 4),
            "recall_at_10_before_updates": round(rec0, 4), "recall_at_10_after_updates": round(rec1, 4),
            "distance_error_vs_fp64": {"before": err0, "after": err1},
            "stage_ms": {"queries_and_filter": round(float(stage[0]), 4), "first_pass": round(float(stage[1]), 4),
                         "merge": round(float(stage[2]), 4), "rerank": round(float(stage[3]), 4)},
            "gemm_rows_per_batch": int(gemm_rows),
            "kernels": {
                "first_pass": {"bound": "mfma",
                               "kernel": f"k_mma_topr ({pname} MFMA GEMM, FP32 accumulation, top-R per row "
                                         f"segment in the epilogue)",
                               "achieved": round(tf, 1), "peak": MFMA_PEAK_TFLOPS[pname], "unit": "TFLOP/s",
                               "frac": round(tf / MFMA_PEAK_TFLOPS[pname], 4), "ms": round(float(stage[1]), 4),
                               "flops_per_launch": int(flops)},
            },
            "merge_ms": round(float(stage[2]), 4),
            "unfiltered_batch": {"ms": round(unf_ms, 4), "gemm_rows": int(unf_rows),
                                 "first_pass_ms": round(float(st_unf[1]), 4), "first_pass_tflops": round(unf_tf, 1),
                                 "stage_ms": [round(float(v), 4) for v in st_unf]},
        }
        dn.close()
        torch.cuda.empty_cache()

    ok = {p: r for p, r in results.items() if r["recall_at_10_before_updates"] >= args.target_recall}
    best = max(ok or results, key=lambda p: (ok or results)[p]["qps"])
    workload = (f"dense first pass over {cfg['kind']}, N={n}, d={d}, k={K}, oversample={OS}, batch={B}, "
                f"{int(cfg['filter'] * 100)}% row filter, {cfg['update_frac'] * 100:g}% append/delete")
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            tr = json.load(f).get(workload, {})
    except (OSError, ValueError):
        tr = {}
    for p, r in results.items():
        for kname, kr in r["kernels"].items():
            kr["traffic"] = (tr.get(f"{p}_{kname}") or {}).get("traffic_bytes_per_launch")
    cpu = None if args.no_cpu else cpu_baseline_dense(idx, q, sel, K, args)
    line = {
        "metric": "QPS at recall@10>=0.95 (dense low-precision first pass + exact FP32 rerank, row filter, "
                  "interleaved append/delete)",
        "value": results[best]["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[best]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{best} GEMM (fp32 accumulation) + f32 rerank",
        "data": "synthetic (seeded Gaussian mixture; index built in-process)",
        "config": {"workload": workload, "global_batch": B, "parallelism": "single GPU", "precision": best},
        "recall_at_10": results[best]["recall_at_10_before_updates"],
        "precisions": results,
        "roofline": dict(results[best]["kernels"]["first_pass"], stage="first_pass"),
        "cpu_baseline": cpu,
        "build_s": round(build_s, 1),
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_dense(idx, q, sel, K, args):
    """The oracle's filtered search over every list (the exhaustive filtered PDX search of the
    reference) on one host core, over a bounded query sample."""
    from oracle import oracle as orc

    o = orc.OracleIndex(data=idx)
    L = idx.l1
    slot = np.zeros(int(sum(c.cap for c in L.clusters)), np.uint8)
    off = 0
    for c in L.clusters:
        slot[off:off + c.n_used] = sel[c.ids[:c.n_used]]
        off += c.cap
    qr = o.preprocess(q[:200])
    t = time.perf_counter()
    n1 = 0
    while n1 < len(qr) and (n1 < 3 or time.perf_counter() - t < args.cpu_seconds):
        o.filtered_search(qr[n1], K, 0, slot, transformed=True)
        n1 += 1
    qps = n1 / (time.perf_counter() - t)
    log(f"cpu oracle ({orc.isa()}): exhaustive filtered search {qps:.2f} QPS over {n1} queries on 1 thread")
    return {"value": round(qps, 3), "unit": "queries/s", "cores": 1, "kind": "port", "isa": orc.isa(),
            "sample": f"{n1} pre-rotated queries, filtered search over every list (nprobe = all), 1 thread"}


def _overlap(a, b):
    """Size of the multiset intersection of two integer lists."""
    from collections import Counter

    return sum((Counter(a.tolist()) & Counter(b.tolist())).values())


def sq8_exhaustive(idx, g, q, gt, s_idist, s_cnt, K, torch, dev, chunk=1 << 19):
    """Separates the two things that can cap recall on an SQ8 index: an exhaustive search over every
    stored code with exact integer distances (no probe, no pruning), compared with the FP64 ground
    truth, shows what SQ8 rounding allows; the pruned GPU search's integer distances compared with
    that exhaustive top-k (as multisets, so ties do not matter) show what the probe and ADSampling
    pruning lose."""
    from pdx_amd import pdxfile as pf

    t0 = time.time()
    L = idx.l1
    codes = np.concatenate([pf.from_pdx(c.data, c.n_used, c.cap, L.d, L.v, L.h, True) for c in L.clusters])
    rows = np.concatenate([c.ids[:c.n_used] for c in L.clusters]).astype(np.int64)
    _, qu = g.preprocess(q)                     # the search's own query codes
    qt = torch.from_numpy(qu.astype(np.float64)).to(dev)
    qn = (qt * qt).sum(1)
    best_d = torch.full((len(q), K), float("inf"), dtype=torch.float64, device=dev)
    best_i = torch.zeros((len(q), K), dtype=torch.int64, device=dev)
    for s in range(0, len(codes), chunk):
        xt = torch.from_numpy(codes[s:s + chunk]).to(dev).to(torch.float64)
        dd = (xt * xt).sum(1)[None, :] - 2.0 * (qt @ xt.T) + qn[:, None]   # exact: integers below 2^53
        vd, vi = torch.topk(dd, min(K, dd.shape[1]), dim=1, largest=False)
        cd = torch.cat([best_d, vd], 1)
        ci = torch.cat([best_i, vi + s], 1)
        o = torch.argsort(cd, dim=1)[:, :K]
        best_d, best_i = torch.gather(cd, 1, o), torch.gather(ci, 1, o)
    ex_d = best_d.cpu().numpy().round().astype(np.int64)
    ex_ids = rows[best_i.cpu().numpy()]
    del codes
    torch.cuda.empty_cache()
    k10 = min(10, K)
    ex_rec = recall(ex_ids[:, :k10].astype(np.uint32), np.full(len(q), k10), gt, k10)
    same = np.mean([_overlap(s_idist[i, :min(k10, s_cnt[i])].astype(np.int64), ex_d[i, :k10]) / k10
                    for i in range(len(q))])
    same_k = np.mean([_overlap(s_idist[i, :s_cnt[i]].astype(np.int64), ex_d[i, :K]) / K for i in range(len(q))])
    log(f"SQ8 exhaustive exact-integer search: recall@10 {ex_rec:.4f} vs FP64; the pruned search returns "
        f"{same:.4f} of its top-10 integer distances ({same_k:.4f} of top-{K}) ({time.time() - t0:.1f}s)")
    return {"queries": len(q), "sq8_exhaustive_recall_at_10": round(float(ex_rec), 4),
            "pruned_search_top10_distances_in_exhaustive_sq8_top10": round(float(same), 4),
            f"pruned_search_top{K}_distances_in_exhaustive_sq8_top{K}": round(float(same_k), 4)}


class Overlap:
    """Batches in flight on streams of their own. Slot j has its own workspace (slot 0 the tuned
    one; the others made by `make` and given its plan, order and meso mode) and output buffers;
    `search(ws, start, outs, stream)` enqueues one batch of the queries from `start` on."""

    def __init__(self, torch, dev, ws0, outs0, make, search, nb, B, tree):
        self.torch, self.dev, self.make, self.search = torch, dev, make, search
        self.nb, self.B, self.tree = nb, B, tree
        self.ws, self.outs, self.st = [ws0], [tuple(outs0)], []

    def _grow(self, ns):
        while len(self.ws) < ns:
            w = self.make()
            w.set_plan(*self.ws[0].plan())
            w.set_order(self.ws[0].order)
            if self.tree:
                w.set_meso_summary(self.ws[0].meso_summary)
            self.ws.append(w)
            self.outs.append(tuple(self.torch.zeros_like(t) for t in self.outs[0]))
        while len(self.st) < ns:
            self.st.append(self.torch.cuda.Stream(device=self.dev))

    def run(self, j, i):
        self.search(self.ws[j], (i % self.nb) * self.B, self.outs[j], self.st[j].cuda_stream)

    def leg(self, ns, n):
        """Seconds for n batches with ns in flight (2 ns untimed batches first)."""
        self._grow(ns)
        for i in range(2 * ns):
            self.run(i % ns, i)
        self.torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(n):
            self.run(i % ns, i)
        self.torch.cuda.synchronize()
        return time.perf_counter() - t

    def _valid(self, outs):
        """Copies of a slot's outputs with the entries past each query's result count zeroed (the
        counts are the last buffer)."""
        torch = self.torch
        c = outs[-1].long()
        keep = torch.arange(outs[0].shape[1], device=c.device)[None, :] < c[:, None]
        return [torch.where(keep, t, torch.zeros_like(t)) for t in outs[:-1]] + [c.clone()]

    def timed(self, slots, steps):
        """(slots used, trial QPS by slot count or None, seconds for `steps` batches, whether each
        slot's last batch equals the same queries searched alone)."""
        ws0 = self.ws[0]
        ws0.enable_timing(False)  # no event records or byte counters between the overlapped batches
        trial = None
        if slots <= 0:
            trial = {ns: round(self.B * 12 / min(self.leg(ns, 12) for _ in range(2)), 1) for ns in (1, 2, 3, 4)}
            slots = max(trial, key=trial.get)
        elapsed = self.leg(slots, steps)
        equal = True
        for i in range(max(0, steps - slots), steps):
            got = self._valid(self.outs[i % slots])
            self.run(0, i)
            self.torch.cuda.synchronize()
            equal = equal and all(bool(self.torch.equal(a, b)) for a, b in zip(got, self._valid(self.outs[0])))
        ws0.enable_timing(True)
        for w in self.ws[1:]:
            w.close()
        self.ws = self.ws[:1]
        self.outs = self.outs[:1]
        return slots, trial, elapsed, equal


def f32_twin_leg(args, cfg, idx, x, q, gt, nprobe_q, torch, dev, _capi, builder, sp):
    """IndexPDXIVFTree (F32) over the clustering of the SQ8 index (same lists, meso level, order of
    members and row ids; rows stored as the normalised + rotated floats): recall@10, nprobe scaled
    up from the configuration's draws until recall@10 reaches the target, QPS of the timed batches,
    the data-level scan's HBM roofline and parity with the oracle on one whole batch."""
    from oracle.oracle import OracleIndex

    t0 = time.time()
    tw = builder.float_twin(idx, x, device=0)
    g = _capi.GpuIndex(tw)
    log(f"f32 twin built in {time.time() - t0:.1f}s, resident {g.device_bytes / 2**30:.2f} GiB")
    B, K, NQ = cfg["batch"], cfg["k"], len(q)
    gtq = len(gt)
    nlist = tw.l1.nc
    qd = torch.from_numpy(q).to(dev)
    ids = torch.zeros((B, K), dtype=torch.int32, device=dev)
    dd = torch.zeros((B, K), dtype=torch.float32, device=dev)
    cnt = torch.zeros(B, dtype=torch.int32, device=dev)
    st = {}

    def make(arr):
        st["npd"] = torch.from_numpy(arr.astype(np.int32)).to(dev)
        st["npm"] = int(arr.max())
        st["ws"] = _capi.Workspace(g, B, K, st["npm"])
        st["ws"].enable_timing(True)

    def run(s, m):
        st["ws"].search(qd[s:s + m].data_ptr(), m, 0, st["npd"][s:s + m].data_ptr(), K, st["npm"], K, ids.data_ptr(),
                        dd.data_ptr(), 0, cnt.data_ptr(), 0, stream=sp, k_all=K)

    def rec10():
        out = np.zeros((gtq, K), np.uint32)
        c = np.zeros(gtq, np.uint32)
        for s in range(0, gtq, B):
            m = min(B, gtq - s)
            run(s, m)
            torch.cuda.synchronize()
            out[s:s + m] = ids[:m].cpu().numpy().view(np.uint32)
            c[s:s + m] = cnt[:m].cpu().numpy()
        k10 = min(10, K)
        return recall(out[:, :k10], np.minimum(c, k10), gt, k10)

    arr = nprobe_q.copy()
    scale = 1
    while True:
        make(arr)
        r10 = rec10()
        log(f"f32 twin: recall@10 {r10:.4f} at nprobe x{scale}")
        if r10 >= args.target_recall or int(arr.min()) >= nlist:
            break
        arr = np.minimum(nlist, arr * 2).astype(np.uint32)
        scale *= 2
    plan = st["ws"].autotune(qd[:B].data_ptr(), B, 0, st["npd"][:B].data_ptr(), K, st["npm"], K, ids.data_ptr(),
                             dd.data_ptr(), 0, cnt.data_ptr(), stream=sp, k_all=K)
    o = OracleIndex(data=tw)
    oi, od, _, oc, _ = o.search(q[:B], K, arr[:B], threads=min(16, os.cpu_count() or 1))
    del o
    run(0, B)
    torch.cuda.synchronize()
    gi = ids.cpu().numpy().view(np.uint32)
    gd = dd.cpu().numpy()
    gc = cnt.cpu().numpy().view(np.uint32)
    par_ok = bool(np.array_equal(gc, oc)) and all(
        np.array_equal(gi[r, :oc[r]], oi[r, :oc[r]]) and np.array_equal(gd[r, :oc[r]].view(np.uint32),
                                                                         od[r, :oc[r]].view(np.uint32))
        for r in range(B))
    nb = max(1, NQ // B)
    for i in range(args.warmup):
        run((i % nb) * B, B)
    torch.cuda.synchronize()
    stage = np.zeros(4)
    sbytes = np.zeros(2)
    t = time.perf_counter()
    for i in range(args.steps):
        run((i % nb) * B, B)
        stage += st["ws"].stage_times()
        sbytes += st["ws"].stage_bytes()
    torch.cuda.synchronize()
    solo_ms = (time.perf_counter() - t) * 1e3 / args.steps
    stage /= args.steps
    sbytes /= args.steps
    gbs = sbytes[1] / (stage[3] * 1e-3) / 1e9

    def slot_search(w, s0, o, stream):
        w.search(qd[s0:s0 + B].data_ptr(), B, 0, st["npd"][s0:s0 + B].data_ptr(), K, st["npm"], K, o[0].data_ptr(),
                 o[1].data_ptr(), 0, o[2].data_ptr(), 0, stream=stream, k_all=K)

    ov = Overlap(torch, dev, st["ws"], (ids, dd, cnt), lambda: _capi.Workspace(g, B, K, st["npm"]), slot_search,
                 nb, B, True)
    n_slots, trial, el, equal = ov.timed(args.slots, args.steps)
    qps = B * args.steps / el
    log(f"f32 twin: {qps:.0f} QPS with {n_slots} batches in flight ({B / solo_ms * 1e3:.0f} one at a time) at "
        f"recall@10 {r10:.4f} (nprobe x{scale}); stage ms {stage.round(4)}; data scan {gbs:.0f} GB/s; parity vs "
        f"oracle on {B} queries: {par_ok}; overlapped equal to one at a time: {equal}")
    out = {"path": "IndexPDXIVFTree (F32) over the SQ8 index's clustering", "qps": round(qps, 1),
           "ms_per_step": round(el * 1e3 / args.steps, 4), "batches_in_flight": n_slots,
           "batches_in_flight_trial_qps": trial, "one_batch_in_flight_qps": round(B / solo_ms * 1e3, 1),
           "overlapped_equal_to_one_at_a_time": equal, "recall_at_10": round(float(r10), 4),
           "nprobe": "U[%d,%d] x %d (capped at %d lists)" % (cfg["nprobe_range"] + (scale, nlist))
           if cfg.get("nprobe_range") else str(int(arr[0])),
           "recall_gate_met": bool(r10 >= args.target_recall), "scan_block": list(plan),
           "meso_two_phase": st["ws"].meso_summary,
           "stage_ms": {"preprocess": round(float(stage[0]), 4), "probe": round(float(stage[1]), 4),
                        "meso_scan": round(float(stage[2]), 4), "scan": round(float(stage[3]), 4)},
           "data_scan": {"bound": "hbm", "kernel": "k_scan<F32> (data-level pruned PDX scan)",
                         "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": int(sbytes[1])},
           "parity_vs_oracle": {"queries": B, "ids_and_float_distances_equal": par_ok},
           "resident_gib": round(g.device_bytes / 2**30, 2)}
    del g, tw
    torch.cuda.empty_cache()
    return out


def check_parity(idx, q, nprobe_q, k_q, n, run, out_ids, out_d, out_i, out_c, torch):
    """GPU results of the first n queries vs the oracle: ids, int32 and float distances bit-exact."""
    from oracle.oracle import OracleIndex

    t0 = time.time()
    o = OracleIndex(data=idx)
    oi, od, oid, oc, _ = o.search(q[:n], k_q[:n], nprobe_q[:n], threads=min(16, os.cpu_count() or 1))
    del o
    run(0, n)
    torch.cuda.synchronize()
    gi = out_ids[:n].cpu().numpy().view(np.uint32)
    gd = out_d[:n].cpu().numpy()
    gid = out_i[:n].cpu().numpy().view(np.uint32)
    gc = out_c[:n].cpu().numpy().view(np.uint32)
    u8 = idx.is_u8
    ok_c = bool(np.array_equal(gc, oc))
    ok_i = ok_c and all(np.array_equal(gi[r, :oc[r]], oi[r, :oc[r]]) for r in range(n))
    ok_d = (ok_c and all(np.array_equal(gid[r, :oc[r]], oid[r, :oc[r]]) for r in range(n))) if u8 else None
    ok_f = ok_c and all(np.array_equal(gd[r, :oc[r]].view(np.uint32), od[r, :oc[r]].view(np.uint32)) for r in range(n))
    log(f"parity vs oracle on {n} queries: ids {ok_i} int32 distances {ok_d} float distances {ok_f} "
        f"({time.time() - t0:.1f}s)")
    return {"queries": n, "ids_equal": ok_i, "int32_distances_equal": ok_d, "float_distances_equal": ok_f}


def serve_leg(_capi, g, q, nprobe_q, k_q, cfg, args, plan, spec):
    """Continuous batching: saturated QPS, then Poisson arrivals at fractions of it (p50/p99).
    spec "S" or "S:P": S slots; with P, a second batch starts beside a running one only when at
    least P requests wait."""
    K = cfg["k"]
    slots, _, pmin = str(spec).partition(":")
    slots, pmin = int(slots), int(pmin or 0)
    budget = int(cfg.get("scratch_gib", 0) * 2**30)  # memory-based admission: slots fit the budget
    sv = _capi.Server(g, max_batch=cfg["batch"], max_k=K, max_nprobe=int(nprobe_q.max()), max_slots=slots,
                      scratch_budget_bytes=budget, scan_block=plan, pipeline_min=pmin)
    rng = np.random.default_rng(7)
    # scan plans per batch size (the batches under load are far smaller than the saturated ones)
    ns = min(len(q), cfg["batch"])
    ladder = sv.autotune(q[:ns], k_q[:ns], nprobe_q[:ns])
    n_sat = min(len(q), 8 * cfg["batch"])
    for _ in range(2):  # second pass: warm
        lat, el = sv.replay(q[:n_sat], k_q[:n_sat], nprobe_q[:n_sat], np.zeros(n_sat))
    sat = n_sat / el
    st0 = sv.stats()
    out = {"saturated_qps": round(sat, 1), "slots": st0["slots"], "pipeline_min": pmin,
           "max_batch": cfg["batch"], "scratch_bytes": st0["scratch_bytes"], "scratch_budget_bytes": budget,
           "summary_capacity_bytes_per_slot": st0["summary_capacity_bytes"],
           "saturated_mean_batch": round(st0["completed"] / max(1, st0["batches"]), 1),
           "footprint_limited_batches": st0["footprint_limited_batches"], "load": {}}
    for frac in cfg.get("loads", (0.1, 0.5, 0.9)):
        rate = frac * sat
        n = int(min(60000, max(2000, rate * 0.4)))
        sel = rng.integers(0, len(q), n)
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        s0 = sv.stats()
        lat, el = sv.replay(q[sel], k_q[sel], nprobe_q[sel], arr)
        s1 = sv.stats()
        nb = max(1, s1["batches"] - s0["batches"])
        per = {f: round((s1[f] - s0[f]) / nb, 4) for f in ("gpu_ms", "host_ms", "turnaround_ms")}
        per["admitted_summary_mib"] = round((s1["admitted_summary_bytes"] - s0["admitted_summary_bytes"]) / nb / 2**20, 1)
        per["footprint_limited"] = s1["footprint_limited_batches"] - s0["footprint_limited_batches"]
        out["load"][f"{int(frac * 100)}%"] = {"offered_qps": round(rate, 1), "served_qps": round(n / el, 1),
                                              "p50_ms": round(float(np.percentile(lat, 50)), 4),
                                              "p99_ms": round(float(np.percentile(lat, 99)), 4),
                                              "mean_batch": round(n / nb, 1), "per_batch_ms": per}
        log(f"serving ({spec} slots) at {int(frac * 100)}% of {sat:.0f} QPS: p50 {np.percentile(lat, 50):.3f} ms "
            f"p99 {np.percentile(lat, 99):.3f} ms; {n / nb:.0f} queries/batch, per batch ms {per}")
    st = sv.stats()
    out["mean_batch"] = round(st["completed"] / max(1, st["batches"]), 1)
    out["plans"] = [list(p) for p in ladder]
    sv.close()
    return out


def cpu_baseline(idx, q, nprobe_q, K, args):
    """K: k for every query, or one k per query."""
    """The oracle (reference semantics, C++) on the host cores, over a bounded query sample."""
    from oracle import oracle as orc

    o = orc.OracleIndex(data=idx)
    nq = min(len(q), 5000)
    qr = o.preprocess(q[:nq])
    npb = nprobe_q[:nq]
    kk = np.broadcast_to(np.asarray(K, np.uint32), (len(q),))[:nq]
    t = time.perf_counter()
    o.search(qr[:20], kk[:20], npb[:20], transformed=True, threads=1)
    per_q = (time.perf_counter() - t) / 20
    n1 = int(max(20, min(nq, args.cpu_seconds / max(per_q, 1e-6))))
    t = time.perf_counter()
    o.search(qr[:n1], kk[:n1], npb[:n1], transformed=True, threads=1)
    qps1 = n1 / (time.perf_counter() - t)
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))
    t = time.perf_counter()
    o.search(qr, kk, npb, transformed=True, threads=cores)
    qpsn = nq / (time.perf_counter() - t)
    log(f"cpu oracle ({orc.isa()}): 1 thread {qps1:.0f} QPS over {n1} queries; {cores} threads {qpsn:.0f} QPS over {nq}")
    return {"value": round(qpsn, 1), "unit": "queries/s", "cores": cores, "kind": "port",
            "single_thread_qps": round(qps1, 1), "isa": orc.isa(),
            "sample": f"{nq} pre-rotated queries on {cores} threads ({n1} on 1 thread), same index and nprobe"}


if __name__ == "__main__":
    main()
