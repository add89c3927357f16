#!/usr/bin/env python3
"""Benchmark of the PDX search path on one MI355X.

Workload (default): IVF float32 with ADSampling. N = 1,000,000 unit-norm vectors of d = 768 from a
1,024-component Gaussian mixture (centres ~ N(0, I) normalised, sigma = 0.15, rows renormalised),
1,000 IVF lists, 10,000 queries from the mixture, k = 10, nprobe = 32, L2 metric.

A step is one batched search of --batch queries (default 1,024) through the device-buffer C-ABI
(pdxg_search_device) with queries, index and outputs resident in HBM. The line also carries the
plain per-query GPU path (one query per call), batch-64 latency percentiles, recall@10 against an
FP64 brute-force ground truth, the roofline of the dominant kernel (the data-level scan) and the CPU
baseline (the oracle, i.e. the reference search semantics restated in C++, on the host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def mixture(n, d, n_centers, sigma, seed, n_queries, per_coord=False):
    """Unit-norm Gaussian mixture rows and queries (centres normalised, rows renormalised).

    sigma is the noise scale relative to the unit-norm centres: per-coordinate std sigma / sqrt(d),
    so the noise vector has norm ~sigma. per_coord=True applies sigma per coordinate instead (noise
    norm sigma * sqrt(d), which swamps unit-norm centres at d = 768)."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n_centers, d), dtype=np.float32)
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    sigma = sigma if per_coord else sigma / np.sqrt(d)
    x = np.empty((n, d), np.float32)
    chunk = 1 << 17
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        lab = rng.integers(0, n_centers, m)
        blk = c[lab] + np.float32(sigma) * rng.standard_normal((m, d), dtype=np.float32)
        blk /= np.linalg.norm(blk, axis=1, keepdims=True)
        x[s:s + m] = blk
    lab = rng.integers(0, n_centers, n_queries)
    q = c[lab] + np.float32(sigma) * rng.standard_normal((n_queries, d), dtype=np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, q.astype(np.float32)


def ground_truth(x, q, k, torch, dev):
    """Exact top-k ids by L2 in float64 on the GPU."""
    xt = torch.from_numpy(x).to(dev, torch.float64)
    xn = (xt * xt).sum(1)
    out = []
    for s in range(0, q.shape[0], 256):
        qt = torch.from_numpy(q[s:s + 256]).to(dev, torch.float64)
        dd = xn[None, :] - 2.0 * (qt @ xt.T) + (qt * qt).sum(1)[:, None]
        out.append(torch.topk(dd, k, dim=1, largest=False).indices.cpu().numpy())
    del xt
    torch.cuda.empty_cache()
    return np.concatenate(out)


def recall(ids, counts, gt, k):
    r = [len(set(ids[i, :counts[i]].tolist()) & set(gt[i, :k].tolist())) / k for i in range(len(gt))]
    return float(np.mean(r))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--d", type=int, default=768)
    ap.add_argument("--centers", type=int, default=1024)
    ap.add_argument("--sigma", type=float, default=0.15)
    ap.add_argument("--nlist", type=int, default=1000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--nprobe", type=int, default=32)
    ap.add_argument("--metric", default="l2sq", choices=["l2sq", "ip"])
    ap.add_argument("--gt-queries", type=int, default=1000)
    ap.add_argument("--plain-queries", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU work per baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-batch", type=int, default=1024)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--sigma-per-coord", action="store_true", help="apply sigma per coordinate (see mixture())")
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="also run a plain device copy of this size (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the index would be "
              "replicated per GPU with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from pdx_amd import _capi, builder

    dev = torch.device("cuda", 0)
    log("device:", _capi.device_info(0))
    t0 = time.time()
    x, q = mixture(args.n, args.d, args.centers, args.sigma, args.seed, args.nq, args.sigma_per_coord)
    log(f"data {x.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    idx = builder.build_index(x, "pdx_f32", num_clusters=args.nlist, normalize=args.metric == "ip",
                              distance_metric=args.metric, seed=42, device=0)
    build_s = time.time() - t0
    sizes = np.array([c.n_live for c in idx.l1.clusters])
    log(f"index built in {build_s:.1f}s; cluster sizes min/mean/max {sizes.min()}/{sizes.mean():.0f}/{sizes.max()}")
    g = _capi.GpuIndex(idx)
    log(f"index resident: {g.device_bytes / 2**30:.2f} GiB")

    gtq = min(args.gt_queries, args.nq)
    t0 = time.time()
    gt = ground_truth(x, q[:gtq], args.k, torch, dev)
    log(f"fp64 ground truth for {gtq} queries in {time.time() - t0:.1f}s")

    if args.copy_calib_mib:
        a = torch.empty(args.copy_calib_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        bcopy = torch.empty_like(a)
        bcopy.copy_(a)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calib_mib} MiB device-to-device copy")
        del a, bcopy

    B, K = args.batch, args.k
    qd = torch.from_numpy(q).to(dev)
    out_ids = torch.zeros((B, K), dtype=torch.int32, device=dev)
    out_d = torch.zeros((B, K), dtype=torch.float32, device=dev)
    out_c = torch.zeros(B, dtype=torch.int32, device=dev)
    out_b = torch.zeros(B, dtype=torch.int64, device=dev)
    ws = _capi.Workspace(g, B, K, 0)
    ws.enable_timing(True)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    nbatches = args.nq // B

    def run(qs, nq, want_bytes=False, w=ws):
        w.search(qs.data_ptr(), nq, 0, 0, K, args.nprobe, K, out_ids.data_ptr(), out_d.data_ptr(), 0,
                 out_c.data_ptr(), out_b.data_ptr() if want_bytes else 0, stream=sp, k_all=K,
                 nprobe_all=args.nprobe)

    # operating point: the configured nprobe, doubled until recall@k on the ground-truth subset
    # reaches the target (the metric is QPS at recall@10 >= 0.95)
    def recall_at(npb):
        ids = np.zeros((gtq, K), np.uint32)
        cnt = np.zeros(gtq, np.uint32)
        for s in range(0, gtq, B):
            m = min(B, gtq - s)
            ws.search(qd[s:s + m].data_ptr(), m, 0, 0, K, npb, K, out_ids.data_ptr(), out_d.data_ptr(), 0,
                      out_c.data_ptr(), 0, stream=sp, k_all=K, nprobe_all=npb)
            torch.cuda.synchronize()
            ids[s:s + m] = out_ids[:m].cpu().numpy().view(np.uint32)
            cnt[s:s + m] = out_c[:m].cpu().numpy()
        return recall(ids, cnt, gt, K)

    sweep = {}
    npb = args.nprobe
    while True:
        sweep[npb] = recall_at(npb)
        log(f"recall@{K} = {sweep[npb]:.4f} at nprobe {npb}")
        if sweep[npb] >= args.target_recall or npb >= args.nlist:
            break
        npb = min(args.nlist, npb * 2)
    args.nprobe = npb
    rec = sweep[npb]

    # algorithmic bytes of the data-level scan per batch (probe bytes subtracted)
    scan_bytes = []
    for b in range(min(nbatches, 4)):
        run(qd[b * B:(b + 1) * B], B, want_bytes=True)
        torch.cuda.synchronize()
        scan_bytes.append(int(out_b.sum().item()) - B * args.nlist * args.d * 4)

    # timed leg: K steps of one batch each, inputs resident in HBM
    for i in range(args.warmup):
        b = i % nbatches
        run(qd[b * B:(b + 1) * B], B)
    torch.cuda.synchronize()
    stage = np.zeros(4)
    t0 = time.perf_counter()
    for i in range(args.steps):
        b = i % nbatches
        run(qd[b * B:(b + 1) * B], B)
        stage += ws.stage_times()  # waits for this batch: one batch in flight at a time
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    qps = B * args.steps / elapsed
    stage /= args.steps
    log(f"batched: {ms_per_step:.3f} ms/step, {qps:.0f} QPS; stage ms preprocess/probe/-/scan = {stage.round(4)}")

    bytes_per_launch = float(np.mean(scan_bytes))
    scan_ms = float(stage[3])
    achieved = bytes_per_launch / (scan_ms * 1e-3) / 1e9
    workload = (f"IVF F32 ADSampling, N={args.n}, d={args.d}, {args.nlist} lists, {args.metric}, k={K}, "
                f"nprobe={args.nprobe}, batch={B}")
    traffic = None  # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE profile of this workload
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            traffic = json.load(f).get(workload, {}).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        pass

    # batch-64 latency and plain per-query path (one query per call)
    ws64 = _capi.Workspace(g, 64, K, args.nprobe)
    lat64 = []
    for i in range(min(64, args.nq // 64)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(qd[i * 64:(i + 1) * 64], 64, w=ws64)
        torch.cuda.synchronize()
        lat64.append(time.perf_counter() - t)
    ws1 = _capi.Workspace(g, 1, K, args.nprobe)
    lat1 = []
    for i in range(args.plain_queries + 5):
        torch.cuda.synchronize()
        t = time.perf_counter()
        run(qd[i:i + 1], 1, w=ws1)
        torch.cuda.synchronize()
        if i >= 5:
            lat1.append(time.perf_counter() - t)
    lat1 = np.array(lat1) * 1e3
    lat64 = np.array(lat64[2:]) * 1e3
    plain_qps = 1e3 / lat1.mean()
    log(f"plain per-query: {plain_qps:.0f} QPS p50 {np.percentile(lat1, 50):.3f} ms p99 {np.percentile(lat1, 99):.3f} ms; "
        f"batch-64 p50 {np.percentile(lat64, 50):.3f} ms")

    serving = None
    if not args.no_serve:
        serving = serve_leg(_capi, g, q, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(idx, q, args)

    line = {
        "metric": "QPS at recall@10>=0.95 (batched GPU search; plain per-query GPU path and p50/p99 reported beside it)",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded unit-norm Gaussian mixture, noise norm sigma; index built in-process)",
        "config": {"workload": workload,
                   "global_batch": B, "parallelism": "single GPU"},
        "recall_at_10": round(rec, 4),
        "nprobe_used": args.nprobe,
        "recall_sweep": {str(k): round(v, 4) for k, v in sweep.items()},
        "plain_per_query_qps": round(plain_qps, 1),
        "batched_over_plain": round(qps / plain_qps, 2),
        "latency_ms": {"plain_p50": round(float(np.percentile(lat1, 50)), 4),
                       "plain_p99": round(float(np.percentile(lat1, 99)), 4),
                       "batch64_p50": round(float(np.percentile(lat64, 50)), 4),
                       "batch64_p99": round(float(np.percentile(lat64, 99)), 4),
                       "batch1024_mean": round(ms_per_step, 4)},
        "stage_ms": {"preprocess": round(float(stage[0]), 4), "probe": round(float(stage[1]), 4),
                     "scan": round(scan_ms, 4)},
        "roofline": {"bound": "hbm", "kernel": "k_scan<F32> (data-level pruned PDX scan)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "serving": serving,
        "cpu_baseline": cpu,
        "build_s": round(build_s, 1),
    }
    print(json.dumps(line), flush=True)


def serve_leg(_capi, g, q, args):
    """Continuous batching: saturated QPS, then Poisson arrivals at fractions of it (p50/p99)."""
    sv = _capi.Server(g, max_batch=args.serve_batch, max_k=args.k, max_nprobe=args.nprobe,
                      max_slots=args.serve_slots)
    rng = np.random.default_rng(7)
    n_sat = min(len(q), 8 * args.serve_batch)
    lat, el = sv.replay(q[:n_sat], args.k, args.nprobe, np.zeros(n_sat))
    sv.replay(q[:n_sat], args.k, args.nprobe, np.zeros(n_sat))  # second pass: warm
    lat, el = sv.replay(q[:n_sat], args.k, args.nprobe, np.zeros(n_sat))
    sat = n_sat / el
    out = {"saturated_qps": round(sat, 1), "slots": sv.stats()["slots"], "max_batch": args.serve_batch,
           "scratch_bytes": sv.stats()["scratch_bytes"], "load": {}}
    for frac in (0.1, 0.5, 0.9):
        rate = frac * sat
        n = int(min(60000, max(2000, rate * 0.4)))
        idx = rng.integers(0, len(q), n)
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        lat, el = sv.replay(q[idx], args.k, args.nprobe, arr)
        out["load"][f"{int(frac * 100)}%"] = {"offered_qps": round(rate, 1), "served_qps": round(n / el, 1),
                                              "p50_ms": round(float(np.percentile(lat, 50)), 4),
                                              "p99_ms": round(float(np.percentile(lat, 99)), 4)}
        log(f"serving at {int(frac * 100)}% of {sat:.0f} QPS: p50 {np.percentile(lat, 50):.3f} ms "
            f"p99 {np.percentile(lat, 99):.3f} ms")
    st = sv.stats()
    out["mean_batch"] = round(st["completed"] / max(1, st["batches"]), 1)
    sv.close()
    return out


def cpu_baseline(idx, q, args):
    """The oracle (reference semantics, C++) on the host cores, over a bounded query sample."""
    from oracle import oracle as orc
    from pdx_amd import pdxfile as pf

    path = os.path.join(tempfile.gettempdir(), f"pdx_bench_{os.getpid()}.pdx")
    pf.save(idx, path)
    o = orc.OracleIndex(path)
    os.remove(path)
    qr = o.preprocess(q[:5000])
    # single thread: calibrate the sample to ~cpu_seconds
    t = time.perf_counter()
    o.search(qr[:20], args.k, args.nprobe, transformed=True, threads=1)
    per_q = (time.perf_counter() - t) / 20
    n1 = int(max(20, min(len(qr), args.cpu_seconds / max(per_q, 1e-6))))
    t = time.perf_counter()
    o.search(qr[:n1], args.k, args.nprobe, transformed=True, threads=1)
    qps1 = n1 / (time.perf_counter() - t)
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))
    nq = int(min(len(qr), max(cores * 20, qps1 * cores * args.cpu_seconds / 4)))
    t = time.perf_counter()
    o.search(qr[:nq], args.k, args.nprobe, transformed=True, threads=cores)
    qpsn = nq / (time.perf_counter() - t)
    log(f"cpu oracle ({orc.isa()}): 1 thread {qps1:.0f} QPS over {n1} queries; {cores} threads {qpsn:.0f} QPS over {nq}")
    return {"value": round(qpsn, 1), "unit": "queries/s", "cores": cores, "kind": "port",
            "single_thread_qps": round(qps1, 1), "isa": orc.isa(),
            "sample": f"{nq} pre-rotated queries on {cores} threads ({n1} on 1 thread), same index, k={args.k}, "
                      f"nprobe={args.nprobe}"}


if __name__ == "__main__":
    main()
