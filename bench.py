"""Benchmark: search QPS on the two-halves dense-space-partition index.

Workload (seeded synthetic data): D=128 vectors from a 64-component Gaussian
mixture rounded to integers in [0,255]; 100k training vectors, 1M base
vectors, 10k queries.  Index: two 64-dim halves with K=256 centroids each,
every vector's cell keyed by its (nearest, second-nearest) centroid pair in
each half (a sparse directory of the non-empty cells of the 256^4 grid), m=8
x 256 PQ codes on the residuals.  Search: cells taken in order of their pair
scores until T=10k candidates, IMI scan, top-100.

One step = one batched search of all 10k queries with the index and queries
resident in HBM.  The headline is the mixed-precision form (FP16 table + the
vectors' stored dot_cr sums + exact re-rank, results bit-identical to the
plain form); the plain FP32 form is timed beside it.  Side lines: the
one-block (nearest, second) index over all 128 dims (fused hipGraph plans),
the inverted multi-index with K=1024 per half, the real-distance form, encode
and k-means timings.  Printed: one JSON line.
"""
import argparse
import hashlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "second-candidate_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "search_scimi_1M_traffic.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def recall_at(gt, ids, R):
    """Fraction of queries whose true nearest neighbour gt[i] is among their
    first R result ids (Evaluation::calc_recall, include/evaluation.h:60-92)."""
    ids = np.asarray(ids)[:, :R]
    return float(np.mean(np.any(ids == np.asarray(gt)[:, None], axis=1)))


def lib_sha16():
    import sc_amd
    h = hashlib.sha256()
    with open(sc_amd.LIB_PATH, "rb") as f:
        h.update(f.read())
    return h.hexdigest()[:16]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def train_halves(S, ctx, mix, ntrain, kc, mp, kp, iters, dim=128):
    """Coarse codebooks per 64-dim half (k-means++ + Lloyd), residuals to the
    nearest centroid of each half, PQ codebooks on the residuals."""
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, ntrain, dim)
    th = xt.numpy()
    h = dim // 2
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, j * h:(j + 1) * h])),
                            kc, iters, seed=11 + j, want_labels=False)[0].numpy() for j in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    bsp = dim // mp
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * bsp:(s + 1) * bsp])),
                            kp, iters, seed=21 + s, want_labels=False)[0].numpy() for s in range(mp)])
    return cq, dcq, pq, S.Dev.from_numpy(ctx, pq)


def timed(fn, ctx, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.sync()
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--ntrain", type=int, default=100_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--kc", type=int, default=256)
    ap.add_argument("--T", type=int, default=10_000)
    ap.add_argument("--R", type=int, default=100)
    ap.add_argument("--w", type=int, default=100_000)
    ap.add_argument("--iters", type=int, default=10, help="Lloyd iterations (setup)")
    ap.add_argument("--ntrain960", type=int, default=500_000,
                    help="training vectors of the D=960 side line")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU work of the oracle baseline (0 = skip)")
    ap.add_argument("--recall-sample", type=int, default=100,
                    help="queries whose GPU ground truth is checked against the oracle "
                         "(0 = no recall)")
    ap.add_argument("--sweep", type=str, default="1000,2000,5000,10000,20000",
                    help="T values of the QPS/recall curve ('' = skip)")
    ap.add_argument("--side", type=int, default=1,
                    help="1: side lines (one-block SC index, K=1024 multi-index, real distances, "
                         "k-means)")
    ap.add_argument("--large", type=int, default=1,
                    help="1: the large-index side line (10M training vectors, K=4096 per half, "
                         "100M base)")
    ap.add_argument("--n-large", type=int, default=100_000_000)
    ap.add_argument("--ntrain-large", type=int, default=10_000_000)
    ap.add_argument("--iters-large", type=int, default=20)
    ap.add_argument("--recall-large", type=int, default=1000)
    ap.add_argument("--large-train", choices=("kmeans", "sample"), default="kmeans",
                    help="codebooks of the large line: k-means on the training vectors, or "
                         "training vectors drawn at random (profiling runs)")
    ap.add_argument("--encode-only", type=int, default=0,
                    help="K > 0: only time the encode of the 1M base with two 64-dim halves of "
                         "K codewords drawn from the training vectors (both forms), then exit")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter scale)")
    args = ap.parse_args()

    if args.copy_calibration_mb > 0:
        # A plain copy kernel of known size, to scale the FETCH_SIZE/WRITE_SIZE
        # counters of the scan kernel: reads and writes exactly this many bytes.
        import torch
        n = args.copy_calibration_mb << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": 3}), flush=True)
        return

    if args.encode_only > 0:
        encode_only(args)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU; this build runs on a single MI355X "
              "(--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import sc_amd as S

    dim, mp, kp, kc = 128, 8, 256, args.kc
    ctx = S.Ctx(0)
    mix = S.mix_d128()

    # ---------------- setup (untimed): train, encode, build
    t0 = time.perf_counter()
    cq, dcq, pq, dpq = train_halves(S, ctx, mix, args.ntrain, kc, mp, kp, args.iters)
    ctx.sync()
    t_train = time.perf_counter() - t0
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, args.n, dim)
    ctx.sync()

    def timed_encode(prec, reps=3):
        best, res = None, None
        for _ in range(reps):
            t0 = time.perf_counter()
            res = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 2, prec)
            ctx.sync()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best, res
    t_enc32, (cid, codes, _) = timed_encode(S.PREC_FP32)
    t_enc16, (cid16, codes16, est16) = timed_encode(S.PREC_BF16)
    enc_same = bool(np.array_equal(cid16.numpy(), cid.numpy()) and
                    np.array_equal(codes16.numpy(), codes.numpy()))
    del cid16, codes16
    t0 = time.perf_counter()
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 2)
    ctx.sync()
    t_build = time.perf_counter() - t0
    ncells = int(ix.shape()[1])
    log("two halves K=%d: trained %.2fs, encode fp32 %.4fs / bf16 %.4fs (identical %s), "
        "%d non-empty cells, built %.3fs" % (kc, t_train, t_enc32, t_enc16, enc_same, ncells,
                                             t_build))
    q = S.gen_rows(ctx, mix, S.STREAM_QUERY, 0, args.nq, dim)
    ctx.sync()

    # ---------------- the two forms, batched (all queries in one call)
    res = {}
    for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
        ids = S.Dev(ctx, (args.nq, args.R), np.int32)
        dists = S.Dev(ctx, (args.nq, args.R), np.float32)
        sums = S.Dev(ctx, (args.nq,), np.int32)

        def step(form=form, ids=ids, dists=dists, sums=sums):
            ix.search(q, S.KIND_SCIMI, args.R, args.w, args.T, form=form, ids=ids,
                      dists=dists, sums=sums, stats=False)
        ms = timed(step, ctx, args.steps, args.warmup) * 1e3
        sts = [ix.search(q, S.KIND_SCIMI, args.R, args.w, args.T, form=form, ids=ids,
                         dists=dists, sums=sums, stats=True)[3] for _ in range(3)]
        res[name] = {"ms": ms, "qps": args.nq / (ms * 1e-3), "ids": ids.numpy(),
                     "dists": dists.numpy(), "sums": sums.numpy(),
                     "scan_ms": float(np.mean([s.ms_scan for s in sts])),
                     "scan_kernel_ms": float(np.mean([s.ms_scan_kernel for s in sts])),
                     "rank_ms": float(np.mean([s.ms_traverse for s in sts])),
                     "lut_ms": float(np.mean([s.ms_lut for s in sts])),
                     "cand": int(sts[-1].candidates), "cells": int(sts[-1].cells)}
        log("%s: %.3f ms/step, %.0f QPS; lut %.3f rank %.3f scan %.3f ms"
            % (name, ms, res[name]["qps"], res[name]["lut_ms"], res[name]["rank_ms"],
               res[name]["scan_ms"]))
    identical = bool(np.array_equal(res["plain"]["ids"], res["mixed"]["ids"]) and
                     np.array_equal(res["plain"]["dists"].view(np.uint32),
                                    res["mixed"]["dists"].view(np.uint32)) and
                     np.array_equal(res["plain"]["sums"], res["mixed"]["sums"]))
    cand = res["plain"]["cand"]

    # ---------------- recall@100: exact nearest neighbours on the GPU (checked
    # against the oracle's brute force on a sample, in the CPU-baseline leg)
    qh = q.numpy()
    recall, gt, gt_info = None, None, None
    if args.recall_sample > 0:
        t0 = time.perf_counter()
        gt = S.groundtruth_nn1(ctx, x, q)[0].numpy()
        t_gt = time.perf_counter() - t0
        recall = recall_at(gt, res["plain"]["ids"], args.R)
        gt_info = {"queries": int(args.nq), "seconds": round(t_gt, 4)}
        log("recall@%d: %.4f (ground truth %.3fs)" % (args.R, recall, t_gt))

    # ---------------- QPS / recall over T (both forms)
    curve = []
    if args.sweep:
        for tv in [int(v) for v in args.sweep.split(",") if v]:
            row = {"T": tv}
            for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
                ids = S.Dev(ctx, (args.nq, args.R), np.int32)
                ms = timed(lambda: ix.search(q, S.KIND_SCIMI, args.R, args.w, tv, form=form,
                                             ids=ids, stats=False), ctx, 3, 1) * 1e3
                row["qps_" + name] = round(args.nq / (ms * 1e-3), 1)
                if form == S.FORM_PLAIN and gt is not None:
                    row["recall_at_%d" % args.R] = round(recall_at(gt, ids.numpy(), args.R), 4)
            curve.append(row)
            log("T=%d: %s" % (tv, row))

    # ---------------- roofline of the dominant kernel: the scan.  Algorithmic
    # bytes per launch: (mp + 4) bytes per candidate (its codes and id, as the
    # reference's scan copies both, include/multi_query.h:209-262), the query
    # vectors, and the R ids and distances written per query.
    scan_bytes = cand * (mp + 4) + args.nq * dim * 4 + args.nq * args.R * 8
    sha = lib_sha16()
    traffic, traffic_src = None, None
    if os.path.exists(TRAFFIC_JSON):
        t = json.load(open(TRAFFIC_JSON))
        wl = t.get("workload", {})
        if t.get("lib_sha16") == sha and (wl.get("n"), wl.get("nq"), wl.get("T"), wl.get("R"),
                                          wl.get("kc")) == (args.n, args.nq, args.T, args.R, kc):
            traffic = {k: int(v) for k, v in t["traffic_bytes"].items()}
            traffic_src = os.path.relpath(TRAFFIC_JSON, ROOT)

    def roof(name):
        # the scan launches alone (HIP events around them on the context
        # stream; the scan of the few re-ranked queries overlaps the main one)
        kms = res[name]["scan_kernel_ms"] or res[name]["scan_ms"]
        ach = scan_bytes / (kms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(name) if traffic else None,
                "kernel": "k_scan_topk (%s IMI scan)" % name,
                "kernel_ms": round(kms, 4),
                "algorithmic_bytes_per_launch": int(scan_bytes),
                "traffic_source": traffic_src}

    # ---------------- CPU baseline: the oracle on a bounded sample
    cpu = None
    if args.cpu_seconds > 0:
        import sco   # the oracle: the CPU baseline and the checker of the GPU ground truth
        e = ix.export()
        ck = ix.cell_keys()
        oix = sco.OracleIndex(dim, 2, kc, mp, kp, cq, pq, e["L"], e["pid"], e["codes"], ckeys=ck)
        threads = sco.nthreads()

        def run_cpu(nthr, seconds):
            n_cal = min(16 * nthr, args.nq)
            t0 = time.perf_counter()
            oix.search_batch(sco.SCIMI, qh[:n_cal], args.R, args.w, args.T, threads=nthr)
            dt = time.perf_counter() - t0
            ns = int(min(args.nq, max(n_cal, n_cal * seconds / max(dt, 1e-6))))
            t0 = time.perf_counter()
            oids, od, osum = oix.search_batch(sco.SCIMI, qh[:ns], args.R, args.w, args.T,
                                              threads=nthr)
            dt = time.perf_counter() - t0
            return ns / dt, ns, bool(np.array_equal(oids, res["plain"]["ids"][:ns]))
        v_all, ns_all, match = run_cpu(threads, args.cpu_seconds)
        v_one, ns_one, match1 = run_cpu(1, max(1.0, args.cpu_seconds / 3))
        cpu = {"value": round(v_all, 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d queries, same index and budget, OpenMP over queries"
                         % (ns_all, args.nq),
               "cpu_model": cpu_model(),
               "one_thread": {"value": round(v_one, 2), "sample_queries": ns_one},
               "ids_match_gpu": bool(match and match1)}
        log("cpu oracle: %.1f QPS on %d threads, %.1f QPS on 1 thread (%s); ids match %s"
            % (v_all, threads, v_one, cpu["cpu_model"], cpu["ids_match_gpu"]))
        if gt is not None:   # the GPU ground truth, checked by the oracle's brute force
            nchk = min(args.recall_sample, args.nq)
            gt_ok = bool(np.array_equal(gt[:nchk], sco.bruteforce_nn1(x.numpy(), qh[:nchk])))
            gt_info.update({"oracle_checked_queries": int(nchk), "matches_oracle": gt_ok})
            cpu["ground_truth_checked"] = gt_ok
            log("ground truth: oracle agrees on %d queries: %s" % (nchk, gt_ok))

    # ---------------- side lines
    side = {}
    if args.side:
        side = side_lines(S, ctx, mix, x, q, gt, args, dim, mp, kp)
    if args.large:
        side["large_two_halves_k4096"] = large_line(S, ctx, args)

    encode_flops = 2.0 * args.n * (kc * dim + mp * kp * (dim // mp))
    out = {
        "metric": "QPS at fixed recall@100, mixed-prec batched vs plain FP32 GPU, 1M×128",
        "value": round(res["mixed"]["qps"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["mixed"]["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16+fp32",
        "data": "synthetic (seeded Philox Gaussian mixture, integer-valued)",
        "config": {
            "workload": "two 64-dim halves, K=256 (nearest, second) per half, PQ 8x256, "
                        "%d base, %d queries, T=%d, w=%d, top-%d" % (args.n, args.nq, args.T,
                                                                     args.w, args.R),
            "value_form": "mixed-precision batched (FP16 diff_qr table + stored dot_cr sums, "
                          "exact FP32 re-rank), one sc_search call per step",
            "qps_plain_fp32_batched": round(res["plain"]["qps"], 1),
            "ms_per_step_plain": round(res["plain"]["ms"], 3),
            "mixed_over_plain": round(res["mixed"]["qps"] / res["plain"]["qps"], 3),
            "identical_to_plain": identical,
            "recall_at_100": recall,
            "ground_truth": gt_info,
            "candidates_per_query": round(cand / args.nq, 1),
            "cells_per_query": round(res["plain"]["cells"] / args.nq, 1),
            "non_empty_cells": ncells,
            "stage_ms": {f: {"lut": round(res[f]["lut_ms"], 4), "rank": round(res[f]["rank_ms"], 4),
                             "scan": round(res[f]["scan_ms"], 4)} for f in ("plain", "mixed")},
            "qps_recall_curve": curve,
            "encode_s": {"fp32_exact": round(t_enc32, 4), "bf16_mfma_recheck": round(t_enc16, 4),
                         "identical": enc_same,
                         "algorithmic_tflops_bf16": round(encode_flops / t_enc16 / 1e12, 2),
                         "rechecked_per_vector": round(est16.rechecked / args.n, 2)},
            "build_s": round(t_build, 4),
            "train_s": round(t_train, 2),
            "library_sha16": sha,
            "side_lines": side,
        },
        "roofline": roof("mixed"),
        "roofline_plain": roof("plain"),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ix.close()
    ctx.close()


def side_lines(S, ctx, mix, x, q, gt, args, dim, mp, kp):
    """The other index families on the same base and queries."""
    out = {}
    nq = args.nq
    # one block of 128 dims, (nearest, second) cells of K=256 (the reference's
    # SCQuery::search_mr_ivf shape): autotuned hipGraph plans of the fused kernel
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, args.ntrain, dim)
    cq1, _ = S.kmeans(ctx, xt, 256, args.iters, seed=1, want_labels=False)
    S.residual(ctx, xt, cq1, 1, 256, 0)
    rh = xt.numpy()
    pq1 = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                             kp, args.iters, seed=2 + s, want_labels=False)[0].numpy()
                    for s in range(mp)])
    dpq1 = S.Dev.from_numpy(ctx, pq1)
    del xt, rh
    cid, codes, _ = S.encode(ctx, x, cq1, dpq1, 1, 256, mp, kp, 2, S.PREC_BF16)
    ix1 = S.Index.build(ctx, cid, codes, dim, cq1, 1, 256, dpq1, mp, kp, 2)
    del cid, codes
    row = {"workload": "one 128-dim block, K=256 (nearest, second) cells, PQ 8x256, "
                       "T=%d, top-%d (SCQuery::search_mr_ivf shape)" % (args.T, args.R)}
    ids1 = None
    for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
        pl = S.Plan(ix1, q, S.KIND_SC, args.R, args.w, args.T, form=form)
        tune = pl.autotune()
        ms = timed(lambda: pl.execute(stats=False), ctx, args.steps, args.warmup) * 1e3
        row["qps_" + name] = round(nq / (ms * 1e-3), 1)
        row["launch_" + name] = "%d threads x %d workgroups, variant %d" % (
            tune["threads"], tune["grid"], tune["variant"])
        if form == S.FORM_PLAIN:
            ids1 = pl.ids.numpy()
            if gt is not None:
                row["recall_at_%d" % args.R] = round(recall_at(gt, ids1, args.R), 4)
        else:
            row["identical"] = bool(np.array_equal(ids1, pl.ids.numpy()))
        pl.close()
    ix1.close()
    out["sc_one_block_k256"] = row
    log("one-block SC K=256: %s" % row)

    # inverted multi-index, two halves with K=1024 each (config 2 shape)
    kci = 1024
    cqi, dcqi, pqi, dpqi = train_halves(S, ctx, mix, args.ntrain, kci, mp, kp, args.iters)
    icid, icodes, _ = S.encode(ctx, x, dcqi, dpqi, 2, kci, mp, kp, 1, S.PREC_BF16)
    iix = S.Index.build(ctx, icid, icodes, dim, dcqi, 2, kci, dpqi, mp, kp, 1)
    del icid, icodes
    row = {"workload": "IMI, two 64-dim halves with K=1024 each, PQ 8x256, T=%d, top-%d"
                       % (args.T, args.R)}
    ref_ids = None
    for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
        ids = S.Dev(ctx, (nq, args.R), np.int32)
        ms = timed(lambda: iix.search(q, S.KIND_IMI, args.R, args.w, args.T, form=form, ids=ids,
                                      stats=False), ctx, args.steps, args.warmup) * 1e3
        row["qps_" + name] = round(nq / (ms * 1e-3), 1)
        if form == S.FORM_PLAIN:
            ref_ids = ids.numpy()
            if gt is not None:
                row["recall_at_%d" % args.R] = round(recall_at(gt, ref_ids, args.R), 4)
            st = iix.search(q, S.KIND_IMI, args.R, args.w, args.T, stats=True)[3]
            row["stage_ms"] = {"lut": round(st.ms_lut, 3), "rank": round(st.ms_rank, 3),
                               "traverse": round(st.ms_traverse, 3), "scan": round(st.ms_scan, 3)}
            row["cells_popped_per_query"] = round(st.visited / nq, 1)
        else:
            row["identical"] = bool(np.array_equal(ref_ids, ids.numpy()))
    iix.close()
    out["imi_k1024"] = row
    log("IMI K=1024: %s" % row)

    # real distances on the two-halves index (candidates ranked by their raw rows)
    kc = args.kc
    cq, dcq, pq, dpq = train_halves(S, ctx, mix, args.ntrain, kc, mp, kp, args.iters)
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 2, S.PREC_BF16)
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 2)
    del cid, codes
    ix.attach_raw(x)
    ids = S.Dev(ctx, (nq, args.R), np.int32)
    sums = S.Dev(ctx, (nq,), np.int32)
    ms = timed(lambda: ix.search(q, S.KIND_SCIMI, args.R, args.w, args.T, form=S.FORM_REAL,
                                 ids=ids, sums=sums, stats=False), ctx, args.steps,
               args.warmup) * 1e3
    rcand = float(sums.numpy().astype(np.int64).sum())
    out["real_distance_two_halves"] = {
        "qps": round(nq / (ms * 1e-3), 1), "ms_per_step": round(ms, 3),
        "raw_gb_per_s": round(rcand * (dim * 4 + 4) / (ms * 1e-3) / 1e9, 1),
        "recall_at_%d" % args.R: round(recall_at(gt, ids.numpy(), args.R), 4) if gt is not None else None}
    log("real distances: %s" % out["real_distance_two_halves"])
    ix.close()

    # k-means at K=4096 on 1M half-vectors (the large-index training shape)
    xk = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 1_000_000, dim)
    xh64 = S.Dev.from_numpy(ctx, np.ascontiguousarray(xk.numpy()[:, :64]))
    del xk

    def timed_kmeans(iters, prec):
        t0 = time.perf_counter()
        c, lab = S.kmeans(ctx, xh64, 4096, iters, seed=3, prec=prec)
        ctx.sync()
        return time.perf_counter() - t0, c, lab
    timed_kmeans(0, S.PREC_FP32)
    t_seed = timed_kmeans(0, S.PREC_FP32)[0]
    t_ex, c_ex, l_ex = timed_kmeans(3, S.PREC_FP32)
    t_mx, c_mx, l_mx = timed_kmeans(3, S.PREC_BF16)
    same = bool(np.array_equal(l_ex.numpy(), l_mx.numpy()) and
                np.array_equal(c_ex.numpy().view(np.uint32), c_mx.numpy().view(np.uint32)))
    flops = 3 * 2.0 * 1_000_000 * 4096 * 64
    out["kmeans_k4096"] = {"workload": "k-means, 1M points x 64 dims, K=4096, 3 Lloyd iterations",
                           "seeding_s": round(t_seed, 3),
                           "lloyd_fp32_s": round(t_ex - t_seed, 4),
                           "lloyd_bf16_s": round(t_mx - t_seed, 4), "identical": same,
                           "lloyd_bf16_algorithmic_tflops": round(flops / max(t_mx - t_seed, 1e-9) / 1e12, 2)}
    log("k-means: %s" % out["kmeans_k4096"])
    del xh64
    out["imi_d960_budgets"] = d960_line(S, ctx, args)
    return out


def d960_line(S, ctx, args):
    """High-dimensional batched search: D=960 (32-component mixture on [0,1]),
    IMI with two 480-dim halves of K=1024, PQ 16x256, 1M base, 20k queries
    with per-query candidate budgets drawn from U[1k, 30k], each form run as
    one replayed hipGraph plan (the staged kernels captured with their
    budgets), recall@100 on a 2000-query sample."""
    dim, kc, mp, kp, n, nq, R = 960, 1024, 16, 256, 1_000_000, 20_000, 100
    mix = S.mix_d960()
    t0 = time.perf_counter()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, args.ntrain960, dim)
    th = xt.numpy()
    cq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 480:(h + 1) * 480])),
                            kc, args.iters, seed=41 + h, want_labels=False)[0].numpy()
                   for h in range(2)])
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 60:(s + 1) * 60])),
                            kp, args.iters, seed=51 + s, want_labels=False)[0].numpy()
                   for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    del xt, th, rh
    ctx.sync()
    t_train = time.perf_counter() - t0
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    t0 = time.perf_counter()
    cid, codes, _ = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 1)   # 60-dim PQ blocks: exact form
    ctx.sync()
    t_enc = time.perf_counter() - t0
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 1)
    del cid, codes
    q = S.gen_rows(ctx, mix, S.STREAM_QUERY, 0, nq, dim)
    T = np.random.default_rng(960).integers(1000, 30001, nq).astype(np.int32)
    dT = S.Dev.from_numpy(ctx, T)
    w = 200_000
    row = {"workload": "D=960 IMI, two 480-dim halves with K=1024 each, PQ 16x256, 1M base, "
                       "20k queries with per-query budgets U[1k,30k], top-100, one hipGraph "
                       "plan per form",
           "train_s": round(t_train, 2), "train_vectors": args.ntrain960,
           "encode_s_fp32_exact": round(t_enc, 3)}
    ids = {}
    for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
        pl = S.Plan(ix, q, S.KIND_IMI, R, w, 30000, form=form, budgets=dT)
        ms = timed(lambda: pl.execute(stats=False), ctx, args.steps, args.warmup) * 1e3
        row["qps_" + name] = round(nq / (ms * 1e-3), 1)
        row["ms_" + name] = round(ms, 3)
        ids[name] = (pl.ids.numpy().copy(), pl.sums.numpy().copy())
        pl.close()
        st = ix.search(q, S.KIND_IMI, R, w, 30000, form=form, budgets=dT, stats=True)[3]
        cand = float(ids[name][1].astype(np.int64).sum())
        algo = cand * (mp + 4) + nq * dim * 4 + nq * R * 8
        row["stage_ms_" + name] = {"lut": round(st.ms_lut, 3), "rank": round(st.ms_rank, 3),
                                   "traverse": round(st.ms_traverse, 3),
                                   "scan": round(st.ms_scan, 3)}
        row["roofline_scan_" + name] = {
            "bound": "hbm", "achieved": round(algo / (st.ms_scan * 1e-3) / 1e9, 1),
            "peak": 8000.0, "unit": "GB/s",
            "frac": round(algo / (st.ms_scan * 1e-3) / 8e12, 4),
            "algorithmic_bytes_per_launch": int(algo)}
    row["candidates_per_query"] = round(float(ids["plain"][1].mean()), 1)
    row["mixed_over_plain"] = round(row["qps_mixed"] / row["qps_plain"], 3)
    row["identical"] = bool(np.array_equal(ids["plain"][0], ids["mixed"][0]) and
                            np.array_equal(ids["plain"][1], ids["mixed"][1]))
    gs = 2000
    gt = S.groundtruth_nn1(ctx, x, S.Dev.from_numpy(ctx, q.numpy()[:gs]))[0].numpy()
    row["recall_at_100_sample"] = round(recall_at(gt, ids["mixed"][0][:gs], R), 4)
    ix.close()
    del x, q
    log("D=960 IMI: %s" % row)
    return row


def encode_only(args):
    """Encode timing: the 1M-vector D=128 base, two 64-dim halves with K
    codewords each drawn from 200k training vectors, PQ 8x256 (sub-vectors of
    the training residuals), both forms three times; one JSON line."""
    import sc_amd as S
    ctx = S.Ctx(0)
    mix = S.mix_d128()
    dim, mp, kp, kc, n = 128, 8, 256, args.encode_only, 1_000_000
    rng = np.random.default_rng(5)
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 200_000, dim)
    th = xt.numpy()
    cq = np.stack([th[rng.choice(th.shape[0], kc, replace=False), h * 64:(h + 1) * 64]
                   for h in range(2)]).astype(np.float32)
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    pq = np.stack([rh[rng.choice(rh.shape[0], kp, replace=False), s * 16:(s + 1) * 16]
                   for s in range(mp)]).astype(np.float32)
    dpq = S.Dev.from_numpy(ctx, pq)
    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, n, dim)
    out = {"workload": "encode 1M x 128, two 64-dim halves, K=%d (nearest, second), PQ 8x256"
                       % kc}
    res = {}
    for prec, name in ((S.PREC_BF16, "bf16_mfma_recheck"), (S.PREC_FP32, "fp32_exact")):
        best = None
        for _ in range(3):
            ctx.sync()
            t0 = time.perf_counter()
            cid, codes, st = S.encode(ctx, x, dcq, dpq, 2, kc, mp, kp, 2, prec)
            ctx.sync()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        res[name] = (cid.numpy(), codes.numpy())
        out[name + "_ms"] = round(best * 1e3, 3)
        if prec == S.PREC_BF16:
            out["rechecked_per_vector"] = round(st.rechecked / n, 2)
    flops = 2.0 * n * (kc * dim + mp * kp * (dim // mp))
    out["algorithmic_tflops_bf16"] = round(flops / (out["bf16_mfma_recheck_ms"] * 1e-3) / 1e12, 2)
    out["identical"] = bool(all(np.array_equal(a, b) for a, b in
                                zip(res["bf16_mfma_recheck"], res["fp32_exact"])))
    print(json.dumps(out), flush=True)
    ctx.close()


def large_line(S, ctx, args):
    """The large dense-partition index on one GPU: D=128 mixture, codebooks
    trained on 10M vectors (k-means++ then 20 Lloyd iterations per codebook,
    the assignment on BF16 MFMA with the exact re-check, FP64 centre sums),
    two 64-dim halves with K=4096 (nearest, second) cells each, PQ 8x256, a
    100M-vector base generated and encoded a 1M-vector chunk at a time (8e8
    bytes of codes, beyond the 256 MB MALL), 10k queries served in batches of
    4096, 64 and 1 (2 x (4096 + 14 x 64 + 8 x 1)), T=10k, top-100, both forms;
    recall@100 against exact nearest neighbours over the whole base (chunked)
    on a sample of the queries."""
    dim, kc, mp, kp, R, T, w = 128, 4096, 8, 256, 100, 10_000, 100_000
    n, ntrain, nq = args.n_large, args.ntrain_large, 10_000
    mix = S.mix_d128()
    row = {"workload": "two 64-dim halves, K=4096 (nearest, second) per half, PQ 8x256, "
                       "%d base streamed in 1M chunks, %d training vectors, %d queries in "
                       "batches of 4096/64/1, T=%d, w=%d, top-%d" % (n, ntrain, nq, T, w, R)}
    # ---- training (untimed setup; reported)
    t0 = time.perf_counter()
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, ntrain, dim)
    th = xt.numpy()
    t_km = 0.0
    cq = []
    rng = np.random.default_rng(71)
    sample = args.large_train == "sample"
    row["codebooks"] = ("k-means" if not sample else
                        "training vectors drawn at random (profiling run)")
    for h in range(2):
        if sample:
            cq.append(th[rng.choice(ntrain, kc, replace=False), h * 64:(h + 1) * 64].copy())
            continue
        xh = S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 64:(h + 1) * 64]))
        ctx.sync()
        t1 = time.perf_counter()
        c, _ = S.kmeans(ctx, xh, kc, args.iters_large, seed=71 + h, prec=S.PREC_BF16,
                        want_labels=False)
        ctx.sync()
        t_km += time.perf_counter() - t1
        cq.append(c.numpy())
        xh.free()
    cq = np.stack(cq)
    dcq = S.Dev.from_numpy(ctx, cq)
    S.residual(ctx, xt, dcq, 2, kc, 0)
    rh = xt.numpy()
    xt.free()
    if sample:
        pq = np.stack([rh[rng.choice(ntrain, kp, replace=False), s * 16:(s + 1) * 16]
                       for s in range(mp)]).astype(np.float32)
    else:
        pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                                kp, args.iters_large, seed=81 + s, prec=S.PREC_BF16,
                                want_labels=False)[0].numpy() for s in range(mp)])
    dpq = S.Dev.from_numpy(ctx, pq)
    del th, rh
    ctx.sync()
    row["train_s"] = round(time.perf_counter() - t0, 2)
    row["coarse_kmeans_s"] = round(t_km, 2)
    log("large: trained in %.1fs (coarse k-means %.1fs)" % (row["train_s"], t_km))
    # ---- streamed encode and build
    def fill(r0, m, buf):
        S.gen_rows(ctx, mix, S.STREAM_BASE, r0, m, dim, out=buf)
    t0 = time.perf_counter()
    cid, codes, rech, exact = S.encode_stream(ctx, fill, n, dim, dcq, dpq, 2, kc, mp, kp, 2,
                                              prec=S.PREC_BF16)
    ctx.sync()
    t_enc = time.perf_counter() - t0
    enc_flops = 2.0 * n * (kc * dim + mp * kp * (dim // mp))
    row["encode_s"] = round(t_enc, 2)
    row["encode_algorithmic_tflops_bf16"] = round(enc_flops / t_enc / 1e12, 2)
    row["encode_rechecked_per_vector"] = round(rech / n, 2)
    t0 = time.perf_counter()
    ix = S.Index.build(ctx, cid, codes, dim, dcq, 2, kc, dpq, mp, kp, 2)
    ctx.sync()
    row["build_s"] = round(time.perf_counter() - t0, 3)
    cid.free()
    codes.free()
    row["non_empty_cells"] = int(ix.shape()[1])
    log("large: encoded %d in %.1fs, %d cells" % (n, t_enc, row["non_empty_cells"]))
    # ---- queries in mixed batch sizes, both forms
    q = S.gen_rows(ctx, mix, S.STREAM_QUERY, 0, nq, dim)
    sched = ([4096] + [64] * 14 + [1] * 8) * 2
    assert sum(sched) == nq
    ids = S.Dev(ctx, (nq, R), np.int32)
    dists = S.Dev(ctx, (nq, R), np.float32)
    sums = S.Dev(ctx, (nq,), np.int32)
    res = {}
    for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
        def serve():
            lat = {}
            q0 = 0
            for bs in sched:
                t1 = time.perf_counter()
                ix.search(q.rows(q0, bs), S.KIND_SCIMI, R, w, T, form=form, ids=ids.rows(q0, bs),
                          dists=dists.rows(q0, bs), sums=sums.rows(q0, bs), stats=False)
                ctx.sync()
                lat.setdefault(bs, []).append(time.perf_counter() - t1)
                q0 += bs
            return lat
        serve()                      # warm-up
        ctx.sync()
        t0 = time.perf_counter()
        lat = serve()
        dt = time.perf_counter() - t0
        st = ix.search(q, S.KIND_SCIMI, R, w, T, form=form, stats=True)[3]
        res[name] = (ids.numpy().copy(), dists.numpy().copy(), sums.numpy().copy())
        cand = float(res[name][2].astype(np.int64).sum())
        algo = cand * (mp + 4) + nq * dim * 4 + nq * R * 8
        row["qps_" + name] = round(nq / dt, 1)
        row["batch_ms_" + name] = {str(bs): round(1e3 * float(np.median(v)), 3)
                                   for bs, v in sorted(lat.items())}
        row["stage_ms_" + name + "_one_batch_of_10k"] = {
            "lut": round(st.ms_lut, 3), "rank": round(st.ms_traverse, 3),
            "scan": round(st.ms_scan, 3)}
        kms = st.ms_scan_kernel or st.ms_scan   # the scan launches alone
        row["roofline_scan_" + name] = {
            "bound": "hbm", "achieved": round(algo / (kms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(algo / (kms * 1e-3) / 8e12, 4),
            "kernel_ms": round(kms, 3),
            "algorithmic_bytes_per_launch": int(algo)}
        row["rank_cells_scored_per_query"] = round(st.visited / nq, 1)
        row["rank_gb_per_s"] = round(st.visited * 4 / (st.ms_traverse * 1e-3) / 1e9, 1)
        log("large %s: %.0f QPS, %s" % (name, nq / dt, row["batch_ms_" + name]))
    row["mixed_over_plain"] = round(row["qps_mixed"] / row["qps_plain"], 3)
    row["identical"] = bool(all(np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                               b.view(np.uint32) if b.dtype == np.float32 else b)
                                for a, b in zip(res["plain"], res["mixed"])))
    row["candidates_per_query"] = round(float(res["plain"][2].mean()), 1)
    # ---- recall@100 on a sample: exact nearest neighbours over the whole base,
    # one regenerated 1M chunk at a time (ties to the lowest id)
    gs = args.recall_large
    if gs > 0:
        qs = S.Dev.from_numpy(ctx, q.numpy()[:gs])
        best_d = np.full(gs, np.inf)
        best_i = np.full(gs, -1, np.int64)
        buf = S.Dev(ctx, (1_000_000, dim), np.float32)
        t0 = time.perf_counter()
        for r0 in range(0, n, 1_000_000):
            m = min(1_000_000, n - r0)
            S.gen_rows(ctx, mix, S.STREAM_BASE, r0, m, dim, out=buf)
            gi, gd = S.groundtruth_nn1(ctx, buf.rows(0, m), qs)
            gi, gd = gi.numpy(), gd.numpy()
            better = gd < best_d
            best_d[better] = gd[better]
            best_i[better] = gi[better] + r0
        row["recall_at_100_sample"] = round(recall_at(best_i, res["mixed"][0][:gs], R), 4)
        row["recall_queries"] = gs
        row["groundtruth_s"] = round(time.perf_counter() - t0, 2)
        buf.free()
    ix.close()
    log("large: %s" % row)
    return row


if __name__ == "__main__":
    main()
