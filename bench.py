"""Benchmark: search QPS on the dense-space-partition (second-candidate) index.

Workload (seeded synthetic data): D=128 vectors from a 64-component Gaussian
mixture rounded to integers in [0,255]; 100k training vectors, 1M base
vectors, 10k queries.  Index: cells keyed by the (nearest, second-nearest)
pair of K=256 coarse centroids over the full 128 dimensions (the reference's
SCQuery::search_mr_ivf shape), m=8 x 256 PQ codes on the residuals.  Search:
multi-sequence traversal until T=10k candidates, top-100.

One step = one batched search of all 10k queries with the index and queries
resident in HBM.  Printed: one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "second-candidate_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--ntrain", type=int, default=100_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--kc", type=int, default=256)
    ap.add_argument("--T", type=int, default=10_000)
    ap.add_argument("--R", type=int, default=100)
    ap.add_argument("--w", type=int, default=100_000)
    ap.add_argument("--iters", type=int, default=10, help="Lloyd iterations (setup)")
    ap.add_argument("--per-query", type=int, default=200,
                    help="queries timed one call at a time (drop-in plain form)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU work for the oracle baseline (0 = skip)")
    ap.add_argument("--recall-sample", type=int, default=100,
                    help="queries whose GPU ground truth is checked against the oracle "
                         "(0 = no recall)")
    ap.add_argument("--plan", type=int, default=1,
                    help="1: time the autotuned hipGraph plan (sc_plan_*); 0: sc_search calls")
    ap.add_argument("--sweep", type=str, default="1000,2000,5000,10000,15000,20000",
                    help="T values of the QPS/recall curve ('' = skip)")
    ap.add_argument("--mixed", type=int, default=1,
                    help="1: also time the mixed-precision form (FP16 table + exact re-rank)")
    ap.add_argument("--real", type=int, default=1,
                    help="1: also time the real-distance form (exact distances from the raw vectors)")
    ap.add_argument("--kmeans", type=int, default=1,
                    help="1: also time k-means training at K=4096 on 1M x 64 (exact and mixed "
                         "Lloyd assignment)")
    ap.add_argument("--imi", type=int, default=1,
                    help="1: also time the inverted multi-index workload (two 64-dim halves, "
                         "K=1024 each, PQ 8x256, same base and queries)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter scale)")
    args = ap.parse_args()

    if args.copy_calibration_mb > 0:
        # A plain copy kernel of known size, to scale the FETCH_SIZE/WRITE_SIZE
        # counters of the search kernel: reads and writes exactly this many bytes.
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
    xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, args.ntrain, dim)
    cq, _ = S.kmeans(ctx, xt, kc, args.iters, seed=1, want_labels=False)
    S.residual(ctx, xt, cq, 1, kc, 0)
    rh = xt.numpy()
    pq = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                            kp, args.iters, seed=2 + s, want_labels=False)[0].numpy()
                   for s in range(mp)])
    del rh, xt
    dpq = S.Dev.from_numpy(ctx, pq)
    ctx.sync()
    t_train = time.perf_counter() - t0
    log("trained codebooks in %.2fs" % t_train)

    x = S.gen_rows(ctx, mix, S.STREAM_BASE, 0, args.n, dim)
    ctx.sync()
    def timed_encode(prec, reps=3):
        """Encodes all base vectors `reps` times; the fastest run's time."""
        best, res = None, None
        for _ in range(reps):
            t0 = time.perf_counter()
            res = S.encode(ctx, x, cq, dpq, 1, kc, mp, kp, 2, prec)
            ctx.sync()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best, res
    t_encode, (cid, codes, _) = timed_encode(S.PREC_FP32)
    # the mixed-precision assignment (BF16 MFMA distances + exact re-check)
    t_encode16, (cid16, codes16, est16) = timed_encode(S.PREC_BF16)
    enc16_same = bool(np.array_equal(cid16.numpy(), cid.numpy()) and
                      np.array_equal(codes16.numpy(), codes.numpy()))
    log("encode: fp32 exact %.4fs, bf16+re-check %.4fs (identical: %s, %.1f re-checked/row)"
        % (t_encode, t_encode16, enc16_same, est16.rechecked / args.n))
    del cid16, codes16
    ix = S.Index.build(ctx, cid, codes, dim, cq, 1, kc, dpq, mp, kp, 2)
    del cid, codes
    q = S.gen_rows(ctx, mix, S.STREAM_QUERY, 0, args.nq, dim)
    ctx.sync()
    log("encoded %d vectors in %.3fs; index built" % (args.n, t_encode))

    ids = S.Dev(ctx, (args.nq, args.R), np.int32)
    dists = S.Dev(ctx, (args.nq, args.R), np.float32)
    sums = S.Dev(ctx, (args.nq,), np.int32)

    # eager calls with statistics: the fused kernel's own duration (HIP events
    # on the context stream around that launch) and its per-phase times
    eager = [ix.search(q, S.KIND_SC, args.R, args.w, args.T, form=S.FORM_PLAIN, ids=ids,
                       dists=dists, sums=sums, stats=True)[3] for _ in range(4)][1:]
    st0 = eager[-1]
    ms_kernel = float(np.mean([e.ms_scan for e in eager]))
    phases = [round(float(v), 4) for v in st0.fused_phase_ms]
    log("eager call: kernel %.3f ms, lut %.3f ms; fused phases per workgroup (ms): "
        "lut %.3f sort %.3f traverse %.3f scan %.3f select %.3f"
        % (st0.ms_scan, st0.ms_lut, *phases))

    plan, tune = None, None
    if args.plan:
        plan = S.Plan(ix, q, S.KIND_SC, args.R, args.w, args.T, form=S.FORM_PLAIN)
        tune = plan.autotune()
        log("autotuned plan: %s" % tune)
        ids, dists, sums = plan.ids, plan.dists, plan.sums

        def step():
            return plan.execute(stats=True)
    else:
        def step():
            return ix.search(q, S.KIND_SC, args.R, args.w, args.T, form=S.FORM_PLAIN, ids=ids,
                             dists=dists, sums=sums, stats=True)[3]

    for _ in range(args.warmup):
        step()
    ctx.sync()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1e3 / args.steps
    qps = args.nq * args.steps / elapsed
    st = stats[-1]
    cand = st.candidates
    fused = st0.ms_rank == 0.0 and st0.ms_traverse == 0.0
    # algorithmic bytes of the dominant kernel: (mp + 4) bytes per candidate
    # (its codes and its id, as the reference's scan copies both,
    # include/sc_query.h:197-239), the query vectors and the R ids and
    # distances written per query
    scan_bytes = cand * (mp + 4) + args.nq * dim * 4 + args.nq * args.R * 8
    achieved = scan_bytes / (ms_kernel * 1e-3) / 1e9
    log("batched: %.1f ms/step, %.0f QPS; stages lut %.3f rank %.3f trav %.3f scan %.3f ms; "
        "%.0f candidates, %.1f cells, %.1f visited per query"
        % (ms_step, qps, st.ms_lut, st.ms_rank, st.ms_traverse, st.ms_scan, cand / args.nq,
           st.cells / args.nq, st.visited / args.nq))

    # ---------------- mixed-precision form, batched (same plan machinery)
    mixed = None
    if args.mixed and args.plan:
        mplan = S.Plan(ix, q, S.KIND_SC, args.R, args.w, args.T, form=S.FORM_MIXED)
        mtune = mplan.autotune()
        for _ in range(args.warmup):
            mplan.execute(stats=False)
        ctx.sync()
        mstats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            mstats.append(mplan.execute(stats=True))
        ctx.sync()
        mel = time.perf_counter() - t0
        same = bool(np.array_equal(mplan.ids.numpy(), ids.numpy()) and
                    np.array_equal(mplan.dists.numpy().view(np.uint32),
                                   dists.numpy().view(np.uint32)))
        mixed = {"qps": round(args.nq * args.steps / mel, 1),
                 "ms_per_step": round(mel * 1e3 / args.steps, 3),
                 "reranked_per_query": round(mstats[-1].reranked / args.nq, 1),
                 "identical_to_plain": same,
                 "launch": "%d threads x %d workgroups, variant %d" % (
                     mtune["threads"], mtune["grid"], mtune["variant"])}
        log("mixed batched: %.1f ms/step, %.0f QPS, %.1f re-ranked/query, identical: %s"
            % (mixed["ms_per_step"], mixed["qps"], mixed["reranked_per_query"], same))
        mplan.close()

    # ---------------- drop-in plain form: one query per call
    pq_n = min(args.per_query, args.nq)
    qh = q.numpy()
    one_ids = S.Dev(ctx, (1, args.R), np.int32)
    one_d = S.Dev(ctx, (1, args.R), np.float32)
    one_s = S.Dev(ctx, (1,), np.int32)
    qps_single = None
    if pq_n > 0:
        qd1 = [S.Dev.from_numpy(ctx, qh[i:i + 1]) for i in range(pq_n)]
        ix.search(qd1[0], S.KIND_SC, args.R, args.w, args.T, ids=one_ids, dists=one_d,
                  sums=one_s, stats=False)
        ctx.sync()
        t0 = time.perf_counter()
        for i in range(pq_n):
            ix.search(qd1[i], S.KIND_SC, args.R, args.w, args.T, ids=one_ids, dists=one_d,
                      sums=one_s, stats=False)
            one_ids.numpy()                  # the reference API returns host arrays per call
        qps_single = pq_n / (time.perf_counter() - t0)
        log("per-query drop-in calls: %.0f QPS" % qps_single)

    # ---------------- recall@100 over all queries: exact nearest neighbours on
    # the GPU (sc_groundtruth_nn1), checked against the oracle's brute force on
    # a sample of the queries
    import sco
    gids = ids.numpy()
    recall = None
    gt_info = None
    if args.recall_sample > 0:
        ns = args.nq
        t0 = time.perf_counter()
        gt = S.groundtruth_nn1(ctx, x, q)[0].numpy()
        t_gt = time.perf_counter() - t0
        nchk = min(args.recall_sample, args.nq)
        xh = x.numpy()
        gt_ok = bool(np.array_equal(gt[:nchk], sco.bruteforce_nn1(xh, qh[:nchk])))
        del xh
        recall = sco.recall(gt, gids, args.R)
        gt_info = {"queries": int(args.nq), "seconds": round(t_gt, 4),
                   "oracle_checked_queries": int(nchk), "matches_oracle": gt_ok}
        log("recall@%d on %d queries: %.4f (GPU ground truth in %.3fs, matches the oracle on "
            "%d queries: %s)" % (args.R, ns, recall, t_gt, nchk, gt_ok))

    # ---------------- real-distance form (real_dist = true): the candidates are
    # ranked by their exact distance to the query, read from the raw vectors
    real = None
    if args.real:
        ix.attach_raw(x)
        rids = S.Dev(ctx, (args.nq, args.R), np.int32)
        rd = S.Dev(ctx, (args.nq, args.R), np.float32)
        rs = S.Dev(ctx, (args.nq,), np.int32)
        for _ in range(max(1, args.warmup)):
            ix.search(q, S.KIND_SC, args.R, args.w, args.T, form=S.FORM_REAL, ids=rids, dists=rd,
                      sums=rs, stats=False)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ix.search(q, S.KIND_SC, args.R, args.w, args.T, form=S.FORM_REAL, ids=rids, dists=rd,
                      sums=rs, stats=False)
        ctx.sync()
        rel = time.perf_counter() - t0
        rcand = float(rs.numpy().astype(np.int64).sum())
        # per candidate: its raw row (dim floats) and its id
        rbytes = rcand * (dim * 4 + 4)
        real = {"qps": round(args.nq * args.steps / rel, 1),
                "ms_per_step": round(rel * 1e3 / args.steps, 3),
                "candidates_per_query": round(rcand / args.nq, 1),
                "raw_gb_per_s": round(rbytes * args.steps / rel / 1e9, 1),
                "recall_at_%d" % args.R: (round(sco.recall(gt, rids.numpy()[:ns], args.R), 4)
                                          if recall is not None else None),
                "launch": "sc_search calls (staged kernels)"}
        log("real-distance form: %.1f ms/step, %.0f QPS, %.0f GB/s of raw rows, recall %s"
            % (real["ms_per_step"], real["qps"], real["raw_gb_per_s"],
               real["recall_at_%d" % args.R]))
        ix.attach_raw(None)
        del rids, rd, rs

    # ---------------- QPS / recall curve over T (same R, both forms, batched)
    curve = []
    if args.sweep and args.plan and args.recall_sample > 0:
        for tv in [int(v) for v in args.sweep.split(",") if v]:
            row = {"T": tv}
            for form, name in ((S.FORM_PLAIN, "plain"), (S.FORM_MIXED, "mixed")):
                pl = S.Plan(ix, q, S.KIND_SC, args.R, args.w, tv, form=form)
                pl.autotune()
                pl.execute(stats=False)
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    pl.execute(stats=False)
                ctx.sync()
                row["qps_" + name] = round(args.nq * args.steps / (time.perf_counter() - t0), 1)
                if form == S.FORM_PLAIN:
                    row["recall_at_%d" % args.R] = round(
                        sco.recall(gt, pl.ids.numpy()[:ns], args.R), 4)
                pl.close()
            curve.append(row)
            log("T=%d: recall %.3f, plain %.0f QPS, mixed %.0f QPS"
                % (tv, row["recall_at_%d" % args.R], row["qps_plain"], row["qps_mixed"]))

    # ---------------- inverted multi-index: two halves with K=1024 centroids
    # each, PQ 8x256 on the residuals, T=10k (the staged kernels)
    imi = None
    if args.imi:
        kci = 1024
        t0 = time.perf_counter()
        xt = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, args.ntrain, dim)
        th = xt.numpy()
        cqi = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(th[:, h * 64:(h + 1) * 64])),
                                 kci, args.iters, seed=11 + h, want_labels=False)[0].numpy()
                        for h in range(2)])
        dcqi = S.Dev.from_numpy(ctx, cqi)
        S.residual(ctx, xt, dcqi, 2, kci, 0)
        rh = xt.numpy()
        pqi = np.stack([S.kmeans(ctx, S.Dev.from_numpy(ctx, np.ascontiguousarray(rh[:, s * 16:(s + 1) * 16])),
                                 kp, args.iters, seed=21 + s, want_labels=False)[0].numpy()
                        for s in range(mp)])
        dpqi = S.Dev.from_numpy(ctx, pqi)
        del xt, th, rh
        t_imi_train = time.perf_counter() - t0
        t0 = time.perf_counter()
        icid, icodes, _ = S.encode(ctx, x, dcqi, dpqi, 2, kci, mp, kp, 1)
        ctx.sync()
        t_imi_enc = time.perf_counter() - t0
        iix = S.Index.build(ctx, icid, icodes, dim, dcqi, 2, kci, dpqi, mp, kp, 1)
        del icid, icodes
        iids = S.Dev(ctx, (args.nq, args.R), np.int32)
        idd = S.Dev(ctx, (args.nq, args.R), np.float32)
        isums = S.Dev(ctx, (args.nq,), np.int32)
        ist = iix.search(q, S.KIND_IMI, args.R, args.w, args.T, ids=iids, dists=idd, sums=isums,
                         stats=True)[3]
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            iix.search(q, S.KIND_IMI, args.R, args.w, args.T, ids=iids, dists=idd, sums=isums,
                       stats=False)
        ctx.sync()
        iel = time.perf_counter() - t0
        imi = {"workload": "IMI, two 64-dim halves with K=1024 centroids each, PQ 8x256, "
                           "%d base, %d queries, w=%d, T=%d, top-%d" % (args.n, args.nq, args.w,
                                                                       args.T, args.R),
               "qps": round(args.nq * args.steps / iel, 1),
               "ms_per_step": round(iel * 1e3 / args.steps, 3),
               "candidates_per_query": round(ist.candidates / args.nq, 1),
               "cells_popped_per_query": round(ist.visited / args.nq, 1),
               "heap_fallback_queries": int(ist.heap_fallback),
               "stage_ms": {"lut": round(ist.ms_lut, 3), "rank": round(ist.ms_rank, 3),
                            "traverse": round(ist.ms_traverse, 3), "scan": round(ist.ms_scan, 3)},
               "encode_s": round(t_imi_enc, 4), "train_s": round(t_imi_train, 2),
               "recall_at_%d" % args.R: (round(sco.recall(gt, iids.numpy(), args.R), 4)
                                         if recall is not None else None)}
        log("imi K=1024: %.1f ms/step, %.0f QPS, %.0f candidates and %.0f cells popped per query, "
            "stages %s, recall %s" % (imi["ms_per_step"], imi["qps"], imi["candidates_per_query"],
                                      imi["cells_popped_per_query"], imi["stage_ms"],
                                      imi["recall_at_%d" % args.R]))
        iix.close()

    # ---------------- k-means at K=4096 (the large-index training shape, on a 1M-point
    # half-vector sample): seeding alone (iters=0), then 3 Lloyd iterations in the
    # exact form and in the mixed form (BF16 MFMA assignment, exact re-check)
    km = None
    if args.kmeans:
        xk = S.gen_rows(ctx, mix, S.STREAM_TRAIN, 0, 1_000_000, dim)
        xh64 = S.Dev.from_numpy(ctx, np.ascontiguousarray(xk.numpy()[:, :64]))
        del xk
        def timed_kmeans(iters, prec):
            t0 = time.perf_counter()
            c, lab = S.kmeans(ctx, xh64, 4096, iters, seed=3, prec=prec)
            ctx.sync()
            return time.perf_counter() - t0, c, lab
        timed_kmeans(0, S.PREC_FP32)                       # warm-up
        t_seed = timed_kmeans(0, S.PREC_FP32)[0]
        t_ex, c_ex, l_ex = timed_kmeans(3, S.PREC_FP32)
        t_mx, c_mx, l_mx = timed_kmeans(3, S.PREC_BF16)
        same = bool(np.array_equal(l_ex.numpy(), l_mx.numpy()) and
                    np.array_equal(c_ex.numpy().view(np.uint32), c_mx.numpy().view(np.uint32)))
        lloyd_flops = 3 * 2.0 * 1_000_000 * 4096 * 64
        km = {"workload": "k-means, 1M points x 64 dims, K=4096, 3 Lloyd iterations",
              "seeding_s": round(t_seed, 3),
              "lloyd_fp32_s": round(t_ex - t_seed, 4), "lloyd_bf16_s": round(t_mx - t_seed, 4),
              "identical": same,
              "lloyd_bf16_algorithmic_tflops": round(lloyd_flops / max(t_mx - t_seed, 1e-9) / 1e12, 2)}
        log("k-means K=4096: seeding %.3fs; 3 Lloyd iterations exact %.3fs, mixed %.3fs "
            "(identical: %s)" % (t_seed, km["lloyd_fp32_s"], km["lloyd_bf16_s"], same))
        del xh64, c_ex, l_ex, c_mx, l_mx

    # ---------------- CPU baseline: the OpenMP oracle on a bounded sample
    cpu = None
    if args.cpu_seconds > 0:
        e = ix.export()
        oix = sco.OracleIndex(dim, 1, kc, mp, kp, cq.numpy(), pq, e["L"], e["pid"], e["codes"])
        threads = sco.nthreads()
        n_cal = min(64, args.nq)
        t0 = time.perf_counter()
        oids, od, osum = oix.search_batch(sco.SC, qh[:n_cal], args.R, args.w, args.T,
                                          threads=threads)
        dt = time.perf_counter() - t0
        ns = int(min(args.nq, max(n_cal, n_cal * args.cpu_seconds / max(dt, 1e-6))))
        t0 = time.perf_counter()
        oids, od, osum = oix.search_batch(sco.SC, qh[:ns], args.R, args.w, args.T,
                                          threads=threads)
        dt = time.perf_counter() - t0
        match = bool(np.array_equal(oids, gids[:ns]))
        cpu = {"value": round(ns / dt, 2), "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": "%d of the %d queries, same index, OpenMP over queries" % (ns, args.nq),
               "ids_match_gpu": match}
        log("cpu oracle: %.1f QPS on %d threads (%d queries), ids match: %s"
            % (ns / dt, threads, ns, match))

    # HBM bytes per launch of the same kernel and workload, measured with
    # rocprofv3 FETCH_SIZE/WRITE_SIZE and scaled by a plain copy of known size
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "search_fused_sc1M_traffic.json")
    if fused and os.path.exists(tpath):
        t = json.load(open(tpath))
        w = t.get("workload", {})
        if (w.get("n"), w.get("nq"), w.get("T"), w.get("R"), w.get("kc")) == \
                (args.n, args.nq, args.T, args.R, kc):
            traffic = int(t["traffic_bytes"])

    encode_flops = 2.0 * args.n * (kc * dim + mp * kp * (dim // mp))
    # MFMAs the BF16 encode issues: per 16 vectors, (K/16 codeword tiles) x
    # (block length / 16) 16x16x16 MFMAs, two per tile step for the coarse
    # block (integer rows are exact in BF16), three for the PQ blocks (the
    # residual rows are split into two BF16 terms); 8192 flops each
    mfma_issued = (args.n / 16.0) * ((kc // 16) * (dim // 16) * 2 +
                                     mp * (kp // 16) * max(1, (dim // mp) // 16) * 3) * 8192.0
    out = {
        "metric": "QPS at fixed recall@100, mixed-prec batched vs plain FP32 GPU, 1M×128",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Philox Gaussian mixture, integer-valued)",
        "config": {
            "workload": "dense-partition SC index, K=256 (nearest, second) cells over D=128, "
                        "PQ 8x256, %d base, %d queries, T=%d, top-%d; form: plain FP32, batched"
                        % (args.n, args.nq, args.T, args.R),
            "form": "plain-fp32-batched",
            "recall_at_100": recall,
            "ground_truth": gt_info,
            "qps_plain_fp32_per_query_calls": round(qps_single, 1) if qps_single else None,
            "mixed_precision_batched": mixed,
            "real_distance": real,
            "imi_k1024": imi,
            "kmeans_k4096": km,
            "candidates_per_query": round(cand / args.nq, 1),
            "stage_ms": {"lut": round(st0.ms_lut, 4), "rank": round(st0.ms_rank, 4),
                         "traverse": round(st0.ms_traverse, 4), "scan_topk": round(ms_kernel, 4)},
            "fused_phase_ms_per_workgroup": dict(zip(
                ["lut", "sort", "traverse", "scan", "select"], phases)),
            "launch": ("hipGraph plan, autotuned: %d threads x %d workgroups, register variant "
                       "%d waves/SIMD" % (tune["threads"], tune["grid"], tune["variant"]))
                      if plan else "sc_search calls",
            "qps_recall_curve": curve,
            "encode_s": round(t_encode, 4),
            "encode_bf16_s": round(t_encode16, 4),
            "encode_bf16_identical": enc16_same,
            "encode_fp64_tflops": round(encode_flops / t_encode / 1e12, 2),
            "encode_bf16_mfma": {
                "algorithmic_tflops": round(encode_flops / t_encode16 / 1e12, 2),
                "issued_tflops": round(mfma_issued / t_encode16 / 1e12, 2),
                "peak_tflops": 2500.0,
                "frac_issued": round(mfma_issued / t_encode16 / 2.5e15, 4),
                "rechecked_per_vector": round(est16.rechecked / args.n, 2),
                "full_exact_rows": int(est16.fallback)},
            "train_s": round(t_train, 2),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_search_fused" if fused else "k_scan_topk",
                     "algorithmic_bytes_per_launch": int(scan_bytes),
                     "traffic_source": "profiles/search_fused_sc1M_hbm_counters.md" if traffic else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if plan:
        plan.close()
    ix.close()
    ctx.close()


if __name__ == "__main__":
    main()
