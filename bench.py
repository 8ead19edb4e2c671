"""IVF-PQ search benchmark on one MI355X.

Workload (default): 10M x 128 base from the seeded Zipf (s=1.1) Gaussian
mixture, IVF with 4096 lists, PQ m=16 x 8 bit, 10,000 queries, w=32, k=100,
plain exact FP32 form.  One step = one search of the whole 10k-query batch
(coarse assignment, probe selection, fused table + ADC scan + top-k, merge)
with queries, index and outputs resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ivfpq_10m_l4096_m16")
    ap.add_argument("--nbase", type=int, default=0, help="override base size (smaller runs)")
    ap.add_argument("--cpu-queries", type=int, default=10000, help="queries in the CPU-baseline sample")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed passes of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gt-queries", type=int, default=10000, help="queries with an FP64 ground truth")
    ap.add_argument("--fast-kl", type=int, default=0, help="shortlist of the reduced-precision form (0: sweep --fast-kls)")
    ap.add_argument("--fast-kls", default="128,160,176,192", help="shortlist sizes of the reduced-precision sweep")
    ap.add_argument("--fast-seed", type=int, default=1, help="lists scanned exactly to seed its filter")
    ap.add_argument("--fast-seed-cap", type=int, default=-1,
                    help="vectors entering the seed bound (0 = whole seed lists, -1 = library default)")
    ap.add_argument("--fast-w", type=int, default=0,
                    help="probes of the reduced-precision form (0 = the workload's w); its recall@100 is "
                         "reported beside the plain form's")
    ap.add_argument("--no-fast", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--batched-cap", type=int, default=0, help="candidate buffer per query of the batched form "
                                                                   "(0 = library default)")
    ap.add_argument("--serve-fracs", default="0.25,0.5,0.75", help="offered load as fractions of the batch QPS")
    ap.add_argument("--serve-reqs", type=int, default=12000)
    ap.add_argument("--profile-only", action="store_true",
                    help="run the timed steps of the search forms and a 1 GiB device copy "
                         "(counter calibration), then exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    import mestrado_amd as M
    from mestrado_amd.workloads import WORKLOADS, D, NCOMP, SIGMA, STREAM_TRAIN, STREAM_BASE, STREAM_QUERY

    cfg = dict(WORKLOADS[args.workload])
    if args.nbase:
        cfg["nbase"] = args.nbase
        cfg["ntrain"] = min(cfg["ntrain"], max(cfg["nbase"] // 10, cfg["nlist"] * 8))
    seed = cfg["seed"]
    t0 = time.time()
    cent = M.synth_centres(seed, NCOMP, D)
    thr = M.zipf_thresholds(seed, NCOMP, cfg["zipf"]) if cfg["zipf"] else None
    train = M.synth_rows(seed, STREAM_TRAIN, 0, cfg["ntrain"], cent, thr, SIGMA)
    base = M.synth_rows(seed, STREAM_BASE, 0, cfg["nbase"], cent, thr, SIGMA)
    query = M.synth_rows(seed, STREAM_QUERY, 0, cfg["nquery"], cent, thr, SIGMA)
    log(f"data generated on device in {time.time() - t0:.1f}s: train {cfg['ntrain']} base {cfg['nbase']}")

    t0 = time.time()
    coarse, pq = M.train(train, cfg["nlist"], cfg["m"], 256, cfg["niter_coarse"], cfg["niter_pq"], seed)
    log(f"trained in {time.time() - t0:.1f}s")
    train.free()
    t0 = time.time()
    ix = M.Index(coarse, pq)
    ix.add(base)
    off, ids_l, codes_l = ix.get_lists()
    log(f"encoded {ix.ntotal} vectors in {time.time() - t0:.1f}s")

    nq, w, k, m = cfg["nquery"], cfg["w"], cfg["k"], cfg["m"]
    out_ids = M.DeviceArray((nq, k), np.int32)
    out_dis = M.DeviceArray((nq, k), np.float32)

    # algorithmic scan bytes: codes of every probed list, once per (query, probe)
    probes, _ = ix.coarse(query, w)
    lens = np.diff(off)
    scan_vectors = int(lens[probes].sum())
    scan_bytes = scan_vectors * m

    for _ in range(args.warmup):
        ix.search(query, w, k, out=(out_ids, out_dis))
    M.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ix.search(query, w, k, out=(out_ids, out_dis))
    M.synchronize()
    elapsed = time.perf_counter() - t_start
    # per-stage times from HIP events in a separate pass (a profiled search
    # waits for its last event, so the timed loop above runs without them)
    ix.set_option("profile", 1)
    stage = np.zeros(8)
    for _ in range(args.steps):
        ix.search(query, w, k, out=(out_ids, out_dis))
        stage += ix.timings()
    ix.set_option("profile", 0)

    # recall against the exact FP64 nearest neighbour
    ng = min(args.gt_queries, nq)
    if not args.profile_only:
        qsub = M.DeviceArray.from_numpy(query.numpy()[:ng])
        gt, gt_d = M.bruteforce_nn(base, qsub)
        base_h = base.numpy()  # host copy for the FP64 distances of returned ids (untimed)
        q64 = qsub.numpy().astype(np.float64)

    def recall(ids_h, dis_h):
        if args.profile_only:
            return {}
        r = {f"recall@{R}": round(100.0 * float(np.mean([gt[i] in ids_h[i, :R] for i in range(ng)])), 3)
             for R in (1, 10, 100) if R <= k}
        # distance error against FP64: (1) the returned distance of each
        # returned id vs its FP64 squared distance, (2) the first result's
        # FP64 distance vs the exact nearest neighbour's
        ne = min(ng, 1000)
        ids_s, dis_s = ids_h[:ne], dis_h[:ne].astype(np.float64)
        ok = ids_s >= 0
        rows = base_h[np.where(ok, ids_s, 0)].astype(np.float64)
        d64 = ((rows - q64[:ne, None, :]) ** 2).sum(axis=2)
        rel = np.abs(dis_s - d64)[ok] / np.maximum(d64[ok], 1e-30)
        first = d64[:, 0][ok[:, 0]]
        gap = (first - gt_d[:ne][ok[:, 0]]) / np.maximum(gt_d[:ne][ok[:, 0]], 1e-30)
        r["dist_rel_err_vs_fp64"] = {"mean": float(f"{rel.mean():.3g}"), "max": float(f"{rel.max():.3g}")}
        r["top1_dist_gap_vs_fp64_nn"] = {"mean": float(f"{gap.mean():.3g}"), "max": float(f"{gap.max():.3g}")}
        return r

    rec = recall(out_ids.numpy(), out_dis.numpy())
    qps = nq * args.steps / elapsed

    fast = None
    if not args.no_fast:
        # The reduced-precision form is reported at the plain form's recall:
        # probes w' and shortlist sizes kl in a short sweep, the fastest
        # whose recall@100 is within 0.5 points of the plain form's (or the
        # --fast-w / --fast-kl given).
        kls = [args.fast_kl] if args.fast_kl else [int(x) for x in args.fast_kls.split(",")]
        if args.profile_only:
            kls = kls[:1]
        f_ids = M.DeviceArray((nq, k), np.int32)
        f_dis = M.DeviceArray((nq, k), np.float32)
        ix.set_option("fast_seed_lists", args.fast_seed)
        if args.fast_seed_cap >= 0:
            ix.set_option("fast_seed_cap", args.fast_seed_cap)
        ix.attach_base(base)  # list-ordered copy of the raw vectors (index build, untimed)
        if args.fast_w:
            sweep_w = [args.fast_w]
        elif args.profile_only:
            sweep_w = [w]
        else:
            sweep_w = sorted({max(1, w - d) for d in (0, 4, 8, 10, 11, 12, 13, 14, 15, 16)}, reverse=True)
        sweep = []
        for kl, fw in [(kl, fw) for kl in kls for fw in sweep_w]:
            for _ in range(args.warmup):
                ix.search_fast(query, fw, kl, k, out=(f_ids, f_dis))
            M.synchronize()
            t_f = time.perf_counter()
            for _ in range(args.steps):
                ix.search_fast(query, fw, kl, k, out=(f_ids, f_dis))
            M.synchronize()
            f_elapsed = time.perf_counter() - t_f
            ix.set_option("profile", 1)  # stage times and candidate counts, separate pass
            fstage = np.zeros(8)
            for _ in range(args.steps):
                ix.search_fast(query, fw, kl, k, out=(f_ids, f_dis))
                fstage += ix.timings()
            ix.set_option("profile", 0)
            fstage /= args.steps
            cands, over = ix.fast_stats()
            f = {"qps": round(nq * args.steps / f_elapsed, 1), "ms_per_step": round(f_elapsed * 1000 / args.steps, 4),
                 "w": fw, "kl": kl, "seed_lists": args.fast_seed, "candidates_per_query": round(cands / nq, 1),
                 "overflowed_queries": over,
                 "stages_ms": {n: round(float(v), 4) for n, v in
                               zip(("coarse", "seed", "pairs", "scan8", "select", "rerank"), fstage)}}
            f.update(recall(f_ids.numpy(), f_dis.numpy()))
            sweep.append(f)
            log(f"fast form w={fw} kl={kl}: qps {f['qps']} recall@100 {f.get('recall@100')}")
        ok = [f for f in sweep if f.get("recall@100", 0.0) >= rec.get("recall@100", 0.0) - 0.5]
        fast = dict(max(ok, key=lambda f: f["qps"]) if ok else sweep[0])
        fast["recall_rule"] = "fastest (probes, shortlist) whose recall@100 is within 0.5 points of the plain form's"
        fast["sweep"] = [{"w": f["w"], "kl": f["kl"], "qps": f["qps"], "recall@100": f.get("recall@100")} for f in sweep]
        log(f"fast form: {fast}")
    uneven = None
    if not args.no_uneven:
        if args.batched_cap:
            ix.set_option("batched_capacity", args.batched_cap)
        # batched uneven exact search: per-query w ~ U{1..64}, k ~ {1, 10, 100}
        rng = np.random.default_rng(seed)
        wq = rng.integers(1, 65, nq).astype(np.int32)
        kq = rng.choice([1, 10, 100], nq).astype(np.int32)
        v_ids = M.DeviceArray((nq, 100), np.int32)
        v_dis = M.DeviceArray((nq, 100), np.float32)
        b_ids = M.DeviceArray((nq, 100), np.int32)
        b_dis = M.DeviceArray((nq, 100), np.float32)
        for _ in range(args.warmup):
            ix.search_var(query, wq, kq, 100, out=(v_ids, v_dis))
        M.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ix.search_var(query, wq, kq, 100, out=(v_ids, v_dis))
        M.synchronize()
        t_var = (time.perf_counter() - t0) / args.steps
        plan = M.Plan(ix, nq, 64, 100, autotune=True)
        for _ in range(max(1, args.warmup)):
            plan.search(query, wq, kq, out=(b_ids, b_dis))
        M.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.search(query, wq, kq, out=(b_ids, b_dis))
        M.synchronize()
        t_b = (time.perf_counter() - t0) / args.steps
        same = bool(np.array_equal(v_ids.numpy(), b_ids.numpy()) and
                    np.array_equal(v_dis.numpy().view(np.uint32), b_dis.numpy().view(np.uint32)))
        uneven = {"w": "U{1..64}", "k": "{1,10,100}", "plain_var_qps": round(nq / t_var, 1),
                  "batched_graph_qps": round(nq / t_b, 1), "speedup": round(t_var / t_b, 3),
                  "bit_identical_to_plain": same, "overflowed_queries": ix.batched_stats(),
                  **{f"tuned_{k2}": v for k2, v in plan.info().items()}}
        plan.close()
        log(f"batched uneven: {uneven}")
        # the same batched exact form on the headline batch (uniform w and k)
        wu = np.full(nq, w, np.int32)
        ku = np.full(nq, k, np.int32)
        u_ids = M.DeviceArray((nq, k), np.int32)
        u_dis = M.DeviceArray((nq, k), np.float32)
        plan = M.Plan(ix, nq, w, k, autotune=True)
        for _ in range(max(1, args.warmup)):
            plan.search(query, wu, ku, out=(u_ids, u_dis))
        M.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.search(query, wu, ku, out=(u_ids, u_dis))
        M.synchronize()
        t_u = (time.perf_counter() - t0) / args.steps
        same_u = bool(np.array_equal(out_ids.numpy(), u_ids.numpy()) and
                      np.array_equal(out_dis.numpy().view(np.uint32), u_dis.numpy().view(np.uint32)))
        uneven["uniform_headline"] = {"w": w, "k": k, "batched_graph_qps": round(nq / t_u, 1),
                                      "ms_per_step": round(t_u * 1000, 4),
                                      "algorithmic_scan_GBps_whole_search": round(scan_bytes / t_u / 1e9, 1),
                                      "bit_identical_to_plain": same_u, "overflowed_queries": ix.batched_stats(),
                                      **{f"tuned_{k2}": v for k2, v in plan.info().items()}}
        info = plan.info()
        plan.close()
        # its scan kernel's time (HIP events, same work-item shape as the plan)
        ix.set_option("batched_group", info["group"])
        ix.set_option("batched_threads", info["threads"])
        ix.set_option("profile", 1)
        bst = np.zeros(8)
        for _ in range(args.steps):
            ix.search_batched(query, wu, ku, k, out=(u_ids, u_dis))
            bst += ix.timings()
        ix.set_option("profile", 0)
        bst /= args.steps
        b_scan = float(bst[2])
        uneven["uniform_headline"]["stages_ms"] = {n: round(float(v), 4) for n, v in
                                                   zip(("coarse", "seed_grouping", "scan", "select"), bst)}
        uneven["uniform_headline"]["scan_kernel_roofline"] = {
            "kernel": "exact_scan_kernel", "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved": round(scan_bytes / (b_scan * 1e-3) / 1e9, 1) if b_scan > 0 else None,
            "frac": round(scan_bytes / (b_scan * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if b_scan > 0 else None,
            "algorithmic_bytes_per_launch": scan_bytes,
            "note": "codes of every (query, probed list) pair, as for the plain scan; each list is read once "
                    "per group of queries"}
        log(f"batched uniform: {uneven['uniform_headline']}")

    serving = None
    if not args.no_serve and not args.profile_only:
        # open-loop Poisson requests (nq log-uniform 1..1024, mean ~148) against the
        # continuous-batching server, exact form, w and k of the headline
        qps_peak = nq * args.steps / elapsed
        mean_nq = 147.6
        pts = []
        for frac in [float(x) for x in args.serve_fracs.split(",")]:
            st = M.serve_bench(ix, query, rate=frac * qps_peak / mean_nq, nreq=args.serve_reqs, seed=11,
                               bmax=1024, w_choices=(w,), k_choices=(k,), slots=3, max_batch=1024)
            pts.append({"load": frac, "offered_qps": round(st["offered_qps"], 1), "qps": round(st["qps"], 1),
                        "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                        "mean_ms": round(st["mean_ms"], 3), "batches": int(st["batches"])})
            log(f"serving at {frac:.2f} of peak: {pts[-1]}")
        serving = {"form": "plain exact FP32 (per-request w, k)", "w": w, "k": k, "slots": 3, "max_batch": 1024,
                   "requests_per_point": args.serve_reqs, "batch_sizes": "log-uniform 1..1024",
                   "arrivals": "Poisson (open loop)", "points": pts}
    if args.profile_only:
        # a plain device-to-device copy of known size: the reference point
        # that scales the HBM byte counters of the search kernels
        src = M.DeviceArray((1 << 28,), np.float32)
        dst = M.DeviceArray((1 << 28,), np.float32)
        M.lib().mivf_memcpy_dtod(dst.ptr, src.ptr, src.nbytes)
        M.synchronize()
        log(f"profile-only: {args.steps} steps, {elapsed * 1000 / args.steps:.3f} ms/step, copy 1 GiB")
        return
    ms_per_step = elapsed * 1000.0 / args.steps
    stage /= args.steps
    qps = nq * args.steps / elapsed
    scan_ms = float(stage[2])
    achieved = scan_bytes / (scan_ms * 1e-3) / 1e9 if scan_ms > 0 else None
    # HBM bytes per launch of the scan kernel from the FETCH_SIZE / WRITE_SIZE
    # counter runs recorded for this workload (null when none is recorded)
    def recorded_traffic(fname, kernel):
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{fname}_{args.workload}.md")
        if os.path.exists(path) and not args.nbase:
            for line in open(path):
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return int(line.split(":")[1])
        return None

    traffic = recorded_traffic("counters", "scan_kernel")
    rb = (uneven or {}).get("uniform_headline", {}).get("scan_kernel_roofline")
    if rb is not None:
        rb["traffic"] = recorded_traffic("batched_exact_scan", "exact_scan_kernel")

    if fast is not None:
        fast["speedup_vs_plain"] = round(fast["qps"] / qps, 3)

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import pyoracle as O
        ncpu = len(os.sched_getaffinity(0))
        nc = min(args.cpu_queries, nq)
        qh = query.numpy()[:nc]
        index = (coarse, pq, off, ids_l, codes_l)
        O.search_phases(index, qh[: min(64, nc)], w, k, nthreads=ncpu)  # warm-up
        times, phases = [], []
        for _ in range(max(1, args.cpu_reps)):
            t0 = time.perf_counter()
            r_ids, r_dis, ph = O.search_phases(index, qh, w, k, nthreads=ncpu)
            times.append(time.perf_counter() - t0)
            phases.append(ph)
        dt = float(np.median(times))
        tot = {n: sum(p[n] for p in phases) for n in phases[0]}
        allp = sum(tot.values())
        cpu = {"value": round(nc / dt, 2), "unit": "queries/s", "cores": ncpu, "kind": "port",
               "sample": f"first {nc} of the {nq} queries, same index, w={w}, k={k}, median of "
                         f"{len(times)} passes ({sum(times):.1f} s of CPU work), OpenMP over queries "
                         f"(oracle restatement of the parallel/ OpenMP search loop: per probe a table, "
                         f"sumidxtab2 into a concatenated array, then one k-selection)",
               "phase_split": {n: round(v / allp, 4) for n, v in tot.items()}}
        log(f"cpu baseline: {cpu}")
        # the oracle's results against the GPU's plain output on the same queries
        g_ids, g_dis = out_ids.numpy()[:nc], out_dis.numpy()[:nc]
        parity = {"queries": nc, "w": w, "k": k,
                  "bit_identical": bool(np.array_equal(g_ids, r_ids) and
                                        np.array_equal(g_dis.view(np.uint32), r_dis.view(np.uint32))),
                  "rows_differing": int(((g_ids != r_ids).any(1) |
                                         (g_dis.view(np.uint32) != r_dis.view(np.uint32)).any(1)).sum())}
        log(f"parity vs oracle: {parity}")

    res = {
        "metric": "QPS at fixed recall@100 (10M x128 IVF4096,PQ16); ADC-scan HBM GB/s; p99 ms",
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
        "data": "synthetic (seeded Gaussian mixture, integer-valued float32)",
        "config": {"workload": args.workload, "nbase": cfg["nbase"], "nlist": cfg["nlist"], "m": m,
                   "nquery": nq, "w": w, "k": k, "form": "plain exact FP32",
                   "kmeans_iterations": {"coarse": cfg["niter_coarse"], "pq": cfg["niter_pq"]}, **rec},
        "stages_ms": {"coarse": round(float(stage[0]), 4), "select": round(float(stage[1]), 4),
                      "scan": round(scan_ms, 4), "merge": round(float(stage[3]), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic,
                     "kernel": "scan_kernel", "algorithmic_bytes_per_launch": scan_bytes},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        # the same exact arithmetic, list-major: the batched form's scan kernel on the headline batch
        "roofline_batched_exact": (uneven or {}).get("uniform_headline", {}).get("scan_kernel_roofline"),
        "fast_form": fast,
        "batched_uneven": uneven,
        "serving": serving,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
