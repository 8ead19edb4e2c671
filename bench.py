"""IVF-PQ search benchmark on one MI355X.

Workload (default): 10M x 128 base from the seeded Zipf (s=1.1) Gaussian
mixture, IVF with 4096 lists, PQ m=16 x 8 bit, 10,000 queries, w=32, k=100,
plain exact FP32 form (ids and distances bit-identical to the reference
arithmetic).  One step = one search of the whole 10k-query batch with
queries, index and outputs resident in HBM; at this batch size mivf_search
runs the list-major kernels (coarse top-w, seed bounds, list-major grouping,
exact scan, select).  The per-query kernels are timed beside it, then the
reduced-precision form (operating point chosen on held-out queries), the
batched uneven form, serving, and the CPU baseline (the oracle on the host
cores), whose results are compared with the GPU's.  Prints one JSON line.
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


def apply_opts(ix, opts):
    for o in opts:
        name, value = o.split("=", 1)
        ix.set_option(name, int(value))
        log(f"option {name} = {value}")


def same_bits(a_ids, a_dis, b_ids, b_dis):
    return bool(np.array_equal(a_ids, b_ids) and np.array_equal(a_dis.view(np.uint32), b_dis.view(np.uint32)))


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
    ap.add_argument("--tune-queries", type=int, default=10000,
                    help="held-out queries (the next rows of the query stream) that choose the "
                         "reduced-precision operating point")
    ap.add_argument("--fast-kl", type=int, default=0, help="shortlist of the reduced-precision form (0: sweep --fast-kls)")
    ap.add_argument("--fast-kls", default="128,160,176,192,224", help="shortlist sizes of the reduced-precision sweep")
    ap.add_argument("--fast-seed", type=int, default=1, help="lists scanned exactly to seed its filter")
    ap.add_argument("--fast-seed-cap", type=int, default=-1,
                    help="vectors entering the seed bound (0 = whole seed lists, -1 = library default)")
    ap.add_argument("--fast-w", type=int, default=0, help="probes of the reduced-precision form (0 = sweep)")
    ap.add_argument("--no-fast", action="store_true")
    ap.add_argument("--fast-float-rows", action="store_true",
                    help="rerank the reduced-precision form from the float32 copy of the base instead of the "
                         "byte copy (the data are integers 0..255, so both give the same distances)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--batched-cap", type=int, default=0, help="candidate buffer per query of the batched form "
                                                                   "(0 = library default)")
    ap.add_argument("--serve-fracs", default="0.25,0.5,0.75", help="offered load as fractions of the batch QPS")
    ap.add_argument("--serve-reqs", type=int, default=12000)
    ap.add_argument("--serve-max-batch", type=int, default=4096, help="queries per serving micro-batch at most")
    ap.add_argument("--serve-slots", type=int, default=3, help="serving micro-batches in flight")
    ap.add_argument("--precision-study", action="store_true",
                    help="the reduced-precision table types on this workload's index (w=32, shortlist 1024, "
                         "k=100): recall and distance error against FP64 per table type")
    ap.add_argument("--opt", action="append", default=[], metavar="NAME=VALUE",
                    help="index option (mivf_set_option) applied before the searches; repeatable")
    ap.add_argument("--quantizer-cache", default="",
                    help="npz file of the trained coarse and PQ centroids: loaded when it exists, else written "
                         "after training (repeated runs on one workload skip the k-means)")
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
    if cfg.get("stream_chunk"):
        return streamed_serving(args, cfg)
    if args.precision_study:
        return precision_study(args, cfg)
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
    if args.quantizer_cache and os.path.exists(args.quantizer_cache):
        qc = np.load(args.quantizer_cache)
        coarse, pq = qc["coarse"], qc["pq"]
        log(f"quantizers loaded from {args.quantizer_cache}")
    else:
        coarse, pq = M.train(train, cfg["nlist"], cfg["m"], 256, cfg["niter_coarse"], cfg["niter_pq"], seed)
        log(f"trained in {time.time() - t0:.1f}s ({cfg['niter_coarse']} coarse, {cfg['niter_pq']} PQ k-means "
            f"iterations)")
        if args.quantizer_cache:
            np.savez(args.quantizer_cache, coarse=coarse, pq=pq)
    train.free()
    t0 = time.time()
    ix = M.Index(coarse, pq)
    apply_opts(ix, args.opt)
    ix.add(base)
    off, ids_l, codes_l = ix.get_lists()
    log(f"encoded {ix.ntotal} vectors in {time.time() - t0:.1f}s")

    nq, w, k, m = cfg["nquery"], cfg["w"], cfg["k"], cfg["m"]
    lens = np.diff(off)
    probes, _ = ix.coarse(query, w)
    # algorithmic scan bytes: codes of every probed list, once per (query, probe)
    scan_vectors = int(lens[probes].sum())
    scan_bytes = scan_vectors * m
    # FP32 table entries read by an exact scan: 4 bytes per (vector, query, subquantizer)
    table_bytes = scan_vectors * m * 4

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        M.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn()
        M.synchronize()
        return (time.perf_counter() - t) / args.steps

    def stages(fn):
        # per-stage times from HIP events in a separate pass (a profiled search
        # waits for its last event, so the timed loops run without them)
        ix.set_option("profile", 1)
        st = np.zeros(8)
        for _ in range(args.steps):
            fn()
            st += ix.timings()
        ix.set_option("profile", 0)
        return st / args.steps

    # ---- plain exact form: mivf_search (list-major kernels at this batch size)
    out_ids = M.DeviceArray((nq, k), np.int32)
    out_dis = M.DeviceArray((nq, k), np.float32)
    ix.set_option("list_major", 1)
    plain_fn = lambda: ix.search(query, w, k, out=(out_ids, out_dis))  # noqa: E731
    t_plain = timed(plain_fn)
    st_plain = stages(plain_fn)
    # ---- the same results from the per-query kernels (option list_major = 0)
    pq_ids = M.DeviceArray((nq, k), np.int32)
    pq_dis = M.DeviceArray((nq, k), np.float32)
    ix.set_option("list_major", 0)
    pq_fn = lambda: ix.search(query, w, k, out=(pq_ids, pq_dis))  # noqa: E731
    t_pq = timed(pq_fn)
    st_pq = stages(pq_fn)
    ix.set_option("list_major", 1)
    out_h = (out_ids.numpy(), out_dis.numpy())
    paths_same = same_bits(*out_h, pq_ids.numpy(), pq_dis.numpy())
    log(f"plain: {nq / t_plain:.0f} q/s (list-major), {nq / t_pq:.0f} q/s (per-query kernels), same {paths_same}")

    # ---- the LDS read ceilings of this device (the scan kernels' table reads)
    lds = {"random16": M.lds_read_rate(16, True), "peak16": M.lds_read_rate(16, False),
           "random4": M.lds_read_rate(4, True), "peak4": M.lds_read_rate(4, False)}
    log(f"LDS read rates GB/s: {lds}")

    # ---- recall against the exact FP64 nearest neighbour
    ng = min(args.gt_queries, nq)
    if not args.profile_only:
        qsub = M.DeviceArray.from_numpy(query.numpy()[:ng])
        gt, gt_d = M.bruteforce_nn(base, qsub)
        base_h = base.numpy()  # host copy for the FP64 distances of returned ids (untimed)
        q64 = qsub.numpy().astype(np.float64)

    def recall_of(ids_h, g):
        n = g.shape[0]
        return {f"recall@{R}": round(100.0 * float((ids_h[:n, :R] == g[:, None]).any(1).mean()), 3)
                for R in (1, 10, 100) if R <= k}

    def recall(ids_h, dis_h):
        if args.profile_only:
            return {}
        r = recall_of(ids_h, gt)
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

    rec = recall(*out_h)

    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs
    # recorded for this workload (null when none is recorded)
    def recorded(fname, key):
        path = os.path.join(ROOT, "profiles", f"{fname}_{args.workload}.md")
        if os.path.exists(path) and not args.nbase:
            for line in open(path):
                if line.startswith(f"{key}:"):
                    return float(line.split(":", 1)[1])
        return None

    def rf(x, nd=4):
        return None if x is None else round(x, nd)

    # ---- reduced-precision form + exact rerank, operating point chosen on
    # held-out queries (the next rows of the query stream)
    fast = None
    if not args.no_fast:
        kls = [args.fast_kl] if args.fast_kl else [int(x) for x in args.fast_kls.split(",")]
        if args.fast_w:
            sweep_w = [args.fast_w]
        elif args.profile_only:
            sweep_w = [w]
        else:
            sweep_w = sorted({max(1, w - d) for d in (0, 4, 8, 10, 11, 12, 13, 14, 15, 16)}, reverse=True)
        if args.profile_only:
            kls = kls[:1]
        ix.set_option("fast_seed_lists", args.fast_seed)
        if args.fast_seed_cap >= 0:
            ix.set_option("fast_seed_cap", args.fast_seed_cap)
        # list-ordered copy of the raw vectors (index build, untimed); one byte
        # per value, exact for this integer-valued data
        if not args.fast_float_rows:
            ix.set_option("base_bytes", 1)
        ix.attach_base(base)
        f_ids = M.DeviceArray((nq, k), np.int32)
        f_dis = M.DeviceArray((nq, k), np.float32)
        sweep = []
        if len(kls) * len(sweep_w) > 1 and not args.profile_only:
            ntune = args.tune_queries
            qt = M.synth_rows(seed, STREAM_QUERY, nq, ntune, cent, thr, SIGMA)
            gtt, _ = M.bruteforce_nn(base, qt)
            t_ids = M.DeviceArray((ntune, k), np.int32)
            t_dis = M.DeviceArray((ntune, k), np.float32)
            ix.search(qt, w, k, out=(t_ids, t_dis))
            plain_tune = recall_of(t_ids.numpy(), gtt)["recall@100"]
            for kl, fw in [(kl, fw) for kl in kls for fw in sweep_w]:
                tt = timed(lambda: ix.search_fast(qt, fw, kl, k, out=(t_ids, t_dis)))
                r = recall_of(t_ids.numpy(), gtt)
                sweep.append({"w": fw, "kl": kl, "qps": round(ntune / tt, 1), "recall@100": r["recall@100"]})
                log(f"fast form (tuning queries) w={fw} kl={kl}: {ntune / tt:.0f} q/s recall@100 {r['recall@100']}")
            ok = [f for f in sweep if f["recall@100"] >= plain_tune - 0.5]
            pick = max(ok, key=lambda f: f["qps"]) if ok else max(sweep, key=lambda f: f["recall@100"])
            fw, kl = pick["w"], pick["kl"]
            qt.free()
        else:
            fw, kl, plain_tune = sweep_w[0], kls[0], None
        fast_fn = lambda: ix.search_fast(query, fw, kl, k, out=(f_ids, f_dis))  # noqa: E731
        t_fast = timed(fast_fn)
        fst = stages(fast_fn)
        cands, over = ix.fast_stats()
        rows_note = None
        if not args.fast_float_rows:
            # the same operating point reranked from the float32 copy: same results, 4x the row bytes
            h_ids, h_dis = f_ids.numpy(), f_dis.numpy()
            ix.set_option("base_bytes", 0)
            ix.attach_base(base)
            t_f32 = timed(fast_fn)
            f32_same = same_bits(h_ids, h_dis, f_ids.numpy(), f_dis.numpy())
            ix.set_option("base_bytes", 1)
            ix.attach_base(base)
            rows_note = {"rerank_rows": "bytes (every base value is an integer in 0..255: the bytes convert to the "
                                        "same floats)", "qps_float32_rows": round(nq / t_f32, 1),
                         "float32_rows_bit_identical": f32_same}
        fast = {"qps": round(nq / t_fast, 1), "ms_per_step": round(t_fast * 1000, 4), "w": fw, "kl": kl,
                "seed_lists": args.fast_seed, "candidates_per_query": round(cands / nq, 1),
                "overflowed_queries": over,
                "stages_ms": {n: round(float(v), 4) for n, v in
                              zip(("coarse", "seed", "pairs", "scan8", "select", "rerank"), fst)}}
        if rows_note:
            fast.update(rows_note)
        fast.update(recall(f_ids.numpy(), f_dis.numpy()))
        if sweep:
            fast["operating_point_rule"] = (
                "fastest (probes, shortlist) whose recall@100 on the held-out tuning queries (rows "
                f"{nq}..{nq + args.tune_queries - 1} of the query stream) is within 0.5 points of the plain "
                "form's on the same queries; qps and recall above are then measured on the evaluation queries")
            fast["plain_recall@100_tuning"] = plain_tune
            fast["tuning_sweep"] = sweep
        # roofline of its dominant kernel (fast_scan_kernel): the list-major
        # (list, <= 16 queries) items past each query's seed lists
        fprobes, _ = ix.coarse(query, fw)
        plen = lens[fprobes]
        csum = np.cumsum(plen, axis=1)
        ns = np.maximum(args.fast_seed, (csum < kl).sum(1) + 1).clip(max=fw)
        mask = np.arange(fw)[None, :] >= ns[:, None]
        f_vectors = int(plen[mask].sum())
        per_list = np.bincount(fprobes[mask], minlength=cfg["nlist"])
        visits = int((np.ceil(per_list / 16.0) * lens).sum())
        f_ms = float(fst[3])
        fast["roofline"] = {
            "kernel": "fast_scan_kernel", "bound": "lds",
            "achieved": rf(visits * m * 16 / (f_ms * 1e-3) / 1e9, 1) if f_ms > 0 else None,
            "peak": round(lds["random16"], 1), "unit": "GB/s",
            "frac": rf(visits * m * 16 / (f_ms * 1e-3) / 1e9 / lds["random16"]) if f_ms > 0 else None,
            "lds_bytes_per_launch": visits * m * 16,
            "algorithmic_code_bytes_per_launch": f_vectors * m,
            "algorithmic_GBps_vs_hbm_peak": rf(f_vectors * m / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS)
            if f_ms > 0 else None,
            "int8_mfma_busy_recorded": recorded("counters", "mfma_busy fast_scan_kernel"),
            "unit_busy_recorded": {u: recorded("counters", f"{u}_busy fast_scan_kernel") for u in ("valu", "lds", "mfma")},
            "traffic": recorded("counters", "traffic_bytes_per_launch fast_scan_kernel"),
            "note": "one 16-byte LDS read per (vector, subquantizer) of a work item serves its 16 queries' "
                    "8-bit entries; the peak is this device's measured rate of random 16-byte LDS reads"}
        log(f"fast form: {fast}")

    # ---- batched uneven exact form
    uneven = None
    if not args.no_uneven:
        if args.batched_cap:
            ix.set_option("batched_capacity", args.batched_cap)
        rng = np.random.default_rng(seed)
        wq = rng.integers(1, 65, nq).astype(np.int32)
        kq = rng.choice([1, 10, 100], nq).astype(np.int32)
        v_ids = M.DeviceArray((nq, 100), np.int32)
        v_dis = M.DeviceArray((nq, 100), np.float32)
        b_ids = M.DeviceArray((nq, 100), np.int32)
        b_dis = M.DeviceArray((nq, 100), np.float32)
        t_var = timed(lambda: ix.search_var(query, wq, kq, 100, out=(v_ids, v_dis)))
        plan = M.Plan(ix, nq, 64, 100, autotune=True)
        plan.search(query, wq, kq, out=(b_ids, b_dis))  # autotune + capture
        t_b = timed(lambda: plan.search(query, wq, kq, out=(b_ids, b_dis)))
        uneven = {"w": "U{1..64}", "k": "{1,10,100}", "per_query_kernels_qps": round(nq / t_var, 1),
                  "batched_graph_qps": round(nq / t_b, 1), "speedup": round(t_var / t_b, 3),
                  "bit_identical_to_per_query": same_bits(v_ids.numpy(), v_dis.numpy(), b_ids.numpy(),
                                                          b_dis.numpy()),
                  "overflowed_queries": ix.batched_stats(),
                  **{f"tuned_{k2}": v for k2, v in plan.info().items()}}
        plan.close()
        log(f"batched uneven: {uneven}")
        # the headline batch through an autotuned graph plan
        wu = np.full(nq, w, np.int32)
        ku = np.full(nq, k, np.int32)
        u_ids = M.DeviceArray((nq, k), np.int32)
        u_dis = M.DeviceArray((nq, k), np.float32)
        plan = M.Plan(ix, nq, w, k, autotune=True)
        plan.search(query, wu, ku, out=(u_ids, u_dis))
        t_u = timed(lambda: plan.search(query, wu, ku, out=(u_ids, u_dis)))
        uneven["uniform_headline_plan"] = {"w": w, "k": k, "qps": round(nq / t_u, 1),
                                           "ms_per_step": round(t_u * 1000, 4),
                                           "bit_identical_to_plain": same_bits(*out_h, u_ids.numpy(),
                                                                               u_dis.numpy()),
                                           **{f"tuned_{k2}": v for k2, v in plan.info().items()}}
        plan.close()
        log(f"batched uniform plan: {uneven['uniform_headline_plan']}")

    # ---- serving: open-loop Poisson requests against the continuous-batching server
    serving = None
    if not args.no_serve and not args.profile_only:
        qps_peak = nq / t_plain
        mean_nq = 147.6  # mean of log-uniform 1..1024
        common = dict(bmax=1024, w_choices=(w,), k_choices=(k,), slots=args.serve_slots,
                      max_batch=args.serve_max_batch)
        sat = M.serve_bench(ix, query, rate=1e9, nreq=min(args.serve_reqs, 4000), seed=21, **common)
        log(f"serving saturated: {sat['qps']:.0f} q/s")
        pts = []
        for frac in [float(x) for x in args.serve_fracs.split(",")]:
            st = M.serve_bench(ix, query, rate=frac * qps_peak / mean_nq, nreq=args.serve_reqs, seed=11, **common)
            pts.append({"load": frac, "offered_qps": round(st["offered_qps"], 1), "qps": round(st["qps"], 1),
                        "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                        "mean_ms": round(st["mean_ms"], 3), "batches": int(st["batches"]),
                        "list_major_batches": int(st["list_major_batches"]),
                        "rerun_queries": int(st["rerun_queries"])})
            log(f"serving at {frac:.2f} of the batch QPS: {pts[-1]}")
        serving = {"form": "plain exact FP32 (per-request w, k)", "w": w, "k": k, "slots": args.serve_slots,
                   "max_batch": args.serve_max_batch, "requests_per_point": args.serve_reqs,
                   "batch_sizes": "log-uniform 1..1024", "arrivals": "Poisson (open loop)",
                   "load_reference_qps": round(qps_peak, 1),
                   "load_reference": "the plain headline batch QPS (value)",
                   "saturated_qps": round(sat["qps"], 1), "points": pts}

    if args.profile_only:
        # a plain device-to-device copy of known size: the reference point
        # that scales the HBM byte counters of the search kernels
        src = M.DeviceArray((1 << 28,), np.float32)
        dst = M.DeviceArray((1 << 28,), np.float32)
        M.lib().mivf_memcpy_dtod(dst.ptr, src.ptr, src.nbytes)
        M.synchronize()
        log(f"profile-only: {args.steps} steps, {t_plain * 1000:.3f} ms/step plain, copy 1 GiB")
        return

    # ---- CPU baseline: the oracle on the host cores, and parity of its results
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
        g_ids, g_dis = out_h[0][:nc], out_h[1][:nc]
        parity = {"queries": nc, "w": w, "k": k, "bit_identical": same_bits(g_ids, g_dis, r_ids, r_dis),
                  "rows_differing": int(((g_ids != r_ids).any(1) |
                                         (g_dis.view(np.uint32) != r_dis.view(np.uint32)).any(1)).sum()),
                  "per_query_kernels_bit_identical": paths_same}
        log(f"parity vs oracle: {parity}")

    qps = nq / t_plain
    scan_ms = float(st_plain[2])
    pq_scan_ms = float(st_pq[2])

    def hbm(ms):
        return scan_bytes / (ms * 1e-3) / 1e9 if ms > 0 else None

    def ldsrate(ms):
        return table_bytes / (ms * 1e-3) / 1e9 if ms > 0 else None

    def busy(kernel):
        # busy fractions of the kernel's units from the recorded counter runs,
        # and the busiest of them (the unit that binds)
        b = {u: recorded("counters", f"{u}_busy {kernel}") for u in ("valu", "lds", "mfma")}
        b = {u: v for u, v in b.items() if v is not None}
        return b, (max(b, key=b.get) if b else None)

    traffic = recorded("counters", "traffic_bytes_per_launch exact_scan_kernel")
    xb, xbind = busy("exact_scan_kernel")
    qb, qbind = busy("scan_kernel")
    roofline = {
        "bound": "hbm", "achieved": rf(hbm(scan_ms), 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": rf(hbm(scan_ms) / HBM_PEAK_GBS) if scan_ms > 0 else None,
        "traffic": int(traffic) if traffic else None,
        "kernel": "exact_scan_kernel", "algorithmic_bytes_per_launch": scan_bytes,
        "binding_unit": xbind, "unit_busy_recorded": xb,
        "lds": {"achieved": rf(ldsrate(scan_ms), 1), "peak": round(lds["random16"], 1), "unit": "GB/s",
                "frac": rf(ldsrate(scan_ms) / lds["random16"]) if scan_ms > 0 else None,
                "conflict_free_peak": round(lds["peak16"], 1), "bytes_per_launch": table_bytes},
        "note": "achieved = codes of every (query, probed list) pair / kernel time; the kernel reads each list "
                "once per work item of 4 queries (HBM traffic far below the algorithmic bytes); its binding unit "
                "is the busiest in the recorded counter runs (the VALU: exact FP32 tables and sums), the LDS "
                "(random 16-byte FP32 table reads) next: lds.peak is this device's measured rate of such reads, "
                "conflict_free_peak the same reads on distinct banks"}
    pq_traffic = recorded("counters", "traffic_bytes_per_launch scan_kernel")
    per_query = {
        "qps": round(nq / t_pq, 1), "ms_per_step": round(t_pq * 1000, 4),
        "stages_ms": {n: round(float(v), 4) for n, v in zip(("coarse", "select", "scan", "merge"), st_pq)},
        "roofline": {"bound": "hbm", "achieved": rf(hbm(pq_scan_ms), 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rf(hbm(pq_scan_ms) / HBM_PEAK_GBS) if pq_scan_ms > 0 else None,
                     "traffic": int(pq_traffic) if pq_traffic else None, "kernel": "scan_kernel",
                     "binding_unit": qbind, "unit_busy_recorded": qb,
                     "lds": {"achieved": rf(ldsrate(pq_scan_ms), 1), "peak": round(lds["random4"], 1),
                             "unit": "GB/s",
                             "frac": rf(ldsrate(pq_scan_ms) / lds["random4"]) if pq_scan_ms > 0 else None,
                             "conflict_free_peak": round(lds["peak4"], 1)}},
        "note": "the per-query kernels (one workgroup per query and probe group, one 4-byte table read per "
                "code byte): mivf_search below 2048 queries, or with option list_major = 0"}
    if fast is not None:
        fast["speedup_vs_plain"] = round(fast["qps"] / qps, 3)
        fast["speedup_vs_per_query_kernels"] = round(fast["qps"] * t_pq / nq, 3)

    res = {
        "metric": "QPS at fixed recall@100 (10M x128 IVF4096,PQ16); ADC-scan HBM GB/s; p99 ms",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_plain * 1000, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian mixture, integer-valued float32)",
        "config": {"workload": args.workload, "nbase": cfg["nbase"], "nlist": cfg["nlist"], "m": m,
                   "nquery": nq, "w": w, "k": k, "form": "plain exact FP32 (mivf_search: list-major kernels)",
                   "kmeans_iterations": {"coarse": cfg["niter_coarse"], "pq": cfg["niter_pq"]}, **rec},
        "stages_ms": {n: round(float(v), 4) for n, v in zip(("coarse", "seed_grouping", "scan", "select"), st_plain)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "per_query_kernels": per_query,
        "fast_form": fast,
        "batched_uneven": uneven,
        "serving": serving,
    }
    print(json.dumps(res), flush=True)


def precision_study(args, cfg):
    """The reduced-precision form on a bit-exact workload's index: BF16
    matrix-core coarse assignment with the FP32 recheck, then each table type
    (8-bit built on the FP16 matrix cores with one scale per (query, list);
    FP16 entries; 8-bit entries with a per-subquantizer offset and scale), a
    shortlist of 1024 per query and the exact FP32 rerank against the raw
    float32 base, at w = 32 and k = 100.  Recall@1/10/100 and the distance
    error are measured against the FP64 brute-force nearest neighbour of 1,000
    queries; QPS over the workload's whole query batch."""
    import mestrado_amd as M
    from mestrado_amd.workloads import D, NCOMP, SIGMA, STREAM_TRAIN, STREAM_BASE, STREAM_QUERY
    from oracle import pyoracle as O

    seed, nq = cfg["seed"], cfg["nquery"]
    w, kl, k = 32, 1024, 100
    t0 = time.time()
    cent = M.synth_centres(seed, NCOMP, D)
    thr = M.zipf_thresholds(seed, NCOMP, cfg["zipf"]) if cfg["zipf"] else None
    train = M.synth_rows(seed, STREAM_TRAIN, 0, cfg["ntrain"], cent, thr, SIGMA)
    coarse, pq = M.train(train, cfg["nlist"], cfg["m"], 256, cfg["niter_coarse"], cfg["niter_pq"], seed)
    train.free()
    base = M.synth_rows(seed, STREAM_BASE, 0, cfg["nbase"], cent, thr, SIGMA)
    query = M.synth_rows(seed, STREAM_QUERY, 0, nq, cent, thr, SIGMA)
    ix = M.Index(coarse, pq)
    apply_opts(ix, args.opt)
    ix.add(base)
    off, ids_l, codes_l = ix.get_lists()
    log(f"index built in {time.time() - t0:.1f}s")
    ng = min(args.gt_queries, 1000)
    qh = query.numpy()
    qg = M.DeviceArray.from_numpy(qh[:ng])
    gt, gt_d = M.bruteforce_nn(base, qg)
    base_h = base.numpy()
    q64 = qh[:ng].astype(np.float64)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        M.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn()
        M.synchronize()
        return (time.perf_counter() - t) / args.steps

    def quality(ids_h, dis_h):
        r = {f"recall@{R}": round(100.0 * float((ids_h[:ng, :R] == gt[:, None]).any(1).mean()), 3)
             for R in (1, 10, 100)}
        ids_s, dis_s = ids_h[:ng], dis_h[:ng].astype(np.float64)
        ok = ids_s >= 0
        rows = base_h[np.where(ok, ids_s, 0)].astype(np.float64)
        d64 = ((rows - q64[:, None, :]) ** 2).sum(axis=2)
        rel = np.abs(dis_s - d64)[ok] / np.maximum(d64[ok], 1e-30)
        gap = (d64[:, 0][ok[:, 0]] - gt_d[ok[:, 0]]) / np.maximum(gt_d[ok[:, 0]], 1e-30)
        r["returned_dist_rel_err_vs_fp64"] = {"mean": float(f"{rel.mean():.3g}"), "max": float(f"{rel.max():.3g}")}
        r["top1_dist_gap_vs_fp64_nn"] = {"mean": float(f"{gap.mean():.3g}"), "max": float(f"{gap.max():.3g}")}
        return r

    oi = M.DeviceArray((nq, k), np.int32)
    od = M.DeviceArray((nq, k), np.float32)
    forms = {}
    # the plain exact form (bit-exact to the reference arithmetic) at w = 32
    t = timed(lambda: ix.search(query, w, k, out=(oi, od)))
    forms["plain_exact"] = {"qps": round(nq / t, 1), **quality(oi.numpy(), od.numpy())}
    # the reference's RERANK path: exact ADC shortlist of 1024, exact FP32 rerank
    rq = M.DeviceArray.from_numpy(qh[:ng])
    t = timed(lambda: ix.search_rerank(rq, w, kl, k, base))
    r_ids, r_dis = ix.search_rerank(rq, w, kl, k, base)
    forms["exact_adc_rerank"] = {"qps_1000_queries": round(ng / t, 1), **quality(r_ids, r_dis)}
    log(f"plain {forms['plain_exact']}, exact-ADC rerank {forms['exact_adc_rerank']}")
    ix.attach_base(base)
    ix.set_option("fast_capacity", 1 << 16)
    names = {0: "int8_tables_per_list_scale_mfma", 1: "fp16_tables", 2: "int8_tables_per_subquantizer_offset_scale"}
    # the approximate ADC of each table type against the exact ADC, over 16
    # queries x the 4 longest lists
    lens = np.diff(off)
    sample_lists = [int(l) for l in np.argsort(-lens)[:4]]
    q16 = qh[:16]
    for tt, name in names.items():
        ix.set_option("fast_tables", tt)
        t = timed(lambda: ix.search_fast(query, w, kl, k, out=(oi, od)))
        ix.set_option("profile", 1)
        ix.search_fast(query, w, kl, k, out=(oi, od))
        cands, over = ix.fast_stats()
        st = ix.timings()
        ix.set_option("profile", 0)
        f = {"qps": round(nq / t, 1), "candidates_per_query": round(cands / nq, 1), "overflowed_queries": over,
             "stages_ms": {n: round(float(v), 4) for n, v in
                           zip(("coarse", "seed", "pairs", "scan", "select", "rerank"), st)}}
        f.update(quality(oi.numpy(), od.numpy()))
        rels = []
        for l in sample_lists:
            cl = codes_l[off[l]:off[l + 1]]
            if tt == 0:
                t8, dl = ix.fast_tables(q16, l)
            else:
                approx = ix.fast_var_adc(q16, l, len(cl), tt)
            for g in range(16):
                res = (q16[g] - coarse[l]).astype(np.float32)
                exact = O.adc(O.tables(res, pq), cl).astype(np.float64)
                if tt == 0:
                    ap_ = dl[g] * t8[np.arange(cfg["m"])[None, :], cl.astype(np.int64), g].astype(np.float64).sum(1)
                else:
                    ap_ = approx[:, g].astype(np.float64)
                rels.append(np.abs(ap_ - exact) / np.maximum(exact, 1e-30))
        rels = np.concatenate(rels)
        f["approx_adc_rel_err"] = {"mean": float(f"{rels.mean():.3g}"), "p99": float(f"{np.quantile(rels, 0.99):.3g}"),
                                   "max": float(f"{rels.max():.3g}"), "pairs": int(rels.size)}
        forms[name] = f
        log(f"{name}: {f}")
    ix.set_option("fast_tables", 0)
    # coarse assignment: BF16 matrix-core distances + FP32 recheck vs the
    # exact FP32 distance tiles
    pm, pmd = ix.coarse(qg, w)
    ix.set_option("coarse_mfma", 0)
    pe, ped = ix.coarse(qg, w)
    ix.set_option("coarse_mfma", 1)
    coarse_same = bool(np.array_equal(pm, pe) and np.array_equal(pmd.view(np.uint32), ped.view(np.uint32)))
    res = {
        "metric": "QPS at fixed recall@100 (10M x128 IVF4096,PQ16); ADC-scan HBM GB/s; p99 ms",
        "value": forms["plain_exact"]["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(nq / forms["plain_exact"]["qps"] * 1000, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian mixture, integer-valued float32)",
        "config": {"workload": args.workload, "study": "reduced-precision table types", "nbase": cfg["nbase"],
                   "nlist": cfg["nlist"], "m": cfg["m"], "nquery": nq, "w": w, "kl": kl, "k": k,
                   "quality_queries": ng, "coarse_bf16_recheck_equals_exact_fp32": coarse_same},
        "forms": forms,
    }
    print(json.dumps(res), flush=True)


def streamed_serving(args, cfg):
    """The serving workload: the base is generated and encoded chunk by chunk
    (ids are global row numbers, as parallel/ivf_pq/ivf_search.cpp:296-329
    fills its lists in 1M-vector chunks) and kept only as PQ codes; then the
    continuous-batching server takes open-loop Poisson requests at 5 rates,
    given as fractions of the server's own saturated throughput."""
    import mestrado_amd as M
    from mestrado_amd.workloads import D, NCOMP, SIGMA, STREAM_TRAIN, STREAM_BASE, STREAM_QUERY

    seed, chunk = cfg["seed"], cfg["stream_chunk"]
    if args.nbase:
        cfg["nbase"] = args.nbase
    nbase, nlist, m, w, k = cfg["nbase"], cfg["nlist"], cfg["m"], cfg["w"], cfg["k"]
    t0 = time.time()
    cent = M.synth_centres(seed, NCOMP, D)
    thr = M.zipf_thresholds(seed, NCOMP, cfg["zipf"]) if cfg["zipf"] else None
    train = M.synth_rows(seed, STREAM_TRAIN, 0, cfg["ntrain"], cent, thr, SIGMA)
    if args.quantizer_cache and os.path.exists(args.quantizer_cache):
        qc = np.load(args.quantizer_cache)
        coarse, pq = qc["coarse"], qc["pq"]
        log(f"quantizers loaded from {args.quantizer_cache}")
    else:
        coarse, pq = M.train(train, nlist, m, 256, cfg["niter_coarse"], cfg["niter_pq"], seed)
        log(f"trained {nlist} lists on {cfg['ntrain']} vectors in {time.time() - t0:.1f}s")
        if args.quantizer_cache:
            np.savez(args.quantizer_cache, coarse=coarse, pq=pq)
    train.free()
    t0 = time.time()
    ix = M.Index(coarse, pq)
    apply_opts(ix, args.opt)
    rows = M.DeviceArray((chunk, D), np.float32)
    for r0 in range(0, nbase, chunk):
        n = min(chunk, nbase - r0)
        M.synth_rows(seed, STREAM_BASE, r0, n, cent, thr, SIGMA, out=rows)
        if n == chunk:
            ix.add(rows, id0=r0)
        else:
            ix.add(M.DeviceArray.from_numpy(rows.head(n)), id0=r0)
        if (r0 // chunk) % 10 == 9:
            log(f"encoded {r0 + n} vectors ({time.time() - t0:.1f}s)")
    t_build = time.time() - t0
    log(f"{ix.ntotal} vectors encoded in {t_build:.1f}s, {ix.ntotal * (m + 4) / 1e9:.2f} GB of codes and ids")

    pool_n = cfg["serve_pool"]
    pool = M.synth_rows(seed, STREAM_QUERY, 0, pool_n, cent, thr, SIGMA)
    nq = cfg["nquery"]
    query = M.DeviceArray.from_numpy(pool.head(nq))
    oi = M.DeviceArray((nq, k), np.int32)
    od = M.DeviceArray((nq, k), np.float32)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        M.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn()
        M.synchronize()
        return (time.perf_counter() - t) / args.steps

    def stages(fn):
        ix.set_option("profile", 1)
        acc = np.zeros(8)
        for _ in range(args.steps):
            fn()
            acc += ix.timings()
        ix.set_option("profile", 0)
        return acc / args.steps

    plain_fn = lambda: ix.search(query, w, k, out=(oi, od))  # noqa: E731
    t_plain = timed(plain_fn)
    st = stages(plain_fn)
    log(f"plain batch: {nq / t_plain:.0f} q/s (list-major), stages {np.round(st[:4], 4)}")
    pi = M.DeviceArray((nq, k), np.int32)
    pd = M.DeviceArray((nq, k), np.float32)
    ix.set_option("list_major", 0)
    pq_fn = lambda: ix.search(query, w, k, out=(pi, pd))  # noqa: E731
    t_pq = timed(pq_fn)
    st_pq = stages(pq_fn)
    ix.set_option("list_major", 1)
    pq_same = same_bits(oi.numpy(), od.numpy(), pi.numpy(), pd.numpy())
    log(f"per-query kernels: {nq / t_pq:.0f} q/s, stages {np.round(st_pq[:4], 4)}, same {pq_same}")

    # recall of 1000 queries against the exact FP64 nearest neighbour over the
    # whole base, regenerated chunk by chunk (ties -> lowest id)
    ng = min(args.gt_queries, 1000)
    qg = M.DeviceArray.from_numpy(pool.head(ng))
    best_d = np.full(ng, np.inf)
    best_i = np.full(ng, -1, np.int64)
    for r0 in range(0, nbase, chunk):
        n = min(chunk, nbase - r0)
        M.synth_rows(seed, STREAM_BASE, r0, n, cent, thr, SIGMA, out=rows)
        src = rows if n == chunk else M.DeviceArray.from_numpy(rows.head(n))
        gi, gd = M.bruteforce_nn(src, qg)
        better = gd < best_d
        best_d[better] = gd[better]
        best_i[better] = gi[better] + r0
    ids_h = oi.numpy()[:ng]
    rec = {f"recall@{R}": round(100.0 * float((ids_h[:, :R] == best_i[:, None]).any(1).mean()), 3)
           for R in (1, 10, 100)}
    log(f"recall on {ng} queries: {rec}")

    # the server's saturated throughput: all requests arrive at once
    mean_nq = 147.6
    common = dict(bmax=1024, w_choices=tuple(cfg["serve_w"]), k_choices=tuple(cfg["serve_k"]),
                  slots=args.serve_slots, max_batch=args.serve_max_batch)
    sat = M.serve_bench(ix, pool, rate=1e9, nreq=min(args.serve_reqs, 4000), seed=21, **common)
    peak = sat["qps"]
    log(f"serving saturated: {peak:.0f} q/s")
    pts = []
    for frac in (0.1, 0.25, 0.5, 0.75, 0.9):
        sb = M.serve_bench(ix, pool, rate=frac * peak / mean_nq, nreq=args.serve_reqs, seed=11, **common)
        pts.append({"load": frac, "offered_qps": round(sb["offered_qps"], 1), "qps": round(sb["qps"], 1),
                    "p50_ms": round(sb["p50_ms"], 3), "p99_ms": round(sb["p99_ms"], 3),
                    "mean_ms": round(sb["mean_ms"], 3), "batches": int(sb["batches"]),
                    "list_major_batches": int(sb["list_major_batches"]),
                    "rerun_queries": int(sb["rerun_queries"])})
        log(f"serving at {frac:.2f} of saturation: {pts[-1]}")
    res = {
        "metric": "QPS at fixed recall@100 (10M x128 IVF4096,PQ16); ADC-scan HBM GB/s; p99 ms",
        "value": round(nq / t_plain, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_plain * 1000, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian mixture, integer-valued float32), streamed in chunks",
        "config": {"workload": args.workload, "nbase": nbase, "nlist": nlist, "m": m, "nquery": nq, "w": w, "k": k,
                   "form": "plain exact FP32 (mivf_search: list-major kernels)",
                   "stream_chunk": chunk, "build_s": round(t_build, 1), "options": args.opt,
                   "kmeans_iterations": {"coarse": cfg["niter_coarse"], "pq": cfg["niter_pq"]},
                   "recall_queries": ng, **rec},
        "stages_ms": {n: round(float(v), 4) for n, v in zip(("coarse", "seed_grouping", "scan", "select"), st)},
        "per_query_kernels": {"qps": round(nq / t_pq, 1), "ms_per_step": round(t_pq * 1000, 4),
                              "stages_ms": {n: round(float(v), 4)
                                            for n, v in zip(("coarse", "select", "scan", "merge"), st_pq)},
                              "bit_identical_to_list_major": pq_same},
        "serving": {"form": "plain exact FP32 (per-request w, k)", "w_choices": list(cfg["serve_w"]),
                    "k_choices": list(cfg["serve_k"]), "slots": args.serve_slots,
                    "max_batch": args.serve_max_batch,
                    "requests_per_point": args.serve_reqs, "batch_sizes": "log-uniform 1..1024",
                    "arrivals": "Poisson (open loop)", "query_pool": pool_n, "saturated_qps": round(peak, 1),
                    "points": pts},
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
