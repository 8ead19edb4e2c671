/*
 * mivf.h -- C ABI of libmivf, the MI355X (gfx950) IVF-PQ search engine.
 *
 * Plain C types only: pointers, sizes and status codes.  Every function
 * returns MIVF_OK (0) or a non-zero status; mivf_last_error() gives the
 * message of the last failure on the calling thread.  Pointers named d_* are
 * device (HBM) pointers; all others are host pointers.  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Each entry point cites the reference interface it replaces (paths are in the
 * reference repository rafaelms101/mestrado).
 *
 * Threading: calls on one index handle must be serialised by the caller (the
 * handle owns a scratch workspace); different handles are independent.
 */
#ifndef MIVF_H
#define MIVF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MIVF_OK = 0,
    MIVF_EINVAL = 1,      /* bad argument or shape                  */
    MIVF_EHIP = 2,        /* HIP runtime failure                    */
    MIVF_ENOMEM = 3,      /* allocation failure / over budget       */
    MIVF_EIO = 4,         /* file missing, truncated or corrupt     */
    MIVF_EUNSUPPORTED = 5 /* shape outside the compiled kernel set  */
};

typedef struct mivf_index mivf_index;

/* Message of the last failure on this thread ("" if none). */
const char *mivf_last_error(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int mivf_version(void);

/* ---- device memory helpers (for callers without a GPU framework) -------- */
int mivf_device_count(int *count);
int mivf_malloc(void **d_ptr, size_t bytes);
int mivf_free(void *d_ptr);
int mivf_memcpy_htod(void *d_dst, const void *src, size_t bytes);
int mivf_memcpy_dtoh(void *dst, const void *d_src, size_t bytes);
int mivf_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes);
int mivf_synchronize(void);

/* ---- training ------------------------------------------------------------
 * Replaces ivfpq_new(coarsek, nsq, vtrain, threads)
 * (parallel/ivf_pq/ivf_training.cpp:60-90; sequential/ivf_pq/ivf_new.cpp:4):
 * coarse k-means with nlist centroids, residuals, then one k-means with ks
 * centroids per subquantizer.  Assignment steps run on the GPU with the
 * reference's distance arithmetic; centroid updates are double-precision
 * sums in point order (deterministic).
 * x: host [n][d] float32.  coarse_out: [nlist][d]; pq_out: [m][ks][d/m].   */
int mivf_train(const float *x, int64_t n, int d, int nlist, int m, int ks,
               int niter_coarse, int niter_pq, uint64_t seed,
               float *coarse_out, float *pq_out);
/* same with x already in HBM */
int mivf_train_device(const float *d_x, int64_t n, int d, int nlist, int m, int ks,
                      int niter_coarse, int niter_pq, uint64_t seed,
                      float *coarse_out, float *pq_out);
/* One k-means (the building block of mivf_train); x host [n][ldx]. */
int mivf_kmeans(const float *x, int64_t n, int d, int ldx, int k, int niter,
                uint64_t seed, float *cent_out);

/* ---- index ----------------------------------------------------------------
 * Replaces preallocate_gpu_mem(pqtipo, ivf_t*, int)
 * (parallel2/ivf_pq/mycuda.h:8): the quantizers and all lists become
 * device-resident and stay so until mivf_index_free
 * (deallocate_gpu_mem, parallel2/ivf_pq/mycuda.h:9).
 * coarse: host [nlist][d]; pq: host [m][ks][d/m]; ks <= 256 (8-bit codes).  */
int mivf_index_create(int d, int nlist, int m, int ks, const float *coarse,
                      const float *pq, mivf_index **out);
int mivf_index_free(mivf_index *ix);
int mivf_index_info(const mivf_index *ix, int *d, int *nlist, int *m, int *ks,
                    int64_t *ntotal);
/* Replace all lists.  offsets [nlist+1] (vector offsets), ids [ntotal],
 * codes [ntotal][m] uint8, list-major.  Host arrays.                       */
int mivf_index_set_lists(mivf_index *ix, const int64_t *offsets, const int32_t *ids,
                         const uint8_t *codes);
/* Same with the reference's int32-per-subquantizer codes (ivf_t.codes,
 * parallel2/ivf_pq/myIVF.h:29-33); every code must be < ks.                */
int mivf_index_set_lists_i32(mivf_index *ix, const int64_t *offsets, const int32_t *ids,
                             const int32_t *codes);
/* Copy the lists back to the host (same layout as set_lists).             */
int mivf_index_get_lists(mivf_index *ix, int64_t *offsets, int32_t *ids, uint8_t *codes);
/* Replaces ivfpq_assign(ivfpq, vbase, ivf) (parallel/ivf_pq/ivf_training.cpp:165-224
 * and parallel/pq-utils/pq_assign.cpp:12-58): coarse argmin, residual and
 * per-subquantizer argmin on the GPU, then a stable append to the lists
 * (ids id0, id0+1, ... in input order; ids ascending within a list).
 * mivf_index_add takes host x, mivf_index_add_device takes d_x.            */
int mivf_index_add(mivf_index *ix, const float *x, int64_t n, int32_t id0);
int mivf_index_add_device(mivf_index *ix, const float *d_x, int64_t n, int32_t id0);
/* Encode only: d_assign [n] int32, d_codes [n][m] uint8 (device).          */
int mivf_index_encode(mivf_index *ix, const float *d_x, int64_t n, int32_t *d_assign,
                      uint8_t *d_codes, void *stream);

/* ---- search ----------------------------------------------------------------
 * Plain exact form: ids and FP32 distances bit-identical to the reference
 * arithmetic; results sorted by (distance, id); missing results are padded
 * with id -1 and distance +inf.  Replaces the sequential ivfpq_search
 * without rerank and the per-rank search loop of parallel_search
 * (parallel/ivf_pq/ivf_search.cpp:97-137).  d_q [nq][d]; outputs [nq][k].
 * Batches of at least "list_major_min_queries" (default 2048) 16-byte
 * aligned queries run on the list-major kernels of mivf_search_batched
 * (each list read once per group of queries; the same results bit for bit);
 * option "list_major" 0 keeps every batch on the per-query kernels.        */
int mivf_search(mivf_index *ix, const float *d_q, int nq, int w, int k,
                int32_t *d_ids, float *d_dis, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mivf_search with its scratch (coarse distance matrix,
 * probes, residual rows, partial top-k lists) carved out of one fixed device
 * array of 64 MiB per GPU, held by the library until the process ends,
 * instead of the index's workspace.  Same results as mivf_search.  It
 * allocates nothing, so (1) the lists must already be on the device: after
 * mivf_index_add / set_lists, one call of any other search uploads them
 * (MIVF_EINVAL until then), and (2) a batch needing more than 64 MiB is
 * rejected with MIVF_ENOMEM (per query about 4 nlist + w (4 d + 8 m + 8)
 * bytes, plus 8 w k when nq < 2048 splits the probes over workgroups; e.g.
 * about 1,000 queries at nlist 4096, w 32, k 100).  Every call on one GPU
 * uses the same array and leaves its contents there: calls that overlap on
 * one GPU (other streams, other threads) race and must be serialised.      */
int mivf_search_noalloc(mivf_index *ix, const float *d_q, int nq, int w, int k,
                        int32_t *d_ids, float *d_dis, void *stream);
/* Per-query w and k (host arrays w[nq], k[nq]); output rows kmax wide.     */
int mivf_search_var(mivf_index *ix, const float *d_q, int nq, const int32_t *w,
                    const int32_t *k, int32_t *d_ids, float *d_dis, int kmax,
                    void *stream);
/* Coarse assignment only: top-w lists of each query, (dist, id) order, as
 * knn_full(2, nq, nlist, d, w, coa, q) in parallel/ivf_pq/ivf_assign.cpp:28. */
int mivf_coarse(mivf_index *ix, const float *d_q, int nq, int w, int32_t *d_idx,
                float *d_dis, void *stream);
/* Residual-input search with core_gpu semantics
 * (parallel2/ivf_pq/mycuda.cu:119-188): d_residual [nq*w][d] and
 * d_list_of [nq*w] (rid_to_ivf); outputs [nq][k].                          */
int mivf_search_residual(mivf_index *ix, const float *d_residual,
                         const int32_t *d_list_of, int nq, int w, int k,
                         int32_t *d_ids, float *d_dis, void *stream);
/* Per-list search of the parallel/ fork, ivfpq_search(ivf_t*, float*,
 * pqtipo, int, double*, double*) (parallel/ivf_pq/ivf_search.cpp:178-228):
 * the m x ks table of one residual d_residual [d] (compute_cross_distances,
 * :206-209) and the ADC distance of each of n codes d_codes [n][m] (uint8,
 * sumidxtab2, :451-468) into d_dis [n], in code order, with the index's PQ.
 * Every code must be < ks (all are when ks = 256).                         */
int mivf_adc(mivf_index *ix, const float *d_residual, const uint8_t *d_codes, int64_t n, float *d_dis,
             void *stream);
/* Replaces my_k_min(dis_t, int, float*, int*) (parallel/ivf_pq/k_min.cpp:3-73)
 * by an exact selection: the k smallest (distance, id) pairs of
 * d_dis/d_ids [n], ascending, ties by lower id, into d_out_ids/d_out_dis [k];
 * entries past n are (-1, +inf).  k <= 1024.  (The reference's heap returns
 * other than the k smallest, e.g. 1, 2, 6 ... for the input 1 ... 7.)       */
int mivf_kmin(const float *d_dis, const int32_t *d_ids, int64_t n, int k, int32_t *d_out_ids, float *d_out_dis,
              void *stream);
/* RERANK path (sequential/ivf_pq/ivf_search.cpp:98-134): ADC shortlist of kl,
 * then exact FP32 distances to the raw base rows d_base [nbase][d], best k
 * by (distance, id).  Distances are returned too.                          */
int mivf_search_rerank(mivf_index *ix, const float *d_q, int nq, int w, int kl, int k,
                       const float *d_base, int64_t nbase, int32_t *d_ids,
                       float *d_dis, void *stream);

/* Keep a copy of the raw vectors d_base [nbase][d] in HBM in list order, so
 * that the rerank of mivf_search_fast reads each query's shortlist from at
 * most w contiguous regions (pass d_base = NULL there to use it).  The copy
 * is dropped when the lists change (add / set_lists).  With option
 * "base_bytes" = 1 the copy holds one byte per value (a quarter of the HBM
 * reads of the rerank); the call then fails with MIVF_EINVAL unless every
 * value is an integer in 0..255 (as data read from bvecs files), so that the
 * bytes convert back to the same floats and the rerank returns the same
 * distances bit for bit.  No reference counterpart: the reference reranks
 * from vbase in host memory.                                                */
int mivf_index_attach_base(mivf_index *ix, const float *d_base, int64_t nbase);

/* Batched reduced-precision form with exact rerank (no reference
 * counterpart; the reference's RERANK path is mivf_search_rerank):
 * (1) coarse top-w, (2) an exact scan of each query's nearest lists (at least
 * "fast_seed_lists", then as many as hold kl vectors) gives the kl-th exact
 * ADC distance tau_q, (3) one persistent kernel scans the other
 * (list, query) pairs list-major, 16 queries per work item, with 8-bit
 * tables (one 16-byte LDS gather serves all 16 queries of a vector) and keeps
 * every vector whose approximate distance can be <= tau_q, (4) the kl best by
 * (approximate distance, list position) are reranked exactly in FP32
 * against d_base (or the attached copy when d_base is NULL).  d_q must be
 * 16-byte aligned.  Returned distances are exact; results sorted by
 * (distance, id).                                                          */
int mivf_search_fast(mivf_index *ix, const float *d_q, int nq, int w, int kl, int k,
                     const float *d_base, int64_t nbase, int32_t *d_ids, float *d_dis,
                     void *stream);
/* Candidates kept by the last profiled mivf_search_fast, and the number of
 * queries whose candidates exceeded the per-query capacity
 * max("fast_capacity" (default 65536), 64 kl); the excess is dropped, which
 * lowers recall and makes that query's shortlist depend on arrival order.
 * The buffer takes nq x capacity x 8 bytes of HBM (5.2 GB at 10k queries
 * and the default); only the kept candidates are written.                 */
int mivf_fast_stats(mivf_index *ix, int64_t *candidates, int64_t *overflowed_queries);

/* The 8-bit tables of mivf_search_fast for queries d_q[0..n-1] (n <= 16)
 * against list `list`: d_t8 [m][256][16] uint8 (entry [s][code][query]) and
 * d_delta [16] (approximate distance = delta * sum of entries).  For tests. */
int mivf_fast_tables(mivf_index *ix, const float *d_q, int n, int list, uint8_t *d_t8,
                     float *d_delta);

/* Test hook of the FP32-built table types (option "fast_tables" 1 = FP16
 * entries, 2 = 8-bit entries with a per-subquantizer offset and scale): the
 * approximate ADC distance of every vector of list `list` for queries
 * d_q[0..n-1] (n <= 16), d_out [len][16], computed by the scan's own loop.  */
int mivf_fast_var_adc(mivf_index *ix, const float *d_q, int n, int list, int tables, float *d_out);

/* The seed thresholds of mivf_search_fast: d_tau [nq] >= the kl-th smallest
 * ADC distance over each query's seed lists (+inf when they hold fewer than
 * kl vectors).  For tests.                                                  */
int mivf_fast_seed_tau(mivf_index *ix, const float *d_q, int nq, int w, int kl, float *d_tau);

/* The shortlist step of mivf_search_fast on its own: for each q < nq, the
 * ids (low 32 bits) of the kl smallest of the keys d_cand[q*capq + i],
 * i < min(d_cnt[q], capq), in any order, padded with -1 into d_short [nq][kl].
 * Keys are (order-preserving distance bits << 32) | position.  For tests.   */
int mivf_fast_select(const uint64_t *d_cand, const int32_t *d_cnt, int capq, int nq, int kl,
                     int32_t *d_short);

/* ---- options and timing ---------------------------------------------------
 * "probes_per_group" (0 = automatic), "profile" (1 = record per-stage HIP
 * event timings of the next searches), "query_order" (1 = scan batches of
 * >= 2048 queries grouped by nearest list; schedule only, default 1),
 * "fast_seed_lists" (lists per query scanned exactly to seed
 * mivf_search_fast, default 1), "fast_seed_cap" (vectors of those lists
 * that enter the seed bound, default 2048, 0 = all), "fast_seed_slack"
 * (8-bit table quanta above the seed bound that still pass the filter of
 * mivf_search_fast, 0..64, default 2; 3m/4 + 2 keeps every vector whose
 * ADC distance is within the seed bound), "fast_capacity",
 * "fast_running_bound" (1 =
 * the 8-bit scan tightens each query's filter as its candidates accumulate;
 * results unchanged, default 1), "batched_group", "batched_threads",
 * "batched_capacity" and "batched_capacity_per_probe" (candidates per query
 * of the list-major exact kernels: at least the first, default 16384, and
 * at least the second times the query batch's largest w, default 256),
 * "list_major" (1 = mivf_search
 * sends large batches to the list-major kernels, default 1),
 * "fast_tables" (table type of mivf_search_fast: 0 = 8-bit entries built on
 * the FP16 matrix cores with one scale per (query, list) and int8
 * matrix-core sums, the default; 1 = FP16 entries; 2 = 8-bit entries with a
 * per-subquantizer offset and scale; 1 and 2 are built from FP32 tables and
 * summed in FP32), "list_major_min_queries" (default 2048), "coarse_mfma"
 * (1 = coarse top-w through BF16 matrix-core distances and the exact
 * recheck, for d = 128 and nlist <= 32768; 0 = exact FP32 distance tiles for
 * every shape; the same probes either way), "exact_seed_cap" (vectors of
 * the nearest lists that set each query's bound in the list-major exact
 * kernels: default 4096, 0 = the whole lists; results unchanged),
 * "base_bytes" (1 = mivf_index_attach_base keeps a byte copy of an
 * integer-valued base, default 0), "fast_front_ranks" (n > 0 = the 8-bit
 * scan takes the work items of each query's n nearest lists first, so that
 * their candidates set the running bounds early; results unchanged; default
 * 0: the extra, less full work items cost more than the candidates saved),
 * "item_order" (1 = the list-major kernels number their work items longest
 * list first; a schedule only, results unchanged; default 0),
 * "exact_running_bound" (1 = the list-major exact scan lowers each query's
 * bound as its kept distances accumulate, from histograms shared by the
 * workgroups; 0 = off; 2 = on when a batch expects more than 3,000
 * candidates per query, w * mean list length * k / exact_seed_cap, the
 * default; results unchanged either way: it costs histogram traffic and
 * saves candidates, overflow reruns and select time),
 * "serve_min_fill" and "serve_max_wait_us" (servers over this index: while
 * a micro-batch is in
 * flight, a free slot launches only once serve_min_fill queries wait or the
 * oldest waiting request has waited serve_max_wait_us; defaults 1024 and
 * 300; serve_min_fill 0 launches as soon as a slot is free).               */
int mivf_set_option(mivf_index *ix, const char *name, int64_t value);
/* Per-stage milliseconds of the last profiled search:
 * [0] coarse distances, [1] probe select, [2] scan, [3] merge (mivf_search on
 * the per-query kernels);
 * [0] coarse, [1] seed scan, [2] wait for the pair grouping (which runs on a
 * second stream beside the seed), [3] 8-bit scan, [4] shortlist, [5] rerank
 * (mivf_search_fast); [0] coarse, [1] probe masking + seed + grouping,
 * [2] exact scan, [3] select (mivf_search_batched, and mivf_search on a
 * batch it sends to the list-major kernels).                                */
int mivf_get_timings(mivf_index *ix, float *ms, int n);

/* ---- batched exact form ----------------------------------------------------
 * Per-query w and k (host arrays w[nq], k[nq], k[i] <= kmax), results
 * identical to mivf_search_var (ids and FP32 distances, rows kmax wide),
 * computed list-major: the (list, query) pairs of the batch are grouped by
 * list into work items of G queries; one persistent kernel builds the G exact
 * FP32 tables of an item in LDS and scans the list once for all of them,
 * keeping the vectors within a per-query bound of the k-th distance (from the
 * query's nearest lists); a select keeps the k best.  Replaces, for a batch
 * with uneven w and k, the per-query loop of parallel_search
 * (parallel/ivf_pq/ivf_search.cpp:97-137).  Options "batched_group" (G),
 * "batched_threads", "batched_capacity".  d_q must be 16-byte aligned.      */
int mivf_search_batched(mivf_index *ix, const float *d_q, int nq, const int32_t *w, const int32_t *k,
                        int kmax, int32_t *d_ids, float *d_dis, void *stream);
/* Queries of the last mivf_search_batched / mivf_plan_search whose
 * candidates exceeded the buffer ("batched_capacity") and were recomputed
 * by the plain kernels.                                                     */
int mivf_batched_stats(mivf_index *ix, int64_t *overflowed_queries);
/* A reusable plan for batches of exactly nq queries with w <= wmax, k <= kmax:
 * the first mivf_plan_search runs the autotuner (when asked: every compiled
 * work-item shape timed on that batch), then captures the whole launch
 * sequence in a hipGraph; later calls replay the graph.  Results are those of
 * mivf_search_batched.                                                       */
typedef struct mivf_plan mivf_plan;
int mivf_plan_create(mivf_index *ix, int nq, int wmax, int kmax, int autotune, mivf_plan **out);
int mivf_plan_search(mivf_plan *p, const float *d_q, const int32_t *w, const int32_t *k, int32_t *d_ids,
                     float *d_dis, void *stream);
/* Test hooks of the batched exact kernel: its FP32 tables of queries
 * d_q[0..n-1] (n <= 8) against list `list`, d_out [m][256][8]; and the ADC
 * distances its scan loop computes for every vector of the list,
 * d_out [len][8] (m = 16).                                                  */
int mivf_exact_tables(mivf_index *ix, const float *d_q, int n, int list, float *d_out);
int mivf_exact_adc(mivf_index *ix, const float *d_q, int n, int list, float *d_out);
/* Chosen work-item shape and the autotuner's best time (ms, 0 if not tuned). */
int mivf_plan_info(const mivf_plan *p, int *g, int *threads, float *tuned_ms);
int mivf_plan_free(mivf_plan *p);

/* ---- serving loop (continuous batching) -----------------------------------
 * Replaces the search stage of the reference's MPI query pipeline (query
 * sender -> search ranks -> aggregator, parallel/ivf_test.cpp) on one GPU.
 * Requests (device queries d_q [nq][d] with their own w and k, device outputs
 * [nq][k]) wait in arrival order; up to `slots` micro-batches run at once,
 * each on its own stream with its own workspace, packed from the waiting
 * requests up to max_batch queries and within mem_budget bytes of per-request
 * device state (0 = 4 GiB).  Searches are exact: a micro-batch of at least
 * the index's list_major_min_queries queries runs the list-major kernels of
 * mivf_search_batched (m in {8, 16}), a smaller one the
 * per-query kernels of mivf_search_var; both give the same rows, so a
 * request's results are the same alone and in any mix.  The server is driven
 * by the calling thread: submit, then progress / wait / drain.                */
typedef struct mivf_server mivf_server;
int mivf_server_create(mivf_index *ix, int slots, int max_batch, int64_t mem_budget, mivf_server **out);
int mivf_server_free(mivf_server *sv);
int mivf_server_submit(mivf_server *sv, const float *d_q, int nq, int w, int k, int32_t *d_ids,
                       float *d_dis, int64_t *ticket);
int mivf_server_progress(mivf_server *sv, int *completed);
int mivf_server_wait(mivf_server *sv, int64_t ticket);
int mivf_server_drain(mivf_server *sv);
/* Submission-to-completion time of a finished request, milliseconds.       */
int mivf_server_latency(mivf_server *sv, int64_t ticket, double *ms);
/* Grow every slot's workspaces before serving, so that no later micro-batch
 * allocates device memory: each slot runs, untimed, micro-batches of w and k
 * (the largest the server will see) at max_batch queries and, when the
 * list-major kernels apply, at list_major_min_queries - 1, reading query rows
 * from d_rows [n_rows][d] (fewer if n_rows is smaller).  Only while no request
 * waits or runs; tickets issued earlier become invalid; counters restart.    */
int mivf_server_warm(mivf_server *sv, const float *d_rows, int64_t n_rows, int w, int k);
/* Counters since creation (or the last warm): micro-batches launched, those that ran the
 * list-major kernels, and queries of those rerun by the per-query kernels
 * because their candidates overflowed (any pointer may be NULL).           */
int mivf_server_stats(mivf_server *sv, int64_t *batches, int64_t *list_major_batches,
                      int64_t *rerun_queries);
/* Open-loop benchmark: nreq requests, Poisson arrivals at `rate` requests/s,
 * nq log-uniform in 1..bmax, w and k drawn from the given choices, queries at
 * random rows of d_pool [pool_n][d].  stats[10] = achieved queries/s, p50 ms,
 * p99 ms, mean ms, queries served, makespan ms, micro-batches, offered
 * queries/s, list-major micro-batches, queries rerun after a candidate
 * overflow (latency counted from each request's scheduled arrival).  Before
 * the clock starts, mivf_server_warm grows the slots' workspaces for the
 * largest w and k.                                                          */
int mivf_serve_bench(mivf_index *ix, int slots, int max_batch, int64_t mem_budget, const float *d_pool,
                     int64_t pool_n, double rate, int nreq, uint64_t seed, int bmax,
                     const int32_t *w_choices, int nw, const int32_t *k_choices, int nk, double *stats);

/* ---- calibration ------------------------------------------------------------
 * LDS read rate of the current device in GB/s: every lane of 512-thread
 * workgroups (two per CU) reads `bytes` (4 or 16) per step from a 64 KiB
 * table laid out as the scan kernels' tables, at random rows (random_rows = 1:
 * data-driven table gathers, bank conflicts included) or at rows on distinct
 * banks (0: the conflict-free peak).  The ceiling of the scan kernels' table
 * reads; no reference counterpart.                                         */
int mivf_lds_read_rate(int bytes, int random_rows, double *gbps);

/* ---- synthetic data and ground truth ------------------------------------ */
/* Seeded Gaussian mixture rows; bit-identical to the CPU generator.
 * d_centres [ncomp][d]; d_thr NULL (uniform components) or [ncomp] Zipf
 * thresholds from mivf_zipf_thresholds.                                     */
int mivf_synth_rows(uint64_t seed, uint64_t stream_id, int64_t row0, int64_t n, int d,
                    const float *d_centres, int ncomp, const uint64_t *d_thr,
                    float sigma, float *d_out, void *stream);
int mivf_synth_centres(uint64_t seed, int ncomp, int d, float *centres);
int mivf_zipf_thresholds(uint64_t seed, int ncomp, double s, uint64_t *thr);
/* Exact FP64 nearest neighbour of each query (ties -> lowest id).          */
int mivf_bruteforce_nn(const float *d_base, int64_t nb, const float *d_q, int nq, int d,
                       int32_t *d_ids, double *d_dis);

/* ---- reference on-disk formats ---------------------------------------------
 * Centroid files: write_cent/read_cent (parallel/ivf_pq/ivf_training.cpp:92-143):
 * header = 8 int32 {nsq, ks, ds, ks*ds, nsq, coarsek, coarsek, d}, cent =
 * float32 [nsq][ks][ds], coa = float32 [coarsek][d].                        */
int mivf_write_cent(const char *header, const char *cent, const char *coa, int nsq, int ks,
                    int ds, int coarsek, int d, const float *pq, const float *coarse);
int mivf_read_cent_header(const char *header, int *nsq, int *ks, int *ds, int *coarsek, int *d);
int mivf_read_cent(const char *header, const char *cent, const char *coa, float *pq,
                   float *coarse);
/* Per-list files <dir>/<j> (parallel2/genivf.cpp:62-86, read_ivf
 * parallel2/ivf_pq/ivf_search.cpp:293-321) and the single-file layout
 * (parallel/ivf_pq/ivf_search.cpp:396-441).  Codes are int32 on disk and
 * uint8 in memory.  Readers always fill offsets [coarsek+1]; pass NULL
 * ids/codes to learn the sizes first; cap = capacity of ids in vectors.    */
int mivf_write_ivf_dir(const char *dir, int coarsek, int m, const int64_t *offsets,
                       const int32_t *ids, const uint8_t *codes);
int mivf_read_ivf_dir(const char *dir, int coarsek, int m, int64_t *offsets, int32_t *ids,
                      uint8_t *codes, int64_t cap);
int mivf_write_ivf_file(const char *path, int coarsek, int m, const int64_t *offsets,
                        const int32_t *ids, const uint8_t *codes);
int mivf_read_ivf_file(const char *path, int coarsek, int m, int64_t *offsets, int32_t *ids,
                       uint8_t *codes, int64_t cap);

/* Vector files: fvecs_read / b2fvecs_read (parallel/yael_needs/vector.c:296-350),
 * ivecs_read and my_bvecs_read (parallel/pq-utils/pq_test_load_vectors.cpp:236-325).
 * Each record is int32 d, then d float32 (fvecs), int32 (ivecs) or uint8
 * (bvecs) values.  vecs_info reads the dimension of the first record and the
 * record count (elem_bytes 4 for fvecs/ivecs, 1 for bvecs; an empty file
 * gives d = n = 0).  The readers copy records [row0, row0 + n) into out
 * [n][d] (bvecs values widened to float32, as my_bvecs_read does); a record
 * of another dimension, or a file shorter than row0 + n records, gives
 * MIVF_EIO.                                                                  */
int mivf_vecs_info(const char *path, int elem_bytes, int *d, int64_t *n);
int mivf_read_fvecs(const char *path, int64_t row0, int64_t n, int d, float *out);
int mivf_read_bvecs(const char *path, int64_t row0, int64_t n, int d, float *out);
int mivf_read_ivecs(const char *path, int64_t row0, int64_t n, int d, int32_t *out);
/* Replaces the load-then-ivfpq_assign loop of the reference's index builders
 * (parallel2/genivf.cpp:33-60, parallel/ivf_pq/ivf_training.cpp:165-224):
 * records [row0, row0 + n) of an fvecs (elem_bytes 4) or bvecs (elem_bytes
 * 1) file of the index's dimension are read `chunk` records at a time
 * (chunk <= 0: 2^20) and added as mivf_index_add would, ids id0, id0+1, ...
 * Host memory holds one chunk; the lists end exactly as one mivf_index_add
 * of all n rows would leave them.                                           */
int mivf_index_add_file(mivf_index *ix, const char *path, int elem_bytes, int64_t row0, int64_t n,
                        int32_t id0, int64_t chunk);

#ifdef __cplusplus
}
#endif
#endif /* MIVF_H */
