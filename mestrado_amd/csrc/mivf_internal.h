// Internal declarations shared by the HIP kernels and the C++ host layer of
// libmivf (MI355X IVF-PQ search).  Not part of the public C ABI.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>

namespace mivf {

// ---------------------------------------------------------------------------
// Errors: every host entry point returns an int status and keeps a
// thread-local message (read through mivf_last_error()).
// ---------------------------------------------------------------------------
enum Status : int {
    kOk = 0,
    kInvalid = 1,     // bad argument / shape
    kHip = 2,         // HIP runtime error
    kNoMem = 3,       // allocation failure or over budget
    kIo = 4,          // file missing, truncated or corrupt
    kUnsupported = 5  // shape outside the compiled kernel set
};

void set_error(const std::string &msg);
const char *last_error();

#define MIVF_CHECK_HIP(call)                                                     \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::mivf::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return ::mivf::kHip;                                                 \
        }                                                                        \
    } while (0)

#define MIVF_REQUIRE(cond, code, msg)  \
    do {                               \
        if (!(cond)) {                 \
            ::mivf::set_error(msg);    \
            return (code);             \
        }                              \
    } while (0)

// Handler list of the function-try-block around every C-ABI body: a C++
// exception (host allocation failure, above all) becomes a status code
// instead of ending the caller's process.
#define MIVF_CATCH_ALL                                                      \
    catch (const std::bad_alloc &) {                                        \
        ::mivf::set_error("out of host memory");                            \
        return ::mivf::kNoMem;                                              \
    }                                                                       \
    catch (const std::exception &e_) {                                      \
        ::mivf::set_error(std::string("internal error: ") + e_.what());     \
        return ::mivf::kInvalid;                                            \
    }

#define MIVF_TRY(expr)              \
    do {                            \
        int rc_ = (expr);           \
        if (rc_ != ::mivf::kOk)     \
            return rc_;             \
    } while (0)

// ---------------------------------------------------------------------------
// Order keys.  A candidate (distance, id) becomes one unsigned 64-bit key
// whose integer order is (distance ascending, id ascending), the order of the
// reference GPU top-k (parallel2/ivf_pq/topk.cu:26-32).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t fkey(float f) {
    f = f + 0.0f;  // -0 and +0 compare equal
    uint32_t b = __builtin_bit_cast(uint32_t, f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline uint64_t ckey(float d, int32_t id) {
    return (uint64_t(fkey(d)) << 32) | uint32_t(id);
}
__host__ __device__ inline float key_dist(uint64_t key) {
    uint32_t h = uint32_t(key >> 32);
    uint32_t b = (h & 0x80000000u) ? (h & 0x7FFFFFFFu) : ~h;
    return __builtin_bit_cast(float, b);
}
__host__ __device__ inline int32_t key_id(uint64_t key) {
    return int32_t(uint32_t(key & 0xFFFFFFFFu));
}
constexpr uint64_t kNoKey = ~uint64_t(0);

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip files).  All take a hipStream_t and
// return hipError_t from the launch.
// ---------------------------------------------------------------------------

// Row norms: f32 accumulator (a side) / f64 accumulator of f32 products (b side).
hipError_t launch_norm_f32(const float *x, int64_t n, int d, int ldx, float *out, hipStream_t s);
hipError_t launch_norm_f64(const float *x, int64_t n, int d, int ldx, double *out, hipStream_t s);

// Exact distances out[j*ldo + i] = dist(a_i, b_j) for i < na, j < nb.
hipError_t launch_dist_store(const float *a, int lda, const float *an, int na,
                             const float *b, int ldb, const double *bn, int64_t nb,
                             int d, float *out, int64_t ldo, hipStream_t s);
// Exact argmin: keys[j] = min_i ckey(dist(a_i, b_j), i).  keys must be preset
// to kNoKey by the caller (fill kernel below).
hipError_t launch_dist_argmin(const float *a, int lda, const float *an, int na,
                              const float *b, int ldb, const double *bn, int64_t nb,
                              int d, uint64_t *keys, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s);
hipError_t launch_keys_to_idx(const uint64_t *keys, int64_t n, int32_t *idx, hipStream_t s);

// Row top-k over a distance matrix (rows of length n, stride ld): sorted
// (dist, col) ascending, k <= 1024.  Writes idx/dis [rows][k].
hipError_t launch_row_topk(const float *D, int64_t rows, int n, int64_t ld, int k,
                           int32_t *idx, float *dis, hipStream_t s);

// Coarse top-w with the exact reference arithmetic: BF16-MFMA approximate
// distances, then an exact recheck of every centroid within the proven error
// margin of the w-th (d == 128, nlist <= 32768, w <= 1024, 16-byte aligned
// queries; otherwise use launch_dist_store + launch_row_topk).
bool coarse_mfma_ok(int d, int nlist, int w);
size_t coarse_select_lds(int nlist);
// chi / clo: BF16 hi and lo terms of the centroids [nlist][d] (launch_coarse_split).
hipError_t launch_coarse_split(const float *coarse, int64_t n, void *hi, void *lo, hipStream_t s);
// Query norms qn [nq] (as launch_norm_f64) and BF16 hi / lo terms of the
// queries qhi, qlo [nq][128].
hipError_t launch_coarse_prep(const float *queries, int nq, double *qn, void *qhi, void *qlo, hipStream_t s);
// approx [nq][lda]: D'; gmin [nq][ldg] (optional, ldg >= ceil(nlist / 16)):
// the minimum of D' over each group of 16 consecutive centroids.  With
// queries non-null the kernel splits the FP32 queries itself (no qhi, qlo,
// qn needed); otherwise it reads launch_coarse_prep's terms and norms.
// zero (or null): set to 0 by the kernel: launch_coarse_select's redo_cnt.
hipError_t launch_coarse_approx(const void *chi, const void *clo, const float *cn, int nlist, const float *queries,
                                const void *qhi, const void *qlo, const double *qn, int nq, float *approx,
                                int64_t lda, float *gmin, int64_t ldg, int *zero, hipStream_t s);
// True when launch_coarse_select takes every query through the group-minima
// select (nlist >= 64, w <= 64, w <= nlist / 16 groups).
bool coarse_group_select_ok(int nlist, int w);
// redo [nq] and redo_cnt [1]: scratch for queries whose recheck set is large
// (redo_cnt cleared beforehand, by launch_coarse_approx);
// gmin: the group minima of launch_coarse_approx, or null.  qn_out non-null
// (the group select only): the select computes the query norms (as
// launch_norm_f64) into qn_out and uses them (qn is then not read).
hipError_t launch_coarse_select(const float *coarse, const float *cn, float maxcn, int nlist, const float *queries,
                                const double *qn, int nq, int w, const float *approx, int64_t lda,
                                const float *gmin, int64_t ldg, int32_t *probes, float *pdis, int32_t *redo,
                                int *redo_cnt, double *qn_out, hipStream_t s);

// Residuals: r[i] = x[i] - c[assign[i]]
hipError_t launch_residual(const float *x, int64_t n, int d, const float *c,
                           const int32_t *assign, float *r, hipStream_t s);
hipError_t launch_codes_from_keys(const uint64_t *keys, int64_t n, int m, int sub,
                                  uint8_t *codes, hipStream_t s);

// Fused plain search: per (query, probe group) residual -> m x 256 table in
// LDS -> ADC scan -> running exact top-k.
struct ScanArgs {
    const float *queries;     // [nq][d], or residuals [nq*w][d] when resid_mode
    const float *coarse;      // [nlist][d]
    const float *pq;          // [m][ks][ds]
    const double *pq_norm;    // [m][ks] (f64-accumulated norms)
    const int64_t *list_off;  // [nlist+1]
    const int32_t *list_ids;  // [N]
    const uint8_t *codes;     // [N][m]
    const int32_t *probes;    // [nq][wstride] list ids (coarse order)
    const int32_t *wq;        // per-query w or nullptr
    const int32_t *kq;        // per-query k or nullptr
    int nq, w, wstride, k, kstride;
    int ks;                   // centroids per subquantizer (<= 256)
    int nlist;                // probes outside [0, nlist) are skipped
    int ppg;                  // probes per group
    int ngroups;
    int resid_mode;           // 1: queries holds residuals, row = q*wstride + p
    int *err;                 // set non-zero by the kernel on a violated precondition
    const int32_t *qorder;    // [nq] scan order of the queries (XCD-blocked), or nullptr
    uint64_t *partial;        // [nq][ngroups][kstride] when ngroups > 1
    int32_t *out_ids;         // [nq][kstride] when ngroups == 1
    float *out_dis;
    float *rres;              // [nq*wstride][d] residual rows (written by the scan prologue; unused in resid_mode)
    double *rnd;              // [nq*wstride][m] f64 sub-vector norms of the residual rows
};
// Prologue: residual rows and their sub-vector norms of every valid
// (query, probe) pair (a.rres unless resid_mode, a.rnd).
hipError_t launch_scan_prologue(const ScanArgs &a, int m, hipStream_t s);
// Fused table + ADC + top-k scan; reads what the prologue wrote.
hipError_t launch_scan(const ScanArgs &a, int m, hipStream_t s);
// Queries ordered by their first probed list (counting sort; the order
// inside a list is arbitrary).  Only the schedule of the scan depends on it.
bool query_order_ok(int nlist);
hipError_t launch_query_order(const int32_t *probes, int wstride, int nq, int nlist, int32_t *order,
                              hipStream_t s);
// Merge per-group partial keys to the final sorted top-k.
hipError_t launch_merge(const uint64_t *partial, int nq, int ngroups, int kstride,
                        const int32_t *kq, int k, int32_t *out_ids, float *out_dis,
                        hipStream_t s);

// Batched reduced-precision scan (list-major work items of up to 16 queries,
// 8-bit tables, candidates appended to per-query buffers).
// Work item of the reduced-precision scan: list l with gc <= 16 of the
// queries that probe it (q[i] = -1 past gc) and their seed thresholds.
// Lists the list-major grouping takes (LDS histograms up to 8192, global
// atomics above).
constexpr int kMaxPairLists = 1 << 20;
struct FastItem {
    int32_t l, gc, len, pad;
    int64_t beg, pad2;
    int32_t q[16];
    float tau[16];
};
struct FastArgs {
    const float *queries;       // [nq][d]
    const float *coarse;        // [nlist][d]
    const void *pb;             // fp16 [m][256][KP] table B operands (launch_fast_prep)
    const float *pmax;          // [m] max_c |p_sc|
    const uint8_t *codes;       // [N][m]
    const FastItem *items;      // work items
    const int *total_items;
    int *next_item;             // work-queue counter (zeroed before launch)
    uint64_t *cand;             // [nq][capq] (approx key, list position)
    int *cand_cnt;              // [nq]
    int capq;
    int ks;
    int *err;
    const float *tau;           // [nq] seed bounds, or nullptr to take them from the item records
    float *runthr;              // [nq] running bound on the kl-th candidate distance (+inf), or nullptr
    int *rhist;                 // [nq][fast_hist_bins()] candidate-distance histograms (zeroed), with runthr
    int kl;
    float slack;                // seed filter slack in table quanta (8-bit tables of fast.hip)
    const float *pq;            // [m][ks][ds] (the FP32-built table types of fastvar.hip)
    const double *pq_norm;      // [m][ks] f64 norms
};
int fast_group_size();
int fast_hist_bins();
size_t fast_pb_bytes(int m);
hipError_t launch_fast_prep(const float *pq, const float *pmax, int m, int ks, void *pb, hipStream_t s);
hipError_t launch_fill_i32(int *p, int n, int v, hipStream_t s);
// Up to six int fills (p[j][0..n[j]) = v[j]) in one kernel launch.
struct FillSet {
    int *p[6];
    int64_t n[6];
    int v[6];
    int count;
};
hipError_t launch_fill_multi(const FillSet &f, hipStream_t s);
// (list, query) pairs of probe ranks in [rlo, rhi), grouped by list, and the
// work items (list, first pair) of at most G = 1..16 pairs, numbered from
// *item_base (0 if item_base is null) list by list in the order lorder gives
// (a permutation of the lists; list order if null); *total_items becomes the
// end of the numbering (item_base may point at total_items, to append a
// second range).  No memset: the counters are cleared by kernels, so the
// sequence can be captured in a graph.
hipError_t launch_pairs(const int32_t *probes, int nq, int w, int rlo, int rhi, const int64_t *list_off,
                        const float *tau, int nlist, int G, int *cnt, int *pair_off, int *item_off, int *total_items,
                        const int *item_base, const int32_t *lorder, int *fill, int32_t *pair_q, FastItem *items,
                        hipStream_t s);
// Batched exact search (exact.hip).
struct ExactArgs {
    const float *queries;       // [nq][d]
    const float *coarse;        // [nlist][d]
    const float *pq;            // [m][ks][ds]
    const double *pq_norm;      // [m][ks] f64 norms (table side of the reference formula)
    const int32_t *list_ids;    // [N]
    const uint8_t *codes;       // [N][m]
    const FastItem *items;      // work items of G queries
    const int *total_items;
    int *next_item;             // work-queue counter (cleared before launch)
    uint64_t *cand;             // [nq][capq] (distance, id) keys
    int *cand_cnt;              // [nq]
    int capq;
    int ks;
    int *err;
    const float *tau;           // [nq] per-query bounds (nullptr: the items' own)
    // running bounds (nullptr: none): runthr [nq] (+inf at launch) and the
    // per-query histograms rhist [nq][kExactBins] of kept distances over
    // [0, tau_q) (zero at launch); kq [nq] per-query k (nullptr: kuni)
    float *runthr;
    int *rhist;
    const int32_t *kq;
    int kuni;
};
constexpr int kExactBins = 16;
size_t exact_lds_bytes(int m, int g);
// Compiled (m, g, threads): (16,4,256) (16,8,256) (16,8,512) (8,4,256) (8,8,256) (8,8,512) (8,16,512).
hipError_t launch_exact_scan(const ExactArgs &a, int m, int g, int nt, int grid, hipStream_t s);
hipError_t launch_mask_probes(int32_t *probes, int nq, int w, const int32_t *wq, hipStream_t s);
// Test hook: exact tables [m][256][8] of queries 0..n-1 (n <= 8) against list l.
hipError_t launch_exact_tables(const ExactArgs &a, int m, int l, int n, float *out, hipStream_t s);
hipError_t launch_exact_adc(const ExactArgs &a, int m, int l, int n, int64_t beg, int len, float *out, hipStream_t s);
// overflow [nq + 1]: per-query flags, then their count (cleared by the caller).
hipError_t launch_exact_select(const uint64_t *cand, const int *cand_cnt, int capq, const int32_t *kq, int kuni,
                               int kstride, int nq, int32_t *out_ids, float *out_dis, int *overflow, hipStream_t s);

struct SeedArgs {
    const float *queries, *coarse, *pq, *pnorm, *pmax;
    const int64_t *list_off;
    const uint8_t *codes;
    const int32_t *probes;  // [nq][w]
    int min_seed;           // seed lists per query: at least this many, then as many as hold kl vectors
    const int32_t *klq;     // [nq] per-query kl or null (uniform kl)
    int nq, w, nlist, ks, kl;
    int cap;                // > 0: at most cap vectors (list prefixes, in probe order) enter the seed
    float *tau;             // [nq]
};
// tau[q] >= the kl-th smallest ADC distance over q's seed lists (+inf when
// they hold fewer than kl vectors).
hipError_t launch_seed_tau(const SeedArgs &a, int m, hipStream_t s);
hipError_t launch_fast_scan(const FastArgs &a, int m, int grid, hipStream_t s);
// The scan with tables built in FP32 and stored as FP16 (tables = 1) or as
// 8-bit entries with a per-subquantizer offset and scale (tables = 2)
// (fastvar.hip); same items, candidates and keys as launch_fast_scan.
hipError_t launch_fast_var_scan(const FastArgs &a, int m, int tables, int grid, hipStream_t s);
// Test hook: approximate ADC distances out [len][16] of queries 0..n-1
// (n <= 16) over list l (positions beg .. beg + len) with table type tables.
hipError_t launch_fast_var_adc(const FastArgs &a, int m, int tables, int l, int n, int64_t beg, int len, float *out,
                               hipStream_t s);
// The 8-bit tables and scales of one work item (queries 0..n-1 of a.queries,
// n <= 16, list l): t8 [m][256][16] uint8, delta [16].  Test hook.
hipError_t launch_fast_tables(const FastArgs &a, int m, int l, int n, uint8_t *t8, float *delta, hipStream_t s);
// The shortlist of launch_fast_select and the exact rerank of
// launch_rerank_pos / launch_rerank_pos_u8 in one kernel (kl <= 256): rows
// are the list-ordered copy (float or bytes) or, with list_ordered false, the
// float base indexed by id.
hipError_t launch_fast_select_rerank(int kl, const uint64_t *cand, const int *cand_cnt, int capq, int nq,
                                     const void *rows, bool rows_u8, bool list_ordered, int64_t nrows,
                                     const int32_t *list_ids, const float *queries, int d, const double *qnorm, int k,
                                     int32_t *out_ids, float *out_dis, hipStream_t s);
hipError_t launch_fast_select(int kl, const uint64_t *cand, const int *cand_cnt, int capq, int nq,
                              int32_t *short_ids, hipStream_t s);

// Exact rerank of a shortlist against raw vectors (a = base row, b = query).
hipError_t launch_rerank(const float *base, int64_t nbase, const float *queries, int nq,
                         int d, const int32_t *short_ids, int kl, int k,
                         int32_t *out_ids, float *out_dis, hipStream_t s);

// Same with a shortlist of list positions (id = list_ids[pos]); rows are
// base[id], or base[pos] of a list-ordered copy when list_ordered.  qnorm:
// the queries' f64-accumulated norms (launch_norm_f64), or nullptr.
hipError_t launch_rerank_pos(const float *rows, int64_t nrows, bool list_ordered, const int32_t *list_ids,
                             const float *queries, int nq, int d, const double *qnorm, const int32_t *short_pos,
                             int kl, int k, int32_t *out_ids, float *out_dis, hipStream_t s);
// out[r] = base[ids[r]] (rows of d floats)
hipError_t launch_gather_rows(const float *base, int d, const int32_t *ids, int64_t n, float *out, hipStream_t s);
// Same as the list-ordered rerank with byte rows (integer-valued data 0..255:
// the same float values, so the same distances bit for bit).
hipError_t launch_rerank_pos_u8(const uint8_t *rows, int64_t nrows, const int32_t *list_ids, const float *queries,
                                int nq, int d, const double *qnorm, const int32_t *short_pos, int kl, int k,
                                int32_t *out_ids, float *out_dis, hipStream_t s);
// out[r] = base[ids[r]] as bytes; *bad |= 1 when a value is not an integer in 0..255
hipError_t launch_gather_rows_u8(const float *base, int d, const int32_t *ids, int64_t n, uint8_t *out, int *bad,
                                 hipStream_t s);

// ADC distances of one residual [m*ds] over n codes [n][m] (uint8), in code
// order, with the reference table arithmetic (adc.hip).
hipError_t launch_adc_codes(const float *resid, int m, int ks, int ds, const float *pq, const double *pq_n64,
                            const uint8_t *codes, int64_t n, float *out, hipStream_t s);
// k smallest (distance, id) of dis/ids [n], ascending, padded with (-1, +inf)
// to k entries; k <= 1024 (select.hip).
hipError_t launch_kmin(const float *dis, const int32_t *ids, int64_t n, int k, int32_t *out_ids, float *out_dis,
                       hipStream_t s);

// Synthetic mixture rows (bit-identical to the CPU generator).
uint64_t synth_stream_base(uint64_t seed, uint64_t stream);
uint64_t synth_draw(uint64_t base, uint64_t ctr);
hipError_t launch_gen_rows(uint64_t seed, uint64_t stream, int64_t row0, int64_t n, int d,
                           const float *centres, int ncomp, const uint64_t *thr,
                           float sigma, float *out, hipStream_t s);

// FP64 brute-force nearest neighbour (ties -> lowest id).
hipError_t launch_bruteforce_nn(const float *base, int64_t nb, const float *q, int nq,
                                int d, int32_t *ids, double *dis, double *part_d,
                                int32_t *part_i, int nchunks, hipStream_t s);

}  // namespace mivf
