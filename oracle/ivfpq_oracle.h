/*
 * CPU oracle for the IVF-PQ search path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the arithmetic of the reference's IVF-PQ
 * search (rafaelms101/mestrado) so that the HIP product can be checked bit for
 * bit.  Only the test-suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it; the product never links or calls it.
 *
 * Arithmetic contract (reference citations are paths in the reference repo):
 *   distance(a_i, b_j)  -- parallel/yael_needs/nn.c:121-150 + netlib SGEMM order
 *       na_i = sum_t fl32(a_it*a_it)         float accumulator, t ascending
 *       nb_j = sum_t (double)fl32(b_jt*b_jt)  double accumulator, t ascending
 *       c    = (float)(nb_j + (double)na_i)
 *       dot  = sum_t fl32(a_it*b_jt)          float accumulator, t ascending, no FMA
 *       D    = fl32(fl32(-2*dot) + c)
 *     coarse:  a = coarse centroid, b = query   (parallel/ivf_pq/ivf_assign.cpp:28,
 *              knn_full arguments nn.c:463, 480-482)
 *     table:   a = residual sub-vector, b = PQ centroid
 *              (parallel/ivf_pq/ivf_search.cpp:206-209)
 *     encode:  a = PQ centroid, b = residual sub-vector
 *              (parallel/pq-utils/pq_assign.cpp:45)
 *     rerank:  a = base vector, b = query (sequential/ivf_pq/ivf_search.cpp:130)
 *   residual  r = fl32(q - c)   (parallel/ivf_pq/ivf_assign.cpp:72-78)
 *   ADC       s = 0f; for j < m: s = fl32(s + T[j][code_j])
 *             (parallel/ivf_pq/ivf_search.cpp:451-468)
 *   select    k smallest by (distance ascending, id ascending), the order of
 *             parallel2/ivf_pq/topk.cu:26-32.  Missing results are padded with
 *             id -1 and distance +inf.
 *
 * Parity pins for each function are listed in tests/test_oracle.py.
 */
#ifndef IVFPQ_ORACLE_H
#define IVFPQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- order keys ---------------------------------------------------------- */
uint32_t orc_fkey(float f);
uint64_t orc_ckey(float d, int32_t id);
float orc_key_dist(uint64_t key);

/* ---- synthetic mixture generator (deterministic, integer + IEEE basic ops) */
void orc_zipf_thresholds(uint64_t seed, int ncomp, double s, uint64_t *thr);
void orc_gen_centres(uint64_t seed, int ncomp, int d, float *centres);
void orc_gen_rows(uint64_t seed, uint64_t stream, int64_t row0, int64_t n, int d,
                  const float *centres, int ncomp, const uint64_t *thr,
                  float sigma, float *out);

/* ---- distance primitive ---------------------------------------------------- */
void orc_cross_dist(int d, int na, int nb, const float *a, int lda,
                    const float *b, int ldb, float *out, int ldo);

/* ---- selection --------------------------------------------------------------- */
void orc_topk_keys(const uint64_t *keys, int64_t n, int k, uint64_t *out);
/* knn_full(2, n1, n2, d, k, mat2, mat1): top-k of mat2 rows for each mat1 row */
void orc_knn(int64_t n1, int n2, int d, int k, const float *mat2, int ld2,
             const float *mat1, int ld1, int32_t *idx, float *dis, int nthreads);

/* ---- search-path pieces ----------------------------------------------------- */
void orc_residual(int d, const float *q, const float *c, float *r);
void orc_tables(int d, int m, int ks, const float *r, const float *pq, float *tab);
void orc_adc(int m, int ks, const float *tab, const uint8_t *codes, int64_t n,
             float *out);

/* full plain search; coaidx/coadis may be NULL */
int orc_search(int d, int nlist, int m, int ks, const float *coarse, const float *pq,
               const int64_t *offsets, const int32_t *ids, const uint8_t *codes,
               const float *queries, int nq, int w, int k,
               int32_t *out_ids, float *out_dis, int32_t *coaidx, float *coadis,
               int nthreads);
/* the same search staged as the parallel/ CPU loop (coarse, tables, ADC into
   a concatenated array, selection), with phase_s[4] = seconds per phase
   summed over threads; results equal orc_search */
int orc_search_phases(int d, int nlist, int m, int ks, const float *coarse, const float *pq,
                      const int64_t *offsets, const int32_t *ids, const uint8_t *codes,
                      const float *queries, int nq, int w, int k,
                      int32_t *out_ids, float *out_dis, double *phase_s, int nthreads);
/* per-query w and k (rows of out_* are kmax wide) */
int orc_search_var(int d, int nlist, int m, int ks, const float *coarse,
                   const float *pq, const int64_t *offsets, const int32_t *ids,
                   const uint8_t *codes, const float *queries, int nq,
                   const int32_t *wq, const int32_t *kq, int kmax,
                   int32_t *out_ids, float *out_dis, int nthreads);
/* search residuals against given lists (parallel2 core_gpu semantics):
   residual [nq*w][d], list_of [nq*w]; outputs [nq][k] */
int orc_search_residual(int d, int m, int ks, const float *pq,
                        const int64_t *offsets, const int32_t *ids,
                        const uint8_t *codes, const float *residual,
                        const int32_t *list_of, int nq, int w, int k,
                        int32_t *out_ids, float *out_dis, int nthreads);
/* shortlist of kl by ADC, exact rerank against the raw base to k */
int orc_search_rerank(int d, int nlist, int m, int ks, const float *coarse,
                      const float *pq, const int64_t *offsets, const int32_t *ids,
                      const uint8_t *codes, const float *base, int64_t nbase,
                      const float *queries, int nq, int w, int kl, int k,
                      int32_t *out_ids, float *out_dis, int nthreads);

/* ---- index build ------------------------------------------------------------- */
int orc_kmeans(const float *x, int64_t n, int d, int ldx, int k, int niter,
               uint64_t seed, float *cent, int nthreads);
int orc_ivfpq_train(const float *x, int64_t n, int d, int nlist, int m, int ks,
                    int niter_coarse, int niter_pq, uint64_t seed,
                    float *coarse, float *pq, int nthreads);
int orc_ivfpq_encode(int d, int nlist, int m, int ks, const float *coarse,
                     const float *pq, const float *x, int64_t n,
                     int32_t *assign, uint8_t *codes, int nthreads);
int orc_ivfpq_bucket(int nlist, int m, const int32_t *assign, const uint8_t *codes,
                     int64_t n, int32_t id0, int64_t *offsets, int32_t *ids,
                     uint8_t *codes_out);

/* ---- ground truth and quality ------------------------------------------------- */
void orc_bruteforce_f64(const float *base, int64_t nb, const float *q, int nq,
                        int d, int k, int32_t *ids, double *dis, int nthreads);
double orc_recall_at(const int32_t *ids, int nq, int k, const int32_t *gt, int R);

#ifdef __cplusplus
}
#endif
#endif
