/*
 * CPU oracle for the IVF-PQ search path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg use it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (no -march), see oracle/Makefile.
 * Every float operation below is a single IEEE operation in program order:
 * no FMA contraction, no reassociation.  Citations are reference paths.
 */
#include "ivfpq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GOLDEN 0x9E3779B97F4A7C15ULL

/* ------------------------------------------------------------------------ */
/* order keys: (distance ascending, id ascending) as one unsigned 64-bit key   */
/* (parallel2/ivf_pq/topk.cu:26-32 orders Img by dist, then by imgid)          */
/* ------------------------------------------------------------------------ */
uint32_t orc_fkey(float f)
{
    f = f + 0.0f; /* -0 -> +0 so that equal distances compare equal */
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

uint64_t orc_ckey(float d, int32_t id)
{
    return ((uint64_t)orc_fkey(d) << 32) | (uint32_t)id;
}

float orc_key_dist(uint64_t key)
{
    uint32_t h = (uint32_t)(key >> 32);
    uint32_t b = (h & 0x80000000u) ? (h & 0x7FFFFFFFu) : ~h;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

/* ------------------------------------------------------------------------ */
/* synthetic mixture generator                                                 */
/* SplitMix64 finaliser as a counter-based hash; normals are Irwin-Hall sums  */
/* of twelve 16-bit uniforms so that every step is integer or a correctly      */
/* rounded IEEE operation (identical on any IEEE machine).                     */
/* ------------------------------------------------------------------------ */
static uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

static uint64_t stream_base(uint64_t seed, uint64_t stream)
{
    return mix64(seed ^ mix64(stream + GOLDEN));
}

static uint64_t draw(uint64_t base, uint64_t ctr)
{
    return mix64(base + (ctr + 1) * GOLDEN);
}

void orc_zipf_thresholds(uint64_t seed, int ncomp, double s, uint64_t *thr)
{
    /* rank of component c is its position in a seeded Fisher-Yates permutation */
    int *perm = (int *)malloc(sizeof(int) * ncomp);
    int *rank = (int *)malloc(sizeof(int) * ncomp);
    uint64_t b = stream_base(seed, 0x7a697066ULL);
    for (int i = 0; i < ncomp; i++) perm[i] = i;
    for (int i = 0; i < ncomp - 1; i++) {
        uint64_t r = draw(b, (uint64_t)i);
        int j = i + (int)(((unsigned __int128)r * (uint64_t)(ncomp - i)) >> 64);
        int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    for (int i = 0; i < ncomp; i++) rank[perm[i]] = i;
    double tot = 0;
    for (int c = 0; c < ncomp; c++) tot += 1.0 / pow((double)(rank[c] + 1), s);
    double cum = 0;
    for (int c = 0; c < ncomp; c++) {
        cum += 1.0 / pow((double)(rank[c] + 1), s);
        double t = floor(cum / tot * 9007199254740992.0);
        thr[c] = (c == ncomp - 1 || t > 9007199254740992.0) ? (1ULL << 53) : (uint64_t)t;
    }
    free(perm);
    free(rank);
}

void orc_gen_centres(uint64_t seed, int ncomp, int d, float *centres)
{
    uint64_t b = stream_base(seed, 1);
    for (int64_t i = 0; i < (int64_t)ncomp * d; i++) {
        uint64_t u = draw(b, (uint64_t)i) >> 11;
        centres[i] = (float)((double)u * (1.0 / 9007199254740992.0) * 128.0);
    }
}

void orc_gen_rows(uint64_t seed, uint64_t stream, int64_t row0, int64_t n, int d,
                  const float *centres, int ncomp, const uint64_t *thr,
                  float sigma, float *out)
{
    uint64_t b = stream_base(seed, stream);
    uint64_t per_row = 1 + 3 * (uint64_t)d;
#pragma omp parallel for schedule(static)
    for (int64_t rr = 0; rr < n; rr++) {
        uint64_t r = (uint64_t)(row0 + rr);
        uint64_t u53 = draw(b, r * per_row) >> 11;
        int comp;
        if (!thr) {
            comp = (int)((u53 * (uint64_t)ncomp) >> 53);
        } else {
            comp = 0;
            while (comp < ncomp - 1 && u53 >= thr[comp]) comp++;
        }
        const float *c = centres + (int64_t)comp * d;
        for (int i = 0; i < d; i++) {
            int64_t s = 0;
            for (int t = 0; t < 3; t++) {
                uint64_t g = draw(b, r * per_row + 1 + 3 * (uint64_t)i + t);
                s += (int64_t)(g & 0xFFFF) + (int64_t)((g >> 16) & 0xFFFF) +
                     (int64_t)((g >> 32) & 0xFFFF) + (int64_t)(g >> 48);
            }
            double z = (double)(s - 393210) / 65536.0;
            double x = (double)c[i] + (double)sigma * z;
            double v = rint(x);
            if (v < 0.0) v = 0.0;
            if (v > 255.0) v = 255.0;
            out[rr * d + i] = (float)v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* distance primitive: parallel/yael_needs/nn.c:121-150 with the netlib SGEMM  */
/* loop (TEMP = TEMP + A(L,I)*B(L,J); C = ALPHA*TEMP + BETA*C), alpha=-2, beta=1 */
/* out[j*ldo + i] = dist(a_i, b_j)                                              */
/* ------------------------------------------------------------------------ */
static float norm_f32(const float *a, int d)
{
    float s = 0;
    for (int t = 0; t < d; t++) s += a[t] * a[t];
    return s;
}

static double norm_f64(const float *b, int d)
{
    double s = 0;
    for (int t = 0; t < d; t++) {
        float p = b[t] * b[t];
        s += (double)p;
    }
    return s;
}

static float yael_dist(const float *a, float na, const float *b, double nb, int d)
{
    float c = (float)(nb + (double)na);
    float dot = 0;
    for (int t = 0; t < d; t++) {
        float p = a[t] * b[t];
        dot = dot + p;
    }
    float m2 = -2.0f * dot;
    return m2 + c;
}

void orc_cross_dist(int d, int na, int nb, const float *a, int lda,
                    const float *b, int ldb, float *out, int ldo)
{
    for (int j = 0; j < nb; j++) {
        double nbj = norm_f64(b + (int64_t)j * ldb, d);
        for (int i = 0; i < na; i++) {
            float nai = norm_f32(a + (int64_t)i * lda, d);
            out[(int64_t)j * ldo + i] =
                yael_dist(a + (int64_t)i * lda, nai, b + (int64_t)j * ldb, nbj, d);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* selection: bounded max-heap of composite keys, then ascending sort          */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t *h;
    int n, cap;
} kheap;

static void kh_push(kheap *H, uint64_t key)
{
    if (H->n < H->cap) {
        int i = H->n++;
        while (i > 0) {
            int p = (i - 1) / 2;
            if (H->h[p] >= key) break;
            H->h[i] = H->h[p];
            i = p;
        }
        H->h[i] = key;
        return;
    }
    if (H->cap == 0 || key >= H->h[0]) return;
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        uint64_t mv = key;
        if (l < H->n && H->h[l] > mv) { m = l; mv = H->h[l]; }
        if (r < H->n && H->h[r] > mv) { m = r; mv = H->h[r]; }
        if (m == i) break;
        H->h[i] = H->h[m];
        i = m;
    }
    H->h[i] = key;
}

static int cmp_u64(const void *x, const void *y)
{
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return (a > b) - (a < b);
}

/* sorted k smallest of keys (fewer if n < k); returns count */
static int kh_finish(kheap *H, uint64_t *out)
{
    qsort(H->h, H->n, sizeof(uint64_t), cmp_u64);
    memcpy(out, H->h, sizeof(uint64_t) * H->n);
    return H->n;
}

void orc_topk_keys(const uint64_t *keys, int64_t n, int k, uint64_t *out)
{
    kheap H = {(uint64_t *)malloc(sizeof(uint64_t) * (k > 0 ? k : 1)), 0, k};
    for (int64_t i = 0; i < n; i++) kh_push(&H, keys[i]);
    int c = kh_finish(&H, out);
    for (int i = c; i < k; i++) out[i] = UINT64_MAX;
    free(H.h);
}

static void emit(const uint64_t *keys, int cnt, int k, int32_t *ids, float *dis)
{
    for (int i = 0; i < k; i++) {
        if (i < cnt) {
            ids[i] = (int32_t)(uint32_t)(keys[i] & 0xFFFFFFFFu);
            dis[i] = orc_key_dist(keys[i]);
        } else {
            ids[i] = -1;
            dis[i] = INFINITY;
        }
    }
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* knn_full(2, n1, n2, d, k, mat2, mat1, ...) -- parallel/yael_needs/nn.c:463-537
   a = mat2 (centroids), b = mat1 (points); ties -> lower index */
void orc_knn(int64_t n1, int n2, int d, int k, const float *mat2, int ld2,
             const float *mat1, int ld1, int32_t *idx, float *dis, int nthreads)
{
    float *n2s = (float *)malloc(sizeof(float) * n2);
    for (int j = 0; j < n2; j++) n2s[j] = norm_f32(mat2 + (int64_t)j * ld2, d);
    set_threads(nthreads);
#pragma omp parallel
    {
        uint64_t *hb = (uint64_t *)malloc(sizeof(uint64_t) * (k > 0 ? k : 1));
        uint64_t *ob = (uint64_t *)malloc(sizeof(uint64_t) * (k > 0 ? k : 1));
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n1; i++) {
            const float *b = mat1 + i * ld1;
            double nb = norm_f64(b, d);
            kheap H = {hb, 0, k};
            for (int j = 0; j < n2; j++) {
                float dd = yael_dist(mat2 + (int64_t)j * ld2, n2s[j], b, nb, d);
                kh_push(&H, orc_ckey(dd, j));
            }
            int c = kh_finish(&H, ob);
            emit(ob, c, k, idx + i * k, dis + i * k);
        }
        free(hb);
        free(ob);
    }
    free(n2s);
}

/* ------------------------------------------------------------------------ */
/* search-path pieces                                                           */
/* ------------------------------------------------------------------------ */
/* bsxfunMINUS: parallel/ivf_pq/ivf_assign.cpp:72-78 */
void orc_residual(int d, const float *q, const float *c, float *r)
{
    for (int i = 0; i < d; i++) r[i] = q[i] - c[i];
}

/* compute_cross_distances(ds, 1, ks, r_s, P_s): parallel/ivf_pq/ivf_search.cpp:206-209
   tab[s*ks + j] = dist(a = r_s, b = P[s][j]) */
void orc_tables(int d, int m, int ks, const float *r, const float *pq, float *tab)
{
    int ds = d / m;
    for (int s = 0; s < m; s++)
        orc_cross_dist(ds, 1, ks, r + s * ds, ds, pq + (int64_t)s * ks * ds, ds,
                       tab + (int64_t)s * ks, 1);
}

/* sumidxtab2: parallel/ivf_pq/ivf_search.cpp:451-468 */
void orc_adc(int m, int ks, const float *tab, const uint8_t *codes, int64_t n,
             float *out)
{
    for (int64_t i = 0; i < n; i++) {
        float s = 0;
        for (int j = 0; j < m; j++) s = s + tab[(int64_t)j * ks + codes[i * m + j]];
        out[i] = s;
    }
}

/* scan the listed probes of one query into heap H */
static void scan_probes(int d, int m, int ks, const float *pq, const int64_t *offsets,
                        const int32_t *ids, const uint8_t *codes, const float *resid,
                        int list, float *tab, kheap *H)
{
    orc_tables(d, m, ks, resid, pq, tab);
    for (int64_t v = offsets[list]; v < offsets[list + 1]; v++) {
        float s = 0;
        const uint8_t *cv = codes + v * m;
        for (int j = 0; j < m; j++) s = s + tab[(int64_t)j * ks + cv[j]];
        kh_push(H, orc_ckey(s, ids[v]));
    }
}

int orc_search_var(int d, int nlist, int m, int ks, const float *coarse,
                   const float *pq, const int64_t *offsets, const int32_t *ids,
                   const uint8_t *codes, const float *queries, int nq,
                   const int32_t *wq, const int32_t *kq, int kmax,
                   int32_t *out_ids, float *out_dis, int nthreads)
{
    float *cn = (float *)malloc(sizeof(float) * nlist);
    for (int c = 0; c < nlist; c++) cn[c] = norm_f32(coarse + (int64_t)c * d, d);
    set_threads(nthreads);
#pragma omp parallel
    {
        uint64_t *ck = (uint64_t *)malloc(sizeof(uint64_t) * nlist);
        uint64_t *hb = (uint64_t *)malloc(sizeof(uint64_t) * (kmax + 1));
        uint64_t *ob = (uint64_t *)malloc(sizeof(uint64_t) * (kmax + 1));
        uint64_t *pb = (uint64_t *)malloc(sizeof(uint64_t) * nlist);
        float *tab = (float *)malloc(sizeof(float) * m * ks);
        float *r = (float *)malloc(sizeof(float) * d);
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < nq; q++) {
            const float *x = queries + (int64_t)q * d;
            int w = wq[q], k = kq[q];
            double nb = norm_f64(x, d);
            kheap C = {pb, 0, w};
            for (int c = 0; c < nlist; c++) {
                float dd = yael_dist(coarse + (int64_t)c * d, cn[c], x, nb, d);
                kh_push(&C, orc_ckey(dd, c));
            }
            int nprobe = kh_finish(&C, ck);
            kheap H = {hb, 0, k};
            for (int p = 0; p < nprobe; p++) {
                int l = (int)(ck[p] & 0xFFFFFFFFu);
                orc_residual(d, x, coarse + (int64_t)l * d, r);
                scan_probes(d, m, ks, pq, offsets, ids, codes, r, l, tab, &H);
            }
            int cnt = kh_finish(&H, ob);
            emit(ob, cnt, k, out_ids + (int64_t)q * kmax, out_dis + (int64_t)q * kmax);
            for (int i = k; i < kmax; i++) {
                out_ids[(int64_t)q * kmax + i] = -1;
                out_dis[(int64_t)q * kmax + i] = INFINITY;
            }
        }
        free(ck); free(hb); free(ob); free(pb); free(tab); free(r);
    }
    free(cn);
    return 0;
}

int orc_search(int d, int nlist, int m, int ks, const float *coarse, const float *pq,
               const int64_t *offsets, const int32_t *ids, const uint8_t *codes,
               const float *queries, int nq, int w, int k,
               int32_t *out_ids, float *out_dis, int32_t *coaidx, float *coadis,
               int nthreads)
{
    if (w < 1 || w > nlist || k < 1 || d % m) return 1;
    int32_t *wq = (int32_t *)malloc(sizeof(int32_t) * nq);
    int32_t *kq = (int32_t *)malloc(sizeof(int32_t) * nq);
    for (int q = 0; q < nq; q++) { wq[q] = w; kq[q] = k; }
    int rc = orc_search_var(d, nlist, m, ks, coarse, pq, offsets, ids, codes, queries,
                            nq, wq, kq, k, out_ids, out_dis, nthreads);
    free(wq);
    free(kq);
    if (coaidx && coadis)
        orc_knn(nq, nlist, d, w, coarse, d, queries, d, coaidx, coadis, nthreads);
    return rc;
}

/* The same search staged as the parallel/ CPU loop is, with a wall-clock
   split per phase (parallel/ivf_pq/ivf_search.cpp:97-137 and its per-thread
   log :150-162): coarse top-w (knn_full, ivf_assign.cpp:28), then per probe
   the residual and table (ivf_search.cpp:206-209) and sumidxtab2 into a
   distance array (:211, :451-468), the arrays of the w probes concatenated
   (:114-128), then the k-selection over the concatenation (:131-137, exact
   (distance, id) order in place of my_k_min).  Results equal orc_search.
   phase_s[4] receives the seconds summed over threads: coarse, residual +
   tables, ADC + concatenation, selection. */
int orc_search_phases(int d, int nlist, int m, int ks, const float *coarse, const float *pq,
                      const int64_t *offsets, const int32_t *ids, const uint8_t *codes,
                      const float *queries, int nq, int w, int k,
                      int32_t *out_ids, float *out_dis, double *phase_s, int nthreads)
{
    if (w < 1 || w > nlist || k < 1 || d % m) return 1;
    int64_t maxlen = 0;
    for (int l = 0; l < nlist; l++)
        if (offsets[l + 1] - offsets[l] > maxlen) maxlen = offsets[l + 1] - offsets[l];
    float *cn = (float *)malloc(sizeof(float) * nlist);
    for (int c = 0; c < nlist; c++) cn[c] = norm_f32(coarse + (int64_t)c * d, d);
    double tot[4] = {0, 0, 0, 0};
    set_threads(nthreads);
#pragma omp parallel
    {
        double t[4] = {0, 0, 0, 0};
        uint64_t *ck = (uint64_t *)malloc(sizeof(uint64_t) * nlist);
        uint64_t *pb = (uint64_t *)malloc(sizeof(uint64_t) * nlist);
        uint64_t *hb = (uint64_t *)malloc(sizeof(uint64_t) * (k + 1));
        uint64_t *ob = (uint64_t *)malloc(sizeof(uint64_t) * (k + 1));
        float *tab = (float *)malloc(sizeof(float) * m * ks);
        float *r = (float *)malloc(sizeof(float) * d);
        int64_t cap = maxlen * w + 1;
        float *cdis = (float *)malloc(sizeof(float) * cap);
        int32_t *cid = (int32_t *)malloc(sizeof(int32_t) * cap);
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < nq; q++) {
            const float *x = queries + (int64_t)q * d;
            double t0 = omp_get_wtime();
            double nb = norm_f64(x, d);
            kheap C = {pb, 0, w};
            for (int c = 0; c < nlist; c++) {
                float dd = yael_dist(coarse + (int64_t)c * d, cn[c], x, nb, d);
                kh_push(&C, orc_ckey(dd, c));
            }
            int nprobe = kh_finish(&C, ck);
            double t1 = omp_get_wtime();
            t[0] += t1 - t0;
            int64_t n = 0;
            for (int p = 0; p < nprobe; p++) {
                int l = (int)(ck[p] & 0xFFFFFFFFu);
                double a = omp_get_wtime();
                orc_residual(d, x, coarse + (int64_t)l * d, r);
                orc_tables(d, m, ks, r, pq, tab);
                double b = omp_get_wtime();
                int64_t len = offsets[l + 1] - offsets[l];
                orc_adc(m, ks, tab, codes + offsets[l] * m, len, cdis + n);
                memcpy(cid + n, ids + offsets[l], sizeof(int32_t) * len);
                n += len;
                double c2 = omp_get_wtime();
                t[1] += b - a;
                t[2] += c2 - b;
            }
            double t2 = omp_get_wtime();
            kheap H = {hb, 0, k};
            for (int64_t i = 0; i < n; i++) kh_push(&H, orc_ckey(cdis[i], cid[i]));
            int cnt = kh_finish(&H, ob);
            emit(ob, cnt, k, out_ids + (int64_t)q * k, out_dis + (int64_t)q * k);
            t[3] += omp_get_wtime() - t2;
        }
#pragma omp critical
        for (int i = 0; i < 4; i++) tot[i] += t[i];
        free(ck); free(pb); free(hb); free(ob); free(tab); free(r); free(cdis); free(cid);
    }
    free(cn);
    if (phase_s)
        for (int i = 0; i < 4; i++) phase_s[i] = tot[i];
    return 0;
}

/* parallel2/ivf_pq/mycuda.cu:119-188 semantics: residuals already computed */
int orc_search_residual(int d, int m, int ks, const float *pq,
                        const int64_t *offsets, const int32_t *ids,
                        const uint8_t *codes, const float *residual,
                        const int32_t *list_of, int nq, int w, int k,
                        int32_t *out_ids, float *out_dis, int nthreads)
{
    set_threads(nthreads);
#pragma omp parallel
    {
        uint64_t *hb = (uint64_t *)malloc(sizeof(uint64_t) * (k + 1));
        uint64_t *ob = (uint64_t *)malloc(sizeof(uint64_t) * (k + 1));
        float *tab = (float *)malloc(sizeof(float) * m * ks);
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < nq; q++) {
            kheap H = {hb, 0, k};
            for (int p = 0; p < w; p++) {
                int64_t rid = (int64_t)q * w + p;
                scan_probes(d, m, ks, pq, offsets, ids, codes, residual + rid * d,
                            list_of[rid], tab, &H);
            }
            int cnt = kh_finish(&H, ob);
            emit(ob, cnt, k, out_ids + (int64_t)q * k, out_dis + (int64_t)q * k);
        }
        free(hb); free(ob); free(tab);
    }
    return 0;
}

/* sequential/ivf_pq/ivf_search.cpp:9-153: ADC shortlist of kl, then exact
   knn_full(2, 1, ktmp, d, k, nearest, query) over the shortlisted raw vectors */
int orc_search_rerank(int d, int nlist, int m, int ks, const float *coarse,
                      const float *pq, const int64_t *offsets, const int32_t *ids,
                      const uint8_t *codes, const float *base, int64_t nbase,
                      const float *queries, int nq, int w, int kl, int k,
                      int32_t *out_ids, float *out_dis, int nthreads)
{
    if (w < 1 || w > nlist || k < 1 || kl < 1 || d % m) return 1;
    int32_t *sid = (int32_t *)malloc(sizeof(int32_t) * (int64_t)nq * kl);
    float *sdis = (float *)malloc(sizeof(float) * (int64_t)nq * kl);
    int rc = orc_search(d, nlist, m, ks, coarse, pq, offsets, ids, codes, queries, nq,
                        w, kl, sid, sdis, NULL, NULL, nthreads);
    if (rc) { free(sid); free(sdis); return rc; }
    set_threads(nthreads);
    int bad = 0;
#pragma omp parallel
    {
        uint64_t *hb = (uint64_t *)malloc(sizeof(uint64_t) * (k + 1));
        uint64_t *ob = (uint64_t *)malloc(sizeof(uint64_t) * (k + 1));
#pragma omp for schedule(dynamic, 4)
        for (int q = 0; q < nq; q++) {
            const float *x = queries + (int64_t)q * d;
            double nb = norm_f64(x, d);
            kheap H = {hb, 0, k};
            for (int i = 0; i < kl; i++) {
                int32_t id = sid[(int64_t)q * kl + i];
                if (id < 0) break;
                if ((int64_t)id >= nbase) {
#pragma omp atomic write
                    bad = 1;
                    break;
                }
                const float *a = base + (int64_t)id * d;
                float dd = yael_dist(a, norm_f32(a, d), x, nb, d);
                kh_push(&H, orc_ckey(dd, id));
            }
            int cnt = kh_finish(&H, ob);
            emit(ob, cnt, k, out_ids + (int64_t)q * k, out_dis + (int64_t)q * k);
        }
        free(hb); free(ob);
    }
    free(sid);
    free(sdis);
    return bad ? 2 : 0;
}

/* ------------------------------------------------------------------------ */
/* index build                                                                   */
/* ------------------------------------------------------------------------ */
/* argmin over centroids (a = centroid, b = point), strict '<' keeps the lowest
   index on ties: parallel/yael_needs/nn.c:404-458 */
static void assign_all(const float *x, int64_t n, int d, int ldx, const float *cent,
                       int k, int32_t *assign)
{
    float *cn = (float *)malloc(sizeof(float) * k);
    for (int j = 0; j < k; j++) cn[j] = norm_f32(cent + (int64_t)j * d, d);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        const float *b = x + i * ldx;
        double nb = norm_f64(b, d);
        uint64_t best = UINT64_MAX;
        for (int j = 0; j < k; j++) {
            uint64_t key = orc_ckey(yael_dist(cent + (int64_t)j * d, cn[j], b, nb, d), j);
            if (key < best) best = key;
        }
        assign[i] = (int32_t)(best & 0xFFFFFFFFu);
    }
    free(cn);
}

/* Lloyd k-means with a seeded sample of distinct training points as the
   initial centroids, double-precision sequential centroid sums and a
   deterministic split of the largest cluster for empty ones. */
int orc_kmeans(const float *x, int64_t n, int d, int ldx, int k, int niter,
               uint64_t seed, float *cent, int nthreads)
{
    if (k < 1 || n < k) return 1;
    set_threads(nthreads);
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * n);
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    uint64_t b = stream_base(seed, 0x6b6d65616e73ULL);
    for (int64_t i = 0; i < k; i++) {
        uint64_t r = draw(b, (uint64_t)i);
        int64_t j = i + (int64_t)(((unsigned __int128)r * (uint64_t)(n - i)) >> 64);
        int64_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
    for (int j = 0; j < k; j++)
        for (int t = 0; t < d; t++) cent[(int64_t)j * d + t] = x[perm[j] * ldx + t];
    free(perm);

    int32_t *assign = (int32_t *)malloc(sizeof(int32_t) * n);
    double *sums = (double *)malloc(sizeof(double) * (int64_t)k * d);
    int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * k);
    for (int it = 0; it < niter; it++) {
        assign_all(x, n, d, ldx, cent, k, assign);
        memset(sums, 0, sizeof(double) * (int64_t)k * d);
        memset(cnt, 0, sizeof(int64_t) * k);
        for (int64_t i = 0; i < n; i++) {
            int c = assign[i];
            cnt[c]++;
            double *s = sums + (int64_t)c * d;
            const float *xi = x + i * ldx;
            for (int t = 0; t < d; t++) s[t] += (double)xi[t];
        }
        for (int c = 0; c < k; c++)
            if (cnt[c] > 0)
                for (int t = 0; t < d; t++)
                    cent[(int64_t)c * d + t] =
                        (float)(sums[(int64_t)c * d + t] / (double)cnt[c]);
        for (int c = 0; c < k; c++) {
            if (cnt[c] > 0) continue;
            int big = 0;
            for (int j = 1; j < k; j++)
                if (cnt[j] > cnt[big]) big = j;
            for (int t = 0; t < d; t++) {
                double v = (double)cent[(int64_t)big * d + t];
                double delta = (1.0 / 1024.0) * (1.0 + fabs(v));
                cent[(int64_t)c * d + t] = (float)(v + delta);
                cent[(int64_t)big * d + t] = (float)(v - delta);
            }
            cnt[c] = cnt[big] / 2;
            cnt[big] -= cnt[c];
        }
    }
    free(assign);
    free(sums);
    free(cnt);
    return 0;
}

/* ivfpq_new: parallel/ivf_pq/ivf_training.cpp:60-90 (coarse k-means, residuals,
   one k-means per subquantizer with ks centroids) */
int orc_ivfpq_train(const float *x, int64_t n, int d, int nlist, int m, int ks,
                    int niter_coarse, int niter_pq, uint64_t seed,
                    float *coarse, float *pq, int nthreads)
{
    if (d % m || n < nlist || n < ks) return 1;
    int ds = d / m;
    int rc = orc_kmeans(x, n, d, d, nlist, niter_coarse, seed, coarse, nthreads);
    if (rc) return rc;
    int32_t *assign = (int32_t *)malloc(sizeof(int32_t) * n);
    assign_all(x, n, d, d, coarse, nlist, assign);
    float *res = (float *)malloc(sizeof(float) * n * d);
    for (int64_t i = 0; i < n; i++)
        orc_residual(d, x + i * d, coarse + (int64_t)assign[i] * d, res + i * d);
    free(assign);
    float *sub = (float *)malloc(sizeof(float) * n * ds);
    for (int s = 0; s < m && !rc; s++) {
        for (int64_t i = 0; i < n; i++)
            memcpy(sub + i * ds, res + i * d + s * ds, sizeof(float) * ds);
        rc = orc_kmeans(sub, n, ds, ds, ks, niter_pq, seed + 1 + (uint64_t)s,
                        pq + (int64_t)s * ks * ds, nthreads);
    }
    free(sub);
    free(res);
    return rc;
}

/* ivfpq_assign: parallel/ivf_pq/ivf_training.cpp:165-224 + pq_assign
   parallel/pq-utils/pq_assign.cpp:41-50 (a = PQ centroid, b = residual) */
int orc_ivfpq_encode(int d, int nlist, int m, int ks, const float *coarse,
                     const float *pq, const float *x, int64_t n,
                     int32_t *assign, uint8_t *codes, int nthreads)
{
    if (d % m || ks > 256) return 1;
    int ds = d / m;
    set_threads(nthreads);
    assign_all(x, n, d, d, coarse, nlist, assign);
    float *pn = (float *)malloc(sizeof(float) * m * ks);
    for (int s = 0; s < m; s++)
        for (int j = 0; j < ks; j++)
            pn[s * ks + j] = norm_f32(pq + ((int64_t)s * ks + j) * ds, ds);
#pragma omp parallel
    {
        float *r = (float *)malloc(sizeof(float) * d);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; i++) {
            orc_residual(d, x + i * d, coarse + (int64_t)assign[i] * d, r);
            for (int s = 0; s < m; s++) {
                const float *b = r + s * ds;
                double nb = norm_f64(b, ds);
                uint64_t best = UINT64_MAX;
                for (int j = 0; j < ks; j++) {
                    float dd = yael_dist(pq + ((int64_t)s * ks + j) * ds, pn[s * ks + j],
                                         b, nb, ds);
                    uint64_t key = orc_ckey(dd, j);
                    if (key < best) best = key;
                }
                codes[i * m + s] = (uint8_t)(best & 0xFF);
            }
        }
        free(r);
    }
    free(pn);
    return 0;
}

/* stable bucketing by list (ivec_sort_index, parallel/yael_needs/sorting.c:264-277;
   ids ascending inside a list as in sequential/ivf_pq/ivf_assign.cpp:3-52) */
int orc_ivfpq_bucket(int nlist, int m, const int32_t *assign, const uint8_t *codes,
                     int64_t n, int32_t id0, int64_t *offsets, int32_t *ids,
                     uint8_t *codes_out)
{
    memset(offsets, 0, sizeof(int64_t) * (nlist + 1));
    for (int64_t i = 0; i < n; i++) {
        if (assign[i] < 0 || assign[i] >= nlist) return 1;
        offsets[assign[i] + 1]++;
    }
    for (int l = 0; l < nlist; l++) offsets[l + 1] += offsets[l];
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * nlist);
    memcpy(pos, offsets, sizeof(int64_t) * nlist);
    for (int64_t i = 0; i < n; i++) {
        int64_t p = pos[assign[i]]++;
        ids[p] = id0 + (int32_t)i;
        memcpy(codes_out + p * m, codes + i * m, m);
    }
    free(pos);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* exact FP64 ground truth and recall                                           */
/* ------------------------------------------------------------------------ */
void orc_bruteforce_f64(const float *base, int64_t nb, const float *q, int nq,
                        int d, int k, int32_t *ids, double *dis, int nthreads)
{
    set_threads(nthreads);
#pragma omp parallel
    {
        /* (distance, id) pairs kept sorted by insertion; k is small */
        double *bd = (double *)malloc(sizeof(double) * (k + 1));
        int32_t *bi = (int32_t *)malloc(sizeof(int32_t) * (k + 1));
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < nq; i++) {
            int cnt = 0;
            const float *x = q + (int64_t)i * d;
            for (int64_t j = 0; j < nb; j++) {
                const float *y = base + j * d;
                double s = 0;
                for (int t = 0; t < d; t++) {
                    double df = (double)x[t] - (double)y[t];
                    s += df * df;
                }
                if (cnt == k && s >= bd[k - 1]) continue;
                int p = cnt < k ? cnt : k - 1;
                while (p > 0 && bd[p - 1] > s) { bd[p] = bd[p - 1]; bi[p] = bi[p - 1]; p--; }
                bd[p] = s;
                bi[p] = (int32_t)j;
                if (cnt < k) cnt++;
            }
            for (int t = 0; t < k; t++) {
                ids[(int64_t)i * k + t] = t < cnt ? bi[t] : -1;
                dis[(int64_t)i * k + t] = t < cnt ? bd[t] : INFINITY;
            }
        }
        free(bd);
        free(bi);
    }
}

/* R@R as parallel/pq-utils/pq_test_compute_stats.cpp:13-72: percentage of
   queries whose ground-truth nearest neighbour is among the first R results */
double orc_recall_at(const int32_t *ids, int nq, int k, const int32_t *gt, int R)
{
    int hit = 0;
    int lim = R < k ? R : k;
    for (int i = 0; i < nq; i++) {
        for (int j = 0; j < lim; j++)
            if (ids[(int64_t)i * k + j] == gt[i]) { hit++; break; }
    }
    return 100.0 * (double)hit / (double)nq;
}
