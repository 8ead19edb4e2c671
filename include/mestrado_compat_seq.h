/*
 * mestrado_compat_seq.h -- drop-in for the reference's single-process search
 * with RERANK (rafaelms101/mestrado, sequential/), implemented on libmivf.
 *
 * Types have the sequential fork's layouts (LP64):
 *   mat / matI   sequential/pq-utils/pq_test_load_vectors.h
 *   pqtipo       sequential/pq-utils/pq_new.h:24-31 (centroids as float*[nsq],
 *                each [ks][ds])
 *   ivfpq_t, ivf_t   sequential/ivf_pq/myIVF.h:16-29
 * Do not include together with mestrado_compat.h: the forks define pqtipo
 * differently, exactly as the reference does.
 */
#ifndef MESTRADO_COMPAT_SEQ_H
#define MESTRADO_COMPAT_SEQ_H

typedef struct {
    float *mat;
    int n, d;
} mat;

typedef struct {
    int *mat;
    int n, d;
} matI;

typedef struct pqtipo {
    int nsq;
    int ks;
    int ds;
    float **centroids;
    int centroidsn;
    int centroidsd;
} pqtipo;

typedef struct ivfpq {
    pqtipo pq;
    int coarsek;
    float *coa_centroids;
    int coa_centroidsn;
    int coa_centroidsd;
} ivfpq_t;

typedef struct ivf {
    int *ids;
    int idstam;
    matI codes;
} ivf_t;

/* sequential/ivf_pq/ivf_search.cpp:9-153: for each query, ADC shortlist of
 * kl over the w nearest lists, then exact FP32 distances to the raw base
 * rows vbase.mat [vbase.n][d], best k by (distance, id) into ids[q*k + c].
 * Deviations from the reference: when a shortlist holds fewer than k
 * vectors the row is padded with id -1 (the reference asserts, defect D7),
 * and when dis is not NULL the exact distances are written to dis[q*k + c]
 * (the reference leaves dis untouched).                                    */
void ivfpq_search(ivfpq_t ivfpq, ivf_t *ivf, mat vquery, mat vbase, int k, int kl, int w, int *ids,
                  float *dis);

#endif /* MESTRADO_COMPAT_SEQ_H */
