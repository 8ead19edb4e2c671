/*
 * mestrado_compat_par.h -- drop-in for the per-list search calls of the
 * reference's MPI/OpenMP CPU fork (rafaelms101/mestrado, parallel/),
 * implemented on libmivf's kernels.
 *
 * Types have the parallel fork's layouts (LP64):
 *   mat / matI   parallel/pq-utils/pq_test_load_vectors.h:19-27
 *   pqtipo       parallel/pq-utils/pq_new.h:26-33 (flat [nsq][ks][ds])
 *   dis_t, ivfpq_t, ivf_t (with the per-vector `dis` field)
 *                parallel/ivf_pq/myIVF.h:16-34
 * Do not include together with mestrado_compat.h or mestrado_compat_seq.h:
 * each fork defines ivf_t (and the sequential fork pqtipo) its own way,
 * exactly as the reference does.  Functions keep C++ linkage, so the mangled
 * names equal the reference's.  Errors print a message to stderr and
 * exit(EXIT_FAILURE).
 *
 * The calls are safe from several host threads at once (the reference calls
 * them from its OpenMP search threads, parallel/ivf_pq/ivf_search.cpp:97-137);
 * the device work is serialised.  Each call is one small launch: for batches
 * of queries the throughput path is mivf_search (include/mivf.h).
 */
#ifndef MESTRADO_COMPAT_PAR_H
#define MESTRADO_COMPAT_PAR_H

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
    float *centroids;
    int centroidsn;
    int centroidsd;
} pqtipo;

typedef struct {
    mat dis;
    matI idx;
} dis_t;

typedef struct ivfpq {
    pqtipo pq;
    int coarsek;
    float *coa_centroids;
    int coa_centroidsn;
    int coa_centroidsd;
} ivfpq_t;

typedef struct ivf {
    int *ids;
    float *dis;
    int idstam;
    matI codes;
} ivf_t;

/* parallel/ivf_pq/ivf_search.cpp:178-228: the ADC distances of every vector
 * of list `centroid_idx` to `residual` (pq.nsq * pq.ds floats): q.dis.mat[i]
 * and q.idx.mat[i] = ivf[centroid_idx].ids[i], in list order, both malloc'd
 * (the caller frees them, as in the reference).  The table is the
 * reference's compute_cross_distances(ds, 1, ks, r_s, P_s) and the sum
 * sumidxtab2's, computed on the GPU with the same FP32 arithmetic.  Codes
 * (int32 per subquantizer) must be < ks <= 256; each list's codes are copied
 * to the device at its first search and kept there, keyed by the address of
 * ivf[centroid_idx].codes.mat, until mestrado_par_release().  *g1 and *g2
 * accumulate microseconds as the reference's timers do: *g1 the residual
 * upload and table + ADC kernel, *g2 the copy of the distances back.      */
dis_t ivfpq_search(ivf_t *ivf, float *residual, pqtipo pq, int centroid_idx, double *g1, double *g2);

/* parallel/ivf_pq/k_min.cpp:3: the ktmp smallest (distance, id) pairs of
 * q, ascending, ties by lower id, into dis[0..ktmp) and ids[0..ktmp).  The
 * reference's heap does not return the ktmp smallest (its sift-down compares
 * against the wrong child and never shrinks the heap); this is the exact
 * selection it was written to be.  ktmp <= min(q.dis.n, 1024).             */
void my_k_min(dis_t q, int ktmp, float *dis, int *ids);

/* No reference counterpart: drop the device copies of the lists and of the
 * PQ kept by ivfpq_search (call it before lists are changed in place).    */
void mestrado_par_release(void);

#endif /* MESTRADO_COMPAT_PAR_H */
