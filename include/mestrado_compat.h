/*
 * mestrado_compat.h -- drop-in C++ entry points with the signatures of the
 * reference's GPU fork (rafaelms101/mestrado, parallel2/), implemented on
 * libmivf.  A caller of the reference includes this header instead of
 * mycuda.h / ivf_training.h and links libmivf.so.
 *
 * Types have the reference's layouts (LP64):
 *   mat / matI          parallel2/pq-utils/pq_test_load_vectors.h:19-27
 *   pqtipo              parallel2/pq-utils/pq_new.h:26-33 (flat [nsq][ks][ds])
 *   ivfpq_t, ivf_t      parallel2/ivf_pq/myIVF.h:21-33
 * Functions keep C++ linkage, as in the reference, so the mangled names match.
 * Errors print a message to stderr and exit(EXIT_FAILURE), as the reference
 * does (parallel2/ivf_pq/mycuda.cu:11-14).
 */
#ifndef MESTRADO_COMPAT_H
#define MESTRADO_COMPAT_H

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

/* Device boundary, parallel2/ivf_pq/mycuda.h:7-9.
 * preallocate_gpu_mem: PQ and every list become device-resident (codes are
 * validated to fit 8 bits and packed).  core_gpu: residual.mat [nq*w][d],
 * rid_to_ivf [nq*w]; per query min(n_q, k) (id, distance) pairs, ascending
 * distance, ties by lower id, written from qid_to_starting_outid[q].       */
void core_gpu(pqtipo PQ, mat residual, ivf_t *ivf, int ivf_size, int *rid_to_ivf,
              int *qid_to_starting_outid, matI idxs, mat dists, int k, int w);
void preallocate_gpu_mem(pqtipo PQ, ivf_t *ivf, int ivf_size);
void deallocate_gpu_mem();

/* Index build, parallel2/ivf_pq/ivf_training.h: ivfpq_new trains the coarse
 * quantizer (50 k-means iterations) and the PQ (100 per subquantizer) with
 * GPU assignment steps.  ks is 256 (8-bit codes); the reference's
 * parallel forks set ks = coarsek (pq_new.cpp:36), which does not fit 8 bits.
 * `threads` is accepted for signature compatibility.                       */
ivfpq_t ivfpq_new(int coarsek, int nsq, mat vtrain, int threads);
void ivfpq_assign(ivfpq_t ivfpq, mat vbase, ivf_t *ivf);
void write_cent(char *file, char *file2, char *file3, ivfpq_t ivfpq);
void read_cent(char *file, char *file2, char *file3, ivfpq_t *ivfpq);
/* Per-list files <ivf_path>/<j> (parallel2/ivf_pq/ivf_search.cpp:293). */
ivf_t *read_ivf(ivfpq_t ivfpq, int tam, int my_rank, char *ivf_path);

#endif /* MESTRADO_COMPAT_H */
