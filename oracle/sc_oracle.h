/*
 * sc_oracle.h -- CPU oracle for the quantize -> encode -> search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry and bench.py's CPU-baseline leg may load it.  The product
 * library (second-candidate_amd/) never links or calls it.
 *
 * Each function restates the reference algorithm it cites (paths are relative
 * to the reference repository root).  Where the reference has undefined or
 * implementation-defined behaviour the restatement fixes one documented
 * choice:
 *   - include/query.h:167,179 store `diff[base] = norm[base++] - d - d`; the
 *     index-aligned meaning (diff[j] = (norm[j] - s) - s) is used.
 *   - cblas_sdot (un-vendored OpenBLAS fork) is replaced by a sequential FP32
 *     dot product with separately rounded multiply and add (no FMA).
 *   - nth_element_id/sort_id (include/sc_algorithm.h:24-75) order ties by
 *     libstdc++ internals; here coarse lists are fully sorted by
 *     (score ASC, centroid ASC) and the top-R by (dist ASC, position ASC),
 *     position = index in the candidate list (traversal order).
 *   - A traversal whose heap runs empty before its budget is met spins forever
 *     in the reference (include/multi_query.h:151, include/sc_query.h:130);
 *     here it stops and reports `exhausted`.
 *
 * Build flags must keep FP contraction off (-ffp-contract=off) so that every
 * float sum is a separately rounded multiply and add.
 */
#ifndef SC_ORACLE_H
#define SC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- synthetic data (Philox4x32-10 counter-based) ---------- */
typedef struct {
    uint32_t seed;      /* Philox key word 0 */
    int32_t ncomp;      /* mixture components */
    int32_t frac_bits;  /* F: mu/sigma are fixed point with F fractional bits */
    int32_t out_bits;   /* O: output grid is 2^-O */
    int64_t mu_lo, mu_span;    /* mu_q = mu_lo + u % (mu_span + 1) */
    int64_t sig_lo, sig_span;  /* sig_q = sig_lo + u % (sig_span + 1) */
    int64_t maxq;       /* clamp of the output integer */
} sco_mix;

void sco_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void sco_gen_rows(const sco_mix *m, uint32_t stream, int64_t row0, int64_t nrows,
                  int32_t dim, float *out, int32_t nthreads);

/* ---------------- distances and encoders ------------------------------- */
double sco_dist_l2sq(const float *x, const float *y, int32_t d);

/* nc = 1: Encoder::encode (coarse argmin + PQ); nc = 2: SCEncoder::encode2.
 * cid layout: nc=1 [N][mc]; nc=2 [N][2*mc] (nearest blocks, then seconds). */
int sco_encode(const float *x, int64_t n, int32_t dim,
               const float *cq, int32_t mc, int32_t kc,
               const float *pq, int32_t mp, int32_t kp, int32_t nc,
               uint16_t *cid, uint8_t *codes, int32_t nthreads);

/* SCEncoder::encode with general nc (1..3): float-rounded distances, the nc
 * nearest by (float dist ASC, index ASC); cid [N][mc][nc]. */
int sco_encode_general(const float *x, int64_t n, int32_t dim,
                       const float *cq, int32_t mc, int32_t kc,
                       const float *pq, int32_t mp, int32_t kp, int32_t nc,
                       uint16_t *cid, uint8_t *codes, int32_t nthreads);

/* Cell key per vector (64-bit, S4); nc = 3 reads the general encoder's
 * layout [N][mc][3]. */
void sco_cell_keys(const uint16_t *cid, int64_t n, int32_t mc, int32_t kc,
                   int32_t nc, uint64_t *keys);

/* Dense inverted file (Encoder::distribution / SCEncoder::distribution and the
 * loaders' inclusive prefix L[]): ids ascending within a cell. */
int sco_build_dense(const uint64_t *keys, int64_t n, int64_t ncells,
                    const uint8_t *codes, int32_t mp, int32_t *L, int32_t *pid,
                    uint8_t *codes_sorted);

/* Sparse inverted file: the distinct keys ascending (ckeys [ncells]), their
 * inclusive prefix L [ncells], ids ascending within a cell (pid [n]) and the
 * codes in cell order.  ckeys/L need room for n entries; returns ncells (or a
 * negative error). */
int64_t sco_build_sparse(const uint64_t *keys, int64_t n, const uint8_t *codes, int32_t mp,
                         uint64_t *ckeys, int32_t *L, int32_t *pid, uint8_t *codes_sorted);

/* PQQuery::pre_compute1 tables. */
void sco_tables(const float *cq, const float *pq, int32_t dim, int32_t mc,
                int32_t kc, int32_t mp, int32_t kp, float *norm_c,
                float *norm_r, float *dot_cr);

/* PQQuery::pre_compute2 LUTs. */
void sco_lut(const float *q, const float *cq, const float *pq,
             const float *norm_c, const float *norm_r, int32_t dim, int32_t mc,
             int32_t kc, int32_t mp, int32_t kp, float *diff_qc, float *diff_qr);

/* ---------------- search ---------------------------------------------- */
typedef struct {
    int32_t dim, mc, kc, mp, kp;
    const float *cq, *pq, *norm_c, *norm_r, *dot_cr;
    int64_t ncells;      /* dense cells: kc^mc (IVF/IMI), kc^2 (SC) or kc^3 (SC3) */
    const int32_t *L;    /* inclusive prefix, [ncells] */
    const int32_t *pid;  /* [N] */
    const uint8_t *codes;/* [N][mp], cell order */
    const uint64_t *ckeys; /* sparse directory: the key of each of the ncells
                              non-empty cells, ascending (NULL: dense cells,
                              cell = key) */
} sco_index;

typedef struct {
    int32_t sum;        /* candidates gathered */
    int32_t ncells;     /* non-empty cells scanned */
    int32_t nvisited;   /* cells popped */
    int32_t empty;      /* empty cells popped (IMI `e`) */
    int32_t exhausted;  /* heap emptied before the budget */
} sco_stats;

/* Search kinds */
#define SCO_IVFADC 0  /* PQQuery::search_ivfadc   (include/query.h:202-325) */
#define SCO_IMI    1  /* MultiQuery::search_multi2 (include/multi_query.h:62-273) */
#define SCO_SC     2  /* SCQuery::search_mr_ivf    (include/sc_query.h:60-256) */
#define SCO_SC3    3  /* SCQuery::search_mr_ivf3   (include/sc_query.h:265-475) */
/* (nearest, second) cells of two half-blocks (mc = 2, nc = 2): the cell key of
 * src/sc_encoder.cpp:39-50 in 64 bits over a sparse directory; non-empty cells
 * are taken in order of the sum of the halves' SC pair scores
 * (include/sc_query.h:113-166 per half, include/multi_query.h:175-203 across
 * them), then scanned as the IMI scans (include/multi_query.h:209-262).  The
 * reference builds no such index (src/sc_encoder.cpp:26-33 aborts) and has no
 * such search; the definition is stated above scimi_rank in sc_oracle.c. */
#define SCO_SCIMI  4

/* SCO_SCIMI directories of at least n cells are ranked by selection (only
 * the cells of the histogram buckets up to the budget are sorted), smaller
 * ones by sorting every cell; both give the same cells.  Default 2^20. */
void sco_scimi_set_select_min(int64_t n);

/* One query.  ids/dists: [R] (padded with -1 / +inf when sum < R, in which
 * case the first `sum` entries keep traversal order like the reference).
 * Optional (may be NULL): cells [w] bid of scanned cells in order, s1/s2 [w],
 * all_ids/all_dists [cap_all] every candidate in scan order.
 * Returns 0, or a negative error code. */
int sco_search(const sco_index *ix, int32_t kind, const float *q, int32_t R,
               int32_t w, int32_t T, int32_t M, int32_t *ids, float *dists,
               sco_stats *st, int32_t *cells, int32_t *s1, int32_t *s2,
               int32_t *all_ids, float *all_dists, int64_t cap_all);

/* real_dist = true: the candidates of sco_search ranked by their S1 distance
 * to the raw vectors raw [n][dim] (by id). */
int sco_search_real(const sco_index *ix, int32_t kind, const float *q, const float *raw,
                    int32_t R, int32_t w, int32_t T, int32_t *ids, float *dists,
                    sco_stats *st);

/* Many queries (OpenMP over queries), ids/dists [nq][R], stats [nq] (NULL ok) */
int sco_search_batch(const sco_index *ix, int32_t kind, const float *q,
                     int64_t nq, int32_t R, int32_t w, int32_t T, int32_t M,
                     int32_t *ids, float *dists, sco_stats *st, int32_t nthreads);

/* Heap replica (include/sc_utilities.h:469-503, 546-585) over
 * (score, h1..h4) entries; exposed for the known-answer tests. */
typedef struct {
    float *v;
    int32_t *h1, *h2, *h3, *h4;
    int32_t pos, cap;
} sco_heap4;
void sco_heap4_insert(sco_heap4 *h, float v, int32_t a, int32_t b, int32_t c,
                      int32_t d);
int sco_heap4_pop(sco_heap4 *h, int32_t *a, int32_t *b, int32_t *c, int32_t *d);

/* Single-id heap of search_ivfadc (include/sc_utilities.h:229-300); pop
 * returns the id, or -1 on an empty heap. */
typedef struct {
    float *v;
    int32_t *id;
    int32_t pos, cap;
} sco_heap1;
void sco_heap1_insert(sco_heap1 *h, float v, int32_t id);
int32_t sco_heap1_pop(sco_heap1 *h);

/* Canonical top-R over a candidate list (dist ASC, position ASC): writes the
 * R selected positions (sorted) to pos_out; returns min(R, n). */
int32_t sco_topr(const float *d, int32_t n, int32_t R, int32_t *pos_out);

/* Exact 1-NN (lowest id on ties) and recall@R (include/evaluation.h:58-92). */
void sco_bruteforce_nn1(const float *base, int64_t n, const float *q, int64_t nq,
                        int32_t dim, int32_t *gt, int32_t nthreads);
double sco_recall(const int32_t *gt, const int32_t *ids, int64_t nq, int32_t R,
                  int32_t stride);

/* ---------------- k-means++ / Lloyd ------------------------------------- */
/* Deterministic k-means++ seeding (D^2 weights floor(d*2^16) in uint64) and
 * `iters` Lloyd iterations with FP64 centroid sums, float-rounded S1
 * distances, ties to the later centroid, SINGLETON empty-cluster repair.
 * lagged=1 reproduces the reference's one-seed lag (k-means.h:250). */
int sco_kmeans(const float *x, int64_t n, int32_t d, int32_t k, int32_t iters,
               uint32_t seed, int32_t lagged, float *centers, int32_t *labels,
               int32_t nthreads);
/* Seeding only: writes k seed indices. */
int sco_kmeans_pp(const float *x, int64_t n, int32_t d, int32_t k, uint32_t seed,
                  int32_t lagged, int64_t *seed_idx, int32_t nthreads);

/* Residuals for PQ training (PQQuantizer / SCQuantizer::calc_residual_vector):
 * which = 0 nearest (float compare, strict >), which = 1 second. In place. */
void sco_residual(float *x, int64_t n, int32_t dim, const float *cq, int32_t mc,
                  int32_t kc, int32_t which, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
