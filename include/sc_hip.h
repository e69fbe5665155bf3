/*
 * sc_hip.h -- C ABI of the MI355X quantize -> encode -> search library
 * (libsc_amd.so).  Plain pointers and sizes, int status codes, no exceptions
 * and no torch types cross this boundary.
 *
 * Pointers named d_* are device (HBM) pointers on the context's device;
 * h_* are host pointers.  Every call is enqueued on the context's stream and
 * is asynchronous unless it says otherwise.
 *
 * Reference interfaces each entry point replaces are cited as path:line
 * relative to the reference repository root.
 */
#ifndef SC_HIP_H
#define SC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SC_OK 0
#define SC_ERR_ARG (-1)         /* bad argument or shape */
#define SC_ERR_NOMEM (-2)       /* device or host allocation failed */
#define SC_ERR_CONFIG (-3)      /* search kind does not match the index
                                   (the reference prints and returns:
                                   include/multi_query.h:71-74,
                                   include/sc_query.h:68-71) */
#define SC_ERR_HIP (-4)         /* HIP runtime error (see sc_last_error) */
#define SC_ERR_TOO_LARGE (-5)   /* index too large (src/sc_encoder.cpp:28-33) */
#define SC_ERR_NO_DEVICE (-6)   /* no gfx950 device */
#define SC_ERR_HEAP_EMPTY (-7)  /* pq_pop on an empty heap
                                   (include/sc_utilities.h:274-275) */

/* ---- search kinds ------------------------------------------------------ */
#define SC_KIND_IVFADC 0  /* PQQuery::search_ivfadc   include/query.h:202-325 */
#define SC_KIND_IMI 1     /* MultiQuery::search_multi2 include/multi_query.h:62-273 */
#define SC_KIND_SC 2      /* SCQuery::search_mr_ivf    include/sc_query.h:60-256 */
#define SC_KIND_SC3 3     /* SCQuery::search_mr_ivf3   include/sc_query.h:265-475 */
#define SC_KIND_SCIMI 4   /* (nearest, second) cells of two half-blocks
                             (mc = 2, nc = 2): the cell key of
                             src/sc_encoder.cpp:39-50 over a sparse directory
                             of the non-empty cells (the reference's dense
                             kc^4 array aborts, src/sc_encoder.cpp:26-33, and
                             it has no search for it).  Cells are taken in
                             order of score = (v0[n0] + v0[s0]) +
                             (v1[n1] + v1[s1]) -- each half's SC pair score
                             (include/sc_query.h:113-166) summed as the IMI
                             sums its lists (include/multi_query.h:175-203),
                             v_h[j] = q_h + diff_qc[h*kc + j] -- ties by cell
                             key, while count < w and sum < T; each cell is
                             scanned as search_multi2 scans
                             (include/multi_query.h:209-262) with the nearest
                             centroid of each half. */

/* ---- precision / forms ------------------------------------------------- */
#define SC_PREC_FP32 0    /* the reference's arithmetic everywhere */
#define SC_PREC_BF16 1    /* MFMA bf16 distance GEMM + exact re-check */
#define SC_FORM_PLAIN 0   /* FP32 LUT scan */
#define SC_FORM_MIXED 1   /* FP16 LUT scan + exact FP32 re-rank: ids and
                             distances bit-identical to SC_FORM_PLAIN.
                             IVF/SC: FP16 combined tables; IMI and SCIMI
                             (mp = 8 or 16): an FP16 diff_qr table plus each
                             vector's stored dot_cr sum.  SC_KIND_SC3 has no
                             mixed scan and refuses it with SC_ERR_ARG rather
                             than run the plain scan */
#define SC_FORM_REAL 2    /* the reference's real_dist = true: every gathered
                             candidate ranked by its exact distance to its raw
                             vector (include/query.h:307-312,
                             include/multi_query.h:255-259,
                             include/sc_query.h:232-237); needs
                             sc_index_attach_raw */

typedef struct sc_ctx sc_ctx;
typedef struct sc_index sc_index;

/* Last error message of the calling thread ("" when none). */
const char *sc_last_error(void);
/* Library build string (arch, flags). */
const char *sc_version(void);

/* ---- context (no reference counterpart: owns the device and stream) ---- */
int sc_ctx_create(int device, sc_ctx **out);
int sc_ctx_destroy(sc_ctx *ctx);
int sc_ctx_sync(sc_ctx *ctx);
/* hipStream_t of the context, as void*. */
void *sc_ctx_stream(sc_ctx *ctx);

/* Device memory for hosts without their own allocator. */
int sc_malloc(sc_ctx *ctx, size_t bytes, void **d_ptr);
int sc_free(sc_ctx *ctx, void *d_ptr);
int sc_memcpy_h2d(sc_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int sc_memcpy_d2h(sc_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int sc_memset(sc_ctx *ctx, void *d_dst, int value, size_t bytes);

/* ---- seeded synthetic data --------------------------------------------- */
/* Gaussian mixture from Philox4x32-10 counters, integer-only arithmetic so
 * that any row range regenerates bit-identically on host or device.  Stands
 * in for the reference's fvecs/bvecs inputs (include/sc_utilities.h:816-886,
 * 1077-1143). */
typedef struct {
    uint32_t seed;
    int32_t ncomp;
    int32_t frac_bits;
    int32_t out_bits;
    int64_t mu_lo, mu_span;
    int64_t sig_lo, sig_span;
    int64_t maxq;
} sc_mix;
int sc_gen_rows(sc_ctx *ctx, const sc_mix *mix, uint32_t stream, int64_t row0,
                int64_t nrows, int32_t dim, float *d_out);

/* ---- encode ------------------------------------------------------------ */
typedef struct {
    int64_t rechecked;   /* SC_PREC_BF16: codewords re-evaluated exactly, summed
                            over rows and blocks */
    int64_t fallback;    /* (row, block) pairs whose candidate set overflowed and
                            were evaluated exactly over the whole codebook */
} sc_encode_stats;

/* Coarse assignment + residual + PQ codes for n vectors.
 *   nc = 1: Encoder::encode<T>   include/encoder.h:69-192
 *           cid [n][mc], nearest per block, lowest index on ties.
 *   nc = 2: SCEncoder::encode2<T> include/sc_encoder.h:63-195
 *           cid [n][2*mc]: nearest of each block, then second of each block;
 *           top-2 by (dist ASC, index DESC).
 * Codebooks: d_cq [mc][kc][dim/mc], d_pq [mp][kp][dim/mp] (float).
 * Codes d_codes [n][mp] uint8.  `prec` SC_PREC_FP32 or SC_PREC_BF16; both
 * produce identical cid and codes.  `st` (host, may be NULL) is filled after
 * an implicit stream synchronisation. */
int sc_encode(sc_ctx *ctx, const float *d_x, int64_t n, int32_t dim,
              const float *d_cq, int32_t mc, int32_t kc, const float *d_pq,
              int32_t mp, int32_t kp, int32_t nc, int32_t prec,
              uint16_t *d_cid, uint8_t *d_codes, sc_encode_stats *st);

/* SCEncoder::encode<T> with general nc (include/sc_encoder.h:205-332), nc in
 * 1..3: every distance is stored in a float, the nc nearest coarse centroids
 * of each block are taken by float distance with ties to the lower index
 * (the reference's nth_element_id + sort_id, whose libstdc++ tie order this
 * replaces), cid [n][mc][nc]; residual against the nearest; PQ argmin of
 * float distances, the lower index on ties.  Exact (FP64 accumulation of
 * the float differences' squares, then rounded to float). */
int sc_encode_general(sc_ctx *ctx, const float *d_x, int64_t n, int32_t dim,
                      const float *d_cq, int32_t mc, int32_t kc, const float *d_pq,
                      int32_t mp, int32_t kp, int32_t nc, uint16_t *d_cid, uint8_t *d_codes);

/* PQ codes against GIVEN coarse cells: Encoder::rencode<T>
 * (include/encoder.h:202-309), which re-encodes the vectors of a loaded
 * inverted file (Encoder::load_encoded_data, src/encoder.cpp:57-146) in its
 * cell order.  Position i encodes raw row d_rows[i] of d_x [nx][dim]
 * (d_rows NULL: row i, n <= nx) against the cells d_cid [n][mc]: float
 * residual per coarse block, then Encoder::encode's PQ argmin (S1
 * distance, first strict minimum).  d_codes [n][mp].  A row outside
 * [0, nx) or a cell id >= kc returns SC_ERR_ARG (the reference reads out
 * of bounds there).  Synchronises the context stream. */
int sc_encode_given(sc_ctx *ctx, const float *d_x, int64_t nx, int32_t dim,
                    const int32_t *d_rows, int64_t n, const uint16_t *d_cid,
                    const float *d_cq, int32_t mc, int32_t kc, const float *d_pq,
                    int32_t mp, int32_t kp, uint8_t *d_codes);

/* ---- index ------------------------------------------------------------- */
/* Builds the device-resident inverted file from encoder outputs:
 * Encoder::distribution (src/encoder.cpp:170-205) / SCEncoder::distribution
 * (src/sc_encoder.cpp:17-81) + the loaders' flattening (src/query.cpp:82-165,
 * src/sc_query.cpp:36-123) + PQQuery::pre_compute1 (include/query.h:91-149).
 * The index copies the codebooks; d_cid/d_codes may be freed afterwards.
 * Dense cells: kc^mc (nc=1), kc^2 (nc=2, mc=1) or kc^3 (nc=3, mc=1, kc <= 1024;
 * d_cid in sc_encode_general's layout [n][3]; SCEncoder::distribution with
 * the per-vector cid stride mc*nc that src/sc_encoder.cpp:41-59 gets wrong
 * for nc != 2).  nc=2 with mc=2 (two halves, SC_KIND_SCIMI; kc <= 65535)
 * keeps only the non-empty cells: a sparse directory of 64-bit keys
 * uid_0*kc^2 + uid_1, uid_h = nearest_h*kc + second_h, ascending, with L
 * the inclusive prefix over those cells (sc_index_shape reports their
 * number, sc_index_cell_keys their keys). */
int sc_index_build(sc_ctx *ctx, const uint16_t *d_cid, const uint8_t *d_codes,
                   int64_t n, int32_t dim, const float *d_cq, int32_t mc,
                   int32_t kc, const float *d_pq, int32_t mp, int32_t kp,
                   int32_t nc, sc_index **out);
/* Loads an index from host arrays already in the loaders' layout:
 * L inclusive prefix [ncells], pid [n], codes [n][mp] in cell order. */
int sc_index_load(sc_ctx *ctx, const int32_t *h_L, int64_t ncells,
                  const int32_t *h_pid, const uint8_t *h_codes, int64_t n,
                  int32_t dim, const float *h_cq, int32_t mc, int32_t kc,
                  const float *h_pq, int32_t mp, int32_t kp, int32_t nc,
                  sc_index **out);
/* The two-halves index (mc = 2, nc = 2) from host arrays: keys [ncells]
 * strictly ascending below kc^4, L [ncells] the inclusive prefix over those
 * (non-empty) cells, pid [n], codes [n][mp] in cell order; h_cq
 * [2][kc][dim/2].  Validated on the host before any device work. */
int sc_index_load_sparse(sc_ctx *ctx, const uint64_t *h_ckeys, const int32_t *h_L,
                         int64_t ncells, const int32_t *h_pid, const uint8_t *h_codes,
                         int64_t n, int32_t dim, const float *h_cq, int32_t kc,
                         const float *h_pq, int32_t mp, int32_t kp, sc_index **out);
/* The sparse directory's keys (two-halves index only), [ncells]. Synchronous. */
int sc_index_cell_keys(sc_ctx *ctx, const sc_index *ix, uint64_t *h_keys);
int sc_index_destroy(sc_index *ix);
/* Gives the index the raw base vectors d_raw [n][dim] (float, in vector id
 * order) for SC_FORM_REAL searches: PQQuery::load_data
 * (include/query.h:83-86).  The index keeps the pointer, not a copy; the
 * buffer must outlive the searches that use it.  d_raw = NULL detaches. */
int sc_index_attach_raw(sc_index *ix, const float *d_raw, int64_t n);
/* Shape: out6 = {n, ncells, mc, kc, mp, kp} */
int sc_index_shape(const sc_index *ix, int64_t *out6);
/* Copies the flattened index back (any pointer may be NULL):
 * L [ncells], pid [n], codes [n][mp], norm_c [mc*kc], norm_r [mp*kp],
 * dot_cr [kc*mp*kp]. Synchronous. */
int sc_index_export(sc_ctx *ctx, const sc_index *ix, int32_t *h_L,
                    int32_t *h_pid, uint8_t *h_codes, float *h_norm_c,
                    float *h_norm_r, float *h_dot_cr);

/* ---- search ------------------------------------------------------------ */
typedef struct {
    int32_t kind;   /* SC_KIND_* */
    int32_t form;   /* SC_FORM_* */
    int32_t R;      /* results per query */
    int32_t w;      /* cell budget (heap capacity; IVF/IMI popped, SC non-empty) */
    int32_t T;      /* candidate budget */
    int32_t M;      /* coarse list length (lists are fully sorted; see notes) */
} sc_search_params;

typedef struct {
    int64_t queries;
    int64_t candidates;   /* sum over queries of gathered candidates */
    int64_t cells;        /* non-empty cells scanned */
    int64_t visited;      /* cells popped (two halves: cells scored by the ranking) */
    int64_t reranked;     /* candidates re-ranked exactly (SC_FORM_MIXED) */
    int64_t heap_fallback;/* queries whose heap exceeded on-chip capacity */
    int64_t exhausted;    /* queries whose heap emptied before the budget */
    /* stage times on the context stream (HIP events), milliseconds */
    double ms_lut, ms_rank, ms_traverse, ms_scan;
    /* fused kernel only: time per phase summed over workgroups, divided by
     * the number of workgroups (milliseconds): 0 LUT load, 1 coarse sort,
     * 2 traversal, 3 scan, 4 final selection and output */
    double fused_phase_ms[5];
    /* staged kinds: the scan kernels alone (HIP events around the scan
     * launches of each pass, with the scan of re-ranked queries that
     * overlaps it), milliseconds; 0 for the fused kernel */
    double ms_scan_kernel;
} sc_search_stats;

/* Batched search of nq queries d_q [nq][dim]; outputs d_ids [nq][R],
 * d_dists [nq][R] sorted by (dist ASC, position ASC); when a query gathers
 * fewer than R candidates the first `sum` entries keep traversal order and
 * the rest are -1 / +inf.  d_sums [nq] (may be NULL) = candidates gathered.
 * Replaces search_ivfadc / search_multi2 / search_mr_ivf called once per
 * query.  `st` (host, may be NULL) is filled after a synchronisation. */
int sc_search(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
              const sc_search_params *p, int32_t *d_ids, float *d_dists,
              int32_t *d_sums, sc_search_stats *st);
/* The same search with a candidate budget per query (no reference
 * counterpart: the reference takes one T per call, include/sc_query.h:60-66;
 * a batch here can mix budgets, e.g. drawn from [1k, 30k]): d_T [nq] int32 on
 * the device; p->T is the largest budget, and each d_T[q] is clamped to
 * [0, p->T].  Query q's results equal a sc_search of that query alone with
 * T = d_T[q]. */
int sc_search_budgets(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                      const sc_search_params *p, const int32_t *d_T, int32_t *d_ids,
                      float *d_dists, int32_t *d_sums, sc_search_stats *st);

/* ---- batched search plans (no reference counterpart) ------------------- */
/* A plan binds one batch shape (nq queries at d_q, outputs at d_ids/d_dists/
 * d_sums) and captures the persistent fused search of that batch (work-queue
 * reset + one persistent launch) in a hipGraph that sc_plan_execute replays.
 * The buffers' contents may change between executions; their addresses may
 * not.  Kinds whose traversal does not fit on chip (IMI) run the staged
 * kernels instead of a graph.  Queries the graph flags as exceeding on-chip
 * capacity are finished by the staged kernels inside sc_plan_execute, which
 * synchronises the context stream.  This batches what the reference does one
 * query per call (include/sc_query.h:60-256, include/multi_query.h:62-273). */
typedef struct sc_plan sc_plan;
int sc_plan_create(sc_ctx *ctx, const sc_index *ix, const sc_search_params *p,
                   const float *d_q, int64_t nq, int32_t *d_ids, float *d_dists,
                   int32_t *d_sums, sc_plan **out);
/* A plan whose queries carry their own budgets d_T [nq] (as in
 * sc_search_budgets); the budgets' contents may change between executions. */
int sc_plan_create_budgets(sc_ctx *ctx, const sc_index *ix, const sc_search_params *p,
                           const float *d_q, const int32_t *d_T, int64_t nq, int32_t *d_ids,
                           float *d_dists, int32_t *d_sums, sc_plan **out);
int sc_plan_execute(sc_plan *plan, sc_search_stats *st);
/* Times the fused kernel over launch shapes (register variant x workgroups in
 * flight) on the plan's own batch, keeps the fastest and re-captures the
 * graph.  out4 (may be NULL) = {threads, grid, best_us, variant}; the variant
 * is the number of waves per SIMD its register allocation allows. */
int sc_plan_autotune(sc_plan *plan, int32_t *out4);
int sc_plan_destroy(sc_plan *plan);

/* ---- k-means (codebook training) --------------------------------------- */
/* k-means++ seeding + `iters` Lloyd iterations on d_x [n][d]:
 * PQQuantizer::create_sub_quantizers -> greg_kmeans
 * (include/quantizer.h:235-272, lib/simple-cluster/include/k-means.h:162-267,
 * 553-912) with a seeded sampler and FP64 centroid sums.  Outputs
 * d_centers [k][d], d_labels [n] (may be NULL). */
int sc_kmeans(sc_ctx *ctx, const float *d_x, int64_t n, int32_t d, int32_t k,
              int32_t iters, uint32_t seed, int32_t lagged, int32_t prec,
              float *d_centers, int32_t *d_labels);

/* Residuals for PQ training, in place on d_x [n][dim]:
 * PQQuantizer::calc_residual_vector (include/quantizer.h:456-510, which=0)
 * and SCQuantizer::calc_residual_vector (include/sc_quantizer.h:72-133,
 * which=1, second-nearest). */
int sc_residual(sc_ctx *ctx, float *d_x, int64_t n, int32_t dim,
                const float *d_cq, int32_t mc, int32_t kc, int32_t which);

/* ---- evaluation ---------------------------------------------------------- */
/* Nearest base vector of each query by the exact S1 distance, the lowest id
 * on ties (the first strict minimum of a scan in id order): the first column
 * of the ground-truth files Evaluation::load_groundtruth reads and
 * calc_recall matches results against (include/evaluation.h:34-92; the
 * reference ships those files, it does not compute them).  d_base [n][dim],
 * d_q [nq][dim] (float); d_ids [nq]; d_dists [nq] (double, may be NULL).
 * Synchronises the context stream. */
int sc_groundtruth_nn1(sc_ctx *ctx, const float *d_base, int64_t n, const float *d_q,
                       int64_t nq, int32_t dim, int32_t *d_ids, double *d_dists);
/* sc_groundtruth_nn1 without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-call partial results live in a fixed 16 MiB device array of the library
 * (one per GPU, held until the process ends), so nq * chunks * 12 bytes must
 * fit: nq is limited to 1398101 queries per call (SC_ERR_ARG beyond), and
 * fewer base chunks are used for large nq.  Results are identical to
 * sc_groundtruth_nn1.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other threads, other streams) race. */
int sc_groundtruth_nn1_noalloc(sc_ctx *ctx, const float *d_base, int64_t n, const float *d_q,
                               int64_t nq, int32_t dim, int32_t *d_ids, double *d_dists);

#ifdef __cplusplus
}
#endif
#endif /* SC_HIP_H */
