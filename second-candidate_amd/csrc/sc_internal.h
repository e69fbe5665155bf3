// sc_internal.h -- shared definitions of libsc_amd (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sc_hip.h"

struct sc_ctx {
    int device;
    hipStream_t stream;
    void *scratch;          // growable device scratch for temporaries
    size_t scratch_bytes;
    int n_cu;               // compute units
    void *fixed_scratch;    // this device's copy of the library's static scratch
    size_t fixed_scratch_bytes;
    // a second stream (default priority) for work that overlaps the context
    // stream's, and two events ordering them; created on first use
    hipStream_t side;
    hipEvent_t ev_fork, ev_join;
};

// The context's second stream and its events (created on first use).
int sc_ctx_side(sc_ctx *ctx);

struct sc_index {
    sc_ctx *ctx;
    int32_t dim, mc, kc, mp, kp, nc;
    int64_t n, ncells;
    float *cq, *pq;                 // codebooks [mc][kc][bsc], [mp][kp][bsp]
    float *cqT, *pqT;               // transposed [mc][bsc][kc], [mp][bsp][kp] (LUT loads)
    float *norm_c, *norm_r, *dot_cr;// pre_compute1 tables
    int32_t *L;                     // inclusive prefix [ncells]
    int32_t *pid;                   // [n]
    uint8_t *codes;                 // [n][mp] in cell order
    int32_t max_cell;               // largest cell size
    int64_t nonempty;               // non-empty cells
    float dot_absmax;               // max |dot_cr| (an index statistic)
    const float *raw;               // raw base vectors [n][dim] by id (not owned) or null
    // Two-halves (mc = 2, nc = 2) index: a sparse directory of the non-empty
    // cells (ncells of them), in key order; L is the inclusive prefix over it.
    bool sparse;
    uint64_t *ckeys;                // [ncells] keys ascending
    uint32_t *crec;                 // [ncells][crec_words] ranking records:
                                    // kc <= 256: {n0 | s0<<8 | n1<<16 | s1<<24, size}
                                    // else {n0 | s0<<16, n1 | s1<<16, size, 0}
    int crec_words;                 // 2 or 4
    // The distinct (nearest, second) pairs of each half that occur in the
    // index (uid_h = n_h*kc + s_h, ascending), and per cell the positions of
    // its two pairs in those lists {i0 | i1 << 16, size} (null when a half
    // has more than 65535 pairs).
    int32_t nu[2];
    uint32_t *upairs[2];            // [nu[h]] n | s << 16
    uint32_t *cpair;                // [ncells][2]
    int32_t *gstart;                // [nu[0] + 1] first cell of each half-0 pair
    int32_t *gvec;                  // [nu[0]] vectors in each half-0 pair's cells
    // Large two-halves indexes (kc <= 4096): the half-0 pairs present
    // ("groups") in key order, each the contiguous run of its cells, and per
    // cell a compact record of its half-1 pair, for the pruned cell ranking.
    int32_t ngroups;
    int32_t *bg_n0;                 // [kc + 1] first group of each nearest-0 centroid
    uint32_t *bg_pair;              // [ngroups] n0 | s0 << 16
    int32_t *bg_start;              // [ngroups + 1] first cell of each group
    uint32_t *brec;                 // [ncells] n1 | s1 << 12 | min(size, 255) << 24
    // Two-block indexes (mc = 2): per vector in cell order, the sequential FP32
    // sum of its dot_cr terms (the query-independent part of its IMI distance),
    // for the mixed-form scan.
    float *evec;                    // [n]
    uint32_t *vcell;                // [n] (mc = 2): n0 | n1 << 16 of each position's cell
    // (mc = 2, mp = 8) per position one 16-byte record for the mixed scan:
    // {codes bytes 0-3, codes bytes 4-7, e_v bits, vcell}, one load each
    uint4 *mrec;                    // [n] or null
};

void sc_set_error(const char *fmt, ...);

#define SC_HIP(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            sc_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),\
                         __FILE__, __LINE__);                                  \
            return SC_ERR_HIP;                                                 \
        }                                                                      \
    } while (0)

#define SC_CHECK_LAUNCH() SC_HIP(hipGetLastError())

#define SC_TRY(expr)                                                           \
    do {                                                                       \
        int rc_ = (expr);                                                      \
        if (rc_ != SC_OK) return rc_;                                          \
    } while (0)

// Context scratch: returns a device buffer of at least `bytes` (contents
// undefined).  Grows by reallocation after a stream synchronisation.
int sc_scratch(sc_ctx *ctx, size_t bytes, void **ptr);

// Mixed-form Lloyd assignment (sc_encode_mfma.hip): labels [n] of the points
// x [n][d] (d a multiple of 16, at most 2048) against the centres c [k][d];
// `scratch` holds kmeans_bf16_scratch(k, d) bytes of device memory.
size_t kmeans_bf16_scratch(int k, int d);
int kmeans_assign_bf16(sc_ctx *ctx, const float *x, int64_t n, int d, const float *c, int k,
                       int32_t *labels, unsigned char *scratch);

// Sets the context's device current on the calling thread.
int sc_ctx_enter(sc_ctx *ctx);

// hipFuncAttributeMaxDynamicSharedMemorySize of kern on the current device
// raised to at least `bytes` (one runtime call per kernel, device and new
// maximum).
int sc_set_max_lds(const void *kern, size_t bytes);

static inline int64_t sc_cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Philox4x32-10 (Salmon et al., SC'11), device and host.
__host__ __device__ inline void sc_philox(uint32_t c0, uint32_t c1, uint32_t c2,
                                          uint32_t c3, uint32_t k0, uint32_t k1,
                                          uint32_t out[4]) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Order-preserving map of a float onto uint32 (ascending floats ->
// ascending keys; -0.0 and +0.0 share a key).
__host__ __device__ inline uint32_t sc_float_key(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if (u == 0x80000000u) u = 0u;  // -0.0 compares equal to +0.0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float sc_key_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, u);
}

// Internal launches shared across translation units.
int sc_launch_tables(sc_index *ix);

// Two-halves cell ranking (sc_rank.hip) of the queries [0, nq): cells taken
// in (score ASC, key ASC) order while count < w and sum < T.  Writes qs
// [nq][4] = {q_sum, q_0, q_1, 0}, the cell lists (cells: starts, clen:
// sizes, cinfo: n0 | n1 << 16; [nq][ccap]) and qinfo [nq][8] = {count, sum,
// count, 0, exhausted, overflow, error, 0}.  Queries whose cells do not fit
// the on-chip list are finished by a device-wide sort in the same call.
struct RankCellsArgs {
    const float *Q;          // [nq][dim]
    const float *dqc;        // [nq][2 kc]
    float *qs;
    int32_t *cells, *clen;
    uint32_t *cinfo;
    int32_t *qinfo;
    const int32_t *Tq;       // per-query budgets or null
    int64_t nq;
    int w, T, ccap;
    int32_t *flag_list;      // device scratch: [nq] queries for the device-wide path
    int32_t *flag_count;     // device scratch: 1 counter
    // Pruned ranking of large indexes (sc_rank_big_cap(ix) > 0): per query
    // `big_cap` appended entries, their bucket-ordered copy and sizes, and
    // four words of per-query state.
    uint64_t *big_list, *big_sorted;
    uint32_t *big_size, *big_aux;
    int big_cap;
    int retry_cap;           // > 0: big_* hold scratch for that many re-ranked queries
                             // (indexed by slot) although big_cap == 0
    uint32_t *big_split;     // [nq][kSplitWords] split-query state (small batches) or null
    int64_t *overflowed;     // host, may be null: += queries that outgrew the first list
    unsigned long long *stats;   // device, may be null: [2] += cells scored
    // Deferred finish (pair-table ranking with the pruned re-ranking of the
    // queries that outgrow its list), when both are set: no host
    // synchronisation; flag_count[0] gets the number of queries that outgrew
    // the first list, defer_list / defer_count[0] the queries still
    // unranked after the re-ranking of the first retry_cap of them (their
    // qinfo says zero cells), *deferred (host) is set to 1 and `overflowed`
    // is left to the caller.  sc_rank_finish completes them.
    int32_t *defer_list = nullptr, *defer_count = nullptr;
    int *deferred = nullptr;
    // With deferral: skip [nq] (device) gets 1 for the queries that outgrew
    // the first list and 0 for the others, and their re-ranking runs on the
    // context's second stream (after ev_fork on the context stream), so that
    // the others can be scanned meanwhile; the caller continues on the
    // second stream and joins it to the context stream before it reads the
    // deferred count.
    int32_t *skip = nullptr;
};
int sc_rank_cells(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a);
// Ranks the n queries of list (device) that a deferred sc_rank_cells left
// unranked, with the same arguments.
int sc_rank_finish(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a, const int32_t *list,
                   int32_t n);
// Entries per query of the pruned ranking's scratch, or 0 when the index is
// ranked by the other kernels.
int sc_rank_big_cap(const sc_index *ix);
// Entries per query of the pruned ranking's scratch when it re-ranks the
// queries whose cells overflow the pair-table ranking's on-chip list, or 0.
int sc_rank_retry_cap(const sc_index *ix);
