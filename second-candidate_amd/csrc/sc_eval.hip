// sc_eval.hip -- exact nearest neighbours for recall evaluation.
//
// The reference measures recall@R against ground-truth files whose first
// column is each query's nearest base vector (Evaluation::load_groundtruth
// and calc_recall, include/evaluation.h:34-92).  sc_groundtruth_nn1 computes
// that column on the GPU: for every query, the base vector of least S1
// distance (lib/simple-cluster/include/utilities.h:194-207: float
// differences, squared and summed in index order in double), the lowest id
// on ties -- the first strict minimum of a scan in id order.
//
// Work split: a workgroup owns TV queries (transposed in LDS) and one chunk
// of consecutive base rows, reduced with the encoder's exact distance tile;
// the chunks' winners ([nq][chunks] partials) are merged by a second kernel.
#include <float.h>

#include "sc_tile.h"

namespace {

using namespace sc_tile;

template <int TV, int TK>
__global__ __launch_bounds__((TV / 4) * (TK / 4)) void k_nn1_partial(
    const float *__restrict__ base, int64_t n, const float *__restrict__ q, int64_t nq, int dim,
    int64_t chunk, int nchunks, double *__restrict__ pd, int32_t *__restrict__ pk) {
    constexpr int NT = (TV / 4) * (TK / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = reinterpret_cast<float *>(smem);                  // [dim][TV]
    float *cs = xs + (size_t)dim * TV;                             // [kTT][TK]
    Best2 *red = reinterpret_cast<Best2 *>(cs + kTT * TK);         // [TK/4][TV]
    int *kb = reinterpret_cast<int *>(red + (TK / 4) * TV);        // [TV]
    double *db = reinterpret_cast<double *>(kb + TV);              // [TV]

    const int tid = threadIdx.x;
    const int64_t q0 = (int64_t)blockIdx.x * TV;
    const int c = (int)blockIdx.y;
    const int64_t b0 = (int64_t)c * chunk;
    const int rows = (int)min(chunk, n - b0);
    const int nv = (int)min((int64_t)TV, nq - q0);
    for (int e = tid; e < TV * dim; e += NT) {
        const int v = e / dim, t = e % dim;
        xs[(size_t)t * TV + v] = v < nv ? q[(size_t)(q0 + v) * dim + t] : 0.0f;
    }
    __syncthreads();
    tile_best<TV, TK, false>(xs, 0, dim, base + (size_t)b0 * dim, rows, cs, red, kb, nullptr, db);
    if (tid < nv) {
        const size_t o = (size_t)(q0 + tid) * nchunks + c;
        pk[o] = kb[tid] >= 0 ? (int32_t)(b0 + kb[tid]) : -1;
        pd[o] = db[tid];
    }
}

// Winner of each query over its chunks, under (distance ASC, id ASC).
__global__ void k_nn1_merge(const double *__restrict__ pd, const int32_t *__restrict__ pk,
                            int64_t nq, int nchunks, int32_t *__restrict__ ids,
                            double *__restrict__ dists) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    double bd = INFINITY;
    int bk = -1;
    for (int c = 0; c < nchunks; ++c) {
        const int32_t k = pk[(size_t)i * nchunks + c];
        const double d = pd[(size_t)i * nchunks + c];
        if (k >= 0 && (bk < 0 ? d < DBL_MAX : precedes<false>(d, k, bd, bk))) {
            bd = d;
            bk = k;
        }
    }
    ids[i] = bk;
    if (dists) dists[i] = bd;
}

// Work split of one call: enough (query tile, chunk) workgroups to fill the
// chip several times, with at most max_partials partial winners in all.
// Returns the number of chunks (0 when even one chunk per query exceeds
// max_partials) and the chunk length in *chunk.
template <int TK>
int64_t nn1_chunks(int64_t n, int64_t nq, int64_t qtiles, int64_t max_partials, int64_t *chunk) {
    int64_t want = std::max<int64_t>(1, sc_cdiv(4096, qtiles));
    want = std::min<int64_t>(want, max_partials / nq);
    if (want < 1) return 0;
    int64_t c = std::max<int64_t>(TK, sc_cdiv(n, want));
    c = sc_cdiv(c, TK) * TK;
    *chunk = c;
    return sc_cdiv(n, c);
}

template <int TV, int TK>
size_t nn1_lds(int dim) {
    return (size_t)dim * TV * 4 + kTT * TK * 4 + (TK / 4) * TV * sizeof(Best2) + TV * 4 + TV * 8;
}

// Launches the partial and merge kernels over pd/pk ([nq][nchunks] partial
// winners) and waits for them.
template <int TV, int TK>
int launch_nn1(sc_ctx *ctx, const float *base, int64_t n, const float *q, int64_t nq, int dim,
               int64_t chunk, int nchunks, double *pd, int32_t *pk, int32_t *ids,
               double *dists) {
    constexpr int NT = (TV / 4) * (TK / 4);
    const size_t lds = nn1_lds<TV, TK>(dim);
    hipStream_t sm = ctx->stream;
    auto kern = k_nn1_partial<TV, TK>;
    int rc = SC_OK;
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        rc = SC_ERR_HIP;
    if (rc == SC_OK) {
        hipLaunchKernelGGL(kern, dim3((unsigned)sc_cdiv(nq, TV), (unsigned)nchunks), dim3(NT),
                           lds, sm, base, n, q, nq, dim, chunk, nchunks, pd, pk);
        hipLaunchKernelGGL(k_nn1_merge, dim3((unsigned)sc_cdiv(nq, 256)), dim3(256), 0, sm, pd,
                           pk, nq, nchunks, ids, dists);
        if (hipGetLastError() != hipSuccess) rc = SC_ERR_HIP;
    }
    if (hipStreamSynchronize(sm) != hipSuccess) rc = SC_ERR_HIP;
    if (rc != SC_OK) sc_set_error("sc_groundtruth_nn1: launch failed");
    return rc;
}

// fixed: partial winners go to the context's fixed scratch (no allocation)
// instead of device memory allocated for the call.
template <int TV, int TK>
int run_nn1(sc_ctx *ctx, const float *base, int64_t n, const float *q, int64_t nq, int dim,
            int32_t *ids, double *dists, bool fixed) {
    if (nn1_lds<TV, TK>(dim) > 160 * 1024) {
        sc_set_error("sc_groundtruth_nn1: dim %d too large for the exact tile", dim);
        return SC_ERR_ARG;
    }
    constexpr int64_t kPartial = 12;                 // one double + one int32
    const int64_t qtiles = sc_cdiv(nq, TV);
    const int64_t max_partials =
        fixed ? (int64_t)(ctx->fixed_scratch_bytes / kPartial) : INT64_MAX / kPartial;
    int64_t chunk = 0;
    const int64_t nchunks = nn1_chunks<TK>(n, nq, qtiles, max_partials, &chunk);
    if (nchunks < 1) {
        sc_set_error("sc_groundtruth_nn1_noalloc: %lld queries exceed the fixed scratch "
                     "(%zu bytes: at most %lld queries per call)", (long long)nq,
                     ctx->fixed_scratch_bytes, (long long)max_partials);
        return SC_ERR_ARG;
    }
    const size_t np = (size_t)nq * (size_t)nchunks;
    double *pd = nullptr;
    int32_t *pk = nullptr;
    if (fixed) {
        pd = reinterpret_cast<double *>(ctx->fixed_scratch);
        pk = reinterpret_cast<int32_t *>(pd + np);
        return launch_nn1<TV, TK>(ctx, base, n, q, nq, dim, chunk, (int)nchunks, pd, pk, ids,
                                  dists);
    }
    if (hipMalloc(&pd, 8 * np) != hipSuccess || hipMalloc(&pk, 4 * np) != hipSuccess) {
        (void)hipFree(pd);
        sc_set_error("sc_groundtruth_nn1: allocation failed");
        return SC_ERR_NOMEM;
    }
    const int rc =
        launch_nn1<TV, TK>(ctx, base, n, q, nq, dim, chunk, (int)nchunks, pd, pk, ids, dists);
    (void)hipFree(pd);
    (void)hipFree(pk);
    return rc;
}

int groundtruth(sc_ctx *ctx, const float *d_base, int64_t n, const float *d_q, int64_t nq,
                int32_t dim, int32_t *d_ids, double *d_dists, bool fixed) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_base || !d_q || !d_ids || n < 1 || n > INT32_MAX || nq < 0 || dim < 1 ||
        dim > 4096) {
        sc_set_error("sc_groundtruth_nn1: bad arguments (n=%lld nq=%lld dim=%d)", (long long)n,
                     (long long)nq, dim);
        return SC_ERR_ARG;
    }
    if (nq == 0) return SC_OK;
    if (dim <= 256) return run_nn1<64, 64>(ctx, d_base, n, d_q, nq, dim, d_ids, d_dists, fixed);
    return run_nn1<16, 256>(ctx, d_base, n, d_q, nq, dim, d_ids, d_dists, fixed);
}

}  // namespace

extern "C" int sc_groundtruth_nn1(sc_ctx *ctx, const float *d_base, int64_t n, const float *d_q,
                                  int64_t nq, int32_t dim, int32_t *d_ids, double *d_dists) {
    return groundtruth(ctx, d_base, n, d_q, nq, dim, d_ids, d_dists, false);
}

// The same computation with the partial winners in the library's fixed
// scratch: no device memory is allocated, so callers that cannot allocate
// (code under graph capture, serving loops) can call it.  The scratch is
// shared by every call on one GPU: calls that overlap on one GPU race.
extern "C" int sc_groundtruth_nn1_noalloc(sc_ctx *ctx, const float *d_base, int64_t n,
                                          const float *d_q, int64_t nq, int32_t dim,
                                          int32_t *d_ids, double *d_dists) {
    return groundtruth(ctx, d_base, n, d_q, nq, dim, d_ids, d_dists, true);
}
