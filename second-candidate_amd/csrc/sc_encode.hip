// sc_encode.hip -- coarse assignment (argmin / nearest + second), residual and
// PQ encoding.
//
// Exact form (SC_PREC_FP32): every (vector, centroid) distance is the
// reference's S1 value (lib/simple-cluster/include/utilities.h:194-207):
// float subtraction, widened to double, squared and accumulated in index
// order in double.  The square of a float is exact in double, so the fused
// multiply-add used below rounds exactly like the reference's separate
// multiply and add.
//
// Tiling: a workgroup owns TV vectors (kept transposed in LDS, [dim][TV]) and
// walks the codebook in chunks of TK centroids x TT dimensions staged in LDS
// ([TT][TK]).  Each thread accumulates a 4x4 (vector x centroid) register
// tile, so every LDS float feeds four FP64 FMAs.
#include <float.h>

#include "sc_tile.h"

namespace {

using namespace sc_tile;

// One workgroup: TV vectors.  NC = 1 (Encoder::encode) or 2 (encode2).
template <int NC, int TV, int TK>
__global__ __launch_bounds__((TV / 4) * (TK / 4)) void k_encode_exact(
    const float *__restrict__ x, int64_t n, int dim, const float *__restrict__ cq, int mc,
    int kc, const float *__restrict__ pq, int mp, int kp, uint16_t *__restrict__ cid,
    uint8_t *__restrict__ codes) {
    constexpr int NT = (TV / 4) * (TK / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = reinterpret_cast<float *>(smem);                       // [dim][TV]
    float *cs = xs + (size_t)dim * TV;                                  // [kTT][TK]
    Best2 *red = reinterpret_cast<Best2 *>(cs + kTT * TK);              // [TK/4][TV]
    int *kb1 = reinterpret_cast<int *>(red + (TK / 4) * TV);            // [TV]
    int *kb2 = kb1 + TV;                                                // [TV]

    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * TV;
    const int nv = (int)min((int64_t)TV, n - v0);
    for (int e = tid; e < TV * dim; e += NT) {
        int v = e / dim, t = e % dim;
        xs[(size_t)t * TV + v] = v < nv ? x[(size_t)(v0 + v) * dim + t] : 0.0f;
    }
    __syncthreads();

    const int bsc = dim / mc, bsp = dim / mp;
    for (int j = 0; j < mc; ++j) {
        const float *cb = cq + (size_t)j * kc * bsc;
        if (NC == 2)
            tile_best<TV, TK, true>(xs, j * bsc, bsc, cb, kc, cs, red, kb1, kb2);
        else
            tile_best<TV, TK, false>(xs, j * bsc, bsc, cb, kc, cs, red, kb1, nullptr);
        if (tid < nv) {
            if (NC == 2) {
                cid[(size_t)(v0 + tid) * 2 * mc + j] = (uint16_t)kb1[tid];
                cid[(size_t)(v0 + tid) * 2 * mc + mc + j] = (uint16_t)kb2[tid];
            } else {
                cid[(size_t)(v0 + tid) * mc + j] = (uint16_t)kb1[tid];
            }
        }
        // residual against the nearest centroid, in place
        for (int e = tid; e < TV * bsc; e += NT) {
            int v = e % TV, t = e / TV;
            float *p = xs + (size_t)(j * bsc + t) * TV + v;
            *p = *p - cb[(size_t)max(kb1[v], 0) * bsc + t];
        }
        __syncthreads();
    }
    for (int s = 0; s < mp; ++s) {
        tile_best<TV, TK, false>(xs, s * bsp, bsp, pq + (size_t)s * kp * bsp, kp, cs, red, kb1,
                                 nullptr);
        if (tid < nv) codes[(size_t)(v0 + tid) * mp + s] = (uint8_t)kb1[tid];
    }
}

template <int NC, int TV, int TK>
int launch_exact(sc_ctx *ctx, const float *x, int64_t n, int dim, const float *cq, int mc,
                 int kc, const float *pq, int mp, int kp, uint16_t *cid, uint8_t *codes) {
    constexpr int NT = (TV / 4) * (TK / 4);
    size_t lds = (size_t)dim * TV * 4 + kTT * TK * 4 + (TK / 4) * TV * sizeof(Best2) + 2 * TV * 4;
    if (lds > 160 * 1024) {
        sc_set_error("sc_encode: dim %d too large for the exact tile", dim);
        return SC_ERR_ARG;
    }
    auto kern = k_encode_exact<NC, TV, TK>;
    SC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    int64_t blocks = sc_cdiv(n, TV);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(NT), lds, ctx->stream, x, n, dim, cq,
                       mc, kc, pq, mp, kp, cid, codes);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

template <int NC>
int encode_exact(sc_ctx *ctx, const float *x, int64_t n, int dim, const float *cq, int mc,
                 int kc, const float *pq, int mp, int kp, uint16_t *cid, uint8_t *codes) {
    if (dim <= 256)
        return launch_exact<NC, 64, 64>(ctx, x, n, dim, cq, mc, kc, pq, mp, kp, cid, codes);
    return launch_exact<NC, 16, 256>(ctx, x, n, dim, cq, mc, kc, pq, mp, kp, cid, codes);
}

// Encoder::rencode<T> (include/encoder.h:202-309): position i encodes raw row
// rows[i] (row i when rows is null) against the GIVEN coarse cells
// cid[i][mc] -- float residual per coarse block, then Encoder::encode's PQ
// argmin (S1 distance, first strict minimum).  A row outside [0, nx) or a
// cell id >= kc sets *bad; such a position reads row 0 / no centroid so that
// every access stays in bounds, and the call reports the error.
template <int TV, int TK>
__global__ __launch_bounds__((TV / 4) * (TK / 4)) void k_encode_given(
    const float *__restrict__ x, int64_t nx, const int32_t *__restrict__ rows, int64_t n,
    int dim, const uint16_t *__restrict__ cid, const float *__restrict__ cq, int mc, int kc,
    const float *__restrict__ pq, int mp, int kp, uint8_t *__restrict__ codes,
    int *__restrict__ bad) {
    constexpr int NT = (TV / 4) * (TK / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = reinterpret_cast<float *>(smem);                       // [dim][TV]
    float *cs = xs + (size_t)dim * TV;                                  // [kTT][TK]
    Best2 *red = reinterpret_cast<Best2 *>(cs + kTT * TK);              // [TK/4][TV]
    int *kb1 = reinterpret_cast<int *>(red + (TK / 4) * TV);            // [TV]
    int64_t *rr = reinterpret_cast<int64_t *>(kb1 + TV);                // [TV]

    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * TV;
    const int nv = (int)min((int64_t)TV, n - v0);
    if (tid < TV) {
        int64_t r = rows ? (tid < nv ? (int64_t)rows[v0 + tid] : 0) : v0 + tid;
        if (tid < nv && (r < 0 || r >= nx)) {
            atomicOr(bad, 1);
            r = 0;
        }
        rr[tid] = r;
    }
    __syncthreads();
    for (int e = tid; e < TV * dim; e += NT) {
        int v = e / dim, t = e % dim;
        xs[(size_t)t * TV + v] = v < nv ? x[(size_t)rr[v] * dim + t] : 0.0f;
    }
    __syncthreads();
    const int bsc = dim / mc, bsp = dim / mp;
    for (int e = tid; e < TV * dim; e += NT) {
        const int v = e % TV, t = e / TV, j = t / bsc;
        if (v >= nv) continue;
        const int c = cid[(size_t)(v0 + v) * mc + j];
        if (c >= kc) {
            if (t % bsc == 0) atomicOr(bad, 2);
            continue;
        }
        float *p = xs + (size_t)t * TV + v;
        *p = *p - cq[((size_t)j * kc + c) * bsc + t % bsc];
    }
    __syncthreads();
    for (int s = 0; s < mp; ++s) {
        tile_best<TV, TK, false>(xs, s * bsp, bsp, pq + (size_t)s * kp * bsp, kp, cs, red, kb1,
                                 nullptr);
        if (tid < nv) codes[(size_t)(v0 + tid) * mp + s] = (uint8_t)kb1[tid];
    }
}

template <int TV, int TK>
int launch_given(sc_ctx *ctx, const float *x, int64_t nx, const int32_t *rows, int64_t n,
                 int dim, const uint16_t *cid, const float *cq, int mc, int kc, const float *pq,
                 int mp, int kp, uint8_t *codes, int *bad) {
    constexpr int NT = (TV / 4) * (TK / 4);
    const size_t lds =
        (size_t)dim * TV * 4 + kTT * TK * 4 + (TK / 4) * TV * sizeof(Best2) + TV * 4 + TV * 8;
    if (lds > 160 * 1024) {
        sc_set_error("sc_encode_given: dim %d too large for the exact tile", dim);
        return SC_ERR_ARG;
    }
    auto kern = k_encode_given<TV, TK>;
    SC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)sc_cdiv(n, TV)), dim3(NT), lds, ctx->stream, x, nx,
                       rows, n, dim, cid, cq, mc, kc, pq, mp, kp, codes, bad);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

// SCEncoder::encode with general nc (include/sc_encoder.h:205-332): the NC
// nearest coarse centroids of each block by float-rounded distance (ties to
// the lower index), cid [n][mc][NC], residual against the nearest, PQ
// argmin of float-rounded distances.
template <int NC, int TV, int TK>
__global__ __launch_bounds__((TV / 4) * (TK / 4)) void k_encode_general(
    const float *__restrict__ x, int64_t n, int dim, const float *__restrict__ cq, int mc,
    int kc, const float *__restrict__ pq, int mp, int kp, uint16_t *__restrict__ cid,
    uint8_t *__restrict__ codes) {
    constexpr int NT = (TV / 4) * (TK / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = reinterpret_cast<float *>(smem);                       // [dim][TV]
    float *cs = xs + (size_t)dim * TV;                                  // [kTT][TK]
    TopF<NC> *red = reinterpret_cast<TopF<NC> *>(cs + kTT * TK);        // [TK/4][TV]
    int *kb = reinterpret_cast<int *>(red + (TK / 4) * TV);             // [TV][NC]
    TopF<1> *red1 = reinterpret_cast<TopF<1> *>(red);

    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * TV;
    const int nv = (int)min((int64_t)TV, n - v0);
    for (int e = tid; e < TV * dim; e += NT) {
        int v = e / dim, t = e % dim;
        xs[(size_t)t * TV + v] = v < nv ? x[(size_t)(v0 + v) * dim + t] : 0.0f;
    }
    __syncthreads();

    const int bsc = dim / mc, bsp = dim / mp;
    for (int j = 0; j < mc; ++j) {
        const float *cb = cq + (size_t)j * kc * bsc;
        tile_top<TV, TK, NC>(xs, j * bsc, bsc, cb, kc, cs, red, kb);
        if (tid < nv)
            for (int t = 0; t < NC; ++t)
                cid[((size_t)(v0 + tid) * mc + j) * NC + t] = (uint16_t)kb[tid * NC + t];
        for (int e = tid; e < TV * bsc; e += NT) {
            int v = e % TV, t = e / TV;
            float *p = xs + (size_t)(j * bsc + t) * TV + v;
            *p = *p - cb[(size_t)max(kb[v * NC], 0) * bsc + t];
        }
        __syncthreads();
    }
    for (int s = 0; s < mp; ++s) {
        tile_top<TV, TK, 1>(xs, s * bsp, bsp, pq + (size_t)s * kp * bsp, kp, cs, red1, kb);
        if (tid < nv) codes[(size_t)(v0 + tid) * mp + s] = (uint8_t)max(kb[tid], 0);
        __syncthreads();
    }
}

template <int NC, int TV, int TK>
int launch_general(sc_ctx *ctx, const float *x, int64_t n, int dim, const float *cq, int mc,
                   int kc, const float *pq, int mp, int kp, uint16_t *cid, uint8_t *codes) {
    constexpr int NT = (TV / 4) * (TK / 4);
    size_t lds = (size_t)dim * TV * 4 + kTT * TK * 4 + (TK / 4) * TV * sizeof(TopF<NC>) +
                 (size_t)TV * NC * 4;
    if (lds > 160 * 1024) {
        sc_set_error("sc_encode_general: dim %d too large for the exact tile", dim);
        return SC_ERR_ARG;
    }
    auto kern = k_encode_general<NC, TV, TK>;
    SC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)sc_cdiv(n, TV)), dim3(NT), lds, ctx->stream, x, n, dim,
                       cq, mc, kc, pq, mp, kp, cid, codes);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

template <int NC>
int encode_general(sc_ctx *ctx, const float *x, int64_t n, int dim, const float *cq, int mc,
                   int kc, const float *pq, int mp, int kp, uint16_t *cid, uint8_t *codes) {
    if (dim <= 256)
        return launch_general<NC, 64, 64>(ctx, x, n, dim, cq, mc, kc, pq, mp, kp, cid, codes);
    return launch_general<NC, 16, 256>(ctx, x, n, dim, cq, mc, kc, pq, mp, kp, cid, codes);
}

}  // namespace

extern "C" int sc_encode_general(sc_ctx *ctx, const float *d_x, int64_t n, int32_t dim,
                                 const float *d_cq, int32_t mc, int32_t kc, const float *d_pq,
                                 int32_t mp, int32_t kp, int32_t nc, uint16_t *d_cid,
                                 uint8_t *d_codes) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_x || !d_cq || !d_pq || !d_cid || !d_codes || n < 0 || dim <= 0 || mc <= 0 ||
        mp <= 0 || dim % mc || dim % mp || kc < 1 || kc > 65536 || kp < 1 || kp > 256 ||
        nc < 1 || nc > 3 || nc > kc || dim > 4096) {
        sc_set_error("sc_encode_general: bad arguments (dim=%d mc=%d kc=%d mp=%d kp=%d nc=%d)",
                     dim, mc, kc, mp, kp, nc);
        return SC_ERR_ARG;
    }
    if (n == 0) return SC_OK;
    if (nc == 3) return encode_general<3>(ctx, d_x, n, dim, d_cq, mc, kc, d_pq, mp, kp, d_cid, d_codes);
    if (nc == 2) return encode_general<2>(ctx, d_x, n, dim, d_cq, mc, kc, d_pq, mp, kp, d_cid, d_codes);
    return encode_general<1>(ctx, d_x, n, dim, d_cq, mc, kc, d_pq, mp, kp, d_cid, d_codes);
}

int sc_encode_bf16(sc_ctx *ctx, const float *x, int64_t n, int dim, const float *cq, int mc,
                   int kc, const float *pq, int mp, int kp, int nc, uint16_t *cid,
                   uint8_t *codes, sc_encode_stats *st);

extern "C" int sc_encode(sc_ctx *ctx, const float *d_x, int64_t n, int32_t dim,
                         const float *d_cq, int32_t mc, int32_t kc, const float *d_pq,
                         int32_t mp, int32_t kp, int32_t nc, int32_t prec, uint16_t *d_cid,
                         uint8_t *d_codes, sc_encode_stats *st) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_x || !d_cq || !d_pq || !d_cid || !d_codes || n < 0 || dim <= 0 || mc <= 0 ||
        mp <= 0 || dim % mc || dim % mp || kc < 1 || kc > 65536 || kp < 1 || kp > 256 ||
        (nc != 1 && nc != 2) || (nc == 2 && kc < 2) || dim > 4096) {
        sc_set_error("sc_encode: bad arguments (dim=%d mc=%d kc=%d mp=%d kp=%d nc=%d)", dim, mc,
                     kc, mp, kp, nc);
        return SC_ERR_ARG;
    }
    if (st) { st->rechecked = 0; st->fallback = 0; }
    if (n == 0) return SC_OK;
    if (prec == SC_PREC_BF16)
        return sc_encode_bf16(ctx, d_x, n, dim, d_cq, mc, kc, d_pq, mp, kp, nc, d_cid, d_codes, st);
    if (prec != SC_PREC_FP32) { sc_set_error("sc_encode: bad precision %d", prec); return SC_ERR_ARG; }
    if (nc == 2) return encode_exact<2>(ctx, d_x, n, dim, d_cq, mc, kc, d_pq, mp, kp, d_cid, d_codes);
    return encode_exact<1>(ctx, d_x, n, dim, d_cq, mc, kc, d_pq, mp, kp, d_cid, d_codes);
}

extern "C" int sc_encode_given(sc_ctx *ctx, const float *d_x, int64_t nx, int32_t dim,
                               const int32_t *d_rows, int64_t n, const uint16_t *d_cid,
                               const float *d_cq, int32_t mc, int32_t kc, const float *d_pq,
                               int32_t mp, int32_t kp, uint8_t *d_codes) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_x || !d_cid || !d_cq || !d_pq || !d_codes || nx < 0 || n < 0 ||
        (!d_rows && n > nx) || dim <= 0 || mc <= 0 || mp <= 0 || dim % mc || dim % mp ||
        kc < 1 || kc > 65536 || kp < 1 || kp > 256 || dim > 4096) {
        sc_set_error("sc_encode_given: bad arguments (nx=%lld n=%lld dim=%d mc=%d kc=%d mp=%d "
                     "kp=%d)", (long long)nx, (long long)n, dim, mc, kc, mp, kp);
        return SC_ERR_ARG;
    }
    if (n == 0) return SC_OK;
    int *bad = nullptr;
    SC_TRY(sc_scratch(ctx, sizeof(int), (void **)&bad));
    SC_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
    const int rc =
        dim <= 256
            ? launch_given<64, 64>(ctx, d_x, nx, d_rows, n, dim, d_cid, d_cq, mc, kc, d_pq, mp,
                                   kp, d_codes, bad)
            : launch_given<16, 256>(ctx, d_x, nx, d_rows, n, dim, d_cid, d_cq, mc, kc, d_pq, mp,
                                    kp, d_codes, bad);
    if (rc != SC_OK) return rc;
    int hbad = 0;
    SC_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    if (hbad) {
        sc_set_error("sc_encode_given: %s%s%s", hbad & 1 ? "a row index outside [0, nx)" : "",
                     hbad == 3 ? " and " : "", hbad & 2 ? "a cell id >= kc" : "");
        return SC_ERR_ARG;
    }
    return SC_OK;
}
