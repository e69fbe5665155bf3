// sc_kmeans.hip -- codebook training: seeded k-means++ and Lloyd iterations
// (PQQuantizer::create_sub_quantizers -> greg_kmeans, include/quantizer.h:
// 235-272, lib/simple-cluster/include/k-means.h:162-267, 553-912), plus the
// residuals PQ training runs on (quantizer.h:456-510, sc_quantizer.h:72-133).
//
// Seeding.  The reference draws from random_device (k-means.h:172-173); here
// the draws come from Philox(key = {seed, 1}) so CPU and GPU agree: D^2
// weights w_i = floor(D2_i * 2^16) summed in uint64 (order-independent), the
// pivot p = u64 % W, and the pick is the first i whose inclusive prefix of w
// exceeds p.  The whole round stays on the device.
//
// Lloyd.  Labels are the argmin of the float-rounded S1 distance with ties to
// the later centroid (`min >= d`, k-means.h:602, 806); sums are FP64, each
// accumulated over its centre's members in index order (the sequential
// order, so the result is the same on every run and for any data), counts
// int; an empty cluster
// takes the point farthest from its own centroid (SINGLETON, k-means.h:
// 845-877 with find_lonely 506-533); centres = (float)(sum / count).
#include <float.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <vector>

#include "sc_tile.h"

namespace {

using namespace sc_tile;

constexpr int kPickBlock = 1024;

__device__ __forceinline__ double s1(const float *a, const float *b, int d) {
    double acc = 0.0;
    for (int t = 0; t < d; ++t) {
        const double dd = (double)(a[t] - b[t]);
        acc = __fma_rn(dd, dd, acc);
    }
    return acc;
}

__device__ __forceinline__ uint64_t weight(double d2) { return (uint64_t)ldexp(d2, 16); }

__device__ __forceinline__ uint64_t pick_u64(uint32_t seed, uint32_t round, uint32_t which) {
    uint32_t o[4];
    sc_philox(round, 0u, 0xA5u, which, seed, 1u, o);
    return ((uint64_t)o[1] << 32) | o[0];
}

__global__ void k_first_seed(int64_t n, uint32_t seed, int64_t *pick) {
    if (threadIdx.x == 0 && blockIdx.x == 0) pick[0] = (int64_t)(pick_u64(seed, 0, 0) % (uint64_t)n);
}

// D2_i = min(D2_i, S1(x_i, x_pick[slot]))  (init: D2_i = S1(x_i, x_pick[0])).
// The workgroup's 256 points are staged 32 dimensions at a time (each row
// slice is one 128-byte read), and every thread accumulates its own point's
// S1 from LDS in dimension order.
constexpr int kD2Dims = 32;
__global__ __launch_bounds__(256) void k_update_d2(const float *__restrict__ x, int64_t n, int d,
                                                   const int64_t *__restrict__ picks, int slot,
                                                   int init, double *__restrict__ d2,
                                                   uint64_t *__restrict__ bsum) {
    __shared__ uint64_t red[256];
    __shared__ float tile[256][kD2Dims + 1];
    __shared__ float cs[kD2Dims];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * 256, i = i0 + tid;
    const float *c = x + (size_t)picks[slot] * d;
    double acc = 0.0;
    for (int t0 = 0; t0 < d; t0 += kD2Dims) {
        const int tt = min(kD2Dims, d - t0);
        __syncthreads();
        for (int e = tid; e < 256 * kD2Dims; e += 256) {
            const int r = e / kD2Dims, col = e % kD2Dims;
            tile[r][col] = (col < tt && i0 + r < n) ? x[(size_t)(i0 + r) * d + t0 + col] : 0.0f;
        }
        if (tid < kD2Dims) cs[tid] = tid < tt ? c[t0 + tid] : 0.0f;
        __syncthreads();
        for (int j = 0; j < tt; ++j) {
            const double dd = (double)(tile[tid][j] - cs[j]);
            acc = __fma_rn(dd, dd, acc);
        }
    }
    uint64_t w = 0;
    if (i < n) {
        const double t = acc;
        double v = init ? t : (d2[i] > t ? t : d2[i]);
        d2[i] = v;
        w = weight(v);
    }
    // per-block sum of weights (blocks of 256 points)
    red[threadIdx.x] = w;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// One workgroup (kPickBlock threads): total W, pivot p, then the first point
// whose inclusive weight prefix exceeds p.  Block sums are split into one
// contiguous range per thread; a scan over the ranges finds the range, the
// owning thread finds the 256-point block, and the workgroup scans that block.
__global__ __launch_bounds__(kPickBlock) void k_pick(const double *__restrict__ d2, int64_t n,
                                                     const uint64_t *__restrict__ bsum,
                                                     int64_t nblocks, uint32_t seed, int round,
                                                     int64_t *__restrict__ picks) {
    __shared__ uint64_t s_part[kPickBlock];
    __shared__ uint64_t s_p, s_base;
    __shared__ int64_t s_blk;
    const int tid = threadIdx.x;
    const int64_t per = (nblocks + kPickBlock - 1) / kPickBlock;
    const int64_t b0 = (int64_t)tid * per, b1 = min(nblocks, b0 + per);
    uint64_t mine = 0;
    for (int64_t b = b0; b < b1; ++b) mine += bsum[b];
    s_part[tid] = mine;
    __syncthreads();
    for (int o = 1; o < kPickBlock; o <<= 1) {          // inclusive scan
        uint64_t t = tid >= o ? s_part[tid - o] : 0;
        __syncthreads();
        s_part[tid] += t;
        __syncthreads();
    }
    const uint64_t total = s_part[kPickBlock - 1];
    if (tid == 0) {
        const uint64_t r = pick_u64(seed, (uint32_t)round, 1);
        s_p = total ? r % total : 0;
        s_blk = -1;
        if (total == 0) picks[round] = (int64_t)(r % (uint64_t)n);
    }
    __syncthreads();
    if (total == 0) return;
    const uint64_t p = s_p;
    const uint64_t excl = s_part[tid] - mine;
    if (excl <= p && p < s_part[tid]) {                  // exactly one thread
        uint64_t run = excl;
        for (int64_t b = b0; b < b1; ++b) {
            if (run + bsum[b] > p) { s_blk = b; s_base = run; break; }
            run += bsum[b];
        }
    }
    __syncthreads();
    if (s_blk < 0) {                                     // p < total: unreachable
        if (tid == 0) picks[round] = n - 1;
        return;
    }
    const int64_t i0 = s_blk * 256;
    const int64_t i = i0 + tid;
    uint64_t w = (tid < 256 && i < n) ? weight(d2[i]) : 0;
    s_part[tid] = w;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint64_t t = (tid < 256 && tid >= o) ? s_part[tid - o] : 0;
        __syncthreads();
        if (tid < 256) s_part[tid] += t;
        __syncthreads();
    }
    if (tid < 256 && i < n) {
        const uint64_t incl = s_base + s_part[tid];
        if (incl > p && incl - w <= p) picks[round] = i;
    }
}

__global__ void k_copy_seeds(const float *__restrict__ x, int d, const int64_t *__restrict__ picks,
                             int k, float *__restrict__ centers) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)k * d) return;
    int64_t j = e / d;
    centers[e] = x[(size_t)picks[j] * d + (e - j * d)];
}

// Lloyd assignment: one workgroup of TV points against all k centres.
template <int TV, int TK>
__global__ __launch_bounds__((TV / 4) * (TK / 4)) void k_assign(
    const float *__restrict__ x, int64_t n, int d, const float *__restrict__ c, int k,
    int32_t *__restrict__ labels) {
    constexpr int NT = (TV / 4) * (TK / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = reinterpret_cast<float *>(smem);
    float *cs = xs + (size_t)d * TV;
    Best2 *red = reinterpret_cast<Best2 *>(cs + kTT * TK);
    int *kb1 = reinterpret_cast<int *>(red + (TK / 4) * TV);
    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * TV;
    const int nv = (int)min((int64_t)TV, n - v0);
    for (int e = tid; e < TV * d; e += NT) {
        int v = e / d, t = e % d;
        xs[(size_t)t * TV + v] = v < nv ? x[(size_t)(v0 + v) * d + t] : 0.0f;
    }
    __syncthreads();
    tile_best<TV, TK, true, true>(xs, 0, d, c, k, cs, red, kb1, nullptr);
    if (tid < nv) labels[v0 + tid] = kb1[tid];
}

// Deterministic Lloyd sums.  The points are ordered by label, stably (a
// radix sort of (label, index) pairs keeps index order within a label), and
// each (centre, dimension) sum is one FP64 accumulation over that centre's
// members in index order: the order of a sequential pass over the points,
// so every sum -- and every centre -- is bitwise the sequential one, with
// no dependence on scheduling.
__global__ void k_label_order(const int32_t *__restrict__ labels, int64_t n,
                              int32_t *__restrict__ idx, int32_t *__restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[i] = (int32_t)i;
    atomicAdd(&counts[labels[i]], 1);   // integer counts: any order is exact
}

// One thread per (centre j, dimension t); members of j are
// order[start[j] .. start[j] + counts[j]).
__global__ void k_member_sums(const float *__restrict__ x, int d, int k,
                              const int32_t *__restrict__ order,
                              const int32_t *__restrict__ start,
                              const int32_t *__restrict__ counts, double *__restrict__ sums) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)k * d) return;
    const int j = (int)(e / d), t = (int)(e - (int64_t)j * d);
    const int32_t *m = order + start[j];
    const int cnt = counts[j];
    double acc = 0.0;
    int i = 0;
    for (; i + 16 <= cnt; i += 16) {   // sixteen independent loads, adds in order
        float f[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) f[u] = x[(size_t)m[i + u] * d + t];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += (double)f[u];
    }
    for (; i < cnt; ++i) acc += (double)x[(size_t)m[i] * d + t];
    sums[e] = acc;
}

// start[j] = counts[0] + ... + counts[j-1], one workgroup.
__global__ void k_count_starts(const int32_t *__restrict__ counts, int k,
                               int32_t *__restrict__ start) {
    __shared__ int32_t part[256];
    __shared__ int32_t carry;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int b = 0; b < k; b += 256) {
        const int v = b + tid < k ? counts[b + tid] : 0;
        part[tid] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int u = tid >= o ? part[tid - o] : 0;
            __syncthreads();
            part[tid] += u;
            __syncthreads();
        }
        if (b + tid < k) start[b + tid] = carry + part[tid] - v;
        __syncthreads();
        if (tid == 255) carry += part[255];
        __syncthreads();
    }
}

// argmax over i of (float)S1(x_i, c[label_i]) (first index on ties):
// per-block best packed as (orderable float key << 32 | ~index).
__global__ void k_lonely(const float *__restrict__ x, int64_t n, int d, const float *__restrict__ c,
                         const int32_t *__restrict__ labels, unsigned long long *__restrict__ best) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long key = 0;
    if (i < n) {
        float t = (float)s1(x + (size_t)i * d, c + (size_t)labels[i] * d, d);
        // find_lonely's `dfst < d` never takes a NaN distance
        if (t == t) key = ((unsigned long long)sc_float_key(t) << 32) | (uint32_t)~(uint32_t)i;
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long u = __shfl_xor(key, o);
        key = u > key ? u : key;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(best, key);
}

__global__ void k_move_center(const float *__restrict__ x, int d, int64_t fst, int j, int old,
                              float *__restrict__ c, double *__restrict__ sums,
                              int32_t *__restrict__ counts, int32_t *__restrict__ labels) {
    for (int t = threadIdx.x; t < d; t += blockDim.x) {
        float v = x[(size_t)fst * d + t];
        c[(size_t)j * d + t] = v;
        sums[(size_t)old * d + t] -= (double)v;
        sums[(size_t)j * d + t] = (double)v;
    }
    if (threadIdx.x == 0) {
        counts[old] -= 1;
        counts[j] = 1;
        labels[fst] = j;
    }
}

__global__ void k_update_centers(const double *__restrict__ sums, const int32_t *__restrict__ counts,
                                 int k, int d, float *__restrict__ c) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)k * d) return;
    c[e] = (float)(sums[e] / (double)counts[e / d]);
}

__global__ void k_fill_small(const float *__restrict__ x, int64_t n, int d, int k,
                             float *__restrict__ c, int32_t *__restrict__ labels) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)k * d) return;
    int64_t j = e / d;
    c[e] = j < n ? x[e] : FLT_MAX;
    if (labels && e % d == 0 && j < n) labels[j] = (int32_t)j;
}

// Residuals: nearest (which=0) or second (which=1) by float-rounded S1 with
// strict `>` (lowest index wins), subtracted in place -- the two smallest by
// (float distance, index).  Tiled form: TV vectors per workgroup against the
// block's codebook in LDS chunks (the general encoder's float-rounded top-NC
// tile), the chosen centroid subtracted from the vectors in global memory.
template <int NC, int TV, int TK>
__global__ __launch_bounds__((TV / 4) * (TK / 4)) void k_residual_tile(
    float *__restrict__ x, int64_t n, int dim, const float *__restrict__ cq, int mc, int kc) {
    constexpr int NT = (TV / 4) * (TK / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    float *xs = reinterpret_cast<float *>(smem);                       // [dim][TV]
    float *cs = xs + (size_t)dim * TV;                                  // [kTT][TK]
    TopF<NC> *red = reinterpret_cast<TopF<NC> *>(cs + kTT * TK);        // [TK/4][TV]
    int *kb = reinterpret_cast<int *>(red + (TK / 4) * TV);             // [TV][NC]
    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * TV;
    const int nv = (int)min((int64_t)TV, n - v0);
    for (int e = tid; e < TV * dim; e += NT) {
        const int v = e / dim, t = e % dim;
        xs[(size_t)t * TV + v] = v < nv ? x[(size_t)(v0 + v) * dim + t] : 0.0f;
    }
    __syncthreads();
    const int bsc = dim / mc;
    for (int j = 0; j < mc; ++j) {
        const float *cb = cq + (size_t)j * kc * bsc;
        tile_top<TV, TK, NC>(xs, j * bsc, bsc, cb, kc, cs, red, kb);
        for (int e = tid; e < nv * bsc; e += NT) {
            const int v = e / bsc, t = e % bsc;
            const int c = max(kb[v * NC + NC - 1], 0);
            x[(size_t)(v0 + v) * dim + j * bsc + t] = xs[(size_t)(j * bsc + t) * TV + v] -
                                                      cb[(size_t)c * bsc + t];
        }
        __syncthreads();
    }
}

template <int NC, int TV, int TK>
size_t residual_lds(int dim) {
    return (size_t)dim * TV * 4 + kTT * TK * 4 + (TK / 4) * TV * sizeof(TopF<NC>) +
           (size_t)TV * NC * 4;
}

template <int NC, int TV, int TK>
int launch_residual(sc_ctx *ctx, float *x, int64_t n, int dim, const float *cq, int mc, int kc) {
    constexpr int NT = (TV / 4) * (TK / 4);
    const size_t lds = residual_lds<NC, TV, TK>(dim);
    auto kern = k_residual_tile<NC, TV, TK>;
    SC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)sc_cdiv(n, TV)), dim3(NT), lds, ctx->stream, x, n, dim,
                       cq, mc, kc);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

// One thread per (vector, block): the form for vectors too long for the tile.
__global__ void k_residual(float *__restrict__ x, int64_t n, int dim, const float *__restrict__ cq,
                           int mc, int kc, int which) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * mc) return;
    const int64_t i = e / mc;
    const int j = (int)(e - i * mc);
    const int bs = dim / mc;
    float *v = x + (size_t)i * dim + (size_t)j * bs;
    const float *cb = cq + (size_t)j * kc * bs;
    float d1 = FLT_MAX, d2 = FLT_MAX;
    int p1 = 0, p2 = 0;
    for (int k = 0; k < kc; ++k) {
        float t = (float)s1(v, cb + (size_t)k * bs, bs);
        if (d1 > t) { d2 = d1; p2 = p1; d1 = t; p1 = k; }
        else if (d2 > t) { d2 = t; p2 = k; }
    }
    const float *c = cb + (size_t)(which ? p2 : p1) * bs;
    for (int t = 0; t < bs; ++t) v[t] -= c[t];
}

}  // namespace

extern "C" int sc_kmeans(sc_ctx *ctx, const float *d_x, int64_t n, int32_t d, int32_t k,
                         int32_t iters, uint32_t seed, int32_t lagged, int32_t prec,
                         float *d_centers, int32_t *d_labels) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_x || !d_centers || n < 1 || d < 1 || d > 1024 || k < 1 || iters < 0 ||
        n >= INT32_MAX) {
        sc_set_error("sc_kmeans: bad arguments (n=%lld d=%d k=%d iters=%d)", (long long)n, d, k,
                     iters);
        return SC_ERR_ARG;
    }
    const bool mixed = prec == SC_PREC_BF16;
    if ((prec != SC_PREC_FP32 && !mixed) || (mixed && (d % 16 || d > 2048))) {
        sc_set_error("sc_kmeans: precision %d not available for d=%d (SC_PREC_BF16 needs d a "
                     "multiple of 16, at most 2048)", prec, d);
        return SC_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    if (n < k) {
        hipLaunchKernelGGL(k_fill_small, dim3((unsigned)sc_cdiv((int64_t)k * d, 256)), dim3(256), 0,
                           s, d_x, n, d, k, d_centers, d_labels);
        SC_CHECK_LAUNCH();
        return SC_OK;
    }
    const int64_t nb = sc_cdiv(n, 256);
    // label bits for the member sort
    unsigned bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < k) ++bits;
    size_t sort_bytes = 0;
    SC_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t *)nullptr,
                                     (uint32_t *)nullptr, (int32_t *)nullptr,
                                     (int32_t *)nullptr, (size_t)n, 0u, bits, s));
    // scratch: d2 [n] f64, bsum [nb] u64, picks [k] i64, sums [k*d] f64,
    // counts [k] i32, labels [n] (when the caller passes none), best u64,
    // member sort: idx [n], order [n], sorted labels [n], starts [k], temp
    auto al = [](size_t o) { return (o + 255) & ~(size_t)255; };
    size_t o_d2 = 0, o_bs = o_d2 + 8 * (size_t)n, o_pk = o_bs + 8 * (size_t)nb,
           o_sum = o_pk + 8 * (size_t)k, o_cnt = o_sum + 8 * (size_t)k * d,
           o_lab = o_cnt + 4 * (size_t)k + 8, o_best = o_lab + 4 * (size_t)n + 8,
           o_idx = al(o_best + 16), o_ord = al(o_idx + 4 * (size_t)n),
           o_skey = al(o_ord + 4 * (size_t)n), o_start = al(o_skey + 4 * (size_t)n),
           o_tmp = al(o_start + 4 * (size_t)k), o_mix = al(o_tmp + sort_bytes),
           total = o_mix + (mixed ? kmeans_bf16_scratch(k, d) : 0);
    unsigned char *S;
    SC_TRY(sc_scratch(ctx, total, (void **)&S));
    double *d2 = (double *)(S + o_d2), *sums = (double *)(S + o_sum);
    uint64_t *bsum = (uint64_t *)(S + o_bs);
    int64_t *picks = (int64_t *)(S + o_pk);
    int32_t *counts = (int32_t *)(S + o_cnt);
    int32_t *labels = d_labels ? d_labels : (int32_t *)(S + o_lab);
    unsigned long long *best = (unsigned long long *)(S + o_best);
    int32_t *idx = (int32_t *)(S + o_idx), *order = (int32_t *)(S + o_ord),
            *start = (int32_t *)(S + o_start);
    uint32_t *skey = (uint32_t *)(S + o_skey);

    // ---- seeding
    hipLaunchKernelGGL(k_first_seed, dim3(1), dim3(1), 0, s, n, seed, picks);
    hipLaunchKernelGGL(k_update_d2, dim3((unsigned)nb), dim3(256), 0, s, d_x, n, d, picks, 0, 1,
                       d2, bsum);
    SC_CHECK_LAUNCH();
    for (int r = 1; r < k; ++r) {
        hipLaunchKernelGGL(k_pick, dim3(1), dim3(kPickBlock), 0, s, d2, n, bsum, nb, seed, r,
                           picks);
        hipLaunchKernelGGL(k_update_d2, dim3((unsigned)nb), dim3(256), 0, s, d_x, n, d, picks,
                           lagged ? r - 1 : r, 0, d2, bsum);
        SC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_copy_seeds, dim3((unsigned)sc_cdiv((int64_t)k * d, 256)), dim3(256), 0, s,
                       d_x, d, picks, k, d_centers);
    SC_CHECK_LAUNCH();

    // ---- Lloyd
    constexpr int TV = 16, TK = 256;
    const size_t lds = (size_t)d * TV * 4 + kTT * TK * 4 + (TK / 4) * TV * sizeof(Best2) + TV * 8;
    auto kern = k_assign<TV, TK>;
    SC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    std::vector<int32_t> hcnt(k);
    for (int it = 0; it < iters; ++it) {
        SC_HIP(hipMemsetAsync(counts, 0, 4 * (size_t)k, s));
        if (mixed) {   // BF16 MFMA distances, exact re-check: the same labels
            SC_TRY(kmeans_assign_bf16(ctx, d_x, n, d, d_centers, k, labels, S + o_mix));
        } else {
            hipLaunchKernelGGL(kern, dim3((unsigned)sc_cdiv(n, TV)), dim3((TV / 4) * (TK / 4)),
                               lds, s, d_x, n, d, d_centers, k, labels);
        }
        SC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_label_order, dim3((unsigned)nb), dim3(256), 0, s, labels, n, idx,
                           counts);
        SC_CHECK_LAUNCH();
        size_t tb = sort_bytes;
        SC_HIP(rocprim::radix_sort_pairs(S + o_tmp, tb, (const uint32_t *)labels, skey, idx,
                                         order, (size_t)n, 0u, bits, s));
        hipLaunchKernelGGL(k_count_starts, dim3(1), dim3(256), 0, s, counts, k, start);
        hipLaunchKernelGGL(k_member_sums, dim3((unsigned)sc_cdiv((int64_t)k * d, 256)),
                           dim3(256), 0, s, d_x, d, k, order, start, counts, sums);
        SC_CHECK_LAUNCH();
        SC_HIP(hipMemcpyAsync(hcnt.data(), counts, 4 * (size_t)k, hipMemcpyDeviceToHost, s));
        SC_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < k; ++j) {
            if (hcnt[j] > 0) continue;
            SC_HIP(hipMemsetAsync(best, 0, 8, s));
            hipLaunchKernelGGL(k_lonely, dim3((unsigned)nb), dim3(256), 0, s, d_x, n, d, d_centers,
                               labels, best);
            SC_CHECK_LAUNCH();
            unsigned long long hb = 0;
            int32_t old = 0;
            SC_HIP(hipMemcpyAsync(&hb, best, 8, hipMemcpyDeviceToHost, s));
            SC_HIP(hipStreamSynchronize(s));
            // no finite distance at all: find_lonely keeps its initial fst = 0
            const int64_t fst = hb ? (int64_t)(uint32_t)~(uint32_t)(hb & 0xFFFFFFFFull) : 0;
            if (fst >= n) {
                sc_set_error("sc_kmeans: farthest-point search returned %lld", (long long)fst);
                return SC_ERR_HIP;
            }
            SC_HIP(hipMemcpyAsync(&old, labels + fst, 4, hipMemcpyDeviceToHost, s));
            SC_HIP(hipStreamSynchronize(s));
            hipLaunchKernelGGL(k_move_center, dim3(1), dim3(256), 0, s, d_x, d, fst, j, old,
                               d_centers, sums, counts, labels);
            SC_CHECK_LAUNCH();
            hcnt[old] -= 1;
            hcnt[j] = 1;
        }
        hipLaunchKernelGGL(k_update_centers, dim3((unsigned)sc_cdiv((int64_t)k * d, 256)),
                           dim3(256), 0, s, sums, counts, k, d, d_centers);
        SC_CHECK_LAUNCH();
    }
    return SC_OK;
}

extern "C" int sc_residual(sc_ctx *ctx, float *d_x, int64_t n, int32_t dim, const float *d_cq,
                           int32_t mc, int32_t kc, int32_t which) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_x || !d_cq || n < 0 || dim <= 0 || mc <= 0 || dim % mc || kc < 1 ||
        (which != 0 && which != 1) || (which == 1 && kc < 2)) {
        sc_set_error("sc_residual: bad arguments");
        return SC_ERR_ARG;
    }
    if (n == 0) return SC_OK;
    if (dim <= 256)
        return which ? launch_residual<2, 64, 64>(ctx, d_x, n, dim, d_cq, mc, kc)
                     : launch_residual<1, 64, 64>(ctx, d_x, n, dim, d_cq, mc, kc);
    if (residual_lds<2, 16, 256>(dim) <= 160 * 1024)
        return which ? launch_residual<2, 16, 256>(ctx, d_x, n, dim, d_cq, mc, kc)
                     : launch_residual<1, 16, 256>(ctx, d_x, n, dim, d_cq, mc, kc);
    hipLaunchKernelGGL(k_residual, dim3((unsigned)sc_cdiv(n * mc, 256)), dim3(256), 0,
                       ctx->stream, d_x, n, dim, d_cq, mc, kc, which);
    SC_CHECK_LAUNCH();
    return SC_OK;
}
