// sc_rank.hip -- cell ranking of the two-halves dense-partition index.
//
// A two-halves cell is keyed by the (nearest, second) centroid pair of each
// 64-dim half (src/sc_encoder.cpp:39-50: uid_h = n_h*kc + s_h, key =
// uid_0*kc^2 + uid_1).  The reference's multi-sequences visit cells in
// ascending order of their rank pairs' scores: per half the SC pair score
// v_h[n_h] + v_h[s_h] (include/sc_query.h:113-166) and across the halves the
// IMI sum (include/multi_query.h:175-203).  Over kc^4 cells (4.3e9 at
// kc = 256) nearly all empty, walking that grid pop by pop on one lane is the
// cost; the index keeps only its non-empty cells (a sparse directory), so
// every one of them is scored here, in parallel, and the cells are taken in
// (score ASC, key ASC) order while count < w and sum < T (the loop test of
// include/sc_query.h:130; w counts non-empty cells):
//
//   score = (v_0[n_0] + v_0[s_0]) + (v_1[n_1] + v_1[s_1])      (FP32)
//   v_h[j] = q_h + diff_qc[h*kc + j],  q_h = sequential FP32 sum of the
//            half's squares (include/multi_query.h:102-122)
//
// One workgroup per query: a 4096-bucket histogram of the scores (count and
// cell sizes per bucket, LDS atomics) finds the bucket where the budget is
// reached, a second histogram inside that bucket narrows it when needed, the
// cells up to it are compacted into LDS as (score key << 32 | cell) and
// sorted, and a prefix sum of their sizes gives the stopping point.  A query
// whose cells up to the stopping point do not fit on chip is finished by a
// device-wide segmented sort of all its cells' keys.
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>

#include "sc_internal.h"

namespace {

constexpr int kRankThreads = 256;
constexpr int kBuckets = 4096;     // 12 bits of the score key per level
constexpr int kCap = 4096;         // cells sorted on chip

struct RankKernelArgs {
    RankCellsArgs a;
    int kc, dim;
    int64_t ncells;
    const uint32_t *rec;
    const int32_t *L;
    const int32_t *qlist;   // fallback: the queries of this launch (null = all)
    int cap;                // on-chip list capacity (at most kCap)
};

template <int W>
struct Rec {
    uint32_t n0, s0, n1, s1, size;
    __device__ __forceinline__ Rec(const uint32_t *rec, int64_t c) {
        if (W == 2) {
            const uint2 r = *reinterpret_cast<const uint2 *>(rec + 2 * c);
            n0 = r.x & 255u; s0 = (r.x >> 8) & 255u; n1 = (r.x >> 16) & 255u; s1 = r.x >> 24;
            size = r.y;
        } else {
            const uint4 r = *reinterpret_cast<const uint4 *>(rec + 4 * c);
            n0 = r.x & 0xFFFFu; s0 = r.x >> 16; n1 = r.y & 0xFFFFu; s1 = r.y >> 16;
            size = r.z;
        }
    }
    // the halves' SC pair scores, then their sum: the FP32 operations of the
    // definition above, in that order
    __device__ __forceinline__ float score(const float *v, int kc) const {
        const float a = v[n0] + v[s0];
        const float b = v[kc + n1] + v[kc + s1];
        return a + b;
    }
};

__device__ __forceinline__ int query_budget(const int32_t *Tq, int64_t q, int T) {
    return Tq ? min(max(Tq[q], 0), T) : T;
}

// Exclusive prefix of one value per thread over the workgroup (256 threads);
// *total gets the sum.  All threads call it.
__device__ uint32_t block_excl_scan(uint32_t x, uint32_t *wsum, uint32_t *total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    __syncthreads();
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int i = 0; i < kRankThreads / 64; ++i) {
        if (i < wv) base += wsum[i];
        all += wsum[i];
    }
    *total = all;
    return base + incl - x;
}

// First bucket at which the running count reaches w or the running size sum
// reaches T, given the counts/sizes before the histogram (pre_cnt, pre_w).
// Returns kBuckets when no bucket does; *bcnt/*bw get the totals before the
// found bucket (or of all buckets).
__device__ int find_bucket(const uint32_t *hcnt, const uint32_t *hwt, uint32_t pre_cnt,
                           uint32_t pre_w, int w, int T, uint32_t *wsum, int *s_b,
                           uint32_t *s_pre) {
    constexpr int per = kBuckets / kRankThreads;
    const int t = threadIdx.x;
    uint32_t c = 0, s = 0;
    for (int i = 0; i < per; ++i) { c += hcnt[t * per + i]; s += hwt[t * per + i]; }
    uint32_t tc, ts;
    const uint32_t bc = block_excl_scan(c, wsum, &tc);
    const uint32_t bs = block_excl_scan(s, wsum + 4, &ts);
    if (t == 0) *s_b = kBuckets;
    __syncthreads();
    uint32_t cc = pre_cnt + bc, cw = pre_w + bs;
    for (int i = 0; i < per; ++i) {
        const int b = t * per + i;
        const uint32_t nc = cc + hcnt[b], nw = cw + hwt[b];
        if ((int64_t)nc >= w || (int64_t)nw >= T) {
            atomicMin(s_b, b);
            break;
        }
        cc = nc;
        cw = nw;
    }
    __syncthreads();
    const int b = *s_b;
    // the owner of the found bucket publishes the totals before it
    if (b < kBuckets && t == b / per) {
        uint32_t pc = pre_cnt + bc, pw = pre_w + bs;
        for (int i = t * per; i < b; ++i) { pc += hcnt[i]; pw += hwt[i]; }
        s_pre[0] = pc;
        s_pre[1] = pw;
    }
    if (b == kBuckets && t == 0) {
        s_pre[0] = pre_cnt + tc;
        s_pre[1] = pre_w + ts;
    }
    __syncthreads();
    return b;
}

template <int W>
__global__ __launch_bounds__(kRankThreads) void k_rank_cells(RankKernelArgs k) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t *hcnt = reinterpret_cast<uint32_t *>(smem);           // [kBuckets]
    uint32_t *hwt = hcnt + kBuckets;                               // [kBuckets]
    uint64_t *lkey = reinterpret_cast<uint64_t *>(smem);           // [cap], after the histograms
    float *v = reinterpret_cast<float *>(smem + 8 * kBuckets);     // [2 kc]
    __shared__ uint32_t wsum[8];
    __shared__ int s_b;
    __shared__ uint32_t s_pre[2];
    __shared__ float s_q[3];
    __shared__ int s_n, s_count;
    __shared__ uint32_t s_sum;

    const RankCellsArgs &a = k.a;
    const int tid = threadIdx.x, kc = k.kc;
    const int64_t q = k.qlist ? k.qlist[blockIdx.x] : blockIdx.x;
    const int T = query_budget(a.Tq, q, a.T), w = a.w;
    if (tid == 0) {
        // include/multi_query.h:102-112: per-half sequential sums, q_sum their sum
        const float *qv = a.Q + (size_t)q * k.dim;
        const int bsc = k.dim / 2;
        float qsum = 0.0f, h[2];
        for (int i = 0; i < 2; ++i) {
            float d1 = 0.0f;
            for (int j = 0; j < bsc; ++j) { const float t = qv[i * bsc + j]; d1 += t * t; }
            qsum += d1;
            h[i] = d1;
        }
        s_q[0] = qsum; s_q[1] = h[0]; s_q[2] = h[1];
        float *qs = a.qs + (size_t)q * 4;
        qs[0] = qsum; qs[1] = h[0]; qs[2] = h[1]; qs[3] = 0.0f;
    }
    for (int i = tid; i < kBuckets; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
    const float *dq = a.dqc + (size_t)q * 2 * kc;
    for (int j = tid; j < 2 * kc; j += kRankThreads) v[j] = s_q[j < kc ? 1 : 2] + dq[j];
    __syncthreads();

    int32_t *qi = a.qinfo + (size_t)q * 8;
    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    if (T <= 0 || k.ncells == 0) {          // the loop test fails before the first cell
        if (tid == 0) {
            qi[0] = 0; qi[1] = 0; qi[2] = 0; qi[3] = 0;
            qi[4] = (k.ncells == 0 && T > 0) ? 1 : 0; qi[5] = 0; qi[6] = 0; qi[7] = 0;
        }
        return;
    }

    // level 1: the top 12 bits of the score key
    for (int64_t c = tid; c < k.ncells; c += kRankThreads) {
        const Rec<W> r(k.rec, c);
        const uint32_t key = sc_float_key(r.score(v, kc));
        atomicAdd(&hcnt[key >> 20], 1u);
        atomicAdd(&hwt[key >> 20], r.size);
    }
    __syncthreads();
    const int b1 = find_bucket(hcnt, hwt, 0u, 0u, w, T, wsum, &s_b, s_pre);
    const uint32_t pre1_cnt = s_pre[0], pre1_w = s_pre[1];
    const uint32_t b1cnt = b1 < kBuckets ? hcnt[b1] : 0u;
    __syncthreads();
    // cells to sort: those of the buckets below b1, plus bucket b1 (or its
    // part up to the level-2 bucket b2)
    int b2 = kBuckets - 1;
    uint32_t need = b1 < kBuckets ? pre1_cnt + b1cnt : pre1_cnt;
    bool overflow = false;
    if (need > (uint32_t)k.cap && b1 < kBuckets) {
        // level 2: the next 12 bits, inside bucket b1
        for (int i = tid; i < kBuckets; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
        __syncthreads();
        for (int64_t c = tid; c < k.ncells; c += kRankThreads) {
            const Rec<W> r(k.rec, c);
            const uint32_t key = sc_float_key(r.score(v, kc));
            if ((int)(key >> 20) != b1) continue;
            atomicAdd(&hcnt[(key >> 8) & 0xFFFu], 1u);
            atomicAdd(&hwt[(key >> 8) & 0xFFFu], r.size);
        }
        __syncthreads();
        const int f = find_bucket(hcnt, hwt, pre1_cnt, pre1_w, w, T, wsum, &s_b, s_pre);
        b2 = f < kBuckets ? f : kBuckets - 1;
        need = f < kBuckets ? s_pre[0] + hcnt[f] : s_pre[0];
        __syncthreads();
        overflow = need > (uint32_t)k.cap;
    } else if (need > (uint32_t)k.cap) {
        overflow = true;     // every cell is taken and there are too many of them
    }
    if (overflow) {
        if (tid == 0) {
            qi[0] = 0; qi[1] = 0; qi[2] = 0; qi[3] = 0; qi[4] = 0; qi[5] = 1; qi[6] = 0;
            qi[7] = 0;
        }
        return;
    }

    // compact the cells up to (b1, b2) into LDS as (score key, cell) pairs
    if (tid == 0) s_n = 0;
    __syncthreads();
    for (int64_t c = tid; c < k.ncells; c += kRankThreads) {
        const Rec<W> r(k.rec, c);
        const uint32_t key = sc_float_key(r.score(v, kc));
        const int kb1 = (int)(key >> 20);
        if (kb1 < b1 || (kb1 == b1 && (int)((key >> 8) & 0xFFFu) <= b2)) {
            const int o = atomicAdd(&s_n, 1);
            lkey[o] = ((uint64_t)key << 32) | (uint32_t)c;
        }
    }
    __syncthreads();
    const int n = s_n;
    int np2 = 64;
    while (np2 < n) np2 <<= 1;
    for (int i = n + tid; i < np2; i += kRankThreads) lkey[i] = ~(uint64_t)0;
    __syncthreads();
    for (int size = 2; size <= np2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < np2 / 2; i += kRankThreads) {
                const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t x = lkey[lo], y = lkey[hi];
                if ((x > y) == up) { lkey[lo] = y; lkey[hi] = x; }
            }
            __syncthreads();
        }
    }

    // stopping point: entry i is taken iff i < w and the sizes before it sum
    // to less than T; the taken entries are a prefix
    const int per = (n + kRankThreads - 1) / kRankThreads;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    uint32_t local = 0;
    for (int i = i0; i < i1; ++i) local += Rec<W>(k.rec, (int64_t)(lkey[i] & 0xFFFFFFFFu)).size;
    uint32_t tot;
    uint32_t run = block_excl_scan(local, wsum, &tot);
    if (tid == 0) { s_count = 0; s_sum = 0; }
    __syncthreads();
    for (int i = i0; i < i1; ++i) {
        const int64_t c = (int64_t)(lkey[i] & 0xFFFFFFFFu);
        const Rec<W> r(k.rec, c);
        if (i < w && (int64_t)run < T) {
            if (i < a.ccap) {
                cells[i] = c > 0 ? k.L[c - 1] : 0;
                clen[i] = (int32_t)r.size;
                cinfo[i] = r.n0 | (r.n1 << 16);
            }
            atomicMax(&s_count, i + 1);
            atomicMax(&s_sum, run + r.size);
        }
        run += r.size;
    }
    __syncthreads();
    if (tid == 0) {
        const int count = s_count;
        qi[0] = count; qi[1] = (int32_t)s_sum; qi[2] = count; qi[3] = 0;
        qi[4] = (count == k.ncells && count < w && (int64_t)s_sum < T) ? 1 : 0;
        qi[5] = 0; qi[6] = count > a.ccap ? 2 : 0; qi[7] = 0;
    }
}

// ---------------------------------------------------------------------------
// Fallback for queries whose cells do not fit on chip: every cell's
// (score key << 32 | cell) into global memory, a segmented radix sort per
// query, and a walk of the sorted keys.
// ---------------------------------------------------------------------------
template <int W>
__global__ void k_rank_keys(RankKernelArgs k, const int32_t *list, uint64_t *keys) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *v = reinterpret_cast<float *>(smem);        // [2 kc]
    const int64_t slot = blockIdx.y;
    const int64_t q = list[slot];
    const int kc = k.kc;
    const float *dq = k.a.dqc + (size_t)q * 2 * kc;
    const float *qs = k.a.qs + (size_t)q * 4;          // written by k_rank_cells
    for (int j = threadIdx.x; j < 2 * kc; j += blockDim.x) v[j] = qs[j < kc ? 1 : 2] + dq[j];
    __syncthreads();
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < k.ncells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const Rec<W> r(k.rec, c);
        keys[(size_t)slot * k.ncells + c] =
            ((uint64_t)sc_float_key(r.score(v, kc)) << 32) | (uint32_t)c;
    }
}

template <int W>
__global__ __launch_bounds__(kRankThreads) void k_rank_walk(RankKernelArgs k, const int32_t *list,
                                                            const uint64_t *skeys) {
    __shared__ uint32_t wsum[8];
    __shared__ int s_stop;
    const RankCellsArgs &a = k.a;
    const int64_t slot = blockIdx.x;
    const int64_t q = list[slot];
    const int tid = threadIdx.x;
    const int T = query_budget(a.Tq, q, a.T), w = a.w;
    const uint64_t *sk = skeys + (size_t)slot * k.ncells;
    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    uint32_t base = 0;           // sizes of the entries before this round
    int count = 0;
    for (int64_t i0 = 0; i0 < k.ncells; i0 += kRankThreads) {
        const int64_t i = i0 + tid;
        uint32_t size = 0;
        int64_t c = -1;
        if (i < k.ncells) {
            c = (int64_t)(sk[i] & 0xFFFFFFFFu);
            size = Rec<W>(k.rec, c).size;
        }
        uint32_t tot;
        const uint32_t ex = base + block_excl_scan(size, wsum, &tot);
        if (tid == 0) s_stop = 0;
        __syncthreads();
        const bool take = c >= 0 && i < w && (int64_t)ex < T;
        if (take) {
            if (i < a.ccap) {
                const Rec<W> r(k.rec, c);
                cells[i] = c > 0 ? k.L[c - 1] : 0;
                clen[i] = (int32_t)r.size;
                cinfo[i] = r.n0 | (r.n1 << 16);
            }
        } else if (i < k.ncells) {
            s_stop = 1;
        }
        const int took = __syncthreads_count(take);
        count += took;
        base += tot;
        if (s_stop) break;
        __syncthreads();
    }
    if (tid == 0) {
        // the taken entries are a prefix: their sizes are those summed so far
        // minus the round's untaken tail, recomputed from the sorted keys
        uint32_t sum = 0;
        for (int i = 0; i < count; ++i) sum += Rec<W>(k.rec, (int64_t)(sk[i] & 0xFFFFFFFFu)).size;
        int32_t *qi = a.qinfo + (size_t)q * 8;
        qi[0] = count; qi[1] = (int32_t)sum; qi[2] = count; qi[3] = 0;
        qi[4] = (count == k.ncells && count < w && (int64_t)sum < T) ? 1 : 0;
        qi[5] = 0; qi[6] = count > a.ccap ? 2 : 0; qi[7] = 0;
    }
}

__global__ void k_rank_flagged(const int32_t *__restrict__ qinfo, int64_t nq,
                               int32_t *__restrict__ list, int32_t *__restrict__ n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq && qinfo[q * 8 + 5]) list[atomicAdd(n, 1)] = (int32_t)q;
}

__global__ void k_segment_offsets(int64_t nseg, int64_t len, int64_t *off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nseg) off[i] = i * len;
}

// SC_RANK_CAP (test hook): a smaller on-chip list, so that tests reach the
// level-2 refinement and the device-wide fallback.
int rank_cap() {
    if (const char *e = getenv("SC_RANK_CAP")) {
        const int v = atoi(e);
        if (v >= 1) return std::min(v, kCap);
    }
    return kCap;
}

template <int W>
int rank_cells(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a) {
    RankKernelArgs k{a, ix->kc, ix->dim, ix->ncells, ix->crec, ix->L, nullptr, rank_cap()};
    const size_t lds = 8 * kBuckets + 8 * (size_t)ix->kc;
    if (lds > 160 * 1024) {
        sc_set_error("sc_search: kc=%d too large for the on-chip cell ranking", ix->kc);
        return SC_ERR_ARG;
    }
    auto kern = k_rank_cells<W>;
    SC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)a.nq), dim3(kRankThreads), lds, ctx->stream, k);
    SC_CHECK_LAUNCH();

    // queries flagged for the device-wide path
    int32_t *list = nullptr, *dn = nullptr;
    const size_t lbytes = (size_t)a.nq * 4 + 256;
    if (hipMalloc(&list, lbytes) != hipSuccess) {
        sc_set_error("sc_search: ranking list allocation failed");
        return SC_ERR_NOMEM;
    }
    dn = list + a.nq;
    int32_t nflag = 0;
    int rc = SC_OK;
    if (hipMemsetAsync(dn, 0, 4, ctx->stream) != hipSuccess) rc = SC_ERR_HIP;
    if (rc == SC_OK) {
        hipLaunchKernelGGL(k_rank_flagged, dim3((unsigned)sc_cdiv(a.nq, 256)), dim3(256), 0,
                           ctx->stream, a.qinfo, a.nq, list, dn);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&nflag, dn, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = SC_ERR_HIP;
    }
    if (rc == SC_OK && nflag > 0) {
        // groups of queries whose keys (two buffers of ncells each) fit 1 GiB
        const size_t per = (size_t)ix->ncells * 16;
        const int64_t group = std::max<int64_t>(1, std::min<int64_t>(nflag, ((size_t)1 << 30) / per));
        size_t tb = 0;
        (void)rocprim::segmented_radix_sort_keys(nullptr, tb, (uint64_t *)nullptr,
                                                 (uint64_t *)nullptr, (unsigned)(group * ix->ncells),
                                                 (unsigned)group, (int64_t *)nullptr,
                                                 (int64_t *)nullptr + 1, 0u, 64u, ctx->stream);
        unsigned char *g = nullptr;
        const size_t kb = per / 2 * (size_t)group;
        const size_t total = 2 * kb + 8 * (size_t)(group + 1) + tb + 512;
        if (hipMalloc(&g, total) != hipSuccess) {
            rc = SC_ERR_NOMEM;
            sc_set_error("sc_search: ranking fallback (%zu bytes) could not be allocated", total);
        } else {
            uint64_t *keys = (uint64_t *)g, *skeys = (uint64_t *)(g + kb);
            int64_t *off = (int64_t *)(g + 2 * kb);
            unsigned char *tmp = g + 2 * kb + 8 * (size_t)(group + 1);
            tmp = (unsigned char *)(((uintptr_t)tmp + 255) & ~(uintptr_t)255);
            for (int64_t g0 = 0; g0 < nflag && rc == SC_OK; g0 += group) {
                const int64_t gn = std::min<int64_t>(group, nflag - g0);
                RankKernelArgs fk = k;
                const unsigned gx = (unsigned)std::min<int64_t>(64, sc_cdiv(ix->ncells, 256));
                hipLaunchKernelGGL(k_rank_keys<W>, dim3(gx, (unsigned)gn), dim3(256),
                                   8 * (size_t)ix->kc, ctx->stream, fk, list + g0, keys);
                hipLaunchKernelGGL(k_segment_offsets, dim3((unsigned)sc_cdiv(gn + 1, 256)),
                                   dim3(256), 0, ctx->stream, gn, ix->ncells, off);
                if (hipGetLastError() != hipSuccess) { rc = SC_ERR_HIP; break; }
                size_t tb2 = tb;
                if (rocprim::segmented_radix_sort_keys(tmp, tb2, keys, skeys,
                                                       (unsigned)(gn * ix->ncells), (unsigned)gn,
                                                       off, off + 1, 0u, 64u,
                                                       ctx->stream) != hipSuccess) {
                    rc = SC_ERR_HIP;
                    break;
                }
                hipLaunchKernelGGL(k_rank_walk<W>, dim3((unsigned)gn), dim3(kRankThreads), 0,
                                   ctx->stream, fk, list + g0, skeys);
                if (hipGetLastError() != hipSuccess) rc = SC_ERR_HIP;
            }
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(g);
            if (rc == SC_ERR_HIP) sc_set_error("sc_search: ranking fallback launch failed");
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(list);
    return rc;
}

}  // namespace

int sc_rank_cells(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a) {
    if (a.nq <= 0) return SC_OK;
    if (!ix->sparse) {
        sc_set_error("sc_search: cell ranking needs the two-halves index");
        return SC_ERR_CONFIG;
    }
    return ix->crec_words == 2 ? rank_cells<2>(ctx, ix, a) : rank_cells<4>(ctx, ix, a);
}
