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
constexpr int kBuckets = 4096;     // histogram buckets of the pruned ranking (below)
// The pair-table and record rankings: score-key histograms of 2048 buckets,
// 11 bits of the key per level (level 1: key >> 21, level 2: the next 11),
// and an on-chip list of 2048 cells (16 KB together), so that four
// workgroups fit a CU; queries whose cells do not fit go to the pruned
// ranking (or, without it, a list four times as long).
constexpr int kPB = 2048;
constexpr int kL1 = 21, kL2 = 10;
constexpr uint32_t kL2Mask = 0x7FFu, kLowBits = 0x3FFu;
constexpr int kCap = 2048;         // cells sorted on chip

struct RankKernelArgs {
    RankCellsArgs a;
    int kc, dim;
    int64_t ncells;
    const uint32_t *rec;
    const int32_t *L;
    const int32_t *qlist;   // fallback: the queries of this launch (null = all)
    int cap;                // on-chip list capacity (at most 2 kCap; 4 kCap for a retry)
    // pair tables (k_rank_pairs)
    int nu0, nu1;
    const uint32_t *up0, *up1, *cpair;
    const int32_t *gstart;
    const int32_t *gvec;    // vectors in each half-0 pair's cells
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
template <int NT = kRankThreads>
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
    for (int i = 0; i < NT / 64; ++i) {
        if (i < wv) base += wsum[i];
        all += wsum[i];
    }
    *total = all;
    return base + incl - x;
}

// The exclusive prefixes of two values per thread at once (one pair of
// barriers); wsum holds 2 NT / 64 words.  All threads call it.
template <int NT = kRankThreads>
__device__ void block_excl_scan2(uint32_t x, uint32_t y, uint32_t *wsum, uint32_t *ex,
                                 uint32_t *ey, uint32_t *tx, uint32_t *ty) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t ix = x, iy = y;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t a = __shfl_up(ix, o), b = __shfl_up(iy, o);
        if (lane >= o) { ix += a; iy += b; }
    }
    __syncthreads();
    if (lane == 63) { wsum[wv] = ix; wsum[NT / 64 + wv] = iy; }
    __syncthreads();
    uint32_t bx = 0, by = 0, ax = 0, ay = 0;
    for (int i = 0; i < NT / 64; ++i) {
        const uint32_t u = wsum[i], v = wsum[NT / 64 + i];
        if (i < wv) { bx += u; by += v; }
        ax += u;
        ay += v;
    }
    *ex = bx + ix - x;
    *ey = by + iy - y;
    *tx = ax;
    *ty = ay;
}

// First bucket at which the running count reaches w or the running size sum
// reaches T, given the counts/sizes before the histogram (pre_cnt, pre_w).
// Returns NB when no bucket does; *bcnt/*bw get the totals before the
// found bucket (or of all buckets).
template <int NT = kRankThreads, int NB = kPB>
__device__ int find_bucket(const uint32_t *hcnt, const uint32_t *hwt, uint32_t pre_cnt,
                           uint32_t pre_w, int w, int T, uint32_t *wsum, int *s_b,
                           uint32_t *s_pre) {
    constexpr int per = NB / NT;
    const int t = threadIdx.x;
    uint32_t c = 0, s = 0;
    for (int i = 0; i < per; ++i) { c += hcnt[t * per + i]; s += hwt[t * per + i]; }
    uint32_t tc, ts, bc, bs;
    block_excl_scan2<NT>(c, s, wsum, &bc, &bs, &tc, &ts);
    if (t == 0) *s_b = NB;
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
    if (b < NB && t == b / per) {
        uint32_t pc = pre_cnt + bc, pw = pre_w + bs;
        for (int i = t * per; i < b; ++i) { pc += hcnt[i]; pw += hwt[i]; }
        s_pre[0] = pc;
        s_pre[1] = pw;
    }
    if (b == NB && t == 0) {
        s_pre[0] = pre_cnt + tc;
        s_pre[1] = pre_w + ts;
    }
    __syncthreads();
    return b;
}

// Shared scalars of one workgroup's ranking.
struct RankShared {
    uint32_t wsum[32];      // two block scans of up to 16 waves
    int b;
    uint32_t pre[2];
    float q[3];
    int n, count, stop;
    uint32_t sum, mink;
    uint64_t stopkey;
};

// A cell's score, size and nearest centroids from its ranking record.
template <int W>
struct RecScorer {
    const uint32_t *rec;
    const float *v;
    int kc;
    __device__ __forceinline__ float score(int64_t c) const { return Rec<W>(rec, c).score(v, kc); }
    __device__ __forceinline__ uint32_t size(int64_t c) const { return Rec<W>(rec, c).size; }
    __device__ __forceinline__ uint32_t nearest(int64_t c) const {
        const Rec<W> r(rec, c);
        return r.n0 | (r.n1 << 16);
    }
};
// The same from the per-query pair tables: st0[i] = v_0[n] + v_0[s] for the
// i-th pair of half 0 (st1 likewise), so score = st0[i0] + st1[i1] performs
// exactly the additions of Rec::score.
struct PairScorer {
    const uint32_t *cpair;
    const float *st0, *st1;
    const uint32_t *up0, *up1;
    __device__ __forceinline__ float score(int64_t c) const {
        const uint32_t p = cpair[2 * c];
        return st0[p & 0xFFFFu] + st1[p >> 16];
    }
    __device__ __forceinline__ uint32_t size(int64_t c) const { return cpair[2 * c + 1]; }
    __device__ __forceinline__ uint32_t nearest(int64_t c) const {
        const uint32_t p = cpair[2 * c];
        return (up0[p & 0xFFFFu] & 0xFFFFu) | ((up1[p >> 16] & 0xFFFFu) << 16);
    }
};

// Bitonic sort of lkey[0 .. n) ascending (padded to a power of two >= 64 with
// ~0; the list has room for it).  All threads call.
__device__ void sort_list(uint64_t *lkey, int n) {
    const int tid = threadIdx.x;
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
}

// The stopping entry of a sorted list of n cells: the first i at which the
// count reaches w or the size sum reaches T (n when none).  All threads call.
template <class Sc>
__device__ int stop_entry(const Sc &sc, const uint64_t *lkey, int n, int w, int T,
                          RankShared &sh) {
    const int tid = threadIdx.x;
    const int per = (n + kRankThreads - 1) / kRankThreads;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    uint32_t local = 0;
    for (int i = i0; i < i1; ++i) local += sc.size((int64_t)(lkey[i] & 0xFFFFFFFFu));
    uint32_t tot;
    uint32_t run = block_excl_scan(local, sh.wsum, &tot);
    if (tid == 0) sh.stop = n;
    __syncthreads();
    for (int i = i0; i < i1; ++i) {
        run += sc.size((int64_t)(lkey[i] & 0xFFFFFFFFu));
        if (i + 1 >= w || (int64_t)run >= T) {
            atomicMin(&sh.stop, i);
            break;
        }
    }
    __syncthreads();
    return sh.stop;
}

// Selection by score histograms (count and sizes per bucket of the score
// key's top 12 bits, then the next 12 inside the crossing bucket) over the
// cells `each` hands out (each(f) calls f(c) once per cell, spread over the
// threads): fills lkey[0 .. sh.n) with every one of those cells up to the
// stopping point, or returns false when they do not fit the list.
template <class Sc, class Each>
__device__ bool hist_select_each(const Sc &sc, Each each, int w, int T, int cap, uint32_t *hcnt,
                                 uint32_t *hwt, uint64_t *lkey, RankShared &sh) {
    const int tid = threadIdx.x;
    for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
    each([&](int64_t c) {
        const uint32_t key = sc_float_key(sc.score(c));
        atomicAdd(&hcnt[key >> kL1], 1u);
        atomicAdd(&hwt[key >> kL1], sc.size(c));
    });
    __syncthreads();
    const int b1 = find_bucket(hcnt, hwt, 0u, 0u, w, T, sh.wsum, &sh.b, sh.pre);
    const uint32_t pre1_cnt = sh.pre[0], pre1_w = sh.pre[1];
    const uint32_t b1cnt = b1 < kPB ? hcnt[b1] : 0u;
    __syncthreads();
    int b2 = kPB - 1;
    uint32_t need = b1 < kPB ? pre1_cnt + b1cnt : pre1_cnt;
    if (need > (uint32_t)cap && b1 < kPB) {
        for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
        __syncthreads();
        each([&](int64_t c) {
            const uint32_t key = sc_float_key(sc.score(c));
            if ((int)(key >> kL1) != b1) return;
            atomicAdd(&hcnt[(key >> kL2) & kL2Mask], 1u);
            atomicAdd(&hwt[(key >> kL2) & kL2Mask], sc.size(c));
        });
        __syncthreads();
        const int f = find_bucket(hcnt, hwt, pre1_cnt, pre1_w, w, T, sh.wsum, &sh.b, sh.pre);
        b2 = f < kPB ? f : kPB - 1;
        need = f < kPB ? sh.pre[0] + hcnt[f] : sh.pre[0];
        __syncthreads();
    }
    if (need > (uint32_t)cap) return false;
    if (tid == 0) sh.n = 0;
    __syncthreads();
    each([&](int64_t c) {
        const uint32_t key = sc_float_key(sc.score(c));
        const int kb1 = (int)(key >> kL1);
        if (kb1 < b1 || (kb1 == b1 && (int)((key >> kL2) & kL2Mask) <= b2)) {
            const int o = atomicAdd(&sh.n, 1);
            lkey[o] = ((uint64_t)key << 32) | (uint32_t)c;
        }
    });
    __syncthreads();
    return true;
}

template <class Sc>
__device__ bool hist_select(const Sc &sc, int64_t ncells, int w, int T, int cap, uint32_t *hcnt,
                            uint32_t *hwt, uint64_t *lkey, RankShared &sh) {
    auto all = [&](auto f) {
        for (int64_t c = threadIdx.x; c < ncells; c += kRankThreads) f(c);
    };
    return hist_select_each(sc, all, w, T, cap, hcnt, hwt, lkey, sh);
}

// The score key of the stopping entry of the list lkey[0 .. n) (unsorted):
// the key at which, in (score, cell) order, the count reaches w or the size
// sum reaches T.  A radix select weighted by count and size, 8 key bits a
// pass, from the highest bit in which the list's keys differ.  Returns false
// when the list never reaches the budget.  All threads call.
template <class Sc>
__device__ bool list_stop(const Sc &sc, const uint64_t *lkey, int n, int w, int T, RankShared &sh,
                          uint32_t *stop, uint32_t *hc, uint32_t *hw) {
    __shared__ uint32_t s_and, s_or;
    const int tid = threadIdx.x;
    // this thread's entries (tid + j * kRankThreads, n <= 2 kCap): their
    // sizes are read from global memory once, for every pass below
    constexpr int kPer = 2 * kCap / kRankThreads;
    uint32_t szr[kPer];
    uint32_t a = ~0u, o = 0u;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = tid + j * kRankThreads;
        szr[j] = 0;
        if (i < n) {
            const uint64_t e = lkey[i];
            const uint32_t key = (uint32_t)(e >> 32);
            a &= key;
            o |= key;
            szr[j] = sc.size((int64_t)(e & 0xFFFFFFFFu));
        }
    }
    for (int m = 32; m > 0; m >>= 1) {
        a &= (uint32_t)__shfl_xor(a, m);
        o |= (uint32_t)__shfl_xor(o, m);
    }
    if (tid == 0) { s_and = ~0u; s_or = 0u; }
    __syncthreads();
    if ((tid & 63) == 0) { atomicAnd(&s_and, a); atomicOr(&s_or, o); }
    __syncthreads();
    const uint32_t diff = s_and ^ s_or;
    uint32_t prefix = s_and & ~diff;      // the bits every key shares
    if (diff == 0) {                      // one key value: stop there if reached
        uint32_t cnt = 0, sz = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (tid + j * kRankThreads < n) {
                cnt += 1;
                sz += szr[j];
            }
        }
        uint32_t tc, ts;
        (void)block_excl_scan(cnt, sh.wsum, &tc);
        (void)block_excl_scan(sz, sh.wsum + 4, &ts);
        *stop = prefix;
        return (int64_t)tc >= w || (int64_t)ts >= T;
    }
    const int hb = 31 - __builtin_clz(diff);   // highest differing bit
    uint32_t pre_cnt = 0, pre_w = 0;
    for (int shift = (hb / 8) * 8; shift >= 0; shift -= 8) {
        hc[tid] = 0;
        hw[tid] = 0;
        __syncthreads();
        const uint32_t hi_mask = shift + 8 >= 32 ? 0u : ~0u << (shift + 8);
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = tid + j * kRankThreads;
            if (i >= n) continue;
            const uint32_t key = (uint32_t)(lkey[i] >> 32);
            if ((key & hi_mask) != (prefix & hi_mask)) continue;
            const uint32_t b = (key >> shift) & 255u;
            atomicAdd(&hc[b], 1u);
            atomicAdd(&hw[b], szr[j]);
        }
        __syncthreads();
        if (tid < 64) {   // wave 0: the first bin where the running totals reach the budget
            uint32_t c4[4], w4[4], tc = 0, tw = 0;
            for (int j = 0; j < 4; ++j) {
                c4[j] = hc[4 * tid + j];
                w4[j] = hw[4 * tid + j];
                tc += c4[j];
                tw += w4[j];
            }
            uint32_t ic = tc, iw = tw;
            for (int m = 1; m < 64; m <<= 1) {
                const uint32_t uc = __shfl_up(ic, m), uw = __shfl_up(iw, m);
                if (tid >= m) { ic += uc; iw += uw; }
            }
            uint32_t cc = pre_cnt + ic - tc, cw = pre_w + iw - tw;
            int found = -1;
            uint32_t fc = 0, fw = 0;
            for (int j = 0; j < 4; ++j) {
                if (found < 0 && ((int64_t)(cc + c4[j]) >= w || (int64_t)(cw + w4[j]) >= T)) {
                    found = 4 * tid + j;
                    fc = cc;
                    fw = cw;
                }
                cc += c4[j];
                cw += w4[j];
            }
            // the lowest lane that found a bin publishes it
            const unsigned long long fl = __ballot(found >= 0);
            if (tid == 0) sh.b = -1;
            if (fl != 0 && tid == __ffsll((long long)fl) - 1) {
                sh.b = found;
                sh.pre[0] = fc;
                sh.pre[1] = fw;
            }
        }
        __syncthreads();
        const int b = sh.b;
        if (b < 0) return false;
        prefix = (prefix & hi_mask) | ((uint32_t)b << shift);
        pre_cnt = sh.pre[0];
        pre_w = sh.pre[1];
        __syncthreads();
    }
    *stop = prefix;
    return true;
}

// Moves the entries of score key at most `stop` to the front of lkey
// (n <= 2 kCap); returns their number.  All threads call.
__device__ int list_keep(uint64_t *lkey, int n, uint32_t stop, RankShared &sh) {
    const int tid = threadIdx.x;
    constexpr int kPer = 2 * kCap / kRankThreads;
    uint64_t mine[kPer];
    int m = 0;
    for (int i = tid; i < n; i += kRankThreads)
        if ((uint32_t)(lkey[i] >> 32) <= stop) mine[m++] = lkey[i];
    if (tid == 0) sh.n = 0;
    __syncthreads();
    for (int j = 0; j < m; ++j) lkey[atomicAdd(&sh.n, 1)] = mine[j];
    __syncthreads();
    return sh.n;
}

// Sorts lkey[0 .. n) and writes the taken prefix to the query's cell list.
template <class Sc>
__device__ void emit_sorted(const Sc &sc, const RankKernelArgs &k, int64_t q, uint64_t *lkey, int n,
                            int w, int T, RankShared &sh) {
    const RankCellsArgs &a = k.a;
    const int tid = threadIdx.x;
    sort_list(lkey, n);
    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    // entry i is taken iff i < w and the sizes before it sum to less than T;
    // the taken entries are a prefix
    const int per = (n + kRankThreads - 1) / kRankThreads;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    constexpr int kPer = kCap / kRankThreads;
    if (per <= kPer) {
        // each entry's size, first position and nearest centroids loaded
        // together, once, and kept in registers
        uint32_t szr[kPer], nr[kPer];
        int32_t str[kPer];
        uint32_t local = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = i0 + j;
            szr[j] = 0; nr[j] = 0; str[j] = 0;
            if (i < i1) {
                const int64_t c = (int64_t)(lkey[i] & 0xFFFFFFFFu);
                szr[j] = sc.size(c);
                str[j] = c > 0 ? k.L[c - 1] : 0;
                nr[j] = sc.nearest(c);
                local += szr[j];
            }
        }
        uint32_t tot;
        uint32_t run = block_excl_scan(local, sh.wsum, &tot);
        if (tid == 0) { sh.count = 0; sh.sum = 0; }
        __syncthreads();
        int cnt = 0;
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = i0 + j;
            if (i < i1 && i < w && (int64_t)run < T) {
                if (i < a.ccap) {
                    cells[i] = str[j];
                    clen[i] = (int32_t)szr[j];
                    cinfo[i] = nr[j];
                }
                cnt = i + 1;
                sum = run + szr[j];
            }
            run += szr[j];
        }
        if (cnt > 0) {
            atomicMax(&sh.count, cnt);
            atomicMax(&sh.sum, sum);
        }
        __syncthreads();
        if (tid == 0) {
            const int count = sh.count;
            int32_t *qi = a.qinfo + (size_t)q * 8;
            qi[0] = count; qi[1] = (int32_t)sh.sum; qi[2] = count; qi[3] = 0;
            qi[4] = (count == k.ncells && count < w && (int64_t)sh.sum < T) ? 1 : 0;
            qi[5] = 0; qi[6] = count > a.ccap ? 2 : 0; qi[7] = 0;
        }
        return;
    }
    uint32_t local = 0;
    for (int i = i0; i < i1; ++i) local += sc.size((int64_t)(lkey[i] & 0xFFFFFFFFu));
    uint32_t tot;
    uint32_t run = block_excl_scan(local, sh.wsum, &tot);
    if (tid == 0) { sh.count = 0; sh.sum = 0; }
    __syncthreads();
    for (int i = i0; i < i1; ++i) {
        const int64_t c = (int64_t)(lkey[i] & 0xFFFFFFFFu);
        const uint32_t size = sc.size(c);
        if (i < w && (int64_t)run < T) {
            if (i < a.ccap) {
                cells[i] = c > 0 ? k.L[c - 1] : 0;
                clen[i] = (int32_t)size;
                cinfo[i] = sc.nearest(c);
            }
            atomicMax(&sh.count, i + 1);
            atomicMax(&sh.sum, run + size);
        }
        run += size;
    }
    __syncthreads();
    if (tid == 0) {
        const int count = sh.count;
        int32_t *qi = a.qinfo + (size_t)q * 8;
        qi[0] = count; qi[1] = (int32_t)sh.sum; qi[2] = count; qi[3] = 0;
        qi[4] = (count == k.ncells && count < w && (int64_t)sh.sum < T) ? 1 : 0;
        qi[5] = 0; qi[6] = count > a.ccap ? 2 : 0; qi[7] = 0;
    }
}

// The query terms of include/multi_query.h:102-112 (per-half sequential sums,
// q_sum their sum) and v_h[j] = q_h + diff_qc[h*kc + j] in LDS.
template <int NT = kRankThreads>
__device__ void query_terms(const RankKernelArgs &k, int64_t q, float *v, RankShared &sh) {
    const RankCellsArgs &a = k.a;
    const int tid = threadIdx.x, kc = k.kc;
    // the query row goes to LDS first (all threads), so that the sequential
    // sums below wait on LDS, not on one global load after another
    float *qrow = v;                               // [dim] <= [2 kc] (checked by the host)
    const float *dq = a.dqc + (size_t)q * 2 * kc;
    // the first entries of this thread's diff_qc, loaded with the query row
    constexpr int kPre = 4;
    float dpre[kPre];
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
        const int j = tid + i * NT;
        dpre[i] = j < 2 * kc ? dq[j] : 0.0f;
    }
    for (int j = tid; j < k.dim; j += NT) qrow[j] = a.Q[(size_t)q * k.dim + j];
    __syncthreads();
    if (tid == 0) {
        const int bsc = k.dim / 2;
        float qsum = 0.0f, h[2];
        for (int i = 0; i < 2; ++i) {
            float d1 = 0.0f;
            for (int j = 0; j < bsc; ++j) { const float t = qrow[i * bsc + j]; d1 += t * t; }
            qsum += d1;
            h[i] = d1;
        }
        sh.q[0] = qsum; sh.q[1] = h[0]; sh.q[2] = h[1];
        float *qs = a.qs + (size_t)q * 4;
        qs[0] = qsum; qs[1] = h[0]; qs[2] = h[1]; qs[3] = 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
        const int j = tid + i * NT;
        if (j < 2 * kc) v[j] = sh.q[j < kc ? 1 : 2] + dpre[i];
    }
    for (int j = tid + kPre * NT; j < 2 * kc; j += NT) v[j] = sh.q[j < kc ? 1 : 2] + dq[j];
    __syncthreads();
}

__device__ bool empty_query(const RankKernelArgs &k, int64_t q, int T) {
    if (T > 0 && k.ncells > 0) return false;
    if (threadIdx.x == 0) {           // the loop test fails before the first cell
        int32_t *qi = k.a.qinfo + (size_t)q * 8;
        qi[0] = 0; qi[1] = 0; qi[2] = 0; qi[3] = 0;
        qi[4] = (k.ncells == 0 && T > 0) ? 1 : 0; qi[5] = 0; qi[6] = 0; qi[7] = 0;
    }
    return true;
}

__device__ void flag_overflow(const RankKernelArgs &k, int64_t q) {
    if (threadIdx.x == 0) {
        int32_t *qi = k.a.qinfo + (size_t)q * 8;
        qi[0] = 0; qi[1] = 0; qi[2] = 0; qi[3] = 0; qi[4] = 0; qi[5] = 1; qi[6] = 0; qi[7] = 0;
    }
}

// LDS of the ranking kernels: region A (the histograms, or the cell list of
// up to cap entries) of k.abytes, then v [2 kc], then the pair tables.
// Floats of the v region of the ranking kernels: v [2 kc] (the query row
// [dim] first), and at least 512 (the pair ranking's list select reuses it).
__host__ __device__ inline int rank_vwords(int kc, int dim) {
    const int m = 2 * kc > dim ? 2 * kc : dim;
    return m > 512 ? m : 512;
}

__host__ __device__ inline size_t rank_abytes(int cap) {
    int np2 = 64;
    while (np2 < cap) np2 <<= 1;
    return 8 * (size_t)(np2 > kPB ? np2 : kPB);
}

// Ranking from the cells' records (any index shape).
template <int W>
__global__ __launch_bounds__(kRankThreads) void k_rank_cells(RankKernelArgs k) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t *hcnt = reinterpret_cast<uint32_t *>(smem);           // [kPB]
    uint32_t *hwt = hcnt + kPB;                                    // [kPB]
    uint64_t *lkey = reinterpret_cast<uint64_t *>(smem);           // [cap], after the histograms
    float *v = reinterpret_cast<float *>(smem + rank_abytes(k.cap));   // [2 kc]
    __shared__ RankShared sh;
    const int64_t q = k.qlist ? k.qlist[blockIdx.x] : blockIdx.x;
    const int T = query_budget(k.a.Tq, q, k.a.T), w = k.a.w;
    query_terms(k, q, v, sh);
    if (empty_query(k, q, T)) return;
    const RecScorer<W> sc{k.rec, v, k.kc};
    if (!hist_select(sc, k.ncells, w, T, k.cap, hcnt, hwt, lkey, sh)) {
        flag_overflow(k, q);
        return;
    }
    emit_sorted(sc, k, q, lkey, sh.n, w, T, sh);
}

// Ranking from the pair tables, with a bound that skips most cells: the cells
// of the half-0 pairs of least st0 whose cells together reach the budget form
// a subset whose own stopping score bounds the full set's (a subset reaches
// the budget no earlier); two histogram levels over the subset's cells give
// that score to 8 bits, thr.  Then only pairs whose least possible cell
// score, st0[g] + min st1, is at most thr are visited, and only cells scoring
// at most thr are listed; when more than the list holds, the histogram
// selection runs over those pairs' cells, and over all cells when the subset
// has too many pairs.
constexpr int kMaxPairs0 = 2048;     // half-0 pairs sorted on chip for the bound
constexpr int kMaxPass = 256;        // pairs whose cells are visited by all threads at once

// Every cell of the groups (half-0 pairs) gbeg/gpre describe is handed to f
// (group slot, cell), all threads at once: gpre[j] is the number of cells in
// the groups before j, gbeg[j] the first cell of group j.
template <class F>
__device__ void for_group_cells(const int32_t *gbeg, const int32_t *gpre, int ng, F f) {
    const int total = gpre[ng];
    for (int t = threadIdx.x; t < total; t += kRankThreads) {
        int lo = 0, hi = ng;           // the last j with gpre[j] <= t
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (gpre[mid] <= t) lo = mid; else hi = mid;
        }
        f(lo, (int64_t)gbeg[lo] + (t - gpre[lo]));
    }
}

// gp[0 .. ng) holds the group sizes on entry and their exclusive prefix on
// exit, gp[ng] the total (ng <= kMaxPass).  All threads call.
__device__ void group_prefix(int32_t *gp, int ng, RankShared &sh) {
    constexpr int per = kMaxPass / kRankThreads;
    const int t = threadIdx.x;
    int32_t sz[per];
    uint32_t loc = 0;
#pragma unroll
    for (int i = 0; i < per; ++i) {
        const int j = t * per + i;
        sz[i] = j < ng ? gp[j] : 0;
        loc += (uint32_t)sz[i];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(loc, sh.wsum, &tot);   // (its barrier orders the reads above)
#pragma unroll
    for (int i = 0; i < per; ++i) {
        const int j = t * per + i;
        if (j < ng) { gp[j] = (int32_t)run; run += (uint32_t)sz[i]; }
    }
    if (t == 0) gp[ng] = (int32_t)tot;
    __syncthreads();
}

// Lists of at most kSmallEmit cells are ordered without sorting: entry i's
// rank is the number of entries of smaller key and its position base the sum
// of their sizes, counted over the whole list (every thread reads the same
// entry at a time, an LDS broadcast).  The sizes sit in LDS after the keys
// (room for kSmallList of them).
constexpr int kSmallList = 1024;
constexpr int kSmallEmit = 256;     // lists up to this long are emitted by counting

template <class Sc>
__device__ void small_ranks(const Sc &sc, const uint64_t *lkey, int n, uint32_t *lsz) {
    for (int i = threadIdx.x; i < n; i += kRankThreads)
        lsz[i] = sc.size((int64_t)(lkey[i] & 0xFFFFFFFFu));
    __syncthreads();
}

// Rank (entries of smaller key) and position base (their sizes' sum) of
// entry i of a small list, counting over all n entries.
__device__ __forceinline__ void small_rank_of(const uint64_t *lkey, const uint32_t *lsz, int n,
                                              uint64_t ki, int *r_out, uint32_t *base_out) {
    int r = 0;
    uint32_t base = 0;
#pragma unroll 8
    for (int j = 0; j < n; ++j) {
        const bool before = lkey[j] < ki;
        r += before;
        base += before ? lsz[j] : 0u;
    }
    *r_out = r;
    *base_out = base;
}

// emit_sorted for a small list: every entry's rank and position base by
// counting, the taken ones written at their rank.
template <class Sc>
__device__ void emit_small(const Sc &sc, const RankKernelArgs &k, int64_t q, uint64_t *lkey, int n,
                           int w, int T, RankShared &sh) {
    const RankCellsArgs &a = k.a;
    uint32_t *lsz = reinterpret_cast<uint32_t *>(lkey + kSmallList);
    small_ranks(sc, lkey, n, lsz);
    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    if (threadIdx.x == 0) { sh.count = 0; sh.sum = 0; }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kRankThreads) {
        const uint64_t ki = lkey[i];
        int r;
        uint32_t base;
        small_rank_of(lkey, lsz, n, ki, &r, &base);
        if (r < w && (int64_t)base < T) {
            const int64_t c = (int64_t)(ki & 0xFFFFFFFFu);
            if (r < a.ccap) {
                cells[r] = c > 0 ? k.L[c - 1] : 0;
                clen[r] = (int32_t)lsz[i];
                cinfo[r] = sc.nearest(c);
            }
            atomicMax(&sh.count, r + 1);
            atomicMax(&sh.sum, base + lsz[i]);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int count = sh.count;
        int32_t *qi = a.qinfo + (size_t)q * 8;
        qi[0] = count; qi[1] = (int32_t)sh.sum; qi[2] = count; qi[3] = 0;
        qi[4] = (count == k.ncells && count < w && (int64_t)sh.sum < T) ? 1 : 0;
        qi[5] = 0; qi[6] = count > a.ccap ? 2 : 0; qi[7] = 0;
    }
}

// An upper bound of the stopping score key of the cells `each` hands out
// (each(f) calls f(c) once per cell, spread over the threads): the key bits
// of the histogram buckets (top 11 bits, then the next 11) in which the
// running count reaches w or the running size reaches T, the remaining bits
// set.  Returns false when those cells never reach the budget.
template <class Sc, class Each>
__device__ bool hist_stop(const Sc &sc, Each each, int w, int T, uint32_t *hcnt, uint32_t *hwt,
                          RankShared &sh, uint32_t *thr) {
    const int tid = threadIdx.x;
    for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
    each([&](int64_t c) {
        const uint32_t key = sc_float_key(sc.score(c));
        atomicAdd(&hcnt[key >> kL1], 1u);
        atomicAdd(&hwt[key >> kL1], sc.size(c));
    });
    __syncthreads();
    const int b1 = find_bucket(hcnt, hwt, 0u, 0u, w, T, sh.wsum, &sh.b, sh.pre);
    const uint32_t pre1_cnt = sh.pre[0], pre1_w = sh.pre[1];
    __syncthreads();
    if (b1 >= kPB) return false;
    for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
    each([&](int64_t c) {
        const uint32_t key = sc_float_key(sc.score(c));
        if ((int)(key >> kL1) != b1) return;
        atomicAdd(&hcnt[(key >> kL2) & kL2Mask], 1u);
        atomicAdd(&hwt[(key >> kL2) & kL2Mask], sc.size(c));
    });
    __syncthreads();
    const int b2 = find_bucket(hcnt, hwt, pre1_cnt, pre1_w, w, T, sh.wsum, &sh.b, sh.pre);
    __syncthreads();
    // the level-1 bucket reaches the budget, so some level-2 bucket in it does
    *thr = ((uint32_t)b1 << kL1) | ((uint32_t)min(b2, kPB - 1) << kL2) | kLowBits;
    return true;
}

// An upper bound of the stopping score key of the cells of the groups
// gbeg/gpre describe (ng of them): the key bits of the histogram buckets
// (top 11 bits, then the next 11) in which the running count reaches w or
// the running size reaches T, the remaining bits set.  Returns false when
// those cells never reach the budget.  A thread keeps the keys and sizes of
// its first kR cells in registers for the second level.  All threads call.
template <class Sc>
__device__ bool subset_stop(const Sc &sc, const int32_t *gbeg, const int32_t *gpre, int ng, int w,
                            int T, uint32_t *hcnt, uint32_t *hwt, RankShared &sh, uint32_t *thr) {
    constexpr int kR = 8;
    const int tid = threadIdx.x;
    const int total = gpre[ng];
    auto cell_of = [&](int t) -> int64_t {   // the last group j with gpre[j] <= t
        int lo = 0, hi = ng;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (gpre[mid] <= t) lo = mid; else hi = mid;
        }
        return (int64_t)gbeg[lo] + (t - gpre[lo]);
    };
    uint32_t kr[kR], sr[kR];
    for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kR; ++j) {
        const int t = tid + j * kRankThreads;
        kr[j] = ~0u;
        sr[j] = 0;
        if (t < total) {
            const int64_t c = cell_of(t);
            kr[j] = sc_float_key(sc.score(c));
            sr[j] = sc.size(c);
            atomicAdd(&hcnt[kr[j] >> kL1], 1u);
            atomicAdd(&hwt[kr[j] >> kL1], sr[j]);
        }
    }
    for (int t = tid + kR * kRankThreads; t < total; t += kRankThreads) {
        const int64_t c = cell_of(t);
        const uint32_t key = sc_float_key(sc.score(c));
        atomicAdd(&hcnt[key >> kL1], 1u);
        atomicAdd(&hwt[key >> kL1], sc.size(c));
    }
    __syncthreads();
    const int b1 = find_bucket(hcnt, hwt, 0u, 0u, w, T, sh.wsum, &sh.b, sh.pre);
    const uint32_t pre1_cnt = sh.pre[0], pre1_w = sh.pre[1];
    __syncthreads();
    if (b1 >= kPB) return false;
    for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kR; ++j) {
        if (kr[j] != ~0u && (int)(kr[j] >> kL1) == b1) {
            atomicAdd(&hcnt[(kr[j] >> kL2) & kL2Mask], 1u);
            atomicAdd(&hwt[(kr[j] >> kL2) & kL2Mask], sr[j]);
        }
    }
    for (int t = tid + kR * kRankThreads; t < total; t += kRankThreads) {
        const int64_t c = cell_of(t);
        const uint32_t key = sc_float_key(sc.score(c));
        if ((int)(key >> kL1) != b1) continue;
        atomicAdd(&hcnt[(key >> kL2) & kL2Mask], 1u);
        atomicAdd(&hwt[(key >> kL2) & kL2Mask], sc.size(c));
    }
    __syncthreads();
    const int b2 = find_bucket(hcnt, hwt, pre1_cnt, pre1_w, w, T, sh.wsum, &sh.b, sh.pre);
    __syncthreads();
    // the level-1 bucket reaches the budget, so some level-2 bucket in it does
    *thr = ((uint32_t)b1 << kL1) | ((uint32_t)min(b2, kPB - 1) << kL2) | kLowBits;
    return true;
}

__global__ __launch_bounds__(kRankThreads) void k_rank_pairs(RankKernelArgs k) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint32_t *hcnt = reinterpret_cast<uint32_t *>(smem);           // [kPB]
    uint32_t *hwt = hcnt + kPB;                                    // [kPB]
    uint64_t *lkey = reinterpret_cast<uint64_t *>(smem);           // [cap]
    float *v = reinterpret_cast<float *>(smem + rank_abytes(k.cap));   // [2 kc]
    const int nu0 = k.nu0, nu1 = k.nu1;
    float *st0 = v + rank_vwords(k.kc, k.dim);                     // [nu0]
    float *st1 = st0 + nu0;                                        // [nu1]
    __shared__ RankShared sh;
    __shared__ int32_t gbeg[kMaxPass], gpre[kMaxPass + 1];   // gpre: sizes, then their prefix
    __shared__ int s_np;
    __shared__ unsigned long long s_acc;
    __shared__ uint32_t s_sub[kMaxPairs0 / 32];   // the subset's pairs, one bit each
    const int tid = threadIdx.x;
    const int64_t q = k.qlist ? k.qlist[blockIdx.x] : blockIdx.x;
    const int T = query_budget(k.a.Tq, q, k.a.T), w = k.a.w;
    query_terms(k, q, v, sh);
    if (empty_query(k, q, T)) return;
    const int kc = k.kc;
    if (tid == 0) sh.mink = ~0u;
    for (int i = tid; i < nu0; i += kRankThreads) {
        const uint32_t p = k.up0[i];
        st0[i] = v[p & 0xFFFFu] + v[p >> 16];
    }
    __syncthreads();
    uint32_t mk = ~0u;
    for (int i = tid; i < nu1; i += kRankThreads) {
        const uint32_t p = k.up1[i];
        const float s1 = v[kc + (p & 0xFFFFu)] + v[kc + (p >> 16)];
        st1[i] = s1;
        mk = min(mk, sc_float_key(s1));
    }
    for (int o = 32; o > 0; o >>= 1) mk = min(mk, (uint32_t)__shfl_xor(mk, o));
    if ((tid & 63) == 0) atomicMin(&sh.mink, mk);
    __syncthreads();
    const PairScorer sc{k.cpair, st0, st1, k.up0, k.up1};
    const float s1min = sc_key_float(sh.mink);

    // the subset: the half-0 pairs of least st0 whose cells together reach
    // the budget (a weighted histogram of st0 over the pairs, weighted by
    // their cells' count and size), and their cells
    for (int i = tid; i < kPB; i += kRankThreads) { hcnt[i] = 0; hwt[i] = 0; }
    __syncthreads();
    for (int g = tid; g < nu0; g += kRankThreads) {
        const int c0 = k.gstart[g], c1 = k.gstart[g + 1];
        const uint32_t b = sc_float_key(st0[g]) >> kL1;
        atomicAdd(&hcnt[b], (uint32_t)(c1 - c0));
        atomicAdd(&hwt[b], (uint32_t)k.gvec[g]);
    }
    __syncthreads();
    const int b0 = find_bucket(hcnt, hwt, 0u, 0u, w, T, sh.wsum, &sh.b, sh.pre);
    // of the crossing bucket, only as many pairs as the budget needs (any of
    // them: every subset that reaches the budget gives a valid bound)
    const int64_t need_c = (int64_t)w - sh.pre[0], need_w = (int64_t)T - sh.pre[1];
    __syncthreads();
    if (tid == 0) { sh.n = 0; s_np = 0; s_acc = 0; }
    for (int i = tid; i < kMaxPairs0 / 32; i += kRankThreads) s_sub[i] = 0;
    __syncthreads();
    for (int g = tid; g < nu0; g += kRankThreads) {
        const int gb = (int)(sc_float_key(st0[g]) >> kL1);
        if (gb > b0) continue;
        if (gb == b0) {
            const int c0 = k.gstart[g], c1 = k.gstart[g + 1];
            const uint32_t gw = (uint32_t)k.gvec[g];
            // one 64-bit counter (cells high, sizes low: neither sum reaches
            // 2^32), so that every pair sees one consistent running total
            const unsigned long long old =
                atomicAdd(&s_acc, ((unsigned long long)(c1 - c0) << 32) | gw);
            if ((int64_t)(old >> 32) >= need_c || (int64_t)(old & 0xFFFFFFFFull) >= need_w)
                continue;
        }
        const int o = atomicAdd(&s_np, 1);
        if (o < kMaxPass) {
            gbeg[o] = k.gstart[g];
            gpre[o] = k.gstart[g + 1] - gbeg[o];
        }
        atomicOr(&s_sub[g >> 5], 1u << (g & 31));
    }
    __syncthreads();
    // the subset's stopping score (to histogram resolution) bounds the full set's
    uint32_t thr = 0;
    bool have = false;
    if (s_np <= kMaxPass) {
        const int np0 = s_np;
        group_prefix(gpre, np0, sh);
        have = subset_stop(sc, gbeg, gpre, np0, w, T, hcnt, hwt, sh, &thr);
    } else {
        // more pairs than the list holds: every cell, those of the marked
        // pairs taken (consecutive threads read consecutive cells, no thread
        // walks a long pair alone)
        auto sub = [&](auto f) {
            for (int64_t c = tid; c < k.ncells; c += kRankThreads) {
                const uint32_t g = k.cpair[2 * c] & 0xFFFFu;
                if ((s_sub[g >> 5] >> (g & 31)) & 1u) f(c);
            }
        };
        have = hist_stop(sc, sub, w, T, hcnt, hwt, sh, &thr);
    }
    __syncthreads();
    if (have) {
        if (tid == 0) { sh.n = 0; s_np = 0; }
        __syncthreads();
        // the pairs that can hold a cell scoring at most thr
        for (int g = tid; g < nu0; g += kRankThreads) {
            if (sc_float_key(st0[g] + s1min) > thr) continue;
            const int o = atomicAdd(&s_np, 1);
            if (o < kMaxPass) {
                gbeg[o] = k.gstart[g];
                gpre[o] = k.gstart[g + 1] - gbeg[o];
            }
        }
        __syncthreads();
        const int np = s_np;
        if (np <= kMaxPass) group_prefix(gpre, np, sh);
        auto cand = [&](auto f) {
            if (np <= kMaxPass) {
                for_group_cells(gbeg, gpre, np, [&](int, int64_t c) { f(c); });
            } else {
                // more pairs than the list holds: every cell whose pair can
                // hold a cell scoring at most thr, consecutive threads on
                // consecutive cells
                for (int64_t c = tid; c < k.ncells; c += kRankThreads) {
                    const uint32_t g = k.cpair[2 * c] & 0xFFFFu;
                    if (sc_float_key(st0[g] + s1min) <= thr) f(c);
                }
            }
        };
        cand([&](int64_t c) {
            const uint32_t key = sc_float_key(sc.score(c));
            if (key <= thr) {
                const int o = atomicAdd(&sh.n, 1);
                if (o < k.cap) lkey[o] = ((uint64_t)key << 32) | (uint32_t)c;
            }
        });
        __syncthreads();
        int n = sh.n;
        if (n <= k.cap) {
            // only the entries up to the exact stopping key are ordered
            uint32_t stop;
            // (the select's two 256-bin histograms in v, free once the pair
            // scores are taken)
            uint32_t *hc = reinterpret_cast<uint32_t *>(v);
            if (n > kSmallEmit && n <= 2 * kCap &&
                list_stop(sc, lkey, n, w, T, sh, &stop, hc, hc + 256))
                n = list_keep(lkey, n, stop, sh);
            if (n <= kSmallEmit) emit_small(sc, k, q, lkey, n, w, T, sh);
            else emit_sorted(sc, k, q, lkey, n, w, T, sh);
            return;
        }
        // more cells score at most thr than the list holds: the histogram
        // selection over the candidate pairs' cells (every cell up to the
        // stopping point is among them)
        __syncthreads();
        if (!hist_select_each(sc, cand, w, T, k.cap, hcnt, hwt, lkey, sh)) {
            flag_overflow(k, q);
            return;
        }
        emit_sorted(sc, k, q, lkey, sh.n, w, T, sh);
        return;
    }
    if (!hist_select(sc, k.ncells, w, T, k.cap, hcnt, hwt, lkey, sh)) {
        flag_overflow(k, q);
        return;
    }
    emit_sorted(sc, k, q, lkey, sh.n, w, T, sh);
}

// ---------------------------------------------------------------------------
// Pruned ranking of large indexes (kc up to 4096, tens of millions of cells):
// scoring every cell for every query is out of reach, so only the cells that
// can lie before the stopping point are visited.
//
// Lower bounds (round-to-nearest addition is monotone in each operand):
//   P_h = v_h[n_h] + v_h[s_h] >= LB_h = m_h1 + m_h2 (the half's two smallest
//   v, as n_h != s_h), so score >= base = LB_0 + LB_1; the cells of a half-0
//   pair score at least P_0 + LB_1, and those of a nearest-0 centroid n at
//   least (v_0[n] + m_01) + LB_1.
// Buckets: d = score - base >= 0 and its top 12 bits (exponent and four
// mantissa bits) order like d, so like the score: a log-spaced histogram
// whose buckets are 1/16 of their value wide, whatever the data's scale.
//
// k_rank_big (one workgroup per query): the nearest-0 centroids are ordered
// by the bucket of their bound (a counting sort) and handed to the waves in
// that order; a wave scores the cells of its centroid's pairs and appends
// every cell whose bucket is at most the current cut to the query's list,
// adding it to a histogram of counts and sizes.  The cut -- the bucket where
// the histogram's running count reaches w or its running size sum reaches T
// -- only falls as cells are added, and it never falls below the cut of the
// full set of cells (a histogram over fewer cells reaches the budget no
// earlier).  A centroid, pair or cell whose bound's bucket exceeds the cut
// is skipped, so every cell whose bucket is at most the final cut has been
// appended and counted: the histogram's running totals are exact up to it
// and its crossing bucket is the full set's.
//
// k_rank_big_emit: the entries up to the final cut, counting-sorted by their
// score key into fine buckets; each entry's rank and position base (the
// sizes of the entries before it) in (score, cell) order are its bucket's
// prefix plus a count over its own bucket; the taken ones (rank < w, base <
// T) are written at their rank.
//
// A query whose list outgrows its capacity is run again starting from the
// final cut of its first run (then only the cells up to that cut are
// appended); if that still overflows, the device-wide sort finishes it.
// ---------------------------------------------------------------------------
constexpr int kBigCap = 32768;       // appended entries per query
constexpr int kBigThreads = 1024;    // threads of the pruned ranking's workgroups
constexpr int kFine = 4096;          // fine buckets of the emit kernel

struct BigArgs {
    RankKernelArgs k;
    const int32_t *bg_n0;
    const uint32_t *bg_pair;
    const int32_t *bg_start;
    const uint32_t *brec;
    const uint64_t *ckeys;
    uint64_t *list, *sorted;
    uint32_t *size, *aux;       // aux [nq][4]: appended, final bucket, base bits, overflow
    int cap;
    int rerun;                  // 1: start from the cut aux holds
    // A query split over `parts` workgroups (small batches): workgroup p
    // takes the nearest-0 centroids n with n % parts == p; they share the
    // query's cut, list counter and (at the end) histograms through split
    // [nq][kSplitWords] = {cut, appended, finished, 0, counts [kBuckets],
    // sizes [kBuckets]}.
    int parts;
    uint32_t *split;
    // 0: the whole ranking; 1: only the warm-up, whose cut goes to split;
    // 2: a query's share (of `parts`) from the cut in split
    int mode;
    // 1: the per-query scratch (aux, list, sorted, size) is indexed by the
    // query's slot in k.qlist rather than by the query (a re-run of a few
    // queries of a batch with scratch for those few only)
    int by_slot;
    // with k.qlist: workgroups at or past *qcount (device) exit at once
    const int32_t *qcount = nullptr;
    // 1 (mode 0 only): the ranking workgroup emits its query's cell list
    // itself (no emission launch; its LDS must hold the emission's)
    int emit = 0;
};
constexpr int kSplitWords = 4 + 2 * 4096;

__device__ __forceinline__ int dbucket(float s, float base) {
    const uint32_t u = __float_as_uint(s - base);
    return (u & 0x80000000u) ? 0 : (int)(u >> 19);
}

__device__ __forceinline__ uint32_t big_size(const BigArgs &b, int64_t c, uint32_t r) {
    const uint32_t sz = r >> 24;
    return sz < 255u ? sz : (uint32_t)(b.k.L[c] - (c > 0 ? b.k.L[c - 1] : 0));
}

// The two smallest values of x[0 .. n) (ascending; equal when the least
// value occurs twice).  All threads call; every thread gets the result.
template <int NT>
__device__ void two_smallest(const float *x, int n, float *scratch, float *lo, float *hi) {
    const int tid = threadIdx.x;
    float a = __builtin_huge_valf(), b = __builtin_huge_valf();
    for (int i = tid; i < n; i += NT) {
        const float t = x[i];
        if (t < a) { b = a; a = t; } else if (t < b) { b = t; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float a2 = __shfl_xor(a, o), b2 = __shfl_xor(b, o);
        const float na = fminf(a, a2), nb = fminf(fmaxf(a, a2), fminf(b, b2));
        a = na;
        b = nb;
    }
    __syncthreads();
    if ((tid & 63) == 0) { scratch[2 * (tid >> 6)] = a; scratch[2 * (tid >> 6) + 1] = b; }
    __syncthreads();
    a = scratch[0];
    b = scratch[1];
    for (int i = 1; i < NT / 64; ++i) {
        const float a2 = scratch[2 * i], b2 = scratch[2 * i + 1];
        const float na = fminf(a, a2), nb = fminf(fmaxf(a, a2), fminf(b, b2));
        a = na;
        b = nb;
    }
    *lo = a;
    *hi = b;
    __syncthreads();
}

// The crossing bucket of the histograms, read by one wave while other waves
// may still add to them: the 64 super-buckets (64 buckets each) locate it,
// then the buckets of that super-bucket.  A cell is added to its bucket
// before its super-bucket, and every value read is at most the count (or
// size sum) of all the cells of its bucket, so the bucket found is never
// below the full set's crossing bucket.  kBuckets when the totals read never
// reach the budget.
__device__ int wave_find_bucket(const uint32_t *hcnt, const uint32_t *hmass,
                                const uint32_t *scnt, const uint32_t *smass, int w, int T) {
    const int lane = threadIdx.x & 63;
    const uint32_t c = scnt[lane], m = smass[lane];
    uint32_t ic = c, im = m;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(ic, o), u = __shfl_up(im, o);
        if (lane >= o) { ic += t; im += u; }
    }
    const unsigned long long hit = __ballot((int64_t)ic >= w || (int64_t)im >= T);
    if (hit == 0) return kBuckets;
    const int l0 = __ffsll((long long)hit) - 1;
    const uint32_t pc = __shfl(ic - c, l0), pm = __shfl(im - m, l0);
    uint32_t fc = hcnt[l0 * 64 + lane], fm = hmass[l0 * 64 + lane];
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(fc, o), u = __shfl_up(fm, o);
        if (lane >= o) { fc += t; fm += u; }
    }
    const unsigned long long hit2 =
        __ballot((int64_t)(pc + fc) >= w || (int64_t)(pm + fm) >= T);
    if (hit2 == 0) return kBuckets;
    return l0 * 64 + __ffsll((long long)hit2) - 1;
}

// LDS of k_rank_big: the histograms [2][kBuckets], v [max(2 kc, dim)], the
// centroid order [kc] (u16).
__host__ __device__ inline size_t big_lds(int kc, int dim) {
    const size_t nv = (size_t)(2 * kc > dim ? 2 * kc : dim);
    return 8 * (size_t)kBuckets + 4 * nv + ((2 * (size_t)kc + 15) & ~(size_t)15);
}

// LDS of the emission: fine-bucket starts, sizes and fill counters.
constexpr size_t big_emit_lds() { return 4 * (3 * (size_t)kFine + 1); }

// The emission of the query of workgroup slot `slot` from its pruned-ranking
// scratch: the entries up to the final cut counting-sorted by fine buckets,
// ordered within their bucket, and the taken prefix written as its cell
// list.  smem holds big_emit_lds() bytes.  All threads call.
__device__ void big_emit(const BigArgs &b, int64_t slot, unsigned char *smem) {
    uint32_t *fstart = reinterpret_cast<uint32_t *>(smem);     // [kFine + 1]
    uint32_t *fmass = fstart + kFine + 1;                      // [kFine]
    uint32_t *ffill = fmass + kFine;                           // [kFine]
    const RankKernelArgs &k = b.k;
    const RankCellsArgs &a = k.a;
    __shared__ RankShared sh;
    __shared__ uint32_t s_lo, s_hi, s_sum;
    __shared__ int s_count;
    const int tid = threadIdx.x;
    const int64_t q = k.qlist ? k.qlist[slot] : slot;
    const int64_t si = b.by_slot ? slot : q;        // the query's scratch
    const int T = query_budget(a.Tq, q, a.T), w = a.w;
    if (empty_query(k, q, T)) return;
    const uint32_t *aux = b.aux + (size_t)si * 4;
    if (aux[3]) {                      // still over capacity: the device-wide path
        flag_overflow(k, q);
        return;
    }
    const int n = (int)aux[0];
    const int bf = (int)aux[1];
    const float base = __uint_as_float(aux[2]);
    const uint64_t *list = b.list + (size_t)si * b.cap;
    uint64_t *S = b.sorted + (size_t)si * b.cap;
    uint32_t *SZ = b.size + (size_t)si * b.cap;
    // the key range of the entries up to the cut
    uint32_t lo = ~0u, hi = 0u;
    for (int i = tid; i < n; i += kBigThreads) {
        const uint32_t key = (uint32_t)(list[i] >> 32);
        if (dbucket(sc_key_float(key), base) <= bf) { lo = min(lo, key); hi = max(hi, key); }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor(lo, o));
        hi = max(hi, (uint32_t)__shfl_xor(hi, o));
    }
    if (tid == 0) { s_lo = ~0u; s_hi = 0u; s_count = 0; s_sum = 0; }
    for (int i = tid; i < kFine; i += kBigThreads) { fstart[i] = 0; fmass[i] = 0; ffill[i] = 0; }
    __syncthreads();
    if ((tid & 63) == 0) { atomicMin(&s_lo, lo); atomicMax(&s_hi, hi); }
    __syncthreads();
    lo = s_lo;
    hi = s_hi;
    int shift = 0;
    while (lo <= hi && ((hi - lo) >> shift) >= (uint32_t)kFine) ++shift;
    const int64_t k2 = (int64_t)k.kc * k.kc;
    // fine histogram (counts, sizes) of the kept entries
    for (int i = tid; i < n; i += kBigThreads) {
        const uint64_t e = list[i];
        const uint32_t key = (uint32_t)(e >> 32);
        if (dbucket(sc_key_float(key), base) > bf) continue;
        const int64_t c = (int64_t)(uint32_t)e;
        const int fb = (int)((key - lo) >> shift);
        atomicAdd(&fstart[fb], 1u);
        atomicAdd(&fmass[fb], big_size(b, c, b.brec[c]));
    }
    __syncthreads();
    {
        constexpr int per = kFine / kBigThreads;
        uint32_t lc = 0, lm = 0;
        for (int i = 0; i < per; ++i) { lc += fstart[tid * per + i]; lm += fmass[tid * per + i]; }
        uint32_t tc, tm, rc, rm;
        block_excl_scan2<kBigThreads>(lc, lm, sh.wsum, &rc, &rm, &tc, &tm);
        for (int i = 0; i < per; ++i) {
            const uint32_t c = fstart[tid * per + i], m = fmass[tid * per + i];
            fstart[tid * per + i] = rc;
            fmass[tid * per + i] = rm;
            rc += c;
            rm += m;
        }
        if (tid == kBigThreads - 1) fstart[kFine] = rc;
    }
    __syncthreads();
    // scatter into fine-bucket order
    for (int i = tid; i < n; i += kBigThreads) {
        const uint64_t e = list[i];
        const uint32_t key = (uint32_t)(e >> 32);
        if (dbucket(sc_key_float(key), base) > bf) continue;
        const int64_t c = (int64_t)(uint32_t)e;
        const int fb = (int)((key - lo) >> shift);
        const uint32_t slot = fstart[fb] + atomicAdd(&ffill[fb], 1u);
        S[slot] = e;
        SZ[slot] = big_size(b, c, b.brec[c]);
    }
    __syncthreads();
    const int nk = (int)fstart[kFine];
    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    for (int i = tid; i < nk; i += kBigThreads) {
        const uint64_t e = S[i];
        const int fb = (int)(((uint32_t)(e >> 32) - lo) >> shift);
        const int j0 = (int)fstart[fb], j1 = j0 + (int)ffill[fb];
        int r = 0;
        uint32_t p = 0;
        for (int j = j0; j < j1; ++j) {
            const bool before = S[j] < e;
            r += before;
            p += before ? SZ[j] : 0u;
        }
        const int rank = j0 + r;
        const uint32_t pos = fmass[fb] + p, sz = SZ[i];
        if (rank < w && (int64_t)pos < T) {
            if (rank < a.ccap) {
                const int64_t c = (int64_t)(uint32_t)e;
                const uint64_t key = b.ckeys[c];
                const int64_t u0 = (int64_t)(key / (uint64_t)k2), u1 = (int64_t)(key % (uint64_t)k2);
                cells[rank] = c > 0 ? k.L[c - 1] : 0;
                clen[rank] = (int32_t)sz;
                cinfo[rank] = (uint32_t)(u0 / k.kc) | ((uint32_t)(u1 / k.kc) << 16);
            }
            atomicMax(&s_count, rank + 1);
            atomicMax(&s_sum, pos + sz);
        }
    }
    __syncthreads();
    if (tid == 0) {
        const int count = s_count;
        int32_t *qi = a.qinfo + (size_t)q * 8;
        qi[0] = count; qi[1] = (int32_t)s_sum; qi[2] = count; qi[3] = 0;
        qi[4] = (count == k.ncells && count < w && (int64_t)s_sum < T) ? 1 : 0;
        qi[5] = 0; qi[6] = count > a.ccap ? 2 : 0; qi[7] = 0;
    }
}

__global__ __launch_bounds__(kBigThreads) void k_rank_big_emit(BigArgs b) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (b.qcount && (int)blockIdx.x >= *b.qcount) return;   // the whole workgroup
    big_emit(b, blockIdx.x, smem);
}


__global__ __launch_bounds__(kBigThreads, 8) void k_rank_big(BigArgs b) {
    extern __shared__ __align__(16) unsigned char smem[];
    const RankKernelArgs &k = b.k;
    const int kc = k.kc;
    uint32_t *hcnt = reinterpret_cast<uint32_t *>(smem);          // [kBuckets]
    uint32_t *hmass = hcnt + kBuckets;                            // [kBuckets]
    float *v = reinterpret_cast<float *>(hmass + kBuckets);       // [max(2 kc, dim)]
    uint16_t *ord = reinterpret_cast<uint16_t *>(v + max(2 * kc, k.dim));   // [kc]
    __shared__ RankShared sh;
    __shared__ float s_min[2 * kBigThreads / 64];
    __shared__ int s_pos, s_bcur, s_napp;
    __shared__ uint32_t s_cnt, s_mass;
    __shared__ uint32_t scnt[64], smass[64];      // super-buckets: 64 buckets each
    __shared__ unsigned long long s_scored;
    if (b.qcount && (int)blockIdx.x >= *b.qcount) return;   // the whole workgroup
    const int tid = threadIdx.x, lane = tid & 63;
    const int P = b.mode == 2 ? b.parts : 1, part = (int)(blockIdx.x % (unsigned)P);
    const int64_t slot = blockIdx.x / P;
    const int64_t q = k.qlist ? k.qlist[slot] : slot;
    const int64_t si = b.by_slot ? slot : q;        // the query's scratch
    const int T = query_budget(k.a.Tq, q, k.a.T), w = k.a.w;
    uint32_t *aux = b.aux + (size_t)si * 4;
    uint32_t *sp = b.mode != 0 ? b.split + (size_t)q * kSplitWords : nullptr;
    const int start_cut = b.rerun ? (int)aux[1] : b.mode == 2 ? (int)sp[0] : kBuckets;
    __shared__ int s_nord, s_last;
    query_terms<kBigThreads>(k, q, v, sh);
    if (T <= 0 || k.ncells == 0) {             // the emission writes the empty result
        if (b.emit && b.mode == 0) big_emit(b, slot, smem);
        return;
    }
    float m0a, m0b, m1a, m1b;
    two_smallest<kBigThreads>(v, kc, s_min, &m0a, &m0b);
    two_smallest<kBigThreads>(v + kc, kc, s_min, &m1a, &m1b);
    const float LB1 = m1a + m1b, base = (m0a + m0b) + LB1;

    // the nearest-0 centroids in order of their bound's bucket
    for (int i = tid; i < kBuckets; i += kBigThreads) hcnt[i] = 0;
    __syncthreads();
    for (int n = tid; n < kc; n += kBigThreads)
        if (n % P == part) atomicAdd(&hcnt[dbucket((v[n] + m0a) + LB1, base)], 1u);
    __syncthreads();
    {
        constexpr int per = kBuckets / kBigThreads;
        uint32_t loc = 0;
        for (int i = 0; i < per; ++i) loc += hcnt[tid * per + i];
        uint32_t tot;
        uint32_t run = block_excl_scan<kBigThreads>(loc, sh.wsum, &tot);
        for (int i = 0; i < per; ++i) {
            const uint32_t c = hcnt[tid * per + i];
            hcnt[tid * per + i] = run;
            run += c;
        }
        if (tid == 0) s_nord = (int)tot;
    }
    __syncthreads();
    const int nord = s_nord;          // this workgroup's centroids
    for (int n = tid; n < kc; n += kBigThreads)
        if (n % P == part) ord[atomicAdd(&hcnt[dbucket((v[n] + m0a) + LB1, base)], 1u)] = (uint16_t)n;
    __syncthreads();
    for (int i = tid; i < kBuckets; i += kBigThreads) { hcnt[i] = 0; hmass[i] = 0; }
    if (tid < 64) { scnt[tid] = 0; smass[tid] = 0; }
    if (tid == 0) {
        s_pos = 0; s_bcur = start_cut; s_napp = 0; s_cnt = 0; s_mass = 0; s_scored = 0;
    }
    __syncthreads();

    uint64_t *list = b.list + (size_t)si * b.cap;
    const float *v1 = v + kc;
    volatile int *bcut = &s_bcur;
    uint32_t since_cnt = 0, since_mass = 0;       // this wave's additions since its last cut
    // Appends the admitted lanes' entries (one LDS atomic per wave).
    auto append = [&](bool adm, uint32_t key, int c) {
        const unsigned long long am = __ballot(adm);
        if (am) {
            int slot0 = 0;
            if (lane == 0)
                slot0 = b.mode == 2 ? (int)atomicAdd(sp + 1, (uint32_t)__popcll(am))
                                    : atomicAdd(&s_napp, (int)__popcll(am));
            slot0 = __shfl(slot0, 0);
            if (adm) {
                const int slot = slot0 + (int)__popcll(am & ((1ull << lane) - 1ull));
                if (slot < b.cap) list[slot] = ((uint64_t)key << 32) | (uint32_t)c;
            }
        }
    };
    constexpr int kU = 4;                          // cell loads in flight per lane
    constexpr int kWarm = 8;                       // centroids of the warm-up phase
    // Phase 0 (warm-up, only without a given cut): the best kWarm centroids
    // counted into the histograms without appending, for a first cut; the
    // histograms are then cleared and phase 1 starts over from that cut with
    // appends, so no cell is counted twice.
    const int phase0 = start_cut < kBuckets || b.mode == 2 ? 1 : 0;
    for (int phase = phase0; phase < (b.mode == 1 ? 1 : 2); ++phase) {
        const bool app = phase == 1;
        const int lim = app ? nord : min(kWarm, nord);
        for (;;) {
            int i = 0;
            if (lane == 0) i = atomicAdd(&s_pos, 1);
            i = __shfl(i, 0);
            if (i >= lim) break;
            const int n0 = ord[i];
            const float a0 = v[n0];
            // one value of the cut per wave (another wave or workgroup may
            // lower it meanwhile)
            if (b.mode == 2 && lane == 0) {
                const int gc = (int)__hip_atomic_load(sp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (gc < *bcut) atomicMin(&s_bcur, gc);
            }
            int bcur = __shfl(*bcut, 0);
            if (dbucket((a0 + m0a) + LB1, base) > bcur) break;   // so are the centroids after it
            const int g0 = b.bg_n0[n0], g1 = b.bg_n0[n0 + 1];
            for (int gb = g0; gb < g1; gb += 64) {
                // this round's pairs (one per lane) that pass the bound
                const int g = gb + lane;
                float p0 = 0.0f;
                int c0 = 0, c1 = 0;
                bool ok = false;
                if (g < g1) {
                    p0 = a0 + v[b.bg_pair[g] >> 16];
                    ok = dbucket(p0 + LB1, base) <= bcur;
                    if (ok) { c0 = b.bg_start[g]; c1 = b.bg_start[g + 1]; }
                }
                unsigned long long gm = __ballot(ok);
                // the pairs one after another, all lanes on a pair's cells; the
                // first kU * 64 cells of the next pair are loaded while the
                // current pair is scored
                uint32_t rn[kU];
                int jn = gm ? __ffsll((long long)gm) - 1 : -1;
                if (jn >= 0) {
                    gm &= gm - 1;
                    const int cb = __shfl(c0, jn), ce = __shfl(c1, jn);
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int c = cb + u * 64 + lane;
                        rn[u] = c < ce ? b.brec[c] : 0u;
                    }
                }
                while (jn >= 0) {
                    const int j = jn;
                    const float P0 = __shfl(p0, j);
                    const int cb = __shfl(c0, j), ce = __shfl(c1, j);
                    uint32_t r[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) r[u] = rn[u];
                    jn = gm ? __ffsll((long long)gm) - 1 : -1;
                    if (jn >= 0) {
                        gm &= gm - 1;
                        const int nb = __shfl(c0, jn), ne = __shfl(c1, jn);
#pragma unroll
                        for (int u = 0; u < kU; ++u) {
                            const int c = nb + u * 64 + lane;
                            rn[u] = c < ne ? b.brec[c] : 0u;
                        }
                    }
                    if (lane == 0) atomicAdd(&s_scored, (unsigned long long)(ce - cb));
                    uint32_t add_cnt = 0, add_mass = 0;         // this lane's additions
                    for (int cs = cb; cs < ce; cs += 64 * kU) {
                        // a pair of more than kU*64 cells: its next kU*64 are
                        // loaded while these are scored
                        const bool more = cs + 64 * kU < ce;    // uniform
                        uint32_t rx[kU];
                        if (more) {
#pragma unroll
                            for (int u = 0; u < kU; ++u) {
                                const int c = cs + (kU + u) * 64 + lane;
                                rx[u] = c < ce ? b.brec[c] : 0u;
                            }
                        }
#pragma unroll
                        for (int u = 0; u < kU; ++u) {
                            if (cs + u * 64 >= ce) break;       // uniform
                            const int c = cs + u * 64 + lane;
                            bool adm = false;
                            uint32_t key = 0;
                            if (c < ce) {
                                const float sc =
                                    P0 + (v1[r[u] & 0xFFFu] + v1[(r[u] >> 12) & 0xFFFu]);
                                const int bk = dbucket(sc, base);
                                if (bk <= bcur) {
                                    adm = true;
                                    key = sc_float_key(sc);
                                    const uint32_t sz = big_size(b, c, r[u]);
                                    atomicAdd(&hcnt[bk], 1u);
                                    atomicAdd(&hmass[bk], sz);
                                    atomicAdd(&scnt[bk >> 6], 1u);
                                    atomicAdd(&smass[bk >> 6], sz);
                                    add_cnt += 1;
                                    add_mass += sz;
                                }
                            }
                            if (app) append(adm, key, c);
                        }
                        if (more) {
#pragma unroll
                            for (int u = 0; u < kU; ++u) r[u] = rx[u];
                        }
                    }
                    // this pair's additions, then a tighter cut when enough was added
                    if (!__ballot(add_cnt > 0)) continue;
                    for (int o = 32; o > 0; o >>= 1) {
                        add_cnt += __shfl_xor(add_cnt, o);
                        add_mass += __shfl_xor(add_mass, o);
                    }
                    uint32_t tc = 0, tm = 0;
                    if (lane == 0) {
                        tc = atomicAdd(&s_cnt, add_cnt) + add_cnt;
                        tm = atomicAdd(&s_mass, add_mass) + add_mass;
                    }
                    tc = __shfl(tc, 0);
                    tm = __shfl(tm, 0);
                    since_cnt += add_cnt;
                    since_mass += add_mass;
                    const bool reach = (int64_t)tc >= w || (int64_t)tm >= T;
                    const bool enough = bcur == kBuckets || (int64_t)since_mass * 32 >= T ||
                                        (int64_t)since_cnt * 32 >= w;
                    if (reach && enough) {
                        const int f = wave_find_bucket(hcnt, hmass, scnt, smass, w, T);
                        if (lane == 0 && f < kBuckets) {
                            atomicMin(&s_bcur, f);
                            if (b.mode == 2) atomicMin(reinterpret_cast<int *>(sp), f);
                        }
                        since_cnt = 0;
                        since_mass = 0;
                    }
                    bcur = __shfl(*bcut, 0);
                }
            }
        }
        if (b.mode == 1) {
            // warm-up only: its exact crossing bucket (a subset's, so a bound)
            // is the starting cut of the query's workgroups
            __syncthreads();
            const int wf = find_bucket<kBigThreads, kBuckets>(hcnt, hmass, 0u, 0u, w, T, sh.wsum, &sh.b,
                                                    sh.pre);
            if (tid == 0) {
                atomicMin(reinterpret_cast<int *>(sp), min(wf, s_bcur));
                if (k.a.stats) atomicAdd(k.a.stats + 2, s_scored);
            }
            return;
        }
        if (phase == 0) {
            // the warm-up's cut stays (its cells are a subset of all cells);
            // its counts are cleared before the cells are visited again
            __syncthreads();
            for (int i = tid; i < kBuckets; i += kBigThreads) { hcnt[i] = 0; hmass[i] = 0; }
            if (tid < 64) { scnt[tid] = 0; smass[tid] = 0; }
            if (tid == 0) { s_pos = 0; s_cnt = 0; s_mass = 0; }
            since_cnt = 0;
            since_mass = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    if (tid == 0 && k.a.stats) atomicAdd(k.a.stats + 2, s_scored);
    if (b.mode == 2) {
        // the workgroups' histograms (disjoint cells) summed into the
        // query's; the last workgroup to finish takes the crossing bucket
        uint32_t *gc = sp + 4, *gm = sp + 4 + kBuckets;
        for (int i = tid; i < kBuckets; i += kBigThreads) {
            if (hcnt[i]) atomicAdd(gc + i, hcnt[i]);
            if (hmass[i]) atomicAdd(gm + i, hmass[i]);
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) s_last = atomicAdd(sp + 2, 1u) == (uint32_t)(P - 1);
        __syncthreads();
        if (!s_last) return;
        __threadfence();
        for (int i = tid; i < kBuckets; i += kBigThreads) {
            hcnt[i] = __hip_atomic_load(gc + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            hmass[i] = __hip_atomic_load(gm + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (tid == 0) s_napp = (int)__hip_atomic_load(sp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
    // the exact crossing bucket of the final histograms
    const int bf = find_bucket<kBigThreads, kBuckets>(hcnt, hmass, 0u, 0u, w, T, sh.wsum, &sh.b, sh.pre);
    if (tid == 0) {
        aux[0] = (uint32_t)s_napp;
        aux[1] = (uint32_t)bf;
        aux[2] = __float_as_uint(base);
        aux[3] = s_napp > b.cap ? 1u : 0u;
    }
    if (b.emit && b.mode == 0) {
        __syncthreads();   // this workgroup's list and aux writes, visible to all its threads
        big_emit(b, slot, smem);
    }
}

// Resets the split state of the queries: cut = none, counters and histograms 0.
__global__ void k_split_reset(uint32_t *__restrict__ split, int64_t words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= words) return;
    split[i] = (i % kSplitWords) == 0 ? (uint32_t)kBuckets : 0u;
}

__global__ void k_big_flagged(const uint32_t *__restrict__ aux, int64_t nq,
                              int32_t *__restrict__ list, int32_t *__restrict__ n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq && aux[q * 4 + 3]) list[atomicAdd(n, 1)] = (int32_t)q;
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
                               int32_t *__restrict__ list, int32_t *__restrict__ n,
                               int32_t *__restrict__ skip) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int32_t f = qinfo[q * 8 + 5];
    if (skip) skip[q] = f ? 1 : 0;
    if (f) list[atomicAdd(n, 1)] = (int32_t)q;
}

__global__ void k_segment_offsets(int64_t nseg, int64_t len, int64_t *off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nseg) off[i] = i * len;
}

// SC_RANK_CAP (test hook): a smaller on-chip list, so that tests reach the
// level-2 refinement and the device-wide fallback.
int rank_cap(const sc_index *ix, const RankCellsArgs &a) {
    if (const char *e = getenv("SC_RANK_CAP")) {
        const int v = atoi(e);
        if (v >= 1) return std::min(v, kCap);
    }
    // twice the list when the budget takes, on average, more than a third
    // of it (T over the vectors per non-empty cell): large budgets would
    // otherwise send many queries to the re-ranking
    const double per = ix->ncells > 0 ? std::max(1.0, (double)ix->n / (double)ix->ncells) : 1.0;
    return 3.0 * (double)a.T / per > (double)kCap ? 2 * kCap : kCap;
}

// SC_RANK_PAIRS=0 (test hook): rank from the cell records only.
bool use_pair_tables(const sc_index *ix) {
    if (const char *e = getenv("SC_RANK_PAIRS"))
        if (atoi(e) == 0) return false;
    return ix->cpair && ix->gstart && ix->gvec && ix->nu[0] <= kMaxPairs0 &&
           ix->nu[0] + ix->nu[1] <= 8192;
}

// Launches the ranking of a.nq queries (or of the queries in qlist) with an
// on-chip list of cap cells.
template <int W>
int launch_rank(sc_ctx *ctx, const sc_index *ix, RankKernelArgs k, const int32_t *qlist,
                int64_t nq, int cap) {
    k.qlist = qlist;
    k.cap = cap;
    const bool pairs = use_pair_tables(ix);
    // v [2 kc] also holds the query row while its sums are taken
    const size_t lds = rank_abytes(cap) + 4 * (size_t)rank_vwords(ix->kc, ix->dim) +
                       (pairs ? 4 * ((size_t)ix->nu[0] + ix->nu[1]) : 0);
    if (lds > 150 * 1024) return SC_ERR_TOO_LARGE;
    const void *kern = pairs ? (const void *)k_rank_pairs : (const void *)k_rank_cells<W>;
    SC_TRY(sc_set_max_lds(kern, lds));
    if (pairs)
        hipLaunchKernelGGL(k_rank_pairs, dim3((unsigned)nq), dim3(kRankThreads), lds, ctx->stream,
                           k);
    else
        hipLaunchKernelGGL(k_rank_cells<W>, dim3((unsigned)nq), dim3(kRankThreads), lds,
                           ctx->stream, k);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

// The queries flagged as overflowing (qinfo[5]) into list; returns their count.
int flagged(sc_ctx *ctx, const RankCellsArgs &a, int32_t *list, int32_t *dn, int32_t *nflag) {
    SC_HIP(hipMemsetAsync(dn, 0, 4, ctx->stream));
    hipLaunchKernelGGL(k_rank_flagged, dim3((unsigned)sc_cdiv(a.nq, 256)), dim3(256), 0,
                       ctx->stream, a.qinfo, a.nq, list, dn, nullptr);
    SC_CHECK_LAUNCH();
    SC_HIP(hipMemcpyAsync(nflag, dn, 4, hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

// The pruned ranking of all a.nq queries: k_rank_big, a second run of the
// queries over capacity, and k_rank_big_emit.  Queries still over capacity
// are flagged (qinfo[5]) for the device-wide path.
int rank_big(sc_ctx *ctx, const sc_index *ix, const RankKernelArgs &k, const RankCellsArgs &a) {
    BigArgs b{k,           ix->bg_n0,    ix->bg_pair, ix->bg_start, ix->brec, ix->ckeys,
              a.big_list,  a.big_sorted, a.big_size,  a.big_aux,    a.big_cap, 0,
              1,           nullptr,      0,           0};
    b.k.qlist = nullptr;
    // batches of a few queries: each query over several workgroups (about
    // two per CU).  The shares start from a warm-up's cut but tighten only
    // from their own cells, so they pay off only where one workgroup per
    // query would leave most of the chip idle (measured on the K=4096 index
    // over 100M vectors: 1 query 1.15 -> 1.07 ms; 64 queries 1.89 -> 2.34 ms).
    if (a.big_split && a.nq <= 4) {
        b.parts = (int)std::min<int64_t>(64, std::max<int64_t>(1, 2 * ctx->n_cu / a.nq));
        if (const char *e = getenv("SC_RANK_PARTS")) {   // test hook
            const int v = atoi(e);
            if (v >= 1) b.parts = std::min(v, 64);
        }
        b.split = a.big_split;
        if (b.parts > 1) {
            const int64_t words = a.nq * (int64_t)kSplitWords;
            hipLaunchKernelGGL(k_split_reset, dim3((unsigned)sc_cdiv(words, 256)), dim3(256), 0,
                               ctx->stream, b.split, words);
            SC_CHECK_LAUNCH();
        }
    }
    const size_t lds0 = big_lds(ix->kc, ix->dim);
    if (lds0 > 150 * 1024) {
        sc_set_error("sc_search: kc=%d too large for the pruned cell ranking", ix->kc);
        return SC_ERR_ARG;
    }
    // one workgroup per query (no split): it emits its query's list itself
    const size_t lds = b.parts > 1 ? lds0 : std::max(lds0, big_emit_lds());
    b.emit = b.parts > 1 ? 0 : 1;
    SC_TRY(sc_set_max_lds((const void *)k_rank_big, std::max(lds0, big_emit_lds())));
    if (b.parts > 1) {
        // the warm-up on one workgroup per query, then the query's shares
        BigArgs wu = b;
        wu.mode = 1;
        hipLaunchKernelGGL(k_rank_big, dim3((unsigned)a.nq), dim3(kBigThreads), lds, ctx->stream,
                           wu);
        SC_CHECK_LAUNCH();
        b.mode = 2;
    }
    hipLaunchKernelGGL(k_rank_big, dim3((unsigned)(a.nq * (b.mode == 2 ? b.parts : 1))),
                       dim3(kBigThreads), lds,
                       ctx->stream, b);
    SC_CHECK_LAUNCH();
    int32_t *list = a.flag_list, *dn = a.flag_count;
    int32_t nflag = 0;
    SC_HIP(hipMemsetAsync(dn, 0, 4, ctx->stream));
    hipLaunchKernelGGL(k_big_flagged, dim3((unsigned)sc_cdiv(a.nq, 256)), dim3(256), 0,
                       ctx->stream, a.big_aux, a.nq, list, dn);
    SC_CHECK_LAUNCH();
    SC_HIP(hipMemcpyAsync(&nflag, dn, 4, hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    if (a.overflowed) *a.overflowed += nflag;
    if (nflag > 0) {
        BigArgs r = b;
        r.k.qlist = list;
        r.rerun = 1;
        r.parts = 1;              // from the exact cut: one workgroup per query
        r.split = nullptr;
        r.mode = 0;
        r.emit = 1;
        hipLaunchKernelGGL(k_rank_big, dim3((unsigned)nflag), dim3(kBigThreads),
                           std::max(lds0, big_emit_lds()), ctx->stream, r);
        SC_CHECK_LAUNCH();
    }
    if (!b.emit) {   // the split queries' lists, from their last workgroups' histograms
        hipLaunchKernelGGL(k_rank_big_emit, dim3((unsigned)a.nq), dim3(kBigThreads),
                           big_emit_lds(), ctx->stream, b);
        SC_CHECK_LAUNCH();
    }
    return SC_OK;
}

// SC_RANK_BIG (test hook): 1 ranks every index that has the group directory
// with the pruned ranking, 0 never uses it.
int big_mode() {
    if (const char *e = getenv("SC_RANK_BIG")) return atoi(e) ? 1 : 0;
    return -1;
}

template <int W>
RankKernelArgs rank_args(const sc_index *ix, const RankCellsArgs &a) {
    return RankKernelArgs{a, ix->kc, ix->dim, ix->ncells, ix->crec, ix->L, nullptr, rank_cap(ix, a),
                          ix->nu[0], ix->nu[1], ix->upairs[0], ix->upairs[1], ix->cpair,
                          ix->gstart, ix->gvec};
}

// The pruned re-ranking is available for queries whose cells outgrow the
// on-chip list (scratch for a.retry_cap queries at a time).
bool retry_ok(const sc_index *ix, const RankCellsArgs &a) {
    return !(a.big_cap > 0 && a.big_list) && a.retry_cap > 0 && a.big_list && ix->brec &&
           ix->bg_n0 && big_lds(ix->kc, ix->dim) <= 150 * 1024;
}

BigArgs retry_args(const sc_index *ix, const RankKernelArgs &k, const RankCellsArgs &a) {
    return BigArgs{k,         ix->bg_n0,    ix->bg_pair, ix->bg_start, ix->brec, ix->ckeys,
                   a.big_list, a.big_sorted, a.big_size,  a.big_aux,    kBigCap,  0,
                   1,          nullptr,      0,           1};
}

// The pruned re-ranking of the n queries of qlist, a.retry_cap at a time; a
// query still over the list's capacity is flagged for the device-wide sort
// by the emit kernel.
int retry_groups(sc_ctx *ctx, const sc_index *ix, const RankKernelArgs &k,
                 const RankCellsArgs &a, const int32_t *qlist, int32_t n) {
    const size_t lds = std::max(big_lds(ix->kc, ix->dim), big_emit_lds());
    SC_TRY(sc_set_max_lds((const void *)k_rank_big, lds));
    for (int32_t g0 = 0; g0 < n; g0 += a.retry_cap) {
        const int32_t gn = std::min<int32_t>(a.retry_cap, n - g0);
        BigArgs b = retry_args(ix, k, a);
        b.k.qlist = qlist + g0;
        b.emit = 1;
        hipLaunchKernelGGL(k_rank_big, dim3((unsigned)gn), dim3(kBigThreads), lds, ctx->stream,
                           b);
        SC_CHECK_LAUNCH();
    }
    return SC_OK;
}

// The device-wide ranking of the nflag queries of list: every cell's key,
// a segmented sort per query and the walk along it.
template <int W>
int device_wide(sc_ctx *ctx, const sc_index *ix, const RankKernelArgs &k, const int32_t *list,
                int32_t nflag) {
    int rc = SC_OK;
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
        sc_set_error("sc_search: ranking fallback (%zu bytes) could not be allocated", total);
        return SC_ERR_NOMEM;
    }
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
        hipLaunchKernelGGL(k_segment_offsets, dim3((unsigned)sc_cdiv(gn + 1, 256)), dim3(256), 0,
                           ctx->stream, gn, ix->ncells, off);
        if (hipGetLastError() != hipSuccess) { rc = SC_ERR_HIP; break; }
        size_t tb2 = tb;
        if (rocprim::segmented_radix_sort_keys(tmp, tb2, keys, skeys, (unsigned)(gn * ix->ncells),
                                               (unsigned)gn, off, off + 1, 0u, 64u,
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
    return rc;
}

template <int W>
int rank_cells(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a) {
    const RankKernelArgs k = rank_args<W>(ix, a);
    const int cap = rank_cap(ix, a);
    int32_t *list = a.flag_list, *dn = a.flag_count;
    int32_t nflag = 0;
    int rc = SC_OK;
    if (a.big_cap > 0 && a.big_list) {
        SC_TRY(rank_big(ctx, ix, k, a));
        SC_TRY(flagged(ctx, a, list, dn, &nflag));
    } else {
        rc = launch_rank<W>(ctx, ix, k, nullptr, a.nq, cap);
        if (rc == SC_ERR_TOO_LARGE) {
            sc_set_error("sc_search: kc=%d too large for the on-chip cell ranking", ix->kc);
            return SC_ERR_ARG;
        }
        if (rc) return rc;
        if (retry_ok(ix, a) && a.defer_list && a.defer_count && a.deferred) {
            // no host synchronisation: the flagged queries listed on the
            // device, the first retry_cap of them re-ranked by workgroups that
            // read the count (the others exit at once), those still flagged
            // listed for the caller
            SC_HIP(hipMemsetAsync(dn, 0, 4, ctx->stream));
            hipLaunchKernelGGL(k_rank_flagged, dim3((unsigned)sc_cdiv(a.nq, 256)), dim3(256), 0,
                               ctx->stream, a.qinfo, a.nq, list, dn, a.skip);
            SC_CHECK_LAUNCH();
            // with skip flags, the re-ranking runs on the second stream
            hipStream_t rs = ctx->stream;
            if (a.skip) {
                SC_TRY(sc_ctx_side(ctx));
                SC_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
                SC_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
                rs = ctx->side;
            }
            const size_t lds = std::max(big_lds(ix->kc, ix->dim), big_emit_lds());
            SC_TRY(sc_set_max_lds((const void *)k_rank_big, lds));
            BigArgs b = retry_args(ix, k, a);
            b.k.qlist = list;
            b.qcount = dn;
            b.emit = 1;
            const unsigned grid = (unsigned)std::min<int64_t>(a.nq, a.retry_cap);
            hipLaunchKernelGGL(k_rank_big, dim3(grid), dim3(kBigThreads), lds, rs, b);
            SC_CHECK_LAUNCH();
            SC_HIP(hipMemsetAsync(a.defer_count, 0, 4, rs));
            hipLaunchKernelGGL(k_rank_flagged, dim3((unsigned)sc_cdiv(a.nq, 256)), dim3(256), 0,
                               rs, a.qinfo, a.nq, a.defer_list, a.defer_count, nullptr);
            SC_CHECK_LAUNCH();
            *a.deferred = 1;
            return SC_OK;
        }
        SC_TRY(flagged(ctx, a, list, dn, &nflag));
        if (a.overflowed) *a.overflowed += nflag;
    }
    // queries whose cells did not fit the on-chip list: the pruned ranking
    // over them where the index has its directory; else once more with a
    // list four times as long (fewer workgroups per CU, only for them), when
    // the LDS allows
    if (nflag > 0 && retry_ok(ix, a)) {
        SC_TRY(retry_groups(ctx, ix, k, a, list, nflag));
        SC_TRY(flagged(ctx, a, list, dn, &nflag));
    } else if (nflag > 0 && cap == kCap && !(a.big_cap > 0 && a.big_list)) {
        rc = launch_rank<W>(ctx, ix, k, list, nflag, 4 * kCap);
        if (rc == SC_OK) SC_TRY(flagged(ctx, a, list, dn, &nflag));
        else if (rc != SC_ERR_TOO_LARGE) return rc;
        rc = SC_OK;
    }
    if (nflag > 0) rc = device_wide<W>(ctx, ix, k, list, nflag);
    return rc;
}

template <int W>
int rank_finish(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a, const int32_t *qlist,
                int32_t n) {
    const RankKernelArgs k = rank_args<W>(ix, a);
    int32_t nflag = 0;
    if (n > 0 && retry_ok(ix, a)) SC_TRY(retry_groups(ctx, ix, k, a, qlist, n));
    SC_TRY(flagged(ctx, a, a.flag_list, a.flag_count, &nflag));
    return nflag > 0 ? device_wide<W>(ctx, ix, k, a.flag_list, nflag) : SC_OK;
}

}  // namespace

int sc_rank_big_cap(const sc_index *ix) {
    if (!ix->sparse || !ix->brec || !ix->bg_n0 || ix->kc > 4096 || ix->ncells <= 0) return 0;
    const int mode = big_mode();
    if (mode == 0) return 0;
    if (mode < 0 && use_pair_tables(ix)) return 0;
    if (big_lds(ix->kc, ix->dim) > 150 * 1024) return 0;
    if (const char *e = getenv("SC_RANK_BIG_CAP")) {   // test hook: a smaller list
        const int v = atoi(e);
        if (v >= 1) return std::min(v, kBigCap);
    }
    return kBigCap;
}

int sc_rank_retry_cap(const sc_index *ix) {
    if (!ix->sparse || !ix->brec || !ix->bg_n0 || ix->kc > 4096 || ix->ncells <= 0) return 0;
    if (big_mode() == 0 || big_lds(ix->kc, ix->dim) > 150 * 1024) return 0;
    return kBigCap;
}

int sc_rank_cells(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a) {
    if (a.nq <= 0) return SC_OK;
    if (!ix->sparse) {
        sc_set_error("sc_search: cell ranking needs the two-halves index");
        return SC_ERR_CONFIG;
    }
    return ix->crec_words == 2 ? rank_cells<2>(ctx, ix, a) : rank_cells<4>(ctx, ix, a);
}

int sc_rank_finish(sc_ctx *ctx, const sc_index *ix, const RankCellsArgs &a, const int32_t *list,
                   int32_t n) {
    if (a.nq <= 0) return SC_OK;
    return ix->crec_words == 2 ? rank_finish<2>(ctx, ix, a, list, n)
                               : rank_finish<4>(ctx, ix, a, list, n);
}
