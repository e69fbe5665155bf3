// sc_search.hip -- batched search: per-query LUTs, coarse ranking, the
// multi-sequence traversal (exact replay of the reference's array heap), the
// ADC code scan and the canonical top-R.
//
// Stages (one launch each, all on the context stream):
//   k_lut       PQQuery::pre_compute2 (include/query.h:154-184), index-aligned
//               stores, sequential FP32 dot products (the defined cblas_sdot).
//   k_rank      coarse scores (query.h:225-236, multi_query.h:102-122,
//               sc_query.h:95-107) and a full sort by (score, centroid).
//   k_traverse  one lane per query: search_ivfadc step 2 (query.h:232-259),
//               search_multi2 step 2 (multi_query.h:133-205), search_mr_ivf
//               step 2 (sc_query.h:111-186) with pq_insert*/pq_pop*
//               (include/sc_utilities.h:229-300, 469-585) replayed exactly.
//   k_scan_topk one workgroup per query: the ADC sequences of query.h:295-305,
//               multi_query.h:231-253, sc_query.h:216-229 and the top-R by
//               (dist ASC, position ASC).
#include <float.h>
#include <stdlib.h>

#include <algorithm>
#include <new>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "sc_internal.h"

namespace {

constexpr int kScanThreads = 256;
constexpr int kMaxFusedThreads = 512;   // fused workgroups run 256 or 512 threads
constexpr int kBufCap = 1536;      // scan: streaming top-R buffer (entries)
constexpr int kScanRound = 512;    // scan: candidates a workgroup takes per round
constexpr int kScanSlots = kScanRound / kScanThreads;   // ... per thread
constexpr int kRound = 1024;       // fused search: candidates a workgroup takes per round
constexpr int kRunCells = 256;     // cells staged per scan run (one per scan thread)
static_assert(kRunCells == kScanThreads, "a scan run loads one cell per thread");
constexpr int kSpec = 16;          // traversal pops run ahead of the L[] loads
constexpr int kWin = 16;           // SC: cell sizes of the kWin x kWin nearest rank pairs

// ----------------------------------------------------------------------------
// Stage 1: LUTs.  out[q][0 .. mc*kc) = diff_qc, out[q][mc*kc .. +mp*kp) = diff_qr
// diff = (norm - s) - s,  s = sequential FP32 dot product.
// ----------------------------------------------------------------------------
__global__ void k_lut(const float *__restrict__ Q, int64_t nq, int dim,
                      const float *__restrict__ cq, const float *__restrict__ pq,
                      const float *__restrict__ norm_c, const float *__restrict__ norm_r, int mc,
                      int kc, int mp, int kp, float *__restrict__ dqc, float *__restrict__ dqr) {
    const int per = mc * kc + mp * kp;
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nq * per) return;
    int64_t q = e / per;
    int j = (int)(e - q * per);
    const float *qv = Q + (size_t)q * dim;
    // cq/pq are the transposed codebooks [blocks][blen][K]: neighbouring
    // threads read neighbouring codewords; each dot still runs t = 0, 1, ...
    if (j < mc * kc) {
        const int bsc = dim / mc, blk = j / kc, k = j - blk * kc;
        const float *a = qv + (size_t)blk * bsc, *b = cq + (size_t)blk * bsc * kc + k;
        float s = 0.0f;
        for (int t = 0; t < bsc; ++t) s += a[t] * b[(size_t)t * kc];
        dqc[(size_t)q * mc * kc + j] = (norm_c[j] - s) - s;
    } else {
        j -= mc * kc;
        const int bsp = dim / mp, sub = j / kp, c = j - sub * kp;
        const float *a = qv + (size_t)sub * bsp, *b = pq + (size_t)sub * bsp * kp + c;
        float s = 0.0f;
        for (int t = 0; t < bsp; ++t) s += a[t] * b[(size_t)t * kp];
        dqr[(size_t)q * mp * kp + j] = (norm_r[j] - s) - s;
    }
}

// Tiled form of k_lut: a workgroup computes a 64-query x 64-codeword tile of
// one block's dot products; each thread owns 4 x 4 of them and runs every dot
// in index order with a separately rounded multiply and add, exactly like
// k_lut.  Queries are staged transposed in LDS ([t][64]), codewords come from
// the transposed codebook ([t][K], contiguous in the codeword).
// grid.x: query tiles, grid.y: output tiles of 64 over mc*kc + mp*kp.
constexpr int kLutQ = 64, kLutC = 64;
__global__ __launch_bounds__(256) void k_lut_tiled(
    const float *__restrict__ Q, int64_t nq, int dim, const float *__restrict__ cqT,
    const float *__restrict__ pqT, const float *__restrict__ norm_c,
    const float *__restrict__ norm_r, int mc, int kc, int mp, int kp, float *__restrict__ dqc,
    float *__restrict__ dqr) {
    __shared__ float qs[32][kLutQ];      // 32 dims at a time
    __shared__ float cs[32][kLutC];
    const int tid = threadIdx.x, tq = tid % 16, tc = tid / 16;
    const int64_t q0 = (int64_t)blockIdx.x * kLutQ;
    const int o0 = blockIdx.y * kLutC;            // first output of the tile
    const int nco = mc * kc;
    // outputs of one tile never straddle a block: kc and kp are multiples of
    // 64 on this path (checked by the caller)
    const bool coarse = o0 < nco;
    const int blk = coarse ? o0 / kc : (o0 - nco) / kp;
    const int first = coarse ? o0 - blk * kc : (o0 - nco) - blk * kp;
    const int K = coarse ? kc : kp;
    const int blen = coarse ? dim / mc : dim / mp;
    const float *cb = (coarse ? cqT : pqT) + (size_t)blk * blen * K + first;
    float acc[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    for (int t0 = 0; t0 < blen; t0 += 32) {
        const int tt = min(32, blen - t0);
        __syncthreads();
        for (int e = tid; e < 32 * kLutQ; e += 256) {
            const int t = e / kLutQ, qq = e % kLutQ;
            qs[t][qq] = (t < tt && q0 + qq < nq) ? Q[(size_t)(q0 + qq) * dim + blk * blen + t0 + t] : 0.0f;
            cs[t][qq] = t < tt ? cb[(size_t)(t0 + t) * K + qq] : 0.0f;
        }
        __syncthreads();
        for (int t = 0; t < tt; ++t) {
            const float4 a = *reinterpret_cast<const float4 *>(&qs[t][tq * 4]);
            const float4 b = *reinterpret_cast<const float4 *>(&cs[t][tc * 4]);
            const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
        }
    }
    for (int i = 0; i < 4; ++i) {
        const int64_t q = q0 + tq * 4 + i;
        if (q >= nq) continue;
        for (int j = 0; j < 4; ++j) {
            const int o = first + tc * 4 + j;          // codeword within the block
            const float s = acc[i][j];
            if (coarse) {
                const int idx = blk * kc + o;
                dqc[(size_t)q * nco + idx] = (norm_c[idx] - s) - s;
            } else {
                const int idx = blk * kp + o;
                dqr[(size_t)q * mp * kp + idx] = (norm_r[idx] - s) - s;
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Stage 2: coarse scores + sort.  One workgroup per query.
// qs[q] = {q_sum, q_half0, q_half1, 0}.  sv/si [q][mc*kc]: sorted scores and
// centroid ids per list (IVF: unsorted scores in centroid order).
// ----------------------------------------------------------------------------
__global__ void k_rank(const float *__restrict__ Q, int dim, int kind, int mc, int kc,
                       const float *__restrict__ dqc, float *__restrict__ qs,
                       float *__restrict__ sv, int32_t *__restrict__ si, int npow2) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint64_t *key = reinterpret_cast<uint64_t *>(smem);
    __shared__ float s_q[3];
    const int64_t q = blockIdx.x;
    const float *qv = Q + (size_t)q * dim;
    if (threadIdx.x == 0) {
        float qsum = 0.0f, h[2] = {0.0f, 0.0f};
        if (kind == SC_KIND_IMI) {
            const int bsc = dim / mc;
            for (int i = 0; i < mc; ++i) {
                float d1 = 0.0f;
                for (int j = 0; j < bsc; ++j) { float t = qv[i * bsc + j]; d1 += t * t; }
                qsum += d1;
                h[i] = d1;
            }
        } else {
            for (int i = 0; i < dim; ++i) { float t = qv[i]; qsum += t * t; }
        }
        s_q[0] = qsum; s_q[1] = h[0]; s_q[2] = h[1];
        qs[q * 4 + 0] = qsum; qs[q * 4 + 1] = h[0]; qs[q * 4 + 2] = h[1]; qs[q * 4 + 3] = 0.0f;
    }
    __syncthreads();
    const float *dq = dqc + (size_t)q * mc * kc;
    if (kind == SC_KIND_IVFADC) {
        for (int j = threadIdx.x; j < kc; j += blockDim.x) {
            sv[(size_t)q * kc + j] = s_q[0] + dq[j];
            si[(size_t)q * kc + j] = j;
        }
        return;
    }
    const int lists = (kind == SC_KIND_IMI) ? 2 : 1;
    for (int h = 0; h < lists; ++h) {
        const float base = (kind == SC_KIND_IMI) ? s_q[1 + h] : s_q[0];
        for (int j = threadIdx.x; j < npow2; j += blockDim.x) {
            if (j < kc) {
                float v = base + dq[h * kc + j];
                key[j] = ((uint64_t)sc_float_key(v) << 32) | (uint32_t)j;
            } else {
                key[j] = ~(uint64_t)0;
            }
        }
        __syncthreads();
        for (int size = 2; size <= npow2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = threadIdx.x; i < npow2 / 2; i += blockDim.x) {
                    int lo = 2 * i - (i & (stride - 1));
                    int hi = lo + stride;
                    bool up = ((lo & size) == 0);
                    uint64_t a = key[lo], b = key[hi];
                    if ((a > b) == up) { key[lo] = b; key[hi] = a; }
                }
                __syncthreads();
            }
        }
        for (int j = threadIdx.x; j < kc; j += blockDim.x) {
            uint64_t k = key[j];
            sv[((size_t)q * lists + h) * kc + j] = sc_key_float((uint32_t)(k >> 32));
            si[((size_t)q * lists + h) * kc + j] = (int32_t)(k & 0xFFFFFFFFu);
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// Stage 3: traversal.  One lane per query; heap and row lengths in LDS
// (interleaved across lanes), or in global memory for the fallback launch.
// The popped cells always form a down-set of the rank grid (a cell is pushed
// only after both its predecessors were popped), so traversed(a, b) of the
// reference is exactly b < rowlen[a].
// qinfo[q*8 + ...]: 0 ncell, 1 sum, 2 visited, 3 empty, 4 exhausted,
//                   5 overflow (on-chip heap too small), 6 error
// ----------------------------------------------------------------------------
struct Heap {
    float *v;
    uint32_t *p;
    int stride;   // element stride
    int pos;
    int cap_ref;  // the reference's capacity (w, or kc for IVF)
    int cap_phys; // storage available
    bool overflow;

    __device__ __forceinline__ float &V(int i) { return v[(size_t)i * stride]; }
    __device__ __forceinline__ uint32_t &P(int i) { return p[(size_t)i * stride]; }

    // Sift-up and sift-down move the entry being placed through a hole kept
    // in registers: one round trip to LDS per level (parent, or both
    // children with their ids) instead of a read-compare-swap chain.  The
    // comparisons and tie rules are exactly those of the array heap of
    // include/sc_utilities.h:469-585, so the pop order is the same.
    __device__ __forceinline__ void insert(float val, uint32_t id) {
        if (pos >= cap_ref || pos < 0) return;           // silently dropped
        if (pos >= cap_phys) { overflow = true; return; }
        int i = pos++;
        while (i > 0) {
            const int j = i >> 1;
            const float vj = V(j);
            const uint32_t pj = P(j);
            if (val < vj) { V(i) = vj; P(i) = pj; i = j; } else break;
        }
        V(i) = val;
        P(i) = id;
    }
    __device__ __forceinline__ uint32_t pop() {
        const uint32_t top = P(0);
        const int last = --pos;
        const float mv = V(last);
        const uint32_t mpid = P(last);
        int i = 0, j = 0;
        while (j + 1 < pos) {      // a lone last child is not compared
            // at the first level the left "child" is the hole itself
            const float a = j == i ? mv : V(j), b = V(j + 1);
            const uint32_t pa = P(j), pb = P(j + 1);
            const bool left = a < b;
            if ((left ? a : b) < mv) {
                V(i) = left ? a : b;
                P(i) = left ? pa : pb;
                i = left ? j : j + 1;
            } else break;
            j = i << 1;
        }
        V(i) = mv;
        P(i) = mpid;
        return top;
    }
};

struct TravArgs {
    int kind, kc, w, T, ccap;
    const float *sv;
    const int32_t *si;
    const int32_t *L;
    int32_t *cells;
    uint32_t *cinfo;
    int32_t *qinfo;
    const int32_t *qlist;   // fallback: queries to redo (nullptr = all)
    int64_t nq;
    float *gheap_v;         // fallback heap storage [nq][w]
    uint32_t *gheap_p;
    int32_t *grow;          // fallback row lengths [nq][kc]
    int32_t *clen;          // [nq][ccap] cell lengths (cells holds cell starts)
    const int32_t *Tq;      // [nq] per-query budgets (at most T) or null
};

// The candidate budget of query q: its own when per-query budgets are given
// (clamped to [0, T]), else T.
__device__ __forceinline__ int query_budget(const int32_t *Tq, int64_t q, int T) {
    return Tq ? min(max(Tq[q], 0), T) : T;
}

__device__ __forceinline__ int32_t cell_len(const int32_t *L, int64_t bid) {
    return bid > 0 ? L[bid] - L[bid - 1] : L[0];
}
__device__ __forceinline__ int32_t cell_start(const int32_t *L, int64_t bid) {
    return bid > 0 ? L[bid - 1] : 0;
}

__global__ void k_traverse(TravArgs a, int hcap, int qpw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int64_t slot = (int64_t)blockIdx.x * qpw + lane;
    if (lane >= qpw || slot >= a.nq) return;
    const int64_t q = a.qlist ? a.qlist[slot] : slot;
    const int kc = a.kc;
    const int T = query_budget(a.Tq, q, a.T);

    Heap h;
    int32_t *rowlen;
    int rstride;
    if (a.gheap_v) {            // fallback: hcap = the reference's capacity
        h.v = a.gheap_v + (size_t)slot * hcap;
        h.p = a.gheap_p + (size_t)slot * hcap;
        h.stride = 1;
        h.cap_phys = hcap;
        rowlen = a.grow + (size_t)slot * kc;
        rstride = 1;
    } else {
        float *hv = reinterpret_cast<float *>(smem);
        uint32_t *hp = reinterpret_cast<uint32_t *>(hv + (size_t)hcap * qpw);
        int32_t *rl = reinterpret_cast<int32_t *>(hp + (size_t)hcap * qpw);
        h.v = hv + lane;
        h.p = hp + lane;
        h.stride = qpw;
        h.cap_phys = hcap;
        rowlen = rl + lane;
        rstride = qpw;
    }
    h.pos = 0;
    h.overflow = false;

    int32_t *cells = a.cells + (size_t)q * a.ccap;      // cell starts
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    int count = 0, count2 = 0, sum = 0, empty = 0, exhausted = 0, err = 0;

    if (a.kind == SC_KIND_IVFADC) {
        const float *v = a.sv + (size_t)q * kc;
        h.cap_ref = kc;
        for (int j = 0; j < kc; ++j) h.insert(v[j], (uint32_t)j);
        for (int i = 0; i < a.w && !h.overflow; ++i) {
            if (h.pos <= 0) { err = 1; break; }      // pq_pop on an empty heap exits
            int bid = (int)h.pop();
            int l = cell_len(a.L, bid);
            cells[count] = cell_start(a.L, bid);
            clen[count] = l;
            cinfo[count] = (uint32_t)bid;
            count++;
            count2++;
            sum += l;
            if (sum >= T) break;
        }
    } else {
        const bool imi = a.kind == SC_KIND_IMI;
        const float *v1 = a.sv + (size_t)q * (imi ? 2 : 1) * kc;
        const float *v2 = imi ? v1 + kc : v1;
        const int32_t *t1 = a.si + (size_t)q * (imi ? 2 : 1) * kc;
        const int32_t *t2 = imi ? t1 + kc : t1;
        h.cap_ref = a.w;
        for (int r = 0; r < kc; ++r) rowlen[(size_t)r * rstride] = 0;
        if (imi) {
            h.insert(v1[0] + v2[0], 0u);
        } else {
            rowlen[0] = 1;                           // (0,0) never enters the SC grid
            float d0 = v1[0] + v1[1];
            h.insert(d0, (0u << 16) | 1u);
            h.insert(d0, (1u << 16) | 0u);
        }
        while ((imi ? count2 : count) < a.w && sum < T) {
            if (h.overflow) break;
            if (h.pos <= 0) { exhausted = 1; break; }
            uint32_t pk = h.pop();
            int h1 = (int)(pk >> 16), h2 = (int)(pk & 0xFFFFu);
            int h3 = t1[h1], h4 = t2[h2];
            int64_t bid = (int64_t)h3 * kc + h4;
            int l = imi ? cell_len(a.L, bid) : (bid > 0 ? cell_len(a.L, bid) : 0);
            if (l > 0) {
                if (count >= a.ccap) { err = 2; break; }
                cells[count] = cell_start(a.L, bid);
                clen[count] = l;
                cinfo[count] = (uint32_t)h3 | ((uint32_t)h4 << 16);
                count++;
                sum += l;
            } else {
                empty++;
            }
            int32_t &rl = rowlen[(size_t)h1 * rstride];
            if (rl != h2) err = 3;                   // not a down-set: cannot happen
            rl = h2 + 1;
            count2++;
            if (sum < T) {
                if (h1 < kc - 1 && (h2 == 0 || (h2 - 1) < rowlen[(size_t)(h1 + 1) * rstride]))
                    h.insert(v1[h1 + 1] + v2[h2], ((uint32_t)(h1 + 1) << 16) | (uint32_t)h2);
                if (h2 < kc - 1 && (h1 == 0 || (h2 + 1) < rowlen[(size_t)(h1 - 1) * rstride]))
                    h.insert(v1[h1] + v2[h2 + 1], ((uint32_t)h1 << 16) | (uint32_t)(h2 + 1));
            }
        }
    }
    int32_t *qi = a.qinfo + (size_t)q * 8;
    qi[0] = count; qi[1] = sum; qi[2] = count2; qi[3] = empty; qi[4] = exhausted;
    qi[5] = h.overflow ? 1 : 0; qi[6] = err; qi[7] = 0;
}

// Multi-sequence traversal of IMI / SC (include/multi_query.h:133-205,
// include/sc_query.h:111-186) with one 64-thread workgroup per query: the
// query's ranked lists, row lengths and heap sit in LDS; thread 0 replays the
// heap and runs up to kSpec pops ahead (the pops and pushes do not depend on
// the cell sizes until the loop stops); the batch's cell sizes are loaded by
// kSpec lanes at once, and thread 0 commits the pops in order, stopping
// exactly where the reference does.  A heap outgrowing hcap is flagged
// (qinfo[5]) for the global-heap lane kernel, as in k_traverse, and so is a
// query whose traversal reaches rank kTravRanks of a list: only the first
// kTravRanks ranks of each list are staged.
// LDS: per list kr scores (float) and centroids (u16), kr row lengths (u16),
// hcap (score, pair) entries -- small, so that many queries are in flight.
constexpr int kTravThreads = 64;
constexpr int kTravRanks = 256;

struct TravWaveLayout {
    size_t st, rowlen, hv, hp, total;
};
__host__ __device__ inline TravWaveLayout trav_wave_layout(int lists, int kc, int hcap) {
    TravWaveLayout L;
    L.st = (size_t)lists * kc * 4;
    L.rowlen = L.st + (size_t)lists * kc * 2;
    L.hv = (L.rowlen + (size_t)kc * 2 + 7) & ~(size_t)7;
    L.hp = L.hv + (size_t)hcap * 4;
    L.total = L.hp + (size_t)hcap * 4;
    return L;
}

__global__ __launch_bounds__(kTravThreads) void k_traverse_wave(TravArgs a, int hcap) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int kc = a.kc, tid = threadIdx.x;
    const int kr = min(kc, kTravRanks);              // ranks staged per list
    const bool imi = a.kind == SC_KIND_IMI;
    const int lists = imi ? 2 : 1;
    const TravWaveLayout lay = trav_wave_layout(lists, kr, hcap);
    float *sv = reinterpret_cast<float *>(smem);                       // [lists][kr]
    uint16_t *st = reinterpret_cast<uint16_t *>(smem + lay.st);       // [lists][kr]
    uint16_t *rowlen = reinterpret_cast<uint16_t *>(smem + lay.rowlen);   // [kr]
    float *hv = reinterpret_cast<float *>(smem + lay.hv);             // [hcap]
    uint32_t *hp = reinterpret_cast<uint32_t *>(smem + lay.hp);       // [hcap]
    __shared__ int64_t s_bid[kSpec];
    __shared__ uint32_t s_info[kSpec];
    __shared__ int32_t s_len[kSpec];
    __shared__ int s_n, s_done;

    const int64_t slot = blockIdx.x;
    if (slot >= a.nq) return;
    const int64_t q = a.qlist ? a.qlist[slot] : slot;
    const int T = query_budget(a.Tq, q, a.T);
    for (int i = tid; i < lists * kr; i += kTravThreads) {
        const int l = i / kr, r = i - l * kr;
        sv[i] = a.sv[((size_t)q * lists + l) * kc + r];
        st[i] = (uint16_t)a.si[((size_t)q * lists + l) * kc + r];
    }
    for (int i = tid; i < kr; i += kTravThreads) rowlen[i] = 0;
    if (tid == 0) s_done = 0;
    __syncthreads();
    const float *v1 = sv, *v2 = imi ? sv + kr : sv;
    const uint16_t *t1 = st, *t2 = imi ? st + kr : st;

    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    // thread 0's state
    Heap h;
    h.v = hv;
    h.p = hp;
    h.stride = 1;
    h.pos = 0;
    h.cap_ref = a.w;
    h.cap_phys = hcap;
    h.overflow = false;
    int count = 0, count2 = 0, sum = 0, empty = 0, exhausted = 0, err = 0, flag = 0;
    bool dry = false;                                // heap empty after the batch (thread 0)
    if (tid == 0) {
        if (imi) {
            h.insert(v1[0] + v2[0], 0u);
        } else {
            rowlen[0] = 1;                           // (0,0) never enters the SC grid
            const float d0 = v1[0] + v1[1];
            h.insert(d0, (0u << 16) | 1u);
            h.insert(d0, (1u << 16) | 0u);
        }
        if (!((imi ? count2 : count) < a.w && sum < T)) s_done = 1;
    }
    __syncthreads();
    while (!s_done) {
        if (tid == 0) {
            // run ahead: up to kSpec pops with their pushes
            int n = 0;
            dry = false;
            while (n < kSpec && !h.overflow) {
                if (h.pos <= 0) { dry = true; break; }
                const uint32_t pk = h.pop();
                const int h1 = (int)(pk >> 16), h2 = (int)(pk & 0xFFFFu);
                const int c1 = t1[h1], c2 = t2[h2];
                const int64_t bid = (int64_t)c1 * kc + c2;
                s_bid[n] = imi ? bid : (bid > 0 ? bid : -1);   // SC: cell 0 counts as empty
                s_info[n] = (uint32_t)c1 | ((uint32_t)c2 << 16);
                ++n;
                if ((int)rowlen[h1] != h2) err = 3;      // not a down-set: cannot happen
                rowlen[h1] = (uint16_t)(h2 + 1);
                // a push past the staged ranks cannot be replayed here: it is
                // treated like a heap overflow (the pops after it are invalid)
                if (h1 < kc - 1) {
                    if (h1 + 1 >= kr) h.overflow = true;
                    else if (h2 == 0 || (h2 - 1) < rowlen[h1 + 1])
                        h.insert(v1[h1 + 1] + v2[h2], ((uint32_t)(h1 + 1) << 16) | (uint32_t)h2);
                }
                if (h2 < kc - 1 && (h1 == 0 || (h2 + 1) < rowlen[h1 - 1])) {
                    if (h2 + 1 >= kr) h.overflow = true;
                    else h.insert(v1[h1] + v2[h2 + 1], ((uint32_t)h1 << 16) | (uint32_t)(h2 + 1));
                }
            }
            s_n = n;
        }
        __syncthreads();
        const int n = s_n;
        if (tid < n) {
            const int64_t b = s_bid[tid];
            s_len[tid] = b >= 0 ? cell_len(a.L, b) : 0;
            s_bid[tid] = b >= 0 ? cell_start(a.L, b) : 0;
        }
        __syncthreads();
        if (tid == 0) {
            // commit in order; the loop test precedes every pop
            int i = 0;
            bool stop = false;
            for (; i < n; ++i) {
                if (!((imi ? count2 : count) < a.w && sum < T)) { stop = true; break; }
                const int l = s_len[i];
                if (l > 0) {
                    if (count >= a.ccap) { err = 2; stop = true; break; }
                    cells[count] = (int32_t)s_bid[i];
                    clen[count] = l;
                    cinfo[count] = s_info[i];
                    count++;
                    sum += l;
                } else {
                    empty++;
                }
                count2++;
            }
            if (!stop && !((imi ? count2 : count) < a.w && sum < T)) stop = true;
            if (!stop) {
                if (dry) { exhausted = 1; stop = true; }             // the heap is empty
                else if (h.overflow) { flag = 1; stop = true; }      // later pops are invalid
            }
            s_done = stop ? 1 : 0;
        }
        __syncthreads();
    }
    if (tid == 0) {
        int32_t *qi = a.qinfo + (size_t)q * 8;
        qi[0] = count; qi[1] = sum; qi[2] = count2; qi[3] = empty; qi[4] = exhausted;
        qi[5] = flag; qi[6] = err; qi[7] = 0;
    }
}

// Three-candidate traversal (SCQuery::search_mr_ivf3, include/sc_query.h:296-402):
// rank triples (h1, h2, h3) of one coarse list, cell (t[h1], t[h2], t[h3]),
// heap of the reference's capacity w in global memory, one lane per query.
// A triple enters the heap once its other predecessors have been popped, so
// the popped set is a down-set of the rank grid and the reference's kc^3
// `traversed` flags reduce to a depth per (h1, h2) column: (h1, h2, h3) has
// been popped iff h3 < depth[h1][h2].
struct Trav3Args {
    int kc, w, T, ccap;
    const float *sv;
    const int32_t *si;
    const int32_t *L;
    int32_t *cells;         // [nq][ccap] cell starts
    int32_t *clen;          // [nq][ccap] cell lengths
    uint32_t *cinfo;        // [nq][ccap] nearest centroid of the cell
    int32_t *qinfo;         // [nq][8]
    int64_t q0, nq;         // the queries [q0, q0 + nq) of this launch
    float *hv;              // [nq][w] heap scores
    uint32_t *hp;           // [nq][w] heap triples (10 bits per rank)
    uint16_t *depth;        // [nq][kc][kc], zero on entry
    const int32_t *Tq;      // per-query budgets (indexed like the queries) or null
};

__global__ void k_traverse3(Trav3Args a) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.nq) return;
    const int64_t q = a.q0 + slot;
    const int kc = a.kc;
    const int T = query_budget(a.Tq, q, a.T);
    const float *v = a.sv + (size_t)q * kc;
    const int32_t *t = a.si + (size_t)q * kc;
    uint16_t *D = a.depth + (size_t)slot * kc * kc;
    auto popped = [&](int x, int y, int z) { return z < (int)D[(size_t)x * kc + y]; };
    auto pack = [](int x, int y, int z) {
        return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
    };
    Heap h;
    h.v = a.hv + (size_t)slot * a.w;
    h.p = a.hp + (size_t)slot * a.w;
    h.stride = 1;
    h.pos = 0;
    h.cap_ref = a.w;
    h.cap_phys = a.w;
    h.overflow = false;
    int32_t *cells = a.cells + (size_t)q * a.ccap;
    int32_t *clen = a.clen + (size_t)q * a.ccap;
    uint32_t *cinfo = a.cinfo + (size_t)q * a.ccap;
    const int64_t grid = (int64_t)kc * kc * kc;
    int count = 0, sum = 0, empty = 0, exhausted = 0, err = 0;
    int64_t count2 = 0;
    h.insert(3.0f * v[0], pack(0, 0, 0));
    while (count < a.w && sum < T && count2 < grid) {
        if (h.pos <= 0) { exhausted = 1; break; }
        const uint32_t pk = h.pop();
        const int h1 = (int)(pk & 1023u), h2 = (int)((pk >> 10) & 1023u), h3 = (int)(pk >> 20);
        const int c1 = t[h1];
        const int64_t bid = ((int64_t)c1 * kc + t[h2]) * kc + t[h3];
        const int l = bid > 0 ? cell_len(a.L, bid) : 0;
        if (l > 0) {
            if (count >= a.ccap) { err = 2; break; }
            cells[count] = cell_start(a.L, bid);
            clen[count] = l;
            cinfo[count] = (uint32_t)c1;
            count++;
            sum += l;
        } else {
            empty++;
        }
        uint16_t &dd = D[(size_t)h1 * kc + h2];
        if ((int)dd != h3) err = 3;                  // not a down-set: cannot happen
        dd = (uint16_t)(h3 + 1);
        count2++;
        if (sum < T) {
            if (h1 < kc - 1 && (h2 == 0 || popped(h1 + 1, h2 - 1, h3)) &&
                (h3 == 0 || popped(h1 + 1, h2, h3 - 1)))
                h.insert((v[h1 + 1] + v[h2]) + v[h3], pack(h1 + 1, h2, h3));
            if (h2 < kc - 1 && (h3 == 0 || popped(h1, h2 + 1, h3 - 1)) &&
                (h1 == 0 || popped(h1 - 1, h2 + 1, h3)))
                h.insert((v[h1] + v[h2 + 1]) + v[h3], pack(h1, h2 + 1, h3));
            if (h3 < kc - 1 && (h2 == 0 || popped(h1, h2 - 1, h3 + 1)) &&
                (h1 == 0 || popped(h1 - 1, h2, h3 + 1)))
                h.insert((v[h1] + v[h2]) + v[h3 + 1], pack(h1, h2, h3 + 1));
        }
    }
    int32_t *qi = a.qinfo + (size_t)q * 8;
    qi[0] = count; qi[1] = sum; qi[2] = (int32_t)count2; qi[3] = empty; qi[4] = exhausted;
    qi[5] = 0; qi[6] = err; qi[7] = 0;
}

// ----------------------------------------------------------------------------
// Stage 4: scan + top-R.  One workgroup per query.
// Candidates are keyed (orderable dist bits << 32 | position); keys are unique
// so the R smallest keys are exactly the canonical (dist, position) top-R.
// A streaming buffer keeps every candidate below the running threshold and is
// compacted to the R best by radix select whenever it nears capacity.
// ----------------------------------------------------------------------------
struct ScanArgs {
    int kind, dim, mc, kc, mp, kp, R, ccap;
    const float *qs, *dqc, *dqr, *dot_cr;
    const int32_t *L, *pid;
    const uint8_t *codes;
    const int32_t *cells;
    const uint32_t *cinfo;
    const int32_t *qinfo;
    int32_t *ids;
    float *dists;
    int32_t *sums;
    const int32_t *qlist;   // rescan: workgroup b handles query qlist[b]
    const int32_t *clen;    // [nq][ccap] cell lengths (cells holds cell starts)
    const float *Q;         // SC_FORM_REAL: the queries [nq][dim]
    const float *raw;       // SC_FORM_REAL: raw base vectors [n][dim] by id, else null
    int imi_mode;           // IMI scan: 0 byte-wise codes, 1 packed codes (mp = 8 or 16)
    // mixed form of the IMI scan (SC_FORM_MIXED): per-vector dot_cr sums, the
    // index's max |dot_cr|, and qinfo (slot 7 flags a query for an exact redo)
    const float *evec;
    float dot_absmax;
    int32_t *redo;
    float margin_scale;     // test hook: widens the keep margin (1 = the bound)
    const uint64_t *ckeys;  // two-halves directory keys (null: dense cells)
    int64_t ncells;
    const int32_t *qcount = nullptr;  // with qlist: workgroups at or past *qcount exit
    const uint32_t *vcell = nullptr;  // [n] n0 | n1 << 16 per position (mc = 2) or null
    const int32_t *skip = nullptr;    // [nq] or null: queries flagged 1 are left to a later scan
    const uint4 *mrec = nullptr;      // mixed form, mp = 8: {codes, e_v, vcell} per position
};

struct ScanShared {
    uint64_t key[kBufCap];
    int32_t idx[kBufCap];        // position in the pid array
    int32_t c_start[kRunCells];  // cell start in pid/codes
    int32_t c_pref[kRunCells + 1];
    uint32_t c_info[kRunCells];
    uint32_t scan_wsum[kScanThreads / 64];
    float c_d0[kRunCells];       // IMI runs: each cell's start term
    int overflow;                // mixed: the kept set outgrew the buffer
    float lscale, twoE;          // mixed: FP16 table scale, keep margin

    uint32_t hist[256];
    int nb;
    uint64_t thr;
    uint64_t sel_prefix;
    int sel_rem;
    int run_cells;
    uint32_t b_lo, b_hi;         // select_bound: the refined range of score keys
    int b_below, b_cnt;          // entries below it, and in it
};

// Radix select: the R-th smallest key among key[0..nb).  All threads call.
__device__ uint64_t select_kth(ScanShared &S, int nb, int R) {
    if (threadIdx.x == 0) { S.sel_prefix = 0; S.sel_rem = R; }
    __syncthreads();
    uint64_t mask = 0;
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) S.hist[b] = 0;
        __syncthreads();
        const uint64_t prefix = S.sel_prefix;
        const uint32_t rem = (uint32_t)S.sel_rem;   // read by all before the update below
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            uint64_t k = S.key[i];
            if ((k & mask) == prefix) atomicAdd(&S.hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            // wave 0: inclusive scan of 256 bins (4 per lane)
            const int l = threadIdx.x;
            uint32_t c0 = S.hist[4 * l], c1 = S.hist[4 * l + 1], c2 = S.hist[4 * l + 2],
                     c3 = S.hist[4 * l + 3];
            uint32_t tot = c0 + c1 + c2 + c3, incl = tot;
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t t = __shfl_up(incl, o);
                if (l >= o) incl += t;
            }
            uint32_t excl = incl - tot;
            if (excl < rem && rem <= incl) {
                uint32_t c[4] = {c0, c1, c2, c3};
                uint32_t acc = excl;
                for (int b = 0; b < 4; ++b) {
                    if (rem <= acc + c[b]) {
                        S.sel_prefix = prefix | ((uint64_t)(4 * l + b) << shift);
                        S.sel_rem = (int)(rem - acc);
                        break;
                    }
                    acc += c[b];
                }
            }
        }
        __syncthreads();
        mask |= (uint64_t)255 << shift;
    }
    return S.sel_prefix;
}

// Keeps only the R smallest keys (all <= kth) in key/idx[0..R).
__device__ void compact(ScanShared &S, int R) {
    const int nb = S.nb;
    const uint64_t kth = select_kth(S, nb, R);
    // gather survivors to the front: entries <= kth; stable order not needed
    __shared__ int s_w;
    if (threadIdx.x == 0) s_w = 0;
    __syncthreads();
    uint64_t mk[kBufCap / kScanThreads];
    int32_t mi[kBufCap / kScanThreads];
    int m = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (S.key[i] <= kth) { mk[m] = S.key[i]; mi[m] = S.idx[i]; ++m; }
    __syncthreads();
    for (int j = 0; j < m; ++j) {
        int o = atomicAdd(&s_w, 1);
        S.key[o] = mk[j];
        S.idx[o] = mi[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) { S.nb = R; S.thr = kth; }
    __syncthreads();
}

// An upper bound on the R-th smallest score (the high key word) among
// key[0..nb) (R <= nb), at most one histogram bucket above it: the score
// keys' range [min, max] is cut into 256 equal buckets, the bucket holding
// the R-th smallest is found and, while it holds more than 32 entries, cut
// again.  Unlike a radix select on fixed bits, the buckets follow the data,
// so the histogram's LDS atomics spread over many addresses.  *n_le gets
// the number of entries whose score key is at most the bound.  All threads
// call.
__device__ uint32_t select_bound(ScanShared &S, int nb, int R, int *n_le) {
    const int tid = threadIdx.x;
    if (tid == 0) { S.b_lo = ~0u; S.b_hi = 0u; S.b_below = 0; }
    __syncthreads();
    uint32_t mn = ~0u, mx = 0u;
    for (int i = tid; i < nb; i += blockDim.x) {
        const uint32_t k = (uint32_t)(S.key[i] >> 32);
        mn = min(mn, k);
        mx = max(mx, k);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, o));
        mx = max(mx, (uint32_t)__shfl_xor(mx, o));
    }
    if ((tid & 63) == 0) { atomicMin(&S.b_lo, mn); atomicMax(&S.b_hi, mx); }
    __syncthreads();
    for (int pass = 0; pass < 4; ++pass) {
        const uint32_t lo = S.b_lo, hi = S.b_hi, width = hi - lo;
        const int below = S.b_below;   // read by all before the update below
        const int shift = width < 256u ? 0 : 32 - __clz(width) - 8;   // (width >> shift) < 256
        for (int b = tid; b < 256; b += blockDim.x) S.hist[b] = 0;
        __syncthreads();
        for (int i = tid; i < nb; i += blockDim.x) {
            const uint32_t k = (uint32_t)(S.key[i] >> 32);
            if (k >= lo && k <= hi) atomicAdd(&S.hist[(k - lo) >> shift], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            // wave 0: inclusive scan of the 256 bins (4 per lane), then the
            // bin where the count reaches R
            const int l = tid;
            const uint32_t c0 = S.hist[4 * l], c1 = S.hist[4 * l + 1], c2 = S.hist[4 * l + 2],
                           c3 = S.hist[4 * l + 3];
            const uint32_t tot = c0 + c1 + c2 + c3;
            uint32_t incl = tot;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o);
                if (l >= o) incl += t;
            }
            const uint32_t rem = (uint32_t)(R - below), excl = incl - tot;
            if (excl < rem && rem <= incl) {
                const uint32_t c[4] = {c0, c1, c2, c3};
                uint32_t acc = excl;
                for (int b = 0; b < 4; ++b) {
                    if (rem <= acc + c[b]) {
                        const uint32_t blo = lo + ((uint32_t)(4 * l + b) << shift);
                        const uint32_t bw = shift == 0 ? 0u : (1u << shift) - 1u;
                        S.b_lo = blo;
                        S.b_hi = hi - blo < bw ? hi : blo + bw;
                        S.b_below = below + (int)acc;
                        S.b_cnt = (int)c[b];
                        break;
                    }
                    acc += c[b];
                }
            }
        }
        __syncthreads();
        if (S.b_cnt <= 32 || S.b_lo == S.b_hi) break;
    }
    *n_le = S.b_below + S.b_cnt;
    const uint32_t ub = S.b_hi;
    __syncthreads();
    return ub;
}

__device__ uint32_t run_excl_scan(uint32_t x, uint32_t *wsum, uint32_t *total);

// Keeps the entries whose score key is at most `keep` (a score key, high
// word) at the front of key/idx, when there are at most min(limit,
// kBufCap / 2) of them; returns false (nothing moved) otherwise.  One
// prefix over the workgroup of each thread's kept count (its entries are
// positions tid + j * kScanThreads) gives the slots.  The entries of the
// first half are in registers before its barrier; the second half's are
// read after it, and no slot (all below kBufCap / 2) overwrites them.  All
// threads call.
__device__ bool keep_at_most(ScanShared &S, uint32_t keep, int limit) {
    constexpr int J = kBufCap / kScanThreads, H = J / 2;
    const int nb = S.nb, tid = threadIdx.x;
    uint64_t mk[H];
    int32_t mi[H];
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int i = tid + j * kScanThreads;
        uint64_t k = ~0ull;
        if (i < nb) k = S.key[i];
        if (j < H) {
            mk[j] = k;
            mi[j] = i < nb ? S.idx[i] : 0;
        }
        if (i < nb && (uint32_t)(k >> 32) <= keep) bits |= 1u << j;
    }
    uint32_t total;
    uint32_t o = run_excl_scan((uint32_t)__popc(bits), S.scan_wsum, &total);
    if ((int)total > limit || total > (uint32_t)(kBufCap / 2)) {
        __syncthreads();
        return false;
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if ((bits >> j) & 1u) {
            if (j < H) {
                S.key[o] = mk[j];
                S.idx[o] = mi[j];
            } else {
                const int i = tid + j * kScanThreads;
                const uint64_t k = S.key[i];
                const int32_t x = S.idx[i];
                S.key[o] = k;
                S.idx[o] = x;
            }
            ++o;
        }
    }
    __syncthreads();
    if (tid == 0) { S.nb = (int)total; S.thr = ((uint64_t)keep << 32) | 0xFFFFFFFFull; }
    __syncthreads();
    return true;
}

// Shrinks the buffer (R < nb) to the entries that can still be among the R
// smallest: score at most a bound on the R-th smallest score, plus `twoE`
// in the mixed form (every entry within 2 E_q of the R-th smallest
// approximate distance, see the mixed IMI scan).  Dropped entries have R
// kept ones of smaller key, so the final top-R is unchanged.  When more
// than min(limit, kBufCap / 2) entries would stay, the plain form keeps the
// R smallest keys exactly; the mixed form, during the scan (limit <
// kBufCap), marks the query for an exact redo and does the same, and after
// it leaves the buffer as it is.  Not inlined: the scan rounds keep the
// next round's loads in registers across it.
__device__ __noinline__ void compact_bound(ScanShared &S, int R, float twoE, bool margin, int limit) {
    int n_le;
    const uint32_t ub = select_bound(S, S.nb, R, &n_le);
    const uint32_t keep = margin ? sc_float_key(sc_key_float(ub) + twoE) : ub;
    if (keep_at_most(S, keep, limit)) return;
    if (margin && limit >= kBufCap) return;
    if (margin && threadIdx.x == 0) S.overflow = 1;
    compact(S, R);
}

// Sorts key/idx[0..n) ascending (n <= kBufCap) by bitonic sort.
__device__ void sort_buffer(ScanShared &S, int n) {
    int np = 1;
    while (np < n) np <<= 1;
    for (int i = n + threadIdx.x; i < np; i += blockDim.x) { S.key[i] = ~(uint64_t)0; S.idx[i] = -1; }
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < np / 2; i += blockDim.x) {
                int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                bool up = ((lo & size) == 0);
                uint64_t a = S.key[lo], b = S.key[hi];
                if ((a > b) == up) {
                    S.key[lo] = b; S.key[hi] = a;
                    int32_t t = S.idx[lo]; S.idx[lo] = S.idx[hi]; S.idx[hi] = t;
                }
            }
            __syncthreads();
        }
    }
}

// Scan + top-R of query q over its cell list: cstart/clen/cinfo [ncell]
// (start of each cell in pid/codes, its length, and h3 | h4 << 16), `sum`
// candidates in all.  dqr/dqc and the cell arrays may live in LDS or global
// memory.  All threads of the workgroup call it; `lut` is LDS of mp*kp floats.
//
// Each round covers kScanRound candidates: a thread takes U = kScanRound / blockDim
// of them (positions tid, tid + blockDim, ...), walking a cell pointer that
// only moves forward, so their code loads are independent and in flight
// together.  Additions keep the reference's order (k ascending).

template <int MPW>   // 32-bit words of codes per candidate (mp = 4 * MPW)
__device__ __forceinline__ void load_codes(const uint8_t *codes, int vidx, uint32_t *w) {
    if (MPW == 2) {
        const uint2 v = *reinterpret_cast<const uint2 *>(codes + (size_t)vidx * 8);
        w[0] = v.x; w[1] = v.y;
    } else if (MPW == 4) {
        const uint4 v = *reinterpret_cast<const uint4 *>(codes + (size_t)vidx * 16);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else {
        const uint32_t *p = reinterpret_cast<const uint32_t *>(codes + (size_t)vidx * 4 * MPW);
        for (int i = 0; i < MPW; ++i) w[i] = p[i];
    }
}

// Any mp; IMI sequences read dot_cr rows from global memory.
__device__ void scan_run_generic(const ScanArgs &a, ScanShared &S, const float *lut,
                                 const float *dqc, float qsum, int rn, int rtotal, int base_pos) {
    const int tid = threadIdx.x, NT = blockDim.x, U = kScanRound / NT;
    const int mp = a.mp, kp = a.kp, kc = a.kc, half = mp >> 1;
    const bool imi = a.kind == SC_KIND_IMI;
    const int bs = kp * mp / a.mc;
    int cur = 0, held = -1;       // the cell whose start term and table rows are held
    float d0 = 0.0f;
    const float *r1 = nullptr, *r2 = nullptr;
    for (int p0 = 0; p0 < rtotal; p0 += kScanRound) {
        const bool full = S.nb > kBufCap - kScanRound;   // read by all before any append
        __syncthreads();
        if (full) compact_bound(S, a.R, 0.0f, false, kBufCap - kScanRound);
        for (int u = 0; u < U; ++u) {
            const int p = p0 + u * NT + tid;
            if (p >= rtotal) continue;
            while (cur + 1 < rn && S.c_pref[cur + 1] <= p) ++cur;
            const int vidx = S.c_start[cur] + (p - S.c_pref[cur]);
            if (cur != held) {      // a thread's positions ascend: once per cell it meets
                const uint32_t ci = S.c_info[cur];
                const int h3 = (int)(ci & 0xFFFFu), h4 = (int)(ci >> 16);
                if (!imi) {
                    d0 = qsum + dqc[h3];
                } else {
                    d0 = (qsum + dqc[h3]) + dqc[kc + h4];
                    r1 = a.dot_cr + (size_t)h3 * bs;
                    r2 = a.dot_cr + (size_t)(h4 + kc - 1) * bs;
                }
                held = cur;
            }
            const uint8_t *cc = a.codes + (size_t)vidx * mp;
            float d = d0;
            if (!imi) {
                for (int k = 0; k < mp; ++k) d += lut[k * kp + cc[k]];
            } else {
                for (int k = 0; k < half; ++k) {
                    int b = k * kp + cc[k];
                    d += lut[b];
                    d += r1[b];
                }
                for (int k = half; k < mp; ++k) {
                    int b = k * kp + cc[k];
                    d += lut[b];
                    d += r2[b];
                }
            }
            const uint64_t key = ((uint64_t)sc_float_key(d) << 32) | (uint32_t)(base_pos + p);
            if (key < S.thr) {
                const int o = atomicAdd(&S.nb, 1);
                S.key[o] = key;
                S.idx[o] = vidx;
            }
        }
        __syncthreads();
    }
}

// SC_FORM_REAL (include/query.h:307-312, include/multi_query.h:255-259,
// include/sc_query.h:232-237): a candidate's distance is S1 between the query
// and its raw vector (lib/simple-cluster/include/utilities.h:194-230: the
// float difference, squared and summed in double in dimension order, stored
// as float; the square is exact, so an FMA gives the same sums).  The raw
// vector is read 16 values at a time so its loads overlap.
__device__ void scan_run_real(const ScanArgs &a, ScanShared &S, const float *qv, int rn,
                              int rtotal, int base_pos) {
    const int tid = threadIdx.x, NT = blockDim.x, U = kScanRound / NT;
    const int dim = a.dim;
    int cur = 0;
    for (int p0 = 0; p0 < rtotal; p0 += kScanRound) {
        const bool full = S.nb > kBufCap - kScanRound;   // read by all before any append
        __syncthreads();
        if (full) compact_bound(S, a.R, 0.0f, false, kBufCap - kScanRound);
        for (int u = 0; u < U; ++u) {
            const int p = p0 + u * NT + tid;
            if (p >= rtotal) continue;
            while (cur + 1 < rn && S.c_pref[cur + 1] <= p) ++cur;
            const int vidx = S.c_start[cur] + (p - S.c_pref[cur]);
            const float *y = a.raw + (size_t)a.pid[vidx] * dim;
            double acc = 0.0;
            int t = 0;
            for (; t + 32 <= dim; t += 32) {     // 128 bytes of the row in flight
                float yv[32];
#pragma unroll
                for (int j = 0; j < 32; ++j) yv[j] = y[t + j];
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    const double d = (double)(qv[t + j] - yv[j]);
                    acc = __fma_rn(d, d, acc);
                }
            }
            for (; t + 16 <= dim; t += 16) {
                float yv[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) yv[j] = y[t + j];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const double d = (double)(qv[t + j] - yv[j]);
                    acc = __fma_rn(d, d, acc);
                }
            }
            for (; t < dim; ++t) {
                const double d = (double)(qv[t] - y[t]);
                acc = __fma_rn(d, d, acc);
            }
            const uint64_t key =
                ((uint64_t)sc_float_key((float)acc) << 32) | (uint32_t)(base_pos + p);
            if (key < S.thr) {
                const int o = atomicAdd(&S.nb, 1);
                S.key[o] = key;
                S.idx[o] = vidx;
            }
        }
        __syncthreads();
    }
}

// Exclusive prefix of one value per thread over a kScanThreads workgroup;
// *total gets the sum.  All threads call it.
__device__ uint32_t run_excl_scan(uint32_t x, uint32_t *wsum, uint32_t *total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int i = 0; i < kScanThreads / 64; ++i) {
        if (i < wv) base += wsum[i];
        all += wsum[i];
    }
    *total = all;
    return base + incl - x;
}

// The cell of run position p (0 <= p < c_pref[rn]): the last cell c with
// c_pref[c] <= p, which is the non-empty cell holding p.
__device__ __forceinline__ int find_cell(const ScanShared &S, int rn, int p) {
    int lo = 0, hi = rn - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.c_pref[mid] <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// One round's candidates of a thread: positions p0 + u * NT + tid, their cell
// (binary search over the run's prefix) and vector index, and their codes
// (and, for the mixed form, e_v) loaded into registers.  A round's loads are
// issued before the previous round is processed, so they are in flight
// while it runs.
template <int MPW, bool EV>
struct RoundRegs {
    uint32_t cw[kScanSlots][MPW];
    int vx[kScanSlots], cl[kScanSlots];
    float ev[kScanSlots];
    __device__ __forceinline__ void fetch(const ScanArgs &a, const ScanShared &S, int rn,
                                          int rtotal, int p0) {
        const int NT = blockDim.x, U = kScanRound / NT;
#pragma unroll
        for (int u = 0; u < kScanSlots; ++u) {
            const int p = p0 + u * NT + threadIdx.x;
            vx[u] = -1;
            if (u < U && p < rtotal) {
                cl[u] = find_cell(S, rn, p);
                vx[u] = S.c_start[cl[u]] + (p - S.c_pref[cl[u]]);
                if (EV && MPW == 2 && a.mrec) {   // codes and e_v in one 16-byte load
                    const uint4 r = a.mrec[vx[u]];
                    cw[u][0] = r.x;
                    cw[u][MPW - 1] = r.y;
                    ev[u] = __uint_as_float(r.z);
                } else {
                    load_codes<MPW>(a.codes, vx[u], cw[u]);
                    if (EV) ev[u] = a.evec[vx[u]];
                }
            }
        }
    }
};

// IVF / SC runs: one combined table row (diff_qr + dot_cr of the run's h3)
// in LDS, d0 = q_sum + diff_qc[h3].
template <int MPW>
__device__ void scan_run_packed(const ScanArgs &a, ScanShared &S, const float *lut, float d0,
                                int rn, int rtotal, int base_pos) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const int kp = a.kp;
    RoundRegs<MPW, false> cur, nxt;
    cur.fetch(a, S, rn, rtotal, 0);
    for (int p0 = 0; p0 < rtotal; p0 += kScanRound) {
        if (p0 + kScanRound < rtotal) nxt.fetch(a, S, rn, rtotal, p0 + kScanRound);
        const bool full = S.nb > kBufCap - kScanRound;   // read by all before any append
        __syncthreads();
        if (full) compact_bound(S, a.R, 0.0f, false, kBufCap - kScanRound);
#pragma unroll
        for (int u = 0; u < kScanSlots; ++u) {
            if (cur.vx[u] < 0) continue;
            float d = d0;
#pragma unroll
            for (int w4 = 0; w4 < MPW; ++w4) {
                const uint32_t wv = cur.cw[u][w4];
#pragma unroll
                for (int b = 0; b < 4; ++b) d += lut[(w4 * 4 + b) * kp + ((wv >> (8 * b)) & 255u)];
            }
            const int p = p0 + u * NT + tid;
            const uint64_t key = ((uint64_t)sc_float_key(d) << 32) | (uint32_t)(base_pos + p);
            if (key < S.thr) {
                const int o = atomicAdd(&S.nb, 1);
                S.key[o] = key;
                S.idx[o] = cur.vx[u];
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

// The IMI sequence over a whole run in traversal order, the rows read from
// dot_cr in global memory (L2), the codes loaded MPW words at a time.
template <int MPW>
__device__ void scan_run_imi_flat(const ScanArgs &a, ScanShared &S, const float *lut,
                                  const float *dqc, float qsum, int rn, int rtotal, int base_pos) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const int kp = a.kp, kc = a.kc, bs = kp * 4 * MPW / a.mc;
    RoundRegs<MPW, false> cur, nxt;
    cur.fetch(a, S, rn, rtotal, 0);
    for (int p0 = 0; p0 < rtotal; p0 += kScanRound) {
        if (p0 + kScanRound < rtotal) nxt.fetch(a, S, rn, rtotal, p0 + kScanRound);
        const bool full = S.nb > kBufCap - kScanRound;   // read by all before any append
        __syncthreads();
        if (full) compact_bound(S, a.R, 0.0f, false, kBufCap - kScanRound);
#pragma unroll
        for (int u = 0; u < kScanSlots; ++u) {
            if (cur.vx[u] < 0) continue;
            const uint32_t ci = S.c_info[cur.cl[u]];
            const int h3 = (int)(ci & 0xFFFFu), h4 = (int)(ci >> 16);
            const float *r1 = a.dot_cr + (size_t)h3 * bs;
            const float *r2 = a.dot_cr + (size_t)(kc + h4) * bs - (size_t)2 * MPW * kp;
            float d = (qsum + dqc[h3]) + dqc[kc + h4];
#pragma unroll
            for (int w4 = 0; w4 < MPW; ++w4) {
                const uint32_t wv = cur.cw[u][w4];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int k = w4 * 4 + b;
                    const int i = k * kp + (int)((wv >> (8 * b)) & 255u);
                    d += lut[i];
                    d += k < 2 * MPW ? r1[i] : r2[i];
                }
            }
            const int p = p0 + u * NT + tid;
            const uint64_t key = ((uint64_t)sc_float_key(d) << 32) | (uint32_t)(base_pos + p);
            if (key < S.thr) {
                const int o = atomicAdd(&S.nb, 1);
                S.key[o] = key;
                S.idx[o] = cur.vx[u];
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

// Mixed form of the IMI scan.  The exact distance E (include/multi_query.h:
// 231-253) is d0 plus, for every subspace k, diff_qr[k][c_k] and dot_cr[row][k]
// [c_k] added one at a time.  The dot_cr terms do not depend on the query, so
// the index keeps their sequential FP32 sum per vector, e_v; the approximate
// distance is
//     A = (d0 + S * sum_k f16(diff_qr[k][c_k] / S)) + e_v
// (8 FP16 LDS lookups and one coalesced load instead of 16 lookups, 8 of them
// L2 gathers).  |A - E| <= E_q for the bound of mixed_bound(); every candidate
// with A within 2 E_q of the running R-th smallest A is kept (compact_bound),
// which contains the exact top-R, and the kept ones are re-ranked with E.
template <int MPW>
__device__ void scan_run_imi_mixed(const ScanArgs &a, ScanShared &S, const _Float16 *lut16,
                                   int rn, int rtotal, int base_pos) {
    const int tid = threadIdx.x, NT = blockDim.x;
    const int kp = a.kp;
    const float scale = S.lscale, twoE = S.twoE;
    RoundRegs<MPW, true> cur, nxt;
    cur.fetch(a, S, rn, rtotal, 0);
    for (int p0 = 0; p0 < rtotal; p0 += kScanRound) {
        if (p0 + kScanRound < rtotal) nxt.fetch(a, S, rn, rtotal, p0 + kScanRound);
        const bool full = S.nb > kBufCap - kScanRound;   // read by all before any append
        __syncthreads();
        if (full) compact_bound(S, a.R, twoE, true, kBufCap - kScanRound);
#pragma unroll
        for (int u = 0; u < kScanSlots; ++u) {
            if (cur.vx[u] < 0) continue;
            float acc = 0.0f;
#pragma unroll
            for (int w4 = 0; w4 < MPW; ++w4) {
                const uint32_t wv = cur.cw[u][w4];
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc += (float)lut16[(w4 * 4 + b) * kp + ((wv >> (8 * b)) & 255u)];
            }
            const float A = (S.c_d0[cur.cl[u]] + scale * acc) + cur.ev[u];
            const int p = p0 + u * NT + tid;
            const uint64_t key = ((uint64_t)sc_float_key(A) << 32) | (uint32_t)(base_pos + p);
            if (key < S.thr) {
                const int o = atomicAdd(&S.nb, 1);
                S.key[o] = key;
                S.idx[o] = cur.vx[u];
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

// Per query, before the mixed IMI scan: the FP16 table f16(diff_qr / S) in
// LDS with S a power of two that keeps every entry below 2^14, and the keep
// margin 2 E_q.  The bound on |A - E| (A, E as above; u = 2^-24):
//   table rounding          2^-11 * sum_k |diff_qr_k|  +  mp * S * u (subnormals)
//   sums of A and of e_v    (7 + 1 + 1 + 7) * u * P
//   the 16 additions of E   16 * u * P
// with P >= |d0| + sum_k |diff_qr_k| + sum_k |dot_cr_k| >= every partial sum,
// |d0| <= |q_sum| + 2 max|diff_qc|; sum_k |.| <= mp * max|.|.  E_q is that
// bound times 1.25 (room for the rounding of the bound itself).
__device__ void mixed_setup(const ScanArgs &a, ScanShared &S, _Float16 *lut16, const float *dqr,
                            const float *dqc, float qsum) {
    const int tid = threadIdx.x, lutn = a.mp * a.kp;
    __shared__ uint32_t s_m[2];
    if (tid == 0) { s_m[0] = 0; s_m[1] = 0; }
    __syncthreads();
    uint32_t m0 = 0, m1 = 0;
    // the table (up to 16 entries per thread) read once, kept in registers
    // for the conversion below
    constexpr int kHeld = 16;
    float held[kHeld];
    const bool hold = lutn <= kHeld * (int)blockDim.x;
#pragma unroll
    for (int j = 0; j < kHeld; ++j) {
        const int i = tid + j * (int)blockDim.x;
        held[j] = hold && i < lutn ? dqr[i] : 0.0f;
        m0 = max(m0, __float_as_uint(fabsf(held[j])));
    }
    if (!hold)
        for (int i = tid; i < lutn; i += blockDim.x) m0 = max(m0, __float_as_uint(fabsf(dqr[i])));
    for (int i = tid; i < 2 * a.kc; i += blockDim.x) m1 = max(m1, __float_as_uint(fabsf(dqc[i])));
    for (int o = 32; o > 0; o >>= 1) {
        m0 = max(m0, (uint32_t)__shfl_xor(m0, o));
        m1 = max(m1, (uint32_t)__shfl_xor(m1, o));
    }
    if ((tid & 63) == 0) { atomicMax(&s_m[0], m0); atomicMax(&s_m[1], m1); }
    __syncthreads();
    const float maxlut = __uint_as_float(s_m[0]), maxqc = __uint_as_float(s_m[1]);
    int ex = 0;
    (void)frexpf(maxlut, &ex);                     // maxlut < 2^ex
    const float scale = maxlut > 0.0f ? ldexpf(1.0f, ex - 14) : 1.0f;
    const float inv = 1.0f / scale;                // exact: a power of two
    if (hold) {
#pragma unroll
        for (int j = 0; j < kHeld; ++j) {
            const int i = tid + j * (int)blockDim.x;
            if (i < lutn) lut16[i] = (_Float16)(held[j] * inv);
        }
    } else {
        for (int i = tid; i < lutn; i += blockDim.x) lut16[i] = (_Float16)(dqr[i] * inv);
    }
    if (tid == 0) {
        const double u = ldexp(1.0, -24), mp = a.mp;
        const double P = (fabs((double)qsum) + 2.0 * maxqc) + mp * ((double)maxlut + a.dot_absmax);
        const double E = ldexp(1.0, -11) * mp * maxlut + mp * scale * u + 32.0 * u * P;
        S.lscale = scale;
        S.twoE = (float)(2.0 * 1.25 * E * a.margin_scale);
    }
    __syncthreads();
}

// The exact IMI distance of the vector at index position vidx.  Its cell's
// nearest centroids (h3, h4) come from the index's per-position record or,
// without one, from the cell (the first whose inclusive prefix L exceeds
// vidx) by its number (dense: h3*kc + h4) or its key (two halves); the
// operation sequence is include/multi_query.h:231-253.
__device__ __forceinline__ float imi_exact(const ScanArgs &a, const float *dqr, const float *dqc,
                                           float qsum, int vidx) {
    const int mp = a.mp, kp = a.kp, kc = a.kc, half = mp >> 1, bs = kp * mp / a.mc;
    if (a.mrec && mp == 8) {       // the position's 16-byte record: codes, e_v, cell
        const uint4 r = a.mrec[vidx];
        const int h3 = (int)(r.w & 0xFFFFu), h4 = (int)(r.w >> 16);
        const float *r1 = a.dot_cr + (size_t)h3 * bs, *r2 = a.dot_cr + (size_t)(h4 + kc - 1) * bs;
        float d = (qsum + dqc[h3]) + dqc[kc + h4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int b = k * kp + (int)((r.x >> (8 * k)) & 255u);
            d += dqr[b];
            d += r1[b];
        }
#pragma unroll
        for (int k = 4; k < 8; ++k) {
            const int b = k * kp + (int)((r.y >> (8 * (k - 4))) & 255u);
            d += dqr[b];
            d += r2[b];
        }
        return d;
    }
    int h3, h4;
    if (a.vcell) {                 // the index's per-position cell record
        const uint32_t v = a.vcell[vidx];
        h3 = (int)(v & 0xFFFFu);
        h4 = (int)(v >> 16);
    } else {
        int64_t lo = 0, hi = a.ncells;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (a.L[mid] <= vidx) lo = mid + 1; else hi = mid;
        }
        if (a.ckeys) {
            const uint64_t k2 = (uint64_t)kc * (uint64_t)kc, key = a.ckeys[lo];
            h3 = (int)(key / k2 / (uint64_t)kc);
            h4 = (int)(key % k2 / (uint64_t)kc);
        } else {
            h3 = (int)(lo / kc);
            h4 = (int)(lo % kc);
        }
    }
    const float *r1 = a.dot_cr + (size_t)h3 * bs, *r2 = a.dot_cr + (size_t)(h4 + kc - 1) * bs;
    const uint8_t *cc = a.codes + (size_t)vidx * mp;
    float d = (qsum + dqc[h3]) + dqc[kc + h4];
    for (int k = 0; k < half; ++k) {
        const int b = k * kp + cc[k];
        d += dqr[b];
        d += r1[b];
    }
    for (int k = half; k < mp; ++k) {
        const int b = k * kp + cc[k];
        d += dqr[b];
        d += r2[b];
    }
    return d;
}

__device__ void scan_query(const ScanArgs &a, ScanShared &S, float *lut, int64_t q,
                           const float *dqr, const float *dqc, float qsum,
                           const int32_t *cstart, const int32_t *clen, const uint32_t *cinfo,
                           int ncell, int sum, uint64_t *t_scanned = nullptr) {
    const int tid = threadIdx.x;
    const int mp = a.mp, kp = a.kp;
    const int lutn = mp * kp;
    const bool imi = a.kind == SC_KIND_IMI;
    const int imi_mode = a.imi_mode;
    const bool real = a.raw != nullptr;   // lut holds the query vector then
    const bool mixed = imi && a.evec && !real && (mp == 8 || mp == 16);
    const int bs = kp * mp / a.mc;
    if (tid == 0) { S.nb = 0; S.thr = ~(uint64_t)0; S.overflow = 0; }
    if (real)
        for (int i = tid; i < a.dim; i += blockDim.x) lut[i] = a.Q[(size_t)q * a.dim + i];
    else if (mixed)
        mixed_setup(a, S, reinterpret_cast<_Float16 *>(lut), dqr, dqc, qsum);
    else if (imi)
        for (int i = tid; i < lutn; i += blockDim.x) lut[i] = dqr[i];
    int lut_row = -1;     // h3 whose combined row is staged (IVF/SC)
    int base_pos = 0;     // position of the first candidate of the run
    __syncthreads();

    int c0 = 0;
    while (c0 < ncell) {
        // a run: consecutive cells sharing h3 (IVF/SC), or up to kRunCells (IMI),
        // loaded by all threads at once (one cell each) with a prefix sum of
        // their sizes; the run ends before the first cell of another h3
        {
            const int navail = min(kRunCells, ncell - c0);
            if (tid == 0) S.run_cells = navail;
            __syncthreads();
            const int i = tid;
            uint32_t ln = 0;
            if (i < navail) {
                const uint32_t ci = cinfo[c0 + i];
                S.c_start[i] = cstart[c0 + i];
                S.c_info[i] = ci;
                ln = (uint32_t)clen[c0 + i];
                if (!imi && (ci & 0xFFFFu) != (cinfo[c0] & 0xFFFFu)) atomicMin(&S.run_cells, i);
            }
            uint32_t tot;
            const uint32_t ex = run_excl_scan(ln, S.scan_wsum, &tot);
            if (i < navail) S.c_pref[i] = (int32_t)ex;
            if (tid == 0) S.c_pref[navail] = (int32_t)tot;
        }
        __syncthreads();

        const int rn = S.run_cells;
        const int rtotal = S.c_pref[rn];
        if (mixed) {
            for (int i = tid; i < rn; i += blockDim.x) {
                const uint32_t ci = S.c_info[i];
                S.c_d0[i] = (qsum + dqc[ci & 0xFFFFu]) + dqc[a.kc + (ci >> 16)];
            }
            __syncthreads();
        }
        float d0 = 0.0f;
        if (!imi && !real) {
            const int h3 = (int)(S.c_info[0] & 0xFFFFu);
            d0 = qsum + dqc[h3];
            if (h3 != lut_row) {
                const float *row = a.dot_cr + (size_t)h3 * bs;
                for (int i = tid; i < lutn; i += blockDim.x) lut[i] = dqr[i] + row[i];
                lut_row = h3;
            }
        }
        __syncthreads();
        if (real)
            scan_run_real(a, S, lut, rn, rtotal, base_pos);
        else if (mixed && mp == 8)
            scan_run_imi_mixed<2>(a, S, reinterpret_cast<const _Float16 *>(lut), rn, rtotal,
                                  base_pos);
        else if (mixed && mp == 16)
            scan_run_imi_mixed<4>(a, S, reinterpret_cast<const _Float16 *>(lut), rn, rtotal,
                                  base_pos);
        else if (imi && imi_mode == 1 && mp == 8)
            scan_run_imi_flat<2>(a, S, lut, dqc, qsum, rn, rtotal, base_pos);
        else if (imi && imi_mode == 1 && mp == 16)
            scan_run_imi_flat<4>(a, S, lut, dqc, qsum, rn, rtotal, base_pos);

        else if (!imi && mp == 8)
            scan_run_packed<2>(a, S, lut, d0, rn, rtotal, base_pos);
        else if (!imi && mp == 16)
            scan_run_packed<4>(a, S, lut, d0, rn, rtotal, base_pos);
        else
            scan_run_generic(a, S, lut, dqc, qsum, rn, rtotal, base_pos);
        base_pos += rtotal;
        c0 += rn;
        __syncthreads();
    }
    if (t_scanned && tid == 0) *t_scanned = wall_clock64();
    if (mixed) {
        if (S.overflow) {              // redone exactly by the caller
            if (tid == 0 && a.redo) a.redo[q * 8 + 7] = 1;
            return;
        }
        // only the entries within the margin of the final R-th smallest
        // approximate distance can be in the exact top-R (a filter worth
        // its passes only when it spares threads a second re-rank)
        if (sum >= a.R && S.nb > a.R && S.nb > kScanThreads)
            compact_bound(S, a.R, S.twoE, true, kBufCap);
        // exact re-rank of the kept candidates (their positions stay)
        for (int i = tid; i < S.nb; i += blockDim.x) {
            const float e = imi_exact(a, dqr, dqc, qsum, S.idx[i]);
            S.key[i] = ((uint64_t)sc_float_key(e) << 32) | (S.key[i] & 0xFFFFFFFFull);
        }
        __syncthreads();
    }
    // final selection: the R smallest keys, ascending
    const int R = a.R;
    int nout;
    if (sum >= R) {
        // (a buffer of at most a sort's 256 entries is sorted whole)
        if (S.nb > R && S.nb > 256) compact_bound(S, R, 0.0f, false, kBufCap);
        nout = R;
        sort_buffer(S, S.nb);
    }
    int32_t *ids = a.ids + (size_t)q * R;
    float *dists = a.dists + (size_t)q * R;
    if (sum < R) {
        // fewer than R candidates: traversal (position) order, like the
        // reference; swap the key halves so the sort runs on position
        for (int i = tid; i < S.nb; i += blockDim.x) {
            uint64_t k = S.key[i];
            S.key[i] = (k << 32) | (k >> 32);
        }
        __syncthreads();
        nout = S.nb;
        sort_buffer(S, nout);
    }
    for (int r = tid; r < R; r += blockDim.x) {
        if (r < nout) {
            ids[r] = a.pid[S.idx[r]];
            const uint64_t k = S.key[r];
            dists[r] = sc_key_float(sum >= R ? (uint32_t)(k >> 32) : (uint32_t)k);
        } else {
            ids[r] = -1;
            dists[r] = INFINITY;
        }
    }
    if (a.sums && tid == 0) a.sums[q] = sum;
}

// Four waves per SIMD: the register budget that lets the four workgroups per
// CU the LDS allows run together (unconstrained, the compiler's allocation
// holds it to two; a few cold-path values spill).
__global__ __launch_bounds__(kScanThreads) __attribute__((amdgpu_waves_per_eu(5, 8)))
void k_scan_topk(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    ScanShared &S = *reinterpret_cast<ScanShared *>(smem);
    float *lut = reinterpret_cast<float *>(smem + sizeof(ScanShared));   // [mp*kp]
    if (a.qcount && (int)blockIdx.x >= *a.qcount) return;   // the whole workgroup
    const int64_t q = a.qlist ? a.qlist[blockIdx.x] : blockIdx.x;
    if (a.skip && a.skip[q]) return;                         // the whole workgroup
    scan_query(a, S, lut, q, a.dqr + (size_t)q * a.mp * a.kp, a.dqc + (size_t)q * a.mc * a.kc,
               a.qs[q * 4], a.cells + (size_t)q * a.ccap, a.clen + (size_t)q * a.ccap,
               a.cinfo + (size_t)q * a.ccap,
               a.qinfo[q * 8 + 0], a.qinfo[q * 8 + 1]);
}

// ----------------------------------------------------------------------------
// Fused persistent search (IVF and SC kinds): one workgroup runs all stages of
// one query in LDS -- LUTs, coarse ranking, traversal (thread 0, exact heap
// replay), scan and top-R -- then takes the next query from a work counter.
// A query whose heap or cell list outgrows its LDS capacity is flagged and
// finished by the staged kernels afterwards.
//
// The scan visits the traversed cells grouped by coarse centroid h3 (a
// counting sort of the cell list), so the combined table diff_qr + dot_cr[h3]
// (include/sc_query.h:216-229, include/query.h:295-305) is built once per
// distinct h3 instead of once per cell.  Every candidate keeps the position it
// has in the reference's traversal order, so the keys (dist, position) and the
// canonical top-R do not depend on the visiting order.
// ----------------------------------------------------------------------------
struct FusedArgs {
    ScanArgs s;
    const float *Q, *cq, *pq, *norm_c, *norm_r;
    int64_t nq;
    int w, T, hcap, ccap, npow2, bufcap;
    int *counter;
    int32_t *flags;     // [nq] 1 = finish with the staged kernels
    int *nflag;
    int32_t *qinfo;     // [nq][8] statistics (same fields as the staged path)
    const float *lut_dqc, *lut_dqr;   // [nq][mc*kc], [nq][mp*kp] or null: compute here
    unsigned long long *phase;        // [6]: wall-clock ticks per phase, workgroups
    float margin_scale;               // mixed form: multiplies the error bound (1)
    int nvec;                         // vectors in the index
    int pop_budget;                   // SC: cells popped before a query is handed on
    const int32_t *Tq;                // [nq] per-query candidate budgets or null
};

struct FusedLayout {
    size_t a, b, c, d, e, f, total;
};

// Region A: top-R buffer (bufcap keys) | heap + row lengths | coarse
// sort keys (kc > 256); B: per-cell start, length, info, position, permuted
// prefix [ccap + 1], permutation (u16); C: sorted scores + centroids [kc]
// (after the traversal: per-centroid counts and the group list); D: diff_qc
// (+ diff_qr when computed here); E: combined table [mp*kp] (FP32, or FP16
// in the mixed form); F: query [dim].
__host__ __device__ inline FusedLayout fused_layout(int dim, int mc, int kc, int mp, int kp,
                                                    int hcap, int ccap, int npow2, int bufcap,
                                                    bool dqr_lds, bool half_table) {
    auto up = [](size_t v) { return (v + 15) & ~(size_t)15; };
    FusedLayout L{};
    size_t a = (size_t)bufcap * 8;
    if ((size_t)kc * 8 > a) a = (size_t)kc * 8;   // the grouping's two counts per centroid
    // heap (value, id) pairs, SC row lengths, then the SC window of
    // (start, length) per rank pair
    const size_t heap = (size_t)hcap * 8 + (size_t)kc * 4 + (size_t)kWin * kWin * 8;
    if (heap > a) a = heap;
    if (npow2 > 256 && (size_t)npow2 * 8 > a) a = (size_t)npow2 * 8;
    L.a = 0;
    L.b = up(a);
    L.c = up(L.b + (size_t)ccap * 18 + ((size_t)ccap + 1) * 4);
    L.d = up(L.c + (size_t)kc * 8);
    L.e = up(L.d + 4 * ((size_t)mc * kc + (dqr_lds ? (size_t)mp * kp : 0)));
    L.f = up(L.e + (half_table ? 2 : 4) * (size_t)mp * kp);
    L.total = up(L.f + 4 * (size_t)dim);
    return L;
}

// Small shared state of the streaming top-R buffer.
struct TopShared {
    uint32_t hist[256];
    uint64_t thr;          // keys >= thr cannot enter the top R any more
    uint64_t sel_prefix, sel_and, sel_or;
    int nb;                // entries in the buffer
    int sel_rem, sel_done, wcount;
    int wsum[kMaxFusedThreads / 64];
    uint32_t amax_r, amax_c;   // mixed form: max |diff_qr|, max |diff_qc| (float bits)
    int ovf, cnt;              // mixed form: margin set outgrew the buffer; a count
};

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int m) {
    const uint32_t lo = __shfl_xor((uint32_t)x, m), hi = __shfl_xor((uint32_t)(x >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Appends the keys of the lanes with `take` at a counter, one LDS atomic per
// wave.  Every lane of the wave must call it.
__device__ __forceinline__ void wave_append(int *counter, uint64_t *key, bool take, uint64_t k) {
    const uint64_t bal = __ballot(take);
    if (bal == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)bal) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(bal));
    base = __shfl(base, leader);
    if (take) {
        const int o = base + __popcll(bal & ((1ull << lane) - 1ull));
        key[o] = k;
    }
}

// Threshold thr with exactly R of the (unique) keys key[0..nb) <= thr,
// R <= nb.  Radix select on 8-bit digits that starts below the bits all keys
// share and stops as soon as the digit bin holding the R-th key is taken
// whole.  All threads call it.
__device__ __forceinline__ uint64_t select_threshold(const uint64_t *key, int nb, int R, TopShared &T) {
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63;
    uint64_t ka = ~0ull, ko = 0;
    for (int i = tid; i < nb; i += NT) { ka &= key[i]; ko |= key[i]; }
    for (int s = 32; s > 0; s >>= 1) { ka &= shfl_xor_u64(ka, s); ko |= shfl_xor_u64(ko, s); }
    if (tid == 0) { T.sel_and = ~0ull; T.sel_or = 0; T.sel_rem = R; T.sel_done = 0; }
    __syncthreads();
    if (lane == 0) { atomicAnd(&T.sel_and, ka); atomicOr(&T.sel_or, ko); }
    __syncthreads();
    const uint64_t kand = T.sel_and, diff = kand ^ T.sel_or;
    if (diff == 0) return kand;                      // a single key (R == nb == 1)
    const int hb = 63 - __clzll((long long)diff);
    int shift = hb - 7;
    uint64_t mask = hb == 63 ? 0ull : (~0ull << (hb + 1));
    uint64_t prefix = kand & mask;
    for (;;) {
        const int sh = shift < 0 ? 0 : shift;
        const int width = shift < 0 ? 8 + shift : 8;
        const uint32_t dmask = (1u << width) - 1u;
        for (int b = tid; b < 256; b += NT) T.hist[b] = 0;
        __syncthreads();
        for (int i = tid; i < nb; i += NT) {
            const uint64_t k = key[i];
            if ((k & mask) == prefix) atomicAdd(&T.hist[(uint32_t)(k >> sh) & dmask], 1u);
        }
        __syncthreads();
        if (tid < 64) {
            const int l = tid;
            const uint32_t c0 = T.hist[4 * l], c1 = T.hist[4 * l + 1], c2 = T.hist[4 * l + 2],
                           c3 = T.hist[4 * l + 3];
            const uint32_t rem = (uint32_t)T.sel_rem;   // read before any lane updates it
            const uint32_t tot = c0 + c1 + c2 + c3;
            uint32_t incl = tot;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o);
                if (l >= o) incl += t;
            }
            const uint32_t excl = incl - tot;
            if (excl < rem && rem <= incl) {
                const uint32_t c[4] = {c0, c1, c2, c3};
                uint32_t acc = excl;
                for (int b = 0; b < 4; ++b) {
                    if (rem <= acc + c[b]) {
                        const uint64_t np = prefix | ((uint64_t)(4 * l + b) << sh);
                        if (rem - acc == c[b]) {      // the whole bin: done
                            T.sel_prefix = np | (sh > 0 ? ((1ull << sh) - 1ull) : 0ull);
                            T.sel_done = 1;
                        } else {
                            T.sel_prefix = np;
                            T.sel_rem = (int)(rem - acc);
                        }
                        break;
                    }
                    acc += c[b];
                }
            }
        }
        __syncthreads();
        prefix = T.sel_prefix;
        if (T.sel_done || sh == 0) return prefix;
        mask |= (uint64_t)dmask << sh;
        shift -= 8;
    }
}

// Keeps exactly the R smallest keys of key[0..T.nb) at the front.  The
// survivors move left chunk by chunk of blockDim.x entries: a chunk is read
// whole before any of its survivors is written, and every write lands below
// the chunk's end, so nothing unread is overwritten.  All threads call it.
__device__ __forceinline__ void compact_top(TopShared &T, uint64_t *key, int R) {
    const int tid = threadIdx.x, NT = blockDim.x;
    __syncthreads();
    const int nb = T.nb;
    const uint64_t th = select_threshold(key, nb, R, T);
    if (tid == 0) T.wcount = 0;
    for (int c0 = 0; c0 < nb; c0 += NT) {
        const int i = c0 + tid;
        const uint64_t k = i < nb ? key[i] : ~0ull;
        __syncthreads();
        wave_append(&T.wcount, key, i < nb && k <= th, k);
    }
    __syncthreads();
    if (tid == 0) { T.nb = R; T.thr = th; }
    __syncthreads();
}

// Mixed form: keeps every key whose approximate distance is within twoE of
// the R-th smallest (at least R keys).  When more than lim keys would stay it
// keeps exactly R and sets T.ovf: the query is then redone exactly.
__device__ __forceinline__ void compact_top_margin(TopShared &T, uint64_t *key, int R, float twoE,
                                   int lim) {
    const int tid = threadIdx.x, NT = blockDim.x;
    __syncthreads();
    const int nb = T.nb;
    const uint64_t kth = select_threshold(key, nb, R, T);
    // one float step above the rounded sum keeps the bound on the safe side
    uint32_t lk = sc_float_key(sc_key_float((uint32_t)(kth >> 32)) + twoE);
    if (lk != 0xFFFFFFFFu) ++lk;
    uint64_t th = ((uint64_t)lk << 32) | 0xFFFFFFFFull;
    if (th < kth) th = kth;
    if (tid == 0) T.cnt = 0;
    __syncthreads();
    int c = 0;
    for (int i = tid; i < nb; i += NT) c += key[i] <= th ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((tid & 63) == 0) atomicAdd(&T.cnt, c);
    __syncthreads();
    const int kept = T.cnt;
    const bool over = kept > lim;
    if (over) th = kth;
    if (tid == 0) T.wcount = 0;
    for (int c0 = 0; c0 < nb; c0 += NT) {
        const int i = c0 + tid;
        const uint64_t k = i < nb ? key[i] : ~0ull;
        __syncthreads();
        wave_append(&T.wcount, key, i < nb && k <= th, k);
    }
    __syncthreads();
    if (tid == 0) {
        T.nb = over ? R : kept;
        T.thr = over ? kth : (th == ~0ull ? th : th + 1);
        if (over) T.ovf = 1;
    }
    __syncthreads();
}

// Bitonic sort of 64 * E keys held E per lane (element lane * E + e), one wave.
template <int E>
__device__ __forceinline__ void wave_bitonic(uint64_t (&k)[E], int32_t (&v)[E]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int size = 2; size <= 64 * E; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= E) {
                const int pl = stride / E;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = lane * E + e;
                    const uint64_t ok = shfl_xor_u64(k[e], pl);
                    const int32_t ov = __shfl_xor(v[e], pl);
                    const bool upw = (i & size) == 0, lower = (i & stride) == 0;
                    const bool sw = (lower == upw) ? (ok < k[e]) : (ok > k[e]);
                    if (sw) { k[e] = ok; v[e] = ov; }
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & stride) continue;
                    const int e2 = e | stride, i = lane * E + e;
                    const bool upw = (i & size) == 0;
                    if ((k[e] > k[e2]) == upw) {
                        const uint64_t tk = k[e]; k[e] = k[e2]; k[e2] = tk;
                        const int32_t tv = v[e]; v[e] = v[e2]; v[e2] = tv;
                    }
                }
            }
        }
    }
}

template <int E>
__device__ __forceinline__ void wave_sort_lds(uint64_t *key, int n) {   // wave 0 only
    const int lane = threadIdx.x & 63;
    uint64_t k[E];
    int32_t v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = lane * E + e;
        k[e] = i < n ? key[i] : ~0ull;
        v[e] = 0;
    }
    wave_bitonic<E>(k, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = lane * E + e;
        if (i < n) key[i] = k[e];
    }
}

// Sorts key[0..n) ascending; the buffer holds at least next_pow2(n)
// entries.  All threads call it.
__device__ __forceinline__ void sort_top(uint64_t *key, int n) {
    const int tid = threadIdx.x, NT = blockDim.x;
    __syncthreads();
    if (n <= 256) {
        if (tid < 64) {
            if (n <= 64) wave_sort_lds<1>(key, n);
            else if (n <= 128) wave_sort_lds<2>(key, n);
            else wave_sort_lds<4>(key, n);
        }
        __syncthreads();
        return;
    }
    int np = 1;
    while (np < n) np <<= 1;
    for (int i = n + tid; i < np; i += NT) key[i] = ~0ull;
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < np / 2; i += NT) {
                const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                const bool upw = (lo & size) == 0;
                const uint64_t x = key[lo], y = key[hi];
                if ((x > y) == upw) { key[lo] = y; key[hi] = x; }
            }
            __syncthreads();
        }
    }
}

// Cells of one scan pass: permuted cell i (of the group) covers positions
// [gp[i], gp[i+1]) of the pass; its vectors start at cstart[order[i]] and its
// first candidate has traversal position cpos[order[i]].
struct GroupCells {
    const int32_t *gp, *cstart, *cpos;
    const uint16_t *order;
};

// One group: candidates at pass positions [p0, p1), first cell i0, table
// `lut` (mp x kp, combined), start value d0 = q_sum + diff_qc[h3].
// MPW = 32-bit words of codes per candidate (mp = 4 * MPW); 0 = any mp.
// MIXED: the table holds FP16 values of (diff_qr + dot_cr) / S; the
// approximate distance d0 + S * sum is kept within a margin of twoE.
// (helpers are force-inlined: the kernel's register-variant attribute keeps
// the inliner from inlining ordinary calls, and calls spill registers)
template <int MPW, bool MIXED>
__device__ __forceinline__ void scan_group(const ScanArgs &a, TopShared &T, uint64_t *key, int bufcap, const float *lut, float d0, int p0, int p1, int i0,
                           const GroupCells &G, float S, float twoE) {
    constexpr int MW = MPW > 0 ? MPW : 1;
    const _Float16 *lut16 = reinterpret_cast<const _Float16 *>(lut);
    const int tid = threadIdx.x, NT = blockDim.x, U = kRound / NT;
    const int mp = a.mp, kp = a.kp;
    int cur = i0, g0 = G.gp[i0], g1 = G.gp[i0 + 1];
    int cs = 0, cp = 0;
    {
        const int c = G.order[i0];
        cs = G.cstart[c];
        cp = G.cpos[c];
    }
    for (int r0 = p0; r0 < p1; r0 += kRound) {
        // every thread reads the fill before any wave appends this round, so
        // all of them take the same compaction decision
        const bool full = T.nb > bufcap - kRound;
        __syncthreads();
        if (full) {
            if (MIXED) compact_top_margin(T, key, a.R, twoE, bufcap - kRound);
            else compact_top(T, key, a.R);
        }
        const uint64_t thr = T.thr;
        uint32_t cw[4][MW];
        int vx[4], ps[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            vx[u] = -1;
            ps[u] = 0;
            const int p = r0 + u * NT + tid;
            if (u < U && p < p1) {
                while (p >= g1) {
                    ++cur;
                    g0 = g1;
                    g1 = G.gp[cur + 1];
                    const int c = G.order[cur];
                    cs = G.cstart[c];
                    cp = G.cpos[c];
                }
                vx[u] = cs + (p - g0);
                ps[u] = cp + (p - g0);
                if (MPW > 0) load_codes<(MPW > 0 ? MPW : 1)>(a.codes, vx[u], cw[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= U) continue;    // U is the same for every thread
            bool take = false;
            uint64_t k = 0;
            if (vx[u] >= 0) {
                float d = d0;
                if (MIXED) {
                    float acc = 0.0f;
                    if (MPW > 0) {
#pragma unroll
                        for (int w4 = 0; w4 < MW; ++w4) {
                            const uint32_t wv = cw[u][w4];
#pragma unroll
                            for (int b = 0; b < 4; ++b)
                                acc += (float)lut16[(w4 * 4 + b) * kp + ((wv >> (8 * b)) & 255u)];
                        }
                    } else {
                        const uint8_t *cc = a.codes + (size_t)vx[u] * mp;
                        for (int j = 0; j < mp; ++j) acc += (float)lut16[j * kp + cc[j]];
                    }
                    d = d0 + S * acc;
                } else if (MPW > 0) {
#pragma unroll
                    for (int w4 = 0; w4 < MW; ++w4) {
                        const uint32_t wv = cw[u][w4];
#pragma unroll
                        for (int b = 0; b < 4; ++b)
                            d += lut[(w4 * 4 + b) * kp + ((wv >> (8 * b)) & 255u)];
                    }
                } else {
                    const uint8_t *cc = a.codes + (size_t)vx[u] * mp;
                    for (int j = 0; j < mp; ++j) d += lut[j * kp + cc[j]];
                }
                k = ((uint64_t)sc_float_key(d) << 32) | (uint32_t)ps[u];
                take = k < thr;
            }
            wave_append(&T.nb, key, take, k);
        }
        __syncthreads();
    }
}

constexpr int kLutRegs = 8;    // table entries per thread kept in registers

// The traversal runs on one lane; left alone, the compiler treats its loaded
// values as wave-uniform and moves each into a scalar register right after its
// load, waiting for the loads one by one.  Passing the address through a
// vector register keeps a batch of loads in flight together.
__device__ __forceinline__ int64_t lane_value(int64_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

// OCC = waves per SIMD the register allocation must allow (3, 4 or 5): more
// workgroups in flight hide the serial phases of one query behind the others
// at the price of register spills; the plan autotuner picks the variant.
// MIXED = the mixed-precision form: FP16 table scan with a rigorous error
// bound, then an exact FP32 re-rank of every candidate the bound cannot
// exclude, so ids and distances equal the plain form's.
template <int OCC, bool MIXED>
__global__ __launch_bounds__(kMaxFusedThreads) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
void k_search_fused(FusedArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ScanArgs &sa = a.s;
    const int dim = sa.dim, mc = sa.mc, kc = sa.kc, mp = sa.mp, kp = sa.kp;
    const bool dqr_lds = a.lut_dqr == nullptr;
    const FusedLayout Lay =
        fused_layout(dim, mc, kc, mp, kp, a.hcap, a.ccap, a.npow2, a.bufcap, dqr_lds, MIXED);
    uint64_t *bkey = reinterpret_cast<uint64_t *>(smem + Lay.a);   // (dist key << 32) | position
    uint64_t *keys = bkey;                                           // coarse sort, kc > 256
    int32_t *cells = reinterpret_cast<int32_t *>(smem + Lay.b);     // cell starts
    int32_t *clen = cells + a.ccap;
    uint32_t *cinfo = reinterpret_cast<uint32_t *>(clen + a.ccap);
    int32_t *cpos = reinterpret_cast<int32_t *>(cinfo + a.ccap);
    int32_t *gp = cpos + a.ccap;                                     // [ccap + 1]
    uint16_t *order = reinterpret_cast<uint16_t *>(gp + a.ccap + 1);
    float *sv = reinterpret_cast<float *>(smem + Lay.c);
    int32_t *si = reinterpret_cast<int32_t *>(sv + kc);
    int32_t *ghist = reinterpret_cast<int32_t *>(sv);                // after the traversal
    uint32_t *gdesc = reinterpret_cast<uint32_t *>(si);
    float *dqc = reinterpret_cast<float *>(smem + Lay.d);
    float *dqr = dqc + mc * kc;                                      // dqr_lds only
    float *lut = reinterpret_cast<float *>(smem + Lay.e);
    float *qv = reinterpret_cast<float *>(smem + Lay.f);
    __shared__ TopShared T;
    __shared__ int64_t s_sbid[kSpec];        // traversal batch (thread 0)
    __shared__ uint32_t s_sinfo[kSpec];
    __shared__ int s_q, s_ncell, s_sum, s_flag, s_ngroup;
    __shared__ float s_qsum;
    const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wid = tid >> 6;
    const bool sc = sa.kind == SC_KIND_SC;
    const int bsc = dim / mc, bsp = dim / mp;
    const int lutn = mp * kp, bs = kp * mp / mc;
    const bool lut_regs = lutn <= kLutRegs * NT;

    // thread 0 keeps per-phase wall-clock ticks (100 MHz) of this workgroup
    __shared__ uint64_t s_tph[6];   // 5 phases + last mark
    if (tid == 0) for (int i = 0; i < 6; ++i) s_tph[i] = 0;
    auto mark = [&](int ph) {
        if (tid == 0) { const uint64_t t = wall_clock64(); s_tph[ph] += t - s_tph[5]; s_tph[5] = t; }
    };
    for (;;) {
        if (tid == 0) {
            s_q = atomicAdd(a.counter, 1);
            T.amax_r = 0u;
            T.amax_c = 0u;
            T.ovf = 0;
        }
        __syncthreads();
        const int64_t q = s_q;
        if (q >= a.nq) break;
        const int Tb = query_budget(a.Tq, q, a.T);   // this query's candidate budget
        if (tid == 0) s_tph[5] = wall_clock64();
        for (int t = tid; t < dim; t += NT) qv[t] = a.Q[(size_t)q * dim + t];
        __syncthreads();
        // LUTs (query.h:154-184) and q_sum (query.h:225-230, sc_query.h:95-100)
        if (tid == 0) {
            float s = 0.0f;
            for (int i = 0; i < dim; ++i) { float t = qv[i]; s += t * t; }
            s_qsum = s;
        }
        if (!dqr_lds) {   // precomputed by k_lut_tiled
            for (int j = tid; j < mc * kc; j += NT) dqc[j] = a.lut_dqc[(size_t)q * mc * kc + j];
        } else for (int j = tid; j < mc * kc + lutn; j += NT) {
            if (j < mc * kc) {     // a.cq / a.pq: transposed codebooks
                const int blk = j / kc, k = j - blk * kc;
                const float *b = a.cq + (size_t)blk * bsc * kc + k;
                const float *x = qv + (size_t)blk * bsc;
                float s = 0.0f;
                for (int t = 0; t < bsc; ++t) s += x[t] * b[(size_t)t * kc];
                dqc[j] = (a.norm_c[j] - s) - s;
            } else {
                const int r = j - mc * kc, sub = r / kp, c = r - sub * kp;
                const float *b = a.pq + (size_t)sub * bsp * kp + c;
                const float *x = qv + (size_t)sub * bsp;
                float s = 0.0f;
                for (int t = 0; t < bsp; ++t) s += x[t] * b[(size_t)t * kp];
                dqr[r] = (a.norm_r[r] - s) - s;
            }
        }
        __syncthreads();
        mark(0);
        const float qsum = s_qsum;
        if (MIXED) {   // max |diff_qc| for the error bound (read after later barriers)
            uint32_t mcq = 0u;
            for (int j = tid; j < mc * kc; j += NT) mcq = max(mcq, __float_as_uint(fabsf(dqc[j])));
            for (int o = 32; o > 0; o >>= 1) mcq = max(mcq, (uint32_t)__shfl_xor(mcq, o));
            if (lane == 0) atomicMax(&T.amax_c, mcq);
        }
        // coarse scores; SC sorts them by (score, centroid)
        if (sc && a.npow2 <= 256) {
            if (tid < 64) {
                auto run = [&](auto tag) {
                    constexpr int E = decltype(tag)::value;
                    uint64_t k[E];
                    int32_t v[E];
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        const int j = lane * E + e;
                        k[e] = j < kc ? (((uint64_t)sc_float_key(qsum + dqc[j]) << 32) | (uint32_t)j)
                                      : ~0ull;
                        v[e] = 0;
                    }
                    wave_bitonic<E>(k, v);
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        const int j = lane * E + e;
                        if (j < kc) {
                            sv[j] = sc_key_float((uint32_t)(k[e] >> 32));
                            si[j] = (int32_t)(k[e] & 0xFFFFFFFFu);
                        }
                    }
                };
                if (a.npow2 <= 64) run(std::integral_constant<int, 1>{});
                else if (a.npow2 <= 128) run(std::integral_constant<int, 2>{});
                else run(std::integral_constant<int, 4>{});
            }
        } else if (sc) {
            for (int j = tid; j < a.npow2; j += NT)
                keys[j] = j < kc ? (((uint64_t)sc_float_key(qsum + dqc[j]) << 32) | (uint32_t)j)
                                 : ~(uint64_t)0;
            __syncthreads();
            for (int size = 2; size <= a.npow2; size <<= 1)
                for (int stride = size >> 1; stride > 0; stride >>= 1) {
                    for (int i = tid; i < a.npow2 / 2; i += NT) {
                        int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                        bool upw = ((lo & size) == 0);
                        uint64_t x = keys[lo], y = keys[hi];
                        if ((x > y) == upw) { keys[lo] = y; keys[hi] = x; }
                    }
                    __syncthreads();
                }
            for (int j = tid; j < kc; j += NT) {
                sv[j] = sc_key_float((uint32_t)(keys[j] >> 32));
                si[j] = (int32_t)(keys[j] & 0xFFFFFFFFu);
            }
        } else {
            for (int j = tid; j < kc; j += NT) { sv[j] = qsum + dqc[j]; si[j] = j; }
        }
        __syncthreads();
        mark(1);
        // traversal: thread 0, heap and row lengths in LDS (region A)
        float *hv = reinterpret_cast<float *>(smem + Lay.a);
        int32_t *rowlen = reinterpret_cast<int32_t *>(hv + 2 * a.hcap);
        int32_t *win = rowlen + kc;    // SC: [kWin][kWin] (start, length) by rank pair
        const int wk = min(kWin, kc);
        if (sc) {
            // all threads: row lengths, and the sizes of the cells of the
            // nearest rank pairs in one round of loads, so the traversal
            // knows them without waiting (cell 0 = (0, 0) is never a pair)
            for (int r = tid; r < kc; r += NT) rowlen[r] = r == 0 ? 1 : 0;
            for (int i = tid; i < wk * wk; i += NT) {
                const int ra = i / wk, rb = i - ra * wk;
                const int64_t b = (int64_t)si[ra] * kc + si[rb];
                const int32_t hi = b > 0 ? sa.L[b] : 0, lo = b > 0 ? sa.L[b - 1] : 0;
                win[2 * i] = lo;
                win[2 * i + 1] = hi - lo;
            }
            __syncthreads();
        }
        if (tid == 0) {
            Heap h;
            h.v = hv;
            h.p = reinterpret_cast<uint32_t *>(hv + a.hcap);
            h.stride = 1;
            h.cap_phys = a.hcap;
            h.pos = 0;
            h.overflow = false;
            int count = 0, count2 = 0, sum = 0, empty = 0, exhausted = 0, flag = 0;
            if (!sc) {
                h.cap_ref = kc;
                for (int j = 0; j < kc; ++j) h.insert(sv[j], (uint32_t)j);
                bool done = h.overflow;
                for (int i0 = 0; i0 < a.w && !done; i0 += kSpec) {
                    const int nsp = min(kSpec, a.w - i0);
                    for (int i = 0; i < nsp; ++i) s_sbid[i] = h.pos > 0 ? (int)h.pop() : -1;
                    int32_t lhi[kSpec], llo[kSpec];
#pragma unroll
                    for (int i = 0; i < kSpec; ++i) {
                        const int64_t b = lane_value(i < nsp ? (int64_t)s_sbid[i] : -1);
                        lhi[i] = b >= 0 ? sa.L[b] : 0;
                        llo[i] = b > 0 ? sa.L[b - 1] : 0;
                    }
#pragma unroll
                    for (int i = 0; i < kSpec; ++i) {
                        if (i >= nsp || done) continue;
                        const int64_t b = s_sbid[i];
                        if (b < 0 || count >= a.ccap) { flag = 1; done = true; continue; }
                        cells[count] = llo[i];
                        clen[count] = lhi[i] - llo[i];
                        cinfo[count] = (uint32_t)b;
                        cpos[count] = sum;
                        count++;
                        count2++;
                        sum += lhi[i] - llo[i];
                        if (sum >= Tb) done = true;
                    }
                }
            } else {
                h.cap_ref = a.w;
                const float d0 = sv[0] + sv[1];
                h.insert(d0, 1u);
                h.insert(d0, 1u << 16);
                // The pops and pushes do not depend on the cell sizes until the
                // loop stops.  Pops of rank pairs inside the window know their
                // cell sizes at once; while every pop of a batch does, the
                // batch ends exactly where the loop stops.  Past the window,
                // pops run ahead in batches of kSpec, their L[] entries load
                // together, and a sequential commit stops exactly where the
                // reference does (pushes past that point are moot).
                bool done = (count >= a.w || sum >= Tb);
                while (!done) {
                    int nsp = 0, kcount = count, ksum = sum;
                    bool dry = false, known = true;
                    while (nsp < kSpec && !h.overflow) {
                        if (h.pos <= 0) { dry = true; break; }
                        const uint32_t pk = h.pop();
                        const int h1 = (int)(pk >> 16), h2 = (int)(pk & 0xFFFFu);
                        const int h3 = si[h1], h4 = si[h2];
                        const bool inw = h1 < wk && h2 < wk;
                        s_sbid[nsp] = inw ? -1 - (int64_t)(h1 * wk + h2) : (int64_t)h3 * kc + h4;
                        s_sinfo[nsp] = (uint32_t)h3 | ((uint32_t)h4 << 16);
                        ++nsp;
                        rowlen[h1] = h2 + 1;
                        if (h1 < kc - 1 && (h2 == 0 || (h2 - 1) < rowlen[h1 + 1]))
                            h.insert(sv[h1 + 1] + sv[h2], ((uint32_t)(h1 + 1) << 16) | (uint32_t)h2);
                        if (h2 < kc - 1 && (h1 == 0 || (h2 + 1) < rowlen[h1 - 1]))
                            h.insert(sv[h1] + sv[h2 + 1], ((uint32_t)h1 << 16) | (uint32_t)(h2 + 1));
                        if (!inw) known = false;
                        if (known) {
                            const int l = win[2 * (h1 * wk + h2) + 1];
                            kcount += l > 0;
                            ksum += l;
                            if (kcount >= a.w || ksum >= Tb) break;    // the loop stops here
                        }
                    }
                    int32_t lhi[kSpec], llo[kSpec];
#pragma unroll
                    for (int i = 0; i < kSpec; ++i) {
                        const int64_t b = lane_value(i < nsp ? s_sbid[i] : 0);
                        if (b < 0) {
                            llo[i] = win[2 * (-1 - b)];
                            lhi[i] = llo[i] + win[2 * (-1 - b) + 1];
                        } else {
                            lhi[i] = b > 0 ? sa.L[b] : 0;
                            llo[i] = b > 0 ? sa.L[b - 1] : 0;
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kSpec; ++i) {
                        if (i >= nsp || done) continue;
                        const int l = lhi[i] - llo[i];
                        if (l > 0) {
                            if (count >= a.ccap) { flag = 1; done = true; continue; }
                            cells[count] = llo[i];
                            clen[count] = l;
                            cinfo[count] = s_sinfo[i];
                            cpos[count] = sum;
                            count++;
                            sum += l;
                        } else {
                            empty++;
                        }
                        count2++;
                        if (count >= a.w || sum >= Tb) done = true;
                        // a traversal this long (mostly empty cells) runs faster in
                        // the staged kernels, which keep more queries in flight
                        else if (count2 >= a.pop_budget) { flag = 1; done = true; }
                    }
                    if (done) break;
                    if (h.overflow) break;                 // flagged below
                    if (dry) { exhausted = 1; break; }
                }
            }
            if (h.overflow) flag = 1;
            s_ncell = count;
            s_sum = sum;
            s_flag = flag;
            int32_t *qi = a.qinfo + (size_t)q * 8;   // a flagged query is counted by its redo
            qi[0] = flag ? 0 : count; qi[1] = flag ? 0 : sum; qi[2] = flag ? 0 : count2;
            qi[3] = flag ? 0 : empty; qi[4] = flag ? 0 : exhausted;
            qi[5] = flag; qi[6] = 0; qi[7] = 0;
            a.flags[q] = flag;
            if (flag) atomicAdd(a.nflag, 1);
        }
        __syncthreads();
        mark(2);
        if (!s_flag) {
            const int ncell = s_ncell, sum = s_sum;
            // group the cells by h3, the groups in order of first appearance
            // in the traversal (the nearest cells come first, so the top-R
            // threshold tightens early and the buffer compacts less often):
            // sizes and first cells; group ranks (wave 0); group offsets
            // (wave 0); permutation; then the pass positions gp of the
            // permuted cells.  The scan order does not change results: every
            // candidate keeps its traversal position.
            int32_t *gfirst = reinterpret_cast<int32_t *>(smem + Lay.a);   // region A is free
            int32_t *gsize = gfirst + kc;                                  // by group rank
            for (int j = tid; j < kc; j += NT) { ghist[j] = 0; gfirst[j] = ncell; }
            __syncthreads();
            for (int c = tid; c < ncell; c += NT) {
                const int h = (int)(cinfo[c] & 0xFFFFu);
                atomicAdd(&ghist[h], 1);
                atomicMin(&gfirst[h], c);
            }
            __syncthreads();
            if (tid < 64) {
                const int per = (ncell + 63) / 64, c0 = min(ncell, tid * per),
                          c1 = min(ncell, c0 + per);
                int nl = 0;
                for (int c = c0; c < c1; ++c) nl += gfirst[cinfo[c] & 0xFFFFu] == c;
                int inc = nl;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(inc, o);
                    if (tid >= o) inc += t;
                }
                int r = inc - nl;
                for (int c = c0; c < c1; ++c) {
                    const int h = (int)(cinfo[c] & 0xFFFFu);
                    if (gfirst[h] == c) { gdesc[r] = (uint32_t)h; gsize[r] = ghist[h]; ++r; }
                }
                if (tid == 63) s_ngroup = inc;
            }
            __syncthreads();
            if (tid < 64) {
                const int ngr = s_ngroup, per = (ngr + 63) / 64, r0 = min(ngr, tid * per),
                          r1 = min(ngr, r0 + per);
                int cnt = 0;
                for (int r = r0; r < r1; ++r) cnt += gsize[r];
                int inc = cnt;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(inc, o);
                    if (tid >= o) inc += t;
                }
                int oc = inc - cnt;
                for (int r = r0; r < r1; ++r) {
                    const uint32_t h = gdesc[r];
                    gdesc[r] = h | ((uint32_t)oc << 16);
                    ghist[h] = oc;
                    oc += gsize[r];
                }
            }
            __syncthreads();
            for (int c = tid; c < ncell; c += NT) {
                const int slot = atomicAdd(&ghist[cinfo[c] & 0xFFFFu], 1);
                order[slot] = (uint16_t)c;
            }
            __syncthreads();
            {
                const int per = (ncell + NT - 1) / NT, i0 = min(ncell, tid * per),
                          i1 = min(ncell, i0 + per);
                int s = 0;
                for (int i = i0; i < i1; ++i) s += clen[order[i]];
                int inc = s;
                for (int o = 1; o < 64; o <<= 1) {
                    const int t = __shfl_up(inc, o);
                    if (lane >= o) inc += t;
                }
                if (lane == 63) T.wsum[wid] = inc;
                if (tid == 0) { gp[ncell] = sum; T.nb = 0; T.thr = ~0ull; }
                __syncthreads();
                int run = inc - s;
                for (int w = 0; w < wid; ++w) run += T.wsum[w];
                for (int i = i0; i < i1; ++i) { gp[i] = run; run += clen[order[i]]; }
            }
            __syncthreads();
            const int ng = s_ngroup;
            const GroupCells G{gp, cells, cpos, order};
            const float *dsrc = dqr_lds ? dqr : a.lut_dqr + (size_t)q * lutn;
            // Mixed form: error bound E of one approximate distance.  With
            // B >= |diff_qr + dot_cr[h3]| for every entry of the query's
            // combined tables (the maximum over its groups' tables, computed
            // with the same FP32 additions as the tables), S = 2^e >= B / 2^14
            // (FP16 range) and M >= |every partial sum|, the FP16 entries are
            // off by <= 2^-11 B (+ S 2^-25 when subnormal), the FP32 sum of mp
            // of them by <= mp^2 2^-24 B, and the exact and approximate sums
            // round by <= (mp + 1) 2^-24 M; E is 1.25x the total.  Every
            // candidate of the exact top R has an approximate distance within
            // 2E of the R-th smallest approximate one.
            float S = 1.0f, twoE = 0.0f;
            if (MIXED) {
                uint32_t mb = 0u;
                for (int g = 0; g < ng; ++g) {
                    const float *row = sa.dot_cr + (size_t)(gdesc[g] & 0xFFFFu) * bs;
                    for (int i = tid; i < lutn; i += NT)
                        mb = max(mb, __float_as_uint(fabsf(dsrc[i] + row[i])));
                }
                for (int o = 32; o > 0; o >>= 1) mb = max(mb, (uint32_t)__shfl_xor(mb, o));
                if (lane == 0) atomicMax(&T.amax_r, mb);
                __syncthreads();
                const float B = __uint_as_float(T.amax_r) * (1.0f + 0x1p-20f);
                int ex = 0;
                if (B > 0.0f) (void)frexpf(B * 0x1p-14f, &ex);
                S = ldexpf(1.0f, ex);
                const float fmp = (float)mp;
                const float M = fabsf(qsum) + __uint_as_float(T.amax_c) + fmp * B;
                const float E = ((fmp + 1.0f) * 0x1p-24f * M + fmp * fmp * 0x1p-24f * B +
                                 fmp * 0x1p-11f * B + fmp * S * 0x1p-25f) * 1.25f * a.margin_scale;
                twoE = 2.0f * E;
            }
            const float invS = 1.0f / S;
            float dq[kLutRegs], rw[kLutRegs];
            if (lut_regs) {
#pragma unroll
                for (int j = 0; j < kLutRegs; ++j) {
                    const int i = tid + j * NT;
                    dq[j] = i < lutn ? dsrc[i] : 0.0f;
                    rw[j] = 0.0f;
                }
                if (ng > 0) {
                    const float *row = sa.dot_cr + (size_t)(gdesc[0] & 0xFFFFu) * bs;
#pragma unroll
                    for (int j = 0; j < kLutRegs; ++j) {
                        const int i = tid + j * NT;
                        if (i < lutn) rw[j] = row[i];
                    }
                }
            }
            for (int g = 0; g < ng; ++g) {
                const uint32_t gd = gdesc[g];
                const int h3 = (int)(gd & 0xFFFFu), i0 = (int)(gd >> 16);
                const int i1 = g + 1 < ng ? (int)(gdesc[g + 1] >> 16) : ncell;
                // the combined table diff_qr + dot_cr[h3]
                _Float16 *lut16 = reinterpret_cast<_Float16 *>(lut);
                if (lut_regs) {
#pragma unroll
                    for (int j = 0; j < kLutRegs; ++j) {
                        const int i = tid + j * NT;
                        if (i < lutn) {
                            const float t = dq[j] + rw[j];
                            if (MIXED) lut16[i] = (_Float16)(t * invS);
                            else lut[i] = t;
                        }
                    }
                } else {
                    const float *row = sa.dot_cr + (size_t)h3 * bs;
                    for (int i = tid; i < lutn; i += NT) {
                        const float t = dsrc[i] + row[i];
                        if (MIXED) lut16[i] = (_Float16)(t * invS);
                        else lut[i] = t;
                    }
                }
                const float d0 = qsum + dqc[h3];
                __syncthreads();
                if (lut_regs && g + 1 < ng) {   // next group's dot_cr row, in flight
                    const float *row = sa.dot_cr + (size_t)(gdesc[g + 1] & 0xFFFFu) * bs;
#pragma unroll
                    for (int j = 0; j < kLutRegs; ++j) {
                        const int i = tid + j * NT;
                        if (i < lutn) rw[j] = row[i];
                    }
                }
                const int p0 = gp[i0], p1 = gp[i1];
                if (mp == 8)
                    scan_group<2, MIXED>(sa, T, bkey, a.bufcap, lut, d0, p0, p1, i0, G, S,
                                         twoE);
                else if (mp == 16)
                    scan_group<4, MIXED>(sa, T, bkey, a.bufcap, lut, d0, p0, p1, i0, G, S,
                                         twoE);
                else
                    scan_group<0, MIXED>(sa, T, bkey, a.bufcap, lut, d0, p0, p1, i0, G, S,
                                         twoE);
                __syncthreads();
            }
            mark(3);
            // final selection: canonical top-R, or traversal order when the
            // query gathered fewer than R candidates
            const int R = sa.R;
            if (MIXED) {
                // keep what the bound cannot exclude, then give every kept
                // candidate its exact distance: q_sum + diff_qc[h3], then
                // + (diff_qr + dot_cr) per sub-quantizer in order, the same
                // operations as the plain form's table and scan
                if (sum >= R && T.nb > R) compact_top_margin(T, bkey, R, twoE, a.bufcap);
                const int m = T.nb;
                if (!T.ovf) {
                    const float *dsrc1 = dqr_lds ? dqr : a.lut_dqr + (size_t)q * lutn;
                    for (int i = tid; i < m; i += NT) {
                        const uint64_t k = bkey[i];
                        const int pos = (int)(uint32_t)k;
                        int lo = 0, hi = ncell;   // last cell with cpos <= pos
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (cpos[mid] <= pos) lo = mid; else hi = mid;
                        }
                        const int h3 = (int)(cinfo[lo] & 0xFFFFu);
                        const int v = min(max(cells[lo] + (pos - cpos[lo]), 0), a.nvec - 1);
                        const uint8_t *cc = sa.codes + (size_t)v * mp;
                        const float *row = sa.dot_cr + (size_t)h3 * bs;
                        float d = qsum + dqc[h3];
                        for (int j = 0; j < mp; ++j) {
                            const int e = j * kp + cc[j];
                            d += dsrc1[e] + row[e];
                        }
                        bkey[i] = ((uint64_t)sc_float_key(d) << 32) | (uint32_t)pos;
                    }
                    if (tid == 0) a.qinfo[(size_t)q * 8 + 7] = m;
                }
                __syncthreads();
            }
            if (MIXED && T.ovf) {
                // the kept set outgrew the buffer: finished exactly afterwards
                if (tid == 0) {
                    int32_t *qi = a.qinfo + (size_t)q * 8;
                    qi[0] = qi[1] = qi[2] = qi[3] = qi[4] = 0;
                    qi[5] = 1;
                    qi[7] = 0;
                    a.flags[q] = 1;
                    atomicAdd(a.nflag, 1);
                }
                __syncthreads();
                mark(4);
                continue;
            }
            int nout;
            if (sum >= R) {
                if (T.nb > R) compact_top(T, bkey, R);
                nout = R;
            } else {
                nout = T.nb;
                for (int i = tid; i < nout; i += NT) {
                    const uint64_t k = bkey[i];
                    bkey[i] = (k << 32) | (k >> 32);
                }
            }
            sort_top(bkey, nout);
            int32_t *ids = sa.ids + (size_t)q * R;
            float *dists = sa.dists + (size_t)q * R;
            for (int r = tid; r < R; r += NT) {
                if (r < nout) {
                    const uint64_t k = bkey[r];
                    // the vector at traversal position pos: last cell with
                    // cpos <= pos, then the offset into it
                    const int pos = (int)(sum >= R ? (uint32_t)k : (uint32_t)(k >> 32));
                    int lo = 0, hi = ncell;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (cpos[mid] <= pos) lo = mid; else hi = mid;
                    }
                    const int v = min(max(cells[lo] + (pos - cpos[lo]), 0), a.nvec - 1);
                    ids[r] = sa.pid[v];
                    dists[r] = sc_key_float(sum >= R ? (uint32_t)(k >> 32) : (uint32_t)k);
                } else {
                    ids[r] = -1;
                    dists[r] = INFINITY;
                }
            }
            if (sa.sums && tid == 0) sa.sums[q] = sum;
        }
        __syncthreads();
        mark(4);
    }
    if (tid == 0 && a.phase) {
        for (int i = 0; i < 5; ++i) atomicAdd(&a.phase[i], (unsigned long long)s_tph[i]);
        atomicAdd(&a.phase[5], 1ull);
    }
}

using FusedKernel = void (*)(FusedArgs);
FusedKernel fused_kernel(int occ, bool mixed) {
    if (mixed)
        return occ >= 5 ? k_search_fused<5, true>
                        : occ == 4 ? k_search_fused<4, true> : k_search_fused<3, true>;
    return occ >= 5 ? k_search_fused<5, false>
                    : occ == 4 ? k_search_fused<4, false> : k_search_fused<3, false>;
}

hipError_t launch_fused(int occ, bool mixed, int grid, int threads, size_t lds, hipStream_t sm,
                        const FusedArgs &fa) {
    FusedArgs arg = fa;
    void *args[] = {&arg};
    return hipLaunchKernel((const void *)fused_kernel(occ, mixed), dim3((unsigned)grid),
                           dim3((unsigned)threads), args, lds, sm);
}

__global__ void k_zero_ints(int *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_gather_rows(const float *__restrict__ Q, int dim, const int32_t *__restrict__ list,
                              int n, float *__restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * dim) return;
    int64_t r = e / dim;
    out[e] = Q[(size_t)list[r] * dim + (e - r * dim)];
}

__global__ void k_list_flags(const int32_t *__restrict__ flags, int64_t nq, int32_t *__restrict__ list,
                             int *__restrict__ n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq && flags[q]) list[atomicAdd(n, 1)] = (int32_t)q;
}

__global__ void k_scatter_results(const int32_t *__restrict__ list, int n, int R,
                                  const int32_t *__restrict__ ids, const float *__restrict__ d,
                                  const int32_t *__restrict__ sums, int32_t *__restrict__ out_ids,
                                  float *__restrict__ out_d, int32_t *__restrict__ out_sums) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * R) return;
    int64_t r = e / R;
    int c = (int)(e - r * R);
    const int64_t q = list[r];
    out_ids[q * R + c] = ids[e];
    out_d[q * R + c] = d[e];
    if (c == 0 && out_sums) out_sums[q] = sums[r];
}

__global__ void k_gather_redo(const int32_t *__restrict__ qinfo, int64_t nq,
                              int32_t *__restrict__ list, int32_t *__restrict__ n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    if (qinfo[q * 8 + 7]) list[atomicAdd(n, 1)] = (int32_t)q;
}
__global__ void k_gather_fallback(const int32_t *__restrict__ qinfo, int64_t nq,
                                   int32_t *__restrict__ list, int32_t *__restrict__ n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    if (qinfo[q * 8 + 5]) list[atomicAdd(n, 1)] = (int32_t)q;
}

// Graph form of the heap fallback: flags[q] = 1 for the queries whose heap
// outgrew the on-chip capacity (finished after the graph), counted in *n.
__global__ void k_flag_fallback(const int32_t *__restrict__ qinfo, int64_t nq,
                                int32_t *__restrict__ flags, int *__restrict__ n) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int f = qinfo[q * 8 + 5] != 0;
    flags[q] = f;
    if (f) atomicAdd(n, 1);
}

// skip_flagged: leave out the queries flagged for the heap fallback (their
// statistics come from the search that finishes them).
__global__ void k_reduce_stats(const int32_t *__restrict__ qinfo, int64_t nq,
                               unsigned long long *__restrict__ out, int skip_flagged = 0) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (q < nq && !(skip_flagged && qinfo[q * 8 + 5])) {
        const int32_t *qi = qinfo + q * 8;
        v[0] = (unsigned long long)qi[1];
        v[1] = (unsigned long long)qi[0];
        v[2] = (unsigned long long)qi[2];
        v[3] = (unsigned long long)qi[4];
        v[4] = (unsigned long long)(qi[6] != 0);
        v[5] = (unsigned long long)qi[5];
        v[6] = (unsigned long long)qi[7];
    }
    for (int i = 0; i < 7; ++i) {
        unsigned long long t = v[i];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(&out[i], t);
    }
}

struct Plan {
    int64_t chunk;      // queries per pass
    int ccap;
    size_t off_qs, off_dqc, off_dqr, off_sv, off_si, off_cells, off_clen, off_cinfo, off_qinfo,
        off_list, off_skip, off_stats, off_okey, off_okey2, off_oid, off_oid2, off_perm, off_otmp,
        otmp_bytes, total, off_big_list, off_big_sorted, off_big_size, off_big_aux, off_big_split;
    int big_cap;        // entries per query of the pruned ranking (0: not used)
    int retry_cap;      // queries the pruned ranking can re-rank per pass (big_cap == 0)
};

Plan make_plan(const sc_index *ix, const sc_search_params *p, int64_t nq) {
    Plan P{};
    const int lists = p->kind == SC_KIND_IMI ? 2 : 1;
    int64_t ccap = p->kind == SC_KIND_IVFADC ? p->w : std::min<int64_t>(p->w, (int64_t)p->T + 1);
    if (ccap < 1) ccap = 1;
    P.ccap = (int)ccap;
    // the pruned ranking of large two-halves indexes: its per-query lists
    const int big = p->kind == SC_KIND_SCIMI ? sc_rank_big_cap(ix) : 0;
    P.big_cap = big;
    const size_t per = 16 + 4ull * ix->mc * ix->kc + 4ull * ix->mp * ix->kp +
                       8ull * lists * ix->kc + 12ull * ccap + 32 + 4 + 20ull * big + 16;
    // up to 4 GiB of per-query scratch (cell lists sized for the budget T),
    // so that a batch of ten thousand queries runs as one pass (8 GiB with
    // the pruned ranking's lists)
    int64_t chunk = (int64_t)((size_t)(big ? 8 : 4) << 30) / (int64_t)per;
    chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, std::max<int64_t>(nq, 1)));
    chunk = std::min<int64_t>(chunk, 65536);
    if (const char *e = getenv("SC_SEARCH_CHUNK")) {   // test hook: smaller passes
        const long long v = atoll(e);
        if (v >= 1) chunk = std::min<int64_t>(chunk, v);
    }
    P.chunk = chunk;
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t r = o; o = (o + bytes + 255) & ~(size_t)255; return r; };
    P.off_qs = take(16ull * chunk);
    P.off_dqc = take(4ull * ix->mc * ix->kc * chunk);
    P.off_dqr = take(4ull * ix->mp * ix->kp * chunk);
    P.off_sv = take(4ull * lists * ix->kc * chunk);
    P.off_si = take(4ull * lists * ix->kc * chunk);
    P.off_cells = take(4ull * ccap * chunk);
    P.off_clen = take(4ull * ccap * chunk);
    P.off_cinfo = take(4ull * ccap * chunk);
    P.off_qinfo = take(32ull * chunk);
    P.off_list = take(8ull * chunk + 16);   // two query lists, then counters
    P.off_skip = take(4ull * chunk);
    P.off_stats = take(8ull * 8);
    // scan order: (first cell, query) pairs sorted, then the permutation
    P.off_okey = take(4ull * chunk);
    P.off_okey2 = take(4ull * chunk);
    P.off_oid = take(4ull * chunk);
    P.off_oid2 = take(4ull * chunk);
    P.off_perm = take(4ull * chunk);
    // without the pruned ranking, scratch for it to re-rank up to 512 queries
    // of a pass whose cells overflow the pair-table ranking
    const int rcap = p->kind == SC_KIND_SCIMI && !big ? sc_rank_retry_cap(ix) : 0;
    P.retry_cap = rcap ? (int)std::min<int64_t>(chunk, 512) : 0;
    const size_t bigq = big ? (size_t)big * chunk : (size_t)rcap * P.retry_cap;
    P.off_big_list = take(8ull * bigq);
    P.off_big_sorted = take(8ull * bigq);
    P.off_big_size = take(4ull * bigq);
    P.off_big_aux = take(16ull * chunk);
    // split-query state for batches of a few queries (4 + 2 x 4096 words each)
    P.off_big_split = take(big && chunk <= 4 ? 4ull * (4 + 2 * 4096) * chunk : 0);
    size_t tb = 0;
    (void)rocprim::radix_sort_pairs(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (int32_t *)nullptr, (int32_t *)nullptr, (size_t)chunk, 0u,
                                    32u, (hipStream_t)0);
    P.otmp_bytes = tb;
    P.off_otmp = take(tb);
    P.total = o;
    return P;
}

// The staged scan's workgroup order.  Each query's first cell (its nearest
// coarse centroid, or pair of them) is its key; queries sorted by key share
// table rows, and the XCD-aware permutation gives each XCD (workgroup i runs
// on XCD i mod nx) a contiguous range of the sorted queries, so that the rows
// they share stay in that XCD's L2.  Results do not depend on the order.
__global__ void k_scan_keys(const uint32_t *__restrict__ cinfo, const int32_t *__restrict__ qinfo,
                            int64_t nq, int ccap, uint32_t *__restrict__ keys,
                            int32_t *__restrict__ ids) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    keys[q] = qinfo[q * 8] > 0 ? cinfo[(size_t)q * ccap] : 0u;
    ids[q] = (int32_t)q;
}
// Before the one-workgroup traversal: a query's first cell is the pair of
// its rank-0 centroids (IMI: one per list; SC: nearest and second nearest).
__global__ void k_trav_keys(const int32_t *__restrict__ si, int64_t nq, int lists, int kc,
                            uint32_t *__restrict__ keys, int32_t *__restrict__ ids) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int32_t *r = si + (size_t)q * lists * kc;
    const int32_t b = lists > 1 ? r[kc] : (kc > 1 ? r[1] : 0);
    keys[q] = (uint32_t)r[0] * (uint32_t)kc + (uint32_t)b;
    ids[q] = (int32_t)q;
}
__global__ void k_xcd_order(const int32_t *__restrict__ sorted, int64_t n, int nx,
                            int32_t *__restrict__ perm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t x = i % nx, j = i / nx;
    int64_t off = 0;                                  // workgroups on the XCDs before x
    for (int64_t y = 0; y < x; ++y) off += (n - y + nx - 1) / nx;
    perm[i] = sorted[off + j];
}

// SC_IMI_SCAN (tuning hook): the IMI scan variant (see ScanArgs::imi_mode).
static int imi_scan_mode() {
    if (const char *e = getenv("SC_IMI_SCAN")) return atoi(e);
    return 1;
}

// SC_SCAN_ORDER (tuning hook): 0 launch order, 1 sorted, 2 sorted per XCD.
// Default: sorted per XCD for the dense kinds, whose plain scan gathers
// table rows that the queries of one XCD share; sorted for the two-halves
// kind (measured on the 1M two-halves index: 1 is 2-4% faster than 2 in
// both forms, and 2 is 7% faster than 1 for the plain K=1024 IMI scan).
static int scan_order_mode(int kind) {
    if (const char *e = getenv("SC_SCAN_ORDER")) return atoi(e);
    return kind == SC_KIND_SCIMI ? 1 : 2;
}
// SC_TRAV_ORDER (tuning hook): 1 (default) orders the one-workgroup traversal
// as well, by the queries' rank-0 centroids; 0 orders only the scan.
static int trav_order_mode() {
    if (const char *e = getenv("SC_TRAV_ORDER")) return atoi(e);
    return 1;
}

}  // namespace

// The staged search as a plan captures it in a graph: the plan's own scratch
// (make_plan's layout), no host synchronisation and no stage events; the
// queries whose heap outgrows the on-chip capacity are flagged (flags [nq],
// counted in *nflag) for the plan to finish after the graph.
struct StagedGraph {
    unsigned char *scratch;
    int32_t *flags;
    int *nflag;
    unsigned long long *stats;   // [8] statistics of the unflagged queries
};

static int search_staged(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                         const sc_search_params *p, const int32_t *d_T, int32_t *d_ids,
                         float *d_dists, int32_t *d_sums, sc_search_stats *st,
                         const StagedGraph *g = nullptr);
static int finish_flagged(sc_ctx *ctx, const sc_index *ix, const sc_search_params *p,
                          const float *d_q, const int32_t *d_T, int64_t nq, const int32_t *flags,
                          int nflag, int32_t *d_ids, float *d_dists, int32_t *d_sums,
                          sc_search_stats *st);

// SC traversals that pop this many cells leave the fused kernel for the
// staged kernels (SC_POP_BUDGET overrides it; a test hook).  Typical queries
// pop tens of cells (at most ~80 on the 1M x 128, K=256 index up to 15k
// candidates); a query that must walk far through empty cells pops
// thousands, and the staged traversal keeps about six times more of them in
// flight.
static int fused_pop_budget() {
    if (const char *e = getenv("SC_POP_BUDGET")) {
        int v = atoi(e);
        if (v >= 1) return v;
    }
    return 128;
}

// Mixed form: factor on its error bound (test hook; 1 otherwise).
static float mixed_margin_scale() {
    if (const char *e = getenv("SC_MIXED_MARGIN")) {
        const float v = (float)atof(e);
        if (v >= 1.0f) return v;
    }
    return 1.0f;
}

// On-chip capacities and LDS layout of the fused kernel for one search.
struct FusedConfig {
    int hcap, ccap, npow2, bufcap;
    bool tiled_lut;     // LUTs precomputed by k_lut_tiled (else in the kernel)
    FusedLayout lay;
    int occ;            // workgroups per CU that the LDS allows
    int variant;        // kernel register variant (waves per SIMD: 3, 4, 5)
};

static FusedConfig fused_config(const sc_index *ix, const sc_search_params *p) {
    FusedConfig f{};
    const int kind = p->kind, kc = ix->kc;
    f.npow2 = 1;
    while (f.npow2 < kc) f.npow2 <<= 1;
    f.hcap = std::min(kind == SC_KIND_IVFADC ? kc : p->w, 1024);
    if (const char *e = getenv("SC_HEAP_ONCHIP")) {   // test hook: force the heap fallback
        int v = atoi(e);
        if (v >= 1) f.hcap = std::min(f.hcap, v);
    }
    const int64_t ccap_ref =
        kind == SC_KIND_IVFADC ? p->w : std::min<int64_t>(p->w, (int64_t)p->T + 1);
    // Cells a query gathers before reaching T: sized from the mean non-empty
    // cell with a wide margin (queries that need more are finished by the
    // staged kernels); smaller lists leave LDS for more workgroups per CU.
    int64_t est = 1024;
    if (kind == SC_KIND_SC && ix->nonempty > 0 && ix->n > 0) {
        const double mean = (double)ix->n / (double)ix->nonempty;
        est = std::max<int64_t>(128, (int64_t)(8.0 * (double)p->T / mean) + 64);
    }
    if (const char *e = getenv("SC_FUSED_CELLS")) {   // test hook: cell list capacity
        int v = atoi(e);
        if (v >= 1) est = v;
    }
    f.ccap = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(ccap_ref, est), 1024));
    // the buffer takes one scan round (kRound) above the R it keeps; at most
    // 8 entries per thread of a 256-thread workgroup (R <= 1024).  The mixed
    // form keeps every candidate within its error margin, up to 5R + 512
    // (3R + 1024 when smaller) before a query is handed to the exact staged
    // kernels.  512 entries of slack above what a compaction keeps let most
    // queries compact once or twice; at R = 100 both forms leave LDS for five
    // workgroups per CU (8-byte keys; the FP16 table takes half the FP32
    // one's room).
    const int keep = p->form == SC_FORM_MIXED ? std::min(5 * p->R + 512, 3 * p->R + 1024)
                                              : p->R + 512;
    f.bufcap = (keep + kRound + 63) & ~63;
    f.tiled_lut = kc % kLutC == 0 && ix->kp % kLutC == 0;
    f.lay = fused_layout(ix->dim, ix->mc, kc, ix->mp, ix->kp, f.hcap, f.ccap, f.npow2, f.bufcap,
                         !f.tiled_lut, p->form == SC_FORM_MIXED);
    // 2 KiB covers the kernel's static LDS
    f.occ = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / (f.lay.total + 2048)));
    f.variant = std::max(3, std::min(f.occ, 5));
    return f;
}

// Fused persistent path for IVF/SC.  Returns SC_ERR_TOO_LARGE when its LDS
// layout does not fit (the caller then uses the staged kernels).
static int search_fused(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                        const sc_search_params *p, const int32_t *d_T, int32_t *d_ids,
                        float *d_dists, int32_t *d_sums, sc_search_stats *st) {
    const int kind = p->kind, kc = ix->kc;
    const FusedConfig F = fused_config(ix, p);
    const FusedLayout &Lay = F.lay;
    const int hcap = F.hcap, ccap = F.ccap, npow2 = F.npow2;
    if (Lay.total > 150 * 1024) return SC_ERR_TOO_LARGE;

    const bool tiled_lut = F.tiled_lut;
    const size_t lut_per = (size_t)ix->mc * kc + (size_t)ix->mp * ix->kp;
    size_t o_cnt = 0, o_flags = 256, o_qinfo = o_flags + ((4 * (size_t)nq + 255) & ~(size_t)255),
           o_stats = o_qinfo + 32 * (size_t)nq, o_lut = o_stats + 256,   // stats: 8 + 6 u64
           total = o_lut + (tiled_lut ? 4 * lut_per * (size_t)nq : 0);
    unsigned char *s;
    SC_TRY(sc_scratch(ctx, total, (void **)&s));
    int *counter = (int *)(s + o_cnt);            // [0] work counter, [1] flagged
    int32_t *flags = (int32_t *)(s + o_flags);
    int32_t *qinfo = (int32_t *)(s + o_qinfo);
    unsigned long long *stats = (unsigned long long *)(s + o_stats);
    hipStream_t sm = ctx->stream;

    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, counter, 2);
    SC_CHECK_LAUNCH();
    const bool mixed = p->form == SC_FORM_MIXED;
    SC_HIP(hipFuncSetAttribute((const void *)fused_kernel(F.variant, mixed),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)Lay.total));
    int64_t grid = (int64_t)ctx->n_cu * std::min(F.occ, F.variant);
    if (const char *e = getenv("SC_FUSED_GRID")) {   // tuning hook: workgroups in flight
        int v = atoi(e);
        if (v >= 1) grid = v;
    }
    grid = std::min<int64_t>(grid, nq);
    FusedArgs fa;
    fa.s = ScanArgs{kind,      ix->dim,   ix->mc,  kc,      ix->mp,  ix->kp, p->R,
                    ccap,      nullptr,   nullptr, nullptr, ix->dot_cr, ix->L, ix->pid,
                    ix->codes, nullptr,   nullptr, nullptr, d_ids,   d_dists, d_sums, nullptr};
    fa.Q = d_q; fa.cq = ix->cqT; fa.pq = ix->pqT; fa.norm_c = ix->norm_c; fa.norm_r = ix->norm_r;
    fa.nq = nq; fa.w = p->w; fa.T = p->T; fa.hcap = hcap; fa.ccap = ccap; fa.npow2 = npow2;
    fa.bufcap = F.bufcap;
    fa.counter = counter; fa.flags = flags; fa.nflag = counter + 1; fa.qinfo = qinfo;
    fa.lut_dqc = fa.lut_dqr = nullptr;
    fa.phase = st ? stats + 8 : nullptr;     // 6 counters after the 8 statistics
    fa.margin_scale = mixed_margin_scale();
    fa.nvec = (int)ix->n;
    fa.pop_budget = fused_pop_budget();
    fa.Tq = d_T;
    if (st) SC_HIP(hipMemsetAsync(stats + 8, 0, 48, sm));
    hipEvent_t e0 = nullptr, e1 = nullptr, el = nullptr;
    if (st) {
        SC_HIP(hipEventCreate(&e0));
        SC_HIP(hipEventCreate(&e1));
        SC_HIP(hipEventCreate(&el));
        SC_HIP(hipEventRecord(el, sm));
    }
    if (tiled_lut) {
        float *lq = (float *)(s + o_lut);
        fa.lut_dqc = lq;
        fa.lut_dqr = lq + (size_t)nq * ix->mc * kc;
        hipLaunchKernelGGL(k_lut_tiled,
                           dim3((unsigned)sc_cdiv(nq, kLutQ), (unsigned)(lut_per / kLutC)),
                           dim3(256), 0, sm, d_q, nq, ix->dim, ix->cqT, ix->pqT, ix->norm_c,
                           ix->norm_r, ix->mc, kc, ix->mp, ix->kp, lq,
                           lq + (size_t)nq * ix->mc * kc);
        SC_CHECK_LAUNCH();
    }
    if (st) SC_HIP(hipEventRecord(e0, sm));
    SC_HIP(launch_fused(F.variant, mixed, (int)grid, kScanThreads, Lay.total, sm, fa));
    if (st) {
        SC_HIP(hipEventRecord(e1, sm));
        SC_HIP(hipMemsetAsync(stats, 0, 64, sm));
        hipLaunchKernelGGL(k_reduce_stats, dim3((unsigned)sc_cdiv(nq, 256)), dim3(256), 0, sm,
                           qinfo, nq, stats);
        SC_CHECK_LAUNCH();
    }
    int nflag = 0;
    unsigned long long h[14] = {0};
    SC_HIP(hipMemcpyAsync(&nflag, counter + 1, 4, hipMemcpyDeviceToHost, sm));
    if (st) SC_HIP(hipMemcpyAsync(h, stats, 112, hipMemcpyDeviceToHost, sm));
    SC_HIP(hipStreamSynchronize(sm));
    if (st) {
        const double wgs = h[13] ? (double)h[13] : 1.0;
        for (int i = 0; i < 5; ++i) st->fused_phase_ms[i] = (double)h[8 + i] / wgs / 1e5;
        float ms = 0.0f, ml = 0.0f;
        SC_HIP(hipEventElapsedTime(&ms, e0, e1));
        SC_HIP(hipEventElapsedTime(&ml, el, e0));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(el);
        st->queries = nq;
        st->candidates = (int64_t)h[0];
        st->cells = (int64_t)h[1];
        st->visited = (int64_t)h[2];
        st->exhausted = (int64_t)h[3];
        st->reranked = (int64_t)h[6];
        st->ms_lut = ml;
        st->ms_scan = ms;
    }
    return finish_flagged(ctx, ix, p, d_q, d_T, nq, flags, nflag, d_ids, d_dists, d_sums, st);
}

// Finishes the queries the fused kernel flagged (on-chip capacity exceeded)
// with the staged kernels and scatters their results into place.
static int finish_flagged(sc_ctx *ctx, const sc_index *ix, const sc_search_params *p,
                          const float *d_q, const int32_t *d_T, int64_t nq, const int32_t *flags,
                          int nflag, int32_t *d_ids, float *d_dists, int32_t *d_sums,
                          sc_search_stats *st) {
    hipStream_t sm = ctx->stream;
    if (nflag > 0) {
        // finish the flagged queries with the staged kernels
        int32_t *list = nullptr, *sub_ids = nullptr, *sub_sums = nullptr, *sub_T = nullptr;
        float *sub_q = nullptr, *sub_d = nullptr;
        int *dn = nullptr;
        bool ok = hipMalloc(&list, 4 * (size_t)nflag) == hipSuccess &&
                  hipMalloc(&dn, 4) == hipSuccess &&
                  hipMalloc(&sub_q, 4 * (size_t)nflag * ix->dim) == hipSuccess &&
                  hipMalloc(&sub_ids, 4 * (size_t)nflag * p->R) == hipSuccess &&
                  hipMalloc(&sub_d, 4 * (size_t)nflag * p->R) == hipSuccess &&
                  hipMalloc(&sub_sums, 4 * (size_t)nflag) == hipSuccess &&
                  (!d_T || hipMalloc(&sub_T, 4 * (size_t)nflag) == hipSuccess);
        int rc = ok ? SC_OK : SC_ERR_NOMEM;
        if (ok) {
            hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, dn, 1);
            hipLaunchKernelGGL(k_list_flags, dim3((unsigned)sc_cdiv(nq, 256)), dim3(256), 0, sm,
                               flags, nq, list, dn);
            hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)sc_cdiv((int64_t)nflag * ix->dim, 256)),
                               dim3(256), 0, sm, d_q, ix->dim, list, nflag, sub_q);
            if (d_T)   // the budgets travel with their queries (4-byte words, copied as is)
                hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)sc_cdiv(nflag, 256)), dim3(256),
                                   0, sm, reinterpret_cast<const float *>(d_T), 1, list, nflag,
                                   reinterpret_cast<float *>(sub_T));
            rc = hipGetLastError() == hipSuccess ? SC_OK : SC_ERR_HIP;
        }
        sc_search_stats sub{};
        if (rc == SC_OK)
            rc = search_staged(ctx, ix, sub_q, nflag, p, sub_T, sub_ids, sub_d, sub_sums,
                               st ? &sub : nullptr);
        if (rc == SC_OK) {
            hipLaunchKernelGGL(k_scatter_results, dim3((unsigned)sc_cdiv((int64_t)nflag * p->R, 256)),
                               dim3(256), 0, sm, list, nflag, p->R, sub_ids, sub_d, sub_sums,
                               d_ids, d_dists, d_sums);
            rc = hipGetLastError() == hipSuccess ? SC_OK : SC_ERR_HIP;
            (void)hipStreamSynchronize(sm);
        }
        for (void *b : {(void *)list, (void *)dn, (void *)sub_q, (void *)sub_ids, (void *)sub_d,
                        (void *)sub_sums, (void *)sub_T})
            if (b) (void)hipFree(b);
        if (rc != SC_OK) {
            if (rc == SC_ERR_NOMEM) sc_set_error("sc_search: fallback allocation failed");
            return rc;
        }
        if (st) {
            st->candidates += sub.candidates;
            st->cells += sub.cells;
            st->visited += sub.visited;
            st->exhausted += sub.exhausted;
            st->heap_fallback = nflag;
            st->ms_traverse = sub.ms_lut + sub.ms_rank + sub.ms_traverse + sub.ms_scan;
        }
    }
    return SC_OK;
}

static int search_impl(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                       const sc_search_params *p, const int32_t *d_T, int32_t *d_ids,
                       float *d_dists, int32_t *d_sums, sc_search_stats *st) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!ix || !p || (nq > 0 && (!d_q || !d_ids || !d_dists)) || nq < 0) {
        sc_set_error("sc_search: bad arguments");
        return SC_ERR_ARG;
    }
    if (p->R < 1 || p->R > kBufCap - kScanRound || p->w < 1 || p->T < 0 ||
        (p->form != SC_FORM_PLAIN && p->form != SC_FORM_MIXED && p->form != SC_FORM_REAL)) {
        sc_set_error("sc_search: bad parameters (R=%d w=%d T=%d form=%d)", p->R, p->w, p->T,
                     p->form);
        return SC_ERR_ARG;
    }
    const int kind = p->kind;
    if ((kind == SC_KIND_IVFADC && (ix->mc != 1 || ix->nc != 1)) ||
        (kind == SC_KIND_IMI && (ix->mc != 2 || ix->nc != 1)) ||
        (kind == SC_KIND_SC && (ix->mc != 1 || ix->nc != 2)) ||
        (kind == SC_KIND_SC3 && (ix->mc != 1 || ix->nc != 3 || ix->kc > 1024)) ||
        (kind == SC_KIND_SCIMI && !ix->sparse) || kind < 0 || kind > 4) {
        sc_set_error("sc_search: kind %d does not match the index (mc=%d nc=%d)", kind, ix->mc,
                     ix->nc);
        return SC_ERR_CONFIG;
    }
    if (kind == SC_KIND_IVFADC && p->w > ix->kc) {
        sc_set_error("sc_search: w=%d > kc=%d pops an empty heap (include/sc_utilities.h:274)",
                     p->w, ix->kc);
        return SC_ERR_HEAP_EMPTY;
    }
    if ((kind == SC_KIND_SC && ix->kc < 2) || (kind == SC_KIND_SC3 && ix->kc < 3)) {
        sc_set_error("sc_search: kind %d needs kc >= %d candidates per list (kc=%d)", kind,
                     kind == SC_KIND_SC3 ? 3 : 2, ix->kc);
        return SC_ERR_ARG;
    }
    if (p->form == SC_FORM_MIXED && kind == SC_KIND_SC3) {
        sc_set_error("sc_search: SC_FORM_MIXED is not available for kind %d; "
                     "use SC_FORM_PLAIN", kind);
        return SC_ERR_ARG;
    }
    if (ix->kc > 8192 || ix->mp * ix->kp > 8192) {
        sc_set_error("sc_search: kc=%d / mp*kp=%d exceed the on-chip ranking limits", ix->kc,
                     ix->mp * ix->kp);
        return SC_ERR_ARG;
    }
    if (p->form == SC_FORM_REAL && !ix->raw) {
        sc_set_error("sc_search: SC_FORM_REAL needs the raw vectors (sc_index_attach_raw)");
        return SC_ERR_ARG;
    }
    if (st) *st = sc_search_stats{};
    if (nq == 0) return SC_OK;
    // real distances run on the staged kernels: the raw vectors are read per
    // candidate, so the scan is bound by those reads, not by the on-chip stages
    if (kind != SC_KIND_IMI && kind != SC_KIND_SC3 && kind != SC_KIND_SCIMI &&
        p->form != SC_FORM_REAL && !getenv("SC_NO_FUSED")) {
        int rc = search_fused(ctx, ix, d_q, nq, p, d_T, d_ids, d_dists, d_sums, st);
        if (rc != SC_ERR_TOO_LARGE) return rc;      // otherwise: staged kernels
        if (st) *st = sc_search_stats{};
    }
    return search_staged(ctx, ix, d_q, nq, p, d_T, d_ids, d_dists, d_sums, st);
}

extern "C" int sc_search(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                         const sc_search_params *p, int32_t *d_ids, float *d_dists,
                         int32_t *d_sums, sc_search_stats *st) {
    return search_impl(ctx, ix, d_q, nq, p, nullptr, d_ids, d_dists, d_sums, st);
}

extern "C" int sc_search_budgets(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                                 const sc_search_params *p, const int32_t *d_T, int32_t *d_ids,
                                 float *d_dists, int32_t *d_sums, sc_search_stats *st) {
    if (nq > 0 && !d_T) {
        sc_set_error("sc_search_budgets: no budgets");
        return SC_ERR_ARG;
    }
    return search_impl(ctx, ix, d_q, nq, p, d_T, d_ids, d_dists, d_sums, st);
}

// The three-candidate traversal of a chunk's cn queries, in groups whose
// global heaps (the reference's capacity w) and depth maps fit 1 GiB.
static int traverse3(sc_ctx *ctx, const sc_search_params *p, int kc, const float *sv,
                     const int32_t *si, const int32_t *L, int32_t *cells, int32_t *clen,
                     uint32_t *cinfo, int32_t *qinfo, int ccap, int64_t cn, const int32_t *Tq) {
    const size_t per = (size_t)p->w * 8 + (size_t)kc * kc * 2;
    const int64_t group =
        std::max<int64_t>(1, std::min<int64_t>(cn, (int64_t)(((size_t)1 << 30) / per)));
    void *g = nullptr;
    if (hipMalloc(&g, per * (size_t)group + 256) != hipSuccess) {
        sc_set_error("sc_search: traversal heaps (%zu bytes) could not be allocated",
                     per * (size_t)group);
        return SC_ERR_NOMEM;
    }
    float *hv = (float *)g;
    uint32_t *hp = (uint32_t *)(hv + (size_t)group * p->w);
    uint16_t *depth = (uint16_t *)(hp + (size_t)group * p->w);
    int rc = SC_OK;
    for (int64_t g0 = 0; g0 < cn && rc == SC_OK; g0 += group) {
        const int64_t gn = std::min<int64_t>(group, cn - g0);
        if (hipMemsetAsync(depth, 0, (size_t)gn * kc * kc * 2, ctx->stream) != hipSuccess) {
            rc = SC_ERR_HIP;
            break;
        }
        Trav3Args ta{kc, p->w, p->T, ccap, sv, si, L, cells, clen, cinfo, qinfo, g0, gn, hv, hp,
                     depth, Tq};
        hipLaunchKernelGGL(k_traverse3, dim3((unsigned)sc_cdiv(gn, 64)), dim3(64), 0, ctx->stream,
                           ta);
        if (hipGetLastError() != hipSuccess) rc = SC_ERR_HIP;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(g);
    if (rc != SC_OK) sc_set_error("sc_search: three-candidate traversal launch failed");
    return rc;
}

static int search_staged(sc_ctx *ctx, const sc_index *ix, const float *d_q, int64_t nq,
                         const sc_search_params *p, const int32_t *d_T, int32_t *d_ids,
                         float *d_dists, int32_t *d_sums, sc_search_stats *st,
                         const StagedGraph *g) {
    const int kind = p->kind;
    const bool three = kind == SC_KIND_SC3;
    const bool halves = kind == SC_KIND_SCIMI;
    // SC3 ranks and scans as SC; the two-halves kind scans as the IMI
    const int scan_kind = three ? SC_KIND_SC : (halves ? SC_KIND_IMI : kind);
    const Plan P = make_plan(ix, p, nq);
    unsigned char *s;
    if (g) {
        if (three || halves) {
            sc_set_error("sc_search: kind %d has no graph form", kind);
            return SC_ERR_ARG;
        }
        s = g->scratch;
        st = nullptr;
    } else {
        SC_TRY(sc_scratch(ctx, P.total, (void **)&s));
    }
    float *qs = (float *)(s + P.off_qs), *dqc = (float *)(s + P.off_dqc),
          *dqr = (float *)(s + P.off_dqr), *sv = (float *)(s + P.off_sv);
    int32_t *si = (int32_t *)(s + P.off_si), *cells = (int32_t *)(s + P.off_cells);
    int32_t *clen = (int32_t *)(s + P.off_clen);
    uint32_t *cinfo = (uint32_t *)(s + P.off_cinfo);
    int32_t *qinfo = (int32_t *)(s + P.off_qinfo), *list = (int32_t *)(s + P.off_list);
    int32_t *list2 = list + P.chunk, *counters = list + 2 * P.chunk;   // [P.chunk] each, then 4
    unsigned long long *stats = (unsigned long long *)(s + P.off_stats);
    if (st) SC_HIP(hipMemsetAsync(stats, 0, 64, ctx->stream));

    const int kc = ix->kc, lists = kind == SC_KIND_IMI ? 2 : 1;
    int npow2 = 1;
    while (npow2 < kc) npow2 <<= 1;
    // traversal launch shape: heap + row lengths per lane in LDS
    const int hcap_ref = kind == SC_KIND_IVFADC ? kc : p->w;
    int hcap = std::min(hcap_ref, 256);
    if (const char *e = getenv("SC_HEAP_ONCHIP")) {   // test hook: smaller on-chip heap
        int v = atoi(e);
        if (v >= 1) hcap = std::min(hcap, v);
    }
    const size_t lane_bytes = (size_t)hcap * 8 + (size_t)kc * 4;
    int qpw = (int)std::min<size_t>(64, (48 * 1024) / lane_bytes);
    if (qpw < 1) { qpw = 1; hcap = 0; }   // row lengths alone do not fit: global heap
    // IMI and SC: one workgroup per query (the first 256 ranks of each list,
    // their row lengths and a heap of up to 256 entries in LDS; a query that
    // needs more is redone by the global-heap lane kernel)
    int hcap_w = 0;
    const int kr = std::min(kc, kTravRanks);
    if (kind != SC_KIND_IVFADC && !three && !halves && kc <= 65535) {
        hcap_w = std::min(hcap_ref, 256);
        if (const char *e = getenv("SC_HEAP_ONCHIP")) {
            int v = atoi(e);
            if (v >= 1) hcap_w = std::min(hcap_w, v);
        }
    }
    const bool wave = hcap_w > 0;
    const int hcap_eff = wave ? hcap_w : hcap;     // on-chip heap of the traversal used
    if (g && hcap_eff == 0 && !three && !halves) {
        sc_set_error("sc_search: no on-chip heap fits; no graph form");
        return SC_ERR_TOO_LARGE;
    }
    int64_t fallback_total = 0;
    // stage timing: five events per chunk on the context stream, and a pair
    // around each chunk's scan launches
    std::vector<hipEvent_t> ev, kev;
    auto mark = [&]() {
        if (!st) return;
        hipEvent_t e;
        if (hipEventCreate(&e) == hipSuccess) {
            (void)hipEventRecord(e, ctx->stream);
            ev.push_back(e);
        }
    };

    for (int64_t q0 = 0; q0 < nq; q0 += P.chunk) {
        const int64_t cn = std::min<int64_t>(P.chunk, nq - q0);
        const float *Q = d_q + (size_t)q0 * ix->dim;
        const int64_t lut_total = cn * (ix->mc * kc + ix->mp * ix->kp);
        mark();
        if (kc % kLutC == 0 && ix->kp % kLutC == 0)   // same values, tiled
            hipLaunchKernelGGL(k_lut_tiled,
                               dim3((unsigned)sc_cdiv(cn, kLutQ),
                                    (unsigned)((ix->mc * kc + ix->mp * ix->kp) / kLutC)),
                               dim3(256), 0, ctx->stream, Q, cn, ix->dim, ix->cqT, ix->pqT,
                               ix->norm_c, ix->norm_r, ix->mc, kc, ix->mp, ix->kp, dqc, dqr);
        else
            hipLaunchKernelGGL(k_lut, dim3((unsigned)sc_cdiv(lut_total, 256)), dim3(256), 0,
                               ctx->stream, Q, cn, ix->dim, ix->cqT, ix->pqT, ix->norm_c,
                               ix->norm_r, ix->mc, kc, ix->mp, ix->kp, dqc, dqr);
        SC_CHECK_LAUNCH();
        mark();
        if (!halves) {      // the two-halves ranking computes its own query terms
            hipLaunchKernelGGL(k_rank, dim3((unsigned)cn), dim3(256), (size_t)npow2 * 8,
                               ctx->stream, Q, ix->dim, scan_kind, ix->mc, kc, dqc, qs, sv, si,
                               npow2);
            SC_CHECK_LAUNCH();
        }
        mark();
        const int32_t *Tq = d_T ? d_T + q0 : nullptr;      // this chunk's budgets
        TravArgs ta{kind, kc, p->w, p->T, P.ccap, sv, si, ix->L, cells, cinfo, qinfo,
                    nullptr, cn, nullptr, nullptr, nullptr, clen, Tq};
        // workgroup order of the traversal and the scan (results do not depend
        // on it); the graph form keeps query order (no sort in the graph)
        const int order = g ? 0 : scan_order_mode(kind);
        int32_t *perm = (int32_t *)(s + P.off_perm);
        bool ordered = false;
        auto sort_order = [&](bool from_ranks) -> int {
            uint32_t *okey = (uint32_t *)(s + P.off_okey), *okey2 = (uint32_t *)(s + P.off_okey2);
            int32_t *oid = (int32_t *)(s + P.off_oid), *oid2 = (int32_t *)(s + P.off_oid2);
            if (from_ranks)
                hipLaunchKernelGGL(k_trav_keys, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256), 0,
                                   ctx->stream, si, cn, lists, kc, okey, oid);
            else
                hipLaunchKernelGGL(k_scan_keys, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256), 0,
                                   ctx->stream, cinfo, qinfo, cn, P.ccap, okey, oid);
            SC_CHECK_LAUNCH();
            size_t tb = P.otmp_bytes;
            SC_HIP(rocprim::radix_sort_pairs(s + P.off_otmp, tb, okey, okey2, oid, oid2,
                                             (size_t)cn, 0u, 32u, ctx->stream));
            const int nx = order == 2 && ctx->n_cu % 32 == 0 ? ctx->n_cu / 32 : 1;
            hipLaunchKernelGGL(k_xcd_order, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256), 0,
                               ctx->stream, oid2, cn, nx, perm);
            SC_CHECK_LAUNCH();
            ordered = true;
            return SC_OK;
        };
        if (order > 0 && cn > 1 && wave && trav_order_mode()) {
            SC_TRY(sort_order(true));
            ta.qlist = perm;
        }
        int deferred = 0;
        RankCellsArgs ra{Q, dqc, qs, cells, clen, cinfo, qinfo, Tq, cn, p->w, p->T, P.ccap,
                         list, counters};
        if (halves) {
            if (P.big_cap > 0 || P.retry_cap > 0) {
                ra.big_list = (uint64_t *)(s + P.off_big_list);
                ra.big_sorted = (uint64_t *)(s + P.off_big_sorted);
                ra.big_size = (uint32_t *)(s + P.off_big_size);
                ra.big_aux = (uint32_t *)(s + P.off_big_aux);
                ra.big_cap = P.big_cap;
                ra.retry_cap = P.retry_cap;
                if (P.chunk <= 4) ra.big_split = (uint32_t *)(s + P.off_big_split);
            }
            ra.overflowed = &fallback_total;
            ra.stats = st ? stats : nullptr;
            if (P.retry_cap > 0 && P.big_cap == 0) {
                ra.defer_list = list2;
                ra.defer_count = counters + 1;
                ra.deferred = &deferred;
                // the re-ranking overlaps the scan of the other queries
                if (!g) ra.skip = (int32_t *)(s + P.off_skip);
            }
            SC_TRY(sc_rank_cells(ctx, ix, ra));
        } else if (three) {
            SC_TRY(traverse3(ctx, p, kc, sv, si, ix->L, cells, clen, cinfo, qinfo, P.ccap, cn, Tq));
        } else if (wave) {
            const size_t lds = trav_wave_layout(lists, kr, hcap_w).total;
            hipLaunchKernelGGL(k_traverse_wave, dim3((unsigned)cn), dim3(kTravThreads), lds,
                               ctx->stream, ta, hcap_w);
            SC_CHECK_LAUNCH();
        } else if (hcap > 0) {
            size_t lds = (size_t)qpw * lane_bytes;
            hipLaunchKernelGGL(k_traverse, dim3((unsigned)sc_cdiv(cn, qpw)), dim3(qpw), lds,
                               ctx->stream, ta, hcap, qpw);
            SC_CHECK_LAUNCH();
        }
        mark();
        ScanArgs sa{scan_kind, ix->dim, ix->mc, kc, ix->mp, ix->kp, p->R, P.ccap, qs, dqc, dqr,
                    ix->dot_cr, ix->L, ix->pid, ix->codes, cells, cinfo, qinfo,
                    d_ids + (size_t)q0 * p->R, d_dists + (size_t)q0 * p->R,
                    d_sums ? d_sums + q0 : nullptr, nullptr, clen,
                    Q, p->form == SC_FORM_REAL ? ix->raw : nullptr, imi_scan_mode(),
                    p->form == SC_FORM_MIXED && scan_kind == SC_KIND_IMI ? ix->evec : nullptr,
                    ix->dot_absmax, qinfo, mixed_margin_scale(),
                    ix->sparse ? ix->ckeys : nullptr, ix->ncells};
        sa.vcell = ix->vcell;
        if (sa.evec) sa.mrec = ix->mrec;
        // LDS: the diff_qr table (or the query row of the real-distance form)
        const size_t lds =
            sizeof(ScanShared) +
            sizeof(float) * std::max<size_t>((size_t)ix->mp * ix->kp,
                                             p->form == SC_FORM_REAL ? ix->dim : 0);
        if (order > 0 && cn > 1) {
            if (!ordered) SC_TRY(sort_order(false));
            sa.qlist = perm;
        }
        if (!g)   // a plan sets it before capture
            SC_TRY(sc_set_max_lds((const void *)k_scan_topk, lds));
        hipEvent_t k0 = nullptr, k1 = nullptr;   // around the scan launches (stats)
        if (st && hipEventCreate(&k0) == hipSuccess) {
            (void)hipEventRecord(k0, ctx->stream);
            kev.push_back(k0);
        }
        if (hcap_eff > 0 || three || halves) {
            ScanArgs first = sa;
            first.skip = deferred ? ra.skip : nullptr;
            hipLaunchKernelGGL(k_scan_topk, dim3((unsigned)cn), dim3(kScanThreads), lds,
                               ctx->stream, first);
            SC_CHECK_LAUNCH();
        }
        if (deferred && ra.skip) {
            // the re-ranked queries on the second stream, after their ranking
            // and alongside the others' scan: a workgroup per query of the
            // chunk, those past the count exit; then the streams join
            ScanArgs later = sa;
            later.qlist = list;
            later.qcount = counters;
            hipLaunchKernelGGL(k_scan_topk, dim3((unsigned)cn), dim3(kScanThreads), lds,
                               ctx->side, later);
            SC_CHECK_LAUNCH();
            SC_HIP(hipEventRecord(ctx->ev_join, ctx->side));
            SC_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
        }
        if (k0 && hipEventCreate(&k1) == hipSuccess) {
            (void)hipEventRecord(k1, ctx->stream);
            kev.push_back(k1);
        }
        if (sa.evec) {
            // mixed IMI scan: queries whose kept set outgrew the buffer are
            // scanned again exactly (their first outputs are overwritten), a
            // workgroup per query of the chunk, those past the device-side
            // count exiting at once (no host synchronisation)
            int32_t *dn = counters + 2;
            hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, ctx->stream, dn, 1);
            hipLaunchKernelGGL(k_gather_redo, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256), 0,
                               ctx->stream, qinfo, cn, list, dn);
            ScanArgs rs = sa;
            rs.evec = nullptr;
            rs.qlist = list;
            rs.qcount = dn;
            hipLaunchKernelGGL(k_scan_topk, dim3((unsigned)cn), dim3(kScanThreads), lds,
                               ctx->stream, rs);
            SC_CHECK_LAUNCH();
        }
        if (deferred) {
            // the one host synchronisation of the chunk: queries the ranking
            // left unranked (past the re-ranking's capacity, or over its
            // list) are ranked now and scanned again exactly (their first,
            // empty outputs are overwritten)
            int32_t h[2] = {0, 0};
            SC_HIP(hipMemcpyAsync(h, counters, 8, hipMemcpyDeviceToHost, ctx->stream));
            SC_HIP(hipStreamSynchronize(ctx->stream));
            fallback_total += h[0];
            if (h[1] > 0) {
                SC_TRY(sc_rank_finish(ctx, ix, ra, list2, h[1]));
                ScanArgs fs = sa;
                fs.evec = nullptr;
                fs.qlist = list2;
                hipLaunchKernelGGL(k_scan_topk, dim3((unsigned)h[1]), dim3(kScanThreads), lds,
                                   ctx->stream, fs);
                SC_CHECK_LAUNCH();
            }
        }
        mark();
        // Queries whose heap outgrew the on-chip capacity are traversed again
        // with a global-memory heap of the reference's capacity and scanned
        // again; the first scan's output for them is overwritten.
        if (g && hcap_eff < hcap_ref) {
            hipLaunchKernelGGL(k_flag_fallback, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256), 0,
                               ctx->stream, qinfo, cn, g->flags + q0, g->nflag);
            SC_CHECK_LAUNCH();
        } else if (!three && !halves && hcap_eff < hcap_ref) {
            int32_t nfb = 0;
            int32_t *dn = counters;
            if (hcap_eff > 0) {
                SC_HIP(hipMemsetAsync(dn, 0, 4, ctx->stream));
                hipLaunchKernelGGL(k_gather_fallback, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256),
                                   0, ctx->stream, qinfo, cn, list, dn);
                SC_CHECK_LAUNCH();
                SC_HIP(hipMemcpyAsync(&nfb, dn, 4, hipMemcpyDeviceToHost, ctx->stream));
                SC_HIP(hipStreamSynchronize(ctx->stream));
            } else {
                nfb = (int32_t)cn;
            }
            if (nfb > 0) {
                size_t gb = (size_t)nfb * ((size_t)hcap_ref * 8 + (size_t)kc * 4);
                void *g = nullptr;
                if (hipMalloc(&g, gb) != hipSuccess) {
                    sc_set_error("sc_search: fallback heap allocation (%zu bytes) failed", gb);
                    return SC_ERR_NOMEM;
                }
                TravArgs fa = ta;
                fa.qlist = hcap_eff > 0 ? list : nullptr;
                fa.nq = nfb;
                fa.gheap_v = (float *)g;
                fa.gheap_p = (uint32_t *)((float *)g + (size_t)nfb * hcap_ref);
                fa.grow = (int32_t *)(fa.gheap_p + (size_t)nfb * hcap_ref);
                hipLaunchKernelGGL(k_traverse, dim3((unsigned)sc_cdiv(nfb, 64)), dim3(64), 0,
                                   ctx->stream, fa, hcap_ref, 64);
                hipError_t le = hipGetLastError();
                ScanArgs fs = sa;
                fs.qlist = hcap_eff > 0 ? list : nullptr;
                fs.evec = nullptr;   // exact scan: the mixed form's redo does not follow
                if (le == hipSuccess) {
                    hipLaunchKernelGGL(k_scan_topk, dim3((unsigned)nfb), dim3(kScanThreads), lds,
                                       ctx->stream, fs);
                    le = hipGetLastError();
                }
                (void)hipStreamSynchronize(ctx->stream);
                (void)hipFree(g);
                if (le != hipSuccess) {
                    sc_set_error("sc_search: fallback launch failed: %s", hipGetErrorString(le));
                    return SC_ERR_HIP;
                }
                fallback_total += nfb;
            }
        }
        if (st || g) {
            hipLaunchKernelGGL(k_reduce_stats, dim3((unsigned)sc_cdiv(cn, 256)), dim3(256), 0,
                               ctx->stream, qinfo, cn, g ? g->stats : stats, g ? 1 : 0);
            SC_CHECK_LAUNCH();
        }
    }
    (void)lists;
    if (st) {
        unsigned long long h[8];
        SC_HIP(hipMemcpyAsync(h, stats, 64, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP(hipStreamSynchronize(ctx->stream));
        st->queries = nq;
        st->candidates = (int64_t)h[0];
        st->cells = (int64_t)h[1];
        st->exhausted = (int64_t)h[3];
        st->heap_fallback = fallback_total;
        st->visited = (int64_t)h[2];
        double acc[4] = {0, 0, 0, 0};
        for (size_t c = 0; c + 4 < ev.size(); c += 5)
            for (int k = 0; k < 4; ++k) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, ev[c + k], ev[c + k + 1]) == hipSuccess) acc[k] += ms;
            }
        st->ms_lut = acc[0]; st->ms_rank = acc[1]; st->ms_traverse = acc[2]; st->ms_scan = acc[3];
        double ks = 0.0;
        for (size_t c = 0; c + 1 < kev.size(); c += 2) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, kev[c], kev[c + 1]) == hipSuccess) ks += ms;
        }
        st->ms_scan_kernel = ks;
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : kev) (void)hipEventDestroy(e);
        if (h[4]) {
            sc_set_error("sc_search: %llu queries hit an internal traversal error", h[4]);
            return SC_ERR_HIP;
        }
    }
    return SC_OK;
}

// ----------------------------------------------------------------------------
// Plans: the fused persistent search of a fixed batch captured in a hipGraph.
// ----------------------------------------------------------------------------
struct sc_plan {
    sc_ctx *ctx;
    const sc_index *ix;
    sc_search_params p;
    const float *d_q;
    const int32_t *d_T;            // per-query budgets or null
    int64_t nq;
    int32_t *d_ids, *d_sums;
    float *d_dists;
    bool fused;
    FusedArgs fa;
    size_t lds;
    int threads, grid, occ, variant;
    bool mixed;                    // SC_FORM_MIXED kernels
    int *counter;                  // [0] work counter, [1] flagged
    int32_t *flags, *qinfo;
    unsigned long long *stats;
    float *lut;                    // [nq][mc*kc + mp*kp] (tiled LUT kernel) or null
    hipGraph_t graph;
    hipGraphExec_t exec;
    unsigned char *sbuf;           // staged graph: the staged kernels' scratch
};

// LDS of the staged scan kernel for this index and form.
static size_t staged_scan_lds(const sc_index *ix, const sc_search_params *p) {
    return sizeof(ScanShared) +
           sizeof(float) * std::max<size_t>((size_t)ix->mp * ix->kp,
                                            p->form == SC_FORM_REAL ? ix->dim : 0);
}

// Captures the staged kernels of the plan's batch (LUTs, coarse ranking,
// traversal, scan, the mixed form's exact redo) in one graph over the plan's
// own scratch.
static int plan_capture_staged(sc_plan *P) {
    if (P->exec) { (void)hipGraphExecDestroy(P->exec); P->exec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
    SC_TRY(sc_set_max_lds((const void *)k_scan_topk, staged_scan_lds(P->ix, &P->p)));
    hipStream_t sm = P->ctx->stream;
    SC_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, P->counter, 2);
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, (int *)P->stats, 16);
    const StagedGraph g{P->sbuf, P->flags, P->counter + 1, P->stats};
    int rc = search_staged(P->ctx, P->ix, P->d_q, P->nq, &P->p, P->d_T, P->d_ids, P->d_dists,
                           P->d_sums, nullptr, &g);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(sm, &P->graph);
    if (rc != SC_OK) return rc;
    if (le != hipSuccess || ce != hipSuccess) {
        sc_set_error("sc_plan: graph capture failed: %s",
                     hipGetErrorString(le != hipSuccess ? le : ce));
        return SC_ERR_HIP;
    }
    SC_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    return SC_OK;
}

static void plan_lut(sc_plan *P) {
    if (!P->lut) return;
    const sc_index *ix = P->ix;
    const int lut_per = ix->mc * ix->kc + ix->mp * ix->kp;
    hipLaunchKernelGGL(k_lut_tiled, dim3((unsigned)sc_cdiv(P->nq, kLutQ), (unsigned)(lut_per / kLutC)),
                       dim3(256), 0, P->ctx->stream, P->d_q, P->nq, ix->dim, ix->cqT, ix->pqT,
                       ix->norm_c, ix->norm_r, ix->mc, ix->kc, ix->mp, ix->kp, P->lut,
                       P->lut + (size_t)P->nq * ix->mc * ix->kc);
}

static int plan_capture(sc_plan *P) {
    if (P->exec) { (void)hipGraphExecDestroy(P->exec); P->exec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
    hipStream_t sm = P->ctx->stream;
    SC_HIP(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
    // the work counter is cleared by a kernel, not a memset node
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, P->counter, 2);
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, (int *)(P->stats + 8), 12);
    plan_lut(P);
    hipError_t le = launch_fused(P->variant, P->mixed, P->grid, P->threads, P->lds, sm, P->fa);
    if (le == hipSuccess) le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(sm, &P->graph);
    if (le != hipSuccess || ce != hipSuccess) {
        sc_set_error("sc_plan: graph capture failed: %s", hipGetErrorString(le != hipSuccess ? le : ce));
        return SC_ERR_HIP;
    }
    SC_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    return SC_OK;
}

static int plan_create(sc_ctx *ctx, const sc_index *ix, const sc_search_params *p,
                       const float *d_q, const int32_t *d_T, int64_t nq, int32_t *d_ids,
                       float *d_dists, int32_t *d_sums, sc_plan **out) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!out || !p || nq < 1 || !d_q || !d_ids || !d_dists || !d_sums) {
        sc_set_error("sc_plan_create: bad arguments");
        return SC_ERR_ARG;
    }
    *out = nullptr;
    // validate with an empty search of the same parameters
    SC_TRY(sc_search(ctx, ix, d_q, 0, p, d_ids, d_dists, d_sums, nullptr));
    sc_plan *P = new (std::nothrow) sc_plan();
    if (!P) return SC_ERR_NOMEM;
    *P = sc_plan{};
    P->ctx = ctx; P->ix = ix; P->p = *p; P->d_q = d_q; P->d_T = d_T; P->nq = nq;
    P->d_ids = d_ids; P->d_dists = d_dists; P->d_sums = d_sums;
    const int kind = p->kind, kc = ix->kc;
    const FusedConfig F = fused_config(ix, p);
    const FusedLayout &Lay = F.lay;
    const int hcap = F.hcap, ccap = F.ccap, npow2 = F.npow2;
    P->fused = kind != SC_KIND_IMI && kind != SC_KIND_SC3 && kind != SC_KIND_SCIMI &&
               p->form != SC_FORM_REAL && Lay.total <= 150 * 1024;
    if (!P->fused) {
        // the staged kernels in a graph (kinds whose staged path needs no
        // host synchronisation); otherwise sc_plan_execute runs sc_search
        if (kind != SC_KIND_SC3 && kind != SC_KIND_SCIMI && !getenv("SC_NO_PLAN_GRAPH")) {
            const Plan SP = make_plan(ix, p, nq);
            bool ok = hipMalloc(&P->counter, 256) == hipSuccess &&
                      hipMalloc(&P->flags, 4 * (size_t)nq) == hipSuccess &&
                      hipMalloc(&P->stats, 8 * 14) == hipSuccess &&
                      hipMalloc(&P->sbuf, SP.total) == hipSuccess;
            if (!ok) {
                sc_plan_destroy(P);
                sc_set_error("sc_plan_create: allocation failed");
                return SC_ERR_NOMEM;
            }
            const int rc = plan_capture_staged(P);
            if (rc == SC_ERR_TOO_LARGE) {   // no graph form: plain staged calls
                for (void *b : {(void *)P->counter, (void *)P->flags, (void *)P->sbuf,
                                (void *)P->stats})
                    (void)hipFree(b);
                P->counter = nullptr; P->flags = nullptr; P->sbuf = nullptr; P->stats = nullptr;
                if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
            } else if (rc != SC_OK) {
                sc_plan_destroy(P);
                return rc;
            }
        }
        *out = P;
        return SC_OK;
    }
    P->lds = Lay.total;
    bool ok = hipMalloc(&P->counter, 256) == hipSuccess &&
              hipMalloc(&P->flags, 4 * (size_t)nq) == hipSuccess &&
              hipMalloc(&P->qinfo, 32 * (size_t)nq) == hipSuccess &&
              hipMalloc(&P->stats, 8 * 14) == hipSuccess;   // 8 statistics + 6 phase counters
    if (ok && F.tiled_lut)
        ok = hipMalloc(&P->lut, 4 * (size_t)nq * (ix->mc * kc + ix->mp * ix->kp)) == hipSuccess;
    if (!ok) { sc_plan_destroy(P); sc_set_error("sc_plan_create: allocation failed"); return SC_ERR_NOMEM; }
    FusedArgs &fa = P->fa;
    fa.s = ScanArgs{kind,      ix->dim, ix->mc,  kc,      ix->mp,     ix->kp, p->R,
                    ccap,      nullptr, nullptr, nullptr, ix->dot_cr, ix->L,  ix->pid,
                    ix->codes, nullptr, nullptr, nullptr, d_ids,      d_dists, d_sums, nullptr};
    fa.Q = d_q; fa.cq = ix->cqT; fa.pq = ix->pqT; fa.norm_c = ix->norm_c; fa.norm_r = ix->norm_r;
    fa.nq = nq; fa.w = p->w; fa.T = p->T; fa.hcap = hcap; fa.ccap = ccap; fa.npow2 = npow2;
    fa.bufcap = F.bufcap;
    fa.counter = P->counter; fa.flags = P->flags; fa.nflag = P->counter + 1; fa.qinfo = P->qinfo;
    fa.lut_dqc = P->lut;
    fa.lut_dqr = P->lut ? P->lut + (size_t)nq * ix->mc * kc : nullptr;
    fa.phase = P->stats + 8;
    fa.margin_scale = mixed_margin_scale();
    fa.nvec = (int)ix->n;
    fa.pop_budget = fused_pop_budget();
    fa.Tq = d_T;
    P->mixed = p->form == SC_FORM_MIXED;
    for (int v = 3; v <= 5; ++v)
        if (hipFuncSetAttribute((const void *)fused_kernel(v, P->mixed),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)P->lds) != hipSuccess) {
            sc_plan_destroy(P);
            return SC_ERR_HIP;
        }
    P->threads = kScanThreads;
    P->occ = F.occ;
    P->variant = F.variant;
    P->grid = (int)std::min<int64_t>((int64_t)ctx->n_cu * std::min(F.occ, F.variant), nq);
    int rc = plan_capture(P);
    if (rc) { sc_plan_destroy(P); return rc; }
    *out = P;
    return SC_OK;
}

extern "C" int sc_plan_create(sc_ctx *ctx, const sc_index *ix, const sc_search_params *p,
                              const float *d_q, int64_t nq, int32_t *d_ids, float *d_dists,
                              int32_t *d_sums, sc_plan **out) {
    return plan_create(ctx, ix, p, d_q, nullptr, nq, d_ids, d_dists, d_sums, out);
}

extern "C" int sc_plan_create_budgets(sc_ctx *ctx, const sc_index *ix,
                                      const sc_search_params *p, const float *d_q,
                                      const int32_t *d_T, int64_t nq, int32_t *d_ids,
                                      float *d_dists, int32_t *d_sums, sc_plan **out) {
    if (!d_T) {
        sc_set_error("sc_plan_create_budgets: no budgets");
        return SC_ERR_ARG;
    }
    return plan_create(ctx, ix, p, d_q, d_T, nq, d_ids, d_dists, d_sums, out);
}

extern "C" int sc_plan_execute(sc_plan *P, sc_search_stats *st) {
    if (!P) return SC_ERR_ARG;
    SC_TRY(sc_ctx_enter(P->ctx));
    if (!P->fused && !P->exec)
        return search_impl(P->ctx, P->ix, P->d_q, P->nq, &P->p, P->d_T, P->d_ids, P->d_dists,
                           P->d_sums, st);
    if (st) *st = sc_search_stats{};
    hipStream_t sm = P->ctx->stream;
    if (!P->fused) {   // the staged graph
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (st) {
            SC_HIP(hipEventCreate(&e0));
            SC_HIP(hipEventCreate(&e1));
            SC_HIP(hipEventRecord(e0, sm));
        }
        SC_HIP(hipGraphLaunch(P->exec, sm));
        int nflag = 0;
        unsigned long long h[8] = {0};
        if (st) SC_HIP(hipEventRecord(e1, sm));
        SC_HIP(hipMemcpyAsync(&nflag, P->counter + 1, 4, hipMemcpyDeviceToHost, sm));
        if (st) SC_HIP(hipMemcpyAsync(h, P->stats, 64, hipMemcpyDeviceToHost, sm));
        SC_HIP(hipStreamSynchronize(sm));
        if (st) {
            float ms = 0.0f;
            SC_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            st->queries = P->nq;
            st->candidates = (int64_t)h[0];
            st->cells = (int64_t)h[1];
            st->visited = (int64_t)h[2];
            st->exhausted = (int64_t)h[3];
            st->reranked = (int64_t)h[6];
            st->ms_scan = ms;
        }
        return finish_flagged(P->ctx, P->ix, &P->p, P->d_q, P->d_T, P->nq, P->flags, nflag,
                              P->d_ids, P->d_dists, P->d_sums, st);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st) {
        SC_HIP(hipEventCreate(&e0));
        SC_HIP(hipEventCreate(&e1));
        SC_HIP(hipEventRecord(e0, sm));
    }
    SC_HIP(hipGraphLaunch(P->exec, sm));
    if (st) {
        SC_HIP(hipEventRecord(e1, sm));
        SC_HIP(hipMemsetAsync(P->stats, 0, 64, sm));
        hipLaunchKernelGGL(k_reduce_stats, dim3((unsigned)sc_cdiv(P->nq, 256)), dim3(256), 0, sm,
                           P->qinfo, P->nq, P->stats);
        SC_CHECK_LAUNCH();
    }
    int nflag = 0;
    unsigned long long h[14] = {0};
    SC_HIP(hipMemcpyAsync(&nflag, P->counter + 1, 4, hipMemcpyDeviceToHost, sm));
    if (st) SC_HIP(hipMemcpyAsync(h, P->stats, 112, hipMemcpyDeviceToHost, sm));
    SC_HIP(hipStreamSynchronize(sm));
    if (st) {
        const double wgs = h[13] ? (double)h[13] : 1.0;
        for (int i = 0; i < 5; ++i) st->fused_phase_ms[i] = (double)h[8 + i] / wgs / 1e5;
        float ms = 0.0f;
        SC_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        st->queries = P->nq;
        st->candidates = (int64_t)h[0];
        st->cells = (int64_t)h[1];
        st->visited = (int64_t)h[2];
        st->exhausted = (int64_t)h[3];
        st->reranked = (int64_t)h[6];
        st->ms_scan = ms;
    }
    return finish_flagged(P->ctx, P->ix, &P->p, P->d_q, P->d_T, P->nq, P->flags, nflag, P->d_ids,
                          P->d_dists, P->d_sums, st);
}

extern "C" int sc_plan_autotune(sc_plan *P, int32_t *out4) {
    if (!P) return SC_ERR_ARG;
    SC_TRY(sc_ctx_enter(P->ctx));
    if (!P->fused) {
        if (out4) { out4[0] = 0; out4[1] = 0; out4[2] = 0; out4[3] = 0; }
        return SC_OK;
    }
    hipStream_t sm = P->ctx->stream;
    const int occ = P->occ;
    hipEvent_t e0, e1;
    SC_HIP(hipEventCreate(&e0));
    SC_HIP(hipEventCreate(&e1));
    float best = INFINITY;
    int bt = P->threads, bg = P->grid, bv = P->variant;
    plan_lut(P);                       // the fused kernel reads this batch's LUTs
    // register variants x workgroups per CU (256 threads); a variant beyond
    // what the LDS allows per CU only adds spills
    for (int v = 3; v <= std::max(3, std::min(occ, 5)); ++v) {
        for (int per_cu = std::max(1, std::min(v, occ) - 1); per_cu <= std::min(v, occ); ++per_cu) {
            const int threads = kScanThreads;
            const int grid = (int)std::min<int64_t>((int64_t)P->ctx->n_cu * per_cu, P->nq);
            float tot = 0.0f;
            for (int rep = 0; rep < 3; ++rep) {
                hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, sm, P->counter, 2);
                SC_HIP(hipEventRecord(e0, sm));
                SC_HIP(launch_fused(v, P->mixed, grid, threads, P->lds, sm, P->fa));
                SC_HIP(hipEventRecord(e1, sm));
                SC_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                SC_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) tot += ms;        // first run warms caches
            }
            if (tot < best) { best = tot; bt = threads; bg = grid; bv = v; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P->threads = bt;
    P->grid = bg;
    P->variant = bv;
    SC_TRY(plan_capture(P));
    if (out4) { out4[0] = bt; out4[1] = bg; out4[2] = (int32_t)(best / 2.0f * 1000.0f); out4[3] = bv; }
    return SC_OK;
}

extern "C" int sc_plan_destroy(sc_plan *P) {
    if (!P) return SC_OK;
    (void)hipSetDevice(P->ctx->device);
    (void)hipStreamSynchronize(P->ctx->stream);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    for (void *b : {(void *)P->counter, (void *)P->flags, (void *)P->qinfo, (void *)P->stats,
                    (void *)P->lut, (void *)P->sbuf})
        if (b) (void)hipFree(b);
    delete P;
    return SC_OK;
}
