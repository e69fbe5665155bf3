// sc_index.hip -- inverted file construction and the query-independent tables.
//
// Keys (S4): IMI/IVF key = sum_j cid_j * kc^(mc-1-j) (vector_base,
// include/sc_utilities.h:156-166); SC key uid_j = nearest_j*kc + second_j,
// key = sum_j uid_j * (kc^2)^(mc-1-j) (src/sc_encoder.cpp:39-50), 64-bit.
// A stable radix sort of (key, id) gives every cell its members in
// ascending id order, exactly the order Bucket::push_back produces
// (src/encoder.cpp:182-186), and the dense inclusive prefix L[] the loaders
// build (src/query.cpp:133-139).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>

#include "sc_internal.h"

namespace {

__global__ void k_keys(const uint16_t *__restrict__ cid, int64_t n, int mc, int kc, int nc,
                       uint64_t *__restrict__ keys, int32_t *__restrict__ ids) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t key = 0;
    if (nc == 1) {
        for (int j = 0; j < mc; ++j) key = key * (uint64_t)kc + cid[(size_t)i * mc + j];
    } else if (nc == 3) {
        // SCEncoder::encode layout [n][mc][3]; uid = c0 kc^2 + c1 kc + c2
        const uint16_t *c = cid + (size_t)i * 3 * mc;
        for (int j = 0; j < mc; ++j)
            key = key * (uint64_t)kc * (uint64_t)kc * (uint64_t)kc +
                  (((uint64_t)c[3 * j] * (uint64_t)kc + c[3 * j + 1]) * (uint64_t)kc + c[3 * j + 2]);
    } else {
        const uint16_t *c = cid + (size_t)i * 2 * mc;
        for (int j = 0; j < mc; ++j)
            key = key * (uint64_t)kc * (uint64_t)kc + ((uint64_t)c[j] * (uint64_t)kc + c[mc + j]);
    }
    keys[i] = key;
    ids[i] = (int32_t)i;
}

// L[c] = number of sorted keys <= c (inclusive prefix of cell sizes).
__global__ void k_dense_prefix(const uint64_t *__restrict__ skeys, int64_t n, int64_t ncells,
                               int32_t *__restrict__ L) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int64_t lo = 0, hi = n;  // first index with key > c
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (skeys[mid] <= (uint64_t)c) lo = mid + 1; else hi = mid;
    }
    L[c] = (int32_t)lo;
}

// Sparse directory of the two-halves index: heads[i] = 1 where sorted key i
// starts a new cell.
__global__ void k_heads(const uint64_t *__restrict__ skeys, int64_t n, int32_t *__restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    h[i] = (i == 0 || skeys[i] != skeys[i - 1]) ? 1 : 0;
}
// From the inclusive scan of the heads (cell number + 1 of every sorted
// entry): each cell's key and the inclusive prefix L (one past its last entry).
__global__ void k_sparse_fill(const uint64_t *__restrict__ skeys, const int32_t *__restrict__ cno,
                              int64_t n, uint64_t *__restrict__ ckeys, int32_t *__restrict__ L) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cno[i] - 1;
    if (i == 0 || skeys[i] != skeys[i - 1]) ckeys[c] = skeys[i];
    if (i == n - 1 || skeys[i + 1] != skeys[i]) L[c] = (int32_t)(i + 1);
}
// Ranking records: the two (nearest, second) pairs of each cell and its size.
__global__ void k_cell_records(const uint64_t *__restrict__ ckeys, const int32_t *__restrict__ L,
                               int64_t ncells, int kc, int words, uint32_t *__restrict__ rec) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint64_t k2 = (uint64_t)kc * (uint64_t)kc, key = ckeys[c];
    const uint64_t u0 = key / k2, u1 = key - u0 * k2;
    const uint32_t n0 = (uint32_t)(u0 / kc), s0 = (uint32_t)(u0 % kc);
    const uint32_t n1 = (uint32_t)(u1 / kc), s1 = (uint32_t)(u1 % kc);
    const uint32_t size = (uint32_t)(L[c] - (c > 0 ? L[c - 1] : 0));
    if (words == 2) {
        rec[2 * c] = n0 | (s0 << 8) | (n1 << 16) | (s1 << 24);
        rec[2 * c + 1] = size;
    } else {
        rec[4 * c] = n0 | (s0 << 16);
        rec[4 * c + 1] = n1 | (s1 << 16);
        rec[4 * c + 2] = size;
        rec[4 * c + 3] = 0;
    }
}

// The halves' pairs: uid_h of every cell (h = 0: key / kc^2, h = 1: key % kc^2).
__global__ void k_cell_uids(const uint64_t *__restrict__ ckeys, int64_t ncells, int kc,
                            uint32_t *__restrict__ u0, uint32_t *__restrict__ u1) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint64_t k2 = (uint64_t)kc * (uint64_t)kc, key = ckeys[c];
    u0[c] = (uint32_t)(key / k2);
    u1[c] = (uint32_t)(key - (key / k2) * k2);
}
// Run heads of a sorted list.
__global__ void k_heads32(const uint32_t *__restrict__ s, int64_t n, int32_t *__restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    h[i] = (i == 0 || s[i] != s[i - 1]) ? 1 : 0;
}
// The distinct values (from the inclusive scan of the heads), stored as
// n | s << 16 for uid = n*kc + s.
__global__ void k_unique_pairs(const uint32_t *__restrict__ s, const int32_t *__restrict__ no,
                               int64_t n, int kc, uint32_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || s[i] != s[i - 1]) out[no[i] - 1] = (s[i] / kc) | ((s[i] % kc) << 16);
}
// Each cell's positions in the two halves' pair lists (binary search of the
// ascending uid lists) and its size.
__global__ void k_cell_pairs(const uint64_t *__restrict__ ckeys, const int32_t *__restrict__ L,
                             int64_t ncells, int kc, const uint32_t *__restrict__ p0, int n0,
                             const uint32_t *__restrict__ p1, int n1, uint32_t *__restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint64_t k2 = (uint64_t)kc * (uint64_t)kc, key = ckeys[c];
    const uint32_t u[2] = {(uint32_t)(key / k2), (uint32_t)(key % k2)};
    const uint32_t *p[2] = {p0, p1};
    const int np[2] = {n0, n1};
    uint32_t pos[2];
    for (int h = 0; h < 2; ++h) {
        int lo = 0, hi = np[h];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const uint32_t v = p[h][mid];
            const uint32_t uv = (v & 0xFFFFu) * (uint32_t)kc + (v >> 16);
            if (uv < u[h]) lo = mid + 1; else hi = mid;
        }
        pos[h] = (uint32_t)lo;
    }
    out[2 * c] = pos[0] | (pos[1] << 16);
    out[2 * c + 1] = (uint32_t)(L[c] - (c > 0 ? L[c - 1] : 0));
}

// First cell of each half-0 pair (cells are in key order, so a pair's cells
// are contiguous); gstart[n0] = ncells.
__global__ void k_group_starts(const uint32_t *__restrict__ cpair, int64_t ncells, int n0,
                               int32_t *__restrict__ gstart) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    if (c == ncells) { gstart[n0] = (int32_t)ncells; return; }
    const uint32_t g = cpair[2 * c] & 0xFFFFu;
    if (c == 0 || (cpair[2 * (c - 1)] & 0xFFFFu) != g) gstart[g] = (int32_t)c;
}

// Vectors in the cells of each half-0 pair (from the inclusive prefix L).
__global__ void k_group_vecs(const int32_t *__restrict__ gstart, const int32_t *__restrict__ L,
                             int n0, int32_t *__restrict__ gvec) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n0) return;
    const int32_t c0 = gstart[g], c1 = gstart[g + 1];
    gvec[g] = L[c1 - 1] - (c0 > 0 ? L[c0 - 1] : 0);
}

// Groups of the pruned ranking: the runs of cells sharing a half-0 pair
// (uid_0 = key / kc^2; cells are in key order).  heads[c] = 1 where a run
// starts.
__global__ void k_group_heads(const uint64_t *__restrict__ ckeys, int64_t ncells, int kc,
                              int32_t *__restrict__ heads) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint64_t k2 = (uint64_t)kc * (uint64_t)kc;
    heads[c] = (c == 0 || ckeys[c] / k2 != ckeys[c - 1] / k2) ? 1 : 0;
}
// From the inclusive scan of the heads (gno = group number + 1 of each cell):
// each group's pair and first cell, and each cell's half-1 record
// n1 | s1 << 12 | min(size, 255) << 24 (kc <= 4096).
__global__ void k_group_fill(const uint64_t *__restrict__ ckeys, const int32_t *__restrict__ L,
                             const int32_t *__restrict__ gno, int64_t ncells, int kc,
                             uint32_t *__restrict__ bg_pair, int32_t *__restrict__ bg_start,
                             uint32_t *__restrict__ brec) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint64_t k2 = (uint64_t)kc * (uint64_t)kc, key = ckeys[c];
    const uint32_t u0 = (uint32_t)(key / k2), u1 = (uint32_t)(key - (uint64_t)u0 * k2);
    if (c == 0 || ckeys[c - 1] / k2 != u0) {
        const int32_t g = gno[c] - 1;
        bg_start[g] = (int32_t)c;
        bg_pair[g] = (u0 / (uint32_t)kc) | ((u0 % (uint32_t)kc) << 16);
    }
    if (c == ncells - 1) bg_start[gno[c]] = (int32_t)ncells;
    const uint32_t size = (uint32_t)(L[c] - (c > 0 ? L[c - 1] : 0));
    brec[c] = (u1 / (uint32_t)kc) | ((u1 % (uint32_t)kc) << 12) | (min(size, 255u) << 24);
}
// bg_n0[n] = the first group whose nearest-0 centroid is at least n
// (n in [0, kc]; bg_n0[kc] = ngroups).
__global__ void k_group_n0(const uint32_t *__restrict__ bg_pair, int32_t ngroups, int kc,
                           int32_t *__restrict__ bg_n0) {
    const int n = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n > kc) return;
    int lo = 0, hi = ngroups;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)(bg_pair[mid] & 0xFFFFu) < n) lo = mid + 1; else hi = mid;
    }
    bg_n0[n] = lo;
}

// Per vector p (cell order) of a two-block index, the sum of its dot_cr terms
// in subspace order, e = sum_k dot_cr[row(k)][k][c_k] with row(k) the cell's
// nearest centroid of the half subspace k lies in (include/multi_query.h:
// 231-253 adds the same terms, interleaved with the query's).
__global__ void k_vec_dot(const uint8_t *__restrict__ codes, const int32_t *__restrict__ L,
                          int64_t ncells, const uint64_t *__restrict__ ckeys, int64_t n, int kc,
                          int mp, int kp, const float *__restrict__ dot_cr, float *__restrict__ evec,
                          uint32_t *__restrict__ vcell, uint4 *__restrict__ mrec) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int64_t lo = 0, hi = ncells;            // the cell of p: first c with L[c] > p
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)L[mid] <= p) lo = mid + 1; else hi = mid;
    }
    int n0, n1;
    if (ckeys) {
        const uint64_t k2 = (uint64_t)kc * (uint64_t)kc, key = ckeys[lo];
        n0 = (int)(key / k2 / (uint64_t)kc);
        n1 = (int)(key % k2 / (uint64_t)kc);
    } else {
        n0 = (int)(lo / kc);
        n1 = (int)(lo % kc);
    }
    const int half = mp >> 1, bs = kp * half;
    const uint8_t *cc = codes + (size_t)p * mp;
    float e = 0.0f;
    for (int k = 0; k < half; ++k) e += dot_cr[(size_t)n0 * bs + k * kp + cc[k]];
    for (int k = half; k < mp; ++k) e += dot_cr[(size_t)(kc + n1) * bs + (k - half) * kp + cc[k]];
    evec[p] = e;
    vcell[p] = (uint32_t)n0 | ((uint32_t)n1 << 16);
    if (mrec) {   // mp == 8
        const uint2 c = *reinterpret_cast<const uint2 *>(cc);
        mrec[p] = make_uint4(c.x, c.y, __float_as_uint(e), vcell[p]);
    }
}

// out = max |x| over x[0..n) as float bits (non-negative floats order like
// their bit patterns)
__global__ void k_absmax(const float *__restrict__ x, int64_t n, uint32_t *out) {
    uint32_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf(x[i])));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// out[0] = largest cell, out[1] = number of non-empty cells
__global__ void k_cell_max(const int32_t *__restrict__ L, int64_t ncells, int32_t *out) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t l = 0;
    if (c < ncells) l = c > 0 ? L[c] - L[c - 1] : L[0];
    const int ne = __popcll(__ballot(l > 0));
    for (int o = 32; o > 0; o >>= 1) l = max(l, __shfl_xor(l, o));
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, l);
        atomicAdd(out + 1, ne);
    }
}

__global__ void k_gather_codes(const uint8_t *__restrict__ codes, const int32_t *__restrict__ ids,
                               int64_t n, int mp, uint8_t *__restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * mp) return;
    int64_t p = e / mp;
    int s = (int)(e - p * mp);
    out[e] = codes[(size_t)ids[p] * mp + s];
}

// PQQuery::pre_compute1 (include/query.h:91-149): sequential FP32 sums with
// separately rounded products.  This file is compiled with
// -ffp-contract=off, so `d += a * b` stays a multiply and an add.
__global__ void k_norms(const float *__restrict__ cb, int64_t rows, int len,
                        float *__restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const float *v = cb + (size_t)i * len;
    float d = 0.0f;
    for (int j = 0; j < len; ++j) { float t = v[j]; d += t * t; }
    out[i] = d;
}

// dot_cr[((i*kc + j)*spb + i1)*kp + j1] = 2 * sum_t c_{i,j}[i1*bsp+t] *
// pq_{i*spb+i1, j1}[t]   (spb = mp/mc sub-quantizers per coarse block)
__global__ void k_dot_cr(const float *__restrict__ cq, const float *__restrict__ pq, int mc,
                         int kc, int mp, int kp, int bsc, int bsp, float *__restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int spb = mp / mc;
    const int64_t total = (int64_t)mc * kc * spb * kp;
    if (e >= total) return;
    const int j1 = (int)(e % kp);
    int64_t r = e / kp;
    const int i1 = (int)(r % spb);
    r /= spb;                                  // r = i*kc + j
    const int i = (int)(r / kc);
    const float *c = cq + (size_t)r * bsc + (size_t)i1 * bsp;
    const float *p = pq + ((size_t)(i * spb + i1) * kp + j1) * bsp;
    float d = 0.0f;
    for (int t = 0; t < bsp; ++t) d += c[t] * p[t];
    out[e] = 2.0f * d;
}

int alloc_index(sc_ctx *ctx, int64_t n, int32_t dim, int32_t mc, int32_t kc, int32_t mp,
                int32_t kp, int32_t nc, int64_t ncells, sc_index **out) {
    sc_index *ix = new (std::nothrow) sc_index();
    if (!ix) return SC_ERR_NOMEM;
    *ix = sc_index{};
    ix->ctx = ctx;
    ix->dim = dim; ix->mc = mc; ix->kc = kc; ix->mp = mp; ix->kp = kp; ix->nc = nc;
    ix->n = n; ix->ncells = ncells;
    const size_t bsc = dim / mc, bsp = dim / mp;
    bool ok = hipMalloc(&ix->cq, sizeof(float) * mc * kc * bsc) == hipSuccess &&
              hipMalloc(&ix->pq, sizeof(float) * mp * kp * bsp) == hipSuccess &&
              hipMalloc(&ix->norm_c, sizeof(float) * mc * kc) == hipSuccess &&
              hipMalloc(&ix->norm_r, sizeof(float) * mp * kp) == hipSuccess &&
              hipMalloc(&ix->dot_cr, sizeof(float) * (size_t)kc * mp * kp) == hipSuccess &&
              hipMalloc(&ix->L, sizeof(int32_t) * (ncells > 0 ? ncells : 1)) == hipSuccess &&
              hipMalloc(&ix->pid, sizeof(int32_t) * (n > 0 ? n : 1)) == hipSuccess &&
              hipMalloc(&ix->codes, (size_t)(n > 0 ? n : 1) * mp) == hipSuccess &&
              hipMalloc(&ix->cqT, sizeof(float) * mc * kc * bsc) == hipSuccess &&
              hipMalloc(&ix->pqT, sizeof(float) * mp * kp * bsp) == hipSuccess;
    if (ok && mc == 2 && nc == 2) {      // two halves: sparse directory
        ix->sparse = true;
        ix->crec_words = kc <= 256 ? 2 : 4;
        ok = hipMalloc(&ix->ckeys, sizeof(uint64_t) * (ncells > 0 ? ncells : 1)) == hipSuccess &&
             hipMalloc(&ix->crec, sizeof(uint32_t) * ix->crec_words * (ncells > 0 ? ncells : 1)) ==
                 hipSuccess;
    }
    if (!ok) {
        sc_index_destroy(ix);
        sc_set_error("index allocation failed (n=%lld cells=%lld)", (long long)n,
                     (long long)ncells);
        return SC_ERR_NOMEM;
    }
    *out = ix;
    return SC_OK;
}

int check_shape(int64_t n, int32_t dim, int32_t mc, int32_t kc, int32_t mp, int32_t kp,
                int32_t nc, int64_t *ncells) {
    if (n < 0 || n >= INT32_MAX || dim <= 0 || mc <= 0 || mp <= 0 || dim % mc || dim % mp ||
        mp % mc || kc < 1 || kc > 65536 || kp < 1 || kp > 256 || nc < 1 || nc > 3 ||
        (nc == 3 && (mc != 1 || kc > 1024))) {
        sc_set_error("index: bad shape (n=%lld dim=%d mc=%d kc=%d mp=%d kp=%d nc=%d)",
                     (long long)n, dim, mc, kc, mp, kp, nc);
        return SC_ERR_ARG;
    }
    if (mc == 2 && nc == 2) {
        // two halves: (nearest, second) per half, 64-bit keys below kc^4 < 2^64,
        // only the non-empty cells are stored (their number is known after
        // the sort)
        if (kc < 2 || kc > 65535) {
            sc_set_error("index: two-halves cells need 2 <= kc <= 65535 (kc=%d)", kc);
            return SC_ERR_ARG;
        }
        *ncells = 0;
        return SC_OK;
    }
    double cells = 1.0;
    for (int j = 0; j < mc * nc; ++j) cells *= kc;
    if (cells >= (double)INT32_MAX) {
        sc_set_error("index: kc^(mc*nc) = %.3g cells is too large for a dense index "
                     "(src/sc_encoder.cpp:28-33)", cells);
        return SC_ERR_TOO_LARGE;
    }
    *ncells = (int64_t)cells;
    return SC_OK;
}

// The halves' distinct pairs and each cell's positions in them (two-halves
// index): the cell ranking then scores a cell with two table lookups.
static int build_pair_lists(sc_index *ix) {
    sc_ctx *ctx = ix->ctx;
    const int64_t nc = ix->ncells;
    size_t sort_bytes = 0, scan_bytes = 0;
    SC_HIP(rocprim::radix_sort_keys(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (size_t)nc, 0u, 32u, ctx->stream));
    SC_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, (int32_t *)nullptr, (int32_t *)nullptr,
                                   (size_t)nc, rocprim::plus<int32_t>(), ctx->stream));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_u1 = al((size_t)nc * 4), o_s = o_u1 + al((size_t)nc * 4),
                 o_h = o_s + al((size_t)nc * 4), o_no = o_h + al((size_t)nc * 4),
                 o_tmp = o_no + al((size_t)nc * 4), total = o_tmp + al(std::max(sort_bytes, scan_bytes));
    unsigned char *s;
    SC_TRY(sc_scratch(ctx, total, (void **)&s));
    uint32_t *u[2] = {(uint32_t *)s, (uint32_t *)(s + o_u1)};
    uint32_t *srt = (uint32_t *)(s + o_s);
    int32_t *heads = (int32_t *)(s + o_h), *no = (int32_t *)(s + o_no);
    const unsigned g = (unsigned)sc_cdiv(nc, 256);
    hipLaunchKernelGGL(k_cell_uids, dim3(g), dim3(256), 0, ctx->stream, ix->ckeys, nc, ix->kc, u[0],
                       u[1]);
    SC_CHECK_LAUNCH();
    for (int h = 0; h < 2; ++h) {
        // h = 0 is already ascending (cells are in key order); sort h = 1
        const uint32_t *sorted = u[h];
        if (h == 1) {
            size_t tb = sort_bytes;
            SC_HIP(rocprim::radix_sort_keys(s + o_tmp, tb, u[1], srt, (size_t)nc, 0u, 32u,
                                            ctx->stream));
            sorted = srt;
        }
        hipLaunchKernelGGL(k_heads32, dim3(g), dim3(256), 0, ctx->stream, sorted, nc, heads);
        SC_CHECK_LAUNCH();
        size_t tb = scan_bytes;
        SC_HIP(rocprim::inclusive_scan(s + o_tmp, tb, heads, no, (size_t)nc,
                                       rocprim::plus<int32_t>(), ctx->stream));
        int32_t cnt = 0;
        SC_HIP(hipMemcpyAsync(&cnt, no + (nc - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP(hipStreamSynchronize(ctx->stream));
        ix->nu[h] = cnt;
        if (cnt > 65535) return SC_OK;          // too many for 16-bit positions: no tables
        if (hipMalloc(&ix->upairs[h], 4 * (size_t)cnt) != hipSuccess) {
            sc_set_error("index: pair list allocation failed");
            return SC_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_unique_pairs, dim3(g), dim3(256), 0, ctx->stream, sorted, no, nc,
                           ix->kc, ix->upairs[h]);
        SC_CHECK_LAUNCH();
    }
    if (hipMalloc(&ix->cpair, 8 * (size_t)nc) != hipSuccess) {
        sc_set_error("index: cell pair allocation failed");
        return SC_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_cell_pairs, dim3(g), dim3(256), 0, ctx->stream, ix->ckeys, ix->L, nc,
                       ix->kc, ix->upairs[0], ix->nu[0], ix->upairs[1], ix->nu[1], ix->cpair);
    SC_CHECK_LAUNCH();
    if (hipMalloc(&ix->gstart, 4 * ((size_t)ix->nu[0] + 1)) != hipSuccess) {
        sc_set_error("index: group start allocation failed");
        return SC_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_group_starts, dim3((unsigned)sc_cdiv(nc + 1, 256)), dim3(256), 0,
                       ctx->stream, ix->cpair, nc, ix->nu[0], ix->gstart);
    SC_CHECK_LAUNCH();
    if (hipMalloc(&ix->gvec, 4 * ((size_t)ix->nu[0] + 1)) != hipSuccess) {
        sc_set_error("index: group size allocation failed");
        return SC_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_group_vecs, dim3((unsigned)sc_cdiv(ix->nu[0], 256)), dim3(256), 0,
                       ctx->stream, ix->gstart, ix->L, ix->nu[0], ix->gvec);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

// The groups and half-1 records of the pruned ranking (two halves, kc <= 4096).
static int build_groups(sc_index *ix) {
    sc_ctx *ctx = ix->ctx;
    const int64_t nc = ix->ncells;
    size_t scan_bytes = 0;
    SC_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, (int32_t *)nullptr, (int32_t *)nullptr,
                                   (size_t)nc, rocprim::plus<int32_t>(), ctx->stream));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_gno = al((size_t)nc * 4), o_tmp = o_gno + al((size_t)nc * 4),
                 total = o_tmp + al(scan_bytes);
    unsigned char *s;
    SC_TRY(sc_scratch(ctx, total, (void **)&s));
    int32_t *heads = (int32_t *)s, *gno = (int32_t *)(s + o_gno);
    const unsigned g = (unsigned)sc_cdiv(nc, 256);
    hipLaunchKernelGGL(k_group_heads, dim3(g), dim3(256), 0, ctx->stream, ix->ckeys, nc, ix->kc,
                       heads);
    SC_CHECK_LAUNCH();
    size_t tb = scan_bytes;
    SC_HIP(rocprim::inclusive_scan(s + o_tmp, tb, heads, gno, (size_t)nc,
                                   rocprim::plus<int32_t>(), ctx->stream));
    int32_t ng = 0;
    SC_HIP(hipMemcpyAsync(&ng, gno + (nc - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    ix->ngroups = ng;
    if (hipMalloc(&ix->bg_pair, 4 * (size_t)ng) != hipSuccess ||
        hipMalloc(&ix->bg_start, 4 * ((size_t)ng + 1)) != hipSuccess ||
        hipMalloc(&ix->bg_n0, 4 * ((size_t)ix->kc + 1)) != hipSuccess ||
        hipMalloc(&ix->brec, 4 * (size_t)nc) != hipSuccess) {
        sc_set_error("index: group directory allocation failed");
        return SC_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_group_fill, dim3(g), dim3(256), 0, ctx->stream, ix->ckeys, ix->L, gno, nc,
                       ix->kc, ix->bg_pair, ix->bg_start, ix->brec);
    SC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_group_n0, dim3((unsigned)sc_cdiv((int64_t)ix->kc + 1, 256)), dim3(256), 0,
                       ctx->stream, ix->bg_pair, ng, ix->kc, ix->bg_n0);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

int finish_index(sc_index *ix) {
    sc_ctx *ctx = ix->ctx;
    SC_TRY(sc_launch_tables(ix));
    if (ix->sparse && ix->ncells > 0) {
        hipLaunchKernelGGL(k_cell_records, dim3((unsigned)sc_cdiv(ix->ncells, 256)), dim3(256), 0,
                           ctx->stream, ix->ckeys, ix->L, ix->ncells, ix->kc, ix->crec_words,
                           ix->crec);
        SC_CHECK_LAUNCH();
        SC_TRY(build_pair_lists(ix));
        if (ix->kc <= 4096) SC_TRY(build_groups(ix));
    }
    int32_t *dmax, h[3] = {0, 0, 0};
    SC_TRY(sc_scratch(ctx, 3 * sizeof(int32_t), (void **)&dmax));
    SC_HIP(hipMemsetAsync(dmax, 0, 3 * sizeof(int32_t), ctx->stream));
    if (ix->ncells > 0) {
        hipLaunchKernelGGL(k_cell_max, dim3((unsigned)sc_cdiv(ix->ncells, 256)), dim3(256), 0,
                           ctx->stream, ix->L, ix->ncells, dmax);
        SC_CHECK_LAUNCH();
    }
    const int64_t ndot = (int64_t)ix->kc * ix->mp * ix->kp;
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)std::min<int64_t>(1024, sc_cdiv(ndot, 256))),
                       dim3(256), 0, ctx->stream, ix->dot_cr, ndot, (uint32_t *)(dmax + 2));
    SC_CHECK_LAUNCH();
    if (ix->mc == 2 && ix->n > 0) {
        if (hipMalloc(&ix->evec, sizeof(float) * (size_t)ix->n) != hipSuccess ||
            hipMalloc(&ix->vcell, sizeof(uint32_t) * (size_t)ix->n) != hipSuccess ||
            (ix->mp == 8 && hipMalloc(&ix->mrec, sizeof(uint4) * (size_t)ix->n) != hipSuccess)) {
            sc_set_error("index: per-vector sums allocation failed");
            return SC_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_vec_dot, dim3((unsigned)sc_cdiv(ix->n, 256)), dim3(256), 0,
                           ctx->stream, ix->codes, ix->L, ix->ncells,
                           ix->sparse ? ix->ckeys : nullptr, ix->n, ix->kc, ix->mp, ix->kp,
                           ix->dot_cr, ix->evec, ix->vcell, ix->mp == 8 ? ix->mrec : nullptr);
        SC_CHECK_LAUNCH();
    }
    SC_HIP(hipMemcpyAsync(h, dmax, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    ix->max_cell = h[0];
    ix->nonempty = h[1];
    ix->dot_absmax = __builtin_bit_cast(float, (uint32_t)h[2]);
    return SC_OK;
}

}  // namespace

// [nb][K][blen] -> [nb][blen][K]
__global__ void k_transpose_cb(const float *__restrict__ cb, int nb, int K, int blen,
                               float *__restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nb * K * blen) return;
    int64_t b = e / ((int64_t)K * blen);
    int r = (int)(e - b * K * blen);
    int k = r / blen, t = r % blen;
    out[(b * blen + t) * K + k] = cb[e];
}

int sc_launch_tables(sc_index *ix) {
    sc_ctx *ctx = ix->ctx;
    const int bsc = ix->dim / ix->mc, bsp = ix->dim / ix->mp;
    hipLaunchKernelGGL(k_transpose_cb, dim3((unsigned)sc_cdiv((int64_t)ix->mc * ix->kc * bsc, 256)),
                       dim3(256), 0, ctx->stream, ix->cq, ix->mc, ix->kc, bsc, ix->cqT);
    hipLaunchKernelGGL(k_transpose_cb, dim3((unsigned)sc_cdiv((int64_t)ix->mp * ix->kp * bsp, 256)),
                       dim3(256), 0, ctx->stream, ix->pq, ix->mp, ix->kp, bsp, ix->pqT);
    SC_CHECK_LAUNCH();
    int64_t rc = (int64_t)ix->mc * ix->kc, rp = (int64_t)ix->mp * ix->kp;
    hipLaunchKernelGGL(k_norms, dim3((unsigned)sc_cdiv(rc, 256)), dim3(256), 0, ctx->stream,
                       ix->cq, rc, bsc, ix->norm_c);
    SC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_norms, dim3((unsigned)sc_cdiv(rp, 256)), dim3(256), 0, ctx->stream,
                       ix->pq, rp, bsp, ix->norm_r);
    SC_CHECK_LAUNCH();
    int64_t total = (int64_t)ix->kc * ix->mp * ix->kp;
    hipLaunchKernelGGL(k_dot_cr, dim3((unsigned)sc_cdiv(total, 256)), dim3(256), 0,
                       ctx->stream, ix->cq, ix->pq, ix->mc, ix->kc, ix->mp, ix->kp, bsc, bsp,
                       ix->dot_cr);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

// The two-halves index (mc = 2, nc = 2): the keys of src/sc_encoder.cpp:39-50
// (uid_h = nearest_h*kc + second_h, key = uid_0*kc^2 + uid_1) in 64 bits, a
// stable radix sort of (key, id) -- ids ascending within a cell, as
// Bucket::push_back leaves them (src/encoder.cpp:182-186) -- and only the
// non-empty cells kept, in key order (the reference's dense kc^4 array,
// src/sc_encoder.cpp:26-33, does not fit).
static int build_sparse(sc_ctx *ctx, const uint16_t *d_cid, const uint8_t *d_codes, int64_t n,
                        int32_t dim, const float *d_cq, int32_t kc, const float *d_pq, int32_t mp,
                        int32_t kp, sc_index **out) {
    const int mc = 2, nc = 2;
    int bits = 1;
    {
        const double cells = (double)kc * kc * kc * kc;
        while (bits < 64 && std::ldexp(1.0, bits) < cells) ++bits;
    }
    // scratch: keys, sorted keys, ids, sorted ids, cell numbers, sort/scan temp
    size_t sort_bytes = 0, scan_bytes = 0;
    if (n > 0) {
        SC_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t *)nullptr,
                                         (uint64_t *)nullptr, (int32_t *)nullptr,
                                         (int32_t *)nullptr, (size_t)n, 0u, (unsigned)bits,
                                         ctx->stream));
        SC_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, (int32_t *)nullptr,
                                       (int32_t *)nullptr, (size_t)n, rocprim::plus<int32_t>(),
                                       ctx->stream));
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_keys2 = al((size_t)n * 8), o_ids = o_keys2 + al((size_t)n * 8),
                 o_ids2 = o_ids + al((size_t)n * 4), o_cno = o_ids2 + al((size_t)n * 4),
                 o_heads = o_cno + al((size_t)n * 4), o_tmp = o_heads + al((size_t)n * 4),
                 total = o_tmp + al(std::max(sort_bytes, scan_bytes));
    unsigned char *s;
    SC_TRY(sc_scratch(ctx, total, (void **)&s));
    uint64_t *keys = (uint64_t *)s, *skeys = (uint64_t *)(s + o_keys2);
    int32_t *ids = (int32_t *)(s + o_ids), *sids = (int32_t *)(s + o_ids2);
    int32_t *cno = (int32_t *)(s + o_cno), *heads = (int32_t *)(s + o_heads);
    int64_t ncells = 0;
    if (n > 0) {
        hipLaunchKernelGGL(k_keys, dim3((unsigned)sc_cdiv(n, 256)), dim3(256), 0, ctx->stream,
                           d_cid, n, mc, kc, nc, keys, ids);
        SC_CHECK_LAUNCH();
        size_t tb = sort_bytes;
        SC_HIP(rocprim::radix_sort_pairs(s + o_tmp, tb, keys, skeys, ids, sids, (size_t)n, 0u,
                                         (unsigned)bits, ctx->stream));
        hipLaunchKernelGGL(k_heads, dim3((unsigned)sc_cdiv(n, 256)), dim3(256), 0, ctx->stream,
                           skeys, n, heads);
        SC_CHECK_LAUNCH();
        tb = scan_bytes;
        SC_HIP(rocprim::inclusive_scan(s + o_tmp, tb, heads, cno, (size_t)n,
                                       rocprim::plus<int32_t>(), ctx->stream));
        int32_t h = 0;
        SC_HIP(hipMemcpyAsync(&h, cno + (n - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP(hipStreamSynchronize(ctx->stream));
        ncells = h;
    }
    sc_index *ix;
    SC_TRY(alloc_index(ctx, n, dim, mc, kc, mp, kp, nc, ncells, &ix));
    auto fail = [&](int rc) { sc_index_destroy(ix); return rc; };
    const size_t bsc = dim / mc, bsp = dim / mp;
    if (hipMemcpyAsync(ix->cq, d_cq, sizeof(float) * mc * kc * bsc, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess ||
        hipMemcpyAsync(ix->pq, d_pq, sizeof(float) * mp * kp * bsp, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess) {
        sc_set_error("index: codebook copy failed");
        return fail(SC_ERR_HIP);
    }
    if (n > 0) {
        if (hipMemcpyAsync(ix->pid, sids, sizeof(int32_t) * n, hipMemcpyDeviceToDevice,
                           ctx->stream) != hipSuccess)
            return fail(SC_ERR_HIP);
        hipLaunchKernelGGL(k_sparse_fill, dim3((unsigned)sc_cdiv(n, 256)), dim3(256), 0,
                           ctx->stream, skeys, cno, n, ix->ckeys, ix->L);
        hipLaunchKernelGGL(k_gather_codes, dim3((unsigned)sc_cdiv(n * mp, 256)), dim3(256), 0,
                           ctx->stream, d_codes, ix->pid, n, mp, ix->codes);
        if (hipGetLastError() != hipSuccess) return fail(SC_ERR_HIP);
    }
    int rc = finish_index(ix);
    if (rc) return fail(rc);
    *out = ix;
    return SC_OK;
}

extern "C" int sc_index_build(sc_ctx *ctx, const uint16_t *d_cid, const uint8_t *d_codes,
                              int64_t n, int32_t dim, const float *d_cq, int32_t mc, int32_t kc,
                              const float *d_pq, int32_t mp, int32_t kp, int32_t nc,
                              sc_index **out) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!out || !d_cq || !d_pq || (n > 0 && (!d_cid || !d_codes))) return SC_ERR_ARG;
    *out = nullptr;
    int64_t ncells;
    SC_TRY(check_shape(n, dim, mc, kc, mp, kp, nc, &ncells));
    if (mc == 2 && nc == 2)
        return build_sparse(ctx, d_cid, d_codes, n, dim, d_cq, kc, d_pq, mp, kp, out);
    sc_index *ix;
    SC_TRY(alloc_index(ctx, n, dim, mc, kc, mp, kp, nc, ncells, &ix));
    const size_t bsc = dim / mc, bsp = dim / mp;
    auto fail = [&](int rc) { sc_index_destroy(ix); return rc; };
    if (hipMemcpyAsync(ix->cq, d_cq, sizeof(float) * mc * kc * bsc, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess ||
        hipMemcpyAsync(ix->pq, d_pq, sizeof(float) * mp * kp * bsp, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess) {
        sc_set_error("index: codebook copy failed");
        return fail(SC_ERR_HIP);
    }
    if (n > 0) {
        int bits = 1;
        while (bits < 64 && ((uint64_t)1 << bits) < (uint64_t)ncells) ++bits;
        size_t tmp_bytes = 0;
        if (rocprim::radix_sort_pairs(nullptr, tmp_bytes, (uint64_t *)nullptr,
                                      (uint64_t *)nullptr, (int32_t *)nullptr,
                                      (int32_t *)nullptr, (size_t)n, 0u, (unsigned)bits,
                                      ctx->stream) != hipSuccess)
            return fail(SC_ERR_HIP);
        size_t off_keys2 = (size_t)n * 8, off_ids = off_keys2 + (size_t)n * 8,
               off_tmp = off_ids + (size_t)n * 4;
        off_tmp = (off_tmp + 255) & ~(size_t)255;
        unsigned char *s;
        int rc = sc_scratch(ctx, off_tmp + tmp_bytes, (void **)&s);
        if (rc) return fail(rc);
        uint64_t *keys = (uint64_t *)s, *skeys = (uint64_t *)(s + off_keys2);
        int32_t *ids = (int32_t *)(s + off_ids);
        hipLaunchKernelGGL(k_keys, dim3((unsigned)sc_cdiv(n, 256)), dim3(256), 0, ctx->stream,
                           d_cid, n, mc, kc, nc, keys, ids);
        if (hipGetLastError() != hipSuccess) return fail(SC_ERR_HIP);
        if (rocprim::radix_sort_pairs(s + off_tmp, tmp_bytes, keys, skeys, ids, ix->pid,
                                      (size_t)n, 0u, (unsigned)bits, ctx->stream) != hipSuccess) {
            sc_set_error("index: radix sort failed");
            return fail(SC_ERR_HIP);
        }
        hipLaunchKernelGGL(k_dense_prefix, dim3((unsigned)sc_cdiv(ncells, 256)), dim3(256), 0,
                           ctx->stream, skeys, n, ncells, ix->L);
        hipLaunchKernelGGL(k_gather_codes, dim3((unsigned)sc_cdiv(n * mp, 256)), dim3(256), 0,
                           ctx->stream, d_codes, ix->pid, n, mp, ix->codes);
        if (hipGetLastError() != hipSuccess) return fail(SC_ERR_HIP);
    } else {
        if (hipMemsetAsync(ix->L, 0, sizeof(int32_t) * ncells, ctx->stream) != hipSuccess)
            return fail(SC_ERR_HIP);
    }
    int rc = finish_index(ix);
    if (rc) return fail(rc);
    *out = ix;
    return SC_OK;
}

static int check_lists(const int32_t *h_L, int64_t ncells, const int32_t *h_pid, int64_t n);

extern "C" int sc_index_load(sc_ctx *ctx, const int32_t *h_L, int64_t ncells,
                             const int32_t *h_pid, const uint8_t *h_codes, int64_t n,
                             int32_t dim, const float *h_cq, int32_t mc, int32_t kc,
                             const float *h_pq, int32_t mp, int32_t kp, int32_t nc,
                             sc_index **out) {
    // host arguments are validated before any device work (and without a
    // device), so a malformed file is refused wherever it is loaded
    if (!out || !h_L || !h_cq || !h_pq || (n > 0 && (!h_pid || !h_codes))) {
        sc_set_error("index_load: null argument");
        return SC_ERR_ARG;
    }
    *out = nullptr;
    int64_t want;
    SC_TRY(check_shape(n, dim, mc, kc, mp, kp, nc, &want));
    if (mc == 2 && nc == 2) {
        sc_set_error("index_load: a two-halves index (mc = 2, nc = 2) is loaded with "
                     "sc_index_load_sparse (cell keys + L over the non-empty cells)");
        return SC_ERR_ARG;
    }
    if (ncells != want) {
        sc_set_error("index_load: %lld cells given, shape needs %lld", (long long)ncells,
                     (long long)want);
        return SC_ERR_ARG;
    }
    // The kernels index codes and ids through L[] and pid[] without bounds
    // checks, so a malformed inverted file is refused here, before upload.
    SC_TRY(check_lists(h_L, ncells, h_pid, n));
    SC_TRY(sc_ctx_enter(ctx));
    sc_index *ix;
    SC_TRY(alloc_index(ctx, n, dim, mc, kc, mp, kp, nc, ncells, &ix));
    const size_t bsc = dim / mc, bsp = dim / mp;
    bool ok =
        hipMemcpyAsync(ix->cq, h_cq, sizeof(float) * mc * kc * bsc, hipMemcpyHostToDevice,
                       ctx->stream) == hipSuccess &&
        hipMemcpyAsync(ix->pq, h_pq, sizeof(float) * mp * kp * bsp, hipMemcpyHostToDevice,
                       ctx->stream) == hipSuccess &&
        hipMemcpyAsync(ix->L, h_L, sizeof(int32_t) * ncells, hipMemcpyHostToDevice,
                       ctx->stream) == hipSuccess &&
        (n == 0 || (hipMemcpyAsync(ix->pid, h_pid, sizeof(int32_t) * n, hipMemcpyHostToDevice,
                                   ctx->stream) == hipSuccess &&
                    hipMemcpyAsync(ix->codes, h_codes, (size_t)n * mp, hipMemcpyHostToDevice,
                                   ctx->stream) == hipSuccess));
    if (!ok) { sc_index_destroy(ix); sc_set_error("index_load: copy failed"); return SC_ERR_HIP; }
    int rc = finish_index(ix);
    if (rc) { sc_index_destroy(ix); return rc; }
    *out = ix;
    return SC_OK;
}

// Host-side checks shared by the loaders: L a non-decreasing inclusive prefix
// from 0 ending at n, every pid in [0, n).
static int check_lists(const int32_t *h_L, int64_t ncells, const int32_t *h_pid, int64_t n) {
    if (ncells > 0 && h_L[ncells - 1] != n) {
        sc_set_error("index_load: L[last]=%d != n=%lld", h_L[ncells - 1], (long long)n);
        return SC_ERR_ARG;
    }
    if (ncells == 0 && n != 0) {
        sc_set_error("index_load: no cells for n=%lld vectors", (long long)n);
        return SC_ERR_ARG;
    }
    for (int64_t c = 0; c < ncells; ++c) {
        const int32_t prev = c > 0 ? h_L[c - 1] : 0;
        if (h_L[c] < prev) {
            sc_set_error("index_load: L is not a non-decreasing inclusive prefix from 0 "
                         "(L[%lld]=%d after %d)", (long long)c, h_L[c], prev);
            return SC_ERR_ARG;
        }
    }
    for (int64_t i = 0; i < n; ++i)
        if (h_pid[i] < 0 || h_pid[i] >= n) {
            sc_set_error("index_load: pid[%lld]=%d outside [0, %lld)", (long long)i, h_pid[i],
                         (long long)n);
            return SC_ERR_ARG;
        }
    return SC_OK;
}

extern "C" int sc_index_load_sparse(sc_ctx *ctx, const uint64_t *h_ckeys, const int32_t *h_L,
                                    int64_t ncells, const int32_t *h_pid, const uint8_t *h_codes,
                                    int64_t n, int32_t dim, const float *h_cq, int32_t kc,
                                    const float *h_pq, int32_t mp, int32_t kp, sc_index **out) {
    const int mc = 2, nc = 2;
    if (!out || !h_cq || !h_pq || ncells < 0 ||
        (ncells > 0 && (!h_ckeys || !h_L)) || (n > 0 && (!h_pid || !h_codes))) {
        sc_set_error("index_load_sparse: null argument");
        return SC_ERR_ARG;
    }
    *out = nullptr;
    int64_t unused;
    SC_TRY(check_shape(n, dim, mc, kc, mp, kp, nc, &unused));
    const double kc4 = (double)kc * kc * kc * kc;
    for (int64_t c = 0; c < ncells; ++c) {
        if ((double)h_ckeys[c] >= kc4 || (c > 0 && h_ckeys[c] <= h_ckeys[c - 1])) {
            sc_set_error("index_load_sparse: cell keys must ascend strictly below kc^4 "
                         "(key[%lld]=%llu)", (long long)c, (unsigned long long)h_ckeys[c]);
            return SC_ERR_ARG;
        }
        if (h_L[c] <= (c > 0 ? h_L[c - 1] : 0)) {
            sc_set_error("index_load_sparse: cell %lld is empty or L decreases", (long long)c);
            return SC_ERR_ARG;
        }
    }
    SC_TRY(check_lists(h_L, ncells, h_pid, n));
    SC_TRY(sc_ctx_enter(ctx));
    sc_index *ix;
    SC_TRY(alloc_index(ctx, n, dim, mc, kc, mp, kp, nc, ncells, &ix));
    const size_t bsc = dim / mc, bsp = dim / mp;
    bool ok =
        hipMemcpyAsync(ix->cq, h_cq, sizeof(float) * mc * kc * bsc, hipMemcpyHostToDevice,
                       ctx->stream) == hipSuccess &&
        hipMemcpyAsync(ix->pq, h_pq, sizeof(float) * mp * kp * bsp, hipMemcpyHostToDevice,
                       ctx->stream) == hipSuccess &&
        (ncells == 0 ||
         (hipMemcpyAsync(ix->L, h_L, sizeof(int32_t) * ncells, hipMemcpyHostToDevice,
                         ctx->stream) == hipSuccess &&
          hipMemcpyAsync(ix->ckeys, h_ckeys, sizeof(uint64_t) * ncells, hipMemcpyHostToDevice,
                         ctx->stream) == hipSuccess)) &&
        (n == 0 || (hipMemcpyAsync(ix->pid, h_pid, sizeof(int32_t) * n, hipMemcpyHostToDevice,
                                   ctx->stream) == hipSuccess &&
                    hipMemcpyAsync(ix->codes, h_codes, (size_t)n * mp, hipMemcpyHostToDevice,
                                   ctx->stream) == hipSuccess));
    if (!ok) {
        sc_index_destroy(ix);
        sc_set_error("index_load_sparse: copy failed");
        return SC_ERR_HIP;
    }
    int rc = finish_index(ix);
    if (rc) { sc_index_destroy(ix); return rc; }
    *out = ix;
    return SC_OK;
}

extern "C" int sc_index_cell_keys(sc_ctx *ctx, const sc_index *ix, uint64_t *h_keys) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!ix || !h_keys || !ix->sparse) {
        sc_set_error("sc_index_cell_keys: the index has no sparse directory (dense cells: "
                     "cell = key)");
        return SC_ERR_ARG;
    }
    if (ix->ncells > 0) {
        SC_HIP(hipMemcpyAsync(h_keys, ix->ckeys, sizeof(uint64_t) * ix->ncells,
                              hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP(hipStreamSynchronize(ctx->stream));
    }
    return SC_OK;
}

extern "C" int sc_index_attach_raw(sc_index *ix, const float *d_raw, int64_t n) {
    if (!ix || (d_raw && n != ix->n)) {
        sc_set_error("sc_index_attach_raw: the raw vectors must match the index (n=%lld)",
                     ix ? (long long)ix->n : 0ll);
        return SC_ERR_ARG;
    }
    ix->raw = d_raw;
    return SC_OK;
}

extern "C" int sc_index_destroy(sc_index *ix) {
    if (!ix) return SC_OK;
    if (ix->ctx) {
        (void)hipSetDevice(ix->ctx->device);
        (void)hipStreamSynchronize(ix->ctx->stream);
    }
    void *bufs[] = {ix->cq,    ix->pq,   ix->norm_c, ix->norm_r, ix->dot_cr,
                    ix->L,     ix->pid,  ix->codes,  ix->cqT,    ix->pqT,
                    ix->ckeys, ix->crec, ix->upairs[0], ix->upairs[1], ix->cpair,
                    ix->gstart, ix->gvec, ix->evec, ix->vcell, ix->mrec, ix->bg_n0,
                    ix->bg_pair, ix->bg_start, ix->brec};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete ix;
    return SC_OK;
}

extern "C" int sc_index_shape(const sc_index *ix, int64_t *out6) {
    if (!ix || !out6) return SC_ERR_ARG;
    out6[0] = ix->n; out6[1] = ix->ncells; out6[2] = ix->mc;
    out6[3] = ix->kc; out6[4] = ix->mp; out6[5] = ix->kp;
    return SC_OK;
}

extern "C" int sc_index_export(sc_ctx *ctx, const sc_index *ix, int32_t *h_L, int32_t *h_pid,
                               uint8_t *h_codes, float *h_norm_c, float *h_norm_r,
                               float *h_dot_cr) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!ix) return SC_ERR_ARG;
    hipStream_t s = ctx->stream;
    if (h_L) SC_HIP(hipMemcpyAsync(h_L, ix->L, sizeof(int32_t) * ix->ncells, hipMemcpyDeviceToHost, s));
    if (h_pid && ix->n) SC_HIP(hipMemcpyAsync(h_pid, ix->pid, sizeof(int32_t) * ix->n, hipMemcpyDeviceToHost, s));
    if (h_codes && ix->n) SC_HIP(hipMemcpyAsync(h_codes, ix->codes, (size_t)ix->n * ix->mp, hipMemcpyDeviceToHost, s));
    if (h_norm_c) SC_HIP(hipMemcpyAsync(h_norm_c, ix->norm_c, sizeof(float) * ix->mc * ix->kc, hipMemcpyDeviceToHost, s));
    if (h_norm_r) SC_HIP(hipMemcpyAsync(h_norm_r, ix->norm_r, sizeof(float) * ix->mp * ix->kp, hipMemcpyDeviceToHost, s));
    if (h_dot_cr) SC_HIP(hipMemcpyAsync(h_dot_cr, ix->dot_cr, sizeof(float) * (size_t)ix->kc * ix->mp * ix->kp, hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    return SC_OK;
}
