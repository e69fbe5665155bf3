// sc_encode_mfma.hip -- mixed-precision assignment (SC_PREC_BF16) with exact
// results.
//
// For every (vector, codeword) pair an approximate distance
//     d~ = ||x'||^2 + ||c'||^2 - 2 x'.c'
// is formed from a BF16 MFMA dot product with FP32 accumulation, where
// x' = x - m and c' = c - m are shifted by a per-block integer mean m (the
// distance is shift-invariant; the shift shrinks the norms that scale the
// error) and every operand is split into two BF16 terms (v = hi + lo,
// |v - hi - lo| <= 2^-16 |v|).  The codeword terms are laid out once in MFMA
// operand order (k_cb_frag), so a lane loads its B operand with one 16-byte
// load; a wave keeps 128 codewords' accumulators in registers at a time.
// d~ is combined in FP32: (float(||x'||^2) + float(||c'||^2)) - 2 x'.c'.
// For a row with X = ||x'|| and C = max_k ||c'_k|| the error obeys
// |d~ - d| <= E with
//     E = 2 * [ 2 X C ((3 * 2^-16 + gamma_n)(1 + 2^-7)) + 2^-22 (X + C)^2 ]
//         + 2^-22 (X + C)^2 + 1e-6,
//     gamma_n = n u / (1 - n u),  u = 2^-23,  n = 3 * blen + 8
// (Cauchy-Schwarz bounds sum |x'_i c'_i| by X C; the outer factor 2 covers the
// shift rounding and the FP64 rounding of the exact distance; the last
// 2^-22 (X + C)^2 covers the three FP32 roundings of the combination, at most
// 2^-24 (2 X^2 + 2 C^2 + (X + C)^2)).  The exact top-2 (or top-1) under the reference's order is then
// contained in {k : d~_k <= d~_(2) + 2E}; the kernel collects a superset of
// it (each group of codewords against the running second smallest, which
// never falls below the final one), and those candidates are re-evaluated
// with the exact S1 arithmetic and ordered with the reference's tie rule, so
// codes and cell ids equal the FP32 form bit for bit.  A row with more than
// kCand candidates is evaluated exactly over all codewords.
//
// Operand layout of v_mfma_f32_16x16x16bf16_1k (one wave): lane l holds
// A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16] for j < 4, and
// receives C[4 (l / 16) + i][l % 16] for i < 4.
#include <float.h>

#include "sc_internal.h"

namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 16;     // vectors per wave
constexpr int kCand = 16;     // exact re-check slots per row
constexpr int kGroup = 8;     // codeword tiles (of 16) per register group

__device__ __forceinline__ unsigned short bf16_bits(float v) {
    return __builtin_bit_cast(unsigned short, (__bf16)v);
}
__device__ __forceinline__ float bf16_val(unsigned short b) {
    return (float)__builtin_bit_cast(__bf16, b);
}

// Per-block codebook preparation: shifted, split BF16 codewords, FP64 norms
// of the shifted codewords and their maximum per block.
// mean[b][t] = round(mean_k c[b][k][t]) when `shift`, else 0.
// One workgroup per (block, dim): strided FP64 partial sums, then a fixed
// tree, so the mean does not depend on scheduling.
__global__ __launch_bounds__(256) void k_cb_mean(const float *__restrict__ cb, int nb, int K,
                                                 int blen, int shift, float *__restrict__ mean) {
    __shared__ double part[256];
    const int e = blockIdx.x;                 // b * blen + t
    const int b = e / blen, t = e % blen;
    double s = 0.0;
    for (int k = threadIdx.x; k < K; k += 256) s += cb[((size_t)b * K + k) * blen + t];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[e] = shift ? (float)rint(part[0] / K) : 0.0f;
}

__global__ void k_cb_norms(const float *__restrict__ cb, int nb, int K, int blen,
                           const float *__restrict__ mean, double *__restrict__ cn2,
                           unsigned long long *__restrict__ cmax_bits) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * K) return;
    int b = e / K;
    const float *c = cb + (size_t)e * blen;
    double s = 0.0;
    for (int t = 0; t < blen; ++t) {
        float v = c[t] - mean[(size_t)b * blen + t];
        s += (double)v * (double)v;
    }
    cn2[e] = s;
    // max of a non-negative double via its bit pattern
    atomicMax(&cmax_bits[b], (unsigned long long)__builtin_bit_cast(unsigned long long, sqrt(s)));
}

// Codebook fragments in MFMA operand order, one 16-byte load per lane per
// MFMA pair: for block b, tile kt of 16 codewords, dims step kk of 16 and
// lane l, the BF16 hi parts then the BF16 lo parts of c'[k][t + j], j < 4,
// with k = 16 kt + l % 16 and t = 16 kk + 4 (l / 16) (zero past K).
// Layout [nb][KT][KK][64] with KT = ceil(K / 16), KK = blen / 16.
__global__ void k_cb_frag(const float *__restrict__ cb, int nb, int K, int blen,
                          const float *__restrict__ mean, uint4 *__restrict__ frag) {
    const int KT = (K + 15) / 16, KK = blen / 16;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nb * KT * KK * 64) return;
    const int l = (int)(e % 64);
    const int64_t f = e / 64;
    const int kk = (int)(f % KK), kt = (int)((f / KK) % KT), b = (int)(f / ((int64_t)KK * KT));
    const int k = 16 * kt + (l & 15), t = 16 * kk + 4 * (l >> 4);
    unsigned short h[4], lo[4];
    for (int j = 0; j < 4; ++j) {
        const float v = k < K ? cb[((size_t)b * K + k) * blen + t + j] - mean[(size_t)b * blen + t + j]
                              : 0.0f;
        h[j] = bf16_bits(v);
        lo[j] = bf16_bits(v - bf16_val(h[j]));
    }
    frag[e] = make_uint4(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16),
                         lo[0] | ((uint32_t)lo[1] << 16), lo[2] | ((uint32_t)lo[3] << 16));
}

// S1 in FP64: exact squares, FMA-accumulated in t order.  b (a codeword in
// global memory) is read 16 values at a time so its loads overlap; the
// accumulation order stays sequential.
__device__ __forceinline__ double s1(const float *a, const float *b, int d) {
    double acc = 0.0;
    int t = 0;
    for (; t + 16 <= d; t += 16) {
        float bv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) bv[j] = b[t + j];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double dd = (double)(a[t + j] - bv[j]);
            acc = __fma_rn(dd, dd, acc);
        }
    }
    for (; t < d; ++t) {
        const double dd = (double)(a[t] - b[t]);
        acc = __fma_rn(dd, dd, acc);
    }
    return acc;
}

struct CbView {
    const float *c;              // [K][blen] original codewords
    const uint4 *frag;           // [KT][KK][64] BF16 hi/lo fragments (k_cb_frag)
    const double *cn2;           // [K]
    double cmax;
    const float *mean;           // [blen]
    int K, blen;
};

// (d, k) precedes (e, j): dist ASC, then index DESC when DESC else ASC.
template <bool DESC>
__device__ __forceinline__ bool precedes(double d, int k, double e, int j) {
    if (d < e) return true;
    if (!(d == e)) return false;   // d > e, or a NaN: the reference's `>`/`>=` never take it
    return DESC ? (k > j) : (k < j);
}

// One wave: best one (TOP2=false) or two (TOP2=true) codewords for the 16
// rows held in xs (row stride `ld`, columns [t0, t0 + blen)).  Writes
// out1/out2 [16].  ROUNDF orders by the distance rounded to float (the
// k-means labels, k-means.h:597-612): the candidate threshold then also
// covers every codeword whose rounded distance can equal the best's
// (within 2^-22 of it), and the re-check rounds before it compares.  Wave-private LDS scratch: cnt [16], ck [16][kCand],
// cd [16][kCand], rown [16] (FP64 norms of shifted rows), sm [blen] (the
// block's shift), wst [2] (the workgroup's re-checked candidates and fully
// exact rows).
template <bool TOP2, bool DESC, bool ROUNDF = false>
__device__ void wave_best(const float *xs, int ld, int t0, int nrows, const CbView cb,
                          int *cnt, int *ck, double *cd, double *rown, float *sm, int *out1,
                          int *out2, unsigned *wst) {
    const int l = threadIdx.x & 63;
    const int blen = cb.blen, K = cb.K;
    for (int t = l; t < blen; t += 64) sm[t] = cb.mean[t];
    __syncthreads();
    // 1. shifted row norms (FP64) and exactness of the BF16 representation
    bool exact_local = true;
    if (l < kRows) {
        double s = 0.0;
#pragma unroll 8
        for (int t = 0; t < blen; ++t) {
            float v = xs[l * ld + t0 + t] - sm[t];
            s += (double)v * (double)v;
            exact_local &= (bf16_val(bf16_bits(v)) == v);
        }
        rown[l] = s;
        cnt[l] = 0;
    }
    const bool x_exact = __all(exact_local);
    __syncthreads();
    const int r4 = 4 * (l >> 4);          // first of this lane's four result rows
    const int col = l & 15;
    double E[4];
    const double u = 1.0 / 8388608.0;     // 2^-23
    const double n = 3.0 * blen + 8.0;
    const double gam = n * u / (1.0 - n * u);
    for (int i = 0; i < 4; ++i) {
        const double X = sqrt(rown[r4 + i]), C = cb.cmax;
        E[i] = 2.0 * (2.0 * X * C * ((3.0 / 65536.0 + gam) * (1.0 + 1.0 / 128.0)) +
                      (X + C) * (X + C) / 4194304.0) +
               (X + C) * (X + C) / 4194304.0 + 1e-6;
    }
    // 2. approximate distances, codewords in groups of kGroup tiles of 16
    //    accumulated in registers over the dims.  After each group the row's
    //    running second (or first) smallest d~ plus 2E is the threshold, and
    //    the group's codewords at or below it become candidates.  A running
    //    threshold never falls below the final one, so the candidates
    //    contain every codeword within 2E of the final second smallest.
    const int KK = blen / 16, KT = (K + 15) / 16;
    const int ar = l & 15;                // A row
    float m1[4], m2[4], rf[4];
    for (int i = 0; i < 4; ++i) { m1[i] = INFINITY; m2[i] = INFINITY; rf[i] = (float)rown[r4 + i]; }
    for (int kt0 = 0; kt0 < KT; kt0 += kGroup) {
        const int gn = min(kGroup, KT - kt0);
        f32x4 acc[kGroup];
#pragma unroll
        for (int g = 0; g < kGroup; ++g) acc[g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const uint4 *fp = cb.frag + (size_t)kt0 * KK * 64 + l;
        for (int kk = 0; kk < KK; ++kk) {
            s16x4 ah, al;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int t = 16 * kk + r4 + j;
                const float v = xs[ar * ld + t0 + t] - sm[t];
                const unsigned short h = bf16_bits(v);
                ah[j] = (short)h;
                al[j] = (short)bf16_bits(v - bf16_val(h));
            }
#pragma unroll
            for (int g = 0; g < kGroup; ++g) {
                if (g < gn) {
                    const uint4 f = fp[((size_t)g * KK + kk) * 64];
                    const s16x4 bh = {(short)(f.x & 0xFFFFu), (short)(f.x >> 16),
                                      (short)(f.y & 0xFFFFu), (short)(f.y >> 16)};
                    const s16x4 bl = {(short)(f.z & 0xFFFFu), (short)(f.z >> 16),
                                      (short)(f.w & 0xFFFFu), (short)(f.w >> 16)};
                    acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[g], 0, 0, 0);
                    acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[g], 0, 0, 0);
                    if (!x_exact)
                        acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[g], 0, 0, 0);
                }
            }
        }
        // d~ in place (+inf past K), the two smallest of this lane's group
        float g1[4], g2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) { g1[i] = INFINITY; g2[i] = INFINITY; }
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
            const int k = 16 * (kt0 + g) + col;
            const float c2 = k < K ? (float)cb.cn2[k] : 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float d = k < K ? (rf[i] + c2) - 2.0f * acc[g][i] : INFINITY;
                acc[g][i] = d;
                if (d < g1[i]) { g2[i] = g1[i]; g1[i] = d; }
                else if (d < g2[i]) g2[i] = d;
            }
        }
        // the group's two smallest over the row's 16 lanes (disjoint sets),
        // merged into the row's running pair, then the threshold
        float thr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            for (int o = 1; o < 16; o <<= 1) {
                float a1 = __shfl_xor(g1[i], o), a2 = __shfl_xor(g2[i], o);
                float lo1 = fminf(g1[i], a1);
                float hi1 = fmaxf(g1[i], a1);
                g2[i] = fminf(hi1, fminf(g2[i], a2));
                g1[i] = lo1;
            }
            const float lo1 = fminf(m1[i], g1[i]), hi1 = fmaxf(m1[i], g1[i]);
            m2[i] = fminf(hi1, fminf(m2[i], g2[i]));
            m1[i] = lo1;
            const double mref = (double)(TOP2 ? m2[i] : m1[i]);
            thr[i] = (float)(mref + 2.0 * E[i] +
                             (ROUNDF ? (fabs(mref) + E[i]) * (1.0 / 4194304.0) : 0.0));
            // rounding the threshold up keeps it an upper bound
            thr[i] = nextafterf(thr[i], INFINITY);
        }
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
            const int k = 16 * (kt0 + g) + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (k < K && acc[g][i] <= thr[i] && r4 + i < nrows) {
                    int s = atomicAdd(&cnt[r4 + i], 1);
                    if (s < kCand) ck[(r4 + i) * kCand + s] = k;
                }
            }
        }
    }
    __syncthreads();
    // 3. exact re-check of the candidates, numbered densely over the rows
    //    (rows with more than kCand are evaluated in full below)
    int total = 0;
    for (int r = 0; r < nrows; ++r) total += cnt[r] <= kCand ? cnt[r] : 0;
    for (int p = l; p < total; p += 64) {
        int r = 0, s = p;
        while (true) {
            const int c = cnt[r] <= kCand ? cnt[r] : 0;
            if (s < c) break;
            s -= c;
            ++r;
        }
        const int k = ck[r * kCand + s];
        const double e = s1(xs + r * ld + t0, cb.c + (size_t)k * blen, blen);
        cd[r * kCand + s] = ROUNDF ? (double)(float)e : e;
    }
    __syncthreads();
    if (l < nrows) {
        double d1 = INFINITY, d2 = INFINITY;
        int k1 = -1, k2 = -1;
        const int c = cnt[l];
        auto push = [&](double d, int k) {
            if (precedes<DESC>(d, k, d1, k1)) { d2 = d1; k2 = k1; d1 = d; k1 = k; }
            else if (precedes<DESC>(d, k, d2, k2)) { d2 = d; k2 = k; }
        };
        if (c <= kCand) {
            for (int s = 0; s < c; ++s) push(cd[l * kCand + s], ck[l * kCand + s]);
            atomicAdd(&wst[0], (unsigned)c);
        } else {
            for (int k = 0; k < K; ++k) {
                const double e = s1(xs + l * ld + t0, cb.c + (size_t)k * blen, blen);
                push(ROUNDF ? (double)(float)e : e, k);
            }
            atomicAdd(&wst[1], 1u);
        }
        out1[l] = k1;
        if (out2) out2[l] = k2;
    }
    __syncthreads();
}

// LDS of one workgroup: rows [16][dim + 1] (float), re-check distances
// [16][kCand] and row norms [16] (double), counts [16], candidates
// [16][kCand], two result rows [16] (int), the block shift [dim] (float).
__host__ __device__ inline size_t encode_lds(int dim) {
    const int ld = dim + 1;
    return sizeof(float) * (kRows * ld + (kRows * ld) % 2) + sizeof(double) * (kRows * kCand + kRows) +
           sizeof(int) * (kRows + kRows * kCand + 2 * kRows) + sizeof(float) * dim;
}

// One wave per 16 vectors; 3 waves per SIMD keep the register tiles
// without spills.
template <int NC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_encode_mfma(
    const float *__restrict__ x, int64_t n, int dim, const float *__restrict__ cq, int mc,
    int kc, const float *__restrict__ pq, int mp, int kp, const uint4 *__restrict__ cfrag,
    const double *__restrict__ cn2,
    const unsigned long long *__restrict__ cmax, const float *__restrict__ cmean,
    const uint4 *__restrict__ pfrag,
    const double *__restrict__ pn2, const unsigned long long *__restrict__ pmax,
    const float *__restrict__ pmean, uint16_t *__restrict__ cid, uint8_t *__restrict__ codes,
    unsigned long long *__restrict__ stats, int coarse_given) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int ld = dim + 1;
    float *xs = reinterpret_cast<float *>(smem);                          // [16][dim+1]
    double *cd = reinterpret_cast<double *>(xs + kRows * ld + (kRows * ld) % 2);
    double *rown = cd + kRows * kCand;
    int *cnt = reinterpret_cast<int *>(rown + kRows);
    int *ck = cnt + kRows;
    int *o1 = ck + kRows * kCand;
    int *o2 = o1 + kRows;
    float *sm = reinterpret_cast<float *>(o2 + kRows);                   // [dim]
    const int l = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * kRows;
    const int nr = (int)min((int64_t)kRows, n - v0);
    const int d4 = dim / 4;                  // dim is a multiple of 16 here
#pragma unroll 4
    for (int e = l; e < kRows * d4; e += 64) {
        const int r = e / d4, c = e - r * d4;
        const float4 v = r < nr ? reinterpret_cast<const float4 *>(x + (size_t)(v0 + r) * dim)[c]
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float *dst = xs + r * ld + 4 * c;
        dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; dst[3] = v.w;
    }
    __shared__ unsigned s_wst[2];   // re-checked candidates, fully exact rows
    if (l == 0) { s_wst[0] = 0u; s_wst[1] = 0u; }
    __syncthreads();
    const int bsc = dim / mc, bsp = dim / mp;
    for (int j = 0; j < mc; ++j) {
        CbView cv{cq + (size_t)j * kc * bsc, cfrag + (size_t)j * ((kc + 15) / 16) * (bsc / 16) * 64,
                  cn2 + (size_t)j * kc,
                  __builtin_bit_cast(double, cmax[j]), cmean + (size_t)j * bsc, kc, bsc};
        if (coarse_given) {          // the nearest centroids are in cid already
            if (l < nr) o1[l] = cid[(size_t)(v0 + l) * NC * mc + j];
        } else {
            wave_best<NC == 2, NC == 2>(xs, ld, j * bsc, nr, cv, cnt, ck, cd, rown, sm, o1,
                                        NC == 2 ? o2 : nullptr, s_wst);
            if (l < nr) {
                if (NC == 2) {
                    cid[(size_t)(v0 + l) * 2 * mc + j] = (uint16_t)o1[l];
                    cid[(size_t)(v0 + l) * 2 * mc + mc + j] = (uint16_t)o2[l];
                } else {
                    cid[(size_t)(v0 + l) * mc + j] = (uint16_t)o1[l];
                }
            }
        }
        __syncthreads();
#pragma unroll 8
        for (int e = l; e < nr * bsc; e += 64) {   // o1 is written for real rows only
            int r = e / bsc, t = e % bsc;
            xs[r * ld + j * bsc + t] -= cv.c[(size_t)max(o1[r], 0) * bsc + t];
        }
        __syncthreads();
    }
    for (int s = 0; s < mp; ++s) {
        CbView pv{pq + (size_t)s * kp * bsp, pfrag + (size_t)s * ((kp + 15) / 16) * (bsp / 16) * 64,
                  pn2 + (size_t)s * kp,
                  __builtin_bit_cast(double, pmax[s]), pmean + (size_t)s * bsp, kp, bsp};
        wave_best<false, false>(xs, ld, s * bsp, nr, pv, cnt, ck, cd, rown, sm, o1, nullptr,
                                s_wst);
        if (l < nr) codes[(size_t)(v0 + l) * mp + s] = (uint8_t)o1[l];
        __syncthreads();
    }
    if (l == 0) {   // one update of the global statistics per workgroup
        if (s_wst[0]) atomicAdd(&stats[0], (unsigned long long)s_wst[0]);
        if (s_wst[1]) atomicAdd(&stats[1], (unsigned long long)s_wst[1]);
    }
}

// Lloyd assignment in the mixed form (SC_PREC_BF16 k-means): the label of
// each point is its nearest centre by the distance rounded to float, the
// later index on ties (k-means.h:597-612, 800-810), from BF16 MFMA distances
// re-checked exactly -- the labels of the exact form.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_assign_mfma(
    const float *__restrict__ x, int64_t n, int d, const float *__restrict__ c, int k,
    const uint4 *__restrict__ cfrag, const double *__restrict__ cn2,
    const unsigned long long *__restrict__ cmax, const float *__restrict__ cmean,
    int32_t *__restrict__ labels, unsigned long long *__restrict__ stats) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int ld = d + 1;
    float *xs = reinterpret_cast<float *>(smem);                          // [16][d+1]
    double *cd = reinterpret_cast<double *>(xs + kRows * ld + (kRows * ld) % 2);
    double *rown = cd + kRows * kCand;
    int *cnt = reinterpret_cast<int *>(rown + kRows);
    int *ck = cnt + kRows;
    int *o1 = ck + kRows * kCand;
    float *sm = reinterpret_cast<float *>(o1 + 2 * kRows);               // [d]
    const int l = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * kRows;
    const int nr = (int)min((int64_t)kRows, n - v0);
    const int d4 = d / 4;                    // d is a multiple of 16 here
#pragma unroll 4
    for (int e = l; e < kRows * d4; e += 64) {
        const int r = e / d4, cc = e - r * d4;
        const float4 v = r < nr ? reinterpret_cast<const float4 *>(x + (size_t)(v0 + r) * d)[cc]
                                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float *dst = xs + r * ld + 4 * cc;
        dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; dst[3] = v.w;
    }
    __shared__ unsigned s_wst[2];
    if (l == 0) { s_wst[0] = 0u; s_wst[1] = 0u; }
    __syncthreads();
    CbView cv{c, cfrag, cn2, __builtin_bit_cast(double, cmax[0]), cmean, k, d};
    wave_best<false, true, true>(xs, ld, 0, nr, cv, cnt, ck, cd, rown, sm, o1, nullptr, s_wst);
    if (l < nr) labels[v0 + l] = o1[l];
    if (l == 0) {
        if (s_wst[0]) atomicAdd(&stats[0], (unsigned long long)s_wst[0]);
        if (s_wst[1]) atomicAdd(&stats[1], (unsigned long long)s_wst[1]);
    }
}

// ---------------------------------------------------------------------------
// Tiled assignment (block lengths 32 and 64): the same approximate distances,
// bound and exact re-check as wave_best, organised for the matrix cores.
// A 256-thread workgroup takes 128 rows (32 per wave, two 16-row tiles); each
// wave keeps its rows' BF16 hi/lo A operands in registers for the whole
// block, and the codebook streams through LDS once per workgroup, 128
// codewords (eight 16-codeword tiles) at a time, in the operand order of
// v_mfma_f32_16x16x32_bf16 (lane l holds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16] for j < 8, and receives C[4 (l / 16) + i][l % 16]
// for i < 4).  Every B fragment read from LDS feeds both row tiles.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTRows = 128;        // rows per workgroup
constexpr int kTWaveRows = 32;     // rows per wave (two 16-row tiles)
constexpr int kTChunk = 8;         // codeword tiles (of 16) per LDS chunk
constexpr int kTCand = 24;         // exact re-check slots per row

// Fragments of a block for the tiled kernel: [KTp][KS][2][64] uint4 with
// KTp = KT rounded up to a multiple of kTChunk (zero past K), KS = blen / 32;
// lane l of (tile kt, step ks) holds codeword k = 16 kt + l % 16, dims
// t = 32 ks + 8 (l / 16) + j, j < 8: the BF16 hi parts ([0]) and lo parts ([1]).
__global__ void k_cb_frag32(const float *__restrict__ cb, int nb, int K, int blen,
                            const float *__restrict__ mean, uint4 *__restrict__ frag) {
    const int KTp = ((K + 15) / 16 + kTChunk - 1) / kTChunk * kTChunk, KS = blen / 32;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nb * KTp * KS * 64) return;
    const int l = (int)(e % 64);
    const int64_t f = e / 64;
    const int ks = (int)(f % KS), kt = (int)((f / KS) % KTp), b = (int)(f / ((int64_t)KS * KTp));
    const int k = 16 * kt + (l & 15), t = 32 * ks + 8 * (l >> 4);
    uint32_t hw[4], lw[4];
    for (int j = 0; j < 8; j += 2) {
        unsigned short h[2], lo[2];
        for (int u = 0; u < 2; ++u) {
            const float v = k < K ? cb[((size_t)b * K + k) * blen + t + j + u] -
                                        mean[(size_t)b * blen + t + j + u]
                                  : 0.0f;
            h[u] = bf16_bits(v);
            lo[u] = bf16_bits(v - bf16_val(h[u]));
        }
        hw[j / 2] = h[0] | ((uint32_t)h[1] << 16);
        lw[j / 2] = lo[0] | ((uint32_t)lo[1] << 16);
    }
    uint4 *o = frag + ((((size_t)b * KTp + kt) * KS + ks) * 2) * 64 + l;
    o[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    o[64] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

struct TileOut {
    uint16_t *cid;          // encode: cid[row * stride + off1] (and + off2 for the second)
    int stride, off1, off2;
    int32_t *labels;        // k-means labels (cid == null)
};

// One coarse block of the rows [128 blockIdx.x, +128): best one (TOP2 false)
// or two (TOP2 true) codewords of the block's codebook, the order of
// wave_best (DESC, ROUNDF as there).  KS = blen / 32 (1 or 2).
template <int KS, bool TOP2, bool DESC, bool ROUNDF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_assign_tiles(
    const float *__restrict__ x, int64_t n, int ldx, int t0, const float *__restrict__ c,
    const uint4 *__restrict__ frag, const double *__restrict__ cn2,
    const unsigned long long *__restrict__ cmaxb, const float *__restrict__ mean, int K,
    TileOut out, unsigned long long *__restrict__ stats) {
    constexpr int blen = 32 * KS;
    constexpr int kChunkU4 = kTChunk * KS * 2 * 64;          // uint4 per chunk
    // the codeword chunk; after the last chunk, the re-check distances
    constexpr int kBufU4 = kChunkU4 > kTRows * kTCand / 2 ? kChunkU4 : kTRows * kTCand / 2;
    __shared__ uint4 bufB[kBufU4];
    __shared__ float c2s[kTChunk * 16];
    __shared__ int cnt[kTRows];
    __shared__ int ck[kTRows * kTCand];
    __shared__ float sm[blen];
    __shared__ unsigned s_wst[2];
    const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * kTRows + wv * kTWaveRows;   // this wave's rows
    const int nr = (int)max((int64_t)0, min((int64_t)kTWaveRows, n - r0));
    for (int t = tid; t < blen; t += 256) sm[t] = mean[t];
    for (int i = tid; i < kTRows; i += 256) cnt[i] = 0;
    if (tid == 0) { s_wst[0] = 0u; s_wst[1] = 0u; }
    __syncthreads();
    int *wcnt = cnt + wv * kTWaveRows;
    int *wck = ck + wv * kTWaveRows * kTCand;
    // A operands (shifted, split) and FP64 norms of this wave's rows
    bf16x8 ah[2][KS], al[2][KS];
    double nrm[2];
    bool exact_local = true;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
        const int rr = rt * 16 + (l & 15);
        const bool ok = rr < nr;
        const float *xr = x + (size_t)(r0 + (ok ? rr : 0)) * ldx + t0;
        double s = 0.0;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int t = 32 * ks + 8 * (l >> 4) + j;
                const float v = ok ? xr[t] - sm[t] : 0.0f;
                const unsigned short h = bf16_bits(v);
                ah[rt][ks][j] = __builtin_bit_cast(__bf16, h);
                al[rt][ks][j] = __builtin_bit_cast(__bf16, bf16_bits(v - bf16_val(h)));
                exact_local &= (bf16_val(h) == v);
                s += (double)v * (double)v;
            }
        }
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        nrm[rt] = s;                      // the norm of row rt * 16 + l % 16
    }
    const bool x_exact = __all(exact_local);
    // this lane's C rows: rt * 16 + 4 (l / 16) + i
    const int col = l & 15, q4 = 4 * (l >> 4);
    float rf[2][4], E[2][4], m1[2][4], m2[2][4];
    const double u = 1.0 / 8388608.0, nn = 3.0 * blen + 8.0, gam = nn * u / (1.0 - nn * u);
    const double C = __builtin_bit_cast(double, cmaxb[0]);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double rn = __shfl(nrm[rt], q4 + i);
            const double X = sqrt(rn);
            const double e = 2.0 * (2.0 * X * C * ((3.0 / 65536.0 + gam) * (1.0 + 1.0 / 128.0)) +
                                    (X + C) * (X + C) / 4194304.0) +
                             (X + C) * (X + C) / 4194304.0 + 1e-6;
            E[rt][i] = nextafterf((float)e, INFINITY);     // rounded up: still a bound
            rf[rt][i] = (float)rn;
            m1[rt][i] = INFINITY;
            m2[rt][i] = INFINITY;
        }
    const int KT = (K + 15) / 16;
    const int KTp = (KT + kTChunk - 1) / kTChunk * kTChunk;
    for (int kc0 = 0; kc0 < KTp; kc0 += kTChunk) {
        __syncthreads();                                      // the previous chunk is consumed
        const uint4 *src = frag + (size_t)kc0 * KS * 2 * 64;
        for (int e = tid; e < kChunkU4; e += 256) bufB[e] = src[e];
        for (int e = tid; e < kTChunk * 16; e += 256) {
            const int k = 16 * kc0 + e;
            c2s[e] = k < K ? (float)cn2[k] : 0.0f;
        }
        __syncthreads();
        f32x4 acc[kTChunk][2];
#pragma unroll
        for (int g = 0; g < kTChunk; ++g) {
            acc[g][0] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            acc[g][1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int g = 0; g < kTChunk; ++g) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const uint4 fh = bufB[((g * KS + ks) * 2) * 64 + l];
                const uint4 fl = bufB[((g * KS + ks) * 2 + 1) * 64 + l];
                const bf16x8 bh = __builtin_bit_cast(bf16x8, fh);
                const bf16x8 bl = __builtin_bit_cast(bf16x8, fl);
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    acc[g][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt][ks], bh, acc[g][rt], 0, 0, 0);
                    acc[g][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rt][ks], bl, acc[g][rt], 0, 0, 0);
                    if (!x_exact)
                        acc[g][rt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rt][ks], bh, acc[g][rt], 0, 0, 0);
                }
            }
        }
        // d~ minus the row's norm (the same for all of a row's codewords, so
        // it changes neither the order nor the differences the threshold
        // uses; fewer roundings keep E a bound) in place, +inf past K; the
        // chunk's two smallest per row, the running pair and the threshold,
        // then the candidates
        const bool full = 16 * (kc0 + kTChunk) <= K;          // uniform
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
            float thr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float g1 = INFINITY, g2 = INFINITY;
#pragma unroll
                for (int g = 0; g < kTChunk; ++g) {
                    const float d0 = c2s[g * 16 + col] - 2.0f * acc[g][rt][i];
                    const float d = full || 16 * (kc0 + g) + col < K ? d0 : INFINITY;
                    acc[g][rt][i] = d;
                    if (d < g1) { g2 = g1; g1 = d; } else if (d < g2) { g2 = d; }
                }
                for (int o = 1; o < 16; o <<= 1) {
                    const float a1 = __shfl_xor(g1, o), a2 = __shfl_xor(g2, o);
                    const float lo1 = fminf(g1, a1), hi1 = fmaxf(g1, a1);
                    g2 = fminf(hi1, fminf(g2, a2));
                    g1 = lo1;
                }
                const float lo1 = fminf(m1[rt][i], g1), hi1 = fmaxf(m1[rt][i], g1);
                m2[rt][i] = fminf(hi1, fminf(m2[rt][i], g2));
                m1[rt][i] = lo1;
                const double mref = (double)(TOP2 ? m2[rt][i] : m1[rt][i]);
                // ROUNDF: the slack for float-rounded distances scales with
                // the full distance, row norm included
                thr[i] = nextafterf((float)(mref + 2.0 * (double)E[rt][i] +
                                            (ROUNDF ? (fabs(mref + (double)rf[rt][i]) +
                                                       (double)E[rt][i]) *
                                                          (1.0 / 4194304.0)
                                                    : 0.0)),
                                    INFINITY);
            }
#pragma unroll
            for (int g = 0; g < kTChunk; ++g) {
                const int k = 16 * (kc0 + g) + col;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = rt * 16 + q4 + i;
                    if (acc[g][rt][i] <= thr[i] && row < nr && (full || k < K)) {
                        const int sl = atomicAdd(&wcnt[row], 1);
                        if (sl < kTCand) wck[row * kTCand + sl] = k;
                    }
                }
            }
        }
    }
    __syncthreads();
    // exact re-check: the candidates' S1 distances (FP64) into the chunk
    // buffer, then each row's best one or two under the reference's order
    double *cd = reinterpret_cast<double *>(bufB) + wv * kTWaveRows * kTCand;
    int total = 0;
    for (int r = 0; r < nr; ++r) total += wcnt[r] <= kTCand ? wcnt[r] : 0;
    for (int p = l; p < total; p += 64) {
        int r = 0, sp = p;
        while (true) {
            const int cc = wcnt[r] <= kTCand ? wcnt[r] : 0;
            if (sp < cc) break;
            sp -= cc;
            ++r;
        }
        const int k = wck[r * kTCand + sp];
        const double e = s1(x + (size_t)(r0 + r) * ldx + t0, c + (size_t)k * blen, blen);
        cd[r * kTCand + sp] = ROUNDF ? (double)(float)e : e;
    }
    __syncthreads();
    if (l < nr) {
        double d1 = INFINITY, d2 = INFINITY;
        int k1 = -1, k2 = -1;
        const int cc = wcnt[l];
        auto push = [&](double d, int k) {
            if (precedes<DESC>(d, k, d1, k1)) { d2 = d1; k2 = k1; d1 = d; k1 = k; }
            else if (precedes<DESC>(d, k, d2, k2)) { d2 = d; k2 = k; }
        };
        const float *xr = x + (size_t)(r0 + l) * ldx + t0;
        if (cc <= kTCand) {
            for (int sp = 0; sp < cc; ++sp) push(cd[l * kTCand + sp], wck[l * kTCand + sp]);
            atomicAdd(&s_wst[0], (unsigned)cc);
        } else {
            for (int k = 0; k < K; ++k) {
                const double e = s1(xr, c + (size_t)k * blen, blen);
                push(ROUNDF ? (double)(float)e : e, k);
            }
            atomicAdd(&s_wst[1], 1u);
        }
        const int64_t row = r0 + l;
        if (out.cid) {
            out.cid[(size_t)row * out.stride + out.off1] = (uint16_t)k1;
            if (TOP2) out.cid[(size_t)row * out.stride + out.off2] = (uint16_t)k2;
        } else {
            out.labels[row] = k1;
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (s_wst[0]) atomicAdd(&stats[0], (unsigned long long)s_wst[0]);
        if (s_wst[1]) atomicAdd(&stats[1], (unsigned long long)s_wst[1]);
    }
}

int prep_codebook(sc_ctx *ctx, const float *cb, int nb, int K, int blen, int shift,
                  unsigned char *base, size_t &off, float **mean, uint4 **frag, double **cn2,
                  unsigned long long **cmax) {
    auto take = [&](size_t bytes) { unsigned char *p = base + off; off = (off + bytes + 255) & ~(size_t)255; return p; };
    const int64_t nfrag = (int64_t)nb * ((K + 15) / 16) * (blen / 16) * 64;
    *mean = (float *)take(sizeof(float) * nb * blen);
    *frag = (uint4 *)take(sizeof(uint4) * (size_t)nfrag);
    *cn2 = (double *)take(8ull * nb * K);
    *cmax = (unsigned long long *)take(8ull * nb);
    if (!base) return SC_OK;   // sizing pass
    SC_HIP(hipMemsetAsync(*cmax, 0, 8ull * nb, ctx->stream));
    hipLaunchKernelGGL(k_cb_mean, dim3((unsigned)(nb * blen)), dim3(256), 0, ctx->stream, cb, nb,
                       K, blen, shift, *mean);
    hipLaunchKernelGGL(k_cb_norms, dim3((unsigned)sc_cdiv((int64_t)nb * K, 256)), dim3(256), 0,
                       ctx->stream, cb, nb, K, blen, *mean, *cn2, *cmax);
    hipLaunchKernelGGL(k_cb_frag, dim3((unsigned)sc_cdiv(nfrag, 256)), dim3(256), 0, ctx->stream,
                       cb, nb, K, blen, *mean, *frag);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

template <bool TOP2, bool DESC, bool ROUNDF>
int launch_tiles(sc_ctx *ctx, const float *x, int64_t n, int ldx, int t0, int blen,
                 const float *c, const uint4 *frag, const double *cn2,
                 const unsigned long long *cmax, const float *mean, int K, TileOut out,
                 unsigned long long *stats) {
    const unsigned g = (unsigned)sc_cdiv(n, kTRows);
    if (blen == 64)
        hipLaunchKernelGGL((k_assign_tiles<2, TOP2, DESC, ROUNDF>), dim3(g), dim3(256), 0,
                           ctx->stream, x, n, ldx, t0, c, frag, cn2, cmax, mean, K, out, stats);
    else
        hipLaunchKernelGGL((k_assign_tiles<1, TOP2, DESC, ROUNDF>), dim3(g), dim3(256), 0,
                           ctx->stream, x, n, ldx, t0, c, frag, cn2, cmax, mean, K, out, stats);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

// Blocks the tiled kernel takes: lengths 32 and 64 (codebooks of any size).
inline bool tiles_ok(int blen) {
    if (const char *e = getenv("SC_ASSIGN_TILES"))   // test hook: 0 = the per-wave kernel
        if (atoi(e) == 0) return false;
    return blen == 32 || blen == 64;
}

int prep_frag32(sc_ctx *ctx, const float *cb, int nb, int K, int blen, const float *mean,
                unsigned char *base, size_t &off, uint4 **frag) {
    const int KTp = ((K + 15) / 16 + kTChunk - 1) / kTChunk * kTChunk;
    const int64_t nf = (int64_t)nb * KTp * (blen / 32) * 64;
    unsigned char *p = base ? base + off : nullptr;
    off = (off + sizeof(uint4) * 2 * (size_t)nf + 255) & ~(size_t)255;
    *frag = (uint4 *)p;
    if (!base) return SC_OK;
    hipLaunchKernelGGL(k_cb_frag32, dim3((unsigned)sc_cdiv(nf, 256)), dim3(256), 0, ctx->stream,
                       cb, nb, K, blen, mean, *frag);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

}  // namespace

// Scratch of kmeans_assign_bf16: the prepared centres and two statistics.
size_t kmeans_bf16_scratch(int k, int d) {
    float *mean;
    uint4 *frag, *frag32;
    double *cn2;
    unsigned long long *cmax;
    size_t off = 0;
    prep_codebook(nullptr, nullptr, 1, k, d, 1, nullptr, off, &mean, &frag, &cn2, &cmax);
    if (tiles_ok(d)) prep_frag32(nullptr, nullptr, 1, k, d, nullptr, nullptr, off, &frag32);
    return off + 256;
}

int kmeans_assign_bf16(sc_ctx *ctx, const float *x, int64_t n, int d, const float *c, int k,
                       int32_t *labels, unsigned char *scratch) {
    float *mean;
    uint4 *frag;
    double *cn2;
    unsigned long long *cmax;
    size_t off = 0;
    SC_TRY(prep_codebook(ctx, c, 1, k, d, 1, scratch, off, &mean, &frag, &cn2, &cmax));
    uint4 *frag32 = nullptr;
    const bool tiles = tiles_ok(d);
    if (tiles) SC_TRY(prep_frag32(ctx, c, 1, k, d, mean, scratch, off, &frag32));
    unsigned long long *stats = (unsigned long long *)(scratch + off);
    SC_HIP(hipMemsetAsync(stats, 0, 16, ctx->stream));
    if (tiles)
        return launch_tiles<false, true, true>(ctx, x, n, d, 0, d, c, frag32, cn2, cmax, mean, k,
                                               TileOut{nullptr, 0, 0, 0, labels}, stats);
    const size_t lds = encode_lds(d);
    SC_HIP(hipFuncSetAttribute((const void *)k_assign_mfma,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_assign_mfma, dim3((unsigned)sc_cdiv(n, kRows)), dim3(64), lds,
                       ctx->stream, x, n, d, c, k, frag, cn2, cmax, mean, labels, stats);
    SC_CHECK_LAUNCH();
    return SC_OK;
}

int sc_encode_bf16(sc_ctx *ctx, const float *x, int64_t n, int dim, const float *cq, int mc,
                   int kc, const float *pq, int mp, int kp, int nc, uint16_t *cid,
                   uint8_t *codes, sc_encode_stats *st) {
    const int bsc = dim / mc, bsp = dim / mp;
    if (bsc % 16 || bsp % 16 || dim > 2048) {
        sc_set_error("sc_encode: SC_PREC_BF16 needs block lengths that are multiples of 16 "
                     "(bsc=%d bsp=%d) and dim <= 2048", bsc, bsp);
        return SC_ERR_ARG;
    }
    float *cmean, *pmean;
    uint4 *cfrag, *pfrag;
    double *cn2, *pn2;
    unsigned long long *cmax, *pmax, *stats;
    size_t off = 0;
    const bool tiles = tiles_ok(bsc);
    uint4 *cfrag32 = nullptr;
    prep_codebook(ctx, cq, mc, kc, bsc, 1, nullptr, off, &cmean, &cfrag, &cn2, &cmax);
    prep_codebook(ctx, pq, mp, kp, bsp, 0, nullptr, off, &pmean, &pfrag, &pn2, &pmax);
    if (tiles) prep_frag32(ctx, cq, mc, kc, bsc, nullptr, nullptr, off, &cfrag32);
    const size_t total = off + 256;
    unsigned char *s;
    SC_TRY(sc_scratch(ctx, total, (void **)&s));
    off = 0;
    SC_TRY(prep_codebook(ctx, cq, mc, kc, bsc, 1, s, off, &cmean, &cfrag, &cn2, &cmax));
    SC_TRY(prep_codebook(ctx, pq, mp, kp, bsp, 0, s, off, &pmean, &pfrag, &pn2, &pmax));
    if (tiles) SC_TRY(prep_frag32(ctx, cq, mc, kc, bsc, cmean, s, off, &cfrag32));
    stats = (unsigned long long *)(s + off);
    SC_HIP(hipMemsetAsync(stats, 0, 16, ctx->stream));
    if (tiles && n > 0) {
        // the coarse blocks on the tiled kernel; the per-wave kernel then
        // takes the residuals and the PQ blocks
        const int KTp = ((kc + 15) / 16 + kTChunk - 1) / kTChunk * kTChunk;
        for (int j = 0; j < mc; ++j) {
            const uint4 *fj = cfrag32 + (size_t)j * KTp * (bsc / 32) * 2 * 64;
            TileOut o{cid, nc * mc, j, mc + j, nullptr};
            if (nc == 2)
                SC_TRY((launch_tiles<true, true, false>(ctx, x, n, dim, j * bsc, bsc,
                                                         cq + (size_t)j * kc * bsc, fj,
                                                         cn2 + (size_t)j * kc, cmax + j,
                                                         cmean + (size_t)j * bsc, kc, o, stats)));
            else
                SC_TRY((launch_tiles<false, false, false>(ctx, x, n, dim, j * bsc, bsc,
                                                           cq + (size_t)j * kc * bsc, fj,
                                                           cn2 + (size_t)j * kc, cmax + j,
                                                           cmean + (size_t)j * bsc, kc, o, stats)));
        }
    }
    const int given = tiles ? 1 : 0;
    const size_t lds = encode_lds(dim);
    const unsigned blocks = (unsigned)sc_cdiv(n, kRows);
    if (nc == 2) {
        SC_HIP(hipFuncSetAttribute((const void *)k_encode_mfma<2>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_encode_mfma<2>, dim3(blocks), dim3(64), lds, ctx->stream, x, n, dim,
                           cq, mc, kc, pq, mp, kp, cfrag, cn2, cmax, cmean, pfrag, pn2,
                           pmax, pmean, cid, codes, stats, given);
    } else {
        SC_HIP(hipFuncSetAttribute((const void *)k_encode_mfma<1>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_encode_mfma<1>, dim3(blocks), dim3(64), lds, ctx->stream, x, n, dim,
                           cq, mc, kc, pq, mp, kp, cfrag, cn2, cmax, cmean, pfrag, pn2,
                           pmax, pmean, cid, codes, stats, given);
    }
    SC_CHECK_LAUNCH();
    if (st) {
        unsigned long long h[2];
        SC_HIP(hipMemcpyAsync(h, stats, 16, hipMemcpyDeviceToHost, ctx->stream));
        SC_HIP(hipStreamSynchronize(ctx->stream));
        st->rechecked = (int64_t)h[0];
        st->fallback = (int64_t)h[1];
    }
    return SC_OK;
}
