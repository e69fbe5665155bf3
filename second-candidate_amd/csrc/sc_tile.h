// sc_tile.h -- exact distance tile shared by the encoder and k-means.
//
// Every (vector, centroid) distance is the reference's S1 value
// (lib/simple-cluster/include/utilities.h:194-207): float subtraction,
// widened to double, squared and accumulated in index order in double.  The
// square of a float is exact in double, so the fused multiply-add below
// rounds exactly like the reference's separate multiply and add.
//
// A workgroup owns TV vectors held transposed in LDS ([dim][TV]) and walks
// the codebook in chunks of TK centroids x kTT dimensions staged in LDS
// ([kTT][TK]); each thread keeps a 4x4 (vector x centroid) register tile.
#pragma once

#include "sc_internal.h"

namespace sc_tile {

constexpr int kTT = 16;  // dimensions per staged codebook slice

struct Best2 {
    double d1, d2;
    int k1, k2;
};

// (d, k) precedes (e, j) under (dist ASC, index DESC) when DESC, else
// (dist ASC, index ASC).
template <bool DESC>
__device__ __forceinline__ bool precedes(double d, int k, double e, int j) {
    if (d < e) return true;
    if (!(d == e)) return false;   // d > e, or a NaN: the reference's `>`/`>=` never take it
    return DESC ? (k > j) : (k < j);
}

template <bool DESC>
__device__ __forceinline__ void push2(Best2 &b, double d, int k) {
    if (precedes<DESC>(d, k, b.d1, b.k1)) {
        b.d2 = b.d1; b.k2 = b.k1; b.d1 = d; b.k1 = k;
    } else if (precedes<DESC>(d, k, b.d2, b.k2)) {
        b.d2 = d; b.k2 = k;
    }
}

// ROUNDF compares the distances rounded to float (greg_kmeans and the
// quantizer residuals store them in a float: k-means.h:597-612,
// quantizer.h:475-482).
// Distances of the TV vectors held in xs (rows [t0, t0+blen) of the
// transposed tile) to K centroids cb[K][blen] (global), reduced to the best
// two (or best one) per vector under the tie order.  Results land in
// out_k1/out_k2 (LDS, TV entries) and, when out_d1 is given, the best
// distance in out_d1.  Must be called by all threads.
template <int TV, int TK, bool DESC, bool ROUNDF = false>
__device__ void tile_best(const float *xs, int t0, int blen, const float *__restrict__ cb,
                          int K, float *cs, Best2 *red, int *out_k1, int *out_k2,
                          double *out_d1 = nullptr) {
    constexpr int NT = (TV / 4) * (TK / 4);
    const int tid = threadIdx.x;
    const int tv = tid % (TV / 4), tk = tid / (TV / 4);
    Best2 b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) { b[i].d1 = b[i].d2 = INFINITY; b[i].k1 = b[i].k2 = -1; }

    for (int k0 = 0; k0 < K; k0 += TK) {
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int s0 = 0; s0 < blen; s0 += kTT) {
            const int tt = min(kTT, blen - s0);
            __syncthreads();
            for (int e = tid; e < TK * kTT; e += NT) {
                int kl = e / kTT, tl = e % kTT, k = k0 + kl;
                float v = 0.0f;
                if (k < K && tl < tt) v = cb[(size_t)k * blen + s0 + tl];
                cs[tl * TK + kl] = v;
            }
            __syncthreads();
            for (int tl = 0; tl < tt; ++tl) {
                const float4 xv = *reinterpret_cast<const float4 *>(xs + (size_t)(t0 + s0 + tl) * TV + tv * 4);
                const float4 cv = *reinterpret_cast<const float4 *>(cs + tl * TK + tk * 4);
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
                const float ca[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double dd = (double)(xa[i] - ca[j]);
                        acc[i][j] = __fma_rn(dd, dd, acc[i][j]);
                    }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + tk * 4 + j;
            if (k < K) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    push2<DESC>(b[i], ROUNDF ? (double)(float)acc[i][j] : acc[i][j], k);
            }
        }
    }
    // cross-thread merge: partial lists of the TK/4 threads sharing a vector
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) red[tk * TV + tv * 4 + i] = b[i];
    __syncthreads();
    if (tid < TV) {
        Best2 m;
        m.d1 = m.d2 = INFINITY;
        m.k1 = m.k2 = -1;
        for (int p = 0; p < TK / 4; ++p) {
            const Best2 o = red[p * TV + tid];
            if (o.k1 >= 0) push2<DESC>(m, o.d1, o.k1);
            if (o.k2 >= 0) push2<DESC>(m, o.d2, o.k2);
        }
        out_k1[tid] = m.k1;
        if (out_k2) out_k2[tid] = m.k2;
        if (out_d1) out_d1[tid] = m.d1;
    }
    __syncthreads();
}

// The NB best (float-rounded distance, index) pairs under (dist ASC, index
// ASC); empty slots (k < 0) come after every candidate.
template <int NB>
struct TopF {
    float d[NB];
    int k[NB];
};

template <int NB>
__device__ __forceinline__ void topf_push(TopF<NB> &t, float d, int k) {
    int p = NB;
    while (p > 0 && (t.k[p - 1] < 0 || d < t.d[p - 1] || (d == t.d[p - 1] && k < t.k[p - 1]))) --p;
    if (p >= NB) return;
    for (int e = NB - 1; e > p; --e) { t.d[e] = t.d[e - 1]; t.k[e] = t.k[e - 1]; }
    t.d[p] = d;
    t.k[p] = k;
}

// tile_best's walk with every distance rounded to float (SCEncoder::encode
// stores them in float arrays, include/sc_encoder.h:276-277, 306-314) and the
// NB best kept per vector; out_k [TV][NB] (LDS).  Must be called by all
// threads.
template <int TV, int TK, int NB>
__device__ void tile_top(const float *xs, int t0, int blen, const float *__restrict__ cb, int K,
                         float *cs, TopF<NB> *red, int *out_k) {
    constexpr int NT = (TV / 4) * (TK / 4);
    const int tid = threadIdx.x;
    const int tv = tid % (TV / 4), tk = tid / (TV / 4);
    TopF<NB> b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < NB; ++e) { b[i].d[e] = INFINITY; b[i].k[e] = -1; }

    for (int k0 = 0; k0 < K; k0 += TK) {
        double acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
        for (int s0 = 0; s0 < blen; s0 += kTT) {
            const int tt = min(kTT, blen - s0);
            __syncthreads();
            for (int e = tid; e < TK * kTT; e += NT) {
                int kl = e / kTT, tl = e % kTT, k = k0 + kl;
                float v = 0.0f;
                if (k < K && tl < tt) v = cb[(size_t)k * blen + s0 + tl];
                cs[tl * TK + kl] = v;
            }
            __syncthreads();
            for (int tl = 0; tl < tt; ++tl) {
                const float4 xv = *reinterpret_cast<const float4 *>(xs + (size_t)(t0 + s0 + tl) * TV + tv * 4);
                const float4 cv = *reinterpret_cast<const float4 *>(cs + tl * TK + tk * 4);
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
                const float ca[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double dd = (double)(xa[i] - ca[j]);
                        acc[i][j] = __fma_rn(dd, dd, acc[i][j]);
                    }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + tk * 4 + j;
            if (k < K) {
#pragma unroll
                for (int i = 0; i < 4; ++i) topf_push<NB>(b[i], (float)acc[i][j], k);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) red[tk * TV + tv * 4 + i] = b[i];
    __syncthreads();
    if (tid < TV) {
        TopF<NB> m;
#pragma unroll
        for (int e = 0; e < NB; ++e) { m.d[e] = INFINITY; m.k[e] = -1; }
        for (int p = 0; p < TK / 4; ++p) {
            const TopF<NB> o = red[p * TV + tid];
#pragma unroll
            for (int e = 0; e < NB; ++e)
                if (o.k[e] >= 0) topf_push<NB>(m, o.d[e], o.k[e]);
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) out_k[tid * NB + e] = m.k[e];
    }
    __syncthreads();
}

}  // namespace sc_tile
