/*
 * sc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see sc_oracle.h).
 *
 * A restatement of the reference algorithms, written from their behaviour;
 * every function cites the reference lines it follows.  Compile with
 * -ffp-contract=off and without -march=native.
 */
#include "sc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SCO_ERR_ARG (-1)
#define SCO_ERR_NOMEM (-2)
#define SCO_ERR_CONFIG (-3)

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11).  Checked against the Random123    */
/* known-answer vectors in tests/test_oracle.py.                           */
/* ---------------------------------------------------------------------- */
void sco_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
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

static int64_t mix_mu(const sco_mix *m, uint32_t c, uint32_t d) {
    uint32_t ctr[4] = {c, d, 0u, 0u}, key[2] = {m->seed, 0u}, o[4];
    sco_philox(ctr, key, o);
    return m->mu_lo + (int64_t)(o[0] % (uint64_t)(m->mu_span + 1));
}

static int64_t mix_sig(const sco_mix *m, uint32_t c) {
    uint32_t ctr[4] = {c, 0xFFFFFFFFu, 0u, 0u}, key[2] = {m->seed, 0u}, o[4];
    sco_philox(ctr, key, o);
    return m->sig_lo + (int64_t)(o[0] % (uint64_t)(m->sig_span + 1));
}

/* One value of the seeded Gaussian mixture.  Integer-only so that the GPU
 * generator reproduces it bit for bit: z is an Irwin-Hall sum of 12 uniform
 * 16-bit draws (unit variance after scaling by 2^-17). */
static float mix_value(const sco_mix *m, uint64_t row, uint32_t stream,
                       uint32_t d, int64_t mu_q, int64_t sig_q) {
    uint32_t key[2] = {m->seed, 0u}, a[4], b[4];
    uint32_t ca[4] = {(uint32_t)row, (uint32_t)(row >> 32), stream, 2u * d};
    uint32_t cb[4] = {(uint32_t)row, (uint32_t)(row >> 32), stream, 2u * d + 1u};
    sco_philox(ca, key, a);
    sco_philox(cb, key, b);
    uint32_t w[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
    int64_t s = 0;
    for (int i = 0; i < 6; ++i) s += (int64_t)(w[i] & 0xFFFFu) + (int64_t)(w[i] >> 16);
    int64_t z2 = 2 * s - 12 * 65535;
    int64_t num = mu_q * ((int64_t)1 << 17) + sig_q * z2;
    int shift = m->frac_bits + 17 - m->out_bits;
    int64_t xq = (num + ((int64_t)1 << (shift - 1))) >> shift;
    if (xq < 0) xq = 0;
    if (xq > m->maxq) xq = m->maxq;
    return (float)ldexp((double)xq, -m->out_bits);
}

void sco_gen_rows(const sco_mix *m, uint32_t stream, int64_t row0, int64_t nrows,
                  int32_t dim, float *out, int32_t nthreads) {
    int64_t *mu = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->ncomp * dim);
    int64_t *sg = (int64_t *)malloc(sizeof(int64_t) * (size_t)m->ncomp);
    for (int c = 0; c < m->ncomp; ++c) {
        sg[c] = mix_sig(m, (uint32_t)c);
        for (int d = 0; d < dim; ++d) mu[(size_t)c * dim + d] = mix_mu(m, (uint32_t)c, (uint32_t)d);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < nrows; ++i) {
        uint64_t row = (uint64_t)(row0 + i);
        uint32_t ctr[4] = {(uint32_t)row, (uint32_t)(row >> 32), stream, 0xFFFFFFFFu};
        uint32_t key[2] = {m->seed, 0u}, o[4];
        sco_philox(ctr, key, o);
        uint32_t c = o[0] % (uint32_t)m->ncomp;
        for (int d = 0; d < dim; ++d)
            out[(size_t)i * dim + d] = mix_value(m, row, stream, (uint32_t)d,
                                                 mu[(size_t)c * dim + d], sg[c]);
    }
    free(mu);
    free(sg);
}

/* ---------------------------------------------------------------------- */
/* S1: lib/simple-cluster/include/utilities.h:194-207 -- float subtraction, */
/* widened to double, squared and summed sequentially in double.           */
/* ---------------------------------------------------------------------- */
double sco_dist_l2sq(const float *x, const float *y, int32_t d) {
    double acc = 0.0;
    for (int32_t i = 0; i < d; ++i) {
        float t = x[i] - y[i];
        acc += (double)t * (double)t;
    }
    return acc;
}

/* ---------------------------------------------------------------------- */
/* Encoders                                                              */
/*  nc=1: Encoder::encode (include/encoder.h:69-192): per block strict `>` */
/*        from DBL_MAX -> lowest index wins a tie.                        */
/*  nc=2: SCEncoder::encode2 (include/sc_encoder.h:63-195): streaming     */
/*        top-2 with `>=` -> top-2 under (dist ASC, index DESC); residual */
/*        against the nearest; PQ strict `>`.                             */
/* ---------------------------------------------------------------------- */
int sco_encode(const float *x, int64_t n, int32_t dim, const float *cq,
               int32_t mc, int32_t kc, const float *pq, int32_t mp, int32_t kp,
               int32_t nc, uint16_t *cid, uint8_t *codes, int32_t nthreads) {
    if (mc <= 0 || mp <= 0 || dim % mc || dim % mp || kc < 1 || kp < 1 ||
        kp > 256 || kc > 65536 || (nc != 1 && nc != 2) || (nc == 2 && kc < 2))
        return SCO_ERR_ARG;
    const int32_t bsc = dim / mc, bsp = dim / mp;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < n; ++i) {
        float r[4096];
        const float *v = x + (size_t)i * dim;
        for (int32_t j = 0; j < mc; ++j) {
            const float *xb = v + (size_t)j * bsc;
            const float *cb = cq + (size_t)j * kc * bsc;
            int32_t nearest = 0;
            if (nc == 1) {
                double d = DBL_MAX;
                for (int32_t k = 0; k < kc; ++k) {
                    double t = sco_dist_l2sq(xb, cb + (size_t)k * bsc, bsc);
                    if (d > t) { d = t; nearest = k; }
                }
                cid[(size_t)i * mc + j] = (uint16_t)nearest;
            } else {
                double d = DBL_MAX, d2 = DBL_MAX;
                int32_t mn = 0, mn2 = 0;
                for (int32_t k = 0; k < kc; ++k) {
                    double t = sco_dist_l2sq(xb, cb + (size_t)k * bsc, bsc);
                    if (d >= t) { d2 = d; d = t; mn2 = mn; mn = k; }
                    else if (d2 >= t) { d2 = t; mn2 = k; }
                }
                nearest = mn;
                cid[(size_t)i * 2 * mc + j] = (uint16_t)mn;
                cid[(size_t)i * 2 * mc + mc + j] = (uint16_t)mn2;
            }
            const float *c = cb + (size_t)nearest * bsc;
            for (int32_t t = 0; t < bsc; ++t) r[j * bsc + t] = xb[t] - c[t];
        }
        for (int32_t s = 0; s < mp; ++s) {
            const float *pb = pq + (size_t)s * kp * bsp;
            double d = DBL_MAX;
            int32_t mn = 0;
            for (int32_t k = 0; k < kp; ++k) {
                double t = sco_dist_l2sq(r + (size_t)s * bsp, pb + (size_t)k * bsp, bsp);
                if (d > t) { d = t; mn = k; }
            }
            codes[(size_t)i * mp + s] = (uint8_t)mn;
        }
    }
    return 0;
}

/* SCEncoder::encode, general nc (include/sc_encoder.h:205-332): each
 * coarse distance is stored in a float (d_tmp[k], :276-277) and the nc
 * nearest are taken by nth_element_id + sort_id (:281-282) -- here the nc
 * smallest by (float dist ASC, index ASC), the canonical order for the
 * libstdc++ tie order; cid [N][mc][nc] (:285-287); residual against the
 * nearest; PQ argmin over float distances with strict `>` from FLT_MAX
 * (:306-318), so the lowest index wins a tie. */
int sco_encode_general(const float *x, int64_t n, int32_t dim, const float *cq,
                       int32_t mc, int32_t kc, const float *pq, int32_t mp, int32_t kp,
                       int32_t nc, uint16_t *cid, uint8_t *codes, int32_t nthreads) {
    if (mc <= 0 || mp <= 0 || dim % mc || dim % mp || kc < 1 || kp < 1 || kp > 256 ||
        kc > 65536 || nc < 1 || nc > 3 || nc > kc || dim > 4096)
        return SCO_ERR_ARG;
    const int32_t bsc = dim / mc, bsp = dim / mp;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < n; ++i) {
        float r[4096];
        const float *v = x + (size_t)i * dim;
        for (int32_t j = 0; j < mc; ++j) {
            const float *xb = v + (size_t)j * bsc;
            const float *cb = cq + (size_t)j * kc * bsc;
            float bd[3];
            int32_t bk[3];
            int32_t nb = 0;
            for (int32_t k = 0; k < kc; ++k) {
                const float d = (float)sco_dist_l2sq(xb, cb + (size_t)k * bsc, bsc);
                /* insert (d, k) into the sorted list of the nc best; k ascends, so
                 * an equal distance goes after the ones already held */
                int32_t p = nb;
                while (p > 0 && d < bd[p - 1]) --p;
                if (p >= nc) continue;
                int32_t e = nb < nc ? nb : nc - 1;
                for (; e > p; --e) { bd[e] = bd[e - 1]; bk[e] = bk[e - 1]; }
                bd[p] = d;
                bk[p] = k;
                if (nb < nc) ++nb;
            }
            for (int32_t t = 0; t < nc; ++t)
                cid[((size_t)i * mc + j) * nc + t] = (uint16_t)bk[t];
            const float *c = cb + (size_t)bk[0] * bsc;
            for (int32_t t = 0; t < bsc; ++t) r[j * bsc + t] = xb[t] - c[t];
        }
        for (int32_t s = 0; s < mp; ++s) {
            const float *pb = pq + (size_t)s * kp * bsp;
            float d = FLT_MAX;
            int32_t mn = 0;
            for (int32_t k = 0; k < kp; ++k) {
                const float t = (float)sco_dist_l2sq(r + (size_t)s * bsp, pb + (size_t)k * bsp, bsp);
                if (d > t) { d = t; mn = k; }
            }
            codes[(size_t)i * mp + s] = (uint8_t)mn;
        }
    }
    return 0;
}

/* S4: IMI key vector_base(cid, mc, kc) (include/sc_utilities.h:156-166,
 * src/encoder.cpp:180); SC key uid_j = nearest_j*kc + second_j,
 * key = sum uid_j (kc^2)^(mc-1-j) (src/sc_encoder.cpp:39-50), in 64 bits. */
void sco_cell_keys(const uint16_t *cid, int64_t n, int32_t mc, int32_t kc,
                   int32_t nc, uint64_t *keys) {
    for (int64_t i = 0; i < n; ++i) {
        uint64_t key = 0;
        if (nc == 1) {
            for (int32_t j = 0; j < mc; ++j) key = key * (uint64_t)kc + cid[(size_t)i * mc + j];
        } else if (nc == 3) {
            /* general encoder layout [N][mc][3]; uid_j = c0 kc^2 + c1 kc + c2
             * (src/sc_encoder.cpp:39-45 with the per-vector stride mc*nc) */
            const uint16_t *c = cid + (size_t)i * 3 * mc;
            for (int32_t j = 0; j < mc; ++j) {
                uint64_t uid = ((uint64_t)c[3 * j] * (uint64_t)kc + c[3 * j + 1]) * (uint64_t)kc +
                               c[3 * j + 2];
                key = key * (uint64_t)kc * (uint64_t)kc * (uint64_t)kc + uid;
            }
        } else {
            const uint16_t *c = cid + (size_t)i * 2 * mc;
            for (int32_t j = 0; j < mc; ++j) {
                uint64_t uid = (uint64_t)c[j] * (uint64_t)kc + c[mc + j];
                key = key * (uint64_t)kc * (uint64_t)kc + uid;
            }
        }
        keys[i] = key;
    }
}

/* Dense buckets in id order (src/encoder.cpp:170-205,
 * src/sc_encoder.cpp:17-60) flattened the way the loaders do
 * (src/query.cpp:82-165): L = inclusive prefix of cell sizes. */
int sco_build_dense(const uint64_t *keys, int64_t n, int64_t ncells,
                    const uint8_t *codes, int32_t mp, int32_t *L, int32_t *pid,
                    uint8_t *codes_sorted) {
    int64_t *cnt = (int64_t *)calloc((size_t)ncells + 1, sizeof(int64_t));
    if (!cnt) return SCO_ERR_NOMEM;
    for (int64_t i = 0; i < n; ++i) {
        if (keys[i] >= (uint64_t)ncells) { free(cnt); return SCO_ERR_ARG; }
        cnt[keys[i] + 1]++;
    }
    for (int64_t c = 0; c < ncells; ++c) cnt[c + 1] += cnt[c];
    for (int64_t c = 0; c < ncells; ++c) L[c] = (int32_t)cnt[c + 1];
    for (int64_t i = 0; i < n; ++i) {
        int64_t p = cnt[keys[i]]++;
        pid[p] = (int32_t)i;
        memcpy(codes_sorted + (size_t)p * mp, codes + (size_t)i * mp, (size_t)mp);
    }
    free(cnt);
    return 0;
}

/* Sparse buckets: the same members and order as sco_build_dense, with only
 * the non-empty cells kept, each named by its key (a stable counting of a
 * sorted copy: keys ascending, ids ascending within a key). */
typedef struct { uint64_t k; int64_t i; } key_id;
static int cmp_key_id(const void *a, const void *b) {
    const key_id *x = (const key_id *)a, *y = (const key_id *)b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    return (x->i > y->i) - (x->i < y->i);
}
int64_t sco_build_sparse(const uint64_t *keys, int64_t n, const uint8_t *codes, int32_t mp,
                         uint64_t *ckeys, int32_t *L, int32_t *pid, uint8_t *codes_sorted) {
    key_id *t = (key_id *)malloc(sizeof(key_id) * ((size_t)n + 1));
    if (!t) return SCO_ERR_NOMEM;
    for (int64_t i = 0; i < n; ++i) { t[i].k = keys[i]; t[i].i = i; }
    qsort(t, (size_t)n, sizeof(key_id), cmp_key_id);
    int64_t nc = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (p == 0 || t[p].k != t[p - 1].k) ckeys[nc++] = t[p].k;
        L[nc - 1] = (int32_t)(p + 1);
        pid[p] = (int32_t)t[p].i;
        memcpy(codes_sorted + (size_t)p * mp, codes + (size_t)t[p].i * mp, (size_t)mp);
    }
    free(t);
    return nc;
}

/* PQQuery::pre_compute1 (include/query.h:91-149): sequential FP32 sums. */
void sco_tables(const float *cq, const float *pq, int32_t dim, int32_t mc,
                int32_t kc, int32_t mp, int32_t kp, float *norm_c,
                float *norm_r, float *dot_cr) {
    const int32_t bsc = dim / mc, bsp = dim / mp, spb = mp / mc;
    for (int64_t i = 0; i < (int64_t)mc * kc; ++i) {
        float d = 0.0f;
        for (int32_t j = 0; j < bsc; ++j) { float t = cq[i * bsc + j]; d += t * t; }
        norm_c[i] = d;
    }
    for (int64_t i = 0; i < (int64_t)mp * kp; ++i) {
        float d = 0.0f;
        for (int32_t j = 0; j < bsp; ++j) { float t = pq[i * bsp + j]; d += t * t; }
        norm_r[i] = d;
    }
    size_t base = 0;
    for (int32_t i = 0; i < mc; ++i)
        for (int32_t j = 0; j < kc; ++j) {
            const float *c = cq + ((size_t)i * kc + j) * bsc;
            for (int32_t i1 = 0; i1 < spb; ++i1) {
                const float *p = pq + ((size_t)(i * spb + i1) * kp) * bsp;
                for (int32_t j1 = 0; j1 < kp; ++j1) {
                    float d = 0.0f;
                    for (int32_t t = 0; t < bsp; ++t) d += c[i1 * bsp + t] * p[(size_t)j1 * bsp + t];
                    dot_cr[base++] = 2.0f * d;
                }
            }
        }
}

/* Sequential FP32 dot product: the defined replacement for cblas_sdot. */
static float sdot_seq(const float *a, const float *b, int32_t n) {
    float s = 0.0f;
    for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* PQQuery::pre_compute2 (include/query.h:154-184), index-aligned stores. */
void sco_lut(const float *q, const float *cq, const float *pq,
             const float *norm_c, const float *norm_r, int32_t dim, int32_t mc,
             int32_t kc, int32_t mp, int32_t kp, float *diff_qc, float *diff_qr) {
    const int32_t bsc = dim / mc, bsp = dim / mp;
    for (int32_t i = 0; i < mc; ++i)
        for (int32_t j = 0; j < kc; ++j) {
            size_t b = (size_t)i * kc + j;
            float s = sdot_seq(q + (size_t)i * bsc, cq + b * bsc, bsc);
            diff_qc[b] = (norm_c[b] - s) - s;
        }
    for (int32_t i = 0; i < mp; ++i)
        for (int32_t j = 0; j < kp; ++j) {
            size_t b = (size_t)i * kp + j;
            float s = sdot_seq(q + (size_t)i * bsp, pq + b * bsp, bsp);
            diff_qr[b] = (norm_r[b] - s) - s;
        }
}

/* ---------------------------------------------------------------------- */
/* Heap replica: pq_insert4 / pq_pop4 (include/sc_utilities.h:469-503,     */
/* 546-585).  Parent of i is i>>1, children of i are {2i, 2i+1}; sift-down  */
/* stops while j+1 < size so a lone last child is never compared; inserts  */
/* into a full heap are dropped.                                           */
/* ---------------------------------------------------------------------- */
#define SWAPI(a, b) do { int32_t _t = (a); (a) = (b); (b) = _t; } while (0)
void sco_heap4_insert(sco_heap4 *h, float v, int32_t a, int32_t b, int32_t c,
                      int32_t d) {
    if (h->pos >= h->cap || h->pos < 0) return;
    int32_t i = h->pos++;
    h->v[i] = v; h->h1[i] = a; h->h2[i] = b; h->h3[i] = c; h->h4[i] = d;
    int32_t j = i >> 1;
    while (i > 0) {
        if (h->v[i] < h->v[j]) {
            float t = h->v[j]; h->v[j] = h->v[i]; h->v[i] = t;
            SWAPI(h->h1[i], h->h1[j]); SWAPI(h->h2[i], h->h2[j]);
            SWAPI(h->h3[i], h->h3[j]); SWAPI(h->h4[i], h->h4[j]);
        } else break;
        i = j;
        j = i >> 1;
    }
}

int sco_heap4_pop(sco_heap4 *h, int32_t *a, int32_t *b, int32_t *c, int32_t *d) {
    if (h->pos <= 0) return 0;
    *a = h->h1[0]; *b = h->h2[0]; *c = h->h3[0]; *d = h->h4[0];
    int32_t last = --h->pos;
    h->v[0] = h->v[last];
    h->h1[0] = h->h1[last]; h->h2[0] = h->h2[last];
    h->h3[0] = h->h3[last]; h->h4[0] = h->h4[last];
    int32_t i = 0, j = 0, mx;
    while (j + 1 < h->pos) {
        mx = (h->v[j] < h->v[j + 1]) ? j : j + 1;
        if (h->v[mx] < h->v[i]) {
            SWAPI(h->h1[i], h->h1[mx]); SWAPI(h->h2[i], h->h2[mx]);
            SWAPI(h->h3[i], h->h3[mx]); SWAPI(h->h4[i], h->h4[mx]);
            float t = h->v[i]; h->v[i] = h->v[mx]; h->v[mx] = t;
            i = mx;
        } else break;
        j = i << 1;
    }
    return 1;
}

/* Single-id heap of search_ivfadc: pq_insert/pq_pop (sc_utilities.h:229-300),
 * same array algorithm. */
void sco_heap1_insert(sco_heap1 *h, float v, int32_t id) {
    if (h->pos >= h->cap || h->pos < 0) return;
    int32_t i = h->pos++;
    h->v[i] = v; h->id[i] = id;
    int32_t j = i >> 1;
    while (i > 0) {
        if (h->v[i] < h->v[j]) {
            float t = h->v[j]; h->v[j] = h->v[i]; h->v[i] = t;
            SWAPI(h->id[i], h->id[j]);
        } else break;
        i = j; j = i >> 1;
    }
}
int32_t sco_heap1_pop(sco_heap1 *h) {
    if (h->pos <= 0) return -1;
    int32_t top = h->id[0];
    int32_t last = --h->pos;
    h->v[0] = h->v[last]; h->id[0] = h->id[last];
    int32_t i = 0, j = 0, mx;
    while (j + 1 < h->pos) {
        mx = (h->v[j] < h->v[j + 1]) ? j : j + 1;
        if (h->v[mx] < h->v[i]) {
            SWAPI(h->id[i], h->id[mx]);
            float t = h->v[i]; h->v[i] = h->v[mx]; h->v[mx] = t;
            i = mx;
        } else break;
        j = i << 1;
    }
    return top;
}

/* ---------------------------------------------------------------------- */
/* Coarse ranking: full sort by (score ASC, centroid ASC).                 */
/* ---------------------------------------------------------------------- */
typedef struct { float v; int32_t i; } fi_pair;
static int cmp_fi(const void *a, const void *b) {
    const fi_pair *x = (const fi_pair *)a, *y = (const fi_pair *)b;
    if (x->v < y->v) return -1;
    if (x->v > y->v) return 1;
    return (x->i > y->i) - (x->i < y->i);
}
static void sort_scores(float *v, int32_t *idx, int32_t n, fi_pair *tmp) {
    for (int32_t i = 0; i < n; ++i) { tmp[i].v = v[i]; tmp[i].i = idx[i]; }
    qsort(tmp, (size_t)n, sizeof(fi_pair), cmp_fi);
    for (int32_t i = 0; i < n; ++i) { v[i] = tmp[i].v; idx[i] = tmp[i].i; }
}

/* Top-R by (dist ASC, position ASC); the reference's nth_element + sort
 * (query.h:320-324) for inputs without distance ties. */
typedef struct { float d; int32_t pos; } dpos;
static int cmp_dpos(const void *a, const void *b) {
    const dpos *x = (const dpos *)a, *y = (const dpos *)b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->pos > y->pos) - (x->pos < y->pos);
}

int32_t sco_topr(const float *d, int32_t n, int32_t R, int32_t *pos_out) {
    dpos *c = (dpos *)malloc(sizeof(dpos) * ((size_t)n + 1));
    if (!c) return SCO_ERR_NOMEM;
    for (int32_t i = 0; i < n; ++i) { c[i].d = d[i]; c[i].pos = i; }
    qsort(c, (size_t)n, sizeof(dpos), cmp_dpos);
    int32_t m = n < R ? n : R;
    for (int32_t i = 0; i < m; ++i) pos_out[i] = c[i].pos;
    free(c);
    return m;
}

static inline int32_t cell_len(const int32_t *L, int64_t bid) {
    return bid > 0 ? L[bid] - L[bid - 1] : L[0];
}

/* ---------------------------------------------------------------------- */
/* Two-halves dense-partition search (SCO_SCIMI).  The reference defines the */
/* cell (src/sc_encoder.cpp:39-50: uid_h = nearest_h*kc + second_h, key =   */
/* uid_0*kc^2 + uid_1) but neither builds (26-33 abort) nor searches it.     */
/* This build's definition:                                                  */
/*  1. per half h: q_h = sequential FP32 sum of the half's squares, q_sum =  */
/*     q_0 + q_1, v_h[j] = q_h + diff_qc[h*kc + j] (include/multi_query.h:   */
/*     102-122);                                                             */
/*  2. every non-empty cell gets the score its rank pairs would have in the  */
/*     multi-sequences: per half the SC pair score v_h[nearest] +            */
/*     v_h[second] (include/sc_query.h:113-124, 151-166), and the IMI sum of */
/*     the two (include/multi_query.h:175-203):                              */
/*        score = (v_0[n_0] + v_0[s_0]) + (v_1[n_1] + v_1[s_1])   (FP32);     */
/*  3. cells are taken in (score ASC, key ASC) order while count < w and    */
/*     sum < T (search_mr_ivf's loop test, include/sc_query.h:130, w counts  */
/*     non-empty cells) -- the order a multi-sequence with an exact priority */
/*     queue would visit them in, without walking the ~kc^4 empty cells;    */
/*  4. the IMI scan (include/multi_query.h:209-262) with h3 = n_0, h4 = n_1: */
/*     each vector's residual is against its nearest centroid of each half   */
/*     (include/sc_encoder.h:159-163).                                       */
/* ---------------------------------------------------------------------- */
typedef struct { float s; uint64_t key; int64_t c; } cell_rank;
static int cmp_cell_rank(const void *a, const void *b) {
    const cell_rank *x = (const cell_rank *)a, *y = (const cell_rank *)b;
    if (x->s < y->s) return -1;
    if (x->s > y->s) return 1;
    return (x->key > y->key) - (x->key < y->key);
}

/* Directories of at least this many cells are ranked by selection (below);
 * smaller ones by sorting every cell.  Settable for tests. */
static int64_t scimi_select_min = (int64_t)1 << 20;
void sco_scimi_set_select_min(int64_t n) { scimi_select_min = n; }

static float scimi_score(const sco_index *ix, const float *v, uint64_t key) {
    const int64_t kc = ix->kc, kc2 = kc * kc;
    const int64_t u0 = (int64_t)(key / (uint64_t)kc2), u1 = (int64_t)(key % (uint64_t)kc2);
    const float a = v[u0 / kc] + v[u0 % kc];
    const float b = v[kc + u1 / kc] + v[kc + u1 % kc];
    return a + b;
}

/* Order-preserving map of a float onto uint32 (-0.0 and +0.0 share a key). */
static uint32_t fkey(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* Steps 2-3 for one query (v [2][kc] by centroid id); cells/s1/s2 [w] get the
 * directory index and the two nearest centroids of each taken cell.
 * Large directories: the taken cells are a prefix of the (score, key) order,
 * so they all lie in the histogram buckets (top 16 bits of the score's
 * order-preserving key) up to the first one at which the running count
 * reaches w or the running size sum reaches T; only the cells of those
 * buckets are sorted.  Small directories: every cell is sorted. */
static int scimi_rank(const sco_index *ix, const float *v, int32_t w, int32_t T,
                      int32_t *prebuck, int32_t *s1, int32_t *s2, int32_t *count_o,
                      int32_t *sum_o) {
    const int64_t kc = ix->kc, kc2 = kc * kc;
    int64_t nr = ix->ncells;
    cell_rank *r;
    if (ix->ncells >= scimi_select_min) {
        enum { NB = 65536 };
        uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)ix->ncells + 1));
        int64_t *hc = (int64_t *)calloc(NB, sizeof(int64_t));
        int64_t *hs = (int64_t *)calloc(NB, sizeof(int64_t));
        if (!keys || !hc || !hs) { free(keys); free(hc); free(hs); return SCO_ERR_NOMEM; }
        for (int64_t c = 0; c < ix->ncells; ++c) {
            keys[c] = fkey(scimi_score(ix, v, ix->ckeys[c]));
            hc[keys[c] >> 16] += 1;
            hs[keys[c] >> 16] += cell_len(ix->L, c);
        }
        int64_t cc = 0, cs = 0;
        uint32_t last = NB - 1;
        for (uint32_t b = 0; b < NB; ++b) {
            cc += hc[b];
            cs += hs[b];
            if (cc >= w || cs >= T) { last = b; break; }
        }
        nr = 0;
        for (int64_t c = 0; c < ix->ncells; ++c) nr += (keys[c] >> 16) <= last;
        r = (cell_rank *)malloc(sizeof(cell_rank) * ((size_t)nr + 1));
        if (!r) { free(keys); free(hc); free(hs); return SCO_ERR_NOMEM; }
        int64_t j = 0;
        for (int64_t c = 0; c < ix->ncells; ++c) {
            if ((keys[c] >> 16) > last) continue;
            r[j].s = scimi_score(ix, v, ix->ckeys[c]);
            r[j].key = ix->ckeys[c];
            r[j].c = c;
            ++j;
        }
        free(keys);
        free(hc);
        free(hs);
    } else {
        r = (cell_rank *)malloc(sizeof(cell_rank) * ((size_t)ix->ncells + 1));
        if (!r) return SCO_ERR_NOMEM;
        for (int64_t c = 0; c < ix->ncells; ++c) {
            r[c].s = scimi_score(ix, v, ix->ckeys[c]);
            r[c].key = ix->ckeys[c];
            r[c].c = c;
        }
    }
    qsort(r, (size_t)nr, sizeof(cell_rank), cmp_cell_rank);
    int32_t count = 0, sum = 0;
    for (int64_t i = 0; i < nr && count < w && sum < T; ++i) {
        const int64_t c = r[i].c;
        const uint64_t key = r[i].key;
        prebuck[count] = (int32_t)c;
        s1[count] = (int32_t)(key / (uint64_t)kc2 / (uint64_t)kc);
        s2[count] = (int32_t)(key % (uint64_t)kc2 / (uint64_t)kc);
        count++;
        sum += cell_len(ix->L, c);
    }
    free(r);
    *count_o = count;
    *sum_o = sum;
    return 0;
}

int sco_search(const sco_index *ix, int32_t kind, const float *q, int32_t R,
               int32_t w, int32_t T, int32_t M, int32_t *ids, float *dists,
               sco_stats *st, int32_t *cells_out, int32_t *s1_out,
               int32_t *s2_out, int32_t *all_ids, float *all_dists,
               int64_t cap_all) {
    const int32_t dim = ix->dim, mc = ix->mc, kc = ix->kc, mp = ix->mp, kp = ix->kp;
    (void)M; /* coarse lists are fully sorted (see header) */
    if (R < 1 || w < 1) return SCO_ERR_ARG;
    if (kind == SCO_IVFADC && mc != 1) return SCO_ERR_CONFIG;
    if (kind == SCO_IMI && mc != 2) return SCO_ERR_CONFIG;
    if (kind == SCO_SC && mc != 1) return SCO_ERR_CONFIG;
    if (kind == SCO_SC3 && (mc != 1 || kc > 1024 || ix->ncells < (int64_t)kc * kc * kc))
        return SCO_ERR_CONFIG;
    if (kind == SCO_SCIMI && (mc != 2 || !ix->ckeys || kc < 2)) return SCO_ERR_CONFIG;
    if (kind != SCO_SCIMI && ix->ckeys) return SCO_ERR_CONFIG;
    if (kind == SCO_IVFADC && w > kc) return SCO_ERR_ARG; /* pq_pop on empty -> exit */

    float *diff_qc = (float *)malloc(sizeof(float) * (size_t)mc * kc);
    float *diff_qr = (float *)malloc(sizeof(float) * (size_t)mp * kp);
    float *v = (float *)malloc(sizeof(float) * ((size_t)mc * kc + (size_t)w + 4));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)mc * kc);
    fi_pair *stmp = (fi_pair *)malloc(sizeof(fi_pair) * (size_t)kc);
    int32_t hcap = (kind == SCO_IVFADC) ? kc : w;
    float *hv = (float *)malloc(sizeof(float) * (size_t)hcap);
    int32_t *hb = (int32_t *)malloc(sizeof(int32_t) * (size_t)hcap * 4);
    int32_t *prebuck = (int32_t *)malloc(sizeof(int32_t) * (size_t)w);
    int32_t *s1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)w);
    int32_t *s2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)w);
    unsigned char *trav = NULL;
    if (kind != SCO_IVFADC && kind != SCO_SCIMI)
        trav = (unsigned char *)calloc((size_t)kc * kc * (kind == SCO_SC3 ? (size_t)kc : 1), 1);
    int rc = 0;
    if (!diff_qc || !diff_qr || !v || !tmp || !stmp || !hv || !hb || !prebuck ||
        !s1 || !s2 || (kind != SCO_IVFADC && kind != SCO_SCIMI && !trav)) {
        rc = SCO_ERR_NOMEM;
        goto done;
    }

    sco_lut(q, ix->cq, ix->pq, ix->norm_c, ix->norm_r, dim, mc, kc, mp, kp, diff_qc, diff_qr);

    float q_sum = 0.0f, qh[2] = {0.0f, 0.0f};
    if (kind == SCO_IMI || kind == SCO_SCIMI) {
        /* multi_query.h:102-112 */
        int32_t bsc = dim / mc;
        for (int32_t i = 0; i < mc; ++i) {
            float d1 = 0.0f;
            for (int32_t j = 0; j < bsc; ++j) { float t = q[i * bsc + j]; d1 += t * t; }
            q_sum += d1;
            qh[i] = d1;
        }
    } else {
        /* query.h:225-230, sc_query.h:95-100 */
        for (int32_t i = 0; i < dim; ++i) { float t = q[i]; q_sum += t * t; }
    }

    int32_t count = 0, count2 = 0, sum = 0, e = 0, exhausted = 0;
    if (kind == SCO_IVFADC) {
        /* query.h:232-259 */
        sco_heap1 h = {hv, hb, 0, kc};
        for (int32_t i = 0; i < kc; ++i) sco_heap1_insert(&h, q_sum + diff_qc[i], i);
        h.pos = kc;
        for (int32_t i = 0; i < w; ++i) {
            int32_t bid = sco_heap1_pop(&h);
            int32_t l = cell_len(ix->L, bid);
            prebuck[count] = bid; s1[count] = bid; s2[count] = 0; count++;
            sum += l;
            count2++;
            if (sum >= T) break;
        }
    } else if (kind == SCO_IMI) {
        /* multi_query.h:114-205 */
        for (int32_t i = 0; i < mc; ++i)
            for (int32_t j = 0; j < kc; ++j) {
                v[i * kc + j] = qh[i] + diff_qc[i * kc + j];
                tmp[i * kc + j] = j;
            }
        sort_scores(v, tmp, kc, stmp);
        sort_scores(v + kc, tmp + kc, kc, stmp);
        const float *v2 = v, *v3 = v + kc;
        const int32_t *t1 = tmp, *t2 = tmp + kc;
        sco_heap4 h = {hv, hb, hb + hcap, hb + 2 * hcap, hb + 3 * hcap, 0, w};
        int32_t h1 = 0, h2 = 0, h3 = t1[0], h4 = t2[0], i = 0;
        sco_heap4_insert(&h, v2[0] + v3[0], 0, 0, h3, h4);
        while (count2 < w && sum < T) {
            if (h.pos > 0) {
                sco_heap4_pop(&h, &h1, &h2, &h3, &h4);
                int64_t bid = (int64_t)h3 * kc + h4;
                i = h1 * kc + h2;
                int32_t l = cell_len(ix->L, bid);
                if (l > 0) {
                    prebuck[count] = (int32_t)bid; s1[count] = h3; s2[count] = h4;
                    count++;
                    sum += l;
                } else e++;
                trav[i] = 1;
                count2++;
            } else { exhausted = 1; break; }
            if (sum < T) {
                if (h1 < kc - 1 && (h2 == 0 || trav[i + kc - 1]))
                    sco_heap4_insert(&h, v2[h1 + 1] + v3[h2], h1 + 1, h2, t1[h1 + 1], t2[h2]);
                if (h2 < kc - 1 && (h1 == 0 || trav[i - kc + 1]))
                    sco_heap4_insert(&h, v2[h1] + v3[h2 + 1], h1, h2 + 1, t1[h1], t2[h2 + 1]);
            }
        }
    } else if (kind == SCO_SCIMI) {
        for (int32_t i = 0; i < mc; ++i)
            for (int32_t j = 0; j < kc; ++j) v[i * kc + j] = qh[i] + diff_qc[i * kc + j];
        rc = scimi_rank(ix, v, w, T, prebuck, s1, s2, &count, &sum);
        if (rc) goto done;
        count2 = count;
        exhausted = count < w && sum < T;   /* every non-empty cell taken */
    } else if (kind == SCO_SC3) {
        /* sc_query.h:296-402: (nearest, second, third) cells over one coarse
         * list; the heap holds rank triples, capacity w */
        for (int32_t j = 0; j < kc; ++j) { v[j] = q_sum + diff_qc[j]; tmp[j] = j; }
        sort_scores(v, tmp, kc, stmp);
        const int64_t k1 = kc, k2 = (int64_t)kc * kc, size2 = k2 * kc;
        sco_heap4 h = {hv, hb, hb + hcap, hb + 2 * hcap, hb + 3 * hcap, 0, w};
        sco_heap4_insert(&h, 3.0f * v[0], 0, 0, 0, 0);
        int32_t h1 = 0, h2 = 0, h3 = 0, unused = 0;
        int64_t i = 0;
        while (count < w && sum < T && count2 < size2) {
            if (h.pos > 0) {
                sco_heap4_pop(&h, &h1, &h2, &h3, &unused);
                int64_t bid = (int64_t)tmp[h1] * k2 + (int64_t)tmp[h2] * k1 + tmp[h3];
                int32_t l = bid > 0 ? ix->L[bid] - ix->L[bid - 1] : 0;
                if (l > 0) {
                    prebuck[count] = (int32_t)bid; s1[count] = tmp[h1]; s2[count] = 0;
                    count++;
                    sum += l;
                } else e++;
                i = k2 * h1 + k1 * h2 + h3;
                trav[i] = 1;
                count2++;
            } else { exhausted = 1; break; }
            if (sum < T) {
                /* a triple enters once its other predecessors have been popped */
                if (h1 < kc - 1 && (h2 == 0 || trav[i + k2 - k1]) && (h3 == 0 || trav[i + k2 - 1]))
                    sco_heap4_insert(&h, v[h1 + 1] + v[h2] + v[h3], h1 + 1, h2, h3, 0);
                if (h2 < kc - 1 && (h3 == 0 || trav[i + k1 - 1]) && (h1 == 0 || trav[i + k1 - k2]))
                    sco_heap4_insert(&h, v[h1] + v[h2 + 1] + v[h3], h1, h2 + 1, h3, 0);
                if (h3 < kc - 1 && (h2 == 0 || trav[i - k1 + 1]) && (h1 == 0 || trav[i - k2 + 1]))
                    sco_heap4_insert(&h, v[h1] + v[h2] + v[h3 + 1], h1, h2, h3 + 1, 0);
            }
        }
    } else {
        /* sc_query.h:102-186 */
        for (int32_t j = 0; j < kc; ++j) { v[j] = q_sum + diff_qc[j]; tmp[j] = j; }
        sort_scores(v, tmp, kc, stmp);
        sco_heap4 h = {hv, hb, hb + hcap, hb + 2 * hcap, hb + 3 * hcap, 0, w};
        float d0 = v[0] + v[1];
        sco_heap4_insert(&h, d0, 0, 1, tmp[0], tmp[1]);
        sco_heap4_insert(&h, d0, 1, 0, tmp[1], tmp[0]);
        int32_t h1 = 0, h2 = 0, h3 = 0, h4 = 0, i = 0;
        while (count < w && sum < T) {
            if (h.pos > 0) {
                sco_heap4_pop(&h, &h1, &h2, &h3, &h4);
                int64_t bid = (int64_t)h3 * kc + h4;
                int32_t l = bid > 0 ? ix->L[bid] - ix->L[bid - 1] : 0;
                if (l > 0) {
                    prebuck[count] = (int32_t)bid; s1[count] = h3; s2[count] = h4;
                    count++;
                    sum += l;
                }
                i = kc * h1 + h2;
                trav[i] = 1;
                count2++;
            } else { exhausted = 1; break; }
            if (sum < T) {
                if (h1 < kc - 1 && (h2 == 0 || trav[i + kc - 1]))
                    sco_heap4_insert(&h, v[h1 + 1] + v[h2], h1 + 1, h2, tmp[h1 + 1], tmp[h2]);
                if (h2 < kc - 1 && (h1 == 0 || trav[i - kc + 1]))
                    sco_heap4_insert(&h, v[h1] + v[h2 + 1], h1, h2 + 1, tmp[h1], tmp[h2 + 1]);
            }
        }
    }

    /* ADC scan: query.h:263-318, multi_query.h:209-262, sc_query.h:197-239 */
    dpos *cand = (dpos *)malloc(sizeof(dpos) * ((size_t)sum + 1));
    int32_t *cid_of = (int32_t *)malloc(sizeof(int32_t) * ((size_t)sum + 1));
    if (!cand || !cid_of) { free(cand); free(cid_of); rc = SCO_ERR_NOMEM; goto done; }
    int32_t pos = 0;
    const int32_t bs = kp * mp / mc;
    for (int32_t c = 0; c < count; ++c) {
        int64_t bid = prebuck[c];
        int32_t l = cell_len(ix->L, bid);
        int64_t start = bid > 0 ? ix->L[bid - 1] : 0;
        const uint8_t *cc = ix->codes + (size_t)start * mp;
        if (kind == SCO_IMI || kind == SCO_SCIMI) {
            int32_t h3 = s1[c], h4 = s2[c], half = mp >> 1;
            float d0 = (q_sum + diff_qc[h3]) + diff_qc[kc + h4];
            const float *r1 = ix->dot_cr + (size_t)h3 * bs;
            const float *r2 = ix->dot_cr + (size_t)(h4 + kc - 1) * bs;
            for (int32_t j = 0; j < l; ++j) {
                float d = d0;
                for (int32_t k = 0; k < half; ++k) {
                    int32_t b = k * kp + cc[k];
                    d += diff_qr[b];
                    d += r1[b];
                }
                for (int32_t k = half; k < mp; ++k) {
                    int32_t b = k * kp + cc[k];
                    d += diff_qr[b];
                    d += r2[b];
                }
                cc += mp;
                cand[pos].d = d; cand[pos].pos = pos; cid_of[pos] = ix->pid[start + j]; pos++;
            }
        } else {
            int32_t h3 = s1[c];
            float d0 = q_sum + diff_qc[h3];
            const float *r1 = ix->dot_cr + (size_t)h3 * bs;
            for (int32_t j = 0; j < l; ++j) {
                float d = d0;
                for (int32_t k = 0; k < mp; ++k) {
                    int32_t b = k * kp + cc[k];
                    d += (diff_qr[b] + r1[b]);
                }
                cc += mp;
                cand[pos].d = d; cand[pos].pos = pos; cid_of[pos] = ix->pid[start + j]; pos++;
            }
        }
        if (pos >= T) break;
    }
    if (all_ids || all_dists)
        for (int32_t p = 0; p < pos && p < cap_all; ++p) {
            if (all_ids) all_ids[p] = cid_of[p];
            if (all_dists) all_dists[p] = cand[p].d;
        }
    if (cells_out) for (int32_t c = 0; c < count; ++c) cells_out[c] = prebuck[c];
    if (s1_out) for (int32_t c = 0; c < count; ++c) s1_out[c] = s1[c];
    if (s2_out) for (int32_t c = 0; c < count; ++c) s2_out[c] = s2[c];

    /* Top-R: query.h:320-324 / multi_query.h:268-272 / sc_query.h:249-252 */
    if (sum >= R) qsort(cand, (size_t)sum, sizeof(dpos), cmp_dpos);
    for (int32_t r = 0; r < R; ++r) {
        if (r < sum) { ids[r] = cid_of[cand[r].pos]; dists[r] = cand[r].d; }
        else { ids[r] = -1; dists[r] = INFINITY; }
    }
    if (st) {
        st->sum = sum; st->ncells = count; st->nvisited = count2; st->empty = e;
        st->exhausted = exhausted;
    }
    free(cand);
    free(cid_of);
done:
    free(diff_qc); free(diff_qr); free(v); free(tmp); free(stmp); free(hv); free(hb);
    free(prebuck); free(s1); free(s2); free(trav);
    return rc;
}

/* The searches with real_dist = true (include/query.h:307-312,
 * include/multi_query.h:255-259, include/sc_query.h:232-237): the same
 * traversal and candidates as sco_search, each candidate's distance the S1
 * distance between the query and its raw vector raw[id] (stored as float),
 * then the same top-R (canonical order; traversal order below R). */
int sco_search_real(const sco_index *ix, int32_t kind, const float *q, const float *raw,
                    int32_t R, int32_t w, int32_t T, int32_t *ids, float *dists,
                    sco_stats *st) {
    sco_stats s0;
    const int64_t n = ix->L[ix->ncells - 1];   /* vectors in the index */
    int32_t *all = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *tid = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    float *td = (float *)malloc(sizeof(float) * (size_t)R);
    if (!all || !tid || !td) { free(all); free(tid); free(td); return SCO_ERR_NOMEM; }
    int rc = sco_search(ix, kind, q, R, w, T, ix->kc, tid, td, &s0, NULL, NULL, NULL, all,
                        NULL, n);
    if (rc == 0) {
        const int32_t sum = s0.sum;
        dpos *cand = (dpos *)malloc(sizeof(dpos) * ((size_t)sum + 1));
        if (!cand) rc = SCO_ERR_NOMEM;
        else {
            for (int32_t p = 0; p < sum; ++p) {
                cand[p].d = (float)sco_dist_l2sq(q, raw + (size_t)all[p] * ix->dim, ix->dim);
                cand[p].pos = p;
            }
            if (sum >= R) qsort(cand, (size_t)sum, sizeof(dpos), cmp_dpos);
            for (int32_t r = 0; r < R; ++r) {
                if (r < sum) { ids[r] = all[cand[r].pos]; dists[r] = cand[r].d; }
                else { ids[r] = -1; dists[r] = INFINITY; }
            }
            if (st) *st = s0;
            free(cand);
        }
    }
    free(all); free(tid); free(td);
    return rc;
}

int sco_search_batch(const sco_index *ix, int32_t kind, const float *q,
                     int64_t nq, int32_t R, int32_t w, int32_t T, int32_t M,
                     int32_t *ids, float *dists, sco_stats *st, int32_t nthreads) {
    int rc = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < nq; ++i) {
        int r = sco_search(ix, kind, q + (size_t)i * ix->dim, R, w, T, M,
                           ids + (size_t)i * R, dists + (size_t)i * R,
                           st ? st + i : NULL, NULL, NULL, NULL, NULL, NULL, 0);
        if (r) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            rc = r;
        }
    }
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Ground truth and recall (include/evaluation.h:58-92: a query counts when */
/* its ground-truth 1-NN is among the first R returned ids).               */
/* ---------------------------------------------------------------------- */
void sco_bruteforce_nn1(const float *base, int64_t n, const float *q, int64_t nq,
                        int32_t dim, int32_t *gt, int32_t nthreads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < nq; ++i) {
        double best = DBL_MAX;
        int32_t bi = -1;
        for (int64_t j = 0; j < n; ++j) {
            double t = sco_dist_l2sq(q + (size_t)i * dim, base + (size_t)j * dim, dim);
            if (t < best) { best = t; bi = (int32_t)j; }
        }
        gt[i] = bi;
    }
}

double sco_recall(const int32_t *gt, const int32_t *ids, int64_t nq, int32_t R,
                  int32_t stride) {
    int64_t hit = 0;
    for (int64_t i = 0; i < nq; ++i)
        for (int32_t r = 0; r < R; ++r)
            if (ids[(size_t)i * stride + r] == gt[i]) { hit++; break; }
    return nq > 0 ? (double)hit / (double)nq : 0.0;
}

/* ---------------------------------------------------------------------- */
/* k-means++ (lib/simple-cluster/include/k-means.h:162-267) with a        */
/* reproducible sampler: weights w_i = floor(D2_i * 2^16) in uint64, pivot */
/* p = u64 % W from Philox(key = {seed, 1}), pick the first i whose        */
/* inclusive prefix exceeds p.                                             */
/* ---------------------------------------------------------------------- */
static uint64_t kpp_u64(uint32_t seed, uint32_t round, uint32_t which) {
    uint32_t ctr[4] = {round, 0u, 0xA5u, which}, key[2] = {seed, 1u}, o[4];
    sco_philox(ctr, key, o);
    return ((uint64_t)o[1] << 32) | o[0];
}

int sco_kmeans_pp(const float *x, int64_t n, int32_t d, int32_t k, uint32_t seed,
                  int32_t lagged, int64_t *seed_idx, int32_t nthreads) {
    if (n < 1 || k < 1 || d < 1) return SCO_ERR_ARG;
    double *d2 = (double *)malloc(sizeof(double) * (size_t)n);
    if (!d2) return SCO_ERR_NOMEM;
    seed_idx[0] = (int64_t)(kpp_u64(seed, 0, 0) % (uint64_t)n);
    const float *c0 = x + (size_t)seed_idx[0] * d;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < n; ++i) d2[i] = sco_dist_l2sq(x + (size_t)i * d, c0, d);
    for (int32_t cnt = 1; cnt < k; ++cnt) {
        uint64_t W = 0;
        for (int64_t i = 0; i < n; ++i) W += (uint64_t)ldexp(d2[i], 16);
        uint64_t r = kpp_u64(seed, (uint32_t)cnt, 1);
        int64_t pick;
        if (W == 0) {
            pick = (int64_t)(r % (uint64_t)n);
        } else {
            uint64_t p = r % W, acc = 0;
            pick = n - 1;
            for (int64_t i = 0; i < n; ++i) {
                acc += (uint64_t)ldexp(d2[i], 16);
                if (acc > p) { pick = i; break; }
            }
        }
        seed_idx[cnt] = pick;
        /* k-means.h:244-262: fold in the newest seed (lagged: the previous) */
        const float *cs = x + (size_t)seed_idx[lagged ? cnt - 1 : cnt] * d;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
        for (int64_t i = 0; i < n; ++i) {
            double t = sco_dist_l2sq(x + (size_t)i * d, cs, d);
            if (d2[i] > t) d2[i] = t;
        }
    }
    free(d2);
    return 0;
}

/* Lloyd iterations: assignment as greg_initialize / greg_kmeans
 * (k-means.h:590-613, 798-812: float-rounded distance, `min >= d` so the
 * later centroid wins a tie), SINGLETON repair (k-means.h:845-877 with
 * find_lonely 506-533), update c = sum / size (update_center 351-374) with
 * FP64 sums. */
int sco_kmeans(const float *x, int64_t n, int32_t d, int32_t k, int32_t iters,
               uint32_t seed, int32_t lagged, float *centers, int32_t *labels,
               int32_t nthreads) {
    if (n < 1 || k < 1 || d < 1 || iters < 0) return SCO_ERR_ARG;
    if (n < k) {
        /* k-means.h:686-701: too few points, copy them */
        for (int32_t i = 0; i < k; ++i) {
            for (int32_t t = 0; t < d; ++t)
                centers[(size_t)i * d + t] = i < n ? x[(size_t)i * d + t] : FLT_MAX;
            if (i < n) labels[i] = i;
        }
        return 0;
    }
    int64_t *sidx = (int64_t *)malloc(sizeof(int64_t) * (size_t)k);
    double *sums = (double *)malloc(sizeof(double) * (size_t)k * d);
    int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * (size_t)k);
    if (!sidx || !sums || !cnt) { free(sidx); free(sums); free(cnt); return SCO_ERR_NOMEM; }
    int rc = sco_kmeans_pp(x, n, d, k, seed, lagged, sidx, nthreads);
    if (rc) goto out;
    for (int32_t i = 0; i < k; ++i)
        memcpy(centers + (size_t)i * d, x + (size_t)sidx[i] * d, sizeof(float) * d);
    for (int32_t it = 0; it < iters; ++it) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
        for (int64_t i = 0; i < n; ++i) {
            float mn = FLT_MAX;
            int32_t lab = -1;
            for (int32_t j = 0; j < k; ++j) {
                float t = (float)sco_dist_l2sq(centers + (size_t)j * d, x + (size_t)i * d, d);
                if (mn >= t) { mn = t; lab = j; }
            }
            labels[i] = lab;
        }
        memset(sums, 0, sizeof(double) * (size_t)k * d);
        memset(cnt, 0, sizeof(int64_t) * (size_t)k);
        for (int64_t i = 0; i < n; ++i) {
            int32_t l = labels[i];
            cnt[l]++;
            for (int32_t t = 0; t < d; ++t) sums[(size_t)l * d + t] += (double)x[(size_t)i * d + t];
        }
        for (int32_t j = 0; j < k; ++j) {
            if (cnt[j] > 0) continue;
            float best = -1.0f;
            int64_t fst = 0;
            for (int64_t i = 0; i < n; ++i) {
                float t = (float)sco_dist_l2sq(x + (size_t)i * d,
                                              centers + (size_t)labels[i] * d, d);
                if (best < t) { best = t; fst = i; }
            }
            int32_t old = labels[fst];
            for (int32_t t = 0; t < d; ++t) {
                float v = x[(size_t)fst * d + t];
                centers[(size_t)j * d + t] = v;
                sums[(size_t)old * d + t] -= (double)v;
                sums[(size_t)j * d + t] = (double)v;
            }
            cnt[old]--;
            cnt[j] = 1;
            labels[fst] = j;
        }
        for (int32_t j = 0; j < k; ++j)
            for (int32_t t = 0; t < d; ++t)
                centers[(size_t)j * d + t] = (float)(sums[(size_t)j * d + t] / (double)cnt[j]);
    }
out:
    free(sidx); free(sums); free(cnt);
    return rc;
}

/* PQQuantizer::calc_residual_vector (include/quantizer.h:456-510): nearest by
 * float-rounded S1 with strict `>`; SCQuantizer (include/sc_quantizer.h:
 * 72-133): second-nearest, float compares, strict `>`, blocks = mc. */
void sco_residual(float *x, int64_t n, int32_t dim, const float *cq, int32_t mc,
                  int32_t kc, int32_t which, int32_t nthreads) {
    const int32_t bs = dim / mc;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = 0; i < n; ++i) {
        float *v = x + (size_t)i * dim;
        for (int32_t j = 0; j < mc; ++j) {
            const float *cb = cq + (size_t)j * kc * bs;
            float d = FLT_MAX, d2 = FLT_MAX;
            int32_t p = 0, p2 = 0;
            for (int32_t k = 0; k < kc; ++k) {
                float t = (float)sco_dist_l2sq(v + j * bs, cb + (size_t)k * bs, bs);
                if (d > t) { d2 = d; p2 = p; d = t; p = k; }
                else if (d2 > t) { d2 = t; p2 = k; }
            }
            const float *c = cb + (size_t)(which ? p2 : p) * bs;
            for (int32_t t = 0; t < bs; ++t) v[j * bs + t] -= c[t];
        }
    }
}
