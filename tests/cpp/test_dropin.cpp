// test_dropin.cpp -- the reference's stage sequence (test/test_quantizer.cpp,
// test/test_sc_encoder.cpp, test/test_sc_query.cpp, test/test_multi_query.cpp,
// test/test_query.cpp) driven through the drop-in classes of
// second-candidate_amd/host/sc_dropin.h, checked against the C oracle
// (oracle/liboracle.so; test infrastructure).  Needs a GPU.
//
// usage: test_dropin <scratch dir>     exit 0 = all checks passed
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../oracle/sc_oracle.h"
#include "../../second-candidate_amd/host/sc_dropin.h"

static int g_fail = 0;
#define EXPECT(c)                                                                  \
    do {                                                                           \
        if (!(c)) {                                                                \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);      \
            ++g_fail;                                                              \
        }                                                                          \
    } while (0)

static void write_fvecs(const std::string &fn, const std::vector<float> &x, int d) {
    FILE *f = std::fopen(fn.c_str(), "wb");
    size_t n = x.size() / d;
    for (size_t i = 0; i < n; ++i) {
        std::fwrite(&d, 4, 1, f);
        std::fwrite(&x[i * d], 4, d, f);
    }
    std::fclose(f);
}

// Rows of d bytes, each preceded by the int d (the .bvecs layout).
static void write_bvecs(const std::string &fn, const std::vector<float> &x, int d) {
    FILE *f = std::fopen(fn.c_str(), "wb");
    size_t n = x.size() / d;
    std::vector<unsigned char> row(d);
    for (size_t i = 0; i < n; ++i) {
        for (int j = 0; j < d; ++j) row[j] = (unsigned char)x[i * d + j];
        std::fwrite(&d, 4, 1, f);
        std::fwrite(row.data(), 1, d, f);
    }
    std::fclose(f);
}

static std::vector<unsigned char> slurp(const std::string &fn) {
    std::vector<unsigned char> b;
    FILE *f = std::fopen(fn.c_str(), "rb");
    if (!f) return b;
    int c;
    while ((c = std::fgetc(f)) != EOF) b.push_back((unsigned char)c);
    std::fclose(f);
    return b;
}

static std::vector<float> gen(uint32_t stream, int64_t n, int d) {
    sco_mix m = {20151001u, 16, 8, 0, 24 * 256, 208 * 256, 8 * 256, 24 * 256, 255};
    std::vector<float> x((size_t)n * d);
    sco_gen_rows(&m, stream, 0, n, d, x.data(), 8);
    return x;
}

static std::vector<float> read_ctr(const std::string &fn, int *k, int *m, int *dim) {
    SC::PQConfig c;
    float *cb = nullptr;
    SC::load_codebook<float>(fn.c_str(), c, cb, 0, false);
    *k = c.kc; *m = c.mc; *dim = c.dim;
    std::vector<float> v(cb, cb + (size_t)c.kc * c.dim);
    ::operator delete(cb);
    return v;
}

// The rows of an .fvecs file (4-byte header per row).
static std::vector<float> xt_of(const std::string &fn, int d) {
    FILE *f = std::fopen(fn.c_str(), "rb");
    std::vector<float> x;
    if (!f) return x;
    std::vector<unsigned char> row(4 + (size_t)d * 4);
    while (std::fread(row.data(), 1, row.size(), f) == row.size()) {
        const size_t at = x.size();
        x.resize(at + d);
        std::memcpy(&x[at], row.data() + 4, (size_t)d * 4);
    }
    std::fclose(f);
    return x;
}

// Trains cq (mc blocks, kc) and pq (8 x 256) on the residuals, like
// test_quantizer.cpp:58-117, and writes <dir>/<tag>_cq.ctr_, _pq.ctr_.
static void train(const std::string &dir, const std::string &tag, const std::string &train_fv,
                  int dim, int mc, int kc, bool second) {
    SC::kmeans_options().iterations = 6;
    SC::PQQuantizer<float> *cq = second ? new SC::SCQuantizer<float>(dim, mc, kc, 4, 1, false)
                                        : new SC::PQQuantizer<float>(dim, mc, kc, 4, 1, false);
    cq->load_data(train_fv.c_str(), false);
    EXPECT(cq->size() == 4000);
    cq->create_sub_quantizers(false);
    const std::string cfn = dir + "/" + tag + "_cq";
    cq->output(cfn.c_str(), false);
    // labels are the argmin of the trained centres (first block)
    const int bs = dim / mc;
    const float *c0 = cq->get_center_at(0, false);
    for (int r = 0; r < 50; ++r) {
        double best = 1e300;
        int lab = -1;
        for (int j = 0; j < kc; ++j) {
            double d = sco_dist_l2sq(cq->data + (size_t)r * dim, c0 + (size_t)j * bs, bs);
            if ((float)d <= best) { best = (float)d; lab = j; }
        }
        EXPECT(cq->get_label_at(0, false)[r] == lab);
    }
    EXPECT(cq->distortion(false) > 0.0);
    cq->set_params(1, 8, 256);
    std::string cpath = cfn + ".ctr_";
    char *names[1] = {&cpath[0]};
    cq->load_codebooks(names, false);
    // calc_residual_vector is not virtual (as in the reference): call the
    // SCQuantizer's own, as test_sc_quantizer.cpp:51-61 does through an
    // SCQuantizer pointer
    if (second)
        static_cast<SC::SCQuantizer<float> *>(cq)->calc_residual_vector(false);
    else
        cq->calc_residual_vector(false);
    // residuals as the reference's quantizer computes them: PQQuantizer over
    // the codebook's mc blocks; SCQuantizer over its 8 part blocks of dim/8
    // (include/sc_quantizer.h:96-124), second nearest
    {
        int k_, m_, d_;
        std::vector<float> cb = read_ctr(cpath, &k_, &m_, &d_);
        std::vector<float> want(xt_of(train_fv, dim));
        EXPECT(k_ == kc && m_ == mc && d_ == dim);
        sco_residual(want.data(), (int64_t)want.size() / dim, dim, cb.data(), second ? 8 : mc,
                     kc, second ? 1 : 0, 8);
        EXPECT(std::memcmp(cq->data, want.data(), want.size() * sizeof(float)) == 0);
    }
    cq->create_sub_quantizers(false);
    const std::string pfn = dir + "/" + tag + "_pq";
    cq->output(pfn.c_str(), false);
    delete cq;
}

struct OracleIx {
    std::vector<float> cq, pq, nc, nr, dcr;
    std::vector<int32_t> L, pid;
    std::vector<uint8_t> codes;
    sco_index s;
};

static void oracle_index(OracleIx &o, const std::vector<float> &x, int dim, int mc, int kc,
                         int nc) {
    const int64_t n = (int64_t)x.size() / dim;
    std::vector<uint16_t> cid((size_t)n * mc * nc);
    std::vector<uint8_t> codes((size_t)n * 8);
    if (nc == 3)   // SCEncoder::encode, general nc
        EXPECT(sco_encode_general(x.data(), n, dim, o.cq.data(), mc, kc, o.pq.data(), 8, 256, nc,
                                  cid.data(), codes.data(), 8) == 0);
    else
        EXPECT(sco_encode(x.data(), n, dim, o.cq.data(), mc, kc, o.pq.data(), 8, 256, nc,
                          cid.data(), codes.data(), 8) == 0);
    std::vector<uint64_t> keys(n);
    sco_cell_keys(cid.data(), n, mc, kc, nc, keys.data());
    int64_t cells = 1;
    for (int j = 0; j < mc * nc; ++j) cells *= kc;
    o.L.resize(cells);
    o.pid.resize(n);
    o.codes.resize((size_t)n * 8);
    EXPECT(sco_build_dense(keys.data(), n, cells, codes.data(), 8, o.L.data(), o.pid.data(),
                           o.codes.data()) == 0);
    o.nc.resize((size_t)mc * kc);
    o.nr.resize(8 * 256);
    o.dcr.resize((size_t)kc * 8 * 256);
    sco_tables(o.cq.data(), o.pq.data(), dim, mc, kc, 8, 256, o.nc.data(), o.nr.data(),
               o.dcr.data());
    o.s = sco_index{dim, mc, kc, 8, 256, o.cq.data(), o.pq.data(), o.nc.data(), o.nr.data(),
                    o.dcr.data(), cells, o.L.data(), o.pid.data(), o.codes.data()};
}

// The .edat_ file holds exactly the oracle's cells.
static void check_edat(const std::string &fn, const OracleIx &o, int64_t n) {
    FILE *f = std::fopen(fn.c_str(), "rb");
    EXPECT(f != nullptr);
    if (!f) return;
    int ne = 0, N = 0;
    EXPECT(std::fread(&ne, 4, 1, f) == 1 && std::fread(&N, 4, 1, f) == 1);
    EXPECT(N == n);
    int nonempty = 0, prev = 0;
    for (size_t c = 0; c < o.L.size(); ++c) {
        int l = 0;
        EXPECT(std::fread(&l, 4, 1, f) == 1);
        EXPECT(l == o.L[c] - prev);
        if (l > 0) {
            ++nonempty;
            std::vector<int> p(l);
            std::vector<uint8_t> cs((size_t)l * 8);
            EXPECT(std::fread(p.data(), 4, l, f) == (size_t)l);
            EXPECT(std::fread(cs.data(), 1, cs.size(), f) == cs.size());
            EXPECT(std::memcmp(p.data(), &o.pid[prev], 4 * (size_t)l) == 0);
            EXPECT(std::memcmp(cs.data(), &o.codes[(size_t)prev * 8], cs.size()) == 0);
        }
        prev = o.L[c];
    }
    EXPECT(ne == nonempty);
    EXPECT(std::fgetc(f) == 0 && std::fgetc(f) == EOF);   // the stretch byte
    std::fclose(f);
}

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const int dim = 128;
    std::vector<float> xt = gen(1, 4000, dim), xb = gen(2, 20000, dim), xq = gen(3, 40, dim);
    write_fvecs(dir + "/train.fvecs", xt, dim);
    write_fvecs(dir + "/base.fvecs", xb, dim);

    // ---------------- SC: mc=1, kc=32, (nearest, second) cells
    train(dir, "sc", dir + "/train.fvecs", dim, 1, 32, true);
    {
        SC::SCEncoder enc;
        enc.load_codebooks((dir + "/sc_cq.ctr_").c_str(), (dir + "/sc_pq.ctr_").c_str(), false);
        enc.encode2<float>((dir + "/base.fvecs").c_str(), 4, false);
        EXPECT(enc.get_config().N == 20000 && enc.get_config().mc == 1 && enc.get_config().mp == 8);
        enc.distribution(false);
        enc.output(dir.c_str(), "base", false);
        OracleIx o;
        int k, m, d;
        o.cq = read_ctr(dir + "/sc_cq.ctr_", &k, &m, &d);
        EXPECT(k == 32 && m == 1 && d == dim);
        o.pq = read_ctr(dir + "/sc_pq.ctr_", &k, &m, &d);
        EXPECT(k == 256 && m == 8);
        oracle_index(o, xb, dim, 1, 32, 2);
        check_edat(dir + "/base_mr2_ivf.edat_", o, 20000);

        SC::SCQuery q;
        q.load_codebooks((dir + "/sc_cq.ctr_").c_str(), (dir + "/sc_pq.ctr_").c_str(), false);
        EXPECT(q.load_encoded_data((dir + "/base_mr2_ivf.edat_").c_str(), false) == 20000);
        q.pre_compute1();
        const int w = 300, T = 1500, R = 100, kc = 32;
        float *v_tmp = new float[kc + w], *dist = nullptr;
        int *it = new int[kc], *result = nullptr;
        int *h1 = new int[w], *h2 = new int[w], *h3 = new int[w], *h4 = new int[w];
        int *s1 = new int[w], *s2 = new int[w], *pb = new int[w], *cache = new int[w];
        bool *trav = new bool[kc * kc]();
        trav[0] = true;
        for (int i = 0; i < 40; ++i) {
            int sum = 0;
            q.search_mr_ivf(&xq[(size_t)i * dim], v_tmp, dist, it, result, h1, h2, h3, h4, s1, s2,
                            pb, cache, trav, sum, R, w, T, kc, false, false);
            EXPECT(trav[0] && !trav[1]);   // test_sc_query.cpp:182-183
            std::vector<int32_t> oid(R);
            std::vector<float> od(R);
            sco_stats st;
            EXPECT(sco_search(&o.s, SCO_SC, &xq[(size_t)i * dim], R, w, T, kc, oid.data(),
                              od.data(), &st, nullptr, nullptr, nullptr, nullptr, nullptr, 0) == 0);
            EXPECT(sum == st.sum);
            for (int r = 0; r < (R < sum ? R : sum); ++r) {
                EXPECT(result[r] == oid[r]);
                EXPECT(std::memcmp(&dist[r], &od[r], 4) == 0);
            }
            ::operator delete(result);
            ::operator delete(dist);
        }
        std::vector<int> bres(40 * R), bs(40);
        std::vector<float> bd(40 * R);
        q.search_batch(xq.data(), 40, SC_KIND_SC, R, w, T, bres.data(), bd.data(), bs.data());
        for (int i = 0; i < 40; ++i) {
            std::vector<int32_t> oid(R);
            std::vector<float> od(R);
            sco_stats st;
            sco_search(&o.s, SCO_SC, &xq[(size_t)i * dim], R, w, T, kc, oid.data(), od.data(),
                       &st, nullptr, nullptr, nullptr, nullptr, nullptr, 0);
            EXPECT(std::memcmp(&bres[(size_t)i * R], oid.data(), 4 * (size_t)R) == 0);
        }
        EXPECT(q.get_full_size() == 32 && q.get_size() > 0 && q.entropy(32 * 32) > 0.0);

        // PQQuery::entropy (src/query.cpp:168-182) is the Shannon entropy in
        // bits of the cell sizes: files with known size distributions give
        // log2(8) = 3 (8 equal cells) and H(1/4, 3/4) = 2 - (3/4) log2 3.
        for (int shape = 0; shape < 2; ++shape) {
            const std::string efn = dir + "/shape_ivf.edat_";
            FILE *f = std::fopen(efn.c_str(), "wb");
            std::vector<int> len(32 * 32, 0);
            if (shape == 0)
                for (int c = 0; c < 8; ++c) len[(size_t)c * 97] = 2500;
            else {
                len[5] = 5000;
                len[1000] = 15000;
            }
            int ne = shape == 0 ? 8 : 2, N = 20000, next = 0;
            std::fwrite(&ne, 4, 1, f);
            std::fwrite(&N, 4, 1, f);
            std::vector<unsigned char> zc((size_t)N * 8, 0);
            for (int c = 0; c < 32 * 32; ++c) {
                std::fwrite(&len[(size_t)c], 4, 1, f);
                for (int v = 0; v < len[(size_t)c]; ++v, ++next) std::fwrite(&next, 4, 1, f);
                std::fwrite(zc.data(), 1, (size_t)len[(size_t)c] * 8, f);
            }
            std::fputc(0, f);
            std::fclose(f);
            SC::SCQuery qs;
            qs.load_codebooks((dir + "/sc_cq.ctr_").c_str(), (dir + "/sc_pq.ctr_").c_str(),
                              false);
            EXPECT(qs.load_encoded_data(efn.c_str(), false) == 20000);
            const double want = shape == 0 ? 3.0 : 2.0 - 0.75 * std::log2(3.0);
            EXPECT(std::fabs(qs.entropy(32 * 32) - want) < 1e-12);
            EXPECT(qs.get_size() == ne);
        }

        // real_dist = true: raw vectors loaded after the index (query.h:83-86)
        q.load_data<float>((dir + "/base.fvecs").c_str(), 4, false);   // 4-byte row header
        for (int i = 0; i < 8; ++i) {
            int sum = 0;
            q.search_mr_ivf(&xq[(size_t)i * dim], v_tmp, dist, it, result, h1, h2, h3, h4, s1, s2,
                            pb, cache, trav, sum, R, w, T, kc, true, false);
            std::vector<int32_t> oid(R);
            std::vector<float> od(R);
            sco_stats st;
            EXPECT(sco_search_real(&o.s, SCO_SC, &xq[(size_t)i * dim], xb.data(), R, w, T,
                                   oid.data(), od.data(), &st) == 0);
            EXPECT(sum == st.sum);
            for (int r = 0; r < (R < sum ? R : sum); ++r) {
                EXPECT(result[r] == oid[r]);
                EXPECT(std::memcmp(&dist[r], &od[r], 4) == 0);
            }
            ::operator delete(result);
            ::operator delete(dist);
        }
    }

    // ---------------- SC, three candidates: nc=3 over the same codebooks
    {
        SC::SCEncoder enc(3);
        enc.load_codebooks((dir + "/sc_cq.ctr_").c_str(), (dir + "/sc_pq.ctr_").c_str(), false);
        enc.encode<float>((dir + "/base.fvecs").c_str(), 4, false);
        enc.distribution(false);
        enc.output(dir.c_str(), "base", false);
        OracleIx o;
        int k, m, d;
        o.cq = read_ctr(dir + "/sc_cq.ctr_", &k, &m, &d);
        o.pq = read_ctr(dir + "/sc_pq.ctr_", &k, &m, &d);
        oracle_index(o, xb, dim, 1, 32, 3);
        check_edat(dir + "/base_mr3_ivf.edat_", o, 20000);
        SC::SCQuery q(3);
        q.load_codebooks((dir + "/sc_cq.ctr_").c_str(), (dir + "/sc_pq.ctr_").c_str(), false);
        EXPECT(q.load_encoded_data((dir + "/base_mr3_ivf.edat_").c_str(), false) == 20000);
        q.pre_compute1();
        const int w = 3000, T = 1500, R = 100, kc = 32;
        float *v_tmp = new float[kc + w], *dist = nullptr;
        int *it = new int[kc], *result = nullptr;
        int *h[6], *s1 = new int[w], *pb = new int[w], *cache = new int[w];
        for (int t = 0; t < 6; ++t) h[t] = new int[w];
        bool *trav = new bool[kc * kc * kc]();
        for (int i = 0; i < 40; ++i) {
            int sum = 0;
            q.search_mr_ivf3(&xq[(size_t)i * dim], v_tmp, dist, it, result, h[0], h[1], h[2], h[3],
                             h[4], h[5], s1, pb, cache, trav, sum, R, w, T, kc, false, false);
            std::vector<int32_t> oid(R);
            std::vector<float> od(R);
            sco_stats st;
            EXPECT(sco_search(&o.s, SCO_SC3, &xq[(size_t)i * dim], R, w, T, kc, oid.data(),
                              od.data(), &st, nullptr, nullptr, nullptr, nullptr, nullptr, 0) == 0);
            EXPECT(sum == st.sum && sum > 0);
            for (int r = 0; r < (R < sum ? R : sum); ++r) {
                EXPECT(result[r] == oid[r]);
                EXPECT(std::memcmp(&dist[r], &od[r], 4) == 0);
            }
            ::operator delete(result);
            ::operator delete(dist);
        }
    }

    // ---------------- IMI: mc=2, kc=16 (MultiQuery::search_multi2)
    train(dir, "imi", dir + "/train.fvecs", dim, 2, 16, false);
    {
        SC::Encoder enc;
        enc.load_codebooks((dir + "/imi_cq.ctr_").c_str(), (dir + "/imi_pq.ctr_").c_str(), false);
        enc.encode<float>((dir + "/base.fvecs").c_str(), 4, false);
        enc.distribution(false);
        enc.output(dir.c_str(), "base", false);
        OracleIx o;
        int k, m, d;
        o.cq = read_ctr(dir + "/imi_cq.ctr_", &k, &m, &d);
        EXPECT(k == 16 && m == 2);
        o.pq = read_ctr(dir + "/imi_pq.ctr_", &k, &m, &d);
        oracle_index(o, xb, dim, 2, 16, 1);
        check_edat(dir + "/base_ivf.edat_", o, 20000);

        // Encoder::statistic (src/encoder.cpp:152-168): ./ivf.txt lists every
        // cell's length, then one line "id:code code ... " per vector
        {
            char cwd[4096];
            EXPECT(getcwd(cwd, sizeof cwd) != nullptr);
            EXPECT(chdir(dir.c_str()) == 0);
            enc.statistic(false);
            EXPECT(chdir(cwd) == 0);
            std::string want;
            int prev = 0;
            char line[64];
            for (size_t c = 0; c < o.L.size(); ++c) {
                std::snprintf(line, sizeof line, "Bucket %zu:%d\n", c, o.L[c] - prev);
                want += line;
                for (int v = prev; v < o.L[c]; ++v) {
                    std::snprintf(line, sizeof line, "%d:", o.pid[v]);
                    want += line;
                    for (int t = 0; t < 8; ++t) {
                        std::snprintf(line, sizeof line, "%d ", (int)o.codes[(size_t)v * 8 + t]);
                        want += line;
                    }
                    want += "\n";
                }
                prev = o.L[c];
            }
            const std::vector<unsigned char> got = slurp(dir + "/ivf.txt");
            EXPECT(std::string(got.begin(), got.end()) == want);
        }

        // test_encoder.cpp:94-104: set_mp(8); load_encoded_data; rencode the
        // raw vectors (as bytes) in the file's cell order; output2.
        for (int variant = 0; variant < 2; ++variant) {
            // variant 0: the encoded vectors themselves -> the same file;
            // variant 1: other vectors of the same count, coded against the
            // loaded cells (residual per block, then the PQ argmin)
            std::vector<float> xr = xb;
            if (variant == 1)
                for (size_t i = 0; i < xr.size(); ++i) xr[i] = (float)(((int)xr[i] * 7 + 37) % 256);
            write_bvecs(dir + "/base.bvecs", xr, dim);
            SC::Encoder re;
            re.load_codebooks((dir + "/imi_cq.ctr_").c_str(), (dir + "/imi_pq.ctr_").c_str(),
                              false);
            re.set_mp(8);
            EXPECT(re.get_config().mp == 8);
            re.load_encoded_data((dir + "/base_ivf.edat_").c_str(), false);
            EXPECT(re.get_config().N == 20000);
            // cells, ids and cell digits as the file holds them
            const int kc = 16;
            int prev = 0;
            bool ok = true;
            for (size_t c = 0; c < o.L.size(); ++c) {
                ok = ok && re.cell_lengths()[c] == o.L[c] - prev;
                for (int v = prev; v < o.L[c]; ++v)
                    ok = ok && re.ids()[v] == o.pid[v] && re.cell_ids()[2 * v] == (int)(c / kc) &&
                         re.cell_ids()[2 * v + 1] == (int)(c % kc);
                prev = o.L[c];
            }
            EXPECT(ok);
            re.rencode<unsigned char>((dir + "/base.bvecs").c_str(), 4, false);
            EXPECT(re.get_config().N == 20000);
            // expected codes: the residual against the loaded cells, coded by
            // the oracle's encoder with one all-zero coarse centroid (x - 0 = x)
            const int bs = dim / 2;
            std::vector<float> res((size_t)20000 * dim), zero(dim, 0.0f);
            for (int i = 0; i < 20000; ++i)
                for (int j = 0; j < 2; ++j) {
                    const int cell = re.cell_ids()[2 * i + j];
                    for (int t = 0; t < bs; ++t)
                        res[(size_t)i * dim + j * bs + t] =
                            xr[(size_t)o.pid[i] * dim + j * bs + t] -
                            o.cq[((size_t)j * kc + cell) * bs + t];
                }
            std::vector<uint16_t> zc(20000);
            OracleIx o2;
            o2.L = o.L;
            o2.pid = o.pid;
            o2.codes.resize((size_t)20000 * 8);
            EXPECT(sco_encode(res.data(), 20000, dim, zero.data(), 1, 1, o.pq.data(), 8, 256, 1,
                              zc.data(), o2.codes.data(), 8) == 0);
            EXPECT(std::memcmp(re.pq_codes(), o2.codes.data(), o2.codes.size()) == 0);
            re.output2(dir.c_str(), "base_r", false);
            check_edat(dir + "/base_r_ivf.edat_", o2, 20000);
            if (variant == 0)
                EXPECT(slurp(dir + "/base_r_ivf.edat_") == slurp(dir + "/base_ivf.edat_"));
            else
                EXPECT(o2.codes != o.codes);
        }

        SC::MultiQuery q;
        q.load_codebooks((dir + "/imi_cq.ctr_").c_str(), (dir + "/imi_pq.ctr_").c_str(), false);
        EXPECT(q.load_encoded_data((dir + "/base_ivf.edat_").c_str(), false) == 20000);
        const int w = 100, T = 2000, R = 50, kc = 16;
        float *v_tmp = new float[2 * kc + w], *dist = nullptr, *qq = new float[2];
        int *it = new int[2 * kc], *result = nullptr;
        int *h1 = new int[w], *h2 = new int[w], *h3 = new int[w], *h4 = new int[w];
        int *s1 = new int[w], *s2 = new int[w], *pb = new int[w], *cache = new int[w];
        bool *trav = new bool[kc * kc]();
        double t1 = 0, t2 = 0;
        for (int i = 0; i < 40; ++i) {
            int sum = 0, e = 0;
            q.search_multi2(&xq[(size_t)i * dim], v_tmp, dist, qq, it, result, h1, h2, h3, h4, s1,
                            s2, pb, cache, trav, sum, R, w, T, kc, e, t1, t2, false, false);
            EXPECT(!trav[0]);   // test_multi_query.cpp:191
            std::vector<int32_t> oid(R);
            std::vector<float> od(R);
            sco_stats st;
            sco_search(&o.s, SCO_IMI, &xq[(size_t)i * dim], R, w, T, kc, oid.data(), od.data(),
                       &st, nullptr, nullptr, nullptr, nullptr, nullptr, 0);
            EXPECT(sum == st.sum && e == st.empty);
            for (int r = 0; r < (R < sum ? R : sum); ++r) EXPECT(result[r] == oid[r]);
            ::operator delete(result);
            ::operator delete(dist);
        }
        EXPECT(t1 > 0 && t2 > 0);
    }
    // ---------------- Evaluation (include/evaluation.h:34-140)
    {
        SC::Evaluation e;
        int *gt = nullptr;
        e.compute_groundtruth(xb.data(), 20000, xq.data(), 40, dim, gt);
        std::vector<int32_t> ogt(40);
        sco_bruteforce_nn1(xb.data(), 20000, xq.data(), 40, dim, ogt.data(), 8);
        EXPECT(std::memcmp(gt, ogt.data(), 4 * 40) == 0);
        // a ground-truth file of 5 ints per row, the first the neighbour
        const std::string gfn = dir + "/gt.ivecs";
        FILE *f = std::fopen(gfn.c_str(), "wb");
        for (int i = 0; i < 40; ++i) {
            int d5 = 5, row[5] = {ogt[i], 7, 8, 9, 10};
            std::fwrite(&d5, 4, 1, f);
            std::fwrite(row, 4, 5, f);
        }
        std::fclose(f);
        int *lg = nullptr, N = 0;
        e.load_groundtruth(gfn.c_str(), lg, 5, N, false);
        EXPECT(N == 40 && std::memcmp(lg, ogt.data(), 4 * 40) == 0);
        // results: the neighbour at rank 3 for even queries, absent for odd
        const int R = 5;
        std::vector<int> res(40 * R, -1), res2(40 * R, -1);
        for (int i = 0; i < 40; i += 2) res[(size_t)i * R + 3] = gt[i];
        for (int i = 1; i < 40; i += 4) res2[(size_t)i * R] = gt[i];
        int rec = -1;
        e.calc_recall(res.data(), gt, R, 40, rec, false);
        EXPECT(rec == 20);
        int *sets[2] = {res.data(), res2.data()};
        e.calc_comb_recall(sets, gt, R, 40, 2, rec, false);
        EXPECT(rec == 30);
        ::operator delete(gt);
        ::operator delete(lg);
    }
    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}
