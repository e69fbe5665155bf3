// sc_dropin.h -- the reference's C++ class API (namespace SC) on top of the
// C ABI in include/sc_hip.h.  Header-only; link with libsc_amd.so.
//
// Names, signatures, argument meaning, file formats and ownership follow the
// reference headers cited per class; every computation runs in the HIP
// library (there is no CPU compute path here).  Errors the reference answers
// with exit(1)/cerr keep that behaviour: I/O failures print and exit, a
// search kind that does not match the codebooks prints and returns with the
// outputs untouched.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <vector>

#include "../../include/sc_hip.h"

namespace SC {

typedef unsigned short ushort;

// include/sc_utilities.h:801-805
typedef struct {
    int N, kc, kp, mc, mp, w, dim, T, L;
    char db_prefix[256];
    char db_path[256];
} PQConfig;

// include/bucket.h:20-24
typedef struct {
    int L = 0;
    std::vector<int> pid;
    std::vector<unsigned char> codes;
} Bucket;

namespace detail {

[[noreturn]] inline void die(const char *what) {
    std::cerr << what << std::endl;
    std::exit(1);
}

inline void check(int rc, const char *what) {
    if (rc != SC_OK) {
        std::cerr << what << " failed (" << rc << "): " << sc_last_error() << std::endl;
        std::exit(1);
    }
}

// One context per process on device SC_DEVICE (default 0).
inline sc_ctx *ctx() {
    static sc_ctx *c = [] {
        sc_ctx *h = nullptr;
        const char *e = std::getenv("SC_DEVICE");
        check(sc_ctx_create(e ? std::atoi(e) : 0, &h), "sc_ctx_create");
        return h;
    }();
    return c;
}

// RAII device buffer.
struct Dev {
    void *p = nullptr;
    size_t bytes = 0;
    Dev() = default;
    explicit Dev(size_t b) { alloc(b); }
    Dev(const void *host, size_t b) { alloc(b); put(host, b); }
    ~Dev() { release(); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
    void alloc(size_t b) {
        release();
        bytes = b;
        check(sc_malloc(ctx(), b ? b : 1, &p), "sc_malloc");
    }
    void put(const void *host, size_t b) { if (b) check(sc_memcpy_h2d(ctx(), p, host, b), "h2d"); }
    void get(void *host, size_t b) const { if (b) check(sc_memcpy_d2h(ctx(), host, p, b), "d2h"); }
    void release() {
        if (p) sc_free(ctx(), p);
        p = nullptr;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

inline size_t file_size(const char *fn) {
    struct stat s;
    if (stat(fn, &s) != 0) return 0;
    return (size_t)s.st_size;
}

// Rows of `d` values of type T, each preceded by `header` bytes
// (include/sc_utilities.h:816-886), converted to float.
template <typename T>
inline int load_rows(const char *filename, float *&out, int header, int d, bool verbose) {
    if (d <= 0 || !filename) die("Defective data!");
    int fd = open(filename, O_RDONLY);
    if (fd < 0) die("Cannot open the file");
    size_t size = file_size(filename);
    const unsigned char *m =
        (const unsigned char *)mmap(nullptr, size ? size : 1, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) die("Cannot map the file");
    size_t row = (size_t)d * sizeof(T) + header;
    size_t n = row ? size / row : 0;
    const size_t cnt = n * (size_t)d;
    out = static_cast<float *>(::operator new(sizeof(float) * (cnt > 0 ? cnt : 1)));
    for (size_t i = 0; i < n; ++i) {
        const unsigned char *r = m + i * row + header;
        for (int j = 0; j < d; ++j) {
            T v;
            std::memcpy(&v, r + (size_t)j * sizeof(T), sizeof(T));
            out[i * d + j] = static_cast<float>(v);
        }
    }
    munmap((void *)m, size ? size : 1);
    close(fd);
    if (verbose) std::cout << "The number of vectors: " << n << std::endl;
    return (int)n;
}

}  // namespace detail

// include/sc_utilities.h:896-980: float header [k, m, dim] then k*dim floats.
// type 0 fills kc/mc (coarse), type 1 kp/mp (product).
template <typename DataType>
inline void load_codebook(const char *filename, PQConfig &config, DataType *&codebook, int type,
                          bool verbose) {
    int fd = open(filename, O_RDONLY);
    if (fd < 0) detail::die("Cannot open the file");
    size_t size = detail::file_size(filename);
    if (size < 3 * sizeof(float)) detail::die("Wrong data");
    const float *m = (const float *)mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) detail::die("Cannot map the file");
    config.dim = (int)m[2];
    int k = (int)m[0], mm = (int)m[1];
    if (type == 0) { config.kc = k; config.mc = mm; }
    else { config.kp = k; config.mp = mm; }
    if (mm == 0) detail::die("Wrong data");
    size_t cnt = (size_t)k * config.dim;
    if ((3 + cnt) * sizeof(float) > size) detail::die("Wrong data");
    codebook = static_cast<DataType *>(::operator new(sizeof(DataType) * cnt));
    for (size_t i = 0; i < cnt; ++i) codebook[i] = (DataType)m[3 + i];
    munmap((void *)m, size);
    close(fd);
    if (verbose)
        std::cout << "Loaded a codebook at " << filename << " with " << mm << " sub codes"
                  << std::endl;
}

// `.ctr_` writer (include/quantizer.h:370-431): header, centres, and the one
// stretch byte the reference's lseek/write leaves at the end.
inline void write_codebook(const char *fn, const float *centers, int nsc, int part, int dim) {
    FILE *f = std::fopen(fn, "wb");
    if (!f) detail::die("Cannot open the file");
    float hdr[3] = {(float)nsc, (float)part, (float)dim};
    std::fwrite(hdr, sizeof(float), 3, f);
    std::fwrite(centers, sizeof(float), (size_t)nsc * dim, f);
    std::fputc(0, f);
    std::fclose(f);
}

// Training knobs with no reference counterpart: the reference's greg_kmeans
// stops after one Lloyd update (k-means.h:363, 906) and seeds from
// random_device; the drop-in runs `iterations` Lloyd updates from a seeded
// k-means++.
struct KmeansOptions {
    int iterations = 25;
    uint32_t seed = 20151001u;
    int lagged = 0;
    // SC_PREC_BF16: Lloyd assignment on BF16 MFMA with an exact re-check
    // (identical labels and centres; sub-vector lengths multiple of 16)
    int prec = SC_PREC_FP32;
};
inline KmeansOptions &kmeans_options() {
    static KmeansOptions o;
    return o;
}

// ---------------------------------------------------------------------------
// include/quantizer.h:46-106
// ---------------------------------------------------------------------------
template <typename DataType>
class PQQuantizer {
protected:
    int dim, part, nsc, offset;
    int N;
    float *centers;
    int *labels;
    float **cq;
    int k;
    int *kc, *mc;

public:
    float *data;

    PQQuantizer(int _dim, int _m, int _nsc, int _os, bool verbose)
        : dim(_dim), part(_m), nsc(_nsc), offset(_os), N(0), centers(nullptr), labels(nullptr),
          cq(nullptr), k(0), kc(nullptr), mc(nullptr), data(nullptr) {
        if (dim <= 0 || part <= 0 || dim % part != 0) {
            if (verbose) std::cerr << "Error at parameters of PQQuantizer" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        centers = static_cast<float *>(::operator new(sizeof(float) * (size_t)nsc * dim));
    }
    PQQuantizer(int _dim, int _m, int _nsc, int _os, int _k, bool verbose)
        : PQQuantizer(_dim, _m, _nsc, _os, verbose) {
        k = _k;
        cq = static_cast<float **>(::operator new(sizeof(float *) * (k > 0 ? k : 1)));
        for (int i = 0; i < k; ++i) cq[i] = nullptr;
        kc = static_cast<int *>(::operator new(sizeof(int) * (k > 0 ? k : 1)));
        mc = static_cast<int *>(::operator new(sizeof(int) * (k > 0 ? k : 1)));
    }
    virtual ~PQQuantizer() {
        ::operator delete(centers);
        ::operator delete(labels);
        ::operator delete(data);
        for (int i = 0; i < k; ++i) ::operator delete(cq[i]);
        ::operator delete(cq);
        ::operator delete(kc);
        ::operator delete(mc);
    }

    // quantizer.h:194-205 (data converted from DataType rows with `offset`
    // header bytes)
    inline void load_data(const char *filename, bool verbose) {
        ::operator delete(data);
        N = detail::load_rows<DataType>(filename, data, offset, dim, verbose);
        ::operator delete(labels);
        labels = static_cast<int *>(::operator new(sizeof(int) * ((size_t)N * part + 1)));
        if (verbose) std::cout << "Initialized " << N << " vectors" << std::endl;
    }
    // quantizer.h:213-219: takes ownership of _data.
    inline void load_data_mat(float *_data, int _N) {
        if (_N > 0) N = _N;
        data = _data;
        if (!labels) labels = static_cast<int *>(::operator new(sizeof(int) * ((size_t)N * part + 1)));
    }
    inline int size() { return N; }

    // quantizer.h:235-272: one k-means per column block, on the GPU.
    inline void create_sub_quantizers(bool verbose) {
        const int bs = dim / part;
        const KmeansOptions &o = kmeans_options();
        detail::Dev dx((size_t)N * bs * sizeof(float)), dc((size_t)nsc * bs * sizeof(float)),
            dl((size_t)N * sizeof(int));
        std::vector<float> strip((size_t)N * bs);
        for (int i = 0; i < part; ++i) {
            for (size_t r = 0; r < (size_t)N; ++r)
                std::memcpy(&strip[r * bs], data + r * dim + (size_t)i * bs, sizeof(float) * bs);
            dx.put(strip.data(), strip.size() * sizeof(float));
            if (verbose) std::cout << "Creating codebook " << i << "/" << part << std::endl;
            detail::check(sc_kmeans(detail::ctx(), dx.as<float>(), N, bs, nsc, o.iterations,
                                    o.seed + (uint32_t)i, o.lagged, o.prec, dc.as<float>(),
                                    dl.as<int32_t>()),
                          "sc_kmeans");
            dc.get(centers + (size_t)i * nsc * bs, (size_t)nsc * bs * sizeof(float));
            dl.get(labels + (size_t)i * N, (size_t)N * sizeof(int));
        }
    }

    // quantizer.h:279-300 with SimpleCluster::distortion (k-means.h:444-464):
    // per part, a float running sum of the S1 distances (each `+=` adds in
    // double and rounds to float) and its float square root; then the double
    // square root of the sum of the parts' squares.
    inline double distortion(bool verbose) {
        (void)verbose;
        const int bs = dim / part;
        double e = 0.0;
        for (int i = 0; i < part; ++i) {
            float et = 0.0f;
            const float *c = centers + (size_t)i * nsc * bs;
            for (size_t r = 0; r < (size_t)N; ++r) {
                const float *x = data + r * dim + (size_t)i * bs;
                const float *y = c + (size_t)labels[(size_t)i * N + r] * bs;
                double acc = 0.0;
                for (int t = 0; t < bs; ++t) {
                    double dd = (double)(x[t] - y[t]);
                    acc += dd * dd;
                }
                et = (float)((double)et + acc);
            }
            const double es = std::sqrt(et);   // float square root, widened
            e += es * es;
        }
        return std::sqrt(e);
    }

    inline float *get_center_at(int id, bool verbose) {
        (void)verbose;
        return centers + (size_t)id * nsc * (dim / part);
    }
    inline int *get_label_at(int id, bool verbose) {
        (void)verbose;
        return labels + (size_t)id * N;
    }

    // quantizer.h:326-362
    inline void set_params(int _k, int _m, int _nsc) {
        for (int i = 0; i < k; ++i) ::operator delete(cq[i]);
        ::operator delete(cq);
        ::operator delete(kc);
        ::operator delete(mc);
        k = _k;
        cq = static_cast<float **>(::operator new(sizeof(float *) * (k > 0 ? k : 1)));
        for (int i = 0; i < k; ++i) cq[i] = nullptr;
        kc = static_cast<int *>(::operator new(sizeof(int) * (k > 0 ? k : 1)));
        mc = static_cast<int *>(::operator new(sizeof(int) * (k > 0 ? k : 1)));
        part = _m;
        nsc = _nsc;
        ::operator delete(centers);
        ::operator delete(labels);
        if (dim <= 0 || part <= 0 || dim % part != 0) {
            std::cerr << "Error at parameters of PQQuantizer" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        centers = static_cast<float *>(::operator new(sizeof(float) * (size_t)nsc * dim));
        labels = static_cast<int *>(::operator new(sizeof(int) * ((size_t)N * part + 1)));
    }

    // quantizer.h:370-431 -> <filename>.ctr_
    inline void output(const char *filename, bool verbose) {
        (void)verbose;
        char fn[512];
        std::snprintf(fn, sizeof fn, "%s.ctr_", filename);
        write_codebook(fn, centers, nsc, part, dim);
    }

    // quantizer.h:437-448
    inline void load_codebooks(char **filename, bool verbose) {
        PQConfig config;
        for (int i = 0; i < k; ++i) {
            ::operator delete(cq[i]);
            load_codebook<float>(filename[i], config, cq[i], 0, verbose);
            kc[i] = config.kc;
            mc[i] = config.mc;
        }
    }

    // quantizer.h:456-510 (nearest) / sc_quantizer.h:72-133 (second), on
    // the GPU, in place on `data`; `which` 0 nearest, 1 second nearest.
    // part_walk: the quantizer's `part` blocks of dim/part (SCQuantizer's
    // walk); otherwise each coarse codebook's own mc blocks.
    inline void residual_pass(int which, bool part_walk) {
        detail::Dev dx(data, (size_t)N * dim * sizeof(float));
        for (int k0 = 0; k0 < k; ++k0) {
            detail::Dev dc(cq[k0], (size_t)kc[k0] * dim * sizeof(float));
            detail::check(sc_residual(detail::ctx(), dx.as<float>(), N, dim, dc.as<float>(),
                                      part_walk ? part : mc[k0], kc[k0], which),
                          "sc_residual");
        }
        dx.get(data, (size_t)N * dim * sizeof(float));
    }
    inline void calc_residual_vector(bool verbose) {
        (void)verbose;
        residual_pass(0, false);
    }
};

// include/sc_quantizer.h:27-37: residuals against the second-nearest centre.
// By default this walks the reference's blocks (sc_quantizer.h:96-124): the
// quantizer's `part` blocks of dim/part, over the coarse codebook's floats
// read as [part][kc][dim/part] -- which is the codebook's own geometry only
// when its mc equals part.  set_codebook_blocks(true) walks the codebook's
// own mc blocks instead (the second-nearest centroid of each coarse block).
template <typename DataType>
class SCQuantizer : public PQQuantizer<DataType> {
    bool codebook_blocks_ = false;

public:
    SCQuantizer(int _dim, int _m, int _nsc, int _os, int _k, bool verbose)
        : PQQuantizer<DataType>(_dim, _m, _nsc, _os, _k, verbose) {}
    virtual ~SCQuantizer() {}
    void set_codebook_blocks(bool on) { codebook_blocks_ = on; }
    inline void calc_residual_vector(bool verbose) {
        (void)verbose;
        this->residual_pass(1, !codebook_blocks_);
    }
};

// ---------------------------------------------------------------------------
// include/encoder.h:30-60, src/encoder.cpp
// ---------------------------------------------------------------------------
class Encoder {
protected:
    PQConfig config;
    float *cq, *pq;
    int size = 0;
    int non_empty_bucket = 0;
    int *L, *pid;
    ushort *cid;
    unsigned char *codes;
    std::vector<Bucket> ivf;
    int old_mp;
    int nc_ = 1;   // coarse candidates per block the last encode wrote

    // general: SCEncoder::encode's float-rounded top-nc (sc_encode_general);
    // otherwise Encoder::encode (nc = 1) / SCEncoder::encode2 (nc = 2)
    inline void encode_device(const float *x, int nc, bool general = false) {
        const size_t n = (size_t)config.N;
        detail::Dev dx(x, n * config.dim * sizeof(float));
        detail::Dev dcq(cq, (size_t)config.kc * config.dim * sizeof(float));
        detail::Dev dpq(pq, (size_t)config.kp * config.dim * sizeof(float));
        detail::Dev dcid(n * config.mc * nc * sizeof(ushort)), dcodes(n * config.mp);
        if (general)
            detail::check(sc_encode_general(detail::ctx(), dx.as<float>(), (int64_t)n, config.dim,
                                            dcq.as<float>(), config.mc, config.kc, dpq.as<float>(),
                                            config.mp, config.kp, nc, dcid.as<uint16_t>(),
                                            dcodes.as<uint8_t>()),
                          "sc_encode_general");
        else
            detail::check(sc_encode(detail::ctx(), dx.as<float>(), (int64_t)n, config.dim,
                                    dcq.as<float>(), config.mc, config.kc, dpq.as<float>(),
                                    config.mp, config.kp, nc, SC_PREC_FP32, dcid.as<uint16_t>(),
                                    dcodes.as<uint8_t>(), nullptr),
                          "sc_encode");
        ::operator delete(cid);
        ::operator delete(codes);
        cid = static_cast<ushort *>(::operator new(n * config.mc * nc * sizeof(ushort) + 2));
        codes = static_cast<unsigned char *>(::operator new(n * config.mp + 1));
        dcid.get(cid, n * config.mc * nc * sizeof(ushort));
        dcodes.get(codes, n * config.mp);
    }

    // Buckets from the device index build (ids ascending per cell).
    inline void distribute(int nc, int64_t ncells) {
        const size_t n = (size_t)config.N;
        detail::Dev dcid(cid, n * config.mc * nc * sizeof(ushort)), dcodes(codes, n * config.mp);
        detail::Dev dcq(cq, (size_t)config.kc * config.dim * sizeof(float));
        detail::Dev dpq(pq, (size_t)config.kp * config.dim * sizeof(float));
        sc_index *ix = nullptr;
        int rc = sc_index_build(detail::ctx(), dcid.as<uint16_t>(), dcodes.as<uint8_t>(),
                                (int64_t)n, config.dim, dcq.as<float>(), config.mc, config.kc,
                                dpq.as<float>(), config.mp, config.kp, nc, &ix);
        if (rc == SC_ERR_TOO_LARGE) {
            std::cerr << "The size of this inverted index is too LARGE!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        detail::check(rc, "sc_index_build");
        std::vector<int32_t> hL(ncells), hpid(n);
        std::vector<uint8_t> hcodes(n * config.mp);
        detail::check(sc_index_export(detail::ctx(), ix, hL.data(), hpid.data(), hcodes.data(),
                                      nullptr, nullptr, nullptr),
                      "sc_index_export");
        sc_index_destroy(ix);
        ivf.assign((size_t)ncells, Bucket());
        non_empty_bucket = 0;
        size_t start = 0;
        for (int64_t c = 0; c < ncells; ++c) {
            Bucket &b = ivf[(size_t)c];
            b.L = hL[(size_t)c] - (int)start;
            b.pid.assign(hpid.begin() + start, hpid.begin() + hL[(size_t)c]);
            b.codes.assign(hcodes.begin() + start * config.mp,
                           hcodes.begin() + (size_t)hL[(size_t)c] * config.mp);
            if (b.L > 0) ++non_empty_bucket;
            start = (size_t)hL[(size_t)c];
        }
    }

    inline void write_ivf(const char *fname) {
        FILE *f = std::fopen(fname, "wb");
        if (!f) detail::die("Cannot open the file");
        int N = config.N;
        std::fwrite(&non_empty_bucket, sizeof(int), 1, f);
        std::fwrite(&N, sizeof(int), 1, f);
        for (const Bucket &b : ivf) {
            std::fwrite(&b.L, sizeof(int), 1, f);
            if (b.L > 0) {
                std::fwrite(b.pid.data(), sizeof(int), b.pid.size(), f);
                std::fwrite(b.codes.data(), 1, b.codes.size(), f);
            }
        }
        std::fputc(0, f);   // the reference's lseek stretch byte
        std::fclose(f);
    }

public:
    Encoder() : cq(nullptr), pq(nullptr), L(nullptr), pid(nullptr), cid(nullptr),
                codes(nullptr), old_mp(0) {
        std::memset(&config, 0, sizeof config);
    }
    virtual ~Encoder() {
        ::operator delete(cq);
        ::operator delete(pq);
        ::operator delete(cid);
        ::operator delete(codes);
        ::operator delete(L);
        ::operator delete(pid);
    }

    // src/encoder.cpp:44-55
    void load_codebooks(const char *cq_path, const char *pq_path, bool verbose) {
        load_codebook<float>(cq_path, config, cq, 0, verbose);
        load_codebook<float>(pq_path, config, pq, 1, verbose);
        size = (int)std::pow((double)config.kc, (double)config.mc);
        std::cout << "--> Settings: (kc,mc,kp,mp)=" << config.kc << " " << config.mc << " "
                  << config.kp << " " << config.mp << std::endl;
    }

    // include/encoder.h:69-192
    template <typename DataType>
    inline void encode(const char *filename, int offset, bool verbose) {
        float *x = nullptr;
        config.N = detail::load_rows<DataType>(filename, x, offset, config.dim, verbose);
        if (config.mc <= 0 || config.mp <= 0) {
            std::cerr << "Nothing to do" << std::endl;
            ::operator delete(x);
            return;
        }
        nc_ = 1;
        encode_device(x, 1);
        ::operator delete(x);
    }

    // src/encoder.cpp:170-205
    void distribution(bool verbose) {
        distribute(1, size);
        if (verbose)
            std::cout << "The number of non-empty buckets: " << non_empty_bucket << std::endl;
    }

    // src/encoder.cpp:207-305 -> <path>/<prefix>_ivf.edat_
    void output(const char *db_path, const char *db_prefix, bool verbose) {
        (void)verbose;
        char fname[600];
        std::snprintf(fname, sizeof fname, "%s/%s_ivf.edat_", db_path, db_prefix);
        write_ivf(fname);
    }

    // src/encoder.cpp:57-146: reads an inverted file written by output() or
    // output2(): per cell (all `size` of them, in cell order) its length, its
    // ids and old_mp code bytes per vector (skipped).  Fills L with the cell
    // lengths (not a prefix), pid with the ids in cell order and cid [N][mc]
    // with each vector's cell digits (the cell index in base kc, most
    // significant first).  A truncated or inconsistent file is an error here
    // (the reference reads past the mapping).
    void load_encoded_data(const char *filename, bool verbose) {
        std::FILE *f = std::fopen(filename, "rb");
        if (!f) {
            if (verbose) std::cerr << "Cannot open the file" << std::endl;
            std::exit(1);
        }
        std::vector<unsigned char> buf;
        {
            unsigned char tmp[1 << 16];
            size_t got;
            while ((got = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
            std::fclose(f);
        }
        size_t at = 0;
        auto take = [&](void *dst, size_t bytes) {
            if (bytes > buf.size() - at) detail::die("load_encoded_data: truncated inverted file");
            std::memcpy(dst, buf.data() + at, bytes);
            at += bytes;
        };
        if (size <= 0 || config.mc <= 0 || config.kc <= 0)
            detail::die("load_encoded_data: load the codebooks first");
        if (old_mp < 0) detail::die("load_encoded_data: negative code length (set_mp)");
        take(&non_empty_bucket, sizeof(int));
        std::cout << "The number of non empty buckets: " << non_empty_bucket << std::endl;
        take(&config.N, sizeof(int));
        if (config.N < 0) detail::die("load_encoded_data: negative vector count");
        ::operator delete(cid);
        ::operator delete(L);
        ::operator delete(pid);
        cid = static_cast<ushort *>(::operator new((size_t)config.N * config.mc * sizeof(ushort) + 2));
        L = static_cast<int *>(::operator new((size_t)size * sizeof(int)));
        pid = static_cast<int *>(::operator new((size_t)config.N * sizeof(int) + 4));
        std::cout << "Old size: " << old_mp << "[byte(s)]" << std::endl;
        std::vector<ushort> c((size_t)config.mc);
        size_t base = 0;
        for (int i = 0; i < size; ++i) {
            int l = 0;
            take(&l, sizeof(int));
            if (l < 0 || (size_t)l > (size_t)config.N - base)
                detail::die("load_encoded_data: cell lengths exceed the vector count");
            L[i] = l;
            if (verbose) std::cout << "This bucket contains " << l << " vector(s)" << std::endl;
            int64_t p = i;
            for (int j = config.mc - 1; j >= 0; --j) {
                c[(size_t)j] = (ushort)(p % config.kc);
                p /= config.kc;
            }
            take(pid + base, (size_t)l * sizeof(int));
            for (int v = 0; v < l; ++v)
                for (int j = 0; j < config.mc; ++j) cid[(base + v) * config.mc + j] = c[(size_t)j];
            base += (size_t)l;
            if ((size_t)l * old_mp > buf.size() - at)
                detail::die("load_encoded_data: truncated inverted file");
            at += (size_t)l * old_mp;
            if (verbose) std::cout << "Read " << l << " vector(s)" << std::endl;
        }
        if (base != (size_t)config.N)
            detail::die("load_encoded_data: cell lengths do not add up to the vector count");
        std::cout << "Read " << config.N << " data  from " << filename << std::endl;
    }

    // include/encoder.h:202-309: PQ codes of the raw vectors of `filename`
    // in the loaded file's cell order -- position i is raw row pid[i],
    // encoded against its loaded cell cid[i] (sc_encode_given).  The raw
    // file must hold exactly the loaded file's N vectors.
    template <typename DataType>
    inline void rencode(const char *filename, int offset, bool verbose) {
        const int nloaded = config.N;
        float *x = nullptr;
        config.N = detail::load_rows<DataType>(filename, x, offset, config.dim, verbose);
        if (config.mc <= 0 || config.mp <= 0) {
            std::cerr << "Nothing to do" << std::endl;
            ::operator delete(x);
            return;
        }
        if (!pid || !cid || config.N != nloaded) {
            ::operator delete(x);
            detail::die("rencode: load_encoded_data first, from a file of the same vectors");
        }
        const size_t n = (size_t)config.N;
        detail::Dev dx(x, n * config.dim * sizeof(float));
        ::operator delete(x);
        detail::Dev drows(pid, n * sizeof(int32_t)), dcid(cid, n * config.mc * sizeof(ushort));
        detail::Dev dcq(cq, (size_t)config.kc * config.dim * sizeof(float));
        detail::Dev dpq(pq, (size_t)config.kp * config.dim * sizeof(float));
        detail::Dev dcodes(n * config.mp);
        detail::check(sc_encode_given(detail::ctx(), dx.as<float>(), (int64_t)n, config.dim,
                                      drows.as<int32_t>(), (int64_t)n, dcid.as<uint16_t>(),
                                      dcq.as<float>(), config.mc, config.kc, dpq.as<float>(),
                                      config.mp, config.kp, dcodes.as<uint8_t>()),
                      "sc_encode_given");
        ::operator delete(codes);
        codes = static_cast<unsigned char *>(::operator new(n * config.mp + 1));
        dcodes.get(codes, n * config.mp);
    }

    // src/encoder.cpp:307-394 -> <path>/<prefix>_ivf.edat_ from the flat
    // arrays load_encoded_data + rencode fill: per cell L[i], its ids and
    // codes, then the lseek stretch byte (N*mp + (N+size+2)*4 + 1 bytes).
    void output2(const char *db_path, const char *db_prefix, bool verbose) {
        if (!L || !pid || !codes) detail::die("output2: nothing loaded");
        char fname[600];
        std::snprintf(fname, sizeof fname, "%s/%s_ivf.edat_", db_path, db_prefix);
        FILE *f = std::fopen(fname, "wb");
        if (!f) {
            if (verbose) std::cerr << "Cannot open the file " << fname << std::endl;
            std::exit(1);
        }
        int N = config.N;
        std::fwrite(&non_empty_bucket, sizeof(int), 1, f);
        std::fwrite(&N, sizeof(int), 1, f);
        size_t base = 0;
        for (int i = 0; i < size; ++i) {
            const int l = L[i];
            std::fwrite(&l, sizeof(int), 1, f);
            std::fwrite(pid + base, sizeof(int), (size_t)l, f);
            std::fwrite(codes + base * config.mp, 1, (size_t)l * config.mp, f);
            base += (size_t)l;
            if (verbose) std::cout << "Exported " << l << " vector(s)" << std::endl;
        }
        std::fputc(0, f);
        const size_t f_size = (size_t)N * config.mp + (size_t)(N + size + 2) * sizeof(int);
        std::fclose(f);
        std::cout << "Read " << f_size << " byte(s) and wrote out " << f_size << " byte(s)"
                  << std::endl;
    }

    PQConfig get_config() { return config; }
    // src/encoder.cpp:396-398: the code length of the file load_encoded_data
    // reads (config.mp, the length rencode writes, stays the codebook's).
    void set_mp(int old) { old_mp = old; }
    void set_size(size_t s) { size = (int)s; }
    // src/encoder.cpp:152-168
    void statistic(bool detail_ = false) {
        (void)detail_;
        FILE *f = std::fopen("./ivf.txt", "w");
        if (!f) return;
        for (size_t i = 0; i < ivf.size(); ++i) {
            std::fprintf(f, "Bucket %zu:%d\n", i, ivf[i].L);
            for (int j = 0; j < ivf[i].L; ++j) {
                std::fprintf(f, "%d:", ivf[i].pid[j]);
                for (int t = 0; t < config.mp; ++t)
                    std::fprintf(f, "%d ", (int)ivf[i].codes[(size_t)j * config.mp + t]);
                std::fprintf(f, "\n");
            }
        }
        std::fclose(f);
    }
    // Read-only views for tests and tools.
    const ushort *cell_ids() const { return cid; }
    const unsigned char *pq_codes() const { return codes; }
    const std::vector<Bucket> &buckets() const { return ivf; }
    const int *cell_lengths() const { return L; }
    const int *ids() const { return pid; }
};

// include/sc_encoder.h:32-54, src/sc_encoder.cpp
class SCEncoder : public Encoder {
protected:
    int nc;

public:
    SCEncoder() : Encoder(), nc(2) {}
    explicit SCEncoder(int _nc) : Encoder(), nc(_nc) { std::cout << "nc = " << nc << std::endl; }
    virtual ~SCEncoder() {}

    // include/sc_encoder.h:63-195: nearest + second per block
    template <typename DataType>
    inline void encode2(const char *filename, int offset, bool verbose) {
        float *x = nullptr;
        config.N = detail::load_rows<DataType>(filename, x, offset, config.dim, verbose);
        if (config.mc <= 0 || config.mp <= 0) {
            std::cerr << "Nothing to do" << std::endl;
            ::operator delete(x);
            return;
        }
        nc_ = 2;
        encode_device(x, 2);
        ::operator delete(x);
    }
    // include/sc_encoder.h:205-332: the nc nearest per block by float-rounded
    // distance (nc = 1..3), cid [N][mc][nc]
    template <typename DataType>
    inline void encode(const char *filename, int offset, bool verbose) {
        if (nc < 1 || nc > 3) {
            std::cerr << "SCEncoder::encode: nc must be 1, 2 or 3" << std::endl;
            return;
        }
        float *x = nullptr;
        config.N = detail::load_rows<DataType>(filename, x, offset, config.dim, verbose);
        if (config.mc <= 0 || config.mp <= 0) {
            std::cerr << "Nothing to do" << std::endl;
            ::operator delete(x);
            return;
        }
        nc_ = nc;
        encode_device(x, nc, true);
        ::operator delete(x);
    }

    // src/sc_encoder.cpp:17-81, with the per-vector cid stride mc*nc (the
    // reference strides 2*mc, which only holds for nc = 2); nc = 3 needs
    // mc = 1 and kc <= 1024
    void distribution(bool verbose) {
        double cells = std::pow((double)config.kc, (double)config.mc * nc_);
        if (cells >= (double)INT32_MAX) {
            std::cerr << "The size of this inverted index is too LARGE!" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        distribute(nc_, (int64_t)cells);
        if (verbose)
            std::cout << "The number of non-empty buckets: " << non_empty_bucket << std::endl;
    }

    // src/sc_encoder.cpp:83-182 -> <path>/<prefix>_mr<nc>_ivf.edat_
    void output(const char *db_path, const char *db_prefix, bool verbose) {
        (void)verbose;
        char fname[600];
        std::snprintf(fname, sizeof fname, "%s/%s_mr%d_ivf.edat_", db_path, db_prefix, nc);
        write_ivf(fname);
    }
};

// ---------------------------------------------------------------------------
// include/query.h:45-80, src/query.cpp
// ---------------------------------------------------------------------------
class PQQuery {
protected:
    int size = 0;
    int not_empty = 0;
    float *cq, *pq;
    int *L;
    int *pid;
    unsigned char *codes;
    PQConfig config;
    float *raw_data;
    int nc_ = 1;
    sc_index *index_ = nullptr;
    detail::Dev q_, ids_, dists_, sums_;
    detail::Dev raw_dev_;         // raw vectors on the device (load_data)
    int64_t raw_n_ = 0;

    // Gives the index the raw vectors when both are loaded and match (a
    // real_dist search reports the mismatch otherwise).
    inline void attach_raw() {
        if (index_ && raw_dev_.p && sc_index_attach_raw(index_, raw_dev_.as<float>(), raw_n_) != SC_OK)
            std::cerr << sc_last_error() << std::endl;
    }
    sc_search_stats last_{};

    // Loads `<file>` (int non_empty, int N, per cell: L, pids, codes) into
    // the host arrays and the device index (src/query.cpp:82-165).
    inline int load_ivf(const char *filename, size_t cells, int nc) {
        int fd = open(filename, O_RDONLY);
        if (fd < 0) detail::die("Cannot open the file");
        size_t fsz = detail::file_size(filename);
        const unsigned char *m =
            (const unsigned char *)mmap(nullptr, fsz, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) detail::die("Cannot map the file");
        // every read is bounded by the file size and every cell by N, so a
        // truncated or mismatched file ends with a message, not an overrun
        const unsigned char *t = m, *end = m + fsz;
        auto bad = [&](const char *why) {
            munmap((void *)m, fsz);
            close(fd);
            detail::die(why);
        };
        if (fsz < 8) bad("Encoded data file too short for its header");
        std::memcpy(&not_empty, t, 4); t += 4;
        std::memcpy(&config.N, t, 4); t += 4;
        if (config.N < 0) bad("Encoded data file: negative N");
        ::operator delete(L);
        ::operator delete(pid);
        ::operator delete(codes);
        L = static_cast<int *>(::operator new(sizeof(int) * cells));
        pid = static_cast<int *>(::operator new(sizeof(int) * ((size_t)config.N + 1)));
        codes = static_cast<unsigned char *>(::operator new((size_t)config.N * config.mp + 1));
        size_t base = 0;
        for (size_t i = 0; i < cells; ++i) {
            int l;
            if ((size_t)(end - t) < 4) bad("Encoded data file truncated (cell length)");
            std::memcpy(&l, t, 4); t += 4;
            if (l < 0 || base + (size_t)l > (size_t)config.N)
                bad("Encoded data file: cell lengths exceed N");
            L[i] = l + (i > 0 ? L[i - 1] : 0);
            if (l > 0) {
                const size_t need = (size_t)l * (4 + (size_t)config.mp);
                if ((size_t)(end - t) < need) bad("Encoded data file truncated (cell body)");
                std::memcpy(pid + base, t, (size_t)l * 4); t += (size_t)l * 4;
                std::memcpy(codes + base * config.mp, t, (size_t)l * config.mp);
                t += (size_t)l * config.mp;
                base += (size_t)l;
            }
        }
        if (base != (size_t)config.N) bad("Encoded data file: cell lengths do not sum to N");
        munmap((void *)m, fsz);
        close(fd);
        if (index_) sc_index_destroy(index_);
        index_ = nullptr;
        nc_ = nc;
        detail::check(sc_index_load(detail::ctx(), L, (int64_t)cells, pid, codes, config.N,
                                    config.dim, cq, config.mc, config.kc, pq, config.mp,
                                    config.kp, nc, &index_),
                      "sc_index_load");
        attach_raw();
        return config.N;
    }

    // One query through the batched device search; fills the reference's
    // callee-allocated result/dist arrays (length sum, first min(R, sum)
    // entries are the answer).
    inline bool run_one(const float *query, int kind, int R, int w, int T, int M, float *&dist,
                        int *&result, int &sum, bool real_dist = false) {
        if (!q_.p) q_.alloc(sizeof(float) * config.dim);
        if (ids_.bytes < sizeof(int) * (size_t)R) {
            ids_.alloc(sizeof(int) * (size_t)R);
            dists_.alloc(sizeof(float) * (size_t)R);
        }
        if (!sums_.p) sums_.alloc(sizeof(int));
        q_.put(query, sizeof(float) * config.dim);
        sc_search_params p{kind, real_dist ? SC_FORM_REAL : SC_FORM_PLAIN, R, w, T, M};
        int rc = sc_search(detail::ctx(), index_, q_.as<float>(), 1, &p, ids_.as<int32_t>(),
                           dists_.as<float>(), sums_.as<int32_t>(), &last_);
        if (rc == SC_ERR_CONFIG) {
            std::cerr << sc_last_error() << std::endl;
            return false;
        }
        detail::check(rc, "sc_search");
        sums_.get(&sum, sizeof(int));
        const size_t len = (size_t)(sum > R ? sum : R);
        result = static_cast<int *>(::operator new(sizeof(int) * len));
        dist = static_cast<float *>(::operator new(sizeof(float) * len));
        ids_.get(result, sizeof(int) * (size_t)R);
        dists_.get(dist, sizeof(float) * (size_t)R);
        return true;
    }

public:
    PQQuery() : cq(nullptr), pq(nullptr), L(nullptr), pid(nullptr), codes(nullptr),
                raw_data(nullptr) {
        std::memset(&config, 0, sizeof config);
    }
    virtual ~PQQuery() {
        if (index_) sc_index_destroy(index_);
        ::operator delete(cq);
        ::operator delete(pq);
        ::operator delete(L);
        ::operator delete(pid);
        ::operator delete(codes);
        ::operator delete(raw_data);
    }

    // src/query.cpp:66-75
    void load_codebooks(const char *cq_path, const char *pq_path, bool verbose) {
        load_codebook<float>(cq_path, config, cq, 0, verbose);
        load_codebook<float>(pq_path, config, pq, 1, verbose);
        size = (int)std::pow((double)config.kc, (double)config.mc);
        std::cout << "--> Settings: (kc,mc,kp,mp)=" << config.kc << " " << config.mc << " "
                  << config.kp << " " << config.mp << std::endl;
    }
    int load_encoded_data(const char *filename, bool verbose) {
        (void)verbose;
        return load_ivf(filename, (size_t)size, 1);
    }
    // query.h:83-86 (raw vectors for real distances), also kept on the device
    template <typename DataType>
    inline void load_data(const char *filename, int offset, bool verbose) {
        ::operator delete(raw_data);
        config.N = detail::load_rows<DataType>(filename, raw_data, offset, config.dim, verbose);
        raw_n_ = config.N;
        raw_dev_.alloc(sizeof(float) * (size_t)raw_n_ * config.dim);
        raw_dev_.put(raw_data, sizeof(float) * (size_t)raw_n_ * config.dim);
        attach_raw();
    }
    // The tables and LUTs live on the device and are built inside the search
    // (query.h:91-194); these keep the call sequence of the reference.
    inline void pre_compute1() {}
    inline void pre_compute2(float *) {}
    inline void pre_compute3(float *) {}

    // query.h:202-325
    inline void search_ivfadc(float *query, float *&v_tmp, float *&dist, int *&result,
                              int *&buckets, int *&prebuck, int &sum, int R, int w, int T,
                              bool real_dist, bool verbose) {
        (void)v_tmp; (void)buckets; (void)prebuck; (void)verbose;
        if (config.mc != 1) {
            std::cerr << "This search method is for IVFADC only" << std::endl;
            return;
        }
        run_one(query, SC_KIND_IVFADC, R, w, T, config.kc, dist, result, sum, real_dist);
    }

    // Batched search (no reference counterpart): nq queries in one call,
    // results [nq][R] on the host.
    inline void search_batch(const float *queries, int nq, int kind, int R, int w, int T,
                             int *result, float *dist, int *sums) {
        detail::Dev dq(queries, sizeof(float) * (size_t)nq * config.dim);
        detail::Dev di(sizeof(int) * (size_t)nq * R), dd(sizeof(float) * (size_t)nq * R),
            ds(sizeof(int) * (size_t)nq);
        sc_search_params p{kind, SC_FORM_PLAIN, R, w, T, config.kc};
        detail::check(sc_search(detail::ctx(), index_, dq.as<float>(), nq, &p, di.as<int32_t>(),
                                dd.as<float>(), ds.as<int32_t>(), &last_),
                      "sc_search");
        di.get(result, sizeof(int) * (size_t)nq * R);
        dd.get(dist, sizeof(float) * (size_t)nq * R);
        if (sums) ds.get(sums, sizeof(int) * (size_t)nq);
    }

    int get_size() { return not_empty; }
    int get_full_size() { return size; }
    // src/query.cpp:168-182
    double entropy(size_t sz) {
        double e = 0.0, Nd = config.N;
        for (size_t i = 0; i < sz; ++i) {
            double l = i > 0 ? L[i] - L[i - 1] : L[0];
            if (l > 0) {
                double x = Nd / l;
                e += std::log2(x) / x;
            }
        }
        return e;
    }
    const sc_search_stats &last_stats() const { return last_; }
};

// include/multi_query.h:36-53
class MultiQuery : public PQQuery {
public:
    MultiQuery() : PQQuery() {}
    virtual ~MultiQuery() {}

    // multi_query.h:62-273; t1/t2 accumulate clock ticks of the LUT and
    // coarse-sort stages like the reference (stage times from HIP events).
    inline void search_multi2(float *query, float *&v_tmp, float *&dist, float *&q, int *tmp,
                              int *&result, int *&hid1, int *&hid2, int *&hid3, int *&hid4,
                              int *&s1, int *&s2, int *&prebuck, int *&cache, bool *traversed,
                              int &sum, int R, int w, int T, int M, int &e, double &t1,
                              double &t2, bool real_dist, bool verbose) {
        (void)v_tmp; (void)q; (void)tmp; (void)hid1; (void)hid2; (void)hid3; (void)hid4;
        (void)s1; (void)s2; (void)prebuck; (void)cache; (void)traversed;
        (void)verbose;
        if (config.mc != 2) {
            std::cerr << "This search method is for Multi-D-ADC-2 only" << std::endl;
            return;
        }
        if (!run_one(query, SC_KIND_IMI, R, w, T, M, dist, result, sum, real_dist)) return;
        e = (int)(last_.visited - last_.cells);
        t1 += last_.ms_lut * CLOCKS_PER_SEC / 1000.0;
        t2 += last_.ms_rank * CLOCKS_PER_SEC / 1000.0;
    }
};

// include/sc_query.h:21-51, src/sc_query.cpp
class SCQuery : public PQQuery {
protected:
    int nc;

public:
    SCQuery() : PQQuery(), nc(2) {}
    explicit SCQuery(int _nc) : PQQuery(), nc(_nc) {}
    virtual ~SCQuery() {}

    // src/sc_query.cpp:36-123: kc^(mc*nc) cells
    int load_encoded_data(const char *filename, bool verbose) {
        (void)verbose;
        size_t cells = (size_t)std::pow((double)size, (double)nc);
        return load_ivf(filename, cells, nc);
    }

    // sc_query.h:60-256 (the top-R comes back sorted by (dist, position))
    inline void search_mr_ivf(float *query, float *&v_tmp, float *&dist, int *tmp, int *&result,
                              int *&hid1, int *&hid2, int *&hid3, int *&hid4, int *&s1, int *&s2,
                              int *&prebuck, int *&cache, bool *traversed, int &sum, int R,
                              int w, int T, int M, bool real_dist, bool verbose) {
        (void)v_tmp; (void)tmp; (void)hid1; (void)hid2; (void)hid3; (void)hid4; (void)s1;
        (void)s2; (void)prebuck; (void)cache; (void)traversed; (void)verbose;
        if (config.mc != 1 || nc != 2) {
            std::cerr << "This search method is for MultiRank IVFADC only" << std::endl;
            return;
        }
        run_one(query, SC_KIND_SC, R, w, T, M, dist, result, sum, real_dist);
    }

    // sc_query.h:265-475: (nearest, second, third) cells (the top-R comes back
    // sorted by (dist, position))
    inline void search_mr_ivf3(float *query, float *&v_tmp, float *&dist, int *tmp, int *&result,
                               int *&hid1, int *&hid2, int *&hid3, int *&hid4, int *&hid5,
                               int *&hid6, int *&s1, int *&prebuck, int *&cache, bool *traversed,
                               int &sum, int R, int w, int T, int M, bool real_dist,
                               bool verbose) {
        (void)v_tmp; (void)tmp; (void)hid1; (void)hid2; (void)hid3; (void)hid4; (void)hid5;
        (void)hid6; (void)s1; (void)prebuck; (void)cache; (void)traversed; (void)verbose;
        if (config.mc != 1 || nc != 3) {
            std::cerr << "This search method is for MultiRank IVFADC-3 only" << std::endl;
            return;
        }
        run_one(query, SC_KIND_SC3, R, w, T, M, dist, result, sum, real_dist);
    }
};

// include/evaluation.h:20-140
class Evaluation {
public:
    // evaluation.h:34-47: the first of the d ints of each row (4-byte row
    // header) of a ground-truth file; gt is callee-allocated (::operator new)
    inline void load_groundtruth(const char *filename, int *&gt, int d, int &N, bool verbose) {
        if (d <= 0 || !filename) detail::die("Defective data!");
        int fd = open(filename, O_RDONLY);
        if (fd < 0) detail::die("Cannot open the file");
        const size_t size = detail::file_size(filename);
        const size_t row = (size_t)d * 4 + 4;
        N = (int)(size / row);
        gt = static_cast<int *>(::operator new(sizeof(int) * ((size_t)N + 1)));
        for (int i = 0; i < N; ++i) {
            if (pread(fd, &gt[i], 4, (off_t)((size_t)i * row + 4)) != 4) detail::die("Cannot read the file");
        }
        close(fd);
        if (verbose) std::cout << "The number of vectors: " << N << std::endl;
    }
    // evaluation.h:58-92: queries whose ground truth is among their R results
    inline void calc_recall(int *result, int *groundtruth, int R, int N, int &recall,
                            bool verbose) {
        recall = 0;
        for (int i = 0; i < N; ++i) {
            const int *r = result + (size_t)i * R;
            bool found = false;
            for (int j = 0; j < R && !found; ++j) found = r[j] == groundtruth[i];
            if (found) recall++;
        }
        if (verbose) std::cout << "Recall@" << R << " = " << recall << std::endl;
    }
    // evaluation.h:103-140: found in any of the S result sets
    inline void calc_comb_recall(int **results, int *groundtruth, int R, int N, int S,
                                 int &recall, bool verbose) {
        recall = 0;
        for (int i = 0; i < N; ++i) {
            bool found = false;
            for (int s = 0; s < S && !found; ++s)
                for (int j = 0; j < R && !found; ++j)
                    found = results[s][(size_t)i * R + j] == groundtruth[i];
            if (found) recall++;
        }
        if (verbose) std::cout << "Recall@" << R << " = " << recall << std::endl;
    }
    // No reference counterpart (the reference reads ground truth from files):
    // the nearest base vector of each query by the exact distance, lowest id
    // on ties, computed on the GPU (sc_groundtruth_nn1).  gt is
    // callee-allocated (::operator new).
    inline void compute_groundtruth(const float *base, int n, const float *queries, int nq,
                                    int dim, int *&gt) {
        detail::Dev b(base, sizeof(float) * (size_t)n * dim);
        detail::Dev q(queries, sizeof(float) * (size_t)nq * dim);
        detail::Dev ids(sizeof(int32_t) * (size_t)(nq > 0 ? nq : 1));
        detail::check(sc_groundtruth_nn1(detail::ctx(), b.as<float>(), n, q.as<float>(), nq, dim,
                                         ids.as<int32_t>(), nullptr),
                      "sc_groundtruth_nn1");
        gt = static_cast<int *>(::operator new(sizeof(int) * ((size_t)nq + 1)));
        ids.get(gt, sizeof(int32_t) * (size_t)nq);
    }
};

}  // namespace SC
