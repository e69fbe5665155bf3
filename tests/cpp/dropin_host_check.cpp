// dropin_host_check.cpp -- host-only parts of the drop-in header
// (second-candidate_amd/host/sc_dropin.h) against the reference's own code
// compiled in place (oracle/_ref/libscref.so; test infrastructure).  No GPU.
//
// PQQuantizer::distortion (include/quantizer.h:279-300) is compared bit for
// bit with the reference's SimpleCluster::distortion
// (lib/simple-cluster/include/k-means.h:444-464) applied part by part.
// usage: dropin_host_check            exit 0 = all checks passed
//        dropin_host_check edat CQ PQ FILE
//            PQQuery::load_codebooks + load_encoded_data on FILE: a malformed
//            inverted file must end the process with a message before any
//            device work (the caller checks the exit status and message)
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../second-candidate_amd/host/sc_dropin.h"

extern "C" float scref_distortion(const float *data, const float *centers, const int *labels,
                                  int d, int N, int k);

namespace {
class Probe : public SC::PQQuantizer<float> {
public:
    Probe(int dim, int part, int nsc) : SC::PQQuantizer<float>(dim, part, nsc, 0, false) {}
    void set(const std::vector<float> &x, int n, const std::vector<float> &c,
             const std::vector<int> &lab) {
        N = n;
        data = static_cast<float *>(::operator new(sizeof(float) * x.size()));
        std::memcpy(data, x.data(), sizeof(float) * x.size());
        std::memcpy(centers, c.data(), sizeof(float) * c.size());
        labels = static_cast<int *>(::operator new(sizeof(int) * lab.size()));
        std::memcpy(labels, lab.data(), sizeof(int) * lab.size());
    }
};

uint32_t lcg(uint32_t &s) { return s = s * 1664525u + 1013904223u; }
}  // namespace

int main(int argc, char **argv) {
    if (argc == 5 && std::strcmp(argv[1], "edat") == 0) {
        SC::PQQuery q;
        q.load_codebooks(argv[2], argv[3], false);
        q.load_encoded_data(argv[4], false);
        std::printf("loaded\n");
        return 0;
    }
    int fails = 0;
    const int cases[][4] = {{32, 2, 16, 5000}, {128, 8, 64, 3000}, {48, 1, 7, 20000}};
    for (const auto &cs : cases) {
        const int dim = cs[0], part = cs[1], nsc = cs[2], n = cs[3], bs = dim / part;
        uint32_t s = (uint32_t)(dim * 131 + n);
        std::vector<float> x((size_t)n * dim), c((size_t)nsc * dim);
        std::vector<int> lab((size_t)part * n);
        for (float &v : x) v = (float)(lcg(s) % 25600) / 100.0f + (float)(lcg(s) % 7) / 3.0f;
        for (float &v : c) v = (float)(lcg(s) % 25600) / 100.0f;
        for (int &l : lab) l = (int)(lcg(s) % (uint32_t)nsc);
        Probe p(dim, part, nsc);
        p.set(x, n, c, lab);
        const double got = p.distortion(false);
        // the reference: strip each part, distortion of the strip, root of
        // the sum of squares (quantizer.h:279-300)
        double e = 0.0;
        std::vector<float> strip((size_t)n * bs);
        for (int i = 0; i < part; ++i) {
            for (int r = 0; r < n; ++r)
                std::memcpy(&strip[(size_t)r * bs], &x[(size_t)r * dim + (size_t)i * bs], 4 * bs);
            const double et = scref_distortion(strip.data(), &c[(size_t)i * nsc * bs],
                                               &lab[(size_t)i * n], bs, n, nsc);
            e += et * et;
        }
        const double want = std::sqrt(e);
        if (std::memcmp(&got, &want, sizeof got) != 0) {
            std::printf("distortion dim=%d part=%d: %.17g vs reference %.17g\n", dim, part, got,
                        want);
            ++fails;
        }
    }
    std::printf("%s: %d failure(s)\n", fails ? "FAILED" : "OK", fails);
    return fails ? 1 : 0;
}
