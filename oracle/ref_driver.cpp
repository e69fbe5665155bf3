// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver over the parts of the reference that compile from
// their own sources without third-party stand-ins: the S1 distance
// (lib/simple-cluster/include/utilities.h:194-230), Encoder::encode<float>
// (include/encoder.h:69-192), Encoder::distribution (src/encoder.cpp:170-205),
// vector_base (include/sc_utilities.h:156-185) and the array heaps
// pq_insert/pq_pop, pq_insert4/pq_pop4, pq_insert6/pq_pop6
// (include/sc_utilities.h:229-300, 469-675) and SimpleCluster::distortion
// (lib/simple-cluster/include/k-means.h:444-464) and the quantizers' residual
// passes PQQuantizer / SCQuantizer::calc_residual_vector
// (include/quantizer.h:456-510, include/sc_quantizer.h:72-133).  The
// reference's query classes and SCEncoder include <cblas.h>
// (OpenBLAS fork, not vendored) and are not built.
//
// Built by oracle/Makefile into oracle/_ref/libscref.so straight from the
// reference tree; used by tests/ to pin the C oracle.
#include <cstdint>
#include <cstring>
#include <vector>

#include "encoder.h"
#include "k-means.h"
#include "quantizer.h"
#include "sc_quantizer.h"
#include "sc_utilities.h"
#include "utilities.h"

namespace {
// Gives the driver read access to the encoder's protected outputs.
class EncoderProbe : public SC::Encoder {
 public:
  const ushort *cid_ptr() const { return cid; }
  const unsigned char *codes_ptr() const { return codes; }
  const std::vector<SC::Bucket> &buckets() const { return ivf; }
  int cells() const { return size; }
  int n() const { return config.N; }
  int mc() const { return config.mc; }
  int mp() const { return config.mp; }
};
}  // namespace

extern "C" {

double scref_dist(const float *x, const float *y, int d) {
  return SimpleCluster::distance_l2_square<float>(const_cast<float *>(x),
                                                  const_cast<float *>(y), d);
}

int scref_vector_base(const uint16_t *code, int size, int base) {
  return SC::vector_base(const_cast<ushort *>(code), size, base);
}

// Runs Encoder::load_codebooks + encode<float> + distribution on files the
// caller wrote (fvecs with a 4-byte row header; codebooks as
// [k, m, dim] float header + k*dim floats).  Outputs: cid [N*mc],
// codes [N*mp], L (inclusive prefix, [cells]), pid [N], codes in cell order.
int scref_encode_imi(const char *fvecs, const char *cq_path, const char *pq_path,
                     uint16_t *cid, uint8_t *codes, int32_t *L, int32_t *pid,
                     uint8_t *codes_sorted, int64_t cap_n, int64_t cap_cells) {
  EncoderProbe enc;
  enc.load_codebooks(cq_path, pq_path, false);
  enc.encode<float>(fvecs, 4, false);
  enc.distribution(false);
  const int64_t n = enc.n(), mc = enc.mc(), mp = enc.mp();
  if (n > cap_n || enc.cells() > cap_cells) return -1;
  std::memcpy(cid, enc.cid_ptr(), sizeof(uint16_t) * n * mc);
  std::memcpy(codes, enc.codes_ptr(), (size_t)(n * mp));
  int64_t acc = 0, p = 0;
  const auto &b = enc.buckets();
  for (int c = 0; c < enc.cells(); ++c) {
    acc += b[c].L;
    L[c] = (int32_t)acc;
    for (int j = 0; j < b[c].L; ++j) {
      pid[p] = b[c].pid[j];
      std::memcpy(codes_sorted + p * mp, &b[c].codes[(size_t)j * mp], (size_t)mp);
      ++p;
    }
  }
  return (int)n;
}

void scref_heap4_insert(float *v, int32_t *h1, int32_t *h2, int32_t *h3,
                        int32_t *h4, float data, int a, int b, int c, int d,
                        int32_t *pos, int cap) {
  int p = *pos;
  SC::pq_insert4(v, h1, h2, h3, h4, data, a, b, c, d, p, cap, false);
  *pos = p;
}

int scref_heap4_pop(float *v, int32_t *h1, int32_t *h2, int32_t *h3,
                    int32_t *h4, int32_t *out4, int32_t *pos) {
  int p = *pos;
  if (p <= 0) return 0;
  int a, b, c, d;
  SC::pq_pop4(v, h1, h2, h3, h4, a, b, c, d, p, false);
  out4[0] = a; out4[1] = b; out4[2] = c; out4[3] = d;
  *pos = p;
  return 1;
}

void scref_heap1_insert(float *v, int32_t *id, float data, int x, int32_t *pos,
                        int cap) {
  int p = *pos;
  SC::pq_insert(v, id, data, x, p, cap, false);
  *pos = p;
}

int scref_heap1_pop(float *v, int32_t *id, int32_t *pos) {
  int p = *pos;
  if (p <= 0) return -1;
  int r = SC::pq_pop(v, id, p, false);
  *pos = p;
  return r;
}

float scref_distortion(const float *data, const float *centers, const int *labels, int d,
                       int N, int k) {
  return SimpleCluster::distortion<float>(const_cast<float *>(data),
                                          const_cast<float *>(centers),
                                          const_cast<int *>(labels),
                                          SimpleCluster::DistanceType::NORM_L2, d, N, k, false);
}

void scref_heap6_insert(float *v, int32_t *h1, int32_t *h2, int32_t *h3, int32_t *h4,
                        int32_t *h5, int32_t *h6, float data, int a, int b, int c, int d,
                        int e, int f, int32_t *pos, int cap) {
  int p = *pos;
  SC::pq_insert6(v, h1, h2, h3, h4, h5, h6, data, a, b, c, d, e, f, p, cap, false);
  *pos = p;
}

int scref_heap6_pop(float *v, int32_t *h1, int32_t *h2, int32_t *h3, int32_t *h4,
                    int32_t *h5, int32_t *h6, int32_t *out6, int32_t *pos) {
  int p = *pos;
  if (p <= 0) return 0;
  int a, b, c, d, e, f;
  SC::pq_pop6(v, h1, h2, h3, h4, h5, h6, a, b, c, d, e, f, p, false);
  out6[0] = a; out6[1] = b; out6[2] = c; out6[3] = d; out6[4] = e; out6[5] = f;
  *pos = p;
  return 1;
}

// calc_residual_vector of PQQuantizer<float> (second = 0) or
// SCQuantizer<float> (second = 1), built with `part` partitions and one
// coarse codebook read from cq_path, in place on x [n][dim].
int scref_residual(float *x, int n, int dim, int part, const char *cq_path, int second) {
  float *d = static_cast<float *>(::operator new(sizeof(float) * (size_t)n * dim));
  std::memcpy(d, x, sizeof(float) * (size_t)n * dim);
  char *names[1] = {const_cast<char *>(cq_path)};
  if (second) {
    SC::SCQuantizer<float> q(dim, part, 1, 0, 1, false);
    q.load_data_mat(d, n);
    q.load_codebooks(names, false);
    q.calc_residual_vector(false);
    std::memcpy(x, q.data, sizeof(float) * (size_t)n * dim);
  } else {
    SC::PQQuantizer<float> q(dim, part, 1, 0, 1, false);
    q.load_data_mat(d, n);
    q.load_codebooks(names, false);
    q.calc_residual_vector(false);
    std::memcpy(x, q.data, sizeof(float) * (size_t)n * dim);
  }
  return 0;
}

}  // extern "C"
