// sc_runtime.hip -- context, errors, device memory and the seeded generator.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <new>
#include <utility>

#include "sc_internal.h"

static thread_local char g_err[512];

// Fixed device scratch for the allocation-free entry points (partial results
// of one call).  Every GPU the library runs on holds its own copy for the
// life of the process.  Host code never passes this name to a kernel: the
// device address of this GPU's copy is looked up once per context (below)
// and kept in sc_ctx::fixed_scratch.
constexpr size_t kFixedScratchBytes = (size_t)16 << 20;   // 16 MiB
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

void sc_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *sc_last_error(void) { return g_err; }

extern "C" const char *sc_version(void) {
    return "libsc_amd gfx950 (-ffp-contract=off, no packed fp32)";
}

int sc_set_max_lds(const void *kern, size_t bytes) {
    // the attribute is set once per kernel and device, to the largest size
    // asked for so far (setting it costs a runtime call per launch otherwise)
    static std::mutex mu;
    static std::map<std::pair<const void *, int>, size_t> set;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    std::lock_guard<std::mutex> lock(mu);
    size_t &cur = set[{kern, dev}];
    if (bytes <= cur) return SC_OK;
    SC_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
    return SC_OK;
}

int sc_ctx_enter(sc_ctx *ctx) {
    if (!ctx) { sc_set_error("null context"); return SC_ERR_ARG; }
    SC_HIP(hipSetDevice(ctx->device));
    return SC_OK;
}

extern "C" int sc_ctx_create(int device, sc_ctx **out) {
    if (!out) { sc_set_error("null out"); return SC_ERR_ARG; }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        sc_set_error("no HIP device");
        return SC_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) { sc_set_error("bad device %d", device); return SC_ERR_ARG; }
    hipDeviceProp_t prop;
    SC_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        sc_set_error("device %d is %s, this build targets gfx950", device, prop.gcnArchName);
        return SC_ERR_NO_DEVICE;
    }
    SC_HIP(hipSetDevice(device));
    sc_ctx *c = new (std::nothrow) sc_ctx();
    if (!c) return SC_ERR_NOMEM;
    c->device = device;
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    c->n_cu = prop.multiProcessorCount;
    // this device's copy of the static scratch (the device is current here)
    c->fixed_scratch = nullptr;
    c->fixed_scratch_bytes = 0;
    hipError_t se = hipGetSymbolAddress(&c->fixed_scratch, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !c->fixed_scratch) {
        delete c;
        sc_set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(se));
        return SC_ERR_HIP;
    }
    c->fixed_scratch_bytes = kFixedScratchBytes;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        sc_set_error("hipStreamCreate failed");
        return SC_ERR_HIP;
    }
    *out = c;
    return SC_OK;
}

int sc_ctx_side(sc_ctx *ctx) {
    if (ctx->side) return SC_OK;
    SC_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    SC_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    SC_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    return SC_OK;
}

extern "C" int sc_ctx_destroy(sc_ctx *ctx) {
    if (!ctx) return SC_OK;
    SC_TRY(sc_ctx_enter(ctx));
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SC_OK;
}

extern "C" int sc_ctx_sync(sc_ctx *ctx) {
    SC_TRY(sc_ctx_enter(ctx));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

extern "C" void *sc_ctx_stream(sc_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int sc_scratch(sc_ctx *ctx, size_t bytes, void **ptr) {
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes < bytes) {
        SC_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->scratch) SC_HIP(hipFree(ctx->scratch));
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        size_t want = bytes + bytes / 4;
        if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
            sc_set_error("scratch allocation of %zu bytes failed", want);
            return SC_ERR_NOMEM;
        }
        ctx->scratch_bytes = want;
    }
    *ptr = ctx->scratch;
    return SC_OK;
}

extern "C" int sc_malloc(sc_ctx *ctx, size_t bytes, void **d_ptr) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!d_ptr) return SC_ERR_ARG;
    if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
        sc_set_error("hipMalloc(%zu) failed", bytes);
        return SC_ERR_NOMEM;
    }
    return SC_OK;
}

extern "C" int sc_free(sc_ctx *ctx, void *d_ptr) {
    SC_TRY(sc_ctx_enter(ctx));
    if (d_ptr) {
        SC_HIP(hipStreamSynchronize(ctx->stream));
        SC_HIP(hipFree(d_ptr));
    }
    return SC_OK;
}

extern "C" int sc_memcpy_h2d(sc_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
    SC_TRY(sc_ctx_enter(ctx));
    SC_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

extern "C" int sc_memcpy_d2h(sc_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
    SC_TRY(sc_ctx_enter(ctx));
    SC_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    SC_HIP(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

extern "C" int sc_memset(sc_ctx *ctx, void *d_dst, int value, size_t bytes) {
    SC_TRY(sc_ctx_enter(ctx));
    SC_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    return SC_OK;
}

// ------------------------------------------------------------------------
// Seeded Gaussian mixture (integer-only; the host oracle draws the same
// integers).  Mixture parameters come from counter (component, dim, 0, 0)
// and (component, ~0, 0, 0); row r of stream s picks its component from
// (r_lo, r_hi, s, ~0) and dimension d's noise from (r_lo, r_hi, s, 2d) and
// (r_lo, r_hi, s, 2d+1): an Irwin-Hall sum of 12 16-bit draws.
// ------------------------------------------------------------------------
__global__ void k_gen_rows(sc_mix m, uint32_t stream, int64_t row0, int64_t nrows,
                           int32_t dim, float *__restrict__ out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t total = nrows * (int64_t)dim;
    if (e >= total) return;
    int64_t i = e / dim;
    uint32_t d = (uint32_t)(e - i * dim);
    uint64_t row = (uint64_t)(row0 + i);
    uint32_t rlo = (uint32_t)row, rhi = (uint32_t)(row >> 32);
    uint32_t o[4], a[4], b[4];
    sc_philox(rlo, rhi, stream, 0xFFFFFFFFu, m.seed, 0u, o);
    uint32_t c = o[0] % (uint32_t)m.ncomp;
    sc_philox(c, d, 0u, 0u, m.seed, 0u, o);
    int64_t mu_q = m.mu_lo + (int64_t)(o[0] % (uint64_t)(m.mu_span + 1));
    sc_philox(c, 0xFFFFFFFFu, 0u, 0u, m.seed, 0u, o);
    int64_t sig_q = m.sig_lo + (int64_t)(o[0] % (uint64_t)(m.sig_span + 1));
    sc_philox(rlo, rhi, stream, 2u * d, m.seed, 0u, a);
    sc_philox(rlo, rhi, stream, 2u * d + 1u, m.seed, 0u, b);
    uint32_t w[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
    int64_t s = 0;
#pragma unroll
    for (int t = 0; t < 6; ++t) s += (int64_t)(w[t] & 0xFFFFu) + (int64_t)(w[t] >> 16);
    int64_t z2 = 2 * s - 12 * 65535;
    int64_t num = mu_q * ((int64_t)1 << 17) + sig_q * z2;
    int shift = m.frac_bits + 17 - m.out_bits;
    int64_t xq = (num + ((int64_t)1 << (shift - 1))) >> shift;
    xq = xq < 0 ? 0 : (xq > m.maxq ? m.maxq : xq);
    out[e] = (float)ldexp((double)xq, -m.out_bits);
}

extern "C" int sc_gen_rows(sc_ctx *ctx, const sc_mix *mix, uint32_t stream, int64_t row0,
                           int64_t nrows, int32_t dim, float *d_out) {
    SC_TRY(sc_ctx_enter(ctx));
    if (!mix || !d_out || nrows < 0 || dim <= 0 || mix->ncomp <= 0 ||
        mix->frac_bits + 17 - mix->out_bits < 1) {
        sc_set_error("sc_gen_rows: bad arguments");
        return SC_ERR_ARG;
    }
    int64_t total = nrows * (int64_t)dim;
    if (total == 0) return SC_OK;
    hipLaunchKernelGGL(k_gen_rows, dim3((unsigned)sc_cdiv(total, 256)), dim3(256), 0,
                       ctx->stream, *mix, stream, row0, nrows, dim, d_out);
    SC_CHECK_LAUNCH();
    return SC_OK;
}
