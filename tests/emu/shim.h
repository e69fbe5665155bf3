// shim.h -- host emulation of the HIP device features the fused search and
// MFMA encode kernels use, so that its device code compiles as ordinary C++ and runs one
// workgroup as std::threads under AddressSanitizer or ThreadSanitizer:
//   * __syncthreads is a pthread barrier over the workgroup's threads;
//   * wave intrinsics (ballot, shuffles, readfirstlane) exchange values through
//     a per-wave barrier, so every lane of a wave must call them, as on the
//     device;
//   * __shared__ variables become statics (one workgroup at a time);
//   * atomics map to the compiler's __atomic builtins.
// Test infrastructure only: nothing in the library includes it.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...) unused
#define __restrict__
#define __align__(x) alignas(x)
#define __shared__ static

struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct uint2 { uint32_t x, y; };
struct uint4 { uint32_t x, y, z, w; };
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline uint2 make_uint2(uint32_t a, uint32_t b) { return {a, b}; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __int_as_float(int32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline int32_t __float_as_int(float f) { int32_t u; memcpy(&u, &f, 4); return u; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline uint64_t wall_clock64() { return 0; }
using std::min; using std::max;

struct Dim3 { unsigned x, y, z; };
extern thread_local Dim3 threadIdx, blockIdx;
extern Dim3 blockDim, gridDim;

// barrier (pthread, so ThreadSanitizer sees the ordering it gives)
#include <pthread.h>
struct Barrier {
    pthread_barrier_t b;
    void init(int k) { pthread_barrier_init(&b, nullptr, (unsigned)k); }
    void wait(const char *) { pthread_barrier_wait(&b); }
};
extern Barrier g_block_bar;
extern Barrier g_wave_bar[16];
extern uint64_t g_wave_x[16][64];
inline void __syncthreads() { g_block_bar.wait("__syncthreads"); }
inline uint64_t wave_exchange(uint64_t v, int src) {   // value of lane src
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    g_wave_x[w][l] = v;
    g_wave_bar[w].wait("wave");
    uint64_t r = g_wave_x[w][src & 63];
    g_wave_bar[w].wait("wave2");
    return r;
}
inline unsigned long long __ballot(int pred) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    g_wave_x[w][l] = pred ? 1 : 0;
    g_wave_bar[w].wait("ballot");
    unsigned long long m = 0;
    for (int i = 0; i < 64; ++i) if (g_wave_x[w][i]) m |= 1ull << i;
    g_wave_bar[w].wait("ballot2");
    return m;
}
template <class T> inline T shfl_any(T v, int src) {
    uint64_t u = 0; memcpy(&u, &v, sizeof(T));
    uint64_t r = wave_exchange(u, src);
    T o; memcpy(&o, &r, sizeof(T)); return o;
}
template <class T> inline T __shfl_xor(T v, int m) { return shfl_any(v, (threadIdx.x % 64) ^ m); }
template <class T> inline T __shfl(T v, int s) { return shfl_any(v, s); }
template <class T> inline T __shfl_up(T v, int d) {
    const int l = threadIdx.x % 64;
    T r = shfl_any(v, l >= d ? l - d : l);
    return r;
}
inline int __builtin_amdgcn_readfirstlane_emu(int v) { return shfl_any(v, 0); }
#define __builtin_amdgcn_readfirstlane(x) __builtin_amdgcn_readfirstlane_emu((int)(x))
inline int __builtin_amdgcn_readlane_emu(int v, int l) { return shfl_any(v, l); }
#define __builtin_amdgcn_readlane(x, l) __builtin_amdgcn_readlane_emu((int)(x), (int)(l))
template <class T, class U> inline T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> inline T atomicAnd(T *p, U v) { return __atomic_fetch_and(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> inline T atomicOr(T *p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <class T, class U> inline T atomicMax(T *p, U v) { T o = __atomic_load_n(p, __ATOMIC_SEQ_CST); while (o < (T)v && !__atomic_compare_exchange_n(p, &o, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {} return o; }
template <class T, class U> inline T atomicMin(T *p, U v) { T o = __atomic_load_n(p, __ATOMIC_SEQ_CST); while (o > (T)v && !__atomic_compare_exchange_n(p, &o, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {} return o; }

inline int __all(int pred) { return __ballot(pred) == __ballot(1); }
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return {a, b, c, d}; }
inline double __fma_rn(double a, double b, double c) { return std::fma(a, b, c); }

// v_mfma_f32_16x16x16bf16_1k of one wave: lane l holds A[l % 16][4 (l / 16) + j]
// and B[4 (l / 16) + j][l % 16] for j < 4 and receives C[4 (l / 16) + i][l % 16];
// products of BF16 values are exact in FP32, the sums are FP32 in k order.
extern short g_mfma_a[16][64][4], g_mfma_b[16][64][4];
inline float emu_bf16_float(short b) { return __uint_as_float((uint32_t)(uint16_t)b << 16); }
template <class A, class C> inline C emu_mfma_bf16(A a, A b, C c, int, int, int) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (int j = 0; j < 4; ++j) { g_mfma_a[w][l][j] = a[j]; g_mfma_b[w][l][j] = b[j]; }
    g_wave_bar[w].wait("mfma");
    C r = c;
    for (int i = 0; i < 4; ++i) {
        const int row = 4 * (l / 16) + i, col = l % 16;
        float s = r[i];
        for (int k = 0; k < 16; ++k)
            s += emu_bf16_float(g_mfma_a[w][row + 16 * (k / 4)][k % 4]) *
                 emu_bf16_float(g_mfma_b[w][col + 16 * (k / 4)][k % 4]);
        r[i] = s;
    }
    g_wave_bar[w].wait("mfma2");
    return r;
}

// v_mfma_f32_16x16x32_bf16 of one wave: lane l holds A[l % 16][8 (l / 16) + j]
// and B[8 (l / 16) + j][l % 16] for j < 8 and receives C[4 (l / 16) + i][l % 16]
// (the layout the device probe confirms); FP32 sums in k order.
extern float g_mfma32_a[16][64][8], g_mfma32_b[16][64][8];
template <class A, class C> inline C emu_mfma32_bf16(A a, A b, C c, int, int, int) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (int j = 0; j < 8; ++j) { g_mfma32_a[w][l][j] = (float)a[j]; g_mfma32_b[w][l][j] = (float)b[j]; }
    g_wave_bar[w].wait("mfma32");
    C r = c;
    for (int i = 0; i < 4; ++i) {
        const int row = 4 * (l / 16) + i, col = l % 16;
        float s = r[i];
        for (int k = 0; k < 32; ++k)
            s += g_mfma32_a[w][row + 16 * (k / 8)][k % 8] * g_mfma32_b[w][col + 16 * (k / 8)][k % 8];
        r[i] = s;
    }
    g_wave_bar[w].wait("mfma32b");
    return r;
}

#define SC_KIND_IVFADC 0
#define SC_KIND_IMI 1
#define SC_KIND_SC 2
inline uint32_t sc_float_key(float f) {
    uint32_t u = __float_as_uint(f);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
inline float sc_key_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}
