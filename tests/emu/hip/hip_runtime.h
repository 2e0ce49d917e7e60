// Test infrastructure: host stand-in for the HIP runtime calls and device
// intrinsics that stress-particle-sph_amd/csrc uses, so that its kernels can
// be compiled by g++ (with AddressSanitizer) and executed on the CPU, one
// block at a time (emu_rt.cpp).  Device memory is host memory; streams,
// events and copies are immediate.  Not part of the product.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#include <algorithm>
#include <ucontext.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

using std::min;
using std::max;
using std::isfinite;

struct dim3e { unsigned x = 0, y = 0, z = 0; };
extern dim3e threadIdx, blockIdx, blockDim, gridDim;
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

struct double2 { double x, y; };
struct double3 { double x, y, z; };
struct double4 { double x, y, z, w; };
struct int2 { int x, y; };
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
inline double2 make_double2(double a, double b) { return {a, b}; }
inline double3 make_double3(double a, double b, double c) { return {a, b, c}; }
inline double4 make_double4(double a, double b, double c, double d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline float2 make_float2(float a, float b) { return {a, b}; }
inline float3 make_float3(float a, float b, float c) { return {a, b, c}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }

namespace emu {
bool needs_sync(const char *name);
void sync_block();
int sync_block_or(int p);
void wave_put(const void *v, int bytes);
void wave_get(int lane, void *v, int bytes);
// run body for every thread of grid x block; sync: the kernel uses barriers
// or wave exchanges (its threads then run as fibers)
void launch(dim3 grid, unsigned block, size_t lds, const std::function<void()> &body, bool sync);
// A launch as `kern<<<grid, block, lds>>>(args...)`: the arguments are
// evaluated once, on the host, and every thread gets the same copies.
template <typename... P, typename... A>
void launch_k(dim3 grid, unsigned block, size_t lds, bool sync, void (*kern)(P...), A... args) {
    launch(grid, block, lds, [=] { kern(args...); }, sync);
}
// Stream capture: launches are recorded (with their argument values) instead
// of run, and replayed by hipGraphLaunch; a memset or copy while capturing
// aborts (none is meant to be recorded).
extern bool capturing;
void check_not_capturing(const char *what);
struct Graph;
Graph *end_capture();
void replay(const Graph *g);
void drop(Graph *g);
extern unsigned char *lds_base;   // the launch's dynamic LDS
extern size_t lds_bytes;
}  // namespace emu

inline void __syncthreads() { emu::sync_block(); }
inline int __syncthreads_or(int p) { return emu::sync_block_or(p); }
template <typename T> T __shfl_up(T v, unsigned d, int w = 64) {
    emu::wave_put(&v, sizeof(T));
    int lane = threadIdx.x & 63, seg = lane & ~(w - 1);
    T r = v;
    if ((int)(lane - d) >= seg) emu::wave_get(lane - d, &r, sizeof(T));
    else emu::wave_get(lane, &r, sizeof(T));
    emu::wave_put(nullptr, 0);
    return r;
}
template <typename T> T __shfl_down(T v, unsigned d, int w = 64) {
    emu::wave_put(&v, sizeof(T));
    int lane = threadIdx.x & 63, seg = lane & ~(w - 1);
    T r = v;
    if ((int)(lane + d) < seg + w) emu::wave_get(lane + d, &r, sizeof(T));
    else emu::wave_get(lane, &r, sizeof(T));
    emu::wave_put(nullptr, 0);
    return r;
}
// lane src of the width-w segment holding this lane
template <typename T> T __shfl(T v, int src, int w = 64) {
    emu::wave_put(&v, sizeof(T));
    int lane = threadIdx.x & 63, seg = lane & ~(w - 1);
    T r = v;
    emu::wave_get(seg + (src & (w - 1)), &r, sizeof(T));
    emu::wave_put(nullptr, 0);
    return r;
}
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }
inline int __ffsll(unsigned long long v) { return __builtin_ffsll((long long)v); }
template <typename T> T atomicAdd(T *p, T v) { T o = *p; *p = o + v; return o; }
// hardware reciprocal and square root (within 1 ulp on the device)
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return sqrtf(x); }
#define __HIP_MEMORY_SCOPE_AGENT 0
inline unsigned __float_as_uint(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(unsigned u) { float f; memcpy(&f, &u, 4); return f; }
inline float __expf(float x) { return expf(x); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
// bit l of the result: lane l of the wave has p set (every live lane takes part)
inline unsigned long long __ballot(int p) {
    const int mine = p != 0;
    emu::wave_put(&mine, sizeof mine);
    unsigned long long m = 0;
    const unsigned w0 = threadIdx.x & ~63u;
    for (unsigned l = 0; l < 64 && w0 + l < blockDim.x; ++l) {
        int v = 0;
        emu::wave_get((int)l, &v, sizeof v);
        if (v) m |= 1ull << l;
    }
    emu::wave_put(nullptr, 0);
    return m;
}
// v_mfma_f32_16x16x16_bf16 of a wave: lane l holds A[l % 16][4 (l / 16) + q]
// and B[4 (l / 16) + q][l % 16] (bfloat16 bits), and gets C[4 (l / 16) + q]
// [l % 16] plus the products, summed in single precision.
typedef short emu_v4s __attribute__((vector_size(8)));
typedef float emu_v4f __attribute__((vector_size(16)));
inline emu_v4f __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(emu_v4s a, emu_v4s b, emu_v4f c, int, int, int) {
    struct AB {
        emu_v4s a, b;
    } mine{a, b};
    emu::wave_put(&mine, sizeof mine);
    const int l = threadIdx.x & 63;
    auto bf = [](short v) { return __uint_as_float((unsigned)(unsigned short)v << 16); };
    emu_v4f d = c;
    for (int q = 0; q < 4; ++q) {
        const int i = 4 * (l / 16) + q, j = l % 16;
        float acc = 0.0f;
        for (int k = 0; k < 16; ++k) {
            AB la, lb;
            emu::wave_get(i + 16 * (k / 4), &la, sizeof la);
            emu::wave_get(j + 16 * (k / 4), &lb, sizeof lb);
            acc += bf(la.a[k % 4]) * bf(lb.b[k % 4]);
        }
        d[q] += acc;
    }
    emu::wave_put(nullptr, 0);
    return d;
}
template <typename T> T __hip_atomic_load(const T *p, int, int) { return *p; }
inline int atomicOr(int *p, int v) { int o = *p; *p = o | v; return o; }
inline int atomicMin(int *p, int v) { int o = *p; *p = std::min(o, v); return o; }
inline unsigned long long atomicMax(unsigned long long *p, unsigned long long v) {
    unsigned long long o = *p; *p = std::max(o, v); return o;
}
inline long long __double_as_longlong(double d) { long long r; memcpy(&r, &d, sizeof r); return r; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorUnknown = 999 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
typedef void *hipStream_t;
typedef void *hipEvent_t;
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
inline hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n ? n : 1); return *p ? hipSuccess : hipErrorUnknown; }
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) { *p = malloc(n); return hipSuccess; }
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    emu::check_not_capturing("hipMemcpyAsync");
    memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    emu::check_not_capturing("hipMemcpy");
    memmove(d, s, n);
    return hipSuccess;
}
typedef emu::Graph *hipGraph_t;
typedef emu::Graph *hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    emu::capturing = true;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    *g = emu::end_capture();
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t) {
    *e = g;   // the recorded launches themselves; hipGraphDestroy of g keeps them
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    emu::drop(e);
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    emu::replay(e);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) {
    emu::check_not_capturing("hipMemsetAsync");
    memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount };
// a small device: a work-queue launch then has few workgroups, each taking many tiles
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 2; return hipSuccess; }
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *nb, const void *, int, size_t) {
    *nb = 4;
    return hipSuccess;
}
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
// a __device__ variable is an ordinary host object here: its address is itself
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t *s) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
// launches run in order here, so a stream waiting for an event has nothing to wait for
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
#define hipEventDisableTiming 2
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated"; }
