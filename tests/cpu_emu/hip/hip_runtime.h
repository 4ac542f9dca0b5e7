// CPU stand-in for the HIP runtime: runs the library's host code and its
// kernels on host threads (one pthread per GPU thread, a pthread barrier for
// __syncthreads, per-wave barriers for wavefront operations), synchronously,
// with graph capture recorded as a list of launches.  For AddressSanitizer
// runs only.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <pthread.h>
#include <thread>
#include <vector>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
struct uint3_ { unsigned x, y, z; };
extern thread_local uint3_ threadIdx, blockIdx, blockDim, gridDim;
extern pthread_barrier_t g_bar;
inline void __syncthreads() { pthread_barrier_wait(&g_bar); }

// Wavefront operations.  The 64 lanes of a wave run as 64 host threads, so a
// wave barrier is a barrier of those threads, and a DPP lane shift passes
// values through a per-wave buffer between two such barriers (every lane of
// the wave must reach it, as every lane of a real wave executes it).
constexpr int EMU_MAX_WAVES = 16;
extern pthread_barrier_t g_wbar[EMU_MAX_WAVES];
extern int g_dpp[EMU_MAX_WAVES][64];
inline void __builtin_amdgcn_wave_barrier() { pthread_barrier_wait(&g_wbar[threadIdx.x >> 6]); }
inline int __builtin_amdgcn_readfirstlane(int x) { return x; }   // callers pass wave-uniform values
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int, int, bool)
{
    const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    g_dpp[w][l] = src;
    pthread_barrier_wait(&g_wbar[w]);
    // 0x138 wave_shr:1 (lane l reads lane l-1), 0x130 wave_shl:1 (lane l+1);
    // a lane with no source keeps old
    const int from = ctrl == 0x138 ? l - 1 : (ctrl == 0x130 ? l + 1 : -1);
    const int v = (from >= 0 && from < 64) ? g_dpp[w][from] : old;
    pthread_barrier_wait(&g_wbar[w]);
    return v;
}
// mov_dpp with bound_ctrl: a lane with no source receives 0
inline int __builtin_amdgcn_mov_dpp(int src, int ctrl, int rm, int bm, bool bound_ctrl)
{
    return __builtin_amdgcn_update_dpp(0, src, ctrl, rm, bm, bound_ctrl);
}
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef std::vector<std::function<void()>>* hipGraph_t;
typedef std::vector<std::function<void()>>* hipGraphExec_t;
enum { hipMemcpyHostToDevice = 1, hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1,
       hipDeviceAttributeMultiprocessorCount = 1 };
extern std::vector<std::function<void()>>* g_capture;
extern long long g_launches;
inline void emu_run(dim3 g, dim3 b, std::function<void()> k)
{
    // one host thread per GPU thread of a workgroup, stepping through the
    // workgroups in order; a barrier between workgroups keeps the static
    // "__shared__" arrays private to the workgroup that runs
    pthread_barrier_t next;
    pthread_barrier_init(&g_bar, nullptr, b.x);
    pthread_barrier_init(&next, nullptr, b.x);
    const unsigned nw = (b.x + 63) / 64;
    for (unsigned w = 0; w < nw && w < (unsigned)EMU_MAX_WAVES; ++w)
        pthread_barrier_init(&g_wbar[w], nullptr, b.x - 64 * w < 64 ? b.x - 64 * w : 64);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < b.x; ++t)
        th.emplace_back([&, t] {
            for (unsigned by = 0; by < g.y; ++by)
                for (unsigned bx = 0; bx < g.x; ++bx) {
                    threadIdx = {t, 0, 0};
                    blockIdx = {bx, by, 0};
                    blockDim = {b.x, 1, 1};
                    gridDim = {g.x, g.y, 1};
                    k();
                    pthread_barrier_wait(&next);
                }
        });
    for (auto& x : th) x.join();
    for (unsigned w = 0; w < nw && w < (unsigned)EMU_MAX_WAVES; ++w) pthread_barrier_destroy(&g_wbar[w]);
    pthread_barrier_destroy(&next);
    pthread_barrier_destroy(&g_bar);
    ++g_launches;
}
inline void emu_launch(dim3 g, dim3 b, std::function<void()> k)
{
    if (g_capture) g_capture->push_back([=] { emu_run(g, b, k); });
    else emu_run(g, b, k);
}
#define hipLaunchKernelGGL(K, G, B, SH, S, ...) emu_launch(dim3(G), dim3(B), [=] { K(__VA_ARGS__); })
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int, hipStream_t) { memcpy(d, s, n); return 0; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { memset(d, v, n); return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipGetLastError() { return 0; }
inline const char* hipGetErrorString(hipError_t) { return "emu"; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = nullptr; return 0; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0; return 0; }
inline hipError_t hipEventDestroy(hipEvent_t) { return 0; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, int) { *s = (void*)1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
// launches run synchronously in enqueue order, so stream order and event waits hold trivially
enum { hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = nullptr; return 0; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return 0; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { g_capture = new std::vector<std::function<void()>>(); return 0; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = g_capture; g_capture = nullptr; return 0; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, void*, void*, size_t) { *x = g; return 0; }
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) { for (auto& f : *x) f(); return 0; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return 0; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return 0; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return 0; }
inline hipError_t hipDeviceGetAttribute(int* v, int, int) { *v = 256; return 0; }
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) { *n = 3; return 0; }
// a __device__ array is host memory here: its "device address" is its address
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return 0; }
