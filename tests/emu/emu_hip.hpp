#include <chrono>
// emu_hip.hpp — CPU emulation of the HIP runtime subset used by colb_amd/csrc, for bounds-checked
// (AddressSanitizer) host builds of the kernels.  TEST INFRASTRUCTURE ONLY.
//
// Kernels run block by block; each GPU thread of a block is a fibre (ucontext).  __shared__
// variables are function statics (blocks run one at a time).  __syncthreads waits for every live
// fibre of the block, wave operations (__shfl_xor, votes, the matrix instruction, wave barriers)
// for every live fibre of the caller's 64-lane wave; fibres that return early drop out of both
// (as exited waves do on the GPU).  A barrier that can never complete aborts with a message.
#pragma once
#include <atomic>
#include <ucontext.h>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3() = default;
    dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct double4 { double x, y, z, w; };
struct float4 { float x, y, z, w; };
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
inline double4 make_double4(double a, double b, double c, double d) { return {a, b, c, d}; }
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline int2 make_int2(int a, int b) { return {a, b}; }
inline int4 make_int4(int a, int b, int c, int d) { return {a, b, c, d}; }

inline dim3 threadIdx, blockIdx, blockDim, gridDim;

// ---- cooperative fibers: one per GPU thread of the running block -----------------------------
struct EmuFiber {
    ucontext_t ctx;
    std::vector<char> stack;
    bool done = false;
    int wait = 0;  // 0 runnable, 1 at a block barrier, 2 at a wave barrier
    int line = 0;  // source line of the block barrier it waits at
};
struct EmuSched {
    ucontext_t main;
    std::vector<EmuFiber> fib;
    unsigned cur = 0;
    std::function<void()> body;
    std::vector<double> xbuf;
    std::vector<float> mbuf;  // matrix-instruction operand exchange, 8 values per lane
    std::vector<double> dbuf;  // the same for the FP64/FP32 16x16x4 instructions, 2 values per lane
    const void* main_bottom = nullptr;
    size_t main_size = 0;
};
inline EmuSched* emu_sched = nullptr;

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_START_SWITCH(fake, bottom, size) __sanitizer_start_switch_fiber(fake, bottom, size)
#define EMU_FINISH_SWITCH(fake, bottom, size) __sanitizer_finish_switch_fiber(fake, bottom, size)
#else
#define EMU_START_SWITCH(fake, bottom, size) ((void)0)
#define EMU_FINISH_SWITCH(fake, bottom, size) ((void)0)
#endif

inline void emu_yield() {
    EmuSched* S = emu_sched;
    EmuFiber& f = S->fib[S->cur];
    void* fake = nullptr;
    EMU_START_SWITCH(&fake, S->main_bottom, S->main_size);
    swapcontext(&f.ctx, &S->main);
    EMU_FINISH_SWITCH(fake, &S->main_bottom, &S->main_size);
}

inline void emu_fiber_entry() {
    EmuSched* S = emu_sched;
    EMU_FINISH_SWITCH(nullptr, &S->main_bottom, &S->main_size);
    S->body();
    S->fib[S->cur].done = true;
    EMU_START_SWITCH(nullptr, S->main_bottom, S->main_size);
}

inline void emu_block_sync(int line = 0) { emu_sched->fib[emu_sched->cur].wait = 1; emu_sched->fib[emu_sched->cur].line = line; emu_yield(); }
inline void emu_wave_sync() { emu_sched->fib[emu_sched->cur].wait = 2; emu_yield(); }
#define __syncthreads() emu_block_sync(__LINE__)
inline void __threadfence() {}

template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    (void)width;
    static_assert(sizeof(T) <= sizeof(double), "shfl of wide type");
    double w = 0;
    std::memcpy(&w, &v, sizeof(T));
    unsigned lane = threadIdx.x;
    emu_sched->xbuf[lane] = w;
    emu_wave_sync();
    double o = emu_sched->xbuf[lane ^ (unsigned)mask];
    emu_wave_sync();
    T r;
    std::memcpy(&r, &o, sizeof(T));
    return r;
}

// read lane `src` (0..63) of the caller's wave
template <class T>
inline T __shfl(T v, int src, int width = 64) {
    (void)width;
    static_assert(sizeof(T) <= sizeof(double), "shfl of wide type");
    double w = 0;
    std::memcpy(&w, &v, sizeof(T));
    unsigned lane = threadIdx.x;
    emu_sched->xbuf[lane] = w;
    emu_wave_sync();
    double o = emu_sched->xbuf[(lane & ~63u) + ((unsigned)src & 63u)];
    emu_wave_sync();
    T r;
    std::memcpy(&r, &o, sizeof(T));
    return r;
}

// wave vote over the lanes of the caller's 64-lane wave (every lane of the wave must call it)
inline int emu_any(int pred) {
    unsigned lane = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    emu_sched->xbuf[lane] = pred ? 1.0 : 0.0;
    emu_wave_sync();
    unsigned w0 = lane & ~63u, nt = blockDim.x * blockDim.y * blockDim.z;
    int r = 0;
    for (unsigned k = w0; k < w0 + 64 && k < nt; ++k)
        if (emu_sched->xbuf[k] != 0.0) r = 1;
    emu_wave_sync();
    return r;
}

inline unsigned __float_as_uint(float f) { unsigned u; std::memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(unsigned u) { float f; std::memcpy(&f, &u, 4); return f; }

// v_mfma_f32_16x16x16_bf16 over the 64 lanes of the caller's wave: lane l supplies A(l%16, 4(l/16)+u)
// and B(4(l/16)+u, l%16), u = 0..3, and receives D(4(l/16)+u, l%16) = C + sum_k A(., k) B(k, .)
template <class S4, class F4>
inline F4 __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(S4 a, S4 b, F4 c, int, int, int) {
    const unsigned lane = threadIdx.x % 64, w0 = threadIdx.x - lane;
    auto bf = [](short h) { return __uint_as_float((unsigned)(unsigned short)h << 16); };
    float* mb = emu_sched->mbuf.data();
    for (int u = 0; u < 4; ++u) { mb[8 * threadIdx.x + u] = bf(a[u]); mb[8 * threadIdx.x + 4 + u] = bf(b[u]); }
    emu_wave_sync();
    F4 d = c;
    const int n = lane % 16;
    for (int u = 0; u < 4; ++u) {
        const int m = 4 * (lane / 16) + u;
        float acc = 0.f;
        for (int k = 0; k < 16; ++k) {
            const float av = mb[8 * (w0 + m + 16 * (k / 4)) + (k % 4)];
            const float bv = mb[8 * (w0 + n + 16 * (k / 4)) + 4 + (k % 4)];
            acc += av * bv;
        }
        d[u] = c[u] + acc;
    }
    emu_wave_sync();
    return d;
}

// v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32 over the 64 lanes of the caller's wave: lane l
// supplies A(l%16, l/16) and B(l/16, l%16) and receives, in slot u = 0..3, D(R, l%16) = C +
// sum_k A(R, k) B(k, l%16) with R = l/16 + 4u (FP64; rows interleaved) or 4(l/16) + u (FP32), the
// layouts measured on the hardware; the products are summed in k order, each rounded to T
template <class T, class V4>
inline V4 emu_mfma_16x16x4(T a, T b, V4 c) {
    const unsigned lane = threadIdx.x % 64, w0 = threadIdx.x - lane;
    double* db = emu_sched->dbuf.data();
    db[2 * threadIdx.x] = (double)a;
    db[2 * threadIdx.x + 1] = (double)b;
    emu_wave_sync();
    V4 d = c;
    const int n = lane % 16;
    for (int u = 0; u < 4; ++u) {
        const int m = sizeof(T) == 8 ? (int)(lane / 16) + 4 * u : 4 * (int)(lane / 16) + u;
        T acc = c[u];
        for (int k = 0; k < 4; ++k) acc = std::fma((T)db[2 * (w0 + m + 16 * k)], (T)db[2 * (w0 + n + 16 * k) + 1], acc);
        d[u] = acc;
    }
    emu_wave_sync();
    return d;
}
template <class V4>
inline V4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, V4 c, int, int, int) {
    return emu_mfma_16x16x4<double>(a, b, c);
}
template <class V4>
inline V4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, V4 c, int, int, int) {
    return emu_mfma_16x16x4<float>(a, b, c);
}

inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline int atomicMax(int* p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline float nextafterf(float a, float b) { return std::nextafter(a, b); }
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) {
    unsigned long long old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
    unsigned long long old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
// device wall clock: the host's monotonic clock in microseconds
inline long long wall_clock64() {
    return std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
#define hipDeviceAttributeWallClockRate 1001
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
inline float __double2float_rd(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float __double2float_ru(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long v) { double r; std::memcpy(&r, &v, 8); return r; }
using std::fmin; using std::fmax; using std::sqrt; using std::fabs; using std::isfinite;
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }

// ---- runtime API ----------------------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorInvalidConfiguration = 9, hipErrorNotSupported = 801 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
inline thread_local hipError_t emu_last_error = hipSuccess;
inline const char* hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline const char* hipGetErrorName(hipError_t e) { return e ? "hipErrorEmulated" : "hipSuccess"; }
inline hipError_t hipGetLastError() { hipError_t e = emu_last_error; emu_last_error = hipSuccess; return e; }
// scoped atomics and cooperative launch: the emulation runs blocks one after another, so a
// cooperative (grid-synchronising) launch is reported as unsupported and callers fall back
#define __HIP_MEMORY_SCOPE_AGENT 4
template <class T> inline T __hip_atomic_load(T* p, int, int) { T r; __atomic_load(p, &r, __ATOMIC_SEQ_CST); return r; }
template <class T> inline void __hip_atomic_store(T* p, T v, int, int) { __atomic_store(p, &v, __ATOMIC_SEQ_CST); }
template <class T, class U> inline T __hip_atomic_fetch_add(T* p, U v, int, int) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
inline double __builtin_amdgcn_rcp(double a) { return 1.0 / a; }
inline double __builtin_amdgcn_rsq(double a) { return 1.0 / std::sqrt(a); }
inline void __builtin_amdgcn_s_sleep(int) { emu_yield(); }
inline void __builtin_amdgcn_s_waitcnt(int) {}
inline void __builtin_amdgcn_wave_barrier() { emu_wave_sync(); }
inline void __builtin_amdgcn_fence(int, const char*) {}
inline hipError_t hipLaunchCooperativeKernel(const void*, dim3, dim3, void**, unsigned, hipStream_t) { return hipErrorNotSupported; }
#define hipDeviceAttributeMultiprocessorCount 0
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int* v, int, int) { *v = 1; return hipSuccess; }
// one co-resident workgroup: cooperative launches of a single block run (emu_launch_cooperative)
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int, size_t) { *n = 1; return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n) { *p = std::malloc(n); if (*p) std::memset(*p, 0xA5, n); return *p ? hipSuccess : 2; }
template <class T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void** p, size_t n, int) { *p = std::malloc(n); return *p ? hipSuccess : 2; }
inline hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { if (n) std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { if (n) std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, hipMemcpyKind,
                                   hipStream_t) {
    for (size_t r = 0; r < h; ++r) std::memcpy((char*)d + r * dp, (const char*)s + r * sp, w);
    return hipSuccess;
}
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
inline hipError_t hipMemset(void* d, int v, size_t n) { if (n) std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)0x1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// no stream capture in the emulation: a capture cannot begin, so the caller launches eagerly
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeRelaxed = 2 };
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1, hipStreamCaptureStatusInvalidated = 2 };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorNotSupported; }
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* s) { *s = hipStreamCaptureStatusNone; return hipSuccess; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, hipGraphNode_t*, char*, size_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (void*)0x1; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }

template <class K, class... Args>
void emu_launch(const char* name, K kernel, dim3 grid, dim3 block, Args... args) {
    const unsigned nt = block.x * block.y * block.z;
    const unsigned nblk = grid.x * grid.y * grid.z;
    if (nt == 0 || nblk == 0 || nt > 1024) { emu_last_error = hipErrorInvalidConfiguration; return; }
    static const bool trace = std::getenv("EMU_TRACE") != nullptr;
    if (trace) std::fprintf(stderr, "launch %s grid %u block %u\n", name, nblk, nt);
    EmuSched S;
    S.xbuf.assign(nt, 0.0);
    S.mbuf.assign((size_t)nt * 8, 0.f);
    S.dbuf.assign((size_t)nt * 2, 0.0);
    S.body = [&]() { kernel(args...); };
    EmuSched* prev = emu_sched;
    emu_sched = &S;
    blockDim = block;
    gridDim = grid;
    S.fib.resize(nt);
    static std::vector<std::vector<char>> pool;  // fiber stacks reused across launches
    if (pool.size() < nt) pool.resize(nt);
    for (unsigned t = 0; t < nt; ++t) {
        if (pool[t].empty()) pool[t].resize(128 * 1024);
        S.fib[t].stack.swap(pool[t]);
    }
    for (unsigned b = 0; b < nblk; ++b) {
        blockIdx = dim3(b % grid.x, (b / grid.x) % grid.y, b / (grid.x * grid.y));
        for (unsigned t = 0; t < nt; ++t) {
            EmuFiber& f = S.fib[t];
            f.done = false;
            getcontext(&f.ctx);
            f.ctx.uc_stack.ss_sp = f.stack.data();
            f.ctx.uc_stack.ss_size = f.stack.size();
            f.ctx.uc_link = &S.main;
            makecontext(&f.ctx, emu_fiber_entry, 0);
        }
        for (unsigned t = 0; t < nt; ++t) S.fib[t].wait = 0;
        static const bool shuffle = std::getenv("EMU_SHUFFLE") != nullptr;
        std::vector<unsigned> order(nt);
        for (unsigned t = 0; t < nt; ++t) order[t] = t;
        static unsigned long long rng = 88172645463325252ull;
        while (true) {
            if (shuffle)
                for (unsigned t = nt - 1; t > 0; --t) {
                    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
                    unsigned j = (unsigned)(rng % (t + 1));
                    unsigned tmp = order[t]; order[t] = order[j]; order[j] = tmp;
                }
            bool ran = false;
            for (unsigned oi = 0; oi < nt; ++oi) {
                unsigned t = order[oi];
                EmuFiber& f = S.fib[t];
                if (f.done || f.wait) continue;
                ran = true;
                S.cur = t;
                threadIdx = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                void* fake = nullptr;
                EMU_START_SWITCH(&fake, f.stack.data(), f.stack.size());
                swapcontext(&S.main, &f.ctx);
                EMU_FINISH_SWITCH(fake, nullptr, nullptr);
            }
            // release the barriers every live fibre of their scope has reached
            unsigned live = 0, at_block = 0;
            for (unsigned t = 0; t < nt; ++t)
                if (!S.fib[t].done) { ++live; at_block += S.fib[t].wait == 1; }
            if (live == 0) break;
            bool released = false;
            if (at_block == live) {
                for (unsigned t = 0; t < nt; ++t) S.fib[t].wait = 0;
                released = true;
            }
            for (unsigned w0 = 0; w0 < nt && !released; w0 += 64) {
                bool all = true, any = false;
                for (unsigned t = w0; t < w0 + 64 && t < nt; ++t)
                    if (!S.fib[t].done) { any = true; all = all && S.fib[t].wait == 2; }
                if (any && all) {
                    for (unsigned t = w0; t < w0 + 64 && t < nt; ++t) S.fib[t].wait = 0;
                    released = true;
                }
            }
            if (!ran && !released) {
                std::fprintf(stderr, "emu: %s block %u: barrier can never complete; fibres (wait, done):", name, b);
                for (unsigned t = 0; t < nt; ++t)
                    std::fprintf(stderr, " %u:%d%s@%d", t, S.fib[t].wait, S.fib[t].done ? "d" : "", S.fib[t].line);
                std::fprintf(stderr, "\n");
                std::abort();
            }
        }
    }
    for (unsigned t = 0; t < nt; ++t) S.fib[t].stack.swap(pool[t]);
    emu_sched = prev;
}
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) emu_launch(#kernel, kernel, dim3(grid), dim3(block), __VA_ARGS__)
