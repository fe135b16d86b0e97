// SIMT emulation on the CPU of the HIP subset the batch kernels use.
//
// TEST INFRASTRUCTURE ONLY (tests/emu): compiles the library's HIP kernel
// source with g++ and runs every kernel on the host.  Each lane is a fiber
// (its own stack, switched by emu_ctx_switch on one host thread); a wavefront
// runs its 64 lanes round-robin up to the next wavefront operation
// (__shfl and its variants, __ballot, readfirstlane, wave_barrier), where
// every lane of the wave must arrive, as converged hardware code does;
// otherwise the emulation stops with a message.  All waves of a block run
// together, interleaved at those points, and __syncthreads waits for every
// thread of the block.  Blocks run one after another.  "Device" memory is
// host memory filled with garbage when allocated; graphs are not emulated.
#pragma once
#include <atomic>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>
#include <math.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __constant__
#define __shared__ static
#define __launch_bounds__(...)
extern double emu_dyn_lds[];  // dynamic LDS of the running block (160 KiB)
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu_dyn_lds;

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx { unsigned x, y, z; };
extern unsigned emu_tid, emu_bid, emu_bdim, emu_gdim, emu_bid_y, emu_gdim_y;
inline emu_idx emu_threadIdx() { return {emu_tid, 0, 0}; }
inline emu_idx emu_blockIdx() { return {emu_bid, emu_bid_y, 0}; }
inline emu_idx emu_blockDim() { return {emu_bdim, 1, 1}; }
inline emu_idx emu_gridDim() { return {emu_gdim, emu_gdim_y, 1}; }
#define threadIdx emu_threadIdx()
#define blockIdx emu_blockIdx()
#define blockDim emu_blockDim()
#define gridDim emu_gridDim()

// x86-64 System V stack switch: saves the callee-saved registers on the
// current stack, stores its pointer in *save_sp and resumes load_sp.
extern "C" void emu_ctx_switch(void **save_sp, void *load_sp);

constexpr int kEmuMaxThreads = 512;  // threads per block

struct EmuWaveSched {  // the running block's threads
    void *main_sp;
    void *sp[kEmuMaxThreads];
    bool done[kEmuMaxThreads];
    int n;    // threads in the block
    int cur;  // threadIdx.x of the running thread
    unsigned long long x[kEmuMaxThreads];
    int bar_arrived;
    unsigned bar_gen;
    void (*body)(void *);
    void *arg;
    std::vector<char> stacks;
};
extern EmuWaveSched *emu_wave;

inline void emu_yield()
{
    EmuWaveSched &S = *emu_wave;
    const int lane = S.cur;
    emu_ctx_switch(&S.sp[lane], S.main_sp);
}
inline unsigned long long emu_exchange(unsigned long long v, int src)
{
    EmuWaveSched &S = *emu_wave;
    const int lane = S.cur;
    S.x[lane] = v;
    emu_yield();  // every lane of the wave has written
    const unsigned long long r = S.x[(lane & ~63) + (src & 63)];
    emu_yield();  // every lane of the wave has read
    return r;
}
inline double __shfl(double v, int src, int = 64)
{
    unsigned long long b;
    memcpy(&b, &v, 8);
    b = emu_exchange(b, src);
    double r;
    memcpy(&r, &b, 8);
    return r;
}
inline int __shfl(int v, int src, int = 64) { return (int)emu_exchange((unsigned long long)(unsigned)v, src); }
inline double __shfl_xor(double v, int m, int = 64) { return __shfl(v, (int)(emu_tid & 63) ^ m); }
inline int __shfl_xor(int v, int m, int = 64) { return __shfl(v, (int)(emu_tid & 63) ^ m); }
inline int __builtin_amdgcn_readfirstlane(int v) { return __shfl(v, 0); }
inline double __shfl_up(double v, unsigned d, int = 64)
{
    const int l = (int)(emu_tid & 63);
    return __shfl(v, l >= (int)d ? l - (int)d : l);
}
inline double __shfl_down(double v, unsigned d, int = 64)
{
    const int l = (int)(emu_tid & 63);
    return __shfl(v, l + (int)d < 64 ? l + (int)d : l);
}
// a wavefront sync point here (every lane arrives before any goes on)
inline void __builtin_amdgcn_wave_barrier() { emu_yield(); }
// memory fences: lanes run one at a time on one host thread, so ordering holds already
inline void __builtin_amdgcn_fence(int, const char *) {}
inline unsigned long long __ballot(int pred)
{
    EmuWaveSched &S = *emu_wave;
    S.x[S.cur] = pred ? 1 : 0;
    emu_yield();  // every lane of the wave has written
    const int base = S.cur & ~63;
    unsigned long long m = 0;
    for (int l = 0; l < 64; l++)
        if (S.x[base + l]) m |= 1ull << l;
    emu_yield();  // every lane of the wave has read
    return m;
}
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }  // the hardware estimate, exactly
// every thread of the block arrives before any goes on
inline void __syncthreads()
{
    EmuWaveSched &S = *emu_wave;
    const unsigned gen = S.bar_gen;
    if (++S.bar_arrived == S.n) {
        S.bar_arrived = 0;
        S.bar_gen++;
    }
    emu_yield();
    while (S.bar_gen == gen) emu_yield();
}

inline int atomicAdd(int *p, int v) { const int o = *p; *p = o + v; return o; }
inline int atomicOr(int *p, int v) { const int o = *p; *p = o | v; return o; }
inline double atomicAdd(double *p, double v) { const double o = *p; *p = o + v; return o; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNotSupported = 801 };
typedef void *hipStream_t;
typedef void *hipEvent_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum { hipDeviceAttributeMultiprocessorCount = 1 };
enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
inline hipError_t hipFuncSetAttribute(const void *, int, int) { return 0; }
enum { hipStreamCaptureModeThreadLocal = 1 };
enum hipMemoryType { hipMemoryTypeHost, hipMemoryTypeDevice, hipMemoryTypeManaged };
struct hipPointerAttribute_t { hipMemoryType type; };
#define HIP_SYMBOL(x) (&(x))
extern int emu_cus, emu_per_cu;
template <class T>
inline hipError_t hipMalloc(T **p, size_t n)
{
    *p = (T *)malloc(n ? n : 1);
    memset((void *)*p, 0xA5, n);  // not zero: kernels must not rely on fresh memory
    return hipSuccess;
}
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, int) { memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemcpyToSymbol(const void *sym, const void *s, size_t n) { memcpy((void *)sym, s, n); return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int *v, int, int) { *v = emu_cus; return hipSuccess; }
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t)
{
    *n = emu_per_cu;
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
// Streams and events: kernels and copies run when issued, so ordering holds trivially.
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (hipEvent_t)1; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) { memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) { *p = malloc(n > 0 ? n : 1); return *p ? hipSuccess : 2; }
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy2DAsync(void *d, size_t dpitch, const void *s, size_t spitch, size_t width, size_t height, int,
                                   hipStream_t)
{
    for (size_t r = 0; r < height; r++) memmove((char *)d + r * dpitch, (const char *)s + r * spitch, width);
    return hipSuccess;
}
inline hipError_t hipMemcpy2D(void *d, size_t dpitch, const void *s, size_t spitch, size_t width, size_t height, int k)
{
    return hipMemcpy2DAsync(d, dpitch, s, spitch, width, height, k, nullptr);
}
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, void *, size_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *) { a->type = hipMemoryTypeDevice; return hipSuccess; }

constexpr size_t kEmuStack = 256 * 1024;

extern "C" void emu_fiber_entry();

// Runs body(arg) on the n threads of one block (threadIdx.x = t), all waves
// together: each pass gives every unfinished thread one turn up to its next
// yield.  The lanes of a wave must yield and finish in the same passes.
inline void emu_run_block(int n, void (*body)(void *), void *arg)
{
    static EmuWaveSched S;
    if (S.stacks.empty()) S.stacks.resize(kEmuMaxThreads * kEmuStack);
    S.n = n;
    S.body = body;
    S.arg = arg;
    S.bar_arrived = 0;
    S.bar_gen = 0;
    for (int l = 0; l < n; l++) {
        S.done[l] = false;
        uintptr_t top = ((uintptr_t)(S.stacks.data() + (l + 1) * kEmuStack)) & ~(uintptr_t)15;
        void **sp = (void **)top;
        *--sp = nullptr;                      // return address of the entry (never used)
        *--sp = (void *)&emu_fiber_entry;     // where the first switch returns to
        for (int r = 0; r < 6; r++) *--sp = nullptr;  // rbp rbx r12 r13 r14 r15
        S.sp[l] = sp;
    }
    emu_wave = &S;
    for (;;) {
        int yielded = 0;
        for (int w0 = 0; w0 < n; w0 += 64) {
            int wy = 0, wf = 0, wdone = 0;
            for (int l = w0; l < w0 + 64; l++) {
                if (S.done[l]) {
                    wdone++;
                    continue;
                }
                S.cur = l;
                emu_tid = l;
                emu_ctx_switch(&S.main_sp, S.sp[l]);
                if (S.done[l]) wf++;
                else wy++;
            }
            if (wy && (wf || wdone)) {
                fprintf(stderr, "emulation: lanes of wave %d diverged at a wavefront operation (%d yielded, %d "
                                "finished now, %d before)\n", w0 / 64, wy, wf, wdone);
                abort();
            }
            yielded += wy;
        }
        if (yielded == 0) break;
        int finished = 0;
        for (int l = 0; l < n; l++) finished += S.done[l];
        if (finished && S.bar_arrived > 0 && S.bar_arrived + finished == n) {
            fprintf(stderr, "emulation: threads finished while others wait at __syncthreads\n");
            abort();
        }
    }
}

template <typename F>
void emu_call(void *p) { (*(F *)p)(); }

template <typename... KA, typename... A>
void emu_launch(void (*k)(KA...), dim3 g, dim3 b, size_t, hipStream_t, A... args)
{
    if (b.x % 64) {
        fprintf(stderr, "emulation: block size %u is not a multiple of 64\n", b.x);
        abort();
    }
    if (b.x > (unsigned)kEmuMaxThreads) {
        fprintf(stderr, "emulation: block size %u above %d\n", b.x, kEmuMaxThreads);
        abort();
    }
    if (b.y != 1 || b.z != 1 || g.z != 1) {
        fprintf(stderr, "emulation: blocks must be one-dimensional and grids at most two-dimensional\n");
        abort();
    }
    emu_bdim = b.x;
    emu_gdim = g.x;
    emu_gdim_y = g.y;
    auto body = [&]() { k(args...); };
    for (unsigned by = 0; by < g.y; by++)
        for (unsigned bx = 0; bx < g.x; bx++) {
            emu_bid = bx;
            emu_bid_y = by;
            emu_run_block((int)b.x, &emu_call<decltype(body)>, &body);
        }
}
#define hipLaunchKernelGGL(k, g, b, sh, st, ...) emu_launch(k, dim3(g), dim3(b), sh, st, ##__VA_ARGS__)
