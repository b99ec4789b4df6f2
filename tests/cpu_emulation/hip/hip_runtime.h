// CPU emulation of the HIP runtime and device intrinsics used by the library,
// for memory-safety checks under AddressSanitizer (never part of the product).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <functional>

#define __HIP_DEVICE_COMPILE__ 1

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotReady = 600, hipErrorNotSupported = 801 };
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph_exec *hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipStreamNonBlocking = 1, hipHostMallocDefault = 0, hipEventDisableTiming = 2 };
typedef int hipStreamCaptureMode;
enum { hipStreamCaptureModeRelaxed = 2 };

hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamCreate(hipStream_t *);
hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipSetDevice(int);
hipError_t hipGetDeviceCount(int *);
hipError_t hipEventCreate(hipEvent_t *);
hipError_t hipEventCreateWithFlags(hipEvent_t *, unsigned);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipEventRecord(hipEvent_t, hipStream_t);
hipError_t hipEventQuery(hipEvent_t);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t);
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode);
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *);
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t);
hipError_t hipGraphExecDestroy(hipGraphExec_t);
hipError_t hipGraphDestroy(hipGraph_t);

// ---- device side -------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __constant__ static
#define HIP_SYMBOL(x) x

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float4 { float x, y, z, w; };
static inline float4 make_float4(float x, float y, float z, float w) { float4 v = {x, y, z, w}; return v; }

namespace emu {
struct idx3 { unsigned x, y, z; };
idx3 &thread_idx();
idx3 &block_idx();
idx3 &block_dim();
idx3 &grid_dim();
int lane();
// wave/block rendezvous: every live lane of the wave (block) deposits a
// value and gets all of them back
const uint64_t *wave_exchange(uint64_t v, int site, const char *what);
void block_barrier(int site);
void wave_barrier(int site);
void fail(const char *msg);
void run_kernel(const char *name, dim3 grid, dim3 block, void (*body)(void *), void *ctx);
// stream-ordered work: run at once, or with EMU_ASYNC_STREAMS queued on the
// stream until the host waits for it
void launch(hipStream_t stream, std::function<void()> op);
// stricter checks (see tests/cpu_emulation/build.py): EMU_POISON refills each
// __shared__ object at every workgroup's start; EMU_REDZONES puts a poisoned
// red zone after each arena sub-allocation
bool redzones();
void poison(void *p, size_t n);
void unpoison(void *p, size_t n);
void shared_region(void *p, size_t n);
}

#define threadIdx (emu::thread_idx())
#define blockIdx (emu::block_idx())
#define blockDim (emu::block_dim())
#define gridDim (emu::grid_dim())
#define __syncthreads() emu::block_barrier(__LINE__)

template <class T> static inline uint64_t emu_bits(T v) { uint64_t b = 0; memcpy(&b, &v, sizeof(T)); return b; }
template <class T> static inline T emu_from(uint64_t b) { T v; memcpy(&v, &b, sizeof(T)); return v; }

#define __ballot(p) emu_ballot((p) ? 1 : 0, __LINE__)
static inline unsigned long long emu_ballot(int p, int site) {
    const uint64_t *v = emu::wave_exchange((uint64_t)p | (1ull << 32), site, "ballot");
    unsigned long long m = 0;
    for (int l = 0; l < 64; ++l)
        if (v[l] == ((1ull << 32) | 1ull))
            m |= 1ull << l;
    return m;
}
template <class T> static inline T emu_shfl(T x, int src, int site) {
    const uint64_t *v = emu::wave_exchange(emu_bits(x), site, "shfl");
    return emu_from<T>(v[src & 63]);
}
template <class T> static inline T emu_shfl_up(T x, int d, int site) {
    const uint64_t *v = emu::wave_exchange(emu_bits(x), site, "shfl_up");
    int l = emu::lane();
    return l - d >= 0 ? emu_from<T>(v[l - d]) : x;
}
template <class T> static inline T emu_shfl_down(T x, int d, int site) {
    const uint64_t *v = emu::wave_exchange(emu_bits(x), site, "shfl_down");
    int l = emu::lane();
    return l + d < 64 ? emu_from<T>(v[l + d]) : x;
}
template <class T> static inline T emu_shfl_xor(T x, int m, int site) {
    const uint64_t *v = emu::wave_exchange(emu_bits(x), site, "shfl_xor");
    return emu_from<T>(v[(emu::lane() ^ m) & 63]);
}
#define __shfl(x, s, ...) emu_shfl((x), (s), __LINE__)
#define __shfl_up(x, d, ...) emu_shfl_up((x), (d), __LINE__)
#define __shfl_down(x, d, ...) emu_shfl_down((x), (d), __LINE__)
#define __shfl_xor(x, m, ...) emu_shfl_xor((x), (m), __LINE__)
static inline int emu_readlane(int x, int l, int site) {
    const uint64_t *v = emu::wave_exchange((uint64_t)(uint32_t)x, site, "readlane");
    return (int)(uint32_t)v[l & 63];
}
#define __builtin_amdgcn_readlane(x, l) emu_readlane((x), (l), __LINE__)
static inline int emu_readfirstlane(int x, int site) {
    const uint64_t *v = emu::wave_exchange((uint64_t)(uint32_t)x | (1ull << 40), site, "readfirstlane");
    for (int l = 0; l < 64; ++l)
        if (v[l] >> 40)
            return (int)(uint32_t)v[l];
    return 0;
}
#define __builtin_amdgcn_readfirstlane(x) emu_readfirstlane((x), __LINE__)

static inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
static inline int __ffsll(long long v) { return __builtin_ffsll(v); }
static inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }
static inline float __int_as_float(int v) { return emu_from<float>((uint32_t)v); }
static inline int __float_as_int(float v) { return (int)(uint32_t)emu_bits(v); }
static inline unsigned __float_as_uint(float v) { return (uint32_t)emu_bits(v); }
static inline float __uint_as_float(unsigned v) { return emu_from<float>(v); }

// Device atomics are host atomics: with EMU_THREADS, workgroups of a kernel
// run on several host threads at once.
template <class T> static inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T> static inline T atomicOr(T *p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T> static inline T atomicMax(T *p, T v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
template <class T> static inline T atomicMin(T *p, T v) {
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}

static inline int min(int a, int b) { return a < b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }
static inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
static inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
static inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
static inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
static inline unsigned long min(unsigned long a, unsigned long b) { return a < b ? a : b; }
static inline unsigned long max(unsigned long a, unsigned long b) { return a > b ? a : b; }
