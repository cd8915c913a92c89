// hip_runtime.h (CPU emulation) - the subset of HIP that liblensfdtd uses, for a CPU
// build of the whole library (host code and kernels) under AddressSanitizer.
//
// TEST INFRASTRUCTURE ONLY: tests/emu/build.py compiles the library's sources against
// this header with g++ so that every out-of-bounds load or store of a kernel becomes an
// ASan report.  Device memory is host memory of the exact requested size; streams run
// their work at once, in call order; a kernel launch runs every workgroup, one after the
// other, with one fiber per work-item, and the fibers of a workgroup meet at every
// cross-lane operation (DPP shifts, readlane, shuffles) and barrier, so the kernels
// compute what they compute on the device.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

// ---------------------------------------------------------------------------- types
struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float2 { float x, y; };
struct double2 { double x, y; };
inline float2 make_float2(float a, float b) { return float2{a, b}; }
inline double2 make_double2(double a, double b) { return double2{a, b}; }

using std::max;
using std::min;
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

inline float __int_as_float(int v) { float f; std::memcpy(&f, &v, 4); return f; }
inline int __float_as_int(float f) { int v; std::memcpy(&v, &f, 4); return v; }

// ---------------------------------------------------------------------------- runtime API
enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2,
                  hipErrorLaunchFailure = 719 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
struct lfemu_stream;
struct lfemu_event;
typedef lfemu_stream *hipStream_t;
typedef lfemu_event *hipEvent_t;
constexpr unsigned hipEventDefault = 0, hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;
struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t totalGlobalMem;
};

hipError_t lfemu_malloc(void **p, size_t n);
hipError_t hipFree(void *p);
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return lfemu_malloc((void **)p, n); }
hipError_t lfemu_host_malloc(void **p, size_t n, unsigned flags);
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned flags = 0) {
    return lfemu_host_malloc((void **)p, n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemset(void *p, int v, size_t n);
hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t s = nullptr);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t lfemu_symbol_address(void **p, const void *sym);
#define hipGetSymbolAddress(p, sym) lfemu_symbol_address((p), (const void *)(sym))

// ---------------------------------------------------------------------------- work-items
namespace lfemu {
struct Idx { unsigned x, y, z; };
struct Item {
    Idx tid;
};
extern thread_local Idx g_block, g_bdim, g_gdim;
Item &cur();
// cross-lane primitives: every active lane of the wave posts a 32-bit value, then reads
// the value that `src(lane)` posted (or `fallback` where that lane is inactive or out
// of range)
int xlane_shift(int old, int v, int delta, bool zero_fill);   // lane + delta
int xlane_read(int v, int lane);
int xlane_first(int v);
int xlane_perm(int v, int mode, int arg);                     // 0: down, 1: up, 2: xor
void barrier();
void launch(dim3 grid, dim3 block, size_t shmem, hipStream_t s, const std::function<void()> &f);
}  // namespace lfemu

#define threadIdx (lfemu::cur().tid)
#define blockIdx (lfemu::g_block)
#define blockDim (lfemu::g_bdim)
#define gridDim (lfemu::g_gdim)
#define warpSize 64

inline void __syncthreads() { lfemu::barrier(); }
inline void lfemu_setprio(int) {}
#define __builtin_amdgcn_s_setprio(x) lfemu_setprio(x)
#define __builtin_amdgcn_readfirstlane(v) lfemu::xlane_first((int)(v))
#define __builtin_amdgcn_readlane(v, l) lfemu::xlane_read((int)(v), (int)(l))
// DPP: 0x130 wave_shl:1 (lane reads lane + 1), 0x138 wave_shr:1 (lane reads lane - 1);
// a lane without a source keeps `old`, or reads 0 with bound_ctrl
inline int lfemu_update_dpp(int old, int src, int ctrl, int, int, bool bound_ctrl) {
    if (ctrl == 0x130) return lfemu::xlane_shift(old, src, 1, bound_ctrl);
    if (ctrl == 0x138) return lfemu::xlane_shift(old, src, -1, bound_ctrl);
    std::fprintf(stderr, "lfemu: DPP control 0x%x is not emulated\n", ctrl);
    std::abort();
}
#define __builtin_amdgcn_update_dpp(o, s, c, r, b, bc) lfemu_update_dpp((o), (s), (c), (r), (b), (bc))

template <typename T> inline T lfemu_perm(T v, int mode, int arg) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit shuffles");
    if constexpr (sizeof(T) == 4) {
        int b; std::memcpy(&b, &v, 4);
        b = lfemu::xlane_perm(b, mode, arg);
        T r; std::memcpy(&r, &b, 4); return r;
    } else {
        int b[2]; std::memcpy(b, &v, 8);
        b[0] = lfemu::xlane_perm(b[0], mode, arg);
        b[1] = lfemu::xlane_perm(b[1], mode, arg);
        T r; std::memcpy(&r, b, 8); return r;
    }
}
template <typename T> inline T __shfl_down(T v, unsigned d, int = 64) { return lfemu_perm(v, 0, (int)d); }
template <typename T> inline T __shfl_up(T v, unsigned d, int = 64) { return lfemu_perm(v, 1, (int)d); }
template <typename T> inline T __shfl_xor(T v, int m, int = 64) { return lfemu_perm(v, 2, m); }

// one workgroup runs at a time, so an atomic is a plain read-modify-write
template <typename T> inline T atomicAdd(T *p, T v) { T o = *p; *p = o + v; return o; }
