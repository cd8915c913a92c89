// reduce.hip - device reductions that replace host read-backs of whole phasor fields.
//
// The steady-state test of a CW run compares two phasor windows A and B, each divided by
// its sample count: change = |A/da - B/db| / |A/da| (fdtd2d_core.py:859-873,
// fdtd3d_core.py:1161-1175, the serving loop's one-period criterion).  Reading both
// windows back costs two full complex fields over PCIe per check; here the two squared
// norms are summed on the device in FP64 and only two doubles leave it.
//
// Two kernels, deterministic for a given n: nd_partials gives each workgroup a fixed chunk
// of kNdChunk elements (16 per work-item, coalesced), reduces it with wavefront shuffles
// (block_sum2_256) and writes one partial pair; nd_final reduces the partials in a fixed
// order and writes (sum |A/da - B/db|^2, sum |A/da|^2) to device memory.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "lensfdtd_dev.h"

using namespace lfd;

namespace {

constexpr int kNdItems = 16;
constexpr long long kNdChunk = 256LL * kNdItems;

template <typename C>
__global__ void __launch_bounds__(256) nd_partials(const C *a, const C *b, long long n, double da,
                                                   double db, double2 *part) {
    double num = 0.0, den = 0.0;
    const long long base = (long long)blockIdx.x * kNdChunk + threadIdx.x;
#pragma unroll 4
    for (int k = 0; k < kNdItems; ++k) {
        const long long i = base + (long long)k * 256;
        if (i < n) {
            const double ar = (double)a[i].x / da, ai = (double)a[i].y / da;
            const double dr = ar - (double)b[i].x / db, di = ai - (double)b[i].y / db;
            num += dr * dr + di * di;
            den += ar * ar + ai * ai;
        }
    }
    block_sum2_256(num, den);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(num, den);
}

__global__ void __launch_bounds__(256) nd_final(const double2 *part, int np, double *out) {
    double num = 0.0, den = 0.0;
    for (int p = threadIdx.x; p < np; p += 256) {
        num += part[p].x;
        den += part[p].y;
    }
    block_sum2_256(num, den);
    if (threadIdx.x == 0) { out[0] = num; out[1] = den; }
}

// library-owned partials, one buffer per (device, stream), grown on demand: calls on one
// stream are ordered, so they may share a buffer; calls on distinct streams do not
std::mutex g_nd_mu;
std::map<std::pair<int, hipStream_t>, std::pair<void *, size_t>> g_nd_buf;

int nd_fail(const char *m) { set_error(m); return -1; }

// Partials of the allocation-free variant: a file-scope device array (1 MiB), one copy on
// each GPU the library runs on, held until the process ends.  It bounds n to
// kNdScratchPairs * kNdChunk = 2^28 elements.  nd_partials writes every pair that
// nd_final reads, so the array needs no clearing between calls.
constexpr int kNdScratchPairs = 65536;
constexpr long long kNdScratchMaxN = (long long)kNdScratchPairs * kNdChunk;
static __device__ __align__(256) double2 scratch[kNdScratchPairs];

// Device address of `scratch` on each device, looked up once per device with that device
// current (each GPU holds its own copy at its own address); host code passes these
// addresses, never the array's name, which is a host-side stand-in.
constexpr int kNdMaxDevices = 64;
std::mutex g_sc_mu;
double2 *g_sc_ptr[kNdMaxDevices] = {};

double2 *scratch_on(int dev) {
    std::lock_guard<std::mutex> lk(g_sc_mu);
    if (g_sc_ptr[dev] == nullptr) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
        g_sc_ptr[dev] = (double2 *)p;
    }
    return g_sc_ptr[dev];
}

}  // namespace

namespace lfd {

int nd_partial_count(long long n) { return (int)((n + kNdChunk - 1) / kNdChunk); }

// Enqueue the two kernels on `st`: partials in `part` (nd_partial_count(n) pairs).
int norm_diff_launch(const void *a, const void *b, long long n, int acc, double da, double db,
                     double2 *part, double *out, hipStream_t st) {
    if (!(da != 0.0) || !(db != 0.0)) return nd_fail("norm_diff: zero divisor");
    if (n <= 0) return nd_fail("norm_diff: empty input");
    const int np = nd_partial_count(n);
    if (acc == LF_ACC_C128)
        nd_partials<double2><<<np, 256, 0, st>>>((const double2 *)a, (const double2 *)b, n, da, db, part);
    else
        nd_partials<float2><<<np, 256, 0, st>>>((const float2 *)a, (const float2 *)b, n, da, db, part);
    nd_final<<<1, 256, 0, st>>>(part, np, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return nd_fail(hipGetErrorString(e));
    return 0;
}

}  // namespace lfd

extern "C" {

int lf_norm_diff(const void *a, const void *b, long long n, int acc, double da, double db,
                 double *out, void *stream) {
    if (!a || !b || !out || n <= 0) return nd_fail("lf_norm_diff: bad arguments");
    if (acc != LF_ACC_C128 && acc != LF_ACC_C64 && acc != LF_ACC_C64P)
        return nd_fail("lf_norm_diff: acc must be LF_ACC_C128, LF_ACC_C64 or LF_ACC_C64P");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nd_fail("lf_norm_diff: no device");
    const hipStream_t st = (hipStream_t)stream;
    const size_t need = (size_t)nd_partial_count(n) * sizeof(double2);
    double2 *part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_nd_mu);
        auto &slot = g_nd_buf[{dev, st}];
        if (slot.second < need) {
            // work queued earlier on this stream may still read the old buffer
            if (slot.first) {
                if (hipStreamSynchronize(st) != hipSuccess || hipFree(slot.first) != hipSuccess)
                    return nd_fail("lf_norm_diff: releasing the partials failed");
                slot = {nullptr, 0};
            }
            if (hipMalloc(&slot.first, need) != hipSuccess) return nd_fail("lf_norm_diff: out of memory");
            slot.second = need;
        }
        part = (double2 *)slot.first;
    }
    return norm_diff_launch(a, b, n, acc, da, db, part, out, st);
}

// lf_norm_diff without allocation: the partials go to the device's `scratch` array, so
// the call makes no hipMalloc / hipFree.  Calls that overlap on one GPU (other streams,
// other host threads) share that array and race.
int lf_norm_diff_noalloc(const void *a, const void *b, long long n, int acc, double da, double db,
                         double *out, void *stream) {
    if (!a || !b || !out || n <= 0) return nd_fail("lf_norm_diff_noalloc: bad arguments");
    if (acc != LF_ACC_C128 && acc != LF_ACC_C64 && acc != LF_ACC_C64P)
        return nd_fail("lf_norm_diff_noalloc: acc must be LF_ACC_C128, LF_ACC_C64 or LF_ACC_C64P");
    if (n > kNdScratchMaxN)
        return nd_fail("lf_norm_diff_noalloc: n exceeds the static scratch (at most 2^28 = "
                       "268435456 elements); use lf_norm_diff");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nd_fail("lf_norm_diff_noalloc: no device");
    if (dev < 0 || dev >= kNdMaxDevices) return nd_fail("lf_norm_diff_noalloc: device index out of range");
    double2 *part = scratch_on(dev);
    if (part == nullptr) return nd_fail("lf_norm_diff_noalloc: hipGetSymbolAddress(scratch) failed");
    return norm_diff_launch(a, b, n, acc, da, db, part, out, (hipStream_t)stream);
}

}  // extern "C"
