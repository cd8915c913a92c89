// lensfdtd_dev.h - device-side types and helpers shared by the translation units of
// liblensfdtd.so (the step kernels in lensfdtd.hip and the 3D blocked pass in yee_tb3.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/lensfdtd.h"

namespace lfd {

// sets the message lf_last_error() returns (thread-local), for the translation units
// whose entry points report their own errors
void set_error(const char *msg);

// steady-state reduction (reduce.hip): partials needed for n elements, and the launch of
// sum |A/da - B/db|^2, sum |A/da|^2 into out[0..1] (device) on stream st
int nd_partial_count(long long n);
int norm_diff_launch(const void *a, const void *b, long long n, int acc, double da, double db,
                     double2 *part, double *out, hipStream_t st);

struct Strip {            // psi slab of one stretch axis: v < lo -> v, v >= hi -> v-hi+lo
    float *psi;
    int lo, hi, len;      // len = lo + (n - hi); psi == nullptr -> no slab
};

struct Prof {             // 1-D CPML profile on one node family of one axis
    const float *b, *c, *k, *km1;
};

__device__ __forceinline__ int strip_pos(const Strip &s, int v) {
    if (s.psi == nullptr) return -1;
    if (v < s.lo) return v;
    if (v >= s.hi) return v - s.hi + s.lo;
    return -1;
}

struct Grid {
    int nx, ny, nz, P;    // logical dims, padded fast axis
    __device__ __forceinline__ long long at(int i, int j, int k) const {
        return ((long long)i * ny + j) * P + k;
    }
};

struct Excite {           // source pair and stitch drive, applied inside the E kernel
    int src_comp, src_ix, src_n;          // plane index = (j*nk + k) or j
    const float *amp, *ph1, *ph2;
    const double *a1c, *a1s, *a2c, *a2s;  // amp*cos/sin(phase) (FP64 phase mode)
    int drv_n, drv_math;                  // overlap cells, LF_DRIVE_*
    const float *wtap;
    const float2 *drv[3];                 // per E component (or nullptr)
    int drv_stride[3];                    // plane size of each driven component
};

struct StepArgs {
    float e32, wt32;
    double ssin, scos, env;
    float dr32, di32;
    double dr, di;
};

// ----------------------------------------------------------------------------- excite helpers
__device__ __forceinline__ float src_term_ref32(float e32, float wt32, float ph, float amp) {
    return (e32 * sinf(wt32 + ph)) * amp;          // vcsel.py:155 (float32)
}

__device__ __forceinline__ float apply_source(float v, int comp, int i, int plane, const Excite &x,
                                              const StepArgs &s, int fp64) {
    if (x.amp == nullptr || comp != x.src_comp) return v;
    if (i == x.src_ix) {
        if (fp64) v = v + (float)(s.ssin * x.a1c[plane] + s.scos * x.a1s[plane]);
        else v = v + src_term_ref32(s.e32, s.wt32, x.ph1[plane], x.amp[plane]);
    } else if (i == x.src_ix + 1) {
        if (fp64) v = v - (float)(s.ssin * x.a2c[plane] + s.scos * x.a2s[plane]);
        else v = v - src_term_ref32(s.e32, s.wt32, x.ph2[plane], x.amp[plane]);
    }
    return v;
}

// The stitch drive blend of one cell: D = the handed-over phasor of the cell, w = the
// taper weight of its plane.
__device__ __forceinline__ float drive_mix(float v, float2 D, float w, int drv_math,
                                           const StepArgs &s) {
    // Re(D * e^{iwt}): NumPy forms complex products as fma(ar, br, -(ai*bi)) on FMA hosts
    // (its AVX2/AVX-512 loops); the same single-rounding form is used here explicitly.
    if (drv_math == LF_DRIVE_C128) {         // fdtd2d_core.py:1426-1427 (float64 blend)
        const double re = fma((double)D.x, s.dr, -((double)D.y * s.di));
        const double din = s.env * re;
        const float keep = (1.0f - w) * v;
        return (float)((double)w * din + (double)keep);
    }
    const float re = fmaf(D.x, s.dr32, -(D.y * s.di32));   // fdtd3d_core.py:1627-1630 (float32)
    return w * (s.e32 * re) + (1.0f - w) * v;
}

__device__ __forceinline__ float apply_drive(float v, int comp, int i, int plane, const Excite &x,
                                             const StepArgs &s) {
    if (i >= x.drv_n || x.drv[comp] == nullptr) return v;
    const float2 D = x.drv[comp][(long long)i * x.drv_stride[comp] + plane];
    return drive_mix(v, D, x.wtap[i], x.drv_math, s);
}

// Whole-wavefront shifts by one lane (DPP wave_shl:1 / wave_shr:1, one VALU op each, no
// LDS round trip).  Lane 63 (shl) and lane 0 (shr) have no source lane and keep their
// own value, as __shfl_down / __shfl_up return there.
__device__ __forceinline__ float lane_next(float v) {          // value of lane + 1
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                      0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_prev(float v) {          // value of lane - 1
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                      0x138, 0xf, 0xf, false));
}
// The same shifts with the sourceless edge lane (63, resp. 0) set to 0 (bound_ctrl): the
// DPP op writes a fresh register, with no copy of the source into it first.  For
// wavefronts whose edge lanes are halo lanes, as in the temporally blocked passes.
__device__ __forceinline__ float lane_next0(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_prev0(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}

// Sum of two doubles over a workgroup of 256 work-items (4 wavefronts): butterfly
// shuffles within each wavefront, then the 4 wavefront sums in a fixed order.  Every
// work-item of the workgroup must call it; the result is valid in work-item 0.  The
// order of the additions depends only on the work-item layout, so sums are reproducible.
__device__ __forceinline__ void block_sum2_256(double &a, double &b) {
    __shared__ double part[2][4];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { part[0][w] = a; part[1][w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = ((part[0][0] + part[0][1]) + part[0][2]) + part[0][3];
        b = ((part[1][0] + part[1][1]) + part[1][2]) + part[1][3];
    }
    __syncthreads();
}

// psi slot of a slab compressed along `axis`, 32-bit element offset (grids < 2^31)
__device__ __forceinline__ unsigned psi_off32(const Strip &s, int axis, int pos, int i, int j, int k,
                                              const Grid &g) {
    if (axis == 0) return ((unsigned)pos * g.ny + (unsigned)j) * g.P + (unsigned)k;
    if (axis == 1) return ((unsigned)i * s.len + (unsigned)pos) * g.P + (unsigned)k;
    return ((unsigned)i * g.ny + (unsigned)j) * s.len + (unsigned)pos;
}

// profile triple of one node of one axis
struct PV { float b, c, m; };
__device__ __forceinline__ PV prof_at(const Prof &p, int n) { return PV{p.b[n], p.c[n], p.km1[n]}; }
// psi <- b psi + c d ; returns psi + (1/kappa - 1) d   (boundaries3d.py:240-251 order)
__device__ __forceinline__ float psi_step(float &psi, const PV &q, float d) {
    float p = psi * q.b;
    p = p + q.c * d;
    psi = p;
    return p + q.m * d;
}

}  // namespace lfd
