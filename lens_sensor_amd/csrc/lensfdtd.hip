// lensfdtd.hip - Yee FDTD step engine for MI355X (gfx950): kernels + C ABI.
//
// Layout in HBM: every field component and every per-cell coefficient of a context
// lives in the same padded grid (Nx, Ny, P) (3D) or (Nx, 1, P) (2D, fast axis y),
// P = Nz (3D) / Ny (2D) rounded up to a multiple of 64, so that one 64-lane wavefront
// reads one contiguous 256-byte row segment and all components share one index
// (i*Ny + j)*P + k.  Cells outside a component's extent stay zero.  CPML psi
// auxiliaries exist only in the slabs of their stretch axis where c != 0 (elsewhere
// psi stays exactly 0 in the reference too), compressed along that axis.
//
// Arithmetic: the file is built with -ffp-contract=off and follows the reference's
// float32 operation order per cell (planar_beads/boundaries.py:186-268,
// planar_3d/boundaries3d.py:240-362), so the float32 fields agree with the NumPy
// reference to the last bit except for the float32 sin of the source phase.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <type_traits>
#include <utility>
#include <string>
#include <vector>

#include "../../include/lensfdtd.h"
#include "lensfdtd_dev.h"
#include "yee_tb3.h"
#include "tb2d.h"
#include "tb2d_plan.h"

using namespace lfd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

}  // namespace

namespace lfd {
void set_error(const char *msg) { g_err = msg; }
}  // namespace lfd

namespace {

#define LF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(-2, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

constexpr int kMaxMon = 32;
constexpr int kBoxBlocks = 1024;   // per-block FP64 partials (deterministic order)

// ----------------------------------------------------------------------------- device structs
// ============================================================================ 2D TE kernels
struct TEArgs {
    Grid g;
    float *Ez, *Hx, *Hy;
    const float *ceE, *ceH;
    float ch;
    Prof xe, xh, ye, yh;
    Strip pHxy, pHyx, pEzx, pEzy;
    int pec_xlo, pec_xhi, pec_ylo, pec_yhi;
};

__global__ void __launch_bounds__(256) te_h_kernel(TEArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int Nx = a.g.nx, Ny = a.g.nz;
    if (i >= Nx || j >= Ny) return;
    const long long c = (long long)i * a.g.P + j;
    if (j < Ny - 1) {                               // Hx (Nx, Ny-1): boundaries.py:189-191
        const float d = a.Ez[c + 1] - a.Ez[c];
        const int s = strip_pos(a.pHxy, j);
        float t;
        if (s >= 0) {
            float *pp = a.pHxy.psi + (long long)i * a.pHxy.len + s;
            const float p = a.yh.b[j] * (*pp) + a.yh.c[j] * d;
            *pp = p;
            t = d / a.yh.k[j] + p;
        } else {
            t = d;
        }
        a.Hx[c] = a.Hx[c] - a.ch * t;
    }
    if (i < Nx - 1) {                               // Hy (Nx-1, Ny): :192-194
        const float d = a.Ez[c + a.g.P] - a.Ez[c];
        const int s = strip_pos(a.pHyx, i);
        float t;
        if (s >= 0) {
            float *pp = a.pHyx.psi + (long long)s * a.g.P + j;
            const float p = a.xh.b[i] * (*pp) + a.xh.c[i] * d;
            *pp = p;
            t = d / a.xh.k[i] + p;
        } else {
            t = d;
        }
        a.Hy[c] = a.Hy[c] + a.ch * t;
    }
}

__global__ void __launch_bounds__(256) te_e_kernel(TEArgs a, Excite x, StepArgs st, int fp64) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int Nx = a.g.nx, Ny = a.g.nz;
    if (i >= Nx || j >= Ny) return;
    const long long c = (long long)i * a.g.P + j;
    float v = a.Ez[c];
    if (i >= 1 && i <= Nx - 2 && j >= 1 && j <= Ny - 2) {   // boundaries.py:199-205
        const float dHy = a.Hy[c] - a.Hy[c - a.g.P];
        const float dHx = a.Hx[c] - a.Hx[c - 1];
        const int sx = strip_pos(a.pEzx, i);
        const int sy = strip_pos(a.pEzy, j);
        float t;
        if (sx >= 0) {
            float *pp = a.pEzx.psi + (long long)sx * a.g.P + j;
            const float p = a.xe.b[i] * (*pp) + a.xe.c[i] * dHy;
            *pp = p;
            t = dHy / a.xe.k[i] + p;
        } else {
            t = dHy;
        }
        if (sy >= 0) {
            float *pp = a.pEzy.psi + (long long)i * a.pEzy.len + sy;
            const float p = a.ye.b[j] * (*pp) + a.ye.c[j] * dHx;
            *pp = p;
            t = (t - dHx / a.ye.k[j]) - p;
        } else {
            t = t - dHx;
        }
        v = a.ceE[c] * v + a.ceH[c] * t;
    }
    v = apply_source(v, LF_EZ, i, j, x, st, fp64);          // vcsel.py:151-156
    v = apply_drive(v, LF_EZ, i, j, x, st);                  // fdtd2d_core.py:1462-1463
    if ((a.pec_xlo && i == 0) || (a.pec_xhi && i == Nx - 1) || (a.pec_ylo && j == 0) ||
        (a.pec_yhi && j == Ny - 1))
        v = 0.0f;                                           // boundaries.py:207-215
    a.Ez[c] = v;
}

// ============================================================================ 2D TM kernels
struct TMArgs {
    Grid g;
    float *Ex, *Ey, *Hz;
    const float *cax, *cay, *dax, *day;
    float ch;
    Prof xe, xh, ye, yh;
    Strip pHzx, pHzy, pExy, pEyx;
    int pec_xlo, pec_xhi, pec_ylo, pec_yhi;
};

__global__ void __launch_bounds__(256) tm_h_kernel(TMArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int Nx = a.g.nx, Ny = a.g.nz;
    if (i >= Nx - 1 || j >= Ny - 1) return;
    const long long c = (long long)i * a.g.P + j;
    const float dEyx = a.Ey[c + a.g.P] - a.Ey[c];           // boundaries.py:243-248
    const float dExy = a.Ex[c + 1] - a.Ex[c];
    const int sx = strip_pos(a.pHzx, i);
    const int sy = strip_pos(a.pHzy, j);
    float tx, ty;
    if (sx >= 0) {
        float *pp = a.pHzx.psi + (long long)sx * a.g.P + j;
        const float p = a.xh.b[i] * (*pp) + a.xh.c[i] * dEyx;
        *pp = p;
        tx = dEyx / a.xh.k[i] + p;
    } else {
        tx = dEyx;
    }
    if (sy >= 0) {
        float *pp = a.pHzy.psi + (long long)i * a.pHzy.len + sy;
        const float p = a.yh.b[j] * (*pp) + a.yh.c[j] * dExy;
        *pp = p;
        ty = dExy / a.yh.k[j] + p;
    } else {
        ty = dExy;
    }
    a.Hz[c] = a.Hz[c] - a.ch * (tx - ty);
}

__global__ void __launch_bounds__(256) tm_e_kernel(TMArgs a, Excite x, StepArgs st, int fp64) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int Nx = a.g.nx, Ny = a.g.nz;
    if (i >= Nx || j >= Ny) return;
    const long long c = (long long)i * a.g.P + j;
    if (i < Nx - 1) {                                       // Ex (Nx-1, Ny)
        float v = a.Ex[c];
        if (a.dax) v = v * a.dax[c];                        // fdtd2d_core.py:796
        if (j >= 1 && j <= Ny - 2) {                        // boundaries.py:253-255
            const float d = a.Hz[c] - a.Hz[c - 1];
            const int s = strip_pos(a.pExy, j);
            float t;
            if (s >= 0) {
                float *pp = a.pExy.psi + (long long)i * a.pExy.len + s;
                const float p = a.ye.b[j] * (*pp) + a.ye.c[j] * d;
                *pp = p;
                t = d / a.ye.k[j] + p;
            } else {
                t = d;
            }
            v = v + a.cax[c] * t;
        }
        if ((a.pec_ylo && j == 0) || (a.pec_yhi && j == Ny - 1)) v = 0.0f;
        a.Ex[c] = v;
    }
    if (j < Ny - 1) {                                       // Ey (Nx, Ny-1)
        float v = a.Ey[c];
        if (a.day) v = v * a.day[c];
        if (i >= 1 && i <= Nx - 2) {                        // boundaries.py:256-258
            const float d = a.Hz[c] - a.Hz[c - a.g.P];
            const int s = strip_pos(a.pEyx, i);
            float t;
            if (s >= 0) {
                float *pp = a.pEyx.psi + (long long)s * a.g.P + j;
                const float p = a.xe.b[i] * (*pp) + a.xe.c[i] * d;
                *pp = p;
                t = d / a.xe.k[i] + p;
            } else {
                t = d;
            }
            v = v + (-a.cay[c]) * t;
        }
        v = apply_source(v, LF_EY, i, j, x, st, fp64);
        v = apply_drive(v, LF_EY, i, j, x, st);
        if ((a.pec_xlo && i == 0) || (a.pec_xhi && i == Nx - 1)) v = 0.0f;
        a.Ey[c] = v;
    }
}


// ---------------------------------------------------------------------------- fused TM step
// One pass per step, barrier free.  Every wavefront is independent: its 64 lanes hold
// consecutive y cells j = jb - 1 + lane (lane 0 is the cell below the wave's 63 owned
// cells) and it marches along an x chunk of `lx` rows.  For row i, Hz^{n+1/2}(i) is formed
// from E^n (Ex(i, j+1) comes from lane+1 by a wavefront shuffle, lane 63 loads it),
// then Ex^{n+1}(i) uses Hz(i, j-1) from lane-1 and Ey^{n+1}(i) uses Hz(i-1) kept in a
// register from the previous row.  Old and new fields (and the psi of Hz) are separate
// buffers (ping-pong): a value recomputed outside a lane's own cells always sees E^n and
// H^{n-1/2}.  Per cell the arithmetic is exactly that of tm_h_kernel + tm_e_kernel.
struct TMFArgs {
    Grid g;
    const float *Ex0, *Ey0, *Hz0;
    float *Ex1, *Ey1, *Hz1;
    const float *cax, *cay, *dax, *day;
    float ch;
    Prof xe, xh, ye, yh;
    Strip pHzx0, pHzy0, pHzx1, pHzy1, pExy, pEyx;
    int pec_xlo, pec_xhi, pec_ylo, pec_yhi;
    int lx, nwy;                      // x rows per wave, waves along y
};

__device__ __forceinline__ float tm_hz_new(const TMFArgs &a, int i, int j, float ey_n, float ey_c,
                                           float ex_n, float ex_c, bool store) {
    const long long c = (long long)i * a.g.P + j;
    const float dEyx = ey_n - ey_c;
    const float dExy = ex_n - ex_c;
    float tx, ty;
    const int sx = strip_pos(a.pHzx0, i);
    if (sx >= 0) {
        const long long q = (long long)sx * a.g.P + j;
        const float p = a.xh.b[i] * a.pHzx0.psi[q] + a.xh.c[i] * dEyx;
        if (store) a.pHzx1.psi[q] = p;
        tx = dEyx / a.xh.k[i] + p;
    } else {
        tx = dEyx;
    }
    const int sy = strip_pos(a.pHzy0, j);
    if (sy >= 0) {
        const long long q = (long long)i * a.pHzy0.len + sy;
        const float p = a.yh.b[j] * a.pHzy0.psi[q] + a.yh.c[j] * dExy;
        if (store) a.pHzy1.psi[q] = p;
        ty = dExy / a.yh.k[j] + p;
    } else {
        ty = dExy;
    }
    return a.Hz0[c] - a.ch * (tx - ty);
}

__global__ void __launch_bounds__(256) tm_fused_kernel(TMFArgs a, Excite x, StepArgs st, int fp64) {
    const int lane = threadIdx.x & 63;
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int wy = wv % a.nwy, wx = wv / a.nwy;
    const int Nx = a.g.nx, Ny = a.g.nz, P = a.g.P;
    const int j = wy * 63 + lane - 1;
    const bool jin = j >= 0 && j < Ny;
    const bool mine = lane >= 1;                 // lanes 1..63 own their cell
    const int i0 = wx * a.lx;
    if (i0 >= Nx) return;                        // whole wave leaves together
    const int i1 = min(i0 + a.lx, Nx);
    const int ia = i0 > 0 ? i0 - 1 : 0;
    float hz_prev = 0.0f;
    float eyc = (jin && j < Ny - 1) ? a.Ey0[(long long)ia * P + j] : 0.0f;
    for (int i = ia; i < i1; ++i) {
        const bool own = i >= i0 && mine && jin;
        const long long c = (long long)i * P + (jin ? j : 0);
        const float exc = (jin && i < Nx - 1) ? a.Ex0[c] : 0.0f;
        const float eyn = (jin && j < Ny - 1 && i < Nx - 1) ? a.Ey0[c + P] : 0.0f;
        float exn = __shfl_down(exc, 1);
        if (lane == 63) exn = (jin && j + 1 < Ny && i < Nx - 1) ? a.Ex0[c + 1] : 0.0f;
        float hz = 0.0f;
        if (jin && i < Nx - 1 && j < Ny - 1) {
            hz = tm_hz_new(a, i, j, eyn, eyc, exn, exc, own);
            if (own) a.Hz1[c] = hz;
        }
        const float hzm = __shfl_up(hz, 1);      // Hz(i, j-1)
        if (own) {
            if (i < Nx - 1) {                    // Ex (Nx-1, Ny)
                float v = exc;
                if (a.dax) v = v * a.dax[c];
                if (j >= 1 && j <= Ny - 2) {
                    const float d = hz - hzm;
                    const int s = strip_pos(a.pExy, j);
                    float tt;
                    if (s >= 0) {
                        float *pp = a.pExy.psi + (long long)i * a.pExy.len + s;
                        const float p = a.ye.b[j] * (*pp) + a.ye.c[j] * d;
                        *pp = p;
                        tt = d / a.ye.k[j] + p;
                    } else {
                        tt = d;
                    }
                    v = v + a.cax[c] * tt;
                }
                if ((a.pec_ylo && j == 0) || (a.pec_yhi && j == Ny - 1)) v = 0.0f;
                a.Ex1[c] = v;
            }
            if (j < Ny - 1) {                    // Ey (Nx, Ny-1)
                float v = eyc;
                if (a.day) v = v * a.day[c];
                if (i >= 1 && i <= Nx - 2) {
                    const float d = hz - hz_prev;
                    const int s = strip_pos(a.pEyx, i);
                    float tt;
                    if (s >= 0) {
                        float *pp = a.pEyx.psi + (long long)s * P + j;
                        const float p = a.xe.b[i] * (*pp) + a.xe.c[i] * d;
                        *pp = p;
                        tt = d / a.xe.k[i] + p;
                    } else {
                        tt = d;
                    }
                    v = v + (-a.cay[c]) * tt;
                }
                v = apply_source(v, LF_EY, i, j, x, st, fp64);
                v = apply_drive(v, LF_EY, i, j, x, st);
                if ((a.pec_xlo && i == 0) || (a.pec_xhi && i == Nx - 1)) v = 0.0f;
                a.Ey1[c] = v;
            }
        }
        hz_prev = hz;
        eyc = eyn;
    }
}

// ---------------------------------------------------------------------------- batched passes
// Many independent 2D domains (sweep instances, served requests) advance T steps in ONE
// launch per wavefront class: a device table holds each instance's pass arguments, a
// prefix table maps the launch's wavefront index to (instance, wavefront within it), and
// the per-step scalars come from a device step table shared by the instances.
__host__ __device__ __forceinline__ StepArgs step_args_of(const lf_step &t) {
    return StepArgs{t.e32, t.wt32, t.ssin, t.scos, t.env, t.dr32, t.di32, t.dr, t.di};
}

struct TBRecTM { TBTMArgs a; Excite x; const lf_step *steps; };   // steps: this pass's first
struct TBRecTE { TBTEArgs a; Excite x; const lf_step *steps; };

__device__ __forceinline__ int batch_instance(const int *start, int ninst, int wv) {
    int lo = 0, hi = ninst - 1;                    // largest b with start[b] <= wv
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= wv) lo = mid; else hi = mid - 1;
    }
    // wave-uniform: the record and step-table reads of the instance become scalar loads
    return __builtin_amdgcn_readfirstlane(lo);
}

template <int T, int MODE>
__global__ void __launch_bounds__(256) tm_tb_batch_kernel(const TBRecTM *__restrict__ rec,
                                                          const int *__restrict__ start, int ninst,
                                                          int fp64) {
    if (MODE != 0) __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wv >= start[ninst]) return;
    const int b = batch_instance(start, ninst, wv);
    // the records are read-only for the whole launch: reads through the constant address
    // space are scalar loads the compiler may repeat instead of holding them in registers
    const TBRecTM __attribute__((address_space(4))) *rc =
        (const TBRecTM __attribute__((address_space(4))) *)rec;
    const lf_step __attribute__((address_space(4))) *steps =
        (const lf_step __attribute__((address_space(4))) *)rc[b].steps;
    StepArgs st[T];
#pragma unroll
    for (int u = 0; u < T; ++u) st[u] = step_args_of(*(const lf_step *)&steps[u]);
    tm_tb_body<T, MODE>(*(const TBTMArgs *)&rc[b].a, *(const Excite *)&rc[b].x, st, wv - start[b], fp64);
}

template <int T, int MODE>
__global__ void __launch_bounds__(256) te_tb_batch_kernel(const TBRecTE *__restrict__ rec,
                                                          const int *__restrict__ start, int ninst,
                                                          int fp64) {
    if (MODE != 0) __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wv >= start[ninst]) return;
    const int b = batch_instance(start, ninst, wv);
    // the records are read-only for the whole launch: reads through the constant address
    // space are scalar loads the compiler may repeat instead of holding them in registers
    const TBRecTE __attribute__((address_space(4))) *rc =
        (const TBRecTE __attribute__((address_space(4))) *)rec;
    const lf_step __attribute__((address_space(4))) *steps =
        (const lf_step __attribute__((address_space(4))) *)rc[b].steps;
    StepArgs st[T];
#pragma unroll
    for (int u = 0; u < T; ++u) st[u] = step_args_of(*(const lf_step *)&steps[u]);
    te_tb_body<T, MODE>(*(const TBTEArgs *)&rc[b].a, *(const Excite *)&rc[b].x, st, wv - start[b], fp64);
}

// ============================================================================ 3D kernels
struct YeeArgs {
    Grid g;
    float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    const float *cex, *cey, *cez, *dax, *day, *daz;
    float ch;
    Prof pe[3], ph[3];        // per axis: full-node (E) and half-node (H) profiles
    // psi slabs: H: Hx_z Hx_y Hy_x Hy_z Hz_y Hz_x ; E: Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y
    Strip sH[6], sE[6];
    int pec[6];               // xlo xhi ylo yhi zlo zhi
};

// psi element of a slab compressed along `axis`
__device__ __forceinline__ float *psi_at(const Strip &s, int axis, int pos, int i, int j, int k,
                                         const Grid &g) {
    if (axis == 0) return s.psi + ((long long)pos * g.ny + j) * g.P + k;
    if (axis == 1) return s.psi + ((long long)i * s.len + pos) * g.P + k;
    return s.psi + ((long long)i * g.ny + j) * s.len + pos;
}

// F += coef * add with add = psi_new + km1 * d  (boundaries3d.py:240-251)
__device__ __forceinline__ float psi_term(const Strip &s, int axis, int pos, const Prof &pr, int n,
                                          float d, int i, int j, int k, const Grid &g) {
    float *pp = psi_at(s, axis, pos, i, j, k, g);
    float p = (*pp) * pr.b[n];
    p = p + pr.c[n] * d;
    *pp = p;
    return p + pr.km1[n] * d;
}

__global__ void __launch_bounds__(256) yee_h_kernel(YeeArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    const Grid &g = a.g;
    if (k >= g.nz || j >= g.ny) return;
    const long long c = g.at(i, j, k);
    const long long sj = g.P, si = (long long)g.ny * g.P;
    const float ch = a.ch;
    if (j < g.ny - 1 && k < g.nz - 1) {                     // Hx: boundaries3d.py:316-320
        const float d1 = a.Ey[c + 1] - a.Ey[c];
        const float d2 = a.Ez[c + sj] - a.Ez[c];
        float h = a.Hx[c] + ch * (d1 - d2);
        int s = strip_pos(a.sH[0], k);
        if (s >= 0) h = h + ch * psi_term(a.sH[0], 2, s, a.ph[2], k, d1, i, j, k, g);
        s = strip_pos(a.sH[1], j);
        if (s >= 0) h = h + (-ch) * psi_term(a.sH[1], 1, s, a.ph[1], j, d2, i, j, k, g);
        a.Hx[c] = h;
    }
    if (i < g.nx - 1 && k < g.nz - 1) {                     // Hy: :321-325
        const float d1 = a.Ez[c + si] - a.Ez[c];
        const float d2 = a.Ex[c + 1] - a.Ex[c];
        float h = a.Hy[c] + ch * (d1 - d2);
        int s = strip_pos(a.sH[2], i);
        if (s >= 0) h = h + ch * psi_term(a.sH[2], 0, s, a.ph[0], i, d1, i, j, k, g);
        s = strip_pos(a.sH[3], k);
        if (s >= 0) h = h + (-ch) * psi_term(a.sH[3], 2, s, a.ph[2], k, d2, i, j, k, g);
        a.Hy[c] = h;
    }
    if (i < g.nx - 1 && j < g.ny - 1) {                     // Hz: :326-330
        const float d1 = a.Ex[c + sj] - a.Ex[c];
        const float d2 = a.Ey[c + si] - a.Ey[c];
        float h = a.Hz[c] + ch * (d1 - d2);
        int s = strip_pos(a.sH[4], j);
        if (s >= 0) h = h + ch * psi_term(a.sH[4], 1, s, a.ph[1], j, d1, i, j, k, g);
        s = strip_pos(a.sH[5], i);
        if (s >= 0) h = h + (-ch) * psi_term(a.sH[5], 0, s, a.ph[0], i, d2, i, j, k, g);
        a.Hz[c] = h;
    }
}

__global__ void __launch_bounds__(256) yee_e_kernel(YeeArgs a, Excite x, StepArgs st, int fp64,
                                                    int nk_src) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    const Grid &g = a.g;
    if (k >= g.nz || j >= g.ny) return;
    const long long c = g.at(i, j, k);
    const long long sj = g.P, si = (long long)g.ny * g.P;
    const int NX = g.nx, NY = g.ny, NZ = g.nz;
    if (i < NX - 1) {                                       // Ex (Nx-1, Ny, Nz)
        float v = a.Ex[c];
        if (a.dax) v = v * a.dax[c];                        // boundaries3d.py:333-336
        if (j >= 1 && j <= NY - 2 && k >= 1 && k <= NZ - 2) {   // :338-343
            const float d1 = a.Hz[c] - a.Hz[c - sj];
            const float d2 = a.Hy[c] - a.Hy[c - 1];
            const float ce = a.cex[c];
            v = v + ce * (d1 - d2);
            int s = strip_pos(a.sE[0], j);
            if (s >= 0) v = v + ce * psi_term(a.sE[0], 1, s, a.pe[1], j, d1, i, j, k, g);
            s = strip_pos(a.sE[1], k);
            if (s >= 0) v = v + (-ce) * psi_term(a.sE[1], 2, s, a.pe[2], k, d2, i, j, k, g);
        }
        if ((a.pec[2] && j == 0) || (a.pec[3] && j == NY - 1) || (a.pec[4] && k == 0) ||
            (a.pec[5] && k == NZ - 1))
            v = 0.0f;                                       // :356-362
        v = apply_drive(v, LF_EX, i, j * NZ + k, x, st);
        a.Ex[c] = v;
    }
    if (j < NY - 1) {                                       // Ey (Nx, Ny-1, Nz)
        float v = a.Ey[c];
        if (a.day) v = v * a.day[c];
        if (i >= 1 && i <= NX - 2 && k >= 1 && k <= NZ - 2) {   // :344-349
            const float d1 = a.Hx[c] - a.Hx[c - 1];
            const float d2 = a.Hz[c] - a.Hz[c - si];
            const float ce = a.cey[c];
            v = v + ce * (d1 - d2);
            int s = strip_pos(a.sE[2], k);
            if (s >= 0) v = v + ce * psi_term(a.sE[2], 2, s, a.pe[2], k, d1, i, j, k, g);
            s = strip_pos(a.sE[3], i);
            if (s >= 0) v = v + (-ce) * psi_term(a.sE[3], 0, s, a.pe[0], i, d2, i, j, k, g);
        }
        if ((a.pec[0] && i == 0) || (a.pec[1] && i == NX - 1) || (a.pec[4] && k == 0) ||
            (a.pec[5] && k == NZ - 1))
            v = 0.0f;
        const int plane = j * NZ + k;                       // Ey plane (Ny-1, Nz)
        v = apply_source(v, LF_EY, i, plane, x, st, fp64);  // fdtd3d_core.py:1127
        v = apply_drive(v, LF_EY, i, plane, x, st);
        a.Ey[c] = v;
    }
    if (k < NZ - 1) {                                       // Ez (Nx, Ny, Nz-1)
        float v = a.Ez[c];
        if (a.daz) v = v * a.daz[c];
        if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2) {   // :350-355
            const float d1 = a.Hy[c] - a.Hy[c - si];
            const float d2 = a.Hx[c] - a.Hx[c - sj];
            const float ce = a.cez[c];
            v = v + ce * (d1 - d2);
            int s = strip_pos(a.sE[4], i);
            if (s >= 0) v = v + ce * psi_term(a.sE[4], 0, s, a.pe[0], i, d1, i, j, k, g);
            s = strip_pos(a.sE[5], j);
            if (s >= 0) v = v + (-ce) * psi_term(a.sE[5], 1, s, a.pe[1], j, d2, i, j, k, g);
        }
        if ((a.pec[0] && i == 0) || (a.pec[1] && i == NX - 1) || (a.pec[2] && j == 0) ||
            (a.pec[3] && j == NY - 1))
            v = 0.0f;
        const int plane = j * (NZ - 1) + k;                 // Ez plane (Ny, Nz-1)
        v = apply_source(v, LF_EZ, i, plane, x, st, fp64);
        v = apply_drive(v, LF_EZ, i, plane, x, st);
        a.Ez[c] = v;
    }
    (void)nk_src;
}


// ---------------------------------------------------------------------------- single-pass 3D step
// One pass per step.  A wavefront holds 64 consecutive z cells k = wz*63 - 1 + lane
// (lane 0 is the cell below the 63 owned cells), owns R y rows j0..j0+R-1 (row slot r
// holds j = j0 - 1 + r; slot 0 is the row below) and marches along an x chunk [X0, X1)
// from X0 - 1.  For plane i it forms H^{n+1/2}(i) on slots 0..R from E^n(i), E^n(i+1) and
// H^{n-1/2}(i), then E^{n+1}(i) on the owned slots from H(i) (row below from the same
// pass, k-1 by a DPP lane shift) and the Hy/Hz of plane i-1 kept in registers.  Old and
// new fields and the H psi slabs are separate buffers (a recomputed halo value always
// reads step-n data); E psi is updated in place by the owner.  Per cell the arithmetic is
// that of yee_h_kernel + yee_e_kernel (boundaries3d.py:311-362, same operation order).
// MODE bits: 1 = lanes may touch a z slab or z edge, 2 = rows may touch a y slab or y
// edge, 4 = planes may touch an x slab or x edge.  A cleared bit drops those psi terms and
// edge tests; the host only gives such waves that provably need neither.
struct SPArgs {
    Grid g;
    const float *Ex0, *Ey0, *Ez0, *Hx0, *Hy0, *Hz0;
    float *Ex1, *Ey1, *Ez1, *Hx1, *Hy1, *Hz1;
    const float *cex, *cey, *cez, *dax, *day, *daz;
    float ch;
    Prof pe[3], ph[3];
    Strip sH0[6], sH1[6], sE[6];   // H: Hx_z Hx_y Hy_x Hy_z Hz_y Hz_x ; E: Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y
    int pec[6];
    int lx;
    int nbox;                      // wavefront boxes in (wz, wy, wx) units, wy fastest
    int bz0[8], bnz[8], by0[8], bny[8], bx0[8], bend[8];
};

// psi update of an H slab with separate old/new buffers (store only when owned)
__device__ __forceinline__ float psi_term_pp(const Strip &s0, const Strip &s1, int axis, int pos,
                                             const Prof &pr, int n, float d, int i, int j, int k,
                                             const Grid &g, bool store) {
    const float *po = psi_at(s0, axis, pos, i, j, k, g);
    float p = (*po) * pr.b[n];
    p = p + pr.c[n] * d;
    if (store) *psi_at(s1, axis, pos, i, j, k, g) = p;
    return p + pr.km1[n] * d;
}

// one x plane of a single-pass wavefront, as loaded from the step-n buffers: Ex and the
// old H of plane i, the lane-63 k+1 values, coefficients, psi of every slab term, and
// Ey / Ez of plane i+1 (the x derivatives of H(i))
template <int R>
struct SPPlane {
    float exA[R + 2], eyB[R + 1], ezB[R + 2], exK[R + 1], eyK[R + 1];
    float hxO[R + 1], hyO[R + 1], hzO[R + 1];
    float cx[R + 1], cy[R + 1], cz[R + 1], dx[R + 1], dy[R + 1], dz[R + 1];
    float pHxz[R + 1], pHxy[R + 1], pHyx[R + 1], pHyz[R + 1], pHzy[R + 1], pHzx[R + 1];
    float pExy[R + 1], pExz[R + 1], pEyz[R + 1], pEyx[R + 1], pEzx[R + 1], pEzy[R + 1];
    int xHy, xHz, xEy, xEz;
    PV xh, xe;
};

template <int R, int MODE>
__global__ void __launch_bounds__(256) yee_sp_kernel(SPArgs a, Excite x, StepArgs st, int fp64) {
    constexpr bool PZ = MODE & 1, PY = MODE & 2, PX = MODE & 4;
    const int lane = threadIdx.x & 63;
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wv >= a.bend[a.nbox - 1]) return;                 // surplus waves of the last block
    int g = 0;
    while (g < a.nbox - 1 && wv >= a.bend[g]) ++g;
    const int loc = wv - (g > 0 ? a.bend[g - 1] : 0);
    const int wy = a.by0[g] + loc % a.bny[g];
    const int rest = loc / a.bny[g];
    const int wz = a.bz0[g] + rest % a.bnz[g];
    const int wx = a.bx0[g] + rest / a.bnz[g];
    const Grid &G = a.g;
    const int NX = G.nx, NY = G.ny, NZ = G.nz;
    const int X0 = wx * a.lx;
    if (X0 >= NX) return;                                 // whole wavefront leaves together
    const int X1 = min(X0 + a.lx, NX);
    const int ia = X0 > 0 ? X0 - 1 : 0;
    const int j0 = wy * R;
    const int k = wz * 63 - 1 + lane;
    const bool kin = !PZ || (k >= 0 && k < NZ);
    const bool kz1 = kin && (!PZ || k < NZ - 1);          // Ez, Hx, Hy exist at this k
    const bool kint = !PZ || (k >= 1 && k <= NZ - 2);      // E curl along z
    const bool kown = kin && lane >= 1;
    const int kc = kin ? k : 0;
    const bool lx_ = a.dax != nullptr, ly_ = a.day != nullptr, lz_ = a.daz != nullptr;
    // 32-bit element offsets (the host launches this kernel only for grids of < 2^31
    // elements): global loads take the SGPR base + one VGPR offset form
    const unsigned P = G.P, si = (unsigned)NY * P;
    auto off = [&](int i, int j) { return ((unsigned)i * NY + (unsigned)j) * P + (unsigned)kc; };
    auto rowok = [&](int j) { return !PY || (j >= 0 && j < NY); };
    const float ch = a.ch;
    // z slabs: positions and profiles of this lane, fixed over the wave's life
    int zHx = -1, zHy = -1, zEx = -1, zEy = -1;
    PV zh{0.f, 0.f, 0.f}, ze{0.f, 0.f, 0.f};
    if (PZ && kin) {
        zHx = k < NZ - 1 ? strip_pos(a.sH0[0], k) : -1;
        zHy = k < NZ - 1 ? strip_pos(a.sH0[3], k) : -1;
        zEx = (k >= 1 && k <= NZ - 2) ? strip_pos(a.sE[1], k) : -1;
        zEy = (k >= 1 && k <= NZ - 2) ? strip_pos(a.sE[2], k) : -1;
        if (zHx >= 0 || zHy >= 0) zh = prof_at(a.ph[2], k);
        if (zEx >= 0 || zEy >= 0) ze = prof_at(a.pe[2], k);
    }
    // y slabs: positions and profiles per row slot, fixed over the wave's life
    int yHx[R + 1], yHz[R + 1], yEx[R + 1], yEz[R + 1];
    PV yh[R + 1], ye[R + 1];
#pragma unroll
    for (int r = 0; r <= R; ++r) {
        const int j = j0 - 1 + r;
        yHx[r] = yHz[r] = yEx[r] = yEz[r] = -1;
        yh[r] = PV{0.f, 0.f, 0.f}; ye[r] = PV{0.f, 0.f, 0.f};
        if (PY && rowok(j)) {
            if (j < NY - 1) { yHx[r] = strip_pos(a.sH0[1], j); yHz[r] = strip_pos(a.sH0[4], j); }
            if (r >= 1 && j >= 1 && j <= NY - 2) { yEx[r] = strip_pos(a.sE[0], j); yEz[r] = strip_pos(a.sE[5], j); }
            if (yHx[r] >= 0 || yHz[r] >= 0) yh[r] = prof_at(a.ph[1], j);
            if (yEx[r] >= 0 || yEz[r] >= 0) ye[r] = prof_at(a.pe[1], j);
        }
    }

    auto load_plane = [&](int i, SPPlane<R> &L) {
        const bool inside = i < X1;
        const bool xin1 = !PX || i < NX - 1;
        const bool xcurl = !PX || (i >= 1 && i <= NX - 2);
        const bool own_i = i >= X0;
        L.xHy = L.xHz = L.xEy = L.xEz = -1;
        L.xh = PV{0.f, 0.f, 0.f}; L.xe = PV{0.f, 0.f, 0.f};
        if (PX && inside) {
            if (i < NX - 1) { L.xHy = strip_pos(a.sH0[2], i); L.xHz = strip_pos(a.sH0[5], i); }
            if (xcurl) { L.xEy = strip_pos(a.sE[3], i); L.xEz = strip_pos(a.sE[4], i); }
            if (L.xHy >= 0 || L.xHz >= 0) L.xh = prof_at(a.ph[0], i);
            if (L.xEy >= 0 || L.xEz >= 0) L.xe = prof_at(a.pe[0], i);
        }
#pragma unroll
        for (int r = 0; r <= R + 1; ++r) {
            const int j = j0 - 1 + r;
            const bool ro = inside && rowok(j);
            const unsigned c = off(i, rowok(j) ? j : 0);
            L.exA[r] = (kin && ro && xin1) ? a.Ex0[c] : 0.f;
            L.ezB[r] = (kz1 && ro && xin1) ? a.Ez0[c + si] : 0.f;
            if (r <= R) {
                const bool jh = ro && j < NY - 1;
                L.eyB[r] = (kin && jh && xin1) ? a.Ey0[c + si] : 0.f;
                const bool k1 = lane == 63 && kin && (!PZ || k + 1 < NZ);
                L.exK[r] = (k1 && ro && xin1) ? a.Ex0[c + 1] : 0.f;
                L.eyK[r] = (k1 && jh) ? a.Ey0[c + 1] : 0.f;
                L.hxO[r] = (kz1 && jh) ? a.Hx0[c] : 0.f;
                L.hyO[r] = (kz1 && ro && xin1) ? a.Hy0[c] : 0.f;
                L.hzO[r] = (kin && jh && xin1) ? a.Hz0[c] : 0.f;
                const bool oe = ro && own_i && r >= 1 && kown;   // owned E cells
                L.cx[r] = (oe && xin1) ? a.cex[c] : 0.f;
                L.cy[r] = (oe && j < NY - 1) ? a.cey[c] : 0.f;
                L.cz[r] = (oe && kz1) ? a.cez[c] : 0.f;
                L.dx[r] = (oe && xin1 && lx_) ? a.dax[c] : 0.f;
                L.dy[r] = (oe && j < NY - 1 && ly_) ? a.day[c] : 0.f;
                L.dz[r] = (oe && kz1 && lz_) ? a.daz[c] : 0.f;
                L.pHxz[r] = L.pHxy[r] = L.pHyx[r] = L.pHyz[r] = L.pHzy[r] = L.pHzx[r] = 0.f;
                L.pExy[r] = L.pExz[r] = L.pEyz[r] = L.pEyx[r] = L.pEzx[r] = L.pEzy[r] = 0.f;
                const int jj = rowok(j) ? j : 0;
                const bool ycurl = !PY || (j >= 1 && j <= NY - 2);
                if (kz1 && jh && zHx >= 0) L.pHxz[r] = a.sH0[0].psi[psi_off32(a.sH0[0], 2, zHx, i, jj, kc, G)];
                if (kz1 && jh && yHx[r] >= 0) L.pHxy[r] = a.sH0[1].psi[psi_off32(a.sH0[1], 1, yHx[r], i, jj, kc, G)];
                if (kz1 && ro && L.xHy >= 0) L.pHyx[r] = a.sH0[2].psi[psi_off32(a.sH0[2], 0, L.xHy, i, jj, kc, G)];
                if (kz1 && ro && xin1 && zHy >= 0) L.pHyz[r] = a.sH0[3].psi[psi_off32(a.sH0[3], 2, zHy, i, jj, kc, G)];
                if (kin && jh && xin1 && yHz[r] >= 0) L.pHzy[r] = a.sH0[4].psi[psi_off32(a.sH0[4], 1, yHz[r], i, jj, kc, G)];
                if (kin && jh && L.xHz >= 0) L.pHzx[r] = a.sH0[5].psi[psi_off32(a.sH0[5], 0, L.xHz, i, jj, kc, G)];
                if (oe) {                                   // E psi of the owned cells
                    if (xin1 && yEx[r] >= 0 && kint) L.pExy[r] = a.sE[0].psi[psi_off32(a.sE[0], 1, yEx[r], i, jj, kc, G)];
                    if (xin1 && zEx >= 0 && ycurl) L.pExz[r] = a.sE[1].psi[psi_off32(a.sE[1], 2, zEx, i, jj, kc, G)];
                    if (jh && zEy >= 0 && xcurl) L.pEyz[r] = a.sE[2].psi[psi_off32(a.sE[2], 2, zEy, i, jj, kc, G)];
                    if (jh && L.xEy >= 0 && kint) L.pEyx[r] = a.sE[3].psi[psi_off32(a.sE[3], 0, L.xEy, i, jj, kc, G)];
                    if (kz1 && L.xEz >= 0 && ycurl) L.pEzx[r] = a.sE[4].psi[psi_off32(a.sE[4], 0, L.xEz, i, jj, kc, G)];
                    if (kz1 && yEz[r] >= 0 && xcurl) L.pEzy[r] = a.sE[5].psi[psi_off32(a.sE[5], 1, yEz[r], i, jj, kc, G)];
                }
            }
        }
    };

    float eyA[R + 1], ezA[R + 2], hyM[R + 1], hzM[R + 1];
#pragma unroll
    for (int r = 0; r <= R; ++r) { hyM[r] = 0.f; hzM[r] = 0.f; }
#pragma unroll
    for (int r = 0; r <= R + 1; ++r) {
        const int j = j0 - 1 + r;
        const unsigned c = off(ia, rowok(j) ? j : 0);
        if (r <= R) eyA[r] = (kin && rowok(j) && j < NY - 1) ? a.Ey0[c] : 0.f;
        ezA[r] = (kz1 && rowok(j)) ? a.Ez0[c] : 0.f;
    }
    SPPlane<R> cur;
    load_plane(ia, cur);
    for (int i = ia; i < X1; ++i) {
        SPPlane<R> nxt;
        load_plane(i + 1, nxt);                            // in flight while plane i runs
        const bool own_i = i >= X0;
        const bool xin1 = !PX || i < NX - 1;               // Ex, Hy, Hz exist on plane i
        const bool xcurl = !PX || (i >= 1 && i <= NX - 2);
        // ---- H^{n+1/2}(i) on slots 0..R
        float hxN[R + 1], hyN[R + 1], hzN[R + 1];
#pragma unroll
        for (int r = 0; r <= R; ++r) {
            const int j = j0 - 1 + r;
            const bool ro = rowok(j);
            const bool jh = ro && (!PY || j < NY - 1);
            const bool st_h = own_i && r >= 1 && kown && ro;
            float ey1 = lane_next(eyA[r]);
            float ex1 = lane_next(cur.exA[r]);
            if (lane == 63) { ey1 = cur.eyK[r]; ex1 = cur.exK[r]; }
            float hx = 0.f, hy = 0.f, hz = 0.f;
            const int jj = ro ? j : 0;
            if (kz1 && jh) {                                                // Hx
                const float d1 = ey1 - eyA[r];
                const float d2 = ezA[r + 1] - ezA[r];
                hx = cur.hxO[r] + ch * (d1 - d2);
                if (zHx >= 0) {
                    hx = hx + ch * psi_step(cur.pHxz[r], zh, d1);
                    if (st_h) a.sH1[0].psi[psi_off32(a.sH1[0], 2, zHx, i, jj, kc, G)] = cur.pHxz[r];
                }
                if (PY && yHx[r] >= 0) {
                    hx = hx + (-ch) * psi_step(cur.pHxy[r], yh[r], d2);
                    if (st_h) a.sH1[1].psi[psi_off32(a.sH1[1], 1, yHx[r], i, jj, kc, G)] = cur.pHxy[r];
                }
            }
            if (kz1 && ro && xin1) {                                        // Hy
                const float d1 = cur.ezB[r] - ezA[r];
                const float d2 = ex1 - cur.exA[r];
                hy = cur.hyO[r] + ch * (d1 - d2);
                if (PX && cur.xHy >= 0) {
                    hy = hy + ch * psi_step(cur.pHyx[r], cur.xh, d1);
                    if (st_h) a.sH1[2].psi[psi_off32(a.sH1[2], 0, cur.xHy, i, jj, kc, G)] = cur.pHyx[r];
                }
                if (zHy >= 0) {
                    hy = hy + (-ch) * psi_step(cur.pHyz[r], zh, d2);
                    if (st_h) a.sH1[3].psi[psi_off32(a.sH1[3], 2, zHy, i, jj, kc, G)] = cur.pHyz[r];
                }
            }
            if (kin && jh && xin1) {                                        // Hz
                const float d1 = cur.exA[r + 1] - cur.exA[r];
                const float d2 = cur.eyB[r] - eyA[r];
                hz = cur.hzO[r] + ch * (d1 - d2);
                if (PY && yHz[r] >= 0) {
                    hz = hz + ch * psi_step(cur.pHzy[r], yh[r], d1);
                    if (st_h) a.sH1[4].psi[psi_off32(a.sH1[4], 1, yHz[r], i, jj, kc, G)] = cur.pHzy[r];
                }
                if (PX && cur.xHz >= 0) {
                    hz = hz + (-ch) * psi_step(cur.pHzx[r], cur.xh, d2);
                    if (st_h) a.sH1[5].psi[psi_off32(a.sH1[5], 0, cur.xHz, i, jj, kc, G)] = cur.pHzx[r];
                }
            }
            if (st_h) {
                const unsigned c = off(i, j);
                if (kz1 && jh) a.Hx1[c] = hx;
                if (kz1 && xin1) a.Hy1[c] = hy;
                if (xin1 && jh) a.Hz1[c] = hz;
            }
            hxN[r] = hx; hyN[r] = hy; hzN[r] = hz;
        }
        // ---- E^{n+1}(i) on the owned slots 1..R
        if (own_i) {
#pragma unroll
            for (int r = 1; r <= R; ++r) {
                const int j = j0 - 1 + r;
                if (!rowok(j)) continue;
                const unsigned c = off(i, j);
                const float hxk = lane_prev(hxN[r]);                       // Hx(k-1)
                const float hyk = lane_prev(hyN[r]);                       // Hy(k-1)
                if (!kown) continue;
                const bool ycurl = !PY || (j >= 1 && j <= NY - 2);
                if (xin1) {                                                // Ex
                    float v = cur.exA[r];
                    if (lx_) v = v * cur.dx[r];
                    if (ycurl && kint) {
                        const float d1 = hzN[r] - hzN[r - 1];
                        const float d2 = hyN[r] - hyk;
                        const float ce = cur.cx[r];
                        v = v + ce * (d1 - d2);
                        if (PY && yEx[r] >= 0) {
                            v = v + ce * psi_step(cur.pExy[r], ye[r], d1);
                            a.sE[0].psi[psi_off32(a.sE[0], 1, yEx[r], i, j, k, G)] = cur.pExy[r];
                        }
                        if (zEx >= 0) {
                            v = v + (-ce) * psi_step(cur.pExz[r], ze, d2);
                            a.sE[1].psi[psi_off32(a.sE[1], 2, zEx, i, j, k, G)] = cur.pExz[r];
                        }
                    }
                    if (PY && ((a.pec[2] && j == 0) || (a.pec[3] && j == NY - 1))) v = 0.0f;
                    if (PZ && ((a.pec[4] && k == 0) || (a.pec[5] && k == NZ - 1))) v = 0.0f;
                    v = apply_drive(v, LF_EX, i, j * NZ + k, x, st);
                    a.Ex1[c] = v;
                }
                if (!PY || j < NY - 1) {                                   // Ey
                    float v = eyA[r];
                    if (ly_) v = v * cur.dy[r];
                    if (xcurl && kint) {
                        const float d1 = hxN[r] - hxk;
                        const float d2 = hzN[r] - hzM[r];
                        const float ce = cur.cy[r];
                        v = v + ce * (d1 - d2);
                        if (zEy >= 0) {
                            v = v + ce * psi_step(cur.pEyz[r], ze, d1);
                            a.sE[2].psi[psi_off32(a.sE[2], 2, zEy, i, j, k, G)] = cur.pEyz[r];
                        }
                        if (PX && cur.xEy >= 0) {
                            v = v + (-ce) * psi_step(cur.pEyx[r], cur.xe, d2);
                            a.sE[3].psi[psi_off32(a.sE[3], 0, cur.xEy, i, j, k, G)] = cur.pEyx[r];
                        }
                    }
                    if (PX && ((a.pec[0] && i == 0) || (a.pec[1] && i == NX - 1))) v = 0.0f;
                    if (PZ && ((a.pec[4] && k == 0) || (a.pec[5] && k == NZ - 1))) v = 0.0f;
                    const int plane = j * NZ + k;
                    v = apply_source(v, LF_EY, i, plane, x, st, fp64);
                    v = apply_drive(v, LF_EY, i, plane, x, st);
                    a.Ey1[c] = v;
                }
                if (kz1) {                                                 // Ez
                    float v = ezA[r];
                    if (lz_) v = v * cur.dz[r];
                    if (xcurl && ycurl) {
                        const float d1 = hyN[r] - hyM[r];
                        const float d2 = hxN[r] - hxN[r - 1];
                        const float ce = cur.cz[r];
                        v = v + ce * (d1 - d2);
                        if (PX && cur.xEz >= 0) {
                            v = v + ce * psi_step(cur.pEzx[r], cur.xe, d1);
                            a.sE[4].psi[psi_off32(a.sE[4], 0, cur.xEz, i, j, k, G)] = cur.pEzx[r];
                        }
                        if (PY && yEz[r] >= 0) {
                            v = v + (-ce) * psi_step(cur.pEzy[r], ye[r], d2);
                            a.sE[5].psi[psi_off32(a.sE[5], 1, yEz[r], i, j, k, G)] = cur.pEzy[r];
                        }
                    }
                    if (PX && ((a.pec[0] && i == 0) || (a.pec[1] && i == NX - 1))) v = 0.0f;
                    if (PY && ((a.pec[2] && j == 0) || (a.pec[3] && j == NY - 1))) v = 0.0f;
                    const int plane = j * (NZ - 1) + k;
                    v = apply_source(v, LF_EZ, i, plane, x, st, fp64);
                    v = apply_drive(v, LF_EZ, i, plane, x, st);
                    a.Ez1[c] = v;
                }
            }
        }
        // ---- carry to plane i+1
#pragma unroll
        for (int r = 0; r <= R; ++r) { hyM[r] = hyN[r]; hzM[r] = hzN[r]; eyA[r] = cur.eyB[r]; }
#pragma unroll
        for (int r = 0; r <= R + 1; ++r) ezA[r] = cur.ezB[r];
        cur = nxt;
    }
}

// ============================================================================ monitors
struct FieldView {          // one component in the padded grid
    const float *p;
    int nx, ny, nz;         // component extent (2D: ny = 1, nz = Ny)
};

struct MonArgs {
    Grid g;
    const float *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
    int kind;               // LF_TE2D ...
};

// 2D node-grid helpers (fdtd2d_core.py:458-486); 2D stores (i, 0, j) with Grid.nz = Ny.
__device__ __forceinline__ float node_ey(const MonArgs &m, int i, int j) {
    const int Ny = m.g.nz;
    const long long r = (long long)i * m.g.P;
    if (j == 0) return m.Ey[r];
    if (j == Ny - 1) return m.Ey[r + Ny - 2];
    return 0.5f * (m.Ey[r + j - 1] + m.Ey[r + j]);
}

__device__ __forceinline__ float node_ex(const MonArgs &m, int i, int j) {
    const int Nx = m.g.nx;
    if (i == 0) return m.Ex[j];
    if (i == Nx - 1) return m.Ex[(long long)(Nx - 2) * m.g.P + j];
    return 0.5f * (m.Ex[(long long)(i - 1) * m.g.P + j] + m.Ex[(long long)i * m.g.P + j]);
}

__device__ __forceinline__ float node_ex2(const MonArgs &m, int i, int j) {
    const int Nx = m.g.nx;
    if (i == 0) { const float s = m.Ex[j]; return s * s; }
    if (i == Nx - 1) { const float s = m.Ex[(long long)(Nx - 2) * m.g.P + j]; return s * s; }
    const float a = m.Ex[(long long)(i - 1) * m.g.P + j];
    const float b = m.Ex[(long long)i * m.g.P + j];
    return 0.5f * (a * a + b * b);
}

__device__ __forceinline__ float intensity2d(const MonArgs &m, int i, int j) {
    if (m.kind == LF_TE2D) {
        const float e = m.Ez[(long long)i * m.g.P + j];
        return e * e;
    }
    const float f = node_ey(m, i, j);
    return f * f + node_ex2(m, i, j);                       // fdtd2d_core.py:812
}

// <|E|^2> accumulated on the driven component's grid (fdtd3d_core.py:438-474)
__device__ float sq3d(const MonArgs &m, int form, int nxa, int i, int j, int k) {
    const Grid &g = m.g;
    const int NY = g.ny, NZ = g.nz;
    const int nF = nxa - 1;
    int a0 = i - 1; a0 = a0 < 0 ? 0 : (a0 > nF - 1 ? nF - 1 : a0);
    int b0 = i;     b0 = b0 < 0 ? 0 : (b0 > nF - 1 ? nF - 1 : b0);
    auto exs = [&](int jj, int kk) {                        // axis-0 averaged Ex^2
        const float u = m.Ex[g.at(a0, jj, kk)];
        const float w = m.Ex[g.at(b0, jj, kk)];
        return 0.5f * (u * u + w * w);
    };
    auto sq = [&](const float *F, int jj, int kk) {
        const float u = F[g.at(i, jj, kk)];
        return u * u;
    };
    float tx, ty, tz;
    if (form == LF_SQ_3D_S) {                               // target Ez grid (Nx, Ny, Nz-1)
        tx = 0.5f * (exs(j, k) + exs(j, k + 1));
        auto eyT = [&](int kk) {
            if (j == 0) return sq(m.Ey, 0, kk);
            if (j == NY - 1) return sq(m.Ey, NY - 2, kk);
            return 0.5f * (sq(m.Ey, j - 1, kk) + sq(m.Ey, j, kk));
        };
        ty = 0.5f * (eyT(k) + eyT(k + 1));
        tz = sq(m.Ez, j, k);
    } else {                                                // target Ey grid (Nx, Ny-1, Nz)
        tx = 0.5f * (exs(j, k) + exs(j + 1, k));
        ty = sq(m.Ey, j, k);
        auto ezM = [&](int kk) { return 0.5f * (sq(m.Ez, j, kk) + sq(m.Ez, j + 1, kk)); };
        if (k == 0) tz = ezM(0);
        else if (k == NZ - 1) tz = ezM(NZ - 2);
        else tz = 0.5f * (ezM(k - 1) + ezM(k));
    }
    return (tx + ty) + tz;
}

__global__ void __launch_bounds__(256) sq_kernel(MonArgs m, float *acc, int form, int nxa, int n1,
                                                 int n2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= n2 || j >= n1 || i >= nxa) return;
    float add;
    if (form == LF_SQ_TE || form == LF_SQ_TM) add = intensity2d(m, i, k);
    else if (form == LF_SQ_TM_EX2) add = node_ex2(m, i, k);
    else add = sq3d(m, form, nxa, i, j, k);
    const long long o = ((long long)i * n1 + j) * n2 + k;
    acc[o] = acc[o] + add;
}

__device__ __forceinline__ float dft_src(const MonArgs &m, int src, int i, int j, int k) {
    if (src == LF_SRC_EY_NODE) return node_ey(m, i, k);
    if (src == LF_SRC_EX_NODE) return node_ex(m, i, k);
    const float *F = src == LF_SRC_EX ? m.Ex : (src == LF_SRC_EY ? m.Ey : m.Ez);
    return F[m.g.at(i, j, k)];
}

struct DftArgs {
    int src, x0, kplane, acc, phasor;
    int n0, n1, n2;          // output extent (rows, mid, fast); kplane: (rows, 1, n2=ny)
    void *buf;
};

__global__ void __launch_bounds__(256) dft_kernel(MonArgs m, DftArgs d, double pr, double pi,
                                                  float fr, float fi) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int r = blockIdx.z;
    if (q >= d.n2 || j >= d.n1 || r >= d.n0) return;
    const int i = r + d.x0;
    float F;
    if (d.kplane >= 0) F = dft_src(m, d.src, i, q, d.kplane);   // plane (x, j) at fixed k
    else F = dft_src(m, d.src, i, j, q);
    const long long o = ((long long)r * d.n1 + j) * d.n2 + q;
    if (d.acc == LF_ACC_C128) {
        double2 *a = (double2 *)d.buf;
        double2 v = a[o];
        v.x = v.x + (double)F * pr;
        v.y = v.y + (double)F * pi;
        a[o] = v;
    } else if (d.acc == LF_ACC_C64P) {                      // complex128 promotion (NumPy 2)
        float2 *a = (float2 *)d.buf;
        float2 v = a[o];
        v.x = (float)((double)v.x + (double)F * pr);
        v.y = (float)((double)v.y + (double)F * pi);
        a[o] = v;
    } else {                                                // complex64 arithmetic
        float2 *a = (float2 *)d.buf;
        float2 v = a[o];
        v.x = v.x + F * fr;
        v.y = v.y + F * fi;
        a[o] = v;
    }
}

struct FluxArgs {
    int il, npair, wide;
    int ecomp[2], hcomp[2];
    int n1[2], n2[2];        // plane extents per pair (2D: n1 = 1)
    void *aE[2], *aH[2];
};

__device__ __forceinline__ const float *comp_ptr(const MonArgs &m, int c) {
    switch (c) {
        case LF_EX: return m.Ex;
        case LF_EY: return m.Ey;
        case LF_EZ: return m.Ez;
        case LF_HX: return m.Hx;
        case LF_HY: return m.Hy;
        default: return m.Hz;
    }
}

__global__ void __launch_bounds__(256) flux_kernel(MonArgs m, FluxArgs f, StepArgs st, double pr,
                                                   double pi, double hr, double hi, float pr32,
                                                   float pi32, float hr32, float hi32) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int p = blockIdx.z;
    if (p >= f.npair || j >= f.n1[p] || q >= f.n2[p]) return;
    const float *E = comp_ptr(m, f.ecomp[p]);
    const float *H = comp_ptr(m, f.hcomp[p]);
    const float e = E[m.g.at(f.il, j, q)];
    const float h = 0.5f * (H[m.g.at(f.il - 1, j, q)] + H[m.g.at(f.il, j, q)]);   // flux.py:42
    const long long o = (long long)j * f.n2[p] + q;
    if (f.wide) {
        double2 *aE = (double2 *)f.aE[p];
        double2 *aH = (double2 *)f.aH[p];
        double2 u = aE[o], w = aH[o];
        u.x = u.x + (double)e * pr; u.y = u.y + (double)e * pi;
        w.x = w.x + (double)h * hr; w.y = w.y + (double)h * hi;
        aE[o] = u; aH[o] = w;
    } else {
        float2 *aE = (float2 *)f.aE[p];
        float2 *aH = (float2 *)f.aH[p];
        float2 u = aE[o], w = aH[o];
        u.x = u.x + e * pr32; u.y = u.y + e * pi32;
        w.x = w.x + h * hr32; w.y = w.y + h * hi32;
        aE[o] = u; aH[o] = w;
    }
    (void)st;
}

struct BoxArgs {
    int ncomp;
    int comp[3];
    int lo[3][3], hi[3][3];
    double *partial;         // kBoxBlocks running partial sums (deterministic order)
};

__global__ void __launch_bounds__(256) box_kernel(MonArgs m, BoxArgs b) {
    double s = 0.0;
    for (int cc = 0; cc < b.ncomp; ++cc) {
        const int n0 = b.hi[cc][0] - b.lo[cc][0];
        const int n1 = b.hi[cc][1] - b.lo[cc][1];
        const int n2 = b.hi[cc][2] - b.lo[cc][2];
        const long long tot = (long long)n0 * n1 * n2;
        for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
             t += (long long)gridDim.x * blockDim.x) {
            const int q = (int)(t % n2);
            const long long rest = t / n2;
            const int j = (int)(rest % n1) + b.lo[cc][1];
            const int i = (int)(rest / n1) + b.lo[cc][0];
            const int k = q + b.lo[cc][2];
            float v;
            if (b.comp[cc] == -2) {                        // 2D TM: node Ey^2 only
                const float e = node_ey(m, i, k);
                v = e * e;
            } else if (b.comp[cc] < 0) {
                v = intensity2d(m, i, k);                  // 2D: I_E on the node grid
            } else {
                const float e = comp_ptr(m, b.comp[cc])[m.g.at(i, j, k)];
                v = e * e;
            }
            s += (double)v;
        }
    }
    double unused = 0.0;
    block_sum2_256(s, unused);                           // wavefront shuffles, fixed order
    if (threadIdx.x == 0) b.partial[blockIdx.x] += s;
}

// ---------------------------------------------------------------- time-domain stitch
// (fdtd2d_core.py:1586-1863, fdtd3d_core.py:1785-2084): plane records, TF/SF corrections,
// instantaneous Poynting sums, max |F|

constexpr int kSxBlocks = 64;       // per-block FP64 partials of a Poynting-sum monitor
constexpr int kMaxAbsBlocks = 512;  // per-block maxima of lf_field_maxabs

// dst[j*n2 + k] = F[il, j, k]: one x plane of a component into one row of a recorder
__global__ void __launch_bounds__(256) rec_plane_kernel(const float *__restrict__ f, Grid g, int il,
                                                        int n1, int n2, float *__restrict__ dst) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n1 * n2) return;
    const int j = q / n2, k = q - (q / n2) * n2;
    dst[q] = f[g.at(il, j, k)];
}

// TF/SF correction on x plane il over [j0,j1) x [k0,k1): F = F + c inc or F = F - c inc,
// c the scalar cs or the coefficient array carr at the cell; inc one row (n1 x n2) of the
// incident series, laid out like the plane
struct TfsfArgs {
    int comp, il, sign, j0, j1, k0, k1, n1, n2, rows;
    int coef;                // LF_* coefficient array id, or -1 for the context's Ch
    float *series;           // rows x n1 x n2
};

__global__ void __launch_bounds__(256) tfsf_kernel(float *__restrict__ f, Grid g, TfsfArgs a,
                                                   const float *__restrict__ carr, float cs,
                                                   int row) {
    const int w = a.k1 - a.k0;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= (a.j1 - a.j0) * w) return;
    const int j = a.j0 + q / w, k = a.k0 + (q - (q / w) * w);
    const long long o = g.at(a.il, j, k);
    const float c = carr ? carr[o] : cs;
    const float p = c * a.series[((long long)row * a.n1 + j) * a.n2 + k];
    f[o] = a.sign > 0 ? f[o] + p : f[o] - p;
}

// Instantaneous net flux in +x through plane il, H averaged onto the E plane
// (fdtd2d_core.py:489-497, fdtd3d_core.py:762-767): float32 products, FP64 sums, one
// partial per block (fixed order), accumulated over the steps of the mask
struct SxArgs {
    int il;
    double *partial;         // kSxBlocks
};

__global__ void __launch_bounds__(256) sx_kernel(MonArgs m, SxArgs a) {
    const Grid &g = m.g;
    double s = 0.0;
    const int stride = gridDim.x * 256;
    const int t0 = blockIdx.x * 256 + threadIdx.x;
    if (m.kind == LF_TE2D) {                    // S_x = -Ez hy
        for (int k = t0; k < g.nz; k += stride) {
            const float hy = 0.5f * (m.Hy[g.at(a.il - 1, 0, k)] + m.Hy[g.at(a.il, 0, k)]);
            s -= (double)(m.Ez[g.at(a.il, 0, k)] * hy);
        }
    } else if (m.kind == LF_TM2D) {             // S_x = Ey hz
        for (int k = t0; k < g.nz - 1; k += stride) {
            const float hz = 0.5f * (m.Hz[g.at(a.il - 1, 0, k)] + m.Hz[g.at(a.il, 0, k)]);
            s += (double)(m.Ey[g.at(a.il, 0, k)] * hz);
        }
    } else {                                    // S_x = Ey hz - Ez hy
        const int ny = g.ny, nz = g.nz;
        double s2 = 0.0;
        for (int q = t0; q < (ny - 1) * nz; q += stride) {
            const int j = q / nz, k = q - (q / nz) * nz;
            const float hz = 0.5f * (m.Hz[g.at(a.il - 1, j, k)] + m.Hz[g.at(a.il, j, k)]);
            s += (double)(m.Ey[g.at(a.il, j, k)] * hz);
        }
        for (int q = t0; q < ny * (nz - 1); q += stride) {
            const int j = q / (nz - 1), k = q - (q / (nz - 1)) * (nz - 1);
            const float hy = 0.5f * (m.Hy[g.at(a.il - 1, j, k)] + m.Hy[g.at(a.il, j, k)]);
            s2 += (double)(m.Ez[g.at(a.il, j, k)] * hy);
        }
        s -= s2;
    }
    double unused = 0.0;
    block_sum2_256(s, unused);
    if (threadIdx.x == 0) a.partial[blockIdx.x] += s;
}

// per-block max |F| over a component's logical extent
__global__ void __launch_bounds__(256) maxabs_kernel(const float *__restrict__ f, Grid g, int ex,
                                                     int ey, int ez, float *__restrict__ part) {
    __shared__ float wmax[4];
    float v = 0.0f;
    const long long tot = (long long)ex * ey * ez;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < tot;
         t += (long long)gridDim.x * 256) {
        const int k = (int)(t % ez);
        const long long r = t / ez;
        const int j = (int)(r % ey), i = (int)(r / ey);
        v = fmaxf(v, fabsf(f[g.at(i, j, k)]));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
}

// Monitor launch shape: rows of 256 along the fast axis when the middle extent is 1
// (2D planes), otherwise 64 x 4 tiles (3D volumes).
static inline dim3 mon_block(int n1) { return n1 == 1 ? dim3(256, 1, 1) : dim3(64, 4, 1); }

// 2D coefficient array from its compact form: row j of column i takes the band value where
// j0 <= j < j0 + rows, else the value of row j in table row cls[i] (rows of P floats)
__global__ void __launch_bounds__(256) coeff_assemble_kernel(float *dst, const int *cls,
                                                             const float *tab, const float *band,
                                                             int ny, int j0, int rows, int P) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= ny) return;
    const int r = j - j0;
    dst[(long long)i * P + j] = (r >= 0 && r < rows) ? band[(long long)i * rows + r]
                                                      : tab[(long long)cls[i] * P + j];
}

__global__ void scatter_kernel(float *dst, const float *src, int nx, int ny, int nz, int NY, int P) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)nx * ny * nz;
    if (t >= tot) return;
    const int k = (int)(t % nz);
    const long long r = t / nz;
    const int j = (int)(r % ny);
    const int i = (int)(r / ny);
    dst[((long long)i * NY + j) * P + k] = src[t];
}

__global__ void gather_kernel(float *dst, const float *src, int nx, int ny, int nz, int NY, int P) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)nx * ny * nz;
    if (t >= tot) return;
    const int k = (int)(t % nz);
    const long long r = t / nz;
    const int j = (int)(r % ny);
    const int i = (int)(r / ny);
    dst[t] = src[((long long)i * NY + j) * P + k];
}

// ---------------------------------------------------------------------------- secondary absorbers
// Graded sponge (3D, fdtd3d_core.py:397-414): the cells within n of an absorbing face are
// multiplied by g[d] per step, d the distance to that face (g[0] the outermost cell), for
// all six fields after the E update and the source.  The faces apply one after another,
// x low, x high, y low, y high, z low, z high, so a cell in several slabs takes the
// products in that order.  One 64-wide block per (i, j) row of the field: rows inside an
// x or y slab run their whole k range, the others only the z slabs.
struct SpongeArgs {
    float *f;
    int ex, ey, ez;          // extents of the field
    Grid g;
    const float *damp;       // n factors, index 0 = outermost cell
    int n;
    unsigned faces;          // bit 2*axis + hi
};

__device__ __forceinline__ float sponge_cell(const SpongeArgs &a, float v, int i, int j, int k) {
    const int n = a.n;
    if ((a.faces & 1u) && i < n) v = v * a.damp[i];
    if ((a.faces & 2u) && i >= a.ex - n) v = v * a.damp[a.ex - 1 - i];
    if ((a.faces & 4u) && j < n) v = v * a.damp[j];
    if ((a.faces & 8u) && j >= a.ey - n) v = v * a.damp[a.ey - 1 - j];
    if ((a.faces & 16u) && k < n) v = v * a.damp[k];
    if ((a.faces & 32u) && k >= a.ez - n) v = v * a.damp[a.ez - 1 - k];
    return v;
}

__global__ void __launch_bounds__(64) sponge_kernel(SpongeArgs a) {
    const int j = blockIdx.x, i = blockIdx.y;
    const int n = a.n;
    const bool row = ((a.faces & 1u) && i < n) || ((a.faces & 2u) && i >= a.ex - n) ||
                     ((a.faces & 4u) && j < n) || ((a.faces & 8u) && j >= a.ey - n);
    float *f = a.f + a.g.at(i, j, 0);
    if (row) {
        for (int k = threadIdx.x; k < a.ez; k += 64) f[k] = sponge_cell(a, f[k], i, j, k);
        return;
    }
    // z slabs only; a cell of both z slabs (n > ez/2) is visited once
    const int lo_end = (a.faces & 16u) ? min(n, a.ez) : 0;
    const int hi_beg = (a.faces & 32u) ? max(a.ez - n, lo_end) : a.ez;
    for (int k = threadIdx.x; k < lo_end; k += 64) f[k] = sponge_cell(a, f[k], i, j, k);
    for (int k = hi_beg + threadIdx.x; k < a.ez; k += 64) f[k] = sponge_cell(a, f[k], i, j, k);
}

// First-order Mur on one face of one tangential E component (boundaries3d.py:99-140;
// 2D inline fdtd2d_core.py:805-810, 831-833): the boundary plane F0 takes
// F1_prev + m (F1_new - F0_old), F1 the plane one cell inside, *_prev / *_old its values
// before the step.  The 3D form takes F0_old from before the step; the 2D inline form
// takes the plane's current value (after the TM loss factor D has scaled it).  Storage
// coordinates: the face axis `ax` (0 x, 1 y, 2 z of the padded grid), the plane spans
// the other two axes (u slow, v fast).
struct MurArgs {
    float *f;
    Grid g;
    int ax, i0, i1;          // face axis, boundary and inner indices along it
    int nu, nv;              // plane extents
    const float *m;          // per-cell coefficient (nullptr: the scalar ms)
    float ms;
    float *prev;             // [F1_prev | F0_prev] planes
    int f0cur;               // 1: F0_old = current value (2D inline form)
    // PEC zeroing of cells on PEC faces where the component is tangential (the PEC the
    // reference applies after this face's update): per storage axis, lo/hi flags
    int pz[3][2];
};

__device__ __forceinline__ long long mur_at(const MurArgs &a, int pos, int u, int v) {
    if (a.ax == 0) return a.g.at(pos, u, v);
    if (a.ax == 1) return a.g.at(u, pos, v);
    return a.g.at(u, v, pos);
}

__global__ void __launch_bounds__(256) mur_capture_kernel(MurArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.nu * a.nv;
    if (t >= n) return;
    const int u = t / a.nv, v = t - u * a.nv;
    a.prev[t] = a.f[mur_at(a, a.i1, u, v)];
    if (!a.f0cur) a.prev[n + t] = a.f[mur_at(a, a.i0, u, v)];
}

__global__ void __launch_bounds__(256) mur_apply_kernel(MurArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.nu * a.nv;
    if (t >= n) return;
    const int u = t / a.nv, v = t - u * a.nv;
    const long long c0 = mur_at(a, a.i0, u, v);
    const float f1 = a.f[mur_at(a, a.i1, u, v)];
    const float f0 = a.f0cur ? a.f[c0] : a.prev[n + t];
    const float m = a.m ? a.m[t] : a.ms;
    float r = a.prev[t] + m * (f1 - f0);
    // the cell's storage coordinates along the two plane axes
    const int au = a.ax == 0 ? 1 : 0, av = a.ax == 2 ? 1 : 2;
    const int nu1 = a.nu - 1, nv1 = a.nv - 1;
    if ((a.pz[au][0] && u == 0) || (a.pz[au][1] && u == nu1) || (a.pz[av][0] && v == 0) ||
        (a.pz[av][1] && v == nv1))
        r = 0.0f;
    a.f[c0] = r;
}

}  // namespace

// ============================================================================ host context
struct Monitor {
    int type = -1;           // 0 box, 1 sq, 2 dft, 3 flux, 4 Poynting sum, 5 plane record
    void *buf = nullptr;
    size_t bytes = 0;
    int form = 0, nxa = 0, n0 = 0, n1 = 0, n2 = 0;
    DftArgs dft{};
    FluxArgs flux{};
    BoxArgs box{};
    SxArgs sx{};
    int rcomp = 0, ril = 0, rphase = 0;  // plane record: component, x plane, phase
};

struct lf_ctx {
    int kind = 0, mode = 0;
    int nx = 0, ny = 0, nz = 0;          // logical dims (2D: nz = 1)
    Grid g{};                            // padded grid; 2D stored as (nx, 1, ny)
    hipStream_t stream = nullptr;
    hipStream_t side[3] = {};            // side streams (default priority) for slab launches
    hipEvent_t ev_fork = nullptr, ev_join[3] = {};
    hipEvent_t ev_lead = nullptr;        // context stream, after the lead class of a 2D pass
    // 2D passes in a row: the side streams of the next pass wait on the previous pass's
    // events directly (one hop) instead of on a join-then-fork through the context stream.
    // Valid only while nothing but passes has been queued since the last pass.
    bool tb_chain = false;
    unsigned tb_chain_sides = 0;         // side streams that recorded ev_join in that pass
    // partials and the 2-double result of the device steady-state test (lf_monitor_change)
    double2 *nd_part = nullptr;
    size_t nd_cap = 0;
    double *nd_out = nullptr;
    float *arr[LF_NARRAYS] = {};
    float *alt[LF_NARRAYS] = {};         // second buffer of ping-pong fields (fused path)
    int variant = 1;                     // 0: single-pass kernels where available, 1: H then E
    int lx = 32;                         // x rows per workgroup in the single-pass kernels
    float *psi_alt[4] = {};              // ping-pong psi of the H slabs (2D fused)
    float *psi_alt3[6] = {};             // ping-pong psi of the H slabs (3D fused)
    float *psi_alt3e[6] = {};            // ping-pong psi of the E slabs (3D blocked)
    // 3D blocked pass: tile plan (boxes of one wavefront class each) for (depth, lx)
    std::vector<TB3Box> tb3_boxes;
    TB3Box *tb3_dbox = nullptr;
    int tb3_T = 0, tb3_lx = 0, tb3_ntile = 0;
    int tb3_nwg = 0;                     // persistent workgroups of a pass (one per CU)
    unsigned *tb3_ticket = nullptr;      // tile ticket counter (device) and the next pass's base
    unsigned tb3_base = 0;
    int tb3_key[4] = {0, 0, 0, 0};       // source plane, drive overlap, rows, monitors of the plan
    // the plan of each depth (1..3), kept so that passes of different depths can follow one
    // another without re-planning (each keeps its own device box table); the fields above
    // are the plan of the depth used last (tb3_dbox aliases that depth's table)
    struct TB3Saved {
        std::vector<TB3Box> boxes;
        TB3Box *dbox = nullptr;
        int lx = 0, ntile = 0, nwg = 0, key[4] = {0, 0, 0, 0};
        bool valid = false;
    } tb3_saved[4];
    int tb3_rows0 = 2;                   // rows per wavefront of the slab-free class
    int ty = 4;                          // owned y rows per wavefront of the 3D single pass
    int tb = 1;                          // temporal-block depth (steps per pass); 1 = off
    int tb_lx = 64;                      // x planes per wavefront chunk of the blocked pass
    int tb_mon = 1;                      // 2D: depth of the passes that sum CW-window monitors
    int tb_lead = 0;                     // 2D: pass class launched first (0 slab-free .. 3)
    float *tbm_stage = nullptr;          // flux-plane staging of those passes
    double last_tb_ms = 0;               // blocked launches of the last lf_run: time, count
    int last_tb_launches = 0;
    float ch = 0.f;
    // profiles per axis: e (k,b,c,km1), h (k,b,c,km1)
    float *prof_e[3][4] = {}, *prof_h[3][4] = {};
    std::vector<float> host_prof_e[3][4], host_prof_h[3][4];
    int have_prof[3] = {0, 0, 0};
    int pec[3][2] = {};
    std::vector<float *> psi_bufs;
    Strip sH[6] = {}, sE[6] = {};        // 3D
    Strip s2[4] = {};                    // 2D: TE Hxy Hyx Ezx Ezy / TM Hzx Hzy Exy Eyx
    // excitation
    Excite ex{};
    std::vector<void *> ex_bufs;
    int nk_src = 0;
    Monitor mon[kMaxMon];
    int nmon = 0;
    // secondary absorbers (selected instead of the CPML on some faces): while either is
    // set, lf_run advances one step per launch and applies them after every step
    int sp_n = 0;                        // sponge cells (3D), 0 = none
    unsigned sp_faces = 0;               // bit 2*axis + hi
    const float *sp_g = nullptr;         // damping factors, index 0 = outermost cell
    std::vector<MurArgs> mur;            // Mur faces, in the order they apply
    std::vector<int> mur_comp;           // their field components (LF_E*)
    // time-domain stitch: TF/SF corrections (each with its incident series on the device)
    // and plane recorders; while either exists lf_run advances one step per launch
    std::vector<TfsfArgs> tfsf;
    int nrec = 0;
    float *mx_part = nullptr;            // per-block maxima of lf_field_maxabs
    // compact forms of the 2D coefficient arrays built by lf_build_coeff, on the host:
    // per-plane table row (planes from the array's extent on: the zero row ncls), table
    // rows of P floats, band rows [j0, j0 + rows) of every plane
    struct CoefForm {
        std::vector<int> cls;
        std::vector<float> tab, band;
        int j0 = 0, rows = 0;
    };
    CoefForm cform[LF_NARRAYS];
    // the forms of a context's arrays merged into one map on the device (CoefMap: one
    // table row per plane and one band for all arrays), which the blocked passes read;
    // on == false: they read the plain arrays through iota (plane p -> row p)
    struct CoefMerged {
        bool on = false;
        int *cls = nullptr;
        float *tab[LF_NARRAYS] = {}, *band[LF_NARRAYS] = {};
        int j0 = 0, rows = 0;
    };
    CoefMerged cmerge;
    int *iota = nullptr;
    bool finalized = false;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double last_update_ms = 0, last_total_ms = 0;
    int last_steps = 0;
};

namespace {

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// component extents in the logical grid
void comp_extent(const lf_ctx *c, int id, int &ex, int &ey, int &ez) {
    const int NX = c->nx, NY = c->ny, NZ = c->nz;
    ex = NX; ey = NY; ez = NZ;
    if (c->kind == LF_YEE3D) {
        switch (id) {
            case LF_EX: case LF_CEX: case LF_DAX: ex = NX - 1; break;
            case LF_EY: case LF_CEY: case LF_DAY: ey = NY - 1; break;
            case LF_EZ: case LF_CEZ: case LF_DAZ: ez = NZ - 1; break;
            case LF_HX: ey = NY - 1; ez = NZ - 1; break;
            case LF_HY: ex = NX - 1; ez = NZ - 1; break;
            case LF_HZ: ex = NX - 1; ey = NY - 1; break;
            default: break;
        }
    } else {
        // 2D stored as (NX, 1, NY): y extent is "ez"
        int a = NX, b = NY;
        switch (id) {
            case LF_HX: b = NY - 1; break;                              // TE Hx (Nx, Ny-1)
            case LF_HY: a = NX - 1; break;                              // TE Hy (Nx-1, Ny)
            case LF_EX: case LF_CAX: case LF_DAX: a = NX - 1; break;    // TM Ex (Nx-1, Ny)
            case LF_EY: case LF_CAY: case LF_DAY: b = NY - 1; break;    // TM Ey (Nx, Ny-1)
            case LF_HZ: a = NX - 1; b = NY - 1; break;                  // TM Hz
            default: break;
        }
        ex = a; ey = 1; ez = b;
    }
}

size_t grid_elems(const lf_ctx *c) { return (size_t)c->g.nx * c->g.ny * c->g.P; }

void drop_coef_merge(lf_ctx *c) {
    auto &m = c->cmerge;
    if (m.cls) (void)hipFree(m.cls);
    for (auto p : m.tab) if (p) (void)hipFree(p);
    for (auto p : m.band) if (p) (void)hipFree(p);
    m = lf_ctx::CoefMerged{};
}

// Whether the blocked 2D passes read coefficients in compact form where one exists
// (LENSFDTD_COMPACT_COEFF=0: always the plain arrays; the values are the same).
bool compact_coeff_on() {
    static const bool on = [] {
        const char *v = std::getenv("LENSFDTD_COMPACT_COEFF");
        return !(v && v[0] == '0');
    }();
    return on;
}

// Merges the compact forms of the context's present coefficient arrays into one map: a
// plane's merged class is the tuple of its classes in every array, and every array's band
// widens to the union of the bands, the added rows taken from its own table.  Element
// (p, j) of every array keeps its value (that of the assembled array, zero outside the
// array's extent).  A present array without a form (uploaded plainly) leaves the map off.
int merge_coef_forms(lf_ctx *c) {
    drop_coef_merge(c);
    const int te[2] = {LF_CEE, LF_CEH}, tm[4] = {LF_CAX, LF_CAY, LF_DAX, LF_DAY};
    const int *ids = c->kind == LF_TE2D ? te : tm;
    const int n = c->kind == LF_TE2D ? 2 : 4;
    std::vector<int> use;
    int J0 = 1 << 30, J1 = 0;
    for (int u = 0; u < n; ++u) {
        if (!c->arr[ids[u]]) continue;
        const auto &f = c->cform[ids[u]];
        if (f.cls.empty()) return 0;
        use.push_back(ids[u]);
        if (f.rows > 0) { J0 = std::min(J0, f.j0); J1 = std::max(J1, f.j0 + f.rows); }
    }
    if (use.empty()) return 0;
    if (J1 == 0) J0 = J1 = 0;
    const int NX = c->g.nx, P = c->g.P, R = J1 - J0;
    std::map<std::vector<int>, int> key;
    std::vector<std::vector<int>> rows_of;            // merged class -> class per array
    std::vector<int> cls(NX);
    for (int i = 0; i < NX; ++i) {
        std::vector<int> k;
        for (int id : use) k.push_back(c->cform[id].cls[i]);
        auto it = key.find(k);
        if (it == key.end()) {
            if (key.size() >= 4096) return 0;          // no layer structure to share
            it = key.emplace(k, (int)rows_of.size()).first;
            rows_of.push_back(k);
        }
        cls[i] = it->second;
    }
    const size_t nc = rows_of.size();
    auto &m = c->cmerge;
    LF_HIP(hipMalloc(&m.cls, NX * sizeof(int)));
    LF_HIP(hipMemcpyAsync(m.cls, cls.data(), NX * sizeof(int), hipMemcpyHostToDevice, c->stream));
    std::vector<float> tab(nc * P), band((size_t)NX * R);
    for (size_t u = 0; u < use.size(); ++u) {
        const int id = use[u];
        const auto &f = c->cform[id];
        for (size_t k = 0; k < nc; ++k)
            std::memcpy(&tab[k * P], &f.tab[(size_t)rows_of[k][u] * P], P * sizeof(float));
        for (int i = 0; i < NX; ++i)
            for (int r = 0; r < R; ++r) {
                const int j = J0 + r, rb = j - f.j0;
                band[(size_t)i * R + r] = (rb >= 0 && rb < f.rows) ? f.band[(size_t)i * f.rows + rb]
                                                                   : f.tab[(size_t)f.cls[i] * P + j];
            }
        LF_HIP(hipMalloc(&m.tab[id], tab.size() * sizeof(float)));
        LF_HIP(hipMemcpy(m.tab[id], tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
        if (R) {
            LF_HIP(hipMalloc(&m.band[id], band.size() * sizeof(float)));
            LF_HIP(hipMemcpy(m.band[id], band.data(), band.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    LF_HIP(hipStreamSynchronize(c->stream));
    m.j0 = J0; m.rows = R;
    m.on = true;
    c->tb_chain = false;
    return 0;
}

// the blocked passes' view of coefficient array id (tab == nullptr when absent) and of
// the map its context's arrays share
Coef2 coef_src(const lf_ctx *c, int id) {
    if (!c->arr[id]) return Coef2{nullptr, nullptr};
    if (c->cmerge.on && compact_coeff_on()) return Coef2{c->cmerge.tab[id], c->cmerge.band[id]};
    return Coef2{c->arr[id], nullptr};
}
CoefMap coef_map(const lf_ctx *c) {
    if (c->cmerge.on && compact_coeff_on()) return CoefMap{c->cmerge.cls, c->cmerge.j0, c->cmerge.rows};
    return CoefMap{c->iota, 0, 0};
}

int ensure_array(lf_ctx *c, int id) {
    if (c->arr[id]) return 0;
    LF_HIP(hipMalloc(&c->arr[id], grid_elems(c) * sizeof(float)));
    LF_HIP(hipMemsetAsync(c->arr[id], 0, grid_elems(c) * sizeof(float), c->stream));
    return 0;
}

Prof prof_of(const lf_ctx *c, int axis, bool half) {
    float *const *p = half ? c->prof_h[axis] : c->prof_e[axis];
    return Prof{p[1], p[2], p[0], p[3]};
}

// Slab extent of a profile over the index range [first, last] that the update touches:
// indices < lo or >= hi carry psi (c != 0 there); lo = hi = n when the slabs meet
// (boundaries3d.py:216-224 scans the sliced profile the same way).
void strip_range(const std::vector<float> &cprof, int first, int last, int &lo, int &hi) {
    const int N = (int)cprof.size();
    const int span = last - first + 1;
    int nlo = 0, nhi = 0;
    while (nlo < span && cprof[first + nlo] != 0.0f) ++nlo;
    while (nhi < span && cprof[last - nhi] != 0.0f) ++nhi;
    if (nlo + nhi >= span) { lo = N; hi = N; return; }
    lo = first + nlo;
    hi = last + 1 - nhi;
}

// allocate a psi slab compressed along `axis` over a field grid of dims (a0, a1, a2);
// positions: v < lo -> v, v >= hi -> v - hi + lo
int make_strip(lf_ctx *c, Strip &s, const std::vector<float> &cprof, int first, int last, int axis,
               long long a0, long long a1, long long a2) {
    int lo, hi;
    strip_range(cprof, first, last, lo, hi);
    const int N = (int)cprof.size();
    const int len = lo + (N - hi);
    s = Strip{nullptr, lo, hi, len};
    if (lo == first && hi == last + 1) return 0;      // no PML cell along this axis
    long long dims[3] = {a0, a1, a2};
    dims[axis] = len;
    const size_t n = (size_t)(dims[0] * dims[1] * dims[2]);
    float *p = nullptr;
    LF_HIP(hipMalloc(&p, n * sizeof(float)));
    LF_HIP(hipMemsetAsync(p, 0, n * sizeof(float), c->stream));
    c->psi_bufs.push_back(p);
    s.psi = p;
    return 0;
}

int upload_vec(lf_ctx *c, const void *host, size_t bytes, void **dev) {
    LF_HIP(hipMalloc(dev, bytes));
    LF_HIP(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, c->stream));
    c->ex_bufs.push_back(*dev);
    return 0;
}

TEArgs te_args(lf_ctx *c) {
    TEArgs a{};
    a.g = c->g;
    a.Ez = c->arr[LF_EZ]; a.Hx = c->arr[LF_HX]; a.Hy = c->arr[LF_HY];
    a.ceE = c->arr[LF_CEE]; a.ceH = c->arr[LF_CEH];
    a.ch = c->ch;
    a.xe = prof_of(c, 0, false); a.xh = prof_of(c, 0, true);
    a.ye = prof_of(c, 1, false); a.yh = prof_of(c, 1, true);
    a.pHxy = c->s2[0]; a.pHyx = c->s2[1]; a.pEzx = c->s2[2]; a.pEzy = c->s2[3];
    a.pec_xlo = c->pec[0][0]; a.pec_xhi = c->pec[0][1];
    a.pec_ylo = c->pec[1][0]; a.pec_yhi = c->pec[1][1];
    return a;
}

TMArgs tm_args(lf_ctx *c) {
    TMArgs a{};
    a.g = c->g;
    a.Ex = c->arr[LF_EX]; a.Ey = c->arr[LF_EY]; a.Hz = c->arr[LF_HZ];
    a.cax = c->arr[LF_CAX]; a.cay = c->arr[LF_CAY];
    a.dax = c->arr[LF_DAX]; a.day = c->arr[LF_DAY];
    a.ch = c->ch;
    a.xe = prof_of(c, 0, false); a.xh = prof_of(c, 0, true);
    a.ye = prof_of(c, 1, false); a.yh = prof_of(c, 1, true);
    a.pHzx = c->s2[0]; a.pHzy = c->s2[1]; a.pExy = c->s2[2]; a.pEyx = c->s2[3];
    a.pec_xlo = c->pec[0][0]; a.pec_xhi = c->pec[0][1];
    a.pec_ylo = c->pec[1][0]; a.pec_yhi = c->pec[1][1];
    return a;
}

YeeArgs yee_args(lf_ctx *c) {
    YeeArgs a{};
    a.g = c->g;
    a.Ex = c->arr[LF_EX]; a.Ey = c->arr[LF_EY]; a.Ez = c->arr[LF_EZ];
    a.Hx = c->arr[LF_HX]; a.Hy = c->arr[LF_HY]; a.Hz = c->arr[LF_HZ];
    a.cex = c->arr[LF_CEX]; a.cey = c->arr[LF_CEY]; a.cez = c->arr[LF_CEZ];
    a.dax = c->arr[LF_DAX]; a.day = c->arr[LF_DAY]; a.daz = c->arr[LF_DAZ];
    a.ch = c->ch;
    for (int ax = 0; ax < 3; ++ax) { a.pe[ax] = prof_of(c, ax, false); a.ph[ax] = prof_of(c, ax, true); }
    for (int q = 0; q < 6; ++q) { a.sH[q] = c->sH[q]; a.sE[q] = c->sE[q]; }
    for (int ax = 0; ax < 3; ++ax) { a.pec[2 * ax] = c->pec[ax][0]; a.pec[2 * ax + 1] = c->pec[ax][1]; }
    return a;
}

MonArgs mon_args(lf_ctx *c) {
    MonArgs m{};
    m.g = c->g;
    m.Ex = c->arr[LF_EX]; m.Ey = c->arr[LF_EY]; m.Ez = c->arr[LF_EZ];
    m.Hx = c->arr[LF_HX]; m.Hy = c->arr[LF_HY]; m.Hz = c->arr[LF_HZ];
    m.kind = c->kind;
    return m;
}

int check_ctx(lf_ctx *c) {
    if (!c) return fail(-1, "null context");
    return 0;
}

}  // namespace

// ============================================================================ C ABI
extern "C" {

const char *lf_last_error(void) { return g_err.c_str(); }

int lf_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; return fail(-2, hipGetErrorString(e)); }
    *count = n;
    return 0;
}

int lf_device_name(char *buf, int len) {
    hipDeviceProp_t p;
    LF_HIP(hipGetDeviceProperties(&p, 0));
    std::snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

int lf_create(int kind, int mode, int nx, int ny, int nz, lf_ctx **out) {
    if (!out) return fail(-1, "null out");
    if (kind < LF_TE2D || kind > LF_YEE3D) return fail(-1, "bad kind");
    if (mode != LF_MODE_REF32 && mode != LF_MODE_FP64ACC) return fail(-1, "bad mode");
    if (nx < 3 || ny < 3 || (kind == LF_YEE3D && nz < 3)) return fail(-1, "grid too small");
    lf_ctx *c = new lf_ctx();
    c->kind = kind; c->mode = mode;
    if (kind != LF_YEE3D) c->tb = 6;             // blocked passes over monitor-free steps
    if (kind == LF_YEE3D) c->tb = 3;
    if (kind == LF_TM2D) c->tb_mon = 4;          // monitored passes (CW-window steps)
    if (kind == LF_TE2D) c->tb_mon = 4;
    c->nx = nx; c->ny = ny; c->nz = (kind == LF_YEE3D) ? nz : 1;
    if (kind == LF_YEE3D) c->g = Grid{nx, ny, nz, round_up(nz, 64)};
    else c->g = Grid{nx, 1, ny, round_up(ny, 64)};
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) { delete c; return fail(-2, hipGetErrorString(e)); }
    const int comps3[6] = {LF_EX, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ};
    const int compsTE[3] = {LF_EZ, LF_HX, LF_HY};
    const int compsTM[3] = {LF_EX, LF_EY, LF_HZ};
    const int *cs = kind == LF_YEE3D ? comps3 : (kind == LF_TE2D ? compsTE : compsTM);
    const int nc = kind == LF_YEE3D ? 6 : 3;
    for (int q = 0; q < nc; ++q) {
        int r = ensure_array(c, cs[q]);
        if (r) { lf_destroy(c); return r; }
    }
    if (kind != LF_YEE3D) {
        std::vector<int> io(c->g.nx);
        for (int i = 0; i < c->g.nx; ++i) io[i] = i;
        e = hipMalloc(&c->iota, io.size() * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpy(c->iota, io.data(), io.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e != hipSuccess) { lf_destroy(c); return fail(-2, hipGetErrorString(e)); }
    }
    *out = c;
    return 0;
}

int lf_destroy(lf_ctx *c) {
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->arr) if (p) (void)hipFree(p);
    for (auto &p : c->alt) if (p) (void)hipFree(p);
    {   // psi buffers: originals, ping-pong partners (pointers swap every fused step)
        std::set<float *> psi(c->psi_bufs.begin(), c->psi_bufs.end());
        for (auto p : c->psi_alt) if (p) psi.insert(p);
        for (auto p : c->psi_alt3) if (p) psi.insert(p);
        for (auto p : c->psi_alt3e) if (p) psi.insert(p);
        for (auto &s : c->s2) if (s.psi) psi.insert(s.psi);
        for (auto &s : c->sH) if (s.psi) psi.insert(s.psi);
        for (auto &s : c->sE) if (s.psi) psi.insert(s.psi);
        for (auto p : psi) (void)hipFree(p);
        c->psi_bufs.clear();
    }
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 4; ++q) {
            if (c->prof_e[a][q]) (void)hipFree(c->prof_e[a][q]);
            if (c->prof_h[a][q]) (void)hipFree(c->prof_h[a][q]);
        }
    for (auto p : c->ex_bufs) (void)hipFree(p);
    for (auto &sv : c->tb3_saved)
        if (sv.dbox) (void)hipFree(sv.dbox);
    if (c->tb3_ticket) (void)hipFree(c->tb3_ticket);
    if (c->tbm_stage) (void)hipFree(c->tbm_stage);
    if (c->nd_part) (void)hipFree(c->nd_part);
    if (c->nd_out) (void)hipFree(c->nd_out);
    for (auto &t : c->tfsf) if (t.series) (void)hipFree(t.series);
    if (c->mx_part) (void)hipFree(c->mx_part);
    drop_coef_merge(c);
    if (c->iota) (void)hipFree(c->iota);
    for (int m = 0; m < c->nmon; ++m) {
        if (c->mon[m].buf) (void)hipFree(c->mon[m].buf);
        if (c->mon[m].type == 3)
            for (int p = 0; p < c->mon[m].flux.npair; ++p) {
                (void)hipFree(c->mon[m].flux.aE[p]);
                (void)hipFree(c->mon[m].flux.aH[p]);
            }
    }
    for (auto e : c->ev) (void)hipEventDestroy(e);
    for (auto st : c->side) if (st) (void)hipStreamSynchronize(st);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_lead) (void)hipEventDestroy(c->ev_lead);
    for (auto e : c->ev_join) if (e) (void)hipEventDestroy(e);
    for (auto st : c->side) if (st) (void)hipStreamDestroy(st);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    (void)hipGetLastError();   // teardown errors must not surface in a later call
    return 0;
}

int lf_upload(lf_ctx *c, int id, const float *host, long long count) {
    if (int r = check_ctx(c)) return r;
    if (id < 0 || id >= LF_NARRAYS) return fail(-1, "bad array id");
    int ex, ey, ez;
    comp_extent(c, id, ex, ey, ez);
    if (count != (long long)ex * ey * ez)
        return fail(-1, "lf_upload: element count " + std::to_string(count) + " != expected " +
                            std::to_string((long long)ex * ey * ez));
    if (int r = ensure_array(c, id)) return r;
    if (id >= LF_CEE && c->kind != LF_YEE3D) {      // the plain array is the data now
        c->cform[id] = lf_ctx::CoefForm{};
        if (int r = merge_coef_forms(c)) return r;
    }
    float *tmp = nullptr;
    LF_HIP(hipMalloc(&tmp, (size_t)count * sizeof(float)));
    LF_HIP(hipMemcpyAsync(tmp, host, (size_t)count * sizeof(float), hipMemcpyHostToDevice, c->stream));
    const int nb = (int)((count + 255) / 256);
    scatter_kernel<<<nb, 256, 0, c->stream>>>(c->arr[id], tmp, ex, ey, ez, c->g.ny, c->g.P);
    LF_HIP(hipGetLastError());
    LF_HIP(hipStreamSynchronize(c->stream));
    LF_HIP(hipFree(tmp));
    return 0;
}

int lf_build_coeff(lf_ctx *c, int id, const unsigned char *xclass, int ncls, const float *table,
                   int j0, int rows, const float *band) {
    if (int r = check_ctx(c)) return r;
    if (c->kind == LF_YEE3D) return fail(-1, "lf_build_coeff: 2D contexts only");
    if (id < LF_CEE || id >= LF_NARRAYS) return fail(-1, "lf_build_coeff: not a coefficient id");
    int ex, ey, ez;
    comp_extent(c, id, ex, ey, ez);
    if (ncls < 1 || ncls > 255 || !xclass || !table) return fail(-1, "lf_build_coeff: bad class table");
    if (rows < 0 || j0 < 0 || j0 + rows > ez || (rows > 0 && !band))
        return fail(-1, "lf_build_coeff: band rows outside the array");
    for (int i = 0; i < ex; ++i)
        if (xclass[i] >= ncls) return fail(-1, "lf_build_coeff: class index out of range");
    if (int r = ensure_array(c, id)) return r;
    // the compact form, kept on the host (merge_coef_forms), equals the assembled array
    // over the whole allocation: planes from ex on (the staggered arrays' padding plane)
    // take the zero table row ncls and zero band rows, table rows are zero from ez to P,
    // as the zero-filled array is
    const int NX = c->g.nx, P = c->g.P;
    lf_ctx::CoefForm f{};
    f.cls.assign(NX, ncls);
    for (int i = 0; i < ex; ++i) f.cls[i] = xclass[i];
    f.tab.assign((size_t)(ncls + 1) * P, 0.f);
    for (int k = 0; k < ncls; ++k)
        std::memcpy(&f.tab[(size_t)k * P], table + (size_t)k * ez, (size_t)ez * sizeof(float));
    f.band.assign((size_t)NX * rows, 0.f);
    if (rows) std::memcpy(f.band.data(), band, (size_t)ex * rows * sizeof(float));
    f.j0 = j0; f.rows = rows;
    int *dcls = nullptr;
    float *dtab = nullptr, *dband = nullptr;
    LF_HIP(hipMalloc(&dcls, f.cls.size() * sizeof(int)));
    LF_HIP(hipMalloc(&dtab, f.tab.size() * sizeof(float)));
    if (rows) LF_HIP(hipMalloc(&dband, f.band.size() * sizeof(float)));
    LF_HIP(hipMemcpyAsync(dcls, f.cls.data(), f.cls.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    LF_HIP(hipMemcpyAsync(dtab, f.tab.data(), f.tab.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (rows)
        LF_HIP(hipMemcpyAsync(dband, f.band.data(), f.band.size() * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
    coeff_assemble_kernel<<<dim3((ez + 255) / 256, ex, 1), 256, 0, c->stream>>>(
        c->arr[id], dcls, dtab, dband, ez, j0, rows, P);
    LF_HIP(hipGetLastError());
    LF_HIP(hipStreamSynchronize(c->stream));
    LF_HIP(hipFree(dcls));
    LF_HIP(hipFree(dtab));
    if (dband) LF_HIP(hipFree(dband));
    c->cform[id] = std::move(f);
    if (int r = merge_coef_forms(c)) return r;
    c->tb_chain = false;
    return 0;
}

int lf_download(lf_ctx *c, int id, float *host, long long count) {
    if (int r = check_ctx(c)) return r;
    if (id < 0 || id >= LF_NARRAYS || !c->arr[id]) return fail(-1, "array not present");
    int ex, ey, ez;
    comp_extent(c, id, ex, ey, ez);
    if (count != (long long)ex * ey * ez) return fail(-1, "lf_download: element count mismatch");
    float *tmp = nullptr;
    LF_HIP(hipMalloc(&tmp, (size_t)count * sizeof(float)));
    const int nb = (int)((count + 255) / 256);
    gather_kernel<<<nb, 256, 0, c->stream>>>(tmp, c->arr[id], ex, ey, ez, c->g.ny, c->g.P);
    LF_HIP(hipGetLastError());
    LF_HIP(hipMemcpyAsync(host, tmp, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(hipStreamSynchronize(c->stream));
    LF_HIP(hipFree(tmp));
    return 0;
}

int lf_set_ch(lf_ctx *c, float ch) {
    if (int r = check_ctx(c)) return r;
    c->ch = ch;
    return 0;
}

int lf_set_cpml(lf_ctx *c, int axis, const float *ek, const float *eb, const float *ec,
                const float *hk, const float *hb, const float *hc, int pec_lo, int pec_hi) {
    if (int r = check_ctx(c)) return r;
    const int ndim = c->kind == LF_YEE3D ? 3 : 2;
    if (axis < 0 || axis >= ndim) return fail(-1, "bad axis");
    if (c->finalized) return fail(-1, "lf_set_cpml after lf_finalize");
    const int N = axis == 0 ? c->nx : (axis == 1 ? c->ny : c->nz);
    const float *src_e[3] = {ek, eb, ec}, *src_h[3] = {hk, hb, hc};
    for (int q = 0; q < 3; ++q) {
        c->host_prof_e[axis][q].assign(src_e[q], src_e[q] + N);
        c->host_prof_h[axis][q].assign(src_h[q], src_h[q] + N - 1);
    }
    // km1 = 1/kappa - 1 in float32 (boundaries3d.py:231)
    c->host_prof_e[axis][3].resize(N);
    c->host_prof_h[axis][3].resize(N - 1);
    for (int n = 0; n < N; ++n) c->host_prof_e[axis][3][n] = 1.0f / ek[n] - 1.0f;
    for (int n = 0; n < N - 1; ++n) c->host_prof_h[axis][3][n] = 1.0f / hk[n] - 1.0f;
    for (int q = 0; q < 4; ++q) {
        if (c->prof_e[axis][q]) (void)hipFree(c->prof_e[axis][q]);
        if (c->prof_h[axis][q]) (void)hipFree(c->prof_h[axis][q]);
        LF_HIP(hipMalloc(&c->prof_e[axis][q], N * sizeof(float)));
        LF_HIP(hipMalloc(&c->prof_h[axis][q], (N - 1) * sizeof(float)));
        LF_HIP(hipMemcpyAsync(c->prof_e[axis][q], c->host_prof_e[axis][q].data(), N * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
        LF_HIP(hipMemcpyAsync(c->prof_h[axis][q], c->host_prof_h[axis][q].data(),
                              (N - 1) * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    LF_HIP(hipStreamSynchronize(c->stream));
    c->have_prof[axis] = 1;
    c->pec[axis][0] = pec_lo ? 1 : 0;
    c->pec[axis][1] = pec_hi ? 1 : 0;
    return 0;
}

int lf_finalize(lf_ctx *c) {
    if (int r = check_ctx(c)) return r;
    const int ndim = c->kind == LF_YEE3D ? 3 : 2;
    for (int a = 0; a < ndim; ++a)
        if (!c->have_prof[a]) return fail(-1, "lf_finalize: CPML profile missing for an axis");
    if (c->finalized) return 0;
    const long long NX = c->g.nx, NY = c->g.ny, P = c->g.P;
    int r = 0;
    if (c->kind == LF_YEE3D) {
        // psi slabs compressed along the stretch axis.  H: half-node profiles over all
        // n-1 nodes; E: full-node profiles over the interior [1, n-2] that the E update
        // touches (boundaries3d.py:286-300).
        const int N3[3] = {c->nx, c->ny, c->nz};
        auto H = [&](Strip &s, int ax) {
            return make_strip(c, s, c->host_prof_h[ax][2], 0, N3[ax] - 2, ax, NX, NY, P);
        };
        auto E = [&](Strip &s, int ax) {
            return make_strip(c, s, c->host_prof_e[ax][2], 1, N3[ax] - 2, ax, NX, NY, P);
        };
        // Hx_z Hx_y Hy_x Hy_z Hz_y Hz_x
        if ((r = H(c->sH[0], 2)) || (r = H(c->sH[1], 1)) || (r = H(c->sH[2], 0)) ||
            (r = H(c->sH[3], 2)) || (r = H(c->sH[4], 1)) || (r = H(c->sH[5], 0)))
            return r;
        // Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y
        if ((r = E(c->sE[0], 1)) || (r = E(c->sE[1], 2)) || (r = E(c->sE[2], 2)) ||
            (r = E(c->sE[3], 0)) || (r = E(c->sE[4], 0)) || (r = E(c->sE[5], 1)))
            return r;
    } else {
        // 2D grid (NX, 1, P): x slabs compress rows (index s*P + j), y slabs compress
        // the fast axis (index i*len + s)
        const int N2[2] = {c->nx, c->ny};
        auto H = [&](Strip &s, int ax) {
            return make_strip(c, s, c->host_prof_h[ax][2], 0, N2[ax] - 2, ax == 0 ? 0 : 2, NX, 1, P);
        };
        auto E = [&](Strip &s, int ax) {
            return make_strip(c, s, c->host_prof_e[ax][2], 1, N2[ax] - 2, ax == 0 ? 0 : 2, NX, 1, P);
        };
        if (c->kind == LF_TE2D) {   // psi_Hxy psi_Hyx psi_Ezx psi_Ezy (boundaries.py:180-184)
            if ((r = H(c->s2[0], 1)) || (r = H(c->s2[1], 0)) || (r = E(c->s2[2], 0)) ||
                (r = E(c->s2[3], 1)))
                return r;
        } else {                    // psi_Hzx psi_Hzy psi_Exy psi_Eyx (boundaries.py:235-238)
            if ((r = H(c->s2[0], 0)) || (r = H(c->s2[1], 1)) || (r = E(c->s2[2], 1)) ||
                (r = E(c->s2[3], 0)))
                return r;
        }
    }
    LF_HIP(hipStreamSynchronize(c->stream));
    c->finalized = true;
    return 0;
}

int lf_set_source(lf_ctx *c, int comp, int ix, int n, const float *amp, const float *ph1,
                  const float *ph2) {
    if (int r = check_ctx(c)) return r;
    int ex, ey, ez;
    comp_extent(c, comp, ex, ey, ez);
    if (n != ey * ez) return fail(-1, "lf_set_source: plane size mismatch");
    if (ix < 0 || ix + 1 >= ex) return fail(-1, "lf_set_source: plane index out of range");
    std::vector<double> a1c(n), a1s(n), a2c(n), a2s(n);
    for (int q = 0; q < n; ++q) {
        const double a = amp[q], p1 = ph1[q], p2 = ph2[q];
        a1c[q] = a * std::cos(p1); a1s[q] = a * std::sin(p1);
        a2c[q] = a * std::cos(p2); a2s[q] = a * std::sin(p2);
    }
    void *d;
    int r;
    if ((r = upload_vec(c, amp, n * sizeof(float), &d))) return r; c->ex.amp = (const float *)d;
    if ((r = upload_vec(c, ph1, n * sizeof(float), &d))) return r; c->ex.ph1 = (const float *)d;
    if ((r = upload_vec(c, ph2, n * sizeof(float), &d))) return r; c->ex.ph2 = (const float *)d;
    if ((r = upload_vec(c, a1c.data(), n * sizeof(double), &d))) return r; c->ex.a1c = (const double *)d;
    if ((r = upload_vec(c, a1s.data(), n * sizeof(double), &d))) return r; c->ex.a1s = (const double *)d;
    if ((r = upload_vec(c, a2c.data(), n * sizeof(double), &d))) return r; c->ex.a2c = (const double *)d;
    if ((r = upload_vec(c, a2s.data(), n * sizeof(double), &d))) return r; c->ex.a2s = (const double *)d;
    c->ex.src_comp = comp;
    c->ex.src_ix = ix;
    c->ex.src_n = n;
    LF_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int lf_set_drive(lf_ctx *c, int ncells, const float *wtap, int ncomp, const int *comps,
                 const float *drives, int math) {
    if (int r = check_ctx(c)) return r;
    if (ncells <= 0) return fail(-1, "lf_set_drive: ncells must be > 0");
    void *d;
    int r;
    if ((r = upload_vec(c, wtap, ncells * sizeof(float), &d))) return r;
    c->ex.wtap = (const float *)d;
    c->ex.drv_n = ncells;
    c->ex.drv_math = math;
    size_t off = 0;
    for (int q = 0; q < ncomp; ++q) {
        if (comps[q] < LF_EX || comps[q] > LF_EZ) return fail(-1, "lf_set_drive: not an E component");
        int ex, ey, ez;
        comp_extent(c, comps[q], ex, ey, ez);
        if (ncells > ex) return fail(-1, "lf_set_drive: overlap larger than the component");
        const int n = ey * ez;
        if ((r = upload_vec(c, drives + off, (size_t)ncells * n * 2 * sizeof(float), &d))) return r;
        off += (size_t)ncells * n * 2;
        c->ex.drv[comps[q]] = (const float2 *)d;
        c->ex.drv_stride[comps[q]] = n;
    }
    LF_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Graded sponge of n cells on the faces in `faces` (bit 2*axis + hi), damping factors
// g[0..n-1] with g[0] at the outermost cell (fdtd3d_core.py:397-414).  3D only.
int lf_set_sponge(lf_ctx *c, int n, const float *g, unsigned faces) {
    if (int r = check_ctx(c)) return r;
    if (c->kind != LF_YEE3D) return fail(-1, "lf_set_sponge: the sponge is a 3D absorber");
    if (n <= 0 || g == nullptr) return fail(-1, "lf_set_sponge: n must be > 0 with n factors");
    if (faces == 0u || faces > 63u) return fail(-1, "lf_set_sponge: faces must be a nonempty subset of 6");
    if (n > c->nx - 1 || n > c->ny - 1 || n > c->nz - 1)
        return fail(-1, "lf_set_sponge: sponge thicker than the grid");
    void *d;
    if (int r = upload_vec(c, g, (size_t)n * sizeof(float), &d)) return r;
    LF_HIP(hipStreamSynchronize(c->stream));
    c->sp_g = (const float *)d;
    c->sp_n = n;
    c->sp_faces = faces;
    return 0;
}

// First-order Mur on face (axis, hi) of the E component `comp`; faces apply in the order
// they are added.  m: n == 1 a scalar coefficient, else one per cell of the component's
// boundary plane (C order of the plane).  flags bit 0: F0_old is the plane's current
// value (the 2D inline form, fdtd2d_core.py:805-810) instead of its value before the step
// (boundaries3d.py:125-140); bit 1: cells of the plane on PEC faces where the component is
// tangential are zeroed after the update (the reference applies that PEC afterwards).
// Call after lf_set_cpml of every axis (the PEC faces are those of lf_set_cpml).
int lf_add_mur(lf_ctx *c, int comp, int axis, int hi, const float *m, long long n, int flags) {
    if (int r = check_ctx(c)) return r;
    const int ndim = c->kind == LF_YEE3D ? 3 : 2;
    for (int a = 0; a < ndim; ++a)
        if (!c->have_prof[a]) return fail(-1, "lf_add_mur: set the profiles of every axis first");
    if (axis < 0 || axis >= ndim) return fail(-1, "lf_add_mur: bad axis");
    const bool ok = c->kind == LF_YEE3D ? (comp >= LF_EX && comp <= LF_EZ)
                  : c->kind == LF_TE2D ? comp == LF_EZ : (comp == LF_EX || comp == LF_EY);
    if (!ok) return fail(-1, "lf_add_mur: not an E component of this kind");
    if (m == nullptr || n < 1) return fail(-1, "lf_add_mur: no coefficient");
    if (c->mur.size() >= 12) return fail(-1, "lf_add_mur: at most 12 faces");
    int ext[3];
    comp_extent(c, comp, ext[0], ext[1], ext[2]);
    const int sa = ndim == 3 ? axis : (axis == 0 ? 0 : 2);     // storage axis
    if (ext[sa] < 2) return fail(-1, "lf_add_mur: component has fewer than 2 planes");
    MurArgs a{};
    a.g = c->g;
    a.ax = sa;
    a.i0 = hi ? ext[sa] - 1 : 0;
    a.i1 = hi ? ext[sa] - 2 : 1;
    a.nu = sa == 0 ? ext[1] : ext[0];
    a.nv = sa == 2 ? ext[1] : ext[2];
    const long long np = (long long)a.nu * a.nv;
    if (n != 1 && n != np) return fail(-1, "lf_add_mur: coefficient count is neither 1 nor the plane size");
    a.f0cur = flags & 1;
    // logical direction of the component (2D TE Ez is normal to the plane of the grid)
    const int dir = ndim == 2 && comp == LF_EZ ? 2 : comp - LF_EX;
    for (int la = 0; la < ndim; ++la) {
        const int s2 = ndim == 3 ? la : (la == 0 ? 0 : 2);
        for (int lh = 0; lh < 2; ++lh)
            a.pz[s2][lh] = (flags & 2) && la != dir && c->pec[la][lh] ? 1 : 0;
    }
    void *d = nullptr;
    if (n == 1) {
        a.ms = m[0];
    } else {
        if (int r = upload_vec(c, m, (size_t)np * sizeof(float), &d)) return r;
        a.m = (const float *)d;
    }
    LF_HIP(hipMalloc(&d, (size_t)2 * np * sizeof(float)));
    c->ex_bufs.push_back(d);
    LF_HIP(hipMemsetAsync(d, 0, (size_t)2 * np * sizeof(float), c->stream));
    a.prev = (float *)d;
    LF_HIP(hipStreamSynchronize(c->stream));
    c->mur.push_back(a);
    c->mur_comp.push_back(comp);
    return 0;
}

}  // extern "C"


namespace {

int ensure_alt(lf_ctx *c, int id) {
    if (c->alt[id]) return 0;
    LF_HIP(hipMalloc(&c->alt[id], grid_elems(c) * sizeof(float)));
    LF_HIP(hipMemsetAsync(c->alt[id], 0, grid_elems(c) * sizeof(float), c->stream));
    c->tb_chain = false;
    return 0;
}

size_t strip_elems_2d(const lf_ctx *c, const Strip &s, bool x_slab) {
    return x_slab ? (size_t)s.len * c->g.P : (size_t)c->g.nx * s.len;
}

// one single-pass TM step: reads arr[], writes alt[], then swaps
int tm_fused_step(lf_ctx *c, const StepArgs &st, int fp64) {
    if (int r = ensure_alt(c, LF_EX)) return r;
    if (int r = ensure_alt(c, LF_EY)) return r;
    if (int r = ensure_alt(c, LF_HZ)) return r;
    for (int q = 0; q < 2; ++q)                       // psi_Hzx (x slab), psi_Hzy (y slab)
        if (c->s2[q].psi && !c->psi_alt[q]) {
            const size_t n = strip_elems_2d(c, c->s2[q], q == 0);
            LF_HIP(hipMalloc(&c->psi_alt[q], n * sizeof(float)));
            LF_HIP(hipMemsetAsync(c->psi_alt[q], 0, n * sizeof(float), c->stream));
            c->tb_chain = false;
        }
    TMFArgs f{};
    f.g = c->g;
    f.Ex0 = c->arr[LF_EX]; f.Ey0 = c->arr[LF_EY]; f.Hz0 = c->arr[LF_HZ];
    f.Ex1 = c->alt[LF_EX]; f.Ey1 = c->alt[LF_EY]; f.Hz1 = c->alt[LF_HZ];
    f.cax = c->arr[LF_CAX]; f.cay = c->arr[LF_CAY]; f.dax = c->arr[LF_DAX]; f.day = c->arr[LF_DAY];
    f.ch = c->ch;
    f.xe = prof_of(c, 0, false); f.xh = prof_of(c, 0, true);
    f.ye = prof_of(c, 1, false); f.yh = prof_of(c, 1, true);
    f.pHzx0 = c->s2[0]; f.pHzy0 = c->s2[1];
    f.pHzx1 = c->s2[0]; f.pHzx1.psi = c->psi_alt[0];
    f.pHzy1 = c->s2[1]; f.pHzy1.psi = c->psi_alt[1];
    f.pExy = c->s2[2]; f.pEyx = c->s2[3];
    f.pec_xlo = c->pec[0][0]; f.pec_xhi = c->pec[0][1];
    f.pec_ylo = c->pec[1][0]; f.pec_yhi = c->pec[1][1];
    f.lx = c->lx;
    f.nwy = (c->g.nz + 62) / 63;                      // 63 owned y cells per wavefront
    const long long nwx = (c->g.nx + c->lx - 1) / c->lx;
    const long long nwaves = (long long)f.nwy * nwx;
    tm_fused_kernel<<<(unsigned)((nwaves + 3) / 4), 256, 0, c->stream>>>(f, c->ex, st, fp64);
    std::swap(c->arr[LF_EX], c->alt[LF_EX]);
    std::swap(c->arr[LF_EY], c->alt[LF_EY]);
    std::swap(c->arr[LF_HZ], c->alt[LF_HZ]);
    for (int q = 0; q < 2; ++q)
        if (c->s2[q].psi) std::swap(c->s2[q].psi, c->psi_alt[q]);
    return 0;
}

}  // namespace

namespace {

StepArgs step_args(const lf_step &t) {
    return StepArgs{t.e32, t.wt32, t.ssin, t.scos, t.env, t.dr32, t.di32, t.dr, t.di};
}

// 2D psi slabs compressed along x: TM Hzx (0), Eyx (3); TE Hyx (1), Ezx (2); the others
// are compressed along y
bool x_slab_2d(int kind, int q) { return kind == LF_TM2D ? (q == 0 || q == 3) : (q == 1 || q == 2); }
bool tm_x_slab(int q) { return x_slab_2d(LF_TM2D, q); }

// the plan's view of a context
TBPlanIn tb_plan_in(const lf_ctx *c) {
    TBPlanIn in{c->g.nx, c->g.nz, 0, c->g.nx, 0, c->g.nz, -1, 0};
    for (int q = 0; q < 4; ++q) {
        if (!c->s2[q].psi) continue;
        if (x_slab_2d(c->kind, q)) { in.xlo = std::max(in.xlo, c->s2[q].lo); in.xhi = std::min(in.xhi, c->s2[q].hi); }
        else { in.ylo = std::max(in.ylo, c->s2[q].lo); in.yhi = std::min(in.yhi, c->s2[q].hi); }
    }
    const int comp = c->kind == LF_TM2D ? LF_EY : LF_EZ;
    if (c->ex.amp && c->ex.src_comp == comp) in.src_ix = c->ex.src_ix;
    if (c->ex.wtap && c->ex.drv[comp]) in.drv_n = c->ex.drv_n;
    return in;
}

TBPlan tb_plan_2d(const lf_ctx *c, int T, int lx) { return tb_plan_2d(tb_plan_in(c), T, lx); }

int ensure_side_streams(lf_ctx *c) {
    for (int u = 0; u < 3; ++u)
        if (!c->side[u]) LF_HIP(hipStreamCreate(&c->side[u]));
    if (!c->ev_fork) {
        LF_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        LF_HIP(hipEventCreateWithFlags(&c->ev_lead, hipEventDisableTiming));
        for (int u = 0; u < 3; ++u) LF_HIP(hipEventCreateWithFlags(&c->ev_join[u], hipEventDisableTiming));
    }
    return 0;
}

// Fork the classes of one 2D pass over the context stream and the three side streams,
// concurrently: disjoint output cells, step-n inputs.  `one` launches a class on a stream.
// Class `lead` goes on the context stream, ahead of the fork event the side streams wait
// for, so its waves are dispatched first.  Which order finishes soonest depends on the
// window: with an x-slab class of few long 256-VGPR waves it pays to start that class
// first (started behind the slab-free waves it found no room on the SIMDs until they
// drained), with a stitch drive over a quarter of the grid the drive classes are the
// long pole instead; lf_autotune measures the choice.
//
// Every class of a pass reads what every class of the previous pass wrote, so a pass waits
// for all of the previous one.  After a plain fork the side streams wait on the context
// stream, which waited on the side streams of the previous pass: two cross-queue hops
// per pass (about 10 us each).  When the previous queued work was a pass (tb_chain), each
// side stream instead waits on the previous pass's lead-class event and on the other side
// streams' events, one hop; its own stream orders its own previous class.
template <typename H, typename F>
int fork_tb_groups(lf_ctx *c, bool interior, int ng, int lead, int solo, H has, F one) {
    if (!interior) {
        one(solo, c->stream);
        c->tb_chain = false;
        return 0;
    }
    if (c->tb_lead == kTBSerial) {       // every group on the context stream, no events
        for (int u = 0; u < ng; ++u)
            if (has(u)) one(u, c->stream);
        c->tb_chain = false;
        return 0;
    }
    if (int e = ensure_side_streams(c)) return e;
    if (lead < 0 || lead >= ng || !has(lead)) lead = 0;
    int side_of[kTBClasses];
    for (int u = 0, n = 0; u < ng; ++u) side_of[u] = u == lead ? -1 : n++;
    const bool chain = c->tb_chain;
    if (!chain) LF_HIP(hipEventRecord(c->ev_fork, c->stream));
    for (int u = 0; u < ng; ++u) {
        const int sd = side_of[u];
        if (sd < 0 || !has(u)) continue;
        if (!chain) {
            LF_HIP(hipStreamWaitEvent(c->side[sd], c->ev_fork, 0));
            continue;
        }
        LF_HIP(hipStreamWaitEvent(c->side[sd], c->ev_lead, 0));
        for (int v = 0; v < 3; ++v)
            if (v != sd && ((c->tb_chain_sides >> v) & 1u))
                LF_HIP(hipStreamWaitEvent(c->side[sd], c->ev_join[v], 0));
    }
    one(lead, c->stream);
    LF_HIP(hipEventRecord(c->ev_lead, c->stream));
    for (int u = 0; u < ng; ++u)
        if (side_of[u] >= 0 && has(u)) one(u, c->side[side_of[u]]);
    unsigned sides = 0;
    for (int u = 0; u < ng; ++u) {
        if (side_of[u] < 0 || !has(u)) continue;
        LF_HIP(hipEventRecord(c->ev_join[side_of[u]], c->side[side_of[u]]));
        LF_HIP(hipStreamWaitEvent(c->stream, c->ev_join[side_of[u]], 0));
        sides |= 1u << side_of[u];
    }
    c->tb_chain = true;
    c->tb_chain_sides = sides;
    return 0;
}

// the four classes of a plan, one launch each (the monitored passes)
template <typename F>
int fork_tb_classes(lf_ctx *c, const TBPlan &pl, int lead, F one) {
    return fork_tb_groups(
        c, pl.interior, kTBClasses, lead, 2, [&](int u) { return pl.cls[u].nseg > 0; },
        [&](int u, hipStream_t st) { one(pl.cls[u], st); });
}

void launch_tb_mode(const TBTMArgs &q, int T, int mode, int exc, const Excite &x, int fp64, hipStream_t st) {
    const unsigned nb = (unsigned)(((long long)q.seg_end[q.nseg - 1] + 3) / 4);
    launch_tm_tb(T, mode, exc, q, x, fp64, st, nb);
}
void launch_tb_mode(const TBTEArgs &q, int T, int mode, int exc, const Excite &x, int fp64, hipStream_t st) {
    const unsigned nb = (unsigned)(((long long)q.seg_end[q.nseg - 1] + 3) / 4);
    launch_te_tb(T, mode, exc, q, x, fp64, st, nb);
}

// Launch the classes of one blocked pass.  Where the slab classes carry no excitation code,
// or only the drive with the float64 blend (the stitch windows after the first), both slab
// classes (1 and 2) go in one slab launch: their bodies together still hold two waves per
// SIMD, and the pass costs the host one launch and one stream's events and waits less
// (queuing a pass takes the host about as long as the device needs for it).  A lead class
// 1 or 2 then leads with the slab launch.  Otherwise the four classes launch apart (the
// source body beside the x-slab body would drop the slab launch to one wave per SIMD).
// Without a slab-free rectangle only the slab classes run, on the context stream.
template <typename A>
int launch_tb_pass(lf_ctx *c, A &f, const TBPlan &pl, int T, int fp64) {
    auto one_class = [&](const TBClass &k, hipStream_t st) {
        if (k.nseg == 0) return;
        A q = f;
        set_class(q, k);
        launch_tb_mode(q, T, k.mode, k.exc, c->ex, fp64, st);
    };
    const bool slab_exc = pl.cls[1].exc || pl.cls[2].exc;
    const bool merge = !slab_exc || (c->ex.amp == nullptr && c->ex.drv_math == LF_DRIVE_C128);
    if (!merge) return fork_tb_classes(c, pl, c->tb_lead, one_class);
    auto has = [&](int u) {
        return u == 1 ? pl.cls[1].nseg + pl.cls[2].nseg > 0 : pl.cls[u == 0 ? 0 : 3].nseg > 0;
    };
    auto one = [&](int u, hipStream_t st) {
        if (u != 1) {
            one_class(pl.cls[u == 0 ? 0 : 3], st);
            return;
        }
        if (!has(1)) return;
        A q = f;
        set_slabs(q, pl.cls[1], pl.cls[2]);
        const unsigned nb = (unsigned)(((long long)q.seg_end[q.nseg - 1] + 3) / 4);
        if constexpr (std::is_same_v<A, TBTMArgs>) launch_tm_tb_slab(T, slab_exc ? 1 : 0, q, c->ex, fp64, st, nb);
        else launch_te_tb_slab(T, slab_exc ? 1 : 0, q, c->ex, fp64, st, nb);
    };
    const int lead = c->tb_lead == 0 ? 0 : c->tb_lead == 3 ? 2 : 1;
    return fork_tb_groups(c, pl.interior, 3, lead, 1, has, one);
}

int ensure_psi_alt_2d(lf_ctx *c) {
    for (int q = 0; q < 4; ++q)
        if (c->s2[q].psi && !c->psi_alt[q]) {
            const size_t n = strip_elems_2d(c, c->s2[q], x_slab_2d(c->kind, q));
            LF_HIP(hipMalloc(&c->psi_alt[q], n * sizeof(float)));
            LF_HIP(hipMemsetAsync(c->psi_alt[q], 0, n * sizeof(float), c->stream));
            c->tb_chain = false;
        }
    return 0;
}

TBTMArgs tm_tb_args(lf_ctx *c) {
    TBTMArgs f{};
    f.g = c->g;
    f.Ex0 = c->arr[LF_EX]; f.Ey0 = c->arr[LF_EY]; f.Hz0 = c->arr[LF_HZ];
    f.Ex1 = c->alt[LF_EX]; f.Ey1 = c->alt[LF_EY]; f.Hz1 = c->alt[LF_HZ];
    f.cax = c->arr[LF_CAX]; f.cay = c->arr[LF_CAY]; f.dax = c->arr[LF_DAX]; f.day = c->arr[LF_DAY];
    f.kax = coef_src(c, LF_CAX); f.kay = coef_src(c, LF_CAY);
    f.kdx = coef_src(c, LF_DAX); f.kdy = coef_src(c, LF_DAY);
    f.cm = coef_map(c);
    f.ch = c->ch;
    f.xe = prof_of(c, 0, false); f.xh = prof_of(c, 0, true);
    f.ye = prof_of(c, 1, false); f.yh = prof_of(c, 1, true);
    for (int q = 0; q < 4; ++q) { f.s0[q] = c->s2[q]; f.s1[q] = c->s2[q]; f.s1[q].psi = c->psi_alt[q]; }
    f.pec_xlo = c->pec[0][0]; f.pec_xhi = c->pec[0][1];
    f.pec_ylo = c->pec[1][0]; f.pec_yhi = c->pec[1][1];
    return f;
}

TBTEArgs te_tb_args(lf_ctx *c) {
    TBTEArgs f{};
    f.g = c->g;
    f.Ez0 = c->arr[LF_EZ]; f.Hx0 = c->arr[LF_HX]; f.Hy0 = c->arr[LF_HY];
    f.Ez1 = c->alt[LF_EZ]; f.Hx1 = c->alt[LF_HX]; f.Hy1 = c->alt[LF_HY];
    f.ceE = c->arr[LF_CEE]; f.ceH = c->arr[LF_CEH];
    f.kE = coef_src(c, LF_CEE); f.kH = coef_src(c, LF_CEH);
    f.cm = coef_map(c);
    f.ch = c->ch;
    f.xe = prof_of(c, 0, false); f.xh = prof_of(c, 0, true);
    f.ye = prof_of(c, 1, false); f.yh = prof_of(c, 1, true);
    for (int q = 0; q < 4; ++q) { f.s0[q] = c->s2[q]; f.s1[q] = c->s2[q]; f.s1[q].psi = c->psi_alt[q]; }
    f.pec_xlo = c->pec[0][0]; f.pec_xhi = c->pec[0][1];
    f.pec_ylo = c->pec[1][0]; f.pec_yhi = c->pec[1][1];
    return f;
}

// fields of the blocked passes (ping-pong) per kind
void swap_tb_buffers(lf_ctx *c) {
    if (c->kind == LF_TM2D) {
        std::swap(c->arr[LF_EX], c->alt[LF_EX]);
        std::swap(c->arr[LF_EY], c->alt[LF_EY]);
        std::swap(c->arr[LF_HZ], c->alt[LF_HZ]);
    } else {
        std::swap(c->arr[LF_EZ], c->alt[LF_EZ]);
        std::swap(c->arr[LF_HX], c->alt[LF_HX]);
        std::swap(c->arr[LF_HY], c->alt[LF_HY]);
    }
    for (int q = 0; q < 4; ++q)
        if (c->s2[q].psi) std::swap(c->s2[q].psi, c->psi_alt[q]);
}

int ensure_tb_buffers(lf_ctx *c) {
    const int ids[2][3] = {{LF_EZ, LF_HX, LF_HY}, {LF_EX, LF_EY, LF_HZ}};
    for (int q = 0; q < 3; ++q)
        if (int r = ensure_alt(c, ids[c->kind == LF_TM2D][q])) return r;
    return ensure_psi_alt_2d(c);
}

// one temporally blocked 2D pass over T steps: reads arr[], writes alt[], then swaps
int tb_chunk_2d(lf_ctx *c, const lf_step *steps, int T, int fp64) {
    if (T != 2 && T != 4 && T != 6 && T != 8)
        return fail(-1, "temporal block depth must be 2, 4, 6 or 8");
    if (int r = ensure_tb_buffers(c)) return r;
    const TBPlan pl = tb_plan_2d(c, T, c->tb_lx);
    if (c->kind == LF_TM2D) {
        TBTMArgs f = tm_tb_args(c);
        for (int s = 0; s < T; ++s) f.st[s] = step_args(steps[s]);
        if (int e = launch_tb_pass(c, f, pl, T, fp64)) return e;
    } else {
        TBTEArgs f = te_tb_args(c);
        for (int s = 0; s < T; ++s) f.st[s] = step_args(steps[s]);
        if (int e = launch_tb_pass(c, f, pl, T, fp64)) return e;
    }
    swap_tb_buffers(c);
    return 0;
}

// ---- monitored 2D passes: the CW-window monitors summed inside the pass
struct TBMonPlan {
    int mon = 0;                      // kernel variant: 9 (one phasor) or 14 (two + intensity), +16 all complex128, +32 TM stitch layout
    TBMon m{};
    int nflux = 0;
    int fid[kTBFlux] = {};            // flux monitor ids in staging order
};

// Can every monitor set on steps[0..T) be summed inside a 2D pass?  Phasors of 2D
// sources (TM: Ex, Ey, node Ey, node Ex; TE: Ez), one intensity sum of a 2D form and up
// to kTBFlux flux planes qualify; detector boxes do not.
bool tb_mon_plan(const lf_ctx *c, const lf_step *steps, int T, TBMonPlan &pl) {
    unsigned any = 0u;
    for (int u = 0; u < T; ++u) any |= steps[u].mask;
    if (any == 0u) return false;
    int did[2] = {-1, -1}, nd = 0, sid = -1;
    pl = TBMonPlan{};
    for (int id = 0; id < c->nmon; ++id) {
        if (!((any >> id) & 1u)) continue;
        const Monitor &mo = c->mon[id];
        if (mo.type == 2) {
            const int src = mo.dft.src;
            const bool ok = c->kind == LF_TM2D
                                ? (src == LF_SRC_EX || src == LF_SRC_EY || src == LF_SRC_EY_NODE ||
                                   src == LF_SRC_EX_NODE)
                                : src == LF_SRC_EZ;
            if (!ok || mo.dft.kplane >= 0 || mo.dft.n1 != 1 || nd == 2) return false;
            did[nd++] = id;
        } else if (mo.type == 1) {
            const bool ok = c->kind == LF_TM2D ? (mo.form == LF_SQ_TM || mo.form == LF_SQ_TM_EX2)
                                               : mo.form == LF_SQ_TE;
            if (!ok || sid >= 0) return false;
            sid = id;
        } else if (mo.type == 3) {
            if (pl.nflux == kTBFlux || mo.flux.npair != 1 || mo.flux.n1[0] != 1) return false;
            pl.fid[pl.nflux++] = id;
        } else {
            return false;
        }
    }
    pl.mon = (nd <= 1 && sid < 0) ? 9 : 14;
    TBMon &m = pl.m;
    for (int k = 0; k < 2; ++k) {
        if (k >= nd) { m.dn2[k] = 0; continue; }
        const DftArgs &d = c->mon[did[k]].dft;
        m.dsrc[k] = d.src; m.dacc[k] = d.acc;
        m.dx0[k] = d.x0; m.dx1[k] = d.x0 + d.n0; m.dn2[k] = d.n2;
        m.dbuf[k] = d.buf;
    }
    bool all128 = nd > 0;
    for (int k = 0; k < nd; ++k) all128 = all128 && m.dacc[k] == LF_ACC_C128;
    if (all128) pl.mon |= 16;            // kernels without the accumulator-format tests
    if (sid >= 0) {
        const Monitor &mo = c->mon[sid];
        m.sform = mo.form; m.sn2 = mo.n2; m.sbuf = (float *)mo.buf;
    }
    m.nflux = pl.nflux;
    for (int f = 0; f < pl.nflux; ++f) {
        const FluxArgs &fa = c->mon[pl.fid[f]].flux;
        m.fil[f] = fa.il; m.fn2[f] = fa.n2[0];
        m.fst[f] = c->tbm_stage ? c->tbm_stage + (size_t)f * kMaxTB * 2 * c->g.P : nullptr;
    }
    for (int u = 0; u < T; ++u) {
        const lf_step &t = steps[u];
        unsigned b = 0u;
        for (int k = 0; k < nd; ++k) b |= ((t.mask >> did[k]) & 1u) << k;
        if (sid >= 0) b |= ((t.mask >> sid) & 1u) << 2;
        for (int f = 0; f < pl.nflux; ++f) b |= ((t.mask >> pl.fid[f]) & 1u) << (3 + f);
        m.lm[u] = b;
        m.pr[u] = t.pr; m.pi[u] = t.pi;
        for (int k = 0; k < 2; ++k) {
            const bool q = k < nd && c->mon[did[k]].dft.phasor == LF_PH_Q;
            m.fr[k][u] = q ? t.qr32 : t.pr32;
            m.fi[k][u] = q ? t.qi32 : t.pi32;
        }
    }
    // the CW-window layout of the TM runs (fdtd2d_core.py:838-842 with the Ex^2 sum of 812):
    // Ey and node-Ex phasors in complex128 and the Ex^2 sum on every step; its kernels test
    // nothing per level
    if (c->kind == LF_TM2D && (pl.mon & 16) && nd == 2 && m.dsrc[0] == LF_SRC_EY &&
        m.dsrc[1] == LF_SRC_EX_NODE && sid >= 0 && m.sform == LF_SQ_TM_EX2) {
        bool every = true;
        for (int u = 0; u < T; ++u) every = every && (m.lm[u] & 7u) == 7u;
        if (every) pl.mon |= 32;
    }
    return true;
}

// Flux planes staged by a monitored pass, added in step order with the arithmetic of
// flux_kernel (2D: one pair per plane, one row of cells).
struct TBFluxPost {
    int nflux, T, wide;
    int n2[kTBFlux];
    unsigned steps[kTBFlux];          // bit s: the plane accumulates on level s
    const float *st[kTBFlux];
    void *aE[kTBFlux], *aH[kTBFlux];
    double pr[kMaxTB], pi[kMaxTB], hr[kMaxTB], hi[kMaxTB];
    float pr32[kMaxTB], pi32[kMaxTB], hr32[kMaxTB], hi32[kMaxTB];
};

__global__ void __launch_bounds__(256) tb_flux_kernel(TBFluxPost f) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    if (k >= f.nflux || j >= f.n2[k]) return;
    const int n2 = f.n2[k];
    if (f.wide) {
        double2 *aE = (double2 *)f.aE[k];
        double2 *aH = (double2 *)f.aH[k];
        double2 u = aE[j], w = aH[j];
        for (int s = 0; s < f.T; ++s) {
            if (!((f.steps[k] >> s) & 1u)) continue;
            const float e = f.st[k][(2 * s) * n2 + j];
            const float h = f.st[k][(2 * s + 1) * n2 + j];
            u.x = u.x + (double)e * f.pr[s]; u.y = u.y + (double)e * f.pi[s];
            w.x = w.x + (double)h * f.hr[s]; w.y = w.y + (double)h * f.hi[s];
        }
        aE[j] = u; aH[j] = w;
    } else {
        float2 *aE = (float2 *)f.aE[k];
        float2 *aH = (float2 *)f.aH[k];
        float2 u = aE[j], w = aH[j];
        for (int s = 0; s < f.T; ++s) {
            if (!((f.steps[k] >> s) & 1u)) continue;
            const float e = f.st[k][(2 * s) * n2 + j];
            const float h = f.st[k][(2 * s + 1) * n2 + j];
            u.x = u.x + e * f.pr32[s]; u.y = u.y + e * f.pi32[s];
            w.x = w.x + h * f.hr32[s]; w.y = w.y + h * f.hi32[s];
        }
        aE[j] = u; aH[j] = w;
    }
}

// one monitored 2D pass over T steps (T in 2..4): reads arr[], writes alt[], sums the
// planned monitors, then swaps
int tb_chunk_2d_mon(lf_ctx *c, const lf_step *steps, int T, TBMonPlan &mp, int fp64) {
    if (T < 2 || T > 4) return fail(-1, "monitored pass depth must be 2, 3 or 4");
    if (int r = ensure_tb_buffers(c)) return r;
    if (mp.nflux > 0 && !c->tbm_stage) {
        const size_t n = (size_t)kTBFlux * kMaxTB * 2 * c->g.P;
        LF_HIP(hipMalloc(&c->tbm_stage, n * sizeof(float)));
        LF_HIP(hipMemsetAsync(c->tbm_stage, 0, n * sizeof(float), c->stream));
        c->tb_chain = false;
        for (int f = 0; f < mp.nflux; ++f) mp.m.fst[f] = c->tbm_stage + (size_t)f * kMaxTB * 2 * c->g.P;
    }
    const TBPlan pl = tb_plan_2d(c, T, c->tb_lx);
    auto pass = [&](auto &f, auto launch) -> int {
        for (int s = 0; s < T; ++s) f.st[s] = step_args(steps[s]);
        auto one = [&](const TBClass &k, hipStream_t st) {
            if (k.nseg == 0) return;
            auto q = f;
            set_class(q, k);
            const unsigned nb = (unsigned)(((long long)q.seg_end[q.nseg - 1] + 3) / 4);
            launch(T, k.mode, k.exc != 0, mp.mon, q, c->ex, mp.m, fp64, st, nb);
        };
        return fork_tb_classes(c, pl, 0, one);
    };
    if (c->kind == LF_TM2D) {
        TBTMArgs f = tm_tb_args(c);
        if (int e = pass(f, launch_tm_tb_mon)) return e;
    } else {
        TBTEArgs f = te_tb_args(c);
        if (int e = pass(f, launch_te_tb_mon)) return e;
    }
    LF_HIP(hipGetLastError());
    if (mp.nflux > 0) {
        TBFluxPost fp{};
        fp.nflux = mp.nflux; fp.T = T; fp.wide = c->mon[mp.fid[0]].flux.wide;
        int n2max = 0;
        for (int f = 0; f < mp.nflux; ++f) {
            const FluxArgs &fa = c->mon[mp.fid[f]].flux;
            fp.n2[f] = fa.n2[0];
            n2max = std::max(n2max, fa.n2[0]);
            fp.st[f] = mp.m.fst[f];
            fp.aE[f] = fa.aE[0]; fp.aH[f] = fa.aH[0];
            unsigned b = 0u;
            for (int s = 0; s < T; ++s) b |= ((mp.m.lm[s] >> (3 + f)) & 1u) << s;
            fp.steps[f] = b;
        }
        for (int s = 0; s < T; ++s) {
            const lf_step &t = steps[s];
            fp.pr[s] = t.pr; fp.pi[s] = t.pi; fp.hr[s] = t.hr; fp.hi[s] = t.hi;
            fp.pr32[s] = t.pr32; fp.pi32[s] = t.pi32; fp.hr32[s] = t.hr32; fp.hi32[s] = t.hi32;
        }
        tb_flux_kernel<<<dim3((n2max + 255) / 256, mp.nflux, 1), 256, 0, c->stream>>>(fp);
        LF_HIP(hipGetLastError());
    }
    swap_tb_buffers(c);
    return 0;
}

}  // namespace

extern "C" int lf_set_temporal_mon(lf_ctx *c, int depth) {
    if (int e = check_ctx(c)) return e;
    if (c->kind == LF_YEE3D) {
        if (depth != 1) return fail(-1, "lf_set_temporal_mon: monitored passes are 2D only");
    } else if (depth < 1 || depth > 4) {
        return fail(-1, "lf_set_temporal_mon: depth must be 1 (off), 2, 3 or 4");
    }
    c->tb_mon = depth;
    return 0;
}

extern "C" int lf_set_temporal(lf_ctx *c, int depth, int lx) {
    if (int e = check_ctx(c)) return e;
    if (c->kind == LF_YEE3D) {
        if (depth < 1 || depth > 3) return fail(-1, "lf_set_temporal: 3D depth must be 1 (off), 2 or 3");
    } else if (depth != 1 && depth != 2 && depth != 4 && depth != 6 && depth != 8) {
        return fail(-1, "lf_set_temporal: depth must be 1 (off), 2, 4, 6 or 8");
    }
    if (lx < 1 || lx > 1 << 20) return fail(-1, "lf_set_temporal: lx out of range");
    c->tb = depth;
    c->tb_lx = lx;
    return 0;
}

extern "C" int lf_set_tb_lead(lf_ctx *c, int cls) {
    if (int e = check_ctx(c)) return e;
    if (cls < 0 || cls > kTBSerial) return fail(-1, "lf_set_tb_lead: class must be 0, 1, 2, 3 or 4");
    if (c->kind == LF_YEE3D && cls != 0) return fail(-1, "lf_set_tb_lead: 3D passes take class 0 only");
    c->tb_lead = cls;
    return 0;
}

extern "C" int lf_tb_stats(lf_ctx *c, double *ms, int *launches, int *depth) {
    if (int e = check_ctx(c)) return e;
    *ms = c->last_tb_ms;
    *launches = c->last_tb_launches;
    *depth = c->tb;
    return 0;
}

extern "C" int lf_set_variant(lf_ctx *c, int variant, int lx) {
    if (int e = check_ctx(c)) return e;
    if (variant < 0 || variant > 1) return fail(-1, "lf_set_variant: variant 0 or 1");
    if (lx < 1 || lx > 4096) return fail(-1, "lf_set_variant: lx out of range");
    c->variant = variant;
    c->lx = lx;
    return 0;
}


namespace {

size_t strip_elems_3d(const lf_ctx *c, const Strip &s, int axis) {
    long long d[3] = {c->g.nx, c->g.ny, c->g.P};
    d[axis] = s.len;
    return (size_t)(d[0] * d[1] * d[2]);
}

int yee_fused_step(lf_ctx *c, const StepArgs &st, int fp64) {
    const int ids[6] = {LF_EX, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ};
    for (int q = 0; q < 6; ++q)
        if (int r = ensure_alt(c, ids[q])) return r;
    const int hax[6] = {2, 1, 0, 2, 1, 0};            // stretch axis of Hx_z Hx_y Hy_x Hy_z Hz_y Hz_x
    for (int q = 0; q < 6; ++q)
        if (c->sH[q].psi && !c->psi_alt3[q]) {
            const size_t n = strip_elems_3d(c, c->sH[q], hax[q]);
            LF_HIP(hipMalloc(&c->psi_alt3[q], n * sizeof(float)));
            LF_HIP(hipMemsetAsync(c->psi_alt3[q], 0, n * sizeof(float), c->stream));
        }
    SPArgs f{};
    f.g = c->g;
    f.Ex0 = c->arr[LF_EX]; f.Ey0 = c->arr[LF_EY]; f.Ez0 = c->arr[LF_EZ];
    f.Hx0 = c->arr[LF_HX]; f.Hy0 = c->arr[LF_HY]; f.Hz0 = c->arr[LF_HZ];
    f.Ex1 = c->alt[LF_EX]; f.Ey1 = c->alt[LF_EY]; f.Ez1 = c->alt[LF_EZ];
    f.Hx1 = c->alt[LF_HX]; f.Hy1 = c->alt[LF_HY]; f.Hz1 = c->alt[LF_HZ];
    f.cex = c->arr[LF_CEX]; f.cey = c->arr[LF_CEY]; f.cez = c->arr[LF_CEZ];
    f.dax = c->arr[LF_DAX]; f.day = c->arr[LF_DAY]; f.daz = c->arr[LF_DAZ];
    f.ch = c->ch;
    for (int ax = 0; ax < 3; ++ax) { f.pe[ax] = prof_of(c, ax, false); f.ph[ax] = prof_of(c, ax, true); }
    for (int q = 0; q < 6; ++q) {
        f.sH0[q] = c->sH[q];
        f.sH1[q] = c->sH[q];
        f.sH1[q].psi = c->psi_alt3[q];
        f.sE[q] = c->sE[q];
    }
    for (int ax = 0; ax < 3; ++ax) { f.pec[2 * ax] = c->pec[ax][0]; f.pec[2 * ax + 1] = c->pec[ax][1]; }
    f.lx = c->lx;
    if ((long long)c->g.nx * c->g.ny * c->g.P >= (1LL << 31))
        return fail(-1, "single-pass 3D kernel: grid exceeds 2^31 elements (use variant 1)");
    const int R = c->ty == 2 ? 2 : 4;                 // owned y rows per wavefront
    const int NX = c->g.nx, NY = c->g.ny, NZ = c->g.nz;
    const int nw[3] = {(NZ + 62) / 63, (NY + R - 1) / R, (NX + f.lx - 1) / f.lx};   // z, y, x
    // per axis (z, y, x): positions outside every psi slab of that axis and inside 1..N-2
    const int N3[3] = {NZ, NY, NX};
    const int axs[3][4] = {{0, 3, 1, 2}, {1, 4, 0, 5}, {2, 5, 3, 4}};   // sH, sH, sE, sE
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = 1; hi[d] = N3[d] - 2;
        for (int u = 0; u < 4; ++u) {
            const Strip &sp = u < 2 ? c->sH[axs[d][u]] : c->sE[axs[d][u]];
            if (!sp.psi) continue;
            lo[d] = std::max(lo[d], sp.lo);
            hi[d] = std::min(hi[d], sp.hi - 1);
        }
    }
    // interval [ia, ib] of wave indices per axis whose touched positions are all interior
    int ia3[3], ib3[3];
    for (int d = 0; d < 3; ++d) {
        ia3[d] = nw[d]; ib3[d] = -1;
        for (int w = 0; w < nw[d]; ++w) {
            int u0, u1;
            if (d == 0) { u0 = w * 63 - 1; u1 = w * 63 + 63; }
            else if (d == 1) { u0 = w * R - 1; u1 = w * R + R; }
            else { u0 = w * f.lx - 1; u1 = std::min(w * f.lx + f.lx, NX); }
            if (u0 >= lo[d] && u1 <= hi[d]) { ia3[d] = std::min(ia3[d], w); ib3[d] = std::max(ib3[d], w); }
        }
    }
    if (int e = ensure_side_streams(c)) return e;
    LF_HIP(hipEventRecord(c->ev_fork, c->stream));
    for (int u = 0; u < 2; ++u) LF_HIP(hipStreamWaitEvent(c->side[u], c->ev_fork, 0));
    for (int mode = 7; mode >= 0; --mode) {
        // intervals per axis: bit clear -> interior interval, bit set -> below / above it
        int iv[3][2][2], niv[3];
        for (int d = 0; d < 3; ++d) {
            const bool edge = mode & (1 << d);
            niv[d] = 0;
            const bool has_int = ib3[d] >= ia3[d];
            if (!edge) {
                if (has_int) { iv[d][0][0] = ia3[d]; iv[d][0][1] = ib3[d] - ia3[d] + 1; niv[d] = 1; }
            } else if (!has_int) {
                iv[d][0][0] = 0; iv[d][0][1] = nw[d]; niv[d] = 1;
            } else {
                if (ia3[d] > 0) { iv[d][niv[d]][0] = 0; iv[d][niv[d]][1] = ia3[d]; ++niv[d]; }
                if (ib3[d] < nw[d] - 1) {
                    iv[d][niv[d]][0] = ib3[d] + 1; iv[d][niv[d]][1] = nw[d] - 1 - ib3[d]; ++niv[d];
                }
            }
        }
        SPArgs q = f;
        q.nbox = 0;
        for (int az = 0; az < niv[0]; ++az)
            for (int ay = 0; ay < niv[1]; ++ay)
                for (int ax = 0; ax < niv[2]; ++ax) {
                    const int b = q.nbox++;
                    q.bz0[b] = iv[0][az][0]; q.bnz[b] = iv[0][az][1];
                    q.by0[b] = iv[1][ay][0]; q.bny[b] = iv[1][ay][1];
                    q.bx0[b] = iv[2][ax][0];
                    const int cnt = iv[0][az][1] * iv[1][ay][1] * iv[2][ax][1];
                    q.bend[b] = (b ? q.bend[b - 1] : 0) + cnt;
                }
        if (q.nbox == 0) continue;
        const unsigned nb = (unsigned)(((long long)q.bend[q.nbox - 1] + 3) / 4);
        hipStream_t sm = mode == 0 ? c->stream : (mode == 1 ? c->side[0] : c->side[1]);
#define LF_SP_CASE(MM)                                                                        \
    case MM:                                                                                \
        if (R == 2) yee_sp_kernel<2, MM><<<nb, 256, 0, sm>>>(q, c->ex, st, fp64);             \
        else yee_sp_kernel<4, MM><<<nb, 256, 0, sm>>>(q, c->ex, st, fp64);                    \
        break;
        switch (mode) {
            LF_SP_CASE(0) LF_SP_CASE(1) LF_SP_CASE(2) LF_SP_CASE(3)
            LF_SP_CASE(4) LF_SP_CASE(5) LF_SP_CASE(6) LF_SP_CASE(7)
        }
#undef LF_SP_CASE
    }
    for (int u = 0; u < 2; ++u) {
        LF_HIP(hipEventRecord(c->ev_join[u], c->side[u]));
        LF_HIP(hipStreamWaitEvent(c->stream, c->ev_join[u], 0));
    }
    for (int q = 0; q < 6; ++q) std::swap(c->arr[ids[q]], c->alt[ids[q]]);
    for (int q = 0; q < 6; ++q)
        if (c->sH[q].psi) std::swap(c->sH[q].psi, c->psi_alt3[q]);
    return 0;
}

}  // namespace

namespace {

// Detector boxes the 3D blocked pass sums itself: the first two box monitors, when the
// tile plan fits their per-tile slots.  Returns their monitor bits.
unsigned tb3_fused_boxes(const lf_ctx *c, int *ids) {
    unsigned bits = 0;
    int n = 0;
    for (int id = 0; id < c->nmon && n < 2; ++id)
        if (c->mon[id].type == 0) { ids[n++] = id; bits |= 1u << id; }
    return bits;
}

// Tile plan of the 3D blocked pass for depth T and x chunk lx.  A tile owns 64 - 2T z
// lanes, kTB3W * rows - 2T y rows and one x chunk, and loads a T-cell halo around them.
// Per axis the owned positions split into at most three intervals: the middle one, whose
// tiles load only positions in 1..N-2 outside every psi slab of that axis, and the
// intervals below and above it (z: whole tiles of a uniform tiling, so that the tiles
// meeting a z slab still carry interior lanes; y and x: position intervals, the x slab
// intervals being one short chunk each).  The middle x interval further splits into runs
// of chunks whose march meets a source or drive plane and runs that do not.  Every
// combination is one box; its class bits say along which axes its tiles may meet a slab
// or an edge.  Slab-free boxes without excitation run c->tb3_rows0 rows per wavefront,
// the others one.
int tb3_plan(lf_ctx *c, int T, int lx) {
    // planes where the E update adds the source pair or the stitch drive
    std::vector<std::pair<int, int>> exc;                 // [a, b) plane ranges
    if (c->ex.amp && (c->ex.src_comp == LF_EY || c->ex.src_comp == LF_EZ))
        exc.push_back({c->ex.src_ix, c->ex.src_ix + 2});
    if (c->ex.wtap && (c->ex.drv[0] || c->ex.drv[1] || c->ex.drv[2])) exc.push_back({0, c->ex.drv_n});
    // x ranges of the detector boxes the pass sums (their owned planes need the box code)
    std::vector<std::pair<int, int>> boxx;
    int fid[2];
    const unsigned fbits = tb3_fused_boxes(c, fid);
    for (int m = 0; m < __builtin_popcount(fbits); ++m) {
        const BoxArgs &b = c->mon[fid[m]].box;
        for (int q = 0; q < 3; ++q) boxx.push_back({b.lo[q][0], b.hi[q][0]});
    }
    const int key_src = c->ex.amp ? c->ex.src_ix : -1, key_drv = c->ex.wtap ? c->ex.drv_n : -1;
    if (T < 1 || T > 3) return fail(-1, "3D blocked pass: depth must be 1, 2 or 3");
    if (c->tb3_dbox && c->tb3_T == T && c->tb3_lx == lx && c->tb3_key[0] == key_src &&
        c->tb3_key[1] == key_drv && c->tb3_key[2] == c->tb3_rows0 && c->tb3_key[3] == c->nmon)
        return 0;
    lf_ctx::TB3Saved &sv = c->tb3_saved[T];
    if (sv.valid && sv.lx == lx && sv.key[0] == key_src && sv.key[1] == key_drv &&
        sv.key[2] == c->tb3_rows0 && sv.key[3] == c->nmon) {
        c->tb3_boxes = sv.boxes;
        c->tb3_dbox = sv.dbox;
        c->tb3_ntile = sv.ntile;
        c->tb3_nwg = sv.nwg;
        c->tb3_T = T;
        c->tb3_lx = lx;
        for (int q = 0; q < 4; ++q) c->tb3_key[q] = sv.key[q];
        return 0;
    }
    const int N3[3] = {c->g.nz, c->g.ny, c->g.nx};                   // z, y, x
    const int axs[3][4] = {{0, 3, 1, 2}, {1, 4, 0, 5}, {2, 5, 3, 4}};   // sH, sH, sE, sE
    struct Iv { int a, b; bool edge, exc, box; };
    std::vector<Iv> iv[3];
    const int WZ = 64 - 2 * T;
    for (int d = 0; d < 3; ++d) {
        const int N = N3[d];
        int lo = 1, hi = N - 2;
        for (int u = 0; u < 4; ++u) {
            const Strip &sp = u < 2 ? c->sH[axs[d][u]] : c->sE[axs[d][u]];
            if (!sp.psi) continue;
            lo = std::max(lo, sp.lo);
            hi = std::min(hi, sp.hi - 1);
        }
        int A, B;                                         // middle interval [A, B)
        if (d == 0) {
            const int nt = (N + WZ - 1) / WZ;
            int ta = nt, tb = -1;
            for (int t = 0; t < nt; ++t)
                if (t * WZ - T >= lo && std::min((t + 1) * WZ, N) + T - 1 <= hi) {
                    ta = std::min(ta, t);
                    tb = std::max(tb, t);
                }
            A = ta * WZ;
            B = tb >= ta ? std::min((tb + 1) * WZ, N) : A;
        } else {
            A = lo + T;
            B = hi - T + 1;
        }
        if (B <= A) {
            iv[d].push_back({0, N, true, false, false});
            continue;
        }
        if (A > 0) iv[d].push_back({0, A, true, false, false});
        if (d < 2) {
            iv[d].push_back({A, B, false, false, false});
        } else {
            // x: the planes of [A, B) by the code their tiles need: 2 excitation (planes
            // whose chunk march can meet a source / drive plane: within T of it), 1 box
            // sums (owned planes of a fused box), 0 neither; runs shorter than 16 planes
            // join the heavier neighbour (a heavier body is a superset), then each run is
            // one interval
            std::vector<int> cls(B - A, 0);
            for (auto &e : exc)
                for (int q = std::max(e.first - T, A); q < std::min(e.second + T, B); ++q) cls[q - A] = 2;
            for (auto &e : boxx)
                for (int q = std::max(e.first, A); q < std::min(e.second, B); ++q)
                    cls[q - A] = std::max(cls[q - A], 1);
            for (bool changed = true; changed;) {
                changed = false;
                int r0 = 0;
                while (r0 < B - A) {
                    int r1 = r0;
                    while (r1 < B - A && cls[r1] == cls[r0]) ++r1;
                    if (r1 - r0 < 16 && (r0 > 0 || r1 < B - A)) {
                        const int left = r0 > 0 ? cls[r0 - 1] : -1, right = r1 < B - A ? cls[r1] : -1;
                        const int to = std::max(left, right);
                        if (to > cls[r0]) {
                            for (int q = r0; q < r1; ++q) cls[q] = to;
                            changed = true;
                        }
                    }
                    r0 = r1;
                }
            }
            int r0 = 0;
            while (r0 < B - A) {
                int r1 = r0;
                while (r1 < B - A && cls[r1] == cls[r0]) ++r1;
                iv[d].push_back({A + r0, A + r1, false, cls[r0] == 2, cls[r0] >= 1});
                r0 = r1;
            }
        }
        if (B < N) iv[d].push_back({B, N, true, false, false});
    }
    for (Iv &v : iv[2])                                  // x slab intervals: one chunk each
        if (v.edge) {
            bool e = false, bx = false;
            for (auto &q : exc) e = e || (v.a - T < q.second && v.b + T > q.first);
            for (auto &q : boxx) bx = bx || (v.a < q.second && v.b > q.first);
            v.exc = e;
            v.box = bx || e;
        }
    std::vector<TB3Box> boxes;
    for (const Iv &vx : iv[2])
        for (const Iv &vy : iv[1])
            for (const Iv &vz : iv[0]) {
                TB3Box b{};
                b.mode = (vz.edge ? 1 : 0) | (vy.edge ? 2 : 0) | (vx.edge ? 4 : 0);
                b.exc = vx.exc ? 1 : 0;
                b.box = vx.box ? 1 : 0;
                b.rows = (b.mode == 0 && !b.exc) ? c->tb3_rows0 : 1;
                b.x0 = vx.a; b.x1 = vx.b;
                const int lenx = vx.b - vx.a;
                b.lx = (vx.edge && lenx + 2 * T <= 64) ? lenx : std::min(lx, lenx);
                b.y0 = vy.a; b.y1 = vy.b; b.z0 = vz.a; b.z1 = vz.b;
                const int WY = kTB3W * b.rows - 2 * T;
                b.nty = (vy.b - vy.a + WY - 1) / WY;
                b.ntz = (vz.b - vz.a + WZ - 1) / WZ;
                boxes.push_back(b);
            }
    // boxes by decreasing tile cost, so that the persistent workgroups end on short
    // tiles: iterations x rows, with every slab axis adding psi work and register spills
    auto cost = [&](const TB3Box &b) {
        const int it = std::min(b.lx, b.x1 - b.x0) + 2 * T;
        int ax = 0;
        for (int m = b.mode; m; m >>= 1) ax += m & 1;
        return (double)it * (b.rows == 2 ? 1.45 : 1.0) * (1.0 + 0.45 * ax * ax) *
               (b.exc ? 1.3 : (b.box ? 1.05 : 1.0));
    };
    std::stable_sort(boxes.begin(), boxes.end(),
                     [&](const TB3Box &p, const TB3Box &q) { return cost(p) > cost(q); });
    int first = 0;
    for (TB3Box &b : boxes) {
        b.first = first;
        const int ntx = (b.x1 - b.x0 + b.lx - 1) / b.lx;
        first += ntx * b.nty * b.ntz;
    }
    if ((int)boxes.size() > kTB3MaxBox) return fail(-1, "3D blocked pass: too many boxes");
    if (!sv.dbox) LF_HIP(hipMalloc(&sv.dbox, kTB3MaxBox * sizeof(TB3Box)));
    c->tb3_dbox = sv.dbox;
    LF_HIP(hipMemcpyAsync(c->tb3_dbox, boxes.data(), boxes.size() * sizeof(TB3Box),
                          hipMemcpyHostToDevice, c->stream));
    LF_HIP(hipStreamSynchronize(c->stream));
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, dev) == hipSuccess && pr.multiProcessorCount >= 1)
            ncu = pr.multiProcessorCount;
    }
    c->tb3_nwg = std::min(ncu, first);
    if (!c->tb3_ticket) {
        LF_HIP(hipMalloc(&c->tb3_ticket, sizeof(unsigned)));
        LF_HIP(hipMemsetAsync(c->tb3_ticket, 0, sizeof(unsigned), c->stream));
        c->tb3_base = 0;
    }
    c->tb3_boxes = boxes;
    c->tb3_ntile = first;
    c->tb3_T = T;
    c->tb3_lx = lx;
    c->tb3_key[0] = key_src; c->tb3_key[1] = key_drv; c->tb3_key[2] = c->tb3_rows0;
    c->tb3_key[3] = c->nmon;
    sv.boxes = boxes;
    sv.lx = lx;
    sv.ntile = first;
    sv.nwg = c->tb3_nwg;
    for (int q = 0; q < 4; ++q) sv.key[q] = c->tb3_key[q];
    sv.valid = true;
    return 0;
}

int ensure_tb3_buffers(lf_ctx *c) {
    const int ids[6] = {LF_EX, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ};
    for (int q = 0; q < 6; ++q)
        if (int r = ensure_alt(c, ids[q])) return r;
    const int hax[6] = {2, 1, 0, 2, 1, 0};            // Hx_z Hx_y Hy_x Hy_z Hz_y Hz_x
    const int eax[6] = {1, 2, 2, 0, 0, 1};            // Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y
    for (int q = 0; q < 6; ++q) {
        if (c->sH[q].psi && !c->psi_alt3[q]) {
            const size_t n = strip_elems_3d(c, c->sH[q], hax[q]);
            LF_HIP(hipMalloc(&c->psi_alt3[q], n * sizeof(float)));
            LF_HIP(hipMemsetAsync(c->psi_alt3[q], 0, n * sizeof(float), c->stream));
        }
        if (c->sE[q].psi && !c->psi_alt3e[q]) {
            const size_t n = strip_elems_3d(c, c->sE[q], eax[q]);
            LF_HIP(hipMalloc(&c->psi_alt3e[q], n * sizeof(float)));
            LF_HIP(hipMemsetAsync(c->psi_alt3e[q], 0, n * sizeof(float), c->stream));
        }
    }
    return 0;
}

// one temporally blocked 3D pass over T steps: reads arr[] and the psi slabs, writes the
// other buffers, then swaps
int tb_chunk_3d(lf_ctx *c, const lf_step *steps, int T, int fp64) {
    if (T < 1 || T > 3) return fail(-1, "3D temporal block depth must be 1, 2 or 3");
    if ((long long)c->g.nx * c->g.ny * c->g.P >= (1LL << 31))
        return fail(-1, "3D blocked pass: grid exceeds 2^31 elements");
    if (int r = ensure_tb3_buffers(c)) return r;
    if (int r = tb3_plan(c, T, c->tb_lx)) return r;
    TB3Args f{};
    f.g = c->g;
    f.Ex0 = c->arr[LF_EX]; f.Ey0 = c->arr[LF_EY]; f.Ez0 = c->arr[LF_EZ];
    f.Hx0 = c->arr[LF_HX]; f.Hy0 = c->arr[LF_HY]; f.Hz0 = c->arr[LF_HZ];
    f.Ex1 = c->alt[LF_EX]; f.Ey1 = c->alt[LF_EY]; f.Ez1 = c->alt[LF_EZ];
    f.Hx1 = c->alt[LF_HX]; f.Hy1 = c->alt[LF_HY]; f.Hz1 = c->alt[LF_HZ];
    f.cex = c->arr[LF_CEX]; f.cey = c->arr[LF_CEY]; f.cez = c->arr[LF_CEZ];
    f.dax = c->arr[LF_DAX]; f.day = c->arr[LF_DAY]; f.daz = c->arr[LF_DAZ];
    f.ch = c->ch;
    for (int ax = 0; ax < 3; ++ax) { f.pe[ax] = prof_of(c, ax, false); f.ph[ax] = prof_of(c, ax, true); }
    for (int q = 0; q < 6; ++q) {
        f.s0[q] = c->sH[q]; f.s1[q] = c->sH[q]; f.s1[q].psi = c->psi_alt3[q];
        f.s0[6 + q] = c->sE[q]; f.s1[6 + q] = c->sE[q]; f.s1[6 + q].psi = c->psi_alt3e[q];
    }
    for (int ax = 0; ax < 3; ++ax) { f.pec[2 * ax] = c->pec[ax][0]; f.pec[2 * ax + 1] = c->pec[ax][1]; }
    f.box = c->tb3_dbox;
    f.nbox = (int)c->tb3_boxes.size();
    f.ntile = c->tb3_ntile;
    f.ticket = c->tb3_ticket;
    f.base = c->tb3_base;
    for (int s = 0; s < T; ++s) { f.st[s] = step_args(steps[s]); f.mask[s] = steps[s].mask; }
    int bid[2];
    const unsigned bbits = c->tb3_ntile <= kTB3BoxSlots ? tb3_fused_boxes(c, bid) : 0u;
    const int nfused = __builtin_popcount(bbits);
    for (int m = 0; m < nfused; ++m) {
        const BoxArgs &bx = c->mon[bid[m]].box;
        f.boxbit[m] = (unsigned)bid[m];
        for (int q = 0; q < 3; ++q)
            for (int ax = 0; ax < 3; ++ax) { f.blo[m][q][ax] = bx.lo[q][ax]; f.bhi[m][q][ax] = bx.hi[q][ax]; }
        f.bpart[m] = bx.partial + kBoxBlocks;
        f.nboxm = m + 1;
    }
    const bool loss = f.dax || f.day || f.daz;
    launch_yee_tb3(T, loss, f, c->ex, fp64, (unsigned)c->tb3_nwg, c->stream);
    LF_HIP(hipGetLastError());
    // every workgroup draws its tiles and one ticket past the end
    c->tb3_base += (unsigned)(c->tb3_ntile + c->tb3_nwg);
    const int ids[6] = {LF_EX, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ};
    for (int q = 0; q < 6; ++q) std::swap(c->arr[ids[q]], c->alt[ids[q]]);
    for (int q = 0; q < 6; ++q) {
        if (c->sH[q].psi) std::swap(c->sH[q].psi, c->psi_alt3[q]);
        if (c->sE[q].psi) std::swap(c->sE[q].psi, c->psi_alt3e[q]);
    }
    return 0;
}

}  // namespace

namespace lfd {
void launch_yee_tb3(int T, bool loss, const TB3Args &f, const Excite &x, int fp64, unsigned nb,
                    hipStream_t st) {
    if (T == 1) (loss ? launch_yee_tb3_1_1 : launch_yee_tb3_1_0)(f, x, fp64, nb, st);
    else if (T == 2) (loss ? launch_yee_tb3_2_1 : launch_yee_tb3_2_0)(f, x, fp64, nb, st);
    else (loss ? launch_yee_tb3_3_1 : launch_yee_tb3_3_0)(f, x, fp64, nb, st);
}
}  // namespace lfd

extern "C" int lf_set_tile(lf_ctx *c, int ty) {
    if (int e = check_ctx(c)) return e;
    if (ty != 2 && ty != 4) return fail(-1, "lf_set_tile: rows per wavefront must be 2 or 4");
    c->ty = ty;
    return 0;
}

// ============================================================================ monitors (C ABI)
namespace {

int new_monitor(lf_ctx *c, int *id, Monitor **m) {
    if (c->nmon >= kMaxMon) return fail(-1, "at most 32 monitors per context");
    *id = c->nmon;
    *m = &c->mon[c->nmon++];
    return 0;
}

int alloc_zero(lf_ctx *c, void **p, size_t bytes) {
    LF_HIP(hipMalloc(p, bytes));
    LF_HIP(hipMemsetAsync(*p, 0, bytes, c->stream));
    return 0;
}

int comp_of_src(int src) {
    return src == LF_SRC_EX ? LF_EX : (src == LF_SRC_EY ? LF_EY : LF_EZ);
}

}  // namespace

extern "C" {

int lf_add_box(lf_ctx *c, const int *r, int *id) {
    if (int e = check_ctx(c)) return e;
    Monitor *m;
    if (int e = new_monitor(c, id, &m)) return e;
    m->type = 0;
    BoxArgs &b = m->box;
    if (c->kind == LF_YEE3D) {
        b.ncomp = 3;
        for (int q = 0; q < 3; ++q) {
            b.comp[q] = LF_EX + q;
            for (int a = 0; a < 3; ++a) { b.lo[q][a] = r[q * 6 + 2 * a]; b.hi[q][a] = r[q * 6 + 2 * a + 1]; }
        }
    } else {                                    // 2D: one node-grid box (x, y) -> (i, 0, k)
        b.ncomp = 1;
        b.comp[0] = -1;
        b.lo[0][0] = r[0]; b.hi[0][0] = r[1];
        b.lo[0][1] = 0;    b.hi[0][1] = 1;
        b.lo[0][2] = r[2]; b.hi[0][2] = r[3];
    }
    for (int q = 0; q < b.ncomp; ++q)
        for (int a = 0; a < 3; ++a)
            if (b.hi[q][a] < b.lo[q][a] || b.lo[q][a] < 0) return fail(-1, "lf_add_box: bad range");
    // 3D: one FP64 slot per tile of the blocked pass follows the per-block partials
    m->bytes = (kBoxBlocks + (c->kind == LF_YEE3D ? kTB3BoxSlots : 0)) * sizeof(double);
    if (int e = alloc_zero(c, &m->buf, m->bytes)) return e;
    b.partial = (double *)m->buf;
    return 0;
}

/* lf_add_box with a 2D form: 0 the node-grid I_E of lf_add_box, 1 (2D TM) the node Ey^2
 * alone (the time-domain stitch detectors, fdtd2d_core.py:1811-1815). */
int lf_add_box_form(lf_ctx *c, const int *r, int form, int *id) {
    if (int e = check_ctx(c)) return e;
    if (form != 0 && (form != 1 || c->kind != LF_TM2D))
        return fail(-1, "lf_add_box_form: form 1 is 2D TM only");
    if (int e = lf_add_box(c, r, id)) return e;
    if (form == 1) c->mon[*id].box.comp[0] = -2;
    return 0;
}

/* Sum over the masked steps of the instantaneous net flux in +x through x plane il
 * (fdtd2d_core.py:489-497, fdtd3d_core.py:762-767), FP64. */
int lf_add_sx(lf_ctx *c, int il, int *id) {
    if (int e = check_ctx(c)) return e;
    if (il < 1 || il > c->nx - 2) return fail(-1, "lf_add_sx: plane index out of range");
    Monitor *m;
    if (int e = new_monitor(c, id, &m)) return e;
    m->type = 4;
    m->sx.il = il;
    m->bytes = kSxBlocks * sizeof(double);
    if (int e = alloc_zero(c, &m->buf, m->bytes)) return e;
    m->sx.partial = (double *)m->buf;
    return 0;
}

/* Plane recorder: x plane il of component comp into row `lf_step.row` of a rows x plane
 * buffer on the steps of its mask bit; phase 0 records at the start of the step (E(n)),
 * phase 1 after the H update and its TF/SF corrections (H(n+1/2)). */
int lf_add_rec(lf_ctx *c, int comp, int il, int phase, int rows, int *id) {
    if (int e = check_ctx(c)) return e;
    if (comp < LF_EX || comp > LF_HZ) return fail(-1, "lf_add_rec: field component expected");
    if (phase != 0 && phase != 1) return fail(-1, "lf_add_rec: phase 0 or 1");
    if (rows < 1) return fail(-1, "lf_add_rec: rows >= 1");
    int ex, ey, ez;
    comp_extent(c, comp, ex, ey, ez);
    if (il < 0 || il >= ex) return fail(-1, "lf_add_rec: plane index out of range");
    Monitor *m;
    if (int e = new_monitor(c, id, &m)) return e;
    m->type = 5;
    m->rcomp = comp; m->ril = il; m->rphase = phase;
    m->n0 = rows; m->n1 = ey; m->n2 = ez;
    m->bytes = (size_t)rows * ey * ez * sizeof(float);
    ++c->nrec;
    return alloc_zero(c, &m->buf, m->bytes);
}

/* TF/SF correction of x plane il of component comp (fdtd2d_core.py:1771-1789,
 * fdtd3d_core.py:1973-1986): F = F + c inc (sign > 0) or F - c inc over [j0,j1) x [k0,k1)
 * of the plane, c = the context's Ch (coef = -1) or a coefficient array (LF_CEH, LF_CAY,
 * LF_CEY, LF_CEZ, ...) at the cell, inc = row `lf_step.row` of a rows x plane series
 * loaded with lf_load_tfsf.  H components act after the H update, E components after the
 * E update (and source) of every step.  Returns the correction's index in *id. */
int lf_add_tfsf(lf_ctx *c, int comp, int il, int sign, int coef, int j0, int j1, int k0, int k1,
                int rows, int *id) {
    if (int e = check_ctx(c)) return e;
    if (comp < LF_EX || comp > LF_HZ) return fail(-1, "lf_add_tfsf: field component expected");
    if (rows < 1) return fail(-1, "lf_add_tfsf: rows >= 1");
    int ex, ey, ez;
    comp_extent(c, comp, ex, ey, ez);
    if (il < 0 || il >= ex) return fail(-1, "lf_add_tfsf: plane index out of range");
    if (j0 < 0 || j1 > ey || j0 >= j1 || k0 < 0 || k1 > ez || k0 >= k1)
        return fail(-1, "lf_add_tfsf: range outside the plane");
    const bool h = comp >= LF_HX;
    if (coef >= 0) {
        if (h || coef < LF_CEE || coef >= LF_NARRAYS || !c->arr[coef])
            return fail(-1, "lf_add_tfsf: coefficient array not uploaded (E components only)");
        int cx, cy, cz;
        comp_extent(c, coef, cx, cy, cz);
        if (il >= cx || j1 > cy || k1 > cz) return fail(-1, "lf_add_tfsf: coefficient array too small");
    } else if (!h) {
        return fail(-1, "lf_add_tfsf: E components take a coefficient array");
    }
    TfsfArgs a{};
    a.comp = comp; a.il = il; a.sign = sign > 0 ? 1 : -1; a.coef = coef;
    a.j0 = j0; a.j1 = j1; a.k0 = k0; a.k1 = k1; a.n1 = ey; a.n2 = ez; a.rows = rows;
    LF_HIP(hipMalloc(&a.series, (size_t)rows * ey * ez * sizeof(float)));
    LF_HIP(hipMemsetAsync(a.series, 0, (size_t)rows * ey * ez * sizeof(float), c->stream));
    *id = (int)c->tfsf.size();
    c->tfsf.push_back(a);
    return 0;
}

/* Rows [0, nrows) of a TF/SF correction's incident series (nrows x plane floats). */
int lf_load_tfsf(lf_ctx *c, int id, const float *rows, int nrows) {
    if (int e = check_ctx(c)) return e;
    if (id < 0 || id >= (int)c->tfsf.size()) return fail(-1, "lf_load_tfsf: bad id");
    TfsfArgs &a = c->tfsf[id];
    if (nrows < 0 || nrows > a.rows || (nrows > 0 && !rows)) return fail(-1, "lf_load_tfsf: bad row count");
    if (nrows == 0) return 0;
    LF_HIP(hipMemcpyAsync(a.series, rows, (size_t)nrows * a.n1 * a.n2 * sizeof(float),
                          hipMemcpyHostToDevice, c->stream));
    LF_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

/* max |F| over a field component (fdtd2d_core.py:1829-1835, fdtd3d_core.py:2000-2004). */
int lf_field_maxabs(lf_ctx *c, int comp, float *out) {
    if (int e = check_ctx(c)) return e;
    if (comp < LF_EX || comp > LF_HZ || !c->arr[comp] || !out)
        return fail(-1, "lf_field_maxabs: bad component");
    if (!c->mx_part) LF_HIP(hipMalloc(&c->mx_part, kMaxAbsBlocks * sizeof(float)));
    int ex, ey, ez;
    comp_extent(c, comp, ex, ey, ez);
    c->tb_chain = false;
    maxabs_kernel<<<kMaxAbsBlocks, 256, 0, c->stream>>>(c->arr[comp], c->g, ex, ey, ez, c->mx_part);
    LF_HIP(hipGetLastError());
    float part[kMaxAbsBlocks];
    LF_HIP(hipMemcpyAsync(part, c->mx_part, sizeof(part), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(hipStreamSynchronize(c->stream));
    float v = 0.0f;
    for (float p : part) v = std::max(v, p);
    *out = v;
    return 0;
}

int lf_add_sq(lf_ctx *c, int form, int nx_acc, int *id) {
    if (int e = check_ctx(c)) return e;
    Monitor *m;
    if (int e = new_monitor(c, id, &m)) return e;
    m->type = 1;
    m->form = form;
    if (c->kind == LF_YEE3D) {
        if (form != LF_SQ_3D_S && form != LF_SQ_3D_P) return fail(-1, "lf_add_sq: 3D form expected");
        if (nx_acc < 2 || nx_acc > c->nx) return fail(-1, "lf_add_sq: bad nx_acc");
        m->nxa = nx_acc;
        m->n1 = form == LF_SQ_3D_S ? c->ny : c->ny - 1;
        m->n2 = form == LF_SQ_3D_S ? c->nz - 1 : c->nz;
    } else {
        if (form > LF_SQ_TM_EX2) return fail(-1, "lf_add_sq: 2D form expected");
        m->nxa = c->nx;
        m->n1 = 1;
        m->n2 = c->ny;
    }
    m->n0 = m->nxa;
    m->bytes = (size_t)m->n0 * m->n1 * m->n2 * sizeof(float);
    return alloc_zero(c, &m->buf, m->bytes);
}

int lf_add_dft(lf_ctx *c, int src, int x0, int x1, int kplane, int phasor, int acc, int *id) {
    if (int e = check_ctx(c)) return e;
    Monitor *m;
    if (int e = new_monitor(c, id, &m)) return e;
    m->type = 2;
    DftArgs &d = m->dft;
    d.src = src; d.x0 = x0; d.kplane = kplane; d.acc = acc; d.phasor = phasor;
    int ex, ey, ez;
    if (src == LF_SRC_EY_NODE || src == LF_SRC_EX_NODE) {
        if (c->kind != LF_TM2D) return fail(-1, "node sources are 2D TM only");
        ex = c->nx; ey = 1; ez = c->ny;
    } else {
        comp_extent(c, comp_of_src(src), ex, ey, ez);
    }
    if (x0 < 0 || x1 > ex || x1 <= x0) return fail(-1, "lf_add_dft: bad x range");
    d.n0 = x1 - x0;
    if (kplane >= 0) {
        if (c->kind != LF_YEE3D || kplane >= ez) return fail(-1, "lf_add_dft: bad kplane");
        d.n1 = 1; d.n2 = ey;                   // stored (rows, ey) ; kernel maps q -> j
    } else {
        d.n1 = ey; d.n2 = ez;
    }
    const size_t el = acc == LF_ACC_C128 ? sizeof(double2) : sizeof(float2);
    m->bytes = (size_t)d.n0 * d.n1 * d.n2 * el;
    if (int e = alloc_zero(c, &m->buf, m->bytes)) return e;
    d.buf = m->buf;
    return 0;
}

int lf_add_flux(lf_ctx *c, int il, int *id) {
    if (int e = check_ctx(c)) return e;
    if (il < 1 || il >= c->nx - 1) return fail(-1, "lf_add_flux: plane index out of range");
    Monitor *m;
    if (int e = new_monitor(c, id, &m)) return e;
    m->type = 3;
    FluxArgs &f = m->flux;
    f.il = il;
    f.wide = c->mode == LF_MODE_FP64ACC;
    if (c->kind == LF_TE2D) { f.npair = 1; f.ecomp[0] = LF_EZ; f.hcomp[0] = LF_HY; }
    else if (c->kind == LF_TM2D) { f.npair = 1; f.ecomp[0] = LF_EY; f.hcomp[0] = LF_HZ; }
    else { f.npair = 2; f.ecomp[0] = LF_EY; f.hcomp[0] = LF_HZ; f.ecomp[1] = LF_EZ; f.hcomp[1] = LF_HY; }
    const size_t el = f.wide ? sizeof(double2) : sizeof(float2);
    m->bytes = 0;
    for (int p = 0; p < f.npair; ++p) {
        int ex, ey, ez;
        comp_extent(c, f.ecomp[p], ex, ey, ez);
        f.n1[p] = ey; f.n2[p] = ez;
        const size_t b = (size_t)ey * ez * el;
        if (int e = alloc_zero(c, &f.aE[p], b)) return e;
        if (int e = alloc_zero(c, &f.aH[p], b)) return e;
        m->bytes += 2 * b;
    }
    return 0;
}

int lf_monitor_bytes(lf_ctx *c, int id, long long *bytes) {
    if (int e = check_ctx(c)) return e;
    if (id < 0 || id >= c->nmon) return fail(-1, "bad monitor id");
    *bytes = (long long)c->mon[id].bytes;
    return 0;
}

int lf_monitor_read(lf_ctx *c, int id, void *dst, long long bytes) {
    if (int e = check_ctx(c)) return e;
    if (id < 0 || id >= c->nmon) return fail(-1, "bad monitor id");
    Monitor &m = c->mon[id];
    if ((size_t)bytes != m.bytes) return fail(-1, "lf_monitor_read: size mismatch");
    LF_HIP(hipStreamSynchronize(c->stream));
    if (m.type == 3) {                          // pairs: aE0 aH0 [aE1 aH1]
        char *o = (char *)dst;
        const size_t el = m.flux.wide ? sizeof(double2) : sizeof(float2);
        for (int p = 0; p < m.flux.npair; ++p) {
            const size_t b = (size_t)m.flux.n1[p] * m.flux.n2[p] * el;
            LF_HIP(hipMemcpy(o, m.flux.aE[p], b, hipMemcpyDeviceToHost)); o += b;
            LF_HIP(hipMemcpy(o, m.flux.aH[p], b, hipMemcpyDeviceToHost)); o += b;
        }
        return 0;
    }
    LF_HIP(hipMemcpy(dst, m.buf, m.bytes, hipMemcpyDeviceToHost));
    return 0;
}

/* Steady-state test on the device: out[0] = sum |A/da - B/db|^2, out[1] = sum |A/da|^2
 * over two DFT monitors of the same shape and format (two doubles to the host instead of
 * two phasor fields). */
int lf_monitor_change(lf_ctx *c, int id_a, int id_b, double da, double db, double *out) {
    if (int e = check_ctx(c)) return e;
    if (id_a < 0 || id_a >= c->nmon || id_b < 0 || id_b >= c->nmon || !out)
        return fail(-1, "lf_monitor_change: bad arguments");
    const Monitor &A = c->mon[id_a], &B = c->mon[id_b];
    if (A.type != 2 || B.type != 2) return fail(-1, "lf_monitor_change: DFT monitors only");
    if (A.bytes != B.bytes || A.dft.acc != B.dft.acc)
        return fail(-1, "lf_monitor_change: the monitors differ in shape or format");
    const size_t el = A.dft.acc == LF_ACC_C128 ? sizeof(double2) : sizeof(float2);
    const long long n = (long long)(A.bytes / el);
    const size_t need = (size_t)nd_partial_count(n) * sizeof(double2);
    if (need > c->nd_cap) {
        LF_HIP(hipStreamSynchronize(c->stream));
        if (c->nd_part) LF_HIP(hipFree(c->nd_part));
        c->nd_part = nullptr;
        c->nd_cap = 0;
        LF_HIP(hipMalloc(&c->nd_part, need));
        c->nd_cap = need;
    }
    if (!c->nd_out) LF_HIP(hipMalloc(&c->nd_out, 2 * sizeof(double)));
    c->tb_chain = false;
    if (int e = norm_diff_launch(A.buf, B.buf, n, A.dft.acc, da, db, c->nd_part, c->nd_out, c->stream))
        return e;
    LF_HIP(hipMemcpyAsync(out, c->nd_out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int lf_monitor_write(lf_ctx *c, int id, const void *src, long long bytes) {
    if (int e = check_ctx(c)) return e;
    if (id < 0 || id >= c->nmon) return fail(-1, "bad monitor id");
    Monitor &m = c->mon[id];
    if ((size_t)bytes != m.bytes) return fail(-1, "lf_monitor_write: size mismatch");
    LF_HIP(hipStreamSynchronize(c->stream));
    if (m.type == 3) {
        const char *o = (const char *)src;
        const size_t el = m.flux.wide ? sizeof(double2) : sizeof(float2);
        for (int p = 0; p < m.flux.npair; ++p) {
            const size_t b = (size_t)m.flux.n1[p] * m.flux.n2[p] * el;
            LF_HIP(hipMemcpy(m.flux.aE[p], o, b, hipMemcpyHostToDevice)); o += b;
            LF_HIP(hipMemcpy(m.flux.aH[p], o, b, hipMemcpyHostToDevice)); o += b;
        }
        return 0;
    }
    if (m.type == 0 || m.type == 4) {           // box / Poynting sum: the value goes to partial[0]
        std::vector<double> part(m.bytes / sizeof(double), 0.0);
        part[0] = *(const double *)src;
        LF_HIP(hipMemcpy(m.buf, part.data(), m.bytes, hipMemcpyHostToDevice));
        return 0;
    }
    LF_HIP(hipMemcpy(m.buf, src, m.bytes, hipMemcpyHostToDevice));
    return 0;
}

int lf_monitor_zero(lf_ctx *c, int id) {
    if (int e = check_ctx(c)) return e;
    if (id < 0 || id >= c->nmon) return fail(-1, "bad monitor id");
    Monitor &m = c->mon[id];
    if (m.type == 3) {
        const size_t el = m.flux.wide ? sizeof(double2) : sizeof(float2);
        for (int p = 0; p < m.flux.npair; ++p) {
            const size_t b = (size_t)m.flux.n1[p] * m.flux.n2[p] * el;
            LF_HIP(hipMemsetAsync(m.flux.aE[p], 0, b, c->stream));
            LF_HIP(hipMemsetAsync(m.flux.aH[p], 0, b, c->stream));
        }
        return 0;
    }
    LF_HIP(hipMemsetAsync(m.buf, 0, m.bytes, c->stream));
    return 0;
}

int lf_set_timing(lf_ctx *c, int on) {
    if (int e = check_ctx(c)) return e;
    c->timing = on != 0;
    return 0;
}

int lf_kernel_ms(lf_ctx *c, double *update_ms, double *total_ms, int *nsteps) {
    if (int e = check_ctx(c)) return e;
    *update_ms = c->last_update_ms;
    *total_ms = c->last_total_ms;
    *nsteps = c->last_steps;
    return 0;
}

int lf_sync(lf_ctx *c) {
    if (int e = check_ctx(c)) return e;
    LF_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

namespace {
int run_impl(lf_ctx *c, const lf_step *steps, int nsteps, bool allow_tb);
}

extern "C" int lf_run(lf_ctx *c, const lf_step *steps, int nsteps) {
    if (int e = check_ctx(c)) return e;
    return run_impl(c, steps, nsteps, true);
}

namespace {

// Mur: the planes each face needs from before the step
int absorb_capture(lf_ctx *c) {
    for (MurArgs &m : c->mur) {
        m.f = c->arr[c->mur_comp[&m - c->mur.data()]];
        const int n = m.nu * m.nv;
        mur_capture_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(m);
    }
    return 0;
}

// sponge on all six fields, then the Mur faces in their order, on the fields of the step
int absorb_apply(lf_ctx *c) {
    if (c->sp_n > 0) {
        const int ids[6] = {LF_EX, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ};
        for (int q = 0; q < 6; ++q) {
            SpongeArgs a{};
            a.f = c->arr[ids[q]];
            comp_extent(c, ids[q], a.ex, a.ey, a.ez);
            a.g = c->g;
            a.damp = c->sp_g;
            a.n = c->sp_n;
            a.faces = c->sp_faces;
            sponge_kernel<<<dim3(a.ey, a.ex, 1), 64, 0, c->stream>>>(a);
        }
    }
    for (size_t q = 0; q < c->mur.size(); ++q) {
        MurArgs &m = c->mur[q];
        m.f = c->arr[c->mur_comp[q]];
        const int n = m.nu * m.nv;
        mur_apply_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(m);
    }
    return 0;
}

// monitors whose bit is set in `mask`, on the fields at the end of step t
int launch_monitors(lf_ctx *c, const MonArgs &ma, const lf_step &t, const StepArgs &st,
                    unsigned mask) {
    for (int id = 0; mask && id < c->nmon; ++id) {
        if (!((mask >> id) & 1u)) continue;
        Monitor &m = c->mon[id];
        if (m.type == 0) {
            // as many blocks as the box fills (at most kBoxBlocks): a fixed grid per box,
            // so its sum keeps one order; the other partial slots stay 0
            long long tot = 0;
            for (int q = 0; q < m.box.ncomp; ++q)
                tot = std::max(tot, (long long)(m.box.hi[q][0] - m.box.lo[q][0]) *
                                        (m.box.hi[q][1] - m.box.lo[q][1]) *
                                        (m.box.hi[q][2] - m.box.lo[q][2]));
            const int nb = (int)std::min<long long>(kBoxBlocks, std::max<long long>(1, (tot + 255) / 256));
            box_kernel<<<nb, 256, 0, c->stream>>>(ma, m.box);
        } else if (m.type == 1) {
            const dim3 mb = mon_block(m.n1);
            dim3 gq((m.n2 + mb.x - 1) / mb.x, (m.n1 + mb.y - 1) / mb.y, m.n0);
            sq_kernel<<<gq, mb, 0, c->stream>>>(ma, (float *)m.buf, m.form, m.nxa, m.n1, m.n2);
        } else if (m.type == 2) {
            const DftArgs &d = m.dft;
            const dim3 mb = mon_block(d.n1);
            dim3 gq((d.n2 + mb.x - 1) / mb.x, (d.n1 + mb.y - 1) / mb.y, d.n0);
            const float fr = d.phasor == LF_PH_Q ? t.qr32 : t.pr32;
            const float fi = d.phasor == LF_PH_Q ? t.qi32 : t.pi32;
            dft_kernel<<<gq, mb, 0, c->stream>>>(ma, d, t.pr, t.pi, fr, fi);
        } else if (m.type == 4) {
            SxArgs a = m.sx;
            const long long n = c->kind == LF_YEE3D ? (long long)c->ny * c->nz : c->ny;
            const int nb = (int)std::min<long long>(kSxBlocks, (n + 255) / 256);
            sx_kernel<<<nb, 256, 0, c->stream>>>(ma, a);
        } else if (m.type == 3) {
            const FluxArgs &f = m.flux;
            const int n1 = f.n1[0] > (f.npair > 1 ? f.n1[1] : 0) ? f.n1[0] : f.n1[1];
            const int n2 = f.n2[0] > (f.npair > 1 ? f.n2[1] : 0) ? f.n2[0] : f.n2[1];
            dim3 gq((n2 + 255) / 256, n1, f.npair);
            flux_kernel<<<gq, 256, 0, c->stream>>>(ma, f, st, t.pr, t.pi, t.hr, t.hi, t.pr32,
                                                   t.pi32, t.hr32, t.hi32);
        }
    }
    return 0;
}

// time-domain stitch: the plane recorders of one phase whose bit is set in the step's
// mask, into row t.row
int td_record(lf_ctx *c, const lf_step &t, int phase) {
    for (int id = 0; id < c->nmon; ++id) {
        const Monitor &m = c->mon[id];
        if (m.type != 5 || m.rphase != phase || !((t.mask >> id) & 1u)) continue;
        const int n = m.n1 * m.n2;
        float *dst = (float *)m.buf + (size_t)t.row * n;
        rec_plane_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(c->arr[m.rcomp], c->g, m.ril,
                                                                m.n1, m.n2, dst);
    }
    return 0;
}

// the TF/SF corrections of the H (h = true) or the E components, row `row` of their series
int td_tfsf(lf_ctx *c, int row, bool h) {
    for (const TfsfArgs &a : c->tfsf) {
        if ((a.comp >= LF_HX) != h) continue;
        const int n = (a.j1 - a.j0) * (a.k1 - a.k0);
        tfsf_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(
            c->arr[a.comp], c->g, a, a.coef >= 0 ? c->arr[a.coef] : nullptr, c->ch, row);
    }
    return 0;
}

// rows the step's recorders and TF/SF series are addressed with must exist
int td_check_row(const lf_ctx *c, const lf_step &t) {
    for (const TfsfArgs &a : c->tfsf)
        if (t.row >= (unsigned)a.rows) return fail(-1, "lf_run: step row outside a TF/SF series");
    for (int id = 0; id < c->nmon; ++id)
        if (c->mon[id].type == 5 && ((t.mask >> id) & 1u) && t.row >= (unsigned)c->mon[id].n0)
            return fail(-1, "lf_run: step row outside a plane recorder");
    return 0;
}

// the step loop of lf_run; allow_tb = false runs one step per launch (the batch engine
// advances the instances it does not block itself this way)
int run_impl(lf_ctx *c, const lf_step *steps, int nsteps, bool allow_tb) {
    if (!c->finalized) return fail(-1, "lf_run before lf_finalize");
    if (nsteps <= 0) return 0;
    if (c->ch == 0.0f) return fail(-1, "lf_run: Ch not set");
    if (c->kind == LF_TE2D && (!c->arr[LF_CEE] || !c->arr[LF_CEH]))
        return fail(-1, "lf_run: Ce_E/Ce_H not uploaded");
    if (c->kind == LF_TM2D && (!c->arr[LF_CAX] || !c->arr[LF_CAY]))
        return fail(-1, "lf_run: Ca_ex/Ca_ey not uploaded");
    if (c->kind == LF_YEE3D && (!c->arr[LF_CEX] || !c->arr[LF_CEY] || !c->arr[LF_CEZ]))
        return fail(-1, "lf_run: ce_x/ce_y/ce_z not uploaded");
    const int fp64 = c->mode == LF_MODE_FP64ACC;
    const bool absorbers = c->sp_n > 0 || !c->mur.empty();
    const bool td = !c->tfsf.empty() || c->nrec > 0;
    if (absorbers || td) allow_tb = false;   // sponge / Mur / TF/SF / records act inside every step
    if (td)
        for (int s = 0; s < nsteps; ++s)
            if (int e = td_check_row(c, steps[s])) return e;
    const Grid &g = c->g;
    dim3 blk(64, 4, 1);
    dim3 grd2((g.nz + 63) / 64, (g.nx + 3) / 4, 1);                    // 2D: (fast y, x)
    dim3 grd3((g.nz + 63) / 64, (g.ny + 3) / 4, g.nx);                 // 3D: (z, y, x)
    TEArgs te{}; TMArgs tm{}; YeeArgs yee{};
    if (c->kind == LF_TE2D) te = te_args(c);
    else if (c->kind == LF_TM2D) tm = tm_args(c);
    else yee = yee_args(c);
    MonArgs ma = mon_args(c);
    c->tb_chain = false;                 // work queued since the last pass (see fork_tb_classes)
    const bool timing = c->timing;
    if (timing) {
        while ((int)c->ev.size() < 2 * nsteps + 2) {
            hipEvent_t e;
            LF_HIP(hipEventCreate(&e));
            c->ev.push_back(e);
        }
        LF_HIP(hipEventRecord(c->ev[2 * nsteps], c->stream));
    }
    // Timed groups: one per one-step update or monitored pass, and one per run of
    // consecutive plain blocked passes (events at the run's ends only, so that a pass's
    // time is the run's start-to-start period, without events between its passes)
    int ngroup = 0;
    std::vector<int> tb_groups;
    int tb_passes = 0;
    bool tb_open = false;
    auto close_run = [&]() {
        if (!tb_open) return hipSuccess;
        tb_open = false;
        tb_groups.push_back(ngroup++);
        return timing ? hipEventRecord(c->ev[2 * (ngroup - 1) + 1], c->stream) : hipSuccess;
    };
    for (int s = 0; s < nsteps; ++s) {
        const lf_step &t = steps[s];
        StepArgs st = step_args(t);
        const int T = allow_tb ? c->tb : 1;
        if (T > 1 && s + T <= nsteps) {
            // 2D: no monitor in the block's steps.  3D: the pass sums the detector boxes
            // itself, and the monitors of its last step run after it on the fields it
            // leaves (they read only the fields at the end of their step)
            int fid[2];
            unsigned fused = 0u;
            if (c->kind == LF_YEE3D && T <= 3) {
                if (int e = tb3_plan(c, T, c->tb_lx)) return e;
                if (c->tb3_ntile <= kTB3BoxSlots) fused = tb3_fused_boxes(c, fid);
            }
            bool quiet = true;
            for (int u = 0; u < T; ++u) {
                const bool last = c->kind == LF_YEE3D && u == T - 1;
                quiet = quiet && (last || (steps[s + u].mask & ~fused) == 0);
            }
            if (quiet) {
                if (!tb_open) {
                    if (timing) LF_HIP(hipEventRecord(c->ev[2 * ngroup], c->stream));
                    tb_open = true;
                }
                if (c->kind == LF_YEE3D) {
                    if (int e = tb_chunk_3d(c, steps + s, T, fp64)) return e;
                } else {
                    if (int e = tb_chunk_2d(c, steps + s, T, fp64)) return e;
                }
                ++tb_passes;
                if (c->kind == LF_TM2D) tm = tm_args(c);
                else if (c->kind == LF_TE2D) te = te_args(c);
                else yee = yee_args(c);
                ma = mon_args(c);
                s += T - 1;
                const lf_step &tl = steps[s];
                if ((tl.mask & ~fused) != 0) {
                    c->tb_chain = false;
                    LF_HIP(close_run());
                }
                if (int e = launch_monitors(c, ma, tl, step_args(tl), tl.mask & ~fused)) return e;
                continue;
            }
        }
        LF_HIP(close_run());
        // 2D steps with monitors: a monitored pass when all of them can be summed inside it
        const int Tm = allow_tb ? c->tb_mon : 1;
        TBMonPlan mp;
        if (c->kind != LF_YEE3D && Tm > 1 && s + Tm <= nsteps && tb_mon_plan(c, steps + s, Tm, mp)) {
            if (timing) LF_HIP(hipEventRecord(c->ev[2 * ngroup], c->stream));
            if (int e = tb_chunk_2d_mon(c, steps + s, Tm, mp, fp64)) return e;
            c->tb_chain = false;         // the flux additions follow the pass on the context stream
            if (timing) LF_HIP(hipEventRecord(c->ev[2 * ngroup + 1], c->stream));
            ++ngroup;
            if (c->kind == LF_TM2D) tm = tm_args(c);
            else te = te_args(c);
            ma = mon_args(c);
            s += Tm - 1;
            continue;
        }
        c->tb_chain = false;
        if (timing) LF_HIP(hipEventRecord(c->ev[2 * ngroup], c->stream));
        if (td) td_record(c, t, 0);
        if (!c->mur.empty())
            if (int e = absorb_capture(c)) return e;
        if (td) {
            // H update, TF/SF on H, records of H(n+1/2), E update with source and PEC, TF/SF
            // on E (fdtd2d_core.py:1768-1794, fdtd3d_core.py:1966-1990)
            if (c->kind == LF_TE2D) te_h_kernel<<<grd2, blk, 0, c->stream>>>(te);
            else if (c->kind == LF_TM2D) tm_h_kernel<<<grd2, blk, 0, c->stream>>>(tm);
            else yee_h_kernel<<<grd3, blk, 0, c->stream>>>(yee);
            td_tfsf(c, (int)t.row, true);
            td_record(c, t, 1);
            if (c->kind == LF_TE2D) te_e_kernel<<<grd2, blk, 0, c->stream>>>(te, c->ex, st, fp64);
            else if (c->kind == LF_TM2D) tm_e_kernel<<<grd2, blk, 0, c->stream>>>(tm, c->ex, st, fp64);
            else yee_e_kernel<<<grd3, blk, 0, c->stream>>>(yee, c->ex, st, fp64, 0);
            td_tfsf(c, (int)t.row, false);
        } else if (c->kind == LF_TE2D) {
            te_h_kernel<<<grd2, blk, 0, c->stream>>>(te);
            te_e_kernel<<<grd2, blk, 0, c->stream>>>(te, c->ex, st, fp64);
        } else if (c->kind == LF_TM2D && c->variant == 0) {
            if (int e = tm_fused_step(c, st, fp64)) return e;
            ma = mon_args(c);
        } else if (c->kind == LF_TM2D) {
            tm_h_kernel<<<grd2, blk, 0, c->stream>>>(tm);
            tm_e_kernel<<<grd2, blk, 0, c->stream>>>(tm, c->ex, st, fp64);
        } else if (allow_tb && c->tb > 1) {
            // 3D steps that no pass of depth c->tb can take (monitors on a step before a
            // pass's last, or too few steps left): a pass of depth 1, the same tile code
            // with one level (one read and one write of every field instead of the two
            // sweeps of the H then E kernels).  It sums the detector boxes itself; the
            // other monitors run after it.
            if (int e = tb3_plan(c, 1, c->tb_lx)) return e;
            int fid[2];
            const unsigned fused = c->tb3_ntile <= kTB3BoxSlots ? tb3_fused_boxes(c, fid) : 0u;
            if (int e = tb_chunk_3d(c, &t, 1, fp64)) return e;
            yee = yee_args(c);
            ma = mon_args(c);
            if (timing) LF_HIP(hipEventRecord(c->ev[2 * ngroup + 1], c->stream));
            ++ngroup;
            if (int e = launch_monitors(c, ma, t, st, t.mask & ~fused)) return e;
            continue;
        } else if (c->variant == 0) {
            if (int e = yee_fused_step(c, st, fp64)) return e;
            ma = mon_args(c);
        } else {
            yee_h_kernel<<<grd3, blk, 0, c->stream>>>(yee);
            yee_e_kernel<<<grd3, blk, 0, c->stream>>>(yee, c->ex, st, fp64, 0);
        }
        if (absorbers)
            if (int e = absorb_apply(c)) return e;
        if (timing) LF_HIP(hipEventRecord(c->ev[2 * ngroup + 1], c->stream));
        ++ngroup;
        if (int e = launch_monitors(c, ma, t, st, t.mask)) return e;
    }
    LF_HIP(close_run());
    LF_HIP(hipGetLastError());
    if (timing) {
        LF_HIP(hipEventRecord(c->ev[2 * nsteps + 1], c->stream));
        LF_HIP(hipEventSynchronize(c->ev[2 * nsteps + 1]));
        double upd = 0.0, tbms = 0.0;
        size_t k = 0;
        for (int gq = 0; gq < ngroup; ++gq) {
            float ms = 0.f;
            LF_HIP(hipEventElapsedTime(&ms, c->ev[2 * gq], c->ev[2 * gq + 1]));
            upd += ms;
            if (k < tb_groups.size() && tb_groups[k] == gq) {
                tbms += ms;
                ++k;
            }
        }
        c->last_tb_ms = tbms;
        c->last_tb_launches = tb_passes;
        float tot = 0.f;
        LF_HIP(hipEventElapsedTime(&tot, c->ev[2 * nsteps], c->ev[2 * nsteps + 1]));
        c->last_update_ms = upd;
        c->last_total_ms = tot;
        c->last_steps = nsteps;
    }
    return 0;
}

}  // namespace

extern "C" {

}  // extern "C"

// ============================================================================ autotuner (C ABI)
namespace {

// device copies of every field and psi slab of a context (the state a blocked pass reads
// and writes), for the autotuner's trial passes
struct Snapshot {
    std::vector<std::pair<float **, size_t>> src;   // where the live buffer pointer sits, floats
    std::vector<float *> copy;
};

int snapshot_take(lf_ctx *c, Snapshot &sn) {
    const int ids3[6] = {LF_EX, LF_EY, LF_EZ, LF_HX, LF_HY, LF_HZ};
    const int idsTE[3] = {LF_EZ, LF_HX, LF_HY}, idsTM[3] = {LF_EX, LF_EY, LF_HZ};
    const int *ids = c->kind == LF_YEE3D ? ids3 : (c->kind == LF_TE2D ? idsTE : idsTM);
    const int n = c->kind == LF_YEE3D ? 6 : 3;
    for (int q = 0; q < n; ++q) sn.src.push_back({&c->arr[ids[q]], grid_elems(c)});
    if (c->kind == LF_YEE3D) {
        const int hax[6] = {2, 1, 0, 2, 1, 0}, eax[6] = {1, 2, 2, 0, 0, 1};
        for (int q = 0; q < 6; ++q) {
            if (c->sH[q].psi) sn.src.push_back({&c->sH[q].psi, strip_elems_3d(c, c->sH[q], hax[q])});
            if (c->sE[q].psi) sn.src.push_back({&c->sE[q].psi, strip_elems_3d(c, c->sE[q], eax[q])});
        }
    } else {
        for (int q = 0; q < 4; ++q)
            if (c->s2[q].psi)
                sn.src.push_back({&c->s2[q].psi, strip_elems_2d(c, c->s2[q], x_slab_2d(c->kind, q))});
    }
    for (auto &e : sn.src) {
        float *d = nullptr;
        LF_HIP(hipMalloc(&d, e.second * sizeof(float)));
        sn.copy.push_back(d);
        LF_HIP(hipMemcpyAsync(d, *e.first, e.second * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    }
    return 0;
}

// copy the snapshot back into the buffers that are live now (the trial passes swap them)
int snapshot_restore(lf_ctx *c, Snapshot &sn) {
    for (size_t u = 0; u < sn.src.size(); ++u)
        LF_HIP(hipMemcpyAsync(*sn.src[u].first, sn.copy[u], sn.src[u].second * sizeof(float),
                              hipMemcpyDeviceToDevice, c->stream));
    LF_HIP(hipStreamSynchronize(c->stream));
    for (float *d : sn.copy) (void)hipFree(d);
    sn.copy.clear();
    return 0;
}

}  // namespace

// Picks the temporal-block depth and x chunk (then, in 2D, the pass class launched first)
// with the shortest measured time per step on this context (its grid, CPML slabs, source,
// drive and fused detector boxes): every candidate runs a warm-up pass and `reps` timed
// passes of monitor-free copies of `row`, then every field and psi slab is restored, so
// the simulation state is unchanged.  The blocked passes give bit-identical results for
// every depth, chunk and launch order, so the choice affects only speed.  No counterpart
// in the reference (its step loop has no blocking).
extern "C" int lf_autotune(lf_ctx *c, const lf_step *row, int reps, int *depth, int *lx,
                           int *lead, double *ms_per_step) {
    if (int e = check_ctx(c)) return e;
    if (!c->finalized) return fail(-1, "lf_autotune before lf_finalize");
    if (!row) return fail(-1, "lf_autotune: null step row");
    if (reps < 1) reps = 1;
    std::vector<int> Ts, lxs;
    if (c->kind == LF_YEE3D) { Ts = {2, 3}; lxs = {32, 48, 64, 96}; }
    else { Ts = {2, 4, 6, 8}; lxs = {32, 64, 128}; }
    const int tb0 = c->tb, lx0 = c->tb_lx, lead0 = c->tb_lead;
    const bool timing0 = c->timing;
    c->timing = false;
    Snapshot sn;
    if (int e = snapshot_take(c, sn)) return e;
    hipEvent_t e0, e1;
    LF_HIP(hipEventCreate(&e0));
    LF_HIP(hipEventCreate(&e1));
    int bestT = tb0, bestLx = lx0, bestLead = 0;
    double best = 1e30;
    auto trial = [&](int T, int L, int ld, double &per, int nrep) -> int {
        std::vector<lf_step> tab((size_t)T * (nrep + 1), *row);
        for (auto &t : tab) t.mask = 0;
        c->tb = T;
        c->tb_lx = L;
        c->tb_lead = ld;
        if (int r = run_impl(c, tab.data(), T, true)) return r;          // warm-up pass
        if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(-2, "lf_autotune: event");
        if (int r = run_impl(c, tab.data() + T, T * nrep, true)) return r;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return fail(-2, "lf_autotune: event");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        per = ms / (double)(T * nrep);
        return 0;
    };
    int rc = 0;
    for (int T : Ts) {
        for (int L : lxs) {
            double per = 0;
            if ((rc = trial(T, L, 0, per, reps)) != 0) break;
            if (per < best) { best = per; bestT = T; bestLx = L; }
        }
        if (rc) break;
    }
    // 2D: the launch order of the pass classes at the chosen depth and chunk, and the
    // serial order on one stream (no events: the choice where queuing a pass costs the
    // host more than the device needs for it).  The orders differ by a few percent, about
    // the noise of one short trial, so every order is timed twice, interleaved, over three
    // times the passes, and keeps its faster time.
    if (rc == 0 && c->kind != LF_YEE3D && bestT > 1) {
        double lt[kTBSerial + 1];
        for (int u = 0; u <= kTBSerial; ++u) lt[u] = 1e30;
        for (int round = 0; round < 2 && rc == 0; ++round)
            for (int ld = 0; ld <= kTBSerial; ++ld) {
                double per = 0;
                if ((rc = trial(bestT, bestLx, ld, per, 3 * reps)) != 0) break;
                lt[ld] = std::min(lt[ld], per);
            }
        if (rc == 0) {
            best = lt[0];
            for (int ld = 1; ld <= kTBSerial; ++ld)
                if (lt[ld] < best) { best = lt[ld]; bestLead = ld; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int rr = snapshot_restore(c, sn);
    c->timing = timing0;
    if (rc) { c->tb = tb0; c->tb_lx = lx0; c->tb_lead = lead0; return rc; }
    if (rr) return rr;
    c->tb = bestT;
    c->tb_lx = bestLx;
    c->tb_lead = bestLead;
    if (depth) *depth = bestT;
    if (lx) *lx = bestLx;
    if (lead) *lead = bestLead;
    if (ms_per_step) *ms_per_step = best;
    return 0;
}

// ============================================================================ batch engine (C ABI)
struct lf_batch {
    std::vector<lf_ctx *> ctx;
    hipStream_t stream = nullptr, side[2] = {};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {};
    int tb = 6, tb_lx = 64;
    // per-pass argument records: a ring of pinned host / device slot pairs; a slot is
    // refilled only after its previous copy has completed (event), and a device slot is
    // rewritten only kRing passes later, after that pass's kernels (stream order)
    static constexpr int kRing = 4;
    struct Stage { char *h = nullptr; char *d = nullptr; size_t cap = 0; hipEvent_t done = nullptr;
                   bool used = false; };
    Stage stage[kRing];
    int ring = 0;
    lf_step *d_steps = nullptr;             // resident step table (device) and its host copy
    size_t steps_cap = 0;
    std::vector<lf_step> h_steps;
    std::vector<lf_step> slice;
    std::vector<hipEvent_t> inst_ev;        // per instance: end of its own-stream work
    hipEvent_t pass_ev = nullptr;           // end of the last blocked pass
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double last_ms = 0.0, last_pass_ms = 0.0;
    int last_passes = 0;
};

namespace {

int batch_check(lf_batch *b) { return b ? 0 : fail(-1, "null batch"); }

template <int T>
void launch_batch_class(int kind, int mode, const void *rec, const int *start, int ninst, int nwave,
                        int fp64, hipStream_t st) {
    const unsigned nb = (unsigned)(((long long)nwave + 3) / 4);
    if (kind == LF_TM2D) {
        const TBRecTM *r = (const TBRecTM *)rec;
        if (mode == 0) tm_tb_batch_kernel<T, 0><<<nb, 256, 0, st>>>(r, start, ninst, fp64);
        else if (mode == 1) tm_tb_batch_kernel<T, 1><<<nb, 256, 0, st>>>(r, start, ninst, fp64);
        else tm_tb_batch_kernel<T, 3><<<nb, 256, 0, st>>>(r, start, ninst, fp64);
    } else {
        const TBRecTE *r = (const TBRecTE *)rec;
        if (mode == 0) te_tb_batch_kernel<T, 0><<<nb, 256, 0, st>>>(r, start, ninst, fp64);
        else if (mode == 1) te_tb_batch_kernel<T, 1><<<nb, 256, 0, st>>>(r, start, ninst, fp64);
        else te_tb_batch_kernel<T, 3><<<nb, 256, 0, st>>>(r, start, ninst, fp64);
    }
}

// MODE of a 2D pass class (classes 0 and 3 are both slab-free)
int pl_mode(int cls) { return cls == 1 ? 1 : (cls == 2 ? 3 : 0); }

// Host-side checks of one batch record before it goes to the device: the buffers the
// body reads and writes exist, its rectangles lie in the grid and match their wave
// counts, the slab-free rectangles (MODE 0, unclamped loads) keep every loaded plane and
// lane inside 1..N-2 and outside the slabs, and its T steps lie inside the resident table.
template <typename A>
int check_batch_rec(const lf_batch *b, const lf_ctx *c, const A &a, const TBClass &k, int T,
                    const lf_step *steps) {
    const TBPlanIn in = tb_plan_in(c);
    const int Nx = c->g.nx, Ny = c->g.nz, W = 64 - 2 * T;
    const int ntab = (int)b->h_steps.size();
    if (steps < b->d_steps || steps + T > b->d_steps + ntab)
        return fail(-1, "batch pass: step rows outside the resident table");
    for (int q = 0; q < 4; ++q)
        if ((a.s0[q].psi == nullptr) != (a.s1[q].psi == nullptr))
            return fail(-1, "batch pass: psi slab without its second buffer");
    int prev = 0;
    for (int g = 0; g < k.nseg; ++g) {
        const int x0 = k.x0[g], x1 = k.x1[g], L = k.lx[g];
        if (x0 < 0 || x1 > Nx || x1 <= x0 || L < 1 || k.nwy[g] < 1 || k.wy0[g] < 0 ||
            (k.wy0[g] + k.nwy[g] - 1) * W - T >= Ny)
            return fail(-1, "batch pass: rectangle outside the grid");
        if ((long long)(x1 - x0 + L - 1) / L * k.nwy[g] != (long long)k.end[g] - prev)
            return fail(-1, "batch pass: rectangle wave count");
        prev = k.end[g];
        if (k.mode == 0) {
            const int jlo = k.wy0[g] * W - T, jhi = (k.wy0[g] + k.nwy[g] - 1) * W - T + 63;
            if (x0 - T < std::max(in.xlo, 1) || x1 + T - 1 > std::min(in.xhi, Nx - 2) ||
                jlo < std::max(in.ylo, 1) || jhi > std::min(in.yhi, Ny - 2) || jhi >= c->g.P)
                return fail(-1, "batch pass: slab-free rectangle reaches a slab or an edge");
        }
    }
    return 0;
}

// one blocked pass of depth T over the listed instances; dsteps[k]: device step table
// position of ids[k] for this pass
int batch_pass(lf_batch *b, const std::vector<int> &ids, const std::vector<const lf_step *> &dsteps,
               int T, int fp64) {
    if (T != 2 && T != 4 && T != 6 && T != 8) return fail(-1, "batch pass: depth must be 2, 4, 6 or 8");
    struct Launch { int kind, cls, mode, ninst, nwave; size_t rec_off, start_off; };
    std::vector<Launch> launches;
    std::vector<char> pack;                        // records and prefix tables of the pass
    auto align = [&](size_t a) { while (pack.size() % a) pack.push_back(0); };
    for (int kind : {LF_TE2D, LF_TM2D}) {
        for (int cls = 0; cls < kTBClasses; ++cls) {
            std::vector<int> start(1, 0);
            std::vector<TBRecTM> rtm;
            std::vector<TBRecTE> rte;
            const int mode = pl_mode(cls);
            for (size_t n = 0; n < ids.size(); ++n) {
                lf_ctx *c = b->ctx[ids[n]];
                if (c->kind != kind) continue;
                const TBPlan pl = tb_plan_2d(c, T, b->tb_lx);
                const TBClass &k = pl.cls[cls];
                if (k.nseg == 0) continue;
                if (kind == LF_TM2D) {
                    TBRecTM r{};
                    r.a = tm_tb_args(c);
                    set_class(r.a, k);
                    r.x = c->ex;
                    r.steps = dsteps[n];
                    if (int e = check_batch_rec(b, c, r.a, k, T, r.steps)) return e;
                    rtm.push_back(r);
                } else {
                    TBRecTE r{};
                    r.a = te_tb_args(c);
                    set_class(r.a, k);
                    r.x = c->ex;
                    r.steps = dsteps[n];
                    if (int e = check_batch_rec(b, c, r.a, k, T, r.steps)) return e;
                    rte.push_back(r);
                }
                if ((long long)start.back() + k.end[k.nseg - 1] > (1ll << 30))
                    return fail(-1, "batch pass: too many wavefronts in one launch");
                start.push_back(start.back() + k.end[k.nseg - 1]);
            }
            const int ninst = (int)start.size() - 1;
            if (ninst == 0) continue;
            Launch L{kind, cls, mode, ninst, start.back(), 0, 0};
            align(256);
            L.rec_off = pack.size();
            const char *rp = kind == LF_TM2D ? (const char *)rtm.data() : (const char *)rte.data();
            const size_t rb = kind == LF_TM2D ? rtm.size() * sizeof(TBRecTM) : rte.size() * sizeof(TBRecTE);
            pack.insert(pack.end(), rp, rp + rb);
            align(256);
            L.start_off = pack.size();
            pack.insert(pack.end(), (const char *)start.data(), (const char *)(start.data() + start.size()));
            launches.push_back(L);
        }
    }
    if (launches.empty()) return 0;
    lf_batch::Stage &sg = b->stage[b->ring];
    b->ring = (b->ring + 1) % lf_batch::kRing;
    if (sg.used) LF_HIP(hipEventSynchronize(sg.done));    // its previous copy has completed
    if (pack.size() > sg.cap) {
        // the slot's previous pass may still be reading its device records: every pass
        // so far has joined the batch stream, so once that stream drains nothing reads it
        LF_HIP(hipStreamSynchronize(b->stream));
        if (sg.h) LF_HIP(hipHostFree(sg.h));
        if (sg.d) LF_HIP(hipFree(sg.d));
        sg.h = nullptr; sg.d = nullptr; sg.cap = 0;
        const size_t cap = pack.size() * 2;
        LF_HIP(hipHostMalloc((void **)&sg.h, cap, hipHostMallocDefault));
        LF_HIP(hipMalloc((void **)&sg.d, cap));
        sg.cap = cap;
    }
    if (!sg.done) LF_HIP(hipEventCreateWithFlags(&sg.done, hipEventDisableTiming));
    std::memcpy(sg.h, pack.data(), pack.size());
    LF_HIP(hipMemcpyAsync(sg.d, sg.h, pack.size(), hipMemcpyHostToDevice, b->stream));
    LF_HIP(hipEventRecord(sg.done, b->stream));
    sg.used = true;
    LF_HIP(hipEventRecord(b->ev_fork, b->stream));
    for (int u = 0; u < 2; ++u) LF_HIP(hipStreamWaitEvent(b->side[u], b->ev_fork, 0));
    for (const Launch &L : launches) {
        hipStream_t st = (L.cls == 0 || L.cls == 3) ? b->stream : b->side[L.cls - 1];
        const void *rec = sg.d + L.rec_off;
        const int *start = (const int *)(sg.d + L.start_off);
        // every depth has its own kernels: a pass planned for depth T must run the depth-T
        // body, whose lanes and halo match the plan's wave layout (64 - 2T owned lanes)
        switch (T) {
            case 2: launch_batch_class<2>(L.kind, L.mode, rec, start, L.ninst, L.nwave, fp64, st); break;
            case 4: launch_batch_class<4>(L.kind, L.mode, rec, start, L.ninst, L.nwave, fp64, st); break;
            case 6: launch_batch_class<6>(L.kind, L.mode, rec, start, L.ninst, L.nwave, fp64, st); break;
            case 8: launch_batch_class<8>(L.kind, L.mode, rec, start, L.ninst, L.nwave, fp64, st); break;
            default: return fail(-1, "batch pass: depth must be 2, 4, 6 or 8");
        }
    }
    LF_HIP(hipGetLastError());
    for (int u = 0; u < 2; ++u) {
        LF_HIP(hipEventRecord(b->ev_join[u], b->side[u]));
        LF_HIP(hipStreamWaitEvent(b->stream, b->ev_join[u], 0));
    }
    for (int id : ids) swap_tb_buffers(b->ctx[id]);
    return 0;
}

}  // namespace

extern "C" {

int lf_batch_create(lf_batch **out) {
    if (!out) return fail(-1, "null out");
    lf_batch *b = new lf_batch();
    hipError_t e = hipStreamCreate(&b->stream);
    for (int u = 0; u < 2 && e == hipSuccess; ++u) e = hipStreamCreate(&b->side[u]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    for (int u = 0; u < 2 && e == hipSuccess; ++u) e = hipEventCreateWithFlags(&b->ev_join[u], hipEventDisableTiming);
    if (e != hipSuccess) { delete b; return fail(-2, hipGetErrorString(e)); }
    *out = b;
    return 0;
}

int lf_batch_add(lf_batch *b, lf_ctx *c) {
    if (int e = batch_check(b)) return e;
    if (int e = check_ctx(c)) return e;
    if (!c->finalized) return fail(-1, "lf_batch_add: context not finalized");
    if (c->kind == LF_YEE3D) return fail(-1, "lf_batch_add: 2D contexts only");
    if (c->sp_n > 0 || !c->mur.empty())
        return fail(-1, "lf_batch_add: contexts with Mur faces run alone (lf_run)");
    if (!c->tfsf.empty() || c->nrec > 0)
        return fail(-1, "lf_batch_add: contexts with TF/SF corrections or plane records run alone");
    if (c->mode != (b->ctx.empty() ? c->mode : b->ctx[0]->mode))
        return fail(-1, "lf_batch_add: all instances need the same precision mode");
    for (lf_ctx *o : b->ctx) if (o == c) return fail(-1, "lf_batch_add: context already in batch");
    LF_HIP(hipStreamSynchronize(c->stream));
    b->ctx.push_back(c);
    return 0;
}

int lf_batch_set_temporal(lf_batch *b, int depth, int lx) {
    if (int e = batch_check(b)) return e;
    if (depth != 1 && depth != 2 && depth != 4 && depth != 6 && depth != 8)
        return fail(-1, "lf_batch_set_temporal: depth must be 1 (off), 2, 4, 6 or 8");
    if (lx < 1 || lx > 1 << 20) return fail(-1, "lf_batch_set_temporal: lx out of range");
    b->tb = depth;
    b->tb_lx = lx;
    return 0;
}

int lf_batch_set_timing(lf_batch *b, int on) {
    if (int e = batch_check(b)) return e;
    b->timing = on != 0;
    return 0;
}

/* Resident step table of the batch (device copy), used by lf_batch_run_at. */
int lf_batch_set_table(lf_batch *b, const lf_step *table, int ntable) {
    if (int e = batch_check(b)) return e;
    if (ntable <= 0 || !table) return fail(-1, "lf_batch_set_table: empty table");
    if ((size_t)ntable > b->steps_cap) {
        if (b->d_steps) LF_HIP(hipFree(b->d_steps));
        LF_HIP(hipMalloc(&b->d_steps, ntable * sizeof(lf_step)));
        b->steps_cap = ntable;
    }
    LF_HIP(hipMemcpy(b->d_steps, table, ntable * sizeof(lf_step), hipMemcpyHostToDevice));
    b->h_steps.assign(table, table + ntable);
    return 0;
}

/* Advance instance i over table[offs[i] + s], s < nact[i] <= nsteps, of the resident
 * table; masks (ninst x nsteps, may be NULL: table masks) are per-instance monitor masks
 * indexed by s. */
int lf_batch_run_at(lf_batch *b, const int *offs, int nsteps, const unsigned *masks, const int *nact) {
    if (int e = batch_check(b)) return e;
    if (nsteps <= 0 || b->ctx.empty()) return 0;
    const int ni = (int)b->ctx.size();
    const int ntab = (int)b->h_steps.size();
    for (int i = 0; i < ni; ++i) {
        const int o = offs ? offs[i] : 0;
        const int na = nact ? std::min(nact[i], nsteps) : nsteps;
        if (o < 0 || o + na > ntab) return fail(-1, "lf_batch_run_at: steps outside the table");
    }
    for (lf_ctx *c : b->ctx) {
        if (c->ch == 0.0f) return fail(-1, "lf_batch_run: Ch not set");
        if (c->kind == LF_TE2D && (!c->arr[LF_CEE] || !c->arr[LF_CEH]))
            return fail(-1, "lf_batch_run: Ce_E/Ce_H not uploaded");
        if (c->kind == LF_TM2D && (!c->arr[LF_CAX] || !c->arr[LF_CAY]))
            return fail(-1, "lf_batch_run: Ca_ex/Ca_ey not uploaded");
        if (int r = ensure_tb_buffers(c)) return r;
    }
    const int fp64 = b->ctx[0]->mode == LF_MODE_FP64ACC;
    for (int i = 0; i < ni; ++i) LF_HIP(hipStreamSynchronize(b->ctx[i]->stream));
    // one-step work runs on each instance's own stream (instances overlap each other and
    // the blocked passes); where an instance moves between its stream and the batch
    // streams, an event orders the two
    if ((int)b->inst_ev.size() < ni) {
        for (int i = (int)b->inst_ev.size(); i < ni; ++i) {
            hipEvent_t e;
            LF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->inst_ev.push_back(e);
        }
    }
    if (!b->pass_ev) LF_HIP(hipEventCreateWithFlags(&b->pass_ev, hipEventDisableTiming));
    std::vector<int> where(ni, 0);                 // 0 idle, 1 own stream, 2 batch streams
    auto off = [&](int i) { return offs ? offs[i] : 0; };
    auto mask_of = [&](int i, int s) {
        return masks ? masks[(size_t)i * nsteps + s] : b->h_steps[off(i) + s].mask;
    };
    auto act = [&](int i) { return nact ? std::min(nact[i], nsteps) : nsteps; };
    const bool timing = b->timing;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
        while (b->ev.size() < 2) { hipEvent_t e; LF_HIP(hipEventCreate(&e)); b->ev.push_back(e); }
        t0 = b->ev[0]; t1 = b->ev[1];
        LF_HIP(hipEventRecord(t0, b->stream));
    }
    int rc = 0, passes = 0;
    int s = 0;
    bool pass_pending = false;                     // pass_ev marks the last pass's end
    std::vector<int> blocked, single;
    std::vector<const lf_step *> dst;
    while (s < nsteps && rc == 0) {
        const int T = b->tb;
        blocked.clear(); single.clear();
        for (int i = 0; i < ni; ++i) {
            if (s >= act(i)) continue;
            bool q = T > 1 && s + T <= act(i);
            for (int u = 0; q && u < T; ++u) q = mask_of(i, s + u) == 0;
            (q ? blocked : single).push_back(i);
        }
        if (blocked.empty() && single.empty()) break;
        const int adv = blocked.empty() ? 1 : T;
        for (int i : single) {                     // one step per launch, own stream, own masks
            lf_ctx *c = b->ctx[i];
            if (where[i] == 2 && pass_pending) LF_HIP(hipStreamWaitEvent(c->stream, b->pass_ev, 0));
            const int n = std::min(adv, act(i) - s);
            b->slice.assign(b->h_steps.begin() + off(i) + s, b->h_steps.begin() + off(i) + s + n);
            for (int u = 0; u < n; ++u) b->slice[u].mask = mask_of(i, s + u);
            if ((rc = run_impl(c, b->slice.data(), n, false)) != 0) break;
            where[i] = 1;
        }
        if (rc == 0 && !blocked.empty()) {
            for (int i : blocked)
                if (where[i] == 1) {
                    LF_HIP(hipEventRecord(b->inst_ev[i], b->ctx[i]->stream));
                    LF_HIP(hipStreamWaitEvent(b->stream, b->inst_ev[i], 0));
                }
            dst.clear();
            for (int i : blocked) dst.push_back(b->d_steps + off(i) + s);
            rc = batch_pass(b, blocked, dst, T, fp64);
            if (rc == 0) {
                LF_HIP(hipEventRecord(b->pass_ev, b->stream));
                pass_pending = true;
                for (int i : blocked) where[i] = 2;
            }
            ++passes;
        }
        s += adv;
    }
    for (int i = 0; i < ni && rc == 0; ++i)         // join every instance stream
        if (where[i] == 1) {
            LF_HIP(hipEventRecord(b->inst_ev[i], b->ctx[i]->stream));
            LF_HIP(hipStreamWaitEvent(b->stream, b->inst_ev[i], 0));
        }
    if (timing && rc == 0) {
        LF_HIP(hipEventRecord(t1, b->stream));
        LF_HIP(hipEventSynchronize(t1));
        float ms = 0.f;
        LF_HIP(hipEventElapsedTime(&ms, t0, t1));
        b->last_ms = ms;
        b->last_passes = passes;
    }
    if (rc == 0) LF_HIP(hipStreamSynchronize(b->stream));
    return rc;
}

/* Advance every instance over the shared step table `steps` (uploaded for this call). */
int lf_batch_run(lf_batch *b, const lf_step *steps, int nsteps, const unsigned *masks,
                 const int *nact) {
    if (int e = batch_check(b)) return e;
    if (nsteps <= 0 || b->ctx.empty()) return 0;
    if (int e = lf_batch_set_table(b, steps, nsteps)) return e;
    return lf_batch_run_at(b, nullptr, nsteps, masks, nact);
}

int lf_batch_stats(lf_batch *b, double *ms, int *passes) {
    if (int e = batch_check(b)) return e;
    *ms = b->last_ms;
    *passes = b->last_passes;
    return 0;
}

int lf_batch_destroy(lf_batch *b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &sg : b->stage) {
        if (sg.h) (void)hipHostFree(sg.h);
        if (sg.d) (void)hipFree(sg.d);
        if (sg.done) (void)hipEventDestroy(sg.done);
    }
    if (b->d_steps) (void)hipFree(b->d_steps);
    for (auto e : b->ev) (void)hipEventDestroy(e);
    for (auto e : b->inst_ev) (void)hipEventDestroy(e);
    if (b->pass_ev) (void)hipEventDestroy(b->pass_ev);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (auto e : b->ev_join) if (e) (void)hipEventDestroy(e);
    for (auto st : b->side) if (st) (void)hipStreamDestroy(st);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    (void)hipGetLastError();
    return 0;
}

}  // extern "C"
