// ops.hip - operator-level CPML updates on caller-owned device arrays in the reference's
// own C-order layouts (the per-operation drop-in for planar_beads/fused2d.py and
// planar_3d/boundaries3d.py:CPML_3D).  The loop-level engine (lf_run) is the fast path;
// these entry points serve callers that keep their own time loop over device arrays.
//
// 2D (fused2d.py:41-94, boundaries.py:186-268), one work-item per updated element,
// consecutive y elements in consecutive lanes:
//   TE: Ez (Nx,Ny), Hx (Nx,Ny-1), Hy (Nx-1,Ny); psi_Hxy (Nx,Ny-1), psi_Hyx (Nx-1,Ny),
//       psi_Ezx = psi_Ezy (Nx-2,Ny-2)
//   TM: Hz (Nx-1,Ny-1), Ex (Nx-1,Ny), Ey (Nx,Ny-1); psi_Hzx = psi_Hzy (Nx-1,Ny-1),
//       psi_Exy (Nx-1,Ny-2), psi_Eyx (Nx-2,Ny-1)
// 3D (boundaries3d.py:311-362): Ex (Nx-1,Ny,Nz), Ey (Nx,Ny-1,Nz), Ez (Nx,Ny,Nz-1),
//   Hx (Nx,Ny-1,Nz-1), Hy (Nx-1,Ny,Nz-1), Hz (Nx-1,Ny-1,Nz); psi in the PML slabs only,
//   owned by an lf_op3d object.
// Per element the float32 operations are the reference's, in its order (no contraction).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "lensfdtd_dev.h"

using namespace lfd;

namespace {

int op_fail(const char *m) {
    set_error(m);
    return -1;
}

#define OP_HIP(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return op_fail(hipGetErrorString(e_));                       \
    } while (0)

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

// ------------------------------------------------------------------------------ 2D TE
__global__ void __launch_bounds__(256) op_te_hx(lf_op2d p, const float *Ez, float *Hx, float ch,
                                                int Nx, int Ny) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)Nx * (Ny - 1)) return;
    const int a = (int)(i / (Ny - 1)), j = (int)(i - (long long)a * (Ny - 1));
    const float d = Ez[(long long)a * Ny + j + 1] - Ez[(long long)a * Ny + j];
    const float q = p.bhy[j] * p.psi0[i] + p.chy[j] * d;      // fused2d.py:44-48
    p.psi0[i] = q;
    Hx[i] = Hx[i] - ch * (d / p.khy[j] + q);
}

__global__ void __launch_bounds__(256) op_te_hy(lf_op2d p, const float *Ez, float *Hy, float ch,
                                                int Nx, int Ny) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)(Nx - 1) * Ny) return;
    const int a = (int)(i / Ny), j = (int)(i - (long long)a * Ny);
    const float d = Ez[(long long)(a + 1) * Ny + j] - Ez[(long long)a * Ny + j];
    const float q = p.bhx[a] * p.psi1[i] + p.chx[a] * d;      // fused2d.py:50-55
    p.psi1[i] = q;
    Hy[i] = Hy[i] + ch * (d / p.khx[a] + q);
}

__global__ void __launch_bounds__(256) op_te_ez(lf_op2d p, float *Ez, const float *Hx, const float *Hy,
                                                const float *CeE, const float *CeH, int Nx, int Ny) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)(Nx - 2) * (Ny - 2)) return;
    const int a = (int)(i / (Ny - 2)), j = (int)(i - (long long)a * (Ny - 2));
    const int ga = a + 1, gb = j + 1;                         // fused2d.py:57-68
    const float dHy = Hy[(long long)ga * Ny + gb] - Hy[(long long)a * Ny + gb];
    const float dHx = Hx[(long long)ga * (Ny - 1) + gb] - Hx[(long long)ga * (Ny - 1) + j];
    const float qx = p.bex[ga] * p.psi0[i] + p.cex[ga] * dHy;
    p.psi0[i] = qx;
    const float qy = p.bey[gb] * p.psi1[i] + p.cey[gb] * dHx;
    p.psi1[i] = qy;
    const long long g = (long long)ga * Ny + gb;
    Ez[g] = CeE[g] * Ez[g] + CeH[g] * (((dHy / p.kex[ga] + qx) - dHx / p.key[gb]) - qy);
}

// ------------------------------------------------------------------------------ 2D TM
__global__ void __launch_bounds__(256) op_tm_hz(lf_op2d p, const float *Ex, const float *Ey, float *Hz,
                                                float ch, int Nx, int Ny) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)(Nx - 1) * (Ny - 1)) return;
    const int a = (int)(i / (Ny - 1)), j = (int)(i - (long long)a * (Ny - 1));
    const float dEyx = Ey[(long long)(a + 1) * (Ny - 1) + j] - Ey[(long long)a * (Ny - 1) + j];
    const float dExy = Ex[(long long)a * Ny + j + 1] - Ex[(long long)a * Ny + j];
    const float qx = p.bhx[a] * p.psi0[i] + p.chx[a] * dEyx;  // fused2d.py:69-78
    p.psi0[i] = qx;
    const float qy = p.bhy[j] * p.psi1[i] + p.chy[j] * dExy;
    p.psi1[i] = qy;
    Hz[i] = Hz[i] - ch * ((dEyx / p.khx[a] + qx) - (dExy / p.khy[j] + qy));
}

__global__ void __launch_bounds__(256) op_tm_ex(lf_op2d p, float *Ex, const float *Hz, const float *Ca,
                                                int Nx, int Ny) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)(Nx - 1) * (Ny - 2)) return;
    const int a = (int)(i / (Ny - 2)), j = (int)(i - (long long)a * (Ny - 2));
    const float d = Hz[(long long)a * (Ny - 1) + j + 1] - Hz[(long long)a * (Ny - 1) + j];
    const float q = p.bey[j + 1] * p.psi0[i] + p.cey[j + 1] * d;   // fused2d.py:80-86
    p.psi0[i] = q;
    const long long g = (long long)a * Ny + j + 1;
    Ex[g] = Ex[g] + Ca[g] * (d / p.key[j + 1] + q);
}

__global__ void __launch_bounds__(256) op_tm_ey(lf_op2d p, float *Ey, const float *Hz, const float *Ca,
                                                int Nx, int Ny) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)(Nx - 2) * (Ny - 1)) return;
    const int a = (int)(i / (Ny - 1)), j = (int)(i - (long long)a * (Ny - 1));
    const float d = Hz[(long long)(a + 1) * (Ny - 1) + j] - Hz[(long long)a * (Ny - 1) + j];
    const float q = p.bex[a + 1] * p.psi1[i] + p.cex[a + 1] * d;   // fused2d.py:87-94
    p.psi1[i] = q;
    const long long g = (long long)(a + 1) * (Ny - 1) + j;
    Ey[g] = Ey[g] + (-Ca[g]) * (d / p.kex[a + 1] + q);
}

// PEC on the PML ends (boundaries.py:207-215, :260-268): one work-item per edge element
__global__ void __launch_bounds__(256) op_pec2d(float *A, int n0, int n1, int lo0, int hi0,
                                                float *B, int m0, int m1, int lo1, int hi1) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n1) {                                           // rows 0 and n0-1 of A
        if (lo0) A[t] = 0.0f;
        if (hi0) A[(long long)(n0 - 1) * n1 + t] = 0.0f;
    }
    if (t < m0) {                                           // columns 0 and m1-1 of B
        if (lo1) B[(long long)t * m1] = 0.0f;
        if (hi1) B[(long long)t * m1 + m1 - 1] = 0.0f;
    }
}

bool op2d_ok(const lf_op2d *p, int Nx, int Ny) {
    return p && Nx >= 3 && Ny >= 3 && p->bex && p->cex && p->kex && p->bhx && p->chx && p->khx &&
           p->bey && p->cey && p->key && p->bhy && p->chy && p->khy && p->psi0 && p->psi1;
}

}  // namespace

// ------------------------------------------------------------------------------ 3D
struct lf_op3d {
    int nx, ny, nz;
    float *prof[3][2][3] = {};        // axis, (E nodes, H nodes), (b, c, 1/kappa - 1)
    Strip sH[6], sE[6];               // Hx_z Hx_y Hy_x Hy_z Hz_y Hz_x ; Ex_y Ex_z Ey_z Ey_x Ez_x Ez_y
    int pec[6];                       // xlo xhi ylo yhi zlo zhi
    std::vector<float> cprof[3][2];   // host c profiles (slab extents)
    bool have[3] = {false, false, false};
    bool finalized = false;
    std::vector<void *> bufs;
};

namespace {

struct Op3 {
    int nx, ny, nz;
    const float *pb[3][2], *pc[3][2], *pm[3][2];
    Strip sH[6], sE[6];
    int pec[6];
};

// psi element of a slab along `axis` for a component of dims (n0, n1, n2)
__device__ __forceinline__ float *psi3(const Strip &s, int axis, int pos, int i, int j, int k, int n1,
                                       int n2) {
    if (axis == 0) return s.psi + ((long long)pos * n1 + j) * n2 + k;
    if (axis == 1) return s.psi + ((long long)i * s.len + pos) * n2 + k;
    return s.psi + ((long long)i * n1 + j) * s.len + pos;
}

// psi <- b psi + c d (in place); returns psi + (1/kappa - 1) d   (boundaries3d.py:240-251)
__device__ __forceinline__ float psi3_term(const Op3 &o, const Strip &s, int axis, int node, int pos,
                                           int i, int j, int k, int n1, int n2, int half, float d) {
    float *pp = psi3(s, axis, pos, i, j, k, n1, n2);
    float p = (*pp) * o.pb[axis][half][node];
    p = p + o.pc[axis][half][node] * d;
    *pp = p;
    return p + o.pm[axis][half][node] * d;
}

__global__ void __launch_bounds__(256) op3d_h(Op3 o, const float *Ex, const float *Ey, const float *Ez,
                                              float *Hx, float *Hy, float *Hz, float ch) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    const int NX = o.nx, NY = o.ny, NZ = o.nz;
    if (k >= NZ || j >= NY) return;
    if (j < NY - 1 && k < NZ - 1) {                    // Hx (Nx, Ny-1, Nz-1): :316-320
        const float d1 = Ey[((long long)i * (NY - 1) + j) * NZ + k + 1] - Ey[((long long)i * (NY - 1) + j) * NZ + k];
        const float d2 = Ez[((long long)i * NY + j + 1) * (NZ - 1) + k] - Ez[((long long)i * NY + j) * (NZ - 1) + k];
        const long long c = ((long long)i * (NY - 1) + j) * (NZ - 1) + k;
        float h = Hx[c] + ch * (d1 - d2);
        int s = strip_pos(o.sH[0], k);
        if (s >= 0) h = h + ch * psi3_term(o, o.sH[0], 2, k, s, i, j, k, NY - 1, NZ - 1, 1, d1);
        s = strip_pos(o.sH[1], j);
        if (s >= 0) h = h + (-ch) * psi3_term(o, o.sH[1], 1, j, s, i, j, k, NY - 1, NZ - 1, 1, d2);
        Hx[c] = h;
    }
    if (i < NX - 1 && k < NZ - 1) {                    // Hy (Nx-1, Ny, Nz-1): :321-325
        const float d1 = Ez[((long long)(i + 1) * NY + j) * (NZ - 1) + k] - Ez[((long long)i * NY + j) * (NZ - 1) + k];
        const float d2 = Ex[((long long)i * NY + j) * NZ + k + 1] - Ex[((long long)i * NY + j) * NZ + k];
        const long long c = ((long long)i * NY + j) * (NZ - 1) + k;
        float h = Hy[c] + ch * (d1 - d2);
        int s = strip_pos(o.sH[2], i);
        if (s >= 0) h = h + ch * psi3_term(o, o.sH[2], 0, i, s, i, j, k, NY, NZ - 1, 1, d1);
        s = strip_pos(o.sH[3], k);
        if (s >= 0) h = h + (-ch) * psi3_term(o, o.sH[3], 2, k, s, i, j, k, NY, NZ - 1, 1, d2);
        Hy[c] = h;
    }
    if (i < NX - 1 && j < NY - 1) {                    // Hz (Nx-1, Ny-1, Nz): :326-330
        const float d1 = Ex[((long long)i * NY + j + 1) * NZ + k] - Ex[((long long)i * NY + j) * NZ + k];
        const float d2 = Ey[((long long)(i + 1) * (NY - 1) + j) * NZ + k] - Ey[((long long)i * (NY - 1) + j) * NZ + k];
        const long long c = ((long long)i * (NY - 1) + j) * NZ + k;
        float h = Hz[c] + ch * (d1 - d2);
        int s = strip_pos(o.sH[4], j);
        if (s >= 0) h = h + ch * psi3_term(o, o.sH[4], 1, j, s, i, j, k, NY - 1, NZ, 1, d1);
        s = strip_pos(o.sH[5], i);
        if (s >= 0) h = h + (-ch) * psi3_term(o, o.sH[5], 0, i, s, i, j, k, NY - 1, NZ, 1, d2);
        Hz[c] = h;
    }
}

__global__ void __launch_bounds__(256) op3d_e(Op3 o, float *Ex, float *Ey, float *Ez, const float *Hx,
                                              const float *Hy, const float *Hz, const float *cex,
                                              const float *cey, const float *cez, const float *dax,
                                              const float *day, const float *daz) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    const int NX = o.nx, NY = o.ny, NZ = o.nz;
    if (k >= NZ || j >= NY) return;
    if (i < NX - 1) {                                  // Ex (Nx-1, Ny, Nz): :333-343
        const long long c = ((long long)i * NY + j) * NZ + k;
        float v = Ex[c];
        if (dax) v = v * dax[c];
        if (j >= 1 && j <= NY - 2 && k >= 1 && k <= NZ - 2) {
            const float d1 = Hz[((long long)i * (NY - 1) + j) * NZ + k] - Hz[((long long)i * (NY - 1) + j - 1) * NZ + k];
            const float d2 = Hy[((long long)i * NY + j) * (NZ - 1) + k] - Hy[((long long)i * NY + j) * (NZ - 1) + k - 1];
            const float ce = cex[c];
            v = v + ce * (d1 - d2);
            int s = strip_pos(o.sE[0], j);
            if (s >= 0) v = v + ce * psi3_term(o, o.sE[0], 1, j, s, i, j, k, NY, NZ, 0, d1);
            s = strip_pos(o.sE[1], k);
            if (s >= 0) v = v + (-ce) * psi3_term(o, o.sE[1], 2, k, s, i, j, k, NY, NZ, 0, d2);
        }
        if ((o.pec[2] && j == 0) || (o.pec[3] && j == NY - 1) || (o.pec[4] && k == 0) ||
            (o.pec[5] && k == NZ - 1))
            v = 0.0f;                                  // :356-362
        Ex[c] = v;
    }
    if (j < NY - 1) {                                  // Ey (Nx, Ny-1, Nz): :344-349
        const long long c = ((long long)i * (NY - 1) + j) * NZ + k;
        float v = Ey[c];
        if (day) v = v * day[c];
        if (i >= 1 && i <= NX - 2 && k >= 1 && k <= NZ - 2) {
            const float d1 = Hx[((long long)i * (NY - 1) + j) * (NZ - 1) + k] - Hx[((long long)i * (NY - 1) + j) * (NZ - 1) + k - 1];
            const float d2 = Hz[((long long)i * (NY - 1) + j) * NZ + k] - Hz[((long long)(i - 1) * (NY - 1) + j) * NZ + k];
            const float ce = cey[c];
            v = v + ce * (d1 - d2);
            int s = strip_pos(o.sE[2], k);
            if (s >= 0) v = v + ce * psi3_term(o, o.sE[2], 2, k, s, i, j, k, NY - 1, NZ, 0, d1);
            s = strip_pos(o.sE[3], i);
            if (s >= 0) v = v + (-ce) * psi3_term(o, o.sE[3], 0, i, s, i, j, k, NY - 1, NZ, 0, d2);
        }
        if ((o.pec[0] && i == 0) || (o.pec[1] && i == NX - 1) || (o.pec[4] && k == 0) ||
            (o.pec[5] && k == NZ - 1))
            v = 0.0f;
        Ey[c] = v;
    }
    if (k < NZ - 1) {                                  // Ez (Nx, Ny, Nz-1): :350-355
        const long long c = ((long long)i * NY + j) * (NZ - 1) + k;
        float v = Ez[c];
        if (daz) v = v * daz[c];
        if (i >= 1 && i <= NX - 2 && j >= 1 && j <= NY - 2) {
            const float d1 = Hy[((long long)i * NY + j) * (NZ - 1) + k] - Hy[((long long)(i - 1) * NY + j) * (NZ - 1) + k];
            const float d2 = Hx[((long long)i * (NY - 1) + j) * (NZ - 1) + k] - Hx[((long long)i * (NY - 1) + j - 1) * (NZ - 1) + k];
            const float ce = cez[c];
            v = v + ce * (d1 - d2);
            int s = strip_pos(o.sE[4], i);
            if (s >= 0) v = v + ce * psi3_term(o, o.sE[4], 0, i, s, i, j, k, NY, NZ - 1, 0, d1);
            s = strip_pos(o.sE[5], j);
            if (s >= 0) v = v + (-ce) * psi3_term(o, o.sE[5], 1, j, s, i, j, k, NY, NZ - 1, 0, d2);
        }
        if ((o.pec[0] && i == 0) || (o.pec[1] && i == NX - 1) || (o.pec[2] && j == 0) ||
            (o.pec[3] && j == NY - 1))
            v = 0.0f;
        Ez[c] = v;
    }
}

// slab extent of a c profile over [first, last] (positions < lo or >= hi carry psi)
void slab_range(const std::vector<float> &c, int first, int last, int &lo, int &hi) {
    const int N = (int)c.size(), span = last - first + 1;
    int nlo = 0, nhi = 0;
    while (nlo < span && c[first + nlo] != 0.0f) ++nlo;
    while (nhi < span && c[last - nhi] != 0.0f) ++nhi;
    if (nlo + nhi >= span) { lo = N; hi = N; return; }
    lo = first + nlo;
    hi = last + 1 - nhi;
}

// psi slab of a component of dims d[3] along `axis`
int op3_strip(lf_op3d *o, Strip &s, int axis, int half, const long long d[3], int first, int last) {
    int lo, hi;
    slab_range(o->cprof[axis][half], first, last, lo, hi);
    const int N = (int)o->cprof[axis][half].size();
    s = Strip{nullptr, lo, hi, lo + (N - hi)};
    if (lo == first && hi == last + 1) return 0;        // no PML cell along this axis
    long long e[3] = {d[0], d[1], d[2]};
    e[axis] = s.len;
    float *p = nullptr;
    const size_t n = (size_t)(e[0] * e[1] * e[2]);
    OP_HIP(hipMalloc(&p, n * sizeof(float)));
    OP_HIP(hipMemset(p, 0, n * sizeof(float)));
    o->bufs.push_back(p);
    s.psi = p;
    return 0;
}

Op3 op3_args(const lf_op3d *o) {
    Op3 a{};
    a.nx = o->nx; a.ny = o->ny; a.nz = o->nz;
    for (int ax = 0; ax < 3; ++ax)
        for (int h = 0; h < 2; ++h) {
            a.pb[ax][h] = o->prof[ax][h][0];
            a.pc[ax][h] = o->prof[ax][h][1];
            a.pm[ax][h] = o->prof[ax][h][2];
        }
    for (int q = 0; q < 6; ++q) { a.sH[q] = o->sH[q]; a.sE[q] = o->sE[q]; a.pec[q] = o->pec[q]; }
    return a;
}

}  // namespace

extern "C" {

int lf_op_te_update_H(const lf_op2d *p, int Nx, int Ny, const float *Ez, float *Hx, float *Hy, float ch,
                      void *stream) {
    if (!op2d_ok(p, Nx, Ny) || !Ez || !Hx || !Hy) return op_fail("lf_op_te_update_H: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    lf_op2d a = *p;
    op_te_hx<<<nblk((long long)Nx * (Ny - 1)), 256, 0, st>>>(a, Ez, Hx, ch, Nx, Ny);
    op_te_hy<<<nblk((long long)(Nx - 1) * Ny), 256, 0, st>>>(a, Ez, Hy, ch, Nx, Ny);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op_te_update_E(const lf_op2d *p, int Nx, int Ny, float *Ez, const float *Hx, const float *Hy,
                      const float *CeE, const float *CeH, void *stream) {
    if (!op2d_ok(p, Nx, Ny) || !Ez || !Hx || !Hy || !CeE || !CeH)
        return op_fail("lf_op_te_update_E: bad arguments");
    op_te_ez<<<nblk((long long)(Nx - 2) * (Ny - 2)), 256, 0, (hipStream_t)stream>>>(*p, Ez, Hx, Hy, CeE,
                                                                                    CeH, Nx, Ny);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op_tm_update_H(const lf_op2d *p, int Nx, int Ny, const float *Ex, const float *Ey, float *Hz,
                      float ch, void *stream) {
    if (!op2d_ok(p, Nx, Ny) || !Ex || !Ey || !Hz) return op_fail("lf_op_tm_update_H: bad arguments");
    op_tm_hz<<<nblk((long long)(Nx - 1) * (Ny - 1)), 256, 0, (hipStream_t)stream>>>(*p, Ex, Ey, Hz, ch,
                                                                                    Nx, Ny);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op_tm_update_E(const lf_op2d *p, int Nx, int Ny, float *Ex, float *Ey, const float *Hz,
                      const float *Ca_ex, const float *Ca_ey, void *stream) {
    if (!op2d_ok(p, Nx, Ny) || !Ex || !Ey || !Hz || !Ca_ex || !Ca_ey)
        return op_fail("lf_op_tm_update_E: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    op_tm_ex<<<nblk((long long)(Nx - 1) * (Ny - 2)), 256, 0, st>>>(*p, Ex, Hz, Ca_ex, Nx, Ny);
    op_tm_ey<<<nblk((long long)(Nx - 2) * (Ny - 1)), 256, 0, st>>>(*p, Ey, Hz, Ca_ey, Nx, Ny);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op_te_terminate(int Nx, int Ny, float *Ez, const int pec[4], void *stream) {
    if (!Ez || !pec || Nx < 2 || Ny < 2) return op_fail("lf_op_te_terminate: bad arguments");
    const int n = Nx > Ny ? Nx : Ny;
    op_pec2d<<<nblk(n), 256, 0, (hipStream_t)stream>>>(Ez, Nx, Ny, pec[0], pec[1], Ez, Nx, Ny, pec[2],
                                                       pec[3]);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op_tm_terminate(int Nx, int Ny, float *Ex, float *Ey, const int pec[4], void *stream) {
    if (!Ex || !Ey || !pec || Nx < 2 || Ny < 2) return op_fail("lf_op_tm_terminate: bad arguments");
    const int n = Nx > Ny ? Nx : Ny;
    // Ey (Nx, Ny-1): x ends; Ex (Nx-1, Ny): y ends
    op_pec2d<<<nblk(n), 256, 0, (hipStream_t)stream>>>(Ey, Nx, Ny - 1, pec[0], pec[1], Ex, Nx - 1, Ny,
                                                       pec[2], pec[3]);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op3d_create(int Nx, int Ny, int Nz, lf_op3d **out) {
    if (!out || Nx < 3 || Ny < 3 || Nz < 3) return op_fail("lf_op3d_create: bad arguments");
    lf_op3d *o = new lf_op3d();
    o->nx = Nx; o->ny = Ny; o->nz = Nz;
    for (int q = 0; q < 6; ++q) { o->sH[q] = Strip{nullptr, 0, 0, 0}; o->sE[q] = Strip{nullptr, 0, 0, 0}; o->pec[q] = 0; }
    *out = o;
    return 0;
}

/* Profiles of one axis: e_* on the N nodes, h_* on the N-1 half nodes, each (kappa, b, c)
 * float32 as the reference builds them (boundaries3d.py:150-194); PEC flags of its two
 * faces. */
int lf_op3d_set_cpml(lf_op3d *o, int axis, const float *e_k, const float *e_b, const float *e_c,
                     const float *h_k, const float *h_b, const float *h_c, int pec_lo, int pec_hi) {
    if (!o || axis < 0 || axis > 2 || o->finalized) return op_fail("lf_op3d_set_cpml: bad arguments");
    const int N = axis == 0 ? o->nx : (axis == 1 ? o->ny : o->nz);
    const float *src[2][3] = {{e_k, e_b, e_c}, {h_k, h_b, h_c}};
    for (int h = 0; h < 2; ++h) {
        const int n = N - h;
        std::vector<float> b(src[h][1], src[h][1] + n), c(src[h][2], src[h][2] + n), m(n);
        for (int u = 0; u < n; ++u) m[u] = 1.0f / src[h][0][u] - 1.0f;   // boundaries3d.py:235
        o->cprof[axis][h] = c;
        const std::vector<float> *v[3] = {&b, &c, &m};
        for (int q = 0; q < 3; ++q) {
            if (!o->prof[axis][h][q]) {
                OP_HIP(hipMalloc(&o->prof[axis][h][q], n * sizeof(float)));
                o->bufs.push_back(o->prof[axis][h][q]);
            }
            OP_HIP(hipMemcpy(o->prof[axis][h][q], v[q]->data(), n * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    o->pec[2 * axis] = pec_lo != 0;
    o->pec[2 * axis + 1] = pec_hi != 0;
    o->have[axis] = true;
    return 0;
}

int lf_op3d_finalize(lf_op3d *o) {
    if (!o) return op_fail("lf_op3d_finalize: null");
    if (o->finalized) return 0;
    if (!o->have[0] || !o->have[1] || !o->have[2]) return op_fail("lf_op3d_finalize: CPML profile missing");
    const long long X = o->nx, Y = o->ny, Z = o->nz;
    const long long dHx[3] = {X, Y - 1, Z - 1}, dHy[3] = {X - 1, Y, Z - 1}, dHz[3] = {X - 1, Y - 1, Z};
    const long long dEx[3] = {X - 1, Y, Z}, dEy[3] = {X, Y - 1, Z}, dEz[3] = {X, Y, Z - 1};
    const int N[3] = {o->nx, o->ny, o->nz};
    // H: half-node profiles over the n-1 nodes; E: full-node profiles over [1, n-2]
    struct S { Strip *s; int axis, half; const long long *d; };
    const S all[12] = {{&o->sH[0], 2, 1, dHx}, {&o->sH[1], 1, 1, dHx}, {&o->sH[2], 0, 1, dHy},
                       {&o->sH[3], 2, 1, dHy}, {&o->sH[4], 1, 1, dHz}, {&o->sH[5], 0, 1, dHz},
                       {&o->sE[0], 1, 0, dEx}, {&o->sE[1], 2, 0, dEx}, {&o->sE[2], 2, 0, dEy},
                       {&o->sE[3], 0, 0, dEy}, {&o->sE[4], 0, 0, dEz}, {&o->sE[5], 1, 0, dEz}};
    for (const S &q : all) {
        const int first = q.half ? 0 : 1, last = N[q.axis] - 2;
        if (int r = op3_strip(o, *q.s, q.axis, q.half, q.d, first, last)) return r;
    }
    o->finalized = true;
    return 0;
}

int lf_op3d_step_H(lf_op3d *o, const float *Ex, const float *Ey, const float *Ez, float *Hx, float *Hy,
                   float *Hz, float ch, void *stream) {
    if (!o || !o->finalized || !Ex || !Ey || !Ez || !Hx || !Hy || !Hz)
        return op_fail("lf_op3d_step_H: bad arguments (finalize first)");
    const dim3 g((o->nz + 63) / 64, (o->ny + 3) / 4, o->nx);
    op3d_h<<<g, 256, 0, (hipStream_t)stream>>>(op3_args(o), Ex, Ey, Ez, Hx, Hy, Hz, ch);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op3d_step_E(lf_op3d *o, float *Ex, float *Ey, float *Ez, const float *Hx, const float *Hy,
                   const float *Hz, const float *ce_x, const float *ce_y, const float *ce_z,
                   const float *da_x, const float *da_y, const float *da_z, void *stream) {
    if (!o || !o->finalized || !Ex || !Ey || !Ez || !Hx || !Hy || !Hz || !ce_x || !ce_y || !ce_z)
        return op_fail("lf_op3d_step_E: bad arguments (finalize first)");
    if ((da_x == nullptr) != (da_y == nullptr) || (da_x == nullptr) != (da_z == nullptr))
        return op_fail("lf_op3d_step_E: the loss factors come as a triple");
    const dim3 g((o->nz + 63) / 64, (o->ny + 3) / 4, o->nx);
    op3d_e<<<g, 256, 0, (hipStream_t)stream>>>(op3_args(o), Ex, Ey, Ez, Hx, Hy, Hz, ce_x, ce_y, ce_z,
                                               da_x, da_y, da_z);
    OP_HIP(hipGetLastError());
    return 0;
}

int lf_op3d_destroy(lf_op3d *o) {
    if (!o) return 0;
    (void)hipDeviceSynchronize();
    for (void *p : o->bufs) (void)hipFree(p);
    delete o;
    return 0;
}

}  // extern "C"
