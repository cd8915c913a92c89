// tb2d.h - the temporally blocked 2D passes (TM and TE): launch arguments, plane loads
// and the wavefront bodies, shared by lensfdtd.hip (plain passes, batched passes) and
// tb2d_mon.hip (passes that also sum the CW-window monitors).
#pragma once
#include "lensfdtd_dev.h"

namespace lfd {

// d / k of a CPML slab term.  The profiles' kappa is 1 wherever kappa_max is 1 (the
// default), and d / 1 is d exactly, so the division runs only where kappa differs from 1:
// for the x profiles the test is wave-uniform (a scalar branch), for the y profiles the
// division runs under the lanes that need it.
__device__ __forceinline__ float div_kappa(float d, float k) {
    float t;
    if (k == 1.0f) t = d;
    else t = d / k;
    return t;
}


// ---------------------------------------------------------------------------- temporal blocking (2D TM)
// T time steps per pass.  A wavefront holds 64 consecutive y cells
// j = wy*(64-2T) - T + lane and marches along its x chunk [X0, X1) from X0-T to X1+T-1.
// Iteration p loads plane p (E^n, Hz^{n-1/2}, psi, coefficients) into pipeline slot 0;
// level s (0 <= s < T) then advances plane q = p-1-s from step n+s to n+s+1: it needs
// Ey^{n+s}(q+1) (this iteration's output of level s-1), E^{n+s}(q) and Hz^{n+s-1/2}(q)
// (level s-1's output of the previous iteration) and its own Hz^{n+s+1/2}(q-1).  Every
// step shrinks the valid y range by one cell on each side (Hz needs Ex(j+1), Ex needs
// Hz(j-1)) and the valid x range by one plane on the low side, so after T steps lanes
// T..63-T of planes X0..X1-1 are exact and are stored.  Inputs and outputs (fields and
// all four psi slabs) are separate buffers.  Per cell and step the arithmetic is that of
// tm_h_kernel + tm_e_kernel (boundaries.py:243-258, fdtd2d_core.py:796).
constexpr int kMaxTB = 8;

// Per-cell coefficient arrays of a 2D grid, as the blocked passes read them.  The arrays
// of one context share a map (CoefMap): value (p, j) of an array is band[p * rows + j - j0]
// inside the band of rows [j0, j0 + rows), else element j of table row cls[p] (rows of P
// floats).  A window's planes of one layer column share their table row, so those loads
// hit the L2 and only the band (the y range of the beads) streams from HBM.  Plain arrays
// are the form cls[p] = p, tab = the array, no band.
struct Coef2 {
    const float *tab;
    const float *band;
};
struct CoefMap {
    const int *cls;
    int j0, rows;
};

// where element (p, j) of a context's coefficient arrays sits: in the band (boff) or in
// table row q (toff); j and the band test are fixed per lane, q is read a plane ahead
struct CoefAt {
    bool inb;
    long long boff, toff;
};
__device__ __forceinline__ CoefAt coef_pos(const CoefMap &m, int p, int q, int j, int P) {
    const int r = j - m.j0;
    return CoefAt{(unsigned)r < (unsigned)m.rows, (long long)p * m.rows + r, (long long)q * P + j};
}
__device__ __forceinline__ float coef_at(Coef2 k, const CoefAt &w) {
    return *(w.inb ? k.band + w.boff : k.tab + w.toff);
}

// table row of plane p (clamped to the grid), as a wave-uniform value
__device__ __forceinline__ int coef_row(const CoefMap &m, int p, int Nx) {
    return __builtin_amdgcn_readfirstlane(m.cls[p < Nx ? p : Nx - 1]);
}

struct TBTMArgs {
    Grid g;
    const float *Ex0, *Ey0, *Hz0;
    float *Ex1, *Ey1, *Hz1;
    const float *cax, *cay, *dax, *day;   // plain arrays (dax / day: nullptr when lossless)
    Coef2 kax, kay, kdx, kdy;             // the same arrays as the slab-free bodies read them
    CoefMap cm;
    float ch;
    Prof xe, xh, ye, yh;
    Strip s0[4], s1[4];       // Hzx Hzy Exy Eyx: read / write buffers
    int pec_xlo, pec_xhi, pec_ylo, pec_yhi;
    // wavefronts of this launch: up to 8 rectangles of (x chunk, y wave) indices; wave
    // wv of segment g (seg_end[g-1] <= wv < seg_end[g], loc = wv - seg_end[g-1]) owns
    // the x chunk [X0, min(X0 + seg_lx, seg_x1)) with X0 = seg_x0 + (loc / seg_nwy) *
    // seg_lx, and the y wave seg_wy0 + loc % seg_nwy
    int nseg;
    int seg_x0[8], seg_x1[8], seg_wy0[8], seg_nwy[8], seg_end[8], seg_lx[8];
    unsigned seg_exc;         // bit g: segment g needs the body with excitation code
    unsigned seg_m3;          // bit g: segment g runs the x-slab body (slab launches)
    StepArgs st[kMaxTB];
};

struct TMPlane {              // one plane of one column, as loaded from the step-n buffers
    float ex, ey, hz, ca, cy, da, dy, ph, py, qy, qx;
};

// MODE bit 1: wavefront lanes may lie in a y slab or on a y edge (psi_Hzy, psi_Exy,
// lane tests); bit 2: planes may lie in an x slab or on an x edge (psi_Hzx, psi_Eyx,
// plane tests).  MODE 0 is the slab-free interior: no psi registers, no edge tests.
template <int MODE, bool CMP = false>
__device__ __forceinline__ TMPlane tm_load_plane(const TBTMArgs &a, int p, int q, int j, bool jin, bool jey,
                                                 int sHzy, int sExy, bool lossy_x, bool lossy_y) {
    constexpr bool PX = MODE & 2, PY = MODE & 1;
    TMPlane v{};
    const int Nx = a.g.nx, P = a.g.P;
    if constexpr (MODE == 0) {
        // slab-free chunks: every loaded lane and plane lies inside the grid, so the
        // loads are unconditional (planes past the grid, loaded ahead and never used,
        // read the last plane).  No merge with zero: a loaded register then flows
        // straight into the iteration that uses it, with no copy that would wait for it.
        const long long c = (long long)(p < Nx ? p : Nx - 1) * P + j;
        v.ex = a.Ex0[c];
        if constexpr (CMP) {
            const CoefAt w = coef_pos(a.cm, p < Nx ? p : Nx - 1, q, j, P);
            v.ca = coef_at(a.kax, w);
            v.da = coef_at(lossy_x ? a.kdx : a.kax, w);      // read only when lossy
            v.cy = coef_at(a.kay, w);
            v.dy = coef_at(lossy_y ? a.kdy : a.kay, w);
        } else {
            v.ca = a.cax[c];
            v.da = (lossy_x ? a.dax : a.cax)[c];
            v.cy = a.cay[c];
            v.dy = (lossy_y ? a.day : a.cay)[c];
        }
        v.ey = a.Ey0[c];
        v.hz = a.Hz0[c];
        return v;
    }
    // slab classes: the loads are unconditional too, from clamped addresses (a dummy
    // array stands in for a missing psi slab); tm_mask_plane, applied where the plane
    // enters the pipeline, zeroes what the grid and the slabs do not hold
    const int pc = p < Nx ? p : Nx - 1;
    const int jc = j < 0 ? 0 : (j < P ? j : P - 1);
    const long long c = (long long)pc * P + jc;
    v.ex = a.Ex0[c];
    // (the slab classes read the plain arrays: their bodies hold more registers, and the
    // compact form's address arithmetic cost them more than its bytes save)
    v.ca = a.cax[c];
    v.da = (lossy_x ? a.dax : a.cax)[c];
    v.ey = a.Ey0[c];
    v.cy = a.cay[c];
    v.dy = (lossy_y ? a.day : a.cay)[c];
    v.hz = a.Hz0[c];
    if (PX) {
        // psi_Hzx lives on the H nodes 0..Nx-2: plane Nx-1 (past the last H node, masked
        // below) reads the slab at the last H node, so the load stays inside the slab
        const int sx0 = strip_pos(a.s0[0], pc < Nx - 1 ? pc : Nx - 2), sx3 = strip_pos(a.s0[3], pc);
        const float *p0 = a.s0[0].psi ? a.s0[0].psi : a.Ex0, *p3 = a.s0[3].psi ? a.s0[3].psi : a.Ex0;
        v.ph = p0[(long long)(sx0 > 0 ? sx0 : 0) * P + jc];
        v.qx = p3[(long long)(sx3 > 0 ? sx3 : 0) * P + jc];
    }
    if (PY) {
        const float *p1 = a.s0[1].psi ? a.s0[1].psi : a.Ex0, *p2 = a.s0[2].psi ? a.s0[2].psi : a.Ex0;
        const long long l1 = a.s0[1].psi ? a.s0[1].len : 0, l2 = a.s0[2].psi ? a.s0[2].len : 0;
        v.py = p1[pc * l1 + (sHzy > 0 ? sHzy : 0)];
        v.qy = p2[pc * l2 + (sExy > 0 ? sExy : 0)];
    }
    return v;
}

// Zeroes, for plane p of a slab class, every value the grid and the slabs do not hold
// (the values the conditional loads of the one-step kernels never read).
template <int MODE>
__device__ __forceinline__ void tm_mask_plane(TMPlane &v, const TBTMArgs &a, int p, bool jin, bool jey,
                                              int sHzy, int sExy) {
    constexpr bool PX = MODE & 2, PY = MODE & 1;
    if constexpr (MODE == 0) return;
    const int Nx = a.g.nx;
    const bool pin = p < Nx;
    const bool xin = pin && (!PX || p < Nx - 1);
    const bool mx = xin && jin, my = pin && jey;
    v.ex = mx ? v.ex : 0.f; v.ca = mx ? v.ca : 0.f; v.da = mx ? v.da : 0.f;
    v.ey = my ? v.ey : 0.f; v.cy = my ? v.cy : 0.f; v.dy = my ? v.dy : 0.f;
    v.hz = (xin && jey) ? v.hz : 0.f;
    if (PX) {
        const int sx0 = xin ? strip_pos(a.s0[0], p) : -1;               // H nodes p <= Nx-2
        const int sx3 = (pin && p >= 1 && p <= Nx - 2) ? strip_pos(a.s0[3], p) : -1;
        v.ph = (sx0 >= 0 && jey) ? v.ph : 0.f;
        v.qx = (sx3 >= 0 && jey) ? v.qx : 0.f;
    } else {
        v.ph = v.qx = 0.f;
    }
    if (PY) {
        v.py = (sHzy >= 0 && xin) ? v.py : 0.f;
        v.qy = (sExy >= 0 && xin) ? v.qy : 0.f;
    } else {
        v.py = v.qy = 0.f;
    }
}

// ---------------------------------------------------------------------------- monitors inside a 2D pass
// The steps of a CW window carry volume phasors (DFT), an intensity sum and flux planes
// (fdtd2d_core.py:832-845, flux.py:253-291).  A monitored pass sums them while its levels
// produce the fields.  The accumulators of a volume monitor travel with their plane
// through the levels in registers: loaded with the plane, one addition per level in step
// order, stored with the plane's final fields.  Every accumulator thus receives exactly
// the additions of the one-step monitor kernels, in the same order and arithmetic.  A
// flux plane is one x plane: the levels stage its E and averaged H per step, and
// tb_flux_kernel adds them in step order after the pass.
//
// MON (template): bits 0-1 the number of volume phasors carried (0..2), bit 2 an
// intensity sum; any MON != 0 also stages flux planes.
constexpr int kTBFlux = 4;

struct TBMon {
    int dsrc[2], dacc[2];        // phasor source (LF_SRC_*), accumulator (LF_ACC_*)
    int dx0[2], dx1[2], dn2[2];  // rows [dx0, dx1) of the accumulator, row length
    void *dbuf[2];
    int sform, sn2;              // intensity form (LF_SQ_*), row length (rows 0..Nx-1)
    float *sbuf;
    int nflux;
    int fil[kTBFlux], fn2[kTBFlux];   // flux plane, cells along y
    float *fst[kTBFlux];              // staging: [level][E, H][fn2]
    unsigned lm[kMaxTB];         // per level: bit d phasor d, bit 2 intensity, bit 3+f flux f
    double pr[kMaxTB], pi[kMaxTB];         // exp(-i w t) of the level's step
    float fr[2][kMaxTB], fi[2][kMaxTB];    // complex64 phasor factor of each phasor
};

template <int T, int MON>
struct TBMonRegs {
    static constexpr int ND = MON & 3, NS = (MON >> 2) & 1;
    static constexpr bool C128 = (MON >> 4) & 1;   // every phasor accumulates in complex128
    // bit 5 (TM): the CW-window layout of the stitch runs, fixed at compile time: phasor 0
    // of Ey, phasor 1 of node Ex, the Ex^2 sum, and every level sums all of them
    static constexpr bool FIX = (MON >> 5) & 1;
    static constexpr int NDA = ND > 0 ? ND : 1, NSA = NS > 0 ? NS : 1;
    // slot L (1..T): accumulators of plane p-L, the plane level L-1 advances in
    // iteration p; the next plane's accumulators wait in dn / sn
    double2 d[NDA][T + 1];
    float s[NSA][T + 1];
    double2 dn[NDA];
    float sn[NSA];
    unsigned used;               // monitors summed on any level of this pass
    bool fwave;                  // a flux plane lies among the wavefront's planes [ps, pe)

    __device__ __forceinline__ void init(const TBMon *m, int ps, int pe) {
        used = 0u;
        fwave = false;
        if constexpr (MON != 0) {
#pragma unroll
            for (int l = 0; l < T; ++l) used |= m->lm[l];
            for (int f = 0; f < m->nflux; ++f) fwave = fwave || (m->fil[f] >= ps && m->fil[f] < pe);
        }
#pragma unroll
        for (int L = 0; L <= T; ++L) {
#pragma unroll
            for (int k = 0; k < NDA; ++k) d[k][L] = make_double2(0.0, 0.0);
#pragma unroll
            for (int k = 0; k < NSA; ++k) s[k][L] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < NDA; ++k) dn[k] = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < NSA; ++k) sn[k] = 0.f;
    }
    // accumulators of plane pl (owned planes X0..X1-1 and lanes only; zero elsewhere)
    __device__ __forceinline__ void load(const TBMon *m, int pl, int j, bool own, int X0, int X1) {
        const bool ok = own && pl >= X0 && pl < X1;
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            double2 v = make_double2(0.0, 0.0);
            if (((used >> k) & 1u) && pl >= m->dx0[k] && pl < m->dx1[k] && ok && j < m->dn2[k]) {
                const long long o = (long long)(pl - m->dx0[k]) * m->dn2[k] + j;
                if (C128 || m->dacc[k] == LF_ACC_C128) {
                    v = ((const double2 *)m->dbuf[k])[o];
                } else {
                    const float2 f = ((const float2 *)m->dbuf[k])[o];
                    v = make_double2((double)f.x, (double)f.y);
                }
            }
            dn[k] = v;
        }
        if constexpr (NS > 0) {
            float v = 0.f;
            if (((used >> 2) & 1u) && ok && j < m->sn2) v = m->sbuf[(long long)pl * m->sn2 + j];
            sn[0] = v;
        }
    }
    __device__ __forceinline__ void take() {         // the prefetched plane enters slot 0
#pragma unroll
        for (int k = 0; k < NDA; ++k) d[k][0] = dn[k];
#pragma unroll
        for (int k = 0; k < NSA; ++k) s[k][0] = sn[k];
    }
    __device__ __forceinline__ void store(const TBMon *m, int q, int j) const {   // slot T
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            if (((used >> k) & 1u) && q >= m->dx0[k] && q < m->dx1[k] && j < m->dn2[k]) {
                const long long o = (long long)(q - m->dx0[k]) * m->dn2[k] + j;
                const double2 v = d[k][T];
                if (C128 || m->dacc[k] == LF_ACC_C128) ((double2 *)m->dbuf[k])[o] = v;
                else ((float2 *)m->dbuf[k])[o] = make_float2((float)v.x, (float)v.y);
            }
        }
        if constexpr (NS > 0) {
            if (((used >> 2) & 1u) && j < m->sn2) m->sbuf[(long long)q * m->sn2 + j] = s[0][T];
        }
    }
    __device__ __forceinline__ void shift() {
#pragma unroll
        for (int L = T; L >= 1; --L) {
#pragma unroll
            for (int k = 0; k < NDA; ++k) d[k][L] = d[k][L - 1];
#pragma unroll
            for (int k = 0; k < NSA; ++k) s[k][L] = s[k][L - 1];
        }
    }
};

// one step's phasor term on an accumulator held as double2 (complex64 kinds hold their
// float32 values exactly); the arithmetic of dft_kernel per kind
__device__ __forceinline__ double2 tb_dft_add(double2 v, float F, int acc, double pr, double pi,
                                              float fr, float fi) {
    if (acc == LF_ACC_C128) {
        v.x = v.x + (double)F * pr;
        v.y = v.y + (double)F * pi;
    } else if (acc == LF_ACC_C64P) {
        v.x = (double)(float)(v.x + (double)F * pr);
        v.y = (double)(float)(v.y + (double)F * pi);
    } else {
        v.x = (double)((float)v.x + F * fr);
        v.y = (double)((float)v.y + F * fi);
    }
    return v;
}

// Monitor work of TM level s at plane q, lane j (fields of step n+s+1): exn, eyn are
// Ex(q), Ey(q); exq1 = Ex(q-1) of the same step; hz, hzq1 = Hz(q), Hz(q-1) of step
// n+s+1/2.  Node averages as node_ey / node_ex / node_ex2 (fdtd2d_core.py:458-486).
// `own`: the cell is an owned cell of the wavefront (owned lane AND owned plane): a halo
// plane below the chunk holds invalid values on the higher levels, so only the owner
// stages a flux plane.
template <int T, int MON, bool XIN = false, bool YIN = false>
__device__ __forceinline__ void tm_mon_level(TBMonRegs<T, MON> &r, const TBMon *m, int s, int q,
                                             int j, int Nx, int Ny, bool own, float exn, float eyn,
                                             float exq1, float hz, float hzq1) {
    constexpr bool FIX = TBMonRegs<T, MON>::FIX;
    const unsigned bits = m->lm[s];
    if (!FIX && bits == 0u) return;
    constexpr int ND = TBMonRegs<T, MON>::ND, NS = TBMonRegs<T, MON>::NS;
    const float eyp = lane_prev(eyn);                           // Ey(q, j-1)
    // XIN / YIN: the wavefront's planes / lanes lie strictly inside the grid, so only the
    // interior averages apply (planes outside the march are never summed into an owned cell)
    auto node_ey = [&]() {
        return YIN ? 0.5f * (eyp + eyn) : j == 0 ? eyn : (j == Ny - 1 ? eyp : 0.5f * (eyp + eyn));
    };
    auto node_ex = [&]() {
        return XIN ? 0.5f * (exq1 + exn) : q == 0 ? exn : (q == Nx - 1 ? exq1 : 0.5f * (exq1 + exn));
    };
    auto node_ex2 = [&]() {
        return XIN ? 0.5f * (exq1 * exq1 + exn * exn)
                   : q == 0 ? exn * exn : (q == Nx - 1 ? exq1 * exq1 : 0.5f * (exq1 * exq1 + exn * exn));
    };
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        if (!FIX && !((bits >> k) & 1u)) continue;
        const int src = FIX ? (k == 0 ? LF_SRC_EY : LF_SRC_EX_NODE) : m->dsrc[k];
        const float F = src == LF_SRC_EY ? eyn
                      : src == LF_SRC_EX ? exn
                      : src == LF_SRC_EY_NODE ? node_ey() : node_ex();
        r.d[k][s + 1] = tb_dft_add(r.d[k][s + 1], F, TBMonRegs<T, MON>::C128 ? LF_ACC_C128 : m->dacc[k],
                                   m->pr[s], m->pi[s], m->fr[k][s], m->fi[k][s]);
    }
    if constexpr (NS > 0) {
        if (FIX || ((bits >> 2) & 1u)) {
            float add;
            if (FIX || m->sform == LF_SQ_TM_EX2) {
                add = node_ex2();
            } else {
                const float f = node_ey();
                add = f * f + node_ex2();
            }
            r.s[0][s + 1] = r.s[0][s + 1] + add;
        }
    }
    if ((bits >> 3) && r.fwave) {       // most wavefronts hold no flux plane
        // a loop, not unrolled: four unrolled copies per level hold their plane indices
        // and pointers in scalar registers and push the body over 150 VGPRs (the
        // monitor block is read through a pointer; the loop copies nothing to scratch)
        for (int f = 0; f < m->nflux; ++f) {
            if (((bits >> (3 + f)) & 1u) && q == m->fil[f] && own && j < m->fn2[f]) {
                float *o = m->fst[f] + (long long)(2 * s) * m->fn2[f] + j;
                o[0] = eyn;                                         // Ey(il, j)
                o[m->fn2[f]] = 0.5f * (hzq1 + hz);                  // Hz averaged to il (flux.py:42)
            }
        }
    }
}

// TE level s: ezn = Ez(q) of step n+s+1; hy, hyq1 = Hy(q), Hy(q-1) of step n+s+1/2
template <int T, int MON>
__device__ __forceinline__ void te_mon_level(TBMonRegs<T, MON> &r, const TBMon *m, int s, int q,
                                             int j, bool own, float ezn, float hy, float hyq1) {
    const unsigned bits = m->lm[s];
    if (bits == 0u) return;
    constexpr int ND = TBMonRegs<T, MON>::ND, NS = TBMonRegs<T, MON>::NS;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        if (!((bits >> k) & 1u)) continue;
        r.d[k][s + 1] = tb_dft_add(r.d[k][s + 1], ezn, TBMonRegs<T, MON>::C128 ? LF_ACC_C128 : m->dacc[k],
                                   m->pr[s], m->pi[s], m->fr[k][s], m->fi[k][s]);
    }
    if constexpr (NS > 0) {
        if ((bits >> 2) & 1u) r.s[0][s + 1] = r.s[0][s + 1] + ezn * ezn;
    }
    if ((bits >> 3) && r.fwave) {
        for (int f = 0; f < m->nflux; ++f) {
            if (((bits >> (3 + f)) & 1u) && q == m->fil[f] && own && j < m->fn2[f]) {
                float *o = m->fst[f] + (long long)(2 * s) * m->fn2[f] + j;
                o[0] = ezn;                                         // Ez(il, j)
                o[m->fn2[f]] = 0.5f * (hyq1 + hy);                  // Hy averaged to il
            }
        }
    }
}

// Body of one wavefront (index wv within its instance's launch segments); st[s] are the
// per-step scalars of step n+s.
template <int T, int MODE, int EXC = 3, int MON = 0>
__device__ __forceinline__ void tm_tb_body(const TBTMArgs &a, const Excite &x, const StepArgs *st,
                                           int wv, int fp64, const TBMon *m = nullptr) {
    constexpr bool PX = MODE & 2, PY = MODE & 1;
    constexpr bool XIN = !PX, YIN = !PY;            // planes / lanes strictly interior
    const int lane = threadIdx.x & 63;
    if (wv >= a.seg_end[a.nseg - 1]) return;       // surplus waves of the last workgroup
    int g = 0;
    while (g < a.nseg - 1 && wv >= a.seg_end[g]) ++g;
    const int loc = wv - (g > 0 ? a.seg_end[g - 1] : 0);
    const int wy = a.seg_wy0[g] + loc % a.seg_nwy[g];
    const int Nx = a.g.nx, Ny = a.g.nz, P = a.g.P;
    const int X0 = a.seg_x0[g] + (loc / a.seg_nwy[g]) * a.seg_lx[g];
    if (X0 >= a.seg_x1[g]) return;                 // whole wavefront leaves together
    const int X1 = min(X0 + a.seg_lx[g], a.seg_x1[g]);
    const int j = wy * (64 - 2 * T) - T + lane;
    // the host launches a MODE only on wavefronts whose loaded lanes (YIN) / planes (XIN)
    // lie in 1..Ny-2 / 1..Nx-2 and outside every slab, so those tests are constant
    const bool jin = YIN || (j >= 0 && j < Ny);
    const bool jey = YIN || (jin && j < Ny - 1);   // Ey and Hz exist in this column
    const bool own = jin && lane >= T && lane <= 63 - T;
    // psi slots exist only where the single-step kernels update psi: Hz slabs for
    // j <= Ny-2 (H nodes), Ex slabs for 1 <= j <= Ny-2 (E nodes)
    const int sHzy = (PY && jey) ? strip_pos(a.s0[1], j) : -1;
    const int sExy = (PY && jin && j >= 1 && j <= Ny - 2) ? strip_pos(a.s0[2], j) : -1;
    float yhb = 0.f, yhc = 0.f, yhk = 1.f, yeb = 0.f, yec = 0.f, yek = 1.f;
    if (sHzy >= 0) { yhb = a.yh.b[j]; yhc = a.yh.c[j]; yhk = a.yh.k[j]; }
    if (sExy >= 0) { yeb = a.ye.b[j]; yec = a.ye.c[j]; yek = a.ye.k[j]; }
    const bool lossy_x = a.dax != nullptr, lossy_y = a.day != nullptr;
    const int ps = X0 - T > 0 ? X0 - T : 0;
    const int pe = X1 + T;
    // x profiles of the wavefront's planes, plane ps + lane in lane `lane`, read per level
    // with v_readlane (no dependent scalar loads inside the level chain)
    // (the plan keeps the chunks of the x-slab modes at most 64 - 2T planes long, so the
    // marched planes fit the 64 lanes; a profile read from memory inside the level chain
    // would make the level wait for every plane load in flight)
    if (PX && pe - ps > 64) return;
    float xhb = 0.f, xhc = 0.f, xhk = 1.f, xeb = 0.f, xec = 0.f, xek = 1.f;
    if (PX) {
        const int pl = ps + lane;
        if (pl < Nx - 1) { xhb = a.xh.b[pl]; xhc = a.xh.c[pl]; xhk = a.xh.k[pl]; }
        if (pl < Nx) { xeb = a.xe.b[pl]; xec = a.xe.c[pl]; xek = a.xe.k[pl]; }
    }
    auto plane_val = [&](float reg, int q) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(reg), q - ps));
    };

    // slot L holds the output of level L-1 (slot 0: loaded plane); C = this iteration,
    // P = previous iteration
    float exC[T + 1], eyC[T + 1], hzC[T + 1], exP[T + 1], eyP[T + 1], hzP[T + 1];
    float hxC[T + 1], hyC[T + 1], eyyC[T + 1], exxC[T + 1];    // psi Hzx Hzy Exy Eyx
    float hxP[T + 1], hyP[T + 1], eyyP[T + 1], exxP[T + 1];
    float cA[T + 1], cY[T + 1], dA[T + 1], dY[T + 1];         // coefficients, by plane age
#pragma unroll
    for (int L = 0; L <= T; ++L) {
        exC[L] = eyC[L] = hzC[L] = exP[L] = eyP[L] = hzP[L] = 0.f;
        hxC[L] = hyC[L] = eyyC[L] = exxC[L] = 0.f;
        hxP[L] = hyP[L] = eyyP[L] = exxP[L] = 0.f;
        cA[L] = cY[L] = dA[L] = dY[L] = 0.f;
    }
    // stitch drive of the driven component (Ey), loaded with its plane and carried by
    // plane age like the coefficients, so the levels wait on no load of their own
    // excitation code: bit 0 source, bit 1 drive (blend chosen at run time), bit 2 drive
    // with the float64 blend only
    constexpr bool SRC = EXC & 1, DRV = EXC & 6, D128 = EXC & 4;
    const float2 *dcomp = DRV ? x.drv[LF_EY] : nullptr;
    float dR[T + 1], dI[T + 1], dW[T + 1], dRn = 0.f, dIn = 0.f, dWn = 0.f;
    auto drv_load = [&](int pl) {
        dRn = dIn = dWn = 0.f;
        if (dcomp != nullptr && pl < x.drv_n) {
            dWn = x.wtap[pl];
            if (jey) {
                const float2 D = dcomp[(long long)pl * x.drv_stride[LF_EY] + j];
                dRn = D.x; dIn = D.y;
            }
        }
    };
#pragma unroll
    for (int L = 0; L <= T; ++L) dR[L] = dI[L] = dW[L] = 0.f;
    if (DRV) drv_load(ps);
    TBMonRegs<T, MON> mr;
    if constexpr (MON != 0) {
        mr.init(m, ps, pe);
        mr.load(m, ps, j, own, X0, X1);
    }
    // planes are prefetched two iterations ahead: a wave's march is a serial chain, and
    // one plane of loads in flight left the waves waiting on memory latency
    // (monitored passes carry their accumulators in registers as well: they keep one
    // plane in flight, in one buffer)
    constexpr int PFD = MON == 0 ? 2 : 1;
    // the coefficient table row of the next plane to load is read one load ahead: its
    // scalar wait would otherwise stand in front of the plane's vector loads
    // (only the slab-free class without excitation or monitors reads the compact form:
    // with it the drive, source and monitored bodies measured slower, their cost being
    // instructions and registers rather than bytes)
    constexpr bool CMP = MODE == 0 && EXC == 0 && MON == 0;
    const int q0 = CMP ? coef_row(a.cm, ps, Nx) : 0, q1 = CMP ? coef_row(a.cm, ps + 1, Nx) : 0;
    TMPlane bufA = tm_load_plane<MODE, CMP>(a, ps, q0, j, jin, jey, sHzy, sExy, lossy_x, lossy_y);
    TMPlane bufB{};
    if constexpr (PFD == 2) bufB = tm_load_plane<MODE, CMP>(a, ps + 1, q1, j, jin, jey, sHzy, sExy, lossy_x, lossy_y);
    int qn = CMP ? coef_row(a.cm, ps + PFD, Nx) : 0;
    // iteration p takes plane p from the buffer loaded two iterations earlier and
    // reloads it with plane p + 2; the loop runs two iterations per trip so that the
    // two buffers alternate without register copies (a copy of a load in flight
    // would wait for it at once)
    auto iter = [&](const int p, TMPlane &nxt) {
        // ---- slot 0: plane p of step n (loaded two iterations ahead)
        tm_mask_plane<MODE>(nxt, a, p, jin, jey, sHzy, sExy);
        exC[0] = nxt.ex; eyC[0] = nxt.ey; hzC[0] = nxt.hz;
        hxC[0] = nxt.ph; hyC[0] = nxt.py; eyyC[0] = nxt.qy; exxC[0] = nxt.qx;
        // a lossless grid multiplies by 1 (exact), chosen once per plane, not per level
        cA[0] = nxt.ca; cY[0] = nxt.cy;
        dA[0] = lossy_x ? nxt.da : 1.0f;
        dY[0] = lossy_y ? nxt.dy : 1.0f;
        if (DRV) {
            dR[0] = dRn; dI[0] = dIn; dW[0] = dWn;
            drv_load(p + 1);
        }
        if constexpr (MON != 0) {
            mr.take();
            mr.load(m, p + 1, j, own, X0, X1);
        }
        nxt = tm_load_plane<MODE, CMP>(a, p + PFD, qn, j, jin, jey, sHzy, sExy, lossy_x, lossy_y);
        if constexpr (CMP) qn = coef_row(a.cm, p + PFD + 1, Nx);
        // ---- levels
#pragma unroll
        for (int s = 0; s < T; ++s) {
            const int q = p - 1 - s;
            // no range test on q: a level's outputs at the march's first planes (q < ps +
            // s + 1) or past the grid depend on planes the march does not hold, and no
            // owned output depends on them (the halo cone), so computing them is harmless
            // and the level chain stays free of per-level branches and copies
            // x profiles of plane q: wave-uniform values, read where every lane is active
            const bool qh = PX && strip_pos(a.s0[0], q) >= 0, qe = PX && strip_pos(a.s0[3], q) >= 0;
            float qhb = 0.f, qhc = 0.f, qhk = 1.f, qeb = 0.f, qec = 0.f, qek = 1.f;
            if (qh) { qhb = plane_val(xhb, q); qhc = plane_val(xhc, q); qhk = plane_val(xhk, q); }
            if (qe) { qeb = plane_val(xeb, q); qec = plane_val(xec, q); qek = plane_val(xek, q); }
            {
                const float ex_o = exP[s], ey_o = eyP[s], hz_o = hzP[s];
                const float ex_jp = lane_next0(ex_o);                // Ex(q, j+1)
                float psx = hxP[s], psy = hyP[s];
                float hz = 0.f;
                if ((XIN || q < Nx - 1) && jey) {                   // Hz(q, j)
                    const float dEyx = eyC[s] - ey_o;
                    const float dExy = ex_jp - ex_o;
                    float tx, ty;
                    if (qh) {
                        psx = qhb * psx + qhc * dEyx;
                        tx = div_kappa(dEyx, qhk) + psx;
                    } else {
                        tx = dEyx;
                    }
                    if (sHzy >= 0) {
                        psy = yhb * psy + yhc * dExy;
                        ty = div_kappa(dExy, yhk) + psy;
                    } else {
                        ty = dExy;
                    }
                    hz = hz_o - a.ch * (tx - ty);
                }
                const float hz_jm = lane_prev0(hz);                  // Hz(q, j-1)
                float qy = eyyP[s];
                float exn = 0.f;
                if ((XIN || q < Nx - 1) && jin) {                   // Ex(q, j)
                    float v = ex_o * dA[s + 1];
                    if (YIN || (j >= 1 && j <= Ny - 2)) {
                        const float d = hz - hz_jm;
                        float t;
                        if (sExy >= 0) {
                            qy = yeb * qy + yec * d;
                            t = div_kappa(d, yek) + qy;
                        } else {
                            t = d;
                        }
                        v = v + cA[s + 1] * t;
                    }
                    if (!YIN && ((a.pec_ylo && j == 0) || (a.pec_yhi && j == Ny - 1))) v = 0.0f;
                    exn = v;
                }
                float qx = exxP[s];
                float eyn = 0.f;
                if (jey) {                                          // Ey(q, j)
                    float v = ey_o * dY[s + 1];
                    if (XIN || (q >= 1 && q <= Nx - 2)) {
                        const float d = hz - hzP[s + 1];            // Hz(q) - Hz(q-1)
                        float t;
                        if (qe) {
                            qx = qeb * qx + qec * d;
                            t = div_kappa(d, qek) + qx;
                        } else {
                            t = d;
                        }
                        v = v + (-cY[s + 1]) * t;
                    }
                    if (SRC) v = apply_source(v, LF_EY, q, j, x, st[s], fp64);
                    if (DRV && dcomp != nullptr && q < x.drv_n)
                        v = drive_mix(v, make_float2(dR[s + 1], dI[s + 1]), dW[s + 1], D128 ? LF_DRIVE_C128 : x.drv_math, st[s]);
                    if (!XIN && ((a.pec_xlo && q == 0) || (a.pec_xhi && q == Nx - 1))) v = 0.0f;
                    eyn = v;
                }
                if constexpr (MON != 0)
                    tm_mon_level<T, MON, XIN, YIN>(mr, m, s, q, j, Nx, Ny, own && q >= X0 && q < X1, exn,
                                         eyn, exP[s + 1], hz, hzP[s + 1]);
                exC[s + 1] = exn; eyC[s + 1] = eyn; hzC[s + 1] = hz;
                hxC[s + 1] = psx; hyC[s + 1] = psy; eyyC[s + 1] = qy; exxC[s + 1] = qx;
            }
            exP[s] = exC[s]; eyP[s] = eyC[s]; hzP[s] = hzC[s];
            hxP[s] = hxC[s]; hyP[s] = hyC[s]; eyyP[s] = eyyC[s]; exxP[s] = exxC[s];
        }
        // ---- store plane p-T of step n+T
        {
            const int q = p - T;
            if (q >= X0 && q < X1 && own) {
                const long long c = (long long)q * P + j;
                if (XIN || q < Nx - 1) {
                    a.Ex1[c] = exC[T];
                    if (jey) a.Hz1[c] = hzC[T];
                }
                if (jey) a.Ey1[c] = eyC[T];
                if (PX) {
                    const int sx0 = q < Nx - 1 ? strip_pos(a.s1[0], q) : -1;
                    if (sx0 >= 0 && jey) a.s1[0].psi[(long long)sx0 * P + j] = hxC[T];
                    const int sx3 = (q >= 1 && q <= Nx - 2) ? strip_pos(a.s1[3], q) : -1;
                    if (sx3 >= 0 && jey) a.s1[3].psi[(long long)sx3 * P + j] = exxC[T];
                }
                if (PY) {
                    if (sHzy >= 0 && (XIN || q < Nx - 1))
                        a.s1[1].psi[(long long)q * a.s1[1].len + sHzy] = hyC[T];
                    if (sExy >= 0 && (XIN || q < Nx - 1))
                        a.s1[2].psi[(long long)q * a.s1[2].len + sExy] = eyyC[T];
                }
                if constexpr (MON != 0) mr.store(m, q, j);
            }
            exP[T] = exC[T]; eyP[T] = eyC[T]; hzP[T] = hzC[T];
        }
#pragma unroll
        for (int L = T; L >= 1; --L) {
            cA[L] = cA[L - 1]; cY[L] = cY[L - 1]; dA[L] = dA[L - 1]; dY[L] = dY[L - 1];
            if (DRV) { dR[L] = dR[L - 1]; dI[L] = dI[L - 1]; dW[L] = dW[L - 1]; }
        }
        if constexpr (MON != 0) mr.shift();
    };
    if constexpr (PFD == 2) {
        for (int p = ps; p < pe; p += 2) {
            iter(p, bufA);
            if (p + 1 < pe) iter(p + 1, bufB);
        }
    } else {
        for (int p = ps; p < pe; ++p) iter(p, bufA);
    }
}

// slab wavefronts are few long serial chains that run beside the interior launch: they
// take issue priority on their SIMD so that they finish with it.  The wave index is read
// into a scalar register, so every x/plane test is a scalar branch.
// EXC: excitation code of the body, bit 0 the source, bit 1 the stitch drive, bit 2 the
// drive with its float64 blend fixed at compile time (no float32 blend code).  EXC = 0:
// no source or drive plane lies within the launch's marches, so the body carries no
// excitation loads (whose waits would also drain the plane prefetch).  A context with
// only a drive (the stitch windows after the first) or only a source takes the body with
// only that code: fewer scalar registers to spill and reload inside the levels.
// The slab-free body with excitation is held to 168 VGPRs (3 waves per SIMD instead of
// 2 at its natural 170): its class carries the stitch drive over a quarter of a window
// and is the long pole of those passes; the two values it then keeps in scratch are
// read only on the source planes.
template <int T, int MODE, int EXC>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 256),
                               amdgpu_waves_per_eu(MODE == 0 && EXC ? 3 : 1)))
tm_tb_kernel(TBTMArgs a, Excite x, int fp64) {
    if (MODE != 0) __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    tm_tb_body<T, MODE, EXC>(a, x, a.st, wv, fp64);
}

// segment of wavefront wv (uniform), -1 past the last
__device__ __forceinline__ int tb_segment(const int *seg_end, int nseg, int wv) {
    if (wv >= seg_end[nseg - 1]) return -1;
    int g = 0;
    while (g < nseg - 1 && wv >= seg_end[g]) ++g;
    return g;
}

// A slab class whose rectangles differ in excitation: each wavefront takes the body its
// rectangle needs (the two bodies hold nearly the same number of waves per SIMD).
template <int T, int MODE, int EXC>
__global__ void __launch_bounds__(256) tm_tb_mixed_kernel(TBTMArgs a, Excite x, int fp64) {
    __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = tb_segment(a.seg_end, a.nseg, wv);
    if (g < 0) return;
    if ((a.seg_exc >> g) & 1u) tm_tb_body<T, MODE, EXC>(a, x, a.st, wv, fp64);
    else tm_tb_body<T, MODE, 0>(a, x, a.st, wv, fp64);
}

// The y-slab rows (MODE 1) and the x-slab chunks (MODE 3) of a pass in one launch: both
// bodies hold two waves per SIMD, so one launch loses no occupancy and saves the host a
// launch and a stream's events and waits per pass.  Each wavefront takes the body of its
// rectangle (seg_m3) with or without excitation code (seg_exc).
template <int T, int EXC>
__global__ void __launch_bounds__(256) tm_tb_slab_kernel(TBTMArgs a, Excite x, int fp64) {
    __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = tb_segment(a.seg_end, a.nseg, wv);
    if (g < 0) return;
    const bool m3 = (a.seg_m3 >> g) & 1u, e = EXC != 0 && ((a.seg_exc >> g) & 1u);
    if (m3) {
        if (e) tm_tb_body<T, 3, EXC>(a, x, a.st, wv, fp64);
        else tm_tb_body<T, 3, 0>(a, x, a.st, wv, fp64);
    } else {
        if (e) tm_tb_body<T, 1, EXC>(a, x, a.st, wv, fp64);
        else tm_tb_body<T, 1, 0>(a, x, a.st, wv, fp64);
    }
}

// ---------------------------------------------------------------------------- temporal blocking (2D TE)
// Same pipeline as the TM pass with the roles of E and H exchanged: level s advances
// plane q = p-1-s from step n+s to n+s+1: Hx^{n+s+1/2}(q) needs Ez(q, j+1) (lane above),
// Hy^{n+s+1/2}(q) needs Ez(q+1) (this iteration's output of level s-1), and
// Ez^{n+s+1}(q) needs Hy(q-1) (own previous output) and Hx(q, j-1) (lane below).  Per
// cell and step the arithmetic is that of te_h_kernel + te_e_kernel
// (boundaries.py:189-215).  psi slabs: Hxy (y), Hyx (x), Ezx (x), Ezy (y).
struct TBTEArgs {
    Grid g;
    const float *Ez0, *Hx0, *Hy0;
    float *Ez1, *Hx1, *Hy1;
    const float *ceE, *ceH;               // plain arrays
    Coef2 kE, kH;                         // as the slab-free bodies read them
    CoefMap cm;
    float ch;
    Prof xe, xh, ye, yh;
    Strip s0[4], s1[4];       // Hxy Hyx Ezx Ezy: read / write buffers
    int pec_xlo, pec_xhi, pec_ylo, pec_yhi;
    int nseg;                 // wavefront rectangles, as in TBTMArgs
    int seg_x0[8], seg_x1[8], seg_wy0[8], seg_nwy[8], seg_end[8], seg_lx[8];
    unsigned seg_exc;
    unsigned seg_m3;
    StepArgs st[kMaxTB];
};

struct TEPlane {
    float ez, hx, hy, ce, cH, pxy, pyx, pzx, pzy;
};

template <int MODE, bool CMP = false>
__device__ __forceinline__ TEPlane te_load_plane(const TBTEArgs &a, int p, int q, int j, bool jin, bool jhx,
                                                 int sHxy, int sEzy) {
    constexpr bool PX = MODE & 2, PY = MODE & 1;
    TEPlane v{};
    const int Nx = a.g.nx, P = a.g.P;
    if constexpr (MODE == 0) {                         // unconditional, as for TM
        const long long c = (long long)(p < Nx ? p : Nx - 1) * P + j;
        v.ez = a.Ez0[c];
        if constexpr (CMP) {
            const CoefAt w = coef_pos(a.cm, p < Nx ? p : Nx - 1, q, j, P);
            v.ce = coef_at(a.kE, w);
            v.cH = coef_at(a.kH, w);
        } else {
            v.ce = a.ceE[c];
            v.cH = a.ceH[c];
        }
        v.hy = a.Hy0[c];
        v.hx = a.Hx0[c];
        return v;
    }
    // slab classes: unconditional loads from clamped addresses, masked by te_mask_plane
    const int pc = p < Nx ? p : Nx - 1;
    const int jc = j < 0 ? 0 : (j < P ? j : P - 1);
    const long long c = (long long)pc * P + jc;
    v.ez = a.Ez0[c];
    v.ce = a.ceE[c];                                  // plain arrays, as for TM
    v.cH = a.ceH[c];
    v.hy = a.Hy0[c];
    v.hx = a.Hx0[c];
    if (PX) {
        // psi_Hyx lives on the H nodes 0..Nx-2, as psi_Hzx in the TM load
        const int sx1 = strip_pos(a.s0[1], pc < Nx - 1 ? pc : Nx - 2), sx2 = strip_pos(a.s0[2], pc);
        const float *p1 = a.s0[1].psi ? a.s0[1].psi : a.Ez0, *p2 = a.s0[2].psi ? a.s0[2].psi : a.Ez0;
        v.pyx = p1[(long long)(sx1 > 0 ? sx1 : 0) * P + jc];
        v.pzx = p2[(long long)(sx2 > 0 ? sx2 : 0) * P + jc];
    }
    if (PY) {
        const float *p0 = a.s0[0].psi ? a.s0[0].psi : a.Ez0, *p3 = a.s0[3].psi ? a.s0[3].psi : a.Ez0;
        const long long l0 = a.s0[0].psi ? a.s0[0].len : 0, l3 = a.s0[3].psi ? a.s0[3].len : 0;
        v.pxy = p0[pc * l0 + (sHxy > 0 ? sHxy : 0)];
        v.pzy = p3[pc * l3 + (sEzy > 0 ? sEzy : 0)];
    }
    return v;
}

template <int MODE>
__device__ __forceinline__ void te_mask_plane(TEPlane &v, const TBTEArgs &a, int p, bool jin, bool jhx,
                                              int sHxy, int sEzy) {
    constexpr bool PX = MODE & 2, PY = MODE & 1;
    if constexpr (MODE == 0) return;
    const int Nx = a.g.nx;
    const bool pin = p < Nx;
    const bool xin = pin && (!PX || p < Nx - 1);                   // Hy exists on this plane
    const bool mj = pin && jin;
    v.ez = mj ? v.ez : 0.f; v.ce = mj ? v.ce : 0.f; v.cH = mj ? v.cH : 0.f;
    v.hy = (xin && jin) ? v.hy : 0.f;
    v.hx = (pin && jhx) ? v.hx : 0.f;
    if (PX) {
        const int sx1 = xin ? strip_pos(a.s0[1], p) : -1;               // H nodes p <= Nx-2
        const int sx2 = (pin && p >= 1 && p <= Nx - 2) ? strip_pos(a.s0[2], p) : -1;
        v.pyx = (sx1 >= 0 && jin) ? v.pyx : 0.f;
        v.pzx = (sx2 >= 0 && jin) ? v.pzx : 0.f;
    } else {
        v.pyx = v.pzx = 0.f;
    }
    if (PY) {
        v.pxy = (pin && sHxy >= 0) ? v.pxy : 0.f;
        v.pzy = (pin && sEzy >= 0 && (!PX || (p >= 1 && p <= Nx - 2))) ? v.pzy : 0.f;
    } else {
        v.pxy = v.pzy = 0.f;
    }
}

template <int T, int MODE, int EXC = 3, int MON = 0>
__device__ __forceinline__ void te_tb_body(const TBTEArgs &a, const Excite &x, const StepArgs *st,
                                           int wv, int fp64, const TBMon *m = nullptr) {
    constexpr bool PX = MODE & 2, PY = MODE & 1;
    constexpr bool XIN = !PX, YIN = !PY;
    const int lane = threadIdx.x & 63;
    if (wv >= a.seg_end[a.nseg - 1]) return;
    int g = 0;
    while (g < a.nseg - 1 && wv >= a.seg_end[g]) ++g;
    const int loc = wv - (g > 0 ? a.seg_end[g - 1] : 0);
    const int wy = a.seg_wy0[g] + loc % a.seg_nwy[g];
    const int Nx = a.g.nx, Ny = a.g.nz, P = a.g.P;
    const int X0 = a.seg_x0[g] + (loc / a.seg_nwy[g]) * a.seg_lx[g];
    if (X0 >= a.seg_x1[g]) return;
    const int X1 = min(X0 + a.seg_lx[g], a.seg_x1[g]);
    const int j = wy * (64 - 2 * T) - T + lane;
    const bool jin = YIN || (j >= 0 && j < Ny);
    const bool jhx = YIN || (jin && j < Ny - 1);   // Hx exists in this column
    const bool jcurl = YIN || (j >= 1 && j <= Ny - 2);
    const bool own = jin && lane >= T && lane <= 63 - T;
    const int sHxy = (PY && jhx) ? strip_pos(a.s0[0], j) : -1;
    const int sEzy = (PY && jin && j >= 1 && j <= Ny - 2) ? strip_pos(a.s0[3], j) : -1;
    float yhb = 0.f, yhc = 0.f, yhk = 1.f, yeb = 0.f, yec = 0.f, yek = 1.f;
    if (sHxy >= 0) { yhb = a.yh.b[j]; yhc = a.yh.c[j]; yhk = a.yh.k[j]; }
    if (sEzy >= 0) { yeb = a.ye.b[j]; yec = a.ye.c[j]; yek = a.ye.k[j]; }
    const int ps = X0 - T > 0 ? X0 - T : 0;
    const int pe = X1 + T;
    // (the plan keeps the chunks of the x-slab modes at most 64 - 2T planes long, so the
    // marched planes fit the 64 lanes; a profile read from memory inside the level chain
    // would make the level wait for every plane load in flight)
    if (PX && pe - ps > 64) return;
    float xhb = 0.f, xhc = 0.f, xhk = 1.f, xeb = 0.f, xec = 0.f, xek = 1.f;
    if (PX) {
        const int pl = ps + lane;
        if (pl < Nx - 1) { xhb = a.xh.b[pl]; xhc = a.xh.c[pl]; xhk = a.xh.k[pl]; }
        if (pl < Nx) { xeb = a.xe.b[pl]; xec = a.xe.c[pl]; xek = a.xe.k[pl]; }
    }
    auto plane_val = [&](float reg, int q) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(reg), q - ps));
    };

    float ezC[T + 1], hxC[T + 1], hyC[T + 1], ezP[T + 1], hxP[T + 1], hyP[T + 1];
    float axyC[T + 1], ayxC[T + 1], azxC[T + 1], azyC[T + 1];   // psi Hxy Hyx Ezx Ezy
    float axyP[T + 1], ayxP[T + 1], azxP[T + 1], azyP[T + 1];
    float cE[T + 1], cH[T + 1];
#pragma unroll
    for (int L = 0; L <= T; ++L) {
        ezC[L] = hxC[L] = hyC[L] = ezP[L] = hxP[L] = hyP[L] = 0.f;
        axyC[L] = ayxC[L] = azxC[L] = azyC[L] = 0.f;
        axyP[L] = ayxP[L] = azxP[L] = azyP[L] = 0.f;
        cE[L] = cH[L] = 0.f;
    }
    // stitch drive of Ez, loaded with its plane and carried by plane age
    // excitation code: bit 0 source, bit 1 drive (blend chosen at run time), bit 2 drive
    // with the float64 blend only
    constexpr bool SRC = EXC & 1, DRV = EXC & 6, D128 = EXC & 4;
    const float2 *dcomp = DRV ? x.drv[LF_EZ] : nullptr;
    float dR[T + 1], dI[T + 1], dW[T + 1], dRn = 0.f, dIn = 0.f, dWn = 0.f;
    auto drv_load = [&](int pl) {
        dRn = dIn = dWn = 0.f;
        if (dcomp != nullptr && pl < x.drv_n) {
            dWn = x.wtap[pl];
            if (jin) {
                const float2 D = dcomp[(long long)pl * x.drv_stride[LF_EZ] + j];
                dRn = D.x; dIn = D.y;
            }
        }
    };
#pragma unroll
    for (int L = 0; L <= T; ++L) dR[L] = dI[L] = dW[L] = 0.f;
    if (DRV) drv_load(ps);
    TBMonRegs<T, MON> mr;
    if constexpr (MON != 0) {
        mr.init(m, ps, pe);
        mr.load(m, ps, j, own, X0, X1);
    }
    constexpr int PFD = MON == 0 ? 2 : 1;                                   // as for TM
    constexpr bool CMP = MODE == 0 && EXC == 0 && MON == 0;                 // as for TM
    const int q0 = CMP ? coef_row(a.cm, ps, Nx) : 0, q1 = CMP ? coef_row(a.cm, ps + 1, Nx) : 0;
    TEPlane bufA = te_load_plane<MODE, CMP>(a, ps, q0, j, jin, jhx, sHxy, sEzy);
    TEPlane bufB{};
    if constexpr (PFD == 2) bufB = te_load_plane<MODE, CMP>(a, ps + 1, q1, j, jin, jhx, sHxy, sEzy);
    int qn = CMP ? coef_row(a.cm, ps + PFD, Nx) : 0;                        // read one load ahead
    // iteration p takes plane p from the buffer loaded two iterations earlier and
    // reloads it with plane p + 2; the loop runs two iterations per trip so that the
    // two buffers alternate without register copies (a copy of a load in flight
    // would wait for it at once)
    auto iter = [&](const int p, TEPlane &nxt) {
        te_mask_plane<MODE>(nxt, a, p, jin, jhx, sHxy, sEzy);
        ezC[0] = nxt.ez; hxC[0] = nxt.hx; hyC[0] = nxt.hy;
        axyC[0] = nxt.pxy; ayxC[0] = nxt.pyx; azxC[0] = nxt.pzx; azyC[0] = nxt.pzy;
        cE[0] = nxt.ce; cH[0] = nxt.cH;
        if (DRV) {
            dR[0] = dRn; dI[0] = dIn; dW[0] = dWn;
            drv_load(p + 1);
        }
        if constexpr (MON != 0) {
            mr.take();
            mr.load(m, p + 1, j, own, X0, X1);
        }
        nxt = te_load_plane<MODE, CMP>(a, p + PFD, qn, j, jin, jhx, sHxy, sEzy);
        if constexpr (CMP) qn = coef_row(a.cm, p + PFD + 1, Nx);
#pragma unroll
        for (int s = 0; s < T; ++s) {
            const int q = p - 1 - s;
            // no range test on q: a level's outputs at the march's first planes (q < ps +
            // s + 1) or past the grid depend on planes the march does not hold, and no
            // owned output depends on them (the halo cone), so computing them is harmless
            // and the level chain stays free of per-level branches and copies
            // x profiles of plane q: wave-uniform values, read where every lane is active
            const bool qh = PX && strip_pos(a.s0[1], q) >= 0, qe = PX && strip_pos(a.s0[2], q) >= 0;
            float qhb = 0.f, qhc = 0.f, qhk = 1.f, qeb = 0.f, qec = 0.f, qek = 1.f;
            if (qh) { qhb = plane_val(xhb, q); qhc = plane_val(xhc, q); qhk = plane_val(xhk, q); }
            if (qe) { qeb = plane_val(xeb, q); qec = plane_val(xec, q); qek = plane_val(xek, q); }
            {
                const float ez_o = ezP[s];
                const float ez_jp = lane_next0(ez_o);                // Ez(q, j+1)
                float qxy = axyP[s], qyx = ayxP[s];
                float hx = 0.f, hy = 0.f;
                if (jhx) {                                          // Hx(q, j)
                    const float d = ez_jp - ez_o;
                    float t;
                    if (sHxy >= 0) {
                        qxy = yhb * qxy + yhc * d;
                        t = div_kappa(d, yhk) + qxy;
                    } else {
                        t = d;
                    }
                    hx = hxP[s] - a.ch * t;
                }
                if ((XIN || q < Nx - 1) && jin) {                   // Hy(q, j)
                    const float d = ezC[s] - ez_o;
                    float t;
                    if (qh) {
                        qyx = qhb * qyx + qhc * d;
                        t = div_kappa(d, qhk) + qyx;
                    } else {
                        t = d;
                    }
                    hy = hyP[s] + a.ch * t;
                }
                const float hx_jm = lane_prev0(hx);                  // Hx(q, j-1)
                float qzx = azxP[s], qzy = azyP[s];
                float ezn = 0.f;
                if (jin) {                                          // Ez(q, j)
                    float v = ez_o;
                    if ((XIN || (q >= 1 && q <= Nx - 2)) && jcurl) {
                        const float dHy = hy - hyP[s + 1];          // Hy(q) - Hy(q-1)
                        const float dHx = hx - hx_jm;
                        float t;
                        if (qe) {
                            qzx = qeb * qzx + qec * dHy;
                            t = div_kappa(dHy, qek) + qzx;
                        } else {
                            t = dHy;
                        }
                        if (sEzy >= 0) {
                            qzy = yeb * qzy + yec * dHx;
                            t = (t - div_kappa(dHx, yek)) - qzy;
                        } else {
                            t = t - dHx;
                        }
                        v = cE[s + 1] * v + cH[s + 1] * t;
                    }
                    if (SRC) v = apply_source(v, LF_EZ, q, j, x, st[s], fp64);
                    if (DRV && dcomp != nullptr && q < x.drv_n)
                        v = drive_mix(v, make_float2(dR[s + 1], dI[s + 1]), dW[s + 1], D128 ? LF_DRIVE_C128 : x.drv_math, st[s]);
                    if (!XIN && ((a.pec_xlo && q == 0) || (a.pec_xhi && q == Nx - 1))) v = 0.0f;
                    if (!YIN && ((a.pec_ylo && j == 0) || (a.pec_yhi && j == Ny - 1))) v = 0.0f;
                    ezn = v;
                }
                if constexpr (MON != 0)
                    te_mon_level<T, MON>(mr, m, s, q, j, own && q >= X0 && q < X1, ezn, hy,
                                         hyP[s + 1]);
                ezC[s + 1] = ezn; hxC[s + 1] = hx; hyC[s + 1] = hy;
                axyC[s + 1] = qxy; ayxC[s + 1] = qyx; azxC[s + 1] = qzx; azyC[s + 1] = qzy;
            }
            ezP[s] = ezC[s]; hxP[s] = hxC[s]; hyP[s] = hyC[s];
            axyP[s] = axyC[s]; ayxP[s] = ayxC[s]; azxP[s] = azxC[s]; azyP[s] = azyC[s];
        }
        {
            const int q = p - T;
            if (q >= X0 && q < X1 && own) {
                const long long c = (long long)q * P + j;
                a.Ez1[c] = ezC[T];
                if (jhx) a.Hx1[c] = hxC[T];
                if (XIN || q < Nx - 1) a.Hy1[c] = hyC[T];
                if (PX) {
                    const int sx1 = q < Nx - 1 ? strip_pos(a.s1[1], q) : -1;
                    if (sx1 >= 0) a.s1[1].psi[(long long)sx1 * P + j] = ayxC[T];
                    const int sx2 = (q >= 1 && q <= Nx - 2) ? strip_pos(a.s1[2], q) : -1;
                    if (sx2 >= 0) a.s1[2].psi[(long long)sx2 * P + j] = azxC[T];
                }
                if (PY) {
                    if (sHxy >= 0) a.s1[0].psi[(long long)q * a.s1[0].len + sHxy] = axyC[T];
                    if (sEzy >= 0 && (XIN || (q >= 1 && q <= Nx - 2)))
                        a.s1[3].psi[(long long)q * a.s1[3].len + sEzy] = azyC[T];
                }
                if constexpr (MON != 0) mr.store(m, q, j);
            }
            ezP[T] = ezC[T]; hxP[T] = hxC[T]; hyP[T] = hyC[T];
        }
#pragma unroll
        for (int L = T; L >= 1; --L) {
            cE[L] = cE[L - 1]; cH[L] = cH[L - 1];
            if (DRV) { dR[L] = dR[L - 1]; dI[L] = dI[L - 1]; dW[L] = dW[L - 1]; }
        }
        if constexpr (MON != 0) mr.shift();
    };
    if constexpr (PFD == 2) {
        for (int p = ps; p < pe; p += 2) {
            iter(p, bufA);
            if (p + 1 < pe) iter(p + 1, bufB);
        }
    } else {
        for (int p = ps; p < pe; ++p) iter(p, bufA);
    }
}

template <int T, int MODE, int EXC>
__global__ void __launch_bounds__(256) te_tb_kernel(TBTEArgs a, Excite x, int fp64) {
    if (MODE != 0) __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    te_tb_body<T, MODE, EXC>(a, x, a.st, wv, fp64);
}

template <int T, int MODE, int EXC>
__global__ void __launch_bounds__(256) te_tb_mixed_kernel(TBTEArgs a, Excite x, int fp64) {
    __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = tb_segment(a.seg_end, a.nseg, wv);
    if (g < 0) return;
    if ((a.seg_exc >> g) & 1u) te_tb_body<T, MODE, EXC>(a, x, a.st, wv, fp64);
    else te_tb_body<T, MODE, 0>(a, x, a.st, wv, fp64);
}

template <int T, int EXC>
__global__ void __launch_bounds__(256) te_tb_slab_kernel(TBTEArgs a, Excite x, int fp64) {
    __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = tb_segment(a.seg_end, a.nseg, wv);
    if (g < 0) return;
    const bool m3 = (a.seg_m3 >> g) & 1u, e = EXC != 0 && ((a.seg_exc >> g) & 1u);
    if (m3) {
        if (e) te_tb_body<T, 3, EXC>(a, x, a.st, wv, fp64);
        else te_tb_body<T, 3, 0>(a, x, a.st, wv, fp64);
    } else {
        if (e) te_tb_body<T, 1, EXC>(a, x, a.st, wv, fp64);
        else te_tb_body<T, 1, 0>(a, x, a.st, wv, fp64);
    }
}

// Monitored passes: the plain pass plus the monitor work of TBMon (instantiated in
// tb2d_mon.hip for depths 2..4 and MON in {9, 14}).
template <int T, int MODE, bool EXC, int MON>
__global__ void __launch_bounds__(256) tm_tb_mon_kernel(TBTMArgs a, Excite x, TBMon m, int fp64) {
    if (MODE != 0) __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    tm_tb_body<T, MODE, EXC ? 3 : 0, MON>(a, x, a.st, wv, fp64, &m);
}

template <int T, int MODE, bool EXC, int MON>
__global__ void __launch_bounds__(256) te_tb_mon_kernel(TBTEArgs a, Excite x, TBMon m, int fp64) {
    if (MODE != 0) __builtin_amdgcn_s_setprio(2);
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    te_tb_body<T, MODE, EXC ? 3 : 0, MON>(a, x, a.st, wv, fp64, &m);
}

// launchers of the plain passes (tb2d_pass.hip): depth T in {2, 4, 6, 8}, mode 0 / 1 /
// 3; exc 0 without excitation code, 1 with it, 2 per rectangle (modes 1 and 3)
void launch_tm_tb_slab(int T, int exc, const TBTMArgs &q, const Excite &x, int fp64, hipStream_t st,
                       unsigned nb);
void launch_te_tb_slab(int T, int exc, const TBTEArgs &q, const Excite &x, int fp64, hipStream_t st,
                       unsigned nb);
void launch_tm_tb(int T, int mode, int exc, const TBTMArgs &q, const Excite &x, int fp64,
                  hipStream_t st, unsigned nb);
void launch_te_tb(int T, int mode, int exc, const TBTEArgs &q, const Excite &x, int fp64,
                  hipStream_t st, unsigned nb);

// launchers (tb2d_mon.hip): depth T in {2, 3, 4}, mode 0 / 1 / 3, mon 9 (one phasor) or
// 14 (two phasors and an intensity sum); flux staging in both
void launch_tm_tb_mon(int T, int mode, bool exc, int mon, const TBTMArgs &q, const Excite &x,
                      const TBMon &m, int fp64, hipStream_t st, unsigned nb);
void launch_te_tb_mon(int T, int mode, bool exc, int mon, const TBTEArgs &q, const Excite &x,
                      const TBMon &m, int fp64, hipStream_t st, unsigned nb);

}  // namespace lfd
