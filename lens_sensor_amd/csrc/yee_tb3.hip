// yee_tb3.hip - kernels of the temporally blocked 3D pass.  Built once per (depth,
// loss) variant with -DLF_TB3_T=<1|2|3> -DLF_TB3_LOSS=<0|1>, each object defining its
// launcher; lensfdtd.hip calls them through launch_yee_tb3.
#include <hip/hip_runtime.h>

#include "yee_tb3.h"

#ifndef LF_TB3_T
#error "LF_TB3_T must be defined (1, 2 or 3)"
#endif
#ifndef LF_TB3_LOSS
#error "LF_TB3_LOSS must be defined (0 or 1)"
#endif

namespace lfd {
namespace {

template <int R>
struct TB3Plane {                    // one x plane of a wavefront, from the step-n buffers
    float ex[R], ey[R], ez[R], hx[R], hy[R], hz[R];
    float cx[R], cy[R], cz[R], dx[R], dy[R], dz[R];
    float p[12][R];
};

// LDS exchange slots: E (Ex, Ez of a wavefront's first row) by iteration parity and
// slot, H (Hx, Hz of its last row) by level
template <int T>
__device__ __forceinline__ int tb3_eidx(int par, int s, int w, int comp, int lane) {
    return (((par * T + s) * kTB3W + w) * 2 + comp) * 64 + lane;
}
template <int T>
__device__ __forceinline__ int tb3_hidx(int s, int w, int comp, int lane) {
    return 2 * T * kTB3W * 128 + ((s * kTB3W + w) * 2 + comp) * 64 + lane;
}

template <int T, int R, int MODE, bool LOSS, bool EXC, bool BOX>
__device__ __forceinline__ void yee_tb_body(const TB3Args &a, const Excite &x, const TB3Box &bx,
                                            int u, float *lds, int fp64, double *bacc) {
    constexpr bool PZ = MODE & 1, PY = MODE & 2, PX = MODE & 4;
    constexpr int WY = kTB3W * R - 2 * T, WZ = 64 - 2 * T;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tz = u % bx.ntz, tr = u / bx.ntz;
    const int ty = tr % bx.nty, tx = tr / bx.nty;
    const int X0 = bx.x0 + tx * bx.lx, X1 = min(X0 + bx.lx, bx.x1);
    const int J0 = bx.y0 + ty * WY, J1 = min(J0 + WY, bx.y1);
    const int K0 = bx.z0 + tz * WZ, K1 = min(K0 + WZ, bx.z1);
    const Grid &G = a.g;
    const int NX = G.nx, NY = G.ny, NZ = G.nz;
    const unsigned P = G.P;
    const float ch = a.ch;
    const int k = K0 - T + lane;
    const bool kuse = k >= 0 && k < min(K1 + T, NZ);
    const bool kz1 = kuse && (!PZ || k < NZ - 1);             // Ez, Hx, Hy exist at this k
    const bool kint = kuse && (!PZ || (k >= 1 && k <= NZ - 2));
    const bool kown = kuse && k >= K0 && k < K1;
    const int kc = kuse ? k : 0;
    const bool lx_ = LOSS && a.dax != nullptr, ly_ = LOSS && a.day != nullptr,
               lz_ = LOSS && a.daz != nullptr;
    // rows (uniform over the wavefront)
    int jr[R];
    bool juse[R], jh[R], ycurl[R], jown[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = J0 - T + w * R + r;
        juse[r] = j >= 0 && j < min(J1 + T, NY);
        jh[r] = juse[r] && (!PY || j < NY - 1);              // Ey, Hx, Hz exist in this row
        ycurl[r] = juse[r] && (!PY || (j >= 1 && j <= NY - 2));
        jown[r] = juse[r] && j >= J0 && j < J1;
        jr[r] = juse[r] ? j : 0;
    }
    auto off = [&](int i, int j) { return ((unsigned)i * NY + (unsigned)j) * P + (unsigned)kc; };
    // z slabs of this lane and y slabs of every row, fixed over the pass
    int zHx = -1, zHy = -1, zEx = -1, zEy = -1;
    PV zh{0.f, 0.f, 0.f}, ze{0.f, 0.f, 0.f};
    if (PZ && kuse) {
        if (k < NZ - 1) { zHx = strip_pos(a.s0[0], k); zHy = strip_pos(a.s0[3], k); }
        if (k >= 1 && k <= NZ - 2) { zEx = strip_pos(a.s0[7], k); zEy = strip_pos(a.s0[8], k); }
        if (zHx >= 0 || zHy >= 0) zh = prof_at(a.ph[2], k);
        if (zEx >= 0 || zEy >= 0) ze = prof_at(a.pe[2], k);
    }
    int yHx[R], yHz[R], yEx[R], yEz[R];
    PV yh[R], ye[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        yHx[r] = yHz[r] = yEx[r] = yEz[r] = -1;
        yh[r] = PV{0.f, 0.f, 0.f}; ye[r] = PV{0.f, 0.f, 0.f};
        if (PY) {
            const int j = jr[r];
            if (jh[r]) { yHx[r] = strip_pos(a.s0[1], j); yHz[r] = strip_pos(a.s0[4], j); }
            if (ycurl[r]) { yEx[r] = strip_pos(a.s0[6], j); yEz[r] = strip_pos(a.s0[11], j); }
            if (yHx[r] >= 0 || yHz[r] >= 0) yh[r] = prof_at(a.ph[1], j);
            if (yEx[r] >= 0 || yEz[r] >= 0) ye[r] = prof_at(a.pe[1], j);
        }
    }
    const int ps = X0 - T > 0 ? X0 - T : 0;
    const int pe = X1 + T;
    // x profiles of the chunk's planes (plane ps + lane in lane `lane`), read per level
    // with v_readlane when the chunk fits one wavefront
    const bool xpre = PX && pe - ps <= 64;
    PV xhl{0.f, 0.f, 0.f}, xel{0.f, 0.f, 0.f};
    if (xpre) {
        const int pl = ps + lane;
        if (pl < NX - 1) xhl = prof_at(a.ph[0], pl);
        if (pl < NX) xel = prof_at(a.pe[0], pl);
    }
    auto rl = [](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
    auto xprof = [&](const PV &reg, const Prof &pr, int q) {
        if (xpre) return PV{rl(reg.b, q - ps), rl(reg.c, q - ps), rl(reg.m, q - ps)};
        return prof_at(pr, q);
    };
    // x slab positions of plane i (-1: none)
    auto xslabs = [&](int i, int &xHy, int &xHz, int &xEy, int &xEz) {
        xHy = xHz = xEy = xEz = -1;
        if (PX) {
            if (i < NX - 1) { xHy = strip_pos(a.s0[2], i); xHz = strip_pos(a.s0[5], i); }
            if (i >= 1 && i <= NX - 2) { xEy = strip_pos(a.s0[9], i); xEz = strip_pos(a.s0[10], i); }
        }
    };
    // psi existence: the cells where the one-step kernels update each slab
    auto psi_on = [&](int q, int r, int i, bool xin1, bool xcurl, int xHy, int xHz, int xEy,
                      int xEz) -> int {
        switch (q) {
            case 0: return (PZ && kz1 && jh[r]) ? zHx : -1;
            case 1: return (PY && kz1 && jh[r]) ? yHx[r] : -1;
            case 2: return (PX && kz1 && juse[r]) ? xHy : -1;
            case 3: return (PZ && kz1 && juse[r] && xin1) ? zHy : -1;
            case 4: return (PY && kuse && jh[r] && xin1) ? yHz[r] : -1;
            case 5: return (PX && kuse && jh[r]) ? xHz : -1;
            case 6: return (PY && kint && xin1) ? yEx[r] : -1;
            case 7: return (PZ && juse[r] && xin1 && ycurl[r]) ? zEx : -1;
            case 8: return (PZ && jh[r] && xcurl) ? zEy : -1;
            case 9: return (PX && kint && jh[r]) ? xEy : -1;
            case 10: return (PX && kz1 && ycurl[r]) ? xEz : -1;
            default: return (PY && kz1 && xcurl) ? yEz[r] : -1;
        }
    };
    constexpr int psi_axis[12] = {2, 1, 0, 2, 1, 0, 1, 2, 2, 0, 0, 1};
    auto load = [&](int i, TB3Plane<R> &L) {
        const bool xok = !PX || i < NX;                    // planes past the grid read as 0
        const bool xin1 = !PX || i < NX - 1;
        const bool xcurl = !PX || (i >= 1 && i <= NX - 2);
        int xHy, xHz, xEy, xEz;
        xslabs(i, xHy, xHz, xEy, xEz);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const unsigned c = off(xok ? i : 0, jr[r]);
            const bool ro = juse[r] && xok;
            L.ex[r] = (kuse && ro && xin1) ? a.Ex0[c] : 0.f;
            const bool rh = jh[r] && xok;
            L.ey[r] = (kuse && rh) ? a.Ey0[c] : 0.f;
            L.ez[r] = (kz1 && ro) ? a.Ez0[c] : 0.f;
            L.hx[r] = (kz1 && rh) ? a.Hx0[c] : 0.f;
            L.hy[r] = (kz1 && ro && xin1) ? a.Hy0[c] : 0.f;
            L.hz[r] = (kuse && rh && xin1) ? a.Hz0[c] : 0.f;
            L.cx[r] = (kuse && ro && xin1) ? a.cex[c] : 0.f;
            L.cy[r] = (kuse && rh) ? a.cey[c] : 0.f;
            L.cz[r] = (kz1 && ro) ? a.cez[c] : 0.f;
            L.dx[r] = (lx_ && kuse && ro && xin1) ? a.dax[c] : 0.f;
            L.dy[r] = (ly_ && kuse && rh) ? a.day[c] : 0.f;
            L.dz[r] = (lz_ && kz1 && ro) ? a.daz[c] : 0.f;
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                L.p[q][r] = 0.f;
                if (!((q == 0 || q == 3 || q == 7 || q == 8) ? PZ
                      : (q == 1 || q == 4 || q == 6 || q == 11) ? PY : PX))
                    continue;
                if (!xok) continue;
                const int pos = psi_on(q, r, i, xin1, xcurl, xHy, xHz, xEy, xEz);
                if (pos >= 0) L.p[q][r] = a.s0[q].psi[psi_off32(a.s0[q], psi_axis[q], pos, i, jr[r], kc, G)];
            }
        }
    };

    // slot L holds the output of level L-1 (slot 0: the loaded plane); p* = previous
    // iteration, c* = this iteration; coefficients and loss factors by plane age
    float pEx[T + 1][R], pEy[T + 1][R], pEz[T + 1][R], pHx[T + 1][R], pHy[T + 1][R], pHz[T + 1][R];
    float cEx[T + 1][R], cEy[T + 1][R], cEz[T + 1][R], cHx[T + 1][R], cHy[T + 1][R], cHz[T + 1][R];
    float pP[12][T + 1][R], cP[12][T + 1][R];
    float aX[T + 1][R], aY[T + 1][R], aZ[T + 1][R], bX[T + 1][R], bY[T + 1][R], bZ[T + 1][R];
#pragma unroll
    for (int L = 0; L <= T; ++L)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pEx[L][r] = pEy[L][r] = pEz[L][r] = pHx[L][r] = pHy[L][r] = pHz[L][r] = 0.f;
            cEx[L][r] = cEy[L][r] = cEz[L][r] = cHx[L][r] = cHy[L][r] = cHz[L][r] = 0.f;
            aX[L][r] = aY[L][r] = aZ[L][r] = bX[L][r] = bY[L][r] = bZ[L][r] = 0.f;
#pragma unroll
            for (int q = 0; q < 12; ++q) pP[q][L][r] = cP[q][L][r] = 0.f;
        }
    TB3Plane<R> nxt;
    load(ps, nxt);
    for (int p = ps; p < pe; ++p) {
        // ---- slot 0: plane p of step n (loaded one iteration ahead)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cEx[0][r] = nxt.ex[r]; cEy[0][r] = nxt.ey[r]; cEz[0][r] = nxt.ez[r];
            cHx[0][r] = nxt.hx[r]; cHy[0][r] = nxt.hy[r]; cHz[0][r] = nxt.hz[r];
            aX[0][r] = nxt.cx[r]; aY[0][r] = nxt.cy[r]; aZ[0][r] = nxt.cz[r];
            bX[0][r] = nxt.dx[r]; bY[0][r] = nxt.dy[r]; bZ[0][r] = nxt.dz[r];
#pragma unroll
            for (int q = 0; q < 12; ++q) cP[q][0][r] = nxt.p[q][r];
        }
        if (p + 1 < pe) load(p + 1, nxt);                  // in flight while plane p runs
        lds[tb3_eidx<T>(p & 1, 0, w, 0, lane)] = cEx[0][0];
        lds[tb3_eidx<T>(p & 1, 0, w, 1, lane)] = cEz[0][0];
        // ---- levels
#pragma unroll
        for (int s = 0; s < T; ++s) {
            const int q = p - 1 - s;
            const bool act = q >= ps && q < NX;            // uniform over the workgroup
            const bool qin1 = !PX || q < NX - 1;           // Ex, Hy, Hz exist on plane q
            const bool qcurl = !PX || (q >= 1 && q <= NX - 2);
            int xHy = -1, xHz = -1, xEy = -1, xEz = -1;
            PV xh{0.f, 0.f, 0.f}, xe{0.f, 0.f, 0.f};
            if (PX && act) {
                xslabs(q, xHy, xHz, xEy, xEz);
                if (xHy >= 0 || xHz >= 0) xh = xprof(xhl, a.ph[0], q);
                if (xEy >= 0 || xEz >= 0) xe = xprof(xel, a.pe[0], q);
            }
            // H(q) of step n+s+1/2
            float hxN[R], hyN[R], hzN[R];
            float exJ = 0.f, ezJ = 0.f;                    // Ex, Ez of the row above the last row
            if (act && w + 1 < kTB3W) {
                exJ = lds[tb3_eidx<T>((p - 1) & 1, s, w + 1, 0, lane)];
                ezJ = lds[tb3_eidx<T>((p - 1) & 1, s, w + 1, 1, lane)];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                hxN[r] = hyN[r] = hzN[r] = 0.f;
#pragma unroll
                for (int qq = 0; qq < 6; ++qq) cP[qq][s + 1][r] = pP[qq][s][r];
                if (!act || !juse[r]) continue;
                const float ex_o = pEx[s][r], ey_o = pEy[s][r], ez_o = pEz[s][r];
                const float ex_k1 = lane_next0(ex_o), ey_k1 = lane_next0(ey_o); // k+1 (lane 63 is halo)
                const float ex_j1 = r + 1 < R ? pEx[s][r + 1 < R ? r + 1 : r] : exJ;   // j+1
                const float ez_j1 = r + 1 < R ? pEz[s][r + 1 < R ? r + 1 : r] : ezJ;
                if (kz1 && jh[r]) {                                             // Hx
                    const float d1 = ey_k1 - ey_o;
                    const float d2 = ez_j1 - ez_o;
                    float h = pHx[s][r] + ch * (d1 - d2);
                    if (PZ && zHx >= 0) h = h + ch * psi_step(cP[0][s + 1][r], zh, d1);
                    if (PY && yHx[r] >= 0) h = h + (-ch) * psi_step(cP[1][s + 1][r], yh[r], d2);
                    hxN[r] = h;
                }
                if (kz1 && qin1) {                                              // Hy
                    const float d1 = cEz[s][r] - ez_o;                          // Ez(q+1) - Ez(q)
                    const float d2 = ex_k1 - ex_o;
                    float h = pHy[s][r] + ch * (d1 - d2);
                    if (PX && xHy >= 0) h = h + ch * psi_step(cP[2][s + 1][r], xh, d1);
                    if (PZ && zHy >= 0) h = h + (-ch) * psi_step(cP[3][s + 1][r], zh, d2);
                    hyN[r] = h;
                }
                if (kuse && jh[r] && qin1) {                                    // Hz
                    const float d1 = ex_j1 - ex_o;
                    const float d2 = cEy[s][r] - ey_o;                          // Ey(q+1) - Ey(q)
                    float h = pHz[s][r] + ch * (d1 - d2);
                    if (PY && yHz[r] >= 0) h = h + ch * psi_step(cP[4][s + 1][r], yh[r], d1);
                    if (PX && xHz >= 0) h = h + (-ch) * psi_step(cP[5][s + 1][r], xh, d2);
                    hzN[r] = h;
                }
            }
            if (act) {
                lds[tb3_hidx<T>(s, w, 0, lane)] = hxN[R - 1];
                lds[tb3_hidx<T>(s, w, 1, lane)] = hzN[R - 1];
            }
            __syncthreads();
            // E(q) of step n+s+1
            float hxJ = 0.f, hzJ = 0.f;                    // Hx, Hz of the row below the first row
            if (act && w >= 1) {
                hxJ = lds[tb3_hidx<T>(s, w - 1, 0, lane)];
                hzJ = lds[tb3_hidx<T>(s, w - 1, 1, lane)];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float exn = 0.f, eyn = 0.f, ezn = 0.f;
#pragma unroll
                for (int qq = 6; qq < 12; ++qq) cP[qq][s + 1][r] = pP[qq][s][r];
                const float hx = hxN[r], hy = hyN[r], hz = hzN[r];
                const float hx_km = lane_prev0(hx), hy_km = lane_prev0(hy);   // k-1 (lane 0 is halo)
                if (act && juse[r]) {
                    const int j = jr[r];
                    const float hx_jm = r > 0 ? hxN[r > 0 ? r - 1 : 0] : hxJ;   // j-1
                    const float hz_jm = r > 0 ? hzN[r > 0 ? r - 1 : 0] : hzJ;
                    const float hy_im = pHy[s + 1][r], hz_im = pHz[s + 1][r]; // plane q-1
                    if (kuse && qin1) {                                         // Ex
                        float v = pEx[s][r];
                        if (lx_) v = v * bX[s + 1][r];
                        if (ycurl[r] && kint) {
                            const float d1 = hz - hz_jm;
                            const float d2 = hy - hy_km;
                            const float ce = aX[s + 1][r];
                            v = v + ce * (d1 - d2);
                            if (PY && yEx[r] >= 0) v = v + ce * psi_step(cP[6][s + 1][r], ye[r], d1);
                            if (PZ && zEx >= 0) v = v + (-ce) * psi_step(cP[7][s + 1][r], ze, d2);
                        }
                        if (PY && ((a.pec[2] && j == 0) || (a.pec[3] && j == NY - 1))) v = 0.0f;
                        if (PZ && ((a.pec[4] && k == 0) || (a.pec[5] && k == NZ - 1))) v = 0.0f;
                        if (EXC) v = apply_drive(v, LF_EX, q, j * NZ + k, x, a.st[s]);
                        exn = v;
                    }
                    if (kuse && jh[r]) {                                        // Ey
                        float v = pEy[s][r];
                        if (ly_) v = v * bY[s + 1][r];
                        if (qcurl && kint) {
                            const float d1 = hx - hx_km;
                            const float d2 = hz - hz_im;
                            const float ce = aY[s + 1][r];
                            v = v + ce * (d1 - d2);
                            if (PZ && zEy >= 0) v = v + ce * psi_step(cP[8][s + 1][r], ze, d1);
                            if (PX && xEy >= 0) v = v + (-ce) * psi_step(cP[9][s + 1][r], xe, d2);
                        }
                        if (PX && ((a.pec[0] && q == 0) || (a.pec[1] && q == NX - 1))) v = 0.0f;
                        if (PZ && ((a.pec[4] && k == 0) || (a.pec[5] && k == NZ - 1))) v = 0.0f;
                        const int plane = j * NZ + k;
                        if (EXC) v = apply_source(v, LF_EY, q, plane, x, a.st[s], fp64);
                        if (EXC) v = apply_drive(v, LF_EY, q, plane, x, a.st[s]);
                        eyn = v;
                    }
                    if (kz1 && juse[r]) {                                       // Ez
                        float v = pEz[s][r];
                        if (lz_) v = v * bZ[s + 1][r];
                        if (qcurl && ycurl[r]) {
                            const float d1 = hy - hy_im;
                            const float d2 = hx - hx_jm;
                            const float ce = aZ[s + 1][r];
                            v = v + ce * (d1 - d2);
                            if (PX && xEz >= 0) v = v + ce * psi_step(cP[10][s + 1][r], xe, d1);
                            if (PY && yEz[r] >= 0) v = v + (-ce) * psi_step(cP[11][s + 1][r], ye[r], d2);
                        }
                        if (PX && ((a.pec[0] && q == 0) || (a.pec[1] && q == NX - 1))) v = 0.0f;
                        if (PY && ((a.pec[2] && j == 0) || (a.pec[3] && j == NY - 1))) v = 0.0f;
                        const int plane = j * (NZ - 1) + k;
                        if (EXC) v = apply_source(v, LF_EZ, q, plane, x, a.st[s], fp64);
                        if (EXC) v = apply_drive(v, LF_EZ, q, plane, x, a.st[s]);
                        ezn = v;
                    }
                    // fused detector boxes: owned cells of this step (fdtd3d_core.py:477-481)
                    if (BOX && a.nboxm > 0 && q >= X0 && q < X1 && jown[r]) {
#pragma unroll
                        for (int m = 0; m < 2; ++m) {
                            if (m >= a.nboxm || !((a.mask[s] >> a.boxbit[m]) & 1u)) continue;
                            const float vv[3] = {exn, eyn, ezn};
                            double add = 0.0;
#pragma unroll
                            for (int cc = 0; cc < 3; ++cc) {
                                const bool in = kown && q >= a.blo[m][cc][0] && q < a.bhi[m][cc][0] &&
                                                j >= a.blo[m][cc][1] && j < a.bhi[m][cc][1] &&
                                                k >= a.blo[m][cc][2] && k < a.bhi[m][cc][2];
                                if (in) add += (double)(vv[cc] * vv[cc]);
                            }
                            bacc[m * 64] += add;            // this lane's own LDS slot
                        }
                    }
                }
                cEx[s + 1][r] = exn; cEy[s + 1][r] = eyn; cEz[s + 1][r] = ezn;
                cHx[s + 1][r] = hx; cHy[s + 1][r] = hy; cHz[s + 1][r] = hz;
            }
            if (s + 1 < T) {                               // first row's Ex, Ez for level s+1
                lds[tb3_eidx<T>(p & 1, s + 1 < T ? s + 1 : 0, w, 0, lane)] = cEx[s + 1][0];
                lds[tb3_eidx<T>(p & 1, s + 1 < T ? s + 1 : 0, w, 1, lane)] = cEz[s + 1][0];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                pEx[s][r] = cEx[s][r]; pEy[s][r] = cEy[s][r]; pEz[s][r] = cEz[s][r];
                pHx[s][r] = cHx[s][r]; pHy[s][r] = cHy[s][r]; pHz[s][r] = cHz[s][r];
#pragma unroll
                for (int qq = 0; qq < 12; ++qq) pP[qq][s][r] = cP[qq][s][r];
            }
        }
        // ---- store plane p-T of step n+T (owned cells)
        {
            const int q = p - T;
            if (q >= X0 && q < X1) {
                const bool qin1 = !PX || q < NX - 1;
                const bool qcurl = !PX || (q >= 1 && q <= NX - 2);
                int xHy, xHz, xEy, xEz;
                xslabs(q, xHy, xHz, xEy, xEz);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (!jown[r] || !kown) continue;
                    const unsigned c = off(q, jr[r]);
                    if (qin1) a.Ex1[c] = cEx[T][r];
                    if (jh[r]) a.Ey1[c] = cEy[T][r];
                    if (kz1) a.Ez1[c] = cEz[T][r];
                    if (kz1 && jh[r]) a.Hx1[c] = cHx[T][r];
                    if (kz1 && qin1) a.Hy1[c] = cHy[T][r];
                    if (jh[r] && qin1) a.Hz1[c] = cHz[T][r];
#pragma unroll
                    for (int qq = 0; qq < 12; ++qq) {
                        if (!((qq == 0 || qq == 3 || qq == 7 || qq == 8) ? PZ
                              : (qq == 1 || qq == 4 || qq == 6 || qq == 11) ? PY : PX))
                            continue;
                        const int pos = psi_on(qq, r, q, qin1, qcurl, xHy, xHz, xEy, xEz);
                        if (pos >= 0)
                            a.s1[qq].psi[psi_off32(a.s1[qq], psi_axis[qq], pos, q, jr[r], kc, G)] = cP[qq][T][r];
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                pEx[T][r] = cEx[T][r]; pEy[T][r] = cEy[T][r]; pEz[T][r] = cEz[T][r];
                pHx[T][r] = cHx[T][r]; pHy[T][r] = cHy[T][r]; pHz[T][r] = cHz[T][r];
            }
        }
#pragma unroll
        for (int L = T; L >= 1; --L)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                aX[L][r] = aX[L - 1][r]; aY[L][r] = aY[L - 1][r]; aZ[L][r] = aZ[L - 1][r];
                bX[L][r] = bX[L - 1][r]; bY[L][r] = bY[L - 1][r]; bZ[L][r] = bZ[L - 1][r];
            }
    }
}

// Persistent workgroups (one per CU) draw the tiles of the cost-ordered list from a
// ticket counter, so that the pass ends on short tiles whatever each tile costs.  The
// ticket is drawn by one thread, passed through LDS and read into a scalar register:
// the loop exit and every branch around a barrier are uniform over the workgroup, and
// no barrier stands under a branch, so all wavefronts meet the same barriers.
template <int T, bool LOSS>
__global__ void __launch_bounds__(1024) yee_tb_kernel(TB3Args a, Excite x, int fp64) {
    __shared__ float lds[3 * T * kTB3W * 128];
    __shared__ double red[kTB3W][2];
    __shared__ double boxacc[kTB3W][2][64];               // per-lane box sums of a tile
    __shared__ int tile_s;
    const TB3Box __attribute__((address_space(4))) *bt =
        (const TB3Box __attribute__((address_space(4))) *)a.box;
    if (threadIdx.x == 0) tile_s = (int)(atomicAdd(a.ticket, 1u) - a.base);
    __syncthreads();
    int t = __builtin_amdgcn_readfirstlane(tile_s);
    while (t < a.ntile) {
        int b = 0;
        while (b + 1 < a.nbox && t >= bt[b + 1].first) ++b;
        const TB3Box bx = *(const TB3Box *)&bt[b];
        const int u = t - bx.first;
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        double *bacc = &boxacc[w][0][lane];
        bacc[0] = 0.0;
        bacc[64] = 0.0;
        // tiles whose march meets a source or drive plane take the bodies with excitation
        // (and box sums); tiles whose planes meet a fused detector box the bodies with box
        // sums; the others carry neither (the excitation loads' waits would also drain
        // the plane prefetch, and both cost registers).  The slab-free class holds two
        // rows per wavefront unless it carries excitation.
#define LF_TB3_BODY(R, M, E, B) yee_tb_body<T, R, M, LOSS, E, B>(a, x, bx, u, lds, fp64, bacc)
#define LF_TB3_MODES(E, B)                                                                   \
    switch (bx.mode) {                                                                       \
        case 0:                                                                              \
            if (bx.rows == 2) LF_TB3_BODY(2, 0, E, B);                                       \
            else LF_TB3_BODY(1, 0, E, B);                                                    \
            break;                                                                           \
        case 1: LF_TB3_BODY(1, 1, E, B); break;                                              \
        case 2: LF_TB3_BODY(1, 2, E, B); break;                                              \
        case 3: LF_TB3_BODY(1, 3, E, B); break;                                              \
        case 4: LF_TB3_BODY(1, 4, E, B); break;                                              \
        case 5: LF_TB3_BODY(1, 5, E, B); break;                                              \
        case 6: LF_TB3_BODY(1, 6, E, B); break;                                              \
        default: LF_TB3_BODY(1, 7, E, B); break;                                             \
    }
        if (bx.exc) {
            switch (bx.mode) {
                case 0: LF_TB3_BODY(1, 0, true, true); break;
                case 1: LF_TB3_BODY(1, 1, true, true); break;
                case 2: LF_TB3_BODY(1, 2, true, true); break;
                case 3: LF_TB3_BODY(1, 3, true, true); break;
                case 4: LF_TB3_BODY(1, 4, true, true); break;
                case 5: LF_TB3_BODY(1, 5, true, true); break;
                case 6: LF_TB3_BODY(1, 6, true, true); break;
                default: LF_TB3_BODY(1, 7, true, true); break;
            }
        } else if (bx.box) {
            LF_TB3_MODES(false, true)
        } else {
            LF_TB3_MODES(false, false)
        }
#undef LF_TB3_MODES
#undef LF_TB3_BODY
        // tile sums of the fused detector boxes: fixed order, one slot per tile.  The
        // barriers run for every tile (no barrier under a branch); the second one keeps
        // the next tile's writes to red[] behind this tile's reads.
        double bacc0 = bacc[0], bacc1 = bacc[64];
        for (int o = 32; o > 0; o >>= 1) {
            bacc0 += __shfl_xor(bacc0, o);
            bacc1 += __shfl_xor(bacc1, o);
        }
        if (lane == 0) { red[w][0] = bacc0; red[w][1] = bacc1; }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (a.nboxm > 0) {
                double s0 = 0.0, s1 = 0.0;
                for (int q = 0; q < kTB3W; ++q) { s0 += red[q][0]; s1 += red[q][1]; }
                a.bpart[0][t] += s0;
                if (a.nboxm > 1) a.bpart[1][t] += s1;
            }
            tile_s = (int)(atomicAdd(a.ticket, 1u) - a.base);   // next tile
        }
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(tile_s);
    }
}

}  // namespace

#define LF_CAT2(a, b, c) a##b##_##c
#define LF_CAT(a, b, c) LF_CAT2(a, b, c)
void LF_CAT(launch_yee_tb3_, LF_TB3_T, LF_TB3_LOSS)(const TB3Args &f, const Excite &x, int fp64,
                                                    unsigned nb, hipStream_t st) {
    yee_tb_kernel<LF_TB3_T, LF_TB3_LOSS != 0><<<nb, 64 * kTB3W, 0, st>>>(f, x, fp64);
}

}  // namespace lfd
