// tb2d_plan.h - the wavefront plan of the temporally blocked 2D passes (host code, no
// device dependencies, so that a CPU test can check it: every owned cell covered once,
// the interior bodies' assumptions met).
#pragma once
#include <algorithm>
#include <tuple>
#include <utility>
#include <vector>

namespace lfd {

// what the plan needs of a 2D context: grid, slab bounds (planes / lanes v < lo or
// v >= hi lie in a slab; lo = 0, hi = n where an axis has none) and excitation planes
struct TBPlanIn {
    int nx, ny;
    int xlo, xhi, ylo, yhi;
    int src_ix;               // first of the two source planes of the stepped component, or -1
    int drv_n;                // planes 0 .. drv_n-1 carry the stitch drive, or 0
};

// Wavefront plan of one temporally blocked 2D pass: up to four classes, each a list of
// at most four rectangles, a rectangle being an x range cut into chunks of its own
// length times a range of y waves.  The slab-free x range [A, B) holds the chunks whose
// loaded planes lie in 1..Nx-2 and outside the x slabs, the slab-free rows the waves
// whose lanes lie in 1..Ny-2 and outside the y slabs; the chunk grid starts at A, so the
// x-slab chunks cover only the planes within T of a slab or an edge.  A chunk whose march
// comes within T planes of a source or drive plane needs the body with excitation code
// (the stitch drive covers the whole 2048-plane window overlap, a quarter of a window);
// the others take the body without it, which holds fewer registers and no excitation
// loads.
//   class 0: MODE 0, slab-free chunks without excitation (chunks of lx)
//   class 1: MODE 1, y-slab rows over [A, B) (chunks of lx/4)
//   class 2: MODE 3, x-slab chunks over all rows (chunks of lx/16: the finest), or
//            every wavefront of a grid without a slab-free rectangle (with excitation);
//            its chunks are at most 64 - 2T planes long, so that a wave's marched planes
//            fit its 64 lanes (the x profiles are held one plane per lane)
//   class 3: MODE 0, slab-free chunks with excitation (chunks of lx/2: their waves
//            carry the excitation work, and few long marches would leave SIMDs idle;
//            with the 2048-plane stitch drive lx/2 measured 3 % faster per driven
//            window than lx/4, lx/8 4 % slower, lx 2 % slower)
// The slab classes choose the body per rectangle when only some of their rectangles
// meet an excitation plane (their bodies with and without excitation code hold the same
// number of waves per SIMD, or nearly).  When the excitation planes scatter so that a
// class would need more than four rectangles, the slab classes take the excitation code
// over all their chunks.
struct TBClass {
    int mode, nseg;
    int exc;                  // 0: no rectangle meets an excitation plane, 1: all, 2: some
    int x0[4], x1[4], lx[4];  // per rectangle: x range, chunk length
    int wy0[4], nwy[4], end[4], sexc[4];
};
constexpr int kTBClasses = 4;
constexpr int kTBSerial = 4;      // lf_set_tb_lead value: every class on the context stream
struct TBPlan {
    TBClass cls[kTBClasses];
    bool interior;
};

inline TBPlan tb_plan_2d(const TBPlanIn &in, int T, int lx) {
    const int W = 64 - 2 * T;
    const int Nx = in.nx, Ny = in.ny;
    const int nwy = (Ny + W - 1) / W;
    const int xlo = in.xlo, xhi = in.xhi, ylo = in.ylo, yhi = in.yhi;
    int iy0 = nwy, iy1 = -1;
    for (int wy = 0; wy < nwy; ++wy) {
        const int j0 = wy * W - T;
        if (j0 >= std::max(ylo, 1) && j0 + 63 < std::min(yhi, Ny - 2) + 1) {
            iy0 = std::min(iy0, wy); iy1 = std::max(iy1, wy);
        }
    }
    // slab-free x range: a chunk [X0, X1) inside [A, B) loads planes X0-T .. X1+T-1, all
    // in 1..Nx-2 and outside the x slabs; the chunk grid starts at A, so the x-slab
    // chunks cover only the planes within T of a slab or edge
    const int A = std::max(xlo, 1) + T, B = std::min(xhi, Nx - 2) - T + 1;
    auto div = [&](int want) {
        for (int cand : {want, 8, 4, 2})
            if (cand <= want && lx % cand == 0 && lx / cand >= 4) return cand;
        return 1;
    };
    // planes whose chunks' marches can meet a source or drive plane: T around each
    std::vector<std::pair<int, int>> zones;
    if (in.src_ix >= 0) zones.push_back({in.src_ix - T, in.src_ix + 2 + T});
    if (in.drv_n > 0) zones.push_back({-T, in.drv_n + T});
    // [u0, u1) in runs of equal excitation (zone edges split it)
    auto runs_of = [&](int u0, int u1) {
        std::vector<int> cut = {u0, u1};
        for (auto &z : zones)
            for (int e : {z.first, z.second})
                if (e > u0 && e < u1) cut.push_back(e);
        std::sort(cut.begin(), cut.end());
        cut.erase(std::unique(cut.begin(), cut.end()), cut.end());
        std::vector<std::tuple<int, int, bool>> out;
        for (size_t i = 0; i + 1 < cut.size(); ++i) {
            const int a = cut[i], b = cut[i + 1];
            bool e = false;
            for (auto &z : zones) e = e || (a < z.second && b > z.first);
            if (!out.empty() && std::get<2>(out.back()) == e) std::get<1>(out.back()) = b;
            else out.push_back({a, b, e});
        }
        return out;
    };
    auto build = [&](bool split_slabs, TBPlan &pl) -> bool {
        pl = TBPlan{};
        pl.interior = B > A && iy1 >= iy0;
        const int modes[kTBClasses] = {0, 1, 3, 0};
        for (int u = 0; u < kTBClasses; ++u) pl.cls[u].mode = modes[u];
        bool ok = true;
        auto add = [&](TBClass &q, int x0, int x1, int L, int wy0, int nwy_s, bool e) {
            if (x1 <= x0 || nwy_s <= 0) return;
            if (q.nseg == 4) { ok = false; return; }
            const int start = q.nseg ? q.end[q.nseg - 1] : 0;
            const int g = q.nseg++;
            q.x0[g] = x0; q.x1[g] = x1; q.lx[g] = L;
            q.wy0[g] = wy0; q.nwy[g] = nwy_s; q.sexc[g] = e ? 1 : 0;
            q.end[g] = start + (x1 - x0 + L - 1) / L * nwy_s;
        };
        auto finish = [&]() {
            for (TBClass &q : pl.cls) {
                int n = 0;
                for (int g = 0; g < q.nseg; ++g) n += q.sexc[g];
                q.exc = n == 0 ? 0 : (n == q.nseg ? 1 : 2);
            }
            return ok;
        };
        // a slab class over [u0, u1) and rows [wy0, wy0 + nwy_s): one rectangle per run
        // (rectangles with excitation take chunks of Lx: their waves run the heavier body)
        auto slab = [&](int cls, int u0, int u1, int L, int Lx, int wy0, int nwy_s) {
            if (!split_slabs) { add(pl.cls[cls], u0, u1, L, wy0, nwy_s, true); return; }
            for (auto &r : runs_of(u0, u1)) {
                const bool e = std::get<2>(r);
                add(pl.cls[cls], std::get<0>(r), std::get<1>(r), e ? Lx : L, wy0, nwy_s, e);
            }
        };
        if (!pl.interior) {
            add(pl.cls[2], 0, Nx, std::min(lx / div(4), W), 0, nwy, true);
            return finish();
        }
        const int ni = iy1 - iy0 + 1;
        for (auto &r : runs_of(A, B)) {                  // slab-free chunks: classes 0 / 3
            if (std::get<2>(r)) add(pl.cls[3], std::get<0>(r), std::get<1>(r), lx / div(2), iy0, ni, true);
            else add(pl.cls[0], std::get<0>(r), std::get<1>(r), lx, iy0, ni, false);
        }
        slab(1, A, B, lx / div(4), lx / div(16), 0, iy0);   // y-slab rows: class 1
        slab(1, A, B, lx / div(4), lx / div(16), iy1 + 1, nwy - 1 - iy1);
        const int L3 = std::min(lx / div(16), W);           // x-slab chunks: class 2
        slab(2, 0, A, L3, L3, 0, nwy);
        slab(2, B, Nx, L3, L3, 0, nwy);
        return finish();
    };
    TBPlan pl;
    if (!build(true, pl)) build(false, pl);
    return pl;
}

// Launch arguments of a class (A: the pass argument struct, seg_* fields of up to eight
// rectangles): its rectangles in order, seg_exc bit g for a rectangle with excitation code.
template <typename A>
void set_class(A &q, const TBClass &k) {
    q.nseg = k.nseg;
    q.seg_exc = 0u;
    q.seg_m3 = 0u;
    for (int u = 0; u < 4; ++u) {
        q.seg_x0[u] = k.x0[u]; q.seg_x1[u] = k.x1[u]; q.seg_lx[u] = k.lx[u];
        q.seg_wy0[u] = k.wy0[u]; q.seg_nwy[u] = k.nwy[u]; q.seg_end[u] = k.end[u];
        if (u < k.nseg) q.seg_exc |= (unsigned)(k.sexc[u] != 0) << u;
    }
}

// the y-slab class (MODE 1) then the x-slab class (MODE 3) as the rectangles of one launch
template <typename A>
void set_slabs(A &q, const TBClass &k1, const TBClass &k3) {
    set_class(q, k1);
    const int base = k1.nseg ? k1.end[k1.nseg - 1] : 0;
    for (int u = 0; u < k3.nseg; ++u) {
        const int g = k1.nseg + u;
        q.seg_x0[g] = k3.x0[u]; q.seg_x1[g] = k3.x1[u]; q.seg_lx[g] = k3.lx[u];
        q.seg_wy0[g] = k3.wy0[u]; q.seg_nwy[g] = k3.nwy[u]; q.seg_end[g] = base + k3.end[u];
        q.seg_exc |= (unsigned)(k3.sexc[u] != 0) << g;
        q.seg_m3 |= 1u << g;
    }
    q.nseg = k1.nseg + k3.nseg;
}


}  // namespace lfd
